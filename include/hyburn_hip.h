/*
 * hyburn_hip.h -- C ABI of the MI355X (gfx950) batched neural-mass integrator.
 *
 * This is the drop-in boundary for hyburn's batched integration path: the
 * `--backend hip` counterpart of BatchHybridEngine<B: Backend>
 * (reference: src/engine/batch_engine/engine.rs:145-860, src/cli.rs:259-271).
 * Plain C types only: pointers, sizes, enums, integer error codes.
 *
 * Units follow the reference: dt and sim_length in ms, BOLD TR in seconds
 * in hb_sim_config (config.rs:219-250) and in ms in hb_instance.
 * State layout is node-major [n_nodes][nvar] per instance (engine.rs:161-170);
 * batched outputs are [n_sweep][n_nodes][nvar] (engine.rs:15-31).
 */
#ifndef HYBURN_HIP_H
#define HYBURN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: mirror SimulationError (src/error.rs:7-23). */
#define HB_OK 0
#define HB_INVALID_CONFIG 1
#define HB_INVALID_STATE 2
#define HB_BACKEND_ERROR 3
#define HB_IO_ERROR 4

/* Models (config.rs:43-50 registry order of the six batch-native models). */
enum { HB_G2DO = 0, HB_MPR = 1, HB_JANSEN_RIT = 2, HB_WILSON_COWAN = 3, HB_KURAMOTO = 4, HB_RWW = 5,
       /* only on the multi-subnetwork path (hb_run_multi) */
       HB_LINEAR = 6, HB_SUP_HOPF = 7, HB_HOPFIELD = 8, HB_COOMBES_BYRNE_2D = 9,
       /* the models the reference pins with trace fixtures (tests/model_reference_test.rs:90-156),
        * multi-subnetwork path only: dfun.rs:952-1870 */
       HB_EPILEPTOR_CODIM3 = 10, HB_EPILEPTOR_CODIM3_SLOWMOD = 11, HB_EPILEPTOR_RS = 12,
       HB_ZERLAUT_FIRST = 13, HB_ZERLAUT_SECOND = 14, HB_KIONEX = 15,
       /* the remaining models of the registry (config.rs:56-75), multi-subnetwork path only:
        * dfun.rs:609-950, 1212-1400 */
       HB_COOMBES_BYRNE = 16, HB_GAST_SD = 17, HB_GAST_SF = 18, HB_LARTER_BREAKSPEAR = 19,
       HB_EPILEPTOR_2D = 20, HB_EPILEPTOR = 21, HB_RWW_EXC_INH = 22, HB_DECO_BALANCED = 23,
       HB_ZETTERBERG_JANSEN = 24, HB_REDUCED_FHN = 25, HB_REDUCED_HR = 26, HB_DUMONT_GUTKIN = 27 };
/* Largest parameter vector of any model (the Zerlaut models: 50). */
#define HB_MAX_PARAMS 64
/* IntegratorKind (integrator.rs:17-30). */
enum { HB_EULER = 0, HB_HEUN = 1, HB_RK4 = 2, HB_EULER_STOCH = 3, HB_HEUN_STOCH = 4, HB_RK4_STOCH = 5 };
/* CouplingFnConfig (coupling.rs:53-71). */
enum { HB_CF_LINEAR = 0, HB_CF_DIFFERENCE = 1, HB_CF_KURAMOTO = 2, HB_CF_SIGMOIDAL = 3,
       HB_CF_SCALED_LINEAR = 4, HB_CF_TANH = 5, HB_CF_SJR = 6, HB_CF_PRESIGMOIDAL = 7 };
/* Weight kinds. SCALAR_MEAN: w * mean_j pre(x_j) (batch engine, engine.rs:602-608);
 * SCALAR_DIAG: w * I (plain engine, construction.rs:639-655). */
enum { HB_W_NONE = 0, HB_W_SCALAR_MEAN = 1, HB_W_SCALAR_DIAG = 2, HB_W_DENSE = 3, HB_W_CSR = 4 };
/* Delay semantics. PER_EDGE: runtime.rs:162-204; SINGLE: one delay per projection
 * with the ring semantics of runtime.rs:214-234; BATCH_COMPAT: the sliding-window
 * read of engine.rs:541-589. */
enum { HB_DELAY_PER_EDGE = 0, HB_DELAY_SINGLE = 1, HB_DELAY_BATCH_COMPAT = 2 };
/* Engine semantics where the two reference engines differ. */
enum { HB_ENGINE_BATCH = 0, HB_ENGINE_PLAIN = 1 };

/* Run options (bit flags). */
#define HB_FLAG_TRAJECTORY 0x1u   /* record [n_steps][n_sweep][n_nodes][nvar] (engine.rs:487-494) */
#define HB_FLAG_EXACT 0x2u        /* no FMA contraction: same f32 rounding as the written expressions */
#define HB_FLAG_NO_GRAPH 0x4u     /* launch directly instead of replaying a captured hipGraph */
#define HB_FLAG_NO_BLOCKING 0x8u  /* disable the delay-blocked schedule (one barrier per step) */
#define HB_FLAG_AUTOTUNE 0x10u    /* time candidate launch shapes per shape class before planning */
#define HB_FLAG_SEPARATE_BOLD 0x20u /* adjoint: BOLD monitor forward in its own kernel, not in helper waves */
#define HB_FLAG_KEEP_NOISE 0x40u  /* plans: generate the noise at the first hb_plan_run only; later runs
                                     reuse it (pre-generated noise resident in HBM, as an input) */
#define HB_FLAG_RESTART 0x80u     /* hb_run_sweep*: start from the configured initial state at step 0
                                     instead of continuing the engine's previous run */
#define HB_FLAG_BOLD_BATCH_MEAN 0x100u /* hb_run_sweep: one BOLD monitor fed var0 averaged over the sweep
                                     points (engine.rs:785-801) instead of one monitor per point */
#define HB_FLAG_NO_MFMA 0x200u    /* dense coupling: keep the per-lane row sums instead of the matrix-core
                                     product over 16-instance tiles */

/* ---------------- config-level API (BatchHybridEngine mirror) ---------------- */

typedef struct {
    int32_t model;        /* HB_G2DO ... */
    int32_t n_nodes;
    int32_t n_params;
    const float* params;  /* model parameters, reference order (Appendix of model files) */
    const float* x0;      /* [n_nodes][nvar] node-major initial state */
} hb_subnetwork;

typedef struct {
    int32_t wkind;        /* HB_W_* */
    float scalar_w;
    const float* dense;   /* [n_nodes][n_nodes], row = target (config.rs:208-216) */
    const int32_t* indptr;   /* CSR by target row, length n_nodes + 1 */
    const int32_t* indices;
    const float* data;
    int32_t nnz;
    const int32_t* delays;   /* per-edge (n_delays == nnz, CSR) or one per projection */
    int32_t n_delays;
    int32_t cfn;          /* HB_CF_* */
    int32_t n_cparams;
    float cparams[8];
} hb_projection;

typedef struct {
    double sim_length;    /* ms */
    double dt;            /* ms */
    int32_t integrator;
    int32_t n_nsig;       /* 0, 1 (scalar) or nvar */
    const float* nsig;
    const hb_subnetwork* subnetworks;
    int32_t n_subnetworks;   /* 1 */
    const hb_projection* projections;
    int32_t n_projections;   /* 0 or 1 (a self-projection 0 -> 0) */
    int32_t bold_period;     /* 0: no BOLD monitor */
    double bold_tr_s;
    int32_t engine;          /* HB_ENGINE_BATCH (BatchHybridEngine) or HB_ENGINE_PLAIN (HybridEngine) */
    int32_t delay_mode;      /* HB_DELAY_* for a CSR projection with per-edge delays */
    uint64_t noise_seed;     /* Philox4x32-10 key for the pre-generated noise */
    int32_t noise_shared;    /* 1: one noise stream broadcast over the batch (integrator.rs:170-198) */
} hb_sim_config;

/* One dimension of a sweep (engine.rs:34-53). kind 0: subnetworks[a].params[b];
 * kind 1: projections[a].coupling_params[b] (an extension: the reference cannot
 * sweep coupling parameters). */
typedef struct {
    int32_t kind;
    int32_t a;
    int32_t b;
    const float* values;
    int32_t n_values;
} hb_sweep_point;

typedef struct {
    int32_t n_sweep, n_nodes, nvar, n_steps;
    float* tavg;             /* [n_sweep][n_nodes][nvar] */
    float* final_states;     /* [n_sweep][n_nodes][nvar] */
    float* trajectories;     /* [n_steps][n_sweep][n_nodes][nvar] or NULL */
    float* bold;             /* [n_sweep][n_bold][n_nodes] or NULL (per instance) */
    int32_t n_bold;
    double elapsed_ms;       /* wall time of the run call (engine.rs:503,825) */
    double kernel_ms;        /* integration kernel time measured with HIP events */
    int32_t bold_series;     /* BOLD series in bold: n_sweep (one monitor per point) or 1 (batch mean) */
    int64_t step;            /* the engine's step counter after the run (self.step) */
} hb_result;

typedef struct hb_engine hb_engine;

/* BatchHybridEngine::from_config (engine.rs:145-149). device = HIP ordinal. */
int hb_engine_create(const hb_sim_config* cfg, int32_t n_sweep, int32_t device, hb_engine** out);
void hb_engine_destroy(hb_engine* e);
/* BatchHybridEngine::run_sweep / run_sweep_with_trajectory (engine.rs:477-494).
 * n_values must equal n_sweep (checked; the reference broadcasts unchecked).
 * Like the reference engine (engine.rs:496-540), a run continues from where
 * the previous one stopped: states, step counter, delay histories and BOLD
 * monitors carry over (HB_FLAG_RESTART starts afresh); tavg averages this
 * run's steps; bold holds every BOLD sample since the engine started,
 * resampled to TR (BoldMonitor::flush, bold_monitor.rs:210-227), as
 * [bold_series][n_bold][n_nodes]. */
int hb_run_sweep(hb_engine* e, int32_t sub_idx, int32_t param_idx, const float* values,
                 int32_t n_values, int64_t n_steps, uint32_t flags, hb_result* out);
/* BatchHybridEngine::run_sweep_multi (engine.rs:845-860): Cartesian product,
 * last dimension fastest. As the reference does (engine.rs:886-909,
 * 1326-1331), every point starts from sweep point 0 of the engine's current
 * state, at step 0, without BOLD monitors; the engine then holds the run's
 * states, step and histories (and no monitors until HB_FLAG_RESTART). */
int hb_run_sweep_multi(hb_engine* e, const hb_sweep_point* points, int32_t n_points,
                       int64_t n_steps, uint32_t flags, hb_result* out);
void hb_result_free(hb_result* r);

/* What an engine carries from one run to the next, for checkpoints. */
typedef struct {
    int64_t step;            /* steps integrated so far (self.step) */
    double dt;
    int32_t integrator;      /* HB_EULER .. HB_RK4_STOCH */
    int32_t n_state;         /* sweep points held; 0: the configured initial state (one point), step 0 */
    int32_t n_nodes, nvar;   /* read-only: the engine's shape */
    int64_t hist_floats;     /* delay history per point ([max_delay][n_nodes][channels], oldest
                                first, hb_history_floats); 0 without delays */
} hb_engine_state;

/* HybridEngine::checkpoint (runtime.rs:519-561) saves step, dt, integrator,
 * nsig and the states; this also returns each point's delay history, so that
 * a resumed run continues bit for bit. Call with null buffers for the sizes,
 * then again with nsig (nvar floats), states (max(n_state, 1) x n_nodes x
 * nvar, node-major) and hist (n_state x hist_floats; ignored when either
 * is 0). */
int hb_engine_get_state(const hb_engine* e, hb_engine_state* st, float* nsig, float* states, float* hist);
/* HybridEngine::resume (runtime.rs:576-742): restores step, dt, integrator,
 * nsig (n_nsig 1 or nvar; 0 keeps the engine's) and the states. n_state 1 on
 * an engine of several sweep points continues every point from that one
 * state. hist null: a zero history, which is what a reference checkpoint
 * (states only) resumes into; else n_state x st->hist_floats floats, and
 * st->hist_floats must match the engine's. BOLD monitors are left as they
 * are, as the reference's resume leaves its monitors. */
int hb_engine_set_state(hb_engine* e, const hb_engine_state* st, const float* nsig, int32_t n_nsig,
                        const float* states, const float* hist);

/* ---------------- instance-level API (ragged batches) ---------------- */

/* One fully specified simulation. Arrays are host memory, read at plan time. */
typedef struct {
    int32_t model, integrator, engine, n_nodes, n_steps;
    float dt;                 /* ms, f32 as the engine holds it (engine.rs:456) */
    double dt_ms;             /* ms, f64 config value (BOLD window times) */
    const float* params;
    int32_t sweep_form;       /* 1: batch-native swept form of the model's sweep index (dfun.rs:49-72) */
    float sweep_val;
    int32_t node_param_idx;   /* -1, or a parameter given per node */
    const float* node_param;  /* [n_nodes] */
    const float* x0;          /* [n_nodes][nvar] */
    const float* nsig;        /* [nvar] or NULL */
    const float* noise;       /* optional explicit raw N(0,1) noise [n_steps][n_nodes][nvar] */
    uint64_t noise_seed;      /* else Philox4x32-10(seed, noise_stream) */
    uint64_t noise_stream;
    int32_t wkind, cfn, n_cparams;
    float cparams[8];
    float scalar_w;
    const float* dense;       /* [n_nodes][n_nodes], row = target */
    const int32_t* indptr;
    const int32_t* indices;
    const float* weights;
    const int32_t* edge_delays;
    int32_t nnz;
    int32_t delay_mode;
    int32_t single_delay;
    int32_t max_delay_hint;   /* ring horizon - 1 if larger than the delays read (BATCH_COMPAT) */
    int32_t bold_period;      /* 0: off */
    double bold_tr_ms;
    /* outputs (host memory, caller-owned) */
    float* tavg;              /* [n_nodes][nvar] */
    float* final_state;       /* [n_nodes][nvar] */
    float* traj;              /* [n_steps][n_nodes][nvar] or NULL */
    float* bold;              /* [bold_cap][n_nodes] or NULL */
    int32_t bold_cap;
    int32_t* bold_n;
    /* continuation of an earlier run (all optional; zero for a fresh run) */
    int64_t step0;            /* steps integrated before this run: delays, the step-0 rule and the
                                 batch window read global steps step0 + t */
    float* history;           /* in/out pre() history, hb_history_floats(inst) floats, [max_delay][n_nodes]
                                 [channels] oldest first: read at plan time when step0 > 0 (the last
                                 max_delay steps before step0), written by hb_plan_download (those
                                 before the run's end) */
    float* bold_state;        /* in/out Balloon-Windkessel s, f, v, q per node [n_nodes][4]: read at
                                 plan time (the monitor's state at the run's start), written by
                                 hb_plan_download */
} hb_instance;

typedef struct {
    int32_t device;
    uint32_t flags;          /* HB_FLAG_EXACT, HB_FLAG_NO_GRAPH */
    int32_t block_size;      /* 0: autotuned */
    int32_t grid_waves;      /* persistent grid size in workgroups, 0: auto */
} hb_run_opts;

typedef struct {
    double kernel_ms;        /* integration kernels, HIP events on the launch stream */
    double total_ms;         /* plan run incl. noise generation, excl. uploads/downloads */
    double node_steps;       /* sum over instances of n_nodes * n_steps */
    int32_t n_launches;
    int32_t block_size;
    int32_t lds_bytes;       /* dynamic LDS per workgroup of the main launch */
} hb_run_stats;

typedef struct hb_plan hb_plan;

/* Upload a ragged set of instances, pack them by cost, pick launch shapes and
 * capture the launch sequence (noise fill + integration) in a hipGraph. */
int hb_plan_create(const hb_instance* inst, int32_t n, const hb_run_opts* opts, hb_plan** out);
/* Run the plan on device (inputs resident in HBM). Outputs stay on device. */
int hb_plan_run(hb_plan* p, hb_run_stats* stats);
/* Copy outputs to the host pointers given in the instances. */
int hb_plan_download(hb_plan* p);
void hb_plan_destroy(hb_plan* p);
/* One-call convenience: create, run, download, destroy. */
int hb_run_instances(const hb_instance* inst, int32_t n, const hb_run_opts* opts, hb_run_stats* stats);

/* Floats of an instance's carried delay history (hb_instance.history): max_delay
 * steps x n_nodes x the pre() channels its ring keeps (Kuramoto 2 per cvar,
 * SigmoidalJansenRit 1, MPR 1 when its cv coefficient is 0, else ncvar);
 * 0 without delays. */
int64_t hb_history_floats(const hb_instance* inst);

/* Philox4x32-10 standard normals (Box-Muller in f64), written as f32 to a host
 * buffer: element k of stream s under key seed. Device-generated. */
int hb_philox_normal(uint64_t seed, uint64_t stream, int64_t n, float* out_host, int32_t device);

/* ---------------- adjoint (gradients) ---------------- */

/* Loss of one instance. FC: sum_ij (FC_ij - target_ij)^2 with FC the Pearson
 * matrix of the native BOLD series (fc.rs:11-57), target [N][N]. BOLD_MSE:
 * mean over rows and nodes of (BOLD - target)^2, target [n_bold][N] native
 * rows. ORDER: (1/T) sum_t (R_t - target_t)^2, R_t = |mean_j exp(i x_{t+1}[j])|
 * (Kuramoto phases after step t), target [n_steps]. */
enum { HB_LOSS_FC = 0, HB_LOSS_BOLD_MSE = 1, HB_LOSS_ORDER = 2 };

typedef struct {
    int32_t loss_kind;
    const float* target;
    int32_t grad_param;       /* model parameter index differentiated per node, or -1 */
    /* outputs (host memory, caller-owned), written by hb_adjoint_download */
    float* loss;              /* [1] */
    float* grad_coupling;     /* [1] dL / d cparams[0] (the global coupling G / a) */
    float* grad_param_out;    /* [n_nodes] dL / d params[grad_param] of each node */
    float* grad_x0;           /* [n_nodes][nvar] or NULL */
    float* final_state;       /* [n_nodes][nvar] or NULL */
} hb_adjoint_spec;

typedef struct {
    int32_t device;
    int32_t checkpoint_every; /* steps per checkpoint segment; 0: the whole trajectory kept in HBM
                                 (no recompute) when every instance's fits a quarter of the free
                                 memory, else ceil(sqrt(n_steps)); >= n_steps: one segment, kept */
    uint32_t flags;           /* HB_FLAG_NO_BLOCKING: one barrier per step even when min delay >= 8;
                                 HB_FLAG_SEPARATE_BOLD: BOLD monitor kernels instead of helper waves */
} hb_adjoint_opts;

typedef struct {
    double forward_ms;        /* forward save + BOLD + loss, HIP events */
    double backward_ms;       /* BOLD adjoint + reverse sweep, HIP events */
    double node_steps;
    int64_t device_bytes;     /* plan arena: coupling log, checkpoints, tapes, buffers */
} hb_adjoint_stats;

typedef struct hb_adjoint_plan hb_adjoint_plan;

/* The differentiable graph of each instance is its forward recursion (the same
 * step order, clamps and per-edge delayed coupling as hb_plan_*); supported:
 * CSR weights with per-edge delays (or none), Linear / ScaledLinear / Kuramoto
 * coupling, all six models and integrators, up to 1024 nodes. No reference
 * counterpart: it is the gradient path added beside the plain forward run. */
int hb_adjoint_plan_create(const hb_instance* inst, const hb_adjoint_spec* spec, int32_t n,
                           const hb_adjoint_opts* opts, hb_adjoint_plan** out);
/* Forward with checkpoints every S steps and the per-step coupling log; loss. */
int hb_adjoint_forward(hb_adjoint_plan* p, hb_adjoint_stats* stats);
/* Reverse sweep: recompute each segment from its checkpoint, step the adjoint back. */
int hb_adjoint_backward(hb_adjoint_plan* p, hb_adjoint_stats* stats);
/* Copy losses and gradients to the spec's host pointers. */
int hb_adjoint_download(hb_adjoint_plan* p);
void hb_adjoint_plan_destroy(hb_adjoint_plan* p);

/* Launch-shape autotuner (HB_FLAG_AUTOTUNE on hb_plan_create). A shape class is
 * (model, ring channels, nodes, delay-blocked, ring in LDS, integrator class,
 * stochastic); its choice is the lanes per node of the delay-blocked kernel
 * (1, 2, 4), the block size of a network larger than one block (256, 512), or,
 * for undelayed dense coupling (blocked field 2), per-lane row sums (1) or the
 * matrix-core product over 16-instance tiles (16),
 * timed on a short sub-plan of up to 32 sampled instances and cached for the
 * process. Instances per workgroup stay 1: every class fills at least one wave.
 * hb_autotune_entries writes up to max_entries rows of 8 int32 (model, np,
 * n_nodes, blocked, lring, integrator class, stochastic, choice) and returns
 * the number of cached classes. No reference counterpart. */
int hb_autotune_entries(int32_t* out, int32_t max_entries);
void hb_autotune_clear(void);

/* ---------------- several subnetworks (BatchHybridEngine with projections) ---------------- */
typedef struct {
    int32_t src, tgt;         /* subnetwork indices */
    int32_t wkind;            /* HB_W_SCALAR_MEAN (w * mean over source nodes), HB_W_DENSE or HB_W_CSR */
    float scalar_w;
    const float* dense_w;     /* [n_tgt][n_src]: row = target node (CSR densified, engine.rs:298-311) */
    int32_t cfn, n_cparams;   /* HB_CF_*, coupling_params */
    float cparams[8];
    int32_t n_map;            /* cvar_map pairs map_src[k]:map_tgt[k] ("0:0" = one pair 0, 0) */
    int32_t map_src[4], map_tgt[4];
    int32_t delay;            /* steps: the projection's first delay (engine.rs:407) */
    /* HB_W_CSR (by target row): densified unless edge_delays is given; with
     * edge_delays (plain mode only) each edge reads its source d_e steps back
     * (runtime.rs:125-205) */
    const int32_t* indptr;
    const int32_t* indices;
    const float* csr_w;
    const int32_t* edge_delays;
    int32_t nnz;
} hb_multi_projection;

/* A stimulus of the plain HybridEngine (src/engine/stimulus.rs:17-128), added to
 * coupling variable 0 of its target subnetwork after the projections
 * (runtime.rs:319-357). pattern: 0 impulse [onset_ms, amplitude], 1 step
 * [onset_ms, offset_ms, amplitude], 2 sinusoid [onset_ms, amplitude, freq_hz,
 * phase_rad], 3 pulse_train [onset_ms, amplitude, period_ms, width_ms]. */
typedef struct {
    int32_t target, pattern, n_params;
    float params[4];
} hb_stimulus;

typedef struct {
    int32_t n_subnets;        /* 1..4, at most 1024 nodes together */
    const hb_subnetwork* subnets;
    int32_t n_projections;    /* 0..8 */
    const hb_multi_projection* projections;
    float dt;
    int32_t integrator;       /* HB_EULER ... HB_RK4_STOCH */
    int32_t hybrid;           /* hybrid_integrator: deterministic Heun for G2DO (engine.rs:698-702) */
    int32_t n_sweep, n_steps;
    int32_t sweep_sub, sweep_param;   /* SweepParam; sweep_sub = -1: none */
    const float* sweep_values;        /* [n_sweep] */
    int32_t max_delay;        /* largest delay of any projection's delay list (history horizon - 1,
                                 engine.rs:253-257); the projections' own delays count too */
    int32_t n_nsig;           /* stochastic integrators: nsig per variable (the reference sizes it by
                                 the first subnetwork's nvar, engine.rs:452) */
    const float* nsig;
    const float* noise;       /* raw N(0,1) draws, per subnetwork [n_steps][n_nodes][nvar], concatenated;
                                 one draw per step shared by every sweep point (engine.rs:726-763) */
    int32_t device;
    float* tavg;              /* out: per subnetwork [n_sweep][n_nodes][nvar], concatenated */
    float* final_state;       /* out: same layout */
    double kernel_ms;         /* out: the kernel's HIP-event time */
    /* resume (all optional): the engine's step counter at the start, per sweep
     * point start states (the final_state layout; null: every point starts from
     * the subnetworks' x0) and the coupling-variable history of the last
     * max_delay + 1 steps, [n_sweep][total nodes][max_delay + 1][2] by step
     * modulo max_delay + 1 (null: zeros). hist_out receives it at the end. */
    int64_t step0;
    const float* state_in;
    const float* hist_in;
    float* hist_out;
    /* plain = 1: the plain HybridEngine's semantics (runtime.rs:46-501): delays
     * read the true ring (step - d), Scalar weights are w * I (construction.rs:
     * 639-654), stimuli apply, the RK4-stochastic order of integrator.rs:296-315,
     * no hybrid_integrator */
    int32_t plain;
    int32_t n_stimuli;        /* 0..8 */
    const hb_stimulus* stimuli;
    double dt_ms;             /* f64 dt for stimulus times (step * dt_ms) */
    float* trajectory;        /* optional out: per subnetwork [n_steps][n_sweep][n_nodes][nvar], concatenated
                                 (run_sweep_with_trajectory, engine.rs:779-783) */
    /* BOLD monitor of subnetwork 0 (the batch engine's monitor target, engine.rs:418-447),
     * per sweep point: bold_period > 0 steps per window, TR bold_tr_ms (<= 0: native
     * window samples); bold receives [n_sweep][hb_bold_rows(...)][n_nodes of subnetwork 0] */
    int32_t bold_period;
    double bold_tr_ms;
    float* bold;
    /* run_sweep_multi over several subnetworks (engine.rs:862-1348; all optional):
     * per subnetwork, point_mode 0 keeps its parameters for every sweep point;
     * 1 sweeps its batch-native parameter in the swept form with the value
     * point_values[point * n_subnets + subnetwork] (one swept dimension on the
     * model's native index, dfun.rs:38-72); 2 gives every point its own parameter
     * vector point_params[(point * n_subnets + subnetwork) * HB_MAX_PARAMS + k]
     * (two or more swept dimensions, dfun_batch_multi_fallback, dfun.rs:99-133).
     * A nonzero point_mode replaces sweep_sub / sweep_param for that run. */
    int32_t point_mode[4];
    const float* point_values;
    const float* point_params;
    /* bold_mean = 1 (with bold_period > 0): the batch engine's own monitor
     * (engine.rs:785-801), one Balloon-Windkessel model per node of subnetwork 0
     * fed var0 averaged over the sweep points. bold_state is its s, f, v, q per
     * node, [n_nodes][4], read and written back (null: from rest, 0 1 1 1), so a
     * caller carries it between runs as the engine's BoldMonitor does; bold
     * receives this run's native window samples, [hb_bold_rows(n_steps,
     * bold_period, dt_ms, 0)][n_nodes], the partial window flushed at the end
     * (bold_monitor.rs:72-107, 207-227). hb_bold_downsample resamples the
     * series of several runs to TR; bold_tr_ms is not used. */
    int32_t bold_mean;
    float* bold_state;
} hb_multi_spec;

/* Rows of BOLD output a run of n_steps produces (bold_monitor.rs:72-227): native
 * windows when tr_ms <= 0 or equal to the window length, else floor(last time / TR). */
int32_t hb_bold_rows(int32_t n_steps, int32_t bold_period, double dt_ms, double tr_ms);

/* BoldMonitor::downsample_to_tr (bold_monitor.rs:123-203): n native samples
 * [n][n_nodes] taken at times[n] (ms, as the monitor forms them) with window
 * length dtb_ms, resampled at multiples of tr_ms (linear between the
 * bracketing samples, nearest outside them; tr_ms <= 0 or equal to dtb_ms:
 * the native samples). Writes the rows into out when out is not null and cap
 * holds them; returns the row count, or -1 when out is too small. */
int32_t hb_bold_downsample(const double* times, int32_t n, const float* native, int32_t n_nodes, double dtb_ms,
                           double tr_ms, float* out, int32_t cap);

/* BatchHybridEngine::run_sweep over several subnetworks joined by projections
 * (src/engine/batch_engine/engine.rs:540-823): per step, each projection's
 * coupling from the current source states (pre, Scalar mean or Dense weights,
 * post, cvar_map remap), summed per target, then every subnetwork integrates
 * with that frozen coupling; tavg = sum of states / n_steps. A projection
 * delay d > 0 reads the batch engine's history window of max_delay + 1 states
 * at slot (step - d + H) % H (engine.rs:541-589), zeros while d > step.
 * Stochastic integrators add raw * sqrt(2 nsig_v) * sqrt(dt) (integrator.rs:
 * 179-191) as the batch engine does. HB_INVALID_CONFIG for unknown models,
 * cvar_map indices out of range (engine.rs:371-388), or an nsig shorter than
 * a stochastic subnetwork's nvar (the reference indexes past its end). */
int hb_run_multi(hb_multi_spec* spec);

/* ---------------- functional-connectivity features ---------------- */
/* Batched FC: series is [n_series][n_steps][n_nodes] row-major (host memory),
 * fc receives [n_series][n_nodes][n_nodes]. Each matrix follows fc_matrix
 * (src/sbi/features/fc.rs:11-57): FP64 means and covariances, population
 * standard deviation (< 1e-12 -> 1), clamp to [-1, 1], f32, symmetric. */
int hb_fc_batch(const float* series, int32_t n_series, int32_t n_steps, int32_t n_nodes, float* fc,
                int32_t device);
/* No counterpart in the reference: the FC entry point for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). d_series [n_series][n_steps][n_nodes] and d_fc
 * [n_series][n_nodes][n_nodes] are device memory on `device`; the kernels of
 * hb_fc_batch run on `stream` (a hipStream_t, NULL: the null stream) and are
 * not waited for. Jobs, column sums and split partials live in a fixed 64 MiB
 * device array per GPU (looked up on the first call on a device: make that
 * call outside graph capture); calls whose scratch exceeds it return
 * HB_INVALID_CONFIG before any GPU work. Every call on a GPU shares its array:
 * calls that overlap on one GPU race. Nothing in it needs clearing. */
int hb_fc_batch_noalloc(const float* d_series, int32_t n_series, int32_t n_steps, int32_t n_nodes, float* d_fc,
                        int32_t device, void* stream);
/* Number of FCD windows, (n_steps - window_len) / stride + 1, as fcd_matrix
 * counts them (src/sbi/features/fc.rs:104); 0 for an invalid window. */
int32_t hb_fcd_windows(int32_t n_steps, int32_t window_len, int32_t stride);
/* Batched FCD (src/sbi/features/fc.rs:94-179 fcd_matrix): per series, the FC
 * of each window of window_len steps every stride steps, then the Pearson
 * correlation of the windows' upper triangles (0 when either is constant),
 * diagonal 1. fcd receives [n_series][W][W], W = hb_fcd_windows(...); when
 * W <= 1 nothing is written (the reference returns an empty matrix). Returns
 * HB_INVALID_CONFIG where fcd_matrix asserts: window_len outside [2, n_steps]
 * or stride < 1. */
int hb_fcd_batch(const float* series, int32_t n_series, int32_t n_steps, int32_t n_nodes, int32_t window_len,
                 int32_t stride, float* fcd, int32_t device);

/* Bins of welch_psd (src/sbi/features/catch22/fft.rs:120-170):
 * next_power_of_2(max(nfft, window_width)) / 2 + 1; 0 for window_width < 2 or
 * a transform longer than 65536 points. */
int32_t hb_welch_bins(int32_t nfft, int32_t window_width);
/* Batched Welch PSD (welch_psd, catch22/fft.rs:120-170) of every node series
 * of series [n_series][n_steps][n_nodes]: segments of window_width samples
 * every window_width / 2, minus the series mean, zero-padded,
 * |DFT|^2 summed over segments, / (n_segments * window_width * fs).
 * psd receives [n_series][n_nodes][hb_welch_bins(nfft, window_width)] (f64);
 * bin k is at frequency k * fs / (2 * (bins - 1)). */
int hb_welch_psd_batch(const float* series, int32_t n_series, int32_t n_steps, int32_t n_nodes, int32_t nfft,
                       double fs, int32_t window_width, double* psd, int32_t device);
/* spectral_features (src/sbi/features/spectral.rs:20-99) of every node series:
 * feats receives [n_series][n_nodes][9] = delta, theta, alpha, beta, gamma
 * power fractions, spectral centroid, spread, skewness, excess kurtosis of
 * the Welch PSD with nfft = next_power_of_2(n_steps) and window
 * clamp(n_steps / 2, 16, 256); NaN x 9 for n_steps < 16 or zero total power. */
int hb_spectral_features_batch(const float* series, int32_t n_series, int32_t n_steps, int32_t n_nodes, double fs,
                               float* feats, int32_t device);

/* ---------------- misc ---------------- */
const char* hb_last_error(void);   /* thread-local message of the last failure */
const char* hb_version(void);
int hb_device_count(void);
/* Wait for all work on a device (hipDeviceSynchronize). */
int hb_device_synchronize(int32_t device);
/* Backend selection, the `--backend hip` value of select_backend (cli.rs:259-271).
 * Returns "hip" for "hip"/"rocm"; anything else returns NULL and sets the last
 * error: this library has no CPU fallback. */
const char* hb_select_backend(const char* name);

#ifdef __cplusplus
}
#endif
#endif /* HYBURN_HIP_H */
