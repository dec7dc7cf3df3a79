/*
 * hb_oracle.c -- CPU parity oracle for the hyburn batched integration path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or as the timed CPU baseline). The product path never calls it.
 *
 * It restates, in plain C, the reference's NdArray<f32> arithmetic:
 *   src/engine/batch_engine/dfun.rs:170-239,271-1870  (the 28 dfuns + clamps)
 *   src/engine/batch_engine/engine.rs:540-823         (time loop, integrators, tavg)
 *   src/engine/integrator.rs:89-315                   (Heun/RK4/stochastic forms)
 *   src/engine/coupling.rs:114-444                    (coupling pre/post)
 *   src/engine/sparse.rs:29-52, runtime.rs:121-211    (CSR, per-edge delays)
 *   src/engine/bold.rs:106-166, bold_monitor.rs:72-227 (Balloon-Windkessel)
 *   src/sbi/features/fc.rs:11-57,94-179               (Pearson FC, FCD)
 * Transcendentals come from the platform libm (expf/sinf/...), as in the
 * reference (Rust f32::exp etc. call libm).
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math (see Makefile). No FMA
 * contraction, so every f32 operation rounds exactly as written.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORC_G2DO = 0, ORC_MPR = 1, ORC_JR = 2, ORC_WC = 3, ORC_KUR = 4, ORC_RWW = 5, ORC_LIN = 6, ORC_SH = 7, ORC_HOP = 8, ORC_CB2 = 9,
       ORC_EC3 = 10, ORC_EC3S = 11, ORC_ERS = 12, ORC_ZL1 = 13, ORC_ZL2 = 14, ORC_KIX = 15,
       ORC_CB = 16, ORC_GSD = 17, ORC_GSF = 18, ORC_LB = 19, ORC_E2D = 20, ORC_EPI = 21, ORC_RWEI = 22,
       ORC_DECO = 23, ORC_ZJ = 24, ORC_RFHN = 25, ORC_RHR = 26, ORC_DG = 27 };
enum { ORC_NMODELS = 28, ORC_MAXP = 64 };
enum { ORC_EULER = 0, ORC_HEUN = 1, ORC_RK4 = 2, ORC_EULER_STOCH = 3, ORC_HEUN_STOCH = 4, ORC_RK4_STOCH = 5 };
enum { ORC_CF_LINEAR = 0, ORC_CF_DIFFERENCE = 1, ORC_CF_KURAMOTO = 2, ORC_CF_SIGMOIDAL = 3,
       ORC_CF_SCALED_LINEAR = 4, ORC_CF_TANH = 5, ORC_CF_SJR = 6, ORC_CF_PRESIGMOIDAL = 7 };
enum { ORC_W_NONE = 0, ORC_W_SCALAR_MEAN = 1, ORC_W_SCALAR_DIAG = 2, ORC_W_DENSE = 3, ORC_W_CSR = 4 };
enum { ORC_DELAY_PER_EDGE = 0, ORC_DELAY_SINGLE = 1, ORC_DELAY_BATCH_COMPAT = 2 };
enum { ORC_ENGINE_BATCH = 0, ORC_ENGINE_PLAIN = 1 };

typedef struct {
    int32_t cfn;
    int32_t n_cparams;
    float cparams[8];
    int32_t n_src;          /* Kuramoto normalisation, engine.rs:282 */
    int32_t cpu_form;       /* 1: pre_cpu op order (per-edge CSR path) */
    int32_t wkind;
    float scalar_w;
    const float* dense_w;   /* [N][N], row = target */
    const int32_t* indptr;  /* CSR by target row */
    const int32_t* indices;
    const float* weights;
    const int32_t* edge_delays;  /* per edge, in steps (per-edge mode) */
    int32_t delay_mode;
    int32_t single_delay;
    int32_t max_delay;      /* horizon - 1 */
} hb_orc_coupling;

typedef struct {
    int32_t model, integrator, engine, n_nodes, n_steps;
    float dt;               /* engine dt, f32 (engine.rs:456) */
    double dt_ms;           /* config dt, f64, used for BOLD window times */
    const float* params;
    int32_t sweep_form;
    float sweep_val;
    int32_t node_param_idx;
    const float* node_param;
    hb_orc_coupling cpl;
    const float* nsig;      /* [NV] */
    const float* noise;     /* raw N(0,1) f32, [T][N][NV] */
    const float* x0;        /* [N][NV] node-major */
    int32_t bold_period;
    int32_t bold_cap;       /* capacity of the BOLD output in samples */
    void* out_tavg;         /* [N][NV] */
    void* out_final;        /* [N][NV] */
    void* out_traj;         /* [T][N][NV] or NULL */
    void* out_bold;         /* [bold_cap][N] or NULL */
    int32_t* out_bold_n;
} hb_orc_problem;

/* model registry, config.rs:43-50 */
static int hb_orc_nvar(int m) {
    static const int v[ORC_NMODELS] = {2, 2, 6, 2, 1, 1, 1, 2, 2, 2, 3, 5, 8, 5, 8, 5, 4, 4, 4, 3, 2, 6, 2, 2, 12, 4, 6, 8};
    return v[m];
}
static int hb_orc_ncvar(int m) {
    static const int v[ORC_NMODELS] = {1, 2, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 3, 1, 1, 1, 4, 4, 4, 1, 1, 2, 1, 1, 1, 2, 2, 4};
    return v[m];
}
/* registry rows of config.rs:43-72 (EpileptorCodim3 ... KIonEx: 10..15,
 * CoombesByrne ... DumontGutkin: 16..27) */
static int hb_orc_nparams(int m) {
    static const int v[ORC_NMODELS] = {12, 7, 13, 22, 1, 8, 1, 2, 3, 4, 13, 27, 28, 50, 50, 14,
                                       5, 9, 9, 32, 12, 17, 19, 20, 18, 8, 12, 14};
    return v[m];
}

int hb_orc_model_info(int m, int32_t* nvar, int32_t* ncvar, int32_t* nparams)
{
    if (m < 0 || m >= ORC_NMODELS) return -1;
    *nvar = hb_orc_nvar(m); *ncvar = hb_orc_ncvar(m); *nparams = hb_orc_nparams(m);
    return 0;
}

/* ---------------- FP32 twin ---------------- */
#define REAL float
#define FN(name) f32_##name
#define EPSR 1.1920929e-07f
#define PIR 3.14159265358979323846f
#define EXPR expf
#define SINR sinf
#define COSR cosf
#define TANHR tanhf
#define POWR powf
#define SQRTR sqrtf
#define FLOORR floorf
#define LOGR logf
#include "hb_oracle_body.h"
#undef REAL
#undef FN
#undef EPSR
#undef PIR
#undef EXPR
#undef SINR
#undef COSR
#undef TANHR
#undef POWR
#undef SQRTR
#undef FLOORR
#undef LOGR

/* ---------------- FP64 twin ---------------- */
#define REAL double
#define FN(name) f64_##name
#define EPSR 1.1920928955078125e-07   /* f32::EPSILON, the same guard value */
#define PIR ((double)3.14159265358979323846f)  /* the f32 constant, widened */
#define EXPR exp
#define SINR sin
#define COSR cos
#define TANHR tanh
#define POWR pow
#define SQRTR sqrt
#define FLOORR floor
#define LOGR log
#include "hb_oracle_body.h"

/* ---------------- exported API ---------------- */

int hb_orc_run_f32(const hb_orc_problem* p)
{
    return f32_run(p, (float*)p->out_tavg, (float*)p->out_final, (float*)p->out_traj,
                   (float*)p->out_bold, p->out_bold_n);
}

int hb_orc_run_f64(const hb_orc_problem* p)
{
    return f64_run(p, (double*)p->out_tavg, (double*)p->out_final, (double*)p->out_traj,
                   (double*)p->out_bold, p->out_bold_n);
}

/* Many independent instances over host threads (OpenMP), the shape of
 * rayon_batch_sweep (engine.rs:1363-1429). Used as the timed CPU baseline. */
int hb_orc_run_many_f32(const hb_orc_problem* ps, int32_t n, int32_t n_threads)
{
    int i, bad = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+:bad)
#endif
    for (i = 0; i < n; ++i) bad += hb_orc_run_f32(&ps[i]) != 0;
    (void)n_threads;
    return bad ? -1 : 0;
}

int hb_orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void hb_orc_dfun_f32(int model, const float* params, int sweep_form, float sweep_val,
                     const float* x, const float* c, float* dx)
{
    f32_mconst k;
    f32_mconst_init(&k, model, params, sweep_form, sweep_val);
    f32_dfun(&k, x, c, dx);
}

void hb_orc_dfun_f64(int model, const float* params, int sweep_form, float sweep_val,
                     const double* x, const double* c, double* dx)
{
    f64_mconst k;
    f64_mconst_init(&k, model, params, sweep_form, sweep_val);
    f64_dfun(&k, x, c, dx);
}

void hb_orc_clamp_f32(int model, float* x) { f32_clamp(model, x); }
void hb_orc_clamp_f64(int model, double* x) { f64_clamp(model, x); }

/* One coupling evaluation without delays, rectangular projection:
 * post_with_target(W @ pre(x_src), x_tgt) (coupling.rs:465-480, sparse.rs:89-128).
 * x_src [n_src][nc], x_tgt [n_tgt][nc], out [n_tgt][nc]. wkind: dense or CSR. */
int hb_orc_coupling_f32(const hb_orc_coupling* cpl, int n_tgt, int n_src, int nc,
                        const float* x_src, const float* x_tgt, float* out)
{
    f32_cconst k;
    float pre[8];
    float* P;
    float g[16];
    int i, j, c, e, np;
    f32_cconst_init(&k, cpl);
    np = f32_npre(&k, nc);
    P = (float*)calloc((size_t)n_src * np, sizeof(float));
    if (!P) return -1;
    for (j = 0; j < n_src; ++j) { f32_pre(&k, x_src + (size_t)j * nc, nc, pre); memcpy(P + (size_t)j * np, pre, sizeof(float) * np); }
    for (i = 0; i < n_tgt; ++i) {
        float rowsum = 0;
        for (c = 0; c < np; ++c) g[c] = 0;
        if (cpl->wkind == ORC_W_DENSE) {
            for (j = 0; j < n_src; ++j) {
                float w = cpl->dense_w[(size_t)i * n_src + j];
                for (c = 0; c < np; ++c) g[c] += w * P[(size_t)j * np + c];
                rowsum += w;
            }
        } else if (cpl->wkind == ORC_W_CSR) {
            for (e = cpl->indptr[i]; e < cpl->indptr[i + 1]; ++e) {
                float w = cpl->weights[e];
                for (c = 0; c < np; ++c) g[c] += w * P[(size_t)cpl->indices[e] * np + c];
                rowsum += w;
            }
        }
        f32_post(&k, g, x_tgt + (size_t)i * nc, rowsum, nc, out + (size_t)i * nc);
    }
    free(P);
    return 0;
}

/* The FP64 twin of hb_orc_coupling_f32: the same pre/post order in double,
 * weights widened from f32. */
int hb_orc_coupling_f64(const hb_orc_coupling* cpl, int n_tgt, int n_src, int nc,
                        const double* x_src, const double* x_tgt, double* out)
{
    f64_cconst k;
    double pre[8];
    double* P;
    double g[16];
    int i, j, c, e, np;
    f64_cconst_init(&k, cpl);
    np = f64_npre(&k, nc);
    P = (double*)calloc((size_t)n_src * np, sizeof(double));
    if (!P) return -1;
    for (j = 0; j < n_src; ++j) { f64_pre(&k, x_src + (size_t)j * nc, nc, pre); memcpy(P + (size_t)j * np, pre, sizeof(double) * np); }
    for (i = 0; i < n_tgt; ++i) {
        double rowsum = 0;
        for (c = 0; c < np; ++c) g[c] = 0;
        if (cpl->wkind == ORC_W_DENSE) {
            for (j = 0; j < n_src; ++j) {
                double w = cpl->dense_w[(size_t)i * n_src + j];
                for (c = 0; c < np; ++c) g[c] += w * P[(size_t)j * np + c];
                rowsum += w;
            }
        } else if (cpl->wkind == ORC_W_CSR) {
            for (e = cpl->indptr[i]; e < cpl->indptr[i + 1]; ++e) {
                double w = cpl->weights[e];
                for (c = 0; c < np; ++c) g[c] += w * P[(size_t)cpl->indices[e] * np + c];
                rowsum += w;
            }
        }
        f64_post(&k, g, x_tgt + (size_t)i * nc, rowsum, nc, out + (size_t)i * nc);
    }
    free(P);
    return 0;
}

/* pre() of n rows of nc cvars -> out [n][npre]; returns npre (coupling.rs:204-321).
 * cpl->cpu_form selects the per-edge pre_cpu op order. */
int hb_orc_pre_f32(const hb_orc_coupling* cpl, int n, int nc, const float* x, float* out)
{
    f32_cconst k;
    int i, np;
    f32_cconst_init(&k, cpl);
    np = f32_npre(&k, nc);
    for (i = 0; i < n; ++i) f32_pre(&k, x + (size_t)i * nc, nc, out + (size_t)i * np);
    return np;
}

/* post_with_target() of n rows: g [n][npre], x_tgt [n][nc] -> out [n][nc]
 * (coupling.rs:323-399); rowsum is the Difference row sum, 0 otherwise. */
void hb_orc_post_f32(const hb_orc_coupling* cpl, int n, int nc, const float* g, const float* x_tgt,
                     float rowsum, float* out)
{
    f32_cconst k;
    int i, np;
    f32_cconst_init(&k, cpl);
    np = f32_npre(&k, nc);
    for (i = 0; i < n; ++i)
        f32_post(&k, g + (size_t)i * np, x_tgt + (size_t)i * nc, rowsum, nc, out + (size_t)i * nc);
}

/* CSR matvec, sparse.rs:29-52 */
void hb_orc_csr_matvec_f32(const float* data, const int32_t* indices, const int32_t* indptr,
                           int n_tgt, const float* x, float* y)
{
    int i, j;
    for (i = 0; i < n_tgt; ++i) {
        float acc = 0.0f;
        for (j = indptr[i]; j < indptr[i + 1]; ++j) acc += data[j] * x[indices[j]];
        y[i] = acc;
    }
}

void hb_orc_bold_step_f32(int n, const float* z, float dt_s, float* s, float* f, float* v, float* q)
{
    f32_bold_step(n, z, dt_s, s, f, v, q);
}

void hb_orc_bold_step_f64(int n, const double* z, double dt_s, double* s, double* f, double* v, double* q)
{
    f64_bold_step(n, z, dt_s, s, f, v, q);
}

void hb_orc_bold_signal_f32(int n, const float* v, const float* q, float* out)
{
    int i;
    for (i = 0; i < n; ++i) out[i] = f32_bold_signal1(v[i], q[i]);
}

void hb_orc_bold_signal_f64(int n, const double* v, const double* q, double* out)
{
    int i;
    for (i = 0; i < n; ++i) out[i] = f64_bold_signal1(v[i], q[i]);
}

/* Downsample native BOLD samples to TR, bold_monitor.rs:123-203.
 * times[n] in ms (f64), data[n][N]. Returns number of output rows written. */
int hb_orc_bold_downsample(const double* times, const float* data, int n, int N,
                           double dt_bold, double tr_ms, float* out, int cap)
{
    int j, k, nn, n_out;
    double max_t;
    if (n == 0) return 0;
    if (tr_ms <= 0.0 || fabs(tr_ms - dt_bold) < 1e-6) {
        int m = n < cap ? n : cap;
        memcpy(out, data, sizeof(float) * (size_t)m * N);
        return n;
    }
    max_t = times[n - 1];
    n_out = max_t >= tr_ms ? (int)floor(max_t / tr_ms) : 0;
    for (j = 0; j < n_out && j < cap; ++j) {
        double tt = (double)(j + 1) * tr_ms;
        int idx = 0;
        float* o = out + (size_t)j * N;
        if (tt <= times[0]) { for (nn = 0; nn < N; ++nn) o[nn] = data[nn]; continue; }
        for (k = 1; k < n; ++k) {
            if (times[k] > tt) { idx = k - 1; break; }
            idx = k;
        }
        if (idx + 1 >= n) { for (nn = 0; nn < N; ++nn) o[nn] = data[(size_t)idx * N + nn]; continue; }
        {
            double t0 = times[idx], t1 = times[idx + 1];
            double w = fabs(t1 - t0) < 1e-12 ? 0.0 : (tt - t0) / (t1 - t0);
            float wf;
            if (w < 0.0) w = 0.0;
            if (w > 1.0) w = 1.0;
            wf = (float)w;
            for (nn = 0; nn < N; ++nn) {
                float v0 = data[(size_t)idx * N + nn], v1 = data[(size_t)(idx + 1) * N + nn];
                o[nn] = v0 + (v1 - v0) * wf;
            }
        }
    }
    return n_out;
}

/* Pearson FC, fc.rs:11-57: f64 accumulators, population std, std<1e-12 -> 1,
 * clamp [-1,1], symmetric. traj [T][N] f32 -> out [N][N] f32. */
int hb_orc_fc(const float* traj, int T, int N, float* out)
{
    double* mean = (double*)calloc((size_t)N, sizeof(double));
    double* sd = (double*)calloc((size_t)N, sizeof(double));
    int t, i, j;
    if (!mean || !sd) { free(mean); free(sd); return -1; }
    for (t = 0; t < T; ++t) for (i = 0; i < N; ++i) mean[i] += (double)traj[(size_t)t * N + i];
    for (i = 0; i < N; ++i) mean[i] /= (double)T;
    for (t = 0; t < T; ++t) for (i = 0; i < N; ++i) { double d = (double)traj[(size_t)t * N + i] - mean[i]; sd[i] += d * d; }
    for (i = 0; i < N; ++i) { sd[i] = sqrt(sd[i] / (double)T); if (sd[i] < 1e-12) sd[i] = 1.0; }
    for (i = 0; i < N; ++i) {
        for (j = 0; j <= i; ++j) {
            double cov = 0.0, r;
            for (t = 0; t < T; ++t)
                cov += ((double)traj[(size_t)t * N + i] - mean[i]) * ((double)traj[(size_t)t * N + j] - mean[j]);
            cov /= (double)T;
            r = cov / (sd[i] * sd[j]);
            if (r < -1.0) r = -1.0;
            if (r > 1.0) r = 1.0;
            out[(size_t)i * N + j] = (float)r;
            if (i != j) out[(size_t)j * N + i] = (float)r;
        }
    }
    free(mean); free(sd);
    return 0;
}

/* Pearson correlation of the i < j entries of two N x N matrices
 * (upper_triangle_correlation, src/sbi/features/fc.rs:143-178). */
static float orc_upper_corr(const float* a, const float* b, int N)
{
    double n = (double)N * (N - 1) / 2, sa = 0.0, sb = 0.0, ma, mb, cov = 0.0, va = 0.0, vb = 0.0, r;
    int i, j;
    for (i = 0; i < N; ++i) for (j = i + 1; j < N; ++j) { sa += a[(size_t)i * N + j]; sb += b[(size_t)i * N + j]; }
    ma = sa / n;
    mb = sb / n;
    for (i = 0; i < N; ++i)
        for (j = i + 1; j < N; ++j) {
            double da = (double)a[(size_t)i * N + j] - ma, db = (double)b[(size_t)i * N + j] - mb;
            cov += da * db;
            va += da * da;
            vb += db * db;
        }
    if (va <= 0.0 || vb <= 0.0) return 0.0f;
    r = cov / (sqrt(va) * sqrt(vb));
    if (r < -1.0) r = -1.0;
    if (r > 1.0) r = 1.0;
    return (float)r;
}

/* FCD matrix (fcd_matrix, src/sbi/features/fc.rs:94-141): the FC of each
 * window of L steps every S steps, then the upper-triangle correlations.
 * Writes W x W into out and returns W; returns 0 (nothing written) when
 * W <= 1, -1 where the reference asserts (L outside [2, T], S < 1). */
int hb_orc_fcd(const float* traj, int T, int N, int L, int S, float* out)
{
    int W, w, a, b;
    float* fcs;
    if (L < 2 || L > T || S < 1) return -1;
    W = (T - L) / S + 1;
    if (W <= 1) return 0;
    fcs = (float*)malloc(sizeof(float) * (size_t)W * N * N);
    if (!fcs) return -1;
    for (w = 0; w < W; ++w) hb_orc_fc(traj + (size_t)w * S * N, L, N, fcs + (size_t)w * N * N);
    for (a = 0; a < W; ++a) {
        out[(size_t)a * W + a] = 1.0f;
        for (b = a + 1; b < W; ++b) {
            float r = orc_upper_corr(fcs + (size_t)a * N * N, fcs + (size_t)b * N * N, N);
            out[(size_t)a * W + b] = r;
            out[(size_t)b * W + a] = r;
        }
    }
    free(fcs);
    return W;
}

/* Effective source time read for delay d at step s (-1: the zero/pre(0) value).
 * ring: HybridEngine ring, runtime.rs:55-85,162-204 (true delay).
 * batch_compat: BatchHybridEngine sliding window read as a ring, engine.rs:541-589. */
int64_t hb_orc_source_time(int mode, int64_t s, int64_t d, int64_t H)
{
    if (d == 0 || s == 0) return s;
    if (d > s) return -1;
    if (mode == ORC_DELAY_BATCH_COMPAT) {
        int64_t t = s - H + 1 + ((s - d + H) % H);
        return t < 0 ? -1 : t;
    }
    return s - d;
}
