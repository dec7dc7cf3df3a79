// hb_engine.cpp -- BatchHybridEngine / HybridEngine mirror over the instance
// planner: config copy + validation, sweep expansion (single, Cartesian),
// batched result layout.
//
// Reference (maedoc/hyburn):
//   BatchHybridEngine::from_config    src/engine/batch_engine/engine.rs:145-471
//   run_sweep / _with_trajectory      engine.rs:477-494, 496-842
//   run_sweep_multi, Cartesian order  engine.rs:56-91, 845-1348
//   SimConfig::validate               src/config.rs:286-428
//   HybridEngine weights/delays       src/engine/construction.rs:543-790
//   checkpoint / resume state         src/engine/runtime.rs:519-561, 576-742
#include "hb_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace hb;

struct hb_engine {
    int device = 0;
    int n_sweep = 0;
    double sim_length = 0.0, dt = 0.0;
    int integrator = HB_HEUN;
    int model = 0, n_nodes = 0, nv = 0;
    std::vector<float> params, x0, nsig;
    bool has_proj = false;
    int wkind_cfg = HB_W_NONE;   // HB_W_SCALAR_MEAN for "scalar", DENSE, CSR
    float scalar_w = 0.0f;
    std::vector<float> dense, data;
    std::vector<int32_t> indptr, indices, delays;
    int cfn = HB_CF_LINEAR;
    std::vector<float> cparams;
    int bold_period = 0;
    double bold_tr_s = 2.0;
    int engine = HB_ENGINE_BATCH;
    int delay_mode = -1;
    uint64_t noise_seed = 0;
    int noise_shared = 1;
    // Carried across runs, as the reference engine mutates self.states,
    // self.step, self.histories and self.bold_monitors (engine.rs:496-540,
    // 1328-1331). n_state == 0: the configured initial state at step 0.
    int64_t step = 0;
    int n_state = 0;
    std::vector<float> state;        // [n_state][N][V]
    std::vector<std::vector<float>> hist;   // per point [max_delay][N][channels]: pre() of the last steps
    // BOLD monitors: one per sweep point, or one fed the batch mean
    // (engine.rs:785-801). Each keeps its Balloon-Windkessel state and every
    // native sample so far; a run returns them resampled to TR over the whole
    // series, as BoldMonitor::flush does (bold_monitor.rs:123-227).
    int n_mon = 0;
    std::vector<float> bw;                    // [n_mon][N][4]: s, f, v, q
    std::vector<std::vector<float>> native;   // per monitor: [rows][N]
    std::vector<double> times;                // native sample times (ms), shared
    bool monitors_dropped = false;            // run_sweep_multi leaves the engine without monitors
    std::vector<float> x0_multi;              // start state of the current run_sweep_multi
};

namespace {

int efail(int code, const std::string& m) { return hb::set_error(code, m); }

// Build the per-point instances of a sweep. Arrays referenced by the instances
// live in `keep` (per-point parameter / coupling copies) or in the engine.
struct Built {
    std::vector<hb_instance> inst;
    std::vector<std::vector<float>> params;
};

hb_instance base_instance(const hb_engine* E, int64_t n_steps) {
    hb_instance I;
    memset(&I, 0, sizeof(I));
    I.model = E->model;
    I.integrator = E->integrator;
    I.engine = E->engine;
    I.n_nodes = E->n_nodes;
    I.n_steps = (int32_t)n_steps;
    I.dt = (float)E->dt;           // engine.rs:456 (dt as f32)
    I.dt_ms = E->dt;
    I.params = E->params.data();
    I.node_param_idx = -1;
    I.x0 = E->x0.data();
    I.nsig = E->nsig.data();
    I.noise_seed = E->noise_seed;
    I.cfn = E->cfn;
    I.n_cparams = (int32_t)std::min<size_t>(E->cparams.size(), 8);
    for (int k = 0; k < I.n_cparams; ++k) I.cparams[k] = E->cparams[k];
    if (E->has_proj) {
        const int ndl = (int)E->delays.size();
        const int d0 = ndl > 0 ? E->delays[0] : 0;
        switch (E->wkind_cfg) {
        case HB_W_SCALAR_MEAN:
            // scalar weights: mean field in the batch engine (engine.rs:602-608),
            // s*I in the plain engine (construction.rs:639-655)
            I.wkind = E->engine == HB_ENGINE_PLAIN ? HB_W_SCALAR_DIAG : HB_W_SCALAR_MEAN;
            I.scalar_w = E->scalar_w;
            break;
        case HB_W_DENSE:
            I.wkind = HB_W_DENSE;
            I.dense = E->dense.data();
            break;
        case HB_W_CSR:
            I.wkind = HB_W_CSR;
            I.indptr = E->indptr.data();
            I.indices = E->indices.data();
            I.weights = E->data.data();
            I.nnz = (int32_t)E->data.size();
            break;
        }
        int mode = E->delay_mode;
        if (mode < 0) {
            if (E->engine == HB_ENGINE_PLAIN)
                mode = (I.wkind == HB_W_CSR && ndl == I.nnz && ndl > 0) ? HB_DELAY_PER_EDGE : HB_DELAY_SINGLE;
            else
                mode = HB_DELAY_BATCH_COMPAT;   // engine.rs:410 + 541-589
        }
        I.delay_mode = mode;
        I.single_delay = d0;
        if (ndl > 0 && ndl == I.nnz && I.wkind == HB_W_CSR) I.edge_delays = E->delays.data();
        // window horizon = max(delays) + 1 although only delays[0] is read
        // (engine.rs:253-270, 410)
        for (int32_t d : E->delays) I.max_delay_hint = std::max(I.max_delay_hint, d);
        if (I.wkind != HB_W_CSR && mode == HB_DELAY_PER_EDGE) I.delay_mode = HB_DELAY_SINGLE;
    }
    I.bold_period = E->bold_period;
    I.bold_tr_ms = E->bold_tr_s * 1000.0;   // bold_monitor.rs:131 (tr * 1000)
    return I;
}

void reset_state(hb_engine* E) {
    E->monitors_dropped = false;
    E->step = 0;
    E->n_state = 0;
    E->state.clear();
    E->hist.clear();
    E->n_mon = 0;
    E->bw.clear();
    E->native.clear();
    E->times.clear();
}

// BoldMonitor::downsample_to_tr (bold_monitor.rs:123-203) over the whole
// native series of one monitor: samples at j * TR, linear between the
// bracketing native samples, nearest before the first and after the last.
// Returns the rows written to out ([rows][N], or the count only when out is null).
int downsample_to_tr(const std::vector<double>& times, const std::vector<float>& data, int N, double dtb,
                     double tr_ms, float* out) {
    const int n = (int)times.size();
    if (n == 0) return 0;
    if (tr_ms <= 0.0 || std::fabs(tr_ms - dtb) < 1e-6) {
        if (out) memcpy(out, data.data(), sizeof(float) * (size_t)n * N);
        return n;
    }
    const double max_t = times[n - 1];
    const int n_out = max_t >= tr_ms ? (int)std::floor(max_t / tr_ms) : 0;
    if (!out) return n_out;
    for (int j = 0; j < n_out; ++j) {
        const double tt = (double)(j + 1) * tr_ms;
        float* o = out + (size_t)j * N;
        if (tt <= times[0]) { memcpy(o, data.data(), sizeof(float) * N); continue; }
        int idx = 0;
        for (int k = 1; k < n; ++k) {
            if (times[k] > tt) { idx = k - 1; break; }
            idx = k;
        }
        if (idx + 1 >= n) { memcpy(o, data.data() + (size_t)idx * N, sizeof(float) * N); continue; }
        const double t0 = times[idx], t1 = times[idx + 1];
        double w = std::fabs(t1 - t0) < 1e-12 ? 0.0 : (tt - t0) / (t1 - t0);
        w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
        const float wf = (float)w;
        const float* v0 = data.data() + (size_t)idx * N;
        const float* v1 = v0 + N;
        for (int q = 0; q < N; ++q) o[q] = v0[q] + (v1[q] - v0[q]) * wf;
    }
    return n_out;
}

void fill_result(const hb_engine* E, int n_sweep, int64_t n_steps, bool traj, hb_result* out) {
    memset(out, 0, sizeof(*out));
    out->n_sweep = n_sweep;
    out->n_nodes = E->n_nodes;
    out->nvar = E->nv;
    out->n_steps = (int32_t)n_steps;
    const size_t st = (size_t)n_sweep * E->n_nodes * E->nv;
    out->tavg = (float*)calloc(st ? st : 1, 4);
    out->final_states = (float*)calloc(st ? st : 1, 4);
    if (traj) out->trajectories = (float*)calloc(st * (size_t)n_steps + 1, 4);
}

// Run the instances of one call and fill the batched result. carry: the
// instances continue the engine's carried state (x0, step, history, BOLD
// monitors), which is updated afterwards; monitors: BOLD is recorded.
int run_built(hb_engine* E, Built& B, int64_t n_steps, uint32_t flags, bool monitors, hb_result* out) {
    const auto t0 = std::chrono::steady_clock::now();
    const int n = (int)B.inst.size();
    const bool traj = (flags & HB_FLAG_TRAJECTORY) != 0;
    const bool mean_bold = (flags & HB_FLAG_BOLD_BATCH_MEAN) != 0;
    const bool with_bold = monitors && E->bold_period > 0;
    fill_result(E, n, n_steps, traj, out);
    const size_t per = (size_t)E->n_nodes * E->nv;
    const int N = E->n_nodes;
    std::vector<float> traj_tmp;
    if (traj) traj_tmp.resize(per * (size_t)n_steps * n + 1);

    // carried delay history: one buffer per sweep point, [L][N][channels]
    std::vector<std::vector<float>> hist(n);
    for (int i = 0; i < n; ++i) {
        hb_instance& I = B.inst[i];
        const int64_t hf = hb_history_floats(&I);
        if (hf <= 0) continue;
        if (I.step0 > 0) {
            if ((int)E->hist.size() != n || (int64_t)E->hist[i].size() != hf) {
                hb_result_free(out);
                return efail(HB_INVALID_STATE, "the carried delay history does not match this run");
            }
            hist[i] = E->hist[i];
        } else {
            hist[i].assign((size_t)hf, 0.0f);
        }
        I.history = hist[i].data();
    }

    // BOLD: native samples of this run per monitor, Balloon-Windkessel state carried
    const int P = E->bold_period;
    const int n_full = P > 0 ? (int)(n_steps / P) : 0, rem = P > 0 ? (int)(n_steps % P) : 0;
    const int rows = n_full + (rem > 0 ? 1 : 0);
    const int n_mon = with_bold ? (mean_bold ? 1 : n) : 0;
    if (with_bold) {
        if (E->n_mon == 0) {
            E->n_mon = n_mon;
            E->bw.assign((size_t)n_mon * N * 4, 0.0f);
            for (size_t k = 0; k < (size_t)n_mon * N; ++k) E->bw[k * 4 + 1] = E->bw[k * 4 + 2] = E->bw[k * 4 + 3] = 1.0f;
            E->native.assign(n_mon, std::vector<float>());
            E->times.clear();
        } else if (E->n_mon != n_mon) {
            hb_result_free(out);
            return efail(HB_INVALID_STATE, "the BOLD monitors of this engine were set up for " +
                                               std::to_string(E->n_mon) + " series, this run asks for " +
                                               std::to_string(n_mon) + " (restart the engine)");
        }
    }
    std::vector<float> nat((size_t)std::max(n_mon, 1) * std::max(rows, 1) * N, 0.0f);
    std::vector<int32_t> bold_n(n, 0);
    for (int i = 0; i < n; ++i) {
        hb_instance& I = B.inst[i];
        I.tavg = out->tavg + (size_t)i * per;
        I.final_state = out->final_states + (size_t)i * per;
        I.traj = traj ? traj_tmp.data() + (size_t)i * per * (size_t)n_steps : nullptr;
        if (with_bold && !mean_bold) {
            I.bold_tr_ms = 0.0;   // native samples; resampled over the whole series below
            I.bold = nat.data() + (size_t)i * rows * N;
            I.bold_cap = rows;
            I.bold_n = &bold_n[i];
            I.bold_state = E->bw.data() + (size_t)i * N * 4;
        } else {
            I.bold_period = 0;
        }
    }
    hb_run_opts o;
    memset(&o, 0, sizeof(o));
    o.device = E->device;
    o.flags = flags & ~(HB_FLAG_BOLD_BATCH_MEAN | HB_FLAG_RESTART);
    if (with_bold && mean_bold) o.flags |= kFlagDeviceTrajectory;
    hb_run_stats st;
    memset(&st, 0, sizeof(st));
    hb_plan* Pn = nullptr;
    int rc = hb_plan_create(B.inst.data(), n, &o, &Pn);
    if (!rc) rc = hb_plan_run(Pn, &st);
    if (!rc && with_bold && mean_bold && rows > 0)
        rc = plan_bold_batch_mean(Pn, P, E->dt, E->bw.data(), nat.data(), o.flags & HB_FLAG_EXACT);
    if (!rc) rc = hb_plan_download(Pn);
    hb_plan_destroy(Pn);
    if (rc) {
        hb_result_free(out);
        return rc;
    }
    if (traj) {   // [B][T][N][V] -> [T][B][N][V] (engine.rs:778-782 appends per step)
        for (int i = 0; i < n; ++i)
            for (int64_t t = 0; t < n_steps; ++t)
                memcpy(out->trajectories + ((size_t)t * n + i) * per, traj_tmp.data() + ((size_t)i * n_steps + t) * per, per * 4);
    }
    // the engine continues from here on the next run
    E->step += n_steps;
    E->n_state = n;
    E->state.assign(out->final_states, out->final_states + per * n);
    E->hist = std::move(hist);
    if (with_bold) {
        // sample times as BoldMonitor forms them: full windows at
        // len * dtb + dtb, the flushed partial window at len * dtb + dt * rem
        // (bold_monitor.rs:100, 214)
        const double dtb = E->dt * (double)P;
        for (int k = 0; k < rows; ++k) {
            const double len = (double)E->times.size();
            E->times.push_back(k < n_full ? len * dtb + dtb : len * dtb + E->dt * (double)rem);
        }
        for (int m = 0; m < n_mon; ++m)
            E->native[m].insert(E->native[m].end(), nat.begin() + (size_t)m * rows * N,
                                nat.begin() + (size_t)(m + 1) * rows * N);
        const double tr_ms = E->bold_tr_s * 1000.0;   // bold_monitor.rs:131 (tr * 1000)
        const int n_out = downsample_to_tr(E->times, E->native[0], N, dtb, tr_ms, nullptr);
        out->n_bold = n_out;
        out->bold_series = n_mon;
        out->bold = (float*)calloc((size_t)n_mon * n_out * N + 1, 4);
        for (int m = 0; m < n_mon; ++m)
            downsample_to_tr(E->times, E->native[m], N, dtb, tr_ms, out->bold + (size_t)m * n_out * N);
    }
    out->step = E->step;
    out->kernel_ms = st.kernel_ms;
    out->elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return HB_OK;
}

}  // namespace

extern "C" {

int hb_engine_create(const hb_sim_config* cfg, int32_t n_sweep, int32_t device, hb_engine** out) {
    if (!out) return HB_INVALID_CONFIG;
    *out = nullptr;
    if (!cfg) return HB_INVALID_CONFIG;
    // SimConfig::validate subset, config.rs:286-428
    if (!(cfg->dt > 0.0)) return efail(HB_INVALID_CONFIG, "dt must be positive");
    if (!(cfg->sim_length > 0.0)) return efail(HB_INVALID_CONFIG, "sim_length must be positive");
    if (cfg->n_subnetworks < 1 || !cfg->subnetworks) return efail(HB_INVALID_CONFIG, "At least one subnetwork required");
    if (cfg->n_subnetworks > 1) return efail(HB_INVALID_CONFIG, "multiple subnetworks are not supported by this backend");
    if (n_sweep < 1) return efail(HB_INVALID_CONFIG, "n_sweep must be >= 1");
    const hb_subnetwork& S = cfg->subnetworks[0];
    if (S.model < 0 || S.model > HB_RWW) return efail(HB_INVALID_CONFIG, "Unknown model");
    if (S.n_params != model_nparams(S.model) || !S.params)
        return efail(HB_INVALID_CONFIG, "model parameter count mismatch");
    if (S.n_nodes < 1 || !S.x0) return efail(HB_INVALID_CONFIG, "subnetwork needs nodes and an initial state");
    if (cfg->integrator < HB_EULER || cfg->integrator > HB_RK4_STOCH) return efail(HB_INVALID_CONFIG, "unknown integrator");
    hb_engine* E = new hb_engine();
    E->device = device;
    E->n_sweep = n_sweep;
    E->sim_length = cfg->sim_length;
    E->dt = cfg->dt;
    E->integrator = cfg->integrator;
    E->model = S.model;
    E->n_nodes = S.n_nodes;
    E->nv = model_nvar(S.model);
    E->params.assign(S.params, S.params + S.n_params);
    E->x0.assign(S.x0, S.x0 + (size_t)S.n_nodes * E->nv);
    E->nsig.assign(E->nv, 0.0f);
    if (cfg->n_nsig == 1 && cfg->nsig) E->nsig.assign(E->nv, cfg->nsig[0]);
    else if (cfg->n_nsig == E->nv && cfg->nsig) E->nsig.assign(cfg->nsig, cfg->nsig + E->nv);
    else if (cfg->n_nsig != 0) { delete E; return efail(HB_INVALID_CONFIG, "nsig per-variable length does not match nvar"); }
    E->bold_period = cfg->bold_period;
    E->bold_tr_s = cfg->bold_tr_s;
    E->engine = cfg->engine;
    E->delay_mode = cfg->delay_mode;
    E->noise_seed = cfg->noise_seed;
    E->noise_shared = cfg->noise_shared;
    if (cfg->n_projections > 1) { delete E; return efail(HB_INVALID_CONFIG, "at most one (self) projection is supported"); }
    if (cfg->n_projections == 1) {
        const hb_projection& Pj = cfg->projections[0];
        const int N = S.n_nodes;
        E->has_proj = true;
        E->cfn = Pj.cfn;
        E->cparams.assign(Pj.cparams, Pj.cparams + std::max(0, std::min(8, (int)Pj.n_cparams)));
        if (Pj.wkind == HB_W_SCALAR_MEAN || Pj.wkind == HB_W_SCALAR_DIAG) {
            E->wkind_cfg = HB_W_SCALAR_MEAN;
            E->scalar_w = Pj.scalar_w;
        } else if (Pj.wkind == HB_W_DENSE) {
            if (!Pj.dense) { delete E; return efail(HB_INVALID_CONFIG, "dense weights missing"); }
            E->wkind_cfg = HB_W_DENSE;
            E->dense.assign(Pj.dense, Pj.dense + (size_t)N * N);
        } else if (Pj.wkind == HB_W_CSR) {
            if (!Pj.indptr || Pj.nnz < 0 || (Pj.nnz > 0 && (!Pj.indices || !Pj.data))) { delete E; return efail(HB_INVALID_CONFIG, "CSR arrays missing"); }
            if (Pj.indptr[N] != Pj.nnz) { delete E; return efail(HB_INVALID_CONFIG, "CSR indptr length != tgt_nnodes + 1"); }
            E->wkind_cfg = HB_W_CSR;
            E->indptr.assign(Pj.indptr, Pj.indptr + N + 1);
            E->indices.assign(Pj.indices, Pj.indices + Pj.nnz);
            E->data.assign(Pj.data, Pj.data + Pj.nnz);
            if (Pj.cfn == HB_CF_DIFFERENCE) {
                // the batch engine densifies CSR before the Difference rewrite
                // (engine.rs:298-350); keep that arithmetic by running it dense
                E->dense.assign((size_t)N * N, 0.0f);
                for (int r = 0; r < N; ++r)
                    for (int e = Pj.indptr[r]; e < Pj.indptr[r + 1]; ++e) E->dense[(size_t)r * N + Pj.indices[e]] = Pj.data[e];
                E->wkind_cfg = HB_W_DENSE;
            }
        } else {
            E->has_proj = false;
        }
        if (Pj.delays && Pj.n_delays > 0) {
            E->delays.assign(Pj.delays, Pj.delays + Pj.n_delays);
            for (int32_t d : E->delays)
                if (d < 0 || d > 65535) { delete E; return efail(HB_INVALID_CONFIG, "delay out of range"); }
        }
    }
    *out = E;
    return HB_OK;
}

void hb_engine_destroy(hb_engine* e) { delete e; }

int hb_run_sweep(hb_engine* E, int32_t sub_idx, int32_t param_idx, const float* values,
                 int32_t n_values, int64_t n_steps, uint32_t flags, hb_result* out) {
    if (!E || !out) return efail(HB_INVALID_STATE, "null engine or result");
    if (n_values != E->n_sweep || !values)
        return efail(HB_INVALID_CONFIG, "number of sweep values must equal n_sweep");
    if (sub_idx != 0) return efail(HB_INVALID_CONFIG, "subnetwork index out of range");
    if (n_steps < 0 || n_steps > INT32_MAX) return efail(HB_INVALID_CONFIG, "n_steps out of range");
    Built B;
    B.inst.reserve(n_values);
    // Only the model's batch-native index is applied; any other index is
    // ignored, as in dfun_batch (dfun.rs:49-72).
    if (flags & HB_FLAG_RESTART) reset_state(E);
    if (E->n_state > 0 && E->n_state != n_values)
        return efail(HB_INVALID_STATE, "the engine holds " + std::to_string(E->n_state) +
                                           " sweep points (left by run_sweep_multi); this run has " +
                                           std::to_string(n_values));
    const bool native = param_idx == model_sweep_index(E->model);
    const size_t per = (size_t)E->n_nodes * E->nv;
    for (int i = 0; i < n_values; ++i) {
        hb_instance I = base_instance(E, n_steps);
        if (native) { I.sweep_form = 1; I.sweep_val = values[i]; }
        I.noise_stream = E->noise_shared ? 0u : (uint64_t)i;
        // continue where the previous run stopped (engine.rs:496-540); fresh
        // noise for the new steps (the reference draws it per step)
        I.step0 = E->step;
        I.noise_seed = E->noise_seed + (uint64_t)E->step;
        if (E->n_state > 0) I.x0 = E->state.data() + (size_t)i * per;
        B.inst.push_back(I);
    }
    return run_built(E, B, n_steps, flags, !E->monitors_dropped, out);
}

int hb_run_sweep_multi(hb_engine* E, const hb_sweep_point* pts, int32_t n_points, int64_t n_steps,
                       uint32_t flags, hb_result* out) {
    if (!E || !out) return efail(HB_INVALID_STATE, "null engine or result");
    if (n_points < 1 || !pts) return efail(HB_INVALID_CONFIG, "empty sweep spec");
    if (n_steps < 0 || n_steps > INT32_MAX) return efail(HB_INVALID_CONFIG, "n_steps out of range");
    int64_t total = 1;
    int n_model = 0;
    for (int d = 0; d < n_points; ++d) {
        if (pts[d].n_values < 1 || !pts[d].values) return efail(HB_INVALID_CONFIG, "sweep dimension without values");
        if (pts[d].kind == 0) {
            if (pts[d].a != 0 || pts[d].b < 0 || pts[d].b >= model_nparams(E->model))
                return efail(HB_INVALID_CONFIG, "sweep parameter index out of range");
            ++n_model;
        } else if (pts[d].kind == 1) {
            if (!E->has_proj || pts[d].a != 0 || pts[d].b < 0 || pts[d].b >= 8)
                return efail(HB_INVALID_CONFIG, "coupling parameter index out of range");
        } else {
            return efail(HB_INVALID_CONFIG, "unknown sweep point kind");
        }
        total *= pts[d].n_values;
        if (total > (int64_t)1 << 24) return efail(HB_INVALID_CONFIG, "sweep too large");
    }
    if (flags & HB_FLAG_RESTART) reset_state(E);
    // run_sweep_multi builds a fresh engine whose points all start from sweep
    // point 0 of the current state, at step 0 with empty histories and no
    // monitors, then takes over its states, step and histories
    // (engine.rs:886-909, 1326-1331)
    std::vector<float> x0 = E->n_state > 0 ? std::vector<float>(E->state.begin(), E->state.begin() + E->n_nodes * E->nv)
                                           : E->x0;
    reset_state(E);
    E->monitors_dropped = true;
    E->x0_multi = std::move(x0);
    Built B;
    B.inst.reserve((size_t)total);
    B.params.resize((size_t)total);
    std::vector<int> idx(n_points, 0);
    for (int64_t s = 0; s < total; ++s) {
        hb_instance I = base_instance(E, n_steps);
        I.x0 = E->x0_multi.data();
        std::vector<float>& p = B.params[(size_t)s];
        p = E->params;
        for (int d = 0; d < n_points; ++d) {
            const float v = pts[d].values[idx[d]];
            if (pts[d].kind == 0) {
                if (n_model == 1) {
                    // dfun_batch_dispatch with one swept parameter: native index only
                    if (pts[d].b == model_sweep_index(E->model)) { I.sweep_form = 1; I.sweep_val = v; }
                } else {
                    p[pts[d].b] = v;   // dfun_batch_multi_fallback: per-point params (dfun.rs:99-133)
                }
            } else {
                const int k = pts[d].b;
                if (k >= I.n_cparams) {
                    for (int z = I.n_cparams; z <= k; ++z) I.cparams[z] = 0.0f;
                    I.n_cparams = k + 1;
                }
                I.cparams[k] = v;
            }
        }
        I.noise_stream = E->noise_shared ? 0u : (uint64_t)s;
        B.inst.push_back(I);
        for (int d = n_points - 1; d >= 0; --d) {   // last dimension fastest (engine.rs:69-90)
            if (++idx[d] < pts[d].n_values) break;
            idx[d] = 0;
        }
    }
    for (int64_t s = 0; s < total; ++s) B.inst[(size_t)s].params = B.params[(size_t)s].data();
    return run_built(E, B, n_steps, flags & ~HB_FLAG_BOLD_BATCH_MEAN, false, out);
}

int hb_engine_get_state(const hb_engine* E, hb_engine_state* st, float* nsig, float* states, float* hist) {
    if (!E || !st) return efail(HB_INVALID_STATE, "null engine or state");
    memset(st, 0, sizeof(*st));
    st->step = E->step;
    st->dt = E->dt;
    st->integrator = E->integrator;
    st->n_state = E->n_state;
    st->n_nodes = E->n_nodes;
    st->nvar = E->nv;
    const hb_instance I = base_instance(E, 0);
    st->hist_floats = hb_history_floats(&I);
    const size_t per = (size_t)E->n_nodes * E->nv;
    if (nsig) memcpy(nsig, E->nsig.data(), (size_t)E->nv * 4);
    if (states) {
        if (E->n_state > 0) memcpy(states, E->state.data(), per * E->n_state * 4);
        else memcpy(states, E->x0.data(), per * 4);
    }
    if (hist && E->n_state > 0 && st->hist_floats > 0) {
        for (int i = 0; i < E->n_state; ++i) {
            float* dst = hist + (size_t)i * st->hist_floats;
            // a run that never used the history (step 0, or a run_sweep_multi
            // left nothing) holds none: that is a zero history
            if (i < (int)E->hist.size() && (int64_t)E->hist[i].size() == st->hist_floats)
                memcpy(dst, E->hist[i].data(), (size_t)st->hist_floats * 4);
            else
                memset(dst, 0, (size_t)st->hist_floats * 4);
        }
    }
    return HB_OK;
}

int hb_engine_set_state(hb_engine* E, const hb_engine_state* st, const float* nsig, int32_t n_nsig,
                        const float* states, const float* hist) {
    if (!E || !st) return efail(HB_INVALID_STATE, "null engine or state");
    if (st->step < 0) return efail(HB_INVALID_STATE, "negative step");
    if (!(st->dt > 0.0)) return efail(HB_INVALID_STATE, "dt must be positive");
    if (st->integrator < HB_EULER || st->integrator > HB_RK4_STOCH)
        return efail(HB_INVALID_STATE, "Unknown integrator kind in checkpoint: " + std::to_string(st->integrator));
    if (st->n_state < 1 || st->n_state > (1 << 24) || !states)
        return efail(HB_INVALID_STATE, "a checkpoint needs at least one state");
    if (st->n_nodes != E->n_nodes || st->nvar != E->nv)
        return efail(HB_INVALID_STATE, "Checkpoint shape mismatch for subnetwork 0: expected (nvar=" +
                                           std::to_string(E->nv) + ",nnodes=" + std::to_string(E->n_nodes) +
                                           ",nmodes=1), got (" + std::to_string(st->nvar) + "," +
                                           std::to_string(st->n_nodes) + ",1)");
    if (n_nsig != 0 && n_nsig != 1 && n_nsig != E->nv)
        return efail(HB_INVALID_STATE, "nsig per-variable length does not match nvar");
    if (n_nsig > 0 && !nsig) return efail(HB_INVALID_STATE, "nsig missing");
    const size_t per = (size_t)E->n_nodes * E->nv;
    // the history layout follows the engine's configuration; dt and the
    // integrator do not enter it
    const hb_instance I = base_instance(E, 0);
    const int64_t hf = hb_history_floats(&I);
    if (hist && st->hist_floats != hf)
        return efail(HB_INVALID_STATE, "checkpoint history holds " + std::to_string(st->hist_floats) +
                                           " floats per point, this engine keeps " + std::to_string(hf));
    const int n = (st->n_state == 1 && E->n_sweep > 1) ? E->n_sweep : st->n_state;
    E->state.resize(per * n);
    E->hist.assign(hf > 0 ? n : 0, std::vector<float>());
    for (int i = 0; i < n; ++i) {
        const int src = st->n_state == 1 ? 0 : i;
        memcpy(E->state.data() + per * i, states + per * src, per * 4);
        if (hf > 0) {
            if (hist) E->hist[i].assign(hist + (size_t)src * hf, hist + (size_t)(src + 1) * hf);
            else E->hist[i].assign((size_t)hf, 0.0f);
        }
    }
    E->n_state = n;
    E->step = st->step;
    E->dt = st->dt;
    E->integrator = st->integrator;
    if (n_nsig == 1) E->nsig.assign(E->nv, nsig[0]);
    else if (n_nsig == E->nv) E->nsig.assign(nsig, nsig + E->nv);
    return HB_OK;
}

void hb_result_free(hb_result* r) {
    if (!r) return;
    free(r->tavg);
    free(r->final_states);
    free(r->trajectories);
    free(r->bold);
    r->tavg = r->final_states = r->trajectories = r->bold = nullptr;
}

int32_t hb_bold_downsample(const double* times, int32_t n, const float* native, int32_t n_nodes, double dtb_ms,
                           double tr_ms, float* out, int32_t cap) {
    if (n <= 0 || n_nodes < 1 || !times || !native) return 0;
    const std::vector<double> tm(times, times + n);
    const std::vector<float> data(native, native + (size_t)n * n_nodes);
    const int rows = downsample_to_tr(tm, data, n_nodes, dtb_ms, tr_ms, nullptr);
    if (!out) return rows;
    if (rows > cap) return -1;
    return downsample_to_tr(tm, data, n_nodes, dtb_ms, tr_ms, out);
}

}  // extern "C"
