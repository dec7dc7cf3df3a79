// hb_host.cpp -- host side of the C ABI: validation, HBM layout, cost-ordered
// work lists, launch shapes, hipGraph capture, timing and the
// BatchHybridEngine-style sweep API.
//
// Reference interfaces mirrored (maedoc/hyburn):
//   BatchHybridEngine::from_config / run_sweep / run_sweep_multi
//       src/engine/batch_engine/engine.rs:145-471, 477-494, 845-1348
//   SimulationError codes            src/error.rs:7-23
//   select_backend                   src/cli.rs:259-271
#include "hb_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

using namespace hb;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

int hb::set_error(int code, const std::string& msg) { return fail(code, msg); }

namespace {

#define HB_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(HB_BACKEND_ERROR, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Arena layout builder: records (offset, size) requests, one allocation later.
struct Arena {
    size_t top = 0;
    size_t take(size_t bytes) {
        size_t off = top;
        top = align_up(top + (bytes ? bytes : 1), 256);
        return off;
    }
};

bool is_stoch(int integ) { return integ >= HB_EULER_STOCH; }

// pre(0) computed with the host libm, as the reference computes pre() on a zero
// tensor for single-delay reads of not-yet-simulated time (engine.rs:580-583).
void pre_zero(int cfn, const float* cc, int classic, int np, float* out) {
    out[0] = out[1] = 0.0f;
    switch (cfn) {
    case HB_CF_KURAMOTO:
        out[0] = sinf(0.0f);
        out[1] = cosf(0.0f);
        break;
    case HB_CF_TANH:
        for (int c = 0; c < np; ++c) out[c] = (tanhf((0.0f * cc[1] + (-cc[2])) / cc[3]) + 1.0f) * cc[0];
        break;
    case HB_CF_SJR: {
        const float den = expf((0.0f + (-cc[4])) * (-cc[3])) + 1.0f;
        out[0] = classic ? (1.0f / den) * cc[2] + cc[1] : (1.0f / den) * cc[1];
        break; }
    case HB_CF_PRESIGMOIDAL:
        for (int c = 0; c < np; ++c) out[c] = (tanhf((0.0f * cc[3] + (-cc[4])) * cc[2]) + cc[1]) * cc[0];
        break;
    default:
        break;
    }
}


}  // namespace

// ============================================================== plan ========

struct hb_plan {
    int device = 0;
    uint32_t flags = 0;
    std::vector<hb_instance> inst;
    std::vector<DevInst> dev;
    char* arena = nullptr;
    size_t arena_bytes = 0;
    DevInst* d_inst = nullptr;
    int32_t* d_order = nullptr;
    int32_t* d_counters = nullptr;
    NoiseJob* d_jobs = nullptr;
    int n_jobs = 0;
    int64_t max_noise = 0;
    std::vector<LaunchDesc> launches;
    hipStream_t stream = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    // launch groups run as parallel branches on up to kLanes side streams (one
    // join event each); lane_of[k] is the lane of launch k, estimated time
    // balanced (longest group first onto the least loaded lane)
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> join;
    std::vector<int> lane_of;
    std::vector<double> launch_cost;
    hipEvent_t fork = nullptr;
    hipGraphExec_t g_noise = nullptr, g_fwd = nullptr;
    hipGraph_t gr_noise = nullptr, gr_fwd = nullptr;
    struct Out {
        size_t tavg, fin, traj, bold;
        size_t n_state, n_traj, n_bold;
    };
    std::vector<Out> outs;
    double node_steps = 0.0;
    int block_main = 0, lds_main = 0;
    bool noise_done = false;   // HB_FLAG_KEEP_NOISE: the noise was generated by an earlier run
    std::vector<size_t> hist_off, bold_off;   // carried history / BOLD state regions (or SIZE_MAX)
};

namespace {

// ---- launch-shape autotuner ------------------------------------------------
// A shape class is (model, ring channels, nodes, delay-blocked, ring in LDS,
// integrator class, stochastic). Its tunable is the lanes per node of the
// delay-blocked kernel (1, 2, 4) or the block size of a network larger than
// one block (256, 512). Choices are timed on a short sub-plan of sampled
// instances and cached for the process.
struct TuneKey {
    int model, np, n_nodes, blocked, lring, integ, stoch;
    bool operator<(const TuneKey& o) const {
        return std::tie(model, np, n_nodes, blocked, lring, integ, stoch) <
               std::tie(o.model, o.np, o.n_nodes, o.blocked, o.lring, o.integ, o.stoch);
    }
};
struct TuneResult {
    int choice;        // lanes per node (blocked) or block size (multi)
    double ms[3];      // measured kernel time of each candidate (0 if not run)
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneResult> g_tuned;
// Forced choice for the sub-plans the tuner times (thread-local: one tuning
// at a time per thread).
thread_local int g_force_choice = 0;

int validate_instance(const hb_instance& I, int idx) {
    char buf[256];
    auto bad = [&](const char* what) {
        snprintf(buf, sizeof(buf), "instance %d: %s", idx, what);
        return fail(HB_INVALID_CONFIG, buf);
    };
    if (I.model < 0 || I.model > HB_RWW) return bad("unknown model");
    if (I.integrator < HB_EULER || I.integrator > HB_RK4_STOCH) return bad("unknown integrator");
    if (I.n_nodes < 1 || I.n_nodes > 65535) return bad("n_nodes must be in [1, 65535]");
    if (I.n_steps < 0) return bad("n_steps must be >= 0");
    if (!(I.dt > 0.0f)) return bad("dt must be positive");
    if (!I.params || !I.x0) return bad("params and x0 are required");
    if (I.wkind < HB_W_NONE || I.wkind > HB_W_CSR) return bad("unknown weight kind");
    if (I.cfn < HB_CF_LINEAR || I.cfn > HB_CF_PRESIGMOIDAL) return bad("unknown coupling function");
    if (const char* why = presigmoidal_unsupported(I.cfn, I.n_cparams, I.cparams)) return bad(why);
    const int nc = model_ncvar(I.model);
    const int np = coupling_npre(I.cfn, nc);
    if (np > 2 || (I.cfn == HB_CF_SJR && nc != 1)) return bad("unsupported model/coupling-function combination");
    if (I.cfn == HB_CF_KURAMOTO && I.model != HB_KURAMOTO && nc != 1) return bad("unsupported model/coupling-function combination");
    if (I.cfn == HB_CF_KURAMOTO && I.model != HB_KURAMOTO) return bad("Kuramoto coupling is only instantiated for the Kuramoto model");
    if (I.wkind == HB_W_DENSE && !I.dense) return bad("dense weights missing");
    if (I.wkind == HB_W_CSR) {
        if (!I.indptr || (I.nnz > 0 && (!I.indices || !I.weights))) return bad("CSR arrays missing");
        if (I.indptr[0] != 0 || I.indptr[I.n_nodes] != I.nnz) return bad("CSR indptr inconsistent with nnz");
        for (int i = 0; i < I.n_nodes; ++i)
            if (I.indptr[i + 1] < I.indptr[i]) return bad("CSR indptr not monotone");
        for (int e = 0; e < I.nnz; ++e) {
            if (I.indices[e] < 0 || I.indices[e] >= I.n_nodes) return bad("CSR index out of range");
            if (I.edge_delays && (I.edge_delays[e] < 0 || I.edge_delays[e] > 65535)) return bad("edge delay out of range [0, 65535]");
        }
    }
    if (I.cfn == HB_CF_DIFFERENCE && I.wkind == HB_W_CSR)
        return bad("Difference coupling with CSR weights is not supported");
    if (I.delay_mode < HB_DELAY_PER_EDGE || I.delay_mode > HB_DELAY_BATCH_COMPAT) return bad("unknown delay mode");
    if (I.single_delay < 0 || I.single_delay > 65535) return bad("delay out of range");
    if (I.max_delay_hint < 0 || I.max_delay_hint > 65535) return bad("delay hint out of range");
    if (I.node_param_idx >= model_nparams(I.model)) return bad("node parameter index out of range");
    if (I.node_param_idx >= 0 && !I.node_param) return bad("node parameter values missing");
    if (is_stoch(I.integrator) && !I.nsig) return bad("stochastic integrator needs nsig");
    if (I.bold_period < 0) return bad("bold_period must be >= 0");
    if (I.step0 < 0) return bad("step0 must be >= 0");
    return HB_OK;
}

// Largest delay the ring must hold. For BATCH_COMPAT the reference's window
// horizon is max(all delays) + 1 even though only delays[0] is read
// (engine.rs:253-270, 410).
int max_delay_of(const hb_instance& I) {
    if (I.wkind == HB_W_NONE) return 0;
    if (I.delay_mode == HB_DELAY_BATCH_COMPAT) {
        int m = std::max(I.single_delay, I.max_delay_hint);
        if (I.edge_delays)
            for (int e = 0; e < I.nnz; ++e) m = std::max(m, (int)I.edge_delays[e]);
        return m;
    }
    if (I.wkind == HB_W_CSR && I.delay_mode == HB_DELAY_PER_EDGE) {
        int m = 0;
        if (I.edge_delays)
            for (int e = 0; e < I.nnz; ++e) m = std::max(m, (int)I.edge_delays[e]);
        return m;
    }
    return I.single_delay;
}

// Floats of the carried delay history: max_delay steps x nodes x ring channels.
int64_t history_floats(const hb_instance& I, int np) {
    return (int64_t)max_delay_of(I) * I.n_nodes * np;
}

TuneKey tune_key(const hb_instance& I, int np, int blocked, int lring) {
    const int ic = (I.integrator == HB_RK4 || I.integrator == HB_RK4_STOCH) ? 2
                   : (I.integrator == HB_HEUN || I.integrator == HB_HEUN_STOCH) ? 1 : 0;
    return TuneKey{I.model, np, I.n_nodes, blocked, lring, ic, I.integrator >= HB_EULER_STOCH ? 1 : 0};
}

}  // namespace

extern "C" int hb_plan_create(const hb_instance* inst, int32_t n, const hb_run_opts* opts, hb_plan** out);
extern "C" int hb_plan_run(hb_plan* p, hb_run_stats* stats);
extern "C" void hb_plan_destroy(hb_plan* p);

namespace {

// Launch every group: one group on the plan stream, several as parallel
// branches (fork from the plan stream onto the lanes' side streams, join
// back), so groups of different launch shapes share the chip.
thread_local std::string g_launch_err;   // the step of launch_groups that failed

hipError_t launch_groups(hb_plan* P, bool exact) {
    auto go = [&](const LaunchDesc& L, hipStream_t st) {
        if (L.mfma) return launch_dense(L, st);
        return exact ? launch_forward_exact(L, st) : launch_forward_fast(L, st);
    };
    g_launch_err.clear();
    if (P->launches.size() <= 1 || P->side.empty()) {
        for (auto& L : P->launches) {
            const hipError_t e = go(L, P->stream);
            if (e != hipSuccess) {
                g_launch_err = "kernel launch (model " + std::to_string(L.model) + ", np " + std::to_string(L.np) +
                               ", block " + std::to_string(L.block) + ", lds " + std::to_string(L.lds_bytes) + ")";
                return e;
            }
        }
        return hipSuccess;
    }
    hipError_t e = hipEventRecord(P->fork, P->stream);
    if (e != hipSuccess) g_launch_err = "fork record";
    for (size_t l = 0; l < P->side.size() && e == hipSuccess; ++l) {
        e = hipStreamWaitEvent(P->side[l], P->fork, 0);
        if (e != hipSuccess) g_launch_err = "branch wait of lane " + std::to_string(l);
        // the lane's groups, longest first
        std::vector<size_t> mine;
        for (size_t k = 0; k < P->launches.size(); ++k)
            if ((size_t)P->lane_of[k] == l) mine.push_back(k);
        std::stable_sort(mine.begin(), mine.end(), [&](size_t a, size_t b) { return P->launch_cost[a] > P->launch_cost[b]; });
        for (size_t k : mine) {
            if (e != hipSuccess) break;
            const LaunchDesc& L = P->launches[k];
            e = go(L, P->side[l]);
            if (e != hipSuccess)
                g_launch_err = "kernel launch of group " + std::to_string(k) + " (model " + std::to_string(L.model) +
                               ", np " + std::to_string(L.np) + ", block " + std::to_string(L.block) + ", grid " +
                               std::to_string(L.grid) + ", lds " + std::to_string(L.lds_bytes) + ")";
        }
        if (e == hipSuccess) {
            e = hipEventRecord(P->join[l], P->side[l]);
            if (e == hipSuccess) e = hipStreamWaitEvent(P->stream, P->join[l], 0);
            if (e != hipSuccess) g_launch_err = "join of lane " + std::to_string(l);
        }
    }
    return e;
}

// Time each candidate of every untuned shape class on a sub-plan of up to 32
// sampled instances truncated to 512 steps (outputs into scratch), keep the
// fastest. Two runs per candidate; the second is timed.
int autotune_plan(const hb_instance* inst, int n, const hb_run_opts* opts, int lds_budget) {
    struct Cls { std::vector<int> members; std::vector<int> cand; };
    std::map<TuneKey, Cls> todo;
    for (int i = 0; i < n; ++i) {
        const hb_instance& I = inst[i];
        const int np = ring_channels(I.model, I.cfn, I.params, I.node_param_idx, -1);
        const bool per_edge = I.wkind == HB_W_CSR && I.delay_mode == HB_DELAY_PER_EDGE && I.edge_delays;
        TuneKey key;
        std::vector<int> cand;
        int dmin = 1 << 30;
        if (per_edge)
            for (int e = 0; e < I.nnz; ++e) dmin = std::min(dmin, (int)I.edge_delays[e]);
        const bool blockable = per_edge && dmin >= 8 && !(opts->flags & HB_FLAG_NO_BLOCKING) &&
                               (int)align_up((size_t)I.n_nodes, 64) <= kWideBlockMax;
        if (I.n_nodes > kBlockMax && !blockable) {
            key = tune_key(I, np, 0, 0);
            cand = {256, 512};
        } else if (I.wkind == HB_W_DENSE && max_delay_of(I) == 0 && !(opts->flags & HB_FLAG_EXACT) &&
                   !(opts->flags & HB_FLAG_NO_MFMA) && I.node_param_idx < 0 && !(I.bold && I.bold_period > 0) &&
                   I.n_nodes <= kDenseMaxNodes) {
            key = tune_key(I, np, 2, 0);
            cand = {kChoiceLanes, kChoiceMfma};
        } else if (blockable && !(opts->flags & HB_FLAG_EXACT)) {
            const int64_t ring_k = (int64_t)(max_delay_of(I) + 16) * I.n_nodes * np * 4;
            key = tune_key(I, np, 1, ring_k <= lds_budget);
            for (int l : {1, 2, 4})
                if ((int)align_up((size_t)I.n_nodes * l, 64) <= kWideBlockMax) cand.push_back(l);
        } else {
            continue;
        }
        if (cand.size() < 2) continue;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            if (g_tuned.count(key)) continue;
        }
        Cls& c = todo[key];
        c.cand = cand;
        if (c.members.size() < 32) c.members.push_back(i);
    }
    for (auto& kv : todo) {
        const Cls& c = kv.second;
        std::vector<hb_instance> sub;
        std::vector<std::vector<float>> scratch;
        for (int i : c.members) {
            hb_instance J = inst[i];
            J.n_steps = std::min(J.n_steps, 512);
            J.traj = nullptr;
            J.bold = nullptr;
            J.bold_period = 0;
            J.bold_n = nullptr;
            scratch.emplace_back((size_t)J.n_nodes * model_nvar(J.model) * 2);
            J.tavg = scratch.back().data();
            J.final_state = scratch.back().data() + (size_t)J.n_nodes * model_nvar(J.model);
            sub.push_back(J);
        }
        hb_run_opts o = *opts;
        o.flags = (opts->flags & ~(uint32_t)HB_FLAG_AUTOTUNE) | HB_FLAG_NO_GRAPH;
        o.block_size = 0;
        TuneResult best{0, {0.0, 0.0, 0.0}};
        double best_ms = 1e300;
        for (size_t k = 0; k < c.cand.size() && k < 3; ++k) {
            g_force_choice = c.cand[k];
            hb_plan* Q = nullptr;
            int rc = hb_plan_create(sub.data(), (int)sub.size(), &o, &Q);
            hb_run_stats st;
            memset(&st, 0, sizeof(st));
            if (!rc) rc = hb_plan_run(Q, &st);
            if (!rc) rc = hb_plan_run(Q, &st);
            hb_plan_destroy(Q);
            g_force_choice = 0;
            if (rc) return rc;
            best.ms[k] = st.kernel_ms;
            if (st.kernel_ms < best_ms) { best_ms = st.kernel_ms; best.choice = c.cand[k]; }
        }
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[kv.first] = best;
    }
    return HB_OK;
}

}  // namespace

extern "C" {

int hb_autotune_entries(int32_t* out, int32_t max_entries) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    int k = 0;
    for (const auto& kv : g_tuned) {
        if (out && k < max_entries) {
            int32_t* r = out + (size_t)k * 8;
            r[0] = kv.first.model; r[1] = kv.first.np; r[2] = kv.first.n_nodes; r[3] = kv.first.blocked;
            r[4] = kv.first.lring; r[5] = kv.first.integ; r[6] = kv.first.stoch; r[7] = kv.second.choice;
        }
        ++k;
    }
    return k;
}

void hb_autotune_clear(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.clear();
}

const char* hb_last_error(void) { return g_err.c_str(); }
const char* hb_version(void) { return "hyburn-hip 0.1 (gfx950)"; }

int hb_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int hb_device_synchronize(int32_t device) {
    HB_HIP(hipSetDevice(device));
    HB_HIP(hipDeviceSynchronize());
    return HB_OK;
}

const char* hb_select_backend(const char* name) {
    if (name && (!strcmp(name, "hip") || !strcmp(name, "rocm"))) return "hip";
    g_err = std::string("backend '") + (name ? name : "(null)") + "' is not provided by this library (HIP only)";
    return nullptr;
}

void hb_plan_destroy(hb_plan* p) {
    if (!p) return;
    if (p->g_noise) (void)hipGraphExecDestroy(p->g_noise);
    if (p->g_fwd) (void)hipGraphExecDestroy(p->g_fwd);
    if (p->gr_noise) (void)hipGraphDestroy(p->gr_noise);
    if (p->gr_fwd) (void)hipGraphDestroy(p->gr_fwd);
    for (auto& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : p->join)
        if (e) (void)hipEventDestroy(e);
    if (p->fork) (void)hipEventDestroy(p->fork);
    for (auto& st : p->side)
        if (st) (void)hipStreamDestroy(st);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    if (p->arena) (void)hipFree(p->arena);
    delete p;
}

int hb_plan_create(const hb_instance* inst, int32_t n, const hb_run_opts* opts, hb_plan** out) {
    if (!out) return fail(HB_INVALID_CONFIG, "null output pointer");
    *out = nullptr;
    if (n < 1 || !inst) return fail(HB_INVALID_CONFIG, "need at least one instance");
    for (int i = 0; i < n; ++i) {
        int rc = validate_instance(inst[i], i);
        if (rc) return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(HB_BACKEND_ERROR, "no HIP device available");
    const int device = opts ? opts->device : 0;
    if (device < 0 || device >= ndev) return fail(HB_INVALID_CONFIG, "device ordinal out of range");
    HB_HIP(hipSetDevice(device));
    int lds_max = 0, n_cu = 0;
    HB_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    HB_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device));
    const int lds_budget = lds_max - 512;   // static LDS of the kernels
    {
        static bool attrs_done[64] = {false};
        if (device < 64 && !attrs_done[device]) {
            HB_HIP(forward_set_attributes_fast(lds_budget));
            HB_HIP(forward_set_attributes_exact(lds_budget));
            HB_HIP(dense_set_attributes(lds_budget));
            attrs_done[device] = true;
        }
    }

    hb_plan* P = new hb_plan();
    P->device = device;
    P->flags = opts ? opts->flags : 0u;
    P->inst.assign(inst, inst + n);
    P->dev.resize(n);
    P->outs.resize(n);
    const int req_block = opts ? opts->block_size : 0;
    const bool exact = (P->flags & HB_FLAG_EXACT) != 0;
    const bool autotune = (P->flags & HB_FLAG_AUTOTUNE) != 0 && g_force_choice == 0;
    if (autotune) {
        const int rc = autotune_plan(inst, n, opts, lds_budget);
        if (rc) { hb_plan_destroy(P); return rc; }
    }

    // ---- per-instance shapes, LDS-vs-HBM ring, block size
    struct Shape {
        int nv, nc, np, depth, block, blocked, lpn;
        bool multi, lds_ring, mfma;
        int lds_bytes;
        double cost;
    };
    std::vector<Shape> sh(n);
    for (int i = 0; i < n; ++i) {
        const hb_instance& I = inst[i];
        Shape& s = sh[i];
        s.nv = model_nvar(I.model);
        s.nc = model_ncvar(I.model);
        s.np = ring_channels(I.model, I.cfn, I.params, I.node_param_idx, -1);
        s.depth = max_delay_of(I) + 2;
        int b = req_block > 0 ? req_block : (int)align_up((size_t)I.n_nodes, 64);
        b = std::max(64, std::min(kBlockMax, b));
        s.block = b;
        s.multi = I.n_nodes > b;
        // delay-blocked schedule: per-edge delays with min delay >= K, one node
        // (or lpn lanes) per thread; networks wider than 512 nodes take the
        // 1024-thread kernel
        s.blocked = 0;
        s.lpn = 1;
        const bool fits_wide = req_block <= 0 && (int)align_up((size_t)I.n_nodes, 64) <= kWideBlockMax;
        if ((!s.multi || fits_wide) && I.wkind == HB_W_CSR && I.delay_mode == HB_DELAY_PER_EDGE && I.edge_delays &&
            !(opts && (opts->flags & HB_FLAG_NO_BLOCKING))) {
            int dmin = 1 << 30;
            for (int e = 0; e < I.nnz; ++e) dmin = std::min(dmin, (int)I.edge_delays[e]);
            const int dmax = max_delay_of(I);
            const bool stoch = I.integrator >= HB_EULER_STOCH;
            // 32-step blocks exist for one-channel rings
            const bool k32 = s.np == 1;
            for (int K : {8, 16, 32}) {   // the kernel's delay-block lengths
                if (dmin < K) continue;
                const int64_t ring_k = (int64_t)(dmax + K + 1 + K - 1) * I.n_nodes * s.np * 4;
                // 8 when the ring fits LDS or the run is stochastic; 16 or 32 for
                // deterministic runs whose ring lives in HBM (64- or 128-byte windows)
                if (K == 8 && ring_k > lds_budget && dmin >= 16 && !stoch) continue;
                if (K == 16 && k32 && dmin >= 32) continue;
                if (K == 32 && !k32) continue;
                {
                    s.blocked = K;
                    s.depth = dmax + K + 1;
                    // lanes per node (FMA mode only: exact mode keeps the row's sequential sum):
                    // forced by the tuner's trials or tuned (up to 1024 threads), else the
                    // widest that fits 512 threads
                    if (!exact && req_block <= 0) {
                        int want = 0;
                        if (g_force_choice > 0) want = g_force_choice;
                        else if (autotune) {
                            std::lock_guard<std::mutex> lk(g_tune_mu);
                            const auto it = g_tuned.find(tune_key(I, s.np, 1, ring_k <= lds_budget));
                            if (it != g_tuned.end()) want = it->second.choice;
                        }
                        const int cap = want > 0 ? kWideBlockMax : kBlockMax;
                        for (int l : {4, 2})
                            if ((want == 0 || l <= want) && (int)align_up((size_t)I.n_nodes * l, 64) <= cap) { s.lpn = l; break; }
                    }
                    s.block = std::max(64, (int)align_up((size_t)I.n_nodes * s.lpn, 64));
                    s.multi = false;
                    // the 1024-thread kernel has 8- and 16-step blocks only
                    if (s.block > kBlockMax && K == 32) {
                        s.blocked = 16;
                        s.depth = dmax + 16 + 1;
                    }
                    break;
                }
            }
        }
        if (s.multi && req_block <= 0) {
            int want = 0;
            if (g_force_choice > 0) want = g_force_choice;
            else if (autotune) {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                const auto it = g_tuned.find(tune_key(I, s.np, 0, 0));
                if (it != g_tuned.end()) want = it->second.choice;
            }
            if (want == 256 || want == 512) {
                s.block = want;
                s.multi = I.n_nodes > s.block;
            }
        }
        const int tab_b = s.multi ? I.n_nodes * (2 * s.nv + 6) * 4 : 0;
        // ring floats: depth slots per node and channel (+ K-1 mirrored slots when blocked)
        const int ring_b = (s.depth + (s.blocked ? s.blocked - 1 : 0)) * I.n_nodes * s.np * 4;
        s.lds_ring = (int64_t)ring_b + tab_b <= lds_budget;
        // undelayed dense coupling of a batch sharing one matrix: the
        // matrix-core product over 16-instance tiles (the group size is
        // checked when the launch groups are formed)
        s.mfma = false;
        if (!exact && !(P->flags & HB_FLAG_NO_MFMA) && I.wkind == HB_W_DENSE && max_delay_of(I) == 0 &&
            I.node_param_idx < 0 && !(I.bold && I.bold_period > 0) && I.n_nodes <= kDenseMaxNodes && !s.multi &&
            s.np <= 2 && req_block <= 0) {
            int want = 0;
            if (g_force_choice > 0) want = g_force_choice;
            else if (autotune) {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                const auto it = g_tuned.find(tune_key(I, s.np, 2, 0));
                if (it != g_tuned.end()) want = it->second.choice;
            }
            s.mfma = want == 0 || want == kChoiceMfma;
        }
        s.lds_bytes = (s.lds_ring ? ring_b : 0) + tab_b;
        if (s.lds_bytes > lds_budget) {
            hb_plan_destroy(P);
            return fail(HB_INVALID_CONFIG, "instance state table exceeds LDS");
        }
        const int stages = (I.integrator == HB_RK4 || I.integrator == HB_RK4_STOCH) ? 4
                           : (I.integrator == HB_HEUN || I.integrator == HB_HEUN_STOCH) ? 2 : 1;
        const double nnz_r = I.wkind == HB_W_CSR ? (double)I.nnz / I.n_nodes
                             : I.wkind == HB_W_DENSE ? (double)I.n_nodes : 1.0;
        s.cost = (double)I.n_steps * ((double)((I.n_nodes + b - 1) / b) * (stages * s.nv * 8.0 + nnz_r * s.np * 2.0) + 64.0);
        P->node_steps += (double)I.n_nodes * (double)I.n_steps;
    }

    // ---- HBM layout (inputs deduplicated by host pointer)
    Arena A;
    std::map<std::tuple<const void*, const void*, const void*, size_t>, size_t> dedup;
    auto shared_off = [&](const void* a, const void* b, const void* c, size_t bytes, bool* fresh) {
        auto key = std::make_tuple(a, b, c, bytes);
        auto it = dedup.find(key);
        if (it != dedup.end()) { *fresh = false; return it->second; }
        size_t off = A.take(bytes);
        dedup[key] = off;
        *fresh = true;
        return off;
    };
    struct Upload { size_t off; std::vector<char> bytes; };
    std::vector<Upload> uploads;
    auto stage = [&](size_t off, const void* src, size_t bytes) {
        Upload u;
        u.off = off;
        u.bytes.assign((const char*)src, (const char*)src + bytes);
        uploads.push_back(std::move(u));
    };
    struct Ptrs {
        size_t x0, node_mc, indptr, edges, dense_t, noise, ring, tavg, fin, traj, bold, hist, bold_io;
        bool has_node_mc, has_noise, has_ring, has_traj, has_bold, has_csr, has_dense, has_hist, has_bold_io;
        int64_t noise_n;
        uint64_t seed, stream;
        bool gen_noise;
    };
    std::vector<Ptrs> pp(n);
    std::map<std::tuple<uint64_t, uint64_t, int64_t>, size_t> noise_dedup;
    std::vector<NoiseJob> jobs_host;  // out pointers filled after allocation (offsets for now)
    std::vector<size_t> job_off;

    for (int i = 0; i < n; ++i) {
        const hb_instance& I = inst[i];
        const Shape& s = sh[i];
        Ptrs& q = pp[i];
        memset(&q, 0, sizeof(q));
        const size_t N = (size_t)I.n_nodes, NV = (size_t)s.nv;
        bool fresh;
        q.x0 = shared_off(I.x0, nullptr, nullptr, N * NV * 4, &fresh);
        if (fresh) stage(q.x0, I.x0, N * NV * 4);
        if (I.node_param_idx >= 0) {
            std::vector<float> mcs(N * kMC);
            std::vector<float> p(I.params, I.params + model_nparams(I.model));
            for (size_t k = 0; k < N; ++k) {
                p[I.node_param_idx] = I.node_param[k];
                model_consts(I.model, p.data(), &mcs[k * kMC]);
            }
            q.has_node_mc = true;
            q.node_mc = A.take(mcs.size() * 4);
            stage(q.node_mc, mcs.data(), mcs.size() * 4);
        }
        if (I.wkind == HB_W_CSR) {
            q.has_csr = true;
            const bool per_edge = I.delay_mode == HB_DELAY_PER_EDGE;
            q.indptr = shared_off(I.indptr, nullptr, nullptr, (N + 1) * 4, &fresh);
            if (fresh) stage(q.indptr, I.indptr, (N + 1) * 4);
            q.edges = shared_off(I.indices, I.weights, per_edge ? (const void*)I.edge_delays : nullptr,
                                 (size_t)std::max(I.nnz, 1) * 8, &fresh);
            if (fresh) {
                std::vector<uint2> ed((size_t)std::max(I.nnz, 1));
                for (int e = 0; e < I.nnz; ++e) {
                    uint32_t d = (per_edge && I.edge_delays) ? (uint32_t)I.edge_delays[e] : 0u;
                    uint32_t wb;
                    memcpy(&wb, &I.weights[e], 4);
                    ed[e] = make_uint2((uint32_t)I.indices[e] | (d << 16), wb);
                }
                stage(q.edges, ed.data(), ed.size() * 8);
            }
        }
        if (I.wkind == HB_W_DENSE) {
            q.has_dense = true;
            q.dense_t = shared_off(I.dense, nullptr, (const void*)(intptr_t)I.cfn, N * N * 4, &fresh);
            if (fresh) {
                std::vector<float> wt(N * N);
                std::vector<float> w(I.dense, I.dense + N * N);
                if (I.cfn == HB_CF_DIFFERENCE) {
                    // construction.rs:694-707 / engine.rs:327-350: W'[i,i] -= sum_j W[i,j]
                    for (size_t r = 0; r < N; ++r) {
                        float rs = 0.0f;
                        for (size_t c = 0; c < N; ++c) rs += w[r * N + c];
                        w[r * N + r] -= rs;
                    }
                }
                for (size_t r = 0; r < N; ++r)
                    for (size_t c = 0; c < N; ++c) wt[c * N + r] = w[r * N + c];
                stage(q.dense_t, wt.data(), wt.size() * 4);
            }
        }
        if (is_stoch(I.integrator)) {
            q.has_noise = true;
            q.noise_n = (int64_t)I.n_steps * (int64_t)N * (int64_t)NV;
            if (I.noise) {
                q.noise = shared_off(I.noise, nullptr, nullptr, (size_t)q.noise_n * 4, &fresh);
                if (fresh) stage(q.noise, I.noise, (size_t)q.noise_n * 4);
            } else {
                auto key = std::make_tuple(I.noise_seed, I.noise_stream, q.noise_n);
                auto it = noise_dedup.find(key);
                if (it != noise_dedup.end()) q.noise = it->second;
                else {
                    q.noise = A.take((size_t)q.noise_n * 4);
                    noise_dedup[key] = q.noise;
                    NoiseJob j;
                    j.out = nullptr;
                    j.n = q.noise_n;
                    j.seed = I.noise_seed;
                    j.stream = I.noise_stream;
                    jobs_host.push_back(j);
                    job_off.push_back(q.noise);
                    P->max_noise = std::max(P->max_noise, q.noise_n);
                }
            }
        }
        if (!s.lds_ring) {
            q.has_ring = true;
            q.ring = A.take((size_t)(s.depth + (s.blocked ? s.blocked - 1 : 0)) * N * s.np * 4);
        }
        P->outs[i].n_state = N * NV;
        q.tavg = A.take(N * NV * 4);
        q.fin = A.take(N * NV * 4);
        P->outs[i].tavg = q.tavg;
        P->outs[i].fin = q.fin;
        if (I.traj || (P->flags & kFlagDeviceTrajectory)) {
            q.has_traj = true;
            P->outs[i].n_traj = (size_t)I.n_steps * N * NV;
            q.traj = A.take(std::max(P->outs[i].n_traj, (size_t)1) * 4);
            P->outs[i].traj = q.traj;
        }
        // a continued run's delay history and BOLD monitor state (in and out)
        const int64_t hf = history_floats(I, s.np);
        if (I.history && hf > 0) {
            q.has_hist = true;
            q.hist = A.take((size_t)hf * 4);
            if (I.step0 > 0) stage(q.hist, I.history, (size_t)hf * 4);
        }
        if (I.bold_state && I.bold_period > 0 && I.bold) {
            q.has_bold_io = true;
            q.bold_io = A.take(N * 4 * 4);
            stage(q.bold_io, I.bold_state, N * 4 * 4);
        }
        P->hist_off.push_back(q.has_hist ? q.hist : SIZE_MAX);
        P->bold_off.push_back(q.has_bold_io ? q.bold_io : SIZE_MAX);
        if (I.bold && I.bold_period > 0) {
            BoldShape b = bold_shape(I.n_steps, I.bold_period, I.dt_ms, I.bold_tr_ms);
            q.has_bold = true;
            P->outs[i].n_bold = (size_t)std::max(b.nout, 0) * N;
            q.bold = A.take(std::max(P->outs[i].n_bold, (size_t)1) * 4);
            P->outs[i].bold = q.bold;
        }
    }
    // instance descriptors, order, counters, noise jobs
    const size_t off_inst = A.take(sizeof(DevInst) * n);
    const size_t off_order = A.take(sizeof(int32_t) * (size_t)n * 17);   // + padding of 16-instance tiles
    const size_t off_cnt = A.take(sizeof(int32_t) * 64);
    const size_t off_jobs = A.take(sizeof(NoiseJob) * std::max<size_t>(jobs_host.size(), 1));

    P->arena_bytes = A.top;
    hipError_t he = hipMalloc((void**)&P->arena, P->arena_bytes);
    if (he != hipSuccess) {
        hb_plan_destroy(P);
        return fail(HB_BACKEND_ERROR, std::string("hipMalloc of plan arena failed: ") + hipGetErrorString(he));
    }
    char* base = P->arena;
    for (auto& u : uploads) {
        he = hipMemcpy(base + u.off, u.bytes.data(), u.bytes.size(), hipMemcpyHostToDevice);
        if (he != hipSuccess) { hb_plan_destroy(P); return fail(HB_BACKEND_ERROR, "upload failed"); }
    }
    uploads.clear();

    // ---- descriptors
    for (int i = 0; i < n; ++i) {
        const hb_instance& I = inst[i];
        const Shape& s = sh[i];
        const Ptrs& q = pp[i];
        DevInst& D = P->dev[i];
        memset(&D, 0, sizeof(D));
        D.model = I.model; D.integrator = I.integrator; D.n_nodes = I.n_nodes; D.n_steps = I.n_steps;
        D.np = s.np; D.wkind = I.wkind;
        D.cfn = I.cfn == HB_CF_DIFFERENCE ? HB_CF_LINEAR : I.cfn;
        D.delay_mode = I.wkind == HB_W_CSR ? I.delay_mode
                       : (I.delay_mode == HB_DELAY_PER_EDGE ? HB_DELAY_SINGLE : I.delay_mode);
        D.single_delay = I.single_delay;
        D.ring_depth = s.depth;
        D.blocked_k = s.blocked;
        D.lpn = s.lpn;
        D.lpn_shift = s.lpn == 4 ? 2 : (s.lpn == 2 ? 1 : 0);
        D.sweep_form = I.sweep_form; D.sweep_val = I.sweep_val;
        D.node_param_idx = q.has_node_mc ? I.node_param_idx : -1;
        D.engine = I.engine;
        D.cpu_form = (I.wkind == HB_W_CSR && I.delay_mode == HB_DELAY_PER_EDGE && I.edge_delays) ? 1 : 0;
        D.dt = I.dt;
        D.dt_half = I.dt * 0.5f;
        D.dt_sixth = I.dt / 6.0f;
        D.inv_nsteps = 1.0f / (float)I.n_steps;
        model_consts(I.model, I.params, D.mc);
        int classic = 0;
        float cp[8];
        memcpy(cp, I.cparams, sizeof(cp));
        int ncp = I.n_cparams;
        coupling_consts(D.cfn, cp, ncp, I.n_nodes, D.cc, &classic);
        if (I.cfn == HB_CF_DIFFERENCE) {   // Difference -> Linear(a, 0) after weight preprocessing
            D.cc[0] = ncp > 0 ? cp[0] : 1.0f;
            D.cc[1] = 0.0f;
        }
        D.sjr_classic = classic;
        if (I.wkind == HB_W_CSR)
            for (int r = 0; r < I.n_nodes; ++r) D.max_row = std::max(D.max_row, I.indptr[r + 1] - I.indptr[r]);
        D.scalar_w = I.scalar_w;
        if (I.cfn == HB_CF_DIFFERENCE && (I.wkind == HB_W_SCALAR_MEAN || I.wkind == HB_W_SCALAR_DIAG))
            D.scalar_w = 0.0f;   // engine.rs:330-333
        float f0[2] = {0.0f, 0.0f};
        const bool per_edge = I.wkind == HB_W_CSR && D.delay_mode == HB_DELAY_PER_EDGE;
        if (!per_edge) pre_zero(D.cfn, D.cc, classic, s.np, f0);
        D.fill0 = f0[0]; D.fill1 = f0[1];
        for (int v = 0; v < s.nv; ++v) {
            const float ns = I.nsig ? I.nsig[v] : 0.0f;
            D.nscale[v] = sqrtf(2.0f * ns) * sqrtf(I.dt);   // integrator.rs:179-191
        }
        D.dt_ms = I.dt_ms;
        if (I.bold_period > 0 && q.has_bold) {
            BoldShape b = bold_shape(I.n_steps, I.bold_period, I.dt_ms, I.bold_tr_ms);
            D.bold_period = I.bold_period;
            D.bold_nout = b.nout;
            D.bold_native = b.native;
            D.tr_ms = I.bold_tr_ms;
            D.dtb_ms = b.dtb;
            D.bold = (float*)(base + q.bold);
        }
        D.x0 = (const float*)(base + q.x0);
        D.node_mc = q.has_node_mc ? (const float*)(base + q.node_mc) : nullptr;
        D.indptr = q.has_csr ? (const int32_t*)(base + q.indptr) : nullptr;
        D.edges = q.has_csr ? (const uint2*)(base + q.edges) : nullptr;
        D.dense_t = q.has_dense ? (const float*)(base + q.dense_t) : nullptr;
        D.noise = q.has_noise ? (const float*)(base + q.noise) : nullptr;
        D.ring = q.has_ring ? (float*)(base + q.ring) : nullptr;
        D.tavg = (float*)(base + q.tavg);
        D.fin = (float*)(base + q.fin);
        D.traj = q.has_traj ? (float*)(base + q.traj) : nullptr;
        D.traj_tstride = (int64_t)I.n_nodes * s.nv;
        D.step0 = (int32_t)std::min<int64_t>(I.step0, (int64_t)1 << 30);
        D.step0_modh = (int32_t)(I.step0 % ((int64_t)max_delay_of(I) + 1));
        if (q.has_hist) {
            D.hist_len = max_delay_of(I);
            D.hist_in = I.step0 > 0 ? (const float*)(base + q.hist) : nullptr;
            D.hist_out = (float*)(base + q.hist);
        }
        if (q.has_bold_io) {
            D.bold_io = (float*)(base + q.bold_io);
            D.bold_carry = 1;
        }
    }
    P->d_inst = (DevInst*)(base + off_inst);
    P->d_order = (int32_t*)(base + off_order);
    P->d_counters = (int32_t*)(base + off_cnt);
    P->d_jobs = (NoiseJob*)(base + off_jobs);
    for (size_t j = 0; j < jobs_host.size(); ++j) jobs_host[j].out = (float*)(base + job_off[j]);
    P->n_jobs = (int)jobs_host.size();
    HB_HIP(hipMemcpy(P->d_inst, P->dev.data(), sizeof(DevInst) * n, hipMemcpyHostToDevice));
    if (!jobs_host.empty())
        HB_HIP(hipMemcpy(P->d_jobs, jobs_host.data(), sizeof(NoiseJob) * jobs_host.size(), hipMemcpyHostToDevice));

    // ---- launch groups: (model, pre channels, block, lds_ring, multi), each a
    // single-model kernel; cost-descending order inside. Groups run as
    // parallel graph branches, so groups of different models share the chip.
    // Dense batches on the matrix cores: instances of one model, size, weight
    // matrix, coupling function, integrator, length and sweep form share a
    // kernel's uniform controls; a batch smaller than kDenseMinBatch keeps
    // the lane loop.
    std::map<std::tuple<int, int, int, const void*, int, int, int, int, int, int>, std::vector<int>> dense_groups;
    for (int i = 0; i < n; ++i)
        if (sh[i].mfma) {
            const hb_instance& I = inst[i];
            dense_groups[std::make_tuple(I.model, sh[i].np, I.n_nodes, (const void*)I.dense, I.cfn, I.integrator,
                                         I.n_steps, I.sweep_form, I.engine, (int)(I.n_cparams >= 5))].push_back(i);
        }
    for (auto& kv : dense_groups)
        if ((int)kv.second.size() < kDenseMinBatch)
            for (int i : kv.second) sh[i].mfma = false;
    // a launch wider than 512 threads runs the 1024-thread kernel, which has
    // only the delay-blocked schedule with 8-step (any ring) or 16-step (HBM
    // ring) blocks
    for (int i = 0; i < n; ++i)
        if (!sh[i].mfma && sh[i].block > kBlockMax &&
            !(sh[i].blocked == 8 || (sh[i].blocked == 16 && !sh[i].lds_ring))) {
            hb_plan_destroy(P);
            return fail(HB_BACKEND_ERROR, "instance " + std::to_string(i) + ": " + std::to_string(sh[i].block) +
                                              "-thread launch without a delay-blocked schedule");
        }
    std::map<std::tuple<int, int, int, int, int>, std::vector<int>> groups;
    for (int i = 0; i < n; ++i)
        if (!sh[i].mfma)
            groups[std::make_tuple(inst[i].model, sh[i].np, sh[i].block, (int)sh[i].lds_ring, (int)sh[i].multi)].push_back(i);
    std::vector<int32_t> order;
    order.reserve((size_t)n * 17);
    int gi = 0;
    int max_lds_all = 0;
    for (auto& kv : dense_groups) {
        auto& v = kv.second;
        if ((int)v.size() < kDenseMinBatch) continue;
        LaunchDesc L;
        memset(&L, 0, sizeof(L));
        L.insts = P->d_inst;
        L.order = P->d_order + order.size();
        const int N = std::get<2>(kv.first);
        const int tiles = ((int)v.size() + 15) / 16;
        for (int k = 0; k < tiles * 16; ++k) order.push_back(k < (int)v.size() ? v[k] : -1);
        L.mfma = 1;
        L.n_nodes = N;
        L.nk = (N + 3) / 4 * 4;
        L.n_work = tiles;
        L.grid = tiles;
        L.block = (N + 15) / 16 * 64;
        L.model = std::get<0>(kv.first);
        L.np = std::get<1>(kv.first);
        L.lds_bytes = dense_lds_bytes(N, L.np);
        max_lds_all = std::max(max_lds_all, L.lds_bytes);
        double cmax = 0.0;
        for (int i : v) cmax = std::max(cmax, sh[i].cost);
        P->launch_cost.push_back(cmax * 0.25);
        P->launches.push_back(L);
        P->block_main = std::max(P->block_main, L.block);
        P->lds_main = std::max(P->lds_main, L.lds_bytes);
    }
    for (auto& kv : groups) {
        auto& v = kv.second;
        std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return sh[a].cost > sh[b].cost; });
        LaunchDesc L;
        memset(&L, 0, sizeof(L));
        L.insts = P->d_inst;
        L.order = P->d_order + order.size();
        L.n_work = (int)v.size();
        L.block = std::get<2>(kv.first);
        L.lds_ring = std::get<3>(kv.first) != 0;
        L.multi = std::get<4>(kv.first) != 0;
        int lds = 0;
        for (int i : v) lds = std::max(lds, sh[i].lds_bytes);
        L.lds_bytes = lds;
        L.model = std::get<0>(kv.first);
        L.np = std::get<1>(kv.first);
        max_lds_all = std::max(max_lds_all, lds);
        for (int i : v) order.push_back(i);
        // persistent when the work exceeds one wave of resident workgroups
        int occ = 0;
        L.grid = L.n_work;
        L.counter = nullptr;
        if (forward_occupancy_fast(L, &occ) == hipSuccess && occ > 0) {
            const int cap = occ * n_cu;
            if (L.n_work > cap && gi < 64) {
                L.grid = cap;
                L.counter = P->d_counters + gi;
            }
        }
        if (opts && opts->grid_waves > 0 && L.counter) L.grid = std::min(L.n_work, opts->grid_waves);
        // estimated time: the longest instance, or the work spread over the grid
        double csum = 0.0, cmax = 0.0;
        for (int i : v) { csum += sh[i].cost; cmax = std::max(cmax, sh[i].cost); }
        P->launch_cost.push_back(std::max(cmax, csum / std::max(1, L.grid)));
        P->launches.push_back(L);
        ++gi;
        P->block_main = std::max(P->block_main, L.block);
        P->lds_main = std::max(P->lds_main, L.lds_bytes);
    }
    HB_HIP(hipMemcpy(P->d_order, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice));

    HB_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    for (auto& e : P->ev) HB_HIP(hipEventCreate(&e));
    if (P->launches.size() > 1) {
        // as many lanes as the device has hardware queues for a process by
        // default (more streams would share queues and serialize behind each other)
        constexpr size_t kLanes = 4;
        const size_t nl = std::min(kLanes, P->launches.size());
        std::vector<size_t> byc(P->launches.size());
        for (size_t k = 0; k < byc.size(); ++k) byc[k] = k;
        std::stable_sort(byc.begin(), byc.end(), [&](size_t a, size_t b) { return P->launch_cost[a] > P->launch_cost[b]; });
        std::vector<double> load(nl, 0.0);
        P->lane_of.assign(P->launches.size(), 0);
        for (size_t k : byc) {
            const size_t l = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
            P->lane_of[k] = (int)l;
            load[l] += P->launch_cost[k];
        }
        HB_HIP(hipEventCreateWithFlags(&P->fork, hipEventDisableTiming));
        P->side.assign(nl, nullptr);
        P->join.assign(nl, nullptr);
        for (size_t k = 0; k < nl; ++k) {
            HB_HIP(hipStreamCreateWithFlags(&P->side[k], hipStreamNonBlocking));   // default priority
            HB_HIP(hipEventCreateWithFlags(&P->join[k], hipEventDisableTiming));
        }
    }

    // ---- capture the launch sequence
    if (!(P->flags & HB_FLAG_NO_GRAPH)) {
        HB_HIP(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e1 = launch_noise(P->d_jobs, P->n_jobs, P->max_noise, P->stream);
        hipError_t e2 = hipStreamEndCapture(P->stream, &P->gr_noise);
        if (e1 != hipSuccess || e2 != hipSuccess) { hb_plan_destroy(P); return fail(HB_BACKEND_ERROR, "noise graph capture failed"); }
        HB_HIP(hipGraphInstantiate(&P->g_noise, P->gr_noise, nullptr, nullptr, 0));
        HB_HIP(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
        hipError_t el = hipSuccess;
        if (gi > 0) el = launch_zero_i32(P->d_counters, std::min(gi, 64), P->stream);
        if (el == hipSuccess) el = launch_groups(P, exact);
        e2 = hipStreamEndCapture(P->stream, &P->gr_fwd);
        if (el != hipSuccess || e2 != hipSuccess) {
            hb_plan_destroy(P);
            return fail(HB_BACKEND_ERROR, std::string("forward graph capture failed: ") +
                                              hipGetErrorString(el != hipSuccess ? el : e2) +
                                              (g_launch_err.empty() ? "" : " at " + g_launch_err));
        }
        HB_HIP(hipGraphInstantiate(&P->g_fwd, P->gr_fwd, nullptr, nullptr, 0));
    }
    *out = P;
    return HB_OK;
}

int hb_plan_run(hb_plan* P, hb_run_stats* stats) {
    if (!P) return fail(HB_INVALID_STATE, "null plan");
    HB_HIP(hipSetDevice(P->device));
    const bool exact = (P->flags & HB_FLAG_EXACT) != 0;
    HB_HIP(hipEventRecord(P->ev[0], P->stream));
    if (!(P->flags & HB_FLAG_KEEP_NOISE) || !P->noise_done) {
        if (P->g_noise) HB_HIP(hipGraphLaunch(P->g_noise, P->stream));
        else HB_HIP(launch_noise(P->d_jobs, P->n_jobs, P->max_noise, P->stream));
        P->noise_done = true;
    }
    HB_HIP(hipEventRecord(P->ev[1], P->stream));
    if (P->g_fwd) HB_HIP(hipGraphLaunch(P->g_fwd, P->stream));
    else {
        HB_HIP(launch_zero_i32(P->d_counters, std::min((int)P->launches.size(), 64), P->stream));
        const hipError_t el = launch_groups(P, exact);
        if (el != hipSuccess)
            return fail(HB_BACKEND_ERROR, std::string("forward launch failed: ") + hipGetErrorString(el) + " at " + g_launch_err);
    }
    HB_HIP(hipEventRecord(P->ev[2], P->stream));
    HB_HIP(hipEventSynchronize(P->ev[2]));
    if (stats) {
        float t01 = 0.0f, t12 = 0.0f;
        HB_HIP(hipEventElapsedTime(&t01, P->ev[0], P->ev[1]));
        HB_HIP(hipEventElapsedTime(&t12, P->ev[1], P->ev[2]));
        stats->kernel_ms = t12;
        stats->total_ms = t01 + t12;
        stats->node_steps = P->node_steps;
        stats->n_launches = (int)P->launches.size();
        stats->block_size = P->block_main;
        stats->lds_bytes = P->lds_main;
    }
    return HB_OK;
}

int hb_plan_download(hb_plan* P) {
    if (!P) return fail(HB_INVALID_STATE, "null plan");
    HB_HIP(hipSetDevice(P->device));
    for (size_t i = 0; i < P->inst.size(); ++i) {
        const hb_instance& I = P->inst[i];
        const auto& o = P->outs[i];
        if (I.tavg) HB_HIP(hipMemcpy(I.tavg, P->arena + o.tavg, o.n_state * 4, hipMemcpyDeviceToHost));
        if (I.final_state) HB_HIP(hipMemcpy(I.final_state, P->arena + o.fin, o.n_state * 4, hipMemcpyDeviceToHost));
        if (I.traj && o.n_traj) HB_HIP(hipMemcpy(I.traj, P->arena + o.traj, o.n_traj * 4, hipMemcpyDeviceToHost));
        if (P->hist_off[i] != SIZE_MAX)
            HB_HIP(hipMemcpy(I.history, P->arena + P->hist_off[i],
                             (size_t)history_floats(I, P->dev[i].np) * 4, hipMemcpyDeviceToHost));
        if (P->bold_off[i] != SIZE_MAX)
            HB_HIP(hipMemcpy(I.bold_state, P->arena + P->bold_off[i], (size_t)I.n_nodes * 16, hipMemcpyDeviceToHost));
        if (I.bold && I.bold_period > 0) {
            const int N = I.n_nodes;
            const size_t rows = N ? o.n_bold / N : 0;
            const size_t keep = std::min(rows, (size_t)std::max(I.bold_cap, 0));
            if (keep) HB_HIP(hipMemcpy(I.bold, P->arena + o.bold, keep * N * 4, hipMemcpyDeviceToHost));
            if (I.bold_n) *I.bold_n = (int32_t)rows;
        }
    }
    return HB_OK;
}

int hb_run_instances(const hb_instance* inst, int32_t n, const hb_run_opts* opts, hb_run_stats* stats) {
    hb_plan* P = nullptr;
    int rc = hb_plan_create(inst, n, opts, &P);
    if (rc) return rc;
    rc = hb_plan_run(P, stats);
    if (!rc) rc = hb_plan_download(P);
    hb_plan_destroy(P);
    return rc;
}

int hb_philox_normal(uint64_t seed, uint64_t stream, int64_t n, float* out_host, int32_t device) {
    if (n < 0 || (n > 0 && !out_host)) return fail(HB_INVALID_CONFIG, "bad arguments");
    if (n == 0) return HB_OK;
    HB_HIP(hipSetDevice(device));
    float* d = nullptr;
    NoiseJob* dj = nullptr;
    HB_HIP(hipMalloc((void**)&d, (size_t)n * 4));
    if (hipMalloc((void**)&dj, sizeof(NoiseJob)) != hipSuccess) { (void)hipFree(d); return fail(HB_BACKEND_ERROR, "hipMalloc failed"); }
    NoiseJob j;
    j.out = d; j.n = n; j.seed = seed; j.stream = stream;
    hipError_t e = hipMemcpy(dj, &j, sizeof(j), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_noise(dj, 1, n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out_host, d, (size_t)n * 4, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    (void)hipFree(dj);
    if (e != hipSuccess) return fail(HB_BACKEND_ERROR, std::string("philox: ") + hipGetErrorString(e));
    return HB_OK;
}

int64_t hb_history_floats(const hb_instance* inst) {
    if (!inst || !inst->params || inst->model < 0 || inst->model > HB_RWW || inst->n_nodes < 1) return 0;
    const hb_instance& I = *inst;
    return history_floats(I, ring_channels(I.model, I.cfn, I.params, I.node_param_idx, -1));
}

}  // extern "C"

namespace hb {

int plan_bold_batch_mean(hb_plan* P, int period, double dt_ms, float* bw, float* native, bool exact) {
    if (!P || P->inst.empty() || period < 1) return fail(HB_INVALID_STATE, "batch-mean BOLD: bad plan or period");
    const int B = (int)P->inst.size();
    const int T = P->inst[0].n_steps, N = P->inst[0].n_nodes, nv = model_nvar(P->inst[0].model);
    std::vector<const float*> trajs(B);
    for (int i = 0; i < B; ++i) {
        const hb_instance& I = P->inst[i];
        if (I.n_steps != T || I.n_nodes != N || !P->outs[i].n_traj)
            return fail(HB_INVALID_STATE, "batch-mean BOLD needs one shape and device trajectories");
        trajs[i] = (const float*)(P->arena + P->outs[i].traj);
    }
    const int rows = T / period + (T % period ? 1 : 0);
    HB_HIP(hipSetDevice(P->device));
    DevBufs D;
    const float** d_trajs = nullptr;
    float *zm = nullptr, *d_bw = nullptr, *d_nat = nullptr;
    hipError_t e = D.alloc(&d_trajs, (size_t)B);
    if (e == hipSuccess) e = D.alloc(&zm, (size_t)T * N);
    if (e == hipSuccess) e = D.alloc(&d_bw, (size_t)N * 4);
    if (e == hipSuccess) e = D.alloc(&d_nat, (size_t)std::max(rows, 1) * N);
    if (e == hipSuccess) e = hipMemcpy(d_trajs, trajs.data(), sizeof(float*) * B, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_bw, bw, sizeof(float) * N * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = exact ? launch_bold_mean_exact(d_trajs, B, T, N, nv, zm, period, dt_ms, d_bw, d_nat, P->stream)
                  : launch_bold_mean_fast(d_trajs, B, T, N, nv, zm, period, dt_ms, d_bw, d_nat, P->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
    if (e == hipSuccess && rows > 0) e = hipMemcpy(native, d_nat, sizeof(float) * rows * N, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(bw, d_bw, sizeof(float) * N * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(HB_BACKEND_ERROR, std::string("batch-mean BOLD: ") + hipGetErrorString(e));
    return HB_OK;
}

}  // namespace hb
