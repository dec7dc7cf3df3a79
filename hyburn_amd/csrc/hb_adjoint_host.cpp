// hb_adjoint_host.cpp -- adjoint plans: validation, HBM layout (coupling log,
// checkpoints, tapes, BOLD buffers), CSC transpose of the connectome, the
// Jacobian of the model constants, launches and downloads.
//
// The C ABI is the gradient counterpart of the forward plan (hb_plan_*): the
// differentiable graph is exactly the forward recursion of one instance, so a
// caller (e.g. a torch.autograd.Function) runs hb_adjoint_forward for the loss
// and hb_adjoint_backward for the gradients.
#include "hb_adjoint.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

using namespace hb;

namespace {

// Forward-mode dual number for d(model constants)/d(parameter).
struct Dual {
    double v, d;
    Dual(float x = 0.0f) : v(x), d(0.0) {}
    Dual(double a, double b) : v(a), d(b) {}
};
Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
bool operator>(Dual a, Dual b) { return a.v > b.v; }
Dual mc_exp(Dual x) { const double e = std::exp(x.v); return Dual(e, e * x.d); }

// d mc[i] / d p[k] at p, from the same expressions as model_consts.
void model_jacobian(int m, const float* p, int np, int k, float* jac) {
    Dual pd[32], mc[kMC];
    for (int i = 0; i < np && i < 32; ++i) pd[i] = Dual((double)p[i], i == k ? 1.0 : 0.0);
    model_consts_t<Dual>(m, pd, mc);
    for (int i = 0; i < kMC; ++i) jac[i] = (float)mc[i].d;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Arena {
    size_t top = 0;
    size_t take(size_t bytes) { const size_t o = top; top += align256(std::max<size_t>(bytes, 1)); return o; }
};

int bad(int idx, const char* what) {
    char buf[256];
    snprintf(buf, sizeof(buf), "instance %d: %s", idx, what);
    return set_error(HB_INVALID_CONFIG, buf);
}

int validate(const hb_instance& I, const hb_adjoint_spec& A, int idx) {
    if (I.model < 0 || I.model > HB_RWW) return bad(idx, "unknown model");
    if (I.integrator < HB_EULER || I.integrator > HB_RK4_STOCH) return bad(idx, "unknown integrator");
    if (I.n_nodes < 1 || I.n_nodes > 1024) return bad(idx, "the adjoint supports 1..1024 nodes");
    if (I.n_steps < 1) return bad(idx, "n_steps must be >= 1");
    if (!(I.dt > 0.0f)) return bad(idx, "dt must be positive");
    if (!I.params || !I.x0) return bad(idx, "params and x0 are required");
    if (I.sweep_form != 0) return bad(idx, "the adjoint takes plain parameters (sweep_form 0)");
    if (I.wkind != HB_W_NONE && I.wkind != HB_W_CSR) return bad(idx, "the adjoint supports CSR weights (or none)");
    if (I.wkind == HB_W_CSR) {
        if (I.delay_mode != HB_DELAY_PER_EDGE) return bad(idx, "the adjoint supports per-edge delays");
        if (I.cfn != HB_CF_LINEAR && I.cfn != HB_CF_SCALED_LINEAR && I.cfn != HB_CF_KURAMOTO)
            return bad(idx, "the adjoint supports Linear, ScaledLinear and Kuramoto coupling");
        if (I.cfn == HB_CF_KURAMOTO && I.model != HB_KURAMOTO) return bad(idx, "Kuramoto coupling needs the Kuramoto model");
        if (!I.indptr || (I.nnz > 0 && (!I.indices || !I.weights))) return bad(idx, "CSR arrays missing");
        if (I.indptr[0] != 0 || I.indptr[I.n_nodes] != I.nnz) return bad(idx, "CSR indptr inconsistent with nnz");
        for (int i = 0; i < I.n_nodes; ++i)
            if (I.indptr[i + 1] < I.indptr[i]) return bad(idx, "CSR indptr not monotone");
        for (int e = 0; e < I.nnz; ++e) {
            if (I.indices[e] < 0 || I.indices[e] >= I.n_nodes) return bad(idx, "CSR index out of range");
            if (I.edge_delays && (I.edge_delays[e] < 0 || I.edge_delays[e] > 65535)) return bad(idx, "edge delay out of range");
        }
    }
    if (I.node_param_idx >= model_nparams(I.model)) return bad(idx, "node parameter index out of range");
    if (I.node_param_idx >= 0 && !I.node_param) return bad(idx, "node parameter values missing");
    if (I.integrator >= HB_EULER_STOCH && !I.nsig) return bad(idx, "stochastic integrator needs nsig");
    if (A.loss_kind < HB_LOSS_FC || A.loss_kind > HB_LOSS_ORDER) return bad(idx, "unknown loss kind");
    if ((A.loss_kind == HB_LOSS_FC || A.loss_kind == HB_LOSS_BOLD_MSE) && I.bold_period <= 0)
        return bad(idx, "BOLD losses need bold_period > 0");
    if (A.loss_kind == HB_LOSS_ORDER && I.model != HB_KURAMOTO) return bad(idx, "the order-parameter loss needs the Kuramoto model");
    if (!A.target) return bad(idx, "loss target missing");
    if (!A.loss || !A.grad_coupling) return bad(idx, "loss and grad_coupling outputs are required");
    if (A.grad_param >= model_nparams(I.model)) return bad(idx, "grad_param out of range");
    if (A.grad_param >= 0 && !A.grad_param_out) return bad(idx, "grad_param_out missing");
    return HB_OK;
}

#define HB_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) return set_error(HB_BACKEND_ERROR, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct OutPtrs {
    size_t loss, gcc, gp, gx0, fin;
    int n_nodes, nv;
};

}  // namespace

struct hb_adjoint_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    char* arena = nullptr;
    size_t bytes = 0;
    AdjInst* d_inst = nullptr;
    std::vector<AdjLaunch> launches;
    std::vector<int> user_of;          // grouped position -> caller index
    std::vector<hb_adjoint_spec> spec; // caller order
    std::vector<hb_instance> inst;     // caller order (host pointers only for outputs / sizes)
    std::vector<OutPtrs> outs;         // caller order
    size_t outs_off = 0, outs_bytes = 0;   // contiguous output region of the arena
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    double node_steps = 0.0;
    bool forward_done = false;
};

extern "C" {

void hb_adjoint_plan_destroy(hb_adjoint_plan* P) {
    if (!P) return;
    if (P->stream) hipStreamSynchronize(P->stream);
    for (auto& e : P->ev)
        if (e) hipEventDestroy(e);
    if (P->arena) hipFree(P->arena);
    if (P->stream) hipStreamDestroy(P->stream);
    delete P;
}

int hb_adjoint_plan_create(const hb_instance* inst, const hb_adjoint_spec* spec, int32_t n,
                           const hb_adjoint_opts* opts, hb_adjoint_plan** out) {
    if (!out) return set_error(HB_INVALID_CONFIG, "null output handle");
    *out = nullptr;
    if (n < 1 || !inst || !spec) return set_error(HB_INVALID_CONFIG, "empty adjoint plan");
    for (int i = 0; i < n; ++i) {
        const int rc = validate(inst[i], spec[i], i);
        if (rc) return rc;
    }
    const int device = opts ? opts->device : 0;
    HB_HIP(hipSetDevice(device));
    int lds_max = 0;
    HB_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    const int lds_budget = lds_max - 2048;   // the adjoint kernels' static LDS is at most 1 KB
    static bool attrs_done[64] = {false};
    if (device >= 0 && device < 64 && !attrs_done[device]) {
        HB_HIP(adjoint_set_attributes(lds_budget));
        attrs_done[device] = true;
    }

    hb_adjoint_plan* P = new hb_adjoint_plan();
    P->device = device;
    P->spec.assign(spec, spec + n);
    P->inst.assign(inst, inst + n);

    // ---- checkpointing: with checkpoint_every = 0 the plan keeps every
    // instance's whole trajectory (one segment: the forward writes it and the
    // reverse reads it, no recompute) when all of them fit a quarter of the
    // free device memory; otherwise segments of ceil(sqrt(n_steps)) steps
    const int ce = opts ? opts->checkpoint_every : 0;
    bool keep_all = false;
    if (ce == 0) {
        size_t free_b = 0, total_b = 0;
        HB_HIP(hipMemGetInfo(&free_b, &total_b));
        double tape_b = 0.0;
        for (int i = 0; i < n; ++i)
            tape_b += ((double)inst[i].n_steps + 1.0) * inst[i].n_nodes * model_nvar(inst[i].model) * 4.0;
        keep_all = tape_b <= 0.25 * (double)free_b;
    }

    // ---- shapes and grouping
    struct Shape { int nv, nc, np, npt, block, Hd, S, nseg, nwin, nnz; bool kur, ring_lds, blocked; size_t ring_b; };
    constexpr int kBlock = 8;   // steps per delay block of the blocked kernels
    std::vector<Shape> sh(n);
    std::map<std::tuple<int, int, int, int, int, int, int>, std::vector<int>> groups;
    for (int i = 0; i < n; ++i) {
        const hb_instance& I = inst[i];
        Shape& s = sh[i];
        s.nv = model_nvar(I.model);
        s.nc = model_ncvar(I.model);
        s.kur = I.cfn == HB_CF_KURAMOTO && I.wkind == HB_W_CSR;
        s.np = I.wkind == HB_W_CSR ? ring_channels(I.model, I.cfn, I.params, I.node_param_idx, spec[i].grad_param)
                                   : s.nc;
        int dmax = 0, dmin = 1 << 30;
        s.nnz = I.wkind == HB_W_CSR ? I.nnz : 0;
        for (int e = 0; e < s.nnz; ++e) {
            const int d = I.edge_delays ? (int)I.edge_delays[e] : 0;
            dmax = std::max(dmax, d);
            dmin = std::min(dmin, d);
        }
        // delay-blocked kernels: one barrier per kBlock steps when every delay >= kBlock,
        // one node per thread up to 1024 (the 1024-thread instantiation past 512)
        s.blocked = s.nnz > 0 && dmin >= kBlock && I.n_nodes <= 1024 && !(opts && (opts->flags & HB_FLAG_NO_BLOCKING));
        s.npt = (s.blocked || I.n_nodes <= 512) ? 1 : 2;
        s.block = s.npt == 1 ? std::max(64, (I.n_nodes + 63) / 64 * 64) : 512;
        s.Hd = s.blocked ? dmax + kBlock + 1 : dmax + 2;
        s.S = ce > 0 ? ce : keep_all ? I.n_steps : std::max(1, (int)std::ceil(std::sqrt((double)I.n_steps)));
        s.S = std::min(s.S, I.n_steps);
        s.nseg = (I.n_steps + s.S - 1) / s.S;
        s.nwin = I.bold_period > 0 ? (I.n_steps + I.bold_period - 1) / I.bold_period : 0;
        s.ring_b = (size_t)(s.Hd + (s.blocked ? kBlock - 1 : 0)) * I.n_nodes * s.np * 4;
        s.ring_lds = s.ring_b <= (size_t)lds_budget;
        // integrator class and (blocked kernels) Linear coupling are compile-time kernel variants
        const int icl = (I.integrator == HB_EULER || I.integrator == HB_EULER_STOCH) ? 0
                        : (I.integrator == HB_HEUN || I.integrator == HB_HEUN_STOCH) ? 1 : 2;
        const int cfk = s.blocked && I.cfn == HB_CF_LINEAR ? (int)HB_CF_LINEAR : -1;
        groups[std::make_tuple(I.model, s.np, s.npt, (int)s.blocked, (int)s.ring_lds, icl, cfk)].push_back(i);
        P->node_steps += (double)I.n_nodes * (double)I.n_steps;
    }
    // The BOLD monitor's forward runs in helper waves of the blocked forward
    // kernel (one helper thread per node, workgroup doubled) when every
    // instance of a blocked group carries a BOLD loss and the doubled
    // workgroup fits 512 threads.
    std::vector<int> bold_fused(n, 0);
    for (auto& kv : groups) {
        bool f = std::get<3>(kv.first) != 0;
        int blk = 64;
        for (int i : kv.second) {
            f = f && inst[i].bold_period > 0 &&
                (spec[i].loss_kind == HB_LOSS_FC || spec[i].loss_kind == HB_LOSS_BOLD_MSE);
            blk = std::max(blk, sh[i].block);
        }
        if (f && 2 * blk <= 512 && !(opts && (opts->flags & HB_FLAG_SEPARATE_BOLD)))
            for (int i : kv.second) bold_fused[i] = 1;
    }
    // With checkpoint segments, the blocked reverse recomputes the segment
    // below the one it reverses in helper waves (one helper thread per node,
    // workgroup doubled, two tapes) when the doubled workgroup fits 512
    // threads; only the last segment's recompute is left on the critical path.
    std::vector<int> rc_fused(n, 0);
    for (auto& kv : groups) {
        bool f = std::get<3>(kv.first) != 0;
        int blk = 64;
        for (int i : kv.second) {
            f = f && sh[i].nseg > 1;
            blk = std::max(blk, sh[i].block);
        }
        if (f && 2 * blk <= 512)
            for (int i : kv.second) rc_fused[i] = 1;
    }

    // ---- arena layout
    struct Off {
        size_t x0, node_mc, node_jac, indptr, edges, indptr_t, edges_t, noise, ring, glog, ckpt, tape,
            zlog, bstate, bold, dsig, bscan, seedw, target, fc, ord, fin, loss, gcc, gp, gx0;
        bool has_node_mc, has_jac, has_noise, has_ring;
    };
    std::vector<Off> off(n);
    Arena A;
    const size_t inst_off = A.take(sizeof(AdjInst) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        const hb_instance& I = inst[i];
        const hb_adjoint_spec& Sp = spec[i];
        const Shape& s = sh[i];
        Off& o = off[i];
        const size_t N = I.n_nodes, T = I.n_steps;
        o.x0 = A.take(N * s.nv * 4);
        o.has_node_mc = I.node_param_idx >= 0;
        o.node_mc = o.has_node_mc ? A.take(N * kMC * 4) : 0;
        o.has_jac = Sp.grad_param >= 0;
        o.node_jac = o.has_jac ? A.take(N * kMC * 4) : 0;
        o.indptr = A.take((N + 1) * 4);
        o.edges = A.take((size_t)s.nnz * 8);
        o.indptr_t = A.take((N + 1) * 4);
        o.edges_t = A.take((size_t)s.nnz * 8);
        o.has_noise = I.integrator >= HB_EULER_STOCH;
        o.noise = o.has_noise ? A.take(T * N * s.nv * 4) : 0;
        o.has_ring = !s.ring_lds;
        o.ring = o.has_ring ? A.take(s.ring_b) : 0;
        o.glog = A.take(T * N * s.np * 4);
        o.ckpt = A.take((size_t)s.nseg * N * s.nv * 4);
        o.tape = A.take((size_t)(rc_fused[i] ? 2 : 1) * (s.S + 1) * N * s.nv * 4);
        const size_t nw = s.nwin;
        o.zlog = A.take(nw * N * 4);
        o.bstate = A.take(nw * N * 32);
        o.bold = A.take(nw * N * 8);
        o.dsig = A.take(nw * N * 8);
        o.seedw = A.take(nw * N * 4);
        o.bscan = nw > 0 ? A.take((size_t)16 * N * 20 * 8) : 0;
        const size_t tgt_n = Sp.loss_kind == HB_LOSS_FC ? N * N : Sp.loss_kind == HB_LOSS_BOLD_MSE ? nw * N : T;
        o.target = A.take(tgt_n * 4);
        o.fc = Sp.loss_kind == HB_LOSS_FC ? A.take((3 * N + 2 * N * N) * 8) : 0;
        o.ord = A.take((size_t)(s.S + 1) * 16);
    }
    // outputs of every instance in one region (one download): per instance
    // loss, dL/dcc0, dL/dp[N], dL/dx0[N][NV], final state[N][NV]
    size_t out_floats = 0;
    std::vector<size_t> out_at(n);
    for (int i = 0; i < n; ++i) {
        out_at[i] = out_floats;
        out_floats += 2 + (size_t)inst[i].n_nodes * (1 + 2 * sh[i].nv);
    }
    const size_t outs_off = A.take(out_floats * 4);
    for (int i = 0; i < n; ++i) {
        Off& o = off[i];
        const size_t N = inst[i].n_nodes, b = outs_off + out_at[i] * 4;
        o.loss = b;
        o.gcc = b + 4;
        o.gp = b + 8;
        o.gx0 = o.gp + N * 4;
        o.fin = o.gx0 + N * sh[i].nv * 4;
    }
    // noise jobs after the instances
    std::vector<NoiseJob> jobs;
    const size_t jobs_off = A.take(sizeof(NoiseJob) * (size_t)n);
    P->bytes = A.top;
    hipError_t he = hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipMalloc((void**)&P->arena, P->bytes);
    for (auto& e : P->ev)
        if (he == hipSuccess) he = hipEventCreate(&e);
    if (he != hipSuccess) {
        hb_adjoint_plan_destroy(P);
        return set_error(HB_BACKEND_ERROR, std::string("adjoint plan allocation failed: ") + hipGetErrorString(he));
    }
    char* base = P->arena;
    auto up = [&](size_t o, const void* src, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpy(base + o, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };

    // ---- per-instance uploads and descriptors (grouped order)
    std::vector<AdjInst> hdev(n);
    std::vector<int> grouped;
    for (auto& kv : groups)
        for (int i : kv.second) grouped.push_back(i);
    P->user_of = grouped;
    P->outs.resize(n);
    for (int gi = 0; gi < n; ++gi) {
        const int i = grouped[gi];
        const hb_instance& I = inst[i];
        const hb_adjoint_spec& Sp = spec[i];
        const Shape& s = sh[i];
        const Off& o = off[i];
        const int N = I.n_nodes, T = I.n_steps, np = model_nparams(I.model);
        AdjInst& D = hdev[gi];
        memset(&D, 0, sizeof(D));
        D.model = I.model; D.integrator = I.integrator; D.n_nodes = N; D.n_steps = T;
        D.np = s.np; D.cfn = I.cfn; D.engine = I.engine;
        D.has_coupling = I.wkind == HB_W_CSR ? 1 : 0;
        D.seg_len = s.S; D.n_seg = s.nseg; D.ring_depth = s.Hd; D.aring_depth = s.Hd;
        D.ring_lds = s.ring_lds; D.aring_lds = s.ring_lds;
        D.bold_period = I.bold_period > 0 ? I.bold_period : 0;
        D.n_win = s.nwin;
        D.loss_kind = Sp.loss_kind;
        D.bold_fused = bold_fused[i];
        D.rc_fused = rc_fused[i];
        D.dt = I.dt; D.dt_half = I.dt * 0.5f; D.dt_sixth = I.dt / 6.0f;
        model_consts(I.model, I.params, D.mc);
        int classic = 0;
        float cp[8];
        memcpy(cp, I.cparams, sizeof(cp));
        coupling_consts(I.cfn, cp, I.n_cparams, N, D.cc, &classic);
        for (int v = 0; v < s.nv; ++v) {
            const float ns = I.nsig ? I.nsig[v] : 0.0f;
            D.nscale[v] = sqrtf(2.0f * ns) * sqrtf(I.dt);   // integrator.rs:179-191
        }
        D.dt_ms = I.dt_ms;
        he = up(o.x0, I.x0, (size_t)N * s.nv * 4);
        std::vector<float> pn(np);
        std::vector<float> nmc, njac;
        if (o.has_node_mc || o.has_jac) {
            if (o.has_node_mc) nmc.resize((size_t)N * kMC);
            if (o.has_jac) njac.resize((size_t)N * kMC);
            for (int k = 0; k < N; ++k) {
                for (int q = 0; q < np; ++q) pn[q] = I.params[q];
                if (I.node_param_idx >= 0) pn[I.node_param_idx] = I.node_param[k];
                if (o.has_node_mc) model_consts(I.model, pn.data(), nmc.data() + (size_t)k * kMC);
                if (o.has_jac) model_jacobian(I.model, pn.data(), np, Sp.grad_param, njac.data() + (size_t)k * kMC);
            }
        }
        if (he == hipSuccess && o.has_node_mc) he = up(o.node_mc, nmc.data(), nmc.size() * 4);
        if (he == hipSuccess && o.has_jac) he = up(o.node_jac, njac.data(), njac.size() * 4);
        // CSR (by target) and CSC (by source), delay packed above the node index
        std::vector<int32_t> ip(N + 1, 0), ipt(N + 1, 0);
        std::vector<uint32_t> ed((size_t)s.nnz * 2), edt((size_t)s.nnz * 2);
        if (s.nnz > 0) {
            for (int r = 0; r <= N; ++r) ip[r] = I.indptr[r];
            for (int e = 0; e < s.nnz; ++e) ipt[I.indices[e] + 1]++;
            for (int r = 0; r < N; ++r) ipt[r + 1] += ipt[r];
            std::vector<int32_t> fillp(ipt.begin(), ipt.end() - 1);
            for (int r = 0; r < N; ++r)
                for (int e = I.indptr[r]; e < I.indptr[r + 1]; ++e) {
                    const uint32_t d = I.edge_delays ? (uint32_t)I.edge_delays[e] : 0u;
                    uint32_t wb;
                    memcpy(&wb, &I.weights[e], 4);
                    ed[2 * e] = (uint32_t)I.indices[e] | (d << 16);
                    ed[2 * e + 1] = wb;
                    const int k = fillp[I.indices[e]]++;
                    edt[2 * k] = (uint32_t)r | (d << 16);
                    edt[2 * k + 1] = wb;
                }
        }
        if (he == hipSuccess) he = up(o.indptr, ip.data(), ip.size() * 4);
        if (he == hipSuccess) he = up(o.indptr_t, ipt.data(), ipt.size() * 4);
        if (he == hipSuccess) he = up(o.edges, ed.data(), ed.size() * 4);
        if (he == hipSuccess) he = up(o.edges_t, edt.data(), edt.size() * 4);
        const size_t tgt_n = Sp.loss_kind == HB_LOSS_FC ? (size_t)N * N
                             : Sp.loss_kind == HB_LOSS_BOLD_MSE ? (size_t)s.nwin * N : (size_t)T;
        if (he == hipSuccess) he = up(o.target, Sp.target, tgt_n * 4);
        if (o.has_noise) {
            if (I.noise) {
                if (he == hipSuccess) he = up(o.noise, I.noise, (size_t)T * N * s.nv * 4);
            } else {
                NoiseJob j;
                j.out = (float*)(base + o.noise);
                j.n = (int64_t)T * N * s.nv;
                j.seed = I.noise_seed;
                j.stream = I.noise_stream;
                jobs.push_back(j);
            }
        }
        if (he != hipSuccess) {
            hb_adjoint_plan_destroy(P);
            return set_error(HB_BACKEND_ERROR, std::string("adjoint upload failed: ") + hipGetErrorString(he));
        }
        D.x0 = (const float*)(base + o.x0);
        D.node_mc = o.has_node_mc ? (const float*)(base + o.node_mc) : nullptr;
        D.node_jac = o.has_jac ? (const float*)(base + o.node_jac) : nullptr;
        D.indptr = (const int32_t*)(base + o.indptr);
        D.edges = (const uint2*)(base + o.edges);
        D.indptr_t = (const int32_t*)(base + o.indptr_t);
        D.edges_t = (const uint2*)(base + o.edges_t);
        D.noise = o.has_noise ? (const float*)(base + o.noise) : nullptr;
        D.ring = o.has_ring ? (float*)(base + o.ring) : nullptr;
        D.glog = (float*)(base + o.glog);
        D.ckpt = (float*)(base + o.ckpt);
        D.tape = (float*)(base + o.tape);
        D.zlog = (float*)(base + o.zlog);
        D.bstate = (double*)(base + o.bstate);
        D.bold = (double*)(base + o.bold);
        D.dsig = (double*)(base + o.dsig);
        D.bscan = s.nwin > 0 ? (double*)(base + o.bscan) : nullptr;
        D.seedw = (float*)(base + o.seedw);
        D.target = (const float*)(base + o.target);
        D.fc_work = Sp.loss_kind == HB_LOSS_FC ? (double*)(base + o.fc) : nullptr;
        D.ord_work = (float*)(base + o.ord);
        D.fin = (float*)(base + o.fin);
        D.out_loss = (float*)(base + o.loss);
        D.out_gcc = (float*)(base + o.gcc);
        D.out_gp = o.has_jac ? (float*)(base + o.gp) : nullptr;
        D.out_gx0 = (float*)(base + o.gx0);
        OutPtrs& op = P->outs[i];
        op.loss = o.loss; op.gcc = o.gcc; op.gp = o.gp; op.gx0 = o.gx0; op.fin = o.fin;
        op.n_nodes = N; op.nv = s.nv;
    }
    P->outs_off = outs_off;
    P->outs_bytes = out_floats * 4;
    P->d_inst = (AdjInst*)(base + inst_off);
    he = hipMemcpy(P->d_inst, hdev.data(), sizeof(AdjInst) * (size_t)n, hipMemcpyHostToDevice);
    if (he == hipSuccess && !jobs.empty()) {
        NoiseJob* dj = (NoiseJob*)(base + jobs_off);
        int64_t mx = 0;
        for (auto& j : jobs) mx = std::max(mx, j.n);
        he = hipMemcpy(dj, jobs.data(), sizeof(NoiseJob) * jobs.size(), hipMemcpyHostToDevice);
        if (he == hipSuccess) he = launch_noise(dj, (int)jobs.size(), mx, P->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(P->stream);
    }
    if (he != hipSuccess) {
        hb_adjoint_plan_destroy(P);
        return set_error(HB_BACKEND_ERROR, std::string("adjoint plan setup failed: ") + hipGetErrorString(he));
    }
    // ---- launch groups
    int pos = 0;
    for (auto& kv : groups) {
        AdjLaunch L;
        memset(&L, 0, sizeof(L));
        L.insts = P->d_inst + pos;
        L.n = (int)kv.second.size();
        L.model = std::get<0>(kv.first);
        L.np = std::get<1>(kv.first);
        L.npt = std::get<2>(kv.first);
        L.blocked = std::get<3>(kv.first) != 0;
        L.lring = std::get<4>(kv.first) != 0;
        L.ic = std::get<5>(kv.first);
        L.cfk = std::get<6>(kv.first);
        int blk = 64, lds = 0;
        for (int i : kv.second) {
            blk = std::max(blk, sh[i].block);
            if (sh[i].ring_lds) lds = std::max(lds, (int)sh[i].ring_b);
        }
        L.bold_fused = bold_fused[kv.second[0]] != 0;
        L.rc_fused = rc_fused[kv.second[0]] != 0;
        L.block = blk;
        L.lds_fwd = lds;
        L.lds_rev = lds;
        P->launches.push_back(L);
        pos += L.n;
    }
    *out = P;
    return HB_OK;
}

int hb_adjoint_forward(hb_adjoint_plan* P, hb_adjoint_stats* st) {
    if (!P) return set_error(HB_INVALID_STATE, "null adjoint plan");
    HB_HIP(hipSetDevice(P->device));
    HB_HIP(hipEventRecord(P->ev[0], P->stream));
    for (int stage = 0; stage <= 2; ++stage)
        for (auto& L : P->launches) HB_HIP(launch_adjoint_stage(stage, L, P->stream));
    HB_HIP(hipEventRecord(P->ev[1], P->stream));
    HB_HIP(hipEventSynchronize(P->ev[1]));
    P->forward_done = true;
    if (st) {
        float ms = 0.0f;
        HB_HIP(hipEventElapsedTime(&ms, P->ev[0], P->ev[1]));
        st->forward_ms = ms;
        st->node_steps = P->node_steps;
        st->device_bytes = (int64_t)P->bytes;
    }
    return HB_OK;
}

int hb_adjoint_backward(hb_adjoint_plan* P, hb_adjoint_stats* st) {
    if (!P) return set_error(HB_INVALID_STATE, "null adjoint plan");
    if (!P->forward_done) return set_error(HB_INVALID_STATE, "hb_adjoint_backward before hb_adjoint_forward");
    HB_HIP(hipSetDevice(P->device));
    HB_HIP(hipEventRecord(P->ev[1], P->stream));
    for (int stage = 3; stage <= 4; ++stage)
        for (auto& L : P->launches) HB_HIP(launch_adjoint_stage(stage, L, P->stream));
    HB_HIP(hipEventRecord(P->ev[2], P->stream));
    HB_HIP(hipEventSynchronize(P->ev[2]));
    if (st) {
        float ms = 0.0f;
        HB_HIP(hipEventElapsedTime(&ms, P->ev[1], P->ev[2]));
        st->backward_ms = ms;
        st->node_steps = P->node_steps;
        st->device_bytes = (int64_t)P->bytes;
    }
    return HB_OK;
}

int hb_adjoint_download(hb_adjoint_plan* P) {
    if (!P) return set_error(HB_INVALID_STATE, "null adjoint plan");
    HB_HIP(hipSetDevice(P->device));
    HB_HIP(hipStreamSynchronize(P->stream));
    std::vector<char> host(P->outs_bytes);
    HB_HIP(hipMemcpy(host.data(), P->arena + P->outs_off, P->outs_bytes, hipMemcpyDeviceToHost));
    const char* base = host.data() - P->outs_off;   // arena offsets into the staged copy
    for (size_t i = 0; i < P->spec.size(); ++i) {
        const hb_adjoint_spec& S = P->spec[i];
        const hb_instance& I = P->inst[i];
        const OutPtrs& o = P->outs[i];
        memcpy(S.loss, base + o.loss, 4);
        float gcc = 0.0f;
        memcpy(&gcc, base + o.gcc, 4);
        // Kuramoto holds a / n_src (coupling.rs:370); the gradient is reported for a
        S.grad_coupling[0] = I.cfn == HB_CF_KURAMOTO ? gcc / (float)I.n_nodes : gcc;
        if (S.grad_param >= 0 && S.grad_param_out) memcpy(S.grad_param_out, base + o.gp, (size_t)o.n_nodes * 4);
        if (S.grad_x0) memcpy(S.grad_x0, base + o.gx0, (size_t)o.n_nodes * o.nv * 4);
        if (S.final_state) memcpy(S.final_state, base + o.fin, (size_t)o.n_nodes * o.nv * 4);
    }
    return HB_OK;
}

}  // extern "C"
