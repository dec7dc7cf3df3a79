// hb_forward.hip -- fused persistent forward integration kernels for gfx950.
//
// One workgroup integrates one simulation instance at a time, end to end:
//   per step: ring write of pre(x_t) -> barrier -> per-node delayed gather
//   (CSR row / dense column / mean field) -> post -> Heun/RK4/Euler stages in
//   registers -> clamp -> tavg/trajectory/BOLD.
// The delay history is a ring of pre(x) values ([depth][N][np] floats) held in
// LDS when it fits (160 KB per CU) and otherwise in a per-instance HBM region.
// Ring depth = max_delay + 2 so a single barrier per step suffices: the slot
// written at step t+1 is never read at step t.
//
// Reference semantics restated here (file:line in maedoc/hyburn):
//   time loop          src/engine/batch_engine/engine.rs:540-823
//   dfuns, clamps      src/engine/batch_engine/dfun.rs:170-239, 271-516
//   integrators        engine.rs:676-773, src/engine/integrator.rs:110-315
//   coupling pre/post  src/engine/coupling.rs:204-444
//   per-edge delays    src/engine/runtime.rs:121-211 (ring: 55-85)
//   BOLD               src/engine/bold.rs:106-166, bold_monitor.rs:72-227
//
// This file is compiled twice: HB_EXACT=0 (FMA contraction allowed) and
// HB_EXACT=1 (-ffp-contract=off: every f32 op rounds as written, so results
// are bit-identical to a correctly rounded restatement of the same ops).
#include "hb_internal.h"

#include <cstring>
#include <type_traits>

#ifndef HB_EXACT
#define HB_EXACT 0
#endif
#if HB_EXACT
#pragma clang fp contract(off)
#define HB_NS exact
#else
#define HB_NS fast
#endif

namespace hb {
namespace HB_NS {

// CSR edges a node keeps in registers for the whole run (longer rows spill
// the rest to the memory path), and the noise prefetch distance in steps.
constexpr int kMaxRegEdges = 32;
constexpr int kNoiseAhead = 4;
// Steps per delay block (needs min delay >= kBlockSteps). Deterministic runs
// whose ring lives in HBM use kLongBlockSteps: each edge's window is then 64
// contiguous bytes (fewer DRAM segments per step) and barriers halve.
constexpr int kBlockSteps = 8, kLongBlockSteps = 16, kLongerBlockSteps = 32;

#include "hb_model.h"

// ---------------------------------------------------------------- view -----
//
// Descriptor fields read once per instance into wave-uniform values (SGPRs),
// with device pointers cast back to the global address space so the loops
// use global (not flat) memory instructions and never re-read the descriptor.
template <typename T> using gp = __attribute__((address_space(1))) T*;

__device__ __forceinline__ int u32(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uf32(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }
__device__ __forceinline__ uint64_t u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double uf64(double x) { return __longlong_as_double((long long)u64((uint64_t)__double_as_longlong(x))); }
template <typename T> __device__ __forceinline__ gp<T> ug(T* p) { return (gp<T>)((T*)u64((uint64_t)p)); }
__device__ __forceinline__ uint2 edge_at(gp<const uint32_t> e, int i) {
    return make_uint2(e[2 * i], e[2 * i + 1]);
}

struct View {
    int model, integrator, n_nodes, n_steps, np, wkind, cfn, delay_mode, single_delay, ring_depth;
    int sweep_form, node_param_idx, bold_period, bold_nout, bold_native, cpu_form, sjr_classic, engine;
    int max_row, blocked_k, lpn, lpn_shift, step0, step0_modh, hist_len, bold_carry;
    float dt, dt_half, dt_sixth, inv_nsteps, sweep_val, scalar_w, fill0, fill1;
    double dt_ms, tr_ms, dtb_ms;
    int64_t traj_tstride;
    gp<const float> x0, node_mc, dense_t, noise;
    gp<const int32_t> indptr;
    gp<const uint32_t> edges;   // 2 words per edge: src | delay << 16, weight bits
    gp<float> ring, tavg, fin, traj, bold;
    gp<const float> hist_in;
    gp<float> hist_out, bold_io;
};

__device__ __forceinline__ View make_view(const DevInst* I) {
    View V;
    V.model = u32(I->model); V.integrator = u32(I->integrator); V.n_nodes = u32(I->n_nodes);
    V.n_steps = u32(I->n_steps); V.np = u32(I->np); V.wkind = u32(I->wkind); V.cfn = u32(I->cfn);
    V.delay_mode = u32(I->delay_mode); V.single_delay = u32(I->single_delay); V.ring_depth = u32(I->ring_depth);
    V.sweep_form = u32(I->sweep_form); V.node_param_idx = u32(I->node_param_idx);
    V.bold_period = u32(I->bold_period); V.bold_nout = u32(I->bold_nout); V.bold_native = u32(I->bold_native);
    V.cpu_form = u32(I->cpu_form); V.sjr_classic = u32(I->sjr_classic); V.engine = u32(I->engine);
    V.max_row = u32(I->max_row); V.blocked_k = u32(I->blocked_k);
    V.lpn = u32(I->lpn); V.lpn_shift = u32(I->lpn_shift);
    V.step0 = u32(I->step0); V.step0_modh = u32(I->step0_modh);
    V.hist_len = u32(I->hist_len); V.bold_carry = u32(I->bold_carry);
    V.dt = uf32(I->dt); V.dt_half = uf32(I->dt_half); V.dt_sixth = uf32(I->dt_sixth); V.inv_nsteps = uf32(I->inv_nsteps);
    V.sweep_val = uf32(I->sweep_val); V.scalar_w = uf32(I->scalar_w); V.fill0 = uf32(I->fill0); V.fill1 = uf32(I->fill1);
    V.dt_ms = uf64(I->dt_ms); V.tr_ms = uf64(I->tr_ms); V.dtb_ms = uf64(I->dtb_ms);
    V.traj_tstride = (int64_t)u64((uint64_t)I->traj_tstride);
    V.x0 = ug(I->x0); V.node_mc = ug(I->node_mc); V.dense_t = ug(I->dense_t); V.noise = ug(I->noise);
    V.indptr = ug(I->indptr); V.edges = ug((const uint32_t*)I->edges);
    V.ring = ug(I->ring); V.tavg = ug(I->tavg); V.fin = ug(I->fin); V.traj = ug(I->traj); V.bold = ug(I->bold);
    V.hist_in = ug(I->hist_in); V.hist_out = ug(I->hist_out); V.bold_io = ug(I->bold_io);
    return V;
}

// Global step of local step t of a continued run, saturated at 2^30 (it only
// enters comparisons with delays <= 65535 and with 0).
__device__ __forceinline__ int global_step(const View& V, int t) {
    return t >= (1 << 30) - V.step0 ? (1 << 30) : t + V.step0;
}

// Sum over the lanes that share a node (1, 2 or 4 adjacent lanes of a quad).
// Both quad swizzles are DPP moves; float addition is commutative, so every
// lane of the group ends with bitwise the same total.
__device__ __forceinline__ float lane_group_sum(float v, int lpn) {
    if (lpn >= 2) v = v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
    if (lpn >= 4) v = v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
    return v;
}

// ---------------------------------------------------------------- BOLD -----

// Emit TR-resampled outputs after native sample k (time tk, value sig) is
// produced; bold_monitor.rs:123-203 evaluated as a stream. Uniform control.
struct BoldClock {
    int k;        // native samples produced so far
    int j;        // next TR row
    double tprev; // time of the previous native sample
};

__device__ __forceinline__ void bold_emit(const View& I, const BoldClock& bc, double tk, float sig,
                                          float prev, gp<float> out_col, int N) {
    if (I.bold_native) {
        if (bc.k < I.bold_nout) out_col[(size_t)bc.k * N] = sig;
        return;
    }
    const double tr = I.tr_ms;
    int j = bc.j;
    if (bc.k == 0) {
        while (j < I.bold_nout && (double)(j + 1) * tr <= tk) { out_col[(size_t)j * N] = sig; ++j; }
    } else {
        while (j < I.bold_nout && (double)(j + 1) * tr < tk) {
            const double tt = (double)(j + 1) * tr;
            const double d = tk - bc.tprev;
            double w = fabs(d) < 1e-12 ? 0.0 : (tt - bc.tprev) / d;
            w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
            out_col[(size_t)j * N] = prev + (sig - prev) * (float)w;
            ++j;
        }
    }
}

// Advance the shared (uniform) clock after native sample k at time tk.
__device__ __forceinline__ void bold_advance(const View& I, BoldClock& bc, double tk) {
    if (!I.bold_native) {
        const double tr = I.tr_ms;
        if (bc.k == 0) { while (bc.j < I.bold_nout && (double)(bc.j + 1) * tr <= tk) bc.j++; }
        else { while (bc.j < I.bold_nout && (double)(bc.j + 1) * tr < tk) bc.j++; }
    }
    bc.k++;
    bc.tprev = tk;
}


// ---------------------------------------------------------------- instance -

// Integrate instance I with the whole workgroup. lds: dynamic LDS base.
template <int M, int NP, bool LRING, bool MULTI>
__device__ __forceinline__ void run_instance(const DevInst* __restrict__ I, float* lds) {
    const DevInst* __restrict__ Ia = I;
    const View V = make_view(I);
    constexpr int NV = Tr<M>::NV, NC = Tr<M>::NC;
    const int N = V.n_nodes, T = V.n_steps, Hd = V.ring_depth;
    const int tid = threadIdx.x, BS = blockDim.x;
    const int NNP = N * NP;
    using RingPtr = std::conditional_t<LRING, float*, gp<float>>;
    RingPtr ring;
    if constexpr (LRING) ring = lds; else ring = V.ring;
    // MULTI state table after the ring: x[N][NV], tavg[N][NV], bold[N][6]
    float* __restrict__ tab = LRING ? lds + (size_t)Hd * NNP : lds;
    float* __restrict__ tab_x = tab;
    float* __restrict__ tab_t = tab + (size_t)N * NV;
    float* __restrict__ tab_b = tab + (size_t)2 * N * NV;

    const int cfn = V.cfn, wkind = V.wkind, dmode = V.delay_mode, integ = V.integrator;
    const bool cpu_form = V.cpu_form != 0, classic = V.sjr_classic != 0;
    const bool sw = V.sweep_form != 0;
    const float sv = V.sweep_val;
    const float dt = V.dt, dth = V.dt_half, dt6 = V.dt_sixth;
    const bool stoch = integ >= HB_EULER_STOCH;
    const bool plain = V.engine == HB_ENGINE_PLAIN;
    StepCtx S;
    S.dt = dt; S.dth = dth; S.dt6 = dt6; S.integ = integ; S.stoch = stoch; S.plain = plain; S.sw = sw; S.sv = sv;
    float cc[kCC];
#pragma unroll
    for (int i = 0; i < kCC; ++i) cc[i] = Ia->cc[i];
    float nsc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) nsc[v] = Ia->nscale[v];
    const float fill[2] = {V.fill0, V.fill1};

    // ring initial fill: 0 (per-edge: exact-zero contributions) or pre(0)
    {
        const int total = Hd * NNP;
        for (int i = tid; i < total; i += BS) ring[i] = (NP == 2 && (i & 1)) ? fill[1] : fill[0];
    }
    // a continued run: the carried pre() of times -L..-1 go to slots Hd-L..Hd-1
    // (after the fill, which another thread may be writing to the same slots)
    if (V.hist_in) {
        __syncthreads();
        const int L = V.hist_len, total = L * NNP;
        for (int i = tid; i < total; i += BS) ring[(size_t)(Hd - L) * NNP + i] = V.hist_in[i];
    }

    const int nper = MULTI ? (N + BS - 1) / BS : 1;
    // model constants (uniform, or per node when a parameter is given per node)
    float mc_u[kMC];
#pragma unroll
    for (int i = 0; i < kMC; ++i) mc_u[i] = Ia->mc[i];
    const int npi = V.node_param_idx;

    // per-thread node state (MULTI=false keeps it in registers)
    float x[NV], tv[NV];
    Bold bw;
    bw.s = 0.0f; bw.f = 1.0f; bw.v = 1.0f; bw.q = 1.0f; bw.acc = 0.0f; bw.prev = 0.0f;
    const bool own = tid < N;
    if (!MULTI && V.bold_carry && own) {
        bw.s = V.bold_io[tid * 4 + 0]; bw.f = V.bold_io[tid * 4 + 1];
        bw.v = V.bold_io[tid * 4 + 2]; bw.q = V.bold_io[tid * 4 + 3];
    }
    // ---- CSR row of this thread's node, folded into ring offsets and kept in
    // registers: offset = delay * N*np - src * np (read address = ws*N*np - offset)
    const bool csr_reg = !MULTI && wkind == HB_W_CSR;
    const bool per_edge = dmode == HB_DELAY_PER_EDGE;
    const int kreg = csr_reg ? min(V.max_row, kMaxRegEdges) : 0;   // uniform
    int eo[kMaxRegEdges];
    float ew[kMaxRegEdges];
    int r_e0 = 0, r_e1 = 0;
    if (csr_reg) {
        if (own) { r_e0 = V.indptr[tid]; r_e1 = V.indptr[tid + 1]; }
#pragma unroll
        for (int k = 0; k < kMaxRegEdges; ++k) {
            eo[k] = -tid * NP;   // padding: zero weight on the node's own slot
            ew[k] = 0.0f;
            if (k < kreg && own && r_e0 + k < r_e1) {
                const uint2 ed = edge_at(V.edges, r_e0 + k);
                const int j = ed.x & 0xFFFF, d = per_edge ? (int)(ed.x >> 16) : 0;
                eo[k] = d * NNP - j * NP;
                ew[k] = __uint_as_float(ed.y);
            }
        }
    }
    // ---- noise prefetch FIFO (kNoiseAhead steps in flight)
    float zq[kNoiseAhead][NV];
    if (!MULTI && stoch) {
#pragma unroll
        for (int q = 0; q < kNoiseAhead; ++q)
#pragma unroll
            for (int v = 0; v < NV; ++v)
                zq[q][v] = (own && q < T) ? V.noise[((size_t)q * N + tid) * NV + v] : 0.0f;
    }
    if (!MULTI) {
#pragma unroll
        for (int v = 0; v < NV; ++v) { x[v] = own ? V.x0[tid * NV + v] : 0.0f; tv[v] = 0.0f; }
    } else {
        for (int n = tid; n < N; n += BS) {
#pragma unroll
            for (int v = 0; v < NV; ++v) { tab_x[n * NV + v] = V.x0[n * NV + v]; tab_t[n * NV + v] = 0.0f; }
            float* b = tab_b + n * 6;
            b[0] = 0.0f; b[1] = 1.0f; b[2] = 1.0f; b[3] = 1.0f; b[4] = 0.0f; b[5] = 0.0f;
            if (V.bold_carry) { b[0] = V.bold_io[n * 4]; b[1] = V.bold_io[n * 4 + 1]; b[2] = V.bold_io[n * 4 + 2]; b[3] = V.bold_io[n * 4 + 3]; }
        }
    }
    float mc_n[kMC];
    if (!MULTI) {
#pragma unroll
        for (int i = 0; i < kMC; ++i) mc_n[i] = (npi >= 0 && own) ? V.node_mc[tid * kMC + i] : mc_u[i];
    }
    const int bper = V.bold_period;
    BoldClock bc;
    bc.k = 0; bc.j = 0; bc.tprev = 0.0;
    int bcount = 0;
    __syncthreads();

    int ws = 0;
    // batch-window position of the global step: (step0 + t) mod (max_delay + 1)
    int tm = V.step0_modh;
    for (int t = 0; t < T; ++t) {
        const int tg = global_step(V, t);
        // ---- phase 1: ring[ws][n] = pre(x_t[n])
        for (int kk = 0; kk < nper; ++kk) {
            const int n = MULTI ? tid + kk * BS : tid;
            if (n < N) {
                float xc[NC], pr[NP];
#pragma unroll
                for (int c = 0; c < NC; ++c) xc[c] = MULTI ? tab_x[n * NV + c] : x[c];
                pre<NC, NP>(cfn, cc, cpu_form, classic, xc, pr);
                auto dst = ring + (size_t)ws * NNP + n * NP;
#pragma unroll
                for (int c = 0; c < NP; ++c) dst[c] = pr[c];
            }
        }
        __syncthreads();

        // uniform read slot for single-delay projections
        int slot_u = ws;
        if (dmode != HB_DELAY_PER_EDGE && tg > 0) {
            const int d = V.single_delay;
            int teff_back = 0;   // how many steps back
            if (d == 0) teff_back = 0;
            else if (d > tg) teff_back = -1;
            else if (dmode == HB_DELAY_BATCH_COMPAT) {
                const int H = Hd - 1;   // reference horizon = max_delay + 1
                int r = tm - d + H;     // (t - d + H) mod H, tm = t mod H
                r -= r >= H ? H : 0;
                const int back = H - 1 - r;
                teff_back = back > tg ? -1 : back;
            } else teff_back = d;
            if (teff_back < 0) slot_u = -1;
            else { slot_u = ws - teff_back; if (slot_u < 0) slot_u += Hd; }
        }
        const int dmask = tg == 0 ? 0 : 0xFFFF;
        // mean-field sum (sequential over sources, identical in every thread)
        float mf[NP];
        if (wkind == HB_W_SCALAR_MEAN) {
#pragma unroll
            for (int c = 0; c < NP; ++c) {
                float s = 0.0f;
                if (slot_u >= 0) {
                    const auto src = ring + (size_t)slot_u * NNP + c;
                    for (int j = 0; j < N; ++j) s += src[j * NP];
                } else {
                    for (int j = 0; j < N; ++j) s += fill[c];
                }
                mf[c] = (s / (float)N) * V.scalar_w;
            }
        }

        // ---- phase 2: gather + integrate, per owned node
        for (int kk = 0; kk < nper; ++kk) {
            const int n = MULTI ? tid + kk * BS : tid;
            if (n < N) {
                float z[NV];
                if (stoch) {
                    if (!MULTI) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) z[v] = zq[0][v] * nsc[v];
#pragma unroll
                        for (int q = 0; q + 1 < kNoiseAhead; ++q)
#pragma unroll
                            for (int v = 0; v < NV; ++v) zq[q][v] = zq[q + 1][v];
                        if (t + kNoiseAhead < T) {
                            const gp<const float> nz = V.noise + ((size_t)(t + kNoiseAhead) * N + n) * NV;
#pragma unroll
                            for (int v = 0; v < NV; ++v) zq[kNoiseAhead - 1][v] = nz[v];
                        }
                    } else {
                        const gp<const float> nz = V.noise + ((size_t)t * N + n) * NV;
#pragma unroll
                        for (int v = 0; v < NV; ++v) z[v] = nz[v] * nsc[v];
                    }
                }
                float g[NP];
#pragma unroll
                for (int c = 0; c < NP; ++c) g[c] = 0.0f;
                if (wkind == HB_W_CSR) {
                    int e_mem0, e_mem1;
                    // register-resident rows (steady state): same summation order as the row
                    const bool reg_ok = csr_reg && (per_edge ? tg > 0 : slot_u >= 0);
                    if (reg_ok) {
                        const int base = (per_edge ? ws : slot_u) * NNP;
                        const int wrap = Hd * NNP;
#pragma unroll
                        for (int k = 0; k < kMaxRegEdges; ++k) {
                            if (k < kreg) {
                                int a = base - eo[k];
                                a += a < 0 ? wrap : 0;
#pragma unroll
                                for (int c = 0; c < NP; ++c) g[c] += ew[k] * ring[a + c];
                            }
                        }
                        e_mem0 = r_e0 + kreg;
                        e_mem1 = r_e1;
                    } else {
                        e_mem0 = csr_reg ? r_e0 : V.indptr[n];
                        e_mem1 = csr_reg ? r_e1 : V.indptr[n + 1];
                    }
                    if (e_mem0 < e_mem1) {
                        if (per_edge) {
                            // 8 edges per batch: edge records, then ring values, then
                            // the FMAs in row order (same sums as one edge at a time)
                            constexpr int kB = 8;
                            int e = e_mem0;
                            for (; e + kB <= e_mem1; e += kB) {
                                uint2 ed[kB];
#pragma unroll
                                for (int k = 0; k < kB; ++k) ed[k] = edge_at(V.edges, e + k);
                                float val[kB][NP];
#pragma unroll
                                for (int k = 0; k < kB; ++k) {
                                    const int j = ed[k].x & 0xFFFF;
                                    const int d = (ed[k].x >> 16) & dmask;
                                    int sl = ws - d;
                                    sl += sl < 0 ? Hd : 0;
                                    const auto src = ring + (size_t)sl * NNP + j * NP;
#pragma unroll
                                    for (int c = 0; c < NP; ++c) val[k][c] = src[c];
                                }
#pragma unroll
                                for (int k = 0; k < kB; ++k) {
                                    const float w = __uint_as_float(ed[k].y);
#pragma unroll
                                    for (int c = 0; c < NP; ++c) g[c] += w * val[k][c];
                                }
                            }
                            for (; e < e_mem1; ++e) {
                                const uint2 ed = edge_at(V.edges, e);
                                const int j = ed.x & 0xFFFF;
                                const int d = (ed.x >> 16) & dmask;
                                const float w = __uint_as_float(ed.y);
                                int sl = ws - d;
                                sl += sl < 0 ? Hd : 0;
                                const auto src = ring + (size_t)sl * NNP + j * NP;
#pragma unroll
                                for (int c = 0; c < NP; ++c) g[c] += w * src[c];
                            }
                        } else {
                            for (int e = e_mem0; e < e_mem1; ++e) {
                                const uint2 ed = edge_at(V.edges, e);
                                const int j = ed.x & 0xFFFF;
                                const float w = __uint_as_float(ed.y);
#pragma unroll
                                for (int c = 0; c < NP; ++c)
                                    g[c] += w * (slot_u >= 0 ? ring[(size_t)slot_u * NNP + j * NP + c] : fill[c]);
                            }
                        }
                    }
                } else if (wkind == HB_W_DENSE) {
                    const gp<const float> wt = V.dense_t;
                    for (int j = 0; j < N; ++j) {
                        const float w = wt[(size_t)j * N + n];
#pragma unroll
                        for (int c = 0; c < NP; ++c)
                            g[c] += w * (slot_u >= 0 ? ring[(size_t)slot_u * NNP + j * NP + c] : fill[c]);
                    }
                } else if (wkind == HB_W_SCALAR_MEAN) {
#pragma unroll
                    for (int c = 0; c < NP; ++c) g[c] = mf[c];
                } else if (wkind == HB_W_SCALAR_DIAG) {
#pragma unroll
                    for (int c = 0; c < NP; ++c)
                        g[c] = V.scalar_w * (slot_u >= 0 ? ring[(size_t)slot_u * NNP + n * NP + c] : fill[c]);
                }
                float cpl[NC];
                if (wkind != HB_W_NONE) {
                    float own_pre[NP];
                    const auto op = ring + (size_t)ws * NNP + n * NP;
#pragma unroll
                    for (int c = 0; c < NP; ++c) own_pre[c] = op[c];
                    post<NC, NP>(cfn, cc, classic, g, own_pre, cpl);
                } else {
#pragma unroll
                    for (int c = 0; c < NC; ++c) cpl[c] = 0.0f;
                }

                float xs[NV];
                float mcl[kMC];
                if (!MULTI) {
#pragma unroll
                    for (int i = 0; i < kMC; ++i) mcl[i] = mc_n[i];
#pragma unroll
                    for (int v = 0; v < NV; ++v) xs[v] = x[v];
                } else {
#pragma unroll
                    for (int v = 0; v < NV; ++v) xs[v] = tab_x[n * NV + v];
                    const float* msrc = npi >= 0 ? (const float*)(V.node_mc + (size_t)n * kMC) : Ia->mc;
#pragma unroll
                    for (int i = 0; i < kMC; ++i) mcl[i] = msrc[i];
                }
                const float* mc = mcl;

                integrate<M>(S, mc, cpl, z, xs);

                // ---- outputs
                if (!MULTI) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) { x[v] = xs[v]; tv[v] = tv[v] + xs[v]; }
                } else {
#pragma unroll
                    for (int v = 0; v < NV; ++v) { tab_x[n * NV + v] = xs[v]; tab_t[n * NV + v] = tab_t[n * NV + v] + xs[v]; }
                }
                if (V.traj) {
                    gp<float> tr = V.traj + (size_t)t * V.traj_tstride + (size_t)n * NV;
#pragma unroll
                    for (int v = 0; v < NV; ++v) tr[v] = xs[v];
                }
                if (bper > 0) {
                    if (!MULTI) bw.acc = bw.acc + xs[0];
                    else tab_b[n * 6 + 4] = tab_b[n * 6 + 4] + xs[0];
                }
            }
        }
        // ---- BOLD window close (uniform), bold_monitor.rs:86-106
        if (bper > 0) {
            bcount++;
            if (bcount >= bper) {
                const double dtb = V.dt_ms * (double)bcount;
                const float dts = (float)(dtb / 1000.0);
                const double tk = __dadd_rn(__dmul_rn((double)bc.k, dtb), dtb);
                for (int kk = 0; kk < nper; ++kk) {
                    const int n = MULTI ? tid + kk * BS : tid;
                    if (n < N) {
                        Bold b;
                        if (!MULTI) b = bw;
                        else { const float* q = tab_b + n * 6; b.s = q[0]; b.f = q[1]; b.v = q[2]; b.q = q[3]; b.acc = q[4]; b.prev = q[5]; }
                        const float sig = bold_step(b, b.acc / (float)bcount, dts);
                        bold_emit(V, bc, tk, sig, b.prev, V.bold + n, N);
                        b.prev = sig;
                        b.acc = 0.0f;
                        if (!MULTI) bw = b;
                        else { float* q = tab_b + n * 6; q[0] = b.s; q[1] = b.f; q[2] = b.v; q[3] = b.q; q[4] = b.acc; q[5] = b.prev; }
                    }
                }
                bold_advance(V, bc, tk);
                bcount = 0;
            }
        }
        ws = (ws + 1 == Hd) ? 0 : ws + 1;
        tm = (tm + 1 == Hd - 1) ? 0 : tm + 1;
    }

    // ---- the carried history: pre() of times T-L..T-1, oldest first
    if (V.hist_out) {
        __syncthreads();
        const int L = V.hist_len, total = L * NNP;
        for (int i = tid; i < total; i += BS) {
            const int j = i / NNP;
            int sl = ws - L + j;
            sl += sl < 0 ? Hd : 0;
            V.hist_out[i] = ring[(size_t)sl * NNP + (i - j * NNP)];
        }
    }

    // ---- BOLD flush of a partial window, then nearest-sample tail rows
    if (bper > 0) {
        double tlast = bc.tprev;
        if (bcount > 0) {
            const double dtb = V.dt_ms * (double)bcount;
            const float dts = (float)(dtb / 1000.0);
            const double tk = __dadd_rn(__dmul_rn((double)bc.k, V.dtb_ms), dtb);
            for (int kk = 0; kk < nper; ++kk) {
                const int n = MULTI ? tid + kk * BS : tid;
                if (n < N) {
                    Bold b;
                    if (!MULTI) b = bw;
                    else { const float* q = tab_b + n * 6; b.s = q[0]; b.f = q[1]; b.v = q[2]; b.q = q[3]; b.acc = q[4]; b.prev = q[5]; }
                    const float sig = bold_step(b, b.acc / (float)bcount, dts);
                    bold_emit(V, bc, tk, sig, b.prev, V.bold + n, N);
                    b.prev = sig;
                    if (!MULTI) bw = b;
                    else tab_b[n * 6 + 5] = sig;
                }
            }
            bold_advance(V, bc, tk);
            tlast = tk;
        }
        (void)tlast;
        if (!V.bold_native && bc.k > 0) {
            for (int kk = 0; kk < nper; ++kk) {
                const int n = MULTI ? tid + kk * BS : tid;
                if (n < N) {
                    const float last = MULTI ? tab_b[n * 6 + 5] : bw.prev;
                    for (int jj = bc.j; jj < V.bold_nout; ++jj) V.bold[(size_t)jj * N + n] = last;
                }
            }
        }
    }

    // ---- the monitor's Balloon-Windkessel state for a continued run
    if (bper > 0 && V.bold_io) {
        for (int kk = 0; kk < nper; ++kk) {
            const int n = MULTI ? tid + kk * BS : tid;
            if (n < N) {
                float q4[4];
                if (!MULTI) { q4[0] = bw.s; q4[1] = bw.f; q4[2] = bw.v; q4[3] = bw.q; }
                else { const float* q = tab_b + n * 6; q4[0] = q[0]; q4[1] = q[1]; q4[2] = q[2]; q4[3] = q[3]; }
#pragma unroll
                for (int k = 0; k < 4; ++k) V.bold_io[n * 4 + k] = q4[k];
            }
        }
    }

    // ---- tavg * (1/n_steps), final state (engine.rs:813-818)
    const float inv = V.inv_nsteps;
    for (int kk = 0; kk < nper; ++kk) {
        const int n = MULTI ? tid + kk * BS : tid;
        if (n < N) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const float xv = MULTI ? tab_x[n * NV + v] : x[v];
                const float tt = MULTI ? tab_t[n * NV + v] : tv[v];
                V.fin[n * NV + v] = xv;
                V.tavg[n * NV + v] = tt * inv;
            }
        }
    }
    __syncthreads();
}


// ------------------------------------------------------- delay-blocked path -
//
// For per-edge CSR delays whose minimum d_min is >= K, the coupling inputs of
// K consecutive steps t0..t0+K-1 read only history times <= t0-1, already in
// the ring when the block starts. One barrier per block suffices, and each
// edge's K values are one contiguous run of the node-major ring
// [np][N][Hd] (Hd >= max_delay + K + 1 so in-block writes never alias reads).
// Per-step arithmetic and the per-row summation order are unchanged.
// CF >= 0 fixes the coupling function at compile time (-1: read per
// instance); OUTS = false drops the trajectory and BOLD outputs. Both remove
// per-step uniform branches from the step loop of the common sweep case.
template <int M, int NP, bool LRING, int K, int IC, int CF, bool OUTS, bool WIDE = false>
__device__ __forceinline__ void run_blocked(const DevInst* __restrict__ I, float* lds) {
    const DevInst* __restrict__ Ia = I;
    const View V = make_view(I);
    constexpr int NV = Tr<M>::NV, NC = Tr<M>::NC;
    const int N = V.n_nodes, T = V.n_steps, Hd = V.ring_depth;
    const int tid = threadIdx.x, BS = blockDim.x;
    const int Hs = Hd + K - 1;      // row stride: ring slots + mirrored copy of slots 0..K-2
    const int NHd = N * Hs;
    // An HBM ring with several channels interleaves them per slot, so each
    // edge's K-step window of every channel is one contiguous run (one or two
    // cache lines instead of one per channel); LDS rings keep channel planes.
    constexpr bool ILV = !LRING && NP > 1;
    auto ridx = [&](int c, int off) { return ILV ? off * NP + c : c * NHd + off; };
    using RingPtr = std::conditional_t<LRING, float*, gp<float>>;
    RingPtr ring;
    if constexpr (LRING) ring = lds; else ring = V.ring;
    const int cfn = CF >= 0 ? CF : V.cfn;
    const bool cpu_form = V.cpu_form != 0, classic = V.sjr_classic != 0;
    StepCtx S;
    S.dt = V.dt; S.dth = V.dt_half; S.dt6 = V.dt_sixth; S.integ = V.integrator;
    S.stoch = S.integ >= HB_EULER_STOCH; S.plain = V.engine == HB_ENGINE_PLAIN;
    S.sw = V.sweep_form != 0; S.sv = V.sweep_val;
    float cc[kCC];
#pragma unroll
    for (int i = 0; i < kCC; ++i) cc[i] = Ia->cc[i];
    float nsc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) nsc[v] = Ia->nscale[v];
    const gp<const float> noise = V.noise;
    const gp<float> traj = OUTS ? V.traj : gp<float>(nullptr);
    const int64_t tstride = V.traj_tstride;
    const int bper = OUTS ? V.bold_period : 0;

    for (int i = tid; i < NP * NHd; i += BS) ring[i] = 0.0f;   // per-edge: exact-zero fill
    // every zero is stored before any lane writes a pre value into the ring:
    // without this wait a fill store of another wave can land after (and
    // erase) a node's step-0 entry
    __syncthreads();
    // a continued run: the carried pre() of times -L..-1 go to slots
    // Hd-L..Hd-1 (>= K + 1, so no mirrored copy)
    if (V.hist_in) {
        const int L = V.hist_len, total = L * N * NP;
        for (int i = tid; i < total; i += BS) {
            const int j = i / (N * NP), r = i - j * (N * NP), nn = r / NP, c = r - nn * NP;
            ring[ridx(c, nn * Hs + (Hd - L + j))] = V.hist_in[i];
        }
    }

    // lanes per node: lane `sub` of node n gathers row edges sub, sub+lpn, ...
    const int lpn = V.lpn, sub = tid & (lpn - 1);
    const int n = tid >> V.lpn_shift;
    const bool own = n < N;
    const bool lead = own && sub == 0;   // writes the node's outputs and ring entries
    float mc[kMC];
    {
        const float* msrc = (V.node_param_idx >= 0 && own) ? (const float*)(V.node_mc + (size_t)n * kMC) : Ia->mc;
#pragma unroll
        for (int i = 0; i < kMC; ++i) mc[i] = msrc[i];
    }
    float x[NV], tv[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { x[v] = own ? V.x0[n * NV + v] : 0.0f; tv[v] = 0.0f; }
    Bold bw;
    bw.s = 0.0f; bw.f = 1.0f; bw.v = 1.0f; bw.q = 1.0f; bw.acc = 0.0f; bw.prev = 0.0f;
    if (V.bold_carry && own) {
        bw.s = V.bold_io[n * 4 + 0]; bw.f = V.bold_io[n * 4 + 1];
        bw.v = V.bold_io[n * 4 + 2]; bw.q = V.bold_io[n * 4 + 3];
    }
    BoldClock bc;
    bc.k = 0; bc.j = 0; bc.tprev = 0.0;
    int bcount = 0;

    // CSR row in registers: source base j*Hd, delay, weight
    // register-resident row edges: fewer with an HBM ring, whose deeper gather
    // pipeline needs the registers (the rest stream through the batched tail)
    // (1024-thread launches have 128 registers per lane: 8 register edges)
    constexpr int RE = WIDE ? 8 : (LRING ? kMaxRegEdges : kMaxRegEdges / 2);
    const int kreg = min((V.max_row + lpn - 1) >> V.lpn_shift, RE);
    int e0 = 0, e1 = 0;
    if (own) { e0 = V.indptr[n]; e1 = V.indptr[n + 1]; }
    uint32_t epk[RE];   // src | delay << 16
    float ewt[RE];
#pragma unroll
    for (int k = 0; k < RE; ++k) {
        epk[k] = own ? (uint32_t)n : 0u; ewt[k] = 0.0f;   // padding: zero weight, own row
        if (k < kreg && own && e0 + sub + k * lpn < e1) {
            const uint2 ed = edge_at(V.edges, e0 + sub + k * lpn);
            epk[k] = ed.x;
            ewt[k] = __uint_as_float(ed.y);
        }
    }
    const int em0 = e0 + sub + kreg * lpn;   // memory tail of long rows (strided)

    // own pre(x_t), kept for Kuramoto's target term
    float own_pre[NP];
    auto write_pre = [&](int slot) {
        float xc[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) xc[c] = x[c];
        pre<NC, NP>(cfn, cc, cpu_form, classic, xc, own_pre);
        if (lead) {
#pragma unroll
            for (int c = 0; c < NP; ++c) {
                ring[ridx(c, n * Hs + slot)] = own_pre[c];
                if (slot < K - 1) ring[ridx(c, n * Hs + Hd + slot)] = own_pre[c];   // mirror
            }
        }
    };
    auto finish_step = [&](int t, const float* g, const float* zraw) {
        float cpl[NC], z[NV];
        post<NC, NP>(cfn, cc, classic, g, own_pre, cpl);
#pragma unroll
        for (int v = 0; v < NV; ++v) z[v] = zraw[v] * nsc[v];
        integrate_c<M, IC>(S, mc, cpl, z, x);
#pragma unroll
        for (int v = 0; v < NV; ++v) tv[v] = tv[v] + x[v];
        if (traj && lead) {
            gp<float> tr = traj + (size_t)t * tstride + (size_t)n * NV;
#pragma unroll
            for (int v = 0; v < NV; ++v) tr[v] = x[v];
        }
        if (bper > 0) {
            bw.acc = bw.acc + x[0];
            if (++bcount >= bper) {
                const double dtb = V.dt_ms * (double)bcount;
                const float dts = (float)(dtb / 1000.0);
                const double tk = __dadd_rn(__dmul_rn((double)bc.k, dtb), dtb);
                const float sig = bold_step(bw, bw.acc / (float)bcount, dts);
                if (lead) bold_emit(V, bc, tk, sig, bw.prev, V.bold + n, N);
                bw.prev = sig;
                bw.acc = 0.0f;
                bold_advance(V, bc, tk);
                bcount = 0;
            }
        }
    };

    // ---- step 0 of a fresh run: every edge reads the current state
    // (runtime.rs:162-164); a continued run has no such step
    const bool fresh = V.step0 == 0;
    write_pre(0);
    __syncthreads();
    if (fresh && T > 0) {
        float g[NP];
#pragma unroll
        for (int c = 0; c < NP; ++c) g[c] = 0.0f;
        for (int e = e0 + sub; e < e1; e += lpn) {
            const uint2 ed = edge_at(V.edges, e);
            const int j = ed.x & 0xFFFF;
            const float w = __uint_as_float(ed.y);
#pragma unroll
            for (int c = 0; c < NP; ++c) g[c] += w * ring[ridx(c, j * Hs)];
        }
#pragma unroll
        for (int c = 0; c < NP; ++c) g[c] = lane_group_sum(g[c], lpn);
        float z0[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) z0[v] = (S.stoch && own) ? noise[(size_t)n * NV + v] : 0.0f;
        finish_step(0, g, z0);
        write_pre(1 % Hd);
    }

    // ---- blocks of K steps starting at t0 = 1 (a continued run: t0 = 0)
    const int tb = fresh ? 1 : 0;
    int ws0 = tb % Hd;   // ring slot of time t0
    // The block's noise (the long block serves deterministic runs only: no
    // noise registers). With up to two state variables the next block's
    // noise is loaded while the current block steps (one block of HBM
    // latency hidden); otherwise it is issued at the block's start, before
    // the gather, so that its latency overlaps the gather only.
    constexpr int KZ = K == kBlockSteps ? K : 1;
    constexpr bool ZAHEAD = KZ == K && NV <= 2;
    float zn[ZAHEAD ? KZ : 1][NV];
    auto load_noise = [&](float (&z)[ZAHEAD ? KZ : 1][NV], int ta) {
        const int ka = min(K, T - ta);
#pragma unroll
        for (int k = 0; k < (ZAHEAD ? KZ : 1); ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v)
                z[k][v] = (S.stoch && own && k < ka) ? noise[((size_t)(ta + k) * N + n) * NV + v] : 0.0f;
    };
    if constexpr (ZAHEAD) load_noise(zn, tb);
    for (int t0 = tb; t0 < T; t0 += K) {
        __syncthreads();
        const int kb = min(K, T - t0);
        float zb[KZ][NV];
        if constexpr (ZAHEAD) {
#pragma unroll
            for (int k = 0; k < KZ; ++k)
#pragma unroll
                for (int v = 0; v < NV; ++v) zb[k][v] = zn[k][v];
            if (t0 + K < T) load_noise(zn, t0 + K);
        } else {
#pragma unroll
            for (int k = 0; k < KZ; ++k)
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    zb[k][v] = (KZ == K && S.stoch && own && k < kb) ? noise[((size_t)(t0 + k) * N + n) * NV + v] : 0.0f;
        }
        float acc[K][NP];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < NP; ++c) acc[k][c] = 0.0f;
        // each edge's K window is contiguous (mirrored tail). Software pipeline:
        // the loads of edge e+1 are issued before edge e is accumulated, so two
        // edges' loads are in flight; accumulation stays in row order per step.
        auto edge_base = [&](int e) {
            int sl = ws0 - (int)(epk[e] >> 16);
            sl += sl < 0 ? Hd : 0;
            return (int)(epk[e] & 0xFFFF) * Hs + sl;
        };
        // D edge windows in flight: 2 from LDS; 4 from an HBM ring when they fit
        // the register budget (HBM latency is what the gather waits on there)
        constexpr int D = WIDE ? (K * NP <= 16 ? 2 : 1) : ((!LRING && K * NP <= 16) ? 4 : (K * NP * NV <= 64 ? 2 : 1));
        float vbuf[D][K][NP];
        auto load_win = [&](float (&vb)[K][NP], int base) {
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int c = 0; c < NP; ++c) vb[k][c] = ring[ridx(c, base + k)];
        };
#pragma unroll
        for (int p = 0; p + 1 < D; ++p)
            if (p < kreg) load_win(vbuf[p], edge_base(p));
#pragma unroll
        for (int e = 0; e < RE; ++e) {
            if (e < kreg) {
                const int en = e + D - 1;
                if (en < RE && en < kreg) load_win(vbuf[en % D], edge_base(en));
                __builtin_amdgcn_sched_barrier(0);
                const float w = ewt[e];
#pragma unroll
                for (int k = 0; k < K; ++k)
#pragma unroll
                    for (int c = 0; c < NP; ++c) acc[k][c] += w * vbuf[e % D][k][c];
            }
        }
        // rows longer than the register budget: D edges per batch (records, then
        // windows, then the sums in row order)
        for (int e = em0; e < e1; e += D * lpn) {
            float w[D];
            int base[D];
#pragma unroll
            for (int b = 0; b < D; ++b) {
                const int eb = e + b * lpn;
                w[b] = 0.0f;
                base[b] = n * Hs;   // padding: zero weight, own row
                if (eb < e1) {
                    const uint2 ed = edge_at(V.edges, eb);
                    int sl = ws0 - (int)(ed.x >> 16);
                    sl += sl < 0 ? Hd : 0;
                    base[b] = (int)(ed.x & 0xFFFF) * Hs + sl;
                    w[b] = __uint_as_float(ed.y);
                }
            }
#pragma unroll
            for (int b = 0; b < D; ++b) load_win(vbuf[b], base[b]);
#pragma unroll
            for (int b = 0; b < D; ++b)
#pragma unroll
                for (int k = 0; k < K; ++k)
#pragma unroll
                    for (int c = 0; c < NP; ++c) acc[k][c] += w[b] * vbuf[b][k][c];
        }
        if (lpn > 1) {
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int c = 0; c < NP; ++c) acc[k][c] = lane_group_sum(acc[k][c], lpn);
        }
        if constexpr (K == kBlockSteps) {
            // 8-step blocks: unrolled (static register slots, no loop or shifts)
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (k < kb) {
                    finish_step(t0 + k, acc[k], zb[k]);
                    int sl = ws0 + k + 1;
                    sl -= sl >= Hd ? Hd : 0;
                    write_pre(sl);
                }
            }
        } else {
            // long blocks: rolled (one integrator copy in the code); slot 0 of
            // the block registers is the current step, shifted down after it
#pragma nounroll
            for (int k = 0; k < kb; ++k) {
                finish_step(t0 + k, acc[0], zb[0]);
                int sl = ws0 + k + 1;
                sl -= sl >= Hd ? Hd : 0;
                write_pre(sl);
#pragma unroll
                for (int q = 0; q + 1 < K; ++q) {
#pragma unroll
                    for (int c = 0; c < NP; ++c) acc[q][c] = acc[q + 1][c];
                }
            }
        }
        ws0 += K;
        ws0 -= ws0 >= Hd ? Hd : 0;
    }

    // ---- the carried history: pre() of times T-L..T-1, oldest first
    if (V.hist_out) {
        __syncthreads();
        const int L = V.hist_len, total = L * N * NP;
        const int wsT = T % Hd;
        for (int i = tid; i < total; i += BS) {
            const int j = i / (N * NP), r = i - j * (N * NP), nn = r / NP, c = r - nn * NP;
            int sl = wsT - L + j;
            sl += sl < 0 ? Hd : 0;
            V.hist_out[i] = ring[ridx(c, nn * Hs + sl)];
        }
    }

    // ---- BOLD flush and TR tail rows
    if (bper > 0) {
        if (bcount > 0) {
            const double dtb = V.dt_ms * (double)bcount;
            const float dts = (float)(dtb / 1000.0);
            const double tk = __dadd_rn(__dmul_rn((double)bc.k, V.dtb_ms), dtb);
            const float sig = bold_step(bw, bw.acc / (float)bcount, dts);
            if (lead) bold_emit(V, bc, tk, sig, bw.prev, V.bold + n, N);
            bw.prev = sig;
            bold_advance(V, bc, tk);
        }
        if (!V.bold_native && bc.k > 0 && lead)
            for (int jj = bc.j; jj < V.bold_nout; ++jj) V.bold[(size_t)jj * N + n] = bw.prev;
        if (V.bold_io && lead) {
            V.bold_io[n * 4 + 0] = bw.s; V.bold_io[n * 4 + 1] = bw.f;
            V.bold_io[n * 4 + 2] = bw.v; V.bold_io[n * 4 + 3] = bw.q;
        }
    }
    if (lead) {
        const float inv = V.inv_nsteps;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            V.fin[n * NV + v] = x[v];
            V.tavg[n * NV + v] = tv[v] * inv;
        }
    }
    __syncthreads();
}

template <int M, int NP, bool LRING, int K, int CF, bool OUTS, bool WIDE>
__device__ __forceinline__ void run_blocked_ic(const DevInst* __restrict__ I, float* lds, int ic) {
    if (ic == 0) run_blocked<M, NP, LRING, K, 0, CF, OUTS, WIDE>(I, lds);
    else if (ic == 1) run_blocked<M, NP, LRING, K, 1, CF, OUTS, WIDE>(I, lds);
    else run_blocked<M, NP, LRING, K, 2, CF, OUTS, WIDE>(I, lds);
}

// compile-time variants: Linear with and without outputs, legacy sigmoidal
// (Jansen-Rit) without outputs; anything else generic
template <int M, int NP, bool LRING, int K, bool WIDE = false>
__device__ __forceinline__ void run_blocked_any(const DevInst* __restrict__ I, float* lds) {
    const int ic = __builtin_amdgcn_readfirstlane(integ_class(I->integrator));
    const int cfn = __builtin_amdgcn_readfirstlane(I->cfn);
    const bool outs = I->traj != nullptr || __builtin_amdgcn_readfirstlane(I->bold_period) > 0;
    if (cfn == HB_CF_LINEAR && !outs) run_blocked_ic<M, NP, LRING, K, HB_CF_LINEAR, false, WIDE>(I, lds, ic);
    else if (cfn == HB_CF_LINEAR) run_blocked_ic<M, NP, LRING, K, HB_CF_LINEAR, true, WIDE>(I, lds, ic);
    else if (cfn == HB_CF_SJR && !outs) run_blocked_ic<M, NP, LRING, K, HB_CF_SJR, false, WIDE>(I, lds, ic);
    else run_blocked_ic<M, NP, LRING, K, -1, true, WIDE>(I, lds, ic);
}

template <int M, int NP, bool LRING, bool MULTI>
__device__ __forceinline__ void run_any(const DevInst* __restrict__ I, float* lds) {
    if constexpr (!MULTI) {
        const int bk = __builtin_amdgcn_readfirstlane(I->blocked_k);
        if (bk == kBlockSteps) {
            run_blocked_any<M, NP, LRING, kBlockSteps>(I, lds);
            return;
        }
        if constexpr (!LRING) {
            if (bk == kLongBlockSteps) {   // deterministic runs with the ring in HBM
                run_blocked_any<M, NP, LRING, kLongBlockSteps>(I, lds);
                return;
            }
            // 128-byte windows (one ring channel)
            if constexpr (NP == 1) {
                if (bk == kLongerBlockSteps) {
                    run_blocked_any<M, NP, LRING, kLongerBlockSteps>(I, lds);
                    return;
                }
            }
        }
    }
    run_instance<M, NP, LRING, MULTI>(I, lds);
}

// The descriptor address is the same in every lane; saying so lets the
// compiler read descriptor fields with scalar loads kept in SGPRs.
__device__ __forceinline__ const DevInst* uniform_inst(const DevInst* p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (const DevInst*)(((uint64_t)hi << 32) | lo);
}

// Work distribution: non-persistent (one workgroup per work item, hardware
// dispatch in cost order) or persistent (workgroups pull items from counter).
__device__ __forceinline__ int next_work(const LaunchDesc& a, int iter) {
    if (!a.counter) return iter == 0 ? (int)blockIdx.x : a.n_work;
    __shared__ int s_w;
    if (threadIdx.x == 0) s_w = atomicAdd(a.counter, 1);
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(s_w);
    __syncthreads();
    return w;
}

template <int M, int NP, bool LRING, bool MULTI>
__global__ __launch_bounds__(512) void k_single(LaunchDesc a) {
    extern __shared__ float lds[];
    for (int it = 0;; ++it) {
        const int w = next_work(a, it);
        if (w >= a.n_work) break;
        run_any<M, NP, LRING, MULTI>(uniform_inst(a.insts + a.order[w]), lds);
    }
}

// One node (times lanes per node) per thread for networks of up to 1024
// threads: the delay-blocked schedule of networks wider than 512 nodes (and
// 2 lanes per node at 400 nodes) needs every node of the instance in one
// workgroup. The launch bound caps registers at 128 per lane, so the gather
// keeps 8 row edges in registers and 1-2 edge windows in flight.
// The planner sends only delay-blocked instances with 8- or 16-step blocks
// here (it checks the block length of every wide launch).
template <int M, int NP, bool LRING>
__global__ __launch_bounds__(1024) void k_wide(LaunchDesc a) {
    extern __shared__ float lds[];
    for (int it = 0;; ++it) {
        const int w = next_work(a, it);
        if (w >= a.n_work) break;
        const DevInst* I = uniform_inst(a.insts + a.order[w]);
        const int bk = __builtin_amdgcn_readfirstlane(I->blocked_k);
        if (bk == kBlockSteps) run_blocked_any<M, NP, LRING, kBlockSteps, true>(I, lds);
        else if constexpr (!LRING) {
            if (bk == kLongBlockSteps) run_blocked_any<M, NP, LRING, kLongBlockSteps, true>(I, lds);
        }
    }
}

template <int M, int NP>
static const void* kptr_single(bool lring, bool multi, bool wide) {
    if (wide) return lring ? (const void*)k_wide<M, NP, true> : (const void*)k_wide<M, NP, false>;
    if (lring) return multi ? (const void*)k_single<M, NP, true, true> : (const void*)k_single<M, NP, true, false>;
    return multi ? (const void*)k_single<M, NP, false, true> : (const void*)k_single<M, NP, false, false>;
}

// This translation unit's kernels: HB_UNIT 0..6 = one (model, pre channels)
// pair.
#ifndef HB_UNIT
#define HB_UNIT 0
#endif
// MPR has a second instantiation with one pre channel: when its second
// coupling coefficient (cv) is zero the V channel only ever adds c1 * 0.
static const void* unit_kernel(bool lring, bool multi, int np, bool wide) {
    (void)np;
    if (wide && multi) return nullptr;   // wide launches hold one node (group) per thread
#if HB_UNIT == 0
    return kptr_single<HB_G2DO, 1>(lring, multi, wide);
#elif HB_UNIT == 1
    return np == 1 ? kptr_single<HB_MPR, 1>(lring, multi, wide) : kptr_single<HB_MPR, 2>(lring, multi, wide);
#elif HB_UNIT == 2
    return kptr_single<HB_JANSEN_RIT, 1>(lring, multi, wide);
#elif HB_UNIT == 3
    return kptr_single<HB_WILSON_COWAN, 1>(lring, multi, wide);
#elif HB_UNIT == 4
    return kptr_single<HB_KURAMOTO, 1>(lring, multi, wide);
#elif HB_UNIT == 5
    return kptr_single<HB_KURAMOTO, 2>(lring, multi, wide);
#else
    return kptr_single<HB_RWW, 1>(lring, multi, wide);
#endif
}

#if HB_UNIT == 0
// ------------------------------------------- batch-averaged BOLD monitor -
//
// The batch engine's BOLD monitor (engine.rs:785-801) is fed, every step,
// var0 averaged over the modes and over the sweep points, and it runs one
// Balloon-Windkessel model per node (bold_monitor.rs:72-107, 207-227). From
// the plan's device trajectories: the mean over the instances per (step,
// node), summed in instance order and divided by their count (mean_dim), then
// one thread per node accumulating windows of `period` steps.
__global__ void k_bold_mean_input(const float* const* __restrict__ trajs, int B, int T, int N, int nv,
                                  float* __restrict__ zm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)T * N) return;
    float s = 0.0f;
    for (int b = 0; b < B; ++b) s += trajs[b][i * nv];
    zm[i] = s / (float)B;
}

__global__ void k_bold_mean_monitor(const float* __restrict__ zm, int T, int N, int period, double dt_ms,
                                    float* __restrict__ bw, float* __restrict__ native) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    Bold b;
    b.s = bw[n * 4 + 0]; b.f = bw[n * 4 + 1]; b.v = bw[n * 4 + 2]; b.q = bw[n * 4 + 3];
    b.acc = 0.0f; b.prev = 0.0f;
    int cnt = 0, k = 0;
    for (int t = 0; t < T; ++t) {
        b.acc = b.acc + zm[(size_t)t * N + n];
        if (++cnt >= period) {
            const float dts = (float)(dt_ms * (double)cnt / 1000.0);
            native[(size_t)k * N + n] = bold_step(b, b.acc / (float)cnt, dts);
            ++k;
            b.acc = 0.0f;
            cnt = 0;
        }
    }
    if (cnt > 0) {   // BoldMonitor::flush of the partial window
        const float dts = (float)(dt_ms * (double)cnt / 1000.0);
        native[(size_t)k * N + n] = bold_step(b, b.acc / (float)cnt, dts);
    }
    bw[n * 4 + 0] = b.s; bw[n * 4 + 1] = b.f; bw[n * 4 + 2] = b.v; bw[n * 4 + 3] = b.q;
}

static hipError_t bold_mean(const float* const* trajs, int n_inst, int T, int N, int nv, float* zm, int period,
                            double dt_ms, float* bw, float* native, hipStream_t s) {
    const int64_t tn = (int64_t)T * N;
    if (tn > 0)
        hipLaunchKernelGGL(k_bold_mean_input, dim3((unsigned)((tn + 255) / 256)), dim3(256), 0, s, trajs, n_inst, T,
                           N, nv, zm);
    hipLaunchKernelGGL(k_bold_mean_monitor, dim3((unsigned)((N + 63) / 64)), dim3(64), 0, s, zm, T, N, period, dt_ms,
                       bw, native);
    return hipGetLastError();
}
#endif

// op 0: launch; 1: occupancy (blocks per CU into *out); 2: set the maximum
// dynamic LDS (arg bytes) on every kernel of the unit.
static hipError_t unit_op(int op, const LaunchDesc* d, hipStream_t s, int arg, int* out) {
    if (op == 2) {
        for (int lr = 0; lr < 2; ++lr)
            for (int mu = 0; mu < 2; ++mu)
                for (int np = 1; np <= 2; ++np)
                    for (int wd = 0; wd < 2; ++wd) {
                        const void* k = unit_kernel(lr != 0, mu != 0, np, wd != 0);
                        if (!k) continue;
                        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, arg);
                        if (e != hipSuccess) return e;
                    }
        return hipSuccess;
    }
    if (d->block > kWideBlockMax || (d->multi && d->block > kBlockMax)) return hipErrorInvalidValue;
    const void* k = unit_kernel(d->lds_ring, d->multi, d->np, d->block > kBlockMax);
    if (!k) return hipErrorInvalidValue;
    if (op == 1) return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k, d->block, (size_t)d->lds_bytes);
    LaunchDesc a = *d;
    void* args[] = {&a};
    return hipLaunchKernel(k, dim3(d->grid), dim3(d->block), args, (size_t)d->lds_bytes, s);
}

}  // namespace HB_NS

#define HB_CAT3(a, b, c) a##b##c
#define HB_XCAT3(a, b, c) HB_CAT3(a, b, c)
#if HB_UNIT == 0
#if HB_EXACT
hipError_t launch_bold_mean_exact(const float* const* trajs, int n_inst, int T, int N, int nv, float* zm, int period,
                                  double dt_ms, float* bw, float* native, hipStream_t s) {
    return exact::bold_mean(trajs, n_inst, T, N, nv, zm, period, dt_ms, bw, native, s);
}
#else
hipError_t launch_bold_mean_fast(const float* const* trajs, int n_inst, int T, int N, int nv, float* zm, int period,
                                 double dt_ms, float* bw, float* native, hipStream_t s) {
    return fast::bold_mean(trajs, n_inst, T, N, nv, zm, period, dt_ms, bw, native, s);
}
#endif
#endif
#if HB_EXACT
hipError_t HB_XCAT3(fwd_unit_exact_, HB_UNIT, )(int op, const LaunchDesc* d, hipStream_t s, int arg, int* out) {
    return exact::unit_op(op, d, s, arg, out);
}
#else
hipError_t HB_XCAT3(fwd_unit_fast_, HB_UNIT, )(int op, const LaunchDesc* d, hipStream_t s, int arg, int* out) {
    return fast::unit_op(op, d, s, arg, out);
}
#endif

}  // namespace hb
