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

template <int M> struct Tr;
template <> struct Tr<HB_G2DO> { static constexpr int NV = 2, NC = 1; };
template <> struct Tr<HB_MPR> { static constexpr int NV = 2, NC = 2; };
template <> struct Tr<HB_JANSEN_RIT> { static constexpr int NV = 6, NC = 1; };
template <> struct Tr<HB_WILSON_COWAN> { static constexpr int NV = 2, NC = 1; };
template <> struct Tr<HB_KURAMOTO> { static constexpr int NV = 1, NC = 1; };
template <> struct Tr<HB_RWW> { static constexpr int NV = 1, NC = 1; };

// ---------------------------------------------------------------- models ---

template <int M>
__device__ __forceinline__ void dfun(const float* mc, bool sw, float sv, const float* x,
                                     const float* c, float* dx);

template <> __device__ __forceinline__ void dfun<HB_G2DO>(const float* mc, bool sw, float sv,
                                                          const float* x, const float* c, float* dx) {
    // dfun.rs:295-318
    const float v = x[0], w = x[1];
    const float iext = sw ? (c[0] + sv) * mc[8] : c[0] * mc[8] + mc[9];
    const float v2 = v * v;
    const float v3 = v * v2;
    dx[0] = ((((w * mc[6] + iext) - v3 * mc[4]) + v2 * mc[3]) + v * mc[5]) * mc[10];
    dx[1] = (((v * mc[1] + v2 * mc[2]) - w * mc[7]) + mc[0]) * mc[11];
}

template <> __device__ __forceinline__ void dfun<HB_MPR>(const float* mc, bool sw, float sv,
                                                         const float* x, const float* c, float* dx) {
    // dfun.rs:455-477
    const float r = x[0], v = x[1];
    dx[0] = ((r * v) * 2.0f + mc[1]) * mc[0];
    const float v2 = v * v;
    const float r2 = r * r;
    float dv;
    if (sw) dv = (((((v2 - r2 * mc[2]) + mc[5]) + sv) + r * mc[3]) + c[0] * mc[6]) + c[1] * mc[7];
    else dv = ((((v2 - r2 * mc[2]) + mc[4]) + r * mc[3]) + c[0] * mc[6]) + c[1] * mc[7];
    dx[1] = dv * mc[0];
}

__device__ __forceinline__ float jr_sigm(const float* mc, float v) {
    // dfun.rs:350-354: 2*nu_max / (exp((v0 - v) * r) + 1)
    return mc[2] / (expf((-v + mc[0]) * mc[1]) + 1.0f);
}

template <> __device__ __forceinline__ void dfun<HB_JANSEN_RIT>(const float* mc, bool sw, float sv,
                                                               const float* x, const float* c, float* dx) {
    // dfun.rs:356-383
    const float s12 = jr_sigm(mc, x[1] - x[2]);
    const float s01 = jr_sigm(mc, x[0] * mc[6]);
    const float s03 = jr_sigm(mc, x[0] * mc[8]);
    const float mu = sw ? sv : mc[13];
    const float mut = (s01 * mc[7] + mu) + c[0];
    dx[0] = x[3]; dx[1] = x[4]; dx[2] = x[5];
    dx[3] = (s12 * mc[3] - x[3] * mc[4]) - x[0] * mc[5];
    dx[4] = (mut * mc[3] - x[4] * mc[4]) - x[1] * mc[5];
    dx[5] = ((s03 * mc[9]) * mc[10] - x[5] * mc[11]) - x[2] * mc[12];
}

template <> __device__ __forceinline__ void dfun<HB_WILSON_COWAN>(const float* mc, bool sw, float sv,
                                                                 const float* x, const float* c, float* dx) {
    // dfun.rs:405-436
    const float e = x[0], i = x[1];
    float xe = ((e * mc[0] - i * mc[1]) - mc[9]) + c[0];
    xe = (xe + (sw ? sv : mc[18])) * mc[20];
    const float xi = (((e * mc[2] - i * mc[3]) + mc[19]) - mc[13]) * mc[21];
    const float ee = expf((xe + (-mc[7])) * (-mc[6]));
    const float se = (1.0f / (ee + 1.0f) - mc[22]) * mc[8];
    const float ei = expf((xi + (-mc[11])) * (-mc[10]));
    const float si = (1.0f / (ei + 1.0f) - mc[23]) * mc[12];
    dx[0] = (-e + (e * (-mc[14]) + mc[16]) * se) * mc[4];
    dx[1] = (-i + (i * (-mc[15]) + mc[17]) * si) * mc[5];
}

template <> __device__ __forceinline__ void dfun<HB_KURAMOTO>(const float* mc, bool sw, float sv,
                                                              const float* x, const float* c, float* dx) {
    // dfun.rs:486-491
    (void)x;
    dx[0] = c[0] + (sw ? sv : mc[0]);
}

template <> __device__ __forceinline__ void dfun<HB_RWW>(const float* mc, bool sw, float sv,
                                                         const float* x, const float* c, float* dx) {
    // dfun.rs:503-515
    const float s = x[0];
    const float xx = sw ? (s * mc[5] + sv) + c[0] * mc[6] : (s * mc[5] + mc[7]) + c[0] * mc[6];
    const float axb = xx * mc[0] + (-mc[1]);
    const float et = expf(axb * (-mc[2]));
    const float den = -et + 1.0f;
    const float h = axb / den;
    dx[0] = (-s) * mc[4] + ((-s + 1.0f) * h) * mc[3];
}

// clamp_batch, dfun.rs:170-239. Ternaries keep NaN like f32::clamp.
template <int M> __device__ __forceinline__ void clampm(float* x) {
    if constexpr (M == HB_WILSON_COWAN) {
        x[0] = x[0] < 0.0f ? 0.0f : (x[0] > 1.0f ? 1.0f : x[0]);
        x[1] = x[1] < 0.0f ? 0.0f : (x[1] > 1.0f ? 1.0f : x[1]);
    } else if constexpr (M == HB_MPR) {
        x[0] = x[0] < 0.0f ? 0.0f : x[0];
    } else if constexpr (M == HB_RWW) {
        x[0] = x[0] < 0.0f ? 0.0f : (x[0] > 1.0f ? 1.0f : x[0]);
    } else if constexpr (M == HB_KURAMOTO) {
        const float tp = 2.0f * kPi;
        x[0] = x[0] - floorf(x[0] / tp) * tp;
    }
}

// ---------------------------------------------------------------- coupling -

// pre(): coupling.rs:204-258 (tensor form) / 268-321 (per-edge CPU form).
template <int NC, int NP>
__device__ __forceinline__ void pre(int cfn, const float* cc, bool cpu, bool classic,
                                    const float* x, float* out) {
    if (cfn == HB_CF_KURAMOTO) {
        if constexpr (NP == 2 * NC) {
#pragma unroll
            for (int c = 0; c < NC; ++c) { out[c] = sinf(x[c]); out[NC + c] = cosf(x[c]); }
        }
    } else if (cfn == HB_CF_TANH) {
#pragma unroll
        for (int c = 0; c < NP; ++c) out[c] = (tanhf((x[c] * cc[1] + (-cc[2])) / cc[3]) + 1.0f) * cc[0];
    } else if (cfn == HB_CF_SJR) {
        float diff = x[0];
        if constexpr (NC >= 2) diff = x[0] - x[1];
        else diff = x[0] - 0.0f;
        const float den = expf((diff + (-cc[4])) * (-cc[3])) + 1.0f;
        if (classic) out[0] = cpu ? cc[1] + cc[2] / den : (1.0f / den) * cc[2] + cc[1];
        else out[0] = cpu ? cc[1] / den : (1.0f / den) * cc[1];
    } else if (cfn == HB_CF_PRESIGMOIDAL) {
#pragma unroll
        for (int c = 0; c < NP; ++c) out[c] = (tanhf((x[c] * cc[3] + (-cc[4])) * cc[2]) + cc[1]) * cc[0];
    } else {
#pragma unroll
        for (int c = 0; c < NP; ++c) out[c] = x[c];
    }
}

// post_with_target: coupling.rs:323-399. sc = (sin, cos) of the target's own
// current state for Kuramoto (the ring already holds exactly those values).
template <int NC, int NP>
__device__ __forceinline__ void post(int cfn, const float* cc, bool classic, const float* g,
                                     const float* own_pre, float* out) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        float r;
        const float gc = g[c < NP ? c : 0];
        switch (cfn) {
        case HB_CF_LINEAR: r = gc * cc[0] + cc[1]; break;
        case HB_CF_SCALED_LINEAR: r = (gc + (-cc[1])) * cc[0]; break;
        case HB_CF_DIFFERENCE: r = gc * cc[0]; break;
        case HB_CF_KURAMOTO:
            if constexpr (NP == 2 * NC) r = (own_pre[NC + c] * g[c] - own_pre[c] * g[NC + c]) * cc[0];
            else r = gc;
            break;
        case HB_CF_SIGMOIDAL: {
            const float sh = (gc + (-cc[2])) / cc[4];
            const float den = expf(-(sh * cc[3])) + 1.0f;
            r = (1.0f / den) * cc[1] + cc[0];
            break; }
        case HB_CF_SJR: r = classic ? gc * cc[0] : gc; break;
        default: r = gc; break;
        }
        out[c] = r;
    }
}

// ---------------------------------------------------------------- BOLD -----

struct Bold {
    float s, f, v, q, acc, prev;
};

// One forward-Euler Balloon-Windkessel step + signal, bold.rs:106-166.
__device__ __forceinline__ float bold_step(Bold& b, float z_raw, float dt_s) {
    const float kappa = 1.0f / 1.54f, gamma = 1.0f / 0.50f, tau_inv = 1.0f / 1.02f;
    const float ialpha = 1.0f / 0.32f, e0 = 0.34f;
    const float z = z_raw > 0.0f ? z_raw : 0.0f;
    const float fs = b.f > 1e-6f ? b.f : 1e-6f;
    const float vs = b.v > 1e-6f ? b.v : 1e-6f;
    const float ds = (z - kappa * b.s) - gamma * (fs - 1.0f);
    const float df = b.s;
    const float pv = powf(vs, ialpha);
    const float dv = tau_inv * (fs - pv);
    const float dq = tau_inv * ((fs * (1.0f - powf(1.0f - e0, 1.0f / fs))) / e0 - pv * (b.q / vs));
    const float sn = b.s + dt_s * ds, fn = fs + dt_s * df, vn = vs + dt_s * dv, qn = b.q + dt_s * dq;
    b.s = sn > 0.0f ? sn : 0.0f;
    b.f = fn > 1e-6f ? fn : 1e-6f;
    b.v = vn > 1e-6f ? vn : 1e-6f;
    b.q = qn > 0.0f ? qn : 0.0f;
    const float vv = b.v > 1e-6f ? b.v : 1e-6f;
    return 4.0f * ((7.0f * (1.0f - b.q) + 2.0f * (1.0f - b.q / vv)) + 2.0f * (1.0f - vv));
}

// Emit TR-resampled outputs after native sample k (time tk, value sig) is
// produced; bold_monitor.rs:123-203 evaluated as a stream. Uniform control.
struct BoldClock {
    int k;        // native samples produced so far
    int j;        // next TR row
    double tprev; // time of the previous native sample
};

__device__ __forceinline__ void bold_emit(const DevInst* I, const BoldClock& bc, double tk, float sig,
                                          float prev, float* out_col, int N) {
    if (I->bold_native) {
        if (bc.k < I->bold_nout) out_col[(size_t)bc.k * N] = sig;
        return;
    }
    const double tr = I->tr_ms;
    int j = bc.j;
    if (bc.k == 0) {
        while (j < I->bold_nout && (double)(j + 1) * tr <= tk) { out_col[(size_t)j * N] = sig; ++j; }
    } else {
        while (j < I->bold_nout && (double)(j + 1) * tr < tk) {
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
__device__ __forceinline__ void bold_advance(const DevInst* I, BoldClock& bc, double tk) {
    if (!I->bold_native) {
        const double tr = I->tr_ms;
        if (bc.k == 0) { while (bc.j < I->bold_nout && (double)(bc.j + 1) * tr <= tk) bc.j++; }
        else { while (bc.j < I->bold_nout && (double)(bc.j + 1) * tr < tk) bc.j++; }
    }
    bc.k++;
    bc.tprev = tk;
}

// ---------------------------------------------------------------- instance -

// Integrate instance I with the whole workgroup. lds: dynamic LDS base.
template <int M, int NP, bool LRING, bool MULTI>
__device__ __forceinline__ void run_instance(const DevInst* __restrict__ I, float* lds) {
    constexpr int NV = Tr<M>::NV, NC = Tr<M>::NC;
    const int N = I->n_nodes, T = I->n_steps, Hd = I->ring_depth;
    const int tid = threadIdx.x, BS = blockDim.x;
    const int NNP = N * NP;
    float* __restrict__ ring = LRING ? lds : I->ring;
    // MULTI state table after the ring: x[N][NV], tavg[N][NV], bold[N][6]
    float* __restrict__ tab = LRING ? lds + (size_t)Hd * NNP : lds;
    float* __restrict__ tab_x = tab;
    float* __restrict__ tab_t = tab + (size_t)N * NV;
    float* __restrict__ tab_b = tab + (size_t)2 * N * NV;

    const int cfn = I->cfn, wkind = I->wkind, dmode = I->delay_mode, integ = I->integrator;
    const bool cpu_form = I->cpu_form != 0, classic = I->sjr_classic != 0;
    const bool sw = I->sweep_form != 0;
    const float sv = I->sweep_val;
    const float dt = I->dt, dth = I->dt_half, dt6 = I->dt_sixth;
    const bool stoch = integ >= HB_EULER_STOCH;
    const bool plain = I->engine == HB_ENGINE_PLAIN;
    float cc[kCC];
#pragma unroll
    for (int i = 0; i < kCC; ++i) cc[i] = I->cc[i];
    float nsc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) nsc[v] = I->nscale[v];
    const float fill[2] = {I->fill0, I->fill1};

    // ring initial fill: 0 (per-edge: exact-zero contributions) or pre(0)
    {
        const int total = Hd * NNP;
        for (int i = tid; i < total; i += BS) ring[i] = fill[NP == 2 ? (i & 1) : 0];
    }

    const int nper = MULTI ? (N + BS - 1) / BS : 1;
    // model constants (uniform, or per node when a parameter is given per node)
    float mc_u[kMC];
#pragma unroll
    for (int i = 0; i < kMC; ++i) mc_u[i] = I->mc[i];
    const int npi = I->node_param_idx;

    // per-thread node state (MULTI=false keeps it in registers)
    float x[NV], tv[NV];
    Bold bw;
    bw.s = 0.0f; bw.f = 1.0f; bw.v = 1.0f; bw.q = 1.0f; bw.acc = 0.0f; bw.prev = 0.0f;
    const bool own = tid < N;
    if (!MULTI) {
#pragma unroll
        for (int v = 0; v < NV; ++v) { x[v] = own ? I->x0[tid * NV + v] : 0.0f; tv[v] = 0.0f; }
    } else {
        for (int n = tid; n < N; n += BS) {
#pragma unroll
            for (int v = 0; v < NV; ++v) { tab_x[n * NV + v] = I->x0[n * NV + v]; tab_t[n * NV + v] = 0.0f; }
            float* b = tab_b + n * 6;
            b[0] = 0.0f; b[1] = 1.0f; b[2] = 1.0f; b[3] = 1.0f; b[4] = 0.0f; b[5] = 0.0f;
        }
    }
    float mc_n[kMC];
    if (!MULTI) {
#pragma unroll
        for (int i = 0; i < kMC; ++i) mc_n[i] = (npi >= 0 && own) ? I->node_mc[tid * kMC + i] : mc_u[i];
    }
    const int bper = I->bold_period;
    BoldClock bc;
    bc.k = 0; bc.j = 0; bc.tprev = 0.0;
    int bcount = 0;
    __syncthreads();

    int ws = 0;
    for (int t = 0; t < T; ++t) {
        // ---- phase 1: ring[ws][n] = pre(x_t[n])
        for (int kk = 0; kk < nper; ++kk) {
            const int n = MULTI ? tid + kk * BS : tid;
            if (n < N) {
                float xc[NC], pr[NP];
#pragma unroll
                for (int c = 0; c < NC; ++c) xc[c] = MULTI ? tab_x[n * NV + c] : x[c];
                pre<NC, NP>(cfn, cc, cpu_form, classic, xc, pr);
                float* dst = ring + (size_t)ws * NNP + n * NP;
#pragma unroll
                for (int c = 0; c < NP; ++c) dst[c] = pr[c];
            }
        }
        __syncthreads();

        // uniform read slot for single-delay projections
        int slot_u = ws;
        if (dmode != HB_DELAY_PER_EDGE && t > 0) {
            const int d = I->single_delay;
            int teff_back = 0;   // how many steps back
            if (d == 0) teff_back = 0;
            else if (d > t) teff_back = -1;
            else if (dmode == HB_DELAY_BATCH_COMPAT) {
                const int H = Hd - 1;   // reference horizon = max_delay + 1
                const int back = H - 1 - ((t - d + H) % H);
                teff_back = back > t ? -1 : back;
            } else teff_back = d;
            if (teff_back < 0) slot_u = -1;
            else { slot_u = ws - teff_back; if (slot_u < 0) slot_u += Hd; }
        }
        const int dmask = t == 0 ? 0 : 0xFFFF;
        // mean-field sum (sequential over sources, identical in every thread)
        float mf[NP];
        if (wkind == HB_W_SCALAR_MEAN) {
#pragma unroll
            for (int c = 0; c < NP; ++c) {
                float s = 0.0f;
                if (slot_u >= 0) {
                    const float* src = ring + (size_t)slot_u * NNP + c;
                    for (int j = 0; j < N; ++j) s += src[j * NP];
                } else {
                    for (int j = 0; j < N; ++j) s += fill[c];
                }
                mf[c] = (s / (float)N) * I->scalar_w;
            }
        }

        // ---- phase 2: gather + integrate, per owned node
        for (int kk = 0; kk < nper; ++kk) {
            const int n = MULTI ? tid + kk * BS : tid;
            if (n < N) {
                float z[NV];
                if (stoch) {
                    const float* nz = I->noise + ((size_t)t * N + n) * NV;
#pragma unroll
                    for (int v = 0; v < NV; ++v) z[v] = nz[v] * nsc[v];
                }
                float g[NP];
#pragma unroll
                for (int c = 0; c < NP; ++c) g[c] = 0.0f;
                if (wkind == HB_W_CSR) {
                    const int e0 = I->indptr[n], e1 = I->indptr[n + 1];
                    if (dmode == HB_DELAY_PER_EDGE) {
                        const float* rbase = ring;
                        for (int e = e0; e < e1; ++e) {
                            const uint2 ed = I->edges[e];
                            const int j = ed.x & 0xFFFF;
                            const int d = (ed.x >> 16) & dmask;
                            const float w = __uint_as_float(ed.y);
                            int sl = ws - d;
                            sl += sl < 0 ? Hd : 0;
                            const float* src = rbase + (size_t)sl * NNP + j * NP;
#pragma unroll
                            for (int c = 0; c < NP; ++c) g[c] += w * src[c];
                        }
                    } else {
                        for (int e = e0; e < e1; ++e) {
                            const uint2 ed = I->edges[e];
                            const int j = ed.x & 0xFFFF;
                            const float w = __uint_as_float(ed.y);
#pragma unroll
                            for (int c = 0; c < NP; ++c)
                                g[c] += w * (slot_u >= 0 ? ring[(size_t)slot_u * NNP + j * NP + c] : fill[c]);
                        }
                    }
                } else if (wkind == HB_W_DENSE) {
                    const float* wt = I->dense_t;
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
                        g[c] = I->scalar_w * (slot_u >= 0 ? ring[(size_t)slot_u * NNP + n * NP + c] : fill[c]);
                }
                float cpl[NC];
                if (wkind != HB_W_NONE) {
                    float own_pre[NP];
                    const float* op = ring + (size_t)ws * NNP + n * NP;
#pragma unroll
                    for (int c = 0; c < NP; ++c) own_pre[c] = op[c];
                    post<NC, NP>(cfn, cc, classic, g, own_pre, cpl);
                } else {
#pragma unroll
                    for (int c = 0; c < NC; ++c) cpl[c] = 0.0f;
                }

                float xs[NV];
                float mcl[kMC];
                const float* mc;
                if (!MULTI) {
                    mc = mc_n;
#pragma unroll
                    for (int v = 0; v < NV; ++v) xs[v] = x[v];
                } else {
#pragma unroll
                    for (int v = 0; v < NV; ++v) xs[v] = tab_x[n * NV + v];
                    if (npi >= 0) {
#pragma unroll
                        for (int i = 0; i < kMC; ++i) mcl[i] = I->node_mc[n * kMC + i];
                        mc = mcl;
                    } else mc = mc_u;
                }

                // ---- integrator stages, engine.rs:698-773
                float k[NV], acc[NV], s[NV];
                dfun<M>(mc, sw, sv, xs, cpl, k);
                if (integ == HB_EULER || integ == HB_EULER_STOCH) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const float y = xs[v] + k[v] * dt;
                        xs[v] = stoch ? y + z[v] : y;
                    }
                    clampm<M>(xs);
                } else if (integ == HB_HEUN || integ == HB_HEUN_STOCH) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const float y = xs[v] + k[v] * dt;
                        s[v] = stoch ? y + z[v] : y;
                        acc[v] = k[v];
                    }
                    clampm<M>(s);
                    dfun<M>(mc, sw, sv, s, cpl, k);
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const float y = xs[v] + (acc[v] + k[v]) * dth;
                        xs[v] = stoch ? y + z[v] : y;
                    }
                    clampm<M>(xs);
                } else {  // RK4 / RK4 stochastic
#pragma unroll
                    for (int v = 0; v < NV; ++v) { acc[v] = k[v]; s[v] = xs[v] + k[v] * dth; }
                    clampm<M>(s);
                    dfun<M>(mc, sw, sv, s, cpl, k);
#pragma unroll
                    for (int v = 0; v < NV; ++v) { acc[v] = acc[v] + k[v] * 2.0f; s[v] = xs[v] + k[v] * dth; }
                    clampm<M>(s);
                    dfun<M>(mc, sw, sv, s, cpl, k);
#pragma unroll
                    for (int v = 0; v < NV; ++v) { acc[v] = acc[v] + k[v] * 2.0f; s[v] = xs[v] + k[v] * dt; }
                    clampm<M>(s);
                    dfun<M>(mc, sw, sv, s, cpl, k);
#pragma unroll
                    for (int v = 0; v < NV; ++v) xs[v] = xs[v] + (acc[v] + k[v]) * dt6;
                    if (stoch) {
                        if (plain) {   // integrator.rs:296-315: clamp(rk4) + Z, no final clamp
                            clampm<M>(xs);
#pragma unroll
                            for (int v = 0; v < NV; ++v) xs[v] = xs[v] + z[v];
                        } else {       // engine.rs:747-773: clamp(det + Z)
#pragma unroll
                            for (int v = 0; v < NV; ++v) xs[v] = xs[v] + z[v];
                            clampm<M>(xs);
                        }
                    } else {
                        clampm<M>(xs);
                    }
                }

                // ---- outputs
                if (!MULTI) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) { x[v] = xs[v]; tv[v] = tv[v] + xs[v]; }
                } else {
#pragma unroll
                    for (int v = 0; v < NV; ++v) { tab_x[n * NV + v] = xs[v]; tab_t[n * NV + v] = tab_t[n * NV + v] + xs[v]; }
                }
                if (I->traj) {
                    float* tr = I->traj + (size_t)t * I->traj_tstride + (size_t)n * NV;
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
                const double dtb = I->dt_ms * (double)bcount;
                const float dts = (float)(dtb / 1000.0);
                const double tk = (double)bc.k * dtb + dtb;
                for (int kk = 0; kk < nper; ++kk) {
                    const int n = MULTI ? tid + kk * BS : tid;
                    if (n < N) {
                        Bold b;
                        if (!MULTI) b = bw;
                        else { const float* q = tab_b + n * 6; b.s = q[0]; b.f = q[1]; b.v = q[2]; b.q = q[3]; b.acc = q[4]; b.prev = q[5]; }
                        const float sig = bold_step(b, b.acc / (float)bcount, dts);
                        bold_emit(I, bc, tk, sig, b.prev, I->bold + n, N);
                        b.prev = sig;
                        b.acc = 0.0f;
                        if (!MULTI) bw = b;
                        else { float* q = tab_b + n * 6; q[0] = b.s; q[1] = b.f; q[2] = b.v; q[3] = b.q; q[4] = b.acc; q[5] = b.prev; }
                    }
                }
                bold_advance(I, bc, tk);
                bcount = 0;
            }
        }
        ws = (ws + 1 == Hd) ? 0 : ws + 1;
    }

    // ---- BOLD flush of a partial window, then nearest-sample tail rows
    if (bper > 0) {
        double tlast = bc.tprev;
        if (bcount > 0) {
            const double dtb = I->dt_ms * (double)bcount;
            const float dts = (float)(dtb / 1000.0);
            const double tk = (double)bc.k * I->dtb_ms + dtb;
            for (int kk = 0; kk < nper; ++kk) {
                const int n = MULTI ? tid + kk * BS : tid;
                if (n < N) {
                    Bold b;
                    if (!MULTI) b = bw;
                    else { const float* q = tab_b + n * 6; b.s = q[0]; b.f = q[1]; b.v = q[2]; b.q = q[3]; b.acc = q[4]; b.prev = q[5]; }
                    const float sig = bold_step(b, b.acc / (float)bcount, dts);
                    bold_emit(I, bc, tk, sig, b.prev, I->bold + n, N);
                    b.prev = sig;
                    if (!MULTI) bw = b;
                    else tab_b[n * 6 + 5] = sig;
                }
            }
            bold_advance(I, bc, tk);
            tlast = tk;
        }
        (void)tlast;
        if (!I->bold_native && bc.k > 0) {
            for (int kk = 0; kk < nper; ++kk) {
                const int n = MULTI ? tid + kk * BS : tid;
                if (n < N) {
                    const float last = MULTI ? tab_b[n * 6 + 5] : bw.prev;
                    for (int jj = bc.j; jj < I->bold_nout; ++jj) I->bold[(size_t)jj * N + n] = last;
                }
            }
        }
    }

    // ---- tavg * (1/n_steps), final state (engine.rs:813-818)
    const float inv = I->inv_nsteps;
    for (int kk = 0; kk < nper; ++kk) {
        const int n = MULTI ? tid + kk * BS : tid;
        if (n < N) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const float xv = MULTI ? tab_x[n * NV + v] : x[v];
                const float tt = MULTI ? tab_t[n * NV + v] : tv[v];
                I->fin[n * NV + v] = xv;
                I->tavg[n * NV + v] = tt * inv;
            }
        }
    }
    __syncthreads();
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
        run_instance<M, NP, LRING, MULTI>(a.insts + a.order[w], lds);
    }
}

template <bool LRING, bool MULTI>
__global__ __launch_bounds__(512) void k_mixed(LaunchDesc a) {
    extern __shared__ float lds[];
    for (int it = 0;; ++it) {
        const int w = next_work(a, it);
        if (w >= a.n_work) break;
        const DevInst* I = a.insts + a.order[w];
        const int key = __builtin_amdgcn_readfirstlane(I->model * 4 + I->np);
        switch (key) {
        case HB_G2DO * 4 + 1: run_instance<HB_G2DO, 1, LRING, MULTI>(I, lds); break;
        case HB_MPR * 4 + 2: run_instance<HB_MPR, 2, LRING, MULTI>(I, lds); break;
        case HB_JANSEN_RIT * 4 + 1: run_instance<HB_JANSEN_RIT, 1, LRING, MULTI>(I, lds); break;
        case HB_WILSON_COWAN * 4 + 1: run_instance<HB_WILSON_COWAN, 1, LRING, MULTI>(I, lds); break;
        case HB_KURAMOTO * 4 + 1: run_instance<HB_KURAMOTO, 1, LRING, MULTI>(I, lds); break;
        case HB_KURAMOTO * 4 + 2: run_instance<HB_KURAMOTO, 2, LRING, MULTI>(I, lds); break;
        case HB_RWW * 4 + 1: run_instance<HB_RWW, 1, LRING, MULTI>(I, lds); break;
        default: break;
        }
    }
}

template <int M, int NP>
static const void* kptr_single(bool lring, bool multi) {
    if (lring) return multi ? (const void*)k_single<M, NP, true, true> : (const void*)k_single<M, NP, true, false>;
    return multi ? (const void*)k_single<M, NP, false, true> : (const void*)k_single<M, NP, false, false>;
}

// The kernel a launch description selects (nullptr: unsupported combination).
static const void* kernel_ptr(const LaunchDesc& d) {
    if (d.model < 0) {
        if (d.lds_ring) return d.multi ? (const void*)k_mixed<true, true> : (const void*)k_mixed<true, false>;
        return d.multi ? (const void*)k_mixed<false, true> : (const void*)k_mixed<false, false>;
    }
    switch (d.model * 4 + d.np) {
    case HB_G2DO * 4 + 1: return kptr_single<HB_G2DO, 1>(d.lds_ring, d.multi);
    case HB_MPR * 4 + 2: return kptr_single<HB_MPR, 2>(d.lds_ring, d.multi);
    case HB_JANSEN_RIT * 4 + 1: return kptr_single<HB_JANSEN_RIT, 1>(d.lds_ring, d.multi);
    case HB_WILSON_COWAN * 4 + 1: return kptr_single<HB_WILSON_COWAN, 1>(d.lds_ring, d.multi);
    case HB_KURAMOTO * 4 + 1: return kptr_single<HB_KURAMOTO, 1>(d.lds_ring, d.multi);
    case HB_KURAMOTO * 4 + 2: return kptr_single<HB_KURAMOTO, 2>(d.lds_ring, d.multi);
    case HB_RWW * 4 + 1: return kptr_single<HB_RWW, 1>(d.lds_ring, d.multi);
    default: return nullptr;
    }
}

static hipError_t launch(const LaunchDesc& d, hipStream_t s) {
    const void* k = kernel_ptr(d);
    if (!k) return hipErrorInvalidValue;
    LaunchDesc a = d;
    void* args[] = {&a};
    return hipLaunchKernel(k, dim3(d.grid), dim3(d.block), args, (size_t)d.lds_bytes, s);
}

static hipError_t occupancy(const LaunchDesc& d, int* blocks) {
    const void* k = kernel_ptr(d);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k, d.block, (size_t)d.lds_bytes);
}

static hipError_t attrs(int b) {
    static const int models[7][2] = {{HB_G2DO, 1}, {HB_MPR, 2}, {HB_JANSEN_RIT, 1}, {HB_WILSON_COWAN, 1},
                                     {HB_KURAMOTO, 1}, {HB_KURAMOTO, 2}, {HB_RWW, 1}};
    for (int m = -1; m < 7; ++m)
        for (int lr = 0; lr < 2; ++lr)
            for (int mu = 0; mu < 2; ++mu) {
                LaunchDesc d;
                memset(&d, 0, sizeof(d));
                d.model = m < 0 ? -1 : models[m][0];
                d.np = m < 0 ? 0 : models[m][1];
                d.lds_ring = lr != 0;
                d.multi = mu != 0;
                const void* k = kernel_ptr(d);
                hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, b);
                if (e != hipSuccess) return e;
            }
    return hipSuccess;
}

}  // namespace HB_NS

#if HB_EXACT
hipError_t launch_forward_exact(const LaunchDesc& d, hipStream_t s) { return exact::launch(d, s); }
hipError_t forward_set_attributes_exact(int max_lds) { return exact::attrs(max_lds); }
#else
hipError_t launch_forward_fast(const LaunchDesc& d, hipStream_t s) { return fast::launch(d, s); }
hipError_t forward_set_attributes_fast(int max_lds) { return fast::attrs(max_lds); }
hipError_t forward_occupancy_fast(const LaunchDesc& d, int* blocks) { return fast::occupancy(d, blocks); }
#endif

}  // namespace hb
