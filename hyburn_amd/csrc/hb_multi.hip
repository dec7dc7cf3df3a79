// hb_multi.hip -- the batch engine's multi-subnetwork step on the GPU.
//
// BatchHybridEngine::run_sweep with several subnetworks and projections
// between them (engine.rs:540-773): every step, each projection reads the
// first ncvar state variables of its source subnetwork, applies the coupling
// function's pre, weights them (Scalar: w * mean over source nodes,
// engine.rs:602-608; Dense: W @ pre, 609-620), applies post (with the
// target's own state where the function needs it), remaps the channels by
// cvar_map (637-664) and adds the result into the target's coupling
// (666-673). Then every subnetwork integrates with that frozen coupling,
// Heun for G2DO and the configured rule for the others when hybrid_integrator
// is set (698-702, dfun.rs:1871), clamps, and adds to its temporal average.
//
// One workgroup per sweep point, one thread per node of every subnetwork;
// the pre values of each projection's sources live in LDS, and each Scalar
// projection's weighted mean is formed once per step. A delayed projection
// reads each source node's own history (global memory, written and read by
// the same thread). Stochastic integrators take one raw N(0,1) draw per
// node, variable and step, shared by every sweep point as the batch engine
// broadcasts it.
#include "hb_internal.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace hb {
namespace multi {

#include "hb_model.h"

constexpr int kMaxSub = 4, kMaxProj = 8, kNV = 12, kMaxNodes = 1024;
constexpr int kMaxLdsFloats = 16384;   // 64 KiB of pre values

struct Sub {
    int32_t model, n, off, nv, integ, sw, stoch;
    float mc[kMCX];        // model_consts (first kMC slots) or model_consts_ext
    float nscale[kNV];     // sqrt(2 nsig_v) * sqrt(dt)
};

struct Proj {
    int32_t src, tgt, wkind, cfn, np, nc, classic, pass_all, n_map, lds_off, delay;
    int32_t map_s[4], map_c[4];
    float w, cc[8];
    const float* W;        // [n_tgt][n_src] (HB_W_DENSE) or null
    const float* rowsum;   // [n_tgt]: Difference over a non-square Dense projection, or null
    const int32_t* ip;     // per-edge CSR (plain mode): row pointers, sources, weights, delays
    const int32_t* ix;
    const float* cw;
    const int32_t* cd;
    int32_t src_off;       // first node of the source subnetwork
};

struct Stim {
    int32_t target, pattern;
    float p[4];
};

// stimulus.rs:63-128, in f64 time t = step * dt_ms
__device__ __attribute__((noinline)) float stim_value(const Stim& S, int64_t step, double dt_ms) {
    const double t = (double)step * dt_ms;
    const double onset = (double)S.p[0];
    switch (S.pattern) {
    case 0:
        return (fabs(t - onset) < 1e-6 || (t > onset && t <= onset + 1e-6)) ? S.p[1] : 0.0f;
    case 1:
        return (t >= onset && t < (double)S.p[1]) ? S.p[2] : 0.0f;
    case 2: {
        if (t < onset) return 0.0f;
        const double raw = sin(2.0 * 3.141592653589793 * (double)S.p[2] * (t - onset) + (double)S.p[3]);
        return S.p[1] * (float)raw;
    }
    default: {
        const double period = (double)S.p[2], width = (double)S.p[3];
        if (t < onset || period <= 0.0 || width <= 0.0) return 0.0f;
        return fmod(t - onset, period) < width ? S.p[1] : 0.0f;
    }
    }
}

struct Args {
    Sub sub[kMaxSub];
    Proj proj[kMaxProj];
    int32_t n_sub, n_proj, n_tot, n_steps, H;   // H: history horizon (0: no delayed projection)
    int32_t x0_per_point;  // x0 is [B][n_tot][kNV] (resume) instead of [n_tot][kNV]
    int32_t plain, n_stim; // the plain HybridEngine's semantics; stimuli
    Stim stim[8];
    double dt_ms;
    int64_t step0;         // the engine's step counter at the first step
    float dt, dth, dt6, inv_nsteps;
    float* hist;           // [B][n_tot][H][2] each node's coupling variables of the last H steps
    const float* noise;    // [n_steps][n_tot][kNV] raw N(0,1), shared by the sweep points, or null
    const float* sweep;    // [B] swept values, or null
    const float* pt_mc;    // [B][n_sub][kMCX] per-point model constants (run_sweep_multi), or null
    const float* pt_sv;    // [B][n_sub] per-point swept-form values, or null
    const float* x0;       // [n_tot][kNV]
    float* tavg;           // [B][n_tot][kNV]
    float* fin;            // [B][n_tot][kNV]
    float* traj;           // [n_steps][B][n_tot][kNV] or null
    float* bold;           // [B][bnout][n of subnetwork 0] or null
    int32_t bper, bnout, bnative;
    double tr_ms, dtb_ms;
    float* v0rec;          // [B][n_steps][n of subnetwork 0] var0 per step (the batch-mean monitor's input) or null
};

__device__ __forceinline__ void pre_rt(const Proj& P, const float* x, float* out) {
    if (P.nc == 4) {   // four cvars (CoombesByrne, the Gast models, DumontGutkin; no Kuramoto)
        if (P.np == 1) pre<4, 1>(P.cfn, P.cc, false, P.classic, x, out);
        else pre<4, 4>(P.cfn, P.cc, false, P.classic, x, out);
    } else if (P.nc == 3) {   // EpileptorRestingState's three cvars (no Kuramoto: the host refuses it)
        if (P.np == 1) pre<3, 1>(P.cfn, P.cc, false, P.classic, x, out);
        else pre<3, 3>(P.cfn, P.cc, false, P.classic, x, out);
    } else if (P.nc == 1) {
        if (P.np == 2) pre<1, 2>(P.cfn, P.cc, false, P.classic, x, out);
        else pre<1, 1>(P.cfn, P.cc, false, P.classic, x, out);
    } else {
        if (P.np == 4) pre<2, 4>(P.cfn, P.cc, false, P.classic, x, out);
        else if (P.np == 1) pre<2, 1>(P.cfn, P.cc, false, P.classic, x, out);
        else pre<2, 2>(P.cfn, P.cc, false, P.classic, x, out);
    }
}

__device__ __forceinline__ void post_rt(const Proj& P, const float* g, const float* own, float* out) {
    if (P.nc == 4) {
        if (P.np == 1) post<4, 1>(P.cfn, P.cc, P.classic, g, own, out);
        else post<4, 4>(P.cfn, P.cc, P.classic, g, own, out);
    } else if (P.nc == 3) {
        if (P.np == 1) post<3, 1>(P.cfn, P.cc, P.classic, g, own, out);
        else post<3, 3>(P.cfn, P.cc, P.classic, g, own, out);
    } else if (P.nc == 1) {
        if (P.np == 2) post<1, 2>(P.cfn, P.cc, P.classic, g, own, out);
        else post<1, 1>(P.cfn, P.cc, P.classic, g, own, out);
    } else {
        if (P.np == 4) post<2, 4>(P.cfn, P.cc, P.classic, g, own, out);
        else if (P.np == 1) post<2, 1>(P.cfn, P.cc, P.classic, g, own, out);
        else post<2, 2>(P.cfn, P.cc, P.classic, g, own, out);
    }
}

// mc: the register copy of the first kMC constants; mcg: the full table in
// global memory, which models 10-27 read directly. EXT 1 holds models 10-27
// but ZetterbergJansen, EXT 2 all of them with twelve state registers; the
// ten models of EXT 0 keep a kernel free of the others' code.
template <int EXT>
__device__ __forceinline__ void step_model(int model, const StepCtx& S, const float* mc, const float* mcg,
                                           const float* cpl, const float* z, float* x) {
    if constexpr (EXT == 2) {
        if (model == HB_ZETTERBERG_JANSEN) { integrate<HB_ZETTERBERG_JANSEN>(S, mcg, cpl, z, x); return; }
    }
    if constexpr (EXT >= 1) {
        switch (model) {
        case HB_COOMBES_BYRNE: integrate<HB_COOMBES_BYRNE>(S, mcg, cpl, z, x); return;
        case HB_GAST_SD: integrate<HB_GAST_SD>(S, mcg, cpl, z, x); return;
        case HB_GAST_SF: integrate<HB_GAST_SF>(S, mcg, cpl, z, x); return;
        case HB_LARTER_BREAKSPEAR: integrate<HB_LARTER_BREAKSPEAR>(S, mcg, cpl, z, x); return;
        case HB_EPILEPTOR_2D: integrate<HB_EPILEPTOR_2D>(S, mcg, cpl, z, x); return;
        case HB_EPILEPTOR: integrate<HB_EPILEPTOR>(S, mcg, cpl, z, x); return;
        case HB_RWW_EXC_INH: integrate<HB_RWW_EXC_INH>(S, mcg, cpl, z, x); return;
        case HB_DECO_BALANCED: integrate<HB_DECO_BALANCED>(S, mcg, cpl, z, x); return;
        case HB_REDUCED_FHN: integrate<HB_REDUCED_FHN>(S, mcg, cpl, z, x); return;
        case HB_REDUCED_HR: integrate<HB_REDUCED_HR>(S, mcg, cpl, z, x); return;
        case HB_DUMONT_GUTKIN: integrate<HB_DUMONT_GUTKIN>(S, mcg, cpl, z, x); return;
        default: break;
        }
        switch (model) {
        case HB_EPILEPTOR_CODIM3: integrate<HB_EPILEPTOR_CODIM3>(S, mcg, cpl, z, x); return;
        case HB_EPILEPTOR_CODIM3_SLOWMOD: integrate<HB_EPILEPTOR_CODIM3_SLOWMOD>(S, mcg, cpl, z, x); return;
        case HB_EPILEPTOR_RS: integrate<HB_EPILEPTOR_RS>(S, mcg, cpl, z, x); return;
        case HB_ZERLAUT_FIRST: integrate<HB_ZERLAUT_FIRST>(S, mcg, cpl, z, x); return;
        case HB_ZERLAUT_SECOND: integrate<HB_ZERLAUT_SECOND>(S, mcg, cpl, z, x); return;
        case HB_KIONEX: integrate<HB_KIONEX>(S, mcg, cpl, z, x); return;
        default: break;
        }
    }
    switch (model) {
    case HB_G2DO: integrate<HB_G2DO>(S, mc, cpl, z, x); break;
    case HB_MPR: integrate<HB_MPR>(S, mc, cpl, z, x); break;
    case HB_JANSEN_RIT: integrate<HB_JANSEN_RIT>(S, mc, cpl, z, x); break;
    case HB_WILSON_COWAN: integrate<HB_WILSON_COWAN>(S, mc, cpl, z, x); break;
    case HB_KURAMOTO: integrate<HB_KURAMOTO>(S, mc, cpl, z, x); break;
    case HB_LINEAR: integrate<HB_LINEAR>(S, mc, cpl, z, x); break;
    case HB_SUP_HOPF: integrate<HB_SUP_HOPF>(S, mc, cpl, z, x); break;
    case HB_HOPFIELD: integrate<HB_HOPFIELD>(S, mc, cpl, z, x); break;
    case HB_COOMBES_BYRNE_2D: integrate<HB_COOMBES_BYRNE_2D>(S, mc, cpl, z, x); break;
    default: integrate<HB_RWW>(S, mc, cpl, z, x); break;
    }
}

// BOLD output rows after native sample bk (time tk, value sig), the stream
// form of bold_monitor.rs:123-203: native samples as they are, or TR rows
// interpolated linearly between the previous sample and this one.
struct BoldClock {
    int k, j;
    double tprev;
};

__device__ __attribute__((noinline)) void bold_emit(const Args* A, const BoldClock& c, double tk, float sig, float prev, float* out, int N0) {
    if (A->bnative) {
        if (c.k < A->bnout) out[(size_t)c.k * N0] = sig;
        return;
    }
    const double tr = A->tr_ms;
    int j = c.j;
    if (c.k == 0) {
        while (j < A->bnout && (double)(j + 1) * tr <= tk) { out[(size_t)j * N0] = sig; ++j; }
    } else {
        while (j < A->bnout && (double)(j + 1) * tr < tk) {
            const double tt = (double)(j + 1) * tr, d = tk - c.tprev;
            double w = fabs(d) < 1e-12 ? 0.0 : (tt - c.tprev) / d;
            w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
            out[(size_t)j * N0] = prev + (sig - prev) * (float)w;
            ++j;
        }
    }
}

__device__ __attribute__((noinline)) void bold_advance(const Args* A, BoldClock& c, double tk) {
    if (!A->bnative) {
        const double tr = A->tr_ms;
        if (c.k == 0) { while (c.j < A->bnout && (double)(c.j + 1) * tr <= tk) c.j++; }
        else { while (c.j < A->bnout && (double)(c.j + 1) * tr < tk) c.j++; }
    }
    c.k++;
    c.tprev = tk;
}

// NVR: state variables held in registers (6 for the ten models of the base
// kernel, 8 with models 10-27 but ZetterbergJansen, 12 with it); memory
// layouts keep the kNV stride.
template <int EXT>
__global__ __launch_bounds__(1024) void k_multi(const Args* __restrict__ A) {
    constexpr int NVR = EXT == 2 ? 12 : (EXT == 1 ? 8 : 6);
    extern __shared__ float pv[];   // pre values of every projection's source nodes
    __shared__ float gmean[kMaxProj * 4];   // w * mean over source nodes, per Scalar projection and channel
    const int t = threadIdx.x, b = blockIdx.x;
    int s = -1;
    for (int q = 0; q < A->n_sub; ++q)
        if (t >= A->sub[q].off && t < A->sub[q].off + A->sub[q].n) s = q;
    const bool live = s >= 0;
    const int i = live ? t - A->sub[s].off : 0;
    const int model = live ? A->sub[s].model : 0, nv = live ? A->sub[s].nv : 0;
    // per-point constants of a run_sweep_multi, else the subnetwork's own
    const float* mcg = live ? (A->pt_mc ? A->pt_mc + ((size_t)b * A->n_sub + s) * kMCX : A->sub[s].mc) : A->sub[0].mc;
    float mc[kMC];
#pragma unroll
    for (int k = 0; k < kMC; ++k) mc[k] = live ? mcg[k] : 0.0f;
    float x[NVR], acc[NVR];
#pragma unroll
    for (int v = 0; v < NVR; ++v) {
        x[v] = (live && v < nv) ? A->x0[((A->x0_per_point ? (size_t)b * A->n_tot : 0) + t) * kNV + v] : 0.0f;
        acc[v] = 0.0f;
    }
    StepCtx S;
    S.dt = A->dt;
    S.dth = A->dth;
    S.dt6 = A->dt6;
    S.integ = live ? A->sub[s].integ : HB_EULER;
    S.stoch = live && A->sub[s].stoch != 0;
    S.plain = A->plain != 0;
    float nscale[NVR];
#pragma unroll
    for (int v = 0; v < NVR; ++v) nscale[v] = live ? A->sub[s].nscale[v] : 0.0f;
    S.sw = live && A->sub[s].sw != 0;
    S.sv = A->pt_sv ? (live ? A->pt_sv[(size_t)b * A->n_sub + s] : 0.0f) : (A->sweep ? A->sweep[b] : 0.0f);
    const int H = A->H;
    float* hist = (live && H > 0) ? A->hist + ((size_t)b * A->n_tot + t) * H * 2 : nullptr;
    // BOLD of subnetwork 0 (bold_monitor.rs:72-106): window sums of variable 0,
    // one Balloon-Windkessel step per closed window; the clock is uniform
    const int bper = A->bold ? A->bper : 0, N0 = A->sub[0].n;
    const bool bnode = live && s == 0 && bper > 0;
    float* bout = bnode ? A->bold + (size_t)b * A->bnout * N0 + i : nullptr;
    Bold bw;
    bw.s = 0.0f; bw.f = 1.0f; bw.v = 1.0f; bw.q = 1.0f; bw.acc = 0.0f; bw.prev = 0.0f;
    BoldClock bc;
    bc.k = 0; bc.j = 0; bc.tprev = 0.0;
    int bcount = 0;
    for (int k = 0; k < A->n_steps; ++k) {
        const int64_t step = A->step0 + k;   // the engine's step counter (delays read by it)
        // the history window gains the current state (engine.rs:541-551); each
        // node keeps its own first two variables by real step, slot step % H
        if (hist) {
            hist[(step % H) * 2] = x[0];
            hist[(step % H) * 2 + 1] = x[1];
        }
        for (int p = 0; p < A->n_proj; ++p) {
            const Proj& P = A->proj[p];
            if (live && s == P.src) {
                float o[4], xd[2];
                const float* xs = x;
                if (P.delay > 0 && step > 0) {
                    // engine.rs:565-589: window slot (step - d + H) % H holds the
                    // state of step step - H + 1 + slot; zeros before step 0 and
                    // while d > step
                    xd[0] = xd[1] = 0.0f;
                    if (P.delay <= step) {
                        // plain: the true ring, state of step - d (runtime.rs:222-233)
                        const int64_t ts = A->plain ? step - P.delay : step - H + 1 + (step - P.delay + H) % H;
                        if (ts >= 0) {
                            xd[0] = hist[(ts % H) * 2];
                            xd[1] = hist[(ts % H) * 2 + 1];
                        }
                    }
                    xs = xd;
                }
                pre_rt(P, xs, o);
                for (int c = 0; c < P.np; ++c) pv[P.lds_off + i * P.np + c] = o[c];
            }
        }
        __syncthreads();
        // Scalar projections: one thread per (projection, channel) sums the
        // source pre values in node order, so each mean is formed once
        if (t < A->n_proj * 4) {
            const Proj& P = A->proj[t >> 2];
            const int c = t & 3;
            if (P.wkind == HB_W_SCALAR_MEAN && c < P.np) {
                const int ns = A->sub[P.src].n, np = P.np;
                const float* src = pv + P.lds_off;
                float sum = 0.0f;
                for (int j = 0; j < ns; ++j) sum += src[j * np + c];
                gmean[t] = (sum / (float)ns) * P.w;
            }
        }
        __syncthreads();
        if (live) {
            float cpl[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int p = 0; p < A->n_proj; ++p) {
                const Proj& P = A->proj[p];
                if (s != P.tgt) continue;
                const int ns = A->sub[P.src].n, np = P.np;
                const float* src = pv + P.lds_off;
                float g[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                if (P.wkind == HB_W_SCALAR_MEAN) {
                    for (int c = 0; c < np; ++c) g[c] = gmean[p * 4 + c];
                } else if (P.wkind == HB_W_CSR) {
                    // runtime.rs:160-203: every edge reads its source d_e steps
                    // back from that node's history (written before the barrier
                    // above); d_e > step contributes a zero pre vector
                    for (int e = P.ip[i]; e < P.ip[i + 1]; ++e) {
                        const int j = P.ix[e], d = P.cd[e];
                        const float w = P.cw[e];
                        float o[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                        if (d == 0 || step == 0 || d <= step) {
                            const int64_t ts = (d == 0 || step == 0) ? step : step - d;
                            const float* hj = A->hist + (((size_t)b * A->n_tot + P.src_off + j) * H + ts % H) * 2;
                            float xd[2] = {hj[0], hj[1]};
                            if (P.nc == 1) {
                                if (np == 2) pre<1, 2>(P.cfn, P.cc, true, P.classic, xd, o);
                                else pre<1, 1>(P.cfn, P.cc, true, P.classic, xd, o);
                            } else {
                                if (np == 4) pre<2, 4>(P.cfn, P.cc, true, P.classic, xd, o);
                                else if (np == 1) pre<2, 1>(P.cfn, P.cc, true, P.classic, xd, o);
                                else pre<2, 2>(P.cfn, P.cc, true, P.classic, xd, o);
                            }
                        }
                        for (int c = 0; c < np; ++c) g[c] += w * o[c];
                    }
                } else {
                    const float* Wr = P.W + (size_t)i * ns;
                    for (int j = 0; j < ns; ++j) {
                        const float w = Wr[j];
                        for (int c = 0; c < np; ++c) g[c] += w * src[j * np + c];
                    }
                }
                if (P.rowsum) {
                    const float rs = P.rowsum[i];
                    for (int c = 0; c < np; ++c) g[c] = g[c] - x[c] * rs;
                }
                float own[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                if (P.cfn == HB_CF_KURAMOTO)
                    for (int c = 0; c < P.nc; ++c) {
                        own[c] = sinf(x[c]);
                        own[P.nc + c] = cosf(x[c]);
                    }
                float out[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                post_rt(P, g, own, out);
                if (P.pass_all) {
                    for (int c = 0; c < P.nc; ++c) cpl[c] += out[c];
                } else {
                    for (int m = 0; m < P.n_map; ++m) cpl[P.map_c[m]] += out[P.map_s[m]];
                }
            }
            for (int q = 0; q < A->n_stim; ++q)   // runtime.rs:319-357: added to coupling variable 0
                if (A->stim[q].target == s) {
                    const float sv = stim_value(A->stim[q], step, A->dt_ms);
                    if (sv != 0.0f) cpl[0] = cpl[0] + sv;
                }
            float z[NVR];
#pragma unroll
            for (int v = 0; v < NVR; ++v) z[v] = 0.0f;
            if (S.stoch) {
                const float* raw = A->noise + ((size_t)k * A->n_tot + t) * kNV;
                for (int v = 0; v < nv; ++v) z[v] = raw[v] * nscale[v];   // integrator.rs:183-191
            }
            step_model<EXT>(model, S, mc, mcg, cpl, z, x);
#pragma unroll
            for (int v = 0; v < NVR; ++v) acc[v] += x[v];
            if (A->traj) {
                float* o = A->traj + (((size_t)k * gridDim.x + b) * A->n_tot + t) * kNV;
                for (int v = 0; v < nv; ++v) o[v] = x[v];
            }
            if (bnode) bw.acc = bw.acc + x[0];
            if (A->v0rec && s == 0) A->v0rec[((size_t)b * A->n_steps + k) * N0 + i] = x[0];
        }
        if (bper > 0 && ++bcount >= bper) {
            const double dtb = A->dt_ms * (double)bcount;
            const float dts = (float)(dtb / 1000.0);
            const double tk = __dadd_rn(__dmul_rn((double)bc.k, dtb), dtb);
            if (bnode) {
                const float sig = bold_step(bw, bw.acc / (float)bcount, dts);
                bold_emit(A, bc, tk, sig, bw.prev, bout, N0);
                bw.prev = sig;
                bw.acc = 0.0f;
            }
            bold_advance(A, bc, tk);
            bcount = 0;
        }
        __syncthreads();
    }
    if (bper > 0) {   // flush of a partial window, then the last value on the remaining TR rows
        if (bcount > 0) {
            const double dtb = A->dt_ms * (double)bcount;
            const float dts = (float)(dtb / 1000.0);
            const double tk = __dadd_rn(__dmul_rn((double)bc.k, A->dtb_ms), dtb);
            if (bnode) {
                const float sig = bold_step(bw, bw.acc / (float)bcount, dts);
                bold_emit(A, bc, tk, sig, bw.prev, bout, N0);
                bw.prev = sig;
            }
            bold_advance(A, bc, tk);
        }
        if (bnode && !A->bnative && bc.k > 0)
            for (int jj = bc.j; jj < A->bnout; ++jj) bout[(size_t)jj * N0] = bw.prev;
    }
    if (live) {
        const size_t o = ((size_t)b * A->n_tot + t) * kNV;
        for (int v = 0; v < nv; ++v) {
            A->tavg[o + v] = acc[v] * A->inv_nsteps;
            A->fin[o + v] = x[v];
        }
    }
}

// Position of state variable t among the model's coupling variables (the
// CVAR constants of src/model/*.rs), 0 when absent (engine.rs:391-393).
int cvar_position(int model, int t) {
    static const int g2do[] = {0}, mpr[] = {0, 1}, one0[] = {0}, jr[] = {1}, ers[] = {0, 3, 6},
                     four[] = {0, 1, 2, 3}, x03[] = {0, 3}, x02[] = {0, 2}, zj[] = {3}, dg[] = {0, 1, 4, 5};
    const int* cv = one0;
    int n = 1;
    if (model == HB_G2DO) cv = g2do;
    else if (model == HB_EPILEPTOR_RS) {
        cv = ers;
        n = 3;
    }
    else if (model == HB_MPR || model == HB_SUP_HOPF || model == HB_HOPFIELD || model == HB_COOMBES_BYRNE_2D) {
        cv = mpr;
        n = 2;
    }
    else if (model == HB_JANSEN_RIT) cv = jr;
    else if (model == HB_COOMBES_BYRNE || model == HB_GAST_SD || model == HB_GAST_SF) { cv = four; n = 4; }
    else if (model == HB_EPILEPTOR || model == HB_REDUCED_HR) { cv = x03; n = 2; }
    else if (model == HB_REDUCED_FHN) { cv = x02; n = 2; }
    else if (model == HB_ZETTERBERG_JANSEN) cv = zj;
    else if (model == HB_DUMONT_GUTKIN) { cv = dg; n = 4; }
    for (int k = 0; k < n; ++k)
        if (cv[k] == t) return k;
    return 0;
}

}  // namespace multi
}  // namespace hb

using namespace hb;

namespace {

int bad(const std::string& m) { return set_error(HB_INVALID_CONFIG, m); }

int hip_fail(hipError_t e, const char* what) {
    return set_error(HB_BACKEND_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

extern "C" int hb_run_multi(hb_multi_spec* spec) {
    using namespace hb::multi;
    if (!spec) return bad("null spec");
    const hb_multi_spec& Q = *spec;
    if (Q.n_subnets < 1 || Q.n_subnets > kMaxSub || !Q.subnets) return bad("n_subnets must be in [1, 4]");
    if (Q.n_projections < 0 || Q.n_projections > kMaxProj || (Q.n_projections > 0 && !Q.projections))
        return bad("n_projections must be in [0, 8]");
    if (Q.n_sweep < 1 || Q.n_steps < 0) return bad("n_sweep must be >= 1 and n_steps >= 0");
    if (Q.integrator < HB_EULER || Q.integrator > HB_RK4_STOCH) return bad("unknown integrator");
    if (Q.sweep_sub < -1 || Q.sweep_sub >= Q.n_subnets) return bad("sweep subnetwork index out of range");
    if (Q.plain && Q.hybrid) return bad("the plain engine has no hybrid_integrator");
    if (Q.n_stimuli < 0 || Q.n_stimuli > 8 || (Q.n_stimuli > 0 && !Q.stimuli)) return bad("n_stimuli must be in [0, 8]");
    if (Q.n_stimuli > 0 && !Q.plain) return bad("stimuli apply to the plain engine only (the batch engine ignores them)");
    if (Q.sweep_sub >= 0 && !Q.sweep_values) return bad("sweep values missing");
    if (!Q.tavg || !Q.final_state) return bad("null output pointer");
    bool any_point_mode = false;
    for (int q = 0; q < Q.n_subnets; ++q) {
        if (Q.point_mode[q] < 0 || Q.point_mode[q] > 2) return bad("point_mode must be 0, 1 or 2");
        if (Q.point_mode[q] == 1 && !Q.point_values) return bad("point_mode 1 needs point_values");
        if (Q.point_mode[q] == 1 && model_sweep_index(Q.subnets[q].model) < 0)
            return bad("point_mode 1: subnetwork " + std::to_string(q) + "'s model has no batch-native swept parameter");
        if (Q.point_mode[q] == 2 && !Q.point_params) return bad("point_mode 2 needs point_params");
        any_point_mode = any_point_mode || Q.point_mode[q] != 0;
    }
    Args A;
    std::memset(&A, 0, sizeof(A));
    int n_tot = 0;
    for (int q = 0; q < Q.n_subnets; ++q) {
        const hb_subnetwork& U = Q.subnets[q];
        if (U.model < HB_G2DO || U.model > HB_DUMONT_GUTKIN) return bad("unknown model in subnetwork " + std::to_string(q));
        if (U.n_nodes < 1 || !U.params || !U.x0) return bad("subnetwork " + std::to_string(q) + ": nodes, params, x0");
        if (model_ext(U.model) && U.n_params != model_nparams(U.model))
            return bad("subnetwork " + std::to_string(q) + ": expected " + std::to_string(model_nparams(U.model)) +
                       " parameters");
        Sub& S = A.sub[q];
        S.model = U.model;
        S.n = U.n_nodes;
        S.off = n_tot;
        S.nv = model_nvar(U.model);
        // engine.rs:698-773: hybrid -> deterministic Heun for G2DO; otherwise the
        // configured integrator, where a configured (deterministic) Heun falls
        // through to Euler
        if (Q.hybrid) S.integ = model_prefers_heun(U.model) ? HB_HEUN : (Q.integrator == HB_HEUN ? HB_EULER : Q.integrator);
        else S.integ = Q.integrator;
        S.stoch = (S.integ == HB_EULER_STOCH || S.integ == HB_HEUN_STOCH || S.integ == HB_RK4_STOCH) ? 1 : 0;
        if (S.stoch) {
            if (!Q.nsig || !Q.noise) return bad("stochastic integration needs nsig and the raw noise");
            if (Q.n_nsig < S.nv)
                return bad("nsig has " + std::to_string(Q.n_nsig) + " entries, subnetwork " + std::to_string(q) +
                           " integrates " + std::to_string(S.nv) + " noisy variables");
            for (int v = 0; v < S.nv; ++v) S.nscale[v] = sqrtf(2.0f * Q.nsig[v]) * sqrtf(Q.dt);
        }
        S.sw = (q == Q.sweep_sub && Q.sweep_param == model_sweep_index(U.model)) ? 1 : 0;
        if (any_point_mode) S.sw = Q.point_mode[q] == 1 ? 1 : 0;
        for (int k = 0; k < kMCX; ++k) S.mc[k] = 0.0f;
        if (model_ext(U.model)) model_consts_ext(U.model, U.params, S.mc);
        else model_consts(U.model, U.params, S.mc);
        n_tot += U.n_nodes;
    }
    if (n_tot > kMaxNodes) return bad("at most 1024 nodes over all subnetworks");
    DevBufs B;
    std::vector<std::vector<float>> host_w((size_t)Q.n_projections), host_rs((size_t)Q.n_projections);
    int lds = 0;
    for (int p = 0; p < Q.n_projections; ++p) {
        const hb_multi_projection& R = Q.projections[p];
        const std::string tag = "projection " + std::to_string(p);
        if (R.src < 0 || R.src >= Q.n_subnets || R.tgt < 0 || R.tgt >= Q.n_subnets) return bad(tag + ": src/tgt out of range");
        if (R.delay < 0) return bad(tag + ": negative delay");
        if (R.wkind != HB_W_SCALAR_MEAN && R.wkind != HB_W_DENSE && R.wkind != HB_W_CSR)
            return bad(tag + ": weights must be Scalar, Dense or CSR");
        if (R.wkind == HB_W_DENSE && !R.dense_w) return bad(tag + ": dense weights missing");
        if (R.wkind == HB_W_CSR) {
            const int ntr = Q.subnets[R.tgt].n_nodes, nsr = Q.subnets[R.src].n_nodes;
            if (!R.indptr || R.nnz < 0 || (R.nnz > 0 && (!R.indices || !R.csr_w))) return bad(tag + ": CSR arrays missing");
            if (R.indptr[0] != 0 || R.indptr[ntr] != R.nnz) return bad(tag + ": CSR indptr inconsistent with nnz");
            for (int r = 0; r < ntr; ++r)
                if (R.indptr[r + 1] < R.indptr[r]) return bad(tag + ": CSR indptr not monotone");
            for (int e = 0; e < R.nnz; ++e) {
                if (R.indices[e] < 0 || R.indices[e] >= nsr) return bad(tag + ": CSR source index out of range");
                if (R.edge_delays && R.edge_delays[e] < 0) return bad(tag + ": negative edge delay");
            }
            if (R.edge_delays && !Q.plain) return bad(tag + ": per-edge delays belong to the plain engine");
            if (R.edge_delays && R.cfn == HB_CF_DIFFERENCE) return bad(tag + ": Difference with per-edge delays");
        }
        if (R.cfn < HB_CF_LINEAR || R.cfn > HB_CF_PRESIGMOIDAL) return bad(tag + ": unknown coupling function");
        if (R.n_cparams < 0 || R.n_cparams > 8) return bad(tag + ": at most 8 coupling parameters");
        if (const char* why = presigmoidal_unsupported(R.cfn, R.n_cparams, R.cparams)) return bad(tag + ": " + why);
        const int ms = Q.subnets[R.src].model, mt = Q.subnets[R.tgt].model;
        const int ncs = model_ncvar(ms), nct = model_ncvar(mt);
        const int ns = Q.subnets[R.src].n_nodes, nt = Q.subnets[R.tgt].n_nodes;
        if (R.n_map < 1 || R.n_map > 4) return bad(tag + ": cvar_map needs 1 to 4 pairs");
        for (int m = 0; m < R.n_map; ++m) {   // engine.rs:371-388
            if (R.map_src[m] < 0 || R.map_src[m] >= ncs)
                return bad(tag + " cvar_map src index " + std::to_string(R.map_src[m]) + " >= src_ncvar " + std::to_string(ncs));
            if (R.map_tgt[m] < 0 || R.map_tgt[m] >= nct)
                return bad(tag + " cvar_map tgt index " + std::to_string(R.map_tgt[m]) + " >= tgt_ncvar " + std::to_string(nct));
        }
        Proj& P = A.proj[p];
        P.src = R.src;
        P.tgt = R.tgt;
        P.wkind = R.wkind;
        P.cfn = R.cfn;
        P.nc = ncs;
        P.np = coupling_npre(R.cfn, ncs);
        if ((R.cfn == HB_CF_KURAMOTO || R.cfn == HB_CF_DIFFERENCE) && ncs != nct)
            return bad(tag + ": Kuramoto and Difference coupling need equal source and target ncvar");
        // three or four coupling variables: the history keeps two per node, and
        // Kuramoto's six or eight pre channels exceed the kernel's four
        if (ncs > 2 && (R.delay > 0 || R.edge_delays))
            return bad(tag + ": delayed projections from a subnetwork of more than two cvars are not supported");
        if (ncs > 2 && R.cfn == HB_CF_KURAMOTO)
            return bad(tag + ": Kuramoto coupling from a subnetwork of more than two cvars is not supported");
        float cp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < R.n_cparams; ++k) cp[k] = R.cparams[k];
        int classic = 0;
        coupling_consts(R.cfn, cp, R.n_cparams, ns, P.cc, &classic);
        P.classic = classic;
        P.w = R.scalar_w;
        if (R.wkind == HB_W_DENSE) host_w[p].assign(R.dense_w, R.dense_w + (size_t)nt * ns);
        if (R.wkind == HB_W_CSR && !R.edge_delays) {   // densified (engine.rs:298-311)
            host_w[p].assign((size_t)nt * ns, 0.0f);
            for (int r = 0; r < nt; ++r)
                for (int e = R.indptr[r]; e < R.indptr[r + 1]; ++e) host_w[p][(size_t)r * ns + R.indices[e]] = R.csr_w[e];
            P.wkind = HB_W_DENSE;
        }
        if (Q.plain && R.wkind == HB_W_SCALAR_MEAN) {   // construction.rs:639-654: w on the diagonal
            if (nt != ns) return bad("Scalar weights require src and tgt nnodes to match");
            host_w[p].assign((size_t)nt * ns, 0.0f);
            for (int r = 0; r < nt; ++r) host_w[p][(size_t)r * ns + r] = R.scalar_w;
            P.wkind = HB_W_DENSE;
        }
        if (R.cfn == HB_CF_DIFFERENCE) {   // engine.rs:325-366
            const float a = R.n_cparams > 0 ? cp[0] : 1.0f;
            if (P.wkind == HB_W_SCALAR_MEAN) {
                P.w = 0.0f;
                P.cfn = HB_CF_LINEAR;
                P.cc[0] = a;
                P.cc[1] = 0.0f;
            } else if (nt == ns) {
                std::vector<float>& W = host_w[p];
                for (int r = 0; r < nt; ++r) {
                    float sum = 0.0f;
                    for (int j = 0; j < ns; ++j) sum += W[(size_t)r * ns + j];
                    W[(size_t)r * ns + r] -= sum;
                }
                P.cfn = HB_CF_LINEAR;
                P.cc[0] = a;
                P.cc[1] = 0.0f;
            } else {
                host_rs[p].assign((size_t)nt, 0.0f);
                for (int r = 0; r < nt; ++r) {
                    float sum = 0.0f;
                    for (int j = 0; j < ns; ++j) sum += host_w[p][(size_t)r * ns + j];
                    host_rs[p][r] = sum;
                }
                P.cc[0] = a;
            }
        }
        P.pass_all = (nct == ncs && R.n_map == 1 && R.map_src[0] == 0 && R.map_tgt[0] == 0) ? 1 : 0;
        P.n_map = R.n_map;
        for (int m = 0; m < R.n_map; ++m) {
            P.map_s[m] = R.map_src[m];
            P.map_c[m] = cvar_position(mt, R.map_tgt[m]);
        }
        P.lds_off = lds;
        P.delay = R.delay;
        lds += ns * P.np;
    }
    if (lds > kMaxLdsFloats) return bad("projection sources exceed 64 KiB of pre values");
    int max_delay = std::max(0, Q.max_delay);
    bool per_edge = false;
    for (int p = 0; p < Q.n_projections; ++p) {
        const hb_multi_projection& R = Q.projections[p];
        max_delay = std::max(max_delay, R.delay);
        if (R.wkind == HB_W_CSR && R.edge_delays) {
            per_edge = true;
            for (int e = 0; e < R.nnz; ++e) max_delay = std::max(max_delay, R.edge_delays[e]);
        }
    }
    if (max_delay > 1000000) return bad("delays above 10^6 steps");
    // engine.rs:253-257 / construction.rs:598-603; per-edge reads always go
    // through the history, so it exists whenever such a projection does
    const int H = max_delay > 0 ? max_delay + 1 : (per_edge ? 1 : 0);
    hipError_t e = hipSetDevice(Q.device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    for (int p = 0; p < Q.n_projections; ++p) {
        if (!host_w[p].empty()) {
            float* d = nullptr;
            e = B.alloc(&d, host_w[p].size());
            if (e == hipSuccess) e = hipMemcpy(d, host_w[p].data(), sizeof(float) * host_w[p].size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(e, "dense weights upload");
            A.proj[p].W = d;
        }
        const hb_multi_projection& R = Q.projections[p];
        A.proj[p].src_off = A.sub[R.src].off;
        if (R.wkind == HB_W_CSR && R.edge_delays) {
            const int ntr = Q.subnets[R.tgt].n_nodes;
            int32_t *dip = nullptr, *dix = nullptr, *dcd = nullptr;
            float* dcw = nullptr;
            e = B.alloc(&dip, (size_t)ntr + 1);
            if (e == hipSuccess) e = B.alloc(&dix, (size_t)R.nnz);
            if (e == hipSuccess) e = B.alloc(&dcd, (size_t)R.nnz);
            if (e == hipSuccess) e = B.alloc(&dcw, (size_t)R.nnz);
            if (e == hipSuccess) e = hipMemcpy(dip, R.indptr, sizeof(int32_t) * (ntr + 1), hipMemcpyHostToDevice);
            if (e == hipSuccess && R.nnz) e = hipMemcpy(dix, R.indices, sizeof(int32_t) * R.nnz, hipMemcpyHostToDevice);
            if (e == hipSuccess && R.nnz) e = hipMemcpy(dcd, R.edge_delays, sizeof(int32_t) * R.nnz, hipMemcpyHostToDevice);
            if (e == hipSuccess && R.nnz) e = hipMemcpy(dcw, R.csr_w, sizeof(float) * R.nnz, hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(e, "CSR upload");
            A.proj[p].ip = dip;
            A.proj[p].ix = dix;
            A.proj[p].cd = dcd;
            A.proj[p].cw = dcw;
        }
        if (!host_rs[p].empty()) {
            float* d = nullptr;
            e = B.alloc(&d, host_rs[p].size());
            if (e == hipSuccess) e = hipMemcpy(d, host_rs[p].data(), sizeof(float) * host_rs[p].size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(e, "row sums upload");
            A.proj[p].rowsum = d;
        }
    }
    A.n_sub = Q.n_subnets;
    A.n_proj = Q.n_projections;
    A.n_tot = n_tot;
    A.n_steps = Q.n_steps;
    A.dt = Q.dt;
    A.dth = Q.dt * 0.5f;
    A.dt6 = Q.dt / 6.0f;
    A.inv_nsteps = 1.0f / (float)Q.n_steps;
    if (Q.step0 < 0) return bad("step0 must be >= 0");
    A.step0 = Q.step0;
    A.x0_per_point = Q.state_in ? 1 : 0;
    std::vector<float> x0((size_t)(Q.state_in ? Q.n_sweep : 1) * n_tot * kNV, 0.0f);
    if (Q.state_in) {   // per subnetwork [n_sweep][n][nv], concatenated
        size_t o = 0;
        for (int q = 0; q < Q.n_subnets; ++q)
            for (int b = 0; b < Q.n_sweep; ++b)
                for (int r = 0; r < A.sub[q].n; ++r)
                    for (int v = 0; v < A.sub[q].nv; ++v, ++o)
                        x0[((size_t)b * n_tot + A.sub[q].off + r) * kNV + v] = Q.state_in[o];
    } else {
        for (int q = 0; q < Q.n_subnets; ++q)
            for (int r = 0; r < A.sub[q].n; ++r)
                for (int v = 0; v < A.sub[q].nv; ++v)
                    x0[(size_t)(A.sub[q].off + r) * kNV + v] = Q.subnets[q].x0[(size_t)r * A.sub[q].nv + v];
    }
    const size_t nout = (size_t)Q.n_sweep * n_tot * kNV;
    float *dx0 = nullptr, *dsw = nullptr, *dtavg = nullptr, *dfin = nullptr;
    Args* dA = nullptr;
    e = B.alloc(&dx0, x0.size());
    if (e == hipSuccess) e = B.alloc(&dtavg, nout);
    if (e == hipSuccess) e = B.alloc(&dfin, nout);
    if (e == hipSuccess) e = B.alloc(&dA, 1);
    if (e == hipSuccess && Q.sweep_sub >= 0) e = B.alloc(&dsw, (size_t)Q.n_sweep);
    float* dhist = nullptr;
    const size_t nhist = (size_t)Q.n_sweep * n_tot * H * 2;
    if (e == hipSuccess && H > 0) {
        e = B.alloc(&dhist, nhist);
        if (e == hipSuccess && Q.hist_in) e = hipMemcpy(dhist, Q.hist_in, sizeof(float) * nhist, hipMemcpyHostToDevice);
        else if (e == hipSuccess) e = hipMemset(dhist, 0, sizeof(float) * nhist);
    }
    bool any_stoch = false;
    for (int q = 0; q < Q.n_subnets; ++q) any_stoch = any_stoch || A.sub[q].stoch;
    float* dnoise = nullptr;
    if (e == hipSuccess && any_stoch) {
        // per subnetwork [n_steps][n][nv] -> [n_steps][n_tot][kNV]
        std::vector<float> zn((size_t)Q.n_steps * n_tot * kNV, 0.0f);
        size_t o = 0;
        for (int q = 0; q < Q.n_subnets; ++q) {
            const Sub& S = A.sub[q];
            for (int st = 0; st < Q.n_steps; ++st)
                for (int r = 0; r < S.n; ++r)
                    for (int v = 0; v < S.nv; ++v, ++o) zn[((size_t)st * n_tot + S.off + r) * kNV + v] = Q.noise[o];
        }
        e = B.alloc(&dnoise, zn.size());
        if (e == hipSuccess) e = hipMemcpy(dnoise, zn.data(), sizeof(float) * zn.size(), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(dx0, x0.data(), sizeof(float) * x0.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && dsw) e = hipMemcpy(dsw, Q.sweep_values, sizeof(float) * Q.n_sweep, hipMemcpyHostToDevice);
    // run_sweep_multi: every point's constants and swept-form values per subnetwork
    float *dpt_mc = nullptr, *dpt_sv = nullptr;
    if (e == hipSuccess && any_point_mode) {
        const size_t nq = (size_t)Q.n_sweep * Q.n_subnets;
        std::vector<float> pmc(nq * kMCX, 0.0f), psv(nq, 0.0f);
        for (int b = 0; b < Q.n_sweep; ++b)
            for (int q = 0; q < Q.n_subnets; ++q) {
                float* row = pmc.data() + ((size_t)b * Q.n_subnets + q) * kMCX;
                const int m = Q.subnets[q].model;
                if (Q.point_mode[q] == 2) {
                    const float* pp = Q.point_params + ((size_t)b * Q.n_subnets + q) * HB_MAX_PARAMS;
                    if (model_ext(m)) model_consts_ext(m, pp, row);
                    else model_consts(m, pp, row);
                } else {
                    std::memcpy(row, A.sub[q].mc, sizeof(float) * kMCX);
                }
                if (Q.point_mode[q] == 1) psv[(size_t)b * Q.n_subnets + q] = Q.point_values[(size_t)b * Q.n_subnets + q];
            }
        e = B.alloc(&dpt_mc, pmc.size());
        if (e == hipSuccess) e = B.alloc(&dpt_sv, psv.size());
        if (e == hipSuccess) e = hipMemcpy(dpt_mc, pmc.data(), sizeof(float) * pmc.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dpt_sv, psv.data(), sizeof(float) * psv.size(), hipMemcpyHostToDevice);
    }
    A.pt_mc = dpt_mc;
    A.pt_sv = dpt_sv;
    if (e != hipSuccess) return hip_fail(e, "multi-subnetwork upload");
    float* dtraj = nullptr;
    const size_t ntraj = (size_t)Q.n_steps * nout;
    if (Q.trajectory) {
        e = B.alloc(&dtraj, ntraj);
        if (e != hipSuccess) return hip_fail(e, "trajectory buffer");
    }
    A.traj = dtraj;
    const BoldShape bs = bold_shape(Q.n_steps, Q.bold_period, Q.dt_ms, Q.bold_tr_ms);
    float* dbold = nullptr;
    const size_t nbold = (size_t)Q.n_sweep * bs.nout * Q.subnets[0].n_nodes;
    // the batch engine's monitor of the sweep mean: var0 of subnetwork 0 is
    // recorded per point and step, then averaged and fed to one monitor below
    const bool bmean = Q.bold_mean != 0 && Q.bold && Q.bold_period > 0;
    const int N0 = Q.subnets[0].n_nodes;
    const BoldShape bsn = bold_shape(Q.n_steps, Q.bold_period, Q.dt_ms, 0.0);
    float *dv0 = nullptr, *dzm = nullptr, *dbw = nullptr, *dnat = nullptr;
    const float** dptr = nullptr;
    if (bmean) {
        if (Q.dt_ms <= 0.0) return bad("BOLD needs dt_ms > 0");
        std::vector<float> bw((size_t)N0 * 4);
        for (int n = 0; n < N0; ++n) {
            for (int c = 0; c < 4; ++c) bw[(size_t)n * 4 + c] = Q.bold_state ? Q.bold_state[(size_t)n * 4 + c] : (c ? 1.0f : 0.0f);
        }
        e = B.alloc(&dv0, (size_t)Q.n_sweep * Q.n_steps * N0);
        if (e == hipSuccess) e = B.alloc(&dzm, (size_t)Q.n_steps * N0);
        if (e == hipSuccess) e = B.alloc(&dbw, (size_t)N0 * 4);
        if (e == hipSuccess) e = B.alloc(&dnat, (size_t)std::max(bsn.nout, 1) * N0);
        if (e == hipSuccess) e = B.alloc(&dptr, (size_t)Q.n_sweep);
        std::vector<const float*> ptrs((size_t)Q.n_sweep);
        for (int b = 0; b < Q.n_sweep; ++b) ptrs[b] = dv0 + (size_t)b * Q.n_steps * N0;
        if (e == hipSuccess) e = hipMemcpy(dptr, ptrs.data(), sizeof(float*) * ptrs.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dbw, bw.data(), sizeof(float) * bw.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "batch-mean BOLD buffers");
        A.v0rec = dv0;
    } else if (Q.bold && Q.bold_period > 0) {
        if (Q.dt_ms <= 0.0) return bad("BOLD needs dt_ms > 0");
        e = B.alloc(&dbold, nbold);
        if (e != hipSuccess) return hip_fail(e, "BOLD buffer");
        A.bold = dbold;
        A.bper = Q.bold_period;
        A.bnout = bs.nout;
        A.bnative = bs.native;
        A.tr_ms = Q.bold_tr_ms;
        A.dtb_ms = bs.dtb;
    }
    A.x0 = dx0;
    A.sweep = dsw;
    A.hist = dhist;
    A.H = H;
    A.noise = dnoise;
    A.plain = Q.plain ? 1 : 0;
    A.dt_ms = Q.dt_ms;
    A.n_stim = 0;
    static const int kStimParams[4] = {2, 3, 3, 4};   // stimulus.rs:31-52
    for (int q = 0; q < Q.n_stimuli; ++q) {
        const hb_stimulus& T0 = Q.stimuli[q];
        if (T0.pattern < 0 || T0.pattern > 3) return bad("Unknown temporal pattern");
        if (T0.n_params < kStimParams[T0.pattern] || T0.n_params > 4)
            return bad("stimulus " + std::to_string(q) + ": too few parameters for its pattern");
        if (T0.target < 0 || T0.target >= Q.n_subnets) continue;   // runtime.rs:320-322 skips it
        Stim& D = A.stim[A.n_stim++];
        D.target = T0.target;
        D.pattern = T0.pattern;
        for (int k = 0; k < 4; ++k) D.p[k] = k < T0.n_params ? T0.params[k] : 0.0f;
    }
    A.tavg = dtavg;
    A.fin = dfin;
    e = hipMemcpy(dA, &A, sizeof(A), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "multi-subnetwork arguments");
    const int threads = std::min(kMaxNodes, (n_tot + 63) / 64 * 64);
    hipEvent_t ev0, ev1;
    (void)hipEventCreate(&ev0);
    (void)hipEventCreate(&ev1);
    (void)hipEventRecord(ev0, nullptr);
    int ext = 0;   // kernel variant: 0 the ten base models, 1 with models 10-27, 2 with twelve state variables
    for (int q = 0; q < Q.n_subnets; ++q) {
        const int m = Q.subnets[q].model;
        if (model_ext(m)) ext = std::max(ext, model_nvar(m) > 8 ? 2 : 1);
    }
    if (ext == 2)
        hipLaunchKernelGGL(k_multi<2>, dim3((unsigned)Q.n_sweep), dim3((unsigned)threads),
                           (size_t)lds * sizeof(float), nullptr, dA);
    else if (ext == 1)
        hipLaunchKernelGGL(k_multi<1>, dim3((unsigned)Q.n_sweep), dim3((unsigned)threads),
                           (size_t)lds * sizeof(float), nullptr, dA);
    else
        hipLaunchKernelGGL(k_multi<0>, dim3((unsigned)Q.n_sweep), dim3((unsigned)threads),
                           (size_t)lds * sizeof(float), nullptr, dA);
    e = hipGetLastError();
    (void)hipEventRecord(ev1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(ev1);
    float ms = 0.0f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, ev0, ev1);
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (e != hipSuccess) return hip_fail(e, "multi-subnetwork kernel");
    spec->kernel_ms = ms;
    if (bmean) {   // mean over the points in order, then the monitor's windows (engine.rs:785-801)
        e = launch_bold_mean_fast(dptr, Q.n_sweep, Q.n_steps, N0, 1, dzm, Q.bold_period, Q.dt_ms, dbw, dnat, nullptr);
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        if (e == hipSuccess && bsn.nout > 0)
            e = hipMemcpy(Q.bold, dnat, sizeof(float) * (size_t)bsn.nout * N0, hipMemcpyDeviceToHost);
        if (e == hipSuccess && Q.bold_state)
            e = hipMemcpy(Q.bold_state, dbw, sizeof(float) * (size_t)N0 * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "batch-mean BOLD");
    }
    std::vector<float> ht(nout), hf(nout);
    e = hipMemcpy(ht.data(), dtavg, sizeof(float) * nout, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hf.data(), dfin, sizeof(float) * nout, hipMemcpyDeviceToHost);
    if (e == hipSuccess && H > 0 && Q.hist_out)
        e = hipMemcpy(Q.hist_out, dhist, sizeof(float) * nhist, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "multi-subnetwork download");
    if (dbold && nbold) {
        e = hipMemcpy(Q.bold, dbold, sizeof(float) * nbold, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "BOLD download");
    }
    if (Q.trajectory && ntraj) {   // per subnetwork [n_steps][n_sweep][n][nv], concatenated
        std::vector<float> tr(ntraj);
        e = hipMemcpy(tr.data(), dtraj, sizeof(float) * ntraj, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "trajectory download");
        size_t o = 0;
        for (int q = 0; q < Q.n_subnets; ++q) {
            const Sub& S = A.sub[q];
            for (int st = 0; st < Q.n_steps; ++st)
                for (int b = 0; b < Q.n_sweep; ++b)
                    for (int r = 0; r < S.n; ++r)
                        for (int v = 0; v < S.nv; ++v, ++o)
                            Q.trajectory[o] = tr[(((size_t)st * Q.n_sweep + b) * n_tot + S.off + r) * kNV + v];
        }
    }
    // per subnetwork, [n_sweep][n_nodes][nvar] (the BatchSweepResult layout), concatenated
    size_t o = 0;
    for (int q = 0; q < Q.n_subnets; ++q) {
        const Sub& S = A.sub[q];
        for (int b = 0; b < Q.n_sweep; ++b)
            for (int r = 0; r < S.n; ++r)
                for (int v = 0; v < S.nv; ++v, ++o) {
                    const size_t src = ((size_t)b * n_tot + S.off + r) * kNV + v;
                    Q.tavg[o] = ht[src];
                    Q.final_state[o] = hf[src];
                }
    }
    return HB_OK;
}

extern "C" int32_t hb_bold_rows(int32_t n_steps, int32_t bold_period, double dt_ms, double tr_ms) {
    if (n_steps < 0 || bold_period <= 0) return 0;
    return hb::bold_shape(n_steps, bold_period, dt_ms, tr_ms).nout;
}
