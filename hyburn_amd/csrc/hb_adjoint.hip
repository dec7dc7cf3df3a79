// hb_adjoint.hip -- discrete adjoint of the batched integration path on gfx950.
//
// Gradients of a per-instance loss with respect to the coupling scale G
// (coupling_params[0]), one model parameter per node and the initial state,
// through the exact discrete recursion the forward kernels run: Euler / Heun /
// RK4 (+ additive noise), clamps, per-edge delayed CSR coupling and, for the
// BOLD losses, the Balloon-Windkessel monitor and Pearson FC.
//
// Kernels, all one workgroup per instance:
//   k_adj_forward   forward integration that also logs, per step, the raw
//                   gathered coupling sums g_t [T][N][np] and writes the state
//                   every S steps (checkpoints), plus window inputs for BOLD.
//   k_bold_forward  Balloon-Windkessel over the window inputs (states kept).
//   k_loss          FC / BOLD-MSE loss and dL/dBOLD.
//   k_bold_reverse  adjoint of the BOLD monitor -> dL/d(window input).
//   k_adj_reverse   for each segment, last first: recompute its states from
//                   the checkpoint and the coupling log (no delay ring and no
//                   barrier needed), then step the adjoint backwards. The
//                   adjoint of the delayed coupling is a gather over each
//                   node's out-edges (CSC) from a ring of coupling adjoints of
//                   later steps: edge (j -> i, delay d) read x_j(s) at step
//                   s + d, so lambda_s[j] += w * lambda_g_{s+d}[i].
//
// Reference semantics differentiated (maedoc/hyburn file:line):
//   time loop, stages, clamps   src/engine/batch_engine/engine.rs:676-773, dfun.rs:170-516
//   per-edge delays             src/engine/runtime.rs:162-204 (d==0 || t==0: current; d>t: 0)
//   coupling pre/post           src/engine/coupling.rs:204-399
//   BOLD                        src/engine/bold.rs:106-166, bold_monitor.rs:72-120
//   FC                          src/sbi/features/fc.rs:11-57
#include "hb_internal.h"
#include "hb_adjoint.h"

// Translation unit: 0..5 = one model's kernels, 6 = shared kernels (see the
// dispatch section at the end).
#ifndef HB_ADJ_UNIT
#define HB_ADJ_UNIT 6
#endif

#include <cmath>
#include <type_traits>

namespace hb {
namespace adj {

#include "hb_model.h"

// ------------------------------------------------------------ clamp masks --
// 1 where clampm<M> passes its argument through, 0 where it pins it.
template <int M> __device__ __forceinline__ void clamp_pass(const float* u, float* m) {
#pragma unroll
    for (int v = 0; v < Tr<M>::NV; ++v) m[v] = 1.0f;
    if constexpr (M == HB_WILSON_COWAN) {
        m[0] = (u[0] < 0.0f || u[0] > 1.0f) ? 0.0f : 1.0f;
        m[1] = (u[1] < 0.0f || u[1] > 1.0f) ? 0.0f : 1.0f;
    } else if constexpr (M == HB_MPR) {
        m[0] = u[0] < 0.0f ? 0.0f : 1.0f;
    } else if constexpr (M == HB_RWW) {
        m[0] = (u[0] < 0.0f || u[0] > 1.0f) ? 0.0f : 1.0f;
    }
}

// ------------------------------------------------------------ dfun VJPs ----
// Given lk = dL/d(dx), accumulate dL/dx into lx, dL/dc into lc, and
// sum_m jac[m] * dL/dmc[m] into dp (jac = d mc / d parameter).
template <int M>
__device__ __forceinline__ void dfun_vjp(const float* mc, const float* x, const float* c, const float* lk,
                                         float* lx, float* lc, const float* jac, float& dp);

template <> __device__ __forceinline__ void dfun_vjp<HB_G2DO>(const float* mc, const float* x, const float* c,
                                                              const float* lk, float* lx, float* lc,
                                                              const float* jac, float& dp) {
    const float v = x[0], w = x[1];
    const float iext = c[0] * mc[8] + mc[9];
    const float v2 = v * v, v3 = v * v2;
    const float A = (((w * mc[6] + iext) - v3 * mc[4]) + v2 * mc[3]) + v * mc[5];
    const float B = ((v * mc[1] + v2 * mc[2]) - w * mc[7]) + mc[0];
    const float a0 = lk[0] * mc[10], a1 = lk[1] * mc[11];
    lx[0] += a0 * ((mc[5] + 2.0f * v * mc[3]) - 3.0f * v2 * mc[4]) + a1 * (mc[1] + 2.0f * v * mc[2]);
    lx[1] += a0 * mc[6] - a1 * mc[7];
    lc[0] += a0 * mc[8];
    dp += jac[0] * a1 + jac[1] * (a1 * v) + jac[2] * (a1 * v2) + jac[3] * (a0 * v2) - jac[4] * (a0 * v3) +
          jac[5] * (a0 * v) + jac[6] * (a0 * w) - jac[7] * (a1 * w) + jac[8] * (a0 * c[0]) + jac[9] * a0 +
          jac[10] * (lk[0] * A) + jac[11] * (lk[1] * B);
}

template <> __device__ __forceinline__ void dfun_vjp<HB_MPR>(const float* mc, const float* x, const float* c,
                                                             const float* lk, float* lx, float* lc,
                                                             const float* jac, float& dp) {
    const float r = x[0], v = x[1];
    const float r2 = r * r;
    const float dv = ((((v * v - r2 * mc[2]) + mc[4]) + r * mc[3]) + c[0] * mc[6]) + c[1] * mc[7];
    const float b0 = lk[0] * mc[0], b1 = lk[1] * mc[0];
    lx[0] += b0 * (2.0f * v) + b1 * (mc[3] - 2.0f * r * mc[2]);
    lx[1] += b0 * (2.0f * r) + b1 * (2.0f * v);
    lc[0] += b1 * mc[6];
    lc[1] += b1 * mc[7];
    dp += jac[0] * (lk[0] * ((r * v) * 2.0f + mc[1]) + lk[1] * dv) + jac[1] * b0 - jac[2] * (b1 * r2) +
          jac[3] * (b1 * r) + jac[4] * b1 + jac[6] * (b1 * c[0]) + jac[7] * (b1 * c[1]);
}

// JR sigmoid mc2 / (exp((mc0 - u) * mc1) + 1) and its partials.
struct Sig { float val, du, dm0, dm1, dm2; };
__device__ __forceinline__ Sig jr_sig(const float* mc, float u) {
    const float e = expf((-u + mc[0]) * mc[1]);
    const float sg = 1.0f / (e + 1.0f);
    const float core = mc[2] * sg * (1.0f - sg);
    Sig s;
    s.val = mc[2] * sg;
    s.du = core * mc[1];
    s.dm0 = -core * mc[1];
    s.dm1 = -core * (mc[0] - u);
    s.dm2 = sg;
    return s;
}

template <> __device__ __forceinline__ void dfun_vjp<HB_JANSEN_RIT>(const float* mc, const float* x, const float* c,
                                                                    const float* lk, float* lx, float* lc,
                                                                    const float* jac, float& dp) {
    const Sig s12 = jr_sig(mc, x[1] - x[2]);
    const Sig s01 = jr_sig(mc, x[0] * mc[6]);
    const Sig s03 = jr_sig(mc, x[0] * mc[8]);
    const float mut = (s01.val * mc[7] + mc[13]) + c[0];
    float g[kMC];
#pragma unroll
    for (int i = 0; i < kMC; ++i) g[i] = 0.0f;
    lx[3] += lk[0]; lx[4] += lk[1]; lx[5] += lk[2];
    // dx3 = s12*mc3 - x3*mc4 - x0*mc5
    const float ls12 = lk[3] * mc[3];
    lx[3] -= lk[3] * mc[4]; lx[0] -= lk[3] * mc[5];
    g[3] += lk[3] * s12.val; g[4] -= lk[3] * x[3]; g[5] -= lk[3] * x[0];
    // dx4 = mut*mc3 - x4*mc4 - x1*mc5
    const float lmut = lk[4] * mc[3];
    lx[4] -= lk[4] * mc[4]; lx[1] -= lk[4] * mc[5];
    g[3] += lk[4] * mut; g[4] -= lk[4] * x[4]; g[5] -= lk[4] * x[1];
    const float ls01 = lmut * mc[7];
    g[7] += lmut * s01.val; g[13] += lmut;
    lc[0] += lmut;
    // dx5 = (s03*mc9)*mc10 - x5*mc11 - x2*mc12
    const float ls03 = lk[5] * mc[9] * mc[10];
    g[9] += lk[5] * s03.val * mc[10]; g[10] += lk[5] * s03.val * mc[9];
    lx[5] -= lk[5] * mc[11]; lx[2] -= lk[5] * mc[12];
    g[11] -= lk[5] * x[5]; g[12] -= lk[5] * x[2];
    // sigmoids
    lx[1] += ls12 * s12.du; lx[2] -= ls12 * s12.du;
    lx[0] += ls01 * s01.du * mc[6]; g[6] += ls01 * s01.du * x[0];
    lx[0] += ls03 * s03.du * mc[8]; g[8] += ls03 * s03.du * x[0];
    g[0] += ls12 * s12.dm0 + ls01 * s01.dm0 + ls03 * s03.dm0;
    g[1] += ls12 * s12.dm1 + ls01 * s01.dm1 + ls03 * s03.dm1;
    g[2] += ls12 * s12.dm2 + ls01 * s01.dm2 + ls03 * s03.dm2;
#pragma unroll
    for (int i = 0; i < 14; ++i) dp += jac[i] * g[i];
}

template <> __device__ __forceinline__ void dfun_vjp<HB_WILSON_COWAN>(const float* mc, const float* x, const float* c,
                                                                      const float* lk, float* lx, float* lc,
                                                                      const float* jac, float& dp) {
    const float e = x[0], i = x[1];
    const float xe0 = ((e * mc[0] - i * mc[1]) - mc[9]) + c[0];
    const float xe = (xe0 + mc[18]) * mc[20];
    const float xi0 = ((e * mc[2] - i * mc[3]) + mc[19]) - mc[13];
    const float xi = xi0 * mc[21];
    const float ee = expf((xe + (-mc[7])) * (-mc[6]));
    const float sge = 1.0f / (ee + 1.0f);
    const float se = (sge - mc[22]) * mc[8];
    const float ei = expf((xi + (-mc[11])) * (-mc[10]));
    const float sgi = 1.0f / (ei + 1.0f);
    const float si = (sgi - mc[23]) * mc[12];
    const float Ae = e * (-mc[14]) + mc[16], Ai = i * (-mc[15]) + mc[17];
    float g[kMC];
#pragma unroll
    for (int k = 0; k < kMC; ++k) g[k] = 0.0f;
    const float a0 = lk[0] * mc[4], a1 = lk[1] * mc[5];
    g[4] += lk[0] * (-e + Ae * se);
    g[5] += lk[1] * (-i + Ai * si);
    lx[0] += -a0 + a0 * (-mc[14]) * se;
    lx[1] += -a1 + a1 * (-mc[15]) * si;
    g[14] -= a0 * e * se; g[16] += a0 * se;
    g[15] -= a1 * i * si; g[17] += a1 * si;
    const float lse = a0 * Ae, lsi = a1 * Ai;
    // se = (sge - mc22) * mc8, sge = 1 / (1 + exp(-(xe - mc7) * mc6))
    const float lsge = lse * mc[8];
    g[8] += lse * (sge - mc[22]); g[22] -= lse * mc[8];
    const float de = lsge * sge * (1.0f - sge);
    const float lxe = de * mc[6];
    g[6] += de * (xe - mc[7]); g[7] -= de * mc[6];
    const float lsgi = lsi * mc[12];
    g[12] += lsi * (sgi - mc[23]); g[23] -= lsi * mc[12];
    const float di = lsgi * sgi * (1.0f - sgi);
    const float lxi = di * mc[10];
    g[10] += di * (xi - mc[11]); g[11] -= di * mc[10];
    // xe = (xe0 + mc18) * mc20
    const float lxe0 = lxe * mc[20];
    g[20] += lxe * (xe0 + mc[18]); g[18] += lxe * mc[20];
    lx[0] += lxe0 * mc[0]; lx[1] -= lxe0 * mc[1];
    g[0] += lxe0 * e; g[1] -= lxe0 * i; g[9] -= lxe0;
    lc[0] += lxe0;
    // xi = xi0 * mc21
    const float lxi0 = lxi * mc[21];
    g[21] += lxi * xi0;
    lx[0] += lxi0 * mc[2]; lx[1] -= lxi0 * mc[3];
    g[2] += lxi0 * e; g[3] -= lxi0 * i; g[19] += lxi0; g[13] -= lxi0;
#pragma unroll
    for (int k = 0; k < 24; ++k) dp += jac[k] * g[k];
}

template <> __device__ __forceinline__ void dfun_vjp<HB_KURAMOTO>(const float* mc, const float* x, const float* c,
                                                                  const float* lk, float* lx, float* lc,
                                                                  const float* jac, float& dp) {
    (void)mc; (void)x; (void)c; (void)lx;
    lc[0] += lk[0];
    dp += jac[0] * lk[0];
}

template <> __device__ __forceinline__ void dfun_vjp<HB_RWW>(const float* mc, const float* x, const float* c,
                                                             const float* lk, float* lx, float* lc,
                                                             const float* jac, float& dp) {
    const float s = x[0];
    const float xx = (s * mc[5] + mc[7]) + c[0] * mc[6];
    const float axb = xx * mc[0] + (-mc[1]);
    const float et = expf(axb * (-mc[2]));
    const float den = -et + 1.0f;
    const float h = axb / den;
    const float dh_daxb = (den - axb * mc[2] * et) / (den * den);
    const float dh_dm2 = -(axb * axb * et) / (den * den);
    const float l = lk[0];
    const float lh = l * (1.0f - s) * mc[3];
    const float laxb = lh * dh_daxb;
    const float lxx = laxb * mc[0];
    lx[0] += l * (-mc[4]) - l * h * mc[3] + lxx * mc[5];
    lc[0] += lxx * mc[6];
    dp += jac[0] * (laxb * xx) - jac[1] * laxb + jac[2] * (lh * dh_dm2) + jac[3] * (l * (1.0f - s) * h) -
          jac[4] * (l * s) + jac[5] * (lxx * s) + jac[6] * (lxx * c[0]) + jac[7] * lxx;
}

// ------------------------------------------------------- coupling VJPs -----
// Coupling functions with an adjoint: Linear (a*g + b), ScaledLinear
// ((g - b) * a) and Kuramoto ((a/n) * (cos x_i * S - sin x_i * C)).
// post_vjp: given mu = dL/dc, return dL/dg (lg), the target's own-state term
// (lx_own, Kuramoto) and dL/d cc[0] (dcc).
template <int NC, int NP>
__device__ __forceinline__ void post_vjp(int cfn, const float* cc, const float* g, const float* own_pre,
                                         const float* mu, float* lg, float* lx_own, float& dcc) {
#pragma unroll
    for (int c = 0; c < NP; ++c) lg[c] = 0.0f;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        lx_own[c] = 0.0f;
        if constexpr (NP == 2 * NC) {
            // Kuramoto: r = (cos_i * gs - sin_i * gc) * cc0
            const float sn = own_pre[c], cs = own_pre[NC + c];
            lg[c] = mu[c] * cc[0] * cs;
            lg[NC + c] = -mu[c] * cc[0] * sn;
            lx_own[c] = mu[c] * cc[0] * (-sn * g[c] - cs * g[NC + c]);
            dcc += mu[c] * (cs * g[c] - sn * g[NC + c]);
        } else if (c < NP) {
            lg[c] = mu[c] * cc[0];
            dcc += cfn == HB_CF_SCALED_LINEAR ? mu[c] * (g[c] - cc[1]) : mu[c] * g[c];
        }
        // c >= NP: a channel whose coefficient is zero (mu[c] == 0), not stored
    }
}

// Source side: dL/dx_j from the gathered coupling adjoints acc[np].
template <int NC, int NP>
__device__ __forceinline__ void pre_vjp(const float* x, const float* acc, float* lx) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        if constexpr (NP == 2 * NC) lx[c] += cosf(x[c]) * acc[c] - sinf(x[c]) * acc[NC + c];
        else if (c < NP) lx[c] += acc[c];
    }
}

// ------------------------------------------------------- step VJP ----------
// Recompute the stages of one step from (x, c, z) and pull lam (dL/dx_{t+1})
// back: lx = dL/dx_t (integrator part), mu = dL/dc_t, dp += parameter term.
// IC >= 0 fixes the integrator class at compile time (-1: from S.integ).
template <int M, int IC = -1>
__device__ __forceinline__ void step_vjp(const StepCtx& S, const float* mc, const float* c, const float* z,
                                         const float* x, const float* lam, float* lx, float* mu,
                                         const float* jac, float& dp) {
    constexpr int NV = Tr<M>::NV, NC = Tr<M>::NC;
    const int ic = IC >= 0 ? IC : integ_class(S.integ);
    float k1[NV], k2[NV], k3[NV], k4[NV], s2[NV], s3[NV], s4[NV], u[NV], m[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) lx[v] = 0.0f;
#pragma unroll
    for (int q = 0; q < NC; ++q) mu[q] = 0.0f;
    dfun<M>(mc, false, 0.0f, x, c, k1);
    if (ic == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) { const float y = x[v] + k1[v] * S.dt; u[v] = S.stoch ? y + z[v] : y; }
        clamp_pass<M>(u, m);
        float lk[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) { const float lu = lam[v] * m[v]; lx[v] += lu; lk[v] = lu * S.dt; }
        dfun_vjp<M>(mc, x, c, lk, lx, mu, jac, dp);
    } else if (ic == 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) { const float y = x[v] + k1[v] * S.dt; s2[v] = S.stoch ? y + z[v] : y; }
        float mp[NV];
        clamp_pass<M>(s2, mp);
        float xp[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) xp[v] = s2[v];
        clampm<M>(xp);
        dfun<M>(mc, false, 0.0f, xp, c, k2);
#pragma unroll
        for (int v = 0; v < NV; ++v) { const float y = x[v] + (k1[v] + k2[v]) * S.dth; u[v] = S.stoch ? y + z[v] : y; }
        clamp_pass<M>(u, m);
        float lk1[NV], lk2[NV], lxp[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const float lu = lam[v] * m[v];
            lx[v] += lu;
            lk1[v] = lu * S.dth;
            lk2[v] = lu * S.dth;
            lxp[v] = 0.0f;
        }
        dfun_vjp<M>(mc, xp, c, lk2, lxp, mu, jac, dp);
#pragma unroll
        for (int v = 0; v < NV; ++v) { const float l1 = lxp[v] * mp[v]; lx[v] += l1; lk1[v] += l1 * S.dt; }
        dfun_vjp<M>(mc, x, c, lk1, lx, mu, jac, dp);
    } else {
        float m2[NV], m3[NV], m4[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) s2[v] = x[v] + k1[v] * S.dth;
        clamp_pass<M>(s2, m2);
        clampm<M>(s2);
        dfun<M>(mc, false, 0.0f, s2, c, k2);
#pragma unroll
        for (int v = 0; v < NV; ++v) s3[v] = x[v] + k2[v] * S.dth;
        clamp_pass<M>(s3, m3);
        clampm<M>(s3);
        dfun<M>(mc, false, 0.0f, s3, c, k3);
#pragma unroll
        for (int v = 0; v < NV; ++v) s4[v] = x[v] + k3[v] * S.dt;
        clamp_pass<M>(s4, m4);
        clampm<M>(s4);
        dfun<M>(mc, false, 0.0f, s4, c, k4);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            float acc = k1[v];
            acc = acc + k2[v] * 2.0f;
            acc = acc + k3[v] * 2.0f;
            const float det = x[v] + (acc + k4[v]) * S.dt6;
            u[v] = (S.stoch && !S.plain) ? det + z[v] : det;
        }
        clamp_pass<M>(u, m);
        float lk[NV], ls[NV];
        // k4 and s4
#pragma unroll
        for (int v = 0; v < NV; ++v) { const float lu = lam[v] * m[v]; lx[v] += lu; u[v] = lu; lk[v] = lu * S.dt6; ls[v] = 0.0f; }
        dfun_vjp<M>(mc, s4, c, lk, ls, mu, jac, dp);
        float lk3[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) { const float l = ls[v] * m4[v]; lx[v] += l; lk3[v] = u[v] * S.dt6 * 2.0f + l * S.dt; ls[v] = 0.0f; }
        dfun_vjp<M>(mc, s3, c, lk3, ls, mu, jac, dp);
        float lk2[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) { const float l = ls[v] * m3[v]; lx[v] += l; lk2[v] = u[v] * S.dt6 * 2.0f + l * S.dth; ls[v] = 0.0f; }
        dfun_vjp<M>(mc, s2, c, lk2, ls, mu, jac, dp);
        float lk1[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) { const float l = ls[v] * m2[v]; lx[v] += l; lk1[v] = u[v] * S.dt6 + l * S.dth; }
        dfun_vjp<M>(mc, x, c, lk1, lx, mu, jac, dp);
    }
}

// ------------------------------------------------------------ helpers ------

__device__ __forceinline__ StepCtx make_ctx(const AdjInst& I) {
    StepCtx S;
    S.dt = I.dt; S.dth = I.dt_half; S.dt6 = I.dt_sixth;
    S.integ = I.integrator;
    S.stoch = I.integrator >= HB_EULER_STOCH;
    S.plain = I.engine == HB_ENGINE_PLAIN;
    S.sw = false; S.sv = 0.0f;
    return S;
}

// One forward step; IC >= 0 fixes the integrator class at compile time.
template <int M, int IC>
__device__ __forceinline__ void integ_step(const StepCtx& S, const float* mc, const float* cpl, const float* z,
                                          float* x) {
    if constexpr (IC >= 0) integrate_c<M, IC>(S, mc, cpl, z, x);
    else integ_step<M, IC>(S, mc, cpl, z, x);
}

__device__ __forceinline__ float block_sum(float v, float* red) {
    // sum over the workgroup, fixed order: wave reductions then wave 0
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    float r = 0.0f;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) r += red[i];
    }
    return r;
}

__device__ __forceinline__ double block_sum_d(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) r += red[i];
    }
    return r;
}

// ------------------------------------------------------- forward save ------
// Same step order as the forward kernels (per-step schedule): ring write of
// pre(x_t), barrier, per-edge gather (slot t-d, t==0 reads current), post,
// stages, clamp. Ring depth dmax + 2.
template <int M, int NP, int NPT, int IC, int CF>
__global__ __launch_bounds__(512) void k_adj_forward(const AdjInst* __restrict__ insts) {
    extern __shared__ float lds[];
    __shared__ float ordp[2][8][2];   // per-wave (sum cos, sum sin) of the last two steps
    const AdjInst& I = insts[blockIdx.x];
    const int cfn = CF >= 0 ? CF : I.cfn;   // coupling function, fixed at compile time when CF >= 0
    constexpr int NV = Tr<M>::NV, NC = Tr<M>::NC;
    const int N = I.n_nodes, T = I.n_steps, Hd = I.ring_depth, Sg = I.seg_len;
    const int tid = threadIdx.x, BS = blockDim.x;
    const int NNP = N * NP;
    // one segment: the forward keeps the whole trajectory and the reverse
    // reads it instead of recomputing
    const bool full_tape = I.n_seg == 1;
    float* ring = I.ring_lds ? lds : I.ring;
    const StepCtx S = make_ctx(I);
    float cc[kCC];
#pragma unroll
    for (int i = 0; i < kCC; ++i) cc[i] = I.cc[i];
    for (int i = tid; i < Hd * NNP; i += BS) ring[i] = 0.0f;

    float x[NPT][NV], mcn[NPT][kMC], bacc[NPT];
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
        const int n = tid + q * BS;
        bacc[q] = 0.0f;
#pragma unroll
        for (int v = 0; v < NV; ++v) x[q][v] = n < N ? I.x0[n * NV + v] : 0.0f;
#pragma unroll
        for (int i = 0; i < kMC; ++i) mcn[q][i] = (I.node_mc && n < N) ? I.node_mc[(size_t)n * kMC + i] : I.mc[i];
        if (n < N) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                I.ckpt[n * NV + v] = x[q][v];
                if (full_tape) I.tape[n * NV + v] = x[q][v];
            }
        }
    }
    const int P = I.bold_period;
    int bcount = 0, kwin = 0;
    int next_ck = Sg > 0 ? Sg : -1;   // the next step count that ends a checkpoint segment
    const bool ord = I.loss_kind == HB_LOSS_ORDER;
    const int nwaves = (BS + 63) >> 6;
    double lsum = 0.0;
    // order parameter R(x_{t+1}) of step t from the wave partials (thread 0)
    auto ord_close = [&](int t) {
        float cs = 0.0f, sn = 0.0f;
        for (int w = 0; w < nwaves; ++w) { cs += ordp[t & 1][w][0]; sn += ordp[t & 1][w][1]; }
        const float C = cs / (float)N, Sn = sn / (float)N;
        const float R = sqrtf(C * C + Sn * Sn);
        const double dd = (double)R - (double)I.target[t];
        lsum += dd * dd;
    };
    __syncthreads();
    for (int t = 0; t < T; ++t) {
        const int ws = t % Hd;
#pragma unroll
        for (int q = 0; q < NPT; ++q) {
            const int n = tid + q * BS;
            if (n < N) {
                float pr[NP];
                pre<NC, NP>(cfn, cc, false, false, x[q], pr);
#pragma unroll
                for (int c = 0; c < NP; ++c) ring[(size_t)ws * NNP + n * NP + c] = pr[c];
            }
        }
        __syncthreads();
        if (ord && tid == 0 && t > 0) ord_close(t - 1);
        ++bcount;
#pragma unroll
        for (int q = 0; q < NPT; ++q) {
            const int n = tid + q * BS;
            if (n >= N) continue;
            float g[NP];
#pragma unroll
            for (int c = 0; c < NP; ++c) g[c] = 0.0f;
            const int e0 = I.indptr[n], e1 = I.indptr[n + 1];
            constexpr int kB = 8;   // edges per batch: records, ring values, then FMAs in row order
            int e = e0;
            for (; e + kB <= e1; e += kB) {
                uint2 ed[kB];
                float val[kB][NP];
#pragma unroll
                for (int k = 0; k < kB; ++k) ed[k] = I.edges[e + k];
#pragma unroll
                for (int k = 0; k < kB; ++k) {
                    const int d = t == 0 ? 0 : (int)(ed[k].x >> 16);
                    int sl = ws - d;
                    sl += sl < 0 ? Hd : 0;
                    const float* src = ring + (size_t)sl * NNP + (ed[k].x & 0xFFFF) * NP;
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
            for (; e < e1; ++e) {
                const uint2 ed = I.edges[e];
                const int j = ed.x & 0xFFFF;
                const int d = t == 0 ? 0 : (int)(ed.x >> 16);
                const float w = __uint_as_float(ed.y);
                int sl = ws - d;
                sl += sl < 0 ? Hd : 0;
                const float* src = ring + (size_t)sl * NNP + j * NP;
#pragma unroll
                for (int c = 0; c < NP; ++c) g[c] += w * src[c];
            }
            float* gl = I.glog + ((size_t)t * N + n) * NP;
#pragma unroll
            for (int c = 0; c < NP; ++c) gl[c] = g[c];
            float own[NP], cpl[NC];
#pragma unroll
            for (int c = 0; c < NP; ++c) own[c] = ring[(size_t)ws * NNP + n * NP + c];
            if (I.has_coupling) post<NC, NP>(cfn, cc, false, g, own, cpl);
            else {
#pragma unroll
                for (int c = 0; c < NC; ++c) cpl[c] = 0.0f;
            }
            float z[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) z[v] = S.stoch ? I.noise[((size_t)t * N + n) * NV + v] * I.nscale[v] : 0.0f;
            integ_step<M, IC>(S, mcn[q], cpl, z, x[q]);
            bacc[q] += x[q][0];
            if (t + 1 == next_ck && t + 1 < T) {   // (t + 1) is a multiple of the segment length
                float* ck = I.ckpt + ((size_t)((t + 1) / Sg) * N + n) * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) ck[v] = x[q][v];
            }
            if (full_tape) {
                float* tp = I.tape + ((size_t)(t + 1) * N + n) * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) tp[v] = x[q][v];
            }
        }
        if (t + 1 == next_ck) next_ck += Sg;
        if (P > 0 && (bcount == P || t == T - 1)) {
#pragma unroll
            for (int q = 0; q < NPT; ++q) {
                const int n = tid + q * BS;
                if (n < N) I.zlog[(size_t)kwin * N + n] = bacc[q] / (float)bcount;
                bacc[q] = 0.0f;
            }
            ++kwin;
            bcount = 0;
        }
        if (ord) {
            float cs = 0.0f, sn = 0.0f;
#pragma unroll
            for (int q = 0; q < NPT; ++q) {
                if (tid + q * BS < N) { cs += cosf(x[q][0]); sn += sinf(x[q][0]); }
            }
            for (int o = 32; o > 0; o >>= 1) { cs += __shfl_down(cs, o, 64); sn += __shfl_down(sn, o, 64); }
            if ((tid & 63) == 0) { ordp[t & 1][tid >> 6][0] = cs; ordp[t & 1][tid >> 6][1] = sn; }
        }
    }
    if (ord) {
        __syncthreads();
        if (tid == 0) {
            ord_close(T - 1);
            I.out_loss[0] = (float)(lsum / (double)T);
        }
    }
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
        const int n = tid + q * BS;
        if (n < N) {
#pragma unroll
            for (int v = 0; v < NV; ++v) I.fin[n * NV + v] = x[q][v];
        }
    }
}

// ------------------------------------------------------- BOLD forward ------
// The adjoint plan evaluates the Balloon-Windkessel monitor in FP64 (the same
// expressions and f32-valued constants as bold_step / bold.rs:106-166): near
// rest v and q sit within a few f32 ulps of 1, where one ulp moves the signal
// by ~2e-6, too coarse to differentiate. Neural inputs stay FP32.
struct BoldD { double s, f, v, q; };
constexpr double kBwKappa = 1.0 / (double)1.54f, kBwGamma = 1.0 / (double)0.50f, kBwTauInv = 1.0 / (double)1.02f;
constexpr double kBwIalpha = 1.0 / (double)0.32f, kBwE0 = (double)0.34f, kBwLo = (double)1e-6f;

// pow(v, 1/alpha) and pow(1 - E0, 1/f) as exp(log(.) * .) (v, f >= 1e-6 > 0):
// a shorter dependent chain than pow, within a few ulps of it.
__device__ __forceinline__ double bw_pv(double vs) { return exp(kBwIalpha * log(vs)); }
__device__ __forceinline__ double bw_ef(double fs, double ln1e) { return exp(ln1e / fs); }

__device__ __forceinline__ double bold_step_d(BoldD& b, double z_raw, double dts, double ln1e) {
    const double z = z_raw > 0.0 ? z_raw : 0.0;
    const double fs = b.f > kBwLo ? b.f : kBwLo;
    const double vs = b.v > kBwLo ? b.v : kBwLo;
    const double ds = (z - kBwKappa * b.s) - kBwGamma * (fs - 1.0);
    const double pv = bw_pv(vs);
    const double dv = kBwTauInv * (fs - pv);
    const double dq = kBwTauInv * ((fs * (1.0 - bw_ef(fs, ln1e))) / kBwE0 - pv * (b.q / vs));
    const double sn = b.s + dts * ds, fn = fs + dts * b.s, vn = vs + dts * dv, qn = b.q + dts * dq;
    b.s = sn > 0.0 ? sn : 0.0;
    b.f = fn > kBwLo ? fn : kBwLo;
    b.v = vn > kBwLo ? vn : kBwLo;
    b.q = qn > 0.0 ? qn : 0.0;
    const double vv = b.v > kBwLo ? b.v : kBwLo;
    return 4.0 * ((7.0 * (1.0 - b.q) + 2.0 * (1.0 - b.q / vv)) + 2.0 * (1.0 - vv));
}

// One window of node n: saves the state entering the window, steps it with the
// window's mean input z and stores the native BOLD sample.
__device__ __forceinline__ void bold_fwd_window(const AdjInst& I, BoldD& b, int k, int n, float z, double ln1e) {
    const int N = I.n_nodes, nw = I.n_win, P = I.bold_period, T = I.n_steps;
    const int cnt = (k == nw - 1) ? T - k * P : P;
    const double dts = (I.dt_ms * (double)cnt) / 1000.0;
    double* st = I.bstate + ((size_t)k * N + n) * 4;
    st[0] = b.s; st[1] = b.f; st[2] = b.v; st[3] = b.q;
    I.bold[(size_t)k * N + n] = bold_step_d(b, (double)z, dts, ln1e);
}

// Each thread runs one node's recursion over the windows; the window inputs of
// the next kBf windows load while the current ones step (the recursion is a
// dependent FP64 chain, so their latency would otherwise add per window).
constexpr int kBf = 8;
#if HB_ADJ_UNIT == 6
__global__ __launch_bounds__(256) void k_bold_forward(const AdjInst* __restrict__ insts) {
    const AdjInst& I = insts[blockIdx.x];
    if (I.bold_period <= 0 || I.bold_fused) return;
    const int N = I.n_nodes, nw = I.n_win;
    const double ln1e = log(1.0 - kBwE0);
    for (int n = threadIdx.x; n < N; n += blockDim.x) {
        BoldD b;
        b.s = 0.0; b.f = 1.0; b.v = 1.0; b.q = 1.0;
        float za[kBf], zb[kBf];
        auto fetch = [&](float (&z)[kBf], int k0) {
#pragma unroll
            for (int m = 0; m < kBf; ++m)
                z[m] = k0 + m < nw ? I.zlog[(size_t)(k0 + m) * N + n] : 0.0f;
        };
        auto run = [&](const float (&z)[kBf], int k0) {
#pragma unroll
            for (int m = 0; m < kBf; ++m) {
                const int k = k0 + m;
                if (k >= nw) break;
                bold_fwd_window(I, b, k, n, z[m], ln1e);
            }
        };
        fetch(za, 0);
        for (int k0 = 0; k0 < nw; k0 += 2 * kBf) {
            fetch(zb, k0 + kBf);
            run(za, k0);
            fetch(za, k0 + 2 * kBf);
            run(zb, k0 + kBf);
        }
    }
}
#endif

// ------------------------------------------------------------- loss --------
// FC (fc.rs:11-57: f64 accumulators, population std, std < 1e-12 -> 1,
// clamp [-1, 1]) of the native BOLD series [nw][N]; loss = sum_ij (FC_ij -
// target_ij)^2 with FC rounded to f32 as the reference returns it. Gradient
// with Z = centred columns / (sqrt(n) s):  dL/dX_ti = (sum_j Gs_ij Z_tj -
// h_i Z_ti) / (sqrt(n) s_i), Gs = G + G^T, G_ij = 2 (r_ij - target_ij) (0 where
// clamped), h_i = sum_j Gs_ij r_ij.
// BOLD MSE: loss = mean (BOLD - target)^2.
// FC path for N <= kFcTile on the FP64 matrix cores (v_mfma_f64_16x16x4f64),
// one workgroup of four waves per instance:
//   covariance C = Zc^T Zc over centred samples staged kFcChunk rows at a time
//   in LDS (lower-triangle 16x16 tiles, dealt round-robin to the waves; the
//   diagonal gives the variances), then r, the loss and Gs = G + G^T in
//   LDS, then dL/dX = (Zs Gs - h .* Zs) / (sqrt(n) s) with Zs = Zc / (sqrt(n) s)
//   as 16-sample x 16-node output tiles of Zs Gs.
// Fragment layout of the 16x16x4 FP64 product (lane l, item r):
//   A[i][k] = a(lane i + 16 k), B[k][j] = b(lane j + 16 k),
//   D[l / 16 + 4 r][l % 16] = d(lane l)[r] (rows interleaved by item).
constexpr int kFcTile = 96, kFcChunk = 32;
constexpr int kFcCs = kFcTile + 1;            // padded chunk row stride (doubles)
constexpr int kFcTiles = 6;                   // lower-triangle tiles per wave: ceil(21 / 4)
constexpr int kFcLds = (kFcTile * kFcTile + kFcChunk * kFcCs + 4 * kFcTile + 4 * kFcTile) * 8;
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kFcStage = kFcChunk * kFcTile / 256;   // chunk elements per thread (256 threads)

__device__ void fc_loss_tiled(const AdjInst& I, double* sh, double* red) {
    const int N = I.n_nodes, nw = I.n_win;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lq = lane >> 4;
    constexpr int nwave = 4;
    const int nt = (N + 15) >> 4, NP = nt * 16;   // node tiles, padded width
    const double* X = I.bold;
    double* gsm = sh;                                  // [NP][NP] (stride kFcTile)
    double* chunk = sh + kFcTile * kFcTile;            // [kFcChunk][kFcCs]
    double* mean = chunk + kFcChunk * kFcCs;           // [kFcTile]
    double* zsc = mean + kFcTile;                      // [kFcTile] 1 / (sqrt(n) s)
    double* hh = zsc + kFcTile;                        // [kFcTile]
    double* rdg = hh + kFcTile;                        // [kFcTile] r_ii
    double* sd = rdg + kFcTile;                        // [kFcTile]
    // This thread's chunk elements: (row, column) packed as row << 8 | column.
    // Each chunk's samples load into registers while the previous chunk is
    // consumed, so the global-load latency overlaps the matrix products.
    int el[kFcStage];
#pragma unroll
    for (int u = 0; u < kFcStage; ++u) {
        const int q = tid + u * 256, k = q / NP, c = q - k * NP;
        el[u] = q < kFcChunk * NP ? (k << 8) | c : -1;
    }
    auto fetch = [&](double (&v)[kFcStage], int k0) {
        const int kc = min(kFcChunk, nw - k0);
#pragma unroll
        for (int u = 0; u < kFcStage; ++u) {
            const int k = el[u] >> 8, c = el[u] & 255;
            v[u] = (el[u] >= 0 && k < kc && c < N) ? X[(size_t)(k0 + k) * N + c] : 0.0;
        }
    };
    // store (v - mean) * scale into the LDS chunk; padding stays zero
    auto stage = [&](const double (&v)[kFcStage], int k0, bool scaled) {
        const int kc = min(kFcChunk, nw - k0);
#pragma unroll
        for (int u = 0; u < kFcStage; ++u) {
            if (el[u] < 0) continue;
            const int k = el[u] >> 8, c = el[u] & 255;
            double z = 0.0;
            if (k < kc && c < N) z = scaled ? (v[u] - mean[c]) * zsc[c] : v[u] - mean[c];
            chunk[k * kFcCs + c] = z;
        }
    };
    double va[kFcStage], vb[kFcStage];
    // Streams the chunks through the two register buffers, two chunks in
    // flight: body(v, k0, refill) consumes chunk k0 from v and calls refill()
    // once v is no longer needed (v then loads the chunk two ahead).
    auto stream = [&](auto&& body) {
        fetch(va, 0);
        if (kFcChunk < nw) fetch(vb, kFcChunk);
        for (int k0 = 0; k0 < nw; k0 += 2 * kFcChunk) {
            body(va, k0, [&]() { if (k0 + 2 * kFcChunk < nw) fetch(va, k0 + 2 * kFcChunk); });
            const int k1 = k0 + kFcChunk;
            if (k1 >= nw) break;
            body(vb, k1, [&]() { if (k1 + 2 * kFcChunk < nw) fetch(vb, k1 + 2 * kFcChunk); });
        }
    };
    // ---- means: per-element partial sums, then per-column sums over the rows
    {
        double ps[kFcStage];
#pragma unroll
        for (int u = 0; u < kFcStage; ++u) ps[u] = 0.0;
        stream([&](double (&v)[kFcStage], int, auto&& refill) {
#pragma unroll
            for (int u = 0; u < kFcStage; ++u) ps[u] += v[u];
            refill();
        });
#pragma unroll
        for (int u = 0; u < kFcStage; ++u)
            if (el[u] >= 0) chunk[(el[u] >> 8) * kFcCs + (el[u] & 255)] = ps[u];
        __syncthreads();
        if (tid < N) {
            double a = 0.0;
            for (int k = 0; k < kFcChunk; ++k) a += chunk[k * kFcCs + tid];
            mean[tid] = a / (double)nw;
        }
    }
    // ---- covariance tiles (ti >= tj) of this wave
    const int ntt = nt * (nt + 1) / 2;
    int tti[kFcTiles], ttj[kFcTiles];
    d4 acc[kFcTiles];
#pragma unroll
    for (int q = 0; q < kFcTiles; ++q) {
        const int t = wave + q * nwave;
        int ti = 0, tj = 0;
        if (t < ntt) {
            while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
            tj = t - ti * (ti + 1) / 2;
        }
        tti[q] = ti; ttj[q] = tj;
        acc[q] = d4{0.0, 0.0, 0.0, 0.0};
    }
    stream([&](double (&v)[kFcStage], int k0, auto&& refill) {
        __syncthreads();
        stage(v, k0, false);
        __syncthreads();
        refill();
#pragma unroll
        for (int kk = 0; kk < kFcChunk; kk += 4) {
            const double* row = chunk + (kk + lq) * kFcCs + lr;
#pragma unroll
            for (int q = 0; q < kFcTiles; ++q)
                if (wave + q * nwave < ntt) acc[q] = mfma_f64(row[tti[q] * 16], row[ttj[q] * 16], acc[q]);
        }
    });
    // variances from the diagonal tiles
    const double sqn = sqrt((double)nw);
#pragma unroll
    for (int q = 0; q < kFcTiles; ++q) {
        if (wave + q * nwave < ntt && tti[q] == ttj[q]) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = tti[q] * 16 + lq + 4 * r, j = ttj[q] * 16 + lr;
                if (i == j && i < N) {
                    double sdev = sqrt(acc[q][r] / (double)nw);
                    sdev = sdev < 1e-12 ? 1.0 : sdev;
                    sd[i] = sdev;
                    zsc[i] = 1.0 / (sqn * sdev);
                }
            }
        }
    }
    __syncthreads();
    // ---- r, loss, Gs: for j < i, gsm[j][i] = Gs_ij (upper) and gsm[i][j] = r_ij
    // (lower); gsm[i][i] = Gs_ii, rdg[i] = r_ii; zero outside N x N
    double lacc = 0.0;
#pragma unroll
    for (int q = 0; q < kFcTiles; ++q) {
        if (wave + q * nwave >= ntt) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = tti[q] * 16 + lq + 4 * r, j = ttj[q] * 16 + lr;
            if (j > i) continue;
            double g = 0.0, rv = 0.0;
            if (i < N && j < N) {
                const double raw = (acc[q][r] / (double)nw) / (sd[i] * sd[j]);
                const bool clamped = raw < -1.0 || raw > 1.0;
                rv = clamped ? (raw < -1.0 ? -1.0 : 1.0) : raw;
                const double rf = (double)(float)rv;
                const double tij = I.target[(size_t)i * N + j], tji = I.target[(size_t)j * N + i];
                g = clamped ? 0.0 : 2.0 * (rf - tij) + 2.0 * (rf - tji);
                lacc += (rf - tij) * (rf - tij);
                if (i != j) lacc += (rf - tji) * (rf - tji);
            }
            if (i == j) { gsm[i * kFcTile + i] = g; rdg[i] = rv; }
            else { gsm[j * kFcTile + i] = g; gsm[i * kFcTile + j] = rv; }
        }
    }
    const double tot = block_sum_d(lacc, red);
    if (tid == 0) I.out_loss[0] = (float)tot;
    __syncthreads();
    // ---- h_i = sum_j Gs_ij r_ij: each off-diagonal product is upper x lower
    if (tid < N) {
        double a = gsm[tid * kFcTile + tid] * rdg[tid];
        for (int j = 0; j < N; ++j)
            if (j != tid) a += gsm[tid * kFcTile + j] * gsm[j * kFcTile + tid];
        hh[tid] = a;
    }
    __syncthreads();
    for (int q = tid; q < NP * NP; q += 256) {   // lower <- upper: Gs symmetric
        const int i = q / NP, j = q - i * NP;
        if (j < i) gsm[i * kFcTile + j] = gsm[j * kFcTile + i];
    }
    // ---- dL/dX per chunk: output tiles (16 samples) x (16 nodes); a wave's
    // tiles advance together so their matrix products are independent
    constexpr int kDt = (kFcChunk / 16) * (kFcTile / 16) / 4;   // output tiles per wave at most
    const int ntile = (kFcChunk / 16) * nt;
    const double* arow[kDt];
    const double* bcol[kDt];
    int oti[kDt];
#pragma unroll
    for (int q = 0; q < kDt; ++q) {
        const int t = wave + q * nwave;
        const int ts = t < ntile ? t / nt : 0, tn = t < ntile ? t - ts * nt : 0;
        oti[q] = t < ntile ? (ts << 8) | tn : -1;
        arow[q] = chunk + (ts * 16 + lr) * kFcCs + lq;        // A[s][j] = Zs[s][j]
        bcol[q] = gsm + lq * kFcTile + tn * 16 + lr;           // B[j][i] = Gs[j][i]
    }
    stream([&](double (&v)[kFcStage], int k0, auto&& refill) {
        const int kc = min(kFcChunk, nw - k0);
        __syncthreads();
        stage(v, k0, true);
        __syncthreads();
        refill();
        d4 d[kDt];
#pragma unroll
        for (int q = 0; q < kDt; ++q) d[q] = d4{0.0, 0.0, 0.0, 0.0};
        for (int j0 = 0; j0 < NP; j0 += 4) {
#pragma unroll
            for (int q = 0; q < kDt; ++q)
                if (oti[q] >= 0) d[q] = mfma_f64(arow[q][j0], bcol[q][j0 * kFcTile], d[q]);
        }
#pragma unroll
        for (int q = 0; q < kDt; ++q) {
            if (oti[q] < 0) continue;
            const int ts = oti[q] >> 8, tn = oti[q] & 255;
            const int i = tn * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int sm = ts * 16 + lq + 4 * r;
                if (sm < kc && i < N)
                    I.dsig[(size_t)(k0 + sm) * N + i] = (d[q][r] - hh[i] * chunk[sm * kFcCs + i]) * zsc[i];
            }
        }
    });
}

#if HB_ADJ_UNIT == 6
__global__ __launch_bounds__(256) void k_loss(const AdjInst* __restrict__ insts) {
    extern __shared__ double shd[];
    __shared__ double red[8];
    const AdjInst& I = insts[blockIdx.x];
    if (I.bold_period <= 0 || I.loss_kind == HB_LOSS_ORDER) return;
    const int N = I.n_nodes, nw = I.n_win;
    const int tid = threadIdx.x, BS = blockDim.x;
    const double* X = I.bold;
    if (I.loss_kind == HB_LOSS_FC && N <= kFcTile) {
        fc_loss_tiled(I, shd, red);
        return;
    }
    if (I.loss_kind == HB_LOSS_BOLD_MSE) {
        const double scale = 1.0 / ((double)nw * (double)N);
        double acc = 0.0;
        for (size_t i = tid; i < (size_t)nw * N; i += BS) {
            const double dd = X[i] - (double)I.target[i];
            acc += dd * dd;
            I.dsig[i] = 2.0 * dd * scale;
        }
        const double tot = block_sum_d(acc, red);
        if (tid == 0) I.out_loss[0] = (float)(tot * scale);
        return;
    }
    double* mean = I.fc_work;                  // [N]
    double* sd = mean + N;                     // [N]
    double* rr = sd + N;                       // [N][N] r (f64)
    double* gs = rr + (size_t)N * N;           // [N][N] G + G^T
    double* h = gs + (size_t)N * N;            // [N]
    for (int n = tid; n < N; n += BS) {
        double m = 0.0;
        for (int k = 0; k < nw; ++k) m += (double)X[(size_t)k * N + n];
        m /= (double)nw;
        double s = 0.0;
        for (int k = 0; k < nw; ++k) { const double d = (double)X[(size_t)k * N + n] - m; s += d * d; }
        s = sqrt(s / (double)nw);
        if (s < 1e-12) s = 1.0;
        mean[n] = m;
        sd[n] = s;
    }
    __syncthreads();
    const int npairs = N * (N + 1) / 2;
    double lacc = 0.0;
    for (int p = tid; p < npairs; p += BS) {
        // p -> (i, j), j <= i
        int i = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
        while (i * (i + 1) / 2 > p) --i;
        while ((i + 1) * (i + 2) / 2 <= p) ++i;
        const int j = p - i * (i + 1) / 2;
        double cov = 0.0;
        const double mi = mean[i], mj = mean[j];
        for (int k = 0; k < nw; ++k) cov += ((double)X[(size_t)k * N + i] - mi) * ((double)X[(size_t)k * N + j] - mj);
        cov /= (double)nw;
        const double raw = cov / (sd[i] * sd[j]);
        const bool clamped = raw < -1.0 || raw > 1.0;
        const double r = clamped ? (raw < -1.0 ? -1.0 : 1.0) : raw;
        const double rf = (double)(float)r;
        const double tij = I.target[(size_t)i * N + j], tji = I.target[(size_t)j * N + i];
        const double gij = clamped ? 0.0 : 2.0 * (rf - tij);
        const double gji = clamped ? 0.0 : 2.0 * (rf - tji);
        rr[(size_t)i * N + j] = r;
        rr[(size_t)j * N + i] = r;
        gs[(size_t)i * N + j] = gij + gji;
        gs[(size_t)j * N + i] = gij + gji;
        lacc += (rf - tij) * (rf - tij);
        if (i != j) lacc += (rf - tji) * (rf - tji);
    }
    const double tot = block_sum_d(lacc, red);
    if (tid == 0) I.out_loss[0] = (float)tot;
    __syncthreads();
    for (int n = tid; n < N; n += BS) {
        double a = 0.0;
        for (int j = 0; j < N; ++j) a += gs[(size_t)n * N + j] * rr[(size_t)n * N + j];
        h[n] = a;
    }
    __syncthreads();
    const double sqn = sqrt((double)nw);
    for (size_t idx = tid; idx < (size_t)nw * N; idx += BS) {
        const int k = (int)(idx / N), i = (int)(idx % N);
        const double* row = X + (size_t)k * N;
        double a = 0.0;
        for (int j = 0; j < N; ++j) a += gs[(size_t)i * N + j] * ((row[j] - mean[j]) / (sqn * sd[j]));
        const double zi = (row[i] - mean[i]) / (sqn * sd[i]);
        I.dsig[idx] = (a - h[i] * zi) / (sqn * sd[i]);
    }
}
#endif

// ------------------------------------------------------- BOLD reverse ------
// Adjoint of bold_step_d per node, windows last to first (FP64). For window k
// the adjoint state a = (ls, lf, lv, lq) maps affinely,
//   a <- B_k D_k (a + g_k),  D_k = clamp masks, g_k = dL/dBOLD_k terms,
// with B_k, D_k, g_k functions of the saved forward state only, and the seed
// dL/d(window input) is linear in the incoming ls. The kernel splits each
// node's windows into C chunks (one thread per node and chunk): every chunk
// composes its affine map (4x4 matrix and offset), each thread combines the
// maps of the chunks after its own to get its incoming state, then reruns its
// chunk from that state and writes the seeds.
constexpr int kBrChunks = 16, kBrMinChunk = 16;
struct BoldRevIn { double s, f, v, q, dsg; float z; };
struct BoldCoef {
    bool ms, mf, mv, mq, mfo, mvo;               // clamp masks
    double gv, gq;                               // dL/dBOLD terms added to lv, lq
    double css, csf, cfs, cfv, cfq, cvv, cvq, cqq, cz;
};
__device__ __forceinline__ BoldRevIn bold_rev_load(const AdjInst& I, int k, int n) {
    BoldRevIn b;
    const int N = I.n_nodes;
    const double4 st = *(const double4*)(I.bstate + ((size_t)k * N + n) * 4);
    b.s = st.x; b.f = st.y; b.v = st.z; b.q = st.w;
    b.z = I.zlog[(size_t)k * N + n];
    b.dsg = I.dsig[(size_t)k * N + n];
    return b;
}
__device__ __forceinline__ BoldCoef bold_rev_coef(const AdjInst& I, int k, const BoldRevIn& in, double ln1e) {
    const int nw = I.n_win, P = I.bold_period, T = I.n_steps;
    const int cnt = (k == nw - 1) ? T - k * P : P;
    const double dts = (I.dt_ms * (double)cnt) / 1000.0;
    const double s = in.s, f = in.f, v = in.v, q = in.q;
    const double zr = (double)in.z;
    const double zz = zr > 0.0 ? zr : 0.0;
    const double fs = f > kBwLo ? f : kBwLo;
    const double vs = v > kBwLo ? v : kBwLo;
    const double ds = (zz - kBwKappa * s) - kBwGamma * (fs - 1.0);
    const double pv = bw_pv(vs);
    const double dv = kBwTauInv * (fs - pv);
    const double ef = bw_ef(fs, ln1e);
    const double dq = kBwTauInv * ((fs * (1.0 - ef)) / kBwE0 - pv * (q / vs));
    const double sn = s + dts * ds, fn = fs + dts * s, vn = vs + dts * dv, qn = q + dts * dq;
    const double v1 = vn > kBwLo ? vn : kBwLo;
    const double q1 = qn > 0.0 ? qn : 0.0;
    BoldCoef c;
    // signal 4 * (7 (1 - q') + 2 (1 - q'/v') + 2 (1 - v'))
    c.gq = in.dsg * 4.0 * (-7.0 - 2.0 / v1);
    c.gv = (v1 > kBwLo) ? in.dsg * 4.0 * (2.0 * q1 / (v1 * v1) - 2.0) : 0.0;
    c.ms = sn > 0.0; c.mf = fn > kBwLo; c.mv = vn > kBwLo; c.mq = qn > 0.0;
    c.mfo = f > kBwLo; c.mvo = v > kBwLo;
    const double dfe = (1.0 - ef) + ef * ln1e / fs;            // d(fs (1 - ef)) / dfs
    const double dpv = kBwIalpha * pv / vs;
    const double dpvq = (kBwIalpha - 1.0) * pv / (vs * vs);    // d(pv / vs) / dvs
    c.css = 1.0 - dts * kBwKappa;
    c.csf = dts;
    c.cfs = -kBwGamma * dts;
    c.cfv = kBwTauInv * dts;
    c.cfq = kBwTauInv * dts * (dfe / kBwE0);
    c.cvv = 1.0 - dts * kBwTauInv * dpv;
    c.cvq = -(dts * kBwTauInv * q * dpvq);
    c.cqq = 1.0 - dts * kBwTauInv * (pv / vs);
    c.cz = zr > 0.0 ? dts / (double)cnt : 0.0;
    return c;
}
// a <- B D (a + g) (offset) or B D a (linear part)
__device__ __forceinline__ void bold_rev_apply(const BoldCoef& c, double (&a)[4], bool offset) {
    const double lsn = c.ms ? a[0] : 0.0;
    const double lfn = c.mf ? a[1] : 0.0;
    const double lvn = c.mv ? (offset ? a[2] + c.gv : a[2]) : 0.0;
    const double lqn = c.mq ? (offset ? a[3] + c.gq : a[3]) : 0.0;
    a[0] = c.css * lsn + c.csf * lfn;
    a[1] = c.mfo ? c.cfs * lsn + lfn + c.cfv * lvn + c.cfq * lqn : 0.0;
    a[2] = c.mvo ? c.cvv * lvn + c.cvq * lqn : 0.0;
    a[3] = c.cqq * lqn;
}
// One window with its seed: dL/d(window input) / count = cz * (masked ls).
__device__ __forceinline__ void bold_rev_window(const AdjInst& I, double (&a)[4], int k, int n, const BoldRevIn& in,
                                                double ln1e) {
    const BoldCoef c = bold_rev_coef(I, k, in, ln1e);
    I.seedw[(size_t)k * I.n_nodes + n] = (float)(c.cz * (c.ms ? a[0] : 0.0));
    bold_rev_apply(c, a, true);
}

#if HB_ADJ_UNIT == 6
__global__ __launch_bounds__(1024) void k_bold_reverse(const AdjInst* __restrict__ insts) {
    const AdjInst& I = insts[blockIdx.x];
    if (I.bold_period <= 0 || I.loss_kind == HB_LOSS_ORDER) return;
    const int N = I.n_nodes, nw = I.n_win, BS = blockDim.x;
    const double ln1e = log(1.0 - kBwE0);
    const int C = max(1, min(min(kBrChunks, BS / N), nw / kBrMinChunk));
    // one window of look-ahead: the next window's inputs load while the
    // current one is differentiated
    auto sweep = [&](double (&a)[4], int n, int khi, int klo) {
        BoldRevIn nx = bold_rev_load(I, khi, n);
        for (int k = khi; k >= klo; --k) {
            const BoldRevIn cu = nx;
            if (k > klo) nx = bold_rev_load(I, k - 1, n);
            bold_rev_window(I, a, k, n, cu, ln1e);
        }
    };
    if (C == 1) {   // whole nodes per thread
        for (int n = threadIdx.x; n < N; n += BS) {
            double a[4] = {0.0, 0.0, 0.0, 0.0};
            sweep(a, n, nw - 1, 0);
        }
        return;
    }
    const int Lc = (nw + C - 1) / C;
    const int n = threadIdx.x % N, c = threadIdx.x / N;
    const bool act = c < C;
    const int klo = c * Lc, khi = min(nw, klo + Lc) - 1;
    double* map = I.bscan;   // [C][N][20]: columns of the 4x4 matrix, then the offset
    // ---- chunk maps (chunk 0's is never needed); one window of look-ahead
    if (act && c > 0 && klo <= khi) {
        double M[4][4], o[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) M[j][r] = j == r ? 1.0 : 0.0;
        BoldRevIn nx = bold_rev_load(I, khi, n);
        for (int k = khi; k >= klo; --k) {
            const BoldRevIn cu = nx;
            if (k > klo) nx = bold_rev_load(I, k - 1, n);
            const BoldCoef cf = bold_rev_coef(I, k, cu, ln1e);
#pragma unroll
            for (int j = 0; j < 4; ++j) bold_rev_apply(cf, M[j], false);
            bold_rev_apply(cf, o, true);
        }
        double* dst = map + ((size_t)c * N + n) * 20;
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[j * 4 + r] = M[j][r];
#pragma unroll
        for (int r = 0; r < 4; ++r) dst[16 + r] = o[r];
    }
    __syncthreads();
    if (!act || klo > khi) return;
    // ---- incoming state: the maps of chunks C-1 .. c+1 applied to zero
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int cc = C - 1; cc > c; --cc) {
        const double* src = map + ((size_t)cc * N + n) * 20;
        double na[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            na[r] = src[16 + r] + src[r] * a[0] + src[4 + r] * a[1] + src[8 + r] * a[2] + src[12 + r] * a[3];
#pragma unroll
        for (int r = 0; r < 4; ++r) a[r] = na[r];
    }
    // ---- rerun the chunk from its incoming state, writing the seeds
    sweep(a, n, khi, klo);
}
#endif

// ------------------------------------------------------- reverse ----------
template <int M, int NP, int NPT, int IC, int CF>
__global__ __launch_bounds__(512) void k_adj_reverse(const AdjInst* __restrict__ insts) {
    extern __shared__ float lds[];
    __shared__ double redd[8];
    const AdjInst& I = insts[blockIdx.x];
    const int cfn = CF >= 0 ? CF : I.cfn;   // coupling function, fixed at compile time when CF >= 0
    constexpr int NV = Tr<M>::NV, NC = Tr<M>::NC;
    const int N = I.n_nodes, T = I.n_steps, Ha = I.aring_depth, Sg = I.seg_len, nseg = I.n_seg;
    const int tid = threadIdx.x, BS = blockDim.x;
    const int NNP = N * NP;
    float* aring = I.aring_lds ? lds : I.ring;
    const StepCtx S = make_ctx(I);
    float cc[kCC];
#pragma unroll
    for (int i = 0; i < kCC; ++i) cc[i] = I.cc[i];
    for (int i = tid; i < Ha * NNP; i += BS) aring[i] = 0.0f;

    float lam[NPT][NV], mcn[NPT][kMC], jac[NPT][kMC];
    // parameter gradients are sums over every step and stage: FP64 accumulators
    double gp[NPT];
    double gcc = 0.0;
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
        const int n = tid + q * BS;
        gp[q] = 0.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) lam[q][v] = 0.0f;
#pragma unroll
        for (int i = 0; i < kMC; ++i) {
            mcn[q][i] = (I.node_mc && n < N) ? I.node_mc[(size_t)n * kMC + i] : I.mc[i];
            jac[q][i] = (I.node_jac && n < N) ? I.node_jac[(size_t)n * kMC + i] : 0.0f;
        }
    }
    const int P = I.bold_period;
    const bool seed_win = P > 0 && (I.loss_kind == HB_LOSS_FC || I.loss_kind == HB_LOSS_BOLD_MSE);
    const bool seed_ord = I.loss_kind == HB_LOSS_ORDER;
    __syncthreads();

    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int t0 = seg * Sg, t1 = min(T, t0 + Sg);
        // ---- recompute states t0..t1 from the checkpoint and the coupling log
        // (with one segment the forward wrote the whole trajectory). Rolled
        // step loop: slot 0 of `cur` is the current step's log row, the window
        // shifts down each step and `nxt` (the following kG rows) loads kG
        // steps ahead of use.
#pragma unroll
        for (int q = 0; q < NPT; ++q) {
            const int n = tid + q * BS;
            if (n >= N || nseg == 1) continue;
            float x[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) { x[v] = I.ckpt[((size_t)seg * N + n) * NV + v]; I.tape[(size_t)n * NV + v] = x[v]; }
            constexpr int kG = 8;
            float cur[kG][NP], nxt[kG][NP];
            auto fetch = [&](float (&gq)[kG][NP], int tg) {
#pragma unroll
                for (int m = 0; m < kG; ++m) {
                    if (tg + m < t1) {
                        const float* gl = I.glog + ((size_t)(tg + m) * N + n) * NP;
#pragma unroll
                        for (int c = 0; c < NP; ++c) gq[m][c] = gl[c];
                    }
                }
            };
            fetch(cur, t0);
            fetch(nxt, t0 + kG);
#pragma nounroll
            for (int t = t0; t < t1; ++t) {
                float own[NP], cpl[NC], z[NV];
                pre<NC, NP>(cfn, cc, false, false, x, own);
                if (I.has_coupling) post<NC, NP>(cfn, cc, false, cur[0], own, cpl);
                else {
#pragma unroll
                    for (int c = 0; c < NC; ++c) cpl[c] = 0.0f;
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) z[v] = S.stoch ? I.noise[((size_t)t * N + n) * NV + v] * I.nscale[v] : 0.0f;
                integ_step<M, IC>(S, mcn[q], cpl, z, x);
                float* tp = I.tape + ((size_t)(t - t0 + 1) * N + n) * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) tp[v] = x[v];
#pragma unroll
                for (int m = 0; m + 1 < kG; ++m)
#pragma unroll
                    for (int c = 0; c < NP; ++c) cur[m][c] = cur[m + 1][c];
                if (((t - t0) & (kG - 1)) == kG - 1) {
#pragma unroll
                    for (int m = 0; m < kG; ++m)
#pragma unroll
                        for (int c = 0; c < NP; ++c) cur[m][c] = nxt[m][c];
                    fetch(nxt, t + 1 + kG);
                }
            }
        }
        __syncthreads();
        // ---- order parameter of states t0..t1: per row (wave-strided)
        if (seed_ord) {
            const int w = tid >> 6, l = tid & 63, nwv = BS >> 6;
            for (int r = w; r <= t1 - t0; r += nwv) {
                float cs = 0.0f, sn = 0.0f;
                for (int n = l; n < N; n += 64) {
                    const float th = I.tape[((size_t)r * N + n) * NV];
                    cs += cosf(th);
                    sn += sinf(th);
                }
                for (int o = 32; o > 0; o >>= 1) { cs += __shfl_down(cs, o, 64); sn += __shfl_down(sn, o, 64); }
                if (l == 0) {
                    const int s = t0 + r;
                    float a = 0.0f, b = 0.0f;
                    if (s >= 1) {
                        const float C = cs / (float)N, Sn = sn / (float)N;
                        const float R = sqrtf(C * C + Sn * Sn);
                        const float diff = R - I.target[s - 1];
                        const float al = R > 0.0f ? 2.0f * diff / ((float)T * (float)N * R) : 0.0f;
                        a = -al * C;
                        b = al * Sn;
                    }
                    I.ord_work[(size_t)r * 4 + 0] = a;
                    I.ord_work[(size_t)r * 4 + 1] = b;
                }
            }
            __syncthreads();
        }
        // ---- seed of the final state x_T
        if (seg == nseg - 1) {
#pragma unroll
            for (int q = 0; q < NPT; ++q) {
                const int n = tid + q * BS;
                if (n >= N) continue;
                if (seed_win) lam[q][0] += I.seedw[(size_t)((T - 1) / P) * N + n];
                if (seed_ord) {
                    const float th = I.tape[((size_t)(t1 - t0) * N + n) * NV];
                    const float* cf = I.ord_work + (size_t)(t1 - t0) * 4;
                    lam[q][0] += cf[0] * sinf(th) + cf[1] * cosf(th);
                }
            }
        }
        // ---- reverse steps t1-1 .. t0; the adjoint ring slot of t and the BOLD
        // window of t, (t - 1) / P, are stepped down with t
        int slot = (t1 - 1) % Ha, win_t = 0, wr_t = 0;
        if (seed_win && t1 >= 2) { win_t = (t1 - 2) / P; wr_t = (t1 - 2) - win_t * P; }
        for (int t = t1 - 1; t >= t0;
             --t, slot = slot == 0 ? Ha - 1 : slot - 1, wr_t = seed_win ? wr_t - 1 : 0) {
            if (wr_t < 0) { wr_t = P - 1; --win_t; }
            float lxp[NPT][NV];
#pragma unroll
            for (int q = 0; q < NPT; ++q) {
                const int n = tid + q * BS;
                if (n >= N) continue;
                float x[NV], g[NP], own[NP], cpl[NC], z[NV], mu[NC], lg[NP], lown[NC];
                const float* tp = I.tape + ((size_t)(t - t0) * N + n) * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) x[v] = tp[v];
                const float* gl = I.glog + ((size_t)t * N + n) * NP;
#pragma unroll
                for (int c = 0; c < NP; ++c) g[c] = gl[c];
                pre<NC, NP>(cfn, cc, false, false, x, own);
                if (I.has_coupling) post<NC, NP>(cfn, cc, false, g, own, cpl);
                else {
#pragma unroll
                    for (int c = 0; c < NC; ++c) cpl[c] = 0.0f;
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) z[v] = S.stoch ? I.noise[((size_t)t * N + n) * NV + v] * I.nscale[v] : 0.0f;
                float dps = 0.0f, dcs = 0.0f;
                step_vjp<M, IC>(S, mcn[q], cpl, z, x, lam[q], lxp[q], mu, jac[q], dps);
                gp[q] += (double)dps;
                if (I.has_coupling) {
                    post_vjp<NC, NP>(cfn, cc, g, own, mu, lg, lown, dcs);
                    gcc += (double)dcs;
#pragma unroll
                    for (int c = 0; c < NC; ++c) lxp[q][c] += lown[c];
                } else {
#pragma unroll
                    for (int c = 0; c < NP; ++c) lg[c] = 0.0f;
                }
                float* ar = aring + (size_t)slot * NNP + n * NP;
#pragma unroll
                for (int c = 0; c < NP; ++c) ar[c] = lg[c];
            }
            __syncthreads();
            // gather over out-edges: d == 0 -> this step; d >= 1 -> step t + d;
            // at t == 0 every edge also read the current state.
#pragma unroll
            for (int q = 0; q < NPT; ++q) {
                const int n = tid + q * BS;
                if (n >= N) continue;
                float acc[NP];
#pragma unroll
                for (int c = 0; c < NP; ++c) acc[c] = 0.0f;
                if (I.has_coupling) {
                    const int e0 = I.indptr_t[n], e1 = I.indptr_t[n + 1];
                    int e = e0;
                    if (t > 0) {
                        constexpr int kB = 8;   // batched as in the forward gather
                        for (; e + kB <= e1; e += kB) {
                            uint2 ed[kB];
                            float val[kB][NP];
#pragma unroll
                            for (int k = 0; k < kB; ++k) ed[k] = I.edges_t[e + k];
#pragma unroll
                            for (int k = 0; k < kB; ++k) {
                                int sl = slot + (int)(ed[k].x >> 16);
                                sl -= sl >= Ha ? Ha : 0;
                                const float* a = aring + (size_t)sl * NNP + (ed[k].x & 0xFFFF) * NP;
#pragma unroll
                                for (int c = 0; c < NP; ++c) val[k][c] = a[c];
                            }
#pragma unroll
                            for (int k = 0; k < kB; ++k) {
                                const float w = __uint_as_float(ed[k].y);
#pragma unroll
                                for (int c = 0; c < NP; ++c) acc[c] += w * val[k][c];
                            }
                        }
                    }
                    for (; e < e1; ++e) {
                        const uint2 ed = I.edges_t[e];
                        const int i = ed.x & 0xFFFF;
                        const int d = (int)(ed.x >> 16);
                        const float w = __uint_as_float(ed.y);
                        int sl = slot + d;
                        sl -= sl >= Ha ? Ha : 0;
                        const float* a = aring + (size_t)sl * NNP + i * NP;
#pragma unroll
                        for (int c = 0; c < NP; ++c) acc[c] += w * a[c];
                        if (t == 0 && d > 0) {
                            const float* a0 = aring + (size_t)slot * NNP + i * NP;
#pragma unroll
                            for (int c = 0; c < NP; ++c) acc[c] += w * a0[c];
                        }
                    }
                }
                float x[NV];
                const float* tp = I.tape + ((size_t)(t - t0) * N + n) * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) x[v] = tp[v];
                pre_vjp<NC, NP>(x, acc, lxp[q]);
#pragma unroll
                for (int v = 0; v < NV; ++v) lam[q][v] = lxp[q][v];
                if (t >= 1) {
                    if (seed_win) lam[q][0] += I.seedw[(size_t)win_t * N + n];
                    if (seed_ord) {
                        const float* cf = I.ord_work + (size_t)(t - t0) * 4;
                        lam[q][0] += cf[0] * sinf(x[0]) + cf[1] * cosf(x[0]);
                    }
                }
            }
        }
        __syncthreads();
    }
    // ---- outputs
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
        const int n = tid + q * BS;
        if (n >= N) continue;
        if (I.out_gx0) {
#pragma unroll
            for (int v = 0; v < NV; ++v) I.out_gx0[n * NV + v] = lam[q][v];
        }
        if (I.out_gp) I.out_gp[n] = (float)gp[q];
    }
    const double tot = block_sum_d(gcc, redd);
    if (tid == 0) I.out_gcc[0] = (float)tot;
}

// ------------------------------------------------- delay-blocked kernels ----
//
// When every edge delay is >= kK, the coupling inputs of kK consecutive steps
// (forward) and the delayed coupling adjoints of kK consecutive reverse steps
// depend only on ring entries written by earlier blocks, so a workgroup needs
// one barrier per block instead of one per step. Ring layout [np][N][Hs],
// Hs = H + kK - 1: slots H..Hs-1 mirror slots 0..kK-2 so that every kK-step
// window of one node and channel is contiguous. H = dmax + kK + 1 (host).

constexpr int kK = 8;

template <typename T> using gp = __attribute__((address_space(1))) T*;
__device__ __forceinline__ int u32(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uf32(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double uf64(double x) {
    const long long v = __double_as_longlong(x);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
template <typename T> __device__ __forceinline__ gp<T> ug(T* p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (gp<T>)((T*)(((uint64_t)hi << 32) | lo));
}

// Uniform per-instance view (scalar registers), read once per kernel.
struct BView {
    int N, T, H, Hs, Sg, nseg, P, cfn, integ, loss, has_c;
    float dt, dth, dt6;
    bool stoch, plain;
    double dt_ms;
    gp<const int32_t> indptr;
    gp<const uint32_t> edges;   // (src|tgt) | delay << 16, weight bits (pairs)
    gp<const float> x0, node_mc, node_jac, noise, target;
    gp<float> ring, glog, ckpt, tape, zlog, fin, loss_out, ord, gx0, gp_out, gcc_out;
    gp<const float> seedw;
};

__device__ __forceinline__ BView bview(const AdjInst* I, bool csc) {
    BView V;
    V.N = u32(I->n_nodes); V.T = u32(I->n_steps); V.H = u32(I->ring_depth); V.Hs = V.H + kK - 1;
    V.Sg = u32(I->seg_len); V.nseg = u32(I->n_seg); V.P = u32(I->bold_period); V.cfn = u32(I->cfn);
    V.integ = u32(I->integrator); V.loss = u32(I->loss_kind); V.has_c = u32(I->has_coupling);
    V.dt = uf32(I->dt); V.dth = uf32(I->dt_half); V.dt6 = uf32(I->dt_sixth);
    V.stoch = V.integ >= HB_EULER_STOCH;
    V.plain = u32(I->engine) == HB_ENGINE_PLAIN;
    V.dt_ms = uf64(I->dt_ms);
    V.indptr = ug(csc ? I->indptr_t : I->indptr);
    V.edges = ug((const uint32_t*)(csc ? I->edges_t : I->edges));
    V.x0 = ug(I->x0); V.node_mc = ug(I->node_mc); V.node_jac = ug(I->node_jac);
    V.noise = ug(I->noise); V.target = ug(I->target);
    V.ring = ug(I->ring); V.glog = ug(I->glog); V.ckpt = ug(I->ckpt); V.tape = ug(I->tape);
    V.zlog = ug(I->zlog); V.fin = ug(I->fin); V.loss_out = ug(I->out_loss); V.ord = ug(I->ord_work);
    V.gx0 = ug(I->out_gx0); V.gp_out = ug(I->out_gp); V.gcc_out = ug(I->out_gcc);
    V.seedw = ug((const float*)I->seedw);
    return V;
}

__device__ __forceinline__ int wrap_slot(int t, int H) { int s = t % H; return s < 0 ? s + H : s; }

// BOLD helper waves of the blocked forward kernel. With AdjInst::bold_fused
// the workgroup is launched twice as wide as the neural threads need; threads
// NB + n (NB = blockDim / 2) run node n's FP64 Balloon-Windkessel recursion on
// other SIMDs while the neural threads integrate, taking part in exactly the
// barriers the neural threads execute. Window inputs cross from the neural
// threads through zlog in global memory, ordered by those barriers: after the
// barrier that closes the block ending at step t, the windows completed by t
// are in zlog, and the helpers step them during the next block.
__device__ __forceinline__ void bold_helper_forward(const AdjInst& I, int n) {
    const int N = I.n_nodes, nw = I.n_win, P = I.bold_period, T = I.n_steps;
    const bool hown = n < N;
    const double ln1e = log(1.0 - kBwE0);
    BoldD b;
    b.s = 0.0; b.f = 1.0; b.v = 1.0; b.q = 1.0;
    int kb = 0;
    auto advance = [&](int kmax) {
        for (; kb < kmax; ++kb)
            if (hown) bold_fwd_window(I, b, kb, n, I.zlog[(size_t)kb * N + n], ln1e);
    };
    auto done_after = [&](int t) { return t >= T - 1 ? nw : (t + 1) / P; };
    __syncthreads();   // before the step-0 ring write
    __syncthreads();   // after it
    __syncthreads();   // after step 0
    int tlast = 0;
    for (int t0 = 1; t0 < T; t0 += kK) {
        advance(done_after(tlast));
        tlast = min(T - 1, t0 + kK - 1);
        __syncthreads();   // end of the block t0 .. tlast
    }
    advance(nw);
}

// Forward save, delay-blocked. One node per thread (N <= blockDim). MAXT:
// the launch bound, 512 or 1024 (networks of 513-1024 nodes; 128 registers
// per lane there, so fewer register-resident row edges).
template <int M, int NP, bool LRING, int IC, int CF, int MAXT>
__global__ __launch_bounds__(MAXT) void k_adj_forward_b(const AdjInst* __restrict__ insts) {
    extern __shared__ float lds[];
    __shared__ float ordp[2][kK][MAXT / 64][2];   // per-wave order-parameter partials, by block parity
    const AdjInst* Ip = insts + blockIdx.x;
    constexpr int NV = Tr<M>::NV, NC = Tr<M>::NC;
    const BView V = bview(Ip, false);
    const int cfn = CF >= 0 ? CF : V.cfn;   // coupling function, fixed at compile time when CF >= 0
    const int N = V.N, T = V.T, H = V.H, Hs = V.Hs;
    const int tid = threadIdx.x, BS = blockDim.x;
    const bool own = tid < N;
    using RingPtr = std::conditional_t<LRING, float*, gp<float>>;
    RingPtr ring;
    if constexpr (LRING) ring = lds; else ring = V.ring;
    StepCtx S;
    S.dt = V.dt; S.dth = V.dth; S.dt6 = V.dt6; S.integ = V.integ; S.stoch = V.stoch; S.plain = V.plain;
    S.sw = false; S.sv = 0.0f;
    float cc[kCC], nsc[NV];
#pragma unroll
    for (int i = 0; i < kCC; ++i) cc[i] = uf32(Ip->cc[i]);
#pragma unroll
    for (int v = 0; v < NV; ++v) nsc[v] = uf32(Ip->nscale[v]);
    for (int i = tid; i < NP * N * Hs; i += BS) ring[i] = 0.0f;
    if (u32(Ip->bold_fused) && tid >= (BS >> 1)) {
        bold_helper_forward(*Ip, tid - (BS >> 1));
        return;
    }
    const bool full_tape = V.nseg == 1;
    float x[NV], mc[kMC];
#pragma unroll
    for (int v = 0; v < NV; ++v) x[v] = own ? V.x0[tid * NV + v] : 0.0f;
#pragma unroll
    for (int i = 0; i < kMC; ++i) mc[i] = (V.node_mc && own) ? V.node_mc[(size_t)tid * kMC + i] : uf32(Ip->mc[i]);
    if (own) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            V.ckpt[tid * NV + v] = x[v];
            if (full_tape) V.tape[tid * NV + v] = x[v];
        }
    }
    // CSR row of this node in registers: source row base, delay, weight;
    // kG register edges per gather group (one at 1024 threads, whose register
    // budget is spent)
    constexpr int kReg = MAXT > 512 ? 8 : 32;
    constexpr int kG = MAXT > 512 ? 1 : 4;
    int en[kReg], ed[kReg];
    float ew[kReg];
    int e0 = 0, e1 = 0;
    if (own) { e0 = V.indptr[tid]; e1 = V.indptr[tid + 1]; }
    const int nreg = min(e1 - e0, kReg);
#pragma unroll
    for (int k = 0; k < kReg; ++k) {
        en[k] = 0; ed[k] = 0; ew[k] = 0.0f;
        if (k < nreg) {
            const uint32_t ex = V.edges[2 * (e0 + k)], ey = V.edges[2 * (e0 + k) + 1];
            en[k] = (int)(ex & 0xFFFF) * Hs;
            ed[k] = (int)(ex >> 16);
            ew[k] = __uint_as_float(ey);
        }
    }
    const bool ord = V.loss == HB_LOSS_ORDER;
    const int nwaves = (BS + 63) >> 6;
    double lsum = 0.0;
    int bcount = 0, kwin = 0;
    int ck_left = V.Sg > 0 ? V.Sg : -1, ck_idx = 0;   // steps to the next checkpoint, its index
    float bacc = 0.0f;
    const size_t chan = (size_t)N * Hs;
    // an HBM ring with several channels interleaves them per slot: one
    // contiguous window per edge (LDS rings keep channel planes)
    constexpr bool ILV = !LRING && NP > 1;
    auto ridx = [&](int c, size_t off) { return ILV ? off * NP + c : (size_t)c * chan + off; };

    auto ring_put = [&](int sl, const float* xs) {   // sl: ring slot of the state's time
        float pr[NP];
        pre<NC, NP>(cfn, cc, false, false, xs, pr);
#pragma unroll
        for (int c = 0; c < NP; ++c) {
            const size_t b = (size_t)tid * Hs;
            ring[ridx(c, b + sl)] = pr[c];
            if (sl < kK - 1) ring[ridx(c, b + H + sl)] = pr[c];
        }
    };
    auto step = [&](int t, const float* g) {
        if (own) {
            gp<float> gl = V.glog + ((size_t)t * N + tid) * NP;
#pragma unroll
            for (int c = 0; c < NP; ++c) gl[c] = g[c];
            float own_pre[NP], cpl[NC], z[NV];
            pre<NC, NP>(cfn, cc, false, false, x, own_pre);
            if (V.has_c) post<NC, NP>(cfn, cc, false, g, own_pre, cpl);
            else {
#pragma unroll
                for (int c = 0; c < NC; ++c) cpl[c] = 0.0f;
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) z[v] = S.stoch ? V.noise[((size_t)t * N + tid) * NV + v] * nsc[v] : 0.0f;
            integ_step<M, IC>(S, mc, cpl, z, x);
            bacc += x[0];
            if (--ck_left == 0) {   // (t + 1) is a multiple of the segment length
                ck_left = V.Sg;
                ++ck_idx;
                if (t + 1 < T) {
                    gp<float> ck = V.ckpt + ((size_t)ck_idx * N + tid) * NV;
#pragma unroll
                    for (int v = 0; v < NV; ++v) ck[v] = x[v];
                }
            }
            if (full_tape) {   // one segment: the whole trajectory, read by the reverse
                gp<float> tp = V.tape + ((size_t)(t + 1) * N + tid) * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) tp[v] = x[v];
            }
        }
        ++bcount;
        if (V.P > 0 && (bcount == V.P || t == T - 1)) {
            if (own) V.zlog[(size_t)kwin * N + tid] = bacc / (float)bcount;
            bacc = 0.0f;
            ++kwin;
            bcount = 0;
        }
    };
    auto ord_part = [&](int buf, int k) {
        float cs = own ? cosf(x[0]) : 0.0f, sn = own ? sinf(x[0]) : 0.0f;
        for (int o = 32; o > 0; o >>= 1) { cs += __shfl_down(cs, o, 64); sn += __shfl_down(sn, o, 64); }
        if ((tid & 63) == 0) { ordp[buf][k][tid >> 6][0] = cs; ordp[buf][k][tid >> 6][1] = sn; }
    };
    auto ord_close = [&](int buf, int k, int t) {
        float cs = 0.0f, sn = 0.0f;
        for (int w = 0; w < nwaves; ++w) { cs += ordp[buf][k][w][0]; sn += ordp[buf][k][w][1]; }
        const float C = cs / (float)N, Sn = sn / (float)N;
        const float R = sqrtf(C * C + Sn * Sn);
        const double dd = (double)R - (double)V.target[t];
        lsum += dd * dd;
    };

    // ---- step 0: every edge reads the current state (block 0, one step)
    __syncthreads();
    if (own) ring_put(0, x);   // slot of time 0
    __syncthreads();
    {
        float g[NP];
#pragma unroll
        for (int c = 0; c < NP; ++c) g[c] = 0.0f;
        if (own) {
            for (int e = e0; e < e1; ++e) {
                const uint32_t ex = V.edges[2 * e], ey = V.edges[2 * e + 1];
                const float w = __uint_as_float(ey);
#pragma unroll
                for (int c = 0; c < NP; ++c) g[c] += w * ring[ridx(c, (size_t)(ex & 0xFFFF) * Hs)];
            }
        }
        step(0, g);
        if (ord) ord_part(0, 0);
    }
    __syncthreads();
    int kprev = 1, tprev = 0, blk = 1;
    // ---- blocks of kK steps from t = 1; reads times t0 - d <= t0 - kK
    for (int t0 = 1; t0 < T; t0 += kK, ++blk) {
        const int kk = min(kK, T - t0);
        const int s0b = wrap_slot(t0, H);   // ring slot of t0 (one division per block)
        float acc[kK][NP];
#pragma unroll
        for (int k = 0; k < kK; ++k)
#pragma unroll
            for (int c = 0; c < NP; ++c) acc[k][c] = 0.0f;
        if (own) {
            const int s0 = s0b;
            // register edges in groups of kG whose windows load together (one
            // round trip per group); a group's padding entries read the own
            // row's base with weight 0. Sums stay in row order.
#pragma unroll
            for (int g = 0; g < kReg; g += kG) {
                if (g < nreg) {
                    float wv[kG][kK][NP];
#pragma unroll
                    for (int j = 0; j < kG; ++j) {
                        int sl = s0 - ed[g + j];
                        sl += sl < 0 ? H : 0;
                        const size_t base = (size_t)en[g + j] + sl;
#pragma unroll
                        for (int c = 0; c < NP; ++c)
#pragma unroll
                            for (int k = 0; k < kK; ++k) wv[j][k][c] = ring[ridx(c, base + k)];
                    }
#pragma unroll
                    for (int j = 0; j < kG; ++j)
#pragma unroll
                        for (int c = 0; c < NP; ++c)
#pragma unroll
                            for (int k = 0; k < kK; ++k) acc[k][c] += ew[g + j] * wv[j][k][c];
                }
            }
            for (int e = e0 + nreg; e < e1; ++e) {
                const uint32_t ex = V.edges[2 * e], ey = V.edges[2 * e + 1];
                const float w = __uint_as_float(ey);
                int sl = s0 - (int)(ex >> 16);
                sl += sl < 0 ? H : 0;
#pragma unroll
                for (int c = 0; c < NP; ++c) {
                    const size_t base = (size_t)(ex & 0xFFFF) * Hs + sl;
#pragma unroll
                    for (int k = 0; k < kK; ++k) acc[k][c] += w * ring[ridx(c, base + k)];
                }
            }
        }
        // the previous block's order partials are complete (barrier at its end)
        if (ord && tid == 0)
            for (int k = 0; k < kprev; ++k) ord_close((blk - 1) & 1, k, tprev + k);
        // unrolled step loop: static register slots
#pragma unroll
        for (int k = 0; k < kK; ++k) {
            if (k < kk) {
                float xs[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) xs[v] = x[v];   // state at the start of step t0 + k
                step(t0 + k, acc[k]);
                if (ord) ord_part(blk & 1, k);
                if (own) {   // slot of t0 + k (read by later blocks only), from the block's first slot
                    int sl = s0b + k;
                    sl -= sl >= H ? H : 0;
                    ring_put(sl, xs);
                }
            }
        }
        kprev = kk;
        tprev = t0;
        __syncthreads();
    }
    if (ord && tid == 0) {
        for (int k = 0; k < kprev; ++k) ord_close((blk - 1) & 1, k, tprev + k);
        V.loss_out[0] = (float)(lsum / (double)T);
    }
    if (own) {
#pragma unroll
        for (int v = 0; v < NV; ++v) V.fin[tid * NV + v] = x[v];
    }
}

// Reverse sweep, delay-blocked. One node per thread; CSC out-edges. The ring
// holds dL/dg (coupling adjoints) by step: step s reads, over out-edge
// (j -> i, delay d >= kK), the adjoint of step s + d -- always in a later
// block. Step 0 (every edge read the current state) runs last, per step.
template <int M, int NP, bool LRING, int IC, int CF, int MAXT>
__global__ __launch_bounds__(MAXT) void k_adj_reverse_b(const AdjInst* __restrict__ insts) {
    extern __shared__ float lds[];
    __shared__ double redd[MAXT / 64];
    const AdjInst* Ip = insts + blockIdx.x;
    constexpr int NV = Tr<M>::NV, NC = Tr<M>::NC;
    const BView V = bview(Ip, true);
    const int cfn = CF >= 0 ? CF : V.cfn;   // coupling function, fixed at compile time when CF >= 0
    const int N = V.N, T = V.T, H = V.H, Hs = V.Hs, Sg = V.Sg, P = V.P;
    const int tid = threadIdx.x, BS = blockDim.x;
    // Recompute helper waves (AdjInst::rc_fused, the workgroup launched twice
    // as wide): thread NB + n recomputes node n's states of the segment below
    // the one being reversed, into the other half of a double tape, and takes
    // part in exactly the barriers the reverse threads execute.
    // (the 1024-thread instantiation never takes them: the doubled workgroup must fit 512)
    const bool rcf = MAXT <= 512 && u32(Ip->rc_fused) != 0;
    const int NB = rcf ? (BS >> 1) : BS;
    const bool helper = tid >= NB;
    const int me = helper ? tid - NB : tid;   // the node this thread works on
    const bool own = !helper && tid < N;      // reverse role
    const bool hown = helper && me < N;       // recompute role
    const size_t chan = (size_t)N * Hs;
    // channels interleaved per slot in an HBM ring, as in the forward save
    constexpr bool ILV = !LRING && NP > 1;
    auto ridx = [&](int c, size_t off) { return ILV ? off * NP + c : (size_t)c * chan + off; };
    using RingPtr = std::conditional_t<LRING, float*, gp<float>>;
    RingPtr ring;
    if constexpr (LRING) ring = lds; else ring = V.ring;
    StepCtx S;
    S.dt = V.dt; S.dth = V.dth; S.dt6 = V.dt6; S.integ = V.integ; S.stoch = V.stoch; S.plain = V.plain;
    S.sw = false; S.sv = 0.0f;
    float cc[kCC], nsc[NV];
#pragma unroll
    for (int i = 0; i < kCC; ++i) cc[i] = uf32(Ip->cc[i]);
#pragma unroll
    for (int v = 0; v < NV; ++v) nsc[v] = uf32(Ip->nscale[v]);
    for (int i = tid; i < NP * N * Hs; i += BS) ring[i] = 0.0f;
    float mc[kMC], jac[kMC], lam[NV];
#pragma unroll
    for (int i = 0; i < kMC; ++i) {
        mc[i] = (V.node_mc && (own || hown)) ? V.node_mc[(size_t)me * kMC + i] : uf32(Ip->mc[i]);
        jac[i] = (V.node_jac && own) ? V.node_jac[(size_t)tid * kMC + i] : 0.0f;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) lam[v] = 0.0f;
    double gpar = 0.0, gcc = 0.0;
    constexpr int kReg = MAXT > 512 ? 8 : 16;   // register-resident out-edges (the rest read from memory)
    constexpr int kG = MAXT > 512 ? 1 : 4;      // register out-edges per gather group
    int en[kReg], ed[kReg];
    float ew[kReg];
    int e0 = 0, e1 = 0;
    if (own) { e0 = V.indptr[tid]; e1 = V.indptr[tid + 1]; }
    const int nreg = min(e1 - e0, kReg);
#pragma unroll
    for (int k = 0; k < kReg; ++k) {
        en[k] = 0; ed[k] = 0; ew[k] = 0.0f;
        if (k < nreg) {
            const uint32_t ex = V.edges[2 * (e0 + k)], ey = V.edges[2 * (e0 + k) + 1];
            en[k] = (int)(ex & 0xFFFF) * Hs;
            ed[k] = (int)(ex >> 16);
            ew[k] = __uint_as_float(ey);
        }
    }
    const bool seed_win = P > 0 && (V.loss == HB_LOSS_FC || V.loss == HB_LOSS_BOLD_MSE);
    const bool seed_ord = V.loss == HB_LOSS_ORDER;

    // one reverse step of own node: lam (dL/dx_{s+1}) -> lx (partial dL/dx_s), lg
    auto vjp = [&](int s, const float* xs, const float* g, const float* zs, float* lx, float* lg) {
        float own_pre[NP], cpl[NC], mu[NC], lown[NC];
        pre<NC, NP>(cfn, cc, false, false, xs, own_pre);
        if (V.has_c) post<NC, NP>(cfn, cc, false, g, own_pre, cpl);
        else {
#pragma unroll
            for (int c = 0; c < NC; ++c) cpl[c] = 0.0f;
        }
        float dps = 0.0f, dcs = 0.0f;
        step_vjp<M, IC>(S, mc, cpl, zs, xs, lam, lx, mu, jac, dps);
        gpar += (double)dps;
        if (V.has_c) {
            post_vjp<NC, NP>(cfn, cc, g, own_pre, mu, lg, lown, dcs);
            gcc += (double)dcs;
#pragma unroll
            for (int c = 0; c < NC; ++c) lx[c] += lown[c];
        } else {
#pragma unroll
            for (int c = 0; c < NP; ++c) lg[c] = 0.0f;
        }
        (void)s;
    };
    // sw: dL/d(window input) / count of step s's BOLD window (prefetched by the caller)
    auto seed = [&](int s, int t0s, const float* xs, float* l, float sw) {
        if (s < 1) return;
        if (seed_win) l[0] += sw;
        if (seed_ord) {
            const gp<float> cf = V.ord + (size_t)(s - t0s) * 4;
            l[0] += cf[0] * sinf(xs[0]) + cf[1] * cosf(xs[0]);
        }
    };
    auto ring_put = [&](int sl, const float* lg) {   // sl: ring slot of the step
#pragma unroll
        for (int c = 0; c < NP; ++c) {
            const size_t b = (size_t)tid * Hs;
            ring[ridx(c, b + sl)] = lg[c];
            if (sl < kK - 1) ring[ridx(c, b + H + sl)] = lg[c];
        }
    };
    __syncthreads();

    // tape rows of segment sg: one buffer, or with helper waves two halves (the
    // segment being reversed and the one being recomputed below it)
    const size_t tape_half = (size_t)(Sg + 1) * N * NV;
    auto tape_of = [&](int sg) { return V.tape + (rcf ? (size_t)(sg & 1) * tape_half : (size_t)0); };
    // helper role: node me's steps [ta, tb) of the segment starting at t0 (at
    // most kK), from x (the state at ta), into tape rows ta - t0 + 1 ...
    auto rc_chunk = [&](gp<float> tq, int t0, int ta, int tb, float (&x)[NV]) {
        float gq[kK][NP];
#pragma unroll
        for (int m = 0; m < kK; ++m) {
            if (ta + m < tb) {
                const gp<float> gl = V.glog + ((size_t)(ta + m) * N + me) * NP;
#pragma unroll
                for (int c = 0; c < NP; ++c) gq[m][c] = gl[c];
            }
        }
#pragma unroll
        for (int m = 0; m < kK; ++m) {
            const int t = ta + m;
            if (t < tb) {
                float own_pre[NP], cpl[NC], z[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) z[v] = S.stoch ? V.noise[((size_t)t * N + me) * NV + v] * nsc[v] : 0.0f;
                pre<NC, NP>(cfn, cc, false, false, x, own_pre);
                if (V.has_c) post<NC, NP>(cfn, cc, false, gq[m], own_pre, cpl);
                else {
#pragma unroll
                    for (int c = 0; c < NC; ++c) cpl[c] = 0.0f;
                }
                integ_step<M, IC>(S, mc, cpl, z, x);
                gp<float> tp = tq + ((size_t)(t - t0 + 1) * N + me) * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) tp[v] = x[v];
            }
        }
    };

    for (int seg = V.nseg - 1; seg >= 0; --seg) {
        const int t0s = seg * Sg, t1s = min(T, t0s + Sg);
        const gp<float> tape = tape_of(seg);
        // ---- recompute states t0s..t1s from the checkpoint and the coupling log
        // (with one segment the forward wrote the whole trajectory); the log
        // and noise rows of the next kK steps are loaded while the current kK
        // steps integrate (double buffer). With helper waves, only the last
        // segment is recomputed here (by the helpers); each other segment was
        // recomputed while the segment above it was reversed.
        if (rcf && hown && seg == V.nseg - 1) {
            float x[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) { x[v] = V.ckpt[((size_t)seg * N + me) * NV + v]; tape[(size_t)me * NV + v] = x[v]; }
            for (int ta = t0s; ta < t1s; ta += kK) rc_chunk(tape, t0s, ta, min(t1s, ta + kK), x);
        }
        if (!rcf && own && V.nseg > 1) {
            float x[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) { x[v] = V.ckpt[((size_t)seg * N + tid) * NV + v]; tape[(size_t)tid * NV + v] = x[v]; }
            float gq0[kK][NP], gq1[kK][NP];
            auto fetch = [&](float (&gq)[kK][NP], int tg) {
#pragma unroll
                for (int m = 0; m < kK; ++m) {
                    const int t = tg + m;
                    if (t < t1s) {
                        const gp<float> gl = V.glog + ((size_t)t * N + tid) * NP;
#pragma unroll
                        for (int c = 0; c < NP; ++c) gq[m][c] = gl[c];
                    }
                }
            };
            auto compute = [&](const float (&gq)[kK][NP], int tg) {
#pragma unroll
                for (int m = 0; m < kK; ++m) {
                    const int t = tg + m;
                    if (t < t1s) {
                        float own_pre[NP], cpl[NC], z[NV];
#pragma unroll
                        for (int v = 0; v < NV; ++v) z[v] = S.stoch ? V.noise[((size_t)t * N + tid) * NV + v] * nsc[v] : 0.0f;
                        pre<NC, NP>(cfn, cc, false, false, x, own_pre);
                        if (V.has_c) post<NC, NP>(cfn, cc, false, gq[m], own_pre, cpl);
                        else {
#pragma unroll
                            for (int c = 0; c < NC; ++c) cpl[c] = 0.0f;
                        }
                        integ_step<M, IC>(S, mc, cpl, z, x);
                        gp<float> tp = tape + ((size_t)(t - t0s + 1) * N + tid) * NV;
#pragma unroll
                        for (int v = 0; v < NV; ++v) tp[v] = x[v];
                    }
                }
            };
            fetch(gq0, t0s);
            for (int tg = t0s; tg < t1s; tg += 2 * kK) {
                if (tg + kK < t1s) fetch(gq1, tg + kK);
                compute(gq0, tg);
                if (tg + kK >= t1s) break;
                if (tg + 2 * kK < t1s) fetch(gq0, tg + 2 * kK);
                compute(gq1, tg + kK);
            }
        }
        __syncthreads();
        if (seed_ord) {
            const int w = tid >> 6, l = tid & 63, nwv = BS >> 6;
            for (int r = w; r <= t1s - t0s; r += nwv) {
                float cs = 0.0f, sn = 0.0f;
                for (int n = l; n < N; n += 64) {
                    const float th = tape[((size_t)r * N + n) * NV];
                    cs += cosf(th);
                    sn += sinf(th);
                }
                for (int o = 32; o > 0; o >>= 1) { cs += __shfl_down(cs, o, 64); sn += __shfl_down(sn, o, 64); }
                if (l == 0) {
                    const int st = t0s + r;
                    float a = 0.0f, b = 0.0f;
                    if (st >= 1) {
                        const float C = cs / (float)N, Sn = sn / (float)N;
                        const float R = sqrtf(C * C + Sn * Sn);
                        const float al = R > 0.0f ? 2.0f * (R - V.target[st - 1]) / ((float)T * (float)N * R) : 0.0f;
                        a = -al * C;
                        b = al * Sn;
                    }
                    V.ord[(size_t)r * 4 + 0] = a;
                    V.ord[(size_t)r * 4 + 1] = b;
                }
            }
            __syncthreads();
        }
        if (seg == V.nseg - 1 && own) {
            float xT[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) xT[v] = tape[((size_t)(t1s - t0s) * N + tid) * NV + v];
            seed(T, t0s, xT, lam, seed_win ? V.seedw[(size_t)((T - 1) / P) * N + tid] : 0.0f);
        }
        // ---- blocks of reverse steps over [max(t0s, 1), t1s - 1], highest first
        const int tlo = max(t0s, 1);
        // the block's states and coupling log rows; the next block's load while
        // this one is differentiated when they fit the register budget
        // (Wilson-Cowan's sigmoid VJPs leave no registers for a second block)
        constexpr bool kPrefetch = (NV + NP) * kK <= 24 && M != HB_WILSON_COWAN;
        // ... and the BOLD seeds of its steps (window (st - 1) / P, stepped up)
        auto load_block = [&](float (&xx)[kK][NV], float (&gg)[kK][NP], float (&ss)[kK], int tb, int nk) {
            if (!own) return;
            int wl = 0, rl = 0;
            if (seed_win) { wl = (tb - 1) / P; rl = (tb - 1) - wl * P; }
#pragma unroll
            for (int m = 0; m < kK; ++m) {
                if (m < nk) {
                    const int st = tb + m;
                    const gp<float> tp = tape + ((size_t)(st - t0s) * N + tid) * NV;
                    const gp<float> gl = V.glog + ((size_t)st * N + tid) * NP;
#pragma unroll
                    for (int v = 0; v < NV; ++v) xx[m][v] = tp[v];
#pragma unroll
                    for (int c = 0; c < NP; ++c) gg[m][c] = gl[c];
                    ss[m] = 0.0f;
                    if (seed_win) {
                        ss[m] = V.seedw[(size_t)wl * N + tid];
                        if (++rl == P) { rl = 0; ++wl; }
                    }
                }
            }
        };
        // helper role: segment seg - 1, one chunk of kK steps per reverse block
        float xr[NV];
        int ra = 0, rb = 0;
        const int r0 = (seg - 1) * Sg;
        const gp<float> tq = tape_of(seg - 1);
        if (rcf && hown && seg > 0) {
            ra = r0;
            rb = seg * Sg;
#pragma unroll
            for (int v = 0; v < NV; ++v) { xr[v] = V.ckpt[((size_t)(seg - 1) * N + me) * NV + v]; tq[(size_t)me * NV + v] = xr[v]; }
        }
        float xb[kK][NV], gb[kK][NP], sdb[kK];
        float xn[kPrefetch ? kK : 1][NV], gn[kPrefetch ? kK : 1][NP], sdn[kPrefetch ? kK : 1];
        bool have = false;   // xb / gb already hold the block (prefetched)
        for (int ttop = t1s - 1; ttop >= tlo; ttop -= kK) {
            if (helper) {   // one barrier per block, as the reverse threads
                if (ra < rb) {
                    const int e = min(rb, ra + kK);
                    rc_chunk(tq, r0, ra, e, xr);
                    ra = e;
                }
                __syncthreads();
                continue;
            }
            const int kk = min(kK, ttop - tlo + 1), tbot = ttop - kk + 1;
            if (!have) load_block(xb, gb, sdb, tbot, kk);
            const int ttop2 = ttop - kK, kk2 = min(kK, ttop2 - tlo + 1), tbot2 = ttop2 - kk2 + 1;
            if constexpr (kPrefetch) {
                if (ttop2 >= tlo) load_block(xn, gn, sdn, tbot2, kk2);
            }
            float acc[kK][NP];
#pragma unroll
            for (int m = 0; m < kK; ++m)
#pragma unroll
                for (int c = 0; c < NP; ++c) acc[m][c] = 0.0f;
            const int sb = wrap_slot(tbot, H);   // ring slot of tbot (one division per block)
            if (own && V.has_c) {
                // register out-edges in groups of kG loaded together, as in the forward
#pragma unroll
                for (int g = 0; g < kReg; g += kG) {
                    if (g < nreg) {
                        float wv[kG][kK][NP];
#pragma unroll
                        for (int j = 0; j < kG; ++j) {
                            int sl = sb + ed[g + j];
                            sl -= sl >= H ? H : 0;
                            const size_t base = (size_t)en[g + j] + sl;
#pragma unroll
                            for (int c = 0; c < NP; ++c)
#pragma unroll
                                for (int m = 0; m < kK; ++m) wv[j][m][c] = ring[ridx(c, base + m)];
                        }
#pragma unroll
                        for (int j = 0; j < kG; ++j)
#pragma unroll
                            for (int c = 0; c < NP; ++c)
#pragma unroll
                                for (int m = 0; m < kK; ++m) acc[m][c] += ew[g + j] * wv[j][m][c];
                    }
                }
                for (int e = e0 + nreg; e < e1; ++e) {
                    const uint32_t ex = V.edges[2 * e], ey = V.edges[2 * e + 1];
                    const float w = __uint_as_float(ey);
                    int sl = sb + (int)(ex >> 16);
                    sl -= sl >= H ? H : 0;
#pragma unroll
                    for (int c = 0; c < NP; ++c) {
                        const size_t base = (size_t)(ex & 0xFFFF) * Hs + sl;
#pragma unroll
                        for (int m = 0; m < kK; ++m) acc[m][c] += w * ring[ridx(c, base + m)];
                    }
                }
            }
            float lgb[kK][NP];
            if (own) {
#pragma unroll
                for (int m = kK - 1; m >= 0; --m) {
                    if (m < kk) {
                        const int st = tbot + m;
                        float lx[NV], zs[NV];
#pragma unroll
                        for (int v = 0; v < NV; ++v) zs[v] = S.stoch ? V.noise[((size_t)st * N + tid) * NV + v] * nsc[v] : 0.0f;
                        vjp(st, xb[m], gb[m], zs, lx, lgb[m]);
                        pre_vjp<NC, NP>(xb[m], acc[m], lx);
                        seed(st, t0s, xb[m], lx, sdb[m]);
#pragma unroll
                        for (int v = 0; v < NV; ++v) lam[v] = lx[v];
                    }
                }
#pragma unroll
                for (int m = 0; m < kK; ++m) {
                    if (m < kk) {
                        int sl = sb + m;
                        sl -= sl >= H ? H : 0;
                        ring_put(sl, lgb[m]);
                    }
                }
            }
            __syncthreads();
            have = false;
            if constexpr (kPrefetch) {
                if (ttop2 >= tlo) {
#pragma unroll
                    for (int m = 0; m < kK; ++m) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) xb[m][v] = xn[m][v];
#pragma unroll
                        for (int c = 0; c < NP; ++c) gb[m][c] = gn[m][c];
                        sdb[m] = sdn[m];
                    }
                    have = true;
                }
            }
        }
        while (ra < rb) {   // a short last segment leaves steps of the one below
            const int e = min(rb, ra + kK);
            rc_chunk(tq, r0, ra, e, xr);
            ra = e;
        }
        if (seg > 0) __syncthreads();
    }
    // ---- step 0: the delayed terms (steps d) plus every edge's current-state read
    {
        float lx[NV], lg[NP], acc[NP];
        float x0v[NV], g0[NP], z0[NV];
        if (own) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                x0v[v] = V.tape[(size_t)tid * NV + v];   // segment 0: the first tape half
                z0[v] = S.stoch ? V.noise[(size_t)tid * NV + v] * nsc[v] : 0.0f;
            }
#pragma unroll
            for (int c = 0; c < NP; ++c) g0[c] = V.glog[(size_t)tid * NP + c];
            vjp(0, x0v, g0, z0, lx, lg);
            ring_put(0, lg);   // slot of step 0
        }
        __syncthreads();
        if (own) {
#pragma unroll
            for (int c = 0; c < NP; ++c) acc[c] = 0.0f;
            if (V.has_c) {
                for (int e = e0; e < e1; ++e) {
                    const uint32_t ex = V.edges[2 * e], ey = V.edges[2 * e + 1];
                    const float w = __uint_as_float(ey);
                    const int d = (int)(ex >> 16), node = (int)(ex & 0xFFFF);
#pragma unroll
                    for (int c = 0; c < NP; ++c) {
                        const size_t b = (size_t)node * Hs;
                        acc[c] += w * ring[ridx(c, b + wrap_slot(d, H))];
                        if (d > 0) acc[c] += w * ring[ridx(c, b)];
                    }
                }
            }
            pre_vjp<NC, NP>(x0v, acc, lx);
#pragma unroll
            for (int v = 0; v < NV; ++v) lam[v] = lx[v];
            if (V.gx0) {
#pragma unroll
                for (int v = 0; v < NV; ++v) V.gx0[tid * NV + v] = lam[v];
            }
            if (V.gp_out) V.gp_out[tid] = (float)gpar;
        }
    }
    const double tot = block_sum_d(gcc, redd);
    if (tid == 0) V.gcc_out[0] = (float)tot;
}

// ------------------------------------------------------------ dispatch -----
//
// hb_adjoint.hip is compiled once per model (HB_ADJ_UNIT 0..5 = the model
// id: that model's forward-save and reverse kernels) plus once for the shared
// kernels and the stage launcher (HB_ADJ_UNIT 6), so the units build in
// parallel. Each model kernel is instantiated per integrator class (IC 0..2);
// the delay-blocked ones also with the Linear coupling fixed (CF).

#if HB_ADJ_UNIT < 6
template <int M, int NP, int IC>
static const void* ptr(bool rev, int npt) {
    if (rev) return npt == 1 ? (const void*)k_adj_reverse<M, NP, 1, IC, -1> : (const void*)k_adj_reverse<M, NP, 2, IC, -1>;
    return npt == 1 ? (const void*)k_adj_forward<M, NP, 1, IC, -1> : (const void*)k_adj_forward<M, NP, 2, IC, -1>;
}

template <int M, int NP, int IC, int CF, int MAXT>
static const void* ptr_bt(bool rev, bool lring) {
    if (rev)
        return lring ? (const void*)k_adj_reverse_b<M, NP, true, IC, CF, MAXT>
                     : (const void*)k_adj_reverse_b<M, NP, false, IC, CF, MAXT>;
    return lring ? (const void*)k_adj_forward_b<M, NP, true, IC, CF, MAXT>
                 : (const void*)k_adj_forward_b<M, NP, false, IC, CF, MAXT>;
}

// blocks wider than 512 threads take the 1024-thread instantiation
template <int M, int NP, int IC, int CF>
static const void* ptr_b(bool rev, bool lring, bool wide) {
    return wide ? ptr_bt<M, NP, IC, CF, 1024>(rev, lring) : ptr_bt<M, NP, IC, CF, 512>(rev, lring);
}

template <int M, int NP>
static const void* pick(const AdjLaunch& L, bool rev) {
    if (L.blocked) {
        const bool lin = L.cfk == HB_CF_LINEAR;
        // the launch width: the forward save of a BOLD-fused group and the
        // reverse of a recompute-fused group double the block
        const bool wide = (rev ? (L.rc_fused ? 2 * L.block : L.block) : (L.bold_fused ? 2 * L.block : L.block)) > 512;
        switch (L.ic) {
        case 0: return lin ? ptr_b<M, NP, 0, HB_CF_LINEAR>(rev, L.lring, wide) : ptr_b<M, NP, 0, -1>(rev, L.lring, wide);
        case 1: return lin ? ptr_b<M, NP, 1, HB_CF_LINEAR>(rev, L.lring, wide) : ptr_b<M, NP, 1, -1>(rev, L.lring, wide);
        default: return lin ? ptr_b<M, NP, 2, HB_CF_LINEAR>(rev, L.lring, wide) : ptr_b<M, NP, 2, -1>(rev, L.lring, wide);
        }
    }
    switch (L.ic) {
    case 0: return ptr<M, NP, 0>(rev, L.npt);
    case 1: return ptr<M, NP, 1>(rev, L.npt);
    default: return ptr<M, NP, 2>(rev, L.npt);
    }
}

// Instantiated (model, pre channels): Kuramoto with and without Kuramoto
// coupling, MPR with two channels or with the zero-coefficient one dropped.
static const void* kernel_of(const AdjLaunch& L, bool rev) {
    constexpr int M = HB_ADJ_UNIT;
    if (L.model != M) return nullptr;
    if constexpr (M == HB_MPR || M == HB_KURAMOTO) {
        if (L.np == 2) return pick<M, 2>(L, rev);
    }
    return L.np == 1 ? pick<M, 1>(L, rev) : nullptr;
}
#endif

}  // namespace adj

#if HB_ADJ_UNIT < 6
#define HB_ADJ_CAT(a, b) a##b
#define HB_ADJ_XCAT(a, b) HB_ADJ_CAT(a, b)
// op 0: launch the forward save; 1: launch the reverse sweep; 2: set the
// maximum dynamic LDS (arg bytes) on every kernel of the unit.
hipError_t HB_ADJ_XCAT(adj_unit_, HB_ADJ_UNIT)(int op, const AdjLaunch* L, hipStream_t s, int arg) {
    if (op == 2) {
        AdjLaunch q{};
        q.model = HB_ADJ_UNIT;
        for (int np = 1; np <= 2; ++np)
            for (int ic = 0; ic < 3; ++ic)
                for (int v = 0; v < 12; ++v) {
                    q.np = np; q.ic = ic;
                    q.blocked = (v % 6) >= 2;
                    q.npt = (v % 6) == 1 ? 2 : 1;
                    q.lring = (v % 6) >= 4;
                    q.cfk = (v & 1) && (v % 6) >= 2 ? HB_CF_LINEAR : -1;
                    q.block = v >= 6 ? 1024 : 64;   // both launch widths of the blocked kernels
                    q.bold_fused = 0;
                    q.rc_fused = 0;
                    for (int rev = 0; rev < 2; ++rev) {
                        const void* k = adj::kernel_of(q, rev != 0);
                        if (!k) continue;
                        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, arg);
                        if (e != hipSuccess) return e;
                    }
                }
        return hipSuccess;
    }
    const void* k = adj::kernel_of(*L, op == 1);
    if (!k) return hipErrorInvalidValue;
    const AdjInst* d = L->insts;
    void* args[] = {(void*)&d};
    // the forward save of a fused group carries the BOLD helper waves, the
    // reverse of a recompute-fused group its recompute helper waves
    const int block = ((op == 0 && L->bold_fused) || (op == 1 && L->rc_fused)) ? 2 * L->block : L->block;
    return hipLaunchKernel(k, dim3(L->n), dim3(block), args, op == 0 ? L->lds_fwd : L->lds_rev, s);
}
#else
namespace {
using AdjUnitFn = hipError_t (*)(int, const AdjLaunch*, hipStream_t, int);
const AdjUnitFn kAdjUnits[6] = {adj_unit_0, adj_unit_1, adj_unit_2, adj_unit_3, adj_unit_4, adj_unit_5};
}  // namespace

hipError_t adjoint_set_attributes(int max_lds) {
    {
        hipError_t e = hipFuncSetAttribute((const void*)adj::k_loss, hipFuncAttributeMaxDynamicSharedMemorySize, adj::kFcLds);
        if (e != hipSuccess) return e;
    }
    for (int m = 0; m < 6; ++m) {
        const hipError_t e = kAdjUnits[m](2, nullptr, nullptr, max_lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_adjoint_stage(int stage, const AdjLaunch& L, hipStream_t s) {
    const AdjInst* d = L.insts;
    if ((stage == 0 || stage == 4) && (L.model < 0 || L.model > 5)) return hipErrorInvalidValue;
    switch (stage) {
    case 0: return kAdjUnits[L.model](0, &L, s, 0);
    case 1:
        if (L.bold_fused) return hipSuccess;   // in the forward kernel's helper waves
        hipLaunchKernelGGL(adj::k_bold_forward, dim3(L.n), dim3(256), 0, s, d);
        return hipGetLastError();
    case 2: hipLaunchKernelGGL(adj::k_loss, dim3(L.n), dim3(256), adj::kFcLds, s, d); return hipGetLastError();
    case 3: hipLaunchKernelGGL(adj::k_bold_reverse, dim3(L.n), dim3(1024), 0, s, d); return hipGetLastError();
    default: return kAdjUnits[L.model](1, &L, s, 0);
    }
}
#endif

}  // namespace hb
