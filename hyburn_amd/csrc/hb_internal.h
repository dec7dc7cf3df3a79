// hb_internal.h -- types shared by the host planner and the gfx950 kernels.
//
// The per-instance device descriptor (hb_dev_inst) carries everything one
// workgroup needs to integrate one simulation: sizes, precomputed model and
// coupling constants (f32, evaluated in the reference's host-side op order),
// and device pointers into the plan's HBM arena.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/hyburn_hip.h"

#define HB_HD __host__ __device__ __forceinline__

namespace hb {

constexpr int kMaxNV = 6;
constexpr int kMC = 28;   // model constant slots
constexpr int kCC = 12;   // coupling constant slots
// Threads per forward workgroup: k_single takes up to 512; k_wide (the
// one-node-per-thread schedules, delay-blocked ones among them) up to 1024.
constexpr int kBlockMax = 512, kWideBlockMax = 1024;

// Model registry (config.rs:43-72): nvar, ncvar, nparams.
HB_HD int model_nvar(int m) {
    switch (m) {
    case HB_JANSEN_RIT: return 6;
    case HB_KURAMOTO: case HB_RWW: case HB_LINEAR: return 1;
    case HB_EPILEPTOR_CODIM3: case HB_LARTER_BREAKSPEAR: return 3;
    case HB_COOMBES_BYRNE: case HB_GAST_SD: case HB_GAST_SF: case HB_REDUCED_FHN: return 4;
    case HB_EPILEPTOR_CODIM3_SLOWMOD: case HB_ZERLAUT_FIRST: case HB_KIONEX: return 5;
    case HB_EPILEPTOR: case HB_REDUCED_HR: return 6;
    case HB_EPILEPTOR_RS: case HB_ZERLAUT_SECOND: case HB_DUMONT_GUTKIN: return 8;
    case HB_ZETTERBERG_JANSEN: return 12;
    default: return 2;
    }
}
HB_HD int model_ncvar(int m) {
    if (m == HB_EPILEPTOR_RS) return 3;
    if (m == HB_COOMBES_BYRNE || m == HB_GAST_SD || m == HB_GAST_SF || m == HB_DUMONT_GUTKIN) return 4;
    return (m == HB_MPR || m == HB_SUP_HOPF || m == HB_HOPFIELD || m == HB_COOMBES_BYRNE_2D || m == HB_EPILEPTOR ||
            m == HB_REDUCED_FHN || m == HB_REDUCED_HR) ? 2 : 1;
}
HB_HD int model_nparams(int m) {
    switch (m) {
    case HB_G2DO: return 12; case HB_MPR: return 7; case HB_JANSEN_RIT: return 13;
    case HB_WILSON_COWAN: return 22; case HB_KURAMOTO: return 1; case HB_LINEAR: return 1;
    case HB_SUP_HOPF: return 2; case HB_HOPFIELD: return 3; case HB_COOMBES_BYRNE_2D: return 4;
    case HB_EPILEPTOR_CODIM3: return 13; case HB_EPILEPTOR_CODIM3_SLOWMOD: return 27;
    case HB_EPILEPTOR_RS: return 28; case HB_ZERLAUT_FIRST: case HB_ZERLAUT_SECOND: return 50;
    case HB_KIONEX: return 14;
    case HB_COOMBES_BYRNE: return 5; case HB_GAST_SD: case HB_GAST_SF: return 9;
    case HB_LARTER_BREAKSPEAR: return 32; case HB_EPILEPTOR_2D: case HB_REDUCED_HR: return 12;
    case HB_EPILEPTOR: return 17; case HB_RWW_EXC_INH: return 19; case HB_DECO_BALANCED: return 20;
    case HB_ZETTERBERG_JANSEN: return 18; case HB_DUMONT_GUTKIN: return 14;
    default: return 8;
    }
}
// Models 10-27 keep their constants in a 64-slot table (model_consts_ext).
HB_HD bool model_ext(int m) { return m >= HB_EPILEPTOR_CODIM3 && m <= HB_DUMONT_GUTKIN; }
constexpr int kMCX = 64;
// hybrid_integrator's deterministic-Heun models (model_prefers_heun, dfun.rs:1871-1873)
HB_HD bool model_prefers_heun(int m) {
    return m == HB_G2DO || m == HB_EPILEPTOR_RS || m == HB_EPILEPTOR_CODIM3 || m == HB_EPILEPTOR_CODIM3_SLOWMOD ||
           m == HB_KIONEX || m == HB_EPILEPTOR || m == HB_EPILEPTOR_2D;
}
// The reference's single batch-sweepable parameter per model (dfun.rs:49-72).
HB_HD int model_sweep_index(int m) {
    switch (m) {
    case HB_G2DO: return 1; case HB_MPR: return 4; case HB_JANSEN_RIT: return 12;
    case HB_WILSON_COWAN: return 18; case HB_KURAMOTO: return 0; case HB_RWW: return 7;
    default: return -1;   // Linear, SupHopf: the batch dfun ignores a swept value (dfun.rs:522-557)
    }
}
// Number of pre() channels the ring stores per node (coupling.rs:204-258).
HB_HD int coupling_npre(int cfn, int nc) {
    if (cfn == HB_CF_KURAMOTO) return 2 * nc;
    if (cfn == HB_CF_SJR) return 1;
    return nc;
}

// Pre channels the ring must hold for this instance: MPR couples through
// c0 * cr + c1 * cv (dfun.rs:455-477); with cv == 0 (the default) the second
// channel only adds c1 * 0, so it is neither stored nor gathered. Per-node
// values of cv, or a gradient with respect to cv, keep it.
HB_HD int ring_channels(int model, int cfn, const float* params, int node_param_idx, int grad_param) {
    const int np = coupling_npre(cfn, model_ncvar(model));
    if (model == HB_MPR && np == 2 && cfn != HB_CF_KURAMOTO && params && params[6] == 0.0f &&
        node_param_idx != 6 && grad_param != 6)
        return 1;
    return np;
}

// f32 epsilon guard used by the reference for time constants (dfun.rs:279).
constexpr float kEps = 1.1920929e-07f;
constexpr float kPi = 3.14159265358979323846f;

// Model constants, f32, evaluated with the reference's host-side op order.
// G2DO  dfun.rs:277-293 | MPR 445-456 | JR 327-366 | WC 392-403 | KUR 486 | RWW 500-502
// Templated on the scalar so the adjoint's host side can evaluate the same
// expressions on dual numbers (d mc / d parameter); S = float is the forward.
HB_HD float mc_exp(float x) { return expf(x); }

template <class S>
HB_HD void model_consts_t(int m, const S* p, S* mc) {
    const S eps = S(kEps), one = S(1.0f), two = S(2.0f), pi = S(kPi);
    for (int i = 0; i < kMC; ++i) mc[i] = S(0.0f);
    switch (m) {
    case HB_G2DO: {
        S tau = p[0] > eps ? p[0] : eps;
        mc[0] = p[2]; mc[1] = p[3]; mc[2] = p[4]; mc[3] = p[6]; mc[4] = p[7]; mc[5] = p[8];
        mc[6] = p[9]; mc[7] = p[10]; mc[8] = p[11];
        mc[9] = p[11] * p[1];          // gamma * I  (add_scalar(gamma * params[1]))
        mc[10] = p[5] * tau;           // d * tau
        mc[11] = p[5] / tau;           // d / tau
        break; }
    case HB_MPR: {
        S tau = p[0] > eps ? p[0] : eps;
        mc[0] = one / tau;
        mc[1] = p[1] / (pi * tau);
        mc[2] = pi * pi * tau * tau;
        mc[3] = p[3] * tau;
        mc[4] = p[2] + p[4];           // eta + i_ext
        mc[5] = p[2];                  // eta (swept form adds I separately)
        mc[6] = p[5]; mc[7] = p[6];
        break; }
    case HB_JANSEN_RIT: {
        S Ap = p[0], Bp = p[1], a = p[2], b = p[3], J = p[7];
        mc[0] = p[4]; mc[1] = p[6]; mc[2] = two * p[5];
        mc[3] = Ap * a; mc[4] = two * a; mc[5] = a * a;
        mc[6] = p[8] * J; mc[7] = p[9] * J; mc[8] = p[10] * J; mc[9] = p[11] * J;
        mc[10] = Bp * b; mc[11] = two * b; mc[12] = b * b; mc[13] = p[12];
        break; }
    case HB_WILSON_COWAN: {
        S te = p[4] > eps ? p[4] : eps, ti = p[5] > eps ? p[5] : eps;
        mc[0] = p[0]; mc[1] = p[1]; mc[2] = p[2]; mc[3] = p[3];
        mc[4] = one / te; mc[5] = one / ti;
        mc[6] = p[6]; mc[7] = p[7]; mc[8] = p[8]; mc[9] = p[9];
        mc[10] = p[10]; mc[11] = p[11]; mc[12] = p[12]; mc[13] = p[13];
        mc[14] = p[14]; mc[15] = p[15]; mc[16] = p[16]; mc[17] = p[17];
        mc[18] = p[18]; mc[19] = p[19]; mc[20] = p[20]; mc[21] = p[21];
        mc[22] = one / (one + mc_exp(p[6] * p[7]));
        mc[23] = one / (one + mc_exp(p[10] * p[11]));
        break; }
    case HB_KURAMOTO:
    case HB_LINEAR:
        mc[0] = p[0];
        break;
    case HB_SUP_HOPF:
        mc[0] = p[0]; mc[1] = p[1];
        break;
    case HB_HOPFIELD: {   // dfun.rs hopfield_dfun_batch: 1 / max(tau, f32::EPSILON), dynamic = p[2] > 0.5
        const S e32 = S(1.1920929e-07f);
        mc[0] = one / (p[0] > e32 ? p[0] : e32);
        mc[1] = one / (p[1] > e32 ? p[1] : e32);
        mc[2] = p[2] > S(0.5f) ? one : S(0.0f);
        break; }
    case HB_COOMBES_BYRNE_2D:   // coombes_byrne2d_dfun_batch: delta / pi, k * pi, pi * pi as f32
        mc[0] = p[0] / pi; mc[1] = p[2] * pi; mc[2] = p[3]; mc[3] = pi * pi; mc[4] = p[1];
        break;
    default: // RWW
        mc[0] = p[0]; mc[1] = p[1]; mc[2] = p[2]; mc[3] = p[3];
        mc[4] = one / p[4];
        mc[5] = p[5] * p[6]; mc[6] = p[6]; mc[7] = p[7];
        break;
    }
}

HB_HD void model_consts(int m, const float* p, float* mc) { model_consts_t<float>(m, p, mc); }

// Constants of models 10-27 (kMCX slots): the parameters, then the products
// the reference forms on the host once per call, in its f32 order.
inline void model_consts_ext(int m, const float* p, float* c) {
    const float eps = kEps;
    for (int i = 0; i < kMCX; ++i) c[i] = 0.0f;
    for (int i = 0; i < model_nparams(m); ++i) c[i] = p[i];
    switch (m) {
    case HB_EPILEPTOR_CODIM3: {   // dfun.rs:958-983: unit start/stop points, e and f of the great circle
        const float st[3] = {p[1], -p[0], p[2]}, sp[3] = {p[4], -p[3], p[5]};
        const float sn = sqrtf(((0.0f + st[0] * st[0]) + st[1] * st[1]) + st[2] * st[2]);
        const float pn = sqrtf(((0.0f + sp[0] * sp[0]) + sp[1] * sp[1]) + sp[2] * sp[2]);
        float a[3], b[3], cr[3];
        for (int i = 0; i < 3; ++i) { a[i] = st[i] / (sn > eps ? sn : eps); b[i] = sp[i] / (pn > eps ? pn : eps); }
        cr[0] = a[1] * b[2] - a[2] * b[1];
        cr[1] = a[2] * b[0] - a[0] * b[2];
        cr[2] = a[0] * b[1] - a[1] * b[0];
        const float cn = sqrtf(((0.0f + cr[0] * cr[0]) + cr[1] * cr[1]) + cr[2] * cr[2]);
        for (int i = 0; i < 3; ++i) c[13 + i] = a[i];
        if (cn > eps) {
            float cca[3];
            for (int i = 0; i < 3; ++i) cca[i] = cr[(i + 1) % 3] * a[(i + 2) % 3] - cr[(i + 2) % 3] * a[(i + 1) % 3];
            float cc2 = sqrtf(((0.0f + cca[0] * cca[0]) + cca[1] * cca[1]) + cca[2] * cca[2]);
            cc2 = cc2 > eps ? cc2 : eps;
            for (int i = 0; i < 3; ++i) c[16 + i] = cca[i] / cc2;
        } else {
            c[16] = 0.0f; c[17] = 1.0f; c[18] = 0.0f;
        }
        break; }
    case HB_EPILEPTOR_CODIM3_SLOWMOD:   // dfun.rs:1063-1074: g*R, h*R, l*R, m*R
        for (int i = 0; i < 3; ++i) {
            c[27 + i] = p[13 + i] * p[7]; c[30 + i] = p[19 + i] * p[7];
            c[33 + i] = p[16 + i] * p[7]; c[36 + i] = p[22 + i] * p[7];
        }
        break;
    case HB_EPILEPTOR_RS:   // dfun.rs:1143-1208
        c[17] = p[17] > eps ? p[17] : eps;
        c[28] = -4.0f * p[2];
        c[29] = p[15] * p[7];
        c[30] = -0.01f * p[15];
        c[31] = p[21] * c[17];
        c[32] = p[21] / c[17];
        c[33] = p[26] * p[18];
        c[34] = p[26] * p[27];
        break;
    case HB_ZERLAUT_FIRST: case HB_ZERLAUT_SECOND:   // dfun.rs:1449-1620
        c[50] = ((1.0f - p[19]) * p[17]) * p[16];
        c[51] = (p[19] * p[18]) * p[16];
        c[52] = p[12] * p[14];
        c[53] = p[13] * p[15];
        c[54] = p[0] * p[1];
        c[55] = p[0] * p[2];
        c[56] = p[0] / p[3];
        c[57] = 1.0f / p[22];
        c[58] = p[16] * (1.0f - p[19]);
        c[59] = p[16] * p[19];
        c[60] = 1.0f / ((2.0f * 1e-4f) * 1e3f);
        c[61] = 1.0f / (((1e-4f * 1e3f) * 1e-4f) * 1e3f);
        break;
    case HB_KIONEX:   // dfun.rs:1788-1849
        c[10] = p[10] > eps ? p[10] : eps;
        c[11] = p[11] > eps ? p[11] : eps;
        c[14] = -1.0f / c[10];
        c[15] = p[6] / kPi;
        c[16] = 2.0f * p[6];
        c[17] = 2.0f * p[8];
        c[18] = -(p[12] / 2160.0f);
        c[19] = 26.64f * logf(112.0f / 5.0f);
        break;
    case HB_COOMBES_BYRNE:   // dfun.rs:609-637: delta/pi, pi*pi, k*pi
        c[5] = p[0] / kPi; c[6] = kPi * kPi; c[7] = p[3] * kPi;
        break;
    case HB_GAST_SD: case HB_GAST_SF:   // dfun.rs:639-716
        c[9] = p[0] > eps ? p[0] : eps;
        c[10] = p[1] > eps ? p[1] : eps;
        c[11] = 1.0f / c[9];
        c[12] = p[4] / (kPi * c[9]);
        c[13] = ((kPi * kPi) * c[9]) * c[9];
        c[14] = p[6] + p[3];
        c[15] = p[5] * c[9];
        c[16] = 1.0f / c[10];
        break;
    case HB_LARTER_BREAKSPEAR:   // dfun.rs:718-766
        c[31] = p[31] > eps ? p[31] : eps;
        c[32] = 1.0f / p[12]; c[33] = 1.0f / p[13]; c[34] = 1.0f / p[14];
        c[35] = 0.5f * p[28]; c[36] = 1.0f / p[15];
        c[37] = 0.5f * p[29]; c[38] = 1.0f / p[16];
        c[39] = (p[23] * p[25]) * p[19];
        c[40] = ((1.0f - p[23]) * p[25]) * p[19];
        c[41] = (1.0f - p[23]) * p[19];
        c[42] = p[23] * p[19];
        c[43] = p[20] * p[24];
        c[44] = ((p[30] * p[4]) * p[22]) / c[31];
        c[45] = p[21] * p[24];
        c[46] = p[30] * p[22];
        break;
    case HB_EPILEPTOR_2D:   // dfun.rs:776-818
        c[12] = p[6] - p[3];
        c[13] = p[5] + p[0];
        c[14] = -4.0f * p[1];
        c[15] = p[10] * p[7];
        break;
    case HB_EPILEPTOR:   // dfun.rs:820-874
        c[17] = -4.0f * p[2];
        c[18] = p[15] * p[7];
        c[19] = -0.01f * p[15];
        break;
    case HB_RWW_EXC_INH: case HB_DECO_BALANCED: {   // dfun.rs:884-950
        const float te = p[4] > eps ? p[4] : eps, ti = p[12] > eps ? p[12] : eps;
        c[19] = p[17] * p[6];
        c[20] = p[7] * p[15] + p[16];
        c[21] = 1.0f / te;
        c[22] = p[14] * p[15];
        c[23] = 1.0f / ti;
        if (m == HB_DECO_BALANCED) {
            c[24] = p[0] * p[19]; c[25] = (-p[1]) * p[19];
            c[26] = p[8] * p[19]; c[27] = (-p[9]) * p[19];
        }
        break; }
    case HB_ZETTERBERG_JANSEN:   // dfun.rs:1212-1268
        c[18] = 2.0f * p[4];
        c[19] = p[0] * p[2]; c[20] = p[1] * p[3];
        c[21] = p[2] * p[2]; c[22] = p[3] * p[3];
        c[23] = c[19] * p[7]; c[24] = p[13] * p[8]; c[25] = p[14] * p[9];
        c[26] = c[19] * p[10]; c[27] = c[20] * p[11];
        c[28] = 2.0f * p[2]; c[29] = 2.0f * p[3];
        break;
    case HB_REDUCED_HR:   // dfun.rs:1302-1348
        c[12] = p[0] * p[5];
        c[13] = (-p[0]) * p[6];
        break;
    case HB_DUMONT_GUTKIN:   // dfun.rs:1350-1400
        c[14] = p[3] > eps ? p[3] : eps;
        c[15] = p[7] > eps ? p[7] : eps;
        c[16] = p[8] > eps ? p[8] : eps;
        c[17] = 1.0f / c[14]; c[18] = 1.0f / c[15]; c[19] = 1.0f / c[16];
        c[20] = p[1] / (kPi * c[14]);
        c[21] = ((kPi * kPi) * c[14]) * c[14];
        c[22] = p[5] / (kPi * c[15]);
        c[23] = ((kPi * kPi) * c[15]) * c[15];
        break;
    default:
        break;
    }
}

// Coupling constants (coupling.rs:114-202 parsing, host-side products as the
// reference forms them). Slot meaning depends on cfn; see pre()/post().
inline void coupling_consts(int cfn, const float* P, int n, int n_src, float* cc, int* classic) {
    for (int i = 0; i < kCC; ++i) cc[i] = 0.0f;
    *classic = 0;
    switch (cfn) {
    case HB_CF_LINEAR: case HB_CF_SCALED_LINEAR:
        cc[0] = n > 0 ? P[0] : 1.0f; cc[1] = n > 1 ? P[1] : 0.0f; break;
    case HB_CF_DIFFERENCE:
        cc[0] = n > 0 ? P[0] : 1.0f; break;
    case HB_CF_KURAMOTO: {
        float a = n > 0 ? P[0] : 1.0f;
        cc[0] = a / (float)n_src;     // coupling.rs:370: a / n_src as f32
        cc[1] = a; break; }
    case HB_CF_SIGMOIDAL: {
        float cmin, cmax, mid, a, sig;
        if (n >= 5) { cmin = P[0]; cmax = P[1]; mid = P[2]; a = P[3]; sig = P[4]; }
        else { cmin = -1.0f; cmax = n > 0 ? P[0] : 1.0f; mid = n > 1 ? P[1] : 0.0f; sig = n > 2 ? P[2] : 1.0f; a = 1.0f; }
        cc[0] = cmin; cc[1] = cmax - cmin; cc[2] = mid; cc[3] = a; cc[4] = sig; break; }
    case HB_CF_TANH:
        cc[0] = n > 0 ? P[0] : 1.0f; cc[1] = n > 1 ? P[1] : 1.0f;
        cc[2] = n > 2 ? P[2] : 0.0f; cc[3] = n > 3 ? P[3] : 1.0f; break;
    case HB_CF_SJR:
        if (n >= 5) { // classic: a, cmin, cmax, r, midpoint
            *classic = 1;
            cc[0] = P[0]; cc[1] = P[1]; cc[2] = P[2] - P[1]; cc[3] = P[3]; cc[4] = P[4];
        } else {      // legacy: a, e0, r, v0
            float a = n > 0 ? P[0] : 1.0f, e0 = n > 1 ? P[1] : 0.005f;
            cc[0] = a; cc[1] = (a * 2.0f) * e0; cc[3] = n > 2 ? P[2] : 0.56f; cc[4] = n > 3 ? P[3] : 6.0f;
        }
        break;
    case HB_CF_PRESIGMOIDAL:
        cc[0] = n > 0 ? P[0] : 1.0f; cc[1] = n > 1 ? P[1] : 1.0f; cc[2] = n > 2 ? P[2] : 1.0f;
        cc[3] = n > 3 ? P[3] : 1.0f; cc[4] = n > 4 ? P[4] : 0.5f; break;
    }
}

// PreSigmoidal with a dynamic threshold (coupling_params[5] != 0,
// coupling.rs:191-198, 238-251) collapses two source cvars into one pre
// channel. The reference rejects it for one-cvar sources (construction.rs:
// 731-740), and for the two-cvar model of this path (MPR) its one-channel
// coupling cannot feed the dfun, which reads coupling channel 1 (dfun.rs:
// 451-452), nor a cvar_map remap, which indexes it with the source's cvar
// stride (engine.rs:637-660). No configuration runs it, so it is rejected
// here rather than ignored. global_t (coupling_params[6]) only acts with a
// dynamic threshold.
inline const char* presigmoidal_unsupported(int cfn, int n_cparams, const float* P) {
    if (cfn == HB_CF_PRESIGMOIDAL && n_cparams > 5 && P[5] != 0.0f)
        return "PreSigmoidal with a dynamic threshold (dynamic != 0) is not supported: it collapses the "
               "source's two cvars into one coupling channel that no model of this path can take";
    return nullptr;
}

// Per-instance device descriptor (one per simulation in a plan).
struct alignas(16) DevInst {
    int32_t model, integrator, n_nodes, n_steps;
    int32_t np, wkind, cfn, delay_mode;
    int32_t single_delay, ring_depth, sweep_form, node_param_idx;
    int32_t bold_period, bold_nout, bold_native, cpu_form;
    int32_t sjr_classic, engine, max_row, blocked_k;   // longest CSR row; delay-block length (0: per step)
    int32_t lpn, lpn_shift, step0, step0_modh;         // lanes per node in the blocked path (1, 2, 4);
                                                       // steps before this run (saturated at 2^30) and
                                                       // step0 % (max_delay + 1) for the batch window
    int32_t hist_len, bold_carry, pad4, pad5;          // carried history steps; BOLD state read from bold_io
    float dt, dt_half, dt_sixth, inv_nsteps;
    float sweep_val, scalar_w, fill0, fill1;
    float mc[kMC];
    float cc[kCC];
    float nscale[8];
    double dt_ms, tr_ms, dtb_ms;
    const float* x0;
    const float* node_mc;      // [N][kMC] per-node model constants, or null
    const int32_t* indptr;
    const uint2* edges;        // per edge: .x = src | delay << 16, .y = weight bits
    const float* dense_t;      // [src][tgt]
    const float* noise;        // raw N(0,1), [T][N][NV]
    float* ring;               // global ring [depth][N][np] when not in LDS
    float* tavg;
    float* fin;
    float* traj;
    int64_t traj_tstride;      // floats between consecutive steps of the trajectory
    float* bold;               // [bold rows][N]
    const float* hist_in;      // [hist_len][N][np] pre() of the steps before step0, oldest first, or null
    float* hist_out;           // same layout for the steps before the run's end, or null
    float* bold_io;            // [N][4] Balloon-Windkessel s, f, v, q in and out, or null
};

// Launch description handed from the planner to a kernel translation unit.
struct LaunchDesc {
    const DevInst* insts;      // device array
    const int32_t* order;      // device array of instance indices, cost-descending
    int32_t n_work;
    int32_t* counter;          // device work counter for persistent launches (or null)
    int32_t grid;              // workgroups
    int32_t block;             // threads per workgroup
    int32_t lds_bytes;         // dynamic LDS per workgroup
    int32_t model;             // model id (every launch holds a single model)
    int32_t np;                // pre channels
    bool lds_ring;             // rings live in LDS
    bool multi;                // nodes per thread > 1 (state table in LDS)
    int32_t mfma;              // 1: dense coupling on the matrix cores, 16 instances per workgroup
                               // (order: tiles of 16 instance indices, -1 padding; n_work tiles)
    int32_t n_nodes, nk;       // mfma: nodes, and nodes padded to a multiple of 4
};

// Dense all-to-all coupling on the matrix cores (hb_dense.hip): instances
// sharing one weight matrix, no delay, FMA mode; at most kDenseMaxNodes nodes,
// and a batch of at least kDenseMinBatch such instances.
constexpr int kDenseMaxNodes = 128, kDenseMinBatch = 8;
// Autotuner choice values of a dense class: the lane loop or the matrix cores.
constexpr int kChoiceLanes = 1, kChoiceMfma = 16;
int dense_lds_bytes(int n_nodes, int np);
hipError_t launch_dense(const LaunchDesc& d, hipStream_t s);
hipError_t dense_set_attributes(int max_lds);

struct BoldShape {
    int native = 1;
    int nout = 0;
    double dtb = 0.0;
};

// Number of BOLD rows a run produces, bold_monitor.rs:72-227 (window times in
// f64, TR resampling count floor(max_t / TR)).
inline BoldShape bold_shape(int T, int P, double dt_ms, double tr_ms) {
    BoldShape b;
    if (P <= 0) return b;
    b.dtb = dt_ms * (double)P;
    const int n_full = T / P, rem = T % P;
    const int n_native = n_full + (rem > 0 ? 1 : 0);
    b.native = (tr_ms <= 0.0 || std::fabs(tr_ms - b.dtb) < 1e-6) ? 1 : 0;
    if (b.native) { b.nout = n_native; return b; }
    if (n_native == 0) { b.nout = 0; return b; }
    double max_t;
    if (rem > 0) max_t = (double)n_full * b.dtb + dt_ms * (double)rem;
    else max_t = (double)(n_full - 1) * b.dtb + b.dtb;
    b.nout = max_t >= tr_ms ? (int)std::floor(max_t / tr_ms) : 0;
    return b;
}

// Owns the device buffers of one host call; frees them on every exit path.
struct DevBufs {
    std::vector<void*> p;
    DevBufs() = default;
    DevBufs(const DevBufs&) = delete;
    DevBufs& operator=(const DevBufs&) = delete;
    ~DevBufs() {
        for (void* q : p)
            if (q) (void)hipFree(q);
    }
    template <typename T> hipError_t alloc(T** out, size_t count) {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, count ? count * sizeof(T) : 1);
        if (e == hipSuccess) p.push_back(q);
        *out = (T*)q;
        return e;
    }
};

// Record the thread-local last-error message (hb_last_error) and return code.
int set_error(int code, const std::string& msg);

// Internal run flag (never set by callers): every instance's trajectory is
// kept on the device (not downloaded) for a follow-up kernel of the plan.
constexpr uint32_t kFlagDeviceTrajectory = 0x40000000u;

// The reference's batch-averaged BOLD monitor (engine.rs:785-801) over a run
// plan whose trajectories are on the device (kFlagDeviceTrajectory): per
// step the mean over the plan's instances of var0, accumulated into windows
// of `period` steps, one Balloon-Windkessel step per window (bw [N][4] host
// in/out), native samples written to native [rows][N] (host).
int plan_bold_batch_mean(hb_plan* P, int period, double dt_ms, float* bw, float* native, bool exact);

// Forward launchers (hb_forward.hip compiled twice: contracted and exact).
hipError_t launch_forward_fast(const LaunchDesc& d, hipStream_t s);
hipError_t launch_forward_exact(const LaunchDesc& d, hipStream_t s);
hipError_t forward_set_attributes_fast(int max_lds);
hipError_t forward_set_attributes_exact(int max_lds);
hipError_t forward_occupancy_fast(const LaunchDesc& d, int* blocks_per_cu);
hipError_t launch_zero_i32(int32_t* p, int n, hipStream_t s);
// Philox normals: for each instance, fill its raw noise buffer.
struct NoiseJob {
    float* out;
    int64_t n;
    uint64_t seed;
    uint64_t stream;
};
hipError_t launch_noise(const NoiseJob* jobs_dev, int n_jobs, int64_t max_n, hipStream_t s);
// Batch-mean BOLD kernels (hb_forward.hip, unit 0): the mean over n_inst
// trajectories (per-instance bases in trajs, [T][N][nv]) of var0 per step
// into zm [T][N], then the windowed monitor per node.
hipError_t launch_bold_mean_fast(const float* const* trajs, int n_inst, int T, int N, int nv, float* zm, int period,
                                 double dt_ms, float* bw, float* native, hipStream_t s);
hipError_t launch_bold_mean_exact(const float* const* trajs, int n_inst, int T, int N, int nv, float* zm, int period,
                                  double dt_ms, float* bw, float* native, hipStream_t s);

}  // namespace hb
