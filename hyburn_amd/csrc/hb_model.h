// hb_model.h -- per-node model arithmetic shared by the forward and adjoint
// kernels: the six dfuns, clamps, coupling pre/post and one Balloon-Windkessel
// step. Included inside a kernel namespace (hb::fast, hb::exact, hb::adj) by
// each translation unit, so every unit compiles it with its own FP contraction
// setting. f32 op order follows the reference expressions (file:line given at
// each function, in maedoc/hyburn).
//
// No include guard: each includer instantiates it in its own namespace.

template <int M> struct Tr;
template <> struct Tr<HB_G2DO> { static constexpr int NV = 2, NC = 1; };
template <> struct Tr<HB_MPR> { static constexpr int NV = 2, NC = 2; };
template <> struct Tr<HB_JANSEN_RIT> { static constexpr int NV = 6, NC = 1; };
template <> struct Tr<HB_WILSON_COWAN> { static constexpr int NV = 2, NC = 1; };
template <> struct Tr<HB_KURAMOTO> { static constexpr int NV = 1, NC = 1; };
template <> struct Tr<HB_RWW> { static constexpr int NV = 1, NC = 1; };
template <> struct Tr<HB_LINEAR> { static constexpr int NV = 1, NC = 1; };
template <> struct Tr<HB_SUP_HOPF> { static constexpr int NV = 2, NC = 2; };
template <> struct Tr<HB_HOPFIELD> { static constexpr int NV = 2, NC = 2; };
template <> struct Tr<HB_COOMBES_BYRNE_2D> { static constexpr int NV = 2, NC = 2; };

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

// dfun.rs:522-532: gamma * x + c0
template <> __device__ __forceinline__ void dfun<HB_LINEAR>(const float* mc, bool sw, float sv,
                                                            const float* x, const float* c, float* dx) {
    dx[0] = x[0] * mc[0] + c[0];
}

// dfun.rs:534-557: amp = x^2 + y^2; dx = (a - amp) x - omega y + c0,
// dy = (a - amp) y + omega x + c1
template <> __device__ __forceinline__ void dfun<HB_SUP_HOPF>(const float* mc, bool sw, float sv,
                                                              const float* x, const float* c, float* dx) {
    const float amp = x[0] * x[0] + x[1] * x[1];
    const float g = -amp + mc[0];
    dx[0] = (g * x[0] - x[1] * mc[1]) + c[0];
    dx[1] = (g * x[1] + x[0] * mc[1]) + c[1];
}

// hopfield_dfun_batch (dfun.rs): dx = (-x + c0) / taux; dtheta = (-theta + c1) / tauT
// when dynamic, else 0
template <> __device__ __forceinline__ void dfun<HB_HOPFIELD>(const float* mc, bool sw, float sv,
                                                              const float* x, const float* c, float* dx) {
    dx[0] = (-x[0] + c[0]) * mc[0];
    dx[1] = mc[2] != 0.0f ? (-x[1] + c[1]) * mc[1] : 0.0f;
}

// coombes_byrne2d_dfun_batch: dr = ((v (2 r) + delta/pi) - r^2 k pi) + c_r;
// dv = ((-v^2 + eta) - r^2 pi^2) + ((-v + v_syn) k pi) r   (c_v unused)
template <> __device__ __forceinline__ void dfun<HB_COOMBES_BYRNE_2D>(const float* mc, bool sw, float sv,
                                                                      const float* x, const float* c,
                                                                      float* dx) {
    const float r = x[0], v = x[1], r2 = r * r;
    dx[0] = ((v * (r * 2.0f) + mc[0]) - r2 * mc[1]) + c[0];
    dx[1] = ((-(v * v) + mc[2]) - r2 * mc[3]) + ((-v + mc[4]) * mc[1]) * r;
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
    } else if constexpr (M == HB_COOMBES_BYRNE_2D) {   // dfun.rs:190-194: r clamped at 0
        x[0] = x[0] < 0.0f ? 0.0f : x[0];
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

// ---------------------------------------------------------------- stepping -

struct StepCtx {
    float dt, dth, dt6;
    int integ;
    bool stoch, plain, sw;
    float sv;
};

// One integration step of one node with frozen coupling (engine.rs:698-773,
// integrator.rs:296-315 for the plain engine's RK4-stochastic order).
template <int M>
__device__ __forceinline__ void integrate(const StepCtx& S, const float* mc, const float* cpl,
                                          const float* z, float* xs) {
    constexpr int NV = Tr<M>::NV;
    float k[NV], acc[NV], s[NV];
    dfun<M>(mc, S.sw, S.sv, xs, cpl, k);
    if (S.integ == HB_EULER || S.integ == HB_EULER_STOCH) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const float y = xs[v] + k[v] * S.dt;
            xs[v] = S.stoch ? y + z[v] : y;
        }
        clampm<M>(xs);
    } else if (S.integ == HB_HEUN || S.integ == HB_HEUN_STOCH) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const float y = xs[v] + k[v] * S.dt;
            s[v] = S.stoch ? y + z[v] : y;
            acc[v] = k[v];
        }
        clampm<M>(s);
        dfun<M>(mc, S.sw, S.sv, s, cpl, k);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const float y = xs[v] + (acc[v] + k[v]) * S.dth;
            xs[v] = S.stoch ? y + z[v] : y;
        }
        clampm<M>(xs);
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) { acc[v] = k[v]; s[v] = xs[v] + k[v] * S.dth; }
        clampm<M>(s);
        dfun<M>(mc, S.sw, S.sv, s, cpl, k);
#pragma unroll
        for (int v = 0; v < NV; ++v) { acc[v] = acc[v] + k[v] * 2.0f; s[v] = xs[v] + k[v] * S.dth; }
        clampm<M>(s);
        dfun<M>(mc, S.sw, S.sv, s, cpl, k);
#pragma unroll
        for (int v = 0; v < NV; ++v) { acc[v] = acc[v] + k[v] * 2.0f; s[v] = xs[v] + k[v] * S.dt; }
        clampm<M>(s);
        dfun<M>(mc, S.sw, S.sv, s, cpl, k);
#pragma unroll
        for (int v = 0; v < NV; ++v) xs[v] = xs[v] + (acc[v] + k[v]) * S.dt6;
        if (S.stoch) {
            if (S.plain) {
                clampm<M>(xs);
#pragma unroll
                for (int v = 0; v < NV; ++v) xs[v] = xs[v] + z[v];
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) xs[v] = xs[v] + z[v];
                clampm<M>(xs);
            }
        } else {
            clampm<M>(xs);
        }
    }
}

// Integrator class at compile time: 0 Euler(-Maruyama), 1 Heun, 2 RK4; the
// stochastic and plain-engine variants are selects on the same code.
constexpr int integ_class(int integ) {
    return (integ == HB_EULER || integ == HB_EULER_STOCH) ? 0 : (integ == HB_HEUN || integ == HB_HEUN_STOCH) ? 1 : 2;
}

template <int M, int IC>
__device__ __forceinline__ void integrate_c(const StepCtx& S, const float* mc, const float* cpl,
                                            const float* z, float* xs) {
    constexpr int NV = Tr<M>::NV;
    float k[NV], acc[NV], s[NV];
    dfun<M>(mc, S.sw, S.sv, xs, cpl, k);
    if constexpr (IC == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const float y = xs[v] + k[v] * S.dt;
            xs[v] = S.stoch ? y + z[v] : y;
        }
        clampm<M>(xs);
    } else if constexpr (IC == 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const float y = xs[v] + k[v] * S.dt;
            s[v] = S.stoch ? y + z[v] : y;
            acc[v] = k[v];
        }
        clampm<M>(s);
        dfun<M>(mc, S.sw, S.sv, s, cpl, k);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const float y = xs[v] + (acc[v] + k[v]) * S.dth;
            xs[v] = S.stoch ? y + z[v] : y;
        }
        clampm<M>(xs);
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) { acc[v] = k[v]; s[v] = xs[v] + k[v] * S.dth; }
        clampm<M>(s);
        dfun<M>(mc, S.sw, S.sv, s, cpl, k);
#pragma unroll
        for (int v = 0; v < NV; ++v) { acc[v] = acc[v] + k[v] * 2.0f; s[v] = xs[v] + k[v] * S.dth; }
        clampm<M>(s);
        dfun<M>(mc, S.sw, S.sv, s, cpl, k);
#pragma unroll
        for (int v = 0; v < NV; ++v) { acc[v] = acc[v] + k[v] * 2.0f; s[v] = xs[v] + k[v] * S.dt; }
        clampm<M>(s);
        dfun<M>(mc, S.sw, S.sv, s, cpl, k);
#pragma unroll
        for (int v = 0; v < NV; ++v) xs[v] = xs[v] + (acc[v] + k[v]) * S.dt6;
        // clamp(det + Z) in the batch engine; clamp(det) + Z (no final clamp) in
        // the plain engine (integrator.rs:296-315); deterministic: clamp(det)
        float zz[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) zz[v] = xs[v] + z[v];
        if (S.stoch && !S.plain) {
#pragma unroll
            for (int v = 0; v < NV; ++v) xs[v] = zz[v];
        }
        clampm<M>(xs);
        if (S.stoch && S.plain) {
#pragma unroll
            for (int v = 0; v < NV; ++v) xs[v] = xs[v] + z[v];
        }
    }
}
