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
template <> struct Tr<HB_EPILEPTOR_CODIM3> { static constexpr int NV = 3, NC = 1; };
template <> struct Tr<HB_EPILEPTOR_CODIM3_SLOWMOD> { static constexpr int NV = 5, NC = 1; };
template <> struct Tr<HB_EPILEPTOR_RS> { static constexpr int NV = 8, NC = 3; };
template <> struct Tr<HB_ZERLAUT_FIRST> { static constexpr int NV = 5, NC = 1; };
template <> struct Tr<HB_ZERLAUT_SECOND> { static constexpr int NV = 8, NC = 1; };
template <> struct Tr<HB_KIONEX> { static constexpr int NV = 5, NC = 1; };
template <> struct Tr<HB_COOMBES_BYRNE> { static constexpr int NV = 4, NC = 4; };
template <> struct Tr<HB_GAST_SD> { static constexpr int NV = 4, NC = 4; };
template <> struct Tr<HB_GAST_SF> { static constexpr int NV = 4, NC = 4; };
template <> struct Tr<HB_LARTER_BREAKSPEAR> { static constexpr int NV = 3, NC = 1; };
template <> struct Tr<HB_EPILEPTOR_2D> { static constexpr int NV = 2, NC = 1; };
template <> struct Tr<HB_EPILEPTOR> { static constexpr int NV = 6, NC = 2; };
template <> struct Tr<HB_RWW_EXC_INH> { static constexpr int NV = 2, NC = 1; };
template <> struct Tr<HB_DECO_BALANCED> { static constexpr int NV = 2, NC = 1; };
template <> struct Tr<HB_ZETTERBERG_JANSEN> { static constexpr int NV = 12, NC = 1; };
template <> struct Tr<HB_REDUCED_FHN> { static constexpr int NV = 4, NC = 2; };
template <> struct Tr<HB_REDUCED_HR> { static constexpr int NV = 6, NC = 2; };
template <> struct Tr<HB_DUMONT_GUTKIN> { static constexpr int NV = 8, NC = 4; };

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
    // dfun.rs:350-354: 2*nu_max / (exp((v0 - v) * r) + 1). Exact builds keep
    // the correctly rounded quotient; FMA builds multiply by the hardware
    // reciprocal (1 ulp; the denominator is >= 1, and an overflowed exp gives 0)
#if HB_EXACT
    return mc[2] / (expf((-v + mc[0]) * mc[1]) + 1.0f);
#else
    return mc[2] * __builtin_amdgcn_rcpf(expf((-v + mc[0]) * mc[1]) + 1.0f);
#endif
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

// ------------------------------------------- models 10-15 (multi path) ---
//
// EpileptorCodim3 (+SlowMod), EpileptorRestingState, Zerlaut first/second
// order and KIonEx, dfun.rs:952-1870. c is the model's kMCX-slot constant
// table (model_consts_ext), read from global memory: slots 0..nparams-1 hold
// the parameters, the host-side products follow.

__device__ __forceinline__ float sgnf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// atan_approx, dfun.rs:1403-1416: |x| > 1 reduced to 1/x, a rational start,
// 12 Newton steps on sin(t) - x cos(t), pi/2 - t for the reduced inputs.
static __device__ __attribute__((noinline)) float atan_ap(float x) {
    const float ax = x < 0.0f ? -x : x;
    const bool red = ax > 1.0f;
    const float xe = red ? 1.0f / x : x;
    const float x2 = xe * xe;
    float th = xe / (x2 * 0.28f + 1.0f);
    for (int it = 0; it < 12; ++it) {
        const float sn = sinf(th), cs = cosf(th);
        th = th - (sn - xe * cs) * cs;
    }
    return red ? -th + 1.57079632679489661923f : th;
}

// acos_approx, dfun.rs:1418-1422
__device__ __forceinline__ float acos_ap(float x) {
    const float xc = x < -0.9999f ? -0.9999f : (x > 0.9999f ? 0.9999f : x);
    return atan_ap(sqrtf((-xc + 1.0f) / (xc + 1.0f))) * 2.0f;
}

// erf_approx, dfun.rs:1424-1446 (Abramowitz-Stegun 7.1.26)
__device__ __forceinline__ float erf_ap(float x) {
    const float ax = x < 0.0f ? -x : x;
    const float t = 1.0f / (ax * 0.3275911f + 1.0f);
    const float t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
    const float y = (((t5 * 1.061405429f + t4 * -1.453152027f) + t3 * 1.421413741f) + t2 * -0.284496736f) +
                    t * 0.254829592f;
    return sgnf(x) * (-(y * expf((ax * -1.0f) * ax)) + 1.0f);
}

// xs on the slow manifold (dfun.rs:996-1019, 1100-1118): trigonometric root
// for a negative discriminant, Cardano otherwise.
static __device__ __attribute__((noinline)) float ec3_xs(float mu1, float mu2, int branch) {
    const float disc = (mu1 * 0.25f) * mu1 - (mu2 * mu2) * (mu2 * (1.0f / 27.0f));
    const float m3 = mu2 / 3.0f;
    const float sr = sqrtf(m3 > kEps ? m3 : kEps);
    const float r32 = m3 * sr;
    float c3 = (mu1 / 2.0f) / (r32 > kEps ? r32 : kEps);
    c3 = c3 < -1.0f ? -1.0f : (c3 > 1.0f ? 1.0f : c3);
    const float th = acos_ap(c3) / 3.0f;
    const float tp3 = 2.0943951023931954923f;   // 2 pi / 3 in f32
    const float xt = (sr * 2.0f) * cosf(branch == 2 ? th - tp3 : (branch == 3 ? th + tp3 : th));
    const float sd = sqrtf(disc > 0.0f ? disc : 0.0f);
    const float h = mu1 / 2.0f, pa = h + sd, na = h - sd, third = 1.0f / 3.0f;
    const float cp = sgnf(pa) * powf(pa < 0.0f ? -pa : pa, third);
    const float cn = sgnf(na) * powf(na < 0.0f ? -na : na, third);
    return disc < 0.0f ? xt : cp + cn;
}

__device__ __forceinline__ float ec3_mod(float z) {   // 0.1 (z - 0.5)^7
    const float zs = z + -0.5f;
    return ((((((zs * zs) * zs) * zs) * zs) * zs) * zs) * 0.1f;
}

// get_fluct_regime_vars, dfun.rs:1449-1488
__device__ __forceinline__ void zl_fluct(const float* __restrict__ c, float fe, float fi, float fe_ext, float fi_ext,
                                         float w, float gle, float& mu_v, float& sigma_v, float& t_v) {
    const float fe_tot = (fe + 1e-6f) * c[50] + fe_ext * c[20];
    const float fi_tot = (fi + 1e-6f) * c[51] + fi_ext * c[21];
    const float ge = fe_tot * c[52], gi = fi_tot * c[53];
    const float mu_g = (ge + gi) + c[0];
    const float t_m = (1.0f / mu_g) * c[3];
    mu_v = (((ge * c[10] + gi * c[11]) + (-w)) + gle) / mu_g;
    const float ue = ((-mu_v + c[10]) * c[12]) / mu_g;
    const float ui = ((-mu_v + c[11]) * c[13]) / mu_g;
    const float uet = ue * c[14], uit = ui * c[15];
    const float ue2 = uet * uet, ui2 = uit * uit;
    const float sv2 = (fe_tot * ue2) / ((t_m + c[14]) * 2.0f) + (fi_tot * ui2) / ((t_m + c[15]) * 2.0f);
    sigma_v = sqrtf(sv2);
    if (!(sigma_v >= 1e-10f)) sigma_v = sigma_v != sigma_v ? sigma_v : 1e-10f;   // clamp keeps NaN
    const float num = fe_tot * ue2 + fi_tot * ui2;
    float den = (fe_tot * ue2) / (t_m + c[14]) + (fi_tot * ui2) / (t_m + c[15]);
    if (!(den >= 1e-10f)) den = den != den ? den : 1e-10f;
    t_v = num / den;
}

// TF: threshold polynomial (dfun.rs:1490-1513), firing rate (1515-1523), 1525-1544
static __device__ __attribute__((noinline)) float zl_tf(const float* __restrict__ c, float fe, float fi, float fe_ext,
                                                 float fi_ext, float w, float gle, int poff) {
    float mu_v, sigma_v, t_v;
    zl_fluct(c, fe, fi, fe_ext, fi_ext, w, gle, mu_v, sigma_v, t_v);
    const float* __restrict__ P = c + poff;
    const float tvn = t_v * c[56];
    const float v = (mu_v + 60.0f) / 10.0f, sg = (sigma_v + -4.0f) / 6.0f, t = tvn + -0.5f;
    float thr = (v * v) * P[4] + (sg * sg) * P[5];
    thr = thr + (t * t) * P[6];
    thr = thr + (v * sg) * P[7];
    thr = thr + (v * t) * P[8];
    thr = thr + (sg * t) * P[9];
    thr = thr + v * P[1];
    thr = thr + sg * P[2];
    thr = thr + t * P[3];
    thr = (thr + P[0]) * 1000.0f;
    const float arg = (thr - mu_v) / (sigma_v * 1.41421356237309504880f);
    float den = t_v * 2.0f;
    den = den > kEps ? den : (den != den ? den : kEps);
    return (-erf_ap(arg) + 1.0f) / den;
}

__device__ __forceinline__ float kx_sig(float u) { return 1.0f / (expf(-u) + 1.0f); }
__device__ __forceinline__ float kx_ln(float u) { const float l = logf(u); return l > 1e-10f ? l : (l != l ? l : 1e-10f); }

template <> __device__ __forceinline__ void dfun<HB_EPILEPTOR_CODIM3>(const float* c, bool, float, const float* x,
                                                                     const float* cpl, float* dx) {
    // dfun.rs:985-1032
    const float cz = cosf(x[2]), sz = sinf(x[2]);
    const float mu2 = (cz * c[13] + sz * c[16]) * c[7];
    const float mu1 = (cz * c[14] + sz * c[17]) * (-c[7]);
    const float nu = (cz * c[15] + sz * c[18]) * c[7];
    const float xs = ec3_xs(mu1, mu2, (int)c[11]);
    const float dd = sqrtf((x[0] - xs) * (x[0] - xs) + x[1] * x[1]);
    float t = -(-dd + c[9]);
    dx[0] = -x[1];
    dx[1] = (((x[0] * x[0]) * x[0] - mu2 * x[0]) - mu1) - x[1] * ((nu + x[0] * c[6]) + x[0] * x[0]);
    if (c[12] > 0.5f) t = t + ec3_mod(x[2]);
    dx[2] = -((t + cpl[0] * c[10]) * c[8]);
}

template <> __device__ __forceinline__ void dfun<HB_EPILEPTOR_CODIM3_SLOWMOD>(const float* c, bool, float,
                                                                             const float* x, const float* cpl,
                                                                             float* dx) {
    // dfun.rs:1035-1134
    const float cua = cosf(x[3]), sua = sinf(x[3]), cub = cosf(x[4]), sub = sinf(x[4]);
    const float a0 = cua * c[27] + sua * c[30], a1 = cua * c[28] + sua * c[31], a2 = cua * c[29] + sua * c[32];
    const float b0 = cub * c[33] + sub * c[36], b1 = cub * c[34] + sub * c[37], b2 = cub * c[35] + sub * c[38];
    float an = sqrtf((a0 * a0 + a1 * a1) + a2 * a2);
    an = an > kEps ? an : (an != an ? an : kEps);
    const float e0 = a0 / an, e1 = a1 / an, e2 = a2 / an;
    const float c0 = a1 * b2 - a2 * b1, c1 = a2 * b0 - a0 * b2, c2 = a0 * b1 - a1 * b0;
    float f0 = c1 * a2 - c2 * a1, f1 = c2 * a0 - c0 * a2, f2 = c0 * a1 - c1 * a0;
    float fn = sqrtf((f0 * f0 + f1 * f1) + f2 * f2);
    fn = fn > kEps ? fn : (fn != fn ? fn : kEps);
    f0 = f0 / fn; f1 = f1 / fn; f2 = f2 / fn;
    const float cz = cosf(x[2]), sz = sinf(x[2]);
    const float mu2 = (cz * e0 + sz * f0) * c[7];
    const float mu1 = -((cz * e1 + sz * f1) * c[7]);
    const float nu = (cz * e2 + sz * f2) * c[7];
    const float xs = ec3_xs(mu1, mu2, (int)c[11]);
    dx[0] = -x[1];
    dx[1] = (((x[0] * x[0]) * x[0] - mu2 * x[0]) - mu1) - x[1] * ((nu + x[0] * c[6]) + x[0] * x[0]);
    const float dd = sqrtf((x[0] - xs) * (x[0] - xs) + x[1] * x[1]);
    const float t = -(-dd + c[9]) + (c[12] > 0.5f ? ec3_mod(x[2]) : 0.0f);
    dx[2] = -((t + cpl[0] * c[10]) * c[8]);
    dx[3] = 0.0f + c[25];
    dx[4] = 0.0f + c[26];
}

template <> __device__ __forceinline__ void dfun<HB_EPILEPTOR_RS>(const float* c, bool, float, const float* x,
                                                                 const float* cpl, float* dx) {
    // dfun.rs:1137-1209
    const float x1 = x[0], y1 = x[1], z = x[2], x2 = x[3], y2 = x[4], g = x[5], xr = x[6], yr = x[7];
    float f1;
    if (x1 < 0.0f) f1 = (-(x1 * x1)) * c[3] + x1 * c[4];
    else { const float z4 = z + -4.0f; f1 = ((-x2) + c[8]) + (z4 * z4) * 0.6f; }
    dx[0] = ((((y1 - z) + cpl[0] * c[12]) + f1 * x1) + c[0]) * c[15];
    dx[1] = (((-y1) + c[5]) - (x1 * x1) * c[6]) * c[15];
    const float zm = z * -0.1f;
    const float yz = z < 0.0f ? -((((((zm * zm) * zm) * zm) * zm) * zm) * zm) : 0.0f;
    const float h = c[16] > 0.5f ? (x1 + 3.0f) / (expf(-((x1 + 0.5f) * 10.0f)) + 1.0f) + yz : (x1 * 4.0f + c[28]) + yz;
    dx[2] = ((h - z) + cpl[0] * c[14]) * c[29];
    dx[3] = ((((((x2 - (x2 * x2) * x2) - y2) + c[1]) + g * c[11]) - (z + -3.5f) * 0.3f) + cpl[1] * c[13]) * c[15];
    const float xp = x2 + 0.25f;
    const float f2 = xp < 0.0f ? 0.0f : xp * c[10];
    dx[4] = ((f2 - y2) * c[15]) / c[9];
    dx[5] = ((-x1) * 0.1f + g) * c[30];
    dx[6] = ((((yr * c[24] - ((xr * xr) * xr) * c[23]) + (xr * xr) * c[22]) + c[33]) + cpl[2] * c[34]) * c[31];
    dx[7] = ((xr * c[20] - yr * c[25]) + c[19]) * c[32];
}

template <> __device__ __forceinline__ void dfun<HB_ZERLAUT_FIRST>(const float* c, bool, float, const float* x,
                                                                  const float* cpl, float* dx) {
    // dfun.rs:1546-1608
    const float e = x[0], i = x[1], we = x[2], wi = x[3], ou = x[4];
    float fx = cpl[0] + ou * c[28];
    if (fx * c[20] < 0.0f) fx = 0.0f;
    const float ee = fx + c[23], ei = 0.0f + c[24], ie = fx + c[25], ii = 0.0f + c[26];
    const float tfe = zl_tf(c, e, i, ee, ei, we, c[54], 30);
    const float tfi = zl_tf(c, e, i, ie, ii, wi, c[55], 40);
    dx[0] = (tfe - e) / c[22];
    dx[1] = (tfi - i) / c[22];
    float mv, s_, t_;
    zl_fluct(c, e, i, ee, ei, we, c[54], mv, s_, t_);
    dx[2] = ((-we) / c[8] + e * c[4]) + ((mv + (-c[1])) * c[5]) / c[8];
    zl_fluct(c, e, i, ie, ii, wi, c[55], mv, s_, t_);
    dx[3] = ((-wi) / c[9] + i * c[6]) + ((mv + (-c[2])) * c[7]) / c[9];
    dx[4] = (-ou) / c[27];
}

template <> __device__ __forceinline__ void dfun<HB_ZERLAUT_SECOND>(const float* c, bool, float, const float* x,
                                                                   const float* cpl, float* dx) {
    // dfun.rs:1610-1780
    const float e = x[0], i = x[1], cee = x[2], cei = x[3], cii = x[4], we = x[5], wi = x[6], ou = x[7];
    float ex = cpl[0] + (ou * c[28] + c[23]);
    if (ex < 0.0f) ex = 0.0f;
    float ix = cpl[0] * c[29] + (ou * c[28] + c[25]);
    if (ix < 0.0f) ix = 0.0f;
    const float iex = 0.0f + c[24], iix = 0.0f + c[26];
    const float df = 1e-4f, ep = e + df, em = e + (-df), ip = i + df, im = i + (-df);
    const float a2 = c[60], q2 = c[61];
    const float tfe = zl_tf(c, e, i, ex, iex, we, c[54], 30);
    const float tfi = zl_tf(c, e, i, ix, iix, wi, c[55], 40);
    const float e_p0 = zl_tf(c, ep, i, ex, iex, we, c[54], 30), e_n0 = zl_tf(c, em, i, ex, iex, we, c[54], 30);
    const float e_0p = zl_tf(c, e, ip, ex, iex, we, c[54], 30), e_0n = zl_tf(c, e, im, ex, iex, we, c[54], 30);
    const float e_pp = zl_tf(c, ep, ip, ex, iex, we, c[54], 30), e_pn = zl_tf(c, ep, im, ex, iex, we, c[54], 30);
    const float e_np = zl_tf(c, em, ip, ex, iex, we, c[54], 30), e_nn = zl_tf(c, em, im, ex, iex, we, c[54], 30);
    const float i_p0 = zl_tf(c, ep, i, ix, iix, wi, c[55], 40), i_n0 = zl_tf(c, em, i, ix, iix, wi, c[55], 40);
    const float i_0p = zl_tf(c, e, ip, ix, iix, wi, c[55], 40), i_0n = zl_tf(c, e, im, ix, iix, wi, c[55], 40);
    const float i_pp = zl_tf(c, ep, ip, ix, iix, wi, c[55], 40), i_pn = zl_tf(c, ep, im, ix, iix, wi, c[55], 40);
    const float i_np = zl_tf(c, em, ip, ix, iix, wi, c[55], 40), i_nn = zl_tf(c, em, im, ix, iix, wi, c[55], 40);
    const float dfe_e = (e_p0 - e_n0) * a2, dfi_e = (e_0p - e_0n) * a2;
    const float dfe_i = (i_p0 - i_n0) * a2, dfi_i = (i_0p - i_0n) * a2;
    const float fefe_e = ((e_p0 - tfe * 2.0f) + e_n0) * q2, fifi_e = ((e_0p - tfe * 2.0f) + e_0n) * q2;
    const float fefi_e = ((e_pp - e_np) * a2 - (e_pn - e_nn) * a2) * a2;
    const float fife_e = ((e_pp - e_pn) * a2 - (e_np - e_nn) * a2) * a2;
    const float fefe_i = ((i_p0 - tfi * 2.0f) + i_n0) * q2, fifi_i = ((i_0p - tfi * 2.0f) + i_0n) * q2;
    const float fefi_i = ((i_pp - i_np) * a2 - (i_pn - i_nn) * a2) * a2;
    const float fife_i = ((i_pp - i_pn) * a2 - (i_np - i_nn) * a2) * a2;
    dx[0] = (((((tfe - e) + cee * (fefe_e * 0.5f)) + cei * (fefi_e * 0.5f)) + cei * (fife_e * 0.5f)) +
             cii * (fifi_e * 0.5f)) / c[22];
    dx[1] = (((((tfi - i) + cee * (fefe_i * 0.5f)) + cei * (fefi_i * 0.5f)) + cei * (fife_i * 0.5f)) +
             cii * (fifi_i * 0.5f)) / c[22];
    dx[2] = ((((tfe * ((c[57] - tfe) / c[58]) + (tfe - e) * (tfe - e)) + cee * (dfe_e * 2.0f)) + cei * (dfi_e * 2.0f)) -
             cee * 2.0f) / c[22];
    dx[3] = ((((((tfe - e) * (tfi - i) + cee * dfe_e) + cei * dfe_i) + cei * dfi_e) + cii * dfi_i) - cei * 2.0f) / c[22];
    dx[4] = ((((tfi * ((c[57] - tfi) / c[59]) + (tfi - i) * (tfi - i)) + cii * (dfi_i * 2.0f)) + cei * (dfe_i * 2.0f)) -
             cii * 2.0f) / c[22];
    float mv, s_, t_;
    zl_fluct(c, e, i, ex, iex, we, c[54], mv, s_, t_);
    dx[5] = ((-we) / c[8] + e * c[4]) + ((mv + (-c[1])) * c[5]) / c[8];
    zl_fluct(c, e, i, ix, iix, wi, c[55], mv, s_, t_);
    dx[6] = ((-wi) / c[9] + i * c[6]) + ((mv + (-c[2])) * c[7]) / c[9];
    dx[7] = (-ou) / c[27];
}

template <> __device__ __forceinline__ void dfun<HB_KIONEX>(const float* c, bool, float, const float* x,
                                                           const float* cpl, float* dx) {
    // dfun.rs:1782-1869
    const float xx = x[0], v = x[1], n = x[2], dki = x[3], kg = x[4];
    const float m_inf = kx_sig((v + 24.0f) * (1.0f / 12.0f));
    const float n_inf = kx_sig((v + 19.0f) * (1.0f / 18.0f));
    const float h_na = -(1.0f / (expf((n + -0.4f) * -8.0f) + 1.0f)) + 1.1f;
    const float bv = 2160.0f / 720.0f;
    const float dna_i = -dki, dna_o = dna_i * (-bv), dk_o = dki * (-bv);
    const float ki = (-dki) + 130.0f, nai = dna_i + 16.0f, nao = dna_o + 138.0f, ko = (dk_o + 4.8f) + kg;
    const float i_k = (n * 22.0f + 0.12f) * (v - (kx_ln(ko) - kx_ln(ki)) * 26.64f);
    const float i_na = ((m_inf * h_na) * 40.0f + 0.02f) * (v - (kx_ln(nao) - kx_ln(nai)) * 26.64f);
    const float i_cl = (v + c[19]) * 7.5f;
    const float i_pump = (kx_sig(nai + -21.0f) * kx_sig(ko + -5.5f)) * 250.0f;
    const float v_dot = (((i_na + i_k) + i_cl) + i_pump) * c[14];
    const float r_x = (xx * c[6]) / kPi;
    const float syn = (cpl[0] * c[15]) * ((-v) + c[0]);
    if (v < c[9]) {
        dx[0] = ((((v + (-c[5])) * c[16]) * xx) + c[4]) - r_x * (xx * c[2]);
        dx[1] = v_dot - xx * (((xx * c[6]) + c[3]) + syn);
    } else {
        dx[0] = ((((v + (-c[7])) * c[17]) * xx) + c[4]) - r_x * (xx * c[2]);
        dx[1] = v_dot - xx * (((xx * c[8]) + c[3]) + syn);
    }
    dx[2] = (n_inf - n) / c[11];
    dx[3] = (i_k - i_pump * 2.0f) * c[18];
    dx[4] = ((-ko) + c[1]) * c[13];
}

// CoombesByrne, dfun.rs:609-637 (c[5] = delta/pi, c[6] = pi^2, c[7] = k pi)
template <> __device__ __forceinline__ void dfun<HB_COOMBES_BYRNE>(const float* c, bool, float, const float* x,
                                                                  const float* cpl, float* dx) {
    const float r = x[0], v = x[1], g = x[2], q = x[3];
    dx[0] = ((((v * r) * 2.0f) + c[5]) - g * r) + cpl[0];
    dx[1] = (((v * v) + c[4]) - r * (r * c[6])) + ((-v) + c[2]) * g;
    dx[2] = q * c[1];
    dx[3] = (((r * c[7]) - g) - q * 2.0f) * c[1];
}

// GastSchmidtKnoscheSD / SF, dfun.rs:639-716. The reference multiplies v by the
// whole bracket (method calls bind before the * operator there); so does this.
template <int M> __device__ __forceinline__ void gast_dfun(const float* c, const float* x, const float* cpl,
                                                           float* dx) {
    const float r = x[0], v = x[1], a = x[2], b = x[3];
    dx[0] = (((r * v) * 2.0f) + c[12]) * c[11];
    float br = (v - r * (r * c[13])) + c[14];
    if constexpr (M == HB_GAST_SD) br = br + (r * c[15]) * ((-a) + 1.0f);
    else br = (br + r * c[15]) - a;
    dx[1] = (v * ((br + cpl[0] * c[7]) + cpl[1] * c[8])) * c[11];
    dx[2] = b * c[16];
    dx[3] = (((b * -2.0f) - a) + r * c[2]) * c[16];
}
template <> __device__ __forceinline__ void dfun<HB_GAST_SD>(const float* c, bool, float, const float* x,
                                                            const float* cpl, float* dx) {
    gast_dfun<HB_GAST_SD>(c, x, cpl, dx);
}
template <> __device__ __forceinline__ void dfun<HB_GAST_SF>(const float* c, bool, float, const float* x,
                                                            const float* cpl, float* dx) {
    gast_dfun<HB_GAST_SF>(c, x, cpl, dx);
}

// LarterBreakspear, dfun.rs:718-766
template <> __device__ __forceinline__ void dfun<HB_LARTER_BREAKSPEAR>(const float* c, bool, float, const float* x,
                                                                      const float* cpl, float* dx) {
    const float V = x[0], W = x[1], Z = x[2];
    const float m_ca = 0.5f * (tanhf((V + (-c[9])) * c[32]) + 1.0f);
    const float m_na = 0.5f * (tanhf((V + (-c[10])) * c[33]) + 1.0f);
    const float m_k = 0.5f * (tanhf((V + (-c[11])) * c[34]) + 1.0f);
    const float qv = c[35] * (tanhf((V + (-c[26])) * c[36]) + 1.0f);
    const float qz = c[37] * (tanhf((Z + (-c[27])) * c[38]) + 1.0f);
    const float tca = (c[0] + qv * c[40]) + cpl[0] * c[39];
    const float tna = ((m_na * c[3]) + qv * c[41]) + cpl[0] * c[42];
    dx[0] = ((((((-tca) * m_ca) * (V + (-c[5])) - (W * c[1]) * (V + (-c[6]))) - (V + (-c[7])) * c[2]) -
              tna * (V + (-c[8]))) - (Z * c[18]) * qz) * c[30] + c[43];
    dx[1] = (m_k - W) * c[44];
    dx[2] = ((V * c[17]) * qv + c[45]) * c[46];
}

// the Epileptor slow-variable term: -(-0.1 z)^7 for z < 0, else 0
__device__ __forceinline__ float epi_ydot(float z) {
    const float zm = z * -0.1f;
    return z < 0.0f ? -((((((zm * zm) * zm) * zm) * zm) * zm) * zm) : 0.0f;
}

// Epileptor2D, dfun.rs:776-818
template <> __device__ __forceinline__ void dfun<HB_EPILEPTOR_2D>(const float* c, bool, float, const float* x,
                                                                 const float* cpl, float* dx) {
    const float x1 = x[0], z = x[1];
    float f1;
    if (x1 < 0.0f) f1 = (x1 * x1) * c[2] + x1 * c[12];
    else { const float z4 = z + -4.0f; f1 = ((z4 * z4) * -0.6f + (-c[4])) + x1 * c[6]; }
    dx[0] = ((((-z) + c[13]) + cpl[0] * c[8]) - f1 * x1) * c[10];
    const float yz = epi_ydot(z);
    const float h = c[11] > 0.5f ? (x1 + 3.0f) / (expf(-((x1 + 0.5f) * 10.0f)) + 1.0f) + yz : (x1 * 4.0f + c[14]) + yz;
    dx[1] = ((h - z) + cpl[0] * c[9]) * c[15];
}

// Epileptor, dfun.rs:820-874
template <> __device__ __forceinline__ void dfun<HB_EPILEPTOR>(const float* c, bool, float, const float* x,
                                                              const float* cpl, float* dx) {
    const float x1 = x[0], y1 = x[1], z = x[2], x2 = x[3], y2 = x[4], g = x[5];
    float f1;
    if (x1 < 0.0f) f1 = (-(x1 * x1)) * c[3] + x1 * c[4];
    else { const float z4 = z + -4.0f; f1 = ((-x2) + c[8]) + (z4 * z4) * 0.6f; }
    dx[0] = ((((y1 - z) + cpl[0] * c[12]) + f1 * x1) + c[0]) * c[15];
    dx[1] = (((-((x1 * x1) * c[6])) + c[5]) - y1) * c[15];
    const float yz = epi_ydot(z);
    const float h = c[16] > 0.5f ? (x1 + 3.0f) / (expf(-((x1 + 0.5f) * 10.0f)) + 1.0f) + yz : (x1 * 4.0f + c[17]) + yz;
    dx[2] = ((h - z) + cpl[0] * c[14]) * c[18];
    dx[3] = ((((((x2 - (x2 * x2) * x2) - y2) + c[1]) + g * c[11]) - (z + -3.5f) * 0.3f) + cpl[1] * c[13]) * c[15];
    const float xp = x2 + 0.25f;
    const float f2 = xp < 0.0f ? 0.0f : xp * c[10];
    dx[4] = ((f2 - y2) * c[15]) / c[9];
    dx[5] = ((-x1) * 0.1f + g) * c[19];
}

// ReducedWongWangExcInh / DecoBalancedExcInh, dfun.rs:876-950
template <int M> __device__ __forceinline__ void rwei_dfun(const float* c, const float* x, const float* cpl,
                                                           float* dx) {
    const float se = x[0], si = x[1], cc = cpl[0] * c[19], jse = se * c[6];
    const float xe = (((jse * c[5]) - si * c[13]) + c[20]) + cc;
    const float xi = ((jse - si) + c[22]) + cc * c[18];
    float ae, ai;
    if constexpr (M == HB_DECO_BALANCED) { ae = xe * c[24] + c[25]; ai = xi * c[26] + c[27]; }
    else { ae = xe * c[0] + (-c[1]); ai = xi * c[8] + (-c[9]); }
    const float he = ae / ((-expf(ae * (-c[2]))) + 1.0f);
    const float hi = ai / ((-expf(ai * (-c[10]))) + 1.0f);
    dx[0] = (-se) * c[21] + (((-se) + 1.0f) * he) * c[3];
    dx[1] = (-si) * c[23] + hi * c[11];
}
template <> __device__ __forceinline__ void dfun<HB_RWW_EXC_INH>(const float* c, bool, float, const float* x,
                                                                const float* cpl, float* dx) {
    rwei_dfun<HB_RWW_EXC_INH>(c, x, cpl, dx);
}
template <> __device__ __forceinline__ void dfun<HB_DECO_BALANCED>(const float* c, bool, float, const float* x,
                                                                  const float* cpl, float* dx) {
    rwei_dfun<HB_DECO_BALANCED>(c, x, cpl, dx);
}

// ZetterbergJansen, dfun.rs:1212-1268; sigmoid 2 e0 / (exp((rho_2 - v) rho_1) + 1)
template <> __device__ __forceinline__ void dfun<HB_ZETTERBERG_JANSEN>(const float* c, bool, float, const float* x,
                                                                      const float* cpl, float* dx) {
    const float s0 = c[18] / (expf(((-x[0]) + c[5]) * c[6]) + 1.0f);
    const float s23 = c[18] / (expf(((-(x[2] - x[4])) + c[5]) * c[6]) + 1.0f);
    dx[0] = x[1];
    dx[1] = ((s23 * c[23]) - x[1] * c[28]) - x[0] * c[21];
    dx[2] = x[3];
    dx[3] = ((((s0 * c[24]) + c[15]) + cpl[0]) * c[19] - x[3] * c[28]) - x[2] * c[21];
    dx[4] = x[5];
    dx[5] = (((s0 * c[25]) + c[17]) * c[20] - x[5] * c[29]) - x[4] * c[22];
    dx[6] = x[7];
    dx[7] = ((s23 * c[26]) - x[7] * c[28]) - x[6] * c[21];
    dx[8] = x[9];
    dx[9] = ((s0 * c[27]) - x[9] * c[29]) - x[8] * c[22];
    dx[10] = x[3] - x[5];
    dx[11] = x[7] - x[9];
}

// ReducedSetFitzHughNagumo, dfun.rs:1270-1300
template <> __device__ __forceinline__ void dfun<HB_REDUCED_FHN>(const float* c, bool, float, const float* x,
                                                                const float* cpl, float* dx) {
    const float xi = x[0], eta = x[1], al = x[2], be = x[3];
    const float xi3 = (xi * xi) * xi, al3 = (al * al) * al;
    dx[0] = ((((xi - xi3 / 3.0f) - eta) * c[0] + (-xi) * c[3]) - (al - xi) * c[4]) + cpl[0] * c[0];
    dx[1] = ((xi - eta * c[2]) + c[1]) / c[0];
    dx[2] = (((al - al3 / 3.0f) - be) * c[0] + (-xi) * c[5]) + cpl[1] * c[0];
    dx[3] = ((al - be * c[2]) + c[1]) / c[0];
}

// ReducedSetHindmarshRose, dfun.rs:1302-1348
template <> __device__ __forceinline__ void dfun<HB_REDUCED_HR>(const float* c, bool, float, const float* x,
                                                               const float* cpl, float* dx) {
    const float xi = x[0], eta = x[1], ts = x[2], al = x[3], be = x[4], ga = x[5];
    const float xi2 = xi * xi, xi3 = (xi * xi) * xi, al2 = al * al, al3 = (al * al) * al;
    dx[0] = (((((eta - xi3 * c[1]) + xi2 * c[2]) - ts) + (-xi) * c[7]) - (al - xi) * c[8]) + cpl[0];
    dx[1] = ((-xi2) * c[4] - eta) + c[3];
    dx[2] = (xi * c[12] - ts * c[0]) + c[13];
    dx[3] = ((((be - al3 * c[1]) + al2 * c[2]) - ga) + (-xi) * c[9]) + cpl[1];
    dx[4] = ((-al2) * c[4] - be) + c[3];
    dx[5] = (al * c[12] - ga * c[0]) + c[13];
}

// DumontGutkin, dfun.rs:1350-1400 (v times the bracket, as the reference writes it)
template <> __device__ __forceinline__ void dfun<HB_DUMONT_GUTKIN>(const float* c, bool, float, const float* x,
                                                                  const float* cpl, float* dx) {
    const float re = x[0], ve = x[1], see = x[2], sei = x[3], ri = x[4], vi = x[5], sie = x[6], sii = x[7];
    dx[0] = (((ve * 2.0f) * re) + c[20]) * c[17];
    dx[1] = (ve * (((((((ve - (re * re) * c[21]) + c[2]) + see * c[14]) - sei * c[14]) + c[0]) + cpl[0]) +
                   cpl[1] * 0.0f)) * c[17];
    dx[2] = (((-see) + re * c[9]) + cpl[2]) * c[19];
    dx[3] = ((-sei) + ri * c[10]) * c[19];
    dx[4] = (((vi * 2.0f) * ri) + c[22]) * c[18];
    dx[5] = (vi * ((((((vi - (ri * ri) * c[23]) + c[6]) + sie * c[15]) - sii * c[15]) + c[4]) + cpl[2] * c[13])) * c[18];
    dx[6] = (((-sie) + re * c[11]) + cpl[2] * c[13]) * c[19];
    dx[7] = ((-sii) + ri * c[12]) * c[19];
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
    } else if constexpr (M == HB_COOMBES_BYRNE_2D || M == HB_KIONEX || M == HB_COOMBES_BYRNE || M == HB_GAST_SD ||
                         M == HB_GAST_SF) {   // dfun.rs:190-209, 232-236: r, x >= 0
        x[0] = x[0] < 0.0f ? 0.0f : x[0];
    } else if constexpr (M == HB_RWW_EXC_INH || M == HB_DECO_BALANCED) {   // dfun.rs:172-176: both in [0, 1]
        x[0] = x[0] < 0.0f ? 0.0f : (x[0] > 1.0f ? 1.0f : x[0]);
        x[1] = x[1] < 0.0f ? 0.0f : (x[1] > 1.0f ? 1.0f : x[1]);
    } else if constexpr (M == HB_DUMONT_GUTKIN) {   // dfun.rs:210-218: r_e, r_i >= 0
        x[0] = x[0] < 0.0f ? 0.0f : x[0];
        x[4] = x[4] < 0.0f ? 0.0f : x[4];
    } else if constexpr (M == HB_ZERLAUT_FIRST || M == HB_ZERLAUT_SECOND) {   // dfun.rs:219-230: E, I in [0, 1]
        x[0] = x[0] < 0.0f ? 0.0f : (x[0] > 1.0f ? 1.0f : x[0]);
        x[1] = x[1] < 0.0f ? 0.0f : (x[1] > 1.0f ? 1.0f : x[1]);
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
#if HB_EXACT
        const float inv = 1.0f / den;
#else
        const float inv = __builtin_amdgcn_rcpf(den);   // FMA builds: the 1-ulp reciprocal (den >= 1)
#endif
        if (classic) out[0] = cpu ? cc[1] + cc[2] / den : inv * cc[2] + cc[1];
        else out[0] = cpu ? cc[1] / den : inv * cc[1];
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
