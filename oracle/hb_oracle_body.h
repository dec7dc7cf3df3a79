/*
 * CPU restatement of hyburn's batched neural-mass integration path.
 * TEST INFRASTRUCTURE ONLY: this is the parity oracle for the HIP library.
 * It is included twice by hb_oracle.c, once with REAL=float (the FP32 twin,
 * same op order as the reference's NdArray<f32> path) and once with
 * REAL=double (the FP64 twin used to state each FP32 result's error).
 *
 * Every function cites the reference file:line it restates.
 */

/* ---- model constants, precomputed exactly as the reference does on host ---- */

typedef struct {
    int model;
    /* G2DO  (src/engine/batch_engine/dfun.rs:271-319) */
    REAL g_a, g_b, g_c, g_e, g_f, g_g, g_alpha, g_beta, g_gamma, g_i, g_dtau, g_dovt;
    /* MPR   (dfun.rs:439-478) */
    REAL m_delta, m_eta, m_j, m_iext, m_cr, m_cv, m_inv_tau, m_dr_c, m_pp, m_jtau, m_eta_i;
    /* JR    (dfun.rs:321-384) */
    REAL j_v0, j_r, j_2nu, j_Aa, j_2a, j_aa, j_a1J, j_a2J, j_a3J, j_a4J, j_Bb, j_2b, j_bb, j_mu;
    /* WC    (dfun.rs:386-437) */
    REAL w_cee, w_cei, w_cie, w_cii, w_inv_te, w_inv_ti, w_ae, w_be, w_ce, w_the, w_ai, w_bi,
         w_ci, w_thi, w_re, w_ri, w_ke, w_ki, w_p, w_q, w_ale, w_ali, w_off_e, w_off_i;
    /* Kuramoto (dfun.rs:480-492) */
    REAL k_omega;
    /* Linear, SupHopf (dfun.rs:522-557), Hopfield (hopfield_dfun_batch) */
    REAL l_gamma, h_a, h_omega, p_itx, p_itt, cb_dpi, cb_kpi, cb_eta, cb_pi2, cb_vsyn;
    int p_dyn;
    /* RWW   (dfun.rs:494-516) */
    REAL r_a, r_b, r_d, r_gamma, r_inv_taus, r_wjn, r_jn, r_io;
    /* EpileptorCodim3 ... KIonEx (models 10-15): raw parameters and the
     * host-side products the reference forms from them, FN(ext_consts) */
    REAL xc[64];
    int sweep_form;   /* 1: reference's batch-native swept form (dfun.rs:49-72) */
    REAL sweep_val;
} FN(mconst);

/* Constants of models 10-15 (slot layout shared with the GPU's constant
 * table): slots 0..nparams-1 are the parameters, the products the reference
 * evaluates once per call on the host follow. */
static void FN(ext_consts)(REAL* c, int model, const REAL* p)
{
    int i;
    const REAL eps = EPSR;
    for (i = 0; i < 64; ++i) c[i] = 0;
    for (i = 0; i < hb_orc_nparams(model); ++i) c[i] = p[i];
    switch (model) {
    case ORC_EC3: { /* dfun.rs:958-983: unit start/stop points, the great circle's e and f */
        REAL st[3], sp[3], a[3], b[3], cr[3], cca[3], sn, pn, cn, cc2;
        st[0] = p[1]; st[1] = -p[0]; st[2] = p[2];
        sp[0] = p[4]; sp[1] = -p[3]; sp[2] = p[5];
        sn = SQRTR(((REAL)0 + st[0] * st[0] + st[1] * st[1]) + st[2] * st[2]);
        pn = SQRTR(((REAL)0 + sp[0] * sp[0] + sp[1] * sp[1]) + sp[2] * sp[2]);
        for (i = 0; i < 3; ++i) { a[i] = st[i] / (sn > eps ? sn : eps); b[i] = sp[i] / (pn > eps ? pn : eps); }
        cr[0] = a[1] * b[2] - a[2] * b[1];
        cr[1] = a[2] * b[0] - a[0] * b[2];
        cr[2] = a[0] * b[1] - a[1] * b[0];
        cn = SQRTR(((REAL)0 + cr[0] * cr[0] + cr[1] * cr[1]) + cr[2] * cr[2]);
        for (i = 0; i < 3; ++i) c[13 + i] = a[i];
        if (cn > eps) {
            for (i = 0; i < 3; ++i) cca[i] = cr[(i + 1) % 3] * a[(i + 2) % 3] - cr[(i + 2) % 3] * a[(i + 1) % 3];
            cc2 = SQRTR(((REAL)0 + cca[0] * cca[0] + cca[1] * cca[1]) + cca[2] * cca[2]);
            cc2 = cc2 > eps ? cc2 : eps;
            for (i = 0; i < 3; ++i) c[16 + i] = cca[i] / cc2;
        } else {
            c[16] = 0; c[17] = 1; c[18] = 0;
        }
        break; }
    case ORC_EC3S: /* dfun.rs:1063-1074: g*R, h*R, l*R, m*R */
        for (i = 0; i < 3; ++i) {
            c[27 + i] = p[13 + i] * p[7]; c[30 + i] = p[19 + i] * p[7];
            c[33 + i] = p[16 + i] * p[7]; c[36 + i] = p[22 + i] * p[7];
        }
        break;
    case ORC_ERS: /* dfun.rs:1143-1208 */
        c[17] = p[17] > eps ? p[17] : eps;
        c[28] = (REAL)-4.0 * p[2];
        c[29] = p[15] * p[7];
        c[30] = (REAL)-0.01 * p[15];
        c[31] = p[21] * c[17];
        c[32] = p[21] / c[17];
        c[33] = p[26] * p[18];
        c[34] = p[26] * p[27];
        break;
    case ORC_ZL1: case ORC_ZL2: /* dfun.rs:1449-1620 host products */
        c[50] = (((REAL)1 - p[19]) * p[17]) * p[16];
        c[51] = (p[19] * p[18]) * p[16];
        c[52] = p[12] * p[14];
        c[53] = p[13] * p[15];
        c[54] = p[0] * p[1];
        c[55] = p[0] * p[2];
        c[56] = p[0] / p[3];
        c[57] = (REAL)1 / p[22];
        c[58] = p[16] * ((REAL)1 - p[19]);
        c[59] = p[16] * p[19];
        c[60] = (REAL)1 / (((REAL)2 * (REAL)1e-4f) * (REAL)1e3f);
        c[61] = (REAL)1 / ((((REAL)1e-4f * (REAL)1e3f) * (REAL)1e-4f) * (REAL)1e3f);
        break;
    case ORC_KIX: { /* dfun.rs:1788-1849 */
        REAL pi = PIR;
        c[10] = p[10] > eps ? p[10] : eps;
        c[11] = p[11] > eps ? p[11] : eps;
        c[14] = (REAL)-1 / c[10];
        c[15] = p[6] / pi;
        c[16] = (REAL)2 * p[6];
        c[17] = (REAL)2 * p[8];
        c[18] = -(p[12] / (REAL)2160);
        c[19] = (REAL)26.64f * (REAL)LOGR((REAL)112 / (REAL)5);
        break; }
    case ORC_CB: { /* dfun.rs:609-637: delta/pi, pi*pi, k*pi */
        REAL pi = PIR;
        c[5] = p[0] / pi; c[6] = pi * pi; c[7] = p[3] * pi;
        break; }
    case ORC_GSD: case ORC_GSF: { /* dfun.rs:639-716: tau, tau_a (eps), 1/tau, delta/(pi tau), pi^2 tau^2, eta+I, J tau, 1/tau_a */
        REAL pi = PIR;
        c[9] = p[0] > eps ? p[0] : eps;
        c[10] = p[1] > eps ? p[1] : eps;
        c[11] = (REAL)1 / c[9];
        c[12] = p[4] / (pi * c[9]);
        c[13] = ((pi * pi) * c[9]) * c[9];
        c[14] = p[6] + p[3];
        c[15] = p[5] * c[9];
        c[16] = (REAL)1 / c[10];
        break; }
    case ORC_LB: /* dfun.rs:718-766 */
        c[31] = p[31] > eps ? p[31] : eps;
        c[32] = (REAL)1 / p[12]; c[33] = (REAL)1 / p[13]; c[34] = (REAL)1 / p[14];
        c[35] = (REAL)0.5f * p[28]; c[36] = (REAL)1 / p[15];
        c[37] = (REAL)0.5f * p[29]; c[38] = (REAL)1 / p[16];
        c[39] = (p[23] * p[25]) * p[19];
        c[40] = (((REAL)1 - p[23]) * p[25]) * p[19];
        c[41] = ((REAL)1 - p[23]) * p[19];
        c[42] = p[23] * p[19];
        c[43] = p[20] * p[24];
        c[44] = ((p[30] * p[4]) * p[22]) / c[31];
        c[45] = p[21] * p[24];
        c[46] = p[30] * p[22];
        break;
    case ORC_E2D: /* dfun.rs:776-818 */
        c[12] = p[6] - p[3];
        c[13] = p[5] + p[0];
        c[14] = (REAL)-4 * p[1];
        c[15] = p[10] * p[7];
        break;
    case ORC_EPI: /* dfun.rs:820-874 */
        c[17] = (REAL)-4 * p[2];
        c[18] = p[15] * p[7];
        c[19] = (REAL)-0.01f * p[15];
        break;
    case ORC_RWEI: case ORC_DECO: { /* dfun.rs:884-950 */
        REAL te = p[4] > eps ? p[4] : eps, ti = p[12] > eps ? p[12] : eps;
        c[19] = p[17] * p[6];
        c[20] = p[7] * p[15] + p[16];
        c[21] = (REAL)1 / te;
        c[22] = p[14] * p[15];
        c[23] = (REAL)1 / ti;
        if (model == ORC_DECO) {
            c[24] = p[0] * p[19]; c[25] = (-p[1]) * p[19];
            c[26] = p[8] * p[19]; c[27] = (-p[9]) * p[19];
        }
        break; }
    case ORC_ZJ: /* dfun.rs:1212-1268 */
        c[18] = (REAL)2 * p[4];
        c[19] = p[0] * p[2]; c[20] = p[1] * p[3];
        c[21] = p[2] * p[2]; c[22] = p[3] * p[3];
        c[23] = c[19] * p[7]; c[24] = p[13] * p[8]; c[25] = p[14] * p[9];
        c[26] = c[19] * p[10]; c[27] = c[20] * p[11];
        c[28] = (REAL)2 * p[2]; c[29] = (REAL)2 * p[3];
        break;
    case ORC_RHR: /* dfun.rs:1302-1348 */
        c[12] = p[0] * p[5];
        c[13] = (-p[0]) * p[6];
        break;
    case ORC_DG: { /* dfun.rs:1350-1400 */
        REAL pi = PIR;
        c[14] = p[3] > eps ? p[3] : eps;
        c[15] = p[7] > eps ? p[7] : eps;
        c[16] = p[8] > eps ? p[8] : eps;
        c[17] = (REAL)1 / c[14]; c[18] = (REAL)1 / c[15]; c[19] = (REAL)1 / c[16];
        c[20] = p[1] / (pi * c[14]);
        c[21] = ((pi * pi) * c[14]) * c[14];
        c[22] = p[5] / (pi * c[15]);
        c[23] = ((pi * pi) * c[15]) * c[15];
        break; }
    }
}

/* Build the constants from a parameter vector (f32 values, as in the config).
 * The FP32 twin evaluates each host-side constant in f32 with the reference's
 * operation order; the FP64 twin uses the same expressions in double. */
static void FN(mconst_init)(FN(mconst)* k, int model, const float* pf, int sweep_form, float sweep_val)
{
    REAL p[ORC_MAXP];
    int i;
    for (i = 0; i < ORC_MAXP; ++i) p[i] = 0;
    for (i = 0; i < hb_orc_nparams(model); ++i) p[i] = (REAL)pf[i];
    memset(k, 0, sizeof(*k));
    k->model = model;
    k->sweep_form = sweep_form;
    k->sweep_val = (REAL)sweep_val;
    switch (model) {
    case ORC_G2DO: { /* dfun.rs:277-293 */
        REAL tau = p[0] > EPSR ? p[0] : EPSR;
        k->g_i = p[1]; k->g_a = p[2]; k->g_b = p[3]; k->g_c = p[4];
        k->g_e = p[6]; k->g_f = p[7]; k->g_g = p[8];
        k->g_alpha = p[9]; k->g_beta = p[10]; k->g_gamma = p[11];
        k->g_dtau = p[5] * tau;
        k->g_dovt = p[5] / tau;
        break; }
    case ORC_MPR: { /* dfun.rs:445-456 */
        REAL tau = p[0] > EPSR ? p[0] : EPSR;
        REAL pi = PIR;
        k->m_delta = p[1]; k->m_eta = p[2]; k->m_j = p[3]; k->m_iext = p[4];
        k->m_cr = p[5]; k->m_cv = p[6];
        k->m_inv_tau = (REAL)1 / tau;
        k->m_dr_c = k->m_delta / (pi * tau);
        k->m_pp = pi * pi * tau * tau;
        k->m_jtau = k->m_j * tau;
        k->m_eta_i = k->m_eta + k->m_iext;
        break; }
    case ORC_JR: { /* dfun.rs:327-366 */
        REAL Ap = p[0], Bp = p[1], a = p[2], b = p[3], J = p[7];
        k->j_v0 = p[4]; k->j_r = p[6];
        k->j_2nu = (REAL)2 * p[5];
        k->j_Aa = Ap * a; k->j_2a = (REAL)2 * a; k->j_aa = a * a;
        k->j_a1J = p[8] * J; k->j_a2J = p[9] * J; k->j_a3J = p[10] * J; k->j_a4J = p[11] * J;
        k->j_Bb = Bp * b; k->j_2b = (REAL)2 * b; k->j_bb = b * b;
        k->j_mu = p[12];
        break; }
    case ORC_WC: { /* dfun.rs:392-403 */
        REAL te = p[4] > EPSR ? p[4] : EPSR;
        REAL ti = p[5] > EPSR ? p[5] : EPSR;
        k->w_cee = p[0]; k->w_cei = p[1]; k->w_cie = p[2]; k->w_cii = p[3];
        k->w_inv_te = (REAL)1 / te; k->w_inv_ti = (REAL)1 / ti;
        k->w_ae = p[6]; k->w_be = p[7]; k->w_ce = p[8]; k->w_the = p[9];
        k->w_ai = p[10]; k->w_bi = p[11]; k->w_ci = p[12]; k->w_thi = p[13];
        k->w_re = p[14]; k->w_ri = p[15]; k->w_ke = p[16]; k->w_ki = p[17];
        k->w_p = p[18]; k->w_q = p[19]; k->w_ale = p[20]; k->w_ali = p[21];
        k->w_off_e = (REAL)1 / ((REAL)1 + EXPR(k->w_ae * k->w_be));
        k->w_off_i = (REAL)1 / ((REAL)1 + EXPR(k->w_ai * k->w_bi));
        break; }
    case ORC_LIN:
        k->l_gamma = p[0];
        break;
    case ORC_SH:
        k->h_a = p[0]; k->h_omega = p[1];
        break;
    case ORC_HOP: {
        REAL e32 = (REAL)1.1920929e-07f;
        k->p_itx = (REAL)1 / (p[0] > e32 ? p[0] : e32);
        k->p_itt = (REAL)1 / (p[1] > e32 ? p[1] : e32);
        k->p_dyn = p[2] > (REAL)0.5;
        break; }
    case ORC_CB2: { /* coombes_byrne2d_dfun_batch, f32 host products */
        REAL pi = PIR;
        k->cb_dpi = p[0] / pi; k->cb_kpi = p[2] * pi; k->cb_eta = p[3]; k->cb_pi2 = pi * pi; k->cb_vsyn = p[1];
        break; }
    case ORC_KUR:
        k->k_omega = p[0];
        break;
    case ORC_RWW: /* dfun.rs:500-502 */
        k->r_a = p[0]; k->r_b = p[1]; k->r_d = p[2]; k->r_gamma = p[3];
        k->r_inv_taus = (REAL)1 / p[4];
        k->r_wjn = p[5] * p[6]; k->r_jn = p[6]; k->r_io = p[7];
        break;
    default:
        FN(ext_consts)(k->xc, model, p);
        break;
    }
}

/* JR sigmoid, dfun.rs:350-354: 2*nu_max / (exp((v0 - v)*r) + 1). */
static REAL FN(jr_sigm)(const FN(mconst)* k, REAL v)
{
    REAL arg = (-v + k->j_v0) * k->j_r;
    REAL den = EXPR(arg) + (REAL)1;
    return k->j_2nu / den;
}

/* ---- models 10-15 (EpileptorCodim3 ... KIonEx), dfun.rs:952-1870 ---- */

static REAL FN(sgn)(REAL x) { return x > 0 ? (REAL)1 : (x < 0 ? (REAL)-1 : (REAL)0); }

/* atan_approx, dfun.rs:1403-1416: reduce |x| > 1 to 1/x, a rational start,
 * 12 Newton steps on sin(t) - x cos(t), pi/2 - t for the reduced inputs. */
static REAL FN(atan_ap)(REAL x)
{
    REAL ax = x < 0 ? -x : x, xe, x2, th;
    int red = ax > (REAL)1, it;
    xe = red ? (REAL)1 / x : x;
    x2 = xe * xe;
    th = xe / (x2 * (REAL)0.28f + (REAL)1);
    for (it = 0; it < 12; ++it) {
        REAL sn = SINR(th), cs = COSR(th);
        th = th - (sn - xe * cs) * cs;
    }
    return red ? -th + (REAL)1.57079632679489661923f : th;
}

/* acos_approx, dfun.rs:1418-1422 */
static REAL FN(acos_ap)(REAL x)
{
    REAL xc = x < (REAL)-0.9999f ? (REAL)-0.9999f : (x > (REAL)0.9999f ? (REAL)0.9999f : x);
    REAL ratio = (-xc + (REAL)1) / (xc + (REAL)1);
    return FN(atan_ap)(SQRTR(ratio)) * (REAL)2;
}

/* erf_approx, dfun.rs:1424-1446 (Abramowitz-Stegun 7.1.26) */
static REAL FN(erf_ap)(REAL x)
{
    REAL ax = x < 0 ? -x : x;
    REAL t = (REAL)1 / (ax * (REAL)0.3275911f + (REAL)1);
    REAL t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
    REAL y = (((t5 * (REAL)1.061405429f + t4 * (REAL)-1.453152027f) + t3 * (REAL)1.421413741f) +
              t2 * (REAL)-0.284496736f) + t * (REAL)0.254829592f;
    REAL ev = EXPR((ax * (REAL)-1) * ax);
    return FN(sgn)(x) * (-(y * ev) + (REAL)1);
}

/* The fixed point xs on the slow manifold (dfun.rs:996-1019, 1100-1118):
 * trigonometric root when the discriminant is negative, Cardano otherwise. */
static REAL FN(ec3_xs)(REAL mu1, REAL mu2, int branch)
{
    REAL disc = (mu1 * (REAL)0.25f) * mu1 - (mu2 * mu2) * (mu2 * ((REAL)1.0f / (REAL)27.0f));
    REAL m3 = mu2 / (REAL)3, sr, r32, c3, th, xt, sd, h, pa, na, third, cp, cn;
    const REAL tp3 = (REAL)2.0943951023931954923f;   /* 2 pi / 3 in f32 */
    sr = SQRTR(m3 > EPSR ? m3 : EPSR);
    r32 = m3 * sr;
    c3 = (mu1 / (REAL)2) / (r32 > EPSR ? r32 : EPSR);
    c3 = c3 < (REAL)-1 ? (REAL)-1 : (c3 > (REAL)1 ? (REAL)1 : c3);
    th = FN(acos_ap)(c3) / (REAL)3;
    if (branch == 2) xt = (sr * (REAL)2) * COSR(th - tp3);
    else if (branch == 3) xt = (sr * (REAL)2) * COSR(th + tp3);
    else xt = (sr * (REAL)2) * COSR(th);
    sd = SQRTR(disc > 0 ? disc : (REAL)0);
    h = mu1 / (REAL)2;
    pa = h + sd;
    na = h - sd;
    third = (REAL)1.0f / (REAL)3.0f;
    cp = FN(sgn)(pa) * POWR(pa < 0 ? -pa : pa, third);
    cn = FN(sgn)(na) * POWR(na < 0 ? -na : na, third);
    return disc < 0 ? xt : cp + cn;
}

/* z - 0.5 to the seventh, times 0.1 (the modification term, dfun.rs:1024, 1124) */
static REAL FN(ec3_mod)(REAL z)
{
    REAL zs = z + (REAL)-0.5f;
    return ((((((zs * zs) * zs) * zs) * zs) * zs) * zs) * (REAL)0.1f;
}

/* get_fluct_regime_vars, dfun.rs:1449-1488 */
static void FN(zl_fluct)(const REAL* c, REAL fe, REAL fi, REAL fe_ext, REAL fi_ext, REAL w, REAL gle,
                         REAL* mu_v, REAL* sigma_v, REAL* t_v)
{
    REAL fe_tot = (fe + (REAL)1e-6f) * c[50] + fe_ext * c[20];
    REAL fi_tot = (fi + (REAL)1e-6f) * c[51] + fi_ext * c[21];
    REAL ge = fe_tot * c[52], gi = fi_tot * c[53];
    REAL mu_g = (ge + gi) + c[0];
    REAL t_m = ((REAL)1 / mu_g) * c[3];
    REAL ue, ui, uet, uit, ue2, ui2, sv2, num, den;
    *mu_v = (((ge * c[10] + gi * c[11]) + (-w)) + gle) / mu_g;
    ue = ((-*mu_v + c[10]) * c[12]) / mu_g;
    ui = ((-*mu_v + c[11]) * c[13]) / mu_g;
    uet = ue * c[14]; uit = ui * c[15];
    ue2 = uet * uet; ui2 = uit * uit;
    sv2 = (fe_tot * ue2) / ((t_m + c[14]) * (REAL)2) + (fi_tot * ui2) / ((t_m + c[15]) * (REAL)2);
    *sigma_v = SQRTR(sv2);
    if (!(*sigma_v >= (REAL)1e-10f)) *sigma_v = *sigma_v != *sigma_v ? *sigma_v : (REAL)1e-10f;
    num = fe_tot * ue2 + fi_tot * ui2;
    den = (fe_tot * ue2) / (t_m + c[14]) + (fi_tot * ui2) / (t_m + c[15]);
    if (!(den >= (REAL)1e-10f)) den = den != den ? den : (REAL)1e-10f;
    *t_v = num / den;
}

/* TF: threshold polynomial (1490-1513), firing rate (1515-1523), dfun.rs:1525-1544 */
static REAL FN(zl_tf)(const REAL* c, REAL fe, REAL fi, REAL fe_ext, REAL fi_ext, REAL w, REAL gle, const REAL* P)
{
    REAL mu_v, sigma_v, t_v, tvn, v, sg, t, thr, arg, den;
    FN(zl_fluct)(c, fe, fi, fe_ext, fi_ext, w, gle, &mu_v, &sigma_v, &t_v);
    tvn = t_v * c[56];
    v = (mu_v + (REAL)60) / (REAL)10;
    sg = (sigma_v + (REAL)-4) / (REAL)6;
    t = tvn + (REAL)-0.5f;
    thr = (v * v) * P[4] + (sg * sg) * P[5];
    thr = thr + (t * t) * P[6];
    thr = thr + (v * sg) * P[7];
    thr = thr + (v * t) * P[8];
    thr = thr + (sg * t) * P[9];
    thr = thr + v * P[1];
    thr = thr + sg * P[2];
    thr = thr + t * P[3];
    thr = thr + P[0];
    thr = thr * (REAL)1000;
    arg = (thr - mu_v) / (sigma_v * (REAL)1.41421356237309504880f);
    den = t_v * (REAL)2;
    den = den > EPSR ? den : (den != den ? den : EPSR);
    return (-FN(erf_ap)(arg) + (REAL)1) / den;
}

/* KIonEx logistic helpers: 1 / (exp(-u) + 1) */
static REAL FN(kx_sig)(REAL u) { return (REAL)1 / (EXPR(-u) + (REAL)1); }
static REAL FN(kx_ln)(REAL u) { REAL l = LOGR(u); return l > (REAL)1e-10f ? l : (l != l ? l : (REAL)1e-10f); }

static void FN(dfun_ext)(const FN(mconst)* k, const REAL* x, const REAL* cpl, REAL* dx)
{
    const REAL* c = k->xc;
    switch (k->model) {
    case ORC_EC3: { /* dfun.rs:985-1032 */
        REAL cz = COSR(x[2]), sz = SINR(x[2]);
        REAL mu2 = (cz * c[13] + sz * c[16]) * c[7];
        REAL mu1 = (cz * c[14] + sz * c[17]) * (-c[7]);
        REAL nu = (cz * c[15] + sz * c[18]) * c[7];
        REAL xs = FN(ec3_xs)(mu1, mu2, (int)c[11]);
        REAL dd = SQRTR((x[0] - xs) * (x[0] - xs) + x[1] * x[1]);
        REAL t = -(-dd + c[9]);
        dx[0] = -x[1];
        dx[1] = ((((x[0] * x[0]) * x[0] - mu2 * x[0]) - mu1) - x[1] * ((nu + x[0] * c[6]) + x[0] * x[0]));
        if (c[12] > (REAL)0.5f) t = t + FN(ec3_mod)(x[2]);
        dx[2] = -((t + cpl[0] * c[10]) * c[8]);
        break; }
    case ORC_EC3S: { /* dfun.rs:1035-1134 */
        REAL cua = COSR(x[3]), sua = SINR(x[3]), cub = COSR(x[4]), sub = SINR(x[4]);
        REAL a0 = cua * c[27] + sua * c[30], a1 = cua * c[28] + sua * c[31], a2 = cua * c[29] + sua * c[32];
        REAL b0 = cub * c[33] + sub * c[36], b1 = cub * c[34] + sub * c[37], b2 = cub * c[35] + sub * c[38];
        REAL an = SQRTR((a0 * a0 + a1 * a1) + a2 * a2), e0, e1, e2, c0, c1, c2, f0, f1, f2, fn_;
        REAL cz, sz, mu2, mu1, nu, xs, dd, t;
        an = an > EPSR ? an : (an != an ? an : EPSR);
        e0 = a0 / an; e1 = a1 / an; e2 = a2 / an;
        c0 = a1 * b2 - a2 * b1; c1 = a2 * b0 - a0 * b2; c2 = a0 * b1 - a1 * b0;
        f0 = c1 * a2 - c2 * a1; f1 = c2 * a0 - c0 * a2; f2 = c0 * a1 - c1 * a0;
        fn_ = SQRTR((f0 * f0 + f1 * f1) + f2 * f2);
        fn_ = fn_ > EPSR ? fn_ : (fn_ != fn_ ? fn_ : EPSR);
        f0 = f0 / fn_; f1 = f1 / fn_; f2 = f2 / fn_;
        cz = COSR(x[2]); sz = SINR(x[2]);
        mu2 = (cz * e0 + sz * f0) * c[7];
        mu1 = -((cz * e1 + sz * f1) * c[7]);
        nu = (cz * e2 + sz * f2) * c[7];
        xs = FN(ec3_xs)(mu1, mu2, (int)c[11]);
        dx[0] = -x[1];
        dx[1] = ((((x[0] * x[0]) * x[0] - mu2 * x[0]) - mu1) - x[1] * ((nu + x[0] * c[6]) + x[0] * x[0]));
        dd = SQRTR((x[0] - xs) * (x[0] - xs) + x[1] * x[1]);
        t = -(-dd + c[9]) + (c[12] > (REAL)0.5f ? FN(ec3_mod)(x[2]) : (REAL)0);
        dx[2] = -((t + cpl[0] * c[10]) * c[8]);
        dx[3] = (REAL)0 + c[25];
        dx[4] = (REAL)0 + c[26];
        break; }
    case ORC_ERS: { /* dfun.rs:1137-1209 */
        REAL x1 = x[0], y1 = x[1], z = x[2], x2 = x[3], y2 = x[4], g = x[5], xr = x[6], yr = x[7];
        REAL f1, z4, zm, yz, h, f2, xp;
        if (x1 < 0) f1 = (-(x1 * x1)) * c[3] + x1 * c[4];
        else { z4 = z + (REAL)-4; f1 = ((-x2) + c[8]) + (z4 * z4) * (REAL)0.6f; }
        dx[0] = ((((y1 - z) + cpl[0] * c[12]) + f1 * x1) + c[0]) * c[15];
        dx[1] = (((-y1) + c[5]) - (x1 * x1) * c[6]) * c[15];
        zm = z * (REAL)-0.1f;
        yz = z < 0 ? -((((((zm * zm) * zm) * zm) * zm) * zm) * zm) : (REAL)0;
        if (c[16] > (REAL)0.5f) h = (x1 + (REAL)3) / (EXPR(-((x1 + (REAL)0.5f) * (REAL)10)) + (REAL)1) + yz;
        else h = (x1 * (REAL)4 + c[28]) + yz;
        dx[2] = ((h - z) + cpl[0] * c[14]) * c[29];
        dx[3] = ((((((x2 - (x2 * x2) * x2) - y2) + c[1]) + g * c[11]) - (z + (REAL)-3.5f) * (REAL)0.3f) +
                 cpl[1] * c[13]) * c[15];
        xp = x2 + (REAL)0.25f;
        f2 = xp < 0 ? (REAL)0 : xp * c[10];
        dx[4] = ((f2 - y2) * c[15]) / c[9];
        dx[5] = ((-x1) * (REAL)0.1f + g) * c[30];
        dx[6] = ((((yr * c[24] - ((xr * xr) * xr) * c[23]) + (xr * xr) * c[22]) + c[33]) + cpl[2] * c[34]) * c[31];
        dx[7] = ((xr * c[20] - yr * c[25]) + c[19]) * c[32];
        break; }
    case ORC_ZL1: { /* dfun.rs:1546-1608 */
        REAL e = x[0], i = x[1], we = x[2], wi = x[3], ou = x[4];
        REAL fx = cpl[0] + ou * c[28], tfe, tfi, mve, mvi, s_, t_;
        if (fx * c[20] < 0) fx = 0;
        tfe = FN(zl_tf)(c, e, i, fx + c[23], (REAL)0 + c[24], we, c[54], c + 30);
        tfi = FN(zl_tf)(c, e, i, fx + c[25], (REAL)0 + c[26], wi, c[55], c + 40);
        dx[0] = (tfe - e) / c[22];
        dx[1] = (tfi - i) / c[22];
        FN(zl_fluct)(c, e, i, fx + c[23], (REAL)0 + c[24], we, c[54], &mve, &s_, &t_);
        dx[2] = ((-we) / c[8] + e * c[4]) + ((mve + (-c[1])) * c[5]) / c[8];
        FN(zl_fluct)(c, e, i, fx + c[25], (REAL)0 + c[26], wi, c[55], &mvi, &s_, &t_);
        dx[3] = ((-wi) / c[9] + i * c[6]) + ((mvi + (-c[2])) * c[7]) / c[9];
        dx[4] = (-ou) / c[27];
        break; }
    case ORC_ZL2: { /* dfun.rs:1610-1780 */
        REAL e = x[0], i = x[1], cee = x[2], cei = x[3], cii = x[4], we = x[5], wi = x[6], ou = x[7];
        REAL ex = cpl[0] + (ou * c[28] + c[23]), ix, iex = (REAL)0 + c[24], iix = (REAL)0 + c[26];
        REAL df = (REAL)1e-4f, ep = e + df, em = e + (-df), ip = i + df, im = i + (-df);
        REAL tfe, tfi, e_p0, e_n0, e_0p, e_0n, e_pp, e_pn, e_np, e_nn, i_p0, i_n0, i_0p, i_0n, i_pp, i_pn, i_np, i_nn;
        REAL a2 = c[60], q2 = c[61], dfe_e, dfi_e, dfe_i, dfi_i, fefe_e, fifi_e, fefi_e, fife_e, fefe_i, fifi_i,
             fefi_i, fife_i, mve, mvi, s_, t_;
        if (ex < 0) ex = 0;
        ix = cpl[0] * c[29] + (ou * c[28] + c[25]);
        if (ix < 0) ix = 0;
        tfe = FN(zl_tf)(c, e, i, ex, iex, we, c[54], c + 30);
        tfi = FN(zl_tf)(c, e, i, ix, iix, wi, c[55], c + 40);
        e_p0 = FN(zl_tf)(c, ep, i, ex, iex, we, c[54], c + 30);
        e_n0 = FN(zl_tf)(c, em, i, ex, iex, we, c[54], c + 30);
        e_0p = FN(zl_tf)(c, e, ip, ex, iex, we, c[54], c + 30);
        e_0n = FN(zl_tf)(c, e, im, ex, iex, we, c[54], c + 30);
        e_pp = FN(zl_tf)(c, ep, ip, ex, iex, we, c[54], c + 30);
        e_pn = FN(zl_tf)(c, ep, im, ex, iex, we, c[54], c + 30);
        e_np = FN(zl_tf)(c, em, ip, ex, iex, we, c[54], c + 30);
        e_nn = FN(zl_tf)(c, em, im, ex, iex, we, c[54], c + 30);
        i_p0 = FN(zl_tf)(c, ep, i, ix, iix, wi, c[55], c + 40);
        i_n0 = FN(zl_tf)(c, em, i, ix, iix, wi, c[55], c + 40);
        i_0p = FN(zl_tf)(c, e, ip, ix, iix, wi, c[55], c + 40);
        i_0n = FN(zl_tf)(c, e, im, ix, iix, wi, c[55], c + 40);
        i_pp = FN(zl_tf)(c, ep, ip, ix, iix, wi, c[55], c + 40);
        i_pn = FN(zl_tf)(c, ep, im, ix, iix, wi, c[55], c + 40);
        i_np = FN(zl_tf)(c, em, ip, ix, iix, wi, c[55], c + 40);
        i_nn = FN(zl_tf)(c, em, im, ix, iix, wi, c[55], c + 40);
        dfe_e = (e_p0 - e_n0) * a2; dfi_e = (e_0p - e_0n) * a2;
        dfe_i = (i_p0 - i_n0) * a2; dfi_i = (i_0p - i_0n) * a2;
        fefe_e = ((e_p0 - tfe * (REAL)2) + e_n0) * q2;
        fifi_e = ((e_0p - tfe * (REAL)2) + e_0n) * q2;
        fefi_e = ((e_pp - e_np) * a2 - (e_pn - e_nn) * a2) * a2;
        fife_e = ((e_pp - e_pn) * a2 - (e_np - e_nn) * a2) * a2;
        fefe_i = ((i_p0 - tfi * (REAL)2) + i_n0) * q2;
        fifi_i = ((i_0p - tfi * (REAL)2) + i_0n) * q2;
        fefi_i = ((i_pp - i_np) * a2 - (i_pn - i_nn) * a2) * a2;
        fife_i = ((i_pp - i_pn) * a2 - (i_np - i_nn) * a2) * a2;
        dx[0] = (((((tfe - e) + cee * (fefe_e * (REAL)0.5f)) + cei * (fefi_e * (REAL)0.5f)) + cei * (fife_e * (REAL)0.5f)) +
                 cii * (fifi_e * (REAL)0.5f)) / c[22];
        dx[1] = (((((tfi - i) + cee * (fefe_i * (REAL)0.5f)) + cei * (fefi_i * (REAL)0.5f)) + cei * (fife_i * (REAL)0.5f)) +
                 cii * (fifi_i * (REAL)0.5f)) / c[22];
        dx[2] = ((((tfe * ((c[57] - tfe) / c[58]) + (tfe - e) * (tfe - e)) + cee * (dfe_e * (REAL)2)) +
                  cei * (dfi_e * (REAL)2)) - cee * (REAL)2) / c[22];
        dx[3] = (((((((tfe - e) * (tfi - i) + cee * dfe_e) + cei * dfe_i) + cei * dfi_e) + cii * dfi_i)) -
                 cei * (REAL)2) / c[22];
        dx[4] = ((((tfi * ((c[57] - tfi) / c[59]) + (tfi - i) * (tfi - i)) + cii * (dfi_i * (REAL)2)) +
                  cei * (dfe_i * (REAL)2)) - cii * (REAL)2) / c[22];
        FN(zl_fluct)(c, e, i, ex, iex, we, c[54], &mve, &s_, &t_);
        dx[5] = ((-we) / c[8] + e * c[4]) + ((mve + (-c[1])) * c[5]) / c[8];
        FN(zl_fluct)(c, e, i, ix, iix, wi, c[55], &mvi, &s_, &t_);
        dx[6] = ((-wi) / c[9] + i * c[6]) + ((mvi + (-c[2])) * c[7]) / c[9];
        dx[7] = (-ou) / c[27];
        break; }
    case ORC_KIX: { /* dfun.rs:1782-1869 */
        REAL xx = x[0], v = x[1], n = x[2], dki = x[3], kg = x[4];
        REAL m_inf = FN(kx_sig)((v + (REAL)24) * ((REAL)1 / (REAL)12));
        REAL n_inf = FN(kx_sig)((v + (REAL)19) * ((REAL)1 / (REAL)18));
        REAL h_na = -((REAL)1 / (EXPR((n + (REAL)-0.4f) * (REAL)-8) + (REAL)1)) + (REAL)1.1f;
        REAL bv = (REAL)2160 / (REAL)720;
        REAL dna_i = -dki, dna_o = dna_i * (-bv), dk_o = dki * (-bv);
        REAL ki = (-dki) + (REAL)130, nai = dna_i + (REAL)16, nao = dna_o + (REAL)138, ko = (dk_o + (REAL)4.8f) + kg;
        REAL i_k = (n * (REAL)22 + (REAL)0.12f) * (v - (FN(kx_ln)(ko) - FN(kx_ln)(ki)) * (REAL)26.64f);
        REAL i_na = ((m_inf * h_na) * (REAL)40 + (REAL)0.02f) * (v - (FN(kx_ln)(nao) - FN(kx_ln)(nai)) * (REAL)26.64f);
        REAL i_cl = (v + c[19]) * (REAL)7.5f;
        REAL i_pump = (FN(kx_sig)(nai + (REAL)-21) * FN(kx_sig)(ko + (REAL)-5.5f)) * (REAL)250;
        REAL v_dot = (((i_na + i_k) + i_cl) + i_pump) * c[14];
        REAL r_x = (xx * c[6]) / PIR;
        REAL syn = (cpl[0] * c[15]) * ((-v) + c[0]);
        if (v < c[9]) {
            dx[0] = ((((v + (-c[5])) * c[16]) * xx) + c[4]) - r_x * (xx * c[2]);
            dx[1] = v_dot - xx * (((xx * c[6]) + c[3]) + syn);
        } else {
            dx[0] = ((((v + (-c[7])) * c[17]) * xx) + c[4]) - r_x * (xx * c[2]);
            dx[1] = v_dot - xx * (((xx * c[8]) + c[3]) + syn);
        }
        dx[2] = (n_inf - n) / c[11];
        dx[3] = (i_k - i_pump * (REAL)2) * c[18];
        dx[4] = ((-ko) + c[1]) * c[13];
        break; }
    case ORC_CB: { /* dfun.rs:609-637 */
        REAL r = x[0], v = x[1], g = x[2], q = x[3];
        dx[0] = ((((v * r) * (REAL)2) + c[5]) - g * r) + cpl[0];
        dx[1] = (((v * v) + c[4]) - r * (r * c[6])) + ((-v) + c[2]) * g;
        dx[2] = q * c[1];
        dx[3] = (((r * c[7]) - g) - q * (REAL)2) * c[1];
        break; }
    case ORC_GSD: case ORC_GSF: { /* dfun.rs:639-716 (v times the bracket, as the reference writes it) */
        REAL r = x[0], v = x[1], a = x[2], b = x[3], br;
        dx[0] = (((r * v) * (REAL)2) + c[12]) * c[11];
        br = (v - r * (r * c[13])) + c[14];
        if (k->model == ORC_GSD) br = br + (r * c[15]) * ((-a) + (REAL)1);
        else br = (br + r * c[15]) - a;
        dx[1] = (v * ((br + cpl[0] * c[7]) + cpl[1] * c[8])) * c[11];
        dx[2] = b * c[16];
        dx[3] = (((b * (REAL)-2) - a) + r * c[2]) * c[16];
        break; }
    case ORC_LB: { /* dfun.rs:718-766 */
        REAL V = x[0], W = x[1], Z = x[2];
        REAL m_ca = (REAL)0.5f * (TANHR((V + (-c[9])) * c[32]) + (REAL)1);
        REAL m_na = (REAL)0.5f * (TANHR((V + (-c[10])) * c[33]) + (REAL)1);
        REAL m_k = (REAL)0.5f * (TANHR((V + (-c[11])) * c[34]) + (REAL)1);
        REAL qv = c[35] * (TANHR((V + (-c[26])) * c[36]) + (REAL)1);
        REAL qz = c[37] * (TANHR((Z + (-c[27])) * c[38]) + (REAL)1);
        REAL tca = (c[0] + qv * c[40]) + cpl[0] * c[39];
        REAL tna = ((m_na * c[3]) + qv * c[41]) + cpl[0] * c[42];
        dx[0] = ((((((-tca) * m_ca) * (V + (-c[5])) - (W * c[1]) * (V + (-c[6]))) - (V + (-c[7])) * c[2]) -
                  tna * (V + (-c[8]))) - (Z * c[18]) * qz) * c[30] + c[43];
        dx[1] = (m_k - W) * c[44];
        dx[2] = ((V * c[17]) * qv + c[45]) * c[46];
        break; }
    case ORC_E2D: { /* dfun.rs:776-818 */
        REAL x1 = x[0], z = x[1], f1, zm, yz, h, z4;
        if (x1 < 0) f1 = (x1 * x1) * c[2] + x1 * c[12];
        else { z4 = z + (REAL)-4; f1 = ((z4 * z4) * (REAL)-0.6f + (-c[4])) + x1 * c[6]; }
        dx[0] = ((((-z) + c[13]) + cpl[0] * c[8]) - f1 * x1) * c[10];
        zm = z * (REAL)-0.1f;
        yz = z < 0 ? -((((((zm * zm) * zm) * zm) * zm) * zm) * zm) : (REAL)0;
        if (c[11] > (REAL)0.5f) h = (x1 + (REAL)3) / (EXPR(-((x1 + (REAL)0.5f) * (REAL)10)) + (REAL)1) + yz;
        else h = (x1 * (REAL)4 + c[14]) + yz;
        dx[1] = ((h - z) + cpl[0] * c[9]) * c[15];
        break; }
    case ORC_EPI: { /* dfun.rs:820-874 */
        REAL x1 = x[0], y1 = x[1], z = x[2], x2 = x[3], y2 = x[4], g = x[5], f1, z4, zm, yz, h, xp, f2;
        if (x1 < 0) f1 = (-(x1 * x1)) * c[3] + x1 * c[4];
        else { z4 = z + (REAL)-4; f1 = ((-x2) + c[8]) + (z4 * z4) * (REAL)0.6f; }
        dx[0] = ((((y1 - z) + cpl[0] * c[12]) + f1 * x1) + c[0]) * c[15];
        dx[1] = (((-((x1 * x1) * c[6])) + c[5]) - y1) * c[15];
        zm = z * (REAL)-0.1f;
        yz = z < 0 ? -((((((zm * zm) * zm) * zm) * zm) * zm) * zm) : (REAL)0;
        if (c[16] > (REAL)0.5f) h = (x1 + (REAL)3) / (EXPR(-((x1 + (REAL)0.5f) * (REAL)10)) + (REAL)1) + yz;
        else h = (x1 * (REAL)4 + c[17]) + yz;
        dx[2] = ((h - z) + cpl[0] * c[14]) * c[18];
        dx[3] = ((((((x2 - (x2 * x2) * x2) - y2) + c[1]) + g * c[11]) - (z + (REAL)-3.5f) * (REAL)0.3f) +
                 cpl[1] * c[13]) * c[15];
        xp = x2 + (REAL)0.25f;
        f2 = xp < 0 ? (REAL)0 : xp * c[10];
        dx[4] = ((f2 - y2) * c[15]) / c[9];
        dx[5] = ((-x1) * (REAL)0.1f + g) * c[19];
        break; }
    case ORC_RWEI: case ORC_DECO: { /* dfun.rs:876-950 */
        REAL se = x[0], si = x[1], cc = cpl[0] * c[19], jse = se * c[6], xe, xi, ae, ai, he, hi;
        xe = (((jse * c[5]) - si * c[13]) + c[20]) + cc;
        xi = ((jse - si) + c[22]) + cc * c[18];
        if (k->model == ORC_DECO) { ae = xe * c[24] + c[25]; ai = xi * c[26] + c[27]; }
        else { ae = xe * c[0] + (-c[1]); ai = xi * c[8] + (-c[9]); }
        he = ae / ((-EXPR(ae * (-c[2]))) + (REAL)1);
        hi = ai / ((-EXPR(ai * (-c[10]))) + (REAL)1);
        dx[0] = (-se) * c[21] + (((-se) + (REAL)1) * he) * c[3];
        dx[1] = (-si) * c[23] + hi * c[11];
        break; }
    case ORC_ZJ: { /* dfun.rs:1212-1268; sigmoid 2 e0 / (exp((rho_2 - v) rho_1) + 1) */
        REAL s0 = c[18] / (EXPR(((-x[0]) + c[5]) * c[6]) + (REAL)1);
        REAL s23 = c[18] / (EXPR(((-(x[2] - x[4])) + c[5]) * c[6]) + (REAL)1);
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
        break; }
    case ORC_RFHN: { /* dfun.rs:1270-1300 */
        REAL xi = x[0], eta = x[1], al = x[2], be = x[3];
        REAL xi3 = (xi * xi) * xi, al3 = (al * al) * al;
        dx[0] = ((((xi - xi3 / (REAL)3) - eta) * c[0] + (-xi) * c[3]) - (al - xi) * c[4]) + cpl[0] * c[0];
        dx[1] = ((xi - eta * c[2]) + c[1]) / c[0];
        dx[2] = (((al - al3 / (REAL)3) - be) * c[0] + (-xi) * c[5]) + cpl[1] * c[0];
        dx[3] = ((al - be * c[2]) + c[1]) / c[0];
        break; }
    case ORC_RHR: { /* dfun.rs:1302-1348 */
        REAL xi = x[0], eta = x[1], ts = x[2], al = x[3], be = x[4], ga = x[5];
        REAL xi2 = xi * xi, xi3 = (xi * xi) * xi, al2 = al * al, al3 = (al * al) * al;
        dx[0] = (((((eta - xi3 * c[1]) + xi2 * c[2]) - ts) + (-xi) * c[7]) - (al - xi) * c[8]) + cpl[0];
        dx[1] = ((-xi2) * c[4] - eta) + c[3];
        dx[2] = (xi * c[12] - ts * c[0]) + c[13];
        dx[3] = ((((be - al3 * c[1]) + al2 * c[2]) - ga) + (-xi) * c[9]) + cpl[1];
        dx[4] = ((-al2) * c[4] - be) + c[3];
        dx[5] = (al * c[12] - ga * c[0]) + c[13];
        break; }
    case ORC_DG: { /* dfun.rs:1350-1400 (v times the bracket, as the reference writes it) */
        REAL re = x[0], ve = x[1], see = x[2], sei = x[3], ri = x[4], vi = x[5], sie = x[6], sii = x[7];
        dx[0] = (((ve * (REAL)2) * re) + c[20]) * c[17];
        dx[1] = (ve * (((((((ve - (re * re) * c[21]) + c[2]) + see * c[14]) - sei * c[14]) + c[0]) + cpl[0]) +
                       cpl[1] * (REAL)0)) * c[17];
        dx[2] = (((-see) + re * c[9]) + cpl[2]) * c[19];
        dx[3] = ((-sei) + ri * c[10]) * c[19];
        dx[4] = (((vi * (REAL)2) * ri) + c[22]) * c[18];
        dx[5] = (vi * ((((((vi - (ri * ri) * c[23]) + c[6]) + sie * c[15]) - sii * c[15]) + c[4]) + cpl[2] * c[13])) *
                c[18];
        dx[6] = (((-sie) + re * c[11]) + cpl[2] * c[13]) * c[19];
        dx[7] = ((-sii) + ri * c[12]) * c[19];
        break; }
    }
}

/* Derivative of one node: x[NV], c[NC] -> dx[NV].
 * Op order follows the tensor expressions in dfun.rs exactly. */
static void FN(dfun)(const FN(mconst)* k, const REAL* x, const REAL* c, REAL* dx)
{
    switch (k->model) {
    case ORC_G2DO: { /* dfun.rs:295-318 */
        REAL v = x[0], w = x[1];
        REAL iext = k->sweep_form ? (c[0] + k->sweep_val) * k->g_gamma
                                  : c[0] * k->g_gamma + k->g_gamma * k->g_i;
        REAL v2 = v * v;
        REAL v3 = v * v2;
        dx[0] = ((((w * k->g_alpha + iext) - v3 * k->g_f) + v2 * k->g_e) + v * k->g_g) * k->g_dtau;
        dx[1] = (((v * k->g_b + v2 * k->g_c) - w * k->g_beta) + k->g_a) * k->g_dovt;
        break; }
    case ORC_MPR: { /* dfun.rs:455-477 */
        REAL r = x[0], v = x[1];
        REAL dv, v2, r2;
        dx[0] = ((r * v) * (REAL)2 + k->m_dr_c) * k->m_inv_tau;
        v2 = v * v;
        r2 = r * r;
        if (k->sweep_form)
            dv = (((((v2 - r2 * k->m_pp) + k->m_eta) + k->sweep_val) + r * k->m_jtau) + c[0] * k->m_cr) + c[1] * k->m_cv;
        else
            dv = ((((v2 - r2 * k->m_pp) + k->m_eta_i) + r * k->m_jtau) + c[0] * k->m_cr) + c[1] * k->m_cv;
        dx[1] = dv * k->m_inv_tau;
        break; }
    case ORC_JR: { /* dfun.rs:356-383 */
        REAL s12 = FN(jr_sigm)(k, x[1] - x[2]);
        REAL s01 = FN(jr_sigm)(k, x[0] * k->j_a1J);
        REAL s03 = FN(jr_sigm)(k, x[0] * k->j_a3J);
        REAL mu = k->sweep_form ? k->sweep_val : k->j_mu;
        REAL mut = (s01 * k->j_a2J + mu) + c[0];
        dx[0] = x[3]; dx[1] = x[4]; dx[2] = x[5];
        dx[3] = (s12 * k->j_Aa - x[3] * k->j_2a) - x[0] * k->j_aa;
        dx[4] = (mut * k->j_Aa - x[4] * k->j_2a) - x[1] * k->j_aa;
        dx[5] = ((s03 * k->j_a4J) * k->j_Bb - x[5] * k->j_2b) - x[2] * k->j_bb;
        break; }
    case ORC_WC: { /* dfun.rs:405-436 */
        REAL e = x[0], i = x[1];
        REAL xe = ((e * k->w_cee - i * k->w_cei) - k->w_the) + c[0];
        REAL xi, ee, ei, se, si;
        xe = (xe + (k->sweep_form ? k->sweep_val : k->w_p)) * k->w_ale;
        xi = (((e * k->w_cie - i * k->w_cii) + k->w_q) - k->w_thi) * k->w_ali;
        ee = EXPR((xe + (-k->w_be)) * (-k->w_ae));
        se = ((REAL)1 / (ee + (REAL)1) - k->w_off_e) * k->w_ce;
        ei = EXPR((xi + (-k->w_bi)) * (-k->w_ai));
        si = ((REAL)1 / (ei + (REAL)1) - k->w_off_i) * k->w_ci;
        dx[0] = (-e + (e * (-k->w_re) + k->w_ke) * se) * k->w_inv_te;
        dx[1] = (-i + (i * (-k->w_ri) + k->w_ki) * si) * k->w_inv_ti;
        break; }
    case ORC_KUR: /* dfun.rs:486-491 */
        dx[0] = c[0] + (k->sweep_form ? k->sweep_val : k->k_omega);
        break;
    case ORC_LIN: /* dfun.rs:522-532 */
        dx[0] = x[0] * k->l_gamma + c[0];
        break;
    case ORC_SH: { /* dfun.rs:534-557 */
        REAL amp = x[0] * x[0] + x[1] * x[1];
        REAL g = -amp + k->h_a;
        dx[0] = (g * x[0] - x[1] * k->h_omega) + c[0];
        dx[1] = (g * x[1] + x[0] * k->h_omega) + c[1];
        break; }
    case ORC_HOP: /* hopfield_dfun_batch */
        dx[0] = (-x[0] + c[0]) * k->p_itx;
        dx[1] = k->p_dyn ? (-x[1] + c[1]) * k->p_itt : (REAL)0;
        break;
    case ORC_CB2: { /* coombes_byrne2d_dfun_batch */
        REAL r = x[0], v = x[1], r2 = r * r;
        dx[0] = ((v * (r * (REAL)2) + k->cb_dpi) - r2 * k->cb_kpi) + c[0];
        dx[1] = ((-(v * v) + k->cb_eta) - r2 * k->cb_pi2) + ((-v + k->cb_vsyn) * k->cb_kpi) * r;
        break; }
    case ORC_EC3: case ORC_EC3S: case ORC_ERS: case ORC_ZL1: case ORC_ZL2: case ORC_KIX:
    case ORC_CB: case ORC_GSD: case ORC_GSF: case ORC_LB: case ORC_E2D: case ORC_EPI: case ORC_RWEI:
    case ORC_DECO: case ORC_ZJ: case ORC_RFHN: case ORC_RHR: case ORC_DG:
        FN(dfun_ext)(k, x, c, dx);
        break;
    case ORC_RWW: { /* dfun.rs:503-515 */
        REAL s = x[0];
        REAL xx = k->sweep_form ? (s * k->r_wjn + k->sweep_val) + c[0] * k->r_jn
                                : (s * k->r_wjn + k->r_io) + c[0] * k->r_jn;
        REAL axb = xx * k->r_a + (-k->r_b);
        REAL et = EXPR(axb * (-k->r_d));
        REAL den = -et + (REAL)1;
        REAL h = axb / den;
        dx[0] = (-s) * k->r_inv_taus + ((-s + (REAL)1) * h) * k->r_gamma;
        break; }
    }
}

/* clamp_batch, dfun.rs:170-239 (six hot models). */
static void FN(clamp)(int model, REAL* x)
{
    switch (model) {
    case ORC_WC:
        x[0] = x[0] < 0 ? 0 : (x[0] > 1 ? 1 : x[0]);
        x[1] = x[1] < 0 ? 0 : (x[1] > 1 ? 1 : x[1]);
        break;
    case ORC_MPR:
        x[0] = x[0] < 0 ? 0 : x[0];
        break;
    case ORC_RWW:
        x[0] = x[0] < 0 ? 0 : (x[0] > 1 ? 1 : x[0]);
        break;
    case ORC_CB2: /* dfun.rs:190-194 */
        x[0] = x[0] < 0 ? 0 : x[0];
        break;
    case ORC_ZL1: case ORC_ZL2: /* dfun.rs:219-230: E, I in [0, 1] */
        x[0] = x[0] < 0 ? 0 : (x[0] > 1 ? 1 : x[0]);
        x[1] = x[1] < 0 ? 0 : (x[1] > 1 ? 1 : x[1]);
        break;
    case ORC_KIX: /* dfun.rs:232-236: x >= 0 */
    case ORC_CB: case ORC_GSD: case ORC_GSF: /* dfun.rs:195-209: r >= 0 */
        x[0] = x[0] < 0 ? 0 : x[0];
        break;
    case ORC_RWEI: case ORC_DECO: /* dfun.rs:172-176: both in [0, 1] */
        x[0] = x[0] < 0 ? 0 : (x[0] > 1 ? 1 : x[0]);
        x[1] = x[1] < 0 ? 0 : (x[1] > 1 ? 1 : x[1]);
        break;
    case ORC_DG: /* dfun.rs:210-218: r_e, r_i >= 0 */
        x[0] = x[0] < 0 ? 0 : x[0];
        x[4] = x[4] < 0 ? 0 : x[4];
        break;
    case ORC_KUR: { /* dfun.rs:188-191 */
        REAL tp = (REAL)2 * PIR;
        x[0] = x[0] - FLOORR(x[0] / tp) * tp;
        break; }
    default:
        break;
    }
}

/* ---- coupling pipeline: coupling.rs:204-345, 360-399 ---- */

typedef struct {
    int cfn;
    REAL a, b, cmin, cmax, mid, sigma, r, e0, v0, h, q, g, p, theta;
    REAL kur_scale;   /* a / n_src, coupling.rs:370 */
    int classic;      /* SigmoidalJansenRit use_classic */
    int cpu_form;     /* pre_cpu op order (runtime.rs per-edge path) */
} FN(cconst);

static void FN(cconst_init)(FN(cconst)* k, const hb_orc_coupling* c)
{
    const float* P = c->cparams;
    int n = c->n_cparams;
    memset(k, 0, sizeof(*k));
    k->cfn = c->cfn;
    k->cpu_form = c->cpu_form;
    /* parameter parsing: coupling.rs:114-202 */
    switch (c->cfn) {
    case ORC_CF_LINEAR: case ORC_CF_SCALED_LINEAR:
        k->a = n > 0 ? P[0] : 1; k->b = n > 1 ? P[1] : 0; break;
    case ORC_CF_DIFFERENCE:
        k->a = n > 0 ? P[0] : 1; break;
    case ORC_CF_KURAMOTO:
        k->a = n > 0 ? P[0] : 1;
        k->kur_scale = (REAL)((float)k->a / (float)c->n_src);
        if (sizeof(REAL) == 8) k->kur_scale = k->a / (REAL)c->n_src;
        break;
    case ORC_CF_SIGMOIDAL:
        if (n >= 5) { k->cmin = P[0]; k->cmax = P[1]; k->mid = P[2]; k->a = P[3]; k->sigma = P[4]; }
        else { k->cmin = -1; k->cmax = n > 0 ? P[0] : 1; k->mid = n > 1 ? P[1] : 0; k->sigma = n > 2 ? P[2] : 1; k->a = 1; }
        break;
    case ORC_CF_TANH:
        k->a = n > 0 ? P[0] : 1; k->b = n > 1 ? P[1] : 1; k->mid = n > 2 ? P[2] : 0; k->sigma = n > 3 ? P[3] : 1;
        break;
    case ORC_CF_SJR:
        if (n >= 5) { k->classic = 1; k->a = P[0]; k->cmin = P[1]; k->cmax = P[2]; k->r = P[3]; k->mid = P[4]; k->e0 = (REAL)0.005f; k->v0 = 6; }
        else { k->classic = 0; k->a = n > 0 ? P[0] : 1; k->e0 = n > 1 ? P[1] : (REAL)0.005f; k->r = n > 2 ? P[2] : (REAL)0.56f; k->v0 = n > 3 ? P[3] : 6; k->cmin = 0; k->cmax = (REAL)0.005f; k->mid = 6; }
        break;
    case ORC_CF_PRESIGMOIDAL:
        k->h = n > 0 ? P[0] : 1; k->q = n > 1 ? P[1] : 1; k->g = n > 2 ? P[2] : 1; k->p = n > 3 ? P[3] : 1; k->theta = n > 4 ? P[4] : (REAL)0.5f;
        break;
    }
}

/* pre(): one source node, nc cvars in -> npre channels out.
 * Tensor form coupling.rs:204-258; CPU form (per-edge path) 268-321. */
static void FN(pre)(const FN(cconst)* k, const REAL* x, int nc, REAL* out)
{
    int c;
    switch (k->cfn) {
    case ORC_CF_KURAMOTO:
        /* tensor form: cat([sin x, cos x], 1) (channel-major);
         * per-edge CPU form interleaves (sin,cos) per cvar. Identical for nc=1. */
        for (c = 0; c < nc; ++c) { out[c] = SINR(x[c]); out[nc + c] = COSR(x[c]); }
        break;
    case ORC_CF_TANH:
        for (c = 0; c < nc; ++c) {
            REAL inner = (x[c] * k->b + (-k->mid)) / k->sigma;
            out[c] = (TANHR(inner) + (REAL)1) * k->a;
        }
        break;
    case ORC_CF_SJR: {
        REAL diff = x[0] - (nc >= 2 ? x[1] : (REAL)0);
        if (k->classic) {
            REAL den = EXPR((diff + (-k->mid)) * (-k->r)) + (REAL)1;
            out[0] = k->cpu_form ? k->cmin + (k->cmax - k->cmin) / den
                                 : ((REAL)1 / den) * (k->cmax - k->cmin) + k->cmin;
        } else {
            REAL den = EXPR((diff + (-k->v0)) * (-k->r)) + (REAL)1;
            out[0] = k->cpu_form ? ((k->a * (REAL)2) * k->e0) / den
                                 : ((REAL)1 / den) * ((k->a * (REAL)2) * k->e0);
        }
        break; }
    case ORC_CF_PRESIGMOIDAL:
        for (c = 0; c < nc; ++c) {
            REAL inner = x[c] * k->p + (-k->theta);
            out[c] = (TANHR(inner * k->g) + k->q) * k->h;
        }
        break;
    default: /* Linear, Sigmoidal, ScaledLinear, Difference: identity */
        for (c = 0; c < nc; ++c) out[c] = x[c];
        break;
    }
}

static int FN(npre)(const FN(cconst)* k, int nc)
{
    if (k->cfn == ORC_CF_KURAMOTO) return 2 * nc;
    if (k->cfn == ORC_CF_SJR) return 1;
    return nc;
}

/* post_with_target: coupling.rs:323-399. g[npre] -> out[nc]; xi = target cvars. */
static void FN(post)(const FN(cconst)* k, const REAL* g, const REAL* xi, REAL rowsum, int nc, REAL* out)
{
    int c;
    switch (k->cfn) {
    case ORC_CF_LINEAR:
        for (c = 0; c < nc; ++c) out[c] = g[c] * k->a + k->b;
        break;
    case ORC_CF_SCALED_LINEAR:
        for (c = 0; c < nc; ++c) out[c] = (g[c] + (-k->b)) * k->a;
        break;
    case ORC_CF_DIFFERENCE: /* non-square rowsum form, coupling.rs:376-380 */
        for (c = 0; c < nc; ++c) out[c] = (g[c] - xi[c] * rowsum) * k->a;
        break;
    case ORC_CF_KURAMOTO: /* coupling.rs:367-374 */
        for (c = 0; c < nc; ++c)
            out[c] = (COSR(xi[c]) * g[c] - SINR(xi[c]) * g[nc + c]) * k->kur_scale;
        break;
    case ORC_CF_SIGMOIDAL: /* coupling.rs:326-331 */
        for (c = 0; c < nc; ++c) {
            REAL sh = (g[c] + (-k->mid)) / k->sigma;
            REAL den = EXPR(-(sh * k->a)) + (REAL)1;
            out[c] = ((REAL)1 / den) * (k->cmax - k->cmin) + k->cmin;
        }
        break;
    case ORC_CF_SJR:
        if (k->classic) { for (c = 0; c < nc; ++c) out[c] = g[c < 1 ? c : 0] * k->a; }
        else { for (c = 0; c < nc; ++c) out[c] = g[c < 1 ? c : 0]; }
        break;
    default: /* Tanh, PreSigmoidal: identity */
        for (c = 0; c < nc; ++c) out[c] = g[c];
        break;
    }
}


/* ---- Balloon-Windkessel: bold.rs:106-166, bold_monitor.rs:72-107 ---- */

/* One forward-Euler BW step of size dt_s seconds for nodes [0,N). bold.rs:106-150 */
static void FN(bold_step)(int N, const REAL* z_in, REAL dt_s, REAL* S, REAL* F, REAL* V, REAL* Q)
{
    const REAL tau_s = (REAL)1.54f, tau_f = (REAL)0.50f, tau_0 = (REAL)1.02f;
    const REAL alpha = (REAL)0.32f, e0 = (REAL)0.34f;
    const REAL kappa = (REAL)1 / tau_s, gamma = (REAL)1 / tau_f, tau_inv = (REAL)1 / tau_0;
    const REAL ialpha = (REAL)1 / alpha;
    int i;
    for (i = 0; i < N; ++i) {
        REAL z = z_in[i] > 0 ? z_in[i] : 0;
        REAL s = S[i], f = F[i], v = V[i], q = Q[i];
        REAL fs = f > (REAL)1e-6f ? f : (REAL)1e-6f;
        REAL vs = v > (REAL)1e-6f ? v : (REAL)1e-6f;
        REAL ds = (z - kappa * s) - gamma * (fs - (REAL)1);
        REAL df = s;
        REAL dv = tau_inv * (fs - POWR(vs, ialpha));
        REAL dq = tau_inv * ((fs * ((REAL)1 - POWR((REAL)1 - e0, (REAL)1 / fs))) / e0
                             - POWR(vs, ialpha) * (q / vs));
        REAL sn = s + dt_s * ds;
        REAL fn = fs + dt_s * df;
        REAL vn = vs + dt_s * dv;
        REAL qn = q + dt_s * dq;
        S[i] = sn > 0 ? sn : 0;
        F[i] = fn > (REAL)1e-6f ? fn : (REAL)1e-6f;
        V[i] = vn > (REAL)1e-6f ? vn : (REAL)1e-6f;
        Q[i] = qn > 0 ? qn : 0;
    }
}

/* BOLD signal, bold.rs:153-166 */
static REAL FN(bold_signal1)(REAL v, REAL q)
{
    const REAL V0 = 4, k1 = 7, k2 = 2, k3 = 2;
    REAL vv = v > (REAL)1e-6f ? v : (REAL)1e-6f;
    return V0 * ((k1 * ((REAL)1 - q) + k2 * ((REAL)1 - q / vv)) + k3 * ((REAL)1 - vv));
}

/* Close one averaging window of `count` neural steps: average, one BW step of
 * dt*count ms, then the signal (bold_monitor.rs:86-106). */
static void FN(bold_window)(const hb_orc_problem* P, int N, const REAL* acc, int count,
                            REAL* S, REAL* F, REAL* V, REAL* Q, REAL* out)
{
    REAL* avg = (REAL*)malloc(sizeof(REAL) * (size_t)N);
    int i;
    double dt_bold = P->dt_ms * (double)count;
    REAL dt_s = (REAL)(dt_bold / 1000.0);
    if (!avg) return;
    for (i = 0; i < N; ++i) avg[i] = acc[i] / (REAL)count;
    FN(bold_step)(N, avg, dt_s, S, F, V, Q);
    for (i = 0; i < N; ++i) out[i] = FN(bold_signal1)(V[i], Q[i]);
    free(avg);
}

/* ---- the time loop: engine.rs:540-804 (batch) / runtime.rs:46-501 (plain) ---- */

/* Run one instance. Outputs are REAL arrays (float for the FP32 twin). */
static int FN(run)(const hb_orc_problem* P, REAL* tavg, REAL* fin, REAL* traj,
                   REAL* bold, int32_t* bold_n)
{
    const int model = P->model, N = P->n_nodes, T = P->n_steps;
    const int NV = hb_orc_nvar(model), NC = hb_orc_ncvar(model);
    const REAL dt = (REAL)P->dt;
    FN(mconst) kbase;
    FN(mconst)* kn = NULL;   /* per-node constants when a node parameter is given */
    FN(cconst) cc;
    int npre, H, t, n, v, e;
    REAL *x = NULL, *hist = NULL, *pre0 = NULL, *cpl = NULL, *tmp = NULL;
    REAL nscale[8];
    REAL *bw_s = NULL, *bw_f = NULL, *bw_v = NULL, *bw_q = NULL, *bw_acc = NULL;
    int bw_count = 0, nb = 0;
    const int stoch = P->integrator >= ORC_EULER_STOCH;
    int rc = 0;

    FN(mconst_init)(&kbase, model, P->params, P->sweep_form, P->sweep_val);
    if (P->node_param_idx >= 0 && P->node_param) {
        float pp[32];
        kn = (FN(mconst)*)malloc(sizeof(FN(mconst)) * (size_t)N);
        for (n = 0; n < N; ++n) {
            memcpy(pp, P->params, sizeof(float) * (size_t)hb_orc_nparams(model));
            pp[P->node_param_idx] = P->node_param[n];
            FN(mconst_init)(&kn[n], model, pp, P->sweep_form, P->sweep_val);
        }
    }
    FN(cconst_init)(&cc, &P->cpl);
    npre = FN(npre)(&cc, NC);

    /* history of pre(x_t) for every t, [T+1][N][npre]: the delayed read for
     * delay d at step s is pre(x_{s-d}), runtime.rs:55-85,176 */
    H = P->cpl.max_delay;
    x = (REAL*)calloc((size_t)N * NV, sizeof(REAL));
    hist = (REAL*)calloc((size_t)(H + 1) * N * npre, sizeof(REAL));
    pre0 = (REAL*)calloc((size_t)npre, sizeof(REAL));
    cpl = (REAL*)calloc((size_t)N * NC, sizeof(REAL));
    tmp = (REAL*)calloc((size_t)npre * 4 + 64, sizeof(REAL));
    if (!x || !hist || !pre0 || !cpl || !tmp) { rc = -1; goto done; }
    {   /* pre(0): the value a single-delay read of "not yet simulated" time yields */
        REAL z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        FN(pre)(&cc, z, NC, pre0);
    }
    for (n = 0; n < N * NV; ++n) x[n] = (REAL)P->x0[n];
    for (n = 0; n < N * NV; ++n) tavg[n] = 0;
    for (v = 0; v < NV; ++v) {
        float nsf = P->nsig ? P->nsig[v] : 0.0f;
        /* integrator.rs:179-191: sqrt(2*nsig) * sqrt(dt) */
        nscale[v] = SQRTR((REAL)2 * (REAL)nsf) * SQRTR(dt);
    }
    if (P->bold_period > 0) {
        bw_s = (REAL*)calloc((size_t)N, sizeof(REAL));
        bw_f = (REAL*)calloc((size_t)N, sizeof(REAL));
        bw_v = (REAL*)calloc((size_t)N, sizeof(REAL));
        bw_q = (REAL*)calloc((size_t)N, sizeof(REAL));
        bw_acc = (REAL*)calloc((size_t)N, sizeof(REAL));
        for (n = 0; n < N; ++n) { bw_f[n] = 1; bw_v[n] = 1; bw_q[n] = 1; }
    }

    {
    const int Hd = H + 1;   /* reference horizon: max_delay + 1 */
    for (t = 0; t < T; ++t) {
        const int ws = t % Hd;
        REAL* cur = hist + (size_t)ws * N * npre;
        /* 0. history write of the current state (pre applied) */
        for (n = 0; n < N; ++n) {
            REAL xc[8];
            for (v = 0; v < NC; ++v) xc[v] = x[n * NV + v];
            FN(pre)(&cc, xc, NC, cur + (size_t)n * npre);
        }
        /* 1. coupling */
        for (n = 0; n < N * NC; ++n) cpl[n] = 0;
        if (P->cpl.wkind != ORC_W_NONE) {
            REAL* g = tmp;           /* [npre] */
            REAL mf[8];
            const REAL* src_single = cur;
            int single_zero = 0;
            if (P->cpl.delay_mode != ORC_DELAY_PER_EDGE) {
                int d = P->cpl.single_delay;
                int teff;
                if (d == 0 || t == 0) teff = t;
                else if (P->cpl.delay_mode == ORC_DELAY_BATCH_COMPAT) {
                    /* engine.rs:541-589: sliding window indexed as a ring */
                    if (d > t) teff = -1;
                    else teff = t - Hd + 1 + ((t - d + Hd) % Hd);
                    (void)0;
                } else teff = d > t ? -1 : t - d;
                if (teff < 0) single_zero = 1;
                else src_single = hist + (size_t)(teff % Hd) * N * npre;
            }
            if (P->cpl.wkind == ORC_W_SCALAR_MEAN) {
                /* engine.rs:602-608: w * mean_j pre(x_j), broadcast */
                int c;
                for (c = 0; c < npre; ++c) {
                    REAL s = 0;
                    for (n = 0; n < N; ++n) s += single_zero ? pre0[c] : src_single[(size_t)n * npre + c];
                    mf[c] = (s / (REAL)N) * (REAL)P->cpl.scalar_w;
                }
            }
            for (n = 0; n < N; ++n) {
                int c;
                REAL xi[8];
                REAL rowsum = 0;
                for (c = 0; c < npre; ++c) g[c] = 0;
                for (c = 0; c < NC; ++c) xi[c] = x[n * NV + c];
                switch (P->cpl.wkind) {
                case ORC_W_SCALAR_MEAN:
                    for (c = 0; c < npre; ++c) g[c] = mf[c];
                    break;
                case ORC_W_SCALAR_DIAG: /* construction.rs:639-655: s * I */
                    for (c = 0; c < npre; ++c)
                        g[c] = (REAL)P->cpl.scalar_w * (single_zero ? pre0[c] : src_single[(size_t)n * npre + c]);
                    break;
                case ORC_W_DENSE: { /* coupling.rs:465-480: W @ pre(x) */
                    int j;
                    const float* wr = P->cpl.dense_w + (size_t)n * N;
                    for (j = 0; j < N; ++j) {
                        for (c = 0; c < npre; ++c)
                            g[c] += (REAL)wr[j] * (single_zero ? pre0[c] : src_single[(size_t)j * npre + c]);
                        rowsum += (REAL)wr[j];
                    }
                    break; }
                case ORC_W_CSR: {
                    for (e = P->cpl.indptr[n]; e < P->cpl.indptr[n + 1]; ++e) {
                        const int j = P->cpl.indices[e];
                        const REAL w = (REAL)P->cpl.weights[e];
                        const REAL* src;
                        if (P->cpl.delay_mode == ORC_DELAY_PER_EDGE) {
                            /* runtime.rs:162-204 */
                            const int d = P->cpl.edge_delays ? P->cpl.edge_delays[e] : 0;
                            if (d == 0 || t == 0) src = cur + (size_t)j * npre;
                            else if (d > t) src = NULL;     /* exact zero contribution */
                            else src = hist + (size_t)((t - d) % Hd) * N * npre + (size_t)j * npre;
                            for (c = 0; c < npre; ++c) g[c] += w * (src ? src[c] : (REAL)0);
                        } else {
                            src = single_zero ? pre0 : src_single + (size_t)j * npre;
                            for (c = 0; c < npre; ++c) g[c] += w * src[c];
                        }
                        rowsum += w;
                    }
                    break; }
                }
                FN(post)(&cc, g, xi, rowsum, NC, cpl + (size_t)n * NC);
            }
        }
        /* 2. integrate every node: engine.rs:676-773 */
        for (n = 0; n < N; ++n) {
            const FN(mconst)* k = kn ? &kn[n] : &kbase;
            REAL* xn = x + (size_t)n * NV;
            const REAL* cn = cpl + (size_t)n * NC;
            REAL k1[8], k2[8], k3[8], k4[8], s[8], z[8];
            if (stoch) {
                const float* raw = P->noise + ((size_t)t * N + n) * NV;
                for (v = 0; v < NV; ++v) z[v] = (REAL)raw[v] * nscale[v];
            }
            FN(dfun)(k, xn, cn, k1);
            switch (P->integrator) {
            case ORC_EULER:
                for (v = 0; v < NV; ++v) xn[v] = xn[v] + k1[v] * dt;
                break;
            case ORC_EULER_STOCH:
                for (v = 0; v < NV; ++v) xn[v] = (xn[v] + k1[v] * dt) + z[v];
                break;
            case ORC_HEUN: case ORC_HEUN_STOCH:
                for (v = 0; v < NV; ++v) s[v] = stoch ? (xn[v] + k1[v] * dt) + z[v] : xn[v] + k1[v] * dt;
                FN(clamp)(model, s);
                FN(dfun)(k, s, cn, k2);
                for (v = 0; v < NV; ++v) {
                    REAL y = xn[v] + (k1[v] + k2[v]) * (dt * (REAL)0.5);
                    xn[v] = stoch ? y + z[v] : y;
                }
                break;
            case ORC_RK4: case ORC_RK4_STOCH:
                for (v = 0; v < NV; ++v) s[v] = xn[v] + k1[v] * (dt / (REAL)2);
                FN(clamp)(model, s);
                FN(dfun)(k, s, cn, k2);
                for (v = 0; v < NV; ++v) s[v] = xn[v] + k2[v] * (dt / (REAL)2);
                FN(clamp)(model, s);
                FN(dfun)(k, s, cn, k3);
                for (v = 0; v < NV; ++v) s[v] = xn[v] + k3[v] * dt;
                FN(clamp)(model, s);
                FN(dfun)(k, s, cn, k4);
                for (v = 0; v < NV; ++v)
                    xn[v] = xn[v] + (((k1[v] + k2[v] * (REAL)2) + k3[v] * (REAL)2) + k4[v]) * (dt / (REAL)6);
                if (P->integrator == ORC_RK4_STOCH) {
                    /* batch: clamp(det + Z) (engine.rs:747-765);
                     * plain: clamp(det) + Z, unclamped (integrator.rs:296-315) */
                    if (P->engine == ORC_ENGINE_PLAIN) {
                        FN(clamp)(model, xn);
                        for (v = 0; v < NV; ++v) xn[v] = xn[v] + z[v];
                        goto no_final_clamp;
                    }
                    for (v = 0; v < NV; ++v) xn[v] = xn[v] + z[v];
                }
                break;
            }
            FN(clamp)(model, xn);
        no_final_clamp:
            for (v = 0; v < NV; ++v) tavg[n * NV + v] += xn[v];
        }
        if (traj) memcpy(traj + (size_t)t * N * NV, x, sizeof(REAL) * (size_t)N * NV);
        /* 3. BOLD: engine.rs:785-801 -> bold_monitor.rs:72-107 (per instance) */
        if (P->bold_period > 0) {
            for (n = 0; n < N; ++n) bw_acc[n] += x[n * NV];
            bw_count++;
            if (bw_count >= P->bold_period) {
                if (nb < P->bold_cap)
                    FN(bold_window)(P, N, bw_acc, bw_count, bw_s, bw_f, bw_v, bw_q, bold + (size_t)nb * N);
                nb++;
                bw_count = 0;
                for (n = 0; n < N; ++n) bw_acc[n] = 0;
            }
        }
    }
    }
    if (P->bold_period > 0 && bw_count > 0) { /* flush, bold_monitor.rs:207-221 */
        if (nb < P->bold_cap)
            FN(bold_window)(P, N, bw_acc, bw_count, bw_s, bw_f, bw_v, bw_q, bold + (size_t)nb * N);
        nb++;
    }
    if (bold_n) *bold_n = nb;
    {   /* engine.rs:813-818: tavg * (1/n_steps) */
        REAL inv = (REAL)1 / (REAL)T;
        if (sizeof(REAL) == 4) inv = (REAL)(1.0f / (float)T);
        for (n = 0; n < N * NV; ++n) tavg[n] = tavg[n] * inv;
    }
    memcpy(fin, x, sizeof(REAL) * (size_t)N * NV);
done:
    free(x); free(hist); free(pre0); free(cpl); free(tmp); free(kn);
    free(bw_s); free(bw_f); free(bw_v); free(bw_q); free(bw_acc);
    return rc;
}
