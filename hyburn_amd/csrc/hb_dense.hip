// hb_dense.hip -- dense all-to-all coupling on the matrix cores for batches of
// instances that share one weight matrix and have no delay.
//
// The batch engine forms a dense projection's coupling as one batched product
// per step, pre(x)[B, Nsrc] x W^T[Nsrc, Ntgt] (src/engine/batch_engine/
// engine.rs:609-620, src/engine/coupling.rs:465-480). Here a workgroup owns a
// tile of 16 instances for the whole run and forms that product every step on
// v_mfma_f32_16x16x4f32:
//
//   C[tgt, inst] = sum_src W[tgt, src] * P[inst, src]
//
// with W (row = target, odd LDS row stride) and the tile's pre values P
// (double-buffered, one barrier per step) in LDS. Wave w owns targets
// 16w..16w+15; lane l owns instance l % 16 and targets 16w + 4(l/16) + r,
// r = 0..3, which is exactly where the instruction leaves C[.., ..] (an
// in-bounds probe on gfx950 gave D[4(l/16) + r][l % 16] in lane l, VGPR r).
// Each lane then applies post() and integrates its four nodes (independent
// chains), so model constants are held once per lane.
//
// FMA-mode only: the product's summation order is the matrix core's, so
// results are checked against the FP64 twin of the oracle (the exact mode
// keeps the sequential row sums of hb_forward.hip).
#include "hb_internal.h"

namespace hb {
namespace dense {

#include "hb_model.h"

typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> using gp = __attribute__((address_space(1))) T*;

template <int M, int NP>
__global__ __launch_bounds__(512) void k_dense(LaunchDesc a) {
    extern __shared__ float lds[];
    constexpr int NV = Tr<M>::NV, NC = Tr<M>::NC;
    const int N = a.n_nodes, Nk = a.nk, Sw = a.nk + 1, nw = blockDim.x >> 6;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, ii = l & 15, q = l >> 4;
    float* __restrict__ Wl = lds;                               // [nw * 16][Sw]
    float* __restrict__ Pl = lds + (size_t)nw * 16 * Sw;        // [2][NP][16][Sw]
    const int tile = blockIdx.x;
    const int inst = a.order[tile * 16 + ii];
    const bool valid = inst >= 0;
    const DevInst* __restrict__ I = a.insts + (valid ? inst : a.order[tile * 16]);
    const DevInst* __restrict__ I0 = a.insts + a.order[tile * 16];

    // W rows of this workgroup's targets (row = target), transposed from the
    // shared [src][tgt] copy; zero beyond N
    {
        const gp<const float> wt = (gp<const float>)I0->dense_t;
        const int total = nw * 16 * Sw;
        for (int i = tid; i < total; i += blockDim.x) {
            const int n = i / Sw, k = i - n * Sw;
            Wl[i] = (n < N && k < N) ? wt[(size_t)k * N + n] : 0.0f;
        }
        // every P entry the products read (columns < Nk, 16 rows, both
        // buffers) is written by a pre() store before the barrier of its step
    }
    // this lane's instance: constants, state of its four nodes
    float mc[kMC], cc[kCC], nsc[NV];
#pragma unroll
    for (int i = 0; i < kMC; ++i) mc[i] = I->mc[i];
#pragma unroll
    for (int i = 0; i < kCC; ++i) cc[i] = I->cc[i];
#pragma unroll
    for (int v = 0; v < NV; ++v) nsc[v] = I->nscale[v];
    const int cfn = I0->cfn, integ = I0->integrator, T = I0->n_steps;
    const bool classic = I0->sjr_classic != 0, cpu_form = I0->cpu_form != 0;
    StepCtx S;
    S.dt = I0->dt; S.dth = I0->dt_half; S.dt6 = I0->dt_sixth; S.integ = integ;
    S.stoch = integ >= HB_EULER_STOCH; S.plain = I0->engine == HB_ENGINE_PLAIN;
    S.sw = I0->sweep_form != 0; S.sv = I->sweep_val;
    const gp<const float> noise = (gp<const float>)I->noise;
    const gp<float> traj = (gp<float>)I->traj;
    const int64_t tstride = I0->traj_tstride;
    const int n0 = w * 16 + 4 * q;   // this lane's first target
    float x[4][NV], tv[4][NV];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const bool own = valid && n0 + r < N;
            x[r][v] = own ? I->x0[(n0 + r) * NV + v] : 0.0f;
            tv[r][v] = 0.0f;
        }
    int buf = 0;
    for (int t = 0; t < T; ++t) {
        // ---- pre(x_t) of this lane's nodes into the P buffer of this step
        float own_pre[4][NP];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float xc[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) xc[c] = x[r][c];
            pre<NC, NP>(cfn, cc, cpu_form, classic, xc, own_pre[r]);
            if (n0 + r < Nk) {
#pragma unroll
                for (int c = 0; c < NP; ++c)
                    Pl[((size_t)(buf * NP + c) * 16 + ii) * Sw + n0 + r] = (valid && n0 + r < N) ? own_pre[r][c] : 0.0f;
            }
        }
        __syncthreads();
        // ---- C[tgt, inst] = sum_src W[tgt, src] P[inst, src] on the matrix cores
        v4f acc[NP];
        const float* __restrict__ wrow = Wl + (size_t)(w * 16 + ii) * Sw + q;
#pragma unroll
        for (int c = 0; c < NP; ++c) {
            const float* __restrict__ prow = Pl + ((size_t)(buf * NP + c) * 16 + ii) * Sw + q;
            v4f s0 = {0.0f, 0.0f, 0.0f, 0.0f}, s1 = {0.0f, 0.0f, 0.0f, 0.0f};
            int k0 = 0;
            for (; k0 + 8 <= Nk; k0 += 8) {   // two accumulators in flight
                s0 = __builtin_amdgcn_mfma_f32_16x16x4f32(wrow[k0], prow[k0], s0, 0, 0, 0);
                s1 = __builtin_amdgcn_mfma_f32_16x16x4f32(wrow[k0 + 4], prow[k0 + 4], s1, 0, 0, 0);
            }
            if (k0 < Nk) s0 = __builtin_amdgcn_mfma_f32_16x16x4f32(wrow[k0], prow[k0], s0, 0, 0, 0);
            acc[c] = s0 + s1;
        }
        // ---- post, integrate the four nodes
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int n = n0 + r;
            float g[NP], cpl[NC], z[NV];
#pragma unroll
            for (int c = 0; c < NP; ++c) g[c] = acc[c][r];
            post<NC, NP>(cfn, cc, classic, g, own_pre[r], cpl);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                z[v] = (S.stoch && valid && n < N) ? noise[((size_t)t * N + n) * NV + v] * nsc[v] : 0.0f;
            integrate<M>(S, mc, cpl, z, x[r]);
#pragma unroll
            for (int v = 0; v < NV; ++v) tv[r][v] = tv[r][v] + x[r][v];
            if (traj && valid && n < N) {
                gp<float> tr = traj + (size_t)t * tstride + (size_t)n * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) tr[v] = x[r][v];
            }
        }
        buf ^= 1;
    }
    if (valid) {
        const float inv = I->inv_nsteps;
        const gp<float> fin = (gp<float>)I->fin, tavg = (gp<float>)I->tavg;
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (n0 + r < N) {
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    fin[(n0 + r) * NV + v] = x[r][v];
                    tavg[(n0 + r) * NV + v] = tv[r][v] * inv;
                }
            }
    }
}

const void* kernel_of(int model, int np) {
    switch (model * 4 + np) {
    case HB_G2DO * 4 + 1: return (const void*)k_dense<HB_G2DO, 1>;
    case HB_MPR * 4 + 1: return (const void*)k_dense<HB_MPR, 1>;
    case HB_MPR * 4 + 2: return (const void*)k_dense<HB_MPR, 2>;
    case HB_JANSEN_RIT * 4 + 1: return (const void*)k_dense<HB_JANSEN_RIT, 1>;
    case HB_WILSON_COWAN * 4 + 1: return (const void*)k_dense<HB_WILSON_COWAN, 1>;
    case HB_KURAMOTO * 4 + 1: return (const void*)k_dense<HB_KURAMOTO, 1>;
    case HB_KURAMOTO * 4 + 2: return (const void*)k_dense<HB_KURAMOTO, 2>;
    case HB_RWW * 4 + 1: return (const void*)k_dense<HB_RWW, 1>;
    default: return nullptr;
    }
}

}  // namespace dense

int dense_lds_bytes(int n_nodes, int np) {
    const int nw = (n_nodes + 15) / 16, nk = (n_nodes + 3) / 4 * 4;
    return (nw * 16 + 2 * np * 16) * (nk + 1) * 4;
}

hipError_t launch_dense(const LaunchDesc& d, hipStream_t s) {
    const void* k = dense::kernel_of(d.model, d.np);
    if (!k) return hipErrorInvalidValue;
    LaunchDesc a = d;
    void* args[] = {&a};
    return hipLaunchKernel(k, dim3(d.grid), dim3(d.block), args, (size_t)d.lds_bytes, s);
}

hipError_t dense_set_attributes(int max_lds) {
    for (int m = 0; m <= HB_RWW; ++m)
        for (int np = 1; np <= 2; ++np) {
            const void* k = dense::kernel_of(m, np);
            if (!k) continue;
            const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

}  // namespace hb
