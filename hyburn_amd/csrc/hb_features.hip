// hb_features.hip -- batched functional-connectivity features on the GPU.
//
//   hb_fc_batch   Pearson FC of each [T][N] series (fc.rs:11-57 fc_matrix):
//                 FP64 means and covariances, population standard deviation
//                 (std < 1e-12 -> 1), clamp to [-1, 1], f32 result, symmetric.
//   hb_fcd_batch  FC dynamics (fc.rs:94-179 fcd_matrix): the FC of every
//                 window of window_len steps taken every stride steps, then the
//                 Pearson correlation of the windows' upper triangles
//                 (upper_triangle_correlation, fc.rs:143-178), diagonal 1.
//
// The cross products Zc^T Zc of the centred columns are the GEMM-shaped part
// and run on the FP64 matrix cores (v_mfma_f64_16x16x4f64 over sample chunks
// staged in LDS), one workgroup of four waves per 96 x 96 block of one
// matrix. The FCD stage is the same kernel over the windows' upper triangles,
// written transposed so that each window is one column.
#include "hb_internal.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

namespace hb {
namespace feat {

// One correlation matrix to compute from the columns of x[t * ldx + n],
// t < T, n < N (FP64 column sums of nseg sample segments in mean[nseg][N],
// filled by k_fc_means).
//   fcd == 0: fc_matrix semantics (fc.rs:11-57): population std, std < 1e-12
//             -> 1, r = (cov / T) / (std_i std_j).
//   fcd == 1: upper_triangle_correlation semantics (fc.rs:143-178): r =
//             cov / (sqrt(var_i) sqrt(var_j)) of the unscaled centred sums, 0
//             when either var <= 0, diagonal 1.
// Either way r is clamped to [-1, 1] and rounded to f32 once. The result goes
// to out[(i * N + j) * ostr] for every (i, j) (upper == 0), or, upper == 1, to
// out[u * ostr] for the i < j entries numbered u in row order.
struct FcJob {
    const float* x;
    float* out;
    const double* mean;
    int32_t T, N, ldx, upper, ostr, fcd;
};

constexpr int kBlk = 96;                  // columns per block of the correlation matrix
constexpr int kChunk = 32;                // samples staged per LDS pass
constexpr int kCs = kBlk + 1;             // padded chunk row stride (doubles)
constexpr int kTilesPerWave = 5;          // 36 tiles of an off-diagonal 96 x 96 block / 8 waves
constexpr int kTilesPerWaveOne = 6;       // 21 lower-triangle tiles of a lone diagonal block / 4 waves
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void put(const FcJob& J, int i, int j, float r) {
    if (J.upper) {
        if (j > i) J.out[((size_t)i * J.N - (size_t)i * (i + 1) / 2 + (j - i - 1)) * J.ostr] = r;
        else if (i > j) J.out[((size_t)j * J.N - (size_t)j * (j + 1) / 2 + (i - j - 1)) * J.ostr] = r;
    } else {
        J.out[((size_t)i * J.N + j) * J.ostr] = r;
        J.out[((size_t)j * J.N + i) * J.ostr] = r;
    }
}

__device__ __forceinline__ double clamp1(double r) { return r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r); }

// Column sums, FP64, over sample segment z of ceil(T / nseg) rows:
// workgroup (job, 64-column group, segment); thread (column c, phase ph < 4)
// sums rows ph, ph + 4, ... of the segment and the four phases are added in
// order into part[z][c] (the job's mean pointer). Consecutive threads read
// consecutive columns of a row.
__global__ __launch_bounds__(256) void k_fc_means(const FcJob* __restrict__ jobs, int nseg) {
    __shared__ double part[4][64];
    const FcJob J = jobs[blockIdx.x];
    const int tid = threadIdx.x, cl = tid & 63, ph = tid >> 6;
    const int c = blockIdx.y * 64 + cl, z = blockIdx.z;
    const int seg = (J.T + nseg - 1) / nseg, t0 = z * seg, t1 = min(J.T, t0 + seg);
    double a = 0.0;
    if (c < J.N)
        for (int t = t0 + ph; t < t1; t += 4) a += (double)J.x[(size_t)t * J.ldx + c];
    part[ph][cl] = a;
    __syncthreads();
    if (tid < 64 && c < J.N)
        ((double*)J.mean)[(size_t)z * J.N + c] = ((part[0][cl] + part[1][cl]) + part[2][cl]) + part[3][cl];
}

// Per-column scale: fc_matrix's population std (< 1e-12 -> 1), or the raw
// sum of squares for upper_triangle_correlation.
__device__ __forceinline__ double col_scale(const FcJob& J, double var) {
    if (J.fcd) return var;
    const double s = sqrt(var / (double)J.T);
    return s < 1e-12 ? 1.0 : s;
}

__device__ __forceinline__ double corr_value(const FcJob& J, int i, int j, double cov, double si, double sj) {
    if (!J.fcd) return clamp1((cov / (double)J.T) / (si * sj));
    if (i == j) return 1.0;
    if (si <= 0.0 || sj <= 0.0) return 0.0;
    return clamp1(cov / (sqrt(si) * sqrt(sj)));
}

// Lower-triangle block pair bp -> (ib, jb), ib >= jb.
__device__ __forceinline__ void block_pair(int bp, int& ib, int& jb) {
    ib = 0;
    while ((ib + 1) * (ib + 2) / 2 <= bp) ++ib;
    jb = bp - ib * (ib + 1) / 2;
}

constexpr int kPart = kBlk * kBlk + 2 * kBlk;   // doubles of one partial: products, then sums of squares

// One 96 x 96 block (ib >= jb) of one job's correlation matrix per
// workgroup (four waves for a lone diagonal block, else eight). Centred samples of the block's row and column
// ranges are staged in LDS 32 at a time; the cross products run on the FP64
// matrix cores in 16 x 16 tiles (the lower triangle of a diagonal block), and
// the per-column sums of squares on the vector units in sample order. With
// nsplit > 1 the samples are split into nsplit chunk-aligned segments, one
// workgroup each, which store raw partial sums for k_fc_finish.
// Fragment layout of the 16x16x4 FP64 product (lane l, item r):
// A[i][k] = a(lane i + 16 k), B[k][j] = b(lane j + 16 k),
// D[l / 16 + 4 r][l % 16] = d(lane l)[r]. ONE: the matrix is a single
// diagonal block (N <= 96), which needs neither the second column-range
// buffers nor more than 6 tiles per wave.
template <bool ONE>
__global__ __launch_bounds__(ONE ? 256 : 512) void k_fc_tiles(const FcJob* __restrict__ jobs, int nbp, int nseg, int nsplit,
                                                  double* __restrict__ partial) {
    constexpr int kThreads = ONE ? 256 : 512, kWaves = kThreads / 64;
    constexpr int kTpw = ONE ? kTilesPerWaveOne : kTilesPerWave;
    __shared__ double cI[kChunk * kCs], cJ[ONE ? 1 : kChunk * kCs];
    __shared__ double mI[kBlk], mJ[kBlk], sI[kBlk], sJ[kBlk];
    const int z = blockIdx.x % nsplit, rest = blockIdx.x / nsplit;
    const int job = rest / nbp, bp = rest - job * nbp;
    const FcJob J = jobs[job];
    const int N = J.N, T = J.T, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lq = lane >> 4;
    const int cps = ((T + kChunk - 1) / kChunk + nsplit - 1) / nsplit;   // chunks per segment
    const int kbeg = min(T, z * cps * kChunk), kend = min(T, kbeg + cps * kChunk);
    int ib, jb;
    block_pair(bp, ib, jb);
    const bool diag = ONE || ib == jb;
    const int I0 = ib * kBlk, J0 = jb * kBlk;
    const int nI = min(kBlk, N - I0), nJ = min(kBlk, N - J0);
    const int tI = (nI + 15) >> 4, tJ = (nJ + 15) >> 4, wI = tI * 16, wJ = tJ * 16;
    const int ntt = diag ? tI * (tI + 1) / 2 : tI * tJ;
    double* bJ = diag ? cI : cJ;
    double* sB = diag ? sI : sJ;
    if (tid < nI || (tid >= 128 && tid - 128 < nJ)) {
        const int c = tid < 128 ? I0 + tid : J0 + tid - 128;
        double a = 0.0;
        for (int z = 0; z < nseg; ++z) a += J.mean[(size_t)z * N + c];
        (tid < 128 ? mI : mJ)[tid & 127] = a / (double)T;
    }
    int tti[kTpw], ttj[kTpw];
    d4 acc[kTpw];
#pragma unroll
    for (int q = 0; q < kTpw; ++q) {
        const int t = wave + kWaves * q;
        int ti = 0, tj = 0;
        if (t < ntt) {
            if (diag) {
                while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
                tj = t - ti * (ti + 1) / 2;
            } else {
                ti = t / tJ;
                tj = t - ti * tJ;
            }
        }
        tti[q] = ti;
        ttj[q] = tj;
        acc[q] = d4{0.0, 0.0, 0.0, 0.0};
    }
    // thread tid < nI owns row column I0 + tid; 128 + c < 128 + nJ owns J0 + c
    const int vc = tid < kBlk ? tid : tid - 128;
    const bool own_i = tid < nI, own_j = !diag && tid >= 128 && vc < nJ;
    double var = 0.0;
    // the next chunk's samples are loaded into registers while the current
    // chunk's products run: slot m of this thread is q = tid + kThreads m
    constexpr int kSlots = kChunk * kBlk / kThreads;
    float pI[kSlots], pJ[ONE ? 1 : kSlots];
    // (row, column) of each slot, packed as row << 8 | column, computed once
    int sI_[kSlots], sJ_[ONE ? 1 : kSlots];
#pragma unroll
    for (int m = 0; m < kSlots; ++m) {
        const int q = tid + kThreads * m, k = q / wI;
        sI_[m] = (k << 8) | (q - k * wI);
        if (!ONE && !diag) {
            const int kj = q / wJ;
            sJ_[m] = (kj << 8) | (q - kj * wJ);
        }
    }
    auto fetch = [&](int k0) {
        const int kc = min(kChunk, T - k0);
#pragma unroll
        for (int m = 0; m < kSlots; ++m) {
            const int k = sI_[m] >> 8, c = sI_[m] & 255;
            pI[m] = (k < kc && c < nI) ? J.x[(size_t)(k0 + k) * J.ldx + I0 + c] : 0.0f;
            if (!ONE && !diag) {
                const int kj = sJ_[m] >> 8, cj = sJ_[m] & 255;
                pJ[m] = (kj < kc && cj < nJ) ? J.x[(size_t)(k0 + kj) * J.ldx + J0 + cj] : 0.0f;
            }
        }
    };
    if (kbeg < kend) fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += kChunk) {
        const int kc = min(kChunk, T - k0);
        __syncthreads();
#pragma unroll
        for (int m = 0; m < kSlots; ++m) {
            const int k = sI_[m] >> 8, c = sI_[m] & 255;
            if (k < kChunk) cI[k * kCs + c] = (k < kc && c < nI) ? (double)pI[m] - mI[c] : 0.0;
            if (!ONE && !diag) {
                const int kj = sJ_[m] >> 8, cj = sJ_[m] & 255;
                if (kj < kChunk) cJ[kj * kCs + cj] = (kj < kc && cj < nJ) ? (double)pJ[m] - mJ[cj] : 0.0;
            }
        }
        __syncthreads();
        if (k0 + kChunk < kend) fetch(k0 + kChunk);
        if (own_i || own_j) {
            const double* col = (own_i ? cI : cJ) + vc;
            for (int k = 0; k < kc; ++k) var += col[k * kCs] * col[k * kCs];
        }
#pragma unroll
        for (int kk = 0; kk < kChunk; kk += 4) {
            const double* rowI = cI + (kk + lq) * kCs + lr;
            const double* rowJ = bJ + (kk + lq) * kCs + lr;
#pragma unroll
            for (int q = 0; q < kTpw; ++q)
                if (wave + kWaves * q < ntt) acc[q] = mfma_f64(rowI[tti[q] * 16], rowJ[ttj[q] * 16], acc[q]);
        }
    }
    if (nsplit > 1) {
        double* out = partial + ((size_t)rest * nsplit + z) * kPart;
        if (own_i || own_j) out[kBlk * kBlk + (own_i ? 0 : kBlk) + vc] = var;
#pragma unroll
        for (int q = 0; q < kTpw; ++q) {
            if (wave + kWaves * q >= ntt) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int il = tti[q] * 16 + lq + 4 * r, jl = ttj[q] * 16 + lr;
                if (il < nI && jl < nJ && !(diag && jl > il)) out[il * kBlk + jl] = acc[q][r];
            }
        }
        return;
    }
    if (own_i || own_j) (own_i ? sI : sJ)[vc] = col_scale(J, var);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kTpw; ++q) {
        if (wave + kWaves * q >= ntt) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int il = tti[q] * 16 + lq + 4 * r, jl = ttj[q] * 16 + lr;
            if (il >= nI || jl >= nJ || (diag && jl > il)) continue;
            put(J, I0 + il, J0 + jl, (float)corr_value(J, I0 + il, J0 + jl, acc[q][r], sI[il], sB[jl]));
        }
    }
}

// Adds the nsplit partials of one block pair in segment order and writes
// the correlations (workgroup = job x block pair x 256-entry slice).
__global__ __launch_bounds__(256) void k_fc_finish(const FcJob* __restrict__ jobs, int nbp, int nsplit,
                                                   const double* __restrict__ partial) {
    __shared__ double sI[kBlk], sJ[kBlk];
    const int rest = blockIdx.x, job = rest / nbp, bp = rest - job * nbp;
    const FcJob J = jobs[job];
    const int N = J.N, tid = threadIdx.x;
    int ib, jb;
    block_pair(bp, ib, jb);
    const bool diag = ib == jb;
    const int I0 = ib * kBlk, J0 = jb * kBlk;
    const int nI = min(kBlk, N - I0), nJ = min(kBlk, N - J0);
    const double* base = partial + (size_t)rest * nsplit * kPart;
    if (tid < nI || (!diag && tid >= 128 && tid - 128 < nJ)) {
        const int off = kBlk * kBlk + (tid < 128 ? tid : kBlk + tid - 128);
        double v = 0.0;
        for (int z = 0; z < nsplit; ++z) v += base[(size_t)z * kPart + off];
        (tid < 128 ? sI : sJ)[tid & 127] = col_scale(J, v);
    }
    __syncthreads();
    const double* sB = diag ? sI : sJ;
    {
        const int e = blockIdx.y * 256 + tid, il = e / kBlk, jl = e - il * kBlk;
        if (il >= nI || jl >= nJ || (diag && jl > il)) return;
        double cov = 0.0;
        for (int z = 0; z < nsplit; ++z) cov += base[(size_t)z * kPart + e];
        put(J, I0 + il, J0 + jl, (float)corr_value(J, I0 + il, J0 + jl, cov, sI[il], sB[jl]));
    }
}

// Per series, v[W][P] -> vt[P][W] through 64 x 64 LDS tiles (grid: P tiles,
// W tiles, series); reads and writes are row-contiguous.
__global__ __launch_bounds__(256) void k_transpose(const float* __restrict__ v, float* __restrict__ vt, int W,
                                                  int P) {
    __shared__ float tile[64][65];
    const int p0 = blockIdx.x * 64, w0 = blockIdx.y * 64, tid = threadIdx.x;
    const size_t base = (size_t)blockIdx.z * W * P;
    const int cx = tid & 63, ry = tid >> 6;
    for (int r = ry; r < 64; r += 4)
        if (w0 + r < W && p0 + cx < P) tile[r][cx] = v[base + (size_t)(w0 + r) * P + p0 + cx];
    __syncthreads();
    for (int r = ry; r < 64; r += 4)
        if (p0 + r < P && w0 + cx < W) vt[base + (size_t)(p0 + r) * W + w0 + cx] = tile[cx][r];
}

// ---------------------------------------------------------------- Welch PSD
// welch_psd (catch22/fft.rs:120-170) of every node series of every batch
// entry: segments of w samples every w / 2, minus the series mean, zero-padded
// to nf points; |DFT|^2 of bins 0..nf/2 summed over segments, / (nseg w fs).
// The DFT of a zero-padded segment is a product with a w x K twiddle table,
// so it runs on the FP64 matrix cores: one wave per 16 nodes x 16 bins,
// looping over segments and accumulating Re^2 + Im^2 in registers.

// tw[i][k] = (cos, -sin)(2 pi ((i k) mod nf) / nf), i < w, k < K.
__global__ __launch_bounds__(256) void k_twiddles(double2* __restrict__ tw, int w, int K, int nf) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)w * K) return;
    const int i = (int)(e / K), k = (int)(e - (int64_t)i * K);
    const int64_t r = ((int64_t)i * k) % nf;
    double sn, cs;
    sincospi(2.0 * (double)r / (double)nf, &sn, &cs);
    tw[e] = double2{cs, -sn};
}

struct WelchArgs {
    const float* y;          // [B][T][N]
    const double* sums;      // [B][nseg_mean][N] column sums (k_fc_means)
    const double2* tw;       // [w][K]
    double* psd;             // [B][N][K]
    int32_t T, N, w, K, nseg, nseg_mean, hop, pad;
    double inv_norm;         // 1 / (nseg_total * w * fs)
};

// grid: (ceil(K / 64), ceil(N / 16), B); wave v of the workgroup owns bins
// [64 bx + 16 v, +16) of nodes [16 by, +16).
__global__ __launch_bounds__(256) void k_welch(WelchArgs a) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lq = lane >> 4;
    const int b = blockIdx.z, n0 = blockIdx.y * 16, k0 = blockIdx.x * 64 + wave * 16;
    const int n = n0 + lr;
    double mean = 0.0;
    if (n < a.N) {
        for (int z = 0; z < a.nseg_mean; ++z) mean += a.sums[((size_t)b * a.nseg_mean + z) * a.N + n];
        mean /= (double)a.T;
    }
    const float* yb = a.y + (size_t)b * a.T * a.N;
    const bool kv = k0 + lr < a.K;
    d4 p = d4{0.0, 0.0, 0.0, 0.0};
    for (int seg = 0; seg < a.nseg; ++seg) {
        const int t0 = seg * a.hop;
        d4 re = d4{0.0, 0.0, 0.0, 0.0}, im = d4{0.0, 0.0, 0.0, 0.0};
        for (int i0 = 0; i0 < a.w; i0 += 4) {
            const int i = i0 + lq;
            const bool iv = i < a.w;
            const double av = (iv && n < a.N) ? (double)yb[(size_t)(t0 + i) * a.N + n] - mean : 0.0;
            const double2 t = (iv && kv) ? a.tw[(size_t)i * a.K + k0 + lr] : double2{0.0, 0.0};
            re = mfma_f64(av, t.x, re);
            im = mfma_f64(av, t.y, im);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) p[r] += re[r] * re[r] + im[r] * im[r];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int nn = n0 + lq + 4 * r;
        if (nn < a.N && kv) a.psd[((size_t)b * a.N + nn) * a.K + k0 + lr] = p[r] * a.inv_norm;
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// spectral_features (spectral.rs:20-99) from each PSD row: one wave per
// (batch entry, node); feats[row][9] = 5 band-power fractions, centroid,
// spread, skewness, excess kurtosis; NaN when the total power is not > 0 or
// not finite.
__global__ __launch_bounds__(256) void k_spectral(const double* __restrict__ psd, float* __restrict__ feats,
                                                  int64_t rows, int K, double df) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= rows) return;
    const double* P = psd + row * K;
    constexpr double lo[5] = {0.5, 4.0, 8.0, 13.0, 30.0}, hi[5] = {4.0, 8.0, 13.0, 30.0, 80.0};
    double tot = 0.0, pf = 0.0, band[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = lane; k < K; k += 64) {
        const double p = P[k], f = (double)k * df;
        tot += p;
        pf += p * f;
#pragma unroll
        for (int q = 0; q < 5; ++q)
            if (f >= lo[q] && f < hi[q]) band[q] += p;
    }
    tot = wave_sum(tot);
    pf = wave_sum(pf);
#pragma unroll
    for (int q = 0; q < 5; ++q) band[q] = wave_sum(band[q]);
    float* o = feats + row * 9;
    if (!(tot > 0.0) || !isfinite(tot)) {
        if (lane < 9) o[lane] = __builtin_nanf("");
        return;
    }
    const double c = pf / tot;
    double m2 = 0.0, m3 = 0.0, m4 = 0.0;
    for (int k = lane; k < K; k += 64) {
        const double p = P[k], d = (double)k * df - c, d2 = d * d;
        m2 += p * d2;
        m3 += p * d2 * d;
        m4 += p * d2 * d2;
    }
    m2 = wave_sum(m2);
    m3 = wave_sum(m3);
    m4 = wave_sum(m4);
    const double spread = sqrt(m2 / tot);
    const double skew = spread > 1e-12 ? m3 / tot / (spread * spread * spread) : 0.0;
    const double kurt = spread > 1e-12 ? m4 / tot / (spread * spread * spread * spread) - 3.0 : 0.0;
    if (lane < 5) {
        double v = band[0];
#pragma unroll
        for (int q = 1; q < 5; ++q)
            if (lane == q) v = band[q];
        o[lane] = (float)(v / tot);
    } else if (lane < 9) {
        const double v = lane == 5 ? c : (lane == 6 ? spread : (lane == 7 ? skew : kurt));
        o[lane] = (float)v;
    }
}

// Launch shape of one FC call: column sums in up to 8 segments of >= 256
// samples (added in order), and the sample split of the tile kernel, chosen
// so that a launch has about two workgroups per CU while each segment keeps
// at least four 32-sample chunks.
struct FcShape {
    int nseg, nb, nbp, nsplit;
    size_t blocks;
};

FcShape fc_shape(size_t nj, int N, int T) {
    FcShape s;
    s.nseg = std::max(1, std::min(8, T / 256));
    s.nb = (N + kBlk - 1) / kBlk;
    s.nbp = s.nb * (s.nb + 1) / 2;
    s.blocks = nj * (size_t)s.nbp;
    const int nch = (T + kChunk - 1) / kChunk;
    int nsplit = (int)std::min<size_t>(16, (512 + s.blocks - 1) / s.blocks);
    s.nsplit = std::max(1, std::min(nsplit, nch / 4));
    return s;
}

// Doubles of the column sums and of the split partials a call needs.
size_t fc_mean_doubles(size_t nj, int N, const FcShape& s) { return nj * (size_t)N * s.nseg; }
size_t fc_partial_doubles(const FcShape& s) { return s.nsplit > 1 ? s.blocks * s.nsplit * kPart : 0; }

// The kernels of one call on stream st: means, block-pair tiles, and the
// ordered sum of the split partials. dj holds nj jobs with their mean
// pointers set; partial holds fc_partial_doubles(s) doubles (or null).
hipError_t launch_fc(const FcJob* dj, size_t nj, int N, const FcShape& s, double* partial, hipStream_t st) {
    if ((uint64_t)nj * (uint64_t)s.nbp * 16u > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fc_means, dim3((unsigned)nj, (unsigned)((N + 63) / 64), (unsigned)s.nseg), dim3(256), 0, st,
                       dj, s.nseg);
    if (s.nb == 1)
        hipLaunchKernelGGL(k_fc_tiles<true>, dim3((unsigned)(s.blocks * s.nsplit)), dim3(256), 0, st, dj, s.nbp,
                           s.nseg, s.nsplit, partial);
    else
        hipLaunchKernelGGL(k_fc_tiles<false>, dim3((unsigned)(s.blocks * s.nsplit)), dim3(512), 0, st, dj, s.nbp,
                           s.nseg, s.nsplit, partial);
    if (s.nsplit > 1)
        hipLaunchKernelGGL(k_fc_finish, dim3((unsigned)s.blocks, (unsigned)((std::min(N, kBlk) * kBlk + 255) / 256)),
                           dim3(256), 0, st, dj, s.nbp, s.nsplit, partial);
    return hipGetLastError();
}

// Every job of one call has the same N. Fills each job's mean pointer from a
// work buffer, uploads the jobs and launches them on the null stream.
hipError_t run_fc_jobs(std::vector<FcJob>& jobs, int N, int T, DevBufs& B) {
    const size_t nj = jobs.size();
    const FcShape s = fc_shape(nj, N, T);
    double* mean = nullptr;
    FcJob* dj = nullptr;
    hipError_t e = B.alloc(&mean, fc_mean_doubles(nj, N, s));
    if (e == hipSuccess) e = B.alloc(&dj, nj);
    if (e != hipSuccess) return e;
    for (size_t q = 0; q < nj; ++q) jobs[q].mean = mean + q * (size_t)N * s.nseg;
    e = hipMemcpy(dj, jobs.data(), sizeof(FcJob) * nj, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    double* partial = nullptr;
    if (fc_partial_doubles(s)) {
        e = B.alloc(&partial, fc_partial_doubles(s));
        if (e != hipSuccess) return e;
    }
    return launch_fc(dj, nj, N, s, partial, nullptr);
}

// The jobs of a full-matrix FC call over B contiguous [T][N] series, written
// on the device (the no-allocation entry point has no host staging).
__global__ void k_fc_make_jobs(FcJob* jobs, const float* x, float* out, double* mean, int B, int T, int N,
                               int nseg) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    FcJob J;
    J.x = x + (size_t)T * N * b;
    J.out = out + (size_t)N * N * b;
    J.mean = mean + (size_t)b * N * nseg;
    J.T = T;
    J.N = N;
    J.ldx = N;
    J.upper = 0;
    J.ostr = 1;
    J.fcd = 0;
    jobs[b] = J;
}

static int next_pow2(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// Device Welch PSD of every node series; psd is allocated in B and filled.
// nf = next_pow2(max(nfft, w)); the caller has validated the sizes.
hipError_t run_welch(const float* dy, int Bn, int T, int N, int nf, int w, double fs, DevBufs& B, double** psd_out) {
    const int K = nf / 2 + 1, hop = w / 2;
    const int nseg_total = T > w ? (T - w) / hop + 1 : 1;
    const int nseg = T >= w ? nseg_total : 0;   // segments that fit (welch_psd breaks on the rest)
    const int nsm = std::max(1, std::min(8, T / 256));
    double *sums = nullptr, *psd = nullptr;
    double2* tw = nullptr;
    FcJob* dj = nullptr;
    hipError_t e = B.alloc(&sums, (size_t)Bn * nsm * N);
    if (e == hipSuccess) e = B.alloc(&psd, (size_t)Bn * N * K);
    if (e == hipSuccess) e = B.alloc(&tw, (size_t)w * K);
    if (e == hipSuccess) e = B.alloc(&dj, (size_t)Bn);
    if (e != hipSuccess) return e;
    std::vector<FcJob> jobs((size_t)Bn);
    for (int b = 0; b < Bn; ++b)
        jobs[b] = FcJob{dy + (size_t)b * T * N, nullptr, sums + (size_t)b * nsm * N, T, N, N, 0, 1, 0};
    e = hipMemcpy(dj, jobs.data(), sizeof(FcJob) * jobs.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fc_means, dim3((unsigned)Bn, (unsigned)((N + 63) / 64), (unsigned)nsm), dim3(256), 0, nullptr,
                       dj, nsm);
    const int64_t nt = (int64_t)w * K;
    hipLaunchKernelGGL(k_twiddles, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, nullptr, tw, w, K, nf);
    WelchArgs a{dy, sums, tw, psd, T, N, w, K, nseg, nsm, hop, 0, 1.0 / ((double)nseg_total * (double)w * fs)};
    hipLaunchKernelGGL(k_welch, dim3((unsigned)((K + 63) / 64), (unsigned)((N + 15) / 16), (unsigned)Bn), dim3(256), 0,
                       nullptr, a);
    *psd_out = psd;
    return hipGetLastError();
}

// Scratch of the no-allocation FC entry point (hb_fc_batch_noalloc): jobs,
// column sums and split partials of one call. Each GPU holds its own copy
// from the first call on it until the process ends; its address comes from
// hipGetSymbolAddress with that device current (the symbol's name is only a
// host-side stand-in). Calls that overlap on one GPU share it and race.
constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

}  // namespace feat
}  // namespace hb

using namespace hb;

namespace {

int check_series(const float* series, int32_t B, int32_t T, int32_t N, const void* out) {
    if (!series || !out) return set_error(HB_INVALID_CONFIG, "null series or output pointer");
    if (B < 1 || T < 1 || N < 1)
        return set_error(HB_INVALID_CONFIG, "n_series, n_steps and n_nodes must be positive");
    return HB_OK;
}

int hip_fail(hipError_t e, const char* what) {
    return set_error(HB_BACKEND_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

// Device address of `scratch` on each device, looked up on the first
// no-allocation call there (with that device current) and kept.
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch;

hipError_t scratch_on(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it != g_scratch.end()) { *out = (unsigned char*)it->second; return hipSuccess; }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(feat::scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = p;
    *out = (unsigned char*)p;
    return hipSuccess;
}

// True when p is device memory of `device` (or managed memory) that the
// kernels may read or write.
bool device_memory_on(const void* p, int device) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (a.type == hipMemoryTypeManaged) return true;
    return a.type == hipMemoryTypeDevice && a.device == device;
}

}  // namespace

extern "C" {

int hb_fc_batch(const float* series, int32_t n_series, int32_t n_steps, int32_t n_nodes, float* fc, int32_t device) {
    const int rc = check_series(series, n_series, n_steps, n_nodes, fc);
    if (rc) return rc;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const size_t per = (size_t)n_steps * n_nodes, nn = (size_t)n_nodes * n_nodes;
    DevBufs B;
    float *dx = nullptr, *dout = nullptr;
    e = B.alloc(&dx, per * n_series);
    if (e == hipSuccess) e = B.alloc(&dout, nn * n_series);
    if (e == hipSuccess) e = hipMemcpy(dx, series, sizeof(float) * per * n_series, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "FC upload");
    std::vector<feat::FcJob> jobs((size_t)n_series);
    for (int32_t b = 0; b < n_series; ++b)
        jobs[b] = feat::FcJob{dx + per * b, dout + nn * b, nullptr, n_steps, n_nodes, n_nodes, 0, 1, 0};
    e = feat::run_fc_jobs(jobs, n_nodes, n_steps, B);
    if (e == hipSuccess) e = hipMemcpy(fc, dout, sizeof(float) * nn * n_series, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "FC kernels");
    return HB_OK;
}

int hb_fc_batch_noalloc(const float* d_series, int32_t n_series, int32_t n_steps, int32_t n_nodes, float* d_fc,
                        int32_t device, void* stream) {
    const int rc = check_series(d_series, n_series, n_steps, n_nodes, d_fc);
    if (rc) return rc;
    const size_t nj = (size_t)n_series;
    const feat::FcShape s = feat::fc_shape(nj, n_nodes, n_steps);
    const size_t jobs_b = (sizeof(feat::FcJob) * nj + 255) & ~(size_t)255;
    const size_t mean_b = (sizeof(double) * feat::fc_mean_doubles(nj, n_nodes, s) + 255) & ~(size_t)255;
    const size_t part_b = sizeof(double) * feat::fc_partial_doubles(s);
    if ((uint64_t)nj * (uint64_t)s.nbp * 16u > 0x7fffffffu || jobs_b + mean_b + part_b > feat::kNoallocScratchBytes)
        return set_error(HB_INVALID_CONFIG, "hb_fc_batch_noalloc: this call needs " +
                                                std::to_string(jobs_b + mean_b + part_b) + " bytes of scratch; the "
                                                "no-allocation entry point holds " +
                                                std::to_string(feat::kNoallocScratchBytes) + " per GPU");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const size_t nx = (size_t)n_steps * n_nodes * nj;
    if (!device_memory_on(d_series, device) || !device_memory_on(d_series + nx - 1, device) ||
        !device_memory_on(d_fc, device) || !device_memory_on(d_fc + (size_t)n_nodes * n_nodes * nj - 1, device))
        return set_error(HB_INVALID_CONFIG, "hb_fc_batch_noalloc: series and fc must be device memory of the "
                                            "given device");
    unsigned char* base = nullptr;
    e = scratch_on(device, &base);
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    feat::FcJob* dj = (feat::FcJob*)base;
    double* mean = (double*)(base + jobs_b);
    double* partial = part_b ? (double*)(base + jobs_b + mean_b) : nullptr;
    const hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(feat::k_fc_make_jobs, dim3((unsigned)((nj + 63) / 64)), dim3(64), 0, st, dj, d_series, d_fc,
                       mean, n_series, n_steps, n_nodes, s.nseg);
    e = hipGetLastError();
    if (e == hipSuccess) e = feat::launch_fc(dj, nj, n_nodes, s, partial, st);
    if (e != hipSuccess) return hip_fail(e, "FC kernels (no-allocation)");
    return HB_OK;
}

int32_t hb_fcd_windows(int32_t n_steps, int32_t window_len, int32_t stride) {
    if (window_len < 2 || window_len > n_steps || stride < 1) return 0;
    return (n_steps - window_len) / stride + 1;
}

int hb_fcd_batch(const float* series, int32_t n_series, int32_t n_steps, int32_t n_nodes, int32_t window_len,
                 int32_t stride, float* fcd, int32_t device) {
    const int rc = check_series(series, n_series, n_steps, n_nodes, fcd);
    if (rc) return rc;
    if (window_len < 2 || window_len > n_steps)
        return set_error(HB_INVALID_CONFIG, "window_len must be in [2, n_steps]");
    if (stride < 1) return set_error(HB_INVALID_CONFIG, "stride must be > 0");
    const int32_t W = hb_fcd_windows(n_steps, window_len, stride);
    if (W <= 1) return HB_OK;   // fcd_matrix returns an empty matrix
    const size_t per = (size_t)n_steps * n_nodes;
    const size_t P = (size_t)n_nodes * (n_nodes - 1) / 2;
    if (P > 0x7fffffffu / 64 || n_series > 65535)
        return set_error(HB_INVALID_CONFIG, "FCD needs n_nodes (n_nodes - 1) / 2 < 2^25 and n_series <= 65535");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // window FCs are written as v[w][p] (W windows x P upper entries per
    // series) and transposed to vt[p][w]: the FCD is the column correlation
    // of that [P][W] series
    DevBufs B;
    float *dx = nullptr, *dv = nullptr, *dvt = nullptr, *dout = nullptr;
    e = B.alloc(&dx, per * n_series);
    if (e == hipSuccess) e = B.alloc(&dv, P * W * n_series);
    if (e == hipSuccess) e = B.alloc(&dvt, P * W * n_series);
    if (e == hipSuccess) e = B.alloc(&dout, (size_t)W * W * n_series);
    if (e == hipSuccess) e = hipMemcpy(dx, series, sizeof(float) * per * n_series, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "FCD upload");
    if (P > 0) {
        std::vector<feat::FcJob> jobs((size_t)n_series * W);
        for (int32_t b = 0; b < n_series; ++b)
            for (int32_t w = 0; w < W; ++w)
                jobs[(size_t)b * W + w] = feat::FcJob{dx + per * b + (size_t)w * stride * n_nodes,
                                                      dv + P * ((size_t)b * W + w), nullptr, window_len, n_nodes,
                                                      n_nodes, 1, 1, 0};
        e = feat::run_fc_jobs(jobs, n_nodes, window_len, B);
        if (e != hipSuccess) return hip_fail(e, "FCD window FC");
        hipLaunchKernelGGL(feat::k_transpose, dim3((unsigned)((P + 63) / 64), (unsigned)((W + 63) / 64),
                                                   (unsigned)n_series), dim3(256), 0, nullptr, dv, dvt, W, (int)P);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "FCD transpose");
    }
    std::vector<feat::FcJob> corr((size_t)n_series);
    for (int32_t b = 0; b < n_series; ++b)
        corr[b] = feat::FcJob{dvt + P * W * b, dout + (size_t)W * W * b, nullptr, (int32_t)P, W, W, 0, 1, 1};
    e = feat::run_fc_jobs(corr, W, (int)P, B);
    if (e == hipSuccess) e = hipMemcpy(fcd, dout, sizeof(float) * (size_t)W * W * n_series, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "FCD kernels");
    return HB_OK;
}

int32_t hb_welch_bins(int32_t nfft, int32_t window_width) {
    if (window_width < 2 || nfft < 0) return 0;
    const int32_t m = std::max(nfft, window_width);
    if (m > 65536) return 0;
    return feat::next_pow2(m) / 2 + 1;
}

int hb_welch_psd_batch(const float* series, int32_t n_series, int32_t n_steps, int32_t n_nodes, int32_t nfft,
                       double fs, int32_t window_width, double* psd, int32_t device) {
    const int rc = check_series(series, n_series, n_steps, n_nodes, psd);
    if (rc) return rc;
    if (window_width < 2) return set_error(HB_INVALID_CONFIG, "window_width must be >= 2 (its half is the hop)");
    const int32_t K = hb_welch_bins(nfft, window_width);
    if (K == 0 || nfft < 0) return set_error(HB_INVALID_CONFIG, "nfft and window_width must be in [0, 65536]");
    if ((int64_t)window_width * K > (1 << 26) || n_series > 65535)
        return set_error(HB_INVALID_CONFIG, "window_width * bins must be <= 2^26 and n_series <= 65535");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const size_t per = (size_t)n_steps * n_nodes;
    DevBufs B;
    float* dy = nullptr;
    double* dpsd = nullptr;
    e = B.alloc(&dy, per * n_series);
    if (e == hipSuccess) e = hipMemcpy(dy, series, sizeof(float) * per * n_series, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = feat::run_welch(dy, n_series, n_steps, n_nodes, 2 * (K - 1), window_width, fs, B, &dpsd);
    if (e == hipSuccess)
        e = hipMemcpy(psd, dpsd, sizeof(double) * (size_t)n_series * n_nodes * K, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "Welch PSD");
    return HB_OK;
}

int hb_spectral_features_batch(const float* series, int32_t n_series, int32_t n_steps, int32_t n_nodes, double fs,
                               float* feats, int32_t device) {
    const int rc = check_series(series, n_series, n_steps, n_nodes, feats);
    if (rc) return rc;
    if (n_steps > 65536 || n_series > 65535)
        return set_error(HB_INVALID_CONFIG, "n_steps must be <= 65536 and n_series <= 65535");
    const size_t rows = (size_t)n_series * n_nodes;
    if (n_steps < 16) {   // spectral_features: too short for a PSD
        for (size_t q = 0; q < rows * 9; ++q) feats[q] = __builtin_nanf("");
        return HB_OK;
    }
    const int nf = feat::next_pow2(n_steps);
    const int w = std::min(std::max(n_steps / 2, 16), 256);
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const size_t per = (size_t)n_steps * n_nodes;
    DevBufs B;
    float *dy = nullptr, *dfeat = nullptr;
    double* dpsd = nullptr;
    e = B.alloc(&dy, per * n_series);
    if (e == hipSuccess) e = B.alloc(&dfeat, rows * 9);
    if (e == hipSuccess) e = hipMemcpy(dy, series, sizeof(float) * per * n_series, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = feat::run_welch(dy, n_series, n_steps, n_nodes, nf, w, fs, B, &dpsd);
    if (e != hipSuccess) return hip_fail(e, "spectral PSD");
    hipLaunchKernelGGL(feat::k_spectral, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, nullptr, dpsd, dfeat,
                       (int64_t)rows, nf / 2 + 1, fs / (double)nf);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(feats, dfeat, sizeof(float) * rows * 9, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "spectral features");
    return HB_OK;
}

}  // extern "C"
