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
// The covariance Zc^T Zc is the GEMM-shaped part: for N <= 96 it runs on the
// FP64 matrix cores (v_mfma_f64_16x16x4f64, lower-triangle 16x16 tiles over
// sample chunks staged in LDS), one workgroup of four waves per series or
// window; wider networks take a pair-per-thread FP64 loop.
#include "hb_internal.h"

#include <vector>

namespace hb {
namespace feat {

// One FC to compute: rows x[t * ldx + n], t < T, n < N; out receives the N x N
// matrix (upper == 0) or the i < j entries in row order (upper == 1).
struct FcJob {
    const float* x;
    float* out;
    int32_t T, N, ldx, upper;
};

constexpr int kTile = 96, kChunk = 32, kCs = kTile + 1;   // padded chunk row stride (doubles)
constexpr int kTilesPerWave = 6;                          // ceil(21 lower-triangle tiles / 4 waves)
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void put(const FcJob& J, int i, int j, float r) {
    if (J.upper) {
        if (j > i) J.out[(size_t)i * J.N - (size_t)i * (i + 1) / 2 + (j - i - 1)] = r;
        else if (i > j) J.out[(size_t)j * J.N - (size_t)j * (j + 1) / 2 + (i - j - 1)] = r;
    } else {
        J.out[(size_t)i * J.N + j] = r;
        J.out[(size_t)j * J.N + i] = r;
    }
}

__device__ __forceinline__ float pearson(double cov, double si, double sj) {
    double r = cov / (si * sj);
    r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
    return (float)r;
}

// Column means over the rows, FP64 (N <= 96): thread (column c, phase ph)
// sums rows ph, ph + nph, ... and the phases are added in order.
__device__ void column_means(const FcJob& J, double* part, double* mean) {
    const int N = J.N, T = J.T, tid = threadIdx.x;
    const int nph = min(4, (int)blockDim.x / N);
    const int c = tid % N, ph = tid / N;
    if (ph < nph) {
        double a = 0.0;
        for (int t = ph; t < T; t += nph) a += (double)J.x[(size_t)t * J.ldx + c];
        part[ph * kTile + c] = a;
    }
    __syncthreads();
    if (tid < N) {
        double a = 0.0;
        for (int q = 0; q < nph; ++q) a += part[q * kTile + tid];
        mean[tid] = a / (double)T;
    }
    __syncthreads();
}

// N <= 96: covariance tiles on the matrix cores. Fragment layout of the
// 16x16x4 FP64 product (lane l, item r): A[i][k] = a(lane i + 16 k),
// B[k][j] = b(lane j + 16 k), D[l / 16 + 4 r][l % 16] = d(lane l)[r].
__global__ __launch_bounds__(256) void k_fc_tiles(const FcJob* __restrict__ jobs) {
    __shared__ double chunk[kChunk * kCs];
    __shared__ double mean[kTile], sd[kTile], part[4 * kTile];
    const FcJob J = jobs[blockIdx.x];
    const int N = J.N, T = J.T, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lq = lane >> 4;
    constexpr int nwave = 4;
    const int nt = (N + 15) >> 4, NP = nt * 16;
    column_means(J, part, mean);
    const int ntt = nt * (nt + 1) / 2;
    int tti[kTilesPerWave], ttj[kTilesPerWave];
    d4 acc[kTilesPerWave];
#pragma unroll
    for (int q = 0; q < kTilesPerWave; ++q) {
        const int t = wave + q * nwave;
        int ti = 0, tj = 0;
        if (t < ntt) {
            while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
            tj = t - ti * (ti + 1) / 2;
        }
        tti[q] = ti; ttj[q] = tj;
        acc[q] = d4{0.0, 0.0, 0.0, 0.0};
    }
    for (int k0 = 0; k0 < T; k0 += kChunk) {
        const int kc = min(kChunk, T - k0);
        __syncthreads();
        for (int q = tid; q < kChunk * NP; q += 256) {
            const int k = q / NP, c = q - k * NP;
            chunk[k * kCs + c] = (k < kc && c < N) ? (double)J.x[(size_t)(k0 + k) * J.ldx + c] - mean[c] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kChunk; kk += 4) {
            const double* row = chunk + (kk + lq) * kCs + lr;
#pragma unroll
            for (int q = 0; q < kTilesPerWave; ++q)
                if (wave + q * nwave < ntt) acc[q] = mfma_f64(row[tti[q] * 16], row[ttj[q] * 16], acc[q]);
        }
    }
    // population standard deviations from the diagonal tiles
#pragma unroll
    for (int q = 0; q < kTilesPerWave; ++q) {
        if (wave + q * nwave < ntt && tti[q] == ttj[q]) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = tti[q] * 16 + lq + 4 * r, j = ttj[q] * 16 + lr;
                if (i == j && i < N) {
                    const double s = sqrt(acc[q][r] / (double)T);
                    sd[i] = s < 1e-12 ? 1.0 : s;
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kTilesPerWave; ++q) {
        if (wave + q * nwave >= ntt) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = tti[q] * 16 + lq + 4 * r, j = ttj[q] * 16 + lr;
            if (j <= i && i < N && j < N) put(J, i, j, pearson(acc[q][r] / (double)T, sd[i], sd[j]));
        }
    }
}

// N > 96: FP64 means and deviations per column, then the lower-triangle
// pairs (i, j <= i) spread over the threads, each an FP64 sum in row order.
__global__ __launch_bounds__(256) void k_fc_pairs(const FcJob* __restrict__ jobs, double* __restrict__ work) {
    const FcJob J = jobs[blockIdx.x];
    const int N = J.N, T = J.T, tid = threadIdx.x;
    double* mean = work + (size_t)blockIdx.x * 2 * N;
    double* sd = mean + N;
    for (int c = tid; c < N; c += blockDim.x) {
        double m = 0.0;
        for (int t = 0; t < T; ++t) m += (double)J.x[(size_t)t * J.ldx + c];
        m /= (double)T;
        double s = 0.0;
        for (int t = 0; t < T; ++t) { const double d = (double)J.x[(size_t)t * J.ldx + c] - m; s += d * d; }
        s = sqrt(s / (double)T);
        mean[c] = m;
        sd[c] = s < 1e-12 ? 1.0 : s;
    }
    __syncthreads();
    const int64_t npairs = (int64_t)N * (N + 1) / 2;
    for (int64_t p = tid; p < npairs; p += blockDim.x) {
        int i = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
        while ((int64_t)i * (i + 1) / 2 > p) --i;
        while ((int64_t)(i + 1) * (i + 2) / 2 <= p) ++i;
        const int j = (int)(p - (int64_t)i * (i + 1) / 2);
        double cov = 0.0;
        for (int t = 0; t < T; ++t)
            cov += ((double)J.x[(size_t)t * J.ldx + i] - mean[i]) * ((double)J.x[(size_t)t * J.ldx + j] - mean[j]);
        put(J, i, j, pearson(cov / (double)T, sd[i], sd[j]));
    }
}

// FCD of one series: W window vectors v[w][P] (f32 upper triangles) ->
// out[W][W]; upper_triangle_correlation (fc.rs:143-178): FP64 means, centred
// cross and self sums, 0 when either variance is 0, clamp, diagonal 1.
__global__ __launch_bounds__(256) void k_fcd_corr(const float* __restrict__ vecs, float* __restrict__ out,
                                                  double* __restrict__ work, int W, int P) {
    const float* v = vecs + (size_t)blockIdx.x * W * P;
    float* o = out + (size_t)blockIdx.x * W * W;
    double* mean = work + (size_t)blockIdx.x * 2 * W;
    double* ss = mean + W;
    const int tid = threadIdx.x;
    for (int w = tid; w < W; w += blockDim.x) {
        double s = 0.0;
        for (int p = 0; p < P; ++p) s += (double)v[(size_t)w * P + p];
        const double m = s / (double)P;
        double q = 0.0;
        for (int p = 0; p < P; ++p) { const double d = (double)v[(size_t)w * P + p] - m; q += d * d; }
        mean[w] = m;
        ss[w] = q;
    }
    __syncthreads();
    const int64_t npairs = (int64_t)W * (W - 1) / 2;
    for (int64_t k = tid; k < npairs; k += blockDim.x) {
        // k -> (a, b), a < b: k = b (b - 1) / 2 + a
        int b = (int)((1.0 + sqrt(1.0 + 8.0 * (double)k)) * 0.5);
        while ((int64_t)b * (b - 1) / 2 > k) --b;
        while ((int64_t)(b + 1) * b / 2 <= k) ++b;
        const int a = (int)(k - (int64_t)b * (b - 1) / 2);
        double r = 0.0;
        if (ss[a] > 0.0 && ss[b] > 0.0) {
            double cov = 0.0;
            for (int p = 0; p < P; ++p)
                cov += ((double)v[(size_t)a * P + p] - mean[a]) * ((double)v[(size_t)b * P + p] - mean[b]);
            r = cov / (sqrt(ss[a]) * sqrt(ss[b]));
            r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
        }
        o[(size_t)a * W + b] = (float)r;
        o[(size_t)b * W + a] = (float)r;
    }
    for (int w = tid; w < W; w += blockDim.x) o[(size_t)w * W + w] = 1.0f;
}

// Owns the device buffers of one call; frees them on every exit path.
struct DevBufs {
    std::vector<void*> p;
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

hipError_t run_fc_jobs(const std::vector<FcJob>& jobs, int N, int T, DevBufs& B) {
    FcJob* dj = nullptr;
    hipError_t e = B.alloc(&dj, jobs.size());
    if (e == hipSuccess) e = hipMemcpy(dj, jobs.data(), sizeof(FcJob) * jobs.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    if (N <= kTile) {
        hipLaunchKernelGGL(k_fc_tiles, dim3((unsigned)jobs.size()), dim3(256), 0, nullptr, dj);
    } else {
        double* work = nullptr;
        e = B.alloc(&work, jobs.size() * 2 * (size_t)N);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_fc_pairs, dim3((unsigned)jobs.size()), dim3(256), 0, nullptr, dj, work);
    }
    (void)T;
    return hipGetLastError();
}

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

}  // namespace

extern "C" {

int hb_fc_batch(const float* series, int32_t n_series, int32_t n_steps, int32_t n_nodes, float* fc, int32_t device) {
    const int rc = check_series(series, n_series, n_steps, n_nodes, fc);
    if (rc) return rc;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const size_t per = (size_t)n_steps * n_nodes, nn = (size_t)n_nodes * n_nodes;
    feat::DevBufs B;
    float *dx = nullptr, *dout = nullptr;
    e = B.alloc(&dx, per * n_series);
    if (e == hipSuccess) e = B.alloc(&dout, nn * n_series);
    if (e == hipSuccess) e = hipMemcpy(dx, series, sizeof(float) * per * n_series, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "FC upload");
    std::vector<feat::FcJob> jobs((size_t)n_series);
    for (int32_t b = 0; b < n_series; ++b)
        jobs[b] = feat::FcJob{dx + per * b, dout + nn * b, n_steps, n_nodes, n_nodes, 0};
    e = feat::run_fc_jobs(jobs, n_nodes, n_steps, B);
    if (e == hipSuccess) e = hipMemcpy(fc, dout, sizeof(float) * nn * n_series, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "FC kernels");
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
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const size_t per = (size_t)n_steps * n_nodes;
    const size_t P = (size_t)n_nodes * (n_nodes - 1) / 2;
    feat::DevBufs B;
    float *dx = nullptr, *dv = nullptr, *dout = nullptr;
    double* work = nullptr;
    e = B.alloc(&dx, per * n_series);
    if (e == hipSuccess) e = B.alloc(&dv, P * W * n_series);
    if (e == hipSuccess) e = B.alloc(&dout, (size_t)W * W * n_series);
    if (e == hipSuccess) e = B.alloc(&work, (size_t)2 * W * n_series);
    if (e == hipSuccess) e = hipMemcpy(dx, series, sizeof(float) * per * n_series, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "FCD upload");
    if (P > 0) {
        std::vector<feat::FcJob> jobs((size_t)n_series * W);
        for (int32_t b = 0; b < n_series; ++b)
            for (int32_t w = 0; w < W; ++w)
                jobs[(size_t)b * W + w] = feat::FcJob{dx + per * b + (size_t)w * stride * n_nodes,
                                                      dv + P * ((size_t)b * W + w), window_len, n_nodes, n_nodes, 1};
        e = feat::run_fc_jobs(jobs, n_nodes, window_len, B);
        if (e != hipSuccess) return hip_fail(e, "FCD window FC");
    }
    hipLaunchKernelGGL(feat::k_fcd_corr, dim3((unsigned)n_series), dim3(256), 0, nullptr, dv, dout, work, W, (int)P);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(fcd, dout, sizeof(float) * (size_t)W * W * n_series, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "FCD kernels");
    return HB_OK;
}

}  // extern "C"
