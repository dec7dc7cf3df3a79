// hb_noise.hip -- pre-generated Gaussian noise for the stochastic integrators.
//
// Counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed by the
// 64-bit seed, counter = (pair index, stream id). Each Philox block gives two
// 53-bit uniforms; Box-Muller in f64 turns them into two standard normals,
// stored as f32. Element k of a stream is normal #(k & 1) of pair k >> 1.
// The reference draws N(0,1) in f64 and casts to f32 (integrator.rs:178-191);
// the stream itself is a new, seeded definition (the reference is unseeded).
#include "hb_internal.h"

namespace hb {

__device__ __forceinline__ void philox4x32_10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                                              uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__global__ __launch_bounds__(256) void k_philox_normal(const NoiseJob* __restrict__ jobs) {
    const NoiseJob J = jobs[blockIdx.y];
    const int64_t npairs = (J.n + 1) >> 1;
    const uint32_t k0 = (uint32_t)J.seed, k1 = (uint32_t)(J.seed >> 32);
    const uint32_t s0 = (uint32_t)J.stream, s1 = (uint32_t)(J.stream >> 32);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += stride) {
        uint32_t c0 = (uint32_t)p, c1 = (uint32_t)((uint64_t)p >> 32), c2 = s0, c3 = s1;
        philox4x32_10(c0, c1, c2, c3, k0, k1);
        const uint64_t a = ((uint64_t)(c0 >> 5) << 26) | (uint64_t)(c1 >> 6);
        const uint64_t b = ((uint64_t)(c2 >> 5) << 26) | (uint64_t)(c3 >> 6);
        const double u1 = (double)(a + 1) * 0x1.0p-53;   // (0, 1]
        const double u2 = (double)b * 0x1.0p-53;         // [0, 1)
        const double r = sqrt(-2.0 * log(u1));
        // (cos, sin)(2 pi u2) with an exact argument reduction
        double sn, cs;
        sincospi(2.0 * u2, &sn, &cs);
        const float z0 = (float)(r * cs);
        const float z1 = (float)(r * sn);
        const int64_t k = p << 1;
        if (k + 1 < J.n) *reinterpret_cast<float2*>(J.out + k) = make_float2(z0, z1);   // k even: 8-byte aligned
        else J.out[k] = z0;
    }
}

__global__ void k_zero_i32(int32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_noise(const NoiseJob* jobs_dev, int n_jobs, int64_t max_n, hipStream_t s) {
    if (n_jobs <= 0 || max_n <= 0) return hipSuccess;
    const int64_t npairs = (max_n + 1) >> 1;
    int64_t bx = (npairs + 255) / 256;
    if (bx > 4096) bx = 4096;
    dim3 g((unsigned)bx, (unsigned)n_jobs), b(256);
    hipLaunchKernelGGL(k_philox_normal, g, b, 0, s, jobs_dev);
    return hipGetLastError();
}

hipError_t launch_zero_i32(int32_t* p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_i32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace hb
