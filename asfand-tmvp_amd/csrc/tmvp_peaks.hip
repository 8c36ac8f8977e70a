// Measured device ceilings for the roofline (libtmvp_peaks.so, used by the
// benchmark only): HBM streaming with the MatrixVectorMul traffic shape, the
// int8 matrix-core rate of both gfx950 int8 MFMA shapes, and the Keccak-f[1600]
// permutation rate of the KEM's lane-pair sponge code on the VALU alone.  Each
// is measured at several occupancies and the best is reported with the shape.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tmvp_keccak.h"

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// Per instance: read 6144 B (A 4608 + s 1536) and write 1536 B, one wave per
// instance at a time, next instance's loads issued before the current one is
// written (the MatrixVectorMul kernel's traffic without its arithmetic).
__global__ __launch_bounds__(256) void stream_kernel(const u32x4 *in, u32x2 *out, int64_t batch)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * 4;
    int64_t id = (int64_t)blockIdx.x * 4 + wave;
    if (id >= batch) return;
    u32x4 v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = in[id * 384 + k * 64 + lane];
    for (;;) {
        const u32x2 r0 = {v[0].x ^ v[3].x, v[1].y ^ v[4].y};
        const u32x2 r1 = {v[2].z ^ v[5].z, v[0].w ^ v[3].w};
        const u32x2 r2 = {v[1].x ^ v[4].z, v[2].y ^ v[5].w};
        const int64_t nid = id + stride;
        if (nid < batch)
#pragma unroll
            for (int k = 0; k < 6; k++) v[k] = in[nid * 384 + k * 64 + lane];
        out[id * 192 + lane] = r0;
        out[id * 192 + 64 + lane] = r1;
        out[id * 192 + 128 + lane] = r2;
        if (nid >= batch) break;
        id = nid;
    }
}

// Eight independent accumulator chains of int8 MFMAs per wave, register
// operands only; the sums are stored so nothing is optimised away.
__global__ __launch_bounds__(256) void mfma_kernel(int iters, int *out)
{
    const int lane = threadIdx.x & 63;
    v4i a = {lane, lane + 1, lane + 2, lane + 3}, b = {3 - lane, lane, 7, -lane};
    v4i c[8];
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = v4i{k, 0, 0, 0};
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 8; k++) c[k] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c[k], 0, 0, 0);
    }
    int s = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) s += c[k][0] + c[k][1] + c[k][2] + c[k][3];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// The same with v_mfma_i32_32x32x32_i8 (four chains of 16 accumulators).
__global__ __launch_bounds__(256) void mfma32_kernel(int iters, int *out)
{
    const int lane = threadIdx.x & 63;
    v4i a = {lane, lane + 1, lane + 2, lane + 3}, b = {3 - lane, lane, 7, -lane};
    v16i c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
#pragma unroll
        for (int e = 0; e < 16; e++) c[k][e] = k + e;
    }
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 4; k++) c[k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c[k], 0, 0, 0);
    }
    int s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
#pragma unroll
        for (int e = 0; e < 16; e++) s += c[k][e];
    }
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Keccak-f[1600] on lane pairs with the state in registers: `iters`
// permutations per pair, the KEM sponges' round code (tmvp_keccak.h) with
// no memory traffic.
template <int U> __global__ __launch_bounds__(256) void keccak_kernel(int iters, uint32_t *out)
{
    const int lane = threadIdx.x & 63, p = lane & 1;
    uint32_t A[25];
#pragma unroll
    for (int k = 0; k < 25; k++) A[k] = (uint32_t)(lane * 25 + k) ^ blockIdx.x;
    for (int i = 0; i < iters; i++) tmvp::keccak::f1600<U>(A, p);
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < 25; k++) x ^= A[k];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = x;
}

template <class F> int best_ms(int reps, F launch, float *best)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -2;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return -2;
    }
    launch();
    float b = 1e30f;
    int rc = 0;
    for (int r = 0; r < 3 && !rc; r++) {
        if (hipEventRecord(e0, nullptr) != hipSuccess) rc = -2;
        for (int i = 0; i < reps && !rc; i++) launch();
        if (!rc && (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)) rc = -2;
        float ms = 0;
        if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = -2;
        if (!rc && ms / reps < b) b = ms / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc && hipGetLastError() != hipSuccess) rc = -2;
    *best = b;
    return rc;
}

} // namespace

extern "C" {

// Best per-launch time (us) and rate (GB/s of the 7680 algorithmic bytes per
// instance) of the streaming kernel over `batch` instances, at `wgs_per_cu`
// workgroups of four waves per CU (0: one, the shape it was first written for).
int tmvp_peak_stream_at(int64_t batch, int reps, int wgs_per_cu, float *us, float *gbs);
int tmvp_peak_stream(int64_t batch, int reps, float *us, float *gbs)
{
    return tmvp_peak_stream_at(batch, reps, 1, us, gbs);
}

int tmvp_peak_stream_at(int64_t batch, int reps, int wgs_per_cu, float *us, float *gbs)
{
    if (batch <= 0 || reps <= 0 || !us || !gbs) return -1;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -2;
    u32x4 *in = nullptr;
    u32x2 *out = nullptr;
    if (hipMalloc(&in, (size_t)batch * 6144) != hipSuccess) return -3;
    if (hipMalloc(&out, (size_t)batch * 1536) != hipSuccess) {
        (void)hipFree(in);
        return -3;
    }
    int rc = hipMemset(in, 1, (size_t)batch * 6144) == hipSuccess ? 0 : -2;
    if (wgs_per_cu < 1) wgs_per_cu = 1;
    const int64_t need = (batch + 3) / 4, resident = (int64_t)cus * wgs_per_cu;
    const unsigned grid = (unsigned)(need < resident ? need : resident);
    float ms = 0;
    if (!rc)
        rc = best_ms(reps, [&] { hipLaunchKernelGGL(stream_kernel, dim3(grid), dim3(256), 0, nullptr, in, out, batch); },
                     &ms);
    (void)hipFree(in);
    (void)hipFree(out);
    if (rc) return rc;
    *us = ms * 1e3f;
    *gbs = (float)((double)batch * 7680.0 / (ms * 1e-3) / 1e9);
    return 0;
}

// Best achieved int8 rate (TOPS, 2 ops per multiply-add) of back-to-back
// int8 MFMAs on every SIMD: shape 0 = v_mfma_i32_16x16x64_i8, 1 =
// v_mfma_i32_32x32x32_i8, at `waves_per_simd` waves on each SIMD.
int tmvp_peak_mfma_i8_shape(int reps, int shape, int waves_per_simd, float *tops)
{
    if (reps <= 0 || !tops || shape < 0 || shape > 1 || waves_per_simd < 1 || waves_per_simd > 8) return -1;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -2;
    const int grid = cus * waves_per_simd, iters = 2048;
    int *out = nullptr;
    if (hipMalloc(&out, (size_t)grid * 256 * sizeof(int)) != hipSuccess) return -3;
    float ms = 0;
    const int rc = best_ms(reps,
                           [&] {
                               if (shape == 0)
                                   hipLaunchKernelGGL(mfma_kernel, dim3(grid), dim3(256), 0, nullptr, iters, out);
                               else
                                   hipLaunchKernelGGL(mfma32_kernel, dim3(grid), dim3(256), 0, nullptr, iters, out);
                           },
                           &ms);
    (void)hipFree(out);
    if (rc) return rc;
    // MACs per wave and iteration: 8 x 16*16*64 or 4 x 32*32*32 (both 131072)
    const double ops = (double)grid * 4 * iters * 131072.0 * 2.0;
    *tops = (float)(ops / (ms * 1e-3) / 1e12);
    return 0;
}

// The 16x16x64 shape at four waves per SIMD (the first form of this probe).
int tmvp_peak_mfma_i8(int reps, float *tops) { return tmvp_peak_mfma_i8_shape(reps, 0, 4, tops); }

// Keccak-f[1600] permutations per second on lane pairs (32 per wave) at
// `waves_per_simd` waves on each SIMD, with the round loop unrolled `unroll`
// (1 or 2) times; the VALU-only ceiling of the KEM's sponges.
int tmvp_peak_keccak(int reps, int waves_per_simd, int unroll, double *perms_per_s)
{
    if (reps <= 0 || !perms_per_s || waves_per_simd < 1 || waves_per_simd > 8 || (unroll != 1 && unroll != 2))
        return -1;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -2;
    const int grid = cus * waves_per_simd, iters = 64;
    uint32_t *out = nullptr;
    if (hipMalloc(&out, (size_t)grid * 256 * sizeof(uint32_t)) != hipSuccess) return -3;
    float ms = 0;
    const int rc = best_ms(reps,
                           [&] {
                               if (unroll == 1)
                                   hipLaunchKernelGGL(keccak_kernel<1>, dim3(grid), dim3(256), 0, nullptr, iters, out);
                               else
                                   hipLaunchKernelGGL(keccak_kernel<2>, dim3(grid), dim3(256), 0, nullptr, iters, out);
                           },
                           &ms);
    (void)hipFree(out);
    if (rc) return rc;
    *perms_per_s = (double)grid * 4 * 32 * iters / (ms * 1e-3);
    return 0;
}

} // extern "C"
