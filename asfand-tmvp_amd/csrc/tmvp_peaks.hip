// Measured device ceilings for the roofline (libtmvp_peaks.so, used by the
// benchmark only): HBM streaming with the MatrixVectorMul traffic shape, and
// the int8 matrix-core rate of v_mfma_i32_16x16x64_i8.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));

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
// instance) of the streaming kernel over `batch` instances.
int tmvp_peak_stream(int64_t batch, int reps, float *us, float *gbs)
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
    int64_t need = (batch + 3) / 4;
    const unsigned grid = (unsigned)(need < cus ? need : cus); // one workgroup per CU streams fastest
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
// v_mfma_i32_16x16x64_i8 on every SIMD (4 waves per SIMD).
int tmvp_peak_mfma_i8(int reps, float *tops)
{
    if (reps <= 0 || !tops) return -1;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -2;
    const int grid = cus * 4, iters = 2048;
    int *out = nullptr;
    if (hipMalloc(&out, (size_t)grid * 256 * sizeof(int)) != hipSuccess) return -3;
    float ms = 0;
    const int rc =
        best_ms(reps, [&] { hipLaunchKernelGGL(mfma_kernel, dim3(grid), dim3(256), 0, nullptr, iters, out); }, &ms);
    (void)hipFree(out);
    if (rc) return rc;
    const double ops = (double)grid * 4 * iters * 8 * 16.0 * 16.0 * 64.0 * 2.0;
    *tops = (float)(ops / (ms * 1e-3) / 1e12);
    return 0;
}

} // extern "C"
