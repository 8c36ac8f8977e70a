// MFMA form: the MatrixVectorMul / InnerProd / polymul launches (single job
// and the mixed persistent launch).  The block decomposition, the operand
// layouts and the per-wave work (mfma_body) are in tmvp_mfma_body.h.
#include <stdlib.h>

#include <atomic>

#include "tmvp_mfma_body.h"

namespace tmvp {

template <int L, int MODE, int EPI, int SHAPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<L, MODE, SHAPE>))) void mfma_kernel(Job J)
{
    // Two slices per wave: term t is staged into slice t & 1 while the term
    // before it is read from the other one.
    __shared__ __attribute__((aligned(16))) uint8_t lds[kLdsBytes<L, MODE>];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    mfma_body<L, MODE, EPI, SHAPE>(J, (int64_t)blockIdx.x * 4 + wave, (int64_t)gridDim.x * 4,
                                   lds + wave * 2 * Slice<L, MODE>::BYTES);
}

// The tile shape of a launch whose job leaves it to the library: 16 unless
// the environment names another (TMVP_MFMA_SHAPE=32, read once).
static int library_shape()
{
    static std::atomic<int> shape{0};
    int v = shape.load(std::memory_order_relaxed);
    if (v == 0) {
        const char *env = getenv("TMVP_MFMA_SHAPE");
        v = (env && atoi(env) == 32) ? 32 : 16;
        shape.store(v, std::memory_order_relaxed);
    }
    return v;
}

// Persistent launch: enough workgroups to fill every CU at the occupancy the
// LDS slice allows, each wave looping over instances with a grid stride.
template <int L, int MODE, int EPI, int SHAPE>
static hipError_t go(const Job &J, hipStream_t stream)
{
    // Workgroups that fit on the device at once; threads that race here
    // compute the same value.
    static std::atomic<int> resident_cache{0};
    int resident = resident_cache.load(std::memory_order_relaxed);
    if (resident == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mfma_kernel<L, MODE, EPI, SHAPE>, 256, 0);
        if (e != hipSuccess) return e;
        if (per_cu < 1) per_cu = 1;
        // Tuning override: workgroups per CU, capped by what fits.
        if (const char *env = getenv("TMVP_WGS_PER_CU")) {
            const int want = atoi(env);
            if (want >= 1 && want < per_cu) per_cu = want;
        }
        resident = cus * per_cu;
        resident_cache.store(resident, std::memory_order_relaxed);
    }
    const int64_t needed = (J.batch + 3) / 4;
    const int64_t blocks = needed < resident ? needed : resident;
    hipLaunchKernelGGL((mfma_kernel<L, MODE, EPI, SHAPE>), dim3((unsigned)blocks), dim3(256), 0, stream, J);
    return hipGetLastError();
}

template <int MODE, int EPI, int SHAPE>
static hipError_t by_l(const Job &J, int l, hipStream_t stream)
{
    switch (l) {
    case 2: return go<2, MODE, EPI, SHAPE>(J, stream);
    case 3: return go<3, MODE, EPI, SHAPE>(J, stream);
    case 4: return go<4, MODE, EPI, SHAPE>(J, stream);
    }
    return hipErrorInvalidValue;
}

template <int SHAPE>
static hipError_t launch_shape(const Job &J, int mode, int l, hipStream_t stream)
{
    switch (mode) {
    case kPoly:
        if (J.epi == kEpiMod) return go<1, kPoly, kEpiMod, SHAPE>(J, stream);
        break;
    case kMatVec:
        if (J.epi == kEpiMod) return by_l<kMatVec, kEpiMod, SHAPE>(J, l, stream);
        if (J.epi == kEpiQ2P) return by_l<kMatVec, kEpiQ2P, SHAPE>(J, l, stream);
        break;
    case kMatVecT:
        if (J.epi == kEpiMod) return by_l<kMatVecT, kEpiMod, SHAPE>(J, l, stream);
        if (J.epi == kEpiQ2P) return by_l<kMatVecT, kEpiQ2P, SHAPE>(J, l, stream);
        break;
    case kInner:
        if (J.epi == kEpiMod) return by_l<kInner, kEpiMod, SHAPE>(J, l, stream);
        if (J.epi == kEpiEnc) return by_l<kInner, kEpiEnc, SHAPE>(J, l, stream);
        if (J.epi == kEpiDec) return by_l<kInner, kEpiDec, SHAPE>(J, l, stream);
        break;
    case kEnc:
        if (J.epi == kEpiQ2P) return by_l<kEnc, kEpiQ2P, SHAPE>(J, l, stream);
        break;
    }
    return hipErrorInvalidValue;
}

hipError_t launch_mfma(const Job &J, int mode, int l, hipStream_t stream)
{
    if (J.batch <= 0) return hipSuccess;
    const int shape = J.shape ? J.shape : library_shape();
    if (shape == 32) return launch_shape<32>(J, mode, l, stream);
    if (shape == 16) return launch_shape<16>(J, mode, l, stream);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------- multi-job launch

namespace {
struct KindInfo {
    int mode, l, epi;
};
constexpr int kCombo[7][2] = {{kMatVec, kEpiMod},  {kMatVec, kEpiQ2P}, {kMatVecT, kEpiMod}, {kMatVecT, kEpiQ2P},
                              {kInner, kEpiMod},   {kInner, kEpiEnc},  {kInner, kEpiDec}};
constexpr KindInfo kind_info(int k)
{
    return k == 21 ? KindInfo{kPoly, 1, kEpiMod} : KindInfo{kCombo[k / 3][0], 2 + k % 3, kCombo[k / 3][1]};
}
// The four-row shapes have the largest slices.
constexpr int kMultiLds = kLdsBytes<4, kMatVec>;
static_assert(kMultiLds >= kLdsBytes<3, kMatVec> && kMultiLds >= kLdsBytes<4, kMatVecT> &&
                  kMultiLds >= kLdsBytes<4, kInner> && kMultiLds >= kLdsBytes<1, kPoly>,
              "LDS bound");

template <int K, int SHAPE>
__device__ __forceinline__ void run_kind(const Job &J, int64_t first, int64_t stride, uint8_t *lds, int wave)
{
    constexpr KindInfo I = kind_info(K);
    mfma_body<I.l, I.mode, I.epi, SHAPE>(J, first, stride, lds + wave * 2 * Slice<I.l, I.mode>::BYTES);
}
} // namespace

// Workgroup b belongs to the segment whose range holds b; its four waves take
// instances first, first + stride, ... of that segment's job.  Three waves per
// SIMD: the register budget of the largest shape (about 150 VGPRs).  SHAPE is
// the MFMA tile of every segment of the launch.
template <int SHAPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SHAPE == 32 ? 2 : 3))) void mfma_multi_kernel(MultiJob M)
{
    __shared__ __attribute__((aligned(16))) uint8_t lds[kMultiLds];
    int k = 0;
    while (k + 1 < M.nseg && (int)blockIdx.x >= M.seg[k + 1].block_begin) k++;
    const Segment &G = M.seg[k];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t first = (int64_t)((int)blockIdx.x - G.block_begin) * 4 + wave;
    const int64_t stride = (int64_t)G.blocks * 4;
    switch (G.kind) {
#define TMVP_KIND(K)                                                                                                 \
    case K: run_kind<K, SHAPE>(G.J, first, stride, lds, wave); break;
        TMVP_KIND(0) TMVP_KIND(1) TMVP_KIND(2) TMVP_KIND(3) TMVP_KIND(4) TMVP_KIND(5) TMVP_KIND(6) TMVP_KIND(7)
        TMVP_KIND(8) TMVP_KIND(9) TMVP_KIND(10) TMVP_KIND(11) TMVP_KIND(12) TMVP_KIND(13) TMVP_KIND(14)
        TMVP_KIND(15) TMVP_KIND(16) TMVP_KIND(17) TMVP_KIND(18) TMVP_KIND(19) TMVP_KIND(20) TMVP_KIND(21)
#undef TMVP_KIND
    default: break;
    }
}

int mfma_kind(int mode, int l, int epi)
{
    for (int k = 0; k < kNumKinds; k++) {
        const KindInfo I = kind_info(k);
        if (I.mode == mode && I.epi == epi && (mode == kPoly || I.l == l)) return k;
    }
    return -1;
}

int mfma_kind_cost(int kind)
{
    // Per-instance time model: each secret term costs its rows' MFMAs plus
    // about two rows' worth of staging and band work.  It matches the measured
    // single-job times of the MatrixVectorMul and InnerProd kinds at L = 2..4
    // to within about 10 %.
    const KindInfo I = kind_info(kind);
    return terms_of(I.mode, I.l) * (rows_of(I.mode, I.l) + 2);
}

int mfma_default_shape() { return library_shape(); }

int mfma_multi_per_cu(int shape)
{
    static std::atomic<int> per_cu[2] = {0, 0}; // racing threads compute the same value
    const int k = shape == 32 ? 1 : 0;
    int n = per_cu[k].load(std::memory_order_relaxed);
    if (n == 0) {
        const hipError_t e = k ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mfma_multi_kernel<32>, 256, 0)
                               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mfma_multi_kernel<16>, 256, 0);
        if (e != hipSuccess || n < 1) n = 1;
        per_cu[k].store(n, std::memory_order_relaxed);
    }
    return n;
}

hipError_t launch_mfma_multi(const MultiJob &M, int grid, int shape, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    if (shape == 32)
        hipLaunchKernelGGL(mfma_multi_kernel<32>, dim3((unsigned)grid), dim3(256), 0, stream, M);
    else if (shape == 16)
        hipLaunchKernelGGL(mfma_multi_kernel<16>, dim3((unsigned)grid), dim3(256), 0, stream, M);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

} // namespace tmvp
