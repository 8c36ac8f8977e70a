// Plain VALU form: one wavefront per Toeplitz-matrix-vector product.
//
// A product c = a * s in Z_{2^16}[x]/(x^256+1) is the matrix-vector product
// c = T(s) a with the negacyclic Toeplitz matrix T(s)[m][k] = ts(m - k),
// ts(d) = s[d] for d >= 0 and -s[d + 256] for d < 0.  The reference builds
// T explicitly in global memory (convertnegacyclictest, S/src/poly.cu:1116)
// or walks its triangle per thread (S/src/poly.cu:281-284).  Here T is never
// materialised: the wave stages the 512-entry band X[n] = ts(n - 256) in LDS,
// lane l owns outputs m = 4l..4l+3, and for each k the four entries
// X[m - k + 256] form a register window that slides by one element per k.
// a[k] is read from LDS at a wave-uniform address (a broadcast).  The
// arithmetic is packed 16-bit multiply-add (two outputs per instruction),
// whose wrap-around is exactly the reference's uint16 semantics.
//
// Launch: 256 threads = 4 waves; every wave owns one (instance, row) unit and
// its own LDS slice, so the waves only meet at the staging barrier.
#include "tmvp_internal.h"

namespace tmvp {

typedef unsigned short us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ us2 as_us2(uint32_t v) { return __builtin_bit_cast(us2, v); }
__device__ __forceinline__ uint32_t as_u32(us2 v) { return __builtin_bit_cast(uint32_t, v); }

// Pair (X[2i-1], X[2i]) from the dwords (X[2i-2],X[2i-1]) and (X[2i],X[2i+1]).
__device__ __forceinline__ uint32_t mid_pair(uint32_t hi, uint32_t lo) { return __builtin_amdgcn_alignbit(hi, lo, 16); }

template <int L, int MODE>
__global__ __launch_bounds__(256) void valu_kernel(Job J)
{
    constexpr int NR = rows_of(MODE, L);
    constexpr int NT = terms_of(MODE, L);
    // Per wave: NT bands X (512 x u16) and NT public polynomials (256 x u16).
    __shared__ __attribute__((aligned(16))) uint32_t lds[4][NT * (256 + 128)];

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t unit = (int64_t)blockIdx.x * 4 + wave;
    const bool valid = unit < J.batch * NR;
    const int64_t inst = valid ? unit / NR : 0;
    const int row = (int)(unit - inst * NR);

    uint32_t *band = lds[wave];            // NT x 256 dwords (X pairs)
    uint32_t *pub = lds[wave] + NT * 256;  // NT x 128 dwords (a pairs)

    if (valid) {
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const uint16_t *ap, *sp;
            if (MODE == kPoly) {
                ap = J.a + inst * kN;
                sp = J.s + inst * kN;
            } else if (MODE == kInner) {
                ap = J.a + (inst * L + j) * kN;
                sp = J.s + (inst * L + j) * kN;
            } else {
                const int pq = (MODE == kMatVec) ? row * L + j : j * L + row;
                ap = J.a + (inst * L * L + pq) * kN;
                sp = J.s + (inst * L + j) * kN;
            }
            const uint2 av = *(const uint2 *)(ap + 4 * lane);
            const uint2 sv = *(const uint2 *)(sp + 4 * lane);
            *(uint2 *)(pub + j * 128 + 2 * lane) = av;
            // X[n] = -s[n] for n < 256, X[n] = s[n - 256] for n >= 256
            *(uint2 *)(band + j * 256 + 128 + 2 * lane) = sv;
            const us2 z = {0, 0};
            uint2 nv;
            nv.x = as_u32(z - as_us2(sv.x));
            nv.y = as_u32(z - as_us2(sv.y));
            *(uint2 *)(band + j * 256 + 2 * lane) = nv;
        }
    }
    __syncthreads();
    if (!valid) return;

    us2 c01 = {0, 0}, c23 = {0, 0};
#pragma unroll 1
    for (int j = 0; j < NT; j++) {
        const uint32_t *X32 = band + j * 256;
        const uint32_t *A32 = pub + j * 128;
        // base = 4*lane + 256 - 4*kappa; window dwords d0..d3 = X[base-4 .. base+3]
        int bd = 2 * lane + 128; // base / 2
        uint32_t d2 = X32[bd], d3 = X32[bd + 1];
#pragma unroll 4
        for (int kappa = 0; kappa < 64; kappa++) {
            // bd - 2 = 2*lane + 126 - 2*kappa >= 0 for kappa <= 63
            const uint2 lo = *(const uint2 *)(X32 + bd - 2);
            const uint32_t d0 = lo.x, d1 = lo.y;
            const uint2 a4 = *(const uint2 *)(A32 + 2 * kappa); // a[4k..4k+3], uniform
            const us2 aa = as_us2(a4.x), ab = as_us2(a4.y);
            const us2 a0 = aa.xx, a1 = aa.yy, a2 = ab.xx, a3 = ab.yy;
            const uint32_t m21 = mid_pair(d2, d1);
            c01 += a0 * as_us2(d2);
            c23 += a0 * as_us2(d3);
            c01 += a1 * as_us2(m21);
            c23 += a1 * as_us2(mid_pair(d3, d2));
            c01 += a2 * as_us2(d1);
            c23 += a2 * as_us2(d2);
            c01 += a3 * as_us2(mid_pair(d1, d0));
            c23 += a3 * as_us2(m21);
            d2 = d0;
            d3 = d1;
            bd -= 2;
        }
    }

    uint16_t *op = J.out + (inst * NR + row) * kN + 4 * lane;
    const uint16_t *auxp = J.aux ? J.aux + inst * kN + 4 * lane : nullptr;
    uint2 oldv = make_uint2(0u, 0u), auxv = make_uint2(0u, 0u);
    if (J.accumulate && J.epi == kEpiMod) oldv = *(const uint2 *)op;
    if (auxp) auxv = *(const uint2 *)auxp;
    const uint32_t r0 = apply_epilogue(J, c01.x, oldv.x & 0xffff, auxv.x & 0xffff);
    const uint32_t r1 = apply_epilogue(J, c01.y, oldv.x >> 16, auxv.x >> 16);
    const uint32_t r2 = apply_epilogue(J, c23.x, oldv.y & 0xffff, auxv.y & 0xffff);
    const uint32_t r3 = apply_epilogue(J, c23.y, oldv.y >> 16, auxv.y >> 16);
    *(uint2 *)op = make_uint2(r0 | (r1 << 16), r2 | (r3 << 16));
}

template <int L, int MODE>
static hipError_t go(const Job &J, hipStream_t stream)
{
    const int64_t units = J.batch * rows_of(MODE, L);
    const int64_t blocks = (units + 3) / 4;
    hipLaunchKernelGGL((valu_kernel<L, MODE>), dim3((unsigned)blocks), dim3(256), 0, stream, J);
    return hipGetLastError();
}

hipError_t launch_valu(const Job &J, int mode, int l, hipStream_t stream)
{
    if (J.batch <= 0) return hipSuccess;
    switch (mode) {
    case kPoly:
        return go<1, kPoly>(J, stream);
    case kMatVec:
        switch (l) {
        case 2: return go<2, kMatVec>(J, stream);
        case 3: return go<3, kMatVec>(J, stream);
        case 4: return go<4, kMatVec>(J, stream);
        }
        break;
    case kMatVecT:
        switch (l) {
        case 2: return go<2, kMatVecT>(J, stream);
        case 3: return go<3, kMatVecT>(J, stream);
        case 4: return go<4, kMatVecT>(J, stream);
        }
        break;
    case kInner:
        switch (l) {
        case 2: return go<2, kInner>(J, stream);
        case 3: return go<3, kInner>(J, stream);
        case 4: return go<4, kInner>(J, stream);
        }
        break;
    }
    return hipErrorInvalidValue;
}

} // namespace tmvp
