// MFMA form: exact negacyclic products on gfx950 int8 matrix cores.
//
// Block decomposition (n = 256 = 8 blocks of 32).  For c = T(s) a with
// T(s)[m][k] = ts(m - k) (ts(d) = s[d] for d >= 0, -s[d + 256] for d < 0),
// write m = 32R + i and k = 32C + jj.  The 32x32 block (R, C) of T(s) is
// eps * B_D with D = (R - C) mod 8, B_D[i][jj] = ts(32D + i - jj) and
// eps = -1 exactly when R < D (then C = R - D + 8 and the entry wraps once).
// So, with the extended block list E[e] = a_{e-8} for e >= 8 and E[e] = -a_e
// for e < 8,
//
//     c[32R + i] = sum_D sum_jj B_D[i][jj] * E[R - D + 8][jj],
//
// one 32 x 256 x 8 matrix product per (row, limb) with no redundant MACs and
// no 256 x 256 matrix anywhere.  The reference instead materialises T in HBM
// (S/src/poly.cu:1116-1129) and multiplies three padded 128 x 128 fp16 tiles
// (S/src/poly.cu:1155-1205), which is inexact for q = 2^13.
//
// Exactness.  The secret is centred mod 2^w into int8 (CBD values are at most
// 5 in magnitude).  The public coefficient (mod 2^w, w <= 14) is split into
// limbs a = lo + 128 hi with lo in [0,127], hi in [0,127]; the negated copies
// stay inside int8.  Each limb product accumulates in int32
// (|sum| <= 4 * 256 * 127 * 127 < 2^31); out = lo + 128 hi is exact mod 2^32.
//
// Mapping onto v_mfma_i32_16x16x64_i8 (M=16, N=16, K=64):
//   M  = 16 rows i of one half (h) of a 32-row block row,
//   K  = 64 = two consecutive D tiles x 32 jj  (four K-chunks p cover D = 0..7),
//   N  = 16 = 8 block columns R x 2 limbs.
// The A operand (Toeplitz tiles of s) is built once per secret polynomial
// into 32 VGPRs and reused by every row and both limbs; the B operand is one
// aligned 16-byte LDS read per MFMA pair (it feeds both halves h).  A lane of
// the A operand needs 16 consecutive band bytes at an unaligned offset whose
// misalignment is fixed per lane, so it is assembled from five dwords with
// v_alignbyte.  K indices are laid out in reversed jj order inside each
// 16-byte group (the E blocks are stored reversed) so that this window is
// ascending in memory.
//
// Launch: 256 threads = 4 waves, one instance per wave (all rows and terms),
// every wave with its own LDS slice.
#include "tmvp_internal.h"

namespace tmvp {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kEBlock = 32;               // bytes per E block (32 coefficients, one limb)
constexpr int kEPoly = 16 * kEBlock;      // e = 0..15 (e = 0 unused)
constexpr int kBand = 320;                // bytes per secret band (288 used)

__device__ __forceinline__ uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

// Bytewise negation of four bytes that are each <= 0x80.
__device__ __forceinline__ uint32_t neg_bytes_small(uint32_t v) { return (0x80808080u - v) ^ 0x80808080u; }

template <int L, int MODE>
__global__ __launch_bounds__(256) void mfma_kernel(Job J)
{
    constexpr int NR = rows_of(MODE, L);
    constexpr int NT = terms_of(MODE, L);
    constexpr int E_BYTES = NR * NT * 2 * kEPoly;
    constexpr int WAVE_BYTES = E_BYTES + NT * kBand;
    __shared__ __attribute__((aligned(16))) uint8_t lds[4 * WAVE_BYTES];

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t inst = (int64_t)blockIdx.x * 4 + wave;
    const bool valid = inst < J.batch;
    uint8_t *E = lds + wave * WAVE_BYTES;
    uint8_t *X = E + E_BYTES;
    const uint32_t wmask = (1u << J.w) - 1u;

    if (valid) {
        // ---- public polynomials -> reversed limb blocks E (and negated copies)
        const uint16_t *abase = (MODE == kPoly) ? J.a + inst * kN
                              : (MODE == kInner) ? J.a + inst * L * kN
                                                 : J.a + inst * L * L * kN;
        for (int c = lane; c < NR * NT * 32; c += 64) {
            const int pq = c >> 5;          // polynomial index in memory order
            const int k0 = (c & 31) * 8;    // first of 8 coefficients
            int r = 0, j = pq;
            if (MODE == kMatVec) { r = pq / L; j = pq % L; }
            if (MODE == kMatVecT) { r = pq % L; j = pq / L; }
            const uint4 v = *(const uint4 *)(abase + pq * kN + k0);
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
            uint32_t lo[2] = {0, 0}, hi[2] = {0, 0};
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const uint32_t x = ((w4[u >> 1] >> (16 * (u & 1))) & 0xffffu) & wmask;
                lo[u >> 2] |= (x & 127u) << (8 * (u & 3));
                hi[u >> 2] |= (x >> 7) << (8 * (u & 3));
            }
            const int blk = k0 >> 5, pos = 24 - (k0 & 31);
            uint8_t *slot = E + (r * NT + j) * 2 * kEPoly;
            const uint2 plo = make_uint2(bswap32(lo[1]), bswap32(lo[0]));
            const uint2 phi = make_uint2(bswap32(hi[1]), bswap32(hi[0]));
            *(uint2 *)(slot + (blk + 8) * kEBlock + pos) = plo;
            *(uint2 *)(slot + kEPoly + (blk + 8) * kEBlock + pos) = phi;
            if (blk > 0) {
                *(uint2 *)(slot + blk * kEBlock + pos) = make_uint2(neg_bytes_small(plo.x), neg_bytes_small(plo.y));
                *(uint2 *)(slot + kEPoly + blk * kEBlock + pos) =
                    make_uint2(neg_bytes_small(phi.x), neg_bytes_small(phi.y));
            }
        }
        // ---- secret polynomials -> int8 bands X[x] = ts(x - 32), x in [0, 288)
        const uint16_t *sbase = (MODE == kPoly) ? J.s + inst * kN : J.s + inst * L * kN;
        const uint32_t half = 1u << (J.w - 1);
        bool bad = false;
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const uint2 v = *(const uint2 *)(sbase + j * kN + 4 * lane);
            const uint32_t w2[2] = {v.x, v.y};
            uint32_t pos = 0, neg = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t x = ((w2[u >> 1] >> (16 * (u & 1))) & 0xffffu) & wmask;
                const int sv = (x >= half) ? (int)x - (int)(wmask + 1u) : (int)x;
                bad |= (sv < -127) | (sv > 127);
                pos |= (uint32_t)(sv & 0xff) << (8 * u);
                neg |= (uint32_t)((-sv) & 0xff) << (8 * u);
            }
            uint8_t *band = X + j * kBand;
            *(uint32_t *)(band + 32 + 4 * lane) = pos;
            if (lane >= 56) *(uint32_t *)(band + 4 * lane - 224) = neg;
            if (lane < 8) *(uint32_t *)(band + 288 + 4 * lane) = 0u; // pad read by the last window
        }
        if (bad) *J.status = 1;
    }
    __syncthreads();
    if (!valid) return;

    const int il = lane & 15, g = lane >> 4;
    const int n = lane & 15, R = n & 7, limb = n >> 3;
    const int sh = (il + 1) & 3;

    v4i acc[NR][2];
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r][0] = acc[r][1] = v4i{0, 0, 0, 0};

#pragma unroll 1
    for (int j = 0; j < NT; j++) {
        // A operand: Toeplitz tiles of s_j for the four K-chunks and both halves.
        const uint32_t *X32 = (const uint32_t *)(X + j * kBand);
        v4i afr[4][2];
#pragma unroll
        for (int p = 0; p < 4; p++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int D = 2 * p + (g >> 1);
                const int x0 = 32 * D + 16 * h - 16 * (g & 1) + il + 17;
                const int q = x0 >> 2;
                const uint32_t w0 = X32[q], w1 = X32[q + 1], w2 = X32[q + 2], w3 = X32[q + 3], w4 = X32[q + 4];
                afr[p][h] = v4i{(int)__builtin_amdgcn_alignbyte(w1, w0, sh), (int)__builtin_amdgcn_alignbyte(w2, w1, sh),
                                (int)__builtin_amdgcn_alignbyte(w3, w2, sh), (int)__builtin_amdgcn_alignbyte(w4, w3, sh)};
            }
        }
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const uint8_t *slot = E + ((r * NT + j) * 2 + limb) * kEPoly + 16 * (1 - (g & 1));
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const int e = R - (2 * p + (g >> 1)) + 8;
                const v4i bfr = *(const v4i *)(slot + e * kEBlock);
                acc[r][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(afr[p][0], bfr, acc[r][0], 0, 0, 0);
                acc[r][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(afr[p][1], bfr, acc[r][1], 0, 0, 0);
            }
        }
    }

    // ---- epilogue: lanes n < 8 own block column R; rows i = 16h + 4g + v.
#pragma unroll
    for (int r = 0; r < NR; r++) {
        uint16_t *op = J.out + (inst * NR + r) * kN;
        const uint16_t *auxp = J.aux ? J.aux + inst * kN : nullptr;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t val[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int hiv = __builtin_amdgcn_mov_dpp(acc[r][h][v], 0x128, 0xf, 0xf, false); // row_ror:8
                val[v] = (uint32_t)acc[r][h][v] + ((uint32_t)hiv << 7);
            }
            if (limb == 0) {
                const int m0 = 32 * R + 16 * h + 4 * g;
                uint2 oldv = make_uint2(0u, 0u), auxv = make_uint2(0u, 0u);
                if (J.accumulate && J.epi == kEpiMod) oldv = *(const uint2 *)(op + m0);
                if (auxp) auxv = *(const uint2 *)(auxp + m0);
                const uint32_t r0 = apply_epilogue(J, val[0], oldv.x & 0xffff, auxv.x & 0xffff);
                const uint32_t r1 = apply_epilogue(J, val[1], oldv.x >> 16, auxv.x >> 16);
                const uint32_t r2 = apply_epilogue(J, val[2], oldv.y & 0xffff, auxv.y & 0xffff);
                const uint32_t r3 = apply_epilogue(J, val[3], oldv.y >> 16, auxv.y >> 16);
                *(uint2 *)(op + m0) = make_uint2(r0 | (r1 << 16), r2 | (r3 << 16));
            }
        }
    }
}

template <int L, int MODE>
static hipError_t go(const Job &J, hipStream_t stream)
{
    const int64_t blocks = (J.batch + 3) / 4;
    hipLaunchKernelGGL((mfma_kernel<L, MODE>), dim3((unsigned)blocks), dim3(256), 0, stream, J);
    return hipGetLastError();
}

hipError_t launch_mfma(const Job &J, int mode, int l, hipStream_t stream)
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
