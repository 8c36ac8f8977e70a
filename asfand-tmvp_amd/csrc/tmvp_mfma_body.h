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
// Exactness.  The secret is used centred mod 2^w as an int8 (CBD values are
// at most 5 in magnitude; for 8 <= w that int8 is simply the low byte).  The
// public coefficient (mod 2^w, 8 <= w <= 14) is split into limbs a = lo + 128 hi,
// lo and hi in [0, 127]; the negated copies stay inside int8.  Each limb
// product accumulates exactly in int32 (|sum| <= 4*256*127*127 < 2^31) and
// out = lo + 128 hi is exact mod 2^32.
//
// Mapping onto v_mfma_i32_16x16x64_i8 (M=16, N=16, K=64):
//   M = 16 rows i of one half (h) of a 32-row block row,
//   K = 64 = two consecutive D tiles x 32 jj (four K-chunks p cover D = 0..7),
//   N = 16 = 8 block columns R x 2 limbs.
// The A operand (Toeplitz tiles of s) is loaded once per secret polynomial
// into 32 VGPRs and reused by every row and both limbs; the B operand is one
// aligned 16-byte LDS read per MFMA pair (it feeds both halves h).
//
// An A-operand lane needs 16 consecutive band bytes at an offset whose
// misalignment (i + 1) mod 4 is fixed per lane.  The band is therefore stored
// four times, shifted by 0..3 bytes, and every lane reads a dword-aligned
// window of its own copy (unaligned LDS loads are exact on gfx950 but many
// times slower).  K indices run in reversed jj order inside each 16-byte group
// (the E blocks are stored byte-reversed) so that these windows ascend in
// memory.  The LDS serves a two-dword read one dword and 32 lanes at a time
// from 32 banks: in each half-wave, each copy's 8 lanes read 8 consecutive
// dwords, and the copies sit at 288 c (+ 4 for c > 0) bytes so that the four
// 8-dword windows fall in distinct banks (at a plain 288-byte stride every
// pass has a two-way conflict).
//
// Launch: persistent, 256 threads = 4 waves per workgroup.  A wave owns one
// instance at a time in a private LDS slice, stages it one secret term at a
// time, walks the batch with a grid stride, and loads each term of the next
// instance into registers as soon as the current one has been staged.
//
// Second tile shape, v_mfma_i32_32x32x32_i8 (M=32, N=32, K=32), which issues
// about 1.7x the int8 MACs per cycle of the 16x16x64 form on gfx950:
//   M = the 32 rows i of a block row,
//   K = the 32 jj of ONE D tile (eight MFMAs per row pair and term),
//   N = 32 = 8 block columns R x 2 limbs x 2 public rows (a row pair rp;
//       an odd row count leaves the second row of the last pair unused).
// A-operand lane (i = lane & 31, kh = lane >> 5) holds the 16 band bytes
// X[32D + i + 1 + 16kh ...] (K index k = 16kh + t is jj = 31 - k, which is
// the byte order of the reversed E blocks), B-operand lane (n = lane & 31,
// kh) the 16 bytes [16kh, 16kh + 16) of E[R - D + 8] of its limb and row.
// The staging, the LDS layouts, the prefetch and the pipeline are shared.
//
// This header holds the device side (mfma_body and its helpers); the launches
// (single job and mixed persistent) are in tmvp_mfma.hip.
#pragma once

#include <type_traits>

#include "tmvp_internal.h"

namespace tmvp {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Per-halfword logical right shift of a packed uint16 pair.
__device__ __forceinline__ uint32_t shr16x2(uint32_t a, int n)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) >> (unsigned short)n);
}
// Per-halfword add of two packed uint16 pairs (wraps mod 2^16).
__device__ __forceinline__ uint32_t add16x2(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b));
}

constexpr int kEBlock = 32;           // bytes per E block (32 coefficients, one limb)
constexpr int kEPoly = 16 * kEBlock;  // blocks e = 0..15 (e = 0 unused)
constexpr int kBand = 288;            // bytes of one band copy: X[x] = ts(x - 32), x < 288
// Byte offset of band copy c (see the bank note above); the four copies take kBandBytes.
__host__ __device__ constexpr int band_copy(int c) { return 288 * c + (c ? 4 : 0); }
constexpr int kBandBytes = 1168; // band_copy(3) + kBand, rounded up to 16
static_assert(band_copy(3) + kBand <= kBandBytes && kBandBytes % 16 == 0, "band copies");

// Bytewise negation of four bytes that are each <= 0x80.
__device__ __forceinline__ uint32_t neg_small(uint32_t v) { return (0x80808080u - v) ^ 0x80808080u; }
// Bytewise two's-complement negation of four arbitrary bytes.
__device__ __forceinline__ uint32_t neg_bytes(uint32_t v)
{
    const uint32_t y = ~v;
    return ((y & 0x7f7f7f7fu) + 0x01010101u) ^ (y & 0x80808080u);
}
__device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel) { return __builtin_amdgcn_perm(hi, lo, sel); }

// Per-halfword helpers on packed uint16 pairs (arithmetic mod 2^16).
__device__ __forceinline__ uint32_t shl16x2(uint32_t a, uint32_t n)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) << __builtin_bit_cast(u16x2, n));
}
__device__ __forceinline__ uint32_t sub16x2(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) - __builtin_bit_cast(u16x2, b));
}

// Epilogue on two outputs at once: y holds (exact sum + h) mod 2^16 for two
// rows (h, the rounding constant, enters through the accumulator), `old` the
// current outputs (accumulate) and `aux` the two aux words.
template <int EPI>
__device__ __forceinline__ uint32_t epi2(const Job &J, uint32_t y, uint32_t old, uint32_t aux)
{
    const uint32_t two = 0x00010001u;
    if (EPI == kEpiQ2P) return shr16x2(y, J.eq - J.ep) & (((1u << J.ep) - 1u) * two);
    if (EPI == kEpiQ2PAcc) return shr16x2(add16x2(y, old), J.eq - J.ep) & (((1u << J.ep) - 1u) * two);
    if (EPI == kEpiEnc) {
        const uint32_t m = shl16x2(aux & two, (uint32_t)(J.ep - 1) * two);
        return shr16x2(sub16x2(y, m), J.ep - J.et) & (((1u << J.et) - 1u) * two);
    }
    if (EPI == kEpiDec) {
        const uint32_t c = shl16x2(aux & (((1u << J.et) - 1u) * two), (uint32_t)(J.ep - J.et) * two);
        return shr16x2(sub16x2(y, c), J.ep - 1) & two;
    }
    return (J.accumulate ? add16x2(y, old) : y) & (((1u << J.w) - 1u) * two);
}

// Per-wave LDS slice: the E blocks of every row for ONE secret term j, and the
// 4 shifted copies of that term's band.  Terms are staged one after another,
// so the staging of term j+1 can overlap the matrix-core work of term j.
template <int L, int MODE> struct Slice {
    static constexpr int NR = rows_of(MODE, L);
    static constexpr int NT = terms_of(MODE, L);
    static constexpr int NPOLY = NR * NT;       // public polynomials per instance
    static constexpr int CJ = NR * 32;          // 8-coefficient chunks of one term
    static constexpr int ITJ = (CJ + 63) / 64;  // staging loads per lane per term
    // a lane past the last chunk stages chunk u0 - CJ, which must exist
    static_assert(CJ % 64 == 0 || 64 * ITJ <= 2 * CJ, "chunk remap");
    static constexpr int E_BYTES = NR * 2 * kEPoly;
    static constexpr int BYTES = E_BYTES + kBandBytes;
};

// DPP wave_shl:1 - lane i receives lane i + 1 (lane 63 receives 0).
__device__ __forceinline__ uint32_t from_next_lane(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xf, 0xf, true);
}

// Register budget: at least 3 waves per SIMD (168 VGPRs) up to 16 products per
// instance; the A fragments are read one K-chunk ahead, which keeps even the
// four-row FireSaber / Sable-3 shapes near 150 VGPRs.  The encryption shapes
// (L + 1 rows) take 2, which they fill without spilling.
// The 32x32x32 tile keeps two sets of 16 accumulators per row pair and takes
// two waves per SIMD everywhere (three would spill).
template <int L, int MODE, int SHAPE = 16>
constexpr int kWavesPerEU = (SHAPE == 32 || MODE == kEnc || rows_of(MODE, L) * terms_of(MODE, L) > 16) ? 2 : 3;

// LDS of one workgroup (four waves, two slices each).
template <int L, int MODE> constexpr int kLdsBytes = 4 * 2 * Slice<L, MODE>::BYTES;

// The work of one wave: instances first, first + stride, ... of job J, with
// the wave's own two LDS slices (2 * Slice::BYTES) at slice0.  SHAPE is the
// MFMA tile: 16 (v_mfma_i32_16x16x64_i8) or 32 (v_mfma_i32_32x32x32_i8).
// first and stride are wave-uniform (the callers derive them from a
// readfirstlane'd wave index), so the instance loop's addresses, bounds and
// branches stay in scalar registers.  STRIDED: instance i's public
// polynomials start at J.a + i * J.astride and its secret at J.s + i *
// J.sstride (a sub-range of a longer row: the one-launch KEM's partial rows).
// PRE2 (16x16x64 only): the wave's range holds at most two instances, and the
// second one's inputs are loaded with the first's at the start instead of one
// term ahead (the one-launch KEM's claims, which are too short to hide a load
// behind one term of work).  QUAD: where eight coefficients per lane would
// leave lanes idle (one public row, or three), each lane stages four.
template <int L, int MODE, int EPI, int SHAPE = 16, bool STRIDED = false, bool PRE2 = false, bool QUAD = true>
__device__ __forceinline__ void mfma_body(const Job &J, int64_t first, int64_t stride, uint8_t *const slice0)
{
    using S = Slice<L, MODE>;
    constexpr int NR = S::NR, NT = S::NT, NPOLY = S::NPOLY, CJ = S::CJ, ITJ = S::ITJ;
    const int lane = threadIdx.x & 63;
    int64_t inst = first;
    if (inst >= J.batch) return; // wave-uniform: each wave owns private LDS slices, no barriers
    int buf = 0; // slice holding the term about to be read
    const uint32_t wmask = (1u << J.w) - 1u;
    const uint32_t lo4 = 0x7f7f7f7fu;                 // lo limb mask of four gathered bytes
    const uint32_t hi4 = (wmask >> 7) * 0x01010101u;  // hi limb mask of four gathered bytes (w >= 8)
    const uint32_t wm2 = wmask * 0x00010001u;

    // lane roles
    const int blk = (lane & 31) >> 2;     // staging: 32-coefficient block of the lane's chunk
    const int bpos = 24 - 8 * (lane & 3); // staging: byte position inside the reversed block
    const int il = lane & 15, g = lane >> 4;
    // R (block column) and limb of the lane's MFMA column, in both shapes
    const int R = lane & 7, limb = (lane >> 3) & 1;
    // 32x32x32: row i and K half of the A operand, row of the pair (rr) of the B operand
    const int i32 = lane & 31, kh = lane >> 5, rr = (lane >> 4) & 1;
    const int sh = SHAPE == 16 ? (il + 1) & 3 : (i32 + 1) & 3;
    // A operand: dword index of the window start x0 - sh in band copy `sh`, without
    // the compile-time part.  16: x0 = 32D + 16h - 16(g&1) + il + 17 with
    // D = 2p + (g>>1), part 16p + 4h; 32: x0 = 32D + i + 1 + 16kh, part 8D.
    // (x0 - sh is a multiple of 4.)
    const int abase = SHAPE == 16 ? band_copy(sh) / 4 + 8 * (g >> 1) - 4 * (g & 1) + ((il + 17 - sh) >> 2)
                                  : band_copy(sh) / 4 + 4 * kh + ((i32 + 1 - sh) >> 2);
    // B operand: byte offset of the lane's E block in its limb plane (and row),
    // 16: E[R - (g>>1) + 8], minus 64 p per K-chunk; 32: E[R + 8] of row rr of
    // the pair, minus 32 D per tile, plus 2 kEPoly per row.
    const int bbase = SHAPE == 16 ? limb * kEPoly + 16 * (1 - (g & 1)) + (R - (g >> 1) + 8) * kEBlock
                                  : rr * 2 * kEPoly + limb * kEPoly + 16 * kh + (R + 8) * kEBlock;

    // Registers holding one instance's inputs, term by term.
    // One public row (InnerProd, polymul): each lane loads and stages four
    // coefficients of the term, so all 64 lanes share the staging work.
    // Three rows (the L = 3 MatrixVectorMul forms, 96 eight-coefficient
    // chunks, which leave half the lanes idle in the second pass): three
    // passes of four coefficients per lane (Saber, 65536 instances: 103.9-
    // 106.7 against 105.4-107.5 us in alternating runs, 120 VGPRs against 128).
    constexpr bool HALF = QUAD && CJ == 32;
    constexpr bool Q3 = QUAD && CJ == 96 && (MODE == kMatVec || MODE == kMatVecT);
    constexpr int ITQ = Q3 ? NR : ITJ;
    using PubT = std::conditional_t<HALF || Q3, u32x2, u32x4>;
    PubT va[NT][ITQ];
    u32x2 vs[NT];
    PubT vb[PRE2 ? NT : 1][ITQ]; // PRE2: the second instance's inputs
    u32x2 vsb[PRE2 ? NT : 1];
    static_assert(!PRE2 || SHAPE == 16, "PRE2 is a 16x16x64 form");
    // Loads of term j of instance i.  Every load is issued on every path (an
    // instance past the end reloads the current one, lanes without a chunk
    // reload another chunk of the same term and ignore it), so the compiler
    // sees one fixed sequence of memory operations per iteration and can wait
    // for exactly the loads it needs; conditional loads make it wait for all.
    // Public polynomials per instance in J.a (kEnc: the L x L matrix; its row
    // L is b, read from J.b).
    constexpr int NPA = (MODE == kEnc) ? L * L : NPOLY;
    auto fetch_to = [&](int64_t i, int j, PubT (&da)[ITQ], u32x2 &dsv) {
        const int64_t ie = i < J.batch ? i : inst;
        int64_t ia = ie; // the instance's matrix (kEnc may read it from a key store)
        if constexpr (MODE == kEnc)
            if (J.akey) ia = J.akey[ie];
        const u32x4 *src = reinterpret_cast<const u32x4 *>(J.a + ia * (STRIDED ? J.astride : (int64_t)NPA * kN));
        if constexpr (HALF) {
            // the term's row is polynomial j of the instance (InnerProd) or its only one (polymul)
            da[0] = reinterpret_cast<const u32x2 *>(src)[j * 64 + lane];
        } else if constexpr (Q3) {
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int pq = (MODE == kMatVec) ? r * L + j : j * L + r;
                da[r] = reinterpret_cast<const u32x2 *>(src)[pq * 64 + lane];
            }
        } else {
#pragma unroll
        for (int it = 0; it < ITJ; it++) {
            const int u0 = lane + 64 * it;                             // chunk of this term: row u >> 5, piece u & 31
            const int u = (CJ % 64 == 0 || u0 < CJ) ? u0 : u0 - CJ;    // always a chunk of this term
            const int r = u >> 5;
            if constexpr (MODE == kEnc) {
                const u32x4 *row = r < L ? src + (r * L + j) * 32
                                         : reinterpret_cast<const u32x4 *>(J.b + ie * (L * kN)) + j * 32;
                da[it] = row[u & 31];
            } else {
                const int pq = (MODE == kMatVec) ? r * L + j : (MODE == kMatVecT) ? j * L + r : j;
                da[it] = src[pq * 32 + (u & 31)];
            }
        }
        }
        dsv = reinterpret_cast<const u32x2 *>(J.s + ie * (STRIDED ? J.sstride : (int64_t)NT * kN))[j * 64 + lane];
    };
    auto fetch_term = [&](int64_t i, int j) { fetch_to(i, j, va[j], vs[j]); };
#pragma unroll
    for (int j = 0; j < NT; j++) fetch_term(inst, j);
    if constexpr (PRE2) {
#pragma unroll
        for (int j = 0; j < NT; j++) fetch_to(inst + stride, j, vb[j], vsb[j]);
    }

    uint32_t bad = 0;
    // Stage term j from va[j] / vs[j] into slice `into`.
    auto stage = [&](int j, int into) {
        uint8_t *const E = slice0 + into * S::BYTES;
        uint8_t *const X = E + S::E_BYTES;
        // ---- public polynomials of term j -> byte-reversed limb blocks E[r][limb][e]
        if constexpr (Q3) {
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const u32x2 v = va[j][r];
                const uint32_t s0 = shr16x2(v.x, 7), s1 = shr16x2(v.y, 7);
                const uint32_t plo = perm(v.y, v.x, 0x00020406u) & lo4;
                const uint32_t phi = perm(s1, s0, 0x00020406u) & hi4;
                uint8_t *slot = E + r * 2 * kEPoly + 28 - 4 * (lane & 7);
                const int b4 = lane >> 3;
                *(uint32_t *)(slot + (b4 + 8) * kEBlock) = plo;
                *(uint32_t *)(slot + kEPoly + (b4 + 8) * kEBlock) = phi;
                *(uint32_t *)(slot + b4 * kEBlock) = neg_small(plo);
                *(uint32_t *)(slot + kEPoly + b4 * kEBlock) = neg_small(phi);
            }
        } else if constexpr (HALF) {
            // Lane l holds coefficients 4l..4l+3 (block l >> 3, piece l & 7);
            // coefficient c of a block goes to byte 31 - c, as below.
            const u32x2 v = va[j][0];
            const uint32_t s0 = shr16x2(v.x, 7), s1 = shr16x2(v.y, 7);
            const uint32_t plo = perm(v.y, v.x, 0x00020406u) & lo4;
            const uint32_t phi = perm(s1, s0, 0x00020406u) & hi4;
            uint8_t *slot = E + 28 - 4 * (lane & 7);
            const int b4 = lane >> 3;
            *(uint32_t *)(slot + (b4 + 8) * kEBlock) = plo;
            *(uint32_t *)(slot + kEPoly + (b4 + 8) * kEBlock) = phi;
            *(uint32_t *)(slot + b4 * kEBlock) = neg_small(plo);
            *(uint32_t *)(slot + kEPoly + b4 * kEBlock) = neg_small(phi);
        } else {
    #pragma unroll
            for (int it = 0; it < ITJ; it++) {
                // Lanes past the last chunk skip the staging.  (A branch-free form,
                // where they rewrite a copy of another lane's chunk, lets the
                // scheduler mix the staging, which waits for the prefetched loads,
                // into the MFMAs, and measured 4 us slower at B = 65536.)
                const int u = lane + 64 * it;
                if (CJ % 64 != 0 && u >= CJ) continue;
                const int r = u >> 5;
                const u32x4 v = va[j][it];
                // hi limb = coefficient >> 7, per 16-bit half
                const uint32_t s0 = shr16x2(v.x, 7), s1 = shr16x2(v.y, 7), s2 = shr16x2(v.z, 7), s3 = shr16x2(v.w, 7);
                // gather the low byte of every coefficient, coefficient u of the chunk to byte 7 - u
                const uint2 plo = make_uint2(perm(v.w, v.z, 0x00020406u) & lo4, perm(v.y, v.x, 0x00020406u) & lo4);
                const uint2 phi = make_uint2(perm(s3, s2, 0x00020406u) & hi4, perm(s1, s0, 0x00020406u) & hi4);
                uint8_t *slot = E + r * 2 * kEPoly + bpos;
                *(uint2 *)(slot + (blk + 8) * kEBlock) = plo;
                *(uint2 *)(slot + kEPoly + (blk + 8) * kEBlock) = phi;
                // negated copy; for blk = 0 it lands in block 0, which no MFMA reads
                *(uint2 *)(slot + blk * kEBlock) = make_uint2(neg_small(plo.x), neg_small(plo.y));
                *(uint2 *)(slot + kEPoly + blk * kEBlock) = make_uint2(neg_small(phi.x), neg_small(phi.y));
            }
        }
        // ---- secret s_j -> int8 band X[x] = ts(x - 32), stored 4x shifted: copy c
        //      holds X[x + c] at byte x.  For w >= 8 the centred int8 is the low byte.
        const u32x2 sv = vs[j];
        const uint32_t pos = perm(sv.y, sv.x, 0x06040200u);
        // range check: centred value in [-127, 127]  <=>  ((s + 127) mod 2^w) + 1 <= 255
        bad |= (add16x2(add16x2(sv.x, 0x007f007fu) & wm2, 0x00010001u) |
                add16x2(add16x2(sv.y, 0x007f007fu) & wm2, 0x00010001u));
        const uint32_t neg = neg_bytes(pos);
        // Band dword k: k = lane + 8 is `pos` of this lane, k = lane - 56 (lanes
        // 56..63) is `neg`; dword 72 is zero padding.
        const uint32_t p1 = from_next_lane(pos);
        // n1: the next lane's neg, and for lane 63 (dword 8 follows dword 7)
        // lane 0's pos.  One DPP move whose out-of-range lane keeps the old
        // value: a select here was compiled into a branch around the DPP move,
        // whose source lane 63 was then disabled and read as 0.
        const uint32_t n1 =
            (uint32_t)__builtin_amdgcn_update_dpp((int)__builtin_amdgcn_readlane(pos, 0), (int)neg, 0x130, 0xf, 0xf, false);
        uint8_t *band = X + 32 + 4 * lane;
        *(uint32_t *)band = pos;
#pragma unroll
        for (int c = 1; c < 4; c++) *(uint32_t *)(band + band_copy(c)) = __builtin_amdgcn_alignbyte(p1, pos, c);
        // The wrapped dwords 0..7 come from lanes 56..63.
        if (lane >= 56) {
            uint8_t *nb = X + 4 * lane - 224;
#pragma unroll
            for (int c = 0; c < 4; c++) *(uint32_t *)(nb + band_copy(c)) = c ? __builtin_amdgcn_alignbyte(n1, neg, c) : neg;
        }
    };
    // The slices are private to this wave and the LDS completes one wave's
    // operations in order: a wave-level fence at the end of every term keeps
    // the compiler from moving a term's staging writes past its reads, or the
    // reads of a slice past the staging that later overwrites it.
    auto wave_fence = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    // Every use of a term's input registers must come before the prefetch that
    // reloads them: otherwise the compiler gives the load another register,
    // copies it back at the loop edge and waits for it there (vmcnt), which
    // exposes the full HBM latency once per instance.  The staging's LDS
    // writes are ordered by the memory clobber; its one pure-register result,
    // the range-check word, is routed through the barrier.
    auto pin_after_stage = [&] { asm volatile("" : "+v"(bad) : : "memory"); };

    // Software pipeline: the staging of the following term (the next term of
    // this instance, or term 0 of the next instance) goes to the other slice
    // and is issued between the MFMAs of the current term.
    stage(0, 0);
    pin_after_stage();
    if constexpr (!PRE2) fetch_term(inst + stride, 0);
    wave_fence();
    // Accumulators start at the rounding constant h in the lo-limb lanes, so
    // the epilogue never adds it.
    const int hinit = (EPI != kEpiMod && limb == 0) ? J.h : 0;
    const v4i init = {hinit, hinit, hinit, hinit};
    v16i init16;
#pragma unroll
    for (int e = 0; e < 16; e++) init16[e] = hinit;
    // Epilogue roles: lo-limb lanes finish rows v = 0, 1 of their group, hi-limb
    // lanes rows v = 2, 3; each sends the partner the other pair.
    const uint32_t lshift = limb ? 0x00070007u : 0u;
    for (;;) {
        const int64_t next = inst + stride;
        if constexpr (SHAPE == 16) {
            v4i acc[NR][2];
    #pragma unroll
            for (int j = 0; j < NT; j++) {
                const uint8_t *const E = slice0 + buf * S::BYTES;
                // A operand: Toeplitz tiles of s_j for the four K-chunks and both halves,
                // 16 bytes from a dword-aligned window each.  The window's byte offset is
                // made opaque so that the 16 reads share one address register and carry
                // their constant parts as instruction offsets.
                int aoff = buf * S::BYTES + S::E_BYTES + 4 * abase;
                asm volatile("" : "+v"(aoff));
                const uint32_t *aw = reinterpret_cast<const uint32_t *>(slice0 + aoff);
                // A fragments of one K-chunk (both halves); chunk p + 1 is read while
                // chunk p is multiplied.
                auto read_a = [&](int p, v4i (&f)[2]) {
    #pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const uint32_t *w = aw + 16 * p + 4 * h;
                        f[h] = v4i{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
                    }
                };
                v4i acur[2];
                read_a(0, acur);
                // B operand of K-chunk 0; chunk p + 1 is read while chunk p is multiplied.
                v4i bcur[NR];
    #pragma unroll
                for (int r = 0; r < NR; r++) bcur[r] = *(const v4i *)(E + r * 2 * kEPoly + bbase);

                if (j + 1 < NT) {
                    stage(j + 1, buf ^ 1);
                    pin_after_stage();
                    if constexpr (!PRE2) fetch_term(next, j + 1);
                } else if (next < J.batch) {
                    if constexpr (PRE2) {
#pragma unroll
                        for (int jj = 0; jj < NT; jj++) {
#pragma unroll
                            for (int it = 0; it < ITQ; it++) va[jj][it] = vb[jj][it];
                            vs[jj] = vsb[jj];
                        }
                    }
                    stage(0, buf ^ 1); // va[0] / vs[0] already hold the next instance's term 0
                    pin_after_stage();
                    if constexpr (!PRE2) fetch_term(next + stride, 0);
                }

                // K-chunk outermost: 2*NR independent accumulators between dependent MFMAs.
    #pragma unroll
                for (int p = 0; p < 4; p++) {
                    v4i bnxt[NR], anxt[2];
                    if (p < 3) {
                        read_a(p + 1, anxt);
    #pragma unroll
                        for (int r = 0; r < NR; r++) bnxt[r] = *(const v4i *)(E + r * 2 * kEPoly + bbase - 64 * (p + 1));
                    }
    #pragma unroll
                    for (int r = 0; r < NR; r++) {
                        const bool first = (j == 0 && p == 0);
                        acc[r][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(acur[0], bcur[r], first ? init : acc[r][0], 0, 0, 0);
                        acc[r][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(acur[1], bcur[r], first ? init : acc[r][1], 0, 0, 0);
                    }
                    if (p < 3) {
                        acur[0] = anxt[0];
                        acur[1] = anxt[1];
    #pragma unroll
                        for (int r = 0; r < NR; r++) bcur[r] = bnxt[r];
                    }
                }
                // Schedule: chunk 0's fragment reads first, then per chunk the next
                // chunk's A and B reads and one MFMA per three staging VALU instructions.
                __builtin_amdgcn_sched_group_barrier(0x100, 4 + NR, 0);
    #pragma unroll
                for (int p = 0; p < 4; p++) {
                    if (p < 3) __builtin_amdgcn_sched_group_barrier(0x100, 4 + NR, 0);
    #pragma unroll
                    for (int k = 0; k < 2 * NR; k++) {
                        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                        __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);
                    }
                }
                buf ^= 1;
                wave_fence();
            }

            // ---- epilogue, in packed 16-bit arithmetic (every output is mod 2^w,
            //      w <= 14): lane (R, limb, g) holds rows i = 16h + 4g + v of block
            //      column R for its limb.  Both lanes of a pair scale their own limb
            //      (hi by 128), keep one row pair and send the other across
            //      row_ror:8, so each finishes two rows: lo lanes v = 0,1, hi v = 2,3.
            //      kEnc: rows 0..L-1 go to J.out (q->p), row L to J.out2 (enc).
            constexpr int NRO = (MODE == kEnc) ? L : NR; // rows in J.out per instance
            const int lpos = 32 * R + 4 * g + 2 * limb;
            uint16_t *obase = J.out + inst * (NRO * kN) + lpos;
            const bool with_aux = EPI == kEpiEnc || EPI == kEpiDec || MODE == kEnc;
            const uint16_t *xbase = with_aux ? J.aux + inst * kN + lpos : nullptr;
    #pragma unroll
            for (int r = 0; r < NR; r++) {
    #pragma unroll
                for (int h = 0; h < 2; h++) {
                    const uint32_t p01 = shl16x2(perm((uint32_t)acc[r][h][1], (uint32_t)acc[r][h][0], 0x05040100u), lshift);
                    const uint32_t p23 = shl16x2(perm((uint32_t)acc[r][h][3], (uint32_t)acc[r][h][2], 0x05040100u), lshift);
                    const uint32_t keep = limb ? p23 : p01;
                    const uint32_t send = limb ? p01 : p23;
                    const uint32_t y = add16x2(keep, (uint32_t)__builtin_amdgcn_mov_dpp((int)send, 0x128, 0xf, 0xf, false));
                    if (MODE == kEnc && r == L) {
                        const uint32_t auxv = *(const uint32_t *)(xbase + 16 * h);
                        *(uint32_t *)(J.out2 + inst * kN + lpos + 16 * h) = epi2<kEpiEnc>(J, y, 0u, auxv);
                        continue;
                    }
                    uint32_t old = 0u, auxv = 0u;
                    if ((EPI == kEpiMod && J.accumulate) || EPI == kEpiQ2PAcc)
                        old = *(const uint32_t *)(obase + r * kN + 16 * h);
                    if (EPI == kEpiEnc || EPI == kEpiDec) auxv = *(const uint32_t *)(xbase + 16 * h);
                    *(uint32_t *)(obase + r * kN + 16 * h) = epi2<EPI>(J, y, old, auxv);
                }
            }
        } else {
            // ---- 32x32x32 tiles: per term, D = 0..7, one A fragment per tile shared
            //      by the NP row pairs, one B fragment per (tile, pair)
            constexpr int NP = (NR + 1) / 2;
            // two accumulator sets (even and odd D tiles), summed in the
            // epilogue: the same accumulator then recurs every 2 NP MFMAs
            v16i acc[2][NP];
#pragma unroll
            for (int j = 0; j < NT; j++) {
                const uint8_t *const E = slice0 + buf * S::BYTES;
                int aoff = buf * S::BYTES + S::E_BYTES + 4 * abase;
                asm volatile("" : "+v"(aoff));
                const uint32_t *aw = reinterpret_cast<const uint32_t *>(slice0 + aoff);
                auto read_a = [&](int D) {
                    const uint32_t *w = aw + 8 * D;
                    return v4i{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
                };
                auto read_b = [&](int D, int rp) { return *(const v4i *)(E + bbase + rp * 4 * kEPoly - kEBlock * D); };
                v4i acur = read_a(0);
                v4i bcur[NP];
#pragma unroll
                for (int rp = 0; rp < NP; rp++) bcur[rp] = read_b(0, rp);

                if (j + 1 < NT) {
                    stage(j + 1, buf ^ 1);
                    pin_after_stage();
                    fetch_term(next, j + 1);
                } else if (next < J.batch) {
                    stage(0, buf ^ 1); // va[0] / vs[0] already hold the next instance's term 0
                    pin_after_stage();
                    fetch_term(next + stride, 0);
                }

                // tile D + 1's fragments are read while tile D is multiplied
#pragma unroll
                for (int D = 0; D < 8; D++) {
                    v4i anxt, bnxt[NP];
                    if (D < 7) {
                        anxt = read_a(D + 1);
#pragma unroll
                        for (int rp = 0; rp < NP; rp++) bnxt[rp] = read_b(D + 1, rp);
                    }
#pragma unroll
                    for (int rp = 0; rp < NP; rp++) {
                        const v16i c0 = (D == 0) ? init16 : v16i{};
                        acc[D & 1][rp] = __builtin_amdgcn_mfma_i32_32x32x32_i8(acur, bcur[rp],
                                                                               (j == 0 && D < 2) ? c0 : acc[D & 1][rp], 0, 0, 0);
                    }
                    if (D < 7) {
                        acur = anxt;
#pragma unroll
                        for (int rp = 0; rp < NP; rp++) bcur[rp] = bnxt[rp];
                    }
                }
                // Schedule: tile 0's fragment reads first, then per tile the next
                // tile's reads and one MFMA per five staging VALU instructions.
                __builtin_amdgcn_sched_group_barrier(0x100, 2 + NP, 0);
#pragma unroll
                for (int D = 0; D < 8; D++) {
                    if (D < 7) __builtin_amdgcn_sched_group_barrier(0x100, 2 + NP, 0);
#pragma unroll
                    for (int k = 0; k < NP; k++) {
                        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                        __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
                    }
                }
                buf ^= 1;
                wave_fence();
            }

            // ---- epilogue: lane (R, limb, rr, kh) of pair rp holds rows
            //      i = 8q + 4kh + v (v = 0..3) in accumulators 4q + v, of public row
            //      2rp + rr.  As in the 16x16 form, each lane scales its limb, keeps
            //      one row pair of every group and sends the other across row_ror:8.
            constexpr int NRO = (MODE == kEnc) ? L : NR; // rows in J.out per instance
#pragma unroll
            for (int rp = 0; rp < NP; rp++) {
                const int r = 2 * rp + rr;
                const v16i sum = acc[0][rp] + acc[1][rp];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t p01 =
                        shl16x2(perm((uint32_t)sum[4 * q + 1], (uint32_t)sum[4 * q], 0x05040100u), lshift);
                    const uint32_t p23 =
                        shl16x2(perm((uint32_t)sum[4 * q + 3], (uint32_t)sum[4 * q + 2], 0x05040100u), lshift);
                    const uint32_t keep = limb ? p23 : p01;
                    const uint32_t send = limb ? p01 : p23;
                    const uint32_t y = add16x2(keep, (uint32_t)__builtin_amdgcn_mov_dpp((int)send, 0x128, 0xf, 0xf, false));
                    if (r >= NR) continue; // the unused row of an odd row count
                    const int pos = 32 * R + 8 * q + 4 * kh + 2 * limb;
                    const bool need_aux = MODE == kEnc ? r == L : (EPI == kEpiEnc || EPI == kEpiDec);
                    const uint32_t auxv = need_aux ? *(const uint32_t *)(J.aux + inst * kN + pos) : 0u;
                    if (MODE == kEnc && r == L) {
                        *(uint32_t *)(J.out2 + inst * kN + pos) = epi2<kEpiEnc>(J, y, 0u, auxv);
                        continue;
                    }
                    uint16_t *o = J.out + inst * (NRO * kN) + r * kN + pos;
                    const uint32_t old = ((EPI == kEpiMod && J.accumulate) || EPI == kEpiQ2PAcc) ? *(const uint32_t *)o : 0u;
                    *(uint32_t *)o = epi2<EPI>(J, y, old, auxv);
                }
            }
        }
        inst = next;
        if (inst >= J.batch) break;
    }
    if (bad & 0xff00ff00u) *J.status = 1;
}

} // namespace tmvp
