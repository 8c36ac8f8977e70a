// Keccak-f[1600] and the FIPS 202 sponges (SHAKE-128, SHA3-256, SHA3-512),
// one state split across a pair of adjacent lanes: lane p = lane & 1 holds
// the low (p = 0) or high (p = 1) 32 bits of each of the 25 words.
//
// A 64-bit rotation by r then needs the partner's half, fetched with one DPP
// swap of adjacent lanes, and one v_alignbit: for 0 < r < 32 the new half is
// alignbit(own, partner, 32 - r) in both lanes, for r > 32 it is
// alignbit(partner, own, 64 - r).  With three-input XORs and chi in v_bitop3,
// a round costs about 125 VALU instructions per lane, against about 290 for a
// whole state in one lane, and a batch gets twice the waves.
//
// Round constants come from the rc(t) LFSR of FIPS 202 Algorithm 5 and the
// rho offsets from the (x, y) -> (y, 2x + 3y) walk of Algorithm 2, both
// evaluated at compile time.  Both lanes of a pair must execute every call
// together (the DPP swap reads the partner's register).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace tmvp {
namespace keccak {

struct Tables {
    uint64_t rc[24];
    int rho[25];
};

constexpr int lfsr_bit(int t)
{
    int m = t % 255;
    if (m == 0) return 1;
    unsigned r = 1;
    for (int i = 0; i < m; i++) {
        r <<= 1;
        if (r & 0x100) r ^= 0x171;
    }
    return (int)(r & 1);
}

constexpr Tables make_tables()
{
    Tables T{};
    for (int ir = 0; ir < 24; ir++) {
        uint64_t rc = 0;
        for (int j = 0; j <= 6; j++)
            if (lfsr_bit(j + 7 * ir)) rc |= 1ull << ((1 << j) - 1);
        T.rc[ir] = rc;
    }
    T.rho[0] = 0;
    int x = 1, y = 0;
    for (int t = 0; t < 24; t++) {
        T.rho[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
        const int nx = y, ny = (2 * x + 3 * y) % 5;
        x = nx;
        y = ny;
    }
    return T;
}

constexpr Tables kT = make_tables();
static_assert(kT.rc[0] == 1ull, "rc(0)");
static_assert(kT.rho[1] == 1 && kT.rho[2] == 62, "rho offsets");

// Round constants as (low, high) halves.
__device__ __constant__ uint32_t kRC2[48] = {
#define TMVP_RC(i) (uint32_t)kT.rc[i], (uint32_t)(kT.rc[i] >> 32)
    TMVP_RC(0),  TMVP_RC(1),  TMVP_RC(2),  TMVP_RC(3),  TMVP_RC(4),  TMVP_RC(5),  TMVP_RC(6),  TMVP_RC(7),
    TMVP_RC(8),  TMVP_RC(9),  TMVP_RC(10), TMVP_RC(11), TMVP_RC(12), TMVP_RC(13), TMVP_RC(14), TMVP_RC(15),
    TMVP_RC(16), TMVP_RC(17), TMVP_RC(18), TMVP_RC(19), TMVP_RC(20), TMVP_RC(21), TMVP_RC(22), TMVP_RC(23)
#undef TMVP_RC
};

// The partner lane's value (DPP quad_perm [1, 0, 3, 2]).
__device__ __forceinline__ uint32_t partner(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xf, 0xf, false);
}

// This lane's half of rotl64(word, R), given its half and the partner's.
template <int R> __device__ __forceinline__ uint32_t rotl2(uint32_t own, uint32_t other)
{
    if constexpr (R == 0) return own;
    else if constexpr (R < 32) return __builtin_amdgcn_alignbit(own, other, 32 - R);
    else if constexpr (R == 32) return other;
    else return __builtin_amdgcn_alignbit(other, own, 64 - R);
}

// rho + pi for lane (x, y): B[y + 5((2x + 3y) mod 5)] = rotl(A[x + 5y], rho[x + 5y])
template <int I> __device__ __forceinline__ void rho_pi(uint32_t (&B)[25], const uint32_t (&A)[25])
{
    constexpr int x = I % 5, y = I / 5;
    B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl2<kT.rho[I]>(A[I], kT.rho[I] ? partner(A[I]) : 0u);
}

template <int... Is>
__device__ __forceinline__ void rho_pi_all(uint32_t (&B)[25], const uint32_t (&A)[25],
                                           std::integer_sequence<int, Is...>)
{
    (rho_pi<Is>(B, A), ...);
}

// a ^ b ^ c in one v_bitop3 (truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// U rounds per loop iteration (1 or 2).  At one or two sponge waves per SIMD,
// 2 makes a lone permutation 6-11 % faster, but whole kernels gain or lose
// with it depending on the code around it, so each kernel chooses.
template <int U = 1> __device__ __forceinline__ void f1600(uint32_t (&A)[25], int p)
{
    // The round constant is read at a uniform index (a scalar load) and the
    // lane's half picked with a mask: a per-lane index would make it a vector
    // load, and its wait would also wait for every store still in flight.
    const uint32_t hi_mask = 0u - (uint32_t)p;
#pragma unroll U
    for (int round = 0; round < 24; round++) {
        uint32_t C[5], R[5], B[25];
        // theta: column parities C, R = rotl(C, 1), A ^= C[x-1] ^ R[x+1]
#pragma unroll
        for (int x = 0; x < 5; x++) C[x] = xor3(xor3(A[x], A[x + 5], A[x + 10]), A[x + 15], A[x + 20]);
#pragma unroll
        for (int x = 0; x < 5; x++) R[x] = rotl2<1>(C[x], partner(C[x]));
#pragma unroll
        for (int x = 0; x < 5; x++)
#pragma unroll
            for (int y = 0; y < 25; y += 5) A[x + y] = xor3(A[x + y], C[(x + 4) % 5], R[(x + 1) % 5]);
        rho_pi_all(B, A, std::make_integer_sequence<int, 25>{});
#pragma unroll
        for (int y = 0; y < 25; y += 5)
#pragma unroll
            for (int x = 0; x < 5; x++) A[x + y] = B[x + y] ^ (~B[(x + 1) % 5 + y] & B[(x + 2) % 5 + y]);
        const uint32_t rc_lo = kRC2[2 * round], rc_hi = kRC2[2 * round + 1];
        A[0] ^= (rc_lo & ~hi_mask) | (rc_hi & hi_mask);
    }
}

constexpr int kShake128Rate = 168;
constexpr int kSha3_256Rate = 136;
constexpr int kSha3_512Rate = 72;
constexpr uint8_t kShakeDS = 0x1F;
constexpr uint8_t kSha3DS = 0x06;

// This lane's halves of the n 64-bit words at p_ (8-byte aligned).
template <int N> __device__ __forceinline__ void load_words(uint32_t (&m)[N], const uint8_t *src, int p)
{
    const uint32_t *w = reinterpret_cast<const uint32_t *>(src);
#pragma unroll
    for (int i = 0; i < N; i++) m[i] = w[2 * i + p];
}
template <int N> __device__ __forceinline__ void store_words(uint8_t *dst, const uint32_t (&m)[N], int p)
{
    uint32_t *w = reinterpret_cast<uint32_t *>(dst);
#pragma unroll
    for (int i = 0; i < N; i++) w[2 * i + p] = m[i];
}

// Absorb `len` bytes (8-byte aligned source) with rate RATE and domain byte
// `ds`, leaving the state ready for the first squeeze permutation.  Every
// state index is a compile-time constant so the state stays in registers.
template <int RATE, int U = 1>
__device__ __forceinline__ void absorb(uint32_t (&A)[25], const uint8_t *in, int len, uint8_t ds, int p)
{
    constexpr int RW = RATE / 8;
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = 0;
    const uint32_t *w = reinterpret_cast<const uint32_t *>(in) + p;
    while (len >= RATE) {
#pragma unroll
        for (int i = 0; i < RW; i++) A[i] ^= w[2 * i];
        f1600<U>(A, p);
        w += 2 * RW;
        len -= RATE;
    }
    const int full = len >> 3, rem = len & 7;
    uint64_t last = (uint64_t)ds << (8 * rem);
    const uint8_t *tail = reinterpret_cast<const uint8_t *>(w - p + 2 * full);
    for (int b = 0; b < rem; b++) last |= (uint64_t)tail[b] << (8 * b);
    const uint32_t lh = p ? (uint32_t)(last >> 32) : (uint32_t)last;
#pragma unroll
    for (int i = 0; i < RW; i++) {
        if (i < full) A[i] ^= w[2 * i];
        else if (i == full) A[i] ^= lh;
    }
    A[RW - 1] ^= p ? 0x80000000u : 0u;
}

// absorb with the next block's words loaded before the current block is
// permuted, so their latency hides behind the permutation (17 more VGPRs).
template <int RATE, int U = 1>
__device__ __forceinline__ void absorb_prefetch(uint32_t (&A)[25], const uint8_t *in, int len, uint8_t ds, int p)
{
    constexpr int RW = RATE / 8;
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = 0;
    const uint32_t *w = reinterpret_cast<const uint32_t *>(in) + p;
    uint32_t nxt[RW];
    if (len >= RATE) {
#pragma unroll
        for (int i = 0; i < RW; i++) nxt[i] = w[2 * i];
    }
    while (len >= RATE) {
#pragma unroll
        for (int i = 0; i < RW; i++) A[i] ^= nxt[i];
        w += 2 * RW;
        len -= RATE;
        if (len >= RATE) {
#pragma unroll
            for (int i = 0; i < RW; i++) nxt[i] = w[2 * i];
        }
        f1600<U>(A, p);
    }
    const int full = len >> 3, rem = len & 7;
    uint64_t last = (uint64_t)ds << (8 * rem);
    const uint8_t *tail = reinterpret_cast<const uint8_t *>(w - p + 2 * full);
    for (int b = 0; b < rem; b++) last |= (uint64_t)tail[b] << (8 * b);
    const uint32_t lh = p ? (uint32_t)(last >> 32) : (uint32_t)last;
#pragma unroll
    for (int i = 0; i < RW; i++) {
        if (i < full) A[i] ^= w[2 * i];
        else if (i == full) A[i] ^= lh;
    }
    A[RW - 1] ^= p ? 0x80000000u : 0u;
}

// Absorption of a message whose bytes become available in pieces (another
// wave is still writing the rest): absorb_upto takes every whole rate block of
// in[done, avail) into A (zeroed by the caller before the first piece),
// advancing `done`; absorb_final then takes the len - done < RATE bytes left
// and the padding.  in is 8-byte aligned and done stays a multiple of RATE.
template <int RATE, int U = 1>
__device__ __forceinline__ void absorb_upto(uint32_t (&A)[25], const uint8_t *in, int &done, int avail, int p)
{
    constexpr int RW = RATE / 8;
    while (done + RATE <= avail) {
        const uint32_t *w = reinterpret_cast<const uint32_t *>(in + done) + p;
        uint32_t m[RW];
#pragma unroll
        for (int i = 0; i < RW; i++) m[i] = w[2 * i];
#pragma unroll
        for (int i = 0; i < RW; i++) A[i] ^= m[i];
        f1600<U>(A, p);
        done += RATE;
    }
}

template <int RATE>
__device__ __forceinline__ void absorb_final(uint32_t (&A)[25], const uint8_t *in, int done, int len, uint8_t ds, int p)
{
    constexpr int RW = RATE / 8;
    const int rest = len - done, full = rest >> 3, rem = rest & 7;
    const uint32_t *w = reinterpret_cast<const uint32_t *>(in + done) + p;
    uint64_t last = (uint64_t)ds << (8 * rem);
    const uint8_t *tail = in + done + 8 * full;
    for (int b = 0; b < rem; b++) last |= (uint64_t)tail[b] << (8 * b);
    const uint32_t lh = p ? (uint32_t)(last >> 32) : (uint32_t)last;
#pragma unroll
    for (int i = 0; i < RW; i++) {
        if (i < full) A[i] ^= w[2 * i];
        else if (i == full) A[i] ^= lh;
    }
    A[RW - 1] ^= p ? 0x80000000u : 0u;
}

// SHA3-256 through absorb_prefetch.
template <int U = 1>
__device__ __forceinline__ void sha3_256_prefetch(uint32_t (&out)[4], const uint8_t *in, int len, int p)
{
    uint32_t A[25];
    absorb_prefetch<kSha3_256Rate, U>(A, in, len, kSha3DS, p);
    f1600<U>(A, p);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = A[i];
}

// Absorb N words held in registers (N * 8 bytes < rate).
template <int N>
__device__ __forceinline__ void absorb_words(uint32_t (&A)[25], const uint32_t (&m)[N], int rate, uint8_t ds, int p)
{
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = 0;
#pragma unroll
    for (int i = 0; i < N; i++) A[i] = m[i];
    A[N] ^= p ? 0u : (uint32_t)ds;
    if (rate == kShake128Rate) A[20] ^= p ? 0x80000000u : 0u;
    else if (rate == kSha3_256Rate) A[16] ^= p ? 0x80000000u : 0u;
    else A[8] ^= p ? 0x80000000u : 0u;
}

// SHA3-256 of an 8-byte-aligned buffer -> four words.
template <int U = 1> __device__ __forceinline__ void sha3_256(uint32_t (&out)[4], const uint8_t *in, int len, int p)
{
    uint32_t A[25];
    absorb<kSha3_256Rate, U>(A, in, len, kSha3DS, p);
    f1600<U>(A, p);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = A[i];
}

template <int N, int U = 1> __device__ __forceinline__ void sha3_256_words(uint32_t (&out)[4], const uint32_t (&m)[N], int p)
{
    uint32_t A[25];
    absorb_words(A, m, kSha3_256Rate, kSha3DS, p);
    f1600<U>(A, p);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = A[i];
}

template <int U = 1> __device__ __forceinline__ void sha3_512_words(uint32_t (&out)[8], const uint32_t (&m)[8], int p)
{
    uint32_t A[25];
    absorb_words(A, m, kSha3_512Rate, kSha3DS, p);
    f1600<U>(A, p);
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = A[i];
}

// Squeeze `nwords` 64-bit words of SHAKE-128 output (absorbed, not yet
// permuted state), calling f(k, half) with this lane's half of word k.
template <int U = 1, class F> __device__ __forceinline__ void squeeze_each(uint32_t (&A)[25], int nwords, int p, F f)
{
    for (int k = 0; k < nwords; k += kShake128Rate / 8) {
        f1600<U>(A, p);
#pragma unroll
        for (int i = 0; i < kShake128Rate / 8; i++)
            if (k + i < nwords) f(k + i, A[i]);
    }
}

} // namespace keccak
} // namespace tmvp
