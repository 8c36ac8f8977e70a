// Keccak-f[1600] and the FIPS 202 sponges (SHAKE-128, SHA3-256, SHA3-512)
// for one thread per sponge, state in registers.
//
// Round constants come from the rc(t) LFSR of FIPS 202 Algorithm 5 and the
// rho offsets from the (x, y) -> (y, 2x + 3y) walk of Algorithm 2, both
// evaluated at compile time.  All sponge inputs and outputs are 8-byte
// aligned per instance (every Saber/Sable key, ciphertext and seed size is a
// multiple of 8), so lanes are moved as whole little-endian 64-bit words.
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

__device__ __constant__ uint64_t kRC[24] = {
    kT.rc[0],  kT.rc[1],  kT.rc[2],  kT.rc[3],  kT.rc[4],  kT.rc[5],  kT.rc[6],  kT.rc[7],
    kT.rc[8],  kT.rc[9],  kT.rc[10], kT.rc[11], kT.rc[12], kT.rc[13], kT.rc[14], kT.rc[15],
    kT.rc[16], kT.rc[17], kT.rc[18], kT.rc[19], kT.rc[20], kT.rc[21], kT.rc[22], kT.rc[23]};

template <int N> __device__ __forceinline__ uint64_t rotl(uint64_t v)
{
    if constexpr (N == 0) return v;
    else return (v << N) | (v >> (64 - N));
}

// rho + pi for lane (x, y): B[y + 5((2x + 3y) mod 5)] = rotl(A[x + 5y], rho[x + 5y])
template <int I> __device__ __forceinline__ void rho_pi(uint64_t (&B)[25], const uint64_t (&A)[25])
{
    constexpr int x = I % 5, y = I / 5;
    B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl<kT.rho[I]>(A[I]);
}

template <int... Is>
__device__ __forceinline__ void rho_pi_all(uint64_t (&B)[25], const uint64_t (&A)[25],
                                           std::integer_sequence<int, Is...>)
{
    (rho_pi<Is>(B, A), ...);
}

__device__ __forceinline__ void f1600(uint64_t (&A)[25])
{
#pragma unroll 1
    for (int round = 0; round < 24; round++) {
        uint64_t C[5], B[25];
#pragma unroll
        for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
#pragma unroll
        for (int x = 0; x < 5; x++) {
            const uint64_t D = C[(x + 4) % 5] ^ rotl<1>(C[(x + 1) % 5]);
#pragma unroll
            for (int y = 0; y < 25; y += 5) A[x + y] ^= D;
        }
        rho_pi_all(B, A, std::make_integer_sequence<int, 25>{});
#pragma unroll
        for (int y = 0; y < 25; y += 5)
#pragma unroll
            for (int x = 0; x < 5; x++) A[x + y] = B[x + y] ^ (~B[(x + 1) % 5 + y] & B[(x + 2) % 5 + y]);
        A[0] ^= kRC[round];
    }
}

// Absorb `len` bytes (8-byte aligned source) with rate RATE and domain byte
// `ds`, leaving the state ready for the first squeeze permutation.  Every
// state index is a compile-time constant so the state stays in registers.
template <int RATE> __device__ __forceinline__ void absorb(uint64_t (&A)[25], const uint8_t *in, int len, uint8_t ds)
{
    constexpr int RW = RATE / 8;
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = 0;
    const uint64_t *w = reinterpret_cast<const uint64_t *>(in);
    while (len >= RATE) {
#pragma unroll
        for (int i = 0; i < RW; i++) A[i] ^= w[i];
        f1600(A);
        w += RW;
        len -= RATE;
    }
    const int full = len >> 3, rem = len & 7;
    uint64_t last = (uint64_t)ds << (8 * rem);
    const uint8_t *tail = reinterpret_cast<const uint8_t *>(w + full);
    for (int b = 0; b < rem; b++) last |= (uint64_t)tail[b] << (8 * b);
#pragma unroll
    for (int i = 0; i < RW; i++) {
        if (i < full) A[i] ^= w[i];
        else if (i == full) A[i] ^= last;
    }
    A[RW - 1] ^= 0x8000000000000000ull;
}

// Absorb a 32-byte value held in four lanes.
__device__ __forceinline__ void absorb32(uint64_t (&A)[25], const uint64_t (&m)[4], int rate, uint8_t ds)
{
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) A[i] = m[i];
    A[4] = ds;
    A[rate / 8 - 1] ^= 0x8000000000000000ull;
}

// Absorb a 64-byte value held in eight lanes.
__device__ __forceinline__ void absorb64(uint64_t (&A)[25], const uint64_t (&m)[8], int rate, uint8_t ds)
{
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) A[i] = m[i];
    A[8] ^= ds;
    A[rate / 8 - 1] ^= 0x8000000000000000ull;
}

constexpr int kShake128Rate = 168;
constexpr int kSha3_256Rate = 136;
constexpr int kSha3_512Rate = 72;
constexpr uint8_t kShakeDS = 0x1F;
constexpr uint8_t kSha3DS = 0x06;

// SHA3-256 of an 8-byte-aligned buffer -> four lanes.
__device__ __forceinline__ void sha3_256(uint64_t (&out)[4], const uint8_t *in, int len)
{
    uint64_t A[25];
    absorb<kSha3_256Rate>(A, in, len, kSha3DS);
    f1600(A);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = A[i];
}

__device__ __forceinline__ void sha3_256_32(uint64_t (&out)[4], const uint64_t (&m)[4])
{
    uint64_t A[25];
    absorb32(A, m, kSha3_256Rate, kSha3DS);
    f1600(A);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = A[i];
}

__device__ __forceinline__ void sha3_256_64(uint64_t (&out)[4], const uint64_t (&m)[8])
{
    uint64_t A[25];
    absorb64(A, m, kSha3_256Rate, kSha3DS);
    f1600(A);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = A[i];
}

__device__ __forceinline__ void sha3_512_64(uint64_t (&out)[8], const uint64_t (&m)[8])
{
    uint64_t A[25];
    absorb64(A, m, kSha3_512Rate, kSha3DS);
    f1600(A);
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = A[i];
}

// Squeeze the sponge (absorbed, not yet permuted) as a stream of `count`
// w-bit values, w <= 16, least significant bit first, calling sink(k, v) for
// value k.  Blocks are consumed as 42 half-lanes in a fully unrolled loop, so
// no state word is indexed at run time.
template <class Sink> __device__ __forceinline__ void squeeze_values(uint64_t (&A)[25], int w, int count, Sink &&sink)
{
    const uint32_t mask = (1u << w) - 1u;
    uint64_t buf = 0;
    int nbits = 0, k = 0;
    while (k < count) {
        f1600(A);
#pragma unroll
        for (int h = 0; h < kShake128Rate / 4; h++) {
            buf |= (uint64_t)(uint32_t)(A[h >> 1] >> (32 * (h & 1))) << nbits;
            nbits += 32;
            while (nbits >= w && k < count) {
                sink(k, (uint32_t)buf & mask);
                buf >>= w;
                nbits -= w;
                k++;
            }
        }
    }
}

// Squeeze `nwords` 64-bit words of output (absorbed, not yet permuted state).
__device__ __forceinline__ void squeeze_words(uint64_t (&A)[25], uint64_t *dst, int nwords)
{
    for (int k = 0; k < nwords; k += kShake128Rate / 8) {
        f1600(A);
#pragma unroll
        for (int i = 0; i < kShake128Rate / 8; i++)
            if (k + i < nwords) dst[k + i] = A[i];
    }
}

} // namespace keccak
} // namespace tmvp
