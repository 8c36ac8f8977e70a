// Batched Saber/Sable KEM on the GPU: sampling (SHAKE-128 GenMatrix and
// centred-binomial GenSecret), codecs, the Fujisaki-Okamoto hashes, and the
// host-side launch chains for key generation, encapsulation and
// decapsulation.  The products and rounding run in the MatrixVectorMul /
// InnerProd kernels (launch_mfma / launch_valu).
//
// Step order follows saber_enc / saber_dec (S/src/cuda_kernel.cu:114-186,
// :341-450) and SABLE_enc / SABLE_dec (SB/src/SABLE_indcpa.cu:133-233,
// :422-563); the decapsulation replaces the pre-key with z before the final
// hash when the re-encryption differs (Saber round 3), which the reference's
// verify_gpu / cmov_gpu (S/src/pack.cu:225-255) never do.  Key generation is
// absent from the reference and follows Saber round 3 (seed_A is hashed with
// SHAKE-128, b = round(A^T s)).
//
// Sponges run on a pair of lanes per instance (state in registers, split in
// 32-bit halves) and write their results in the form the products read: A
// (through an LDS ring), s, b of the public key and the message bits.  The
// lane pairs also pack the secret into the sk (key generation) and the
// ciphertext before hashing it (encapsulation); the other codecs run one
// thread per 64 coefficients.
#include <string.h>

#include <type_traits>
#include <utility>

#include "tmvp_internal.h"
#include "tmvp_keccak.h"
#include "tmvp_kem.h"
#include "tmvp_mfma_body.h"

namespace tmvp {

using namespace keccak;

// ---------------------------------------------------------------- sampling
//
// Every sponge runs on a pair of adjacent lanes (tmvp_keccak.h): thread g of a
// sponge kernel works on pair t = g >> 1, half p = g & 1, and both lanes of a
// pair always take the same branch.

// Each sponge kernel picks the Keccak round unroll U (tmvp_keccak.h) that
// measured faster for it at B = 16384: 2 for the encapsulation front,
// key-store GenMatrix and the hash kernels; 1 for the key-generation front and
// both decapsulation sponges (dec_pre shares the SIMDs with the decryption
// chain, and its larger unrolled body slowed that chain more than it gained).

// The compile-time shape of a parameter set's sampling: EQ (matrix bits), MU
// (CBD width) and EP (bits of b).  Each sponge kernel is instantiated per
// shape, so it holds only its own set's code paths (a kernel holding every
// set's variants measured 19 us slower on Saber encapsulation).
template <int EQ_, int MU_, int EP_> struct Shape {
    static constexpr int EQ = EQ_, MU = MU_, EP = EP_;
};

#define TMVP_PAIR(t0)                                                                                                \
    const int64_t t = (((int64_t)blockIdx.x * 64 + threadIdx.x) >> 1) + (t0);                                       \
    const int p = (int)(threadIdx.x & 1)

// GenMatrix (S/src/pack.cu:145-153 with the 13-bit unpacking of :51-65;
// SB/src/pack.cu:22-50, 11 bits): the SHAKE-128(seed) stream read as EQ-bit
// little-endian coefficients, in the order of A[i][j][k].  64 coefficients
// are exactly EQ 64-bit words of the stream, so the squeezed words go through
// a ring of three such groups in LDS (per instance; the pair's lanes write the
// low and high dwords of each word) and every group the stream has completed
// is unpacked at once, lane p taking coefficients 32p..32p+31 of it, which are
// exactly its EQ dwords.  A block of 21 words leaves at most EQ - 1 words of
// an unfinished group behind, so three groups always hold what is still
// needed.  The ring's instance stride (78 dwords) keeps the 64 lanes' reads
// and writes in distinct banks.
constexpr int kRingDwords = 6 * 13; // per instance, for EQ <= 13

// The 32 W-bit coefficients packed LSB-first in d[0..W) to dst (16-byte aligned).
template <int W> __device__ __forceinline__ void unpack32(const uint32_t (&d)[W], uint16_t *dst)
{
    constexpr uint32_t mask = (1u << W) - 1u;
    uint32_t o[16];
#pragma unroll
    for (int j = 0; j < 32; j++) {
        const int bit = W * j, i = bit >> 5, sh = bit & 31;
        const uint32_t v =
            (sh + W <= 32 ? (d[i] >> sh) : __builtin_amdgcn_alignbit(d[i + 1 < W ? i + 1 : i], d[i], sh)) & mask;
        if (j & 1) o[j >> 1] |= v << 16;
        else o[j >> 1] = v;
    }
    uint4 *q4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
    for (int q = 0; q < 4; q++) q4[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
}

// A packed W-bit vector of ncoef coefficients (ncoef a multiple of 256) at
// src, unpacked to dst by a lane pair: 64 coefficients are W 64-bit words, and
// lane p takes coefficients 32p..32p+31 of each group, its W dwords.  Four
// groups' loads are issued before their stores (stores in flight hold up the
// next loads' wait).
template <int W> __device__ __forceinline__ void unpack_pair(const uint8_t *src, uint16_t *dst, int ncoef, int p)
{
    for (int g0 = 0; g0 < ncoef / 64; g0 += 4) {
        uint32_t d[4][W];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t *s = reinterpret_cast<const uint32_t *>(src + 8 * W * (g0 + u)) + W * p;
#pragma unroll
            for (int i = 0; i < W; i++) d[u][i] = s[i];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) unpack32<W>(d[u], dst + 64 * (g0 + u) + 32 * p);
    }
}

// The vector b of a public key (its first pvc bytes, EP bits per coefficient).
template <int EP> __device__ __forceinline__ void unpack_b(const KemParams &P, const uint8_t *pk, uint16_t *dst, int p)
{
    unpack_pair<EP>(pk, dst, P.l * kN, p);
}

// The message bits as one 0/1 coefficient each (bit k of m is coefficient k):
// this lane's halves of the four words cover coefficients 64w + 32p + j.
__device__ __forceinline__ void message_bits(const uint32_t (&m)[4], uint16_t *dst, int p)
{
#pragma unroll
    for (int w = 0; w < 4; w++) {
        uint32_t o[16];
#pragma unroll
        for (int q = 0; q < 16; q++) o[q] = ((m[w] >> (2 * q)) & 1u) | (((m[w] >> (2 * q + 1)) & 1u) << 16);
        uint4 *q4 = reinterpret_cast<uint4 *>(dst + 64 * w + 32 * p);
#pragma unroll
        for (int q = 0; q < 4; q++) q4[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    }
}

// The SHAKE-128(seed) stream of nwords words, handed to f(g, d) one group of
// G words at a time: d holds this lane's G dwords of group g (dwords G p ..
// G p + G - 1 of the group's 2G).  The squeezed words go through a ring of RG
// groups in LDS (the instance's kRingDwords; the pair's lanes write the low
// and high dwords of each word), and every group the stream has completed is
// handed over at once.  A block of 21 words leaves at most G - 1 words of an
// unfinished group behind, so RG groups of at least 21 + G - 1 words always
// hold what is still needed.
template <int G, int RG, int U, class F>
__device__ __forceinline__ void stream_groups(const uint32_t (&seed)[4], int nwords, uint32_t *ring, int p, F f)
{
    constexpr int RW = RG * G; // ring words
    constexpr int RB = kShake128Rate / 8;
    static_assert(RW >= RB + G - 1 && 2 * RW <= kRingDwords, "ring size");
    uint32_t st[25];
    absorb_words(st, seed, kShake128Rate, kShakeDS, p);
    int g = 0;  // next group to hand over
    int kb = 0; // ring word of the block's first word
    for (int k = 0; k < nwords; k += RB) {
        f1600<U>(st, p);
#pragma unroll
        for (int i = 0; i < RB; i++) {
            if (k + i < nwords) {
                int w = kb + i;
                if (w >= RW) w -= RW;
                ring[2 * w + p] = st[i];
            }
        }
        kb += RB;
        if (kb >= RW) kb -= RW;
        // The pair's writes above are read by both lanes below; the wave's LDS
        // operations execute in order, so only code motion must be stopped.
        __builtin_amdgcn_wave_barrier();
        const int avail = k + RB < nwords ? k + RB : nwords;
        for (; G * (g + 1) <= avail; g++) {
            const uint32_t *src = ring + 2 * G * (g % RG) + G * p;
            uint32_t d[G];
#pragma unroll
            for (int i = 0; i < G; i++) d[i] = src[i];
            f(g, d);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// GenMatrix: 64 coefficients of EQ bits are exactly EQ words of the stream.
template <int EQ, int U>
__device__ __forceinline__ void matrix_to_a(const uint32_t (&seed)[4], uint16_t *arow, int nwords, uint32_t *ring,
                                            int p)
{
    stream_groups<EQ, 3, U>(seed, nwords, ring, p,
                         [&](int g, const uint32_t(&d)[EQ]) { unpack32<EQ>(d, arow + 64 * g + 32 * p); });
}

// A[t] from the 32-byte seed; `ring` is the instance's kRingDwords in LDS.
template <int EQ, int U>
__device__ __forceinline__ void gen_matrix(const KemParams &P, const uint32_t (&seed)[4], uint16_t *arow,
                                           uint32_t *ring, int p)
{
    matrix_to_a<EQ, U>(seed, arow, P.araw / 8, ring, p);
}

#define TMVP_RING                                                                                                    \
    __shared__ uint32_t ring_lds[32 * kRingDwords];                                                                  \
    uint32_t *const ring = ring_lds + (threadIdx.x >> 1) * kRingDwords

// CBD_mu (S/src/poly.cu:836-916, SB/src/fips202.cu:625-651): coefficient c
// is the popcount of stream bits [c mu, c mu + mu/2) minus that of the next
// mu/2 bits, stored as int16 (Saber) or masked mod q (Sable) by qmask.  For
// mu = 8 and mu = 2 every coefficient lies inside one 32-bit half of a word
// (bit i of this lane's half of word k is stream bit 64k + 32p + i) and is
// sampled straight from the squeeze; for mu = 6 and 10 the coefficients
// straddle the halves, and 64 of them are exactly mu words, so they go through
// the LDS ring.  S is the instance's row of L * 256 coefficients.

__device__ __forceinline__ void cbd8_half(uint16_t *S, int k, uint32_t h, int p, uint32_t qmask)
{
    uint32_t x = h - ((h >> 1) & 0x55555555u);
    x = (x & 0x33333333u) + ((x >> 2) & 0x33333333u); // popcount of every nibble
    uint32_t e[4];
#pragma unroll
    for (int j = 0; j < 4; j++) e[j] = (((x >> (8 * j)) & 0xfu) - ((x >> (8 * j + 4)) & 0xfu)) & qmask;
    *reinterpret_cast<uint2 *>(S + 8 * k + 4 * p) = make_uint2(e[0] | (e[1] << 16), e[2] | (e[3] << 16));
}

__device__ __forceinline__ void cbd2_half(uint16_t *S, int k, uint32_t h, int p, uint32_t qmask)
{
    const uint32_t a = h & 0x55555555u, b = (h >> 1) & 0x55555555u;
    uint32_t o[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const uint32_t e0 = (((a >> (4 * q)) & 1u) - ((b >> (4 * q)) & 1u)) & qmask;
        const uint32_t e1 = (((a >> (4 * q + 2)) & 1u) - ((b >> (4 * q + 2)) & 1u)) & qmask;
        o[q] = e0 | (e1 << 16);
    }
    uint4 *d = reinterpret_cast<uint4 *>(S + 32 * k + 16 * p);
    d[0] = make_uint4(o[0], o[1], o[2], o[3]);
    d[1] = make_uint4(o[4], o[5], o[6], o[7]);
}

// The 32 coefficients whose MU-bit fields are packed LSB-first in d[0..MU).
template <int MU> __device__ __forceinline__ void cbd32(const uint32_t (&d)[MU], uint16_t *dst, uint32_t qmask)
{
    constexpr int H = MU / 2;
    constexpr uint32_t lo = (1u << H) - 1u;
    uint32_t o[16];
#pragma unroll
    for (int j = 0; j < 32; j++) {
        const int bit = MU * j, i = bit >> 5, sh = bit & 31;
        const uint32_t x = sh + MU <= 32 ? (d[i] >> sh) : __builtin_amdgcn_alignbit(d[i + 1 < MU ? i + 1 : i], d[i], sh);
        const uint32_t v = ((uint32_t)__builtin_popcount(x & lo) - (uint32_t)__builtin_popcount((x >> H) & lo)) & qmask;
        if (j & 1) o[j >> 1] |= v << 16;
        else o[j >> 1] = v;
    }
    uint4 *q4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
    for (int q = 0; q < 4; q++) q4[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
}

// The secret of one instance from its 32-byte seed, sampled into S.
template <int MU, int U>
__device__ __forceinline__ void secret_stream(const KemParams &P, const uint32_t (&seed)[4], uint16_t *S,
                                              uint32_t *ring, int p)
{
    const int nwords = P.sraw / 8;
    const uint32_t qmask = P.sable ? (1u << P.eq) - 1u : 0xffffu;
    if constexpr (MU == 8 || MU == 2) {
        uint32_t A[25];
        absorb_words(A, seed, kShake128Rate, kShakeDS, p);
        if constexpr (MU == 8) squeeze_each<U>(A, nwords, p, [&](int k, uint32_t h) { cbd8_half(S, k, h, p, qmask); });
        else squeeze_each<U>(A, nwords, p, [&](int k, uint32_t h) { cbd2_half(S, k, h, p, qmask); });
    } else {
        static_assert(MU == 6 || MU == 10, "CBD width");
        constexpr int RG = MU == 6 ? 5 : 3; // groups in the ring (at least 21 + MU - 1 words)
        stream_groups<MU, RG, U>(seed, nwords, ring, p,
                              [&](int g, const uint32_t(&d)[MU]) { cbd32<MU>(d, S + 64 * g + 32 * p, qmask); });
    }
}

// One instance's polynomials packed by its lane pair (the codecs of
// S/src/pack.cu:1-213 and SB/src/pack.cu:182-467): chunk c of 64 coefficients
// is W 64-bit words of the LSB-first W-bit code, and lane p's 32 coefficients
// 32p..32p+31 of it are exactly dwords pW..pW+W-1.  SK2: the Sable secret
// code s & 3 (SB/src/pack.cu:366-377).
//
// Callers order these accesses against the pair's other ones with a
// workgroup-scope fence (the workgroup is this one wave): one after the stores
// that wrote the source, one before loads of what was packed.  An agent-scope
// fence would also write back the L2, which made key generation 24 us slower.
__device__ __forceinline__ void wave_fence_global()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <int W, bool SK2>
__device__ __forceinline__ void pack_half(const uint16_t *S, uint8_t *dst, int nchunks, int p, uint8_t *dst2 = nullptr)
{
    constexpr uint32_t mask = SK2 ? 3u : (1u << W) - 1u;
    constexpr int G = 4; // chunks whose loads are in flight together (nchunks is a multiple of 4)
    for (int c0 = 0; c0 < nchunks; c0 += G) {
        uint4 v[G][4];
#pragma unroll
        for (int g = 0; g < G; g++) {
            const uint4 *src = reinterpret_cast<const uint4 *>(S + 64 * (c0 + g) + 32 * p);
#pragma unroll
            for (int q = 0; q < 4; q++) v[g][q] = src[q];
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            uint32_t e[32];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t h[4] = {v[g][q].x, v[g][q].y, v[g][q].z, v[g][q].w};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    e[8 * q + 2 * k] = h[k] & mask;
                    e[8 * q + 2 * k + 1] = (h[k] >> 16) & mask;
                }
            }
            uint32_t w[W];
#pragma unroll
            for (int i = 0; i < W; i++) w[i] = 0;
#pragma unroll
            for (int j = 0; j < 32; j++) {
                const int bit = W * j, i = bit >> 5, o = bit & 31;
                w[i] |= e[j] << o;
                if (o + W > 32) w[i + 1] |= e[j] >> (32 - o);
            }
            uint32_t *d = reinterpret_cast<uint32_t *>(dst + (int64_t)(c0 + g) * W * 8) + p * W;
#pragma unroll
            for (int i = 0; i < W; i++) d[i] = w[i];
            if (dst2) {
                uint32_t *d2 = reinterpret_cast<uint32_t *>(dst2 + (int64_t)(c0 + g) * W * 8) + p * W;
#pragma unroll
                for (int i = 0; i < W; i++) d2[i] = w[i];
            }
        }
    }
}

// pack_half's code of one instance's polynomials, compared with the bytes at
// dst instead of written there: nonzero when they differ (this lane's half).
template <int W>
__device__ __forceinline__ uint32_t pack_compare_half(const uint16_t *S, const uint8_t *dst, int nchunks, int p)
{
    constexpr uint32_t mask = (1u << W) - 1u;
    uint32_t x = 0;
    for (int c = 0; c < nchunks; c++) {
        const uint4 *src = reinterpret_cast<const uint4 *>(S + 64 * c + 32 * p);
        const uint32_t *d = reinterpret_cast<const uint32_t *>(dst + (int64_t)c * W * 8) + p * W;
        uint4 v[4];
        uint32_t have[W];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = src[q];
#pragma unroll
        for (int i = 0; i < W; i++) have[i] = d[i];
        uint32_t w[W];
#pragma unroll
        for (int i = 0; i < W; i++) w[i] = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t h[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
#pragma unroll
                for (int e2 = 0; e2 < 2; e2++) {
                    const int j = 8 * q + 2 * kk + e2;
                    const uint32_t e = (e2 ? h[kk] >> 16 : h[kk]) & mask;
                    const int bit = W * j, i = bit >> 5, o = bit & 31;
                    w[i] |= e << o;
                    if (o + W > 32) w[i + 1] |= e >> (32 - o);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < W; i++) x |= w[i] ^ have[i];
    }
    return x;
}

// G(m | H(pk)) -> buf = m | H(pk), kr; returns the coins kr[32:64]; the
// shared first step of encapsulation and of the re-encryption in
// decapsulation (S/src/cuda_kernel.cu:116-118, :378-379).
template <int U>
__device__ __forceinline__ void hash_g(const uint32_t (&m)[4], const uint32_t (&hpk)[4], uint8_t *buf, uint8_t *kr,
                                       uint32_t (&coins)[4], int p)
{
    uint32_t b[8], k[8];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        b[i] = m[i];
        b[4 + i] = hpk[i];
    }
    sha3_512_words<U>(k, b, p);
    store_words(buf, b, p);
    store_words(kr, k, p);
#pragma unroll
    for (int i = 0; i < 4; i++) coins[i] = k[4 + i];
}

// The front kernels run the long matrix sponge (23 permutations for Saber;
// pairs [0, B)) beside the rest of the instance's hashing and its secret
// stream (pairs [B, 2B)), one wave per SIMD at B = 16384: a sponge wave alone
// is latency-bound, and separate launches of the two roles share SIMDs and
// slow the matrix sponge down.  A kernel starting at pair t0 = B skips the
// matrix.

// Key generation: seeds[t] = seed_A | seed_s | z.  Matrix role: seed_A' =
// SHAKE-128(seed_A) goes to the pk tail (and the sk's copy of it) and seeds
// GenMatrix.  Secret role: seed_s seeds the secret, which goes to S and,
// packed, to the head of the sk; z goes to the sk tail.
template <class SH, int U>
__device__ __forceinline__ void kp_matrix_role(const KemParams &P, const uint8_t *seeds, uint8_t *pk, uint8_t *sk,
                                               uint16_t *A, int64_t t, uint32_t *ring, int p)
{
    uint32_t m[4], st[25];
    load_words(m, seeds + t * 96, p);
    absorb_words(st, m, kShake128Rate, kShakeDS, p);
    f1600<U>(st, p);
    const uint32_t sa[4] = {st[0], st[1], st[2], st[3]};
    store_words(pk + t * P.pk + P.pvc, sa, p);
    store_words(sk + t * P.sk + P.indcpa_sk + P.pvc, sa, p);
    gen_matrix<SH::EQ, U>(P, sa, A + t * P.l * P.l * kN, ring, p);
}

template <class SH, int U>
__device__ __forceinline__ void kp_secret_role(const KemParams &P, const uint8_t *seeds, uint8_t *sk, uint16_t *S,
                                               int64_t t, uint32_t *ring, int p)
{
    uint32_t m[4];
    load_words(m, seeds + t * 96 + 32, p);
    secret_stream<SH::MU, U>(P, m, S + t * P.l * kN, ring, p);
    wave_fence_global(); // the sampling wrote S with the pair's lanes interleaved
    if constexpr (SH::MU == 2) pack_half<2, true>(S + t * P.l * kN, sk + t * P.sk, 4 * P.l, p);
    else pack_half<SH::EQ, false>(S + t * P.l * kN, sk + t * P.sk, 4 * P.l, p);
    load_words(m, seeds + t * 96 + 64, p);
    store_words(sk + t * P.sk + P.sk - 32, m, p);
}

// Encapsulation.  Matrix role: GenMatrix from the pk tail.  Secret role:
// m = H(rnd), kr = G(m | H(pk)), the secret from kr[32:64], and the InnerProd
// operands: b of the pk to V and m's bits to bits.  hpk (if not null) holds
// H(pk) per instance.
template <class SH, int U>
__device__ __forceinline__ void enc_matrix_role(const KemParams &P, const uint8_t *pk, uint16_t *A, int64_t t,
                                                uint32_t *ring, int p)
{
    uint32_t m[4];
    load_words(m, pk + t * P.pk + P.pvc, p);
    gen_matrix<SH::EQ, U>(P, m, A + t * P.l * P.l * kN, ring, p);
}

template <class SH, int U>
__device__ __forceinline__ void enc_secret_role(const KemParams &P, const uint8_t *rnd, const uint8_t *pk,
                                                const uint8_t *hpk, uint8_t *buf, uint8_t *kr, uint16_t *S,
                                                uint16_t *V, uint16_t *bits, int64_t t, int64_t kt, uint32_t *ring,
                                                int p)
{
    // t: the instance; kt: its row of pk / hpk (a key id when they are a key store)
    uint32_t r[4], m[4], h[4], coins[4];
    load_words(r, rnd + t * 32, p);
    sha3_256_words<4, U>(m, r, p);
    if (hpk) load_words(h, hpk + kt * 32, p);
    else sha3_256<U>(h, pk + kt * P.pk, P.pk, p);
    hash_g<U>(m, h, buf + t * 64, kr + t * 64, coins, p);
    secret_stream<SH::MU, U>(P, coins, S + t * P.l * kN, ring, p);
    message_bits(m, bits + t * kN, p);
    unpack_b<SH::EP>(P, pk + kt * P.pk, V + t * P.l * kN, p);
}

template <class SH>
__global__ __launch_bounds__(64) void kp_front_kernel(KemParams P, const uint8_t *seeds, uint8_t *pk, uint8_t *sk,
                                                      uint16_t *A, uint16_t *S, int64_t B)
{
    TMVP_PAIR(0);
    TMVP_RING;
    if (t >= 2 * B) return;
    if (t < B) kp_matrix_role<SH, 1>(P, seeds, pk, sk, A, t, ring, p);
    else kp_secret_role<SH, 1>(P, seeds, sk, S, t - B, ring, p);
}

// KEYED (A resident, t0 = B): pk and hpk are key stores, read at keys[i].  A
// template flag rather than a null test of keys: with the runtime select the
// non-keyed Saber kernel measured 176 us at B = 16384 against 159 us without.
template <class SH, bool KEYED>
__global__ __launch_bounds__(64) void enc_front_kernel(KemParams P, const uint8_t *rnd, const uint8_t *pk,
                                                       const uint8_t *hpk, uint8_t *buf, uint8_t *kr, uint16_t *A,
                                                       uint16_t *S, uint16_t *V, uint16_t *bits,
                                                       int64_t B, int64_t t0, const int32_t *keys)
{
    TMVP_PAIR(t0);
    TMVP_RING;
    if (t >= 2 * B) return;
    if (t < B) {
        enc_matrix_role<SH, 2>(P, pk, A, t, ring, p);
    } else {
        const int64_t i = t - B;
        enc_secret_role<SH, 2>(P, rnd, pk, hpk, buf, kr, S, V, bits, i, KEYED ? (int64_t)keys[i] : i, ring, p);
    }
}

// Decapsulation, before the decryption and beside it (on the auxiliary
// stream): the matrix stream from the pk inside the sk (pairs [0, B)) and
// H(ct) (pairs [B, 2B)).
template <class SH>
__global__ __launch_bounds__(64) void dec_pre_kernel(KemParams P, const uint8_t *ct, const uint8_t *sk, uint16_t *A,
                                                     uint8_t *hct, int64_t B)
{
    TMVP_PAIR(0);
    TMVP_RING;
    if (t >= 2 * B) return;
    if (t < B) {
        uint32_t m[4];
        load_words(m, sk + t * P.sk + P.indcpa_sk + P.pvc, p);
        gen_matrix<SH::EQ, 1>(P, m, A + t * P.l * P.l * kN, ring, p);
    } else {
        const int64_t i = t - B;
        uint32_t h[4];
        sha3_256<1>(h, ct + i * P.ct, P.ct, p);
        store_words(hct + i * 32, h, p);
    }
}

// Decapsulation, after the decryption: m' from its bits, kr = G(m' | H(pk))
// with H(pk) from the sk, secret from kr[32:64], b of the pk inside the sk to
// V for the re-encryption, and the instance's mismatch flag cleared (pairs
// [0, B)).  With hct set, pairs [B, 2B) compute H(ct) beside it (the
// decapsulation of keys whose A is resident has no matrix sponge to run it
// next to).  With skey set, instance t's sk is row skey[t] of sk.
template <class SH>
__global__ __launch_bounds__(64) void dec_front_kernel(KemParams P, const uint16_t *mbits, const uint8_t *sk,
                                                       uint8_t *buf, uint8_t *kr, uint16_t *S, uint16_t *V,
                                                       uint32_t *diff, int64_t B, const uint8_t *ct, uint8_t *hct,
                                                       const int32_t *skey)
{
    TMVP_PAIR(0);
    TMVP_RING;
    if (t >= (hct ? 2 * B : B)) return;
    if (t >= B) {
        const int64_t i = t - B;
        uint32_t h[4];
        sha3_256<1>(h, ct + i * P.ct, P.ct, p);
        store_words(hct + i * 32, h, p);
        return;
    }
    if (p == 0) diff[t] = 0; // set by the re-encryption's compare-packs below
    uint32_t m[4];
    // this lane's 32 bits of each message word: bits 64w + 32p + j
    const uint32_t *mb = reinterpret_cast<const uint32_t *>(mbits + t * kN + 32 * p); // bit k in uint16 k
#pragma unroll
    for (int w = 0; w < 4; w++) {
        uint32_t x = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint32_t two = mb[32 * w + j];
            x |= ((two & 1u) << (2 * j)) | (((two >> 16) & 1u) << (2 * j + 1));
        }
        m[w] = x;
    }
    uint32_t h[4], coins[4];
    const uint8_t *skt = sk + (skey ? (int64_t)skey[t] : t) * P.sk;
    load_words(h, skt + P.sk - 64, p);
    hash_g<1>(m, h, buf + t * 64, kr + t * 64, coins, p);
    secret_stream<SH::MU, 1>(P, coins, S + t * P.l * kN, ring, p);
    unpack_b<SH::EP>(P, skt + P.indcpa_sk, V + t * P.l * kN, p);
}

// ---------------------------------------------------------------- codecs

// One thread per (instance, chunk of 64 coefficients), which is exactly W
// 64-bit words of a W-bit little-endian bit stream (LSB first, as every
// reference codec: S/src/pack.cu:1-213, SB/src/pack.cu:182-467).  W is a
// template argument so every shift and word index is a constant.
// kXfSk2 is the Sable secret-key code on unpacking, ((x ^ 2) - 2)
// (SB/src/pack.cu:366-377).
enum Xf : int { kXfNone = 0, kXfSk2 = 1 };

// One chunk: 64 coefficients at s -> W 64-bit words.
template <int W> __device__ __forceinline__ void pack_chunk(const uint4 *s, uint64_t (&words)[W])
{
    constexpr uint32_t mask = (1u << W) - 1u;
#pragma unroll
    for (int k = 0; k < W; k++) words[k] = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const uint4 v = s[q];
        const uint32_t e[8] = {v.x & 0xffffu, v.x >> 16, v.y & 0xffffu, v.y >> 16,
                               v.z & 0xffffu, v.z >> 16, v.w & 0xffffu, v.w >> 16};
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int bit = (8 * q + j) * W, wi = bit >> 6, o = bit & 63;
            const uint64_t x = e[j] & mask;
            words[wi] |= x << o;
            if (o + W > 64) words[wi + 1] |= x >> (64 - o);
        }
    }
}

// With diff set, nothing is written: the packed words are compared with those
// already at dst, and a mismatch sets diff[t] to 1 (decapsulation's check of
// the re-encryption against the ciphertext).
template <int W>
__global__ __launch_bounds__(256) void pack_kernel(const uint16_t *src, int64_t src_stride, uint8_t *dst,
                                                   int64_t dst_stride, int dst_off, int chunks, int64_t B,
                                                   uint32_t *diff)
{
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= B * chunks) return;
    const int64_t t = gid / chunks;
    const int c = (int)(gid - t * chunks);
    uint64_t words[W];
    pack_chunk<W>(reinterpret_cast<const uint4 *>(src + t * src_stride + c * 64), words);
    uint64_t *d = reinterpret_cast<uint64_t *>(dst + t * dst_stride + dst_off + (int64_t)c * W * 8);
    if (diff) {
        uint64_t x = 0;
#pragma unroll
        for (int k = 0; k < W; k++) x |= d[k] ^ words[k];
        if (x) atomicOr(diff + t, 1u);
        return;
    }
#pragma unroll
    for (int k = 0; k < W; k++) d[k] = words[k];
}

// Decapsulation's check of the whole re-encryption in one launch: 4L chunks
// of b' (bp, EP bits) against ct[0, pvc) and 4 chunks of v' (vp, ET bits)
// against ct[pvc, ...); a mismatch sets diff[t] (S/src/cuda_kernel.cu:414,
// :442, :445).
template <int EP, int ET>
__global__ __launch_bounds__(256) void dec_compare_kernel(const uint16_t *bp, const uint16_t *vp, const uint8_t *ct,
                                                          int64_t ct_stride, int pvc, uint32_t *diff, int l, int64_t B)
{
    const int cv = 4 * l;
    const int chunks = cv + 4;
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= B * chunks) return;
    const int64_t t = gid / chunks;
    const int c = (int)(gid - t * chunks);
    uint64_t x = 0;
    if (c < cv) {
        uint64_t w[EP];
        pack_chunk<EP>(reinterpret_cast<const uint4 *>(bp + (t * cv + c) * 64), w);
        const uint64_t *d = reinterpret_cast<const uint64_t *>(ct + t * ct_stride + (int64_t)c * EP * 8);
#pragma unroll
        for (int k = 0; k < EP; k++) x |= d[k] ^ w[k];
    } else {
        uint64_t w[ET];
        pack_chunk<ET>(reinterpret_cast<const uint4 *>(vp + t * kN + (c - cv) * 64), w);
        const uint64_t *d = reinterpret_cast<const uint64_t *>(ct + t * ct_stride + pvc + (int64_t)(c - cv) * ET * 8);
#pragma unroll
        for (int k = 0; k < ET; k++) x |= d[k] ^ w[k];
    }
    if (x) atomicOr(diff + t, 1u);
}

// One chunk: W 64-bit words at s -> 64 coefficients at d.
template <int W, int XF> __device__ __forceinline__ void unpack_chunk(const uint64_t *s, uint4 *d)
{
    constexpr uint32_t mask = (1u << W) - 1u;
    uint64_t words[W];
#pragma unroll
    for (int k = 0; k < W; k++) words[k] = s[k];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        uint32_t e[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int bit = (8 * q + j) * W, wi = bit >> 6, o = bit & 63;
            uint64_t x = words[wi] >> o;
            if (o + W > 64) x |= words[wi + 1] << (64 - o);
            uint32_t v = (uint32_t)x & mask;
            if (XF == kXfSk2) v = ((v ^ 2u) - 2u) & 0xffffu;
            e[j] = v;
        }
        d[q] = make_uint4(e[0] | (e[1] << 16), e[2] | (e[3] << 16), e[4] | (e[5] << 16), e[6] | (e[7] << 16));
    }
}

template <int W, int XF>
__global__ __launch_bounds__(256) void unpack_kernel(const uint8_t *src, int64_t src_stride, int src_off,
                                                     uint16_t *dst, int64_t dst_stride, int chunks, int64_t B)
{
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= B * chunks) return;
    const int64_t t = gid / chunks;
    const int c = (int)(gid - t * chunks);
    unpack_chunk<W, XF>(reinterpret_cast<const uint64_t *>(src + t * src_stride + src_off + (int64_t)c * W * 8),
                        reinterpret_cast<uint4 *>(dst + t * dst_stride + c * 64));
}

// Decapsulation's three decodes in one launch (S/src/cuda_kernel.cu:341-352;
// SB/src/SABLE_indcpa.cu:422-440): per instance 4L chunks of s from the sk (WS
// bits, or the Sable 2-bit code), 4L chunks of b' (WB bits) and 4 chunks of
// c_m (WC bits) from the ct.  Used where nothing runs beside the decryption
// (A resident); beside the matrix sponge, three launches measured faster.
// With skey set, instance t's sk is row skey[t] of sk (a key store).
template <int WS, int XS, int WB, int WC>
__global__ __launch_bounds__(256) void dec_unpack_kernel(const uint8_t *sk, int64_t sk_stride, const uint8_t *ct,
                                                         int64_t ct_stride, int pvc, uint16_t *S, uint16_t *Bp,
                                                         uint16_t *Cm, int l, int64_t B, const int32_t *skey)
{
    const int cv = 4 * l; // chunks of one polynomial vector
    const int chunks = 2 * cv + 4;
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= B * chunks) return;
    const int64_t t = gid / chunks;
    const int c = (int)(gid - t * chunks);
    if (c < cv)
        unpack_chunk<WS, XS>(reinterpret_cast<const uint64_t *>(sk + (skey ? (int64_t)skey[t] : t) * sk_stride +
                                                                (int64_t)c * WS * 8),
                             reinterpret_cast<uint4 *>(S + (t * cv + c) * 64));
    else if (c < 2 * cv)
        unpack_chunk<WB, kXfNone>(reinterpret_cast<const uint64_t *>(ct + t * ct_stride + (int64_t)(c - cv) * WB * 8),
                                  reinterpret_cast<uint4 *>(Bp + (t * cv + c - cv) * 64));
    else
        unpack_chunk<WC, kXfNone>(
            reinterpret_cast<const uint64_t *>(ct + t * ct_stride + pvc + (int64_t)(c - 2 * cv) * WC * 8),
            reinterpret_cast<uint4 *>(Cm + t * kN + (c - 2 * cv) * 64));
}

// Up to three gathers of one batch in one launch: segment g copies words[g]
// 8-byte words per instance, dst[g] + t * dst_stride[g] <- src[g] +
// keys[t] * src_stride[g] (key ids checked by the caller).  Unused segments
// have zero words.
struct GatherSet {
    const uint8_t *src[3];
    int64_t src_stride[3];
    uint8_t *dst[3];
    int64_t dst_stride[3];
    int words[3];
};

__global__ __launch_bounds__(256) void gather_kernel(GatherSet G, const int32_t *keys, int64_t B)
{
    const int total = G.words[0] + G.words[1] + G.words[2];
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= B * total) return;
    const int64_t t = gid / total;
    int k = (int)(gid - t * total);
    const uint8_t *src;
    uint8_t *dst;
    const int64_t key = keys[t];
    if (k < G.words[0]) {
        src = G.src[0] + key * G.src_stride[0];
        dst = G.dst[0] + t * G.dst_stride[0];
    } else if (k < G.words[0] + G.words[1]) {
        k -= G.words[0];
        src = G.src[1] + key * G.src_stride[1];
        dst = G.dst[1] + t * G.dst_stride[1];
    } else {
        k -= G.words[0] + G.words[1];
        src = G.src[2] + key * G.src_stride[2];
        dst = G.dst[2] + t * G.dst_stride[2];
    }
    reinterpret_cast<uint64_t *>(dst)[k] = reinterpret_cast<const uint64_t *>(src)[k];
}

// ---------------------------------------------------------------- hashing steps (one lane pair per instance)

// H(pk) of each public key (for a key store).
__global__ __launch_bounds__(64) void hash_pk_kernel(KemParams P, const uint8_t *pk, uint8_t *hpk, int64_t B)
{
    TMVP_PAIR(0);
    if (t >= B) return;
    uint32_t h[4];
    sha3_256<2>(h, pk + t * P.pk, P.pk, p);
    store_words(hpk + t * 32, h, p);
}

// GenMatrix from the 32-byte seed at seed + t * stride (for a key store).
template <class SH>
__global__ __launch_bounds__(64) void matrix_kernel(KemParams P, const uint8_t *seed, int64_t stride, uint16_t *A,
                                                    int64_t B)
{
    TMVP_PAIR(0);
    TMVP_RING;
    if (t >= B) return;
    uint32_t m[4];
    load_words(m, seed + t * stride, p);
    gen_matrix<SH::EQ, 2>(P, m, A + t * P.l * P.l * kN, ring, p);
}

// Key generation, last step: H(pk) into the sk.
// Key generation, last step: b (bp, EP bits) packed into the pk and the sk's
// copy of it, then H(pk) into the sk.
template <int EP>
__global__ __launch_bounds__(64) void kp_hash_kernel(KemParams P, uint8_t *pk, uint8_t *sk, int64_t B,
                                                     const uint16_t *bp)
{
    TMVP_PAIR(0);
    if (t >= B) return;
    pack_half<EP, false>(bp + t * P.l * kN, pk + t * P.pk, 4 * P.l, p, sk + t * P.sk + P.indcpa_sk);
    wave_fence_global(); // H(pk) reads both lanes' words
    uint32_t h[4];
    sha3_256_prefetch<2>(h, pk + t * P.pk, P.pk, p);
    store_words(sk + t * P.sk + P.sk - 64, h, p);
}

// Encapsulation, last step: with EP > 0 the ciphertext is first packed from
// b' (bp, EP bits) and v' (vp, ET bits) (S/src/cuda_kernel.cu:133, :178);
// then kr[32:] = H(ct), ss = H(kr) (:185-186).
template <int EP, int ET>
__global__ __launch_bounds__(64) void enc_hash2_kernel(KemParams P, uint8_t *ct, uint8_t *kr, uint8_t *ss, int64_t B,
                                                       const uint16_t *bp, const uint16_t *vp)
{
    TMVP_PAIR(0);
    if (t >= B) return;
    if constexpr (EP > 0) {
        pack_half<EP, false>(bp + t * P.l * kN, ct + t * P.ct, 4 * P.l, p);
        pack_half<ET, false>(vp + t * kN, ct + t * P.ct + P.pvc, 4, p);
        wave_fence_global(); // H(ct) reads both lanes' words
    }
    uint32_t h[4], k[8], s[4], k0[4];
    sha3_256_prefetch<2>(h, ct + t * P.ct, P.ct, p);
    load_words(k0, kr + t * 64, p);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        k[i] = k0[i];
        k[4 + i] = h[i];
    }
    sha3_256_words<8, 2>(s, k, p);
    store_words(ss + t * 32, s, p);
}

// Decapsulation, last step: diff[t] is the outcome of comparing ct with the
// re-encryption (set while packing it), kr[32:] = H(ct) (computed
// beforehand); replace kr[0:32] by z on mismatch, ss = H(kr).  fail[t]
// records the outcome.  KEYED: instance t's sk is row skey[t] of sk (a template
// flag, as in enc_front_kernel: the runtime test cost the non-keyed kernel
// 0.7 us at 16384 instances).
template <bool KEYED>
__global__ __launch_bounds__(64) void dec_fin_kernel(KemParams P, const uint32_t *diff, const uint8_t *sk,
                                                     const uint8_t *hct, const uint8_t *kr, uint8_t *ss, uint8_t *fail,
                                                     int64_t B, const int32_t *skey)
{
    TMVP_PAIR(0);
    if (t >= B) return;
    const uint32_t mask = 0u - (uint32_t)(diff[t] != 0); // all ones on mismatch, in both lanes
    uint32_t h[4], k[8], s[4], k0[4], z[4];
    load_words(h, hct + t * 32, p);
    load_words(k0, kr + t * 64, p);
    load_words(z, sk + (KEYED ? (int64_t)skey[t] : t) * P.sk + P.sk - 32, p);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        k[i] = k0[i] ^ (mask & (k0[i] ^ z[i]));
        k[4 + i] = h[i];
    }
    sha3_256_words<8, 1>(s, k, p);
    store_words(ss + t * 32, s, p);
    if (fail && p == 0) fail[t] = (uint8_t)(mask & 1u);
}

// ---------------------------------------------------------------- one-launch key generation and encapsulation
//
// The MFMA backend runs key generation and encapsulation as ONE launch in
// which the products and the final hash overlap the matrix sponge.  A
// workgroup of 16 waves takes 64 instances (the waves go round-robin over
// the CU's four SIMDs, so each SIMD holds one sponge wave and three MFMA
// waves):
//   waves 0, 1   the matrix sponges (GenMatrix) of instances 0-31 and 32-63;
//   waves 2, 3   the rest of those instances' hashing: the secret, and at the
//                end the hash over the packed output (H(pk) / H(ct)),
//                absorbed row by row as the products finish the rows;
//   waves 4-15   the products on the matrix cores, one output row at a time
//                (mfma_body as an InnerProd of the row's terms with the
//                secret), each as soon as its polynomials and the secret
//                exist.  In encapsulation waves 6 and 7, which share their
//                SIMDs with the secret waves rather than the matrix sponges,
//                first hash the public keys.
// A row's polynomials come out of the matrix stream in order (encapsulation,
// A s': row i is A[i][0..L); key generation, A^T s: row j is A[0..L)[j], so
// the rows finish with the stream's last L polynomials), and the hash over
// the output starts with the first finished row while the stream still runs.
// Saber, 16384 instances (one workgroup per CU): key generation 224 us and
// encapsulation 225 us, against 254 and 261 us for the three-launch chains
// below (same box, alternating runs).  With four MFMA waves per workgroup the
// products, latency-bound at one MFMA wave per SIMD, became the critical
// path (each row of 64 instances took 22-48 us), and the launch was slower
// than the chains.
//
// The matrix still goes through memory: a workgroup's share is 64 L^2 x 512 B
// (288 KB for Saber) against 160 KB of LDS, and in encapsulation most of it
// exists before the secret that multiplies it.  It is written in the layout
// the row products read, [row][instance][term][256], and read back from the
// L2 right after.
//
// Progress goes through LDS counters, each raised by one wave or counted up
// by the MFMA waves (wait_for / publish / fused_leave).  Every wait is on a
// counter raised by waves that do not themselves wait on the waiter, and
// every wave raises its counters to their final values even when it has no
// instances, so every wait ends; a bound on each wait (about 0.2 s) sets the
// status word instead of hanging if that ever failed.
constexpr int kFusedInst = 64;
constexpr int kFusedMfma = 12;                        // MFMA waves per workgroup
constexpr int kFusedThreads = 64 * (4 + kFusedMfma); // four sponge waves, then the MFMA waves
constexpr int kSpinLimit = 1 << 22;
constexpr int kDone = 1 << 20; // a counter's value when its wave has no instances

constexpr int kFusedJobs = 8; // product jobs: at most 2 L (key generation's heads and tails)
constexpr int kFusedGrab = 2; // instances an MFMA wave takes from a job at a time

struct FusedSync {
    int mat[2];            // polynomials of A written by matrix wave w (published at the milestones below)
    int sec[2];            // secret wave w: 1 once the secret (and in encapsulation b and m's bits) is written
    int hpk[2];            // encapsulation: 1 once H(pk) of group w's instances is in LDS
    int next[kFusedJobs];  // next unclaimed instance of each product job
    int done[kFusedJobs];  // MFMA waves that have left each job (all its instances finished at kFusedMfma)
    int unp;               // decapsulation: MFMA waves that have decoded their share of the inputs
};

__device__ __forceinline__ void wait_for(const int *f, int v, int *status)
{
    int n = 0;
    while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < v) {
        __builtin_amdgcn_s_sleep(2);
        if (++n == kSpinLimit) {
            *status = 1;
            break;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ void publish(int *f, int v)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Stream polynomials written before a product job may start.  Key
// generation: the heads of every row (terms i < L - 1, A[i][r]) need rows
// 0..L-2 of the stream, (L - 1) L polynomials; row r's tail (A[L-1][r]) needs
// (L - 1) L + r + 1.  Encapsulation: row i needs (i + 1) L.
template <int L> __device__ __forceinline__ constexpr bool milestone(bool keygen, int done)
{
    return keygen ? done >= (L - 1) * L : done % L == 0;
}

// GenMatrix of instance t into A in row-product layout; publishes `flag` at
// every polynomial count some row waits for.
template <class SH, int L, bool KEYGEN, int U>
__device__ __forceinline__ void fused_matrix(const KemParams &P, const uint32_t (&seed)[4], uint16_t *A, int64_t B,
                                             int64_t t, uint32_t *ring, int p, int *flag)
{
    stream_groups<SH::EQ, 3, U>(seed, P.araw / 8, ring, p, [&](int g, const uint32_t(&d)[SH::EQ]) {
        const int pi = g >> 2, q = g & 3, i = pi / L, j = pi % L; // stream polynomial pi is A[i][j]
        const int row = KEYGEN ? j : i, term = KEYGEN ? i : j;
        unpack32<SH::EQ>(d, A + (((int64_t)row * B + t) * L + term) * kN + 64 * q + 32 * p);
        if (q == 3 && milestone<L>(KEYGEN, pi + 1)) publish(flag, pi + 1);
    });
}

// Product job `job`: a product of LT terms for each of the workgroup's
// instances, one output row (J's pointers are already at its first
// instance; STRIDED: a sub-range of each instance's terms).  The MFMA waves
// claim kFusedGrab instances at a time (fused_claim), so waves sharing a SIMD
// with a busy sponge take fewer; a wave that finds nothing left counts itself
// out of the job (fused_leave), and the job is finished when all kFusedMfma
// have.
template <int LT, int EPI, bool STRIDED, bool PRE = false, bool QUAD = true>
__device__ __forceinline__ bool fused_claim(FusedSync &sy, int job, int nloc, Job &J, uint8_t *slices)
{
    int i = 0;
    if ((threadIdx.x & 63) == 0)
        i = __hip_atomic_fetch_add(&sy.next[job], kFusedGrab, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    i = __builtin_amdgcn_readfirstlane(i);
    if (i >= nloc) return false;
    J.batch = i + kFusedGrab < nloc ? i + kFusedGrab : nloc;
    static_assert(kFusedGrab <= 2, "PRE loads a claim whole");
    mfma_body<LT, kInner, EPI, 16, STRIDED, PRE, QUAD>(J, i, 1, slices);
    return true;
}

__device__ __forceinline__ void fused_leave(FusedSync &sy, int job)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(&sy.done[job], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// J with the pointers of one row job; astride / sstride > 0 select a
// sub-range of each instance's L terms (the strided form).
__device__ __forceinline__ Job row_job_of(Job J, const uint16_t *a, const uint16_t *s, uint16_t *out,
                                          const uint16_t *aux, int w, int epi, int64_t astride = 0, int64_t sstride = 0)
{
    J.a = a;
    J.s = s;
    J.out = out;
    J.aux = aux;
    J.w = w;
    J.epi = epi;
    J.astride = astride;
    J.sstride = sstride;
    return J;
}

template <int LT, int EPI, bool STRIDED, bool PRE = false, bool QUAD = true>
__device__ __forceinline__ void fused_row(FusedSync &sy, int job, int nloc, Job J, uint8_t *slices)
{
    while (fused_claim<LT, EPI, STRIDED, PRE, QUAD>(sy, job, nloc, J, slices)) {
    }
    fused_leave(sy, job);
}

// A key-generation row (a column of A) whose last term arrives late is
// split: the head (terms 0..L-2) sums into the row's output mod 2^eq as soon
// as its polynomials exist, and the tail (term L-1) adds the last product and
// rounds (kEpiQ2PAcc), so only one term's work is left once the stream has
// written the column.  A tail reads what its head wrote, so it starts only
// when every wave has left the head.
template <int L>
__device__ __forceinline__ Job tail_job_of(const Job &J, const uint16_t *a, const uint16_t *s, uint16_t *out, int eq)
{
    return row_job_of(J, a + (L - 1) * kN, s + (L - 1) * kN, out, nullptr, eq, kEpiQ2PAcc, L * kN, L * kN);
}

template <int L> struct FusedLds {
    FusedSync sync;
    uint32_t hpk[kFusedInst * 8];                            // H(pk) per local instance (encapsulation)
    uint32_t ring[4][32 * kRingDwords];                      // matrix waves 0, 1; secret waves 2, 3 (MU = 6, 10)
    __attribute__((aligned(16))) uint8_t mfma[kFusedMfma][2 * Slice<L, kInner>::BYTES]; // MFMA waves' slices
};

template <class SH, int L>
__global__ __launch_bounds__(kFusedThreads) void kp_fused_kernel(KemParams P, const uint8_t *seeds, uint8_t *pk,
                                                                 uint8_t *sk, KemWorkspace W, int64_t B, int *status)
{
    constexpr int EP = SH::EP;
    __shared__ FusedLds<L> S;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t base = (int64_t)blockIdx.x * kFusedInst;
    const int nloc = (int)(B - base < kFusedInst ? B - base : kFusedInst);
    if (threadIdx.x < sizeof(FusedSync) / 4) reinterpret_cast<int *>(&S.sync)[threadIdx.x] = 0;
    __syncthreads();
    const int p = lane & 1;
    if (wave < 4) {
        const int grp = wave & 1, li = 32 * grp + (lane >> 1);
        const int64_t t = base + li;
        uint32_t *ring = S.ring[wave] + (lane >> 1) * kRingDwords;
        if (32 * grp >= nloc) {
            publish(wave < 2 ? &S.sync.mat[grp] : &S.sync.sec[grp], kDone);
            return;
        }
        if (li >= nloc) return; // lane 0 of the wave has an instance and publishes
        if (wave < 2) {
            // seed_A' = SHAKE-128(seed_A) to the pk tail and the sk's copy of the pk, then GenMatrix
            uint32_t m[4], st[25];
            load_words(m, seeds + t * 96, p);
            absorb_words(st, m, kShake128Rate, kShakeDS, p);
            f1600<1>(st, p);
            const uint32_t sa[4] = {st[0], st[1], st[2], st[3]};
            store_words(pk + t * P.pk + P.pvc, sa, p);
            store_words(sk + t * P.sk + P.indcpa_sk + P.pvc, sa, p);
            // two rounds per loop iteration: keypair 217-218 against 220 us
            // (Saber), 349.5 against 354 us (FireSaber) at 16384 instances
            fused_matrix<SH, L, true, 2>(P, sa, W.A, B, t, ring, p, &S.sync.mat[grp]);
            return;
        }
        // the secret, packed into the sk, and z
        uint32_t m[4];
        load_words(m, seeds + t * 96 + 32, p);
        uint16_t *s = W.s + t * L * kN;
        secret_stream<SH::MU, 1>(P, m, s, ring, p);
        wave_fence_global();
        if constexpr (SH::MU == 2) pack_half<2, true>(s, sk + t * P.sk, 4 * L, p);
        else pack_half<SH::EQ, false>(s, sk + t * P.sk, 4 * L, p);
        load_words(m, seeds + t * 96 + 64, p);
        store_words(sk + t * P.sk + P.sk - 32, m, p);
        publish(&S.sync.sec[grp], 1);
        // b row by row into the pk and the sk's copy, and H(pk) over it as it grows
        uint8_t *pkt = pk + t * P.pk;
        uint32_t st[25];
#pragma unroll
        for (int i = 0; i < 25; i++) st[i] = 0;
        int done = 0;
        for (int r = 0; r < L; r++) {
            wait_for(&S.sync.done[L + r], kFusedMfma, status);
            pack_half<EP, false>(W.v1 + ((int64_t)r * B + t) * kN, pkt + r * EP * 32, 4, p,
                                 sk + t * P.sk + P.indcpa_sk + r * EP * 32);
            wave_fence_global();
            absorb_upto<kSha3_256Rate, 2>(st, pkt, done, (r + 1) * EP * 32, p);
        }
        // the pk tail (seed_A') was written by the matrix wave before the rows' products began
        absorb_upto<kSha3_256Rate, 2>(st, pkt, done, P.pk, p);
        absorb_final<kSha3_256Rate>(st, pkt, done, P.pk, kSha3DS, p);
        f1600<2>(st, p);
        const uint32_t h[4] = {st[0], st[1], st[2], st[3]};
        store_words(sk + t * P.sk + P.sk - 64, h, p);
        return;
    }
    // products: b = round(A^T s) row by row
    const int k = wave - 4;
    Job J;
    memset(&J, 0, sizeof(J));
    J.status = status;
    J.batch = nloc;
    J.eq = P.eq;
    J.ep = P.ep;
    J.et = P.et;
    J.h = P.h1;
    // Jobs: the heads of rows 0..L-1 (jobs 0..L-1), then the tails (jobs L..2L-1).
    const uint16_t *s = W.s + base * L * kN;
    auto a_of = [&](int r) { return W.A + ((int64_t)r * B + base) * L * kN; };
    auto out_of = [&](int r) { return W.v1 + ((int64_t)r * B + base) * kN; };
    for (int g = 0; g < 2; g++) wait_for(&S.sync.sec[g], 1, status);
    for (int g = 0; g < 2; g++) wait_for(&S.sync.mat[g], (L - 1) * L, status);
    // Order: head 0, head 1, tail 0, head 2, tail 1, ..., tail L-1, so that
    // the hash over the pk can take row r while the later heads run.  Against
    // all heads, then the tails: FireSaber 377 against 387 us, LightSaber 124
    // against 128, Saber 221 against 222 at 16384 instances.  (Taking a tail
    // as soon as it is ready, checked between the claims of the heads,
    // measured 7 us slower on Saber.)  The claims load both of their instances up
    // at the start (PRE): FireSaber 353-354 against 376-377 us, the other
    // sets within 2 us.  They keep eight-coefficient staging on half the
    // lanes (QUAD off): with four coefficients per lane key generation took
    // 229 against 221 us (Saber) and 372 against 354 us (FireSaber), while
    // decapsulation, which keeps it, gained 6 and 28 us.
    auto tail = [&](int r) {
        for (int g = 0; g < 2; g++) wait_for(&S.sync.mat[g], (L - 1) * L + r + 1, status);
        wait_for(&S.sync.done[r], kFusedMfma, status);
        fused_row<1, kEpiQ2PAcc, true, true, false>(S.sync, L + r, nloc, tail_job_of<L>(J, a_of(r), s, out_of(r), P.eq),
                                                    S.mfma[k]);
    };
    for (int r = 0; r < L; r++) {
        fused_row<L - 1, kEpiMod, true, true, false>(
            S.sync, r, nloc, row_job_of(J, a_of(r), s, out_of(r), nullptr, P.eq, kEpiMod, L * kN, L * kN), S.mfma[k]);
        if (r >= 1) tail(r - 1);
    }
    tail(L - 1);
}

// Encapsulation.  Products: rows 0..L-2 of A s', v' = b . s' (b of the pk,
// with the message folded in by the encryption epilogue), and row L-1, the
// last one the matrix stream finishes, as a head and a tail.
template <class SH, int L, int ET>
__global__ __launch_bounds__(kFusedThreads) void enc_fused_kernel(KemParams P, const uint8_t *pk, const uint8_t *rnd,
                                                                  uint8_t *ct, uint8_t *ss, KemWorkspace W, int64_t B,
                                                                  int *status)
{
    constexpr int EP = SH::EP;
    __shared__ FusedLds<L> S;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t base = (int64_t)blockIdx.x * kFusedInst;
    const int nloc = (int)(B - base < kFusedInst ? B - base : kFusedInst);
    if (threadIdx.x < sizeof(FusedSync) / 4) reinterpret_cast<int *>(&S.sync)[threadIdx.x] = 0;
    __syncthreads();
    const int p = lane & 1;
    if (wave < 4) {
        const int grp = wave & 1, li = 32 * grp + (lane >> 1);
        const int64_t t = base + li;
        uint32_t *ring = S.ring[wave] + (lane >> 1) * kRingDwords;
        if (32 * grp >= nloc) {
            publish(wave < 2 ? &S.sync.mat[grp] : &S.sync.sec[grp], kDone);
            return;
        }
        if (li >= nloc) return; // lane 0 of the wave has an instance and publishes
        if (wave < 2) {
            uint32_t m[4];
            load_words(m, pk + t * P.pk + P.pvc, p);
            fused_matrix<SH, L, false, 2>(P, m, W.A, B, t, ring, p, &S.sync.mat[grp]);
            return;
        }
        // m = H(rnd), (K, r) = G(m | H(pk)), the secret from r, b of the pk and m's bits
        uint32_t r[4], m[4], kb[8], kk[8];
        load_words(r, rnd + t * 32, p);
        sha3_256_words<4, 2>(m, r, p);
        wait_for(&S.sync.hpk[grp], 1, status);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            kb[i] = m[i];
            kb[4 + i] = S.hpk[li * 8 + 2 * i + p];
        }
        sha3_512_words<2>(kk, kb, p);
        const uint32_t coins[4] = {kk[4], kk[5], kk[6], kk[7]};
        secret_stream<SH::MU, 2>(P, coins, W.s + t * L * kN, ring, p);
        message_bits(m, W.bits + t * kN, p);
        unpack_b<EP>(P, pk + t * P.pk, W.v2 + t * L * kN, p);
        publish(&S.sync.sec[grp], 1);
        // the ciphertext row by row, and H(ct) over it as it grows
        uint8_t *c = ct + t * P.ct;
        uint32_t st[25];
#pragma unroll
        for (int i = 0; i < 25; i++) st[i] = 0;
        int done = 0;
        for (int row = 0; row < L; row++) {
            wait_for(&S.sync.done[row < L - 1 ? row : L], kFusedMfma, status);
            pack_half<EP, false>(W.v1 + ((int64_t)row * B + t) * kN, c + row * EP * 32, 4, p);
            wave_fence_global();
            absorb_upto<kSha3_256Rate, 2>(st, c, done, (row + 1) * EP * 32, p);
        }
        wait_for(&S.sync.done[L - 1], kFusedMfma, status); // v' (already finished: every wave left it before row L - 1)
        pack_half<ET, false>(W.poly + t * kN, c + P.pvc, 4, p);
        wave_fence_global();
        absorb_upto<kSha3_256Rate, 2>(st, c, done, P.ct, p);
        absorb_final<kSha3_256Rate>(st, c, done, P.ct, kSha3DS, p);
        f1600<2>(st, p);
        // ss = H(K | H(ct))
#pragma unroll
        for (int i = 0; i < 4; i++) kb[4 + i] = st[i];
#pragma unroll
        for (int i = 0; i < 4; i++) kb[i] = kk[i];
        uint32_t sv[4];
        sha3_256_words<8, 2>(sv, kb, p);
        store_words(ss + t * 32, sv, p);
        return;
    }
    const int k = wave - 4;
    if (k == 2 || k == 3) {
        // H(pk) of group k - 2's instances, for the secret waves.  Waves 6 and
        // 7 share their SIMDs with the secret waves (waves go round-robin over
        // the four SIMDs), which wait for this, rather than with the matrix
        // sponges.
        const int li = 32 * (k - 2) + (lane >> 1);
        if (li < nloc) {
            uint32_t h[4];
            sha3_256<2>(h, pk + (base + li) * P.pk, P.pk, p);
#pragma unroll
            for (int i = 0; i < 4; i++) S.hpk[li * 8 + 2 * i + p] = h[i];
        }
        publish(&S.sync.hpk[k - 2], 1);
    }
    Job J;
    memset(&J, 0, sizeof(J));
    J.status = status;
    J.batch = nloc;
    J.eq = P.eq;
    J.ep = P.ep;
    J.et = P.et;
    J.h = P.h1;
    for (int g = 0; g < 2; g++) wait_for(&S.sync.sec[g], 1, status);
    // Claims keep one-term-ahead loads (PRE off: Saber 236 against 224 us)
    // and eight-coefficient staging (QUAD off: Saber 233 against 223 us,
    // FireSaber 360 against 330 us).
    // Jobs: rows 0..L-2 (jobs 0..L-2), v' (job L-1), row L-1 (job L).  Row
    // L - 1 is not split into a head and a tail as in key generation: the
    // head's products then ran beside the last permutations of the matrix
    // stream, which slowed down, and encapsulation took 11 us longer.
    const uint16_t *s = W.s + base * L * kN;
    auto row_of = [&](int r) {
        return row_job_of(J, W.A + ((int64_t)r * B + base) * L * kN, s, W.v1 + ((int64_t)r * B + base) * kN, nullptr,
                          P.eq, kEpiQ2P);
    };
    auto streamed = [&](int r, bool wait) {
        bool ok = true;
        for (int g = 0; g < 2; g++) {
            if (wait) wait_for(&S.sync.mat[g], (r + 1) * L, status);
            else ok = ok && __hip_atomic_load(&S.sync.mat[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >=
                                (r + 1) * L;
        }
        return __builtin_amdgcn_readfirstlane((int)ok) != 0;
    };
    for (int r = 0; r < L - 1; r++) {
        streamed(r, true);
        fused_row<L, kEpiQ2P, false, false, false>(S.sync, r, nloc, row_of(r), S.mfma[k]);
    }
    // Row L - 1 is on the critical path: the stream finishes it last, and the
    // hash over the ciphertext waits for it.  v' is needed only for the
    // ciphertext's last bytes, so a wave works on v' only while row L - 1 is
    // not streamed yet, checking before every claim.
    Job Jv = row_job_of(J, W.v2 + base * L * kN, s, W.poly + base * kN, W.bits + base * kN, P.ep, kEpiEnc);
    bool v_left = true;
    while (v_left && !streamed(L - 1, false)) {
        if (!fused_claim<L, kEpiEnc, false, false, false>(S.sync, L - 1, nloc, Jv, S.mfma[k])) {
            fused_leave(S.sync, L - 1);
            v_left = false;
        }
    }
    streamed(L - 1, true);
    fused_row<L, kEpiQ2P, false, false, false>(S.sync, L, nloc, row_of(L - 1), S.mfma[k]);
    if (v_left) fused_row<L, kEpiEnc, false, false, false>(S.sync, L - 1, nloc, Jv, S.mfma[k]);
}

// Decapsulation in one launch.  The MFMA waves first decode the sk's secret,
// b' and c_m (waves 6 and 7 then hash the ciphertext), then run the
// decryption (job 0: m' = dec(b' . s, c_m)), and after the secret waves have
// turned m' into the re-encryption's secret s'', its products: rows 0..L-2 of
// A s'' (jobs 1..L-1), v'' = b . s'' with m' folded in (job L) and row L-1
// (job L + 1), as in encapsulation.  The secret waves compare each finished
// row with the ciphertext (packing it in registers), and the final hash
// takes K' or z by the outcome.  The sk's pk supplies the matrix seed and b.
template <class SH, int L, int ET>
__global__ __launch_bounds__(kFusedThreads) void dec_fused_kernel(KemParams P, const uint8_t *ct, const uint8_t *sk,
                                                                  uint8_t *ss, uint8_t *fail, KemWorkspace W, int64_t B,
                                                                  int *status)
{
    constexpr int EP = SH::EP;
    // Claims load both instances up front for L >= 3 (measured at 16384
    // instances: Saber 206-208 against 211-212 us; LightSaber 144-145 against
    // 140-141 us, so its claims keep the one-term-ahead loads).
    constexpr bool kDecPre = L >= 3;
    __shared__ FusedLds<L> S;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t base = (int64_t)blockIdx.x * kFusedInst;
    const int nloc = (int)(B - base < kFusedInst ? B - base : kFusedInst);
    if (threadIdx.x < sizeof(FusedSync) / 4) reinterpret_cast<int *>(&S.sync)[threadIdx.x] = 0;
    __syncthreads();
    const int p = lane & 1;
    if (wave < 4) {
        const int grp = wave & 1, li = 32 * grp + (lane >> 1);
        const int64_t t = base + li;
        uint32_t *ring = S.ring[wave] + (lane >> 1) * kRingDwords;
        if (32 * grp >= nloc) {
            publish(wave < 2 ? &S.sync.mat[grp] : &S.sync.sec[grp], kDone);
            return;
        }
        if (li >= nloc) return; // lane 0 of the wave has an instance and publishes
        const uint8_t *skt = sk + t * P.sk;
        if (wave < 2) {
            uint32_t m[4];
            load_words(m, skt + P.indcpa_sk + P.pvc, p);
            fused_matrix<SH, L, false, 2>(P, m, W.A, B, t, ring, p, &S.sync.mat[grp]);
            return;
        }
        // m' from its bits (bit k of m' is coefficient k of the decryption)
        wait_for(&S.sync.done[0], kFusedMfma, status);
        uint32_t m[4];
        const uint32_t *mb = reinterpret_cast<const uint32_t *>(W.bits + t * kN + 32 * p);
#pragma unroll
        for (int w = 0; w < 4; w++) {
            uint32_t x = 0;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const uint32_t two = mb[32 * w + j];
                x |= ((two & 1u) << (2 * j)) | (((two >> 16) & 1u) << (2 * j + 1));
            }
            m[w] = x;
        }
        // (K', r') = G(m' | H(pk)), the secret s'' from r', b of the pk
        uint32_t hpk[4], kb[8], kk[8];
        load_words(hpk, skt + P.sk - 64, p);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            kb[i] = m[i];
            kb[4 + i] = hpk[i];
        }
        sha3_512_words<2>(kk, kb, p);
        const uint32_t coins[4] = {kk[4], kk[5], kk[6], kk[7]};
        secret_stream<SH::MU, 2>(P, coins, W.s + t * L * kN, ring, p);
        unpack_b<EP>(P, skt + P.indcpa_sk, W.v2 + t * L * kN, p);
        publish(&S.sync.sec[grp], 1);
        // the re-encryption against the ciphertext, row by row
        const uint8_t *c = ct + t * P.ct;
        uint32_t diff = 0;
        for (int row = 0; row < L; row++) {
            wait_for(&S.sync.done[row < L - 1 ? row + 1 : L + 1], kFusedMfma, status);
            diff |= pack_compare_half<EP>(W.v1 + ((int64_t)row * B + t) * kN, c + row * EP * 32, 4, p);
        }
        wait_for(&S.sync.done[L], kFusedMfma, status);
        diff |= pack_compare_half<ET>(W.poly + t * kN, c + P.pvc, 4, p);
        diff |= partner(diff);
        // ss = H(K | H(ct)) with K = K' on a match, z otherwise (constant time)
        wait_for(&S.sync.hpk[grp], 1, status);
        uint32_t z[4], sv[4];
        load_words(z, skt + P.sk - 32, p);
        const uint32_t mask = 0u - (uint32_t)(diff != 0);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            kb[i] = kk[i] ^ (mask & (kk[i] ^ z[i]));
            kb[4 + i] = S.hpk[li * 8 + 2 * i + p];
        }
        sha3_256_words<8, 2>(sv, kb, p);
        store_words(ss + t * 32, sv, p);
        if (fail && p == 0) fail[t] = (uint8_t)(mask & 1u);
        return;
    }
    const int k = wave - 4;
    // decode s (sk), b' and c_m (ct), one chunk of 64 coefficients per thread
    {
        constexpr int CH = 8 * L + 4; // chunks per instance
        for (int u = (int)threadIdx.x - 256; u < nloc * CH; u += 64 * kFusedMfma) {
            const int li = u / CH, c = u - li * CH;
            const int64_t t = base + li;
            if (c < 4 * L) {
                const uint64_t *src = reinterpret_cast<const uint64_t *>(sk + t * P.sk);
                uint4 *dst = reinterpret_cast<uint4 *>(W.s + t * L * kN + c * 64);
                if constexpr (SH::MU == 2) unpack_chunk<2, kXfSk2>(src + c * 2, dst);
                else unpack_chunk<SH::EQ, kXfNone>(src + c * SH::EQ, dst);
            } else if (c < 8 * L) {
                unpack_chunk<EP, kXfNone>(reinterpret_cast<const uint64_t *>(ct + t * P.ct) + (c - 4 * L) * EP,
                                          reinterpret_cast<uint4 *>(W.v2 + t * L * kN + (c - 4 * L) * 64));
            } else {
                unpack_chunk<ET, kXfNone>(reinterpret_cast<const uint64_t *>(ct + t * P.ct + P.pvc) + (c - 8 * L) * ET,
                                          reinterpret_cast<uint4 *>(W.poly + t * kN + (c - 8 * L) * 64));
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __hip_atomic_fetch_add(&S.sync.unp, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (k == 2 || k == 3) {
        // H(ct) of group k - 2's instances, for the final hash
        const int li = 32 * (k - 2) + (lane >> 1);
        if (li < nloc) {
            uint32_t h[4];
            sha3_256<2>(h, ct + (base + li) * P.ct, P.ct, p);
#pragma unroll
            for (int i = 0; i < 4; i++) S.hpk[li * 8 + 2 * i + p] = h[i];
        }
        publish(&S.sync.hpk[k - 2], 1);
    }
    Job J;
    memset(&J, 0, sizeof(J));
    J.status = status;
    J.batch = nloc;
    J.eq = P.eq;
    J.ep = P.ep;
    J.et = P.et;
    // decryption: m' bits to W.bits
    wait_for(&S.sync.unp, kFusedMfma, status);
    {
        Job Jd = row_job_of(J, W.v2 + base * L * kN, W.s + base * L * kN, W.bits + base * kN, W.poly + base * kN, P.ep,
                            kEpiDec);
        Jd.h = P.h2;
        fused_row<L, kEpiDec, false, kDecPre>(S.sync, 0, nloc, Jd, S.mfma[k]);
    }
    // the re-encryption
    J.h = P.h1;
    for (int g = 0; g < 2; g++) wait_for(&S.sync.sec[g], 1, status);
    const uint16_t *s = W.s + base * L * kN;
    auto row_of = [&](int r) {
        return row_job_of(J, W.A + ((int64_t)r * B + base) * L * kN, s, W.v1 + ((int64_t)r * B + base) * kN, nullptr,
                          P.eq, kEpiQ2P);
    };
    auto streamed = [&](int r, bool wait) {
        bool ok = true;
        for (int g = 0; g < 2; g++) {
            if (wait) wait_for(&S.sync.mat[g], (r + 1) * L, status);
            else ok = ok && __hip_atomic_load(&S.sync.mat[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >=
                                (r + 1) * L;
        }
        return __builtin_amdgcn_readfirstlane((int)ok) != 0;
    };
    for (int r = 0; r < L - 1; r++) {
        streamed(r, true);
        fused_row<L, kEpiQ2P, false, kDecPre>(S.sync, r + 1, nloc, row_of(r), S.mfma[k]);
    }
    Job Jv = row_job_of(J, W.v2 + base * L * kN, s, W.poly + base * kN, W.bits + base * kN, P.ep, kEpiEnc);
    bool v_left = true;
    while (v_left && !streamed(L - 1, false)) {
        if (!fused_claim<L, kEpiEnc, false, kDecPre>(S.sync, L, nloc, Jv, S.mfma[k])) {
            fused_leave(S.sync, L);
            v_left = false;
        }
    }
    streamed(L - 1, true);
    fused_row<L, kEpiQ2P, false, kDecPre>(S.sync, L + 1, nloc, row_of(L - 1), S.mfma[k]);
    if (v_left) fused_row<L, kEpiEnc, false, kDecPre>(S.sync, L, nloc, Jv, S.mfma[k]);
}

// The parameter set as compile-time values: the sampling shape, L and ET.
template <class SH, int L_, int ET_> struct SetT {
    using Sh = SH;
    static constexpr int L = L_, ET = ET_;
};

template <class F> static hipError_t by_set(const KemParams &P, F f)
{
    auto is = [&](int l, int eq, int mu, int ep, int et) {
        return P.l == l && P.eq == eq && P.mu == mu && P.ep == ep && P.et == et;
    };
    if (is(2, 13, 10, 10, 3)) return f(SetT<Shape<13, 10, 10>, 2, 3>{});
    if (is(3, 13, 8, 10, 4)) return f(SetT<Shape<13, 8, 10>, 3, 4>{});
    if (is(4, 13, 6, 10, 6)) return f(SetT<Shape<13, 6, 10>, 4, 6>{});
    if (is(2, 11, 2, 9, 3)) return f(SetT<Shape<11, 2, 9>, 2, 3>{});
    if (is(3, 11, 2, 9, 5)) return f(SetT<Shape<11, 2, 9>, 3, 5>{});
    if (is(4, 11, 2, 10, 3)) return f(SetT<Shape<11, 2, 10>, 4, 3>{});
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------- host chains

static inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

static Job make_job(const KemParams &P, int mode, int epi, uint16_t *out, const uint16_t *a, const uint16_t *s,
                    const uint16_t *aux, int64_t B, int *status)
{
    Job J;
    memset(&J, 0, sizeof(J));
    J.out = out;
    J.a = a;
    J.s = s;
    J.aux = aux;
    J.status = status;
    J.batch = B;
    J.eq = P.eq;
    J.ep = P.ep;
    J.et = P.et;
    J.epi = epi;
    J.w = (mode == kInner) ? P.ep : P.eq;
    J.h = (epi == kEpiDec) ? P.h2 : P.h1;
    return J;
}

static hipError_t product(const KemParams &P, int backend, int mode, int epi, uint16_t *out, const uint16_t *a,
                          const uint16_t *s, const uint16_t *aux, int64_t B, int *status, hipStream_t st)
{
    const Job J = make_job(P, mode, epi, out, a, s, aux, B, status);
    return backend == 0 ? launch_mfma(J, mode, P.l, st) : launch_valu(J, mode, P.l, st);
}

#define TRY(x)                                                                                                       \
    do {                                                                                                             \
        hipError_t e_ = (x);                                                                                         \
        if (e_ != hipSuccess) return e_;                                                                             \
        e_ = hipGetLastError();                                                                                      \
        if (e_ != hipSuccess) return e_;                                                                             \
    } while (0)

// Codec launch for a run-time width: finds the instantiation with W == w.
template <int W>
static hipError_t pack_w(const uint16_t *src, int64_t src_stride, uint8_t *dst, int64_t dst_stride, int dst_off, int w,
                         int chunks, int64_t B, hipStream_t st, uint32_t *diff)
{
    if constexpr (W > 13) {
        return hipErrorInvalidValue;
    } else {
        if (w != W) return pack_w<W + 1>(src, src_stride, dst, dst_stride, dst_off, w, chunks, B, st, diff);
        hipLaunchKernelGGL((pack_kernel<W>), dim3(blocks_for(B * chunks, 256)), dim3(256), 0, st, src, src_stride,
                           dst, dst_stride, dst_off, chunks, B, diff);
        return hipGetLastError();
    }
}

template <int W>
static hipError_t unpack_w(const uint8_t *src, int64_t src_stride, int src_off, uint16_t *dst, int64_t dst_stride,
                           int w, int chunks, int xf, int64_t B, hipStream_t st)
{
    if constexpr (W > 13) {
        return hipErrorInvalidValue;
    } else {
        if (w != W) return unpack_w<W + 1>(src, src_stride, src_off, dst, dst_stride, w, chunks, xf, B, st);
        const dim3 grid(blocks_for(B * chunks, 256));
        if (xf == kXfSk2)
            hipLaunchKernelGGL((unpack_kernel<W, kXfSk2>), grid, dim3(256), 0, st, src, src_stride, src_off, dst,
                               dst_stride, chunks, B);
        else
            hipLaunchKernelGGL((unpack_kernel<W, kXfNone>), grid, dim3(256), 0, st, src, src_stride, src_off, dst,
                               dst_stride, chunks, B);
        return hipGetLastError();
    }
}

// ncoef is a multiple of 64; byte offsets and strides are multiples of 8.
static hipError_t pack(const uint16_t *src, int64_t src_stride, uint8_t *dst, int64_t dst_stride, int dst_off, int w,
                       int ncoef, int64_t B, hipStream_t st, uint32_t *diff = nullptr)
{
    return pack_w<1>(src, src_stride, dst, dst_stride, dst_off, w, ncoef / 64, B, st, diff);
}

static hipError_t unpack(const uint8_t *src, int64_t src_stride, int src_off, uint16_t *dst, int64_t dst_stride,
                         int w, int ncoef, int xf, int64_t B, hipStream_t st)
{
    return unpack_w<1>(src, src_stride, src_off, dst, dst_stride, w, ncoef / 64, xf, B, st);
}

template <int WS, int XS, int WB, int WC>
static hipError_t dec_unpack_w(const KemParams &P, const uint8_t *sk, const uint8_t *ct, uint16_t *S, uint16_t *Bp,
                               uint16_t *Cm, int64_t B, hipStream_t st, const int32_t *skey)
{
    hipLaunchKernelGGL((dec_unpack_kernel<WS, XS, WB, WC>), dim3(blocks_for(B * (8 * P.l + 4), 256)), dim3(256), 0,
                       st, sk, (int64_t)P.sk, ct, (int64_t)P.ct, P.pvc, S, Bp, Cm, P.l, B, skey);
    return hipGetLastError();
}

template <int EP, int ET>
static hipError_t dec_compare_w(const KemParams &P, const uint16_t *bp, const uint16_t *vp, const uint8_t *ct,
                                uint32_t *diff, int64_t B, hipStream_t st)
{
    hipLaunchKernelGGL((dec_compare_kernel<EP, ET>), dim3(blocks_for(B * (4 * P.l + 4), 256)), dim3(256), 0, st, bp, vp,
                       ct, (int64_t)P.ct, P.pvc, diff, P.l, B);
    return hipGetLastError();
}

// The re-encryption (b' and v') against ct in one launch.
static hipError_t dec_compare(const KemParams &P, const uint16_t *bp, const uint16_t *vp, const uint8_t *ct,
                              uint32_t *diff, int64_t B, hipStream_t st)
{
    if (P.ep == 10 && P.et == 3) return dec_compare_w<10, 3>(P, bp, vp, ct, diff, B, st);
    if (P.ep == 10 && P.et == 4) return dec_compare_w<10, 4>(P, bp, vp, ct, diff, B, st);
    if (P.ep == 10 && P.et == 6) return dec_compare_w<10, 6>(P, bp, vp, ct, diff, B, st);
    if (P.ep == 9 && P.et == 3) return dec_compare_w<9, 3>(P, bp, vp, ct, diff, B, st);
    if (P.ep == 9 && P.et == 5) return dec_compare_w<9, 5>(P, bp, vp, ct, diff, B, st);
    return hipErrorInvalidValue;
}

// s (from the sk), b' and c_m (from the ct) of a decapsulation in one launch,
// for the six parameter sets' widths.
static hipError_t dec_unpack(const KemParams &P, const uint8_t *sk, const uint8_t *ct, uint16_t *S, uint16_t *Bp,
                             uint16_t *Cm, int64_t B, hipStream_t st, const int32_t *skey)
{
    if (!P.sable && P.eq == 13 && P.ep == 10) {
        if (P.et == 3) return dec_unpack_w<13, kXfNone, 10, 3>(P, sk, ct, S, Bp, Cm, B, st, skey);
        if (P.et == 4) return dec_unpack_w<13, kXfNone, 10, 4>(P, sk, ct, S, Bp, Cm, B, st, skey);
        if (P.et == 6) return dec_unpack_w<13, kXfNone, 10, 6>(P, sk, ct, S, Bp, Cm, B, st, skey);
    }
    if (P.sable) {
        if (P.ep == 9 && P.et == 3) return dec_unpack_w<2, kXfSk2, 9, 3>(P, sk, ct, S, Bp, Cm, B, st, skey);
        if (P.ep == 9 && P.et == 5) return dec_unpack_w<2, kXfSk2, 9, 5>(P, sk, ct, S, Bp, Cm, B, st, skey);
        if (P.ep == 10 && P.et == 3) return dec_unpack_w<2, kXfSk2, 10, 3>(P, sk, ct, S, Bp, Cm, B, st, skey);
    }
    return hipErrorInvalidValue;
}

// Calls f(Shape<EQ, MU, EP>{}) for the parameter set's sampling shape.
template <class F> static hipError_t by_shape(const KemParams &P, F f)
{
    if (P.eq == 13 && P.ep == 10) {
        if (P.mu == 8) return f(Shape<13, 8, 10>{});
        if (P.mu == 10) return f(Shape<13, 10, 10>{});
        if (P.mu == 6) return f(Shape<13, 6, 10>{});
    } else if (P.eq == 11 && P.mu == 2) {
        if (P.ep == 9) return f(Shape<11, 2, 9>{});
        if (P.ep == 10) return f(Shape<11, 2, 10>{});
    }
    return hipErrorInvalidValue;
}

// Threads for n lane pairs.
static inline unsigned pair_blocks(int64_t n) { return blocks_for(2 * n, 64); }

// IND-CPA encryption (S/src/cuda_kernel.cu:124-178) in two halves.  The
// InnerProd half needs s', b of the pk (unpacked in W.v2) and the message's
// bits (W.bits, one per coefficient: S/src/pack.cu:86-99), all written by the
// front kernels, and writes pack_ET(round(b . s' - m)) to the ct tail; the
// MatrixVectorMul half needs A and writes pack_EP(round(A s')) to the ct head.
// With diff set (decapsulation's re-encryption), the halves compare with the
// ciphertext at ct instead of writing it.
static hipError_t enc_inner_half(const KemParams &P, const KemWorkspace &W, int backend, uint8_t *ct,
                                 uint32_t *diff, int64_t B, int *status, hipStream_t st)
{
    TRY(product(P, backend, kInner, kEpiEnc, W.poly, W.v2, W.s, W.bits, B, status, st));
    TRY(pack(W.poly, kN, ct, P.ct, P.pvc, P.et, kN, B, st, diff));
    return hipSuccess;
}

static hipError_t enc_matvec_half(const KemParams &P, const KemWorkspace &W, int backend, uint8_t *ct,
                                  uint32_t *diff, int64_t B, int *status, hipStream_t st)
{
    const int L = P.l;
    TRY(product(P, backend, kMatVec, kEpiQ2P, W.v1, W.A, W.s, nullptr, B, status, st));
    TRY(pack(W.v1, L * kN, ct, P.ct, 0, P.ep, L * kN, B, st, diff));
    return hipSuccess;
}

// Encryption's products and the final hashes.  MFMA: A s' and b . s' as one
// product, and the ciphertext packed inside the hash kernel; VALU: the two
// halves, each packing its part, then the hash kernel.
// A / akey: where the MFMA product reads each instance's matrix (default:
// W.A, instance i at i; keyed calls: the key store at the key id).
static hipError_t enc_finish(const KemParams &P, const KemWorkspace &W, int backend, uint8_t *ct, uint8_t *ss,
                             int64_t B, int *status, hipStream_t st, const uint16_t *A = nullptr,
                             const int32_t *akey = nullptr)
{
    const dim3 grid(pair_blocks(B));
    if (backend != 0) {
        TRY(enc_matvec_half(P, W, backend, ct, nullptr, B, status, st));
        TRY(enc_inner_half(P, W, backend, ct, nullptr, B, status, st));
        hipLaunchKernelGGL((enc_hash2_kernel<0, 0>), grid, dim3(64), 0, st, P, ct, W.kr, ss, B, nullptr, nullptr);
        return hipGetLastError();
    }
    Job J = make_job(P, kMatVec, kEpiQ2P, W.v1, A ? A : W.A, W.s, W.bits, B, status);
    J.b = W.v2;
    J.out2 = W.poly;
    J.akey = akey;
    TRY(launch_mfma(J, kEnc, P.l, st));
    auto go = [&](auto k) {
        constexpr int EP = decltype(k)::first_type::value, ET = decltype(k)::second_type::value;
        hipLaunchKernelGGL((enc_hash2_kernel<EP, ET>), grid, dim3(64), 0, st, P, ct, W.kr, ss, B, W.v1, W.poly);
        return hipGetLastError();
    };
    using std::integral_constant;
    using std::pair;
    if (P.ep == 10 && P.et == 3) return go(pair<integral_constant<int, 10>, integral_constant<int, 3>>{});
    if (P.ep == 10 && P.et == 4) return go(pair<integral_constant<int, 10>, integral_constant<int, 4>>{});
    if (P.ep == 10 && P.et == 6) return go(pair<integral_constant<int, 10>, integral_constant<int, 6>>{});
    if (P.ep == 9 && P.et == 3) return go(pair<integral_constant<int, 9>, integral_constant<int, 3>>{});
    if (P.ep == 9 && P.et == 5) return go(pair<integral_constant<int, 9>, integral_constant<int, 5>>{});
    return hipErrorInvalidValue;
}

// Key generation and encapsulation run on one stream: the front kernel keeps
// one sponge wave per SIMD, and the steps after it are short and mostly
// bandwidth-bound, so running them beside each other gains less than the
// event waits cost.
hipError_t kem_keypair(const KemParams &P, const KemWorkspace &W, int backend, const uint8_t *seeds, uint8_t *pk,
                       uint8_t *sk, int64_t B, int *status, hipStream_t st)
{
    if (B <= 0) return hipSuccess;
    if (backend == 0) {
        // one launch (the one-launch section above)
        return by_set(P, [&](auto set) {
            using T = decltype(set);
            hipLaunchKernelGGL((kp_fused_kernel<typename T::Sh, T::L>), dim3(blocks_for(B, kFusedInst)),
                               dim3(kFusedThreads), 0, st, P, seeds, pk, sk, W, B, status);
            return hipGetLastError();
        });
    }
    TRY(by_shape(P, [&](auto sh) {
        hipLaunchKernelGGL(kp_front_kernel<decltype(sh)>, dim3(pair_blocks(2 * B)), dim3(64), 0, st, P, seeds, pk, sk,
                           W.A, W.s, B);
        return hipGetLastError();
    }));
    TRY(product(P, backend, kMatVecT, kEpiQ2P, W.v1, W.A, W.s, nullptr, B, status, st));
    // the front kernel packed s into the sk; the hash kernel packs b into the
    // pk and the sk's copy of it, then hashes the pk
    if (P.ep == 10)
        hipLaunchKernelGGL(kp_hash_kernel<10>, dim3(pair_blocks(B)), dim3(64), 0, st, P, pk, sk, B, W.v1);
    else if (P.ep == 9)
        hipLaunchKernelGGL(kp_hash_kernel<9>, dim3(pair_blocks(B)), dim3(64), 0, st, P, pk, sk, B, W.v1);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t kem_encaps(const KemParams &P, const KemWorkspace &W, int backend, const uint8_t *pk, const uint8_t *rnd,
                      uint8_t *ct, uint8_t *ss, int64_t B, int *status, hipStream_t st)
{
    if (B <= 0) return hipSuccess;
    if (backend == 0) {
        return by_set(P, [&](auto set) {
            using T = decltype(set);
            hipLaunchKernelGGL((enc_fused_kernel<typename T::Sh, T::L, T::ET>), dim3(blocks_for(B, kFusedInst)),
                               dim3(kFusedThreads), 0, st, P, pk, rnd, ct, ss, W, B, status);
            return hipGetLastError();
        });
    }
    TRY(by_shape(P, [&](auto sh) {
        hipLaunchKernelGGL((enc_front_kernel<decltype(sh), false>), dim3(pair_blocks(2 * B)), dim3(64), 0, st, P, rnd, pk,
                           nullptr, W.buf, W.kr, W.A, W.s, W.v2, W.bits, B, (int64_t)0, nullptr);
        return hipGetLastError();
    }));
    return enc_finish(P, W, backend, ct, ss, B, status, st);
}

// Decapsulation; with `resident` set, the matrix A of every instance is
// already in W.A and is not generated again.  The matrix stream (and its
// unpacking) and H(ct) depend only on the inputs and run on W.aux beside the
// decryption and the InnerProd half of the re-encryption.
// resident: A is already in place (W.A, or the key store A at akey[i] for the
// MFMA backend), so the matrix sponge is skipped; skey (resident only): instance
// i's sk is row skey[i] of sk (a key store).
static hipError_t decaps_chain(const KemParams &P, const KemWorkspace &W, int backend, const uint8_t *ct,
                               const uint8_t *sk, uint8_t *ss, uint8_t *fail, int64_t B, int *status, hipStream_t st,
                               bool resident, const uint16_t *A = nullptr, const int32_t *akey = nullptr,
                               const int32_t *skey = nullptr)
{
    const int L = P.l;
    // Without resident A, the matrix sponge and H(ct) (dec_pre_kernel) run on
    // the side stream beside the decryption; with it, H(ct) runs inside the
    // re-encryption's front kernel and the call stays on one stream.
    if (!resident) {
        TRY(hipEventRecord(W.fork, st));
        TRY(hipStreamWaitEvent(W.aux, W.fork, 0));
        TRY(by_shape(P, [&](auto sh) {
            hipLaunchKernelGGL(dec_pre_kernel<decltype(sh)>, dim3(pair_blocks(2 * B)), dim3(64), 0, W.aux, P, ct, sk,
                               W.A, W.hct, B);
            return hipGetLastError();
        }));
        TRY(hipEventRecord(W.join, W.aux));
    }
    // IND-CPA decryption (S/src/cuda_kernel.cu:341-370; SB/src/SABLE_indcpa.cu:422-463)
    if (resident) {
        TRY(dec_unpack(P, sk, ct, W.s, W.v2, W.poly, B, st, skey));
    } else {
        if (P.sable)
            TRY(unpack(sk, P.sk, 0, W.s, L * kN, 2, L * kN, kXfSk2, B, st));
        else
            TRY(unpack(sk, P.sk, 0, W.s, L * kN, P.eq, L * kN, kXfNone, B, st));
        TRY(unpack(ct, P.ct, 0, W.v2, L * kN, P.ep, L * kN, kXfNone, B, st));
        TRY(unpack(ct, P.ct, P.pvc, W.poly, kN, P.et, kN, kXfNone, B, st));
    }
    // m' bits to W.bits, where the re-encryption's InnerProd reads them
    TRY(product(P, backend, kInner, kEpiDec, W.bits, W.v2, W.s, W.poly, B, status, st));
    // re-encryption of m' under the pk stored in the sk
    TRY(by_shape(P, [&](auto sh) {
        hipLaunchKernelGGL(dec_front_kernel<decltype(sh)>, dim3(pair_blocks(resident ? 2 * B : B)), dim3(64), 0, st, P,
                           W.bits, sk, W.buf, W.kr, W.s, W.v2, W.diff, B, ct, resident ? W.hct : nullptr, skey);
        return hipGetLastError();
    }));
    // ct is only read: the compare-packs never write through it
    uint8_t *ctc = const_cast<uint8_t *>(ct);
    if (resident && backend == 0) {
        // A is there from the start: both products of the re-encryption as
        // one L + 1 row product (as in encapsulation), then both compares
        Job J = make_job(P, kMatVec, kEpiQ2P, W.v1, A ? A : W.A, W.s, W.bits, B, status);
        J.b = W.v2;
        J.out2 = W.poly;
        J.akey = akey;
        TRY(launch_mfma(J, kEnc, L, st));
        TRY(dec_compare(P, W.v1, W.poly, ct, W.diff, B, st));
    } else {
        TRY(enc_inner_half(P, W, backend, ctc, W.diff, B, status, st));
        if (!resident) TRY(hipStreamWaitEvent(st, W.join, 0));
        TRY(enc_matvec_half(P, W, backend, ctc, W.diff, B, status, st));
    }
    if (skey) {
        hipLaunchKernelGGL(dec_fin_kernel<true>, dim3(pair_blocks(B)), dim3(64), 0, st, P, W.diff, sk, W.hct, W.kr,
                           ss, fail, B, skey);
    } else {
        hipLaunchKernelGGL(dec_fin_kernel<false>, dim3(pair_blocks(B)), dim3(64), 0, st, P, W.diff, sk, W.hct, W.kr,
                           ss, fail, B, skey);
    }
    return hipGetLastError();
}

hipError_t kem_decaps(const KemParams &P, const KemWorkspace &W, int backend, const uint8_t *ct, const uint8_t *sk,
                      uint8_t *ss, uint8_t *fail, int64_t B, int *status, hipStream_t st)
{
    if (B <= 0) return hipSuccess;
    if (backend == 0) {
        return by_set(P, [&](auto set) {
            using T = decltype(set);
            hipLaunchKernelGGL((dec_fused_kernel<typename T::Sh, T::L, T::ET>), dim3(blocks_for(B, kFusedInst)),
                               dim3(kFusedThreads), 0, st, P, ct, sk, ss, fail, W, B, status);
            return hipGetLastError();
        });
    }
    return decaps_chain(P, W, backend, ct, sk, ss, fail, B, status, st, false);
}

// One launch for the gathers in G (bytes per row given in words[] as bytes,
// multiples of 8).
static hipError_t gather(GatherSet G, const int32_t *keys, int64_t B, hipStream_t st)
{
    for (int g = 0; g < 3; g++) G.words[g] /= 8;
    const int total = G.words[0] + G.words[1] + G.words[2];
    hipLaunchKernelGGL(gather_kernel, dim3(blocks_for(B * total, 256)), dim3(256), 0, st, G, keys, B);
    return hipGetLastError();
}
hipError_t kem_encaps_keyed(const KemParams &P, const KemWorkspace &W, int backend, const KeyStore &K,
                            const int32_t *keys, const uint8_t *rnd, uint8_t *ct, uint8_t *ss, int64_t B, int *status,
                            hipStream_t st)
{
    if (B <= 0) return hipSuccess;
    // The front kernel reads pk and H(pk) from the key store at the key id,
    // and the MFMA product reads A there too; the VALU backend gets a
    // gathered copy of A.
    if (backend != 0) {
        const int abytes = P.l * P.l * kN * 2;
        TRY(gather(GatherSet{{reinterpret_cast<const uint8_t *>(K.A), nullptr, nullptr},
                             {abytes, 0, 0},
                             {reinterpret_cast<uint8_t *>(W.A), nullptr, nullptr},
                             {abytes, 0, 0},
                             {abytes, 0, 0}},
                   keys, B, st));
    }
    TRY(by_shape(P, [&](auto sh) {
        hipLaunchKernelGGL((enc_front_kernel<decltype(sh), true>), dim3(pair_blocks(B)), dim3(64), 0, st, P, rnd, K.pk, K.hpk,
                           W.buf, W.kr, W.A, W.s, W.v2, W.bits, B, B, keys);
        return hipGetLastError();
    }));
    if (backend == 0) return enc_finish(P, W, backend, ct, ss, B, status, st, K.A, keys);
    return enc_finish(P, W, backend, ct, ss, B, status, st);
}

hipError_t kem_decaps_keyed(const KemParams &P, const KemWorkspace &W, int backend, const KeyStore &K,
                            const int32_t *keys, const uint8_t *ct, uint8_t *ss, uint8_t *fail, int64_t B,
                            int *status, hipStream_t st)
{
    if (B <= 0) return hipSuccess;
    // The decode, front and final kernels read the sk from the key store at
    // the key id; as in kem_encaps_keyed, the MFMA product reads A there too
    // and the VALU backend gets a gathered copy of A.
    if (backend != 0) {
        const int abytes = P.l * P.l * kN * 2;
        TRY(gather(GatherSet{{reinterpret_cast<const uint8_t *>(K.A), nullptr, nullptr},
                             {abytes, 0, 0},
                             {reinterpret_cast<uint8_t *>(W.A), nullptr, nullptr},
                             {abytes, 0, 0},
                             {abytes, 0, 0}},
                   keys, B, st));
        return decaps_chain(P, W, backend, ct, K.sk, ss, fail, B, status, st, true, nullptr, nullptr, keys);
    }
    return decaps_chain(P, W, backend, ct, K.sk, ss, fail, B, status, st, true, K.A, keys, keys);
}

hipError_t kem_store_derive(const KemParams &P, const uint8_t *pk, int64_t n, uint16_t *A, uint8_t *hpk,
                            hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    TRY(by_shape(P, [&](auto sh) {
        hipLaunchKernelGGL(matrix_kernel<decltype(sh)>, dim3(pair_blocks(n)), dim3(64), 0, st, P, pk + P.pvc,
                           (int64_t)P.pk, A, n);
        return hipGetLastError();
    }));
    hipLaunchKernelGGL(hash_pk_kernel, dim3(pair_blocks(n)), dim3(64), 0, st, P, pk, hpk, n);
    return hipGetLastError();
}

// The matrix A is the last piece, so a workspace without it (keyed MFMA
// calls read A from the key store) is the same layout cut short.
size_t kem_workspace_bytes(const KemParams &P, int64_t B, bool with_A)
{
    const size_t L = (size_t)P.l, n = (size_t)B;
    size_t b = 0;
    b += n * L * kN * 2;     // s
    b += n * L * kN * 2;     // v1
    b += n * L * kN * 2;     // v2
    b += n * kN * 2;         // bits
    b += n * kN * 2;         // poly
    b += n * 64;             // buf
    b += n * 64;             // kr
    b += n * 4;              // diff
    b += n * 32;             // hct
    b += 9 * 256;            // each of the pieces above rounded up to 256 bytes (kem_workspace_carve)
    if (with_A) b += n * L * L * kN * 2 + 256;
    return b;
}

void kem_workspace_carve(const KemParams &P, int64_t B, uint8_t *base, KemWorkspace &W, bool with_A)
{
    const size_t L = (size_t)P.l, n = (size_t)B;
    auto take = [&](size_t bytes) {
        uint8_t *p = base;
        base += (bytes + 255) & ~(size_t)255;
        return p;
    };
    W.s = reinterpret_cast<uint16_t *>(take(n * L * kN * 2));
    W.v1 = reinterpret_cast<uint16_t *>(take(n * L * kN * 2));
    W.v2 = reinterpret_cast<uint16_t *>(take(n * L * kN * 2));
    W.bits = reinterpret_cast<uint16_t *>(take(n * kN * 2));
    W.poly = reinterpret_cast<uint16_t *>(take(n * kN * 2));
    W.buf = take(n * 64);
    W.kr = take(n * 64);
    W.diff = reinterpret_cast<uint32_t *>(take(n * 4));
    W.hct = take(n * 32);
    W.A = with_A ? reinterpret_cast<uint16_t *>(take(n * L * L * kN * 2)) : nullptr;
}

} // namespace tmvp
