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
// Sponges run one thread per instance (state in registers); codecs run one
// thread per group of 8 coefficients, which is exactly w bytes of a w-bit
// stream (LSB first, as every reference codec).
#include <string.h>

#include <utility>

#include "tmvp_internal.h"
#include "tmvp_keccak.h"
#include "tmvp_kem.h"

namespace tmvp {

using namespace keccak;

// ---------------------------------------------------------------- sampling

// Bit-stream reader over the output of a SHAKE-128 squeeze.
struct ShakeReader {
    uint64_t A[25];
    int lane;     // next lane of the current block
    uint64_t buf; // pending bits, LSB first
    int nbits;

    __device__ __forceinline__ void start(const uint64_t (&seed)[4])
    {
        absorb32(A, seed, kShake128Rate, kShakeDS);
        f1600(A);
        lane = 0;
        buf = 0;
        nbits = 0;
    }
    // Make at least `need` (<= 32) bits available.
    __device__ __forceinline__ void refill(int need)
    {
        while (nbits < need) {
            if (lane == kShake128Rate / 8 * 2) { // 21 lanes = 42 half-lanes per block
                f1600(A);
                lane = 0;
            }
            const uint32_t half = (uint32_t)(A[lane >> 1] >> (32 * (lane & 1)));
            buf |= (uint64_t)half << nbits;
            nbits += 32;
            lane++;
        }
    }
    __device__ __forceinline__ uint32_t take(int w)
    {
        refill(w);
        const uint32_t v = (uint32_t)(buf & ((1ull << w) - 1));
        buf >>= w;
        nbits -= w;
        return v;
    }
};

__device__ __forceinline__ void load32(uint64_t (&m)[4], const uint8_t *p)
{
    const uint64_t *w = reinterpret_cast<const uint64_t *>(p);
#pragma unroll
    for (int i = 0; i < 4; i++) m[i] = w[i];
}
__device__ __forceinline__ void store32(uint8_t *p, const uint64_t (&m)[4])
{
    uint64_t *w = reinterpret_cast<uint64_t *>(p);
#pragma unroll
    for (int i = 0; i < 4; i++) w[i] = m[i];
}

// Threads [0, B): GenMatrix of instance t from the 32-byte seed at
// seed_a + t * a_stride  -> A[t] = L*L EQ-bit polynomials, row-major
// (S/src/pack.cu:145-153, SB/src/pack.cu:41-50).
// Threads [B, 2B): GenSecret of instance t - B from seed_s + t * s_stride
// -> S[t] = L CBD_mu polynomials (S/src/poly.cu:836-916, :1062-1073;
// SB/src/fips202.cu:625-668), Saber as int16, Sable masked mod q.
__global__ __launch_bounds__(64) void gen_kernel(KemParams P, const uint8_t *seed_a, int a_stride,
                                                 const uint8_t *seed_s, int s_stride, uint16_t *A, uint16_t *S,
                                                 int64_t B)
{
    const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (t >= 2 * B) return;
    ShakeReader rd;
    uint64_t seed[4];
    if (t < B) {
        load32(seed, seed_a + t * a_stride);
        rd.start(seed);
        const int n = P.l * P.l * kN;
        uint2 *dst = reinterpret_cast<uint2 *>(A + t * n);
        for (int k = 0; k < n; k += 4) {
            const uint32_t c0 = rd.take(P.eq), c1 = rd.take(P.eq), c2 = rd.take(P.eq), c3 = rd.take(P.eq);
            dst[k >> 2] = make_uint2(c0 | (c1 << 16), c2 | (c3 << 16));
        }
    } else {
        const int64_t i = t - B;
        load32(seed, seed_s + i * s_stride);
        rd.start(seed);
        const int n = P.l * kN, half = P.mu / 2;
        const uint32_t qmask = (1u << P.eq) - 1u;
        uint2 *dst = reinterpret_cast<uint2 *>(S + i * n);
        for (int k = 0; k < n; k += 4) {
            uint32_t c[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t bits = rd.take(P.mu);
                const int v = __builtin_popcount(bits & ((1u << half) - 1u)) - __builtin_popcount(bits >> half);
                c[u] = P.sable ? ((uint32_t)v & qmask) : ((uint32_t)v & 0xffffu);
            }
            dst[k >> 2] = make_uint2(c[0] | (c[1] << 16), c[2] | (c[3] << 16));
        }
    }
}

// ---------------------------------------------------------------- codecs

// One thread per (instance, group of 8 coefficients):
//   pack:   8 w-bit values -> w bytes;  unpack: w bytes -> 8 values.
// kXfSk2 is the Sable secret-key code: s & 3 on packing, ((x ^ 2) - 2) on
// unpacking (SB/src/pack.cu:366-377).
enum Xf : int { kXfNone = 0, kXfSk2 = 1 };

__global__ __launch_bounds__(256) void pack_kernel(const uint16_t *src, int64_t src_stride, uint8_t *dst,
                                                   int64_t dst_stride, int dst_off, int w, int groups, int xf,
                                                   int64_t B)
{
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= B * groups) return;
    const int64_t t = gid / groups;
    const int g = (int)(gid - t * groups);
    const uint16_t *s = src + t * src_stride + g * 8;
    const uint32_t mask = (1u << w) - 1u;
    uint64_t lo = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint32_t v = s[i];
        if (xf == kXfSk2) v &= 3u;
        v &= mask;
        const int bit = i * w;
        if (bit < 64) {
            lo |= (uint64_t)v << bit;
            if (bit + w > 64) hi |= (uint64_t)v >> (64 - bit);
        } else {
            hi |= (uint64_t)v << (bit - 64);
        }
    }
    uint8_t *d = dst + t * dst_stride + dst_off + (int64_t)g * w;
    for (int b = 0; b < w; b++) d[b] = (uint8_t)(b < 8 ? (lo >> (8 * b)) : (hi >> (8 * (b - 8))));
}

__global__ __launch_bounds__(256) void unpack_kernel(const uint8_t *src, int64_t src_stride, int src_off,
                                                     uint16_t *dst, int64_t dst_stride, int w, int groups, int xf,
                                                     int64_t B)
{
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= B * groups) return;
    const int64_t t = gid / groups;
    const int g = (int)(gid - t * groups);
    const uint8_t *s = src + t * src_stride + src_off + (int64_t)g * w;
    uint64_t lo = 0, hi = 0;
    for (int b = 0; b < w; b++) {
        if (b < 8) lo |= (uint64_t)s[b] << (8 * b);
        else hi |= (uint64_t)s[b] << (8 * (b - 8));
    }
    const uint32_t mask = (1u << w) - 1u;
    uint16_t *d = dst + t * dst_stride + g * 8;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int bit = i * w;
        uint64_t v;
        if (bit + w <= 64) v = lo >> bit;
        else if (bit >= 64) v = hi >> (bit - 64);
        else v = (lo >> bit) | (hi << (64 - bit));
        uint32_t x = (uint32_t)v & mask;
        if (xf == kXfSk2) x = ((x ^ 2u) - 2u) & 0xffffu;
        d[i] = (uint16_t)x;
    }
}

// Copy `len` bytes (multiple of 8) per instance.
__global__ __launch_bounds__(256) void copy_kernel(const uint8_t *src, int64_t src_stride, int src_off, uint8_t *dst,
                                                   int64_t dst_stride, int dst_off, int len, int64_t B)
{
    const int words = len / 8;
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= B * words) return;
    const int64_t t = gid / words;
    const int k = (int)(gid - t * words);
    reinterpret_cast<uint64_t *>(dst + t * dst_stride + dst_off)[k] =
        reinterpret_cast<const uint64_t *>(src + t * src_stride + src_off)[k];
}

// ---------------------------------------------------------------- hashing steps (one thread per instance)

// Key generation, before sampling: seed_A' = SHAKE-128(seed_A) into the pk
// tail, z into the sk tail.  seeds[t] = seed_A | seed_s | z.
__global__ __launch_bounds__(64) void kp_seed_kernel(KemParams P, const uint8_t *seeds, uint8_t *pk, uint8_t *sk,
                                                     int64_t B)
{
    const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (t >= B) return;
    uint64_t m[4], A[25];
    load32(m, seeds + t * 96);
    absorb32(A, m, kShake128Rate, kShakeDS);
    f1600(A);
    uint64_t o[4] = {A[0], A[1], A[2], A[3]};
    store32(pk + t * P.pk + P.pvc, o);
    load32(m, seeds + t * 96 + 64);
    store32(sk + t * P.sk + P.sk - 32, m);
}

// Key generation, last step: H(pk) into the sk.
__global__ __launch_bounds__(64) void kp_hash_kernel(KemParams P, const uint8_t *pk, uint8_t *sk, int64_t B)
{
    const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (t >= B) return;
    uint64_t h[4];
    sha3_256(h, pk + t * P.pk, P.pk);
    store32(sk + t * P.sk + P.sk - 64, h);
}

// Encapsulation, first step (S/src/cuda_kernel.cu:116-118):
// m = H(rnd), buf = m | H(pk), kr = G(buf).
__global__ __launch_bounds__(64) void enc_hash1_kernel(KemParams P, const uint8_t *rnd, const uint8_t *pk,
                                                       uint8_t *buf, uint8_t *kr, int64_t B)
{
    const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (t >= B) return;
    uint64_t r[4], m[4], h[4], b[8], k[8];
    load32(r, rnd + t * 32);
    sha3_256_32(m, r);
    sha3_256(h, pk + t * P.pk, P.pk);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        b[i] = m[i];
        b[4 + i] = h[i];
    }
    sha3_512_64(k, b);
    uint64_t *bo = reinterpret_cast<uint64_t *>(buf + t * 64);
    uint64_t *ko = reinterpret_cast<uint64_t *>(kr + t * 64);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        bo[i] = b[i];
        ko[i] = k[i];
    }
}

// Message bits of buf[t][0:32] (BS2POLmsg_gpu, S/src/pack.cu:86-99): one
// thread per (instance, byte).
__global__ __launch_bounds__(256) void msg_bits_kernel(const uint8_t *buf, uint16_t *bits, int64_t B)
{
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= B * 32) return;
    const int64_t t = gid / 32;
    const int byte = (int)(gid - t * 32);
    const uint32_t v = buf[t * 64 + byte];
    uint2 lo, hi;
    lo.x = (v & 1u) | (((v >> 1) & 1u) << 16);
    lo.y = ((v >> 2) & 1u) | (((v >> 3) & 1u) << 16);
    hi.x = ((v >> 4) & 1u) | (((v >> 5) & 1u) << 16);
    hi.y = ((v >> 6) & 1u) | (((v >> 7) & 1u) << 16);
    uint2 *d = reinterpret_cast<uint2 *>(bits + t * kN + byte * 8);
    d[0] = lo;
    d[1] = hi;
}

// Encapsulation, last step (S/src/cuda_kernel.cu:185-186): kr[32:] = H(ct), ss = H(kr).
__global__ __launch_bounds__(64) void enc_hash2_kernel(KemParams P, const uint8_t *ct, uint8_t *kr, uint8_t *ss,
                                                       int64_t B)
{
    const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (t >= B) return;
    uint64_t h[4], k[8], s[4];
    sha3_256(h, ct + t * P.ct, P.ct);
    const uint64_t *ki = reinterpret_cast<const uint64_t *>(kr + t * 64);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        k[i] = ki[i];
        k[4 + i] = h[i];
    }
    sha3_256_64(s, k);
    store32(ss + t * 32, s);
}

// Decapsulation, after the decryption (S/src/cuda_kernel.cu:369-379):
// m' from its bits, buf = m' | H(pk) (from the sk), kr = G(buf).
__global__ __launch_bounds__(64) void dec_hash1_kernel(KemParams P, const uint16_t *mbits, const uint8_t *sk,
                                                       uint8_t *buf, uint8_t *kr, int64_t B)
{
    const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (t >= B) return;
    uint64_t b[8], k[8];
    const uint16_t *mb = mbits + t * kN;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        uint64_t w = 0;
        for (int bit = 0; bit < 64; bit++) w |= (uint64_t)(mb[64 * i + bit] & 1u) << bit;
        b[i] = w;
    }
    const uint64_t *h = reinterpret_cast<const uint64_t *>(sk + t * P.sk + P.sk - 64);
#pragma unroll
    for (int i = 0; i < 4; i++) b[4 + i] = h[i];
    sha3_512_64(k, b);
    uint64_t *bo = reinterpret_cast<uint64_t *>(buf + t * 64);
    uint64_t *ko = reinterpret_cast<uint64_t *>(kr + t * 64);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        bo[i] = b[i];
        ko[i] = k[i];
    }
}

// Decapsulation, last step: compare ct with the re-encryption, kr[32:] = H(ct),
// replace kr[0:32] by z on mismatch, ss = H(kr).  fail[t] records the outcome.
__global__ __launch_bounds__(64) void dec_fin_kernel(KemParams P, const uint8_t *ct, const uint8_t *ctp,
                                                     const uint8_t *sk, uint8_t *kr, uint8_t *ss, uint8_t *fail,
                                                     int64_t B)
{
    const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (t >= B) return;
    const uint64_t *a = reinterpret_cast<const uint64_t *>(ct + t * P.ct);
    const uint64_t *b = reinterpret_cast<const uint64_t *>(ctp + t * P.ct);
    uint64_t diff = 0;
    for (int i = 0; i < P.ct / 8; i++) diff |= a[i] ^ b[i];
    const uint64_t mask = 0ull - (uint64_t)(diff != 0); // all ones on mismatch
    uint64_t h[4], k[8], s[4];
    sha3_256(h, ct + t * P.ct, P.ct);
    const uint64_t *ki = reinterpret_cast<const uint64_t *>(kr + t * 64);
    const uint64_t *z = reinterpret_cast<const uint64_t *>(sk + t * P.sk + P.sk - 32);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        k[i] = ki[i] ^ (mask & (ki[i] ^ z[i]));
        k[4 + i] = h[i];
    }
    sha3_256_64(s, k);
    store32(ss + t * 32, s);
    if (fail) fail[t] = (uint8_t)(mask & 1u);
}

// ---------------------------------------------------------------- host chains

static inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

static hipError_t product(const KemParams &P, int backend, int mode, int epi, uint16_t *out, const uint16_t *a,
                          const uint16_t *s, const uint16_t *aux, int64_t B, int *status, hipStream_t st)
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
    return backend == 0 ? launch_mfma(J, mode, P.l, st) : launch_valu(J, mode, P.l, st);
}

#define TRY(x)                                                                                                       \
    do {                                                                                                             \
        hipError_t e_ = (x);                                                                                         \
        if (e_ != hipSuccess) return e_;                                                                             \
        e_ = hipGetLastError();                                                                                      \
        if (e_ != hipSuccess) return e_;                                                                             \
    } while (0)

static hipError_t pack(const uint16_t *src, int64_t src_stride, uint8_t *dst, int64_t dst_stride, int dst_off, int w,
                       int ncoef, int xf, int64_t B, hipStream_t st)
{
    const int groups = ncoef / 8;
    hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(B * groups, 256)), dim3(256), 0, st, src, src_stride, dst,
                       dst_stride, dst_off, w, groups, xf, B);
    return hipGetLastError();
}

static hipError_t unpack(const uint8_t *src, int64_t src_stride, int src_off, uint16_t *dst, int64_t dst_stride,
                         int w, int ncoef, int xf, int64_t B, hipStream_t st)
{
    const int groups = ncoef / 8;
    hipLaunchKernelGGL(unpack_kernel, dim3(blocks_for(B * groups, 256)), dim3(256), 0, st, src, src_stride, src_off,
                       dst, dst_stride, w, groups, xf, B);
    return hipGetLastError();
}

// IND-CPA encryption of buf[t][0:32] with coins kr[t][32:64] under pk (at
// pk + t * pk_stride) into ct (S/src/cuda_kernel.cu:124-178).
static hipError_t indcpa_enc(const KemParams &P, const KemWorkspace &W, int backend, const uint8_t *pk,
                             int64_t pk_stride, uint8_t *ct, int64_t B, int *status, hipStream_t st)
{
    const int L = P.l;
    hipLaunchKernelGGL(gen_kernel, dim3(blocks_for(2 * B, 64)), dim3(64), 0, st, P, pk + P.pvc, (int)pk_stride,
                       W.kr + 32, 64, W.A, W.s, B);
    TRY(hipGetLastError());
    TRY(product(P, backend, kMatVec, kEpiQ2P, W.v1, W.A, W.s, nullptr, B, status, st));
    TRY(pack(W.v1, L * kN, ct, P.ct, 0, P.ep, L * kN, kXfNone, B, st));
    TRY(unpack(pk, pk_stride, 0, W.v2, L * kN, P.ep, L * kN, kXfNone, B, st));
    hipLaunchKernelGGL(msg_bits_kernel, dim3(blocks_for(B * 32, 256)), dim3(256), 0, st, W.buf, W.bits, B);
    TRY(hipGetLastError());
    TRY(product(P, backend, kInner, kEpiEnc, W.poly, W.v2, W.s, W.bits, B, status, st));
    TRY(pack(W.poly, kN, ct, P.ct, P.pvc, P.et, kN, kXfNone, B, st));
    return hipSuccess;
}

hipError_t kem_keypair(const KemParams &P, const KemWorkspace &W, int backend, const uint8_t *seeds, uint8_t *pk,
                       uint8_t *sk, int64_t B, int *status, hipStream_t st)
{
    if (B <= 0) return hipSuccess;
    const int L = P.l;
    hipLaunchKernelGGL(kp_seed_kernel, dim3(blocks_for(B, 64)), dim3(64), 0, st, P, seeds, pk, sk, B);
    TRY(hipGetLastError());
    hipLaunchKernelGGL(gen_kernel, dim3(blocks_for(2 * B, 64)), dim3(64), 0, st, P, pk + P.pvc, P.pk, seeds + 32, 96,
                       W.A, W.s, B);
    TRY(hipGetLastError());
    TRY(product(P, backend, kMatVecT, kEpiQ2P, W.v1, W.A, W.s, nullptr, B, status, st));
    TRY(pack(W.v1, L * kN, pk, P.pk, 0, P.ep, L * kN, kXfNone, B, st));
    if (P.sable)
        TRY(pack(W.s, L * kN, sk, P.sk, 0, 2, L * kN, kXfSk2, B, st));
    else
        TRY(pack(W.s, L * kN, sk, P.sk, 0, P.eq, L * kN, kXfNone, B, st));
    hipLaunchKernelGGL(copy_kernel, dim3(blocks_for(B * (P.pk / 8), 256)), dim3(256), 0, st, pk, (int64_t)P.pk, 0, sk,
                       (int64_t)P.sk, P.indcpa_sk, P.pk, B);
    TRY(hipGetLastError());
    hipLaunchKernelGGL(kp_hash_kernel, dim3(blocks_for(B, 64)), dim3(64), 0, st, P, pk, sk, B);
    return hipGetLastError();
}

hipError_t kem_encaps(const KemParams &P, const KemWorkspace &W, int backend, const uint8_t *pk, const uint8_t *rnd,
                      uint8_t *ct, uint8_t *ss, int64_t B, int *status, hipStream_t st)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(enc_hash1_kernel, dim3(blocks_for(B, 64)), dim3(64), 0, st, P, rnd, pk, W.buf, W.kr, B);
    TRY(hipGetLastError());
    TRY(indcpa_enc(P, W, backend, pk, P.pk, ct, B, status, st));
    hipLaunchKernelGGL(enc_hash2_kernel, dim3(blocks_for(B, 64)), dim3(64), 0, st, P, ct, W.kr, ss, B);
    return hipGetLastError();
}

hipError_t kem_decaps(const KemParams &P, const KemWorkspace &W, int backend, const uint8_t *ct, const uint8_t *sk,
                      uint8_t *ss, uint8_t *fail, int64_t B, int *status, hipStream_t st)
{
    if (B <= 0) return hipSuccess;
    const int L = P.l;
    // IND-CPA decryption (S/src/cuda_kernel.cu:341-370; SB/src/SABLE_indcpa.cu:422-463)
    if (P.sable)
        TRY(unpack(sk, P.sk, 0, W.s, L * kN, 2, L * kN, kXfSk2, B, st));
    else
        TRY(unpack(sk, P.sk, 0, W.s, L * kN, P.eq, L * kN, kXfNone, B, st));
    TRY(unpack(ct, P.ct, 0, W.v2, L * kN, P.ep, L * kN, kXfNone, B, st));
    TRY(unpack(ct, P.ct, P.pvc, W.bits, kN, P.et, kN, kXfNone, B, st));
    TRY(product(P, backend, kInner, kEpiDec, W.poly, W.v2, W.s, W.bits, B, status, st));
    hipLaunchKernelGGL(dec_hash1_kernel, dim3(blocks_for(B, 64)), dim3(64), 0, st, P, W.poly, sk, W.buf, W.kr, B);
    TRY(hipGetLastError());
    // re-encryption under the pk stored in the sk
    TRY(indcpa_enc(P, W, backend, sk + P.indcpa_sk, P.sk, W.ctp, B, status, st));
    hipLaunchKernelGGL(dec_fin_kernel, dim3(blocks_for(B, 64)), dim3(64), 0, st, P, ct, W.ctp, sk, W.kr, ss, fail, B);
    return hipGetLastError();
}

size_t kem_workspace_bytes(const KemParams &P, int64_t B)
{
    const size_t L = (size_t)P.l, n = (size_t)B;
    size_t b = 0;
    b += n * L * L * kN * 2; // A
    b += n * L * kN * 2;     // s
    b += n * L * kN * 2;     // v1
    b += n * L * kN * 2;     // v2
    b += n * kN * 2;         // bits
    b += n * kN * 2;         // poly
    b += n * 64;             // buf
    b += n * 64;             // kr
    b += n * (size_t)P.ct;   // ctp
    return b + 16 * 256;
}

void kem_workspace_carve(const KemParams &P, int64_t B, uint8_t *base, KemWorkspace &W)
{
    const size_t L = (size_t)P.l, n = (size_t)B;
    auto take = [&](size_t bytes) {
        uint8_t *p = base;
        base += (bytes + 255) & ~(size_t)255;
        return p;
    };
    W.A = reinterpret_cast<uint16_t *>(take(n * L * L * kN * 2));
    W.s = reinterpret_cast<uint16_t *>(take(n * L * kN * 2));
    W.v1 = reinterpret_cast<uint16_t *>(take(n * L * kN * 2));
    W.v2 = reinterpret_cast<uint16_t *>(take(n * L * kN * 2));
    W.bits = reinterpret_cast<uint16_t *>(take(n * kN * 2));
    W.poly = reinterpret_cast<uint16_t *>(take(n * kN * 2));
    W.buf = take(n * 64);
    W.kr = take(n * 64);
    W.ctp = take(n * (size_t)P.ct);
}

} // namespace tmvp
