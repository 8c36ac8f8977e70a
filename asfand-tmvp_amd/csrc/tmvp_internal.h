// Internal declarations shared by the HIP kernels and the C-ABI host layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tmvp {

constexpr int kN = 256;  // polynomial degree bound

// Which product a launch computes.
enum Mode : int {
    kPoly = 0,   // c = a * s                        (one product)
    kMatVec = 1, // out[i] = sum_j A[i][j] s[j]      (L rows x L terms)
    kMatVecT = 2,// out[i] = sum_j A[j][i] s[j]      (L rows x L terms)
    kInner = 3,  // v = sum_j b[j] s[j]              (1 row x L terms)
    kEnc = 4     // encryption: out[i] = sum_j A[i][j] s[j] (q->p) and
                 // out2 = sum_j b[j] s[j] (enc epilogue), L + 1 rows x L terms
};

// Fused epilogues applied to the exact sum x.
enum Epilogue : int {
    kEpiMod = 0,  // (x [+ old]) mod 2^w
    kEpiQ2P = 1,  // ((x + h) mod 2^eq) >> (eq - ep)
    kEpiEnc = 2,  // ((x + h - ((aux & 1) << (ep-1))) mod 2^ep) >> (ep - et)
    kEpiDec = 3,  // ((x + h - ((aux mod 2^et) << (ep-et))) mod 2^ep) >> (ep - 1)
    kEpiQ2PAcc = 4 // ((x + old + h) mod 2^eq) >> (eq - ep): the last terms of a row
                   // whose first terms' sum (mod 2^eq) is already in out
};

// Everything a product launch needs; passed by value as the kernel argument.
struct Job {
    uint16_t *out;
    const uint16_t *a;   // public operand (A, b or a)
    const uint16_t *b;   // kEnc: b of the public key [batch][L][256]
    uint16_t *out2;      // kEnc: the InnerProd row's output [batch][256]
    const int32_t *akey; // kEnc: instance i reads the matrix a + akey[i] (null: a + i)
    const uint16_t *s;   // secret operand
    const uint16_t *aux; // message bits / scale polynomial (epilogues 2, 3)
    int *status;         // set to 1 when an MFMA launch meets an out-of-range secret
    int64_t batch;
    int w;               // output modulus bits: results are taken mod 2^w
    int epi;
    int accumulate;
    int h;               // h1 (q2p, enc) or h2 (dec)
    int eq, ep, et;
    int shape;           // MFMA tile: 16 (16x16x64), 32 (32x32x32), 0 = the library's choice
    // Elements between consecutive instances of a and s, read only by the
    // strided form of the MFMA body (the one-launch KEM's partial rows);
    // every other launch uses the packed layout of its mode.
    int64_t astride, sstride;
};

__host__ __device__ inline constexpr int rows_of(int mode, int l)
{
    return (mode == kMatVec || mode == kMatVecT) ? l : mode == kEnc ? l + 1 : 1;
}
__host__ __device__ inline constexpr int terms_of(int mode, int l) { return mode == kPoly ? 1 : l; }

// Epilogue on one coefficient.  x is the exact sum reduced mod 2^16 or mod
// 2^32 (both are >= 2^w, so every branch below only reads its low w bits).
__device__ __forceinline__ uint32_t apply_epilogue(const Job &J, uint32_t x, uint32_t old, uint32_t aux)
{
    switch (J.epi) {
    case kEpiQ2P:
        return ((x + (uint32_t)J.h) & ((1u << J.eq) - 1u)) >> (J.eq - J.ep);
    case kEpiEnc:
        return ((x + (uint32_t)J.h - ((aux & 1u) << (J.ep - 1))) & ((1u << J.ep) - 1u)) >> (J.ep - J.et);
    case kEpiDec:
        return ((x + (uint32_t)J.h - ((aux & ((1u << J.et) - 1u)) << (J.ep - J.et))) & ((1u << J.ep) - 1u)) >>
               (J.ep - 1);
    case kEpiQ2PAcc:
        return ((x + old + (uint32_t)J.h) & ((1u << J.eq) - 1u)) >> (J.eq - J.ep);
    default:
        return (J.accumulate ? x + old : x) & ((1u << J.w) - 1u);
    }
}

// Host launchers (each defined next to its kernels).
hipError_t launch_valu(const Job &J, int mode, int l, hipStream_t stream);
hipError_t launch_mfma(const Job &J, int mode, int l, hipStream_t stream);

// ---- one persistent MFMA launch over several jobs of different shapes ----
// Kinds enumerate the (L, mode, epilogue) instantiations: 7 product/epilogue
// combinations x L = 2, 3, 4, then the single polymul.
constexpr int kNumKinds = 22;
constexpr int kMaxSegments = 16;
// Kind of (mode, l, epi), or -1 when there is no such instantiation.
int mfma_kind(int mode, int l, int epi);
// Relative time per instance of a kind (terms x (rows + 2)).
int mfma_kind_cost(int kind);
// A segment: workgroups [block_begin, block_begin + blocks) run job J.
struct Segment {
    Job J;
    int kind;
    int block_begin;
    int blocks;
};
struct MultiJob {
    Segment seg[kMaxSegments];
    int nseg;
};
// The MFMA tile shape (16 or 32) launches use when a job or plan leaves it open.
int mfma_default_shape();
// Workgroups of the multi-job kernel of tile `shape` that fit on one CU.
int mfma_multi_per_cu(int shape);
hipError_t launch_mfma_multi(const MultiJob &M, int grid, int shape, hipStream_t stream);

} // namespace tmvp
