/*
 * ORACLE - test infrastructure only.
 *
 * CPU restatement of the Saber/Sable polynomial arithmetic of the reference
 * (Muhammad-Asfand/asfand-tmvp).  It exists to check the GPU library and to
 * provide the CPU baseline timing; nothing in the shipped library links it,
 * calls it or falls back to it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 *
 * Parity unpinned: or_indcpa_keypair's seed_A = SHAKE-128(seed_A) step and
 * the KEM's ciphertext / shared-key bytes.  The reference has no key
 * generation and holds no KEM known-answer vectors, so those are checked
 * only through their pinned parts (GenMatrix, GenSecret, the transposed
 * product and its rounding against the reference's public keys, the exact
 * decoding window of the encryption rounding) and by round trips.
 *
 * Paths are relative to the reference root:
 *   S/  = TMVP-Saber/SABER-GPU/
 *   SB/ = TMVP_SABLE/
 */
#ifndef TMVP_ORACLE_H
#define TMVP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_N 256

/* Parameter-set ids (same numbering as the GPU library). */
enum {
    OR_LIGHTSABER = 0, /* S/include/params.h:11-13  (L=2, mu=10, ET=3) */
    OR_SABER = 1,      /* S/include/params.h:14-16  (L=3, mu=8,  ET=4) */
    OR_FIRESABER = 2,  /* S/include/params.h:17-19  (L=4, mu=6,  ET=6) */
    OR_SABLE1 = 3,     /* SB/include/SABLE_params.h:8-12  (K=2, p=2^9,  ET=3) */
    OR_SABLE2 = 4,     /* SB/include/SABLE_params.h:14-18 (K=3, p=2^9,  ET=5) */
    OR_SABLE3 = 5,     /* SB/include/SABLE_params.h:20-24 (K=4, p=2^10, ET=3) */
    OR_NUM_PARAM_SETS = 6
};

typedef struct {
    int l, eq, ep, et, mu, sable;
    int h1, h2;
    int polyvec_compressed_bytes; /* L * EP * 32 */
    int scale_bytes;              /* ET * 32 */
    int indcpa_pk_bytes;
    int indcpa_sk_bytes;
    int pk_bytes, sk_bytes, ct_bytes;
    int matrix_stream_bytes;      /* L*L*EQ*32 */
    int secret_stream_bytes;      /* L*MU*32 */
} or_params;

int or_get_params(int ps, or_params *out);

/* ---- FIPS 202 (written from the standard) ---- */
void or_keccak_f1600(uint64_t st[25]);
void or_shake128(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void or_sha3_256(uint8_t out[32], const uint8_t *in, size_t inlen);
void or_sha3_512(uint8_t out[64], const uint8_t *in, size_t inlen);

/* ---- negacyclic polynomial products in Z_{2^16}[x]/(x^256+1) ---- */
void or_polymul_schoolbook(const uint16_t a[OR_N], const uint16_t s[OR_N], uint16_t c[OR_N]);
void or_polymul_toom4(const uint16_t a[OR_N], const uint16_t s[OR_N], uint16_t c[OR_N]);

/* method: 0 = schoolbook, 1 = Toom-4.  Outputs are raw sums mod 2^16,
 * accumulated into out (out += ...), exactly like the reference kernels. */
void or_matvec(int l, const uint16_t *A, const uint16_t *s, uint16_t *out, int transpose, int method);
void or_innerprod(int l, const uint16_t *b, const uint16_t *s, uint16_t *out, int method);

/* Mean ns of one single-instance MatrixVectorMul (l <= 4), repeated reps times. */
double or_bench_matvec1(int l, const uint16_t *A, const uint16_t *s, int method, int reps);

/* Batched drivers over independent instances, nthreads >= 1 (pthreads). */
void or_polymul_batch(const uint16_t *a, const uint16_t *s, uint16_t *c, int64_t batch,
                      int method, int nthreads);
void or_matvec_batch(int l, const uint16_t *A, const uint16_t *s, uint16_t *out, int64_t batch,
                     int transpose, int method, int nthreads);
void or_innerprod_batch(int l, const uint16_t *b, const uint16_t *s, uint16_t *out, int64_t batch,
                        int method, int nthreads);

/* ---- rounding / shift steps (outputs masked to their wire width) ---- */
uint16_t or_round_q2p(int ps, uint16_t x);
uint16_t or_round_enc_v(int ps, uint16_t v, uint16_t mbit);
uint16_t or_round_dec_m(int ps, uint16_t v, uint16_t c);

/* ---- codecs: LSB-first bit packing of w-bit coefficients ---- */
void or_pack_bits(const uint16_t *data, int n, int w, uint8_t *bytes);
void or_unpack_bits(const uint8_t *bytes, int n, int w, uint16_t *data);

/* ---- sampling ---- */
void or_cbd(int ps, const uint8_t *buf, uint16_t s[OR_N]);
void or_gen_matrix(int ps, const uint8_t seed[32], uint16_t *A);
void or_gen_secret(int ps, const uint8_t seed[32], uint16_t *s);

/* ---- IND-CPA and KEM ---- */
void or_indcpa_keypair(int ps, const uint8_t seed_a[32], const uint8_t seed_s[32],
                       uint8_t *pk, uint8_t *sk);
void or_indcpa_enc(int ps, const uint8_t m[32], const uint8_t coins[32], const uint8_t *pk,
                   uint8_t *ct);
void or_indcpa_dec(int ps, const uint8_t *sk, const uint8_t *ct, uint8_t m[32]);
void or_kem_keypair(int ps, const uint8_t seed_a[32], const uint8_t seed_s[32],
                    const uint8_t z[32], uint8_t *pk, uint8_t *sk);
void or_kem_enc(int ps, const uint8_t rnd[32], const uint8_t *pk, uint8_t *ct, uint8_t ss[32]);
int or_kem_dec(int ps, const uint8_t *ct, const uint8_t *sk, uint8_t ss[32]);

void or_kem_keypair_batch(int ps, const uint8_t *seeds, uint8_t *pk, uint8_t *sk, int64_t batch,
                          int nthreads);
void or_kem_enc_batch(int ps, const uint8_t *rnd, const uint8_t *pk, uint8_t *ct, uint8_t *ss,
                      int64_t batch, int nthreads);
void or_kem_dec_batch(int ps, const uint8_t *ct, const uint8_t *sk, uint8_t *ss, int64_t batch,
                      int nthreads);

#ifdef __cplusplus
}
#endif
#endif
