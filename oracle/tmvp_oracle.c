/*
 * ORACLE - test infrastructure only (see tmvp_oracle.h).
 *
 * A plain-C restatement of the reference's Saber/Sable arithmetic, used to
 * check the GPU library bit for bit and as the CPU baseline.  The shipped
 * library never links or calls this file.
 *
 * Reference citations are relative to the reference root:
 *   S/  = TMVP-Saber/SABER-GPU/      SB/ = TMVP_SABLE/
 */
#define _POSIX_C_SOURCE 200809L /* clock_gettime */
#include "tmvp_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Parameter sets: S/include/params.h:1-50, SB/include/SABLE_params.h:1-73.  */
/* ------------------------------------------------------------------------ */

static const int k_ps_table[OR_NUM_PARAM_SETS][6] = {
    /* l, eq, ep, et, mu, sable */
    {2, 13, 10, 3, 10, 0}, /* LightSaber  S/include/params.h:11-13 */
    {3, 13, 10, 4, 8, 0},  /* Saber       S/include/params.h:14-16 */
    {4, 13, 10, 6, 6, 0},  /* FireSaber   S/include/params.h:17-19 */
    {2, 11, 9, 3, 2, 1},   /* Sable t1    SB/include/SABLE_params.h:8-12, :26-27 */
    {3, 11, 9, 5, 2, 1},   /* Sable t2    SB/include/SABLE_params.h:14-18 */
    {4, 11, 10, 3, 2, 1},  /* Sable t3    SB/include/SABLE_params.h:20-24 */
};

int or_get_params(int ps, or_params *p)
{
    if (ps < 0 || ps >= OR_NUM_PARAM_SETS || !p) return -1;
    const int *t = k_ps_table[ps];
    p->l = t[0];
    p->eq = t[1];
    p->ep = t[2];
    p->et = t[3];
    p->mu = t[4];
    p->sable = t[5];
    /* h1, h2: S/src/poly.cu:15-16, SB/src/pack.cu:15-16 */
    p->h1 = 1 << (p->eq - p->ep - 1);
    p->h2 = (1 << (p->ep - 2)) - (1 << (p->ep - p->et - 1)) + p->h1;
    p->polyvec_compressed_bytes = p->l * p->ep * OR_N / 8;
    p->scale_bytes = p->et * OR_N / 8;
    p->indcpa_pk_bytes = p->polyvec_compressed_bytes + 32;
    /* Saber stores s as 13-bit words (S/include/params.h:45);
     * Sable as 2-bit codes (SB/include/SABLE_params.h:59). */
    p->indcpa_sk_bytes = p->sable ? p->l * 2 * OR_N / 8 : p->l * p->eq * OR_N / 8;
    p->pk_bytes = p->indcpa_pk_bytes;
    p->sk_bytes = p->indcpa_sk_bytes + p->indcpa_pk_bytes + 32 + 32;
    p->ct_bytes = p->polyvec_compressed_bytes + p->scale_bytes;
    p->matrix_stream_bytes = p->l * p->l * p->eq * OR_N / 8;
    p->secret_stream_bytes = p->l * p->mu * OR_N / 8;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FIPS 202.  Written from the standard (sections 3.2 and 4-6): the round    */
/* constants come from the rc(t) LFSR of Algorithm 5 and the rho offsets    */
/* from the (x,y) -> (y, 2x+3y) walk of Algorithm 2.  The reference's own   */
/* Keccak (S/src/poly.cu:457-797) computes the same permutation; it is      */
/* checked here against Python's hashlib in the tests instead.               */
/* ------------------------------------------------------------------------ */

static uint64_t g_rc[24];
static int g_rho[25];
static pthread_once_t g_keccak_once = PTHREAD_ONCE_INIT;

static int lfsr_rc_bit(int t)
{
    int m = t % 255;
    unsigned r = 1; /* R = 10000000 : bit i of r is R[i] */
    if (m == 0) return 1;
    for (int i = 0; i < m; i++) {
        r <<= 1;
        if (r & 0x100) r ^= 0x171; /* R[0],R[4],R[5],R[6] ^= R[8]; drop R[8] */
    }
    return (int)(r & 1);
}

static void keccak_tables(void)
{
    for (int ir = 0; ir < 24; ir++) {
        uint64_t rc = 0;
        for (int j = 0; j <= 6; j++)
            if (lfsr_rc_bit(j + 7 * ir)) rc |= 1ULL << ((1 << j) - 1);
        g_rc[ir] = rc;
    }
    g_rho[0] = 0;
    int x = 1, y = 0;
    for (int t = 0; t < 24; t++) {
        g_rho[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
        int nx = y, ny = (2 * x + 3 * y) % 5;
        x = nx;
        y = ny;
    }
}

static inline uint64_t rotl64(uint64_t v, int n) { return n ? (v << n) | (v >> (64 - n)) : v; }

void or_keccak_f1600(uint64_t a[25])
{
    pthread_once(&g_keccak_once, keccak_tables);
    uint64_t c[5], d[5], b[25];
    for (int round = 0; round < 24; round++) {
        for (int x = 0; x < 5; x++) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        for (int x = 0; x < 5; x++) d[x] = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) a[i] ^= d[i % 5];
        /* rho + pi: B[y, 2x+3y] = rot(A[x,y], r[x,y]) */
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(a[x + 5 * y], g_rho[x + 5 * y]);
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++)
                a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
        a[0] ^= g_rc[round];
    }
}

static void sponge(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen, size_t rate,
                   uint8_t dsbyte)
{
    uint64_t st[25] = {0};
    uint8_t blk[200];
    while (inlen >= rate) {
        for (size_t i = 0; i < rate / 8; i++) {
            uint64_t w = 0;
            for (int k = 0; k < 8; k++) w |= (uint64_t)in[8 * i + k] << (8 * k);
            st[i] ^= w;
        }
        or_keccak_f1600(st);
        in += rate;
        inlen -= rate;
    }
    memset(blk, 0, sizeof(blk));
    memcpy(blk, in, inlen);
    blk[inlen] ^= dsbyte;    /* domain bits + first pad bit */
    blk[rate - 1] ^= 0x80;   /* last pad bit (pad10*1) */
    for (size_t i = 0; i < rate / 8; i++) {
        uint64_t w = 0;
        for (int k = 0; k < 8; k++) w |= (uint64_t)blk[8 * i + k] << (8 * k);
        st[i] ^= w;
    }
    while (outlen) {
        or_keccak_f1600(st);
        size_t take = outlen < rate ? outlen : rate;
        for (size_t i = 0; i < take; i++) out[i] = (uint8_t)(st[i / 8] >> (8 * (i % 8)));
        out += take;
        outlen -= take;
    }
}

void or_shake128(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen)
{
    sponge(out, outlen, in, inlen, 168, 0x1F);
}
void or_sha3_256(uint8_t out[32], const uint8_t *in, size_t inlen) { sponge(out, 32, in, inlen, 136, 0x06); }
void or_sha3_512(uint8_t out[64], const uint8_t *in, size_t inlen) { sponge(out, 64, in, inlen, 72, 0x06); }

/* ------------------------------------------------------------------------ */
/* Schoolbook negacyclic product with the reference's uint16 wrap-around.   */
/* Restates the inner loops of MatVecMul_gpu_shared (S/src/poly.cu:281-284, */
/* repeated to :375; SB/src/pack.cu:67-70):                                  */
/*   c[m] = sum_{k<=m} s[m-k] a[k] - sum_{k=1}^{255-m} s[m+k] a[256-k]       */
/* evaluated mod 2^16.  Loops run over k outermost so that the compiler can */
/* vectorise the contiguous m loops (16-bit lanes wrap exactly mod 2^16).   */
/* ------------------------------------------------------------------------ */

void or_polymul_schoolbook(const uint16_t a[OR_N], const uint16_t s[OR_N], uint16_t c[OR_N])
{
    uint16_t acc[OR_N];
    memset(acc, 0, sizeof(acc));
    for (int k = 0; k < OR_N; k++) {
        const uint16_t ak = a[k];
        /* terms with m >= k: + a[k] s[m-k] */
        for (int m = k; m < OR_N; m++) acc[m] = (uint16_t)(acc[m] + (uint16_t)((uint32_t)ak * s[m - k]));
        /* terms with m < k: x^(k + (m-k+256)) = -x^m */
        for (int m = 0; m < k; m++)
            acc[m] = (uint16_t)(acc[m] - (uint16_t)((uint32_t)ak * s[m - k + OR_N]));
    }
    memcpy(c, acc, sizeof(acc));
}

/* ------------------------------------------------------------------------ */
/* Toom-Cook-4 over two-level Karatsuba, restated from the unlaunched device */
/* code SB/src/pack.cu:505-732 (karatsuba_simple :512-603, toom_cook_4way  */
/* :607-722, pol_mul :726-732).  Evaluation points 0, 1, -1, 1/2, -1/2, 2,  */
/* inf; the interpolation uses the inverses of 3, 9 and 15 mod 2^16 and    */
/* shifts, so the result is exact mod 2^13 only.  The int32 promotions and */
/* arithmetic right shifts of the reference are kept.                      */
/* ------------------------------------------------------------------------ */

#define KS_N 64
static inline uint16_t m16(uint32_t x, uint32_t y) { return (uint16_t)(x * y); }

static void karatsuba64(const uint16_t *a, const uint16_t *b, uint16_t *r /* 127 */)
{
    /* Two levels: 64 = (32 | 32), each 32 = (16 | 16).  a_q = quarter q. */
    const int Q = KS_N / 4;
    uint16_t d01[2 * Q - 1], d23[2 * Q - 1], d0123[2 * Q - 1], mid[KS_N - 1];
    memset(r, 0, (2 * KS_N - 1) * sizeof(uint16_t));
    memset(d01, 0, sizeof(d01));
    memset(d23, 0, sizeof(d23));
    memset(d0123, 0, sizeof(d0123));
    memset(mid, 0, sizeof(mid));
    for (int i = 0; i < Q; i++) {
        uint16_t a0 = a[i], a1 = a[i + Q], a2 = a[i + 2 * Q], a3 = a[i + 3 * Q];
        for (int j = 0; j < Q; j++) {
            uint16_t b0 = b[j], b1 = b[j + Q], b2 = b[j + 2 * Q], b3 = b[j + 3 * Q];
            r[i + j] += m16(a0, b0);
            r[i + j + 2 * Q] += m16(a1, b1);
            d01[i + j] += m16((uint16_t)(a0 + a1), (uint16_t)(b0 + b1));
            r[i + j + 4 * Q] += m16(a2, b2);
            r[i + j + 6 * Q] += m16(a3, b3);
            d23[i + j] += m16((uint16_t)(a2 + a3), (uint16_t)(b2 + b3));
            uint16_t a02 = a0 + a2, b02 = b0 + b2, a13 = a1 + a3, b13 = b1 + b3;
            mid[i + j] += m16(a02, b02);
            mid[i + j + 2 * Q] += m16(a13, b13);
            d0123[i + j] += m16((uint16_t)(a02 + a13), (uint16_t)(b02 + b13));
        }
    }
    for (int i = 0; i < 2 * Q - 1; i++) {
        d0123[i] = (uint16_t)(d0123[i] - mid[i] - mid[i + 2 * Q]);
        d01[i] = (uint16_t)(d01[i] - r[i] - r[i + 2 * Q]);
        d23[i] = (uint16_t)(d23[i] - r[i + 4 * Q] - r[i + 6 * Q]);
    }
    for (int i = 0; i < 2 * Q - 1; i++) {
        mid[i + Q] += d0123[i];
        r[i + Q] += d01[i];
        r[i + 5 * Q] += d23[i];
    }
    for (int i = 0; i < KS_N - 1; i++) mid[i] = (uint16_t)(mid[i] - r[i] - r[i + KS_N]);
    for (int i = 0; i < KS_N - 1; i++) r[i + KS_N / 2] += mid[i];
}

static inline uint16_t asr16(int32_t v, int sh) { return (uint16_t)(v >> sh); }
static inline int32_t wmul(int32_t x, int32_t y) { return (int32_t)((uint32_t)x * (uint32_t)y); }

static void toom4_full(const uint16_t *a, const uint16_t *b, uint16_t *c /* 512, += */)
{
    const int NB = OR_N / 4; /* 64 */
    const int32_t inv3 = 43691, inv9 = 36409, inv15 = 61167;
    uint16_t aw[7][NB], bw[7][NB], w[7][2 * NB - 1];
    const uint16_t *ops[2] = {a, b};
    uint16_t(*outs[2])[NB] = {aw, bw};
    for (int t = 0; t < 2; t++) {
        const uint16_t *x = ops[t];
        uint16_t(*o)[NB] = outs[t];
        for (int j = 0; j < NB; j++) {
            uint16_t x0 = x[j], x1 = x[j + NB], x2 = x[j + 2 * NB], x3 = x[j + 3 * NB];
            uint16_t e = x0 + x2, f = x1 + x3;
            o[0][j] = x3;                                              /* inf */
            o[1][j] = (uint16_t)((x3 << 3) + (x2 << 2) + (x1 << 1) + x0); /* 2 */
            o[2][j] = (uint16_t)(e + f);                               /* 1 */
            o[3][j] = (uint16_t)(e - f);                               /* -1 */
            uint16_t g = (uint16_t)(((x0 << 2) + x2) << 1), h = (uint16_t)((x1 << 2) + x3);
            o[4][j] = (uint16_t)(g + h);                               /* 8 * x(1/2) */
            o[5][j] = (uint16_t)(g - h);                               /* 8 * x(-1/2) */
            o[6][j] = x0;                                              /* 0 */
        }
    }
    for (int k = 0; k < 7; k++) karatsuba64(aw[k], bw[k], w[k]);
    for (int i = 0; i < 2 * NB - 1; i++) {
        uint16_t r0 = w[0][i], r1 = w[1][i], r2 = w[2][i], r3 = w[3][i], r4 = w[4][i], r5 = w[5][i],
                 r6 = w[6][i];
        r1 = (uint16_t)(r1 + r4);
        r5 = (uint16_t)(r5 - r4);
        r3 = asr16((int32_t)r3 - (int32_t)r2, 1);
        r4 = (uint16_t)(r4 - r0);
        r4 = (uint16_t)(r4 - (r6 << 6));
        r4 = (uint16_t)((r4 << 1) + r5);
        r2 = (uint16_t)(r2 + r3);
        r1 = (uint16_t)(r1 - (r2 << 6) - r2);
        r2 = (uint16_t)(r2 - r6);
        r2 = (uint16_t)(r2 - r0);
        r1 = (uint16_t)(r1 + 45 * r2);
        r4 = asr16(wmul((int32_t)r4 - ((int32_t)r2 << 3), inv3), 3);
        r5 = (uint16_t)(r5 + r1);
        r1 = asr16(wmul((int32_t)r1 + ((int32_t)r3 << 4), inv9), 1);
        r3 = (uint16_t)(-(int32_t)(r3 + r1));
        r5 = asr16(wmul(30 * (int32_t)r1 - (int32_t)r5, inv15), 2);
        r2 = (uint16_t)(r2 - r4);
        r1 = (uint16_t)(r1 - r5);
        c[i] += r6;
        c[i + 64] += r5;
        c[i + 128] += r4;
        c[i + 192] += r3;
        c[i + 256] += r2;
        c[i + 320] += r1;
        c[i + 384] += r0;
    }
}

void or_polymul_toom4(const uint16_t a[OR_N], const uint16_t s[OR_N], uint16_t c[OR_N])
{
    uint16_t full[2 * OR_N];
    memset(full, 0, sizeof(full));
    toom4_full(a, s, full);
    for (int i = 0; i < OR_N; i++) c[i] = (uint16_t)(full[i] - full[i + OR_N]);
}

typedef void (*mulfn)(const uint16_t *, const uint16_t *, uint16_t *);
static mulfn pick(int method) { return method == 1 ? or_polymul_toom4 : or_polymul_schoolbook; }

/* MatrixVectorMul: r[i] += sum_j A[i][j] s[j]  (S/src/poly.cu:269-376, L=3
 * unrolled; SB/src/pack.cu:54-163).  transpose != 0 uses A[j][i] (the key-
 * generation form, which the reference does not contain). */
void or_matvec(int l, const uint16_t *A, const uint16_t *s, uint16_t *out, int transpose, int method)
{
    uint16_t t[OR_N];
    mulfn f = pick(method);
    for (int i = 0; i < l; i++)
        for (int j = 0; j < l; j++) {
            const uint16_t *a = A + (size_t)(transpose ? (j * l + i) : (i * l + j)) * OR_N;
            f(a, s + (size_t)j * OR_N, t);
            for (int k = 0; k < OR_N; k++) out[i * OR_N + k] = (uint16_t)(out[i * OR_N + k] + t[k]);
        }
}

/* InnerProd: r += sum_j b_j s_j  (S/src/poly.cu:227-248; SB/src/pack.cu:262-286
 * masks mod p after every j, which is the same value mod p). */
void or_innerprod(int l, const uint16_t *b, const uint16_t *s, uint16_t *out, int method)
{
    uint16_t t[OR_N];
    mulfn f = pick(method);
    for (int j = 0; j < l; j++) {
        f(b + (size_t)j * OR_N, s + (size_t)j * OR_N, t);
        for (int k = 0; k < OR_N; k++) out[k] = (uint16_t)(out[k] + t[k]);
    }
}

/* Mean wall time (ns) of one single-instance MatrixVectorMul, out zeroed and
 * recomputed `reps` times on this thread: the batch-1 CPU timing (no call
 * overhead from the caller per instance). */
double or_bench_matvec1(int l, const uint16_t *A, const uint16_t *s, int method, int reps)
{
    uint16_t out[4 * OR_N];
    volatile uint16_t sink = 0;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int r = 0; r < reps; r++) {
        memset(out, 0, sizeof(out));
        or_matvec(l, A, s, out, 0, method);
        sink = (uint16_t)(sink + out[r % (l * OR_N)]);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    (void)sink;
    return ((double)(t1.tv_sec - t0.tv_sec) * 1e9 + (double)(t1.tv_nsec - t0.tv_nsec)) / (reps > 0 ? reps : 1);
}

/* ---- pthread batch drivers ---- */

typedef struct {
    int kind, l, transpose, method;
    const uint16_t *x, *y;
    uint16_t *out;
    int64_t lo, hi;
} job_t;

static void *run_job(void *arg)
{
    job_t *j = (job_t *)arg;
    const int64_t n = OR_N, l = j->l;
    for (int64_t b = j->lo; b < j->hi; b++) {
        if (j->kind == 0)
            pick(j->method)(j->x + b * n, j->y + b * n, j->out + b * n);
        else if (j->kind == 1)
            or_matvec(j->l, j->x + b * l * l * n, j->y + b * l * n, j->out + b * l * n, j->transpose,
                      j->method);
        else
            or_innerprod(j->l, j->x + b * l * n, j->y + b * l * n, j->out + b * n, j->method);
    }
    return NULL;
}

static void run_batch(job_t proto, int64_t batch, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (nthreads > batch) nthreads = (int)(batch > 0 ? batch : 1);
    pthread_t th[256];
    job_t jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = proto;
        jobs[t].lo = batch * t / nthreads;
        jobs[t].hi = batch * (t + 1) / nthreads;
    }
    if (nthreads == 1) {
        run_job(&jobs[0]);
        return;
    }
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, run_job, &jobs[t]);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

void or_polymul_batch(const uint16_t *a, const uint16_t *s, uint16_t *c, int64_t batch, int method,
                      int nthreads)
{
    job_t j = {0, 1, 0, method, a, s, c, 0, 0};
    run_batch(j, batch, nthreads);
}

void or_matvec_batch(int l, const uint16_t *A, const uint16_t *s, uint16_t *out, int64_t batch,
                     int transpose, int method, int nthreads)
{
    job_t j = {1, l, transpose, method, A, s, out, 0, 0};
    run_batch(j, batch, nthreads);
}

void or_innerprod_batch(int l, const uint16_t *b, const uint16_t *s, uint16_t *out, int64_t batch,
                        int method, int nthreads)
{
    job_t j = {2, l, 0, method, b, s, out, 0, 0};
    run_batch(j, batch, nthreads);
}

/* ------------------------------------------------------------------------ */
/* Rounding / shift steps, in each family's own form.                       */
/* Saber: post_process / post_process2 / post_process3, S/src/poly.cu:381-403 */
/*        (int arithmetic, arithmetic shift, masked later by the codecs     */
/*        S/src/pack.cu:28-32, :119, :208).                                  */
/* Sable: post_process, post_process3+msg shift+post_process4, post_process7 */
/*        SB/src/pack.cu:167-178, :290-335, :436-442.                        */
/* Outputs are masked to the wire width (EP, ET or 1 bit).                  */
/* ------------------------------------------------------------------------ */

uint16_t or_round_q2p(int ps, uint16_t x)
{
    or_params p;
    or_get_params(ps, &p);
    int r;
    if (!p.sable)
        r = ((int)x + p.h1) >> (p.eq - p.ep);
    else
        r = (((int)x + p.h1) & ((1 << p.eq) - 1)) >> (p.eq - p.ep);
    return (uint16_t)(r & ((1 << p.ep) - 1));
}

uint16_t or_round_enc_v(int ps, uint16_t v, uint16_t mbit)
{
    or_params p;
    or_get_params(ps, &p);
    int r;
    if (!p.sable) {
        r = ((int)v - ((int)(mbit & 1) << (p.ep - 1)) + p.h1) >> (p.ep - p.et);
    } else {
        int pm = (1 << p.ep) - 1;
        int vv = ((int)v + p.h1) & pm;                      /* post_process3 */
        int mm = (int)(mbit & 1) << (p.ep - 1);             /* msg_unpack_post_encode */
        r = ((vv - mm) & pm) >> (p.ep - p.et);              /* post_process4 */
    }
    return (uint16_t)(r & ((1 << p.et) - 1));
}

uint16_t or_round_dec_m(int ps, uint16_t v, uint16_t c)
{
    or_params p;
    or_get_params(ps, &p);
    int r;
    if (!p.sable)
        r = ((int)v + p.h2 - ((int)c << (p.ep - p.et))) >> (p.ep - 1);
    else
        r = (((int)v + p.h2 - ((int)c << (p.ep - p.et))) & ((1 << p.ep) - 1)) >> (p.ep - 1);
    return (uint16_t)(r & 1);
}

/* ------------------------------------------------------------------------ */
/* Codecs.  Every reference codec is an LSB-first bit stream of w-bit words: */
/* 13-bit S/src/poly.cu:799-825, 10-bit S/src/pack.cu:6-32, 4-bit :101-132,  */
/* 1-bit :86-99 and :201-212, 11-bit SB/src/pack.cu:22-38, 9-bit :182-244,   */
/* 5-bit :339-349 and :405-422, 2-bit :366-377.                              */
/* ------------------------------------------------------------------------ */

void or_pack_bits(const uint16_t *data, int n, int w, uint8_t *bytes)
{
    memset(bytes, 0, (size_t)(n * w + 7) / 8);
    for (int i = 0; i < n; i++)
        for (int b = 0; b < w; b++)
            if ((data[i] >> b) & 1) {
                int bit = i * w + b;
                bytes[bit >> 3] |= (uint8_t)(1u << (bit & 7));
            }
}

void or_unpack_bits(const uint8_t *bytes, int n, int w, uint16_t *data)
{
    for (int i = 0; i < n; i++) {
        unsigned v = 0;
        for (int b = 0; b < w; b++) {
            int bit = i * w + b;
            v |= (unsigned)((bytes[bit >> 3] >> (bit & 7)) & 1) << b;
        }
        data[i] = (uint16_t)v;
    }
}

/* ------------------------------------------------------------------------ */
/* Centred binomial sampling.  Coefficient k reads mu bits starting at bit   */
/* k*mu of the stream: popcount of the first mu/2 minus popcount of the     */
/* next mu/2 (cbd_gpu, S/src/poly.cu:836-916 for mu = 6/8/10; SB/src/       */
/* fips202.cu:625-651 for mu = 2).  Saber stores the uint16 two's-complement */
/* value (S/src/poly.cu:868-871), Sable the value masked mod q              */
/* (SB/src/fips202.cu:646-649).                                              */
/* ------------------------------------------------------------------------ */

static int popc_bits(const uint8_t *buf, int start, int count)
{
    int c = 0;
    for (int b = start; b < start + count; b++) c += (buf[b >> 3] >> (b & 7)) & 1;
    return c;
}

void or_cbd(int ps, const uint8_t *buf, uint16_t s[OR_N])
{
    or_params p;
    or_get_params(ps, &p);
    int half = p.mu / 2;
    for (int k = 0; k < OR_N; k++) {
        int v = popc_bits(buf, k * p.mu, half) - popc_bits(buf, k * p.mu + half, half);
        s[k] = p.sable ? (uint16_t)(v & ((1 << p.eq) - 1)) : (uint16_t)v;
    }
}

/* GenMatrix: SHAKE-128(seed) unpacked as L*L EQ-bit polynomials, row-major
 * (shake128_gpu + BS2POLVECq_gpu2, S/src/pack.cu:145-153; SB/src/pack.cu:41-50). */
void or_gen_matrix(int ps, const uint8_t seed[32], uint16_t *A)
{
    or_params p;
    or_get_params(ps, &p);
    uint8_t *buf = (uint8_t *)malloc((size_t)p.matrix_stream_bytes);
    or_shake128(buf, (size_t)p.matrix_stream_bytes, seed, 32);
    int pb = p.eq * OR_N / 8;
    for (int ij = 0; ij < p.l * p.l; ij++) or_unpack_bits(buf + ij * pb, OR_N, p.eq, A + ij * OR_N);
    free(buf);
}

/* GenSecret: SHAKE-128(seed) -> L CBD polynomials (S/src/poly.cu:1062-1073;
 * SB/src/fips202.cu:655-668). */
void or_gen_secret(int ps, const uint8_t seed[32], uint16_t *s)
{
    or_params p;
    or_get_params(ps, &p);
    uint8_t buf[4 * 10 * OR_N / 8];
    or_shake128(buf, (size_t)p.secret_stream_bytes, seed, 32);
    int cb = p.mu * OR_N / 8;
    for (int i = 0; i < p.l; i++) or_cbd(ps, buf + i * cb, s + i * OR_N);
}

/* ------------------------------------------------------------------------ */
/* IND-CPA.  Encryption and decryption follow the step order of saber_enc / */
/* saber_dec (S/src/cuda_kernel.cu:114-178, :341-370) and SABLE_enc /       */
/* SABLE_dec (SB/src/SABLE_indcpa.cu:133-225, :422-463).  Key generation is */
/* absent from the reference; it follows Saber round 3 (seed_A is hashed    */
/* with SHAKE-128, b = round(A^T s)).                                        */
/* ------------------------------------------------------------------------ */

#define MAXL 4

static void unpack_secret_from_sk(const or_params *p, const uint8_t *sk, uint16_t *s)
{
    if (!p->sable) {
        or_unpack_bits(sk, p->l * OR_N, p->eq, s); /* BS2POLVECq_gpu, S/src/pack.cu:136-143 */
    } else {
        uint16_t codes[MAXL * OR_N];
        or_unpack_bits(sk, p->l * OR_N, 2, codes); /* BS2POLp_d_gpu, SB/src/pack.cu:366-377 */
        for (int i = 0; i < p->l * OR_N; i++) s[i] = (uint16_t)((codes[i] ^ 2) - 2);
    }
}

void or_indcpa_keypair(int ps, const uint8_t seed_a[32], const uint8_t seed_s[32], uint8_t *pk,
                       uint8_t *sk)
{
    or_params p;
    or_get_params(ps, &p);
    uint8_t sa[32];
    uint16_t A[MAXL * MAXL * OR_N], s[MAXL * OR_N], b[MAXL * OR_N];
    or_shake128(sa, 32, seed_a, 32);
    or_gen_matrix(ps, sa, A);
    or_gen_secret(ps, seed_s, s);
    memset(b, 0, sizeof(b));
    or_matvec(p.l, A, s, b, 1, 0);
    for (int i = 0; i < p.l * OR_N; i++) b[i] = or_round_q2p(ps, b[i]);
    or_pack_bits(b, p.l * OR_N, p.ep, pk);
    memcpy(pk + p.polyvec_compressed_bytes, sa, 32);
    if (!p.sable) {
        uint16_t sq[MAXL * OR_N];
        for (int i = 0; i < p.l * OR_N; i++) sq[i] = (uint16_t)(s[i] & ((1 << p.eq) - 1));
        or_pack_bits(sq, p.l * OR_N, p.eq, sk);
    } else {
        uint16_t codes[MAXL * OR_N];
        for (int i = 0; i < p.l * OR_N; i++) codes[i] = (uint16_t)(s[i] & 3);
        or_pack_bits(codes, p.l * OR_N, 2, sk);
    }
}

void or_indcpa_enc(int ps, const uint8_t m[32], const uint8_t coins[32], const uint8_t *pk,
                   uint8_t *ct)
{
    or_params p;
    or_get_params(ps, &p);
    uint16_t A[MAXL * MAXL * OR_N], sp[MAXL * OR_N], bp[MAXL * OR_N], b[MAXL * OR_N];
    uint16_t vp[OR_N], mp[OR_N];
    or_gen_matrix(ps, pk + p.polyvec_compressed_bytes, A);
    or_gen_secret(ps, coins, sp);
    memset(bp, 0, sizeof(bp));
    or_matvec(p.l, A, sp, bp, 0, 0);
    for (int i = 0; i < p.l * OR_N; i++) bp[i] = or_round_q2p(ps, bp[i]);
    or_pack_bits(bp, p.l * OR_N, p.ep, ct);
    or_unpack_bits(pk, p.l * OR_N, p.ep, b);
    if (p.sable) /* post_process2, SB/src/pack.cu:248-258 */
        for (int i = 0; i < p.l * OR_N; i++) sp[i] = (uint16_t)(sp[i] & ((1 << p.ep) - 1));
    memset(vp, 0, sizeof(vp));
    or_innerprod(p.l, b, sp, vp, 0);
    or_unpack_bits(m, OR_N, 1, mp);
    for (int i = 0; i < OR_N; i++) vp[i] = or_round_enc_v(ps, vp[i], mp[i]);
    or_pack_bits(vp, OR_N, p.et, ct + p.polyvec_compressed_bytes);
}

void or_indcpa_dec(int ps, const uint8_t *sk, const uint8_t *ct, uint8_t m[32])
{
    or_params p;
    or_get_params(ps, &p);
    uint16_t s[MAXL * OR_N], b[MAXL * OR_N], v[OR_N], cm[OR_N];
    unpack_secret_from_sk(&p, sk, s);
    or_unpack_bits(ct, p.l * OR_N, p.ep, b);
    memset(v, 0, sizeof(v));
    or_innerprod(p.l, b, s, v, 0);
    or_unpack_bits(ct + p.polyvec_compressed_bytes, OR_N, p.et, cm);
    for (int i = 0; i < OR_N; i++) v[i] = or_round_dec_m(ps, v[i], cm[i]);
    or_pack_bits(v, OR_N, 1, m);
}

/* KEM with the Fujisaki-Okamoto transform of Saber round 3.  The reference
 * computes the same hashes in the same order (S/src/cuda_kernel.cu:116-118,
 * :185-186, :378-379, :445-450) but its verify/cmov never reject
 * (S/src/pack.cu:225-255); here a mismatching re-encryption replaces the
 * pre-key with z before the final hash. */
void or_kem_keypair(int ps, const uint8_t seed_a[32], const uint8_t seed_s[32], const uint8_t z[32],
                    uint8_t *pk, uint8_t *sk)
{
    or_params p;
    or_get_params(ps, &p);
    or_indcpa_keypair(ps, seed_a, seed_s, pk, sk);
    memcpy(sk + p.indcpa_sk_bytes, pk, (size_t)p.indcpa_pk_bytes);
    or_sha3_256(sk + p.sk_bytes - 64, pk, (size_t)p.indcpa_pk_bytes);
    memcpy(sk + p.sk_bytes - 32, z, 32);
}

void or_kem_enc(int ps, const uint8_t rnd[32], const uint8_t *pk, uint8_t *ct, uint8_t ss[32])
{
    or_params p;
    or_get_params(ps, &p);
    uint8_t buf[64], kr[64];
    or_sha3_256(buf, rnd, 32);
    or_sha3_256(buf + 32, pk, (size_t)p.indcpa_pk_bytes);
    or_sha3_512(kr, buf, 64);
    or_indcpa_enc(ps, buf, kr + 32, pk, ct);
    or_sha3_256(kr + 32, ct, (size_t)p.ct_bytes);
    or_sha3_256(ss, kr, 64);
}

int or_kem_dec(int ps, const uint8_t *ct, const uint8_t *sk, uint8_t ss[32])
{
    or_params p;
    or_get_params(ps, &p);
    uint8_t buf[64], kr[64], cmp[2048];
    const uint8_t *pk = sk + p.indcpa_sk_bytes;
    or_indcpa_dec(ps, sk, ct, buf);
    memcpy(buf + 32, sk + p.sk_bytes - 64, 32);
    or_sha3_512(kr, buf, 64);
    or_indcpa_enc(ps, buf, kr + 32, pk, cmp);
    uint8_t diff = 0;
    for (int i = 0; i < p.ct_bytes; i++) diff |= (uint8_t)(ct[i] ^ cmp[i]);
    uint8_t mask = (uint8_t)(-(int)((diff | (uint8_t)-diff) >> 7)); /* 0xFF on mismatch */
    or_sha3_256(kr + 32, ct, (size_t)p.ct_bytes);
    for (int i = 0; i < 32; i++) kr[i] ^= mask & (kr[i] ^ sk[p.sk_bytes - 32 + i]);
    or_sha3_256(ss, kr, 64);
    return mask ? 1 : 0;
}

/* ---- batched KEM drivers ---- */

typedef struct {
    int kind, ps;
    const uint8_t *x, *y;
    uint8_t *o1, *o2;
    int64_t lo, hi;
} kjob_t;

static void *run_kjob(void *arg)
{
    kjob_t *j = (kjob_t *)arg;
    or_params p;
    or_get_params(j->ps, &p);
    for (int64_t b = j->lo; b < j->hi; b++) {
        if (j->kind == 0)
            or_kem_keypair(j->ps, j->x + b * 96, j->x + b * 96 + 32, j->x + b * 96 + 64,
                           j->o1 + b * p.pk_bytes, j->o2 + b * p.sk_bytes);
        else if (j->kind == 1)
            or_kem_enc(j->ps, j->x + b * 32, j->y + b * p.pk_bytes, j->o1 + b * p.ct_bytes,
                       j->o2 + b * 32);
        else
            or_kem_dec(j->ps, j->x + b * p.ct_bytes, j->y + b * p.sk_bytes, j->o1 + b * 32);
    }
    return NULL;
}

static void run_kbatch(kjob_t proto, int64_t batch, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (nthreads > batch) nthreads = (int)(batch > 0 ? batch : 1);
    pthread_t th[256];
    kjob_t jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = proto;
        jobs[t].lo = batch * t / nthreads;
        jobs[t].hi = batch * (t + 1) / nthreads;
    }
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, run_kjob, &jobs[t]);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* seeds: batch x 96 bytes (seed_A | seed_s | z) */
void or_kem_keypair_batch(int ps, const uint8_t *seeds, uint8_t *pk, uint8_t *sk, int64_t batch,
                          int nthreads)
{
    kjob_t j = {0, ps, seeds, NULL, pk, sk, 0, 0};
    run_kbatch(j, batch, nthreads);
}

void or_kem_enc_batch(int ps, const uint8_t *rnd, const uint8_t *pk, uint8_t *ct, uint8_t *ss,
                      int64_t batch, int nthreads)
{
    kjob_t j = {1, ps, rnd, pk, ct, ss, 0, 0};
    run_kbatch(j, batch, nthreads);
}

void or_kem_dec_batch(int ps, const uint8_t *ct, const uint8_t *sk, uint8_t *ss, int64_t batch,
                      int nthreads)
{
    kjob_t j = {2, ps, ct, sk, ss, NULL, 0, 0};
    run_kbatch(j, batch, nthreads);
}
