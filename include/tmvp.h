/*
 * tmvp.h - C ABI of the MI355X (gfx950) Saber/Sable TMVP library (libtmvp.so).
 *
 * Drop-in for the polynomial-arithmetic hot path of Muhammad-Asfand/asfand-tmvp:
 * the negacyclic products in Z_q[x]/(x^256+1) behind MatrixVectorMul (A.s and
 * A^T.s) and InnerProd, with the rounding/shift steps around them.  Reference
 * citations are relative to the reference root, with
 *   S/  = TMVP-Saber/SABER-GPU/      SB/ = TMVP_SABLE/
 *
 * Conventions (all entry points):
 *  - Return 0 on success or a negative TMVP_E* code; nothing aborts or exits.
 *  - Device pointers are plain HIP device addresses, owned by the caller.
 *    Work is enqueued on `stream` (NULL = the legacy default stream) and the
 *    call returns without waiting, as the reference's kernel launches do.
 *    The kernels access polynomials 16 bytes and byte strings 4 or 8 bytes
 *    at a time, so device buffers must be 16-byte aligned (hipMalloc's bases
 *    are 256-byte aligned).
 *  - Layouts are the reference's: polynomials are 256 consecutive uint16
 *    coefficients; A is [batch][L][L][256] row-major in (i, j); vectors are
 *    [batch][L][256]; one result polynomial is [batch][256].
 *  - Secret polynomials may be stored either as uint16 two's complement
 *    (Saber, S/src/poly.cu:868-871) or masked mod q (Sable,
 *    SB/src/fips202.cu:646-649): only their value modulo the output modulus
 *    is used.  The MFMA backend needs that value, centred, to lie in
 *    [-128, 127] (every CBD secret does); a violation is reported by
 *    tmvp_ctx_status().
 *  - Outputs are exact.  Raw products are returned reduced modulo the output
 *    modulus 2^W (W = EQ for MatrixVectorMul and polymul, EP for InnerProd):
 *    the reference's uint16 results agree with them mod 2^W, which is the only
 *    part any later step reads (rounding masks it, S/src/pack.cu:28-32).
 *  - Thread-safe per context; a context may be used from one host thread at
 *    a time.
 */
#ifndef TMVP_H
#define TMVP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t without requiring the HIP headers. */
typedef struct ihipStream_t *tmvp_stream_t;

/* ---- status codes ---- */
#define TMVP_OK 0
#define TMVP_EARG (-1)         /* bad argument (null pointer, size, enum) */
#define TMVP_EHIP (-2)         /* HIP runtime error */
#define TMVP_ENOMEM (-3)       /* device or host allocation failed */
#define TMVP_EUNSUPPORTED (-4) /* combination not supported by this backend */
#define TMVP_ERANGE (-5)       /* a secret coefficient was outside the MFMA range */
#define TMVP_ENODEV (-6)       /* no gfx950 device */

/* ---- parameter sets (S/include/params.h:6-19, SB/include/SABLE_params.h:6-24) ---- */
enum tmvp_param_set {
    TMVP_LIGHTSABER = 0, /* L=2, q=2^13, p=2^10, T=2^3, mu=10 */
    TMVP_SABER = 1,      /* L=3, q=2^13, p=2^10, T=2^4, mu=8  */
    TMVP_FIRESABER = 2,  /* L=4, q=2^13, p=2^10, T=2^6, mu=6  */
    TMVP_SABLE1 = 3,     /* K=2, q=2^11, p=2^9,  T=2^3, mu=2  */
    TMVP_SABLE2 = 4,     /* K=3, q=2^11, p=2^9,  T=2^5, mu=2  */
    TMVP_SABLE3 = 5,     /* K=4, q=2^11, p=2^10, T=2^3, mu=2  */
    TMVP_NUM_PARAM_SETS = 6
};

/* ---- kernel forms; the reference's `mode` (S/main.cu:25-33):
 *      0 "TMVP-Tensor" -> TMVP_BACKEND_MFMA (exact int8 split-limb MFMA; the
 *        reference's fp16 WMMA form, S/src/poly.cu:1155-1205, is inexact)
 *      1 "TMVP-CUDA"   -> TMVP_BACKEND_VALU
 *      2 "Schoolbook"  -> TMVP_BACKEND_VALU (same product, one wave per TMVP) */
enum tmvp_backend {
    TMVP_BACKEND_MFMA = 0,
    TMVP_BACKEND_VALU = 1,
    /* An alias of TMVP_BACKEND_VALU, not a third kernel: it runs the same
     * VALU product, and exists so that reference mode 2 maps one to one. */
    TMVP_BACKEND_SCHOOLBOOK = 2
};

/* ---- flags ---- */
#define TMVP_ACCUMULATE 0x1 /* out = (out + result) mod 2^W, the reference's `r += sum` */
/* MFMA backend only: the int8 matrix-core tile, v_mfma_i32_16x16x64_i8 or
 * v_mfma_i32_32x32x32_i8 (results are the same).  Without either flag the
 * library's default tile is used (the 16x16x64 one, which measures faster
 * for every shape here); in a work list every descriptor that names a tile
 * must name the same one, and it overrides the plan's. */
#define TMVP_TILE_16 0x2
#define TMVP_TILE_32 0x4

/* ---- InnerProd epilogues ---- */
enum tmvp_ip_epilogue {
    TMVP_IP_RAW = 0, /* v mod p */
    TMVP_IP_ENC = 1, /* ((v + h1 - (m << (EP-1))) mod p) >> (EP-ET); aux = message bits */
    TMVP_IP_DEC = 2  /* ((v + h2 - (c << (EP-ET))) mod p) >> (EP-1);  aux = ct scale poly */
};

typedef struct {
    int l, eq, ep, et, mu, sable, h1, h2;
    int pk_bytes, sk_bytes, ct_bytes;
} tmvp_params;

typedef struct tmvp_ctx tmvp_ctx;

/* Library and context management (the reference allocates per call and
 * never frees, S/src/cuda_kernel.cu:43-94; a context owns its scratch). */
int tmvp_init(tmvp_ctx **ctx, int device);
int tmvp_destroy(tmvp_ctx *ctx);
/* Waits for `stream` and returns TMVP_ERANGE if any MFMA launch since the
 * last call saw an out-of-range secret coefficient, or if a one-launch KEM
 * call's internal wait between its waves ran out (a bound of about 0.2 s that
 * correct launches never reach; that call's outputs are then not valid).
 * Clears the flag. */
int tmvp_ctx_status(tmvp_ctx *ctx, tmvp_stream_t stream);
const char *tmvp_strerror(int code);
int tmvp_get_params(int param_set, tmvp_params *out);
int tmvp_version(void);

/* Raw negacyclic products c[b] = a[b] * s[b] mod (x^256+1, 2^eq_bits).
 * Replaces one TMVP of the reference: the convertnegacyclictest ->
 * submatrix_cuda -> matvecp_cuda -> matvecout_cuda chain
 * (S/src/poly.cu:1116-1129, :1295-1392) and the per-(i,j) triangle loops of
 * MatVecMul_gpu_shared (S/src/poly.cu:281-284).  eq_bits in [1,16] (VALU) or
 * [8,14] (MFMA). */
int tmvp_polymul(tmvp_ctx *ctx, int backend, uint16_t *d_c, const uint16_t *d_a,
                 const uint16_t *d_s, int64_t batch, int eq_bits, int flags,
                 tmvp_stream_t stream);

/* MatrixVectorMul: out[b][i] = sum_j A[b][i][j] * s[b][j]            (transpose = 0)
 *                  out[b][i] = sum_j A[b][j][i] * s[b][j]            (transpose = 1)
 * Replaces MatVecMul_gpu_shared (S/src/poly.cu:269-376; SB/src/pack.cu:54-163),
 * launched at S/src/cuda_kernel.cu:129, :411 and SB/src/SABLE_indcpa.cu:186,
 * :506, and the mode-1 chain S/src/cuda_kernel.cu:389-408.  The reference has
 * no transposed form (key generation); it is the one its test-vector public
 * keys were made with.  round_to_p != 0 fuses post_process
 * (S/src/poly.cu:381-389; SB/src/pack.cu:167-178): out = ((x + h1) mod q) >> (EQ-EP).
 * Otherwise out = x mod q (TMVP_ACCUMULATE adds into out first). */
int tmvp_matvec(tmvp_ctx *ctx, int param_set, int backend, int transpose, uint16_t *d_out,
                const uint16_t *d_A, const uint16_t *d_s, int64_t batch, int round_to_p,
                int flags, tmvp_stream_t stream);

/* InnerProd: v[b] = sum_j b[b][j] * s[b][j] mod p, then the epilogue.
 * Replaces VecVecMul_gpu_shared (S/src/poly.cu:227-248, launched
 * S/src/cuda_kernel.cu:162, :366, :438), VecVecMul_Inner_gpu
 * (SB/src/pack.cu:262-286) and the mode-1 chain S/src/cuda_kernel.cu:150-159;
 * the epilogues fuse post_process2 (S/src/poly.cu:391-395; SB/src/pack.cu:290-335)
 * and post_process3 (S/src/poly.cu:398-403; SB/src/pack.cu:436-442).
 * d_aux is [batch][256] (message bits for TMVP_IP_ENC, scale polynomial for
 * TMVP_IP_DEC) and may be NULL for TMVP_IP_RAW. */
int tmvp_innerprod(tmvp_ctx *ctx, int param_set, int backend, int epilogue, uint16_t *d_out,
                   const uint16_t *d_b, const uint16_t *d_s, const uint16_t *d_aux,
                   int64_t batch, int flags, tmvp_stream_t stream);

/* ---- mixed batches: one persistent launch over many jobs ----
 * The reference launches one kernel per product and parameter set with one
 * block per instance (S/src/cuda_kernel.cu:129, :162; SB/src/SABLE_indcpa.cu:186,
 * :217).  A work list instead runs any mix of products, parameter sets and
 * epilogues in ONE launch of a persistent MFMA kernel: its workgroups are
 * split between the jobs in proportion to their work, with the split and
 * the workgroups per CU taken from a plan that tmvp_batch_autotune() measures
 * for the list's shape (kinds and batch sizes); without one, the split follows
 * a per-instance cost model (terms x (rows + 2)) at full occupancy.  The VALU backends issue one
 * launch per job.  Each descriptor has the meaning of the matching single
 * call: tmvp_polymul (eq_bits), tmvp_matvec (epilogue 0 = mod q, 1 = round to
 * p) or tmvp_innerprod (epilogue = tmvp_ip_epilogue). */
enum tmvp_op {
    TMVP_OP_POLYMUL = 0,
    TMVP_OP_MATVEC = 1,
    TMVP_OP_MATVEC_T = 2,
    TMVP_OP_INNERPROD = 3
};
#define TMVP_MAX_WORK 16

typedef struct {
    int op;        /* tmvp_op */
    int param_set; /* not read for TMVP_OP_POLYMUL */
    int epilogue;
    int eq_bits;   /* TMVP_OP_POLYMUL only */
    int flags;     /* TMVP_ACCUMULATE (raw outputs only) */
    int64_t batch;
    uint16_t *out;
    const uint16_t *a; /* A, b or a */
    const uint16_t *s;
    const uint16_t *aux; /* InnerProd ENC / DEC */
} tmvp_work_desc;

/* Every descriptor is checked before anything is launched. */
int tmvp_batch_launch(tmvp_ctx *ctx, int backend, const tmvp_work_desc *work, int n, tmvp_stream_t stream);

/* Times candidate plans for this work list on its own buffers (outputs are
 * overwritten; TMVP_ACCUMULATE is refused), keeps the fastest in the context
 * for later launches of lists of the same shape, and returns its time per
 * launch in *best_ms (may be NULL).  The candidates cover the MFMA tile shape
 * (16x16x64 or 32x32x32 int8 tiles), the workgroups per CU and the split
 * between the jobs.  Synchronises `stream`. */
int tmvp_batch_autotune(tmvp_ctx *ctx, const tmvp_work_desc *work, int n, tmvp_stream_t stream, float *best_ms);

/* The plan launches of this work list's shape use: the MFMA tile shape
 * (16 = v_mfma_i32_16x16x64_i8 tiles, 32 = v_mfma_i32_32x32x32_i8 tiles) and
 * the workgroups per CU.  Returns TMVP_OK for a plan tmvp_batch_autotune
 * measured, 1 for the default plan (cost-model split, full occupancy). */
int tmvp_batch_plan(tmvp_ctx *ctx, const tmvp_work_desc *work, int n, int *shape, int *wgs_per_cu);

/* ---- hipGraph capture of a sequence of library calls on one stream ----
 * Between tmvp_graph_begin and tmvp_graph_end every tmvp call on `stream` is
 * recorded instead of run; the graph replays it with the same buffers.  A
 * KEM call whose workspace would have to grow fails with TMVP_EUNSUPPORTED
 * during capture: reserve it first with tmvp_kem_reserve.  The legacy NULL
 * stream cannot be captured.  A graph holds the KEM workspace buffer of the
 * stream it was captured on until tmvp_graph_destroy: later calls that grow
 * that workspace, or hand it to another stream, move to a new buffer and
 * leave the graph's in place.  The graph shares the buffer with eager KEM
 * calls on the capture stream, so replay it on that stream (or order the
 * two with events): a replay overlapping such a call races on the scratch. */
typedef struct tmvp_graph tmvp_graph;
int tmvp_graph_begin(tmvp_ctx *ctx, tmvp_stream_t stream);
int tmvp_graph_end(tmvp_ctx *ctx, tmvp_stream_t stream, tmvp_graph **graph);
int tmvp_graph_launch(tmvp_graph *graph, tmvp_stream_t stream);
int tmvp_graph_destroy(tmvp_graph *graph);

/* ---- batched KEM (SHAKE-128 GenMatrix and CBD GenSecret in the sponges,
 * MatrixVectorMul / InnerProd with their rounding, codecs, Fujisaki-Okamoto
 * hashes).  With TMVP_BACKEND_MFMA each call is one launch in which the
 * products and hashes overlap the matrix sponge (64 instances per workgroup
 * of 16 waves); with TMVP_BACKEND_VALU a chain of kernels.  Both give the
 * same bytes. ----
 * Byte layouts are the reference's: pk = pack_EP(b) | seed_A,
 * sk = s | pk | H(pk) | z (Saber s as 13-bit words, Sable as 2-bit codes),
 * ct = pack_EP(b') | pack_ET(v').  Buffers are [batch][size] device arrays.
 * Scratch is owned by the context, one workspace per stream; it grows on
 * demand (which synchronises the device once) and is reused afterwards. */

/* Key generation from per-instance seeds [batch][96] = seed_A | seed_s | z.
 * The reference has no key generation; this is Saber round 3
 * (seed_A' = SHAKE-128(seed_A), b = round(A^T s)), the form the reference's
 * test-vector keys (S/include/tv.h:8-10, SB/include/tv.h:15-19) satisfy. */
int tmvp_kem_keypair(tmvp_ctx *ctx, int param_set, int backend, uint8_t *d_pk, uint8_t *d_sk,
                     const uint8_t *d_seeds, int64_t batch, tmvp_stream_t stream);

/* Encapsulation with per-instance randomness d_rnd [batch][32].
 * Replaces saber_enc's device chain (S/src/cuda_kernel.cu:114-186) and
 * SABLE_enc's (SB/src/SABLE_indcpa.cu:133-233): m = H(rnd),
 * kr = G(m | H(pk)), ct = Enc(pk, m; kr[32:]), ss = H(kr[0:32] | H(ct)). */
int tmvp_kem_encaps(tmvp_ctx *ctx, int param_set, int backend, uint8_t *d_ct, uint8_t *d_ss,
                    const uint8_t *d_pk, const uint8_t *d_rnd, int64_t batch, tmvp_stream_t stream);

/* Decapsulation.  Replaces saber_dec's device chain
 * (S/src/cuda_kernel.cu:341-450) and SABLE_dec's (SB/src/SABLE_indcpa.cu:422-563),
 * with a working implicit rejection: if the re-encryption differs from ct,
 * ss = H(z | H(ct)).  d_fail (optional, [batch] bytes) receives 1 for a
 * rejected ciphertext. */
int tmvp_kem_decaps(tmvp_ctx *ctx, int param_set, int backend, uint8_t *d_ss, const uint8_t *d_ct,
                    const uint8_t *d_sk, uint8_t *d_fail, int64_t batch, tmvp_stream_t stream);

/* Encapsulation that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as tmvp_kem_encaps, but its scratch is a fixed
 * 32 MiB device array of the library (one per GPU, held from load until the
 * process ends) instead of the context's workspace: the call makes no
 * hipMalloc / hipFree, creates no stream or event, and runs on `stream` only.
 * A batch must fit the array: at most 5318 instances for L = 2 (LightSaber,
 * Sable-1), 3224 for L = 3 (Saber, Sable-2) and 2161 for L = 4 (FireSaber,
 * Sable-3); a larger batch returns TMVP_ENOMEM without running anything.
 * Every call on one GPU uses that GPU's one array: calls that overlap on a
 * GPU (from any context, thread or stream) race on it. */
int tmvp_kem_encaps_noalloc(tmvp_ctx *ctx, int param_set, int backend, uint8_t *d_ct, uint8_t *d_ss,
                            const uint8_t *d_pk, const uint8_t *d_rnd, int64_t batch, tmvp_stream_t stream);

/* Grow the KEM workspace of `stream` to `batch` instances of `param_set` now
 * (synchronises the device), so later KEM calls on it allocate nothing. */
int tmvp_kem_reserve(tmvp_ctx *ctx, int param_set, int64_t batch, tmvp_stream_t stream);

/* ---- serving: continuous batching of KEM requests over resident keys ----
 * Not in the reference (its host functions run one fixed batch per call).
 * Keys live in HBM per parameter set, with the matrix A and H(pk) derived
 * once per key.  Requests (one operation, one parameter set, 1..max_batch
 * instances, each under its own key id) are queued per (set, operation); a
 * scheduler thread forms each batch from the queue whose oldest request has
 * waited longest, in arrival order up to max_batch instances and up to the
 * HBM budget left, and runs it on a free slot (a stream with its own pinned
 * staging).  Each batch reserves exactly the device bytes its parameter set,
 * operation and size need (workspace, key ids, inputs, outputs), and the key
 * stores count against the same budget: small batches run side by side where
 * one large batch would take the budget, and a request that does not fit yet
 * waits (nothing behind it overtakes it).  Batches on different slots
 * overlap on the device.  Host buffers of a request must stay
 * valid until its completion is polled; buffers in page-locked memory from
 * hipHostMalloc are read and written by the device in place; other buffers
 * go through the server's pinned staging, copied by the scheduler thread.  Results are those of
 * tmvp_kem_encaps / tmvp_kem_decaps for the same keys and inputs, whatever
 * the batch a request lands in.  Run the serving process with
 * GPU_MAX_HW_QUEUES=8 (HIP's default is 4) so that every slot's stream has
 * a hardware queue of its own; with four queues latency depends on how the
 * runtime places the process's streams on them. */
typedef struct tmvp_server tmvp_server;

typedef struct {
    int64_t hbm_budget; /* device bytes for the key stores and the batches in
                           flight together; 0 = 4 GiB.  Must hold one batch of
                           max_batch instances of the largest set, else
                           tmvp_srv_create returns TMVP_ENOMEM. */
    int64_t max_batch;  /* instances per batch; 0 = 16384 */
    int slots;          /* concurrent batches (streams) at most; 0 = 4 */
    int backend;        /* tmvp_backend */
} tmvp_srv_config;

#define TMVP_SRV_ENCAPS 0
#define TMVP_SRV_DECAPS 1

typedef struct {
    int op;              /* TMVP_SRV_ENCAPS / TMVP_SRV_DECAPS */
    int param_set;
    int64_t count;       /* instances */
    const int32_t *keys; /* [count] key ids of this parameter set */
    const uint8_t *in;   /* randomness [count][32] (encaps) / ciphertexts [count][ct_bytes] (decaps) */
    uint8_t *ct;         /* encaps output [count][ct_bytes] */
    uint8_t *ss;         /* output [count][32] */
    uint8_t *fail;       /* decaps output [count] (1 = rejected), may be NULL */
    uint64_t tag;        /* returned in the completion */
} tmvp_srv_request;

typedef struct {
    uint64_t id, tag;
    int status;              /* TMVP_OK or an error code */
    int batch_requests;      /* requests in the batch that carried it */
    int64_t batch_instances; /* instances in that batch */
    double submit_us, start_us, done_us; /* server clock */
} tmvp_srv_completion;

typedef struct {
    int64_t requests_submitted, requests_done, instances_done, batches;
    int64_t slots;              /* streams: batches in flight at most */
    int64_t slot_bytes;         /* device bytes of one max_batch batch of the largest set */
    int64_t max_concurrent;     /* most batches in flight at once so far */
    int64_t hbm_budget;         /* the budget the key stores and batches share */
    int64_t store_bytes;        /* device bytes the key stores hold */
    int64_t arena_bytes;        /* device bytes reserved for batches */
    int64_t max_inflight_bytes; /* most batch bytes in flight at once so far */
} tmvp_srv_stats;

int tmvp_srv_create(tmvp_ctx *ctx, const tmvp_srv_config *config, tmvp_server **server);
/* Appends n keys (sk may be NULL: encapsulation only); ids first_id..
 * When h_sk is given, sk i must hold pk i (sk = s | pk | H(pk) | z), else
 * TMVP_EARG: keyed decapsulation takes A from the store's pk and b, H(pk)
 * and z from the sk.  TMVP_ENOMEM (nothing added) when the grown store
 * would leave the HBM budget less than one max_batch batch.  Waits for the
 * batches in flight. */
int tmvp_srv_add_keys(tmvp_server *server, int param_set, const uint8_t *h_pk, const uint8_t *h_sk, int64_t n,
                      int64_t *first_id);
/* Checks the request's key ids and copies them: the caller may reuse
 * request->keys as soon as this returns (the in/out buffers must stay valid). */
int tmvp_srv_submit(tmvp_server *server, const tmvp_srv_request *request, uint64_t *id);
/* Up to max completions, waiting up to timeout_us for the first; returns the count. */
int tmvp_srv_poll(tmvp_server *server, tmvp_srv_completion *out, int max, int timeout_us);
int tmvp_srv_drain(tmvp_server *server);
int tmvp_srv_get_stats(tmvp_server *server, tmvp_srv_stats *out);
int tmvp_srv_destroy(tmvp_server *server);

/* ---- the reference's host entry points (same names and signatures) ----
 * saber_enc / saber_dec (S/include/cuda_kernel.cuh:6-7) and SABLE_enc /
 * SABLE_dec (SB/include/SABLE_indcpa.cuh:8-10) over the batched KEM above.
 * The reference fixes the batch (BATCH: 1 for Saber, S/include/cuda_kernel.cuh:1;
 * 512 for Sable, SB/include/SABLE_indcpa.cuh:6) and the parameter set at compile
 * time and reads its keys and randomness from tv.h; here the batch and set
 * come from tmvp_host_configure() (defaults: device 0, Saber L=3 x 1,
 * Sable type 2 x 512), keys from the caller's buffers, and randomness from
 * tmvp_host_set_randomness() (used by the next encapsulation only) or else
 * from the operating system.  mode: 0 = MFMA, 1 = VALU, 2 = schoolbook VALU
 * (S/main.cu:25-33).  Host buffers are [batch][size]: h_pk pk_bytes, h_sk
 * sk_bytes, h_c ct_bytes, h_k 32.  h_pk is not read by the decapsulations (the
 * sk holds the public key).  The functions return nothing, as in the
 * reference; tmvp_host_status() gives the status of the last one.  Calls are
 * serialised on one library-owned context and stream. */
int tmvp_host_configure(int device, int64_t saber_batch, int saber_set, int64_t sable_batch, int sable_set);
int tmvp_host_set_randomness(const uint8_t *h_rnd, int64_t batch);
int tmvp_host_status(void);
int tmvp_host_release(void);
void saber_enc(uint8_t mode, uint8_t *h_k, uint8_t *h_pk, uint8_t *h_c);
void saber_dec(uint8_t mode, uint8_t *h_c, uint8_t *h_pk, uint8_t *h_sk, uint8_t *h_k);
void SABLE_enc(uint8_t *h_k, uint8_t *h_pk, uint8_t *h_c, int mode);
void SABLE_dec(uint8_t *h_c, uint8_t *h_pk, uint8_t *h_sk, uint8_t *h_k, int mode);

#ifdef __cplusplus
}
#endif
#endif /* TMVP_H */
