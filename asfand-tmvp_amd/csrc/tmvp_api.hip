// C-ABI host layer of libtmvp.so: argument checks, parameter-set table,
// backend dispatch.  Every entry point returns a status code; none aborts.
#include <string.h>

#include <new>

#include "../../include/tmvp.h"
#include "tmvp_internal.h"
#include "tmvp_kem.h"

using namespace tmvp;

namespace {
constexpr int kMaxWorkspaces = 16;
struct Workspace {
    hipStream_t stream;
    void *ptr;
    size_t bytes;
};
} // namespace

struct tmvp_ctx {
    int device;
    int *d_status; // device word set by MFMA launches on out-of-range secrets
    Workspace ws[kMaxWorkspaces];
    int nws;
};

namespace {

struct PS {
    int l, eq, ep, et, mu, sable;
};

// S/include/params.h:6-24, SB/include/SABLE_params.h:6-30
constexpr PS kParams[TMVP_NUM_PARAM_SETS] = {
    {2, 13, 10, 3, 10, 0}, {3, 13, 10, 4, 8, 0}, {4, 13, 10, 6, 6, 0},
    {2, 11, 9, 3, 2, 1},   {3, 11, 9, 5, 2, 1},  {4, 11, 10, 3, 2, 1},
};

inline int h1_of(const PS &p) { return 1 << (p.eq - p.ep - 1); }
inline int h2_of(const PS &p) { return (1 << (p.ep - 2)) - (1 << (p.ep - p.et - 1)) + h1_of(p); }

int hip_status(hipError_t e) { return e == hipSuccess ? TMVP_OK : TMVP_EHIP; }

int set_device(const tmvp_ctx *ctx)
{
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return TMVP_EHIP;
    if (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess) return TMVP_EHIP;
    return TMVP_OK;
}

int run(tmvp_ctx *ctx, int backend, const Job &J, int mode, int l, tmvp_stream_t stream)
{
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    switch (backend) {
    case TMVP_BACKEND_MFMA:
        if (J.w < 8 || J.w > 14) return TMVP_EUNSUPPORTED;
        return hip_status(launch_mfma(J, mode, l, st));
    case TMVP_BACKEND_VALU:
    case TMVP_BACKEND_SCHOOLBOOK:
        return hip_status(launch_valu(J, mode, l, st));
    default:
        return TMVP_EARG;
    }
}

KemParams kem_params(int ps)
{
    const PS &p = kParams[ps];
    KemParams P;
    P.l = p.l;
    P.eq = p.eq;
    P.ep = p.ep;
    P.et = p.et;
    P.mu = p.mu;
    P.sable = p.sable;
    P.h1 = h1_of(p);
    P.h2 = h2_of(p);
    P.pvc = p.l * p.ep * 32;
    P.indcpa_sk = p.sable ? p.l * 64 : p.l * p.eq * 32;
    P.pk = P.pvc + 32;
    P.sk = P.indcpa_sk + P.pk + 64;
    P.ct = P.pvc + p.et * 32;
    P.araw = p.l * p.l * p.eq * 32;
    return P;
}

// Workspace of at least `bytes` for `stream`; growing it synchronises the device.
int workspace(tmvp_ctx *ctx, hipStream_t stream, size_t bytes, uint8_t **out)
{
    Workspace *w = nullptr;
    for (int i = 0; i < ctx->nws; i++)
        if (ctx->ws[i].stream == stream) w = &ctx->ws[i];
    if (!w) {
        if (ctx->nws == kMaxWorkspaces) return TMVP_ENOMEM;
        w = &ctx->ws[ctx->nws++];
        w->stream = stream;
        w->ptr = nullptr;
        w->bytes = 0;
    }
    if (w->bytes < bytes) {
        if (hipDeviceSynchronize() != hipSuccess) return TMVP_EHIP;
        if (w->ptr) (void)hipFree(w->ptr);
        w->ptr = nullptr;
        w->bytes = 0;
        if (hipMalloc(&w->ptr, bytes) != hipSuccess) return TMVP_ENOMEM;
        w->bytes = bytes;
    }
    *out = static_cast<uint8_t *>(w->ptr);
    return TMVP_OK;
}

bool kem_args_ok(const tmvp_ctx *ctx, int ps, int backend, int64_t batch)
{
    return ctx && ps >= 0 && ps < TMVP_NUM_PARAM_SETS && batch >= 0 &&
           (backend == TMVP_BACKEND_MFMA || backend == TMVP_BACKEND_VALU || backend == TMVP_BACKEND_SCHOOLBOOK);
}

int kem_setup(tmvp_ctx *ctx, int ps, int64_t batch, tmvp_stream_t stream, KemParams &P, KemWorkspace &W)
{
    int rc = set_device(ctx);
    if (rc) return rc;
    P = kem_params(ps);
    uint8_t *base = nullptr;
    rc = workspace(ctx, reinterpret_cast<hipStream_t>(stream), kem_workspace_bytes(P, batch), &base);
    if (rc) return rc;
    kem_workspace_carve(P, batch, base, W);
    return TMVP_OK;
}

Job base_job(tmvp_ctx *ctx, uint16_t *out, const uint16_t *a, const uint16_t *s, int64_t batch)
{
    Job J;
    memset(&J, 0, sizeof(J));
    J.out = out;
    J.a = a;
    J.s = s;
    J.aux = nullptr;
    J.status = ctx->d_status;
    J.batch = batch;
    J.epi = kEpiMod;
    return J;
}

} // namespace

extern "C" {

int tmvp_version(void) { return 1; }

const char *tmvp_strerror(int code)
{
    switch (code) {
    case TMVP_OK: return "ok";
    case TMVP_EARG: return "invalid argument";
    case TMVP_EHIP: return "HIP runtime error";
    case TMVP_ENOMEM: return "out of memory";
    case TMVP_EUNSUPPORTED: return "unsupported by this backend";
    case TMVP_ERANGE: return "secret coefficient outside the int8 MFMA range";
    case TMVP_ENODEV: return "no gfx950 device";
    default: return "unknown error";
    }
}

int tmvp_get_params(int ps, tmvp_params *out)
{
    if (ps < 0 || ps >= TMVP_NUM_PARAM_SETS || !out) return TMVP_EARG;
    const PS &p = kParams[ps];
    out->l = p.l;
    out->eq = p.eq;
    out->ep = p.ep;
    out->et = p.et;
    out->mu = p.mu;
    out->sable = p.sable;
    out->h1 = h1_of(p);
    out->h2 = h2_of(p);
    const int pvc = p.l * p.ep * 32;
    out->pk_bytes = pvc + 32;
    out->sk_bytes = (p.sable ? p.l * 64 : p.l * p.eq * 32) + out->pk_bytes + 64;
    out->ct_bytes = pvc + p.et * 32;
    return TMVP_OK;
}

int tmvp_init(tmvp_ctx **ctx, int device)
{
    if (!ctx) return TMVP_EARG;
    *ctx = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return TMVP_ENODEV;
    if (device < 0 || device >= count) return TMVP_EARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return TMVP_EHIP;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return TMVP_ENODEV;
    if (hipSetDevice(device) != hipSuccess) return TMVP_EHIP;
    tmvp_ctx *c = new (std::nothrow) tmvp_ctx;
    if (!c) return TMVP_ENOMEM;
    c->device = device;
    c->d_status = nullptr;
    c->nws = 0;
    if (hipMalloc(&c->d_status, sizeof(int)) != hipSuccess) {
        delete c;
        return TMVP_ENOMEM;
    }
    if (hipMemset(c->d_status, 0, sizeof(int)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(c->d_status);
        delete c;
        return TMVP_EHIP;
    }
    *ctx = c;
    return TMVP_OK;
}

int tmvp_destroy(tmvp_ctx *ctx)
{
    if (!ctx) return TMVP_EARG;
    int rc = set_device(ctx);
    if (hipDeviceSynchronize() != hipSuccess) rc = TMVP_EHIP;
    for (int i = 0; i < ctx->nws; i++)
        if (ctx->ws[i].ptr) (void)hipFree(ctx->ws[i].ptr);
    if (ctx->d_status) (void)hipFree(ctx->d_status);
    delete ctx;
    return rc;
}

int tmvp_ctx_status(tmvp_ctx *ctx, tmvp_stream_t stream)
{
    if (!ctx) return TMVP_EARG;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int flag = 0;
    if (hipMemcpyAsync(&flag, ctx->d_status, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess)
        return TMVP_EHIP;
    if (hipStreamSynchronize(st) != hipSuccess) return TMVP_EHIP;
    if (flag) {
        if (hipMemsetAsync(ctx->d_status, 0, sizeof(int), st) != hipSuccess) return TMVP_EHIP;
        if (hipStreamSynchronize(st) != hipSuccess) return TMVP_EHIP;
        return TMVP_ERANGE;
    }
    return TMVP_OK;
}

int tmvp_polymul(tmvp_ctx *ctx, int backend, uint16_t *d_c, const uint16_t *d_a, const uint16_t *d_s,
                 int64_t batch, int eq_bits, int flags, tmvp_stream_t stream)
{
    if (!ctx || batch < 0 || eq_bits < 1 || eq_bits > 16 || (flags & ~TMVP_ACCUMULATE)) return TMVP_EARG;
    if (batch == 0) return TMVP_OK;
    if (!d_c || !d_a || !d_s) return TMVP_EARG;
    Job J = base_job(ctx, d_c, d_a, d_s, batch);
    J.w = eq_bits;
    J.accumulate = (flags & TMVP_ACCUMULATE) ? 1 : 0;
    return run(ctx, backend, J, kPoly, 1, stream);
}

int tmvp_matvec(tmvp_ctx *ctx, int ps, int backend, int transpose, uint16_t *d_out, const uint16_t *d_A,
                const uint16_t *d_s, int64_t batch, int round_to_p, int flags, tmvp_stream_t stream)
{
    if (!ctx || ps < 0 || ps >= TMVP_NUM_PARAM_SETS || batch < 0 || (flags & ~TMVP_ACCUMULATE)) return TMVP_EARG;
    if (round_to_p && (flags & TMVP_ACCUMULATE)) return TMVP_EARG;
    if (batch == 0) return TMVP_OK;
    if (!d_out || !d_A || !d_s) return TMVP_EARG;
    const PS &p = kParams[ps];
    Job J = base_job(ctx, d_out, d_A, d_s, batch);
    J.w = p.eq;
    J.eq = p.eq;
    J.ep = p.ep;
    J.et = p.et;
    J.accumulate = (flags & TMVP_ACCUMULATE) ? 1 : 0;
    if (round_to_p) {
        J.epi = kEpiQ2P;
        J.h = h1_of(p);
    }
    return run(ctx, backend, J, transpose ? kMatVecT : kMatVec, p.l, stream);
}

int tmvp_innerprod(tmvp_ctx *ctx, int ps, int backend, int epilogue, uint16_t *d_out, const uint16_t *d_b,
                   const uint16_t *d_s, const uint16_t *d_aux, int64_t batch, int flags, tmvp_stream_t stream)
{
    if (!ctx || ps < 0 || ps >= TMVP_NUM_PARAM_SETS || batch < 0 || (flags & ~TMVP_ACCUMULATE)) return TMVP_EARG;
    if (epilogue < TMVP_IP_RAW || epilogue > TMVP_IP_DEC) return TMVP_EARG;
    if (epilogue != TMVP_IP_RAW && (flags & TMVP_ACCUMULATE)) return TMVP_EARG;
    if (batch == 0) return TMVP_OK;
    if (!d_out || !d_b || !d_s || (epilogue != TMVP_IP_RAW && !d_aux)) return TMVP_EARG;
    const PS &p = kParams[ps];
    Job J = base_job(ctx, d_out, d_b, d_s, batch);
    J.w = p.ep;
    J.eq = p.eq;
    J.ep = p.ep;
    J.et = p.et;
    J.accumulate = (flags & TMVP_ACCUMULATE) ? 1 : 0;
    if (epilogue == TMVP_IP_ENC) {
        J.epi = kEpiEnc;
        J.h = h1_of(p);
        J.aux = d_aux;
    } else if (epilogue == TMVP_IP_DEC) {
        J.epi = kEpiDec;
        J.h = h2_of(p);
        J.aux = d_aux;
    }
    return run(ctx, backend, J, kInner, p.l, stream);
}

int tmvp_kem_keypair(tmvp_ctx *ctx, int ps, int backend, uint8_t *d_pk, uint8_t *d_sk, const uint8_t *d_seeds,
                     int64_t batch, tmvp_stream_t stream)
{
    if (!kem_args_ok(ctx, ps, backend, batch)) return TMVP_EARG;
    if (batch == 0) return TMVP_OK;
    if (!d_pk || !d_sk || !d_seeds) return TMVP_EARG;
    KemParams P;
    KemWorkspace W;
    int rc = kem_setup(ctx, ps, batch, stream, P, W);
    if (rc) return rc;
    return hip_status(kem_keypair(P, W, backend == TMVP_BACKEND_MFMA ? 0 : 1, d_seeds, d_pk, d_sk, batch,
                                  ctx->d_status, reinterpret_cast<hipStream_t>(stream)));
}

int tmvp_kem_encaps(tmvp_ctx *ctx, int ps, int backend, uint8_t *d_ct, uint8_t *d_ss, const uint8_t *d_pk,
                    const uint8_t *d_rnd, int64_t batch, tmvp_stream_t stream)
{
    if (!kem_args_ok(ctx, ps, backend, batch)) return TMVP_EARG;
    if (batch == 0) return TMVP_OK;
    if (!d_ct || !d_ss || !d_pk || !d_rnd) return TMVP_EARG;
    KemParams P;
    KemWorkspace W;
    int rc = kem_setup(ctx, ps, batch, stream, P, W);
    if (rc) return rc;
    return hip_status(kem_encaps(P, W, backend == TMVP_BACKEND_MFMA ? 0 : 1, d_pk, d_rnd, d_ct, d_ss, batch,
                                 ctx->d_status, reinterpret_cast<hipStream_t>(stream)));
}

int tmvp_kem_decaps(tmvp_ctx *ctx, int ps, int backend, uint8_t *d_ss, const uint8_t *d_ct, const uint8_t *d_sk,
                    uint8_t *d_fail, int64_t batch, tmvp_stream_t stream)
{
    if (!kem_args_ok(ctx, ps, backend, batch)) return TMVP_EARG;
    if (batch == 0) return TMVP_OK;
    if (!d_ss || !d_ct || !d_sk) return TMVP_EARG;
    KemParams P;
    KemWorkspace W;
    int rc = kem_setup(ctx, ps, batch, stream, P, W);
    if (rc) return rc;
    return hip_status(kem_decaps(P, W, backend == TMVP_BACKEND_MFMA ? 0 : 1, d_ct, d_sk, d_ss, d_fail, batch,
                                 ctx->d_status, reinterpret_cast<hipStream_t>(stream)));
}

} // extern "C"
