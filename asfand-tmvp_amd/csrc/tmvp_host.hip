// The reference's host entry points, same names and signatures, over the
// batched GPU KEM.  The reference (S/include/cuda_kernel.cuh:6-7,
// SB/include/SABLE_indcpa.cuh:8-10) takes its batch size and parameter set
// from compile-time macros (BATCH = 1 for Saber, 512 for Sable type 2) and
// its randomness from a test vector; here they come from
// tmvp_host_configure() / tmvp_host_set_randomness(), with the reference's
// values as defaults and fresh OS randomness per call.  The functions are
// void, as in the reference; tmvp_host_status() returns the status of the
// last call.  Host buffers are [batch][size] arrays in the reference layouts.
#include <hip/hip_runtime.h>
#include <string.h>
#include <sys/random.h>

#include <mutex>

#include "../../include/tmvp.h"

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    int reserve(size_t bytes)
    {
        if (n >= bytes) return TMVP_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) return TMVP_ENOMEM;
        n = bytes;
        return TMVP_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    uint8_t *u8() const { return static_cast<uint8_t *>(p); }
};

struct HostState {
    std::mutex mu;
    int device = 0;
    int64_t saber_batch = 1, sable_batch = 512;
    int saber_set = TMVP_SABER, sable_set = TMVP_SABLE2;
    tmvp_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    DevBuf a, b, c, d, r;
    uint8_t *rnd = nullptr; // caller randomness for the next encapsulation
    size_t rnd_bytes = 0;
    int status = TMVP_OK;
};

HostState &state()
{
    static HostState s;
    return s;
}

int ensure_ctx(HostState &S)
{
    if (S.ctx) return TMVP_OK;
    int rc = tmvp_init(&S.ctx, S.device);
    if (rc) {
        S.ctx = nullptr;
        return rc;
    }
    if (hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking) != hipSuccess) {
        tmvp_destroy(S.ctx);
        S.ctx = nullptr;
        return TMVP_EHIP;
    }
    return TMVP_OK;
}

int backend_of(int mode)
{
    switch (mode) {
    case 0: return TMVP_BACKEND_MFMA;
    case 1: return TMVP_BACKEND_VALU;
    case 2: return TMVP_BACKEND_SCHOOLBOOK;
    default: return -1;
    }
}

int fill_random(uint8_t *dst, size_t n)
{
    size_t got = 0;
    while (got < n) {
        ssize_t k = getrandom(dst + got, n - got, 0);
        if (k <= 0) return TMVP_EARG;
        got += (size_t)k;
    }
    return TMVP_OK;
}

int copy_in(void *dst, const void *src, size_t n, hipStream_t st)
{
    return hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st) == hipSuccess ? TMVP_OK : TMVP_EHIP;
}

int copy_out(void *dst, const void *src, size_t n, hipStream_t st)
{
    return hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, st) == hipSuccess ? TMVP_OK : TMVP_EHIP;
}

int finish(HostState &S)
{
    if (hipStreamSynchronize(S.stream) != hipSuccess) return TMVP_EHIP;
    return tmvp_ctx_status(S.ctx, S.stream);
}

// Encapsulation of `batch` instances: h_pk in, h_c and h_k out.
int host_enc(HostState &S, int ps, int64_t batch, int mode, uint8_t *h_k, const uint8_t *h_pk, uint8_t *h_c)
{
    const int backend = backend_of(mode);
    if (backend < 0 || !h_k || !h_pk || !h_c) return TMVP_EARG;
    int rc = ensure_ctx(S);
    if (rc) return rc;
    tmvp_params P;
    tmvp_get_params(ps, &P);
    const size_t B = (size_t)batch;
    if ((rc = S.a.reserve(B * P.pk_bytes)) || (rc = S.b.reserve(B * 32)) || (rc = S.c.reserve(B * P.ct_bytes)) ||
        (rc = S.d.reserve(B * 32)))
        return rc;
    uint8_t *rnd = S.rnd;
    uint8_t *fresh = nullptr;
    if (rnd && S.rnd_bytes != B * 32) {
        // randomness set for another batch size: an error, not a silent
        // switch to fresh randomness (the caller expects its own coins)
        delete[] S.rnd;
        S.rnd = nullptr;
        S.rnd_bytes = 0;
        return TMVP_EARG;
    }
    if (!rnd) {
        fresh = new uint8_t[B * 32];
        if ((rc = fill_random(fresh, B * 32))) {
            delete[] fresh;
            return rc;
        }
        rnd = fresh;
    }
    rc = copy_in(S.a.p, h_pk, B * P.pk_bytes, S.stream);
    if (!rc) rc = copy_in(S.b.p, rnd, B * 32, S.stream);
    if (!rc) rc = tmvp_kem_encaps(S.ctx, ps, backend, S.c.u8(), S.d.u8(), S.a.u8(), S.b.u8(), batch, S.stream);
    if (!rc) rc = copy_out(h_c, S.c.p, B * P.ct_bytes, S.stream);
    if (!rc) rc = copy_out(h_k, S.d.p, B * 32, S.stream);
    const int rc2 = finish(S); // also waits for the copy of `rnd`
    delete[] fresh;
    delete[] S.rnd; // randomness is used once
    S.rnd = nullptr;
    S.rnd_bytes = 0;
    return rc ? rc : rc2;
}

// Decapsulation of `batch` instances: h_c and h_sk in, h_k out.  The public
// key is read from the sk, which holds a copy (S/include/params.h:45-46).
int host_dec(HostState &S, int ps, int64_t batch, int mode, const uint8_t *h_c, const uint8_t *h_sk, uint8_t *h_k)
{
    const int backend = backend_of(mode);
    if (backend < 0 || !h_c || !h_sk || !h_k) return TMVP_EARG;
    int rc = ensure_ctx(S);
    if (rc) return rc;
    tmvp_params P;
    tmvp_get_params(ps, &P);
    const size_t B = (size_t)batch;
    if ((rc = S.a.reserve(B * P.sk_bytes)) || (rc = S.c.reserve(B * P.ct_bytes)) || (rc = S.d.reserve(B * 32)))
        return rc;
    rc = copy_in(S.c.p, h_c, B * P.ct_bytes, S.stream);
    if (!rc) rc = copy_in(S.a.p, h_sk, B * P.sk_bytes, S.stream);
    if (!rc) rc = tmvp_kem_decaps(S.ctx, ps, backend, S.d.u8(), S.c.u8(), S.a.u8(), nullptr, batch, S.stream);
    if (!rc) rc = copy_out(h_k, S.d.p, B * 32, S.stream);
    const int rc2 = finish(S);
    return rc ? rc : rc2;
}

bool is_saber(int ps) { return ps == TMVP_LIGHTSABER || ps == TMVP_SABER || ps == TMVP_FIRESABER; }
bool is_sable(int ps) { return ps == TMVP_SABLE1 || ps == TMVP_SABLE2 || ps == TMVP_SABLE3; }

} // namespace

extern "C" {

int tmvp_host_configure(int device, int64_t saber_batch, int saber_set, int64_t sable_batch, int sable_set)
{
    if (device < 0 || saber_batch <= 0 || sable_batch <= 0 || !is_saber(saber_set) || !is_sable(sable_set))
        return TMVP_EARG;
    HostState &S = state();
    std::lock_guard<std::mutex> g(S.mu);
    if (S.ctx && device != S.device) return TMVP_EARG; // release first to move to another device
    S.device = device;
    S.saber_batch = saber_batch;
    S.saber_set = saber_set;
    S.sable_batch = sable_batch;
    S.sable_set = sable_set;
    return TMVP_OK;
}

int tmvp_host_set_randomness(const uint8_t *h_rnd, int64_t batch)
{
    HostState &S = state();
    std::lock_guard<std::mutex> g(S.mu);
    delete[] S.rnd;
    S.rnd = nullptr;
    S.rnd_bytes = 0;
    if (!h_rnd) return TMVP_OK;
    if (batch <= 0) return TMVP_EARG;
    S.rnd_bytes = (size_t)batch * 32;
    S.rnd = new uint8_t[S.rnd_bytes];
    memcpy(S.rnd, h_rnd, S.rnd_bytes);
    return TMVP_OK;
}

int tmvp_host_status(void)
{
    HostState &S = state();
    std::lock_guard<std::mutex> g(S.mu);
    return S.status;
}

int tmvp_host_release(void)
{
    HostState &S = state();
    std::lock_guard<std::mutex> g(S.mu);
    int rc = TMVP_OK;
    if (S.ctx) {
        if (S.stream && hipStreamSynchronize(S.stream) != hipSuccess) rc = TMVP_EHIP;
        S.a.release();
        S.b.release();
        S.c.release();
        S.d.release();
        S.r.release();
        if (S.stream) (void)hipStreamDestroy(S.stream);
        S.stream = nullptr;
        const int rc2 = tmvp_destroy(S.ctx);
        S.ctx = nullptr;
        if (!rc) rc = rc2;
    }
    delete[] S.rnd;
    S.rnd = nullptr;
    S.rnd_bytes = 0;
    return rc;
}

void saber_enc(uint8_t mode, uint8_t *h_k, uint8_t *h_pk, uint8_t *h_c)
{
    HostState &S = state();
    std::lock_guard<std::mutex> g(S.mu);
    S.status = host_enc(S, S.saber_set, S.saber_batch, mode, h_k, h_pk, h_c);
}

void saber_dec(uint8_t mode, uint8_t *h_c, uint8_t *h_pk, uint8_t *h_sk, uint8_t *h_k)
{
    (void)h_pk;
    HostState &S = state();
    std::lock_guard<std::mutex> g(S.mu);
    S.status = host_dec(S, S.saber_set, S.saber_batch, mode, h_c, h_sk, h_k);
}

void SABLE_enc(uint8_t *h_k, uint8_t *h_pk, uint8_t *h_c, int mode)
{
    HostState &S = state();
    std::lock_guard<std::mutex> g(S.mu);
    S.status = host_enc(S, S.sable_set, S.sable_batch, mode, h_k, h_pk, h_c);
}

void SABLE_dec(uint8_t *h_c, uint8_t *h_pk, uint8_t *h_sk, uint8_t *h_k, int mode)
{
    (void)h_pk;
    HostState &S = state();
    std::lock_guard<std::mutex> g(S.mu);
    S.status = host_dec(S, S.sable_set, S.sable_batch, mode, h_c, h_sk, h_k);
}

} // extern "C"
