// C-ABI host layer of libtmvp.so: argument checks, parameter-set table,
// backend dispatch.  Every entry point returns a status code; none aborts.
#include <string.h>

#include <mutex>
#include <new>

#include "../../include/tmvp.h"
#include "tmvp_internal.h"
#include "tmvp_kem.h"

using namespace tmvp;

namespace {
constexpr int kMaxWorkspaces = 16;
// Device scratch of a workspace.  A graph captured on the workspace's stream
// holds a reference: its kernels point into the buffer, so the workspace
// never frees or hands out a buffer that a live graph still holds, and moves
// to a new one instead when it has to grow or serve another stream.
struct Buf {
    void *ptr;
    size_t bytes;
    int refs;
    int device;
};
void buf_release(Buf *b)
{
    if (!b || --b->refs > 0) return;
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != b->device) (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize(); // a replay may still run on some stream
    (void)hipFree(b->ptr);
    if (cur >= 0 && cur != b->device) (void)hipSetDevice(cur);
    delete b;
}
struct Workspace {
    hipStream_t stream;
    Buf *buf;          // null until the first KEM call on the stream
    uint64_t used;     // last-use stamp
    hipStream_t aux;   // side stream for decapsulation steps that depend only on inputs
    hipEvent_t fork, join;
};
// A measured plan for the multi-job MFMA launch of one work-list shape.
constexpr int kMaxPlans = 32;
struct Plan {
    uint64_t sig; // 0 = empty
    int shape;    // MFMA tile: 16 or 32 (0 = the library's default)
    int wgs_per_cu;
    float weight[kNumKinds];
};
} // namespace

struct tmvp_ctx {
    int device;
    int cus;
    int *d_status; // device word set by MFMA launches on out-of-range secrets
    uint8_t *scratch; // this device's copy of `scratch` (tmvp_kem_encaps_noalloc)
    Workspace ws[kMaxWorkspaces];
    int nws;
    uint64_t ws_clock;
    Plan plans[kMaxPlans];
    int next_plan;
};

struct tmvp_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
    Buf *ws; // the workspace buffer the captured KEM kernels use, or null
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

// Scratch of tmvp_kem_encaps_noalloc: a fixed device array instead of the
// context's workspace, so that call allocates nothing.  Every GPU the library
// runs on holds its own copy from module load until the process ends.  Host
// code never passes the name `scratch` to a kernel (on the host it is only a
// stand-in for the symbol): it uses the device address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device by noalloc_scratch().
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch[kMaxDevices]; // device address of `scratch` per device, null until looked up

// This device's address of `scratch` (the device must be current).
uint8_t *noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
        g_scratch[device] = static_cast<uint8_t *>(ptr);
    }
    return g_scratch[device];
}

inline int h1_of(const PS &p) { return 1 << (p.eq - p.ep - 1); }
inline int h2_of(const PS &p) { return (1 << (p.ep - 2)) - (1 << (p.ep - p.et - 1)) + h1_of(p); }

int hip_status(hipError_t e) { return e == hipSuccess ? TMVP_OK : TMVP_EHIP; }

constexpr int kKnownFlags = TMVP_ACCUMULATE | TMVP_TILE_16 | TMVP_TILE_32;
// The MFMA tile a call's flags name: 16, 32, 0 for none, -1 for both.
int tile_of(int flags)
{
    const bool t16 = flags & TMVP_TILE_16, t32 = flags & TMVP_TILE_32;
    return t16 && t32 ? -1 : t16 ? 16 : t32 ? 32 : 0;
}

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
    P.sraw = p.l * p.mu * 32;
    return P;
}

bool capturing(hipStream_t stream)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return stream && hipStreamIsCapturing(stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Workspace of at least `bytes` for `stream`; growing it synchronises the
// device.  A stream seen for the first time when all slots are taken gets the
// least recently used slot, after a device synchronisation.
Workspace *workspace(tmvp_ctx *ctx, hipStream_t stream, size_t bytes, int *rc)
{
    Workspace *w = nullptr;
    *rc = TMVP_OK;
    for (int i = 0; i < ctx->nws; i++)
        if (ctx->ws[i].stream == stream) w = &ctx->ws[i];
    if (!w) {
        if (capturing(stream)) { // reserve before capturing
            *rc = TMVP_EUNSUPPORTED;
            return nullptr;
        }
        if (ctx->nws < kMaxWorkspaces) {
            w = &ctx->ws[ctx->nws];
            w->buf = nullptr;
            if (hipStreamCreateWithFlags(&w->aux, hipStreamNonBlocking) != hipSuccess) {
                *rc = TMVP_EHIP;
                return nullptr;
            }
            if (hipEventCreateWithFlags(&w->fork, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&w->join, hipEventDisableTiming) != hipSuccess) {
                (void)hipStreamDestroy(w->aux);
                *rc = TMVP_EHIP;
                return nullptr;
            }
            ctx->nws++;
        } else {
            w = &ctx->ws[0];
            for (int i = 1; i < kMaxWorkspaces; i++)
                if (ctx->ws[i].used < w->used) w = &ctx->ws[i];
            if (hipDeviceSynchronize() != hipSuccess) {
                *rc = TMVP_EHIP;
                return nullptr;
            }
            // the evicted stream's buffer is reused unless a graph holds it
            if (w->buf && w->buf->refs > 1) {
                buf_release(w->buf);
                w->buf = nullptr;
            }
        }
        w->stream = stream;
    }
    w->used = ++ctx->ws_clock;
    if (!w->buf || w->buf->bytes < bytes) {
        if (capturing(stream)) { // no allocation inside a graph capture
            *rc = TMVP_EUNSUPPORTED;
            return nullptr;
        }
        if (hipDeviceSynchronize() != hipSuccess) {
            *rc = TMVP_EHIP;
            return nullptr;
        }
        buf_release(w->buf); // freed here unless a graph still holds it
        w->buf = nullptr;
        Buf *b = new (std::nothrow) Buf{nullptr, bytes, 1, ctx->device};
        if (!b) {
            *rc = TMVP_ENOMEM;
            return nullptr;
        }
        if (hipMalloc(&b->ptr, bytes) != hipSuccess) {
            delete b;
            *rc = TMVP_ENOMEM;
            return nullptr;
        }
        w->buf = b;
    }
    return w;
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
    Workspace *w = workspace(ctx, reinterpret_cast<hipStream_t>(stream), kem_workspace_bytes(P, batch), &rc);
    if (!w) return rc;
    kem_workspace_carve(P, batch, static_cast<uint8_t *>(w->buf->ptr), W);
    W.aux = w->aux;
    W.fork = w->fork;
    W.join = w->join;
    return TMVP_OK;
}

} // namespace

namespace tmvp {
KemParams kem_params_of(int ps) { return kem_params(ps); }
int kem_backend_ok(int backend)
{
    return backend == TMVP_BACKEND_MFMA || backend == TMVP_BACKEND_VALU || backend == TMVP_BACKEND_SCHOOLBOOK;
}
} // namespace tmvp

namespace {

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

// Checks one work descriptor and builds its job, exactly as the matching
// single-job entry point does.  Pointers are only required for batch > 0.
int job_from_desc(tmvp_ctx *ctx, const tmvp_work_desc &d, Job &J, int &mode, int &l)
{
    if (d.batch < 0 || (d.flags & ~kKnownFlags) || tile_of(d.flags) < 0) return TMVP_EARG;
    const int acc = (d.flags & TMVP_ACCUMULATE) ? 1 : 0;
    J = base_job(ctx, d.out, d.a, d.s, d.batch);
    J.accumulate = acc;
    J.shape = tile_of(d.flags);
    bool need_aux = false;
    if (d.op == TMVP_OP_POLYMUL) {
        if (d.eq_bits < 1 || d.eq_bits > 16 || d.epilogue != 0) return TMVP_EARG;
        J.w = d.eq_bits;
        mode = kPoly;
        l = 1;
    } else {
        if (d.param_set < 0 || d.param_set >= TMVP_NUM_PARAM_SETS) return TMVP_EARG;
        const PS &p = kParams[d.param_set];
        J.eq = p.eq;
        J.ep = p.ep;
        J.et = p.et;
        l = p.l;
        if (d.op == TMVP_OP_MATVEC || d.op == TMVP_OP_MATVEC_T) {
            if (d.epilogue != 0 && d.epilogue != 1) return TMVP_EARG;
            if (d.epilogue && acc) return TMVP_EARG;
            J.w = p.eq;
            if (d.epilogue) {
                J.epi = kEpiQ2P;
                J.h = h1_of(p);
            }
            mode = d.op == TMVP_OP_MATVEC ? kMatVec : kMatVecT;
        } else if (d.op == TMVP_OP_INNERPROD) {
            if (d.epilogue < TMVP_IP_RAW || d.epilogue > TMVP_IP_DEC) return TMVP_EARG;
            if (d.epilogue != TMVP_IP_RAW && acc) return TMVP_EARG;
            J.w = p.ep;
            if (d.epilogue == TMVP_IP_ENC) {
                J.epi = kEpiEnc;
                J.h = h1_of(p);
                J.aux = d.aux;
                need_aux = true;
            } else if (d.epilogue == TMVP_IP_DEC) {
                J.epi = kEpiDec;
                J.h = h2_of(p);
                J.aux = d.aux;
                need_aux = true;
            }
            mode = kInner;
        } else {
            return TMVP_EARG;
        }
    }
    if (d.batch > 0 && (!d.out || !d.a || !d.s || (need_aux && !d.aux))) return TMVP_EARG;
    return TMVP_OK;
}

uint64_t fnv(uint64_t h, uint64_t v)
{
    for (int i = 0; i < 8; i++) {
        h ^= (v >> (8 * i)) & 0xff;
        h *= 0x100000001b3ull;
    }
    return h;
}

// A checked MFMA work list: the non-empty jobs, their kinds, and the shape
// signature plans are cached under.
struct WorkList {
    Job J[TMVP_MAX_WORK];
    int kind[TMVP_MAX_WORK];
    int n;
    uint64_t sig;
    int tile; // the tile the descriptors name (0: none)
};

int check_list(tmvp_ctx *ctx, int backend, const tmvp_work_desc *work, int n, WorkList &W, int *modes, int *ls)
{
    if (!ctx || n < 0 || n > TMVP_MAX_WORK || (n > 0 && !work)) return TMVP_EARG;
    W.n = 0;
    W.sig = 0xcbf29ce484222325ull;
    W.tile = 0;
    for (int i = 0; i < n; i++) {
        Job J;
        int mode = 0, l = 0;
        const int rc = job_from_desc(ctx, work[i], J, mode, l);
        if (rc) return rc;
        if (J.shape) {
            if (W.tile && W.tile != J.shape) return TMVP_EARG; // one tile per launch
            W.tile = J.shape;
        }
        if (backend == TMVP_BACKEND_MFMA && (J.w < 8 || J.w > 14)) return TMVP_EUNSUPPORTED;
        if (J.batch == 0) continue;
        const int k = W.n++;
        W.J[k] = J;
        W.kind[k] = mfma_kind(mode, l, J.epi);
        if (W.kind[k] < 0) return TMVP_EUNSUPPORTED;
        if (modes) modes[k] = mode;
        if (ls) ls[k] = l;
        W.sig = fnv(fnv(W.sig, (uint64_t)W.kind[k]), (uint64_t)J.batch);
    }
    if (W.sig == 0) W.sig = 1;
    return TMVP_OK;
}

const Plan *find_plan(const tmvp_ctx *ctx, uint64_t sig)
{
    for (int i = 0; i < kMaxPlans; i++)
        if (ctx->plans[i].sig == sig) return &ctx->plans[i];
    return nullptr;
}

Plan default_plan(uint64_t sig)
{
    Plan P;
    P.sig = sig;
    P.shape = 0;      // the library's default tile
    P.wgs_per_cu = 0; // the kernel's full occupancy
    for (int k = 0; k < kNumKinds; k++) P.weight[k] = 1.0f;
    return P;
}

// Splits the resident workgroups between the jobs in proportion to
// batch x per-instance cost x weight; each job gets at least one workgroup and
// no more than it has groups of four instances.
int plan_shape(const Plan &P) { return P.shape ? P.shape : mfma_default_shape(); }

int build_multi(const tmvp_ctx *ctx, const WorkList &W, const Plan &P, MultiJob &M)
{
    int per_cu = mfma_multi_per_cu(plan_shape(P));
    if (P.wgs_per_cu >= 1 && P.wgs_per_cu < per_cu) per_cu = P.wgs_per_cu;
    const double resident = (double)ctx->cus * per_cu;
    double total = 0.0;
    double work[TMVP_MAX_WORK];
    for (int i = 0; i < W.n; i++) {
        work[i] = (double)W.J[i].batch * mfma_kind_cost(W.kind[i]) * P.weight[W.kind[i]];
        total += work[i];
    }
    int begin = 0;
    M.nseg = W.n;
    for (int i = 0; i < W.n; i++) {
        const int64_t cap = (W.J[i].batch + 3) / 4;
        int64_t b = (int64_t)(resident * work[i] / total + 0.5);
        if (b < 1) b = 1;
        if (b > cap) b = cap;
        M.seg[i].J = W.J[i];
        M.seg[i].kind = W.kind[i];
        M.seg[i].block_begin = begin;
        M.seg[i].blocks = (int)b;
        begin += (int)b;
    }
    return begin;
}

int launch_list(tmvp_ctx *ctx, const WorkList &W, const Plan &P, hipStream_t st)
{
    if (W.n == 0) return TMVP_OK;
    MultiJob M;
    Plan use = P;
    if (W.tile && W.tile != plan_shape(P)) { // the descriptors name another tile than the plan
        use.shape = W.tile;
        use.wgs_per_cu = 0;
    }
    const int grid = build_multi(ctx, W, use, M);
    return hip_status(launch_mfma_multi(M, grid, plan_shape(use), st));
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
    c->scratch = noalloc_scratch(device); // set-up of tmvp_kem_encaps_noalloc, once per device
    if (!c->scratch) {
        delete c;
        return TMVP_EHIP;
    }
    c->nws = 0;
    c->ws_clock = 0;
    c->next_plan = 0;
    for (int i = 0; i < kMaxPlans; i++) c->plans[i].sig = 0;
    if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c->cus < 1) {
        delete c;
        return TMVP_EHIP;
    }
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
    for (int i = 0; i < ctx->nws; i++) {
        buf_release(ctx->ws[i].buf); // graphs still alive keep theirs
        (void)hipEventDestroy(ctx->ws[i].fork);
        (void)hipEventDestroy(ctx->ws[i].join);
        (void)hipStreamDestroy(ctx->ws[i].aux);
    }
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
    if (!ctx || batch < 0 || eq_bits < 1 || eq_bits > 16 || (flags & ~kKnownFlags) || tile_of(flags) < 0)
        return TMVP_EARG;
    if (batch == 0) return TMVP_OK;
    if (!d_c || !d_a || !d_s) return TMVP_EARG;
    Job J = base_job(ctx, d_c, d_a, d_s, batch);
    J.w = eq_bits;
    J.accumulate = (flags & TMVP_ACCUMULATE) ? 1 : 0;
    J.shape = tile_of(flags);
    return run(ctx, backend, J, kPoly, 1, stream);
}

int tmvp_matvec(tmvp_ctx *ctx, int ps, int backend, int transpose, uint16_t *d_out, const uint16_t *d_A,
                const uint16_t *d_s, int64_t batch, int round_to_p, int flags, tmvp_stream_t stream)
{
    if (!ctx || ps < 0 || ps >= TMVP_NUM_PARAM_SETS || batch < 0 || (flags & ~kKnownFlags) || tile_of(flags) < 0)
        return TMVP_EARG;
    if (round_to_p && (flags & TMVP_ACCUMULATE)) return TMVP_EARG;
    if (batch == 0) return TMVP_OK;
    if (!d_out || !d_A || !d_s) return TMVP_EARG;
    const PS &p = kParams[ps];
    Job J = base_job(ctx, d_out, d_A, d_s, batch);
    J.shape = tile_of(flags);
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
    if (!ctx || ps < 0 || ps >= TMVP_NUM_PARAM_SETS || batch < 0 || (flags & ~kKnownFlags) || tile_of(flags) < 0)
        return TMVP_EARG;
    if (epilogue < TMVP_IP_RAW || epilogue > TMVP_IP_DEC) return TMVP_EARG;
    if (epilogue != TMVP_IP_RAW && (flags & TMVP_ACCUMULATE)) return TMVP_EARG;
    if (batch == 0) return TMVP_OK;
    if (!d_out || !d_b || !d_s || (epilogue != TMVP_IP_RAW && !d_aux)) return TMVP_EARG;
    const PS &p = kParams[ps];
    Job J = base_job(ctx, d_out, d_b, d_s, batch);
    J.shape = tile_of(flags);
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

int tmvp_kem_encaps_noalloc(tmvp_ctx *ctx, int ps, int backend, uint8_t *d_ct, uint8_t *d_ss, const uint8_t *d_pk,
                            const uint8_t *d_rnd, int64_t batch, tmvp_stream_t stream)
{
    if (!kem_args_ok(ctx, ps, backend, batch)) return TMVP_EARG;
    if (batch == 0) return TMVP_OK;
    if (!d_ct || !d_ss || !d_pk || !d_rnd) return TMVP_EARG;
    const KemParams P = kem_params(ps);
    if (kem_workspace_bytes(P, batch) > kNoAllocScratchBytes) return TMVP_ENOMEM; // the array does not hold it
    int rc = set_device(ctx);
    if (rc) return rc;
    if (!ctx->scratch) return TMVP_EHIP;
    // The whole chain runs on `stream` (no side stream, no events), and every
    // scratch word it reads it has written earlier in the same call, so
    // nothing needs clearing between calls.
    KemWorkspace W;
    kem_workspace_carve(P, batch, ctx->scratch, W);
    W.aux = nullptr;
    W.fork = W.join = nullptr;
    return hip_status(kem_encaps(P, W, backend == TMVP_BACKEND_MFMA ? 0 : 1, d_pk, d_rnd, d_ct, d_ss, batch,
                                 ctx->d_status, reinterpret_cast<hipStream_t>(stream)));
}

int tmvp_kem_reserve(tmvp_ctx *ctx, int ps, int64_t batch, tmvp_stream_t stream)
{
    if (!kem_args_ok(ctx, ps, TMVP_BACKEND_MFMA, batch)) return TMVP_EARG;
    if (batch == 0) return TMVP_OK;
    KemParams P;
    KemWorkspace W;
    return kem_setup(ctx, ps, batch, stream, P, W);
}

int tmvp_batch_launch(tmvp_ctx *ctx, int backend, const tmvp_work_desc *work, int n, tmvp_stream_t stream)
{
    if (backend != TMVP_BACKEND_MFMA && backend != TMVP_BACKEND_VALU && backend != TMVP_BACKEND_SCHOOLBOOK)
        return TMVP_EARG;
    WorkList W;
    int modes[TMVP_MAX_WORK], ls[TMVP_MAX_WORK];
    int rc = check_list(ctx, backend, work, n, W, modes, ls);
    if (rc) return rc;
    if ((rc = set_device(ctx))) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (backend != TMVP_BACKEND_MFMA) {
        for (int i = 0; i < W.n; i++)
            if ((rc = hip_status(launch_valu(W.J[i], modes[i], ls[i], st)))) return rc;
        return TMVP_OK;
    }
    const Plan *P = find_plan(ctx, W.sig);
    return launch_list(ctx, W, P ? *P : default_plan(W.sig), st);
}

int tmvp_batch_autotune(tmvp_ctx *ctx, const tmvp_work_desc *work, int n, tmvp_stream_t stream, float *best_ms)
{
    WorkList W;
    int rc = check_list(ctx, TMVP_BACKEND_MFMA, work, n, W, nullptr, nullptr);
    if (rc) return rc;
    for (int i = 0; i < W.n; i++)
        if (W.J[i].accumulate) return TMVP_EARG;
    if ((rc = set_device(ctx))) return rc;
    if (W.n == 0) {
        if (best_ms) *best_ms = 0.0f;
        return TMVP_OK;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return TMVP_EHIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return TMVP_EHIP;
    }
    // Time per launch of a plan: the best of three runs of five launches.
    auto measure = [&](const Plan &P, float &ms) -> int {
        int r;
        for (int i = 0; i < 2; i++)
            if ((r = launch_list(ctx, W, P, st))) return r;
        ms = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            if (hipEventRecord(e0, st) != hipSuccess) return TMVP_EHIP;
            for (int i = 0; i < 5; i++)
                if ((r = launch_list(ctx, W, P, st))) return r;
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return TMVP_EHIP;
            float t = 0.0f;
            if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) return TMVP_EHIP;
            if (t / 5 < ms) ms = t / 5;
        }
        return TMVP_OK;
    };
    Plan best = default_plan(W.sig);
    float best_t = 0.0f;
    rc = measure(best, best_t);
    // tile shape x workgroups per CU
    const Plan base = best;
    for (int shape : {16, 32}) {
        for (int w = 1; !rc && w <= mfma_multi_per_cu(shape); w++) {
            Plan c = base;
            c.shape = shape;
            c.wgs_per_cu = w;
            float t = 0.0f;
            rc = measure(c, t);
            if (!rc && t < best_t) {
                best = c;
                best_t = t;
            }
        }
    }
    if (!best.shape) best.shape = mfma_default_shape(); // the plan records the tile it was measured with
    // per-kind weights, coordinate descent
    bool present[kNumKinds] = {};
    int kinds = 0;
    for (int i = 0; i < W.n; i++) {
        kinds += present[W.kind[i]] ? 0 : 1;
        present[W.kind[i]] = true;
    }
    const float mult[4] = {0.7f, 0.85f, 1.2f, 1.4f};
    for (int pass = 0; pass < 2 && !rc && kinds > 1; pass++)
        for (int k = 0; k < kNumKinds && !rc; k++) {
            if (!present[k]) continue;
            for (int m = 0; m < 4 && !rc; m++) {
                Plan c = best;
                c.weight[k] *= mult[m];
                float t = 0.0f;
                rc = measure(c, t);
                if (!rc && t < best_t) {
                    best = c;
                    best_t = t;
                }
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    Plan *slot = nullptr;
    for (int i = 0; i < kMaxPlans; i++)
        if (ctx->plans[i].sig == W.sig) slot = &ctx->plans[i];
    if (!slot) {
        slot = &ctx->plans[ctx->next_plan];
        ctx->next_plan = (ctx->next_plan + 1) % kMaxPlans;
    }
    *slot = best;
    if (best_ms) *best_ms = best_t;
    return TMVP_OK;
}

int tmvp_batch_plan(tmvp_ctx *ctx, const tmvp_work_desc *work, int n, int *shape, int *wgs_per_cu)
{
    WorkList W;
    int rc = check_list(ctx, TMVP_BACKEND_MFMA, work, n, W, nullptr, nullptr);
    if (rc) return rc;
    const Plan *P = find_plan(ctx, W.sig);
    const Plan D = default_plan(W.sig);
    const Plan &use = P ? *P : D;
    int per_cu = mfma_multi_per_cu(plan_shape(use));
    if (use.wgs_per_cu >= 1 && use.wgs_per_cu < per_cu) per_cu = use.wgs_per_cu;
    if (shape) *shape = W.tile ? W.tile : plan_shape(use);
    if (wgs_per_cu) *wgs_per_cu = per_cu;
    return P ? TMVP_OK : 1;
}

int tmvp_graph_begin(tmvp_ctx *ctx, tmvp_stream_t stream)
{
    if (!ctx || !stream) return TMVP_EARG;
    int rc = set_device(ctx);
    if (rc) return rc;
    // Resolve the launch-shape queries the kernels make on first use outside the capture.
    (void)mfma_multi_per_cu(16);
    (void)mfma_multi_per_cu(32);
    return hip_status(hipStreamBeginCapture(reinterpret_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal));
}

int tmvp_graph_end(tmvp_ctx *ctx, tmvp_stream_t stream, tmvp_graph **out)
{
    if (!ctx || !stream || !out) return TMVP_EARG;
    *out = nullptr;
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(reinterpret_cast<hipStream_t>(stream), &g) != hipSuccess || !g) return TMVP_EHIP;
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return TMVP_EHIP;
    }
    tmvp_graph *h = new (std::nothrow) tmvp_graph;
    if (!h) {
        (void)hipGraphExecDestroy(x);
        (void)hipGraphDestroy(g);
        return TMVP_ENOMEM;
    }
    h->graph = g;
    h->exec = x;
    h->ws = nullptr;
    // The captured KEM kernels point into the workspace of this stream: the
    // graph keeps that buffer alive until it is destroyed.
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    for (int i = 0; i < ctx->nws; i++)
        if (ctx->ws[i].stream == st && ctx->ws[i].buf) {
            h->ws = ctx->ws[i].buf;
            h->ws->refs++;
        }
    *out = h;
    return TMVP_OK;
}

int tmvp_graph_launch(tmvp_graph *graph, tmvp_stream_t stream)
{
    if (!graph) return TMVP_EARG;
    return hip_status(hipGraphLaunch(graph->exec, reinterpret_cast<hipStream_t>(stream)));
}

int tmvp_graph_destroy(tmvp_graph *graph)
{
    if (!graph) return TMVP_EARG;
    int rc = TMVP_OK;
    if (hipGraphExecDestroy(graph->exec) != hipSuccess) rc = TMVP_EHIP;
    if (hipGraphDestroy(graph->graph) != hipSuccess) rc = TMVP_EHIP;
    buf_release(graph->ws);
    delete graph;
    return rc;
}

} // extern "C"
