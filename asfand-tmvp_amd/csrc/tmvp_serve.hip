// Serving loop: continuous batching of KEM encapsulation / decapsulation
// requests over device-resident keys.  The reference has no serving layer
// (its host functions run one fixed batch per call, S/src/cuda_kernel.cu:25,
// SB/src/SABLE_indcpa.cu:24); this is the request path around the same
// kernels.
//
// Design:
//  - Key store per parameter set, in HBM: pk, sk, and the matrix A and H(pk)
//    derived once per key, so a request never runs the GenMatrix sponge.
//  - HBM budget: the key stores and the batches in flight share it.  Each
//    batch reserves exactly the device bytes its (parameter set, operation,
//    instances) needs - workspace, key ids, inputs, outputs - from one arena
//    allocated at creation, and gives them back when it retires.  A batch is
//    admitted when a stream is free and its bytes fit the budget left: small
//    batches run side by side where one large batch would take it all, and
//    a request that does not fit yet waits in its queue.
//  - Slots: each slot is a stream with its own pinned host staging; their
//    number bounds the batches in flight.
//  - A scheduler thread forms each batch from the queue (parameter set,
//    operation) whose oldest request waited longest, taking requests in
//    arrival order up to max_batch instances and up to the budget left, and
//    retires finished slots by polling their events.  Requests arriving
//    meanwhile join the next batch that a free slot takes.
//  - Data crosses the bus through a copy kernel that reads or writes
//    page-locked host memory directly, once per batch each way: the key ids
//    and pageable inputs from the slot's pinned staging, pinned inputs from
//    the caller's own buffer; results into the caller's pinned buffers, else
//    into the staging.  Submitting DMA copies instead occasionally blocked the
//    submitting thread for 7-15 ms inside the runtime, which stalled every
//    slot.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/tmvp.h"
#include "tmvp_internal.h"
#include "tmvp_kem.h"

using namespace tmvp;

namespace tmvp {
KemParams kem_params_of(int ps); // tmvp_api.hip
int kem_backend_ok(int backend);
} // namespace tmvp

namespace {

using Clock = std::chrono::steady_clock;

struct Store {
    uint8_t *pk = nullptr, *sk = nullptr, *hpk = nullptr;
    uint16_t *A = nullptr;
    bool has_sk = true;
    int64_t n = 0, cap = 0;
    int64_t bytes = 0; // device bytes held (cap keys)
};

struct Pending {
    uint64_t id;
    tmvp_srv_request req;       // req.keys is not read after submit:
    std::vector<int32_t> keys;  // the ids are copied here as they are checked
    double t_submit;
    bool pinned_in, pinned_out; // host buffers the device can reach directly
    const uint8_t *dev_in;      // pinned_in: the device's address of req.in
    uint8_t *dev_ct, *dev_ss, *dev_fail; // pinned_out: the device's addresses of the outputs
};

// Whether p lies in page-locked host memory from hipHostMalloc; *dev gets
// the device's address of p when it does.
bool is_pinned(const void *p, const void **dev = nullptr)
{
    if (!p) return true;
    hipPointerAttribute_t a;
    const hipError_t e = hipPointerGetAttributes(&a, p);
    (void)hipGetLastError(); // pageable memory reports an error: clear it
    const bool ok = e == hipSuccess && a.type == hipMemoryTypeHost && a.devicePointer;
    if (ok && dev) *dev = a.devicePointer;
    return ok;
}

// One piece of a batch's data: `bytes` from src to dst, one of them in
// page-locked host memory (read or written over the bus).
struct CopyDesc {
    const uint8_t *src;
    uint8_t *dst;
    uint32_t bytes;
};
// Bytes per descriptor, one workgroup each: small pieces put many workgroups
// on a request's bytes (64 KiB pieces raised p50 by about 30 us).
constexpr uint32_t kPiece = 8u << 10;

// The workgroups take the pieces b = blockIdx.x, + gridDim.x, ...  16-byte
// accesses when both ends and the size allow, else 4-byte, else bytes; each
// thread keeps four loads in flight.  The grid is capped at kCopyGrid
// workgroups: a copy over the bus holds its CU slots while it waits for the
// host, and at saturation a workgroup per piece (about 2000 per batch of
// 15000 instances) kept the slots from the KEM kernels.  Serving mix, four
// slots: saturation 22.5k against 18.5k requests/s with the cap at 64 (16:
// 18.3k, 32: 21.7k, 256: 20.3k); p50 at 2000 requests/s 0.277 against 0.267 ms.
constexpr int kCopyGrid = 64;

__global__ __launch_bounds__(256) void copy_kernel(const CopyDesc *descs, int n)
{
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        const CopyDesc d = descs[b];
        const uintptr_t al = (uintptr_t)d.src | (uintptr_t)d.dst | d.bytes;
        if ((al & 15) == 0) {
            const uint4 *src = reinterpret_cast<const uint4 *>(d.src);
            uint4 *dst = reinterpret_cast<uint4 *>(d.dst);
            const uint32_t m = d.bytes / 16;
            for (uint32_t i = threadIdx.x; i < m; i += 4 * 256) {
                uint4 v[4];
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (i + k * 256 < m) v[k] = src[i + k * 256];
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (i + k * 256 < m) dst[i + k * 256] = v[k];
            }
        } else if ((al & 3) == 0) {
            const uint32_t *src = reinterpret_cast<const uint32_t *>(d.src);
            uint32_t *dst = reinterpret_cast<uint32_t *>(d.dst);
            for (uint32_t i = threadIdx.x; i < d.bytes / 4; i += 256) dst[i] = src[i];
        } else {
            for (uint32_t i = threadIdx.x; i < d.bytes; i += 256) d.dst[i] = d.src[i];
        }
    }
}

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    // device memory of the batch in flight, a range of the server's arena:
    // workspace | keys | inputs | outputs (ct | ss | fail)
    size_t off = 0, bytes = 0;
    uint8_t *ws = nullptr, *d_in = nullptr, *d_out = nullptr;
    int *d_status = nullptr;
    uint8_t *h_in = nullptr, *h_out = nullptr; // pinned staging
    const uint8_t *h_in_dev = nullptr;         // the device's addresses of h_in,
    uint8_t *h_out_dev = nullptr;              // h_out
    int *h_status_dev = nullptr;               // and h_status
    CopyDesc *h_desc = nullptr;                // pinned: the copy pieces of a batch
    const CopyDesc *h_desc_dev = nullptr;
    int *h_status = nullptr;
    bool busy = false;
    int op = 0, ps = 0;
    int64_t instances = 0;
    double t_start = 0;
    std::vector<Pending> reqs;
};

} // namespace

struct tmvp_server {
    tmvp_ctx *ctx;
    int device, backend;
    int64_t max_batch;
    size_t in_bytes, out_bytes; // pinned staging per slot (max_batch of the largest set)
    int64_t desc_cap; // copy-in pieces one batch can need
    // HBM: the budget covers the key stores and the batches in flight
    int64_t budget = 0, store_bytes = 0, inflight_bytes = 0;
    int64_t max_batch_bytes = 0; // one batch of max_batch instances, largest set and operation
    uint8_t *arena = nullptr;
    size_t arena_bytes = 0;
    Clock::time_point t0;
    Store store[TMVP_NUM_PARAM_SETS];
    std::vector<Slot> slots;
    std::deque<Pending> queue[TMVP_NUM_PARAM_SETS][2];
    std::deque<tmvp_srv_completion> done;
    std::mutex mu;
    std::condition_variable wake, completed;
    std::thread worker;
    bool stop = false;
    uint64_t next_id = 1;
    int64_t pending_requests = 0, inflight = 0;
    tmvp_srv_stats stats{};
    int fatal = TMVP_OK;

    double now_us() const { return std::chrono::duration<double, std::micro>(Clock::now() - t0).count(); }
};

namespace {

constexpr size_t kAlign = 256;
size_t up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// Keyed MFMA calls read A from the key store; the VALU backend gathers a copy.
bool ws_with_A(int backend) { return backend != TMVP_BACKEND_MFMA; }

// Device bytes of one batch: workspace | keys | inputs | outputs.
struct BatchLayout {
    size_t ws, keys, in, out, total;
};
BatchLayout batch_layout(const KemParams &P, int op, int64_t B, int backend)
{
    BatchLayout b;
    const size_t n = (size_t)B;
    b.ws = up(kem_workspace_bytes(P, B, ws_with_A(backend)));
    b.keys = up(n * 4);
    b.in = up(n * (op == TMVP_SRV_ENCAPS ? 32 : (size_t)P.ct));
    b.out = up(op == TMVP_SRV_ENCAPS ? n * (P.ct + 32) : n * 33);
    b.total = b.ws + b.keys + b.in + b.out;
    return b;
}

// Pinned staging sizes for max_batch instances of the largest parameter set,
// and the device bytes of the largest max_batch batch.
void max_sizes(int64_t B, int backend, size_t &in, size_t &out, size_t &batch)
{
    in = out = batch = 0;
    for (int ps = 0; ps < TMVP_NUM_PARAM_SETS; ps++) {
        const KemParams P = kem_params_of(ps);
        const size_t i = up((size_t)B * 4) + up((size_t)B * (size_t)(P.ct > 32 ? P.ct : 32));
        const size_t o = (size_t)B * (P.ct + 32 + 1);
        in = i > in ? i : in;
        out = o > out ? o : out;
        for (int op = 0; op < 2; op++) {
            const size_t t = batch_layout(P, op, B, backend).total;
            batch = t > batch ? t : batch;
        }
    }
}

// Arena bytes with key stores of `store` bytes: the budget left, but never
// more than `slots` batches of max_batch instances can use at once.
int64_t arena_size(const tmvp_server *S, int64_t store)
{
    const int64_t most = S->max_batch_bytes * (int64_t)S->slots.size();
    const int64_t left = S->budget - store;
    return left < most ? left : most;
}

// Device bytes of a key store of `cap` keys.
int64_t store_bytes_of(const KemParams &P, int64_t cap, bool with_sk)
{
    return cap * ((int64_t)P.pk + 32 + (int64_t)P.l * P.l * kN * 2 + (with_sk ? (int64_t)P.sk : 0));
}

void free_slot(Slot &s)
{
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.d_status) (void)hipFree(s.d_status);
    if (s.h_in) (void)hipHostFree(s.h_in);
    if (s.h_out) (void)hipHostFree(s.h_out);
    if (s.h_status) (void)hipHostFree(s.h_status);
    if (s.h_desc) (void)hipHostFree(s.h_desc);
    s = Slot();
}

int alloc_slot(tmvp_server *S, Slot &s)
{
    if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess) return TMVP_EHIP;
    if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) return TMVP_EHIP;
    if (hipMalloc(&s.d_status, sizeof(int)) != hipSuccess) return TMVP_ENOMEM;
    if (hipHostMalloc(&s.h_in, S->in_bytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&s.h_out, S->out_bytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&s.h_status, sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&s.h_desc, (size_t)S->desc_cap * sizeof(CopyDesc), hipHostMallocDefault) != hipSuccess)
        return TMVP_ENOMEM;
    const void *dv = nullptr;
    if (!is_pinned(s.h_in, &dv)) return TMVP_EHIP;
    s.h_in_dev = static_cast<const uint8_t *>(dv);
    if (!is_pinned(s.h_out, &dv)) return TMVP_EHIP;
    s.h_out_dev = static_cast<uint8_t *>(const_cast<void *>(dv));
    if (!is_pinned(s.h_status, &dv)) return TMVP_EHIP;
    s.h_status_dev = static_cast<int *>(const_cast<void *>(dv));
    if (!is_pinned(s.h_desc, &dv)) return TMVP_EHIP;
    s.h_desc_dev = static_cast<const CopyDesc *>(dv);
    if (hipMemset(s.d_status, 0, sizeof(int)) != hipSuccess) return TMVP_EHIP;
    return TMVP_OK;
}

// Launch the requests in s.reqs on slot s (lock held).
int launch_slot(tmvp_server *S, Slot &s)
{
    const KemParams P = kem_params_of(s.ps);
    const int64_t B = s.instances;
    // Key ids, and the inputs of requests in pageable memory, go through the
    // pinned staging; one copy-in kernel then reads them, and the pinned
    // requests' inputs in place, into d_in (keys | inputs).
    int32_t *hkeys = reinterpret_cast<int32_t *>(s.h_in);
    const size_t in_off = up((size_t)B * 4);
    uint8_t *din_base = s.d_in + in_off;
    const size_t in_each = s.op == TMVP_SRV_ENCAPS ? 32 : (size_t)P.ct;
    int nd = 0;
    auto add = [&](const uint8_t *src, uint8_t *dst, size_t bytes) {
        for (size_t o = 0; o < bytes; o += kPiece) {
            if (nd >= S->desc_cap) return false;
            const size_t b = bytes - o < kPiece ? bytes - o : kPiece;
            s.h_desc[nd++] = CopyDesc{src + o, dst + o, (uint32_t)b};
        }
        return true;
    };
    int64_t off = 0;
    for (const Pending &p : s.reqs) {
        memcpy(hkeys + off, p.keys.data(), (size_t)p.req.count * 4);
        off += p.req.count;
    }
    bool fits = add(s.h_in_dev, s.d_in, (size_t)B * 4);
    off = 0;
    for (const Pending &p : s.reqs) {
        const size_t bytes = (size_t)p.req.count * in_each, at = (size_t)off * in_each;
        if (p.pinned_in) {
            fits = fits && add(p.dev_in, din_base + at, bytes);
        } else {
            memcpy(s.h_in + in_off + at, p.req.in, bytes);
            fits = fits && add(s.h_in_dev + in_off + at, din_base + at, bytes);
        }
        off += p.req.count;
    }
    if (!fits) return TMVP_EARG; // cannot happen: desc_cap covers max_batch requests
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)(nd < kCopyGrid ? nd : kCopyGrid)), dim3(256), 0, s.stream,
                       s.h_desc_dev, nd);
    if (hipGetLastError() != hipSuccess) return TMVP_EHIP;
    KemWorkspace W;
    kem_workspace_carve(P, B, s.ws, W, ws_with_A(S->backend));
    // keyed calls run on the slot's one stream (A is resident, so the
    // decapsulation has no side-stream work)
    W.aux = nullptr;
    W.fork = W.join = nullptr;
    const Store &st = S->store[s.ps];
    KeyStore K{st.pk, st.sk, st.A, st.hpk};
    const int32_t *dkeys = reinterpret_cast<const int32_t *>(s.d_in);
    const uint8_t *din = s.d_in + up((size_t)B * 4);
    // outputs, packed: encapsulation ct | ss, decapsulation ss | fail
    const size_t out_total = s.op == TMVP_SRV_ENCAPS ? (size_t)B * (P.ct + 32) : (size_t)B * 33;
    const int backend = S->backend == TMVP_BACKEND_MFMA ? 0 : 1;
    hipError_t e;
    if (s.op == TMVP_SRV_ENCAPS)
        e = kem_encaps_keyed(P, W, backend, K, dkeys, din, s.d_out, s.d_out + (size_t)B * P.ct, B, s.d_status,
                             s.stream);
    else
        e = kem_decaps_keyed(P, W, backend, K, dkeys, din, s.d_out, s.d_out + (size_t)B * 32, B, s.d_status,
                             s.stream);
    if (e != hipSuccess) return TMVP_EHIP;
    // results (and the range flag): one copy kernel writes them straight into
    // pinned request buffers, else into the staging
    const int nd_in = nd;
    bool any_staged = false;
    const bool enc = s.op == TMVP_SRV_ENCAPS;
    const uint8_t *dss = enc ? s.d_out + (size_t)B * P.ct : s.d_out, *dfail = s.d_out + (size_t)B * 32;
    off = 0;
    for (const Pending &p : s.reqs) {
        const size_t n = (size_t)p.req.count;
        if (!p.pinned_out) {
            any_staged = true;
        } else {
            if (enc) fits = fits && add(s.d_out + off * P.ct, p.dev_ct, n * P.ct);
            fits = fits && add(dss + off * 32, p.dev_ss, n * 32);
            if (!enc && p.req.fail) fits = fits && add(dfail + off, p.dev_fail, n);
        }
        off += (int64_t)n;
    }
    if (any_staged) fits = fits && add(s.d_out, s.h_out_dev, out_total);
    fits = fits && add(reinterpret_cast<const uint8_t *>(s.d_status), reinterpret_cast<uint8_t *>(s.h_status_dev),
                       sizeof(int));
    if (!fits) return TMVP_EARG; // cannot happen: desc_cap covers max_batch requests
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)(nd - nd_in < kCopyGrid ? nd - nd_in : kCopyGrid)), dim3(256), 0,
                       s.stream, s.h_desc_dev + nd_in, nd - nd_in);
    if (hipGetLastError() != hipSuccess || hipEventRecord(s.done, s.stream) != hipSuccess) return TMVP_EHIP;
    return TMVP_OK;
}

// Copy a finished slot's results to the requests' buffers and describe the
// completions (worker thread, without the lock: the slot is the worker's).
void copy_out(tmvp_server *S, Slot &s, int rc, std::vector<tmvp_srv_completion> &out)
{
    const KemParams P = kem_params_of(s.ps);
    const int64_t B = s.instances;
    if (!rc && *s.h_status) {
        rc = TMVP_ERANGE;
        *s.h_status = 0;
        (void)hipMemsetAsync(s.d_status, 0, sizeof(int), s.stream);
    }
    const bool enc = s.op == TMVP_SRV_ENCAPS;
    const uint8_t *hct = s.h_out, *hss = enc ? s.h_out + (size_t)B * P.ct : s.h_out, *hfail = s.h_out + (size_t)B * 32;
    int64_t off = 0;
    for (const Pending &p : s.reqs) {
        const int64_t n = p.req.count;
        if (!rc && !p.pinned_out) {
            if (enc) memcpy(p.req.ct, hct + off * P.ct, (size_t)n * P.ct);
            memcpy(p.req.ss, hss + off * 32, (size_t)n * 32);
            if (!enc && p.req.fail) memcpy(p.req.fail, hfail + off, (size_t)n);
        }
        tmvp_srv_completion c;
        c.id = p.id;
        c.tag = p.req.tag;
        c.status = rc;
        c.batch_requests = (int)s.reqs.size();
        c.batch_instances = B;
        c.submit_us = p.t_submit;
        c.start_us = s.t_start;
        c.done_us = 0;
        out.push_back(c);
        off += n;
    }
}

// Publish a retired slot's completions (lock held).
void publish(tmvp_server *S, Slot &s, std::vector<tmvp_srv_completion> &cs)
{
    const double t = S->now_us();
    for (tmvp_srv_completion &c : cs) {
        c.done_us = t;
        S->done.push_back(c);
    }
    S->stats.requests_done += (int64_t)s.reqs.size();
    S->stats.instances_done += s.instances;
    S->stats.batches += 1;
    S->inflight -= 1;
    S->inflight_bytes -= (int64_t)s.bytes;
    s.bytes = 0;
    s.busy = false;
    s.reqs.clear();
}

// Free arena bytes for one more batch: the budget left after the key stores
// and the batches in flight, and the largest gap between those batches'
// ranges of the arena (lock held).  *at receives the gap's offset.
size_t room(const tmvp_server *S, size_t *at)
{
    std::vector<std::pair<size_t, size_t>> used;
    for (const Slot &o : S->slots)
        if (o.busy && o.bytes) used.emplace_back(o.off, o.bytes);
    std::sort(used.begin(), used.end());
    size_t best = 0, best_at = 0, cur = 0;
    for (const auto &u : used) {
        if (u.first - cur > best) {
            best = u.first - cur;
            best_at = cur;
        }
        cur = u.first + u.second;
    }
    if (S->arena_bytes - cur > best) {
        best = S->arena_bytes - cur;
        best_at = cur;
    }
    const int64_t left = S->budget - S->store_bytes - S->inflight_bytes;
    if (left <= 0) return 0;
    *at = best_at;
    return best < (size_t)left ? best : (size_t)left;
}

// Form the next batch (lock held): from the (set, op) queue whose head waited
// longest, requests in arrival order while the batch stays within max_batch
// instances and its device bytes within the room left.  Returns false, and
// leaves every queue as it was, when not even that queue's first request fits
// now: it keeps its place, and nothing behind it overtakes it.
bool pick_batch(tmvp_server *S, Slot &s)
{
    int bps = -1, bop = -1;
    double oldest = 0;
    for (int ps = 0; ps < TMVP_NUM_PARAM_SETS; ps++)
        for (int op = 0; op < 2; op++) {
            const auto &q = S->queue[ps][op];
            if (!q.empty() && (bps < 0 || q.front().t_submit < oldest)) {
                bps = ps;
                bop = op;
                oldest = q.front().t_submit;
            }
        }
    if (bps < 0) return false;
    auto &q = S->queue[bps][bop];
    const KemParams P = kem_params_of(bps);
    size_t at = 0;
    const size_t avail = room(S, &at);
    int64_t n = 0;
    size_t k = 0;
    while (k < q.size() && n + q[k].req.count <= S->max_batch &&
           batch_layout(P, bop, n + q[k].req.count, S->backend).total <= avail) {
        n += q[k].req.count;
        k++;
    }
    if (k == 0) return false;
    s.ps = bps;
    s.op = bop;
    s.instances = n;
    s.reqs.clear();
    for (size_t i = 0; i < k; i++) {
        s.reqs.push_back(std::move(q.front()));
        q.pop_front();
    }
    const BatchLayout b = batch_layout(P, bop, n, S->backend);
    s.off = at;
    s.bytes = b.total;
    s.ws = S->arena + at;
    s.d_in = s.ws + b.ws;                 // keys | inputs
    s.d_out = s.d_in + b.keys + b.in;
    S->inflight_bytes += (int64_t)b.total;
    if (S->inflight_bytes > S->stats.max_inflight_bytes) S->stats.max_inflight_bytes = S->inflight_bytes;
    S->pending_requests -= (int64_t)k;
    return true;
}

// The scheduler.  Under the lock it only moves requests and completions
// between queues; staging copies, launches and result copies run unlocked,
// so submit() and poll() never wait behind a transfer.
void worker_main(tmvp_server *S)
{
    (void)hipSetDevice(S->device);
    std::vector<Slot *> finished, starting;
    std::vector<int> ferr;
    std::vector<std::vector<tmvp_srv_completion>> outs(S->slots.size());
    std::unique_lock<std::mutex> lk(S->mu);
    while (!S->stop) {
        finished.clear();
        ferr.clear();
        starting.clear();
        for (Slot &s : S->slots) {
            if (!s.busy) continue;
            const hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            finished.push_back(&s);
            ferr.push_back(q == hipSuccess ? TMVP_OK : TMVP_EHIP);
        }
        for (Slot &s : S->slots) {
            if (s.busy || S->pending_requests == 0) continue;
            if (!pick_batch(S, s)) break; // waits for budget: batches in flight give it back
            s.busy = true;
            s.t_start = S->now_us();
            S->inflight += 1;
            if (S->inflight > S->stats.max_concurrent) S->stats.max_concurrent = S->inflight;
            starting.push_back(&s);
        }
        if (finished.empty() && starting.empty()) {
            if (S->inflight == 0)
                S->wake.wait(lk, [S] { return S->stop || S->pending_requests > 0; });
            else {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(10));
                lk.lock();
            }
            continue;
        }
        lk.unlock();
        for (size_t i = 0; i < finished.size(); i++) {
            const size_t k = (size_t)(finished[i] - S->slots.data());
            outs[k].clear();
            copy_out(S, *finished[i], ferr[i], outs[k]);
        }
        std::vector<Slot *> failed;
        for (Slot *s : starting) {
            if (launch_slot(S, *s)) {
                (void)hipStreamSynchronize(s->stream);
                const size_t k = (size_t)(s - S->slots.data());
                outs[k].clear();
                copy_out(S, *s, TMVP_EHIP, outs[k]);
                failed.push_back(s);
            }
        }
        lk.lock();
        for (Slot *s : finished) publish(S, *s, outs[(size_t)(s - S->slots.data())]);
        for (Slot *s : failed) publish(S, *s, outs[(size_t)(s - S->slots.data())]);
        if (!finished.empty() || !failed.empty()) S->completed.notify_all();
    }
}

// Wait until no batch is in flight and nothing is queued (lock held via lk).
void drain_locked(tmvp_server *S, std::unique_lock<std::mutex> &lk)
{
    S->completed.wait(lk, [S] { return S->inflight == 0 && S->pending_requests == 0; });
}

// New capacity for `need` keys: exactly `need` for a store's first keys,
// else at least double the old capacity (fewer copies as keys trickle in).
int64_t grown_cap(const Store &st, int64_t need)
{
    if (need <= st.cap) return st.cap;
    return st.cap == 0 ? need : (need > 2 * st.cap ? need : 2 * st.cap);
}

int grow_store(Store &st, const KemParams &P, int64_t need, bool with_sk)
{
    if (need <= st.cap) return TMVP_OK;
    const int64_t cap = grown_cap(st, need);
    const size_t abytes = (size_t)P.l * P.l * kN * 2;
    uint8_t *pk = nullptr, *sk = nullptr, *hpk = nullptr;
    uint16_t *A = nullptr;
    if (hipMalloc(&pk, (size_t)cap * P.pk) != hipSuccess || hipMalloc(&hpk, (size_t)cap * 32) != hipSuccess ||
        hipMalloc(&A, (size_t)cap * abytes) != hipSuccess || (with_sk && hipMalloc(&sk, (size_t)cap * P.sk) != hipSuccess)) {
        (void)hipFree(pk);
        (void)hipFree(hpk);
        (void)hipFree(A);
        (void)hipFree(sk);
        return TMVP_ENOMEM;
    }
    if (st.n) {
        (void)hipMemcpy(pk, st.pk, (size_t)st.n * P.pk, hipMemcpyDeviceToDevice);
        (void)hipMemcpy(hpk, st.hpk, (size_t)st.n * 32, hipMemcpyDeviceToDevice);
        (void)hipMemcpy(A, st.A, (size_t)st.n * abytes, hipMemcpyDeviceToDevice);
        if (with_sk) (void)hipMemcpy(sk, st.sk, (size_t)st.n * P.sk, hipMemcpyDeviceToDevice);
    }
    (void)hipFree(st.pk);
    (void)hipFree(st.hpk);
    (void)hipFree(st.A);
    (void)hipFree(st.sk);
    st.pk = pk;
    st.sk = sk;
    st.hpk = hpk;
    st.A = A;
    st.cap = cap;
    st.bytes = store_bytes_of(P, cap, with_sk);
    return TMVP_OK;
}

} // namespace

extern "C" {

int tmvp_srv_create(tmvp_ctx *ctx, const tmvp_srv_config *cfg, tmvp_server **out)
{
    if (!ctx || !out) return TMVP_EARG;
    *out = nullptr;
    tmvp_srv_config c = cfg ? *cfg : tmvp_srv_config{};
    if (c.hbm_budget < 0 || c.max_batch < 0 || c.slots < 0 || !kem_backend_ok(c.backend)) return TMVP_EARG;
    if (c.hbm_budget == 0) c.hbm_budget = (int64_t)4 << 30;
    if (c.max_batch == 0) c.max_batch = 16384;
    // Four concurrent batches by default: with a hardware queue per slot
    // stream (GPU_MAX_HW_QUEUES=8), four slots at 6000 req/s gave p50 0.35 ms /
    // p99 0.67 ms, three 0.35 / 0.76 ms and six 0.37 / 0.97 ms.
    if (c.slots == 0) c.slots = 4;
    if (c.max_batch > ((int64_t)1 << 24)) return TMVP_EARG;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return TMVP_EHIP;
    tmvp_server *S = new (std::nothrow) tmvp_server;
    if (!S) return TMVP_ENOMEM;
    S->ctx = ctx;
    S->device = dev;
    S->backend = c.backend;
    S->max_batch = c.max_batch;
    S->t0 = Clock::now();
    size_t batch_max = 0;
    max_sizes(c.max_batch, c.backend, S->in_bytes, S->out_bytes, batch_max);
    S->max_batch_bytes = (int64_t)batch_max;
    S->budget = c.hbm_budget;
    // every request has at least one instance: inputs take at most one piece
    // per request plus in_bytes / kPiece and the keys' pieces, outputs three
    // per request plus out_bytes / kPiece, the staging and the flag
    S->desc_cap = 5 * c.max_batch + (int64_t)((S->in_bytes + S->out_bytes) / kPiece) + 8;
    if (c.hbm_budget < S->max_batch_bytes) {
        delete S;
        return TMVP_ENOMEM; // the budget does not cover one batch of max_batch instances
    }
    const int nslots = c.slots;
    S->slots.resize(nslots);
    S->arena_bytes = (size_t)arena_size(S, 0);
    if (hipMalloc(&S->arena, S->arena_bytes) != hipSuccess) {
        delete S;
        return TMVP_ENOMEM;
    }
    for (Slot &s : S->slots) {
        const int rc = alloc_slot(S, s);
        if (rc) {
            for (Slot &f : S->slots) free_slot(f);
            (void)hipFree(S->arena);
            delete S;
            return rc;
        }
    }
    S->stats.slots = nslots;
    S->stats.slot_bytes = S->max_batch_bytes;
    S->stats.hbm_budget = S->budget;
    S->stats.arena_bytes = (int64_t)S->arena_bytes;
    try {
        S->worker = std::thread(worker_main, S);
    } catch (...) {
        for (Slot &f : S->slots) free_slot(f);
        (void)hipFree(S->arena);
        delete S;
        return TMVP_ENOMEM;
    }
    *out = S;
    return TMVP_OK;
}

int tmvp_srv_add_keys(tmvp_server *S, int ps, const uint8_t *h_pk, const uint8_t *h_sk, int64_t n, int64_t *first_id)
{
    if (!S || ps < 0 || ps >= TMVP_NUM_PARAM_SETS || n < 0 || (n > 0 && !h_pk)) return TMVP_EARG;
    std::unique_lock<std::mutex> lk(S->mu);
    drain_locked(S, lk); // batches in flight read the store
    Store &st = S->store[ps];
    if (st.n > 0 && st.has_sk != (h_sk != nullptr)) return TMVP_EARG; // all keys of a set with or without sk
    if (st.n + n > ((int64_t)1 << 31) - 1) return TMVP_EARG;
    if (first_id) *first_id = st.n;
    if (n == 0) return TMVP_OK;
    const KemParams P = kem_params_of(ps);
    // Keyed decapsulation reads A and H(pk) derived from h_pk, and b, H(pk)
    // and z from the sk: each sk must hold its own pk (sk = s | pk | H(pk) | z).
    if (h_sk)
        for (int64_t i = 0; i < n; i++)
            if (memcmp(h_sk + (size_t)i * P.sk + P.indcpa_sk, h_pk + (size_t)i * P.pk, (size_t)P.pk) != 0)
                return TMVP_EARG;
    const bool with_sk = st.n == 0 ? h_sk != nullptr : st.has_sk;
    // The stores count against the HBM budget, which must still hold one
    // batch of max_batch instances after them.
    const int64_t grown = store_bytes_of(P, grown_cap(st, st.n + n), with_sk);
    if (S->store_bytes - st.bytes + grown + S->max_batch_bytes > S->budget) return TMVP_ENOMEM;
    // No batch is in flight: shrink the arena first, so that the arena and
    // the stores together never hold more device memory than the budget.
    const int64_t new_arena = arena_size(S, S->store_bytes - st.bytes + grown);
    if ((size_t)new_arena < S->arena_bytes) {
        (void)hipFree(S->arena);
        S->arena = nullptr;
        S->arena_bytes = 0;
        if (hipMalloc(&S->arena, (size_t)new_arena) != hipSuccess) return TMVP_ENOMEM;
        S->arena_bytes = (size_t)new_arena;
        S->stats.arena_bytes = new_arena;
    }
    st.has_sk = with_sk;
    const int64_t before = st.bytes;
    int rc = grow_store(st, P, st.n + n, st.has_sk);
    if (rc) return rc;
    S->store_bytes += st.bytes - before;
    S->stats.store_bytes = S->store_bytes;
    const size_t abytes = (size_t)P.l * P.l * kN * 2;
    rc = TMVP_OK;
    if (hipMemcpy(st.pk + (size_t)st.n * P.pk, h_pk, (size_t)n * P.pk, hipMemcpyHostToDevice) != hipSuccess ||
        (h_sk && hipMemcpy(st.sk + (size_t)st.n * P.sk, h_sk, (size_t)n * P.sk, hipMemcpyHostToDevice) != hipSuccess))
        rc = TMVP_EHIP;
    if (!rc && kem_store_derive(P, st.pk + (size_t)st.n * P.pk, n, st.A + (size_t)st.n * (abytes / 2),
                                st.hpk + (size_t)st.n * 32, nullptr) != hipSuccess)
        rc = TMVP_EHIP;
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = TMVP_EHIP;
    if (!rc) st.n += n;
    return rc;
}

int tmvp_srv_submit(tmvp_server *S, const tmvp_srv_request *r, uint64_t *id)
{
    if (!S || !r) return TMVP_EARG;
    if (r->op != TMVP_SRV_ENCAPS && r->op != TMVP_SRV_DECAPS) return TMVP_EARG;
    if (r->param_set < 0 || r->param_set >= TMVP_NUM_PARAM_SETS) return TMVP_EARG;
    if (r->count < 1 || !r->keys || !r->in || !r->ss || (r->op == TMVP_SRV_ENCAPS && !r->ct)) return TMVP_EARG;
    // outside the lock: pointer queries of the caller's buffers
    const void *dev_in = nullptr, *dev_ct = nullptr, *dev_ss = nullptr, *dev_fail = nullptr;
    const bool pinned_in = is_pinned(r->in, &dev_in);
    const bool pinned_out = is_pinned(r->ss, &dev_ss) && (r->op != TMVP_SRV_ENCAPS || is_pinned(r->ct, &dev_ct)) &&
                            (r->op != TMVP_SRV_DECAPS || is_pinned(r->fail, &dev_fail));
    if (r->count > S->max_batch) return TMVP_EARG; // fixed at creation
    // The ids are checked and copied in one pass, so what the kernels index
    // with is exactly what was checked, whatever the caller does with its array.
    Pending p;
    p.keys.resize((size_t)r->count);
    std::lock_guard<std::mutex> g(S->mu);
    const Store &st = S->store[r->param_set];
    if (r->op == TMVP_SRV_DECAPS && !st.has_sk) return TMVP_EARG;
    for (int64_t i = 0; i < r->count; i++) {
        const int32_t k = r->keys[i];
        if (k < 0 || k >= st.n) return TMVP_EARG;
        p.keys[(size_t)i] = k;
    }
    p.id = S->next_id++;
    p.req = *r;
    p.req.keys = nullptr;
    p.t_submit = S->now_us();
    p.pinned_in = pinned_in;
    p.pinned_out = pinned_out;
    p.dev_in = static_cast<const uint8_t *>(dev_in);
    p.dev_ct = static_cast<uint8_t *>(const_cast<void *>(dev_ct));
    p.dev_ss = static_cast<uint8_t *>(const_cast<void *>(dev_ss));
    p.dev_fail = static_cast<uint8_t *>(const_cast<void *>(dev_fail));
    S->queue[r->param_set][r->op].push_back(std::move(p));
    S->pending_requests += 1;
    S->stats.requests_submitted += 1;
    if (id) *id = p.id;
    S->wake.notify_one();
    return TMVP_OK;
}

int tmvp_srv_poll(tmvp_server *S, tmvp_srv_completion *out, int max, int timeout_us)
{
    if (!S || max < 0 || (max > 0 && !out)) return TMVP_EARG;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->done.empty() && timeout_us > 0)
        S->completed.wait_for(lk, std::chrono::microseconds(timeout_us), [S] { return !S->done.empty(); });
    int k = 0;
    while (k < max && !S->done.empty()) {
        out[k++] = S->done.front();
        S->done.pop_front();
    }
    return k;
}

int tmvp_srv_drain(tmvp_server *S)
{
    if (!S) return TMVP_EARG;
    std::unique_lock<std::mutex> lk(S->mu);
    drain_locked(S, lk);
    return TMVP_OK;
}

int tmvp_srv_get_stats(tmvp_server *S, tmvp_srv_stats *out)
{
    if (!S || !out) return TMVP_EARG;
    std::lock_guard<std::mutex> g(S->mu);
    *out = S->stats;
    return TMVP_OK;
}

int tmvp_srv_destroy(tmvp_server *S)
{
    if (!S) return TMVP_EARG;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        drain_locked(S, lk);
        S->stop = true;
        S->wake.notify_all();
    }
    S->worker.join();
    for (Slot &s : S->slots) free_slot(s);
    (void)hipFree(S->arena);
    for (Store &st : S->store) {
        (void)hipFree(st.pk);
        (void)hipFree(st.sk);
        (void)hipFree(st.hpk);
        (void)hipFree(st.A);
    }
    delete S;
    return TMVP_OK;
}

} // extern "C"
