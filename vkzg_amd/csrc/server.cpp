// Request server: continuous batching of proof and recovery requests of any
// size on one prover (include/vkzg_amd.h, vkzg_amd_server_*).
//
// Requests queue FIFO.  A dispatcher thread packs the oldest request's kind
// into device batches of up to `batch_blobs` blobs (a large request spans
// several batches, several small ones share one) and launches each batch on a
// free slot: a working set, sub-streams and pinned staging of its own, so up
// to `slots` batches are in flight at once and requests that arrive while
// others run are admitted as soon as a slot frees.  The number of slots is
// bounded by the HBM budget.  A completer thread waits for batches in launch
// order and scatters results back; blob errors are tracked per blob, so an
// invalid blob fails only its own request.  Tables are the prover's,
// read-only and shared.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "../../include/vkzg_amd.h"
#include "runtime.h"

using namespace vkzg_amd;

namespace {

using Clock = std::chrono::steady_clock;

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               Clock::now().time_since_epoch())
        .count();
}

enum class Kind { Proofs, Recovery };

struct Request {
    uint64_t ticket = 0;
    Kind kind = Kind::Proofs;
    size_t n = 0;
    const uint8_t *in = nullptr;       // blobs or cells
    const uint8_t *present = nullptr;  // recovery mask
    uint8_t *out = nullptr;
    size_t placed = 0;    // blobs already put into batches
    size_t finished = 0;  // blobs whose results are back
    uint32_t batches = 0;
    vkzg_result status = VKZG_OK;
    bool done = false;
    uint64_t t_submit = 0, t_start = 0, t_done = 0;
};

// Blobs [first, first + count) of `req`, at offset `off` of a batch.
struct Piece {
    Request *req;
    size_t first;
    uint32_t count;
    uint32_t off;
};

struct Slot {
    WorkSet ws;
    StreamSet ss;
    GraphCache graphs;
    hipStream_t stream = nullptr;
    hipEvent_t done_ev = nullptr;
    uint8_t *h_in = nullptr, *h_out = nullptr, *h_present = nullptr;
    uint32_t *h_err = nullptr;
    Kind kind = Kind::Proofs;
    uint32_t nb = 0;
    std::vector<Piece> pieces;
    bool busy = false;
    bool failed = false;
};

}  // namespace

struct vkzg_amd_server {
    vkzg_prover *p = nullptr;
    uint32_t cap = 0;
    uint64_t slot_bytes = 0;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv_dispatch, cv_complete, cv_done;
    std::deque<Request *> pending;
    std::deque<Slot *> inflight;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t next_ticket = 1;
    uint32_t launching = 0;  // batches formed but not yet in `inflight`
    bool stop = false;
    std::thread dispatcher, completer;
    uint64_t n_requests = 0, n_blobs = 0, n_batches = 0;

    ~vkzg_amd_server();
    void dispatch_loop();
    void complete_loop();
    bool launch(Slot &slot);
};

namespace {

// Fills `slot` with pieces of pending requests of the oldest request's kind,
// up to cap blobs.  Called with the server mutex held.
bool form_batch(vkzg_amd_server *sv, Slot &slot) {
    if (sv->pending.empty()) return false;
    slot.kind = sv->pending.front()->kind;
    slot.pieces.clear();
    slot.nb = 0;
    for (auto it = sv->pending.begin(); it != sv->pending.end() && slot.nb < sv->cap;) {
        Request *r = *it;
        if (r->kind != slot.kind) {
            ++it;
            continue;
        }
        const uint32_t take = (uint32_t)std::min<size_t>(r->n - r->placed, sv->cap - slot.nb);
        if (r->placed == 0) r->t_start = now_ns();
        slot.pieces.push_back(Piece{r, r->placed, take, slot.nb});
        slot.nb += take;
        r->placed += take;
        r->batches++;
        if (r->placed == r->n) it = sv->pending.erase(it);
        else ++it;
    }
    return slot.nb > 0;
}

}  // namespace

bool vkzg_amd_server::launch(Slot &slot) {
    const size_t in_bytes = slot.kind == Kind::Proofs ? kBlobBytes : kCellsBytes;
    const size_t out_bytes = slot.kind == Kind::Proofs ? kProofBytes : kCellsBytes;
    for (const Piece &pc : slot.pieces) {
        std::memcpy(slot.h_in + (size_t)pc.off * in_bytes, pc.req->in + pc.first * in_bytes,
                    (size_t)pc.count * in_bytes);
        if (slot.kind == Kind::Recovery)
            std::memcpy(slot.h_present + (size_t)pc.off * 128, pc.req->present + pc.first * 128,
                        (size_t)pc.count * 128);
    }
    try {
        DeviceGuard guard(p->device);
        hipStream_t s = slot.stream;
        StageClock clk(p, s, false);
        const Tuning tn = tuning_for(p, slot.nb);
        if (slot.kind == Kind::Proofs) {
            ck(hipMemcpyAsync(slot.ws.blobs, slot.h_in, slot.nb * in_bytes, hipMemcpyHostToDevice, s));
            enqueue_proofs(p, slot.ws, slot.ss, slot.ws.proofs, slot.ws.blobs, slot.nb, s, clk, tn,
                           &slot.graphs);
            ck(hipMemcpyAsync(slot.h_err, slot.ws.err, slot.nb * 4, hipMemcpyDeviceToHost, s));
            ck(hipMemcpyAsync(slot.h_out, slot.ws.proofs, slot.nb * out_bytes, hipMemcpyDeviceToHost, s));
        } else {
            ck(hipMemcpyAsync(slot.ws.cells, slot.h_in, slot.nb * in_bytes, hipMemcpyHostToDevice, s));
            ck(hipMemcpyAsync(slot.ws.present, slot.h_present, slot.nb * 128, hipMemcpyHostToDevice, s));
            enqueue_recovery(p, slot.ws, slot.ss, slot.ws.cells_out, slot.ws.cells, slot.ws.present,
                             slot.nb, s, clk, tn, &slot.graphs);
            ck(hipMemcpyAsync(slot.h_out, slot.ws.cells_out, slot.nb * out_bytes, hipMemcpyDeviceToHost, s));
        }
        ck(hipEventRecord(slot.done_ev, s));
    } catch (HipFail &) {
        return false;
    }
    return true;
}

void vkzg_amd_server::dispatch_loop() {
    (void)hipSetDevice(p->device);
    std::unique_lock<std::mutex> lock(mu);
    for (;;) {
        Slot *free_slot = nullptr;
        cv_dispatch.wait(lock, [&] {
            if (stop && pending.empty()) return true;
            if (pending.empty()) return false;
            for (auto &s : slots)
                if (!s->busy) {
                    free_slot = s.get();
                    return true;
                }
            return false;
        });
        if (stop && pending.empty()) return;
        if (!free_slot || !form_batch(this, *free_slot)) continue;
        free_slot->busy = true;
        n_batches++;
        launching++;
        lock.unlock();
        // Staging copies and launches run without the lock: request buffers
        // stay valid until their request completes.
        free_slot->failed = !launch(*free_slot);
        lock.lock();
        launching--;
        inflight.push_back(free_slot);
        cv_complete.notify_one();
    }
}

void vkzg_amd_server::complete_loop() {
    (void)hipSetDevice(p->device);
    std::unique_lock<std::mutex> lock(mu);
    for (;;) {
        const auto drained = [&] { return stop && pending.empty() && launching == 0; };
        cv_complete.wait(lock, [&] { return !inflight.empty() || drained(); });
        if (inflight.empty()) {
            if (drained()) return;
            continue;
        }
        Slot *slot = inflight.front();
        lock.unlock();
        bool ok = !slot->failed && hipEventSynchronize(slot->done_ev) == hipSuccess;
        const size_t out_bytes = slot->kind == Kind::Proofs ? kProofBytes : kCellsBytes;
        // Scatter results (request outputs are disjoint per piece).
        for (const Piece &pc : slot->pieces) {
            if (!ok) continue;
            std::memcpy(pc.req->out + pc.first * out_bytes, slot->h_out + (size_t)pc.off * out_bytes,
                        (size_t)pc.count * out_bytes);
        }
        lock.lock();
        const uint64_t t = now_ns();
        for (const Piece &pc : slot->pieces) {
            Request *r = pc.req;
            if (!ok) r->status = VKZG_ERR_GPU;
            else if (slot->kind == Kind::Proofs)
                for (uint32_t b = 0; b < pc.count; b++)
                    if (slot->h_err[pc.off + b] && r->status == VKZG_OK) r->status = VKZG_ERR_INVALID_BLOB;
            r->finished += pc.count;
            if (r->finished == r->n) {
                r->done = true;
                r->t_done = t;
            }
        }
        inflight.pop_front();
        slot->busy = false;
        cv_done.notify_all();
        cv_dispatch.notify_one();
    }
}

vkzg_amd_server::~vkzg_amd_server() {
    {
        std::lock_guard<std::mutex> lock(mu);
        stop = true;
    }
    cv_dispatch.notify_all();
    cv_complete.notify_all();
    if (dispatcher.joinable()) dispatcher.join();
    if (completer.joinable()) completer.join();
    DeviceGuard guard(p->device);
    for (auto &s : slots) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        if (s->done_ev) (void)hipEventDestroy(s->done_ev);
        s->graphs.clear();
        s->ss.destroy();
        if (s->stream) (void)hipStreamDestroy(s->stream);
    }
    slots.clear();  // frees the working sets while the device is current
}

namespace {

vkzg_result submit(vkzg_amd_server *sv, Kind kind, uint8_t *out, const uint8_t *in,
                   const uint8_t *present, size_t n, uint64_t *ticket) {
    if (!sv || !out || !in || !ticket || (kind == Kind::Recovery && !present)) return VKZG_ERR_BADARGS;
    auto r = std::make_unique<Request>();
    r->kind = kind;
    r->n = n;
    r->in = in;
    r->present = present;
    r->out = out;
    r->t_submit = now_ns();
    if (kind == Kind::Recovery && !chunk_missing_ok(present, (uint32_t)n)) return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(sv->mu);
    if (sv->stop) return VKZG_ERR_BADARGS;
    r->ticket = sv->next_ticket++;
    *ticket = r->ticket;
    sv->n_requests++;
    sv->n_blobs += n;
    if (n == 0) {
        r->done = true;
        r->t_start = r->t_done = r->t_submit;
    } else {
        sv->pending.push_back(r.get());
    }
    sv->requests.emplace(r->ticket, std::move(r));
    sv->cv_dispatch.notify_one();
    return VKZG_OK;
}

}  // namespace

extern "C" {

void vkzg_amd_server_options_default(vkzg_amd_server_options *o) {
    if (!o) return;
    o->batch_blobs = 0;
    o->slots = 0;
    o->hbm_budget_bytes = 0;
}

vkzg_result vkzg_amd_server_new(vkzg_amd_server **out, vkzg_prover *p,
                                const vkzg_amd_server_options *opts) {
    if (!out || !p) return VKZG_ERR_BADARGS;
    *out = nullptr;
    auto sv = std::make_unique<vkzg_amd_server>();
    sv->p = p;
    sv->cap = opts && opts->batch_blobs ? opts->batch_blobs : p->max_batch;
    uint32_t nslots = opts && opts->slots ? opts->slots : 2;
    const uint64_t budget = opts ? opts->hbm_budget_bytes : 0;
    const uint64_t per_blob = WorkSet::bytes_per_blob();
    if (budget) {
        // Admission by HBM budget: the prover's resident phase-A tables count
        // against it; then as many slots as fit, and smaller batches when
        // even one full slot would not.
        uint64_t tables = 0;
        {
            std::lock_guard<std::mutex> lock(p->mu);
            tables = p->table_bytes;
            for (const auto &t : p->extra_tables) tables += t.bytes;
        }
        if (budget <= tables) return VKZG_ERR_MALLOC;
        const uint64_t fit_blobs = (budget - tables) / per_blob;
        if (fit_blobs == 0) return VKZG_ERR_MALLOC;
        if (fit_blobs < sv->cap) sv->cap = (uint32_t)fit_blobs;
        nslots = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(nslots, fit_blobs / sv->cap));
    }
    sv->slot_bytes = per_blob * sv->cap;
    try {
        DeviceGuard guard(p->device);
        for (uint32_t i = 0; i < nslots; i++) {
            auto s = std::make_unique<Slot>();
            s->ws.allocate(sv->cap);
            s->ss.create();
            ck(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
            ck(hipEventCreateWithFlags(&s->done_ev, hipEventBlockingSync | hipEventDisableTiming));
            s->h_in = s->ws.arena.halloc<uint8_t>((size_t)sv->cap * kCellsBytes);
            s->h_out = s->ws.arena.halloc<uint8_t>((size_t)sv->cap * kCellsBytes);
            s->h_present = s->ws.arena.halloc<uint8_t>((size_t)sv->cap * 128);
            s->h_err = s->ws.arena.halloc<uint32_t>(sv->cap);
            sv->slots.push_back(std::move(s));
        }
    } catch (HipFail &) {
        return VKZG_ERR_MALLOC;
    }
    vkzg_amd_server *raw = sv.get();
    raw->dispatcher = std::thread([raw] { raw->dispatch_loop(); });
    raw->completer = std::thread([raw] { raw->complete_loop(); });
    *out = sv.release();
    return VKZG_OK;
}

void vkzg_amd_server_free(vkzg_amd_server *s) { delete s; }

vkzg_result vkzg_amd_server_submit_proofs(vkzg_amd_server *s, uint8_t *proofs, const uint8_t *blobs,
                                          size_t num_blobs, uint64_t *ticket) {
    return submit(s, Kind::Proofs, proofs, blobs, nullptr, num_blobs, ticket);
}

vkzg_result vkzg_amd_server_submit_recovery(vkzg_amd_server *s, uint8_t *cells_out,
                                            const uint8_t *cells, const uint8_t *cell_present,
                                            size_t num_blobs, uint64_t *ticket) {
    return submit(s, Kind::Recovery, cells_out, cells, cell_present, num_blobs, ticket);
}

vkzg_result vkzg_amd_server_poll(vkzg_amd_server *s, uint64_t ticket, int *done) {
    if (!s || !done) return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) return VKZG_ERR_BADARGS;
    *done = it->second->done ? 1 : 0;
    return VKZG_OK;
}

vkzg_result vkzg_amd_server_wait(vkzg_amd_server *s, uint64_t ticket, vkzg_amd_request_info *info) {
    if (!s) return VKZG_ERR_BADARGS;
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) return VKZG_ERR_BADARGS;
    Request *r = it->second.get();
    s->cv_done.wait(lock, [&] { return r->done; });
    const vkzg_result status = r->status;
    if (info) {
        info->submit_ns = r->t_submit;
        info->start_ns = r->t_start;
        info->done_ns = r->t_done;
        info->batches = r->batches;
    }
    s->requests.erase(it);
    return status;
}

vkzg_result vkzg_amd_server_stats(vkzg_amd_server *s, vkzg_amd_server_counters *out) {
    if (!s || !out) return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(s->mu);
    out->requests = s->n_requests;
    out->blobs = s->n_blobs;
    out->batches = s->n_batches;
    out->slots = (uint32_t)s->slots.size();
    out->batch_blobs = s->cap;
    out->slot_bytes = s->slot_bytes;
    return VKZG_OK;
}

}  // extern "C"
