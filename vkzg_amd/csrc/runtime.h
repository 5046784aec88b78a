// Internal host-side types shared by the prover (prover.cpp) and the request
// server (server.cpp).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vkzg.h"
#include "curve.h"
#include "kernels.h"

namespace vkzg_amd {

constexpr size_t kBlobBytes = VKZG_BYTES_PER_BLOB;
constexpr size_t kProofBytes = (size_t)VKZG_NUM_CELL_PROOFS * VKZG_BYTES_PER_PROOF;
constexpr size_t kCellsBytes = (size_t)VKZG_NUM_CELL_PROOFS * VKZG_BYTES_PER_CELL;
constexpr int kMaxSplits = 4;  // sub-streams one batch may be spread over
constexpr uint32_t kMaxGroups = 20;  // phase-A window groups per MSM

// Launch shape of the proof pipeline for one batch-size class.  Every choice
// gives bit-identical proofs; only the time differs.
struct Tuning {
    uint32_t groups = 1;         // phase-A window groups (tiles) per MSM, 1..kMaxGroups
    uint32_t blocks_per_cu = 2;  // persistent phase-A grid = CUs x this, 1..2
    uint32_t norm_chunk = 0;     // points per batched inversion in the normalisation (0: fitted to the chip)
    uint32_t splits = 1;         // sub-streams, 1..kMaxSplits
    uint32_t rec_waves = 1;      // recovery work items (waves) per block: 1, 2 or 4
    uint32_t rec_radix = 2;      // recovery: 2 (three kernels, 4 values per lane), 4 or 8 (five kernels)
    uint32_t ladder_splits = 1;  // ladder segments (phase A also computes their starts): 1, 2 or 4
    uint32_t reduce_lanes = 8;   // bucket-reduction lanes per output: 4, 8, 16 or 32 (one per bucket
                                 // range), or 64 (16 ranges of a quad each)
    uint32_t window_bits = 0;    // phase-A table: 0 = the prover's main table, else a resident
                                 // smaller one of this window (10..13, below the main one)
    // Not a tuning: the test switch that sends phase A, phase B and the bucket
    // reduction down their checked recompute (vkzg_amd_debug_force_checked).
    // Set by tuning_for from the prover; kept here so graph keys include it.
    uint32_t force_checked = 0;
};

// A resident phase-A multiples table of window c: W = ceil(256 / c) windows
// of H = 2^(c-1) multiples (table[128][64][W][H], one 128-byte line each).
struct PhaseATable {
    uint32_t c = 0, W = 0, H = 0, k = 0;  // window plan (c, W, k), H = 2^(c-1)
    uint64_t bytes = 0;
    G1AT *table = nullptr;
    WindowPlan plan() const { return WindowPlan{c, W, k}; }
};
// Smaller tables the autotuner may keep beside the main one, by window.
constexpr uint32_t kExtraWindowMin = 10, kExtraWindowMax = 13;

// Batch-size classes: <= 1, <= 4, <= 16, <= 64 and more blobs.
constexpr int kTuneClasses = 5;
constexpr uint32_t kTuneClassTop[kTuneClasses] = {1, 4, 16, 64, 0xffffffffu};
inline int tune_class(uint32_t nb) {
    int c = 0;
    while (c + 1 < kTuneClasses && nb > kTuneClassTop[c]) c++;
    return c;
}

struct HipFail {};

inline void ck(hipError_t e) {
    if (e != hipSuccess) throw HipFail{};
}

struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) {
        (void)hipGetDevice(&prev);
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
};

// Device allocations owned by one object, freed together.
struct DeviceArena {
    std::vector<void *> dev, host;
    template <typename T>
    T *dalloc(size_t count) {
        void *ptr = nullptr;
        ck(hipMalloc(&ptr, count * sizeof(T) ? count * sizeof(T) : 16));
        dev.push_back(ptr);
        return static_cast<T *>(ptr);
    }
    template <typename T>
    T *halloc(size_t count) {
        void *ptr = nullptr;
        ck(hipHostMalloc(&ptr, count * sizeof(T) ? count * sizeof(T) : 16, hipHostMallocDefault));
        host.push_back(ptr);
        return static_cast<T *>(ptr);
    }
    void dfree(void *ptr) {
        for (auto &a : dev)
            if (a == ptr) {
                (void)hipFree(a);
                a = nullptr;
            }
    }
    void release() {
        for (void *a : dev)
            if (a) (void)hipFree(a);
        for (void *a : host)
            if (a) (void)hipHostFree(a);
        dev.clear();
        host.clear();
    }
};

// Per-batch device buffers for `cap` blobs.  Every buffer is indexed by blob,
// so a sub-batch at blob offset `off` uses slices at that offset.
struct WorkSet {
    uint32_t cap = 0;
    DeviceArena arena;
    // proofs
    uint8_t *blobs = nullptr, *proofs = nullptr;
    uint32_t *err = nullptr;
    uint32_t *queue = nullptr;    // phase-A tile counter of the sub-batch at each blob offset
    uint32_t *queue_b = nullptr;  // phase-B work counter, likewise
    uint32_t *queue_r = nullptr;  // bucket-reduction work counter, likewise
    Fr *scal = nullptr;
    G1X *u = nullptr, *buckets = nullptr, *cvals = nullptr;
    uint8_t *bucket_flags = nullptr;  // phase-B buckets to recompute with checked operations
    uint8_t *cval_flags = nullptr;    // reduced outputs to recompute likewise
    G1J *lad = nullptr;
    G1A *lad_aff = nullptr, *lad_phi = nullptr;
    Fp *prefix = nullptr;
    // recovery
    uint8_t *cells = nullptr, *present = nullptr, *cells_out = nullptr;
    Fr *bufA = nullptr, *bufB = nullptr;

    void allocate(uint32_t blobs_cap);
    ~WorkSet() { arena.release(); }
    static uint64_t bytes_per_blob();
};

// Sub-streams a batch is spread over, with the events that fork them from
// and join them back into the caller's stream.
struct StreamSet {
    hipStream_t sub[kMaxSplits] = {};
    hipEvent_t fork = nullptr, join[kMaxSplits] = {};
    void create() {
        for (auto &q : sub) ck(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
        ck(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        for (auto &e : join) ck(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    void destroy() {
        for (auto &e : join)
            if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        for (auto &q : sub)
            if (q) (void)hipStreamDestroy(q);
    }
};

// Instantiated HIP graphs of whole pipelines, keyed by what the captured
// launches depend on (kind, blob count, split count, buffer addresses).  A
// pipeline is captured the first time its key is seen and replayed after
// that, which turns 10-20 kernel launches into one graph launch.  The
// captured pipelines hold kernels and event dependencies only: no memset or
// memcpy nodes.
struct GraphKey {
    uint64_t v[6];
    bool operator==(const GraphKey &o) const {
        for (int i = 0; i < 6; i++)
            if (v[i] != o.v[i]) return false;
        return true;
    }
};

struct GraphCache {
    struct Entry {
        GraphKey key;
        hipGraphExec_t exec;
        uint64_t last_use;
    };
    std::vector<Entry> entries;
    uint64_t tick = 0;
    uint64_t captures = 0, replays = 0;
    static constexpr size_t kCapacity = 256;  // every batch size of both kinds up to 128

    // Enqueue `body(s)` on s through a cached graph.
    template <typename Body>
    void run(const GraphKey &key, hipStream_t s, Body body) {
        hipGraphExec_t exec = nullptr;
        for (auto &e : entries)
            if (e.key == key) {
                exec = e.exec;
                e.last_use = ++tick;
            }
        if (!exec) {
            ck(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            hipGraph_t g = nullptr;
            try {
                body(s);
            } catch (HipFail &) {
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            ck(hipStreamEndCapture(s, &g));
            const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            ck(ie);
            if (entries.size() >= kCapacity) {
                size_t old = 0;
                for (size_t i = 1; i < entries.size(); i++)
                    if (entries[i].last_use < entries[old].last_use) old = i;
                (void)hipGraphExecDestroy(entries[old].exec);
                entries.erase(entries.begin() + (long)old);
            }
            entries.push_back({key, exec, ++tick});
            captures++;
        } else {
            replays++;
        }
        ck(hipGraphLaunch(exec, s));
    }
    void clear() {
        for (auto &e : entries) (void)hipGraphExecDestroy(e.exec);
        entries.clear();
    }
};

// Verification state (verify.cpp): G2 setup lines and per-call buffers.
struct VerifyState;
void destroy_verify_state(VerifyState *v);

}  // namespace vkzg_amd

struct vkzg_prover {
    int device = 0;
    std::string name;
    uint32_t cus = 0;
    hipStream_t stream = nullptr;
    vkzg_amd::StreamSet streams;
    std::mutex mu;
    uint32_t max_batch = 256;
    double setup_ms = 0;
    uint32_t last_splits = 1;
    int forced_splits = 0;  // $VKZG_AMD_SPLITS, 0 = automatic
    std::atomic<bool> use_graphs{true};  // $VKZG_AMD_GRAPHS=0 launches kernel by kernel
    // Tunings per batch-size class.  Written under tune_mu (and mu); read under
    // tune_mu by tuning_for, which the server's dispatcher thread calls
    // without holding mu.
    mutable std::mutex tune_mu;
    vkzg_amd::Tuning tune[vkzg_amd::kTuneClasses];
    std::atomic<uint32_t> force_checked{0};  // vkzg_amd_debug_force_checked
    bool autotuned = false;
    double autotune_ms = 0;
    vkzg_amd::GraphCache graphs;  // graphs of the synchronous entry points

    // Resident tables.
    vkzg_amd::DeviceArena arena;
    vkzg_amd::Fr *roots = nullptr, *roots_inv = nullptr, *shift_fwd = nullptr, *shift_inv = nullptr;
    uint4 *frl_tabs[4] = {};  // the four tables above as FrL entries (radix-4 recovery)
    vkzg_amd::G1A *srs = nullptr, *xext = nullptr;
    vkzg_amd::G1AT *table = nullptr;  // the main phase-A table (window c)
    uint32_t c = 0, W = 0, H = 0, wide = 0;  // its window plan (c, W, wide)
    uint64_t table_bytes = 0;
    // Smaller phase-A tables kept by the autotuner (window kExtraWindowMin + i).
    // A slot is filled under mu before any tuning names it and is freed only
    // if no tuning ever named it, so a tuning copied by another thread always
    // finds its table.
    vkzg_amd::PhaseATable extra_tables[vkzg_amd::kExtraWindowMax - vkzg_amd::kExtraWindowMin + 1];
    bool extra_named[vkzg_amd::kExtraWindowMax - vkzg_amd::kExtraWindowMin + 1] = {};
    uint32_t *kb_items = nullptr, *kb_offsets = nullptr;
    uint32_t kb_count = 0;
    uint32_t kb_per_blob = 0;  // items the 128 phase-B outputs of one blob add up

    // Working set of the synchronous entry points and its pinned staging.
    vkzg_amd::WorkSet ws;
    uint8_t *h_in = nullptr, *h_out = nullptr, *h_present = nullptr;
    uint32_t *h_err = nullptr;

    // Batch verification and commitments (created on first use).
    vkzg_amd::VerifyState *vfy = nullptr;

    // Per-stage timing of the last call (events on the stream of the first
    // sub-batch), accumulated over its chunks.
    bool timing = false;
    static constexpr int kMaxStages = 16;
    hipEvent_t ev[kMaxStages + 1] = {};
    float stage_ms[kMaxStages] = {};
    int nstages = 0;
    const char *stage_names[kMaxStages] = {};

    ~vkzg_prover();
};

namespace vkzg_amd {

// Stage boundaries of one sub-batch pipeline when timing is on.
struct StageClock {
    vkzg_prover *p;
    hipStream_t s;
    bool on;
    int n = 0;
    StageClock(vkzg_prover *pp, hipStream_t ss, bool enable) : p(pp), s(ss), on(enable && pp->timing) {
        if (on) ck(hipEventRecord(p->ev[0], s));
    }
    void mark(const char *name) {
        if (!on || n >= vkzg_prover::kMaxStages) return;
        p->stage_names[n] = name;
        ck(hipEventRecord(p->ev[++n], s));
    }
    // After the stream has been synchronised.
    void finish() {
        if (!on) return;
        p->nstages = n;
        for (int i = 0; i < n; i++) {
            float ms = 0;
            ck(hipEventElapsedTime(&ms, p->ev[i], p->ev[i + 1]));
            p->stage_ms[i] += ms;
        }
    }
};

// The tuning for a batch of nb blobs ($VKZG_AMD_SPLITS overrides its splits).
Tuning tuning_for(const vkzg_prover *p, uint32_t nb);

// The phase-A table a tuning names (window_bits 0: the main one).
PhaseATable phase_a_table(const vkzg_prover *p, uint32_t window_bits);

// Enqueue a pipeline through `cache` when graphs are on and stage timing is
// off (timing records events between stages), else launch it directly.
// A stream the caller is already capturing gets the launches themselves, so
// they land in the caller's graph.
template <typename Body>
void enqueue_maybe_graph(vkzg_prover *p, GraphCache &cache, const GraphKey &key, hipStream_t s,
                         const StageClock &clk, Body body) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    ck(hipStreamIsCapturing(s, &cs));
    if (p->use_graphs && !clk.on && cs == hipStreamCaptureStatusNone) cache.run(key, s, body);
    else body(s);
}

// Enqueue the whole proof / recovery pipeline for nb blobs on stream s, using
// working set ws (capacity >= nb).  The batch is split over sub-streams when
// that overlaps latency-bound stages with throughput-bound ones; the split is
// forked from and joined back into s.  `clk` times the first sub-batch.
// With `graphs` non-null the pipeline goes through that graph cache.
void enqueue_proofs(vkzg_prover *p, WorkSet &ws, StreamSet &ss, uint8_t *d_proofs,
                    const uint8_t *d_blobs, uint32_t nb, hipStream_t s, StageClock &clk,
                    const Tuning &tn, GraphCache *graphs);
void enqueue_recovery(vkzg_prover *p, WorkSet &ws, StreamSet &ss, uint8_t *d_out,
                      const uint8_t *d_cells, const uint8_t *d_present, uint32_t nb, hipStream_t s,
                      StageClock &clk, const Tuning &tn, GraphCache *graphs);

// The recovery pipeline of nb blobs on stream s with caller-provided scratch:
// A and B hold nb x 8192 Fr each.  Every kernel writes each scratch element
// before any kernel reads it, so the scratch needs no clearing.  The cell,
// mask and output buffers must be 16-byte aligned (vector accesses).
void enqueue_recovery_into(vkzg_prover *p, Fr *A, Fr *B, uint8_t *d_out, const uint8_t *d_cells,
                           const uint8_t *d_present, uint32_t nb, hipStream_t s, StageClock &clk,
                           uint32_t radix, uint32_t waves);

bool chunk_missing_ok(const uint8_t *present, uint32_t nb);

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

// Clears the per-stage times before a call (stage timing accumulates over chunks).
void reset_stage_times(vkzg_prover *p);

// Resolves, once per device and with that device current, the address of the
// file-scope scratch array of vkzg_amd_recover_cells_device_noalloc
// (noalloc.cpp).  Called at prover creation.
void noalloc_scratch_setup(int device);

}  // namespace vkzg_amd
