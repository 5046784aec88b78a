// Cell recovery on device buffers with no allocation in the call
// (vkzg_amd_recover_cells_device_noalloc, include/vkzg_amd.h).
//
// The recovery kernels need scratch beyond the caller's inputs and outputs:
// the two ping-pong transform buffers (8192 Fr each per blob).  Here that
// scratch is one file-scope device array, `scratch`, sized for
// kNoAllocMaxBlobs blobs;
// every GPU the library runs on holds its own copy until the process ends.
// Host code never uses the array's name as an address: the device address of
// each GPU's copy is resolved once per device by hipGetSymbolAddress (at
// prover creation, with that device current) and kept in `g_scratch`.
//
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU (two provers, or two host threads on different streams) race.  Each
// kernel writes every scratch element before reading it, so nothing needs
// clearing between calls.
#include <mutex>

#include "../../include/vkzg_amd.h"
#include "runtime.h"

using namespace vkzg_amd;

namespace {

constexpr uint32_t kNoAllocMaxBlobs = 16;
constexpr size_t kBufBytes = (size_t)kNoAllocMaxBlobs * 8192 * sizeof(Fr);  // A, B: 4 MiB each
constexpr size_t kScratchBytes = 2 * kBufBytes;                             // 8 MiB
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device `device`.
void *scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) throw HipFail{};
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        ck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw HipFail{};
        g_scratch[device] = ptr;
    }
    return g_scratch[device];
}

}  // namespace

namespace vkzg_amd {

void noalloc_scratch_setup(int device) { (void)scratch_ptr(device); }

}  // namespace vkzg_amd

extern "C" vkzg_result vkzg_amd_recover_cells_device_noalloc(vkzg_prover *p, void *d_cells_out,
                                                             const void *d_cells, const void *d_present,
                                                             const uint8_t *h_present, size_t num_blobs,
                                                             void *hip_stream) {
    if (!p || !d_cells_out || !d_cells || !d_present || !h_present) return VKZG_ERR_BADARGS;
    if (num_blobs == 0) return VKZG_OK;
    if (num_blobs > kNoAllocMaxBlobs) return VKZG_ERR_BADARGS;
    if (!chunk_missing_ok(h_present, (uint32_t)num_blobs)) return VKZG_ERR_BADARGS;
    if (!aligned16(d_cells_out) || !aligned16(d_cells) || !aligned16(d_present)) return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    reset_stage_times(p);
    hipStream_t s = hip_stream ? (hipStream_t)hip_stream : p->stream;
    try {
        uint8_t *base = static_cast<uint8_t *>(scratch_ptr(p->device));
        Fr *A = reinterpret_cast<Fr *>(base);
        Fr *B = reinterpret_cast<Fr *>(base + kBufBytes);
        const uint32_t nb = (uint32_t)num_blobs;
        StageClock clk(p, s, true);
        const Tuning tn = tuning_for(p, nb);
        enqueue_recovery_into(p, A, B, static_cast<uint8_t *>(d_cells_out),
                              static_cast<const uint8_t *>(d_cells), static_cast<const uint8_t *>(d_present),
                              nb, s, clk, tn.rec_radix, tn.rec_waves);
        ck(hipStreamSynchronize(s));
        clk.finish();
    } catch (HipFail &) {
        return VKZG_ERR_GPU;
    }
    return VKZG_OK;
}
