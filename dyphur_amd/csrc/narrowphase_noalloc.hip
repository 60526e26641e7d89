// narrowphase_noalloc.hip — the narrowphase pass for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Same kernel and results as dy_narrowphase_run_device_count;
// the contact store and counts are the caller's, and the pass's only scratch
// (one 8-byte word per 256-pair tile) is a fixed device array of this file.
#include <mutex>

#include "narrowphase_pass.hpp"

namespace {

// 65 536 tile words (512 KiB): at most 16 777 216 pairs per call.  Each GPU the
// library runs on holds its own copy until the process ends.  Zero at load,
// which no run's epoch takes.
constexpr uint32_t kMaxTiles = 1u << 16;
static __device__ __align__(256) uint64_t scratch[kMaxTiles];

struct DeviceScratch {
    uint64_t *ptr = nullptr;     // this device's address of `scratch`
    uint32_t epoch = 0;          // the last call's epoch on this device
};
// Fixed per-device table (no host allocation either).
constexpr int kMaxDevices = 64;
std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

// The current device's scratch address (hipGetSymbolAddress on its first call
// there: in host code `scratch` names a host-side stand-in, not the device
// array) and a fresh epoch for this call.  wrapped: the epochs have run out
// and the array must be cleared before use.
int acquire(uint64_t **ptr, uint32_t *epoch, bool *wrapped) {
    int dev = 0;
    DY_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return DY_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceScratch &d = g_dev[dev];
    if (!d.ptr) {
        void *p = nullptr;
        DY_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        d.ptr = static_cast<uint64_t *>(p);
    }
    *wrapped = ++d.epoch == 0;
    if (*wrapped) d.epoch = 1;
    *ptr = d.ptr;
    *epoch = d.epoch;
    return DY_OK;
}

}  // namespace

extern "C" int dy_narrowphase_run_noalloc(dy_stream_t s, const dy_contact_pair *d_pairs, const uint32_t *d_n_pairs,
                                          uint32_t max_pairs, const dy_body_view *bodies, const dy_shape_view *shapes,
                                          const dy_contact_view *store, uint32_t *d_counts, uint32_t *d_features) {
    if (!bodies || !shapes || !store || !d_counts || (max_pairs && !d_pairs)) return DY_ERR_INVALID_ARG;
    if (store->capacity && (!store->body_a || !store->body_b || !store->pos_x || !store->pos_y || !store->pos_z ||
                            !store->norm_x || !store->norm_y || !store->norm_z || !store->depth))
        return DY_ERR_INVALID_ARG;
    if (dy::ceil_div(max_pairs, dy::kNpTilePairs) > kMaxTiles) return DY_ERR_INVALID_ARG;
    uint64_t *sp = nullptr;
    uint32_t epoch = 0;
    bool wrapped = false;
    DY_TRY_STATUS(acquire(&sp, &epoch, &wrapped));
    hipStream_t st = dy::hs(s);
    if (wrapped) DY_TRY(hipMemsetAsync(sp, 0, sizeof(uint64_t) * kMaxTiles, st));
    dy_contact_view cv = *store;
    cv.n = d_counts;
    return dy::np_enqueue(st, d_pairs, max_pairs, d_n_pairs, *bodies, *shapes, sp, epoch,
                          dy::NpOut{cv, d_counts, d_features});
}
