// noalloc.hip — the batched edge check without device allocation (mrcc_check_edges_noalloc).
//
// The queue of mrcc_check_edges needs a workspace (per-edge records, the sorted order, the pass
// tables, the interpolation steps, counters).  mrcc_check_edges allocates it per call; this entry
// point uses one file-scope device array instead, `scratch`, for callers that cannot allocate:
// code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  Every GPU the
// library runs on holds its own copy of the array until the process ends.  Host code reaches a
// device's copy only through the address hipGetSymbolAddress returns with that device current (the
// array's name in host code is a host-side stand-in, not a device address); that address is looked
// up once per device and kept.
//
// Every call on one GPU shares that GPU's array and leaves its contents behind: the queue build
// clears what must start at zero (histogram, counters, tickets) with a fill kernel, so calls on one
// stream follow each other safely, and calls that overlap on one GPU (other streams, other threads)
// race.
#include <mutex>

#include "engine.hpp"

// 16 MiB: per edge 16 + 4 + 4 + 8 D bytes (record, order, long-edge list, interpolation step),
// plus about 264 KiB of pass tables and counters: about 65,000 edges of 28 coordinates.
#define MRCC_NOALLOC_SCRATCH_BYTES (16u << 20)
#define MRCC_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[MRCC_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_dev[MRCC_NOALLOC_MAX_DEVICES];  // device address of `scratch` on each GPU

// This device's copy of `scratch` (looked up with the device current, once per device).
int scratch_address(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return mrcc_internal::fail(MRCC_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= MRCC_NOALLOC_MAX_DEVICES)
        return mrcc_internal::fail(MRCC_ERR_UNSUPPORTED, "device %d: the scratch table holds %d devices", dev,
                                   MRCC_NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return mrcc_internal::fail(MRCC_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return MRCC_OK;
}

__global__ void k_store_count(const long long* src, int64_t* dst) {
    if (threadIdx.x == 0) *dst = src ? (int64_t)*src : 0;
}

}  // namespace

// Called by mrcc_init with the device current: looks the scratch array up at set-up.
int mrcc_noalloc_setup() {
    void* p = nullptr;
    return scratch_address(&p);
}

extern "C" int mrcc_check_edges_noalloc(const mrcc_scene* scene, const mrcc_edge_batch* batch,
                                        const mrcc_tune* tune, int64_t* n_invalid, void* stream) {
    int rc = mrcc_internal::injected();
    if (rc != MRCC_OK) return rc;
    if (!scene) return mrcc_internal::fail(MRCC_ERR_INVALID, "scene is null");
    const SceneRec* sc = mrcc_internal::scene_rec(scene);
    rc = mrcc_internal::check_batch(sc, batch);
    if (rc != MRCC_OK) return rc;
    mrcc_tune t;
    rc = mrcc_internal::tune_defaults(sc, tune, &t);
    if (rc != MRCC_OK) return rc;
    if (!t.persistent || t.collect_stats || t.time_kernels)
        return mrcc_internal::fail(MRCC_ERR_UNSUPPORTED,
                                   "the no-allocation check runs the persistent queue without statistics or timing");
    const size_t need = mrcc_internal::plan_layout(batch->n_edges, sc->D, nullptr, nullptr);
    if (need > MRCC_NOALLOC_SCRATCH_BYTES)
        return mrcc_internal::fail(MRCC_ERR_INVALID,
                                   "%lld edges of %d coordinates need %zu bytes of workspace; the no-allocation "
                                   "check holds %u (about %lld edges of this size)",
                                   (long long)batch->n_edges, sc->D, need, MRCC_NOALLOC_SCRATCH_BYTES,
                                   (long long)((MRCC_NOALLOC_SCRATCH_BYTES - (300u << 10)) / (24 + 8 * sc->D)));
    void* base = nullptr;
    rc = scratch_address(&base);
    if (rc != MRCC_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    PlanWork W;
    mrcc_internal::plan_layout(batch->n_edges, sc->D, (unsigned char*)base, &W);
    W.stats = nullptr;
    if (batch->n_edges > 0) {
        const EdgeBatchView B = mrcc_internal::batch_view(batch);
        mrcc_internal::launch_plan_sequence(*sc, B, W, __builtin_ctz((unsigned)t.slots_per_pass), t.slots_per_ticket,
                                            mrcc_internal::plan_grid(sc, t), st);
    }
    if (n_invalid) k_store_count<<<1, 64, 0, st>>>(batch->n_edges > 0 ? W.meta + 2 : nullptr, n_invalid);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mrcc_internal::fail(MRCC_ERR_HIP, "launch: %s", hipGetErrorString(e));
    return MRCC_OK;
}
