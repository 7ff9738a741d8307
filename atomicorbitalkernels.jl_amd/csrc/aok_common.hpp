// aok_common.hpp -- context, status plumbing and device helpers shared by the
// libaok.so translation units (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <deque>
#include <string>
#include <vector>

#include "../../include/aok.h"

namespace aok {
struct Ovl2cPlanState;
}
struct aok_graph_s;

struct aok_context {
    int device = 0;
    int num_cu = 0;  // compute units of the device (read on first use)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string last_error;
    // Grow-only device workspace (partials, bucket tables, primitive-pair tables).
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // Workspace of the most recently destroyed 2-centre plan, handed to the next one.
    void *plan_cache = nullptr;
    size_t plan_cache_bytes = 0;
    // Live plans of this context: destroying the context releases their workspaces and
    // detaches them (a detached plan only accepts aok_ovl2c_plan_destroy).
    std::vector<aok::Ovl2cPlanState *> plans;
    // Error word in host-mapped pinned memory: h_err (host view), d_err (device view).
    unsigned *d_err = nullptr;
    unsigned *h_err = nullptr;
    // Blocking (default): every entry point waits for its work and reports device-side
    // errors.  Stream-ordered: entry points return once the work is enqueued; device-side
    // error flags accumulate in the error word until aok_sync.
    bool blocking = true;
    // Optional timing of the dominant kernel of each call: an event pair per timed launch.
    // Pairs wait in `pend` until read (at the closing synchronisation of a blocking call,
    // or at aok_sync / aok_kernel_times / aok_last_kernel_ms); read times go to `done`.
    struct Timed {
        const char *name;
        hipEvent_t a, b;
    };
    bool timing = false;
    bool open = false;  // time_begin recorded pend.back().a, time_end not yet
    std::vector<Timed> pend;
    std::vector<hipEvent_t> evpool;
    std::deque<std::pair<const char *, double>> done;
    double last_ms = 0.0;
    const char *last_kernel = "";
    // HIP-graph capture (aok_graph_begin .. aok_graph_end, or any capture of this stream):
    // calls enqueue without synchronising, timing or clearing the error word, and may not
    // grow the workspace.  A call that fails while capturing marks the capture failed.
    bool capturing = false;
    bool capture_failed = false;
    // Graphs of this context still alive: device buffers they may reference (workspace,
    // plan buffers) are retired instead of freed until the last one is destroyed.
    std::vector<aok_graph_s *> graphs;
    std::vector<void *> retired;
    // Tuning knobs (0 = built-in default), set by aok_tune or the autotuner.
    struct {
        int ao_ppl = 0;   // AO values-only: consecutive points per lane (1, 2, 4)
        int ao_grid = 0;  // AO: persistent grid size (blocks)
        int ovl_cap = 0;  // 2c: LDS stripe capacity (floats per lane)
        int ovl_cap_pb = 0;  // 2c pullback / position cotangents only (0: ovl_cap)
        int ovl_split = 0;  // 2c pullback: light chunks in the l <= 2 instance (0/1 yes, 2 no)
        int ovl_order = 0;  // 2c workgroup order: 0 auto, 1 tile-major, 2 XCD-grouped
        int ovl3_parts = 0;  // 3c pullback parts: 0 auto, 1 none, 2 by C shell, 3 by (C, B) shells
    } tune;
};

namespace aok {

constexpr unsigned ERR_SPECIES = 1u;

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Status helpers ---------------------------------------------------------------
inline aok_status fail(aok_ctx ctx, aok_status st, const std::string &msg)
{
    if (ctx) {
        ctx->last_error = msg;
        if (ctx->capturing) ctx->capture_failed = true;
    }
    return st;
}

#define AOK_HIP(ctx, call)                                                                \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::aok::fail((ctx), AOK_ERR_DEVICE,                                     \
                               std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

// Ensure at least `bytes` of workspace; returns the base pointer or nullptr.
aok_status ensure_ws(aok_ctx ctx, size_t bytes);

// True while the context stream is being captured into a HIP graph.
bool in_capture(aok_ctx ctx);
// Give back a device buffer of the context: freed once the stream has drained, or kept
// until the last live graph is destroyed (a graph may reference it).
aok_status release_buffer(aok_ctx ctx, void *buf);

// Clear the error word (host store, before the call's first launch).
aok_status clear_err(aok_ctx ctx);
// Synchronise the stream, read the error word, translate into a status.
aok_status finish(aok_ctx ctx, const char *what);

// Event-bracketed launch of the dominant kernel (only when timing is enabled).
void time_begin(aok_ctx ctx, const char *name);
void time_end(aok_ctx ctx);
// Wait for every pending timed launch and move its time to ctx->done.
void collect_times(aok_ctx ctx);

// Device helpers ----------------------------------------------------------------
// 1-based species index (int64 or int32) -> 0-based sigma; NULL input means species 1.
__device__ __forceinline__ int load_sigma(const void *sidx, int itype, int64_t i)
{
    if (sidx == nullptr) return 0;
    if (itype == AOK_I32) return static_cast<const int32_t *>(sidx)[i] - 1;
    return static_cast<int>(static_cast<const int64_t *>(sidx)[i] - 1);
}

template <typename T>
__device__ __forceinline__ void store_nt(T *p, T v)
{
    __builtin_nontemporal_store(v, p);
}

}  // namespace aok
