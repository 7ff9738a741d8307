// api.hip -- extern "C" entry points of libaok.so (declared in include/aok.h).
#include <cstring>
#include <new>

#include "aok_common.hpp"
#include "ao_plan.hpp"
#include "ovl_plan.hpp"

namespace aok {

bool in_capture(aok_ctx ctx)
{
    if (ctx->capturing) return true;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(ctx->stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive;
}

aok_status release_buffer(aok_ctx ctx, void *buf)
{
    if (!buf) return AOK_OK;
    if (!ctx->graphs.empty() || in_capture(ctx)) {
        ctx->retired.push_back(buf);
        return AOK_OK;
    }
    AOK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    AOK_HIP(ctx, hipFree(buf));
    return AOK_OK;
}

aok_status ensure_ws(aok_ctx ctx, size_t bytes)
{
    if (bytes <= ctx->ws_bytes) return AOK_OK;
    if (in_capture(ctx))
        return fail(ctx, AOK_ERR_ARG,
                    "a workspace of " + std::to_string(bytes) +
                        " bytes is needed inside a graph capture: run the same calls once before capturing");
    size_t want = std::max(bytes, ctx->ws_bytes + ctx->ws_bytes / 2);
    want = (want + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
    if (ctx->ws) {
        void *old = ctx->ws;
        ctx->ws = nullptr;
        ctx->ws_bytes = 0;
        aok_status st = release_buffer(ctx, old);
        if (st != AOK_OK) return st;
    }
    if (hipMalloc(&ctx->ws, want) != hipSuccess) {
        ctx->ws = nullptr;
        return fail(ctx, AOK_ERR_NOMEM, "workspace allocation of " + std::to_string(want) + " bytes failed");
    }
    ctx->ws_bytes = want;
    return AOK_OK;
}

// The error word lives in host-mapped pinned memory: kernels store flags into it (plain
// stores, the same value from every writer) and the host reads it after a
// synchronisation.  Blocking calls clear it before their first launch and read it after
// their closing synchronisation; in stream-ordered mode it accumulates the flags of
// every call until aok_sync reads and clears it.
aok_status clear_err(aok_ctx ctx)
{
    if (ctx->blocking && !in_capture(ctx)) {
        *reinterpret_cast<volatile unsigned *>(ctx->h_err) = 0u;
        // a blocking call reports its own timed kernel (none: 0)
        ctx->last_ms = 0.0;
        ctx->last_kernel = "";
    }
    return AOK_OK;
}

static hipEvent_t take_event(aok_ctx ctx)
{
    if (!ctx->evpool.empty()) {
        hipEvent_t e = ctx->evpool.back();
        ctx->evpool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}

void time_begin(aok_ctx ctx, const char *name)
{
    ctx->open = false;
    if (!ctx->timing || in_capture(ctx)) return;
    hipEvent_t a = take_event(ctx), b = a ? take_event(ctx) : nullptr;
    if (!a || !b || hipEventRecord(a, ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        if (a) ctx->evpool.push_back(a);
        if (b) ctx->evpool.push_back(b);
        return;  // this launch goes untimed
    }
    ctx->pend.push_back({name, a, b});
    ctx->open = true;
}

void time_end(aok_ctx ctx)
{
    if (!ctx->open) return;
    ctx->open = false;
    if (hipEventRecord(ctx->pend.back().b, ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        ctx->evpool.push_back(ctx->pend.back().a);
        ctx->evpool.push_back(ctx->pend.back().b);
        ctx->pend.pop_back();
    }
}

void collect_times(aok_ctx ctx)
{
    constexpr size_t DONE_CAP = 1 << 16;
    for (const auto &t : ctx->pend) {
        float ms = 0.f;
        if (hipEventSynchronize(t.b) == hipSuccess && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            ctx->last_ms = ms;
            ctx->last_kernel = t.name;
            ctx->done.emplace_back(t.name, static_cast<double>(ms));
            if (ctx->done.size() > DONE_CAP) ctx->done.pop_front();
        } else {
            (void)hipGetLastError();  // do not leave a timing failure for the caller's next HIP call
        }
        ctx->evpool.push_back(t.a);
        ctx->evpool.push_back(t.b);
    }
    ctx->pend.clear();
}

aok_status finish(aok_ctx ctx, const char *what)
{
    // stream-ordered: return with the work enqueued (timed launches are read later);
    // under graph capture the work is only recorded
    if (!ctx->blocking || in_capture(ctx)) return AOK_OK;
    AOK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    collect_times(ctx);
    if (*reinterpret_cast<volatile unsigned *>(ctx->h_err) & ERR_SPECIES)
        return fail(ctx, AOK_ERR_SPECIES, std::string(what) + ": species index outside 1..NZ");
    return AOK_OK;
}

static bool bad_dtype(aok_dtype d) { return d != AOK_F32 && d != AOK_F64; }
static bool bad_itype(aok_itype t) { return t != AOK_I64 && t != AOK_I32; }

}  // namespace aok

// A recorded graph of one context.  Destroying the context first detaches it: the
// executable is released and the handle then only accepts aok_graph_destroy.
struct aok_graph_s {
    aok_ctx ctx;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

static void detach_graph(aok_graph_s *g)
{
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->exec = nullptr, g->graph = nullptr, g->ctx = nullptr;
}

using namespace aok;

#define AOK_CHECK_CTX(ctx) \
    if (!(ctx)) return AOK_ERR_ARG
#define AOK_REQUIRE(ctx, cond, msg) \
    if (!(cond)) return ::aok::fail((ctx), AOK_ERR_ARG, (msg))

extern "C" {

aok_status aok_create(aok_ctx *out, int device, void *stream)
{
    if (!out) return AOK_ERR_ARG;
    *out = nullptr;
    aok_context *c = new (std::nothrow) aok_context();
    if (!c) return AOK_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return AOK_ERR_DEVICE;
    }
    if (stream) {
        c->stream = static_cast<hipStream_t>(stream);
    } else {
        // a blocking stream: ordered after work the caller queued on the legacy default
        // stream (inputs produced there are complete before the library reads them)
        if (hipStreamCreate(&c->stream) != hipSuccess) {
            delete c;
            return AOK_ERR_DEVICE;
        }
        c->own_stream = true;
    }
    if (hipHostMalloc(&c->h_err, sizeof(unsigned), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void **>(&c->d_err), c->h_err, 0) != hipSuccess) {
        aok_destroy(c);
        return AOK_ERR_DEVICE;
    }
    *reinterpret_cast<volatile unsigned *>(c->h_err) = 0u;  // pinned memory comes uncleared
    // this device's copy of the allocation-free entry point's scratch array; a failure
    // here is reported by aok_ovl2c_noalloc, which repeats the lookup
    (void)ovl2c_noalloc_setup(c);
    c->last_error.clear();
    *out = c;
    return AOK_OK;
}

aok_status aok_destroy(aok_ctx ctx)
{
    if (!ctx) return AOK_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ws) (void)hipFree(ctx->ws);
    if (ctx->plan_cache) (void)hipFree(ctx->plan_cache);
    for (Ovl2cPlanState *p : ctx->plans) {
        if (p->buf) (void)hipFree(p->buf);
        p->buf = nullptr, p->bytes = 0, p->ctx = nullptr;
    }
    for (aok_graph_s *g : ctx->graphs) detach_graph(g);
    for (void *r : ctx->retired) (void)hipFree(r);
    if (ctx->h_err) (void)hipHostFree(ctx->h_err);
    for (const auto &t : ctx->pend) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (hipEvent_t e : ctx->evpool) (void)hipEventDestroy(e);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return AOK_OK;
}

const char *aok_last_error(aok_ctx ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

void *aok_stream(aok_ctx ctx) { return ctx ? static_cast<void *>(ctx->stream) : nullptr; }

const char *aok_version(void) { return "aok-mi355x 0.1.0 (gfx950)"; }

double aok_last_kernel_ms(aok_ctx ctx)
{
    if (!ctx) return 0.0;
    collect_times(ctx);
    return ctx->last_ms;
}

const char *aok_last_kernel_name(aok_ctx ctx)
{
    if (!ctx) return "";
    collect_times(ctx);
    return ctx->last_kernel;
}

aok_status aok_kernel_times(aok_ctx ctx, int64_t cap, double *ms, const char **names, int64_t *count)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, cap >= 0 && count != nullptr && (cap == 0 || (ms != nullptr && names != nullptr)),
                "aok_kernel_times: bad arguments");
    collect_times(ctx);
    int64_t n = 0;
    while (n < cap && !ctx->done.empty()) {
        names[n] = ctx->done.front().first;
        ms[n] = ctx->done.front().second;
        ctx->done.pop_front();
        ++n;
    }
    *count = n;
    return AOK_OK;
}

aok_status aok_tune(aok_ctx ctx, const char *key, int64_t value)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, key != nullptr, "key is NULL");
    const std::string k(key);
    if (k == "ao_ppl") {
        AOK_REQUIRE(ctx, value == 0 || value == 1 || value == 2 || value == 4, "ao_ppl must be 0, 1, 2 or 4");
        ctx->tune.ao_ppl = static_cast<int>(value);
    } else if (k == "ao_grid") {
        AOK_REQUIRE(ctx, value >= 0 && value <= (1 << 20), "ao_grid out of range");
        ctx->tune.ao_grid = static_cast<int>(value);
    } else if (k == "ovl_cap") {
        AOK_REQUIRE(ctx, value == 0 || (value >= 16 && value <= 1024), "ovl_cap must be 0 or in 16..1024");
        ctx->tune.ovl_cap = static_cast<int>(value);
    } else if (k == "ovl_cap_pb") {
        AOK_REQUIRE(ctx, value == 0 || (value >= 16 && value <= 1024), "ovl_cap_pb must be 0 or in 16..1024");
        ctx->tune.ovl_cap_pb = static_cast<int>(value);
    } else if (k == "ovl_order") {
        AOK_REQUIRE(ctx, value >= 0 && value <= OVL_MAXCH + 1, "ovl_order must be in 0..161");
        ctx->tune.ovl_order = static_cast<int>(value);
    } else if (k == "ovl_split") {
        AOK_REQUIRE(ctx, value >= 0 && value <= 2, "ovl_split must be 0, 1 or 2");
        ctx->tune.ovl_split = static_cast<int>(value);
    } else if (k == "ovl3_parts") {
        AOK_REQUIRE(ctx, value >= 0 && value <= 3, "ovl3_parts must be 0, 1, 2 or 3");
        ctx->tune.ovl3_parts = static_cast<int>(value);
    } else {
        return fail(ctx, AOK_ERR_ARG, "unknown tuning key '" + k + "'");
    }
    return AOK_OK;
}

aok_status aok_tune_get(aok_ctx ctx, const char *key, int64_t *value)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, key != nullptr && value != nullptr, "NULL argument");
    const std::string k(key);
    if (k == "ao_ppl") *value = ctx->tune.ao_ppl;
    else if (k == "ao_grid") *value = ctx->tune.ao_grid;
    else if (k == "ovl_cap") *value = ctx->tune.ovl_cap;
    else if (k == "ovl_cap_pb") *value = ctx->tune.ovl_cap_pb;
    else if (k == "ovl_split") *value = ctx->tune.ovl_split;
    else if (k == "ovl_order") *value = ctx->tune.ovl_order;
    else if (k == "ovl3_parts") *value = ctx->tune.ovl3_parts;
    else return fail(ctx, AOK_ERR_ARG, "unknown tuning key '" + k + "'");
    return AOK_OK;
}

aok_status aok_set_timing(aok_ctx ctx, int enable)
{
    AOK_CHECK_CTX(ctx);
    ctx->timing = enable != 0;
    return AOK_OK;
}

aok_status aok_sync(aok_ctx ctx)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, !in_capture(ctx), "aok_sync inside a graph capture");
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    AOK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    collect_times(ctx);
    volatile unsigned *e = reinterpret_cast<volatile unsigned *>(ctx->h_err);
    const unsigned flags = *e;
    *e = 0u;
    if (flags & ERR_SPECIES) return fail(ctx, AOK_ERR_SPECIES, "species index outside 1..NZ (an earlier call)");
    return AOK_OK;
}

aok_status aok_set_blocking(aok_ctx ctx, int blocking)
{
    AOK_CHECK_CTX(ctx);
    // leaving stream-ordered mode reports what the enqueued work flagged; entering it
    // starts from a clear error word (blocking calls have reported their own flags)
    const aok_status st = ctx->blocking ? AOK_OK : aok_sync(ctx);
    if (ctx->blocking && !blocking) *reinterpret_cast<volatile unsigned *>(ctx->h_err) = 0u;
    ctx->blocking = blocking != 0;
    return st;
}

int aok_get_blocking(aok_ctx ctx) { return ctx ? (ctx->blocking ? 1 : 0) : -1; }

// ---- H1 --------------------------------------------------------------------------------
static aok_status ao_common(aok_ctx ctx, aok_dtype dtype, const void *X, aok_itype itype, int64_t nX,
                            const void *zeta, const void *D, int64_t nRad, int64_t K, int64_t NZ,
                            const int64_t *poly, const int64_t *iR, const int64_t *iY, int64_t nOrb, int L,
                            aok_decay decay, AoRunTable &runs)
{
    AOK_REQUIRE(ctx, !bad_dtype(dtype), "dtype must be AOK_F32 or AOK_F64");
    AOK_REQUIRE(ctx, !bad_itype(itype), "itype must be AOK_I64 or AOK_I32");
    AOK_REQUIRE(ctx, nX >= 0, "nX must be >= 0");
    AOK_REQUIRE(ctx, nX == 0 || X != nullptr, "X is NULL");
    AOK_REQUIRE(ctx, zeta && D, "zeta/D are NULL");
    AOK_REQUIRE(ctx, nRad >= 1 && K >= 1 && NZ >= 1, "nRad, K, NZ must be >= 1");
    AOK_REQUIRE(ctx, NZ * nRad * K <= 4096, "parameter table too large (NZ*nRad*K > 4096)");
    AOK_REQUIRE(ctx, nOrb >= 1 && iR && iY, "orbital maps iR/iY required");
    AOK_REQUIRE(ctx, decay == AOK_DECAY_GAUSSIAN || decay == AOK_DECAY_SLATER, "unknown decay");
    if (L < 0 || L > AOK_AO_LMAX)
        return fail(ctx, AOK_ERR_LMAX, "L = " + std::to_string(L) + " outside 0.." + std::to_string(AOK_AO_LMAX));
    const std::string msg = build_runs(poly, iR, iY, nOrb, nRad, L, runs);
    if (!msg.empty()) return fail(ctx, AOK_ERR_ARG, msg);
    return AOK_OK;
}

aok_status aok_ao_eval(aok_ctx ctx, aok_dtype dtype, const void *X, const void *sidx, aok_itype itype, int64_t nX,
                       const void *zeta, const void *D, int64_t nRad, int64_t K, int64_t NZ, const int64_t *poly,
                       const int64_t *iR, const int64_t *iY, int64_t nOrb, int L, double lengthscale,
                       aok_decay decay, void *P, void *dP)
{
    AOK_CHECK_CTX(ctx);
    AoRunTable runs;
    aok_status st = ao_common(ctx, dtype, X, itype, nX, zeta, D, nRad, K, NZ, poly, iR, iY, nOrb, L, decay, runs);
    if (st != AOK_OK) return st;
    AOK_REQUIRE(ctx, nX == 0 || P != nullptr, "P is NULL");
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    if ((st = clear_err(ctx)) != AOK_OK) return st;
    AoArgs a{X, sidx, itype, nX, zeta, D, (int)nRad, (int)K, (int)NZ, L, lengthscale, decay, P, dP, nullptr};
    st = dtype == AOK_F32 ? ao_eval_launch<float>(ctx, a, runs) : ao_eval_launch<double>(ctx, a, runs);
    if (st != AOK_OK) return st;
    return finish(ctx, "aok_ao_eval");
}

aok_status aok_ao_pullback(aok_ctx ctx, aok_dtype dtype, const void *X, const void *sidx, aok_itype itype,
                           int64_t nX, const void *zeta, const void *D, int64_t nRad, int64_t K, int64_t NZ,
                           const int64_t *poly, const int64_t *iR, const int64_t *iY, int64_t nOrb, int L,
                           double lengthscale, aok_decay decay, const void *dPin, void *dX, void *dzeta, void *dD)
{
    AOK_CHECK_CTX(ctx);
    AoRunTable runs;
    aok_status st = ao_common(ctx, dtype, X, itype, nX, zeta, D, nRad, K, NZ, poly, iR, iY, nOrb, L, decay, runs);
    if (st != AOK_OK) return st;
    AOK_REQUIRE(ctx, nX == 0 || dPin != nullptr, "dPin is NULL");
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    if ((st = clear_err(ctx)) != AOK_OK) return st;
    AoArgs a{X, sidx, itype, nX, zeta, D, (int)nRad, (int)K, (int)NZ, L, lengthscale, decay, nullptr, nullptr, dPin};
    st = dtype == AOK_F32 ? ao_pullback_launch<float>(ctx, a, runs, dX, dzeta, dD)
                          : ao_pullback_launch<double>(ctx, a, runs, dX, dzeta, dD);
    if (st != AOK_OK) return st;
    return finish(ctx, "aok_ao_pullback");
}

// ---- H2 normalisation -------------------------------------------------------------------
aok_status aok_compile_coef(aok_ctx ctx, aok_dtype dtype, const void *zeta, const void *D, int64_t ns, int64_t K,
                            int64_t NZ, const int64_t *ls, void *coef)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, !bad_dtype(dtype), "dtype must be AOK_F32 or AOK_F64");
    AOK_REQUIRE(ctx, zeta && D && coef && ls, "NULL argument");
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    aok_status st = clear_err(ctx);
    if (st != AOK_OK) return st;
    st = dtype == AOK_F32 ? compile_coef_run<float>(ctx, zeta, D, ns, K, NZ, ls, coef)
                          : compile_coef_run<double>(ctx, zeta, D, ns, K, NZ, ls, coef);
    if (st != AOK_OK) return st;
    return finish(ctx, "aok_compile_coef");
}

aok_status aok_compile_coef_pb(aok_ctx ctx, aok_dtype dtype, const void *dcoef, const void *zeta, const void *D,
                               int64_t ns, int64_t K, int64_t NZ, const int64_t *ls, void *dD, void *dzeta_norm)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, !bad_dtype(dtype), "dtype must be AOK_F32 or AOK_F64");
    AOK_REQUIRE(ctx, dcoef && zeta && D && ls && dD && dzeta_norm, "NULL argument");
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    aok_status st = clear_err(ctx);
    if (st != AOK_OK) return st;
    st = dtype == AOK_F32 ? compile_coef_pb_run<float>(ctx, dcoef, zeta, D, ns, K, NZ, ls, dD, dzeta_norm)
                          : compile_coef_pb_run<double>(ctx, dcoef, zeta, D, ns, K, NZ, ls, dD, dzeta_norm);
    if (st != AOK_OK) return st;
    return finish(ctx, "aok_compile_coef_pb");
}

// ---- H2 2-centre ------------------------------------------------------------------------
// op: 0 forward, 1 parameter pullback, 2 position cotangents
static aok_status ovl2c_entry(aok_ctx ctx, aok_dtype dtype, Ovl2cCall &c, int op, const char *what)
{
    // op: 0 forward, 1 (coef, alpha) pullback, 2 position cotangents, 3 allocation-free forward
    const bool pullback = op == 1 || op == 2;
    AOK_REQUIRE(ctx, !bad_dtype(dtype), "dtype must be AOK_F32 or AOK_F64");
    AOK_REQUIRE(ctx, !bad_itype(static_cast<aok_itype>(c.itype)), "itype must be AOK_I64 or AOK_I32");
    AOK_REQUIRE(ctx, (c.flags & ~AOK_PLAIN) == 0u, "unknown flags");
    AOK_REQUIRE(ctx, c.ls != nullptr, "ls is NULL");
    AOK_REQUIRE(ctx, c.B >= 0, "B must be >= 0");
    AOK_REQUIRE(ctx, c.B == 0 || (c.posA && c.posB), "positions are NULL");
    AOK_REQUIRE(ctx, c.alpha != nullptr, "exponents are NULL");
    AOK_REQUIRE(ctx, c.mode == 0 ? c.coef != nullptr : c.D != nullptr, "coefficients are NULL");
    if (!pullback) AOK_REQUIRE(ctx, c.B == 0 || c.out != nullptr, "out is NULL");
    if (pullback) AOK_REQUIRE(ctx, c.B == 0 || c.dS != nullptr, "dS is NULL");
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    aok_status st = clear_err(ctx);
    if (st != AOK_OK) return st;
    if (op == 2)
        st = dtype == AOK_F32 ? ovl2c_dx_run<float>(ctx, c) : ovl2c_dx_run<double>(ctx, c);
    else if (op == 3)
        st = dtype == AOK_F32 ? ovl2c_run_noalloc<float>(ctx, c) : ovl2c_run_noalloc<double>(ctx, c);
    else
        st = dtype == AOK_F32 ? ovl2c_run<float>(ctx, c, pullback) : ovl2c_run<double>(ctx, c, pullback);
    if (st != AOK_OK) return st;
    return finish(ctx, what);
}

aok_status aok_ovl2c(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *posA, const void *posB,
                     const void *sidxA, const void *sidxB, aok_itype itype, int64_t B, int64_t ns, int64_t K,
                     int64_t NZ, const int64_t *ls, const void *coef, const void *alpha, double lengthscale,
                     void *out)
{
    AOK_CHECK_CTX(ctx);
    Ovl2cCall c{};
    c.flags = flags;
    c.posA = posA, c.posB = posB, c.sidxA = sidxA, c.sidxB = sidxB, c.itype = itype, c.B = B;
    c.ls = ls, c.ns = ns, c.K = K, c.NZ = NZ, c.coef = coef, c.alpha = alpha, c.mode = 0, c.scale = lengthscale;
    c.out = out;
    return ovl2c_entry(ctx, dtype, c, 0, "aok_ovl2c");
}

aok_status aok_ovl2c_noalloc(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *posA, const void *posB,
                             const void *sidxA, const void *sidxB, aok_itype itype, int64_t B, int64_t ns,
                             int64_t K, int64_t NZ, const int64_t *ls, const void *coef, const void *alpha,
                             double lengthscale, void *out)
{
    AOK_CHECK_CTX(ctx);
    Ovl2cCall c{};
    c.flags = flags;
    c.posA = posA, c.posB = posB, c.sidxA = sidxA, c.sidxB = sidxB, c.itype = itype, c.B = B;
    c.ls = ls, c.ns = ns, c.K = K, c.NZ = NZ, c.coef = coef, c.alpha = alpha, c.mode = 0, c.scale = lengthscale;
    c.out = out;
    return ovl2c_entry(ctx, dtype, c, 3, "aok_ovl2c_noalloc");
}

aok_status aok_ovl2c_ps(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *posA, const void *posB,
                        const void *sidxA, const void *sidxB, aok_itype itype, int64_t B, int64_t ns, int64_t K,
                        int64_t NZ, const int64_t *ls, const void *zeta, const void *D, double lengthscale,
                        void *out)
{
    AOK_CHECK_CTX(ctx);
    Ovl2cCall c{};
    c.flags = flags;
    c.posA = posA, c.posB = posB, c.sidxA = sidxA, c.sidxB = sidxB, c.itype = itype, c.B = B;
    c.ls = ls, c.ns = ns, c.K = K, c.NZ = NZ, c.alpha = zeta, c.D = D, c.mode = 1, c.scale = lengthscale;
    c.out = out;
    return ovl2c_entry(ctx, dtype, c, 0, "aok_ovl2c_ps");
}

aok_status aok_ovl2c_pullback(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *dS, const void *posA,
                              const void *posB, const void *sidxA, const void *sidxB, aok_itype itype, int64_t B,
                              int64_t ns, int64_t K, int64_t NZ, const int64_t *ls, const void *coef,
                              const void *alpha, double lengthscale, void *dcoef, void *dalpha)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, dcoef && dalpha, "dcoef/dalpha are NULL");
    Ovl2cCall c{};
    c.flags = flags;
    c.posA = posA, c.posB = posB, c.sidxA = sidxA, c.sidxB = sidxB, c.itype = itype, c.B = B;
    c.ls = ls, c.ns = ns, c.K = K, c.NZ = NZ, c.coef = coef, c.alpha = alpha, c.mode = 0, c.scale = lengthscale;
    c.dS = dS, c.o1 = dcoef, c.o2 = dalpha, c.vjp = false;
    return ovl2c_entry(ctx, dtype, c, 1, "aok_ovl2c_pullback");
}

aok_status aok_ovl2c_pullback_x(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *dS, const void *posA,
                                const void *posB, const void *sidxA, const void *sidxB, aok_itype itype, int64_t B,
                                int64_t ns, int64_t K, int64_t NZ, const int64_t *ls, const void *coef,
                                const void *alpha, double lengthscale, void *dA, void *dB)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, B == 0 || dA != nullptr, "dA is NULL");
    Ovl2cCall c{};
    c.flags = flags;
    c.posA = posA, c.posB = posB, c.sidxA = sidxA, c.sidxB = sidxB, c.itype = itype, c.B = B;
    c.ls = ls, c.ns = ns, c.K = K, c.NZ = NZ, c.coef = coef, c.alpha = alpha, c.mode = 0, c.scale = lengthscale;
    c.dS = dS, c.dA = dA, c.dB = dB;
    return ovl2c_entry(ctx, dtype, c, 2, "aok_ovl2c_pullback_x");
}

aok_status aok_ovl2c_pullback_ps(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *dS, const void *posA,
                                 const void *posB, const void *sidxA, const void *sidxB, aok_itype itype, int64_t B,
                                 int64_t ns, int64_t K, int64_t NZ, const int64_t *ls, const void *zeta,
                                 const void *D, double lengthscale, void *dzeta, void *dD)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, dzeta && dD, "dzeta/dD are NULL");
    Ovl2cCall c{};
    c.flags = flags;
    c.posA = posA, c.posB = posB, c.sidxA = sidxA, c.sidxB = sidxB, c.itype = itype, c.B = B;
    c.ls = ls, c.ns = ns, c.K = K, c.NZ = NZ, c.alpha = zeta, c.D = D, c.mode = 1, c.scale = lengthscale;
    c.dS = dS, c.o1 = dzeta, c.o2 = dD, c.vjp = true;
    return ovl2c_entry(ctx, dtype, c, 1, "aok_ovl2c_pullback_ps");
}

// ---- H2 2-centre plans -------------------------------------------------------------------
struct aok_ovl2c_plan_s {
    aok::Ovl2cPlanState st;
};

aok_status aok_ovl2c_plan_create(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *sidxA, const void *sidxB,
                                 aok_itype itype, int64_t B, int64_t ns, int64_t K, int64_t NZ, const int64_t *ls,
                                 int params, const void *p1, const void *p2, double lengthscale,
                                 aok_ovl2c_plan *plan)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, plan != nullptr, "plan is NULL");
    *plan = nullptr;
    AOK_REQUIRE(ctx, params == AOK_PARAMS_COEF || params == AOK_PARAMS_PS, "params must be AOK_PARAMS_COEF or _PS");
    AOK_REQUIRE(ctx, !bad_dtype(dtype), "dtype must be AOK_F32 or AOK_F64");
    AOK_REQUIRE(ctx, !bad_itype(itype), "itype must be AOK_I64 or AOK_I32");
    AOK_REQUIRE(ctx, (flags & ~AOK_PLAIN) == 0u, "unknown flags");
    AOK_REQUIRE(ctx, ls != nullptr, "ls is NULL");
    AOK_REQUIRE(ctx, ns >= 1 && ns <= OVL_MAXSH, "ns out of range");
    AOK_REQUIRE(ctx, B >= 0, "B must be >= 0");
    AOK_REQUIRE(ctx, p1 && p2, "parameters are NULL");
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    auto *pl = new (std::nothrow) aok_ovl2c_plan_s;
    if (!pl) return fail(ctx, AOK_ERR_NOMEM, "plan allocation failed");
    Ovl2cPlanState &p = pl->st;
    p.ctx = ctx;
    p.dtype = dtype;
    for (int64_t s = 0; s < ns; ++s) p.ls[s] = ls[s];
    Ovl2cCall c{};
    c.flags = flags;
    c.sidxA = sidxA, c.sidxB = sidxB, c.itype = itype, c.B = B;
    c.ls = p.ls, c.ns = ns, c.K = K, c.NZ = NZ, c.scale = lengthscale;
    if (params == AOK_PARAMS_COEF) {
        c.coef = p1, c.alpha = p2, c.mode = 0, c.vjp = false;
    } else {
        c.alpha = p1, c.D = p2, c.mode = 1, c.vjp = true;
    }
    // Take over the workspace of the last destroyed plan of this context.
    p.buf = ctx->plan_cache, p.bytes = ctx->plan_cache_bytes;
    ctx->plan_cache = nullptr, ctx->plan_cache_bytes = 0;
    aok_status st = clear_err(ctx);
    if (st == AOK_OK)
        st = dtype == AOK_F32 ? ovl2c_plan_build<float>(ctx, c, p) : ovl2c_plan_build<double>(ctx, c, p);
    if (st == AOK_OK) st = finish(ctx, "aok_ovl2c_plan_create");
    if (st != AOK_OK) {
        aok_ovl2c_plan_destroy(pl);
        return st;
    }
    ctx->plans.push_back(&p);
    *plan = pl;
    return AOK_OK;
}

static aok_status plan_exec(aok_ovl2c_plan plan, const Ovl2cCall &io, bool pullback, const char *what)
{
    aok_ctx ctx = plan->st.ctx;
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    aok_status st = clear_err(ctx);
    if (st != AOK_OK) return st;
    st = plan->st.dtype == AOK_F32 ? ovl2c_plan_exec<float>(ctx, plan->st, io, pullback)
                                   : ovl2c_plan_exec<double>(ctx, plan->st, io, pullback);
    if (st != AOK_OK) return st;
    return finish(ctx, what);
}

aok_status aok_ovl2c_plan_forward(aok_ovl2c_plan plan, const void *posA, const void *posB, void *out)
{
    if (!plan || !plan->st.ctx) return AOK_ERR_ARG;
    aok_ctx ctx = plan->st.ctx;
    const int64_t B = plan->st.call.B;
    AOK_REQUIRE(ctx, B == 0 || (posA && posB), "positions are NULL");
    AOK_REQUIRE(ctx, B == 0 || out != nullptr, "out is NULL");
    Ovl2cCall io{};
    io.posA = posA, io.posB = posB, io.out = out, io.scale = plan->st.call.scale;
    return plan_exec(plan, io, false, "aok_ovl2c_plan_forward");
}

aok_status aok_ovl2c_plan_pullback(aok_ovl2c_plan plan, const void *dS, const void *posA, const void *posB,
                                   void *d1, void *d2)
{
    if (!plan || !plan->st.ctx) return AOK_ERR_ARG;
    aok_ctx ctx = plan->st.ctx;
    const int64_t B = plan->st.call.B;
    AOK_REQUIRE(ctx, B == 0 || (posA && posB), "positions are NULL");
    AOK_REQUIRE(ctx, B == 0 || dS != nullptr, "dS is NULL");
    AOK_REQUIRE(ctx, d1 && d2, "parameter cotangents are NULL");
    Ovl2cCall io{};
    io.posA = posA, io.posB = posB, io.dS = dS, io.o1 = d1, io.o2 = d2, io.scale = plan->st.call.scale;
    return plan_exec(plan, io, true, "aok_ovl2c_plan_pullback");
}

aok_status aok_ovl2c_plan_pullback_x(aok_ovl2c_plan plan, const void *dS, const void *posA, const void *posB,
                                     void *dA, void *dB)
{
    if (!plan || !plan->st.ctx) return AOK_ERR_ARG;
    aok_ctx ctx = plan->st.ctx;
    const int64_t B = plan->st.call.B;
    AOK_REQUIRE(ctx, B == 0 || (posA && posB), "positions are NULL");
    AOK_REQUIRE(ctx, B == 0 || dS != nullptr, "dS is NULL");
    AOK_REQUIRE(ctx, B == 0 || dA != nullptr, "dA is NULL");
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    aok_status st = clear_err(ctx);
    if (st != AOK_OK) return st;
    Ovl2cCall io{};
    io.posA = posA, io.posB = posB, io.dS = dS, io.dA = dA, io.dB = dB, io.scale = plan->st.call.scale;
    st = plan->st.dtype == AOK_F32 ? ovl2c_plan_dx<float>(ctx, plan->st, io) : ovl2c_plan_dx<double>(ctx, plan->st, io);
    if (st != AOK_OK) return st;
    return finish(ctx, "aok_ovl2c_plan_pullback_x");
}

aok_status aok_ovl2c_plan_update(aok_ovl2c_plan plan)
{
    if (!plan || !plan->st.ctx) return AOK_ERR_ARG;
    aok_ctx ctx = plan->st.ctx;
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    aok_status st = clear_err(ctx);
    if (st != AOK_OK) return st;
    st = plan->st.dtype == AOK_F32 ? ovl2c_plan_refresh<float>(ctx, plan->st)
                                   : ovl2c_plan_refresh<double>(ctx, plan->st);
    if (st != AOK_OK) return st;
    return finish(ctx, "aok_ovl2c_plan_update");
}

aok_status aok_ovl2c_plan_destroy(aok_ovl2c_plan plan)
{
    if (!plan) return AOK_OK;
    aok_ctx ctx = plan->st.ctx;
    void *buf = plan->st.buf;
    const size_t bytes = plan->st.bytes;
    if (ctx) {
        auto &v = ctx->plans;
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i] == &plan->st) {
                v.erase(v.begin() + static_cast<std::ptrdiff_t>(i));
                break;
            }
    }
    delete plan;
    if (!buf) return AOK_OK;  // detached plans hold no workspace
    if (!ctx) {
        (void)hipFree(buf);
        return AOK_OK;
    }
    // Keep the larger of the two buffers for the next plan (whose work is ordered after
    // this plan's on the context stream).  The other is freed once the stream has drained:
    // in stream-ordered mode enqueued work may still read it.
    (void)hipSetDevice(ctx->device);
    void *drop = buf;
    if (bytes > ctx->plan_cache_bytes) {
        drop = ctx->plan_cache;
        ctx->plan_cache = buf, ctx->plan_cache_bytes = bytes;
    }
    (void)release_buffer(ctx, drop);
    return AOK_OK;
}

aok_status aok_ovl3c(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *posA, const void *posB,
                     const void *posC, const void *sidxA, const void *sidxB, const void *sidxC, aok_itype itype,
                     int64_t B, int64_t ns, int64_t K, int64_t NZ, const int64_t *ls, const void *coef,
                     const void *alpha, double lengthscale, void *out)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, !bad_dtype(dtype), "dtype must be AOK_F32 or AOK_F64");
    AOK_REQUIRE(ctx, !bad_itype(itype), "itype must be AOK_I64 or AOK_I32");
    AOK_REQUIRE(ctx, (flags & ~AOK_PLAIN) == 0u, "unknown flags");
    AOK_REQUIRE(ctx, ls && coef && alpha, "NULL basis argument");
    AOK_REQUIRE(ctx, B >= 0, "B must be >= 0");
    AOK_REQUIRE(ctx, B == 0 || (posA && posB && posC && out), "positions/out are NULL");
    Ovl3cCall c{flags, posA, posB, posC, sidxA, sidxB, sidxC, itype, B, ls, ns, K, NZ, coef, alpha, lengthscale, out,
                nullptr, nullptr, nullptr};
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    aok_status st = clear_err(ctx);
    if (st != AOK_OK) return st;
    st = dtype == AOK_F32 ? ovl3c_run<float>(ctx, c) : ovl3c_run<double>(ctx, c);
    if (st != AOK_OK) return st;
    return finish(ctx, "aok_ovl3c");
}

aok_status aok_ovl3c_pullback(aok_ctx ctx, aok_dtype dtype, unsigned flags, const void *dT, const void *posA,
                              const void *posB, const void *posC, const void *sidxA, const void *sidxB,
                              const void *sidxC, aok_itype itype, int64_t B, int64_t ns, int64_t K, int64_t NZ,
                              const int64_t *ls, const void *coef, const void *alpha, double lengthscale,
                              void *dcoef, void *dalpha)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, !bad_dtype(dtype), "dtype must be AOK_F32 or AOK_F64");
    AOK_REQUIRE(ctx, !bad_itype(itype), "itype must be AOK_I64 or AOK_I32");
    AOK_REQUIRE(ctx, (flags & ~AOK_PLAIN) == 0u, "unknown flags");
    AOK_REQUIRE(ctx, ls && coef && alpha, "NULL basis argument");
    AOK_REQUIRE(ctx, dcoef && dalpha, "dcoef/dalpha are NULL");
    AOK_REQUIRE(ctx, B >= 0, "B must be >= 0");
    AOK_REQUIRE(ctx, B == 0 || (posA && posB && posC && dT), "positions/dT are NULL");
    Ovl3cCall c{flags, posA, posB, posC, sidxA, sidxB, sidxC, itype, B, ls, ns, K, NZ, coef, alpha, lengthscale,
                nullptr, dT, dcoef, dalpha};
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    aok_status st = clear_err(ctx);
    if (st != AOK_OK) return st;
    st = dtype == AOK_F32 ? ovl3c_pullback_run<float>(ctx, c) : ovl3c_pullback_run<double>(ctx, c);
    if (st != AOK_OK) return st;
    return finish(ctx, "aok_ovl3c_pullback");
}


aok_status aok_graph_begin(aok_ctx ctx)
{
    AOK_CHECK_CTX(ctx);
    AOK_REQUIRE(ctx, !in_capture(ctx), "the context stream is already being captured");
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    AOK_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    ctx->capturing = true;
    ctx->capture_failed = false;
    return AOK_OK;
}

aok_status aok_graph_end(aok_ctx ctx, aok_graph *graph)
{
    AOK_CHECK_CTX(ctx);
    if (graph) *graph = nullptr;
    AOK_REQUIRE(ctx, ctx->capturing, "no capture in progress (aok_graph_begin)");
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->capturing = false;
    const bool failed = ctx->capture_failed;
    ctx->capture_failed = false;
    if (e != hipSuccess || failed || !graph) {
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess)
            return fail(ctx, AOK_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        if (!graph) return fail(ctx, AOK_ERR_ARG, "graph is NULL");
        return fail(ctx, AOK_ERR_ARG, "a call failed during the capture: " + ctx->last_error);
    }
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipGraphDestroy(g);
        return fail(ctx, AOK_ERR_DEVICE, "hipGraphInstantiate failed");
    }
    auto *gr = new (std::nothrow) aok_graph_s{ctx, g, x};
    if (!gr) {
        (void)hipGraphExecDestroy(x);
        (void)hipGraphDestroy(g);
        return fail(ctx, AOK_ERR_NOMEM, "graph allocation failed");
    }
    ctx->graphs.push_back(gr);
    *graph = gr;
    return AOK_OK;
}

aok_status aok_graph_launch(aok_graph graph)
{
    if (!graph || !graph->ctx) return AOK_ERR_ARG;  // NULL or detached
    aok_ctx ctx = graph->ctx;
    AOK_REQUIRE(ctx, !in_capture(ctx), "cannot launch a graph into a stream under capture");
    AOK_HIP(ctx, hipSetDevice(ctx->device));
    aok_status st = clear_err(ctx);
    if (st != AOK_OK) return st;
    AOK_HIP(ctx, hipGraphLaunch(graph->exec, ctx->stream));
    return finish(ctx, "aok_graph_launch");
}

aok_status aok_graph_destroy(aok_graph graph)
{
    if (!graph) return AOK_OK;
    aok_ctx ctx = graph->ctx;
    if (!ctx) {  // detached by aok_destroy
        delete graph;
        return AOK_OK;
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);  // a launched replay may still run
    detach_graph(graph);
    auto &v = ctx->graphs;
    for (size_t i = 0; i < v.size(); ++i)
        if (v[i] == graph) {
            v.erase(v.begin() + static_cast<std::ptrdiff_t>(i));
            break;
        }
    delete graph;
    if (v.empty() && !in_capture(ctx)) {
        for (void *r : ctx->retired) (void)hipFree(r);
        ctx->retired.clear();
    }
    return AOK_OK;
}

}  // extern "C"
