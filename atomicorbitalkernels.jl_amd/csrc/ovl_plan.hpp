// ovl_plan.hpp -- shared layout of the overlap kernels' launch plans and workspace.
//
// Work decomposition of the fast 2-centre kernels:
//  * batch elements b are stably bucketed by species pair (sigma_A, sigma_B) on the
//    device, and every wave takes 64 b's of one bucket, so primitive data, shell
//    structure and loop trip counts are wave-uniform;
//  * per (species pair, shell pair) the primitive-pair constants are precomputed once
//    per call in FP64 ("records"); the kernels read them as uniform data;
//  * the N x N block of each b is produced shell-column by shell-column ("stripes":
//    all rows of shells sa in [sa0, sa1) times the nbf(sb) columns of shell sb), staged
//    in LDS and written (or, for the cotangent, read) with coalesced accesses, because
//    out[:, :, b] is contiguous in the Julia layout.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "../../include/aok.h"

namespace aok {

constexpr int OVL_MAXSH = 64;    // shells per species
constexpr int OVL_MAXCH = 160;   // stripe chunks per call
constexpr int OVL_MAXK = 32;     // primitive slots per shell
constexpr int OVL_MAXNZ = 16;    // species
constexpr int OVL_TILE = 64;     // batch elements per wave

// One stripe chunk: bra shells sa0..sa1-1 (rows row0..row0+nrows-1) x ket shell sb.
struct OvlChunk {
    int sb, sa0, sa1, row0, nrows;
};

// Chunks travel packed in two 32-bit words (sb | sa0 << 8 | sa1 << 16, row0 | nrows << 16):
// the shape is a kernel argument indexed by wave-uniform counters, and 32-bit fields are
// what the kernels can read with scalar loads.
struct OvlChunkP {
    uint32_t s, r;
};

__host__ __device__ inline OvlChunkP pack_chunk(const OvlChunk &c)
{
    return OvlChunkP{static_cast<uint32_t>(c.sb) | (static_cast<uint32_t>(c.sa0) << 8) |
                         (static_cast<uint32_t>(c.sa1) << 16),
                     static_cast<uint32_t>(c.row0) | (static_cast<uint32_t>(c.nrows) << 16)};
}

__host__ __device__ inline OvlChunk unpack_chunk(OvlChunkP p)
{
    return OvlChunk{static_cast<int>(p.s & 0xffu), static_cast<int>((p.s >> 8) & 0xffu),
                    static_cast<int>((p.s >> 16) & 0xffu), static_cast<int>(p.r & 0xffffu),
                    static_cast<int>(p.r >> 16)};
}

// Chunks whose shells all have l <= OVL_LIGHT_L come first (nlight of them, stripe length
// strl) and run in a kernel instance compiled for l <= OVL_LIGHT_L (fewer registers, more
// waves per SIMD); the others run in the instance for the basis' largest l.  str is the
// longest stripe of all chunks.
constexpr int OVL_LIGHT_L = 2;

struct OvlShape {
    int ns, K, NZ, N;
    int nchunk, nlight;
    int str, strl;  // LDS floats per lane (odd): all chunks, light chunks
    int ls[OVL_MAXSH];
    int off[OVL_MAXSH + 1];
    OvlChunkP ch[OVL_MAXCH];
};

// One primitive pair (record) of the fast kernels.  mu2 = alpha*beta/p (times log2(e)
// in the FP32 variant); aop = alpha/p, bop = beta/p, i2p = 1/(2p); Kab = ca cb (pi/p)^{3/2};
// w1 = cb (pi/p)^{3/2}, w2 = ca (pi/p)^{3/2} (pullback weights).
struct alignas(32) PairRecF {
    double mu2;
    float aop, bop, i2p, Kab, w1, w2;
};
struct alignas(64) PairRecD {
    double mu2, aop, bop, i2p, Kab, w1, w2, pad;
};
template <typename T>
struct RecOf;
template <>
struct RecOf<float> {
    using type = PairRecF;
};
template <>
struct RecOf<double> {
    using type = PairRecD;
};

// Pullback-only companion of a record: the coefficients of the pullback's exponent
// tables (pair_contract), c_i = (alpha - (2i+1) beta)/(4 alpha p) = ca - i sa and
// d_j = (beta - (2j+1) alpha)/(4 beta p) = cb - j sb, from ca = (alpha - beta)/(4 alpha p),
// cb = -(alpha - beta)/(4 beta p), sa = beta/(2 alpha p), sb = alpha/(2 beta p) (FP64,
// rounded once).
template <typename T>
struct alignas(4 * sizeof(T)) PairPb {
    T ca, cb, sa, sb;
};

__host__ __device__ inline int nbf_of(int l) { return (l + 1) * (l + 2) / 2; }

// Stripe plan: for each ket shell sb, consecutive bra shells grouped while
// rows * nbf(sb) fits the per-lane LDS capacity `cap`.
inline std::string build_shape(int64_t ns, int64_t K, int64_t NZ, const int64_t *ls, int lmax_allowed, int cap,
                               OvlShape &s, bool split_light = false)
{
    if (ns < 1 || ns > OVL_MAXSH) return "number of shells must be in 1.." + std::to_string(OVL_MAXSH);
    if (K < 1 || K > OVL_MAXK) return "primitive slots K must be in 1.." + std::to_string(OVL_MAXK);
    if (NZ < 1 || NZ > OVL_MAXNZ) return "number of species must be in 1.." + std::to_string(OVL_MAXNZ);
    s.ns = static_cast<int>(ns);
    s.K = static_cast<int>(K);
    s.NZ = static_cast<int>(NZ);
    s.off[0] = 0;
    for (int64_t q = 0; q < ns; ++q) {
        if (ls[q] < 0 || ls[q] > lmax_allowed) return "shell l out of range (max " + std::to_string(lmax_allowed) + ")";
        s.ls[q] = static_cast<int>(ls[q]);
        s.off[q + 1] = s.off[q] + nbf_of(static_cast<int>(ls[q]));
    }
    s.N = s.off[ns];
    // split (asked for by split_light and used when some shell is heavy): pass 0 emits the light chunks (ket and bra shells all l <= OVL_LIGHT_L), pass 1
    // the rest, and bra runs never mix light and heavy shells.  When the split needs more
    // than OVL_MAXCH chunks (shell lists that alternate light and heavy), every chunk runs
    // in the wide instance instead (nlight = 0).
    auto plan = [&](bool split) -> bool {
        int maxl = 0, maxh = 0;
        s.nchunk = 0;
        s.nlight = 0;
        auto light = [&](int q) { return !split || s.ls[q] <= OVL_LIGHT_L; };
        for (int pass = split ? 0 : 1; pass < 2; ++pass) {
            for (int sb = 0; sb < ns; ++sb) {
                const int nb = nbf_of(s.ls[sb]);
                int sa = 0;
                while (sa < ns) {
                    const bool cls = light(sa);
                    int sa1 = sa, rows = 0;
                    while (sa1 < ns && light(sa1) == cls && (rows + nbf_of(s.ls[sa1])) * nb <= cap)
                        rows += nbf_of(s.ls[sa1++]);
                    if (sa1 == sa) {  // a single shell pair larger than cap still gets its own chunk
                        rows = nbf_of(s.ls[sa]);
                        sa1 = sa + 1;
                    }
                    const bool is_light = split && cls && light(sb);
                    if (is_light == (pass == 0)) {
                        if (s.nchunk >= OVL_MAXCH) return false;
                        s.ch[s.nchunk++] = pack_chunk(OvlChunk{sb, sa, sa1, s.off[sa], rows});
                        int &m = is_light ? maxl : maxh;
                        m = std::max(m, rows * nb);
                    }
                    sa = sa1;
                }
            }
            if (pass == 0) s.nlight = s.nchunk;
        }
        s.str = std::max(maxl, maxh) | 1;
        s.strl = split ? (maxl | 1) : s.str;
        return true;
    };
    bool any_heavy = false;
    for (int q = 0; q < s.ns; ++q) any_heavy |= s.ls[q] > OVL_LIGHT_L;
    if (!(split_light && any_heavy && plan(true)) && !plan(false)) return "too many stripe chunks";
    return std::string();
}

// Workspace map of one fast 2c call (byte offsets, 256-aligned).
struct Ovl2cWs {
    size_t coef64, alpha64, nreal, nlist, lidx, pcls, rec, rec2;  // rec2: pullback maps only, else 0
    size_t blk_hist, rank, blk_off, bcnt, boff, toff, perm, tile_key;
    size_t partial, part2;
    size_t total;
    int nblk_h, NK, maxtiles;
};


inline Ovl2cWs ovl2c_ws_map(int64_t B, int ns, int K, int NZ, int nchunk, size_t rec_bytes, size_t pb_bytes,
                            bool pullback)
{
    Ovl2cWs w{};
    const size_t npar = static_cast<size_t>(ns) * K * NZ;
    w.NK = NZ * NZ;
    w.nblk_h = static_cast<int>((B + 255) / 256);
    w.maxtiles = static_cast<int>((B + OVL_TILE - 1) / OVL_TILE) + w.NK;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    w.coef64 = take(sizeof(double) * npar);
    w.alpha64 = take(sizeof(double) * npar);
    w.nreal = take(sizeof(int) * NZ * ns);
    w.nlist = take(sizeof(int) * NZ * ns);
    w.lidx = take(sizeof(int) * npar);
    w.pcls = take(sizeof(int) * npar);
    w.rec = take(rec_bytes * static_cast<size_t>(w.NK) * ns * ns * K * K);
    w.blk_hist = take(sizeof(int) * static_cast<size_t>(w.nblk_h) * w.NK);
    w.rank = take(sizeof(int) * static_cast<size_t>(B > 0 ? B : 1));
    w.blk_off = take(sizeof(int) * static_cast<size_t>(w.nblk_h) * w.NK);
    w.bcnt = take(sizeof(int) * w.NK);
    w.boff = take(sizeof(int) * w.NK);
    w.toff = take(sizeof(int) * (w.NK + 1));
    w.perm = take(sizeof(int) * static_cast<size_t>(B > 0 ? B : 1));
    w.tile_key = take(sizeof(int) * static_cast<size_t>(w.maxtiles));
    if (pullback) {
        w.rec2 = take(pb_bytes * static_cast<size_t>(w.NK) * ns * ns * K * K);
        w.partial = take(sizeof(double) * static_cast<size_t>(w.maxtiles) * nchunk * (ns + 1) * K * 2);
        w.part2 = take(sizeof(double) * 4 * static_cast<size_t>(NZ) * npar);
    }
    w.total = o;
    return w;
}

// One 2-centre call as seen by the kernels' host side.
struct Ovl2cCall {
    unsigned flags;
    const void *posA, *posB, *sidxA, *sidxB;
    int itype;
    int64_t B;
    const int64_t *ls;
    int64_t ns, K, NZ;
    const void *coef;   // mode 0
    const void *alpha;  // exponents (zeta)
    const void *D;      // mode 1
    int mode;
    double scale;
    void *out;
    const void *dS;
    void *o1, *o2;  // pullback outputs: (dcoef, dalpha) or (dzeta, dD)
    bool vjp;
    void *dA, *dB;  // position cotangents [3*B] (dB may be NULL)
};

// One 3-centre call.
struct Ovl3cCall {
    unsigned flags;
    const void *posA, *posB, *posC, *sidxA, *sidxB, *sidxC;
    int itype;
    int64_t B;
    const int64_t *ls;
    int64_t ns, K, NZ;
    const void *coef, *alpha;
    double scale;
    void *out;
    const void *dT;  // pullback: cotangent [N*N*N*B]
    void *o1, *o2;   // pullback outputs (dcoef, dalpha)
};

// A reusable 2-centre plan (owned device workspace + the call description).
struct Ovl2cPlanState {
    aok_ctx ctx = nullptr;
    int dtype = 0;
    int64_t ls[OVL_MAXSH] = {};  // host copy of the shell l's (call.ls points here)
    Ovl2cCall call{};
    OvlShape sh{};   // pullback / position-cotangent chunks (light ones split off)
    OvlShape shf{};  // forward chunks
    Ovl2cWs w{};
    void *buf = nullptr;
    size_t bytes = 0;
};

template <typename T>
aok_status ovl2c_run(aok_ctx ctx, const Ovl2cCall &c, bool pullback);
template <typename T>
aok_status ovl2c_run_noalloc(aok_ctx ctx, const Ovl2cCall &c);
// Look up this device's copy of the fixed scratch array (device current); once per device.
aok_status ovl2c_noalloc_setup(aok_ctx ctx);
template <typename T>
aok_status ovl2c_plan_build(aok_ctx ctx, const Ovl2cCall &c, Ovl2cPlanState &p);
template <typename T>
aok_status ovl2c_plan_refresh(aok_ctx ctx, Ovl2cPlanState &p);
template <typename T>
aok_status ovl2c_plan_exec(aok_ctx ctx, const Ovl2cPlanState &p, const Ovl2cCall &io, bool pullback);
template <typename T>
aok_status ovl2c_dx_run(aok_ctx ctx, const Ovl2cCall &c);
template <typename T>
aok_status ovl2c_plan_dx(aok_ctx ctx, const Ovl2cPlanState &p, const Ovl2cCall &io);
template <typename T>
aok_status ovl3c_run(aok_ctx ctx, const Ovl3cCall &c);
template <typename T>
aok_status ovl3c_pullback_run(aok_ctx ctx, const Ovl3cCall &c);
template <typename T>
aok_status compile_coef_run(aok_ctx ctx, const void *zeta, const void *D, int64_t ns, int64_t K, int64_t NZ,
                            const int64_t *ls, void *coef);
template <typename T>
aok_status compile_coef_pb_run(aok_ctx ctx, const void *dcoef, const void *zeta, const void *D, int64_t ns,
                               int64_t K, int64_t NZ, const int64_t *ls, void *dD, void *dzeta_norm);

}  // namespace aok
