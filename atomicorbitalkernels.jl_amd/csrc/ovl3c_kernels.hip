// ovl3c_kernels.hip -- H2: batched 3-centre Cartesian-Gaussian overlaps on gfx950.
//
// V[mu, nu, lambda](b) = int phi_mu(r - A_b) phi_nu(r - B_b) phi_lambda(r - C_b) dr with the
// semantics of batch_V_kernel! (src/integrals/kernels_3c.jl:12-162): contracted values
// accumulated at lin = index1 + N1 index2 + N1 N2 index3 (N1 = 2l_a+1, N2 = 2l_b+1) and
// read back as out[off_a+i, off_b+j, off_c+k] = blk[i + nbf_a (j + nbf_b k)]; generalised to
// a species per centre with the orbital path's padded [ns x K x NZ] primitive layout.
//
// Fast path: batch elements are bucketed by species triple; a wave takes 64 b's of one
// bucket.  Per primitive triple the constants are FP64 records built on the device:
//   exponent  (ab |AB|^2 + ag |AC|^2 + bg |BC|^2) / q   (times log2 e), evaluated per lane
//             in FP64 from the pairwise distances and range-reduced for v_exp_f32;
//   geometry  PA = -(beta/p) AB, PB = (alpha/p) AB, PC = PA + AC,
//             QA = PA - (gamma/q) PC, QB = PB - (gamma/q) PC, QC = (p/q) PC;
//   1-D Obara-Saika three-centre overlaps s_ijk (s_000 = 1) per axis,
//   s_{i+1,j,k} = QA s_ijk + (i s_{i-1,j,k} + j s_{i,j-1,k} + k s_{i,j,k-1}) / (2q).
// Plain path: one work item per (b, s_a, s_b, s_c), runtime l loops, private E3 scratch
// (the reference decomposition; the E3 table is zero-filled beyond t = i+j+k).
#include "aok_common.hpp"
#include "ovl_plan.hpp"
#include "ovl_prep.hpp"

namespace aok {

constexpr int OVL3_MAXNZ = 8;
constexpr int OVL3_FAST_LMAX = 2;
constexpr size_t OVL3_ROWS_CAP = size_t(1) << 30;  // pullback partial rows, bytes (part choice)

struct alignas(16) TripRecF {
    double cab, cac, cbc;  // (pairwise exponent products / q) * log2(e)
    float aop, bop, gq, pq, i2q, Kabc;
};
struct alignas(16) TripRecD {
    double cab, cac, cbc, aop, bop, gq, pq, i2q, Kabc, pad;
};
template <typename T>
struct Rec3Of;
template <>
struct Rec3Of<float> {
    using type = TripRecF;
};
template <>
struct Rec3Of<double> {
    using type = TripRecD;
};

struct Ovl3Ws {
    size_t coef64, alpha64, nreal, lidx, grp_off, rec;
    size_t blk_hist, rank, blk_off, bcnt, boff, toff, perm, tile_key, tri;
    size_t total;
    int nblk_h, NK, maxtiles;
    int64_t ngrp, maxrec;
};

static Ovl3Ws ovl3_ws_map(int64_t B, int ns, int K, int NZ, size_t rec_bytes)
{
    Ovl3Ws w{};
    const size_t npar = static_cast<size_t>(ns) * K * NZ;
    w.NK = NZ * NZ * NZ;
    w.nblk_h = static_cast<int>((B + 255) / 256);
    // tiles padded to a multiple of 8: the forward's grid maps tile t to XCD t mod 8 (both
    // grid forms), so its blocks' partial lines meet in that XCD's L2
    w.maxtiles = ((static_cast<int>((B + OVL_TILE - 1) / OVL_TILE) + w.NK) + 7) & ~7;
    w.ngrp = static_cast<int64_t>(w.NK) * ns * ns * ns;
    w.maxrec = w.ngrp * K * K * K;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    w.coef64 = take(sizeof(double) * npar);
    w.alpha64 = take(sizeof(double) * npar);
    w.nreal = take(sizeof(int) * NZ * ns);
    w.lidx = take(sizeof(int) * npar);
    w.grp_off = take(sizeof(int64_t) * (w.ngrp + 1));
    w.rec = take(rec_bytes * static_cast<size_t>(w.maxrec));
    w.blk_hist = take(sizeof(int) * static_cast<size_t>(w.nblk_h) * w.NK);
    w.rank = take(sizeof(int) * static_cast<size_t>(B > 0 ? B : 1));
    w.blk_off = take(sizeof(int) * static_cast<size_t>(w.nblk_h) * w.NK);
    w.bcnt = take(sizeof(int) * w.NK);
    w.boff = take(sizeof(int) * w.NK);
    w.toff = take(sizeof(int) * (w.NK + 1));
    w.perm = take(sizeof(int) * static_cast<size_t>(B > 0 ? B : 1));
    w.tile_key = take(sizeof(int) * static_cast<size_t>(w.maxtiles));
    w.tri = take(sizeof(int) * static_cast<size_t>(ns) * ns * ns);
    w.total = o;
    return w;
}

__device__ __forceinline__ int key3(const void *sA, const void *sB, const void *sC, int itype, int64_t i, int NZ,
                                    unsigned *err)
{
    int a = load_sigma(sA, itype, i), b = load_sigma(sB, itype, i), c = load_sigma(sC, itype, i);
    const bool bad = a < 0 || a >= NZ || b < 0 || b >= NZ || c < 0 || c >= NZ;
    if (bad && err) *err = ERR_SPECIES;
    a = (a < 0 || a >= NZ) ? 0 : a;
    b = (b < 0 || b >= NZ) ? 0 : b;
    c = (c < 0 || c >= NZ) ? 0 : c;
    return (a * NZ + b) * NZ + c;
}

__global__ __launch_bounds__(256) void ovl3_hist_kernel(const void *__restrict__ sA, const void *__restrict__ sB,
                                                        const void *__restrict__ sC, int itype, int64_t B, int NZ,
                                                        int *__restrict__ rank, int *__restrict__ blk_hist,
                                                        unsigned *__restrict__ err)
{
    __shared__ int whist[4 * OVL3_MAXNZ * OVL3_MAXNZ * OVL3_MAXNZ];
    const int NK = NZ * NZ * NZ;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int key = i < B ? key3(sA, sB, sC, itype, i, NZ, err) : -1;
    const int r = block_rank(key, NK, whist, blk_hist + static_cast<int64_t>(blockIdx.x) * NK);
    if (i < B) rank[i] = r;
}

// One block: real-primitive lists, bucket offsets, and the exclusive scan of the
// per-(species triple, shell triple) record counts.
template <typename T>
__global__ __launch_bounds__(256) void ovl3_plan_kernel(const OvlShape sh, const T *__restrict__ coefT,
                                                        const T *__restrict__ alphaT, double *__restrict__ coef64,
                                                        double *__restrict__ alpha64, int *__restrict__ nreal,
                                                        int *__restrict__ lidx, int64_t *__restrict__ grp_off,
                                                        int nblk_h, const int *__restrict__ blk_hist,
                                                        int *__restrict__ blk_off, int *__restrict__ bcnt,
                                                        int *__restrict__ boff, int *__restrict__ toff,
                                                        int *__restrict__ tri, int light_l)
{
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ, NK = NZ * NZ * NZ;
    for (int idx = threadIdx.x; idx < NZ * ns; idx += blockDim.x) {
        const int sg = idx / ns, s = idx - sg * ns;
        int nr = 0;
        for (int k = 0; k < K; ++k) {
            const int64_t q = s + static_cast<int64_t>(ns) * (k + static_cast<int64_t>(K) * sg);
            const double c = static_cast<double>(coefT[q]);
            coef64[q] = c;
            alpha64[q] = static_cast<double>(alphaT[q]);
            if (coefT[q] != T(0)) lidx[static_cast<int64_t>(idx) * K + nr++] = k;
        }
        nreal[idx] = nr;
    }
    bucket_offsets(NK, nblk_h, blk_hist, blk_off, bcnt, boff, toff);
    // the forward's shell triples: those of light shells only (l <= light_l) first, then
    // the others, each part in (c, b, a) order (measured on C4a: 0.427 ms, against 0.44-0.45
    // with the heavy part ordered by its number of d shells).  Stream compaction in rounds
    // of 256 triples: wave ballots give each triple its rank within its part, the running
    // part totals sit in LDS.
    {
        __shared__ int part_tot[2], wcnt[4][2];
        const int ns3 = ns * ns * ns;
        int nL = 0;
        for (int q = 0; q < ns; ++q) nL += sh.ls[q] <= light_l;
        const int poff[2] = {0, nL * nL * nL};
        if (threadIdx.x < 2) part_tot[threadIdx.x] = 0;
        __syncthreads();
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int q0 = 0; q0 < ns3; q0 += 256) {
            const int q = q0 + static_cast<int>(threadIdx.x);
            int m = -1, sa = 0, sb = 0, sc = 0;
            if (q < ns3) {
                sa = q % ns;
                sb = (q / ns) % ns;
                sc = q / (ns * ns);
                m = (sh.ls[sa] > light_l || sh.ls[sb] > light_l || sh.ls[sc] > light_l) ? 1 : 0;
            }
            int rank = 0;
            for (int k = 0; k < 2; ++k) {
                const unsigned long long bal = __ballot(m == k);
                if (m == k) rank = __popcll(bal & ((1ull << lane) - 1ull));
                if (lane == 0) wcnt[wave][k] = __popcll(bal);
            }
            __syncthreads();
            if (m >= 0) {
                int pos = poff[m] + part_tot[m] + rank;
                for (int w = 0; w < wave; ++w) pos += wcnt[w][m];
                tri[pos] = sa | (sb << 8) | (sc << 16);
            }
            __syncthreads();
            if (threadIdx.x < 2)
                part_tot[threadIdx.x] += wcnt[0][threadIdx.x] + wcnt[1][threadIdx.x] + wcnt[2][threadIdx.x] +
                                         wcnt[3][threadIdx.x];
            __syncthreads();
        }
    }
    // record offsets: contiguous ranges of groups per thread, then a scan of the partials
    __shared__ int64_t wtot[4];
    const int64_t ngrp = static_cast<int64_t>(NK) * ns * ns * ns;
    const int64_t per = (ngrp + blockDim.x - 1) / blockDim.x;
    const int64_t g0 = per * threadIdx.x, g1 = min(ngrp, g0 + per);
    auto count = [&](int64_t g) -> int64_t {
        const int sa = static_cast<int>(g % ns), sb = static_cast<int>((g / ns) % ns);
        const int sc = static_cast<int>((g / (static_cast<int64_t>(ns) * ns)) % ns);
        const int key = static_cast<int>(g / (static_cast<int64_t>(ns) * ns * ns));
        const int a = key / (NZ * NZ), b = (key / NZ) % NZ, c = key % NZ;
        return static_cast<int64_t>(nreal[a * ns + sa]) * nreal[b * ns + sb] * nreal[c * ns + sc];
    };
    __syncthreads();  // nreal of every (species, shell) is written
    int64_t s = 0;
    for (int64_t g = g0; g < g1; ++g) s += count(g);
    int64_t total = 0;
    s = block_excl_scan64(s, wtot, total);
    if (threadIdx.x == 0) grp_off[ngrp] = total;
    for (int64_t g = g0; g < g1; ++g) {
        grp_off[g] = s;
        s += count(g);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void ovl3_fill_kernel(const OvlShape sh, const void *__restrict__ sA,
                                                        const void *__restrict__ sB, const void *__restrict__ sC,
                                                        int itype, int64_t B, int nblk_perm, int nblk_tile,
                                                        const int *__restrict__ rank, const int *__restrict__ blk_off,
                                                        const int *__restrict__ toff, int *__restrict__ perm,
                                                        int *__restrict__ tile_key, int maxtiles,
                                                        const double *__restrict__ coef64,
                                                        const double *__restrict__ alpha64,
                                                        const int *__restrict__ nreal, const int *__restrict__ lidx,
                                                        const int64_t *__restrict__ grp_off,
                                                        typename Rec3Of<T>::type *__restrict__ rec)
{
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ, NK = NZ * NZ * NZ;
    const int bid = blockIdx.x;
    if (bid < nblk_perm) {
        const int64_t i = static_cast<int64_t>(bid) * 256 + threadIdx.x;
        if (i >= B) return;
        const int key = key3(sA, sB, sC, itype, i, NZ, nullptr);
        perm[blk_off[static_cast<int64_t>(i / 256) * NK + key] + rank[i]] = static_cast<int>(i);
        return;
    }
    if (bid < nblk_perm + nblk_tile) {
        const int t = (bid - nblk_perm) * 256 + threadIdx.x;
        if (t >= maxtiles) return;
        int key = -1;
        if (t < toff[NK]) {
            key = 0;
            while (key + 1 < NK && toff[key + 1] <= t) ++key;
        }
        tile_key[t] = key;
        return;
    }
    const int64_t gid = static_cast<int64_t>(bid - nblk_perm - nblk_tile) * 256 + threadIdx.x;
    const int64_t K3 = static_cast<int64_t>(K) * K * K;
    const int64_t ngrp = static_cast<int64_t>(NK) * ns * ns * ns;
    if (gid >= ngrp * K3) return;
    const int64_t g = gid / K3;
    const int rem = static_cast<int>(gid - g * K3);
    const int i = rem / (K * K), j = (rem / K) % K, k = rem % K;
    const int sa = static_cast<int>(g % ns), sb = static_cast<int>((g / ns) % ns);
    const int sc = static_cast<int>((g / (static_cast<int64_t>(ns) * ns)) % ns);
    const int key = static_cast<int>(g / (static_cast<int64_t>(ns) * ns * ns));
    const int a = key / (NZ * NZ), b = (key / NZ) % NZ, c = key % NZ;
    const int na = nreal[a * ns + sa], nb = nreal[b * ns + sb], nc = nreal[c * ns + sc];
    if (i >= na || j >= nb || k >= nc) return;
    const int ka = lidx[(a * ns + sa) * K + i], kb = lidx[(b * ns + sb) * K + j], kc = lidx[(c * ns + sc) * K + k];
    const int64_t qa = sa + static_cast<int64_t>(ns) * (ka + static_cast<int64_t>(K) * a);
    const int64_t qb = sb + static_cast<int64_t>(ns) * (kb + static_cast<int64_t>(K) * b);
    const int64_t qc = sc + static_cast<int64_t>(ns) * (kc + static_cast<int64_t>(K) * c);
    const double al = alpha64[qa], be = alpha64[qb], ga = alpha64[qc];
    const double p = al + be, q = p + ga, iq = 1.0 / q, ip = 1.0 / p;
    const double l2e = std::is_same<T, float>::value ? 1.4426950408889634074 : 1.0;
    typename Rec3Of<T>::type r;
    r.cab = al * be * iq * l2e;
    r.cac = al * ga * iq * l2e;
    r.cbc = be * ga * iq * l2e;
    r.aop = static_cast<T>(al * ip);
    r.bop = static_cast<T>(be * ip);
    r.gq = static_cast<T>(ga * iq);
    r.pq = static_cast<T>(p * iq);
    r.i2q = static_cast<T>(0.5 * iq);
    r.Kabc = static_cast<T>(coef64[qa] * coef64[qb] * coef64[qc] * 5.568327996831707845 * iq * sqrt(iq));
    rec[grp_off[g] + (static_cast<int64_t>(i) * nb + j) * nc + k] = r;
}

// Unroll factor of the forward's primitive-triple loop (records of several triples
// loaded together).
#ifndef AOK_F3_UNROLL
#define AOK_F3_UNROLL 2
#endif
#ifndef AOK_PRAGMA
#define AOK_PRAGMA_(x) _Pragma(#x)
#define AOK_PRAGMA(x) AOK_PRAGMA_(x)
#endif

template <int L>
struct Sh3 {
    static constexpr int nbf = (L + 1) * (L + 2) / 2;
};

// 1-D three-centre Obara-Saika table, s[0][0][0] = 1.
template <typename T, int IA, int IB, int IC>
__device__ __forceinline__ void os3(T QA, T QB, T QC, T i2q, T (&s)[IA + 1][IB + 1][IC + 1])
{
    s[0][0][0] = T(1);
#pragma unroll
    for (int i = 1; i <= IA; ++i)
        s[i][0][0] = QA * s[i - 1][0][0] + (i > 1 ? T(i - 1) * i2q * s[i > 1 ? i - 2 : 0][0][0] : T(0));
#pragma unroll
    for (int j = 1; j <= IB; ++j)
#pragma unroll
        for (int i = 0; i <= IA; ++i) {
            T v = QB * s[i][j - 1][0];
            if (i > 0) v += T(i) * i2q * s[i > 0 ? i - 1 : 0][j - 1][0];
            if (j > 1) v += T(j - 1) * i2q * s[i][j > 1 ? j - 2 : 0][0];
            s[i][j][0] = v;
        }
#pragma unroll
    for (int k = 1; k <= IC; ++k)
#pragma unroll
        for (int j = 0; j <= IB; ++j)
#pragma unroll
            for (int i = 0; i <= IA; ++i) {
                T v = QC * s[i][j][k - 1];
                if (i > 0) v += T(i) * i2q * s[i > 0 ? i - 1 : 0][j][k - 1];
                if (j > 0) v += T(j) * i2q * s[i][j > 0 ? j - 1 : 0][k - 1];
                if (k > 1) v += T(k - 1) * i2q * s[i][j][k > 1 ? k - 2 : 0];
                s[i][j][k] = v;
            }
}

__device__ __forceinline__ float trip_exp(const TripRecF &r, double rab, double rac, double rbc)
{
    const double y = r.cab * rab + r.cac * rac + r.cbc * rbc;
    const double n = rint(y);
    return ldexpf(__builtin_amdgcn_exp2f(-static_cast<float>(y - n)), -static_cast<int>(n));
}
__device__ __forceinline__ double trip_exp(const TripRecD &r, double rab, double rac, double rbc)
{
    return exp(-(r.cab * rab + r.cac * rac + r.cbc * rbc));
}

struct Geo3 {
    double rab, rac, rbc;
};

template <typename T, int LA, int LB, int LC>
__device__ __forceinline__ void triple_fwd(const typename Rec3Of<T>::type *__restrict__ recs, int na, int nb, int nc,
                                           const Geo3 &geo, const T (&ab)[3], const T (&ac)[3], T *__restrict__ out,
                                           int N, int offa, int offb, int offc, bool active)
{
    constexpr int NBA = Sh3<LA>::nbf, NBB = Sh3<LB>::nbf, NBC = Sh3<LC>::nbf;
    constexpr int N1 = 2 * LA + 1, N2 = 2 * LB + 1;
    constexpr int NLIN = (NBA - 1) + N1 * (NBB - 1) + N1 * N2 * (NBC - 1) + 1;
    T acc[NLIN];
#pragma unroll
    for (int q = 0; q < NLIN; ++q) acc[q] = T(0);
    const int ntrip = na * nb * nc;
    AOK_PRAGMA(unroll AOK_F3_UNROLL)
    for (int t = 0; t < ntrip; ++t) {
        const typename Rec3Of<T>::type r = recs[t];
        const T G = r.Kabc * trip_exp(r, geo.rab, geo.rac, geo.rbc);
        if constexpr (LA + LB + LC == 0) {
            acc[0] += G;
        } else {
            T sx[LA + 1][LB + 1][LC + 1], sy[LA + 1][LB + 1][LC + 1], sz[LA + 1][LB + 1][LC + 1];
            {
                const T PA = -r.bop * ab[0], PB = r.aop * ab[0], PC = PA + ac[0];
                os3<T, LA, LB, LC>(PA - r.gq * PC, PB - r.gq * PC, r.pq * PC, r.i2q, sx);
            }
            {
                const T PA = -r.bop * ab[1], PB = r.aop * ab[1], PC = PA + ac[1];
                os3<T, LA, LB, LC>(PA - r.gq * PC, PB - r.gq * PC, r.pq * PC, r.i2q, sy);
            }
            {
                const T PA = -r.bop * ab[2], PB = r.aop * ab[2], PC = PA + ac[2];
                os3<T, LA, LB, LC>(PA - r.gq * PC, PB - r.gq * PC, r.pq * PC, r.i2q, sz);
            }
            int index1 = 0;
#pragma unroll
            for (int ll1 = LA; ll1 >= 0; --ll1)
#pragma unroll
                for (int n1 = 0; n1 <= LA - ll1; ++n1) {
                    const int m1 = LA - ll1 - n1;
                    int index2 = 0;
#pragma unroll
                    for (int ll2 = LB; ll2 >= 0; --ll2)
#pragma unroll
                        for (int n2 = 0; n2 <= LB - ll2; ++n2) {
                            const int m2 = LB - ll2 - n2;
                            int index3 = 0;
#pragma unroll
                            for (int ll3 = LC; ll3 >= 0; --ll3)
#pragma unroll
                                for (int n3 = 0; n3 <= LC - ll3; ++n3) {
                                    const int m3 = LC - ll3 - n3;
                                    acc[index1 + N1 * index2 + N1 * N2 * index3] +=
                                        G * sx[ll1][ll2][ll3] * sy[m1][m2][m3] * sz[n1][n2][n3];
                                    ++index3;
                                }
                            ++index2;
                        }
                    ++index1;
                }
        }
    }
    if (!active) return;
#pragma unroll
    for (int k = 0; k < NBC; ++k)
#pragma unroll
        for (int j = 0; j < NBB; ++j)
#pragma unroll
            for (int i = 0; i < NBA; ++i) {
                const int q = i + NBA * (j + NBB * k);
                out[(offa + i) + static_cast<int64_t>(N) * ((offb + j) + static_cast<int64_t>(N) * (offc + k))] =
                    q < NLIN ? acc[q < NLIN ? q : 0] : T(0);
            }
}

#define AOK_L3_CASES(X)                                                                                         \
    X(0, 0, 0) X(0, 0, 1) X(0, 0, 2) X(0, 1, 0) X(0, 1, 1) X(0, 1, 2) X(0, 2, 0) X(0, 2, 1) X(0, 2, 2) X(1, 0, 0) \
        X(1, 0, 1) X(1, 0, 2) X(1, 1, 0) X(1, 1, 1) X(1, 1, 2) X(1, 2, 0) X(1, 2, 1) X(1, 2, 2) X(2, 0, 0)       \
            X(2, 0, 1) X(2, 0, 2) X(2, 1, 0) X(2, 1, 1) X(2, 1, 2) X(2, 2, 0) X(2, 2, 1) X(2, 2, 2)

// The FP32 l <= 2 instance is compiled for 2 waves/SIMD (256 VGPRs instead of 1 wave).
template <typename T, int LMAXB>
constexpr int ovl3_fwd_waves()
{
    return std::is_same<T, float>::value && LMAXB == 2 ? 2 : 1;
}

// One workgroup per (tile, shell triple of this launch's range tri[tri0 ...]), so a heavy
// triple (9 x 9 x 9 primitive triples) and a light one (1 x 1 x 1) are separate workgroups
// that the dispatcher balances.
template <typename T, int LMAXB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(ovl3_fwd_waves<T, LMAXB>(), 8))) void ovl3c_fwd_fast(const OvlShape sh, const T *__restrict__ posA,
                                                     const T *__restrict__ posB, const T *__restrict__ posC, T scale,
                                                     const int *__restrict__ perm, const int *__restrict__ tile_key,
                                                     const int *__restrict__ boff, const int *__restrict__ bcnt,
                                                     const int *__restrict__ toff,
                                                     const typename Rec3Of<T>::type *__restrict__ rec,
                                                     const int64_t *__restrict__ grp_off,
                                                     const int *__restrict__ nreal, const int *__restrict__ tri,
                                                     int tri0, int tiles8, T *__restrict__ out)
{
    const int ns = sh.ns, NZ = sh.NZ, N = sh.N, NK = NZ * NZ * NZ;
    // 1-D grid, XCD-affine and triple-fastest: workgroup w runs on XCD w % 8; its sequence
    // number s = w / 8 on that XCD walks tile q * 8 + w % 8 (q = s / count) through all
    // `count` triples of the launch before the XCD's next tile, so a tile's output lines are
    // completed in one L2
    const int count = static_cast<int>(gridDim.x >> 3) / tiles8;
    const int sq = static_cast<int>(blockIdx.x >> 3), q = sq / count;
    const int t = q * 8 + static_cast<int>(blockIdx.x & 7u), j = sq - q * count;
    if (t >= toff[NK]) return;
    const int code = tri[tri0 + j];
    const int sa = code & 0xff, sb = (code >> 8) & 0xff, sc = code >> 16;
    const int key = tile_key[t];
    const int first = boff[key] + (t - toff[key]) * OVL_TILE;
    const int nact = min(OVL_TILE, boff[key] + bcnt[key] - first);
    const int lane = threadIdx.x;
    const bool active = lane < nact;
    const int64_t b = perm[first + (active ? lane : 0)];
    const int sgA = key / (NZ * NZ), sgB = (key / NZ) % NZ, sgC = key % NZ;
    const int la = sh.ls[sa], lb = sh.ls[sb], lc = sh.ls[sc];
    const int na = nreal[sgA * ns + sa], nb = nreal[sgB * ns + sb], nc = nreal[sgC * ns + sc];
    T A[3], Bv[3], C[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        A[ax] = scale * posA[3 * b + ax];
        Bv[ax] = scale * posB[3 * b + ax];
        C[ax] = scale * posC[3 * b + ax];
    }
    Geo3 geo{0.0, 0.0, 0.0};
    T ab[3], ac[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const double dab = static_cast<double>(A[ax]) - static_cast<double>(Bv[ax]);
        const double dac = static_cast<double>(A[ax]) - static_cast<double>(C[ax]);
        const double dbc = static_cast<double>(Bv[ax]) - static_cast<double>(C[ax]);
        geo.rab += dab * dab;
        geo.rac += dac * dac;
        geo.rbc += dbc * dbc;
        ab[ax] = static_cast<T>(dab);
        ac[ax] = static_cast<T>(dac);
    }
    T *ob = out + static_cast<int64_t>(N) * N * N * b;
    const int64_t g = ((static_cast<int64_t>(key) * ns + sc) * ns + sb) * ns + sa;
    const typename Rec3Of<T>::type *recs = rec + grp_off[g];
#define AOK_F3_CASE(A_, B_, C_)                                                                                 \
    if constexpr (A_ <= LMAXB && B_ <= LMAXB && C_ <= LMAXB)                                                    \
        if (la == A_ && lb == B_ && lc == C_) {                                                                 \
            triple_fwd<T, A_, B_, C_>(recs, na, nb, nc, geo, ab, ac, ob, N, sh.off[sa], sh.off[sb], sh.off[sc], \
                                      active);                                                                  \
            return;                                                                                             \
        }
    AOK_L3_CASES(AOK_F3_CASE)
#undef AOK_F3_CASE
}

// ---- plain path: one item per (b, s_a, s_b, s_c) --------------------------------------------
template <typename T, int LMAXP>
__global__ __launch_bounds__(128) void ovl3c_fwd_plain(const OvlShape sh, int64_t B, const T *__restrict__ coef,
                                                       const T *__restrict__ alpha, const void *__restrict__ sA,
                                                       const void *__restrict__ sB, const void *__restrict__ sC,
                                                       int itype, const T *__restrict__ posA,
                                                       const T *__restrict__ posB, const T *__restrict__ posC,
                                                       T scale, T *__restrict__ out, unsigned *__restrict__ err)
{
    constexpr int NBF = (LMAXP + 1) * (LMAXP + 2) / 2;
    constexpr int TT = 3 * LMAXP + 4;
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ, N = sh.N;
    const int64_t item = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nsh3 = static_cast<int64_t>(ns) * ns * ns;
    if (item >= B * nsh3) return;
    const int64_t b = item % B;
    const int64_t s3 = item / B;
    const int sa = static_cast<int>(s3 % ns), sb = static_cast<int>((s3 / ns) % ns), sc = static_cast<int>(s3 / (ns * ns));
    int ga = load_sigma(sA, itype, b), gb = load_sigma(sB, itype, b), gc = load_sigma(sC, itype, b);
    if (ga < 0 || ga >= NZ || gb < 0 || gb >= NZ || gc < 0 || gc >= NZ) {
        *err = ERR_SPECIES;
        return;
    }
    const int la = sh.ls[sa], lb = sh.ls[sb], lc = sh.ls[sc];
    const int na = nbf_of(la), nb = nbf_of(lb), nc = nbf_of(lc);
    const int N1 = 2 * la + 1, N2 = 2 * lb + 1;
    T E3[TT][LMAXP + 1][LMAXP + 1][LMAXP + 1][3];
    T blk[NBF * NBF * NBF];
    for (int q = 0; q < na * nb * nc; ++q) blk[q] = T(0);
    T A[3], Bv[3], C[3];
    for (int ax = 0; ax < 3; ++ax) {
        A[ax] = scale * posA[3 * b + ax];
        Bv[ax] = scale * posB[3 * b + ax];
        C[ax] = scale * posC[3 * b + ax];
    }
    const T pi32 = T(5.568327996831707845);
    for (int ip = 0; ip < K; ++ip)
        for (int jp = 0; jp < K; ++jp)
            for (int kp = 0; kp < K; ++kp) {
                const int64_t qa = sa + static_cast<int64_t>(ns) * (ip + static_cast<int64_t>(K) * ga);
                const int64_t qb = sb + static_cast<int64_t>(ns) * (jp + static_cast<int64_t>(K) * gb);
                const int64_t qc = sc + static_cast<int64_t>(ns) * (kp + static_cast<int64_t>(K) * gc);
                const T ca = coef[qa], cb = coef[qb], cc = coef[qc];
                if (ca == T(0) || cb == T(0) || cc == T(0)) continue;
                const T aa = alpha[qa], ab = alpha[qb], ac = alpha[qc];
                const T p = aa + ab, q = p + ac, mu = aa * ab / p, nu = p * ac / q, oo2q = T(0.5) / q;
                const T prefac = ca * cb * cc * pi32 / (q * sqrt(q));
                for (int ax = 0; ax < 3; ++ax) {
                    const T P = (aa * A[ax] + ab * Bv[ax]) / p;
                    const T Q = (p * P + ac * C[ax]) / q;
                    const T AB = A[ax] - Bv[ax], PC = P - C[ax];
                    const T QA = Q - A[ax], QB = Q - Bv[ax], QC = Q - C[ax];
                    for (int k = 0; k <= lc; ++k)
                        for (int j = 0; j <= lb; ++j)
                            for (int i = 0; i <= la; ++i)
                                for (int t = 0; t < TT; ++t) E3[t][i][j][k][ax] = T(0);
                    E3[0][0][0][0][ax] = exp(-mu * AB * AB - nu * PC * PC);
                    for (int k = 0; k <= lc; ++k)
                        for (int j = 0; j <= lb; ++j)
                            for (int i = 0; i <= la; ++i) {
                                if (i == 0 && j == 0 && k == 0) continue;
                                int pi_ = i, pj = j, pk = k;
                                T X;
                                if (i > 0) { pi_ = i - 1; X = QA; }
                                else if (j > 0) { pj = j - 1; X = QB; }
                                else { pk = k - 1; X = QC; }
                                for (int t = 0; t <= i + j + k; ++t) {
                                    const T em = t > 0 ? E3[t - 1][pi_][pj][pk][ax] : T(0);
                                    E3[t][i][j][k][ax] = X * E3[t][pi_][pj][pk][ax] +
                                                         T(t + 1) * E3[t + 1][pi_][pj][pk][ax] + oo2q * em;
                                }
                            }
                }
                int index1 = 0;
                for (int ll1 = la; ll1 >= 0; --ll1)
                    for (int n1 = 0; n1 <= la - ll1; ++n1) {
                        const int m1 = la - ll1 - n1;
                        int index2 = 0;
                        for (int ll2 = lb; ll2 >= 0; --ll2)
                            for (int n2 = 0; n2 <= lb - ll2; ++n2) {
                                const int m2 = lb - ll2 - n2;
                                int index3 = 0;
                                for (int ll3 = lc; ll3 >= 0; --ll3)
                                    for (int n3 = 0; n3 <= lc - ll3; ++n3) {
                                        const int m3 = lc - ll3 - n3;
                                        blk[index1 + N1 * index2 + N1 * N2 * index3] +=
                                            E3[0][ll1][ll2][ll3][0] * E3[0][m1][m2][m3][1] * E3[0][n1][n2][n3][2] *
                                            prefac;
                                        ++index3;
                                    }
                                ++index2;
                            }
                        ++index1;
                    }
            }
    T *ob = out + static_cast<int64_t>(N) * N * N * b;
    for (int k = 0; k < nc; ++k)
        for (int j = 0; j < nb; ++j)
            for (int i = 0; i < na; ++i)
                ob[(sh.off[sa] + i) + static_cast<int64_t>(N) * ((sh.off[sb] + j) + static_cast<int64_t>(N) * (sh.off[sc] + k))] =
                    blk[i + na * (j + nb * k)];
}

// ---------------------------------------------------------------------------------
// Pullback (coef, alpha) of the 3-centre overlap.  No reference counterpart; the same
// rules as the 2-centre pullback: every primitive slot receives its coefficient
// cotangent (a padded slot's is the derivative at coef = 0), exponent cotangents by the
// raised-momentum identity d/dalpha x^l e^{-alpha x^2} = -x^{l+2} e^{-alpha x^2}.
// Per (species, shell) a list holds the real slots, then one representative per
// distinct padded exponent (pcls maps every slot to its list entry).  One thread per
// (batch element, part) walks the part's shell triples and every list triple with at
// most one padded entry (two padded entries contribute nothing), accumulating into its
// own FP64 row [side][shell][list entry][coef, exponent]; ovl3_bwd_reduce then sums the
// rows of each slot over the batch and the parts in a fixed order.  Deterministic, no
// atomics.  Parts split a triple's shell loops over the grid's y dimension so that a
// batch of a few thousand triples still fills the chip: by third-centre shell (the row
// then holds that one C shell) and optionally by second-centre shell too (one B shell).
// Rows are stored element-major, batch innermost (element e of (b, part p) at
// rows[(p * rlen + e) * Bc + b - b0] for the chunk [b0, b0 + Bc) of the batch): consecutive
// lanes touch consecutive doubles.  Split rows are folded per chunk into the unsplit
// layout over the whole batch (ovl3_parts_sum), so the split rows need only one chunk.
// ---------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void widen_pair(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                  double *__restrict__ x64, double *__restrict__ y64)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    x64[i] = static_cast<double>(x[i]);
    y64[i] = static_cast<double>(y[i]);
}

__global__ __launch_bounds__(64) void ovl3_pb_lists(const OvlShape sh, const double *__restrict__ coef64,
                                                    const double *__restrict__ alpha64, int *__restrict__ nreal,
                                                    int *__restrict__ nlist, int *__restrict__ lidx,
                                                    int *__restrict__ pcls, double *__restrict__ lcoef,
                                                    double *__restrict__ lalpha)
{
    const int ns = sh.ns, K = sh.K;
    const int idx = static_cast<int>(blockIdx.x) * 64 + threadIdx.x;  // sigma * ns + s
    if (idx >= sh.NZ * ns) return;
    const int sg = idx / ns, s = idx - sg * ns;
    auto q = [&](int k) { return s + static_cast<int64_t>(ns) * (k + static_cast<int64_t>(K) * sg); };
    int n = 0;
    for (int k = 0; k < K; ++k)
        if (coef64[q(k)] != 0.0) {
            pcls[static_cast<int64_t>(idx) * K + k] = n;
            lidx[static_cast<int64_t>(idx) * K + n++] = k;
        }
    nreal[idx] = n;
    for (int k = 0; k < K; ++k) {
        if (coef64[q(k)] != 0.0) continue;
        int rep = -1;
        for (int r = nreal[idx]; r < n; ++r)
            if (alpha64[q(lidx[static_cast<int64_t>(idx) * K + r])] == alpha64[q(k)]) rep = r;
        if (rep < 0) {
            rep = n;
            lidx[static_cast<int64_t>(idx) * K + n++] = k;
        }
        pcls[static_cast<int64_t>(idx) * K + k] = rep;
    }
    nlist[idx] = n;
    // (coef, alpha) in list order: the pullback reads them without the index indirection
    for (int r = 0; r < n; ++r) {
        const int64_t qk = q(lidx[static_cast<int64_t>(idx) * K + r]);
        lcoef[static_cast<int64_t>(idx) * K + r] = coef64[qk];
        lalpha[static_cast<int64_t>(idx) * K + r] = alpha64[qk];
    }
}

template <typename T>
__device__ __forceinline__ double exp_neg3(double x);
template <>
__device__ __forceinline__ double exp_neg3<float>(double x)
{
    const double y = x * 1.4426950408889634074;  // log2(e)
    const double n = rint(y);
    return static_cast<double>(ldexpf(__builtin_amdgcn_exp2f(-static_cast<float>(y - n)), -static_cast<int>(n)));
}
template <>
__device__ __forceinline__ double exp_neg3<double>(double x)
{
    return exp(-x);
}

// pi^{3/2} q^{-3/2} and 1/q.  FP32 kernels: 1/q from v_rcp_f64 and two Newton steps
// (FP64 accuracy, not correctly rounded), q^{-1/2} in FP32 (the prefactor multiplies FP32
// terms); FP64 kernels: correctly rounded division and square root.
template <typename T>
__device__ __forceinline__ double inv_q(double q, double &w0);
template <>
__device__ __forceinline__ double inv_q<float>(double q, double &w0)
{
    double r = __builtin_amdgcn_rcp(q);
    r = fma(fma(-q, r, 1.0), r, r);
    r = fma(fma(-q, r, 1.0), r, r);
    w0 = 5.568327996831707845 * r * static_cast<double>(__builtin_sqrtf(static_cast<float>(r)));
    return r;
}
template <>
__device__ __forceinline__ double inv_q<double>(double q, double &w0)
{
    const double r = 1.0 / q;
    w0 = 5.568327996831707845 * r * sqrt(r);
    return r;
}

template <typename T, int LA, int LB, int LC>
__device__ __forceinline__ void triple_bwd(const OvlShape &sh, int sa, int sb, int sc, const int (&sg)[3],
                                           const double *__restrict__ lcoef, const double *__restrict__ lalpha,
                                           const int *__restrict__ nreal, const int *__restrict__ nlist,
                                           const Geo3 &geo, const T (&ab)[3],
                                           const T (&ac)[3], const T *__restrict__ g, double *__restrict__ rA,
                                           double *__restrict__ rB, double *__restrict__ rC, int64_t rs,
                                           double *__restrict__ cacc, double *__restrict__ cval, float *__restrict__ caccf,
                                           float *__restrict__ cvalf, T *__restrict__ gls)
{
    constexpr int NBA = Sh3<LA>::nbf, NBB = Sh3<LB>::nbf;
    constexpr int N1 = 2 * LA + 1, N2 = 2 * LB + 1;
    const int ns = sh.ns, K = sh.K, N = sh.N;
    const int ia = sg[0] * ns + sa, ib = sg[1] * ns + sb, ic = sg[2] * ns + sc;
    const int nra = nreal[ia], nrb = nreal[ib], nrc = nreal[ic];
    const int nla = nlist[ia], nlb = nlist[ib], nlc = nlist[ic];
    // small blocks: the cotangent block of this shell triple in registers, read once
    constexpr int NBC = Sh3<LC>::nbf, NG = NBA * NBB * NBC;
    constexpr bool GREG = NG <= 27;
    T gr[GREG ? NG : 1];
    // larger blocks: staged once per shell triple into the lane's LDS column (element e at
    // gls[64 e]) instead of re-read from global memory per primitive triple, where 64 lanes'
    // blocks 13.5 KB apart do not stay in the caches (C4a two-d triples 769 -> 255 us; the
    // (d, p, p) triples, the only one-d blocks above 27 values, run in their own launch so
    // the other one-d triples keep the occupancy the staging area would cost)
    // (FP64 (d, d, d) blocks stay in global memory: 110 KB of staging plus the 2 KB x K
    // accumulator columns would not fit the 160 KB LDS for K > 24)
    constexpr bool GLDS = !GREG && (std::is_same<T, float>::value || NG <= 108);
    // FP32 instance: prefactor, exponential and per-k sums in FP32 (C4a rows 3.56 -> 3.27 ms),
    // and the C-side sums and C-list values in LDS in FP32 (half the LDS traffic of FP64
    // ones: 3.27 -> 3.17 ms); FP32 error against FP64 9.8e-8 -> 1.2e-7 of the largest
    // cotangent.  The (i, j) sums and the rows stay FP64.
    constexpr bool F32P = std::is_same<T, float>::value;
    constexpr bool F32L = F32P;
    if constexpr (GREG || GLDS) {
#pragma unroll
        for (int kk = 0; kk < NBC; ++kk)
#pragma unroll
            for (int jj = 0; jj < NBB; ++jj)
#pragma unroll
                for (int ii = 0; ii < NBA; ++ii) {
                    const T v = g[(sh.off[sa] + ii) +
                                  static_cast<int64_t>(N) * ((sh.off[sb] + jj) + static_cast<int64_t>(N) * (sh.off[sc] + kk))];
                    if constexpr (GREG)
                        gr[ii + NBA * (jj + NBB * kk)] = v;
                    else
                        gls[64 * (ii + NBA * (jj + NBB * kk))] = v;
                }
    }
    // bra sums over (j, k) and ket-B sums over k in registers, ket-C sums over (i, j) in the
    // lane's LDS column (cacc[(2 k + c) * 64]); one row update each
    // the C shell's (coef, alpha) list in the lane's LDS column too (cval[(2 k + c) * 64])
    for (int k = 0; k < nlc; ++k) {
        if constexpr (F32L) {
            caccf[64 * (2 * k)] = 0.f;
            caccf[64 * (2 * k + 1)] = 0.f;
            cvalf[64 * (2 * k)] = static_cast<float>(lcoef[static_cast<int64_t>(ic) * K + k]);
            cvalf[64 * (2 * k + 1)] = static_cast<float>(lalpha[static_cast<int64_t>(ic) * K + k]);
        } else {
            cacc[64 * (2 * k)] = 0.0;
            cacc[64 * (2 * k + 1)] = 0.0;
            cval[64 * (2 * k)] = lcoef[static_cast<int64_t>(ic) * K + k];
            cval[64 * (2 * k + 1)] = lalpha[static_cast<int64_t>(ic) * K + k];
        }
    }
    for (int i = 0; i < nla; ++i) {
        double sA0 = 0.0, sA1 = 0.0;
        const double ca = lcoef[static_cast<int64_t>(ia) * K + i], al = lalpha[static_cast<int64_t>(ia) * K + i];
        for (int j = 0; j < nlb; ++j) {
            double sB0 = 0.0, sB1 = 0.0;
            const double cb = lcoef[static_cast<int64_t>(ib) * K + j], be = lalpha[static_cast<int64_t>(ib) * K + j];
            const double p = al + be, ip = 1.0 / p, eab = al * be * geo.rab;
            const T aop = static_cast<T>(al * ip), bop = static_cast<T>(be * ip);
            // FP32 instance: the per-(i, j) pieces of the exponent argument (times log2 e) and
            // per-k FP32 sums t0 = sum_k cc w H, ta = sum_k cc w Ha, tb = sum_k cc w Hb
            [[maybe_unused]] float pf = 0.f, E0 = 0.f, E1 = 0.f, t0 = 0.f, ta = 0.f, tb = 0.f;
            const double cab = ca * cb;
            [[maybe_unused]] const float cabf = static_cast<float>(cab);
            if constexpr (F32P) {
                pf = static_cast<float>(p);
                E0 = static_cast<float>(eab * 1.4426950408889634074);
                E1 = static_cast<float>((al * geo.rac + be * geo.rbc) * 1.4426950408889634074);
            }
            for (int k = 0; k < nlc; ++k) {
                if ((i >= nra) + (j >= nrb) + (k >= nrc) > 1) continue;
                double cc, ga;
                if constexpr (F32L) {
                    cc = cvalf[64 * (2 * k)];
                    ga = cvalf[64 * (2 * k + 1)];
                } else {
                    cc = cval[64 * (2 * k)];
                    ga = cval[64 * (2 * k + 1)];
                }
                T gq, pq, i2q;
                [[maybe_unused]] double w = 0.0;
                [[maybe_unused]] float wf = 0.f;
                if constexpr (F32P) {
                    // pullback terms are summed over the batch and judged against the largest
                    // cotangent, so 1/q, q^{-3/2} and exp(-(ab rab + ag rac + bg rbc)/q) in FP32
                    // suffice here (the forward keeps its FP64 argument for per-block accuracy)
                    const float gaf = static_cast<float>(ga);
                    const float iqf = __builtin_amdgcn_rcpf(pf + gaf);
                    wf = 5.568327997f * iqf * __builtin_sqrtf(iqf) *
                         __builtin_amdgcn_exp2f(-(fmaf(gaf, E1, E0) * iqf));
                    gq = gaf * iqf;
                    pq = pf * iqf;
                    i2q = 0.5f * iqf;
                } else {
                    double w0;
                    const double iq = inv_q<T>(p + ga, w0);
                    // exp(-(ab rab + ag rac + bg rbc) / q): the argument in FP64, the
                    // exponential in the kernel precision
                    w = w0 * exp_neg3<T>((eab + ga * (al * geo.rac + be * geo.rbc)) * iq);
                    gq = static_cast<T>(ga * iq);
                    pq = static_cast<T>(p * iq);
                    i2q = static_cast<T>(0.5 * iq);
                }
                T sx[LA + 3][LB + 3][LC + 3], sy[LA + 3][LB + 3][LC + 3], sz[LA + 3][LB + 3][LC + 3];
                {
                    const T PA = -bop * ab[0], PB = aop * ab[0], PC = PA + ac[0];
                    os3<T, LA + 2, LB + 2, LC + 2>(PA - gq * PC, PB - gq * PC, pq * PC, i2q, sx);
                }
                {
                    const T PA = -bop * ab[1], PB = aop * ab[1], PC = PA + ac[1];
                    os3<T, LA + 2, LB + 2, LC + 2>(PA - gq * PC, PB - gq * PC, pq * PC, i2q, sy);
                }
                {
                    const T PA = -bop * ab[2], PB = aop * ab[2], PC = PA + ac[2];
                    os3<T, LA + 2, LB + 2, LC + 2>(PA - gq * PC, PB - gq * PC, pq * PC, i2q, sz);
                }
                T H = T(0), Ha = T(0), Hb = T(0), Hc = T(0);
                // Factored over the x exponents: for each (lx_a, lx_b, lx_c) the y-z sums
                // t = sum g s_y s_z and, per raised centre c, ty_c = sum g s_y' s_z,
                // tz_c = sum g s_y s_z' (' = that centre's exponent raised by 2) run over the
                // elements sharing those x exponents; then H += s_x t and
                // H_c += s_x' t + s_x (ty_c + tz_c): 2 mul + 7 fma per element instead of about
                // 30 operations (C4a pullback 4.29 -> 4.18 ms, FP32 error unchanged).
                int base1 = 0;
#pragma unroll
                for (int ll1 = LA; ll1 >= 0; --ll1) {
                    int base2 = 0;
#pragma unroll
                    for (int ll2 = LB; ll2 >= 0; --ll2) {
                        int base3 = 0;
#pragma unroll
                        for (int ll3 = LC; ll3 >= 0; --ll3) {
                            T t = T(0), tya = T(0), tyb = T(0), tyc = T(0), tza = T(0), tzb = T(0), tzc = T(0);
#pragma unroll
                            for (int n1 = 0; n1 <= LA - ll1; ++n1) {
                                const int m1 = LA - ll1 - n1;
#pragma unroll
                                for (int n2 = 0; n2 <= LB - ll2; ++n2) {
                                    const int m2 = LB - ll2 - n2;
#pragma unroll
                                    for (int n3 = 0; n3 <= LC - ll3; ++n3) {
                                        const int m3 = LC - ll3 - n3;
                                        const int lin = (base1 + n1) + N1 * (base2 + n2) + N1 * N2 * (base3 + n3);
                                        const int ii = lin % NBA, jj = (lin / NBA) % NBB, kk = lin / (NBA * NBB);
                                        T gv;
                                        if constexpr (GREG)
                                            gv = gr[ii + NBA * (jj + NBB * kk)];
                                        else if constexpr (GLDS)
                                            gv = gls[64 * (ii + NBA * (jj + NBB * kk))];
                                        else
                                            gv = g[(sh.off[sa] + ii) +
                                                   static_cast<int64_t>(N) *
                                                       ((sh.off[sb] + jj) + static_cast<int64_t>(N) * (sh.off[sc] + kk))];
                                        const T y = sy[m1][m2][m3], z = sz[n1][n2][n3];
                                        const T u = gv * z, v = gv * y;
                                        t = fma(u, y, t);
                                        tya = fma(u, sy[m1 + 2][m2][m3], tya);
                                        tyb = fma(u, sy[m1][m2 + 2][m3], tyb);
                                        tyc = fma(u, sy[m1][m2][m3 + 2], tyc);
                                        tza = fma(v, sz[n1 + 2][n2][n3], tza);
                                        tzb = fma(v, sz[n1][n2 + 2][n3], tzb);
                                        tzc = fma(v, sz[n1][n2][n3 + 2], tzc);
                                    }
                                }
                            }
                            const T x = sx[ll1][ll2][ll3];
                            H = fma(x, t, H);
                            Ha = fma(sx[ll1 + 2][ll2][ll3], t, fma(x, tya + tza, Ha));
                            Hb = fma(sx[ll1][ll2 + 2][ll3], t, fma(x, tyb + tzb, Hb));
                            Hc = fma(sx[ll1][ll2][ll3 + 2], t, fma(x, tyc + tzc, Hc));
                            base3 += LC - ll3 + 1;
                        }
                        base2 += LB - ll2 + 1;
                    }
                    base1 += LA - ll1 + 1;
                }
                if constexpr (F32P) {
                    const float wc = static_cast<float>(cc) * wf;
                    t0 = fmaf(wc, H, t0);
                    ta = fmaf(wc, Ha, ta);
                    tb = fmaf(wc, Hb, tb);
                    if constexpr (F32L) {
                        caccf[64 * (2 * k)] = fmaf(cabf, wf * H, caccf[64 * (2 * k)]);
                        caccf[64 * (2 * k + 1)] = fmaf(-cabf, wc * Hc, caccf[64 * (2 * k + 1)]);
                    } else {
                        cacc[64 * (2 * k)] += cab * static_cast<double>(wf * H);
                        cacc[64 * (2 * k + 1)] -= cab * static_cast<double>(wc * Hc);
                    }
                } else {
                    const double wH = w * static_cast<double>(H);
                    const double cw = cab * cc * w;
                    sA0 += cb * cc * wH;
                    sB0 += ca * cc * wH;
                    cacc[64 * (2 * k)] += cab * wH;
                    sA1 -= cw * static_cast<double>(Ha);
                    sB1 -= cw * static_cast<double>(Hb);
                    cacc[64 * (2 * k + 1)] -= cw * static_cast<double>(Hc);
                }
            }
            if constexpr (F32P) {
                sA0 += cb * static_cast<double>(t0);
                sB0 += ca * static_cast<double>(t0);
                sA1 -= cab * static_cast<double>(ta);
                sB1 -= cab * static_cast<double>(tb);
            }
            rB[(2 * j) * rs] += sB0;
            rB[(2 * j + 1) * rs] += sB1;
        }
        rA[(2 * i) * rs] += sA0;
        rA[(2 * i + 1) * rs] += sA1;
    }
    for (int k = 0; k < 2 * nlc; ++k) rC[k * rs] += F32L ? static_cast<double>(caccf[64 * k]) : cacc[64 * k];
}

// Class of a shell triple (shells up to d): its number of d shells, except that the
// one-d triples whose cotangent block exceeds 27 values (d, p, p) form class 4, so that
// only their launch carries the LDS staging area.
__host__ __device__ constexpr int triple_class(int la, int lb, int lc)
{
    const int nd = (la == 2) + (lb == 2) + (lc == 2);
    const int ng = (la + 1) * (la + 2) / 2 * ((lb + 1) * (lb + 2) / 2) * ((lc + 1) * (lc + 2) / 2);
    return nd == 1 && ng > 27 ? 4 : nd;
}

// CLS >= 0: only the shell triples of class CLS (one instance per class, so that each is
// compiled for the registers of its own cases; CLS 0 clears the rows and runs first),
// CLS < 0: every triple.
// Waves per SIMD the FP32 instances are compiled for: 2 for the triples with no d shell
// (224 VGPRs), with one d (189) and for the (d, p, p) class (248); forcing 3 on the
// no-d class (212-byte spill) or the one-d class (84-byte spill: 852 -> 967 us), or 2 on
// the two-d class measured slower.  Other instances keep the compiler's choice.
template <typename T, int LMAXB, int CLS>
constexpr int ovl3_pb_waves()
{
    return std::is_same<T, float>::value && LMAXB == 2 && (CLS == 0 || CLS == 1 || CLS == 4) ? 2 : 1;
}

template <typename T, int LMAXB, int CLS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(ovl3_pb_waves<T, LMAXB, CLS>(), 8))) void ovl3c_bwd_rows(const OvlShape sh, int64_t b0, int64_t Bc,
                                                     const T *__restrict__ dT,
                                                     const T *__restrict__ posA, const T *__restrict__ posB,
                                                     const T *__restrict__ posC, T scale, const void *__restrict__ sA,
                                                     const void *__restrict__ sB, const void *__restrict__ sC,
                                                     int itype, const double *__restrict__ lcoef,
                                                     const double *__restrict__ lalpha,
                                                     const int *__restrict__ nreal, const int *__restrict__ nlist,
                                                     int splitB, int splitC,
                                                     double *__restrict__ rows, unsigned *__restrict__ err)
{
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ, N = sh.N;
    const int64_t bl = static_cast<int64_t>(blockIdx.x) * 64 + threadIdx.x;
    if (bl >= Bc) return;
    const int64_t b = b0 + bl;
    // part p = pc * nbp + pb, last parts (shells listed last, usually the largest l) first
    const int nbp = splitB ? ns : 1;
    const int p = static_cast<int>(gridDim.y) - 1 - static_cast<int>(blockIdx.y);
    const int pc = p / nbp, pb = p - pc * nbp;
    const int nB = splitB ? 1 : ns, nC = splitC ? 1 : ns;
    const int64_t rlen = 2 * static_cast<int64_t>(K) * (ns + nB + nC);
    double *row = rows + static_cast<int64_t>(p) * rlen * Bc + bl;
    if constexpr (CLS <= 0)
        for (int64_t q = 0; q < rlen; ++q) row[q * Bc] = 0.0;
    int sg[3] = {load_sigma(sA, itype, b), load_sigma(sB, itype, b), load_sigma(sC, itype, b)};
    for (int c = 0; c < 3; ++c)
        if (sg[c] < 0 || sg[c] >= NZ) {
            *err = ERR_SPECIES;
            return;
        }
    T A[3], Bv[3], C[3];
    Geo3 geo{0.0, 0.0, 0.0};
    T ab[3], ac[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        A[ax] = scale * posA[3 * b + ax];
        Bv[ax] = scale * posB[3 * b + ax];
        C[ax] = scale * posC[3 * b + ax];
        const double dab = static_cast<double>(A[ax]) - static_cast<double>(Bv[ax]);
        const double dac = static_cast<double>(A[ax]) - static_cast<double>(C[ax]);
        const double dbc = static_cast<double>(Bv[ax]) - static_cast<double>(C[ax]);
        geo.rab += dab * dab;
        geo.rac += dac * dac;
        geo.rbc += dbc * dbc;
        ab[ax] = static_cast<T>(dab);
        ac[ax] = static_cast<T>(dac);
    }
    const T *g = dT + static_cast<int64_t>(N) * N * N * b;
    extern __shared__ double cacc_all[];  // [2 K][64] sums, [2 K][64] C-list values, [NG][64] T cotangents
    double *cacc = cacc_all + threadIdx.x, *cval = cacc_all + 64 * 2 * K + threadIdx.x;
    float *caccf = reinterpret_cast<float *>(cacc_all) + threadIdx.x;  // FP32 layout of the same area
    float *cvalf = reinterpret_cast<float *>(cacc_all) + 64 * 2 * K + threadIdx.x;
    T *gls = reinterpret_cast<T *>(cacc_all + 64 * 4 * K) + threadIdx.x;
    const int sc0 = splitC ? pc : 0, sc1 = splitC ? pc + 1 : ns;
    const int sb0 = splitB ? pb : 0, sb1 = splitB ? pb + 1 : ns;
    const int64_t K2 = 2 * static_cast<int64_t>(K);
    for (int sc = sc0; sc < sc1; ++sc)
        for (int sb = sb0; sb < sb1; ++sb)
            for (int sa = 0; sa < ns; ++sa) {
                const int la = sh.ls[sa], lb = sh.ls[sb], lc = sh.ls[sc];
                if constexpr (CLS >= 0)
                    if (triple_class(la, lb, lc) != CLS) continue;
                double *rA = row + K2 * sa * Bc;
                double *rB = row + K2 * (ns + (splitB ? 0 : sb)) * Bc;
                double *rC = row + K2 * (ns + nB + (splitC ? 0 : sc)) * Bc;
#define AOK_B3_CASE(A_, B_, C_)                                                                                  \
    if constexpr (A_ <= LMAXB && B_ <= LMAXB && C_ <= LMAXB &&                                                   \
                  (CLS < 0 || triple_class(A_, B_, C_) == CLS))                                                  \
        if (la == A_ && lb == B_ && lc == C_) {                                                                  \
            triple_bwd<T, A_, B_, C_>(sh, sa, sb, sc, sg, lcoef, lalpha, nreal, nlist, geo, ab, ac, g, rA, rB,  \
                                      rC, Bc, cacc, cval, caccf, cvalf, gls);                                     \
            continue;                                                                                            \
        }
                AOK_L3_CASES(AOK_B3_CASE)
#undef AOK_B3_CASE
            }
}

// Fold the parts' partial rows into one row per batch element, in the unsplit layout
// ([side][shell][list entry][coef, exponent], batch innermost): thread per (b, element),
// a fixed-order sum over the parts that hold the element.  Fully parallel, so the
// per-slot reduction below then reads one row per element instead of one per part.
__global__ __launch_bounds__(256) void ovl3_parts_sum(const OvlShape sh, int64_t B, int64_t b0, int64_t Bc,
                                                      int splitB, int splitC, const double *__restrict__ rows,
                                                      double *__restrict__ rows1)
{
    const int64_t bl = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (bl >= Bc) return;
    const int ns = sh.ns;
    const int K2 = 2 * sh.K, E = 3 * ns * K2;
    const int nbp = splitB ? ns : 1, ncp = splitC ? ns : 1;
    const int nB = splitB ? 1 : ns, nC = splitC ? 1 : ns;
    const int64_t rlen = static_cast<int64_t>(K2) * (ns + nB + nC);
    for (int e1 = blockIdx.y; e1 < E; e1 += gridDim.y) {
        const int side = e1 / (ns * K2), r = e1 - side * ns * K2, s = r / K2, w = r - s * K2;
        const int64_t e = side == 0   ? static_cast<int64_t>(K2) * s + w
                          : side == 1 ? static_cast<int64_t>(K2) * (ns + (splitB ? 0 : s)) + w
                                      : static_cast<int64_t>(K2) * (ns + nB + (splitC ? 0 : s)) + w;
        const int c0 = side == 2 && splitC ? s : 0, c1 = side == 2 && splitC ? s + 1 : ncp;
        const int p0 = side == 1 && splitB ? s : 0, p1 = side == 1 && splitB ? s + 1 : nbp;
        double acc = 0.0;
        for (int pc = c0; pc < c1; ++pc)
            for (int pb = p0; pb < p1; ++pb) acc += rows[(static_cast<int64_t>(pc * nbp + pb) * rlen + e) * Bc + bl];
        rows1[static_cast<int64_t>(e1) * B + b0 + bl] = acc;
    }
}

// Block per slot (sigma, s, k): sum over the batch of the unsplit rows (parts already
// folded) whose species on that side is sigma, sides A, B, C; thread-strided partial
// sums in a fixed order, then a block tree.
template <typename T>
__global__ __launch_bounds__(256) void ovl3_bwd_reduce(const OvlShape sh, int64_t B, const void *__restrict__ sA,
                                                       const void *__restrict__ sB, const void *__restrict__ sC,
                                                       int itype, const int *__restrict__ pcls,
                                                       const double *__restrict__ rows, T *__restrict__ dcoef,
                                                       T *__restrict__ dalpha)
{
    __shared__ double red[8];
    const int ns = sh.ns, K = sh.K;
    const int q = blockIdx.x;  // [s + ns * (k + K * sg)], the Julia index
    const int s = q % ns, k = (q / ns) % K, sg = q / (ns * K);
    const int li = pcls[static_cast<int64_t>(sg * ns + s) * K + k];
    const int64_t K2 = 2 * static_cast<int64_t>(K);
    const void *sides[3] = {sA, sB, sC};
    double c = 0.0, a = 0.0;
    for (int64_t b = threadIdx.x; b < B; b += 256)
        for (int side = 0; side < 3; ++side) {
            if (load_sigma(sides[side], itype, b) != sg) continue;
            const double *r = rows + (K2 * (side * ns + s) + 2 * li) * B + b;
            c += r[0];
            a += r[B];
        }
    block_sum256x2(c, a, red);
    if (threadIdx.x == 0) {
        dcoef[q] = static_cast<T>(c);
        dalpha[q] = static_cast<T>(a);
    }
}

// ---------------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------------
template <typename T, int LMAXB>
static aok_status fast3_launch(aok_ctx ctx, const Ovl3cCall &c, const OvlShape &sh)
{
    using R = typename Rec3Of<T>::type;
    const Ovl3Ws w = ovl3_ws_map(c.B, sh.ns, sh.K, sh.NZ, sizeof(R));
    aok_status st = ensure_ws(ctx, w.total);
    if (st != AOK_OK) return st;
    char *ws = static_cast<char *>(ctx->ws);
    hipLaunchKernelGGL(ovl3_hist_kernel, dim3(w.nblk_h), dim3(256), 0, ctx->stream, c.sidxA, c.sidxB, c.sidxC,
                       c.itype, c.B, sh.NZ, reinterpret_cast<int *>(ws + w.rank),
                       reinterpret_cast<int *>(ws + w.blk_hist), ctx->d_err);
    AOK_HIP(ctx, hipGetLastError());
    // light shells: those the l <= 1 instance takes (all shells when the basis has no d)
    const int light_l = LMAXB == 2 ? 1 : LMAXB;
    int nL = 0;
    for (int q = 0; q < sh.ns; ++q) nL += sh.ls[q] <= light_l;
    hipLaunchKernelGGL((ovl3_plan_kernel<T>), dim3(1), dim3(256), 0, ctx->stream, sh,
                       static_cast<const T *>(c.coef), static_cast<const T *>(c.alpha),
                       reinterpret_cast<double *>(ws + w.coef64), reinterpret_cast<double *>(ws + w.alpha64),
                       reinterpret_cast<int *>(ws + w.nreal), reinterpret_cast<int *>(ws + w.lidx),
                       reinterpret_cast<int64_t *>(ws + w.grp_off), w.nblk_h,
                       reinterpret_cast<const int *>(ws + w.blk_hist), reinterpret_cast<int *>(ws + w.blk_off),
                       reinterpret_cast<int *>(ws + w.bcnt), reinterpret_cast<int *>(ws + w.boff),
                       reinterpret_cast<int *>(ws + w.toff), reinterpret_cast<int *>(ws + w.tri), light_l);
    AOK_HIP(ctx, hipGetLastError());
    const int nblk_tile = (w.maxtiles + 255) / 256;
    const int nblk_rec = static_cast<int>((w.maxrec + 255) / 256);
    hipLaunchKernelGGL((ovl3_fill_kernel<T>), dim3(w.nblk_h + nblk_tile + nblk_rec), dim3(256), 0, ctx->stream, sh,
                       c.sidxA, c.sidxB, c.sidxC, c.itype, c.B, w.nblk_h, nblk_tile,
                       reinterpret_cast<const int *>(ws + w.rank), reinterpret_cast<const int *>(ws + w.blk_off),
                       reinterpret_cast<const int *>(ws + w.toff), reinterpret_cast<int *>(ws + w.perm),
                       reinterpret_cast<int *>(ws + w.tile_key), w.maxtiles,
                       reinterpret_cast<const double *>(ws + w.coef64), reinterpret_cast<const double *>(ws + w.alpha64),
                       reinterpret_cast<const int *>(ws + w.nreal), reinterpret_cast<const int *>(ws + w.lidx),
                       reinterpret_cast<const int64_t *>(ws + w.grp_off), reinterpret_cast<R *>(ws + w.rec));
    AOK_HIP(ctx, hipGetLastError());
    // With d shells, the triples of s and p shells alone run in the l <= 1 instance (64
    // VGPRs, 8 waves/SIMD) and the rest in the l <= 2 one (its (d, d, d) case holds 216
    // accumulators: 250 VGPRs, 2 waves/SIMD).  Measured on C4a: one launch per number of d
    // shells (126 / 170 / 250 VGPRs) 0.44-0.45 ms against 0.43 ms for these two.
    time_begin(ctx, "ovl3c_fwd_fast");
    const int ns3 = sh.ns * sh.ns * sh.ns, nL3 = nL * nL * nL;
    auto launch = [&](auto kern, int t0, int count) {
        if (count <= 0) return;
        // XCD-affine, triple-fastest 1-D grid (C4a: 0.423 -> 0.409 ms, writes 778 -> 586 MB for
        // the 222 MB output; what remains are lines shared by the light and heavy launches).
        // maxtiles * count < 2^31 for every output that fits in device memory (N >= ns, so
        // the output holds at least 64 * ns^3 floats per tile); checked in ovl3c_run.
        hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(w.maxtiles * count)), dim3(OVL_TILE), 0, ctx->stream, sh,
                           static_cast<const T *>(c.posA), static_cast<const T *>(c.posB),
                           static_cast<const T *>(c.posC), static_cast<T>(c.scale),
                           reinterpret_cast<const int *>(ws + w.perm), reinterpret_cast<const int *>(ws + w.tile_key),
                           reinterpret_cast<const int *>(ws + w.boff), reinterpret_cast<const int *>(ws + w.bcnt),
                           reinterpret_cast<const int *>(ws + w.toff), reinterpret_cast<const R *>(ws + w.rec),
                           reinterpret_cast<const int64_t *>(ws + w.grp_off),
                           reinterpret_cast<const int *>(ws + w.nreal), reinterpret_cast<const int *>(ws + w.tri), t0,
                           w.maxtiles / 8, static_cast<T *>(c.out));
    };
    if constexpr (LMAXB == 2) {
        launch(&ovl3c_fwd_fast<T, 1>, 0, nL3);
        launch(&ovl3c_fwd_fast<T, 2>, nL3, ns3 - nL3);
    } else {
        launch(&ovl3c_fwd_fast<T, LMAXB>, 0, ns3);
    }
    time_end(ctx);
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

template <typename T, int LMAXP>
static aok_status plain3_launch(aok_ctx ctx, const Ovl3cCall &c, const OvlShape &sh)
{
    const int64_t items = c.B * sh.ns * sh.ns * sh.ns;
    if (items == 0) return AOK_OK;
    time_begin(ctx, "ovl3c_fwd_plain");
    hipLaunchKernelGGL((ovl3c_fwd_plain<T, LMAXP>), dim3(static_cast<unsigned>((items + 127) / 128)), dim3(128), 0,
                       ctx->stream, sh, c.B, static_cast<const T *>(c.coef), static_cast<const T *>(c.alpha), c.sidxA,
                       c.sidxB, c.sidxC, c.itype, static_cast<const T *>(c.posA), static_cast<const T *>(c.posB),
                       static_cast<const T *>(c.posC), static_cast<T>(c.scale), static_cast<T *>(c.out), ctx->d_err);
    time_end(ctx);
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

template <typename T>
aok_status ovl3c_run(aok_ctx ctx, const Ovl3cCall &c)
{
    OvlShape sh;
    const std::string msg = build_shape(c.ns, c.K, c.NZ, c.ls, AOK_OVL_LMAX, 1 << 20, sh);
    if (!msg.empty()) return fail(ctx, msg.find("shell l") != std::string::npos ? AOK_ERR_LMAX : AOK_ERR_ARG, msg);
    if (c.B < 0 || c.B > (int64_t(1) << 28)) return fail(ctx, AOK_ERR_ARG, "B out of range");
    // one workgroup per (64-triple tile, shell triple) (tiles padded by up to 8 x 8 x 8 species
    // triples and to a multiple of 8), or 128 items of (triple, shell triple)
    const int64_t ns = sh.ns;
    if (((c.B + OVL_TILE - 1) / OVL_TILE + 520) * ns * ns * ns > INT32_MAX || (c.B * ns * ns * ns + 127) / 128 > INT32_MAX)
        return fail(ctx, AOK_ERR_ARG, "B x shells^3 too large for one call: split the batch");
    int lmax = 0;
    for (int q = 0; q < sh.ns; ++q) lmax = std::max<int>(lmax, sh.ls[q]);
    const bool fast = !(c.flags & AOK_PLAIN) && lmax <= OVL3_FAST_LMAX && sh.NZ <= OVL3_MAXNZ && sh.ns <= 24;
    if (!fast) {
        switch (lmax) {
            case 0: return plain3_launch<T, 0>(ctx, c, sh);
            case 1: return plain3_launch<T, 1>(ctx, c, sh);
            case 2: return plain3_launch<T, 2>(ctx, c, sh);
            case 3: return plain3_launch<T, 3>(ctx, c, sh);
            default: return plain3_launch<T, 4>(ctx, c, sh);
        }
    }
    if (c.B == 0) return AOK_OK;
    switch (lmax) {
        case 0: return fast3_launch<T, 0>(ctx, c, sh);
        case 1: return fast3_launch<T, 1>(ctx, c, sh);
        default: return fast3_launch<T, 2>(ctx, c, sh);
    }
}

// (coef, alpha) pullback: lists, per-element rows, fixed-order reduction.
template <typename T, int LMAXB>
static aok_status bwd3_launch(aok_ctx ctx, const Ovl3cCall &c, const OvlShape &sh)
{
    const size_t npar = static_cast<size_t>(sh.ns) * sh.K * sh.NZ;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    const size_t o_c = take(sizeof(double) * npar), o_a = take(sizeof(double) * npar);
    const size_t o_nr = take(sizeof(int) * sh.NZ * sh.ns), o_nl = take(sizeof(int) * sh.NZ * sh.ns);
    const size_t o_li = take(sizeof(int) * npar), o_pc = take(sizeof(int) * npar);
    const size_t o_lc = take(sizeof(double) * npar), o_la = take(sizeof(double) * npar);
    // parts: by (third, second) shell, else by third shell, whichever is the finest whose
    // split rows for min(B, 4096) triples fit OVL3_ROWS_CAP, else none; split rows are
    // processed in batch chunks that fit the cap.  The "ovl3_parts" knob forces a split
    // (1 none, 2 third, 3 third and second).
    const int64_t Bn = c.B > 0 ? c.B : 1;
    const size_t nsz = sh.ns;
    auto row_bytes = [&](int sb_, int sc_) {  // per triple
        const size_t parts = (sb_ ? nsz : 1) * (sc_ ? nsz : 1);
        return sizeof(double) * 2 * sh.K * (nsz + (sb_ ? 1 : nsz) + (sc_ ? 1 : nsz)) * parts;
    };
    auto chunk_of = [&](int sb_, int sc_) {
        return std::max<int64_t>(1, static_cast<int64_t>(OVL3_ROWS_CAP / row_bytes(sb_, sc_)));
    };
    int splitB = 1, splitC = 1;
    if (ctx->tune.ovl3_parts != 0) {
        splitC = ctx->tune.ovl3_parts >= 2;
        splitB = ctx->tune.ovl3_parts >= 3;
    } else if (chunk_of(1, 1) < std::min<int64_t>(Bn, 4096)) {
        splitB = 0;
        if (chunk_of(0, 1) < std::min<int64_t>(Bn, 4096)) splitC = 0;
    }
    const int nparts = (splitB ? sh.ns : 1) * (splitC ? sh.ns : 1);
    const int64_t Bc = nparts > 1 ? std::min<int64_t>(Bn, chunk_of(splitB, splitC)) : Bn;
    const size_t o_rows = take(row_bytes(splitB, splitC) * static_cast<size_t>(Bc));
    const size_t o_rows1 = nparts > 1 ? take(row_bytes(0, 0) * static_cast<size_t>(Bn)) : o_rows;  // parts folded
    aok_status st = ensure_ws(ctx, o);
    if (st != AOK_OK) return st;
    char *ws = static_cast<char *>(ctx->ws);
    double *c64 = reinterpret_cast<double *>(ws + o_c), *a64 = reinterpret_cast<double *>(ws + o_a);
    hipLaunchKernelGGL((widen_pair<T>), dim3(static_cast<unsigned>((npar + 255) / 256)), dim3(256), 0, ctx->stream,
                       static_cast<int64_t>(npar), static_cast<const T *>(c.coef), static_cast<const T *>(c.alpha), c64,
                       a64);
    AOK_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(ovl3_pb_lists, dim3((sh.NZ * sh.ns + 63) / 64), dim3(64), 0, ctx->stream, sh, c64, a64,
                       reinterpret_cast<int *>(ws + o_nr), reinterpret_cast<int *>(ws + o_nl),
                       reinterpret_cast<int *>(ws + o_li), reinterpret_cast<int *>(ws + o_pc),
                       reinterpret_cast<double *>(ws + o_lc), reinterpret_cast<double *>(ws + o_la));
    AOK_HIP(ctx, hipGetLastError());
    for (int64_t b0 = 0; b0 < c.B; b0 += Bc) {
        const int64_t n = std::min<int64_t>(Bc, c.B - b0);
        time_begin(ctx, "ovl3c_bwd_rows");
        // ngl: the largest cotangent block the instance stages in LDS (more than 27 values;
        // K <= 32, so the total stays within 160 KB)
        auto rows_launch = [&](auto kern, int ngl) -> aok_status {
            const size_t shm = sizeof(double) * 64 * 4 * sh.K + sizeof(T) * 64 * ngl;
            if (shm > 64 * 1024)
                AOK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm)));
            hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>((n + 63) / 64), static_cast<unsigned>(nparts)),
                               dim3(64), shm, ctx->stream, sh, b0, n,
                               static_cast<const T *>(c.dT), static_cast<const T *>(c.posA),
                               static_cast<const T *>(c.posB), static_cast<const T *>(c.posC), static_cast<T>(c.scale),
                               c.sidxA, c.sidxB, c.sidxC, c.itype, reinterpret_cast<const double *>(ws + o_lc),
                               reinterpret_cast<const double *>(ws + o_la), reinterpret_cast<const int *>(ws + o_nr),
                               reinterpret_cast<const int *>(ws + o_nl), splitB, splitC,
                               reinterpret_cast<double *>(ws + o_rows), ctx->d_err);
            return AOK_OK;
        };
        // with d shells: one launch per number of d shells in the triple (0 first: it clears)
        if constexpr (LMAXB == 2) {
            const int ngl3 = std::is_same<T, float>::value ? 216 : 0;
            aok_status rs = rows_launch(&ovl3c_bwd_rows<T, 2, 0>, 0);
            if (rs == AOK_OK) rs = rows_launch(&ovl3c_bwd_rows<T, 2, 1>, 0);
            if (rs == AOK_OK) rs = rows_launch(&ovl3c_bwd_rows<T, 2, 4>, 54);
            if (rs == AOK_OK) rs = rows_launch(&ovl3c_bwd_rows<T, 2, 2>, 108);
            if (rs == AOK_OK) rs = rows_launch(&ovl3c_bwd_rows<T, 2, 3>, ngl3);
            if (rs != AOK_OK) {
                time_end(ctx);
                return rs;
            }
        } else {
            const aok_status rs = rows_launch(&ovl3c_bwd_rows<T, LMAXB, -1>, 0);  // l <= 1: blocks in registers
            if (rs != AOK_OK) {
                time_end(ctx);
                return rs;
            }
        }
        time_end(ctx);
        AOK_HIP(ctx, hipGetLastError());
        if (nparts > 1) {
            const int E = 6 * sh.ns * sh.K;
            hipLaunchKernelGGL(ovl3_parts_sum, dim3(static_cast<unsigned>((n + 255) / 256), std::min(E, 65535)),
                               dim3(256), 0, ctx->stream, sh, c.B, b0, n, splitB, splitC,
                               reinterpret_cast<const double *>(ws + o_rows), reinterpret_cast<double *>(ws + o_rows1));
            AOK_HIP(ctx, hipGetLastError());
        }
    }
    hipLaunchKernelGGL((ovl3_bwd_reduce<T>), dim3(static_cast<unsigned>(npar)), dim3(256), 0, ctx->stream, sh, c.B,
                       c.sidxA, c.sidxB, c.sidxC, c.itype, reinterpret_cast<const int *>(ws + o_pc),
                       reinterpret_cast<const double *>(ws + o_rows1), static_cast<T *>(c.o1),
                       static_cast<T *>(c.o2));
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

template <typename T>
aok_status ovl3c_pullback_run(aok_ctx ctx, const Ovl3cCall &c)
{
    OvlShape sh;
    const std::string msg = build_shape(c.ns, c.K, c.NZ, c.ls, OVL3_FAST_LMAX, 1 << 20, sh);
    if (!msg.empty()) return fail(ctx, msg.find("shell l") != std::string::npos ? AOK_ERR_LMAX : AOK_ERR_ARG, msg);
    if (c.B < 0 || c.B > (int64_t(1) << 28)) return fail(ctx, AOK_ERR_ARG, "B out of range");
    int lmax = 0;
    for (int q = 0; q < sh.ns; ++q) lmax = std::max<int>(lmax, sh.ls[q]);
    switch (lmax) {
        case 0: return bwd3_launch<T, 0>(ctx, c, sh);
        case 1: return bwd3_launch<T, 1>(ctx, c, sh);
        default: return bwd3_launch<T, 2>(ctx, c, sh);
    }
}

template aok_status ovl3c_pullback_run<float>(aok_ctx, const Ovl3cCall &);
template aok_status ovl3c_pullback_run<double>(aok_ctx, const Ovl3cCall &);
template aok_status ovl3c_run<float>(aok_ctx, const Ovl3cCall &);
template aok_status ovl3c_run<double>(aok_ctx, const Ovl3cCall &);

}  // namespace aok
