// ovl2c_kernels.hip -- H2: batched 2-centre Cartesian-Gaussian overlaps and their
// (coef, alpha) pullback on gfx950.
//
// Semantics follow batch_S_orbital_kernel! (src/integrals/orbital/overlap_2c.jl:165-281)
// and batch_S_orbital_pb_kernel! (src/integrals/orbital/overlap_2c_grad.jl:71-214):
// Cartesian components ordered (x,y,z) = (ll, l-ll-n, n) for ll = l..0, n = 0..l-ll;
// contracted values accumulated at lin = index1 + (2 l_a + 1) index2 and read back
// as out[off_a + i, off_b + j] = blk[i + nbf_a j] (the reference's l >= 2 aliasing);
// forward skips primitive pairs with a zero coefficient, the pullback does not.
//
// Fast path (AOK_FAST): see ovl_plan.hpp.  Per primitive pair and lane the work is
//   e = exp(-mu |AB|^2) with the exponent argument formed and range-reduced in FP64
//       (y = mu log2(e) |AB|^2, y = n + f, e = 2^-n * v_exp_f32(-f)), so the FP32 result
//       carries no error proportional to the argument;
//   1-D Obara-Saika overlaps s_ij (s_00 = 1) per axis in registers, from
//       PA = -(beta/p) AB and PB = (alpha/p) AB;
//   contraction G s_x s_y s_z into compile-time-indexed accumulators.
// This is the same quantity as the MD E^{ij}_0 product (E^{ij}_0 = exp(-mu X_AB^2) s_ij).
//
// Plain path (AOK_PLAIN): the reference's own decomposition -- one work item per
// (b, s_a, s_b), runtime l loops, private E scratch, three exps per primitive pair,
// float atomics in the pullback -- kept as the baseline the fast path is measured against.
#include <mutex>

#include "aok_common.hpp"
#include "ovl_plan.hpp"
#include "ovl_prep.hpp"

namespace aok {

// ---------------------------------------------------------------------------------
// FP64 helpers (device)
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double cart_cl(int l)
{
    double df = 1.0;
    for (int k = 2 * l - 1; k > 1; k -= 2) df *= k;
    return 5.568327996831707845 * df / static_cast<double>(1 << l);  // pi^{3/2} (2l-1)!! / 2^l
}

// x^-(l + 3/2) by integer powers and one square root
__device__ __forceinline__ double inv_pow_l32(double x, int l)
{
    double t = x * sqrt(x);
    for (int q = 0; q < l; ++q) t *= x;
    return 1.0 / t;
}

// normsq of one (shell, species) contraction (cartesian_gto_basis.jl:74-83)
__device__ double cart_normsq(const double *Dv, const double *zv, int K, int l)
{
    double s = 0.0;
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j) s += Dv[i] * Dv[j] * inv_pow_l32(zv[i] + zv[j], l);
    return cart_cl(l) * s;
}

// One (species, shell) handled by a whole block: zeta, D of its K slots in FP64 and the
// fixed-order row sums
//   g[m] = sum_j D_j (zeta_m + zeta_j)^-(l+3/2),  h[m] = sum_j D_j (zeta_m + zeta_j)^-(l+5/2),
// so that normsq = c_l sum_m D_m g[m] (cartesian_gto_basis.jl:74-83) and g, h are the
// sums of the normalisation VJP (overlap_2c_grad.jl:251-286).  Terms are spread over the
// block's threads; each row is then summed by one thread in j order.
struct ShellNorm {
    double z[OVL_MAXK], d[OVL_MAXK], g[OVL_MAXK], h[OVL_MAXK];
    double tg[OVL_MAXK * OVL_MAXK], th[OVL_MAXK * OVL_MAXK];
    double nsq;
};

template <typename T>
__device__ void shell_norm(const T *__restrict__ zetaT, const T *__restrict__ DT, int64_t q0, int64_t qstep, int K,
                           int l, bool want_h, ShellNorm &S)
{
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        S.z[k] = static_cast<double>(zetaT[q0 + qstep * k]);
        S.d[k] = static_cast<double>(DT[q0 + qstep * k]);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < K * K; t += blockDim.x) {
        const int m = t / K, j = t - m * K;
        const double x = S.z[m] + S.z[j];
        const double pe = inv_pow_l32(x, l);
        S.tg[t] = S.d[j] * pe;
        if (want_h) S.th[t] = S.d[j] * pe / x;
    }
    __syncthreads();
    for (int m = threadIdx.x; m < K; m += blockDim.x) {
        double g = 0.0, h = 0.0;
        for (int j = 0; j < K; ++j) {
            g += S.tg[m * K + j];
            if (want_h) h += S.th[m * K + j];
        }
        S.g[m] = g;
        S.h[m] = h;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int m = 0; m < K; ++m) a += S.d[m] * S.g[m];
        S.nsq = cart_cl(l) * a;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------
// Prep 1: species-pair keys, stable in-block ranks, per-block histograms.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void ovl_hist_kernel(const void *__restrict__ sidxA, const void *__restrict__ sidxB,
                                                       int itype, int64_t B, int NZ, int *__restrict__ rank,
                                                       int *__restrict__ blk_hist, unsigned *__restrict__ err)
{
    __shared__ int whist[4 * OVL_MAXNZ * OVL_MAXNZ];
    const int NK = NZ * NZ;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    int key = -1;
    if (i < B) {
        int sa = load_sigma(sidxA, itype, i), sb = load_sigma(sidxB, itype, i);
        if (sa < 0 || sa >= NZ || sb < 0 || sb >= NZ) {
            *err = ERR_SPECIES;
            sa = sa < 0 || sa >= NZ ? 0 : sa;
            sb = sb < 0 || sb >= NZ ? 0 : sb;
        }
        key = sa * NZ + sb;
    }
    const int r = block_rank(key, NK, whist, blk_hist + static_cast<int64_t>(blockIdx.x) * NK);
    if (i < B) rank[i] = r;
}

// ---------------------------------------------------------------------------------
// Prep 2 (one block): primitive lists per (species, shell) and bucket offsets.
//   mode 0: coef given (cast to FP64); mode 1: coef = D / sqrt(normsq) from (zeta, D).
// List of a (species, shell): its real primitives (coef != 0) in slot order, then one
// representative per distinct exponent among its padded (coef == 0) slots; pcls maps
// every slot to its list entry.  Padded slots with the same exponent have identical
// cotangents, so the pullback computes each once.
// ---------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(1024) void ovl_plan_kernel(const OvlShape sh, int mode, const T *__restrict__ coefT,
                                                       const T *__restrict__ zetaT, const T *__restrict__ DT,
                                                       double *__restrict__ coef64, double *__restrict__ alpha64,
                                                       int *__restrict__ nreal, int *__restrict__ nlist,
                                                       int *__restrict__ lidx, int *__restrict__ pcls, int nblk_h,
                                                       const int *__restrict__ blk_hist, int *__restrict__ blk_off,
                                                       int *__restrict__ bcnt, int *__restrict__ boff,
                                                       int *__restrict__ toff)
{
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ;
    if (blockIdx.x == 0) {  // block 0: bucket offsets of the batch
        bucket_offsets(NZ * NZ, nblk_h, blk_hist, blk_off, bcnt, boff, toff);
        return;
    }
    // blocks 1..NZ*ns: one (species, shell) each
    __shared__ ShellNorm S;
    __shared__ double cv[OVL_MAXK];
    const int idx = blockIdx.x - 1;
    const int sg = idx / ns, s = idx - sg * ns;
    const int64_t q0 = s + static_cast<int64_t>(ns) * K * sg;
    if (mode == 1) {
        shell_norm(zetaT, DT, q0, ns, K, sh.ls[s], false, S);
        const double nsq = S.nsq;
        const double inv = nsq > 0.0 ? 1.0 / sqrt(nsq) : 0.0;
        for (int k = threadIdx.x; k < K; k += blockDim.x) cv[k] = nsq > 0.0 ? S.d[k] * inv : 0.0;
    } else {
        for (int k = threadIdx.x; k < K; k += blockDim.x) {
            S.z[k] = static_cast<double>(zetaT[q0 + static_cast<int64_t>(ns) * k]);
            cv[k] = static_cast<double>(coefT[q0 + static_cast<int64_t>(ns) * k]);
        }
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    // one wave builds the primitive list: real slots in slot order, then the first slot
    // of every distinct padded exponent (K <= 32 lanes)
    const int k = threadIdx.x;
    const bool valid = k < K;
    // the forward's skip test is made in the kernel precision (overlap_2c.jl:202)
    const bool real = valid && static_cast<T>(cv[k]) != T(0);
    const unsigned long long mr = __ballot(real);
    int first = k;
    if (valid && !real)
        for (int k2 = 0; k2 < k; ++k2)
            if (!((mr >> k2) & 1ull) && S.z[k2] == S.z[k]) {
                first = k2;
                break;
            }
    const bool rep = valid && !real && first == k;
    const unsigned long long mp = __ballot(rep);
    const int nr = __popcll(mr);
    if (valid) {
        const int64_t q = q0 + static_cast<int64_t>(ns) * k;
        coef64[q] = cv[k];
        alpha64[q] = S.z[k];
        const int pos = real ? __popcll(mr & ((1ull << k) - 1ull)) : nr + __popcll(mp & ((1ull << first) - 1ull));
        pcls[static_cast<int64_t>(idx) * K + k] = pos;
        if (real || rep) lidx[static_cast<int64_t>(idx) * K + pos] = k;
    }
    if (k == 0) {
        nreal[idx] = nr;
        nlist[idx] = nr + __popcll(mp);
    }
}

// ---------------------------------------------------------------------------------
// Prep 3 (grid, three roles): stable scatter of b into bucket order, tile -> bucket
// keys, and the FP64 primitive-pair records of every (species pair, shell pair).
// ---------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void make_record(double a, double b, double ca, double cb, typename RecOf<T>::type &r)
{
    const double p = a + b;
    const double ip = 1.0 / p;
    const double mu = a * b * ip;
    const double pref = 5.568327996831707845 * ip * sqrt(ip);  // (pi/p)^{3/2}
    if constexpr (std::is_same<T, float>::value) {
        r.mu2 = mu * 1.4426950408889634074;
    } else {
        r.mu2 = mu;
    }
    r.aop = static_cast<T>(a * ip);
    r.bop = static_cast<T>(b * ip);
    r.i2p = static_cast<T>(0.5 * ip);
    r.Kab = static_cast<T>(ca * cb * pref);
    r.w1 = static_cast<T>(cb * pref);
    r.w2 = static_cast<T>(ca * pref);
}
template <typename T>
__device__ __forceinline__ void make_pb_record(double a, double b, double ca, double cb, bool ss, PairPb<T> &r)
{
    // A non-positive exponent has no L2 norm to take out (n = (2l + 3)/(4 alpha) is
    // undefined): its side keeps the plain raised table, c_i = (i + 1)/(2p) = c0 - i s with
    // c0 = 1/(2p), s = -1/(2p), and ovl2c_bwd_reduce adds nothing back for that slot.
    const double p = a + b, d = a - b;
    const double c0a = a > 0.0 ? 0.25 * d / (a * p) : 0.5 / p;
    const double c0b = b > 0.0 ? -0.25 * d / (b * p) : 0.5 / p;
    if (ss) {
        // s-s pairs close their exponent sums over the axes (pair_contract): the companion
        // holds Kab (beta/p)^2, 3 Kab c0a, Kab (alpha/p)^2, 3 Kab c0b, so that
        // Kab Ha = g (ca r2 + cb) and Kab Hb = g (sa r2 + sb) cost one FMA each
        const double kab = ca * cb * 5.568327996831707845 / (p * sqrt(p));
        r.ca = static_cast<T>(kab * (b / p) * (b / p));
        r.cb = static_cast<T>(3.0 * kab * c0a);
        r.sa = static_cast<T>(kab * (a / p) * (a / p));
        r.sb = static_cast<T>(3.0 * kab * c0b);
        return;
    }
    r.ca = static_cast<T>(c0a);
    r.cb = static_cast<T>(c0b);
    r.sa = static_cast<T>(a > 0.0 ? 0.5 * b / (a * p) : -0.5 / p);
    r.sb = static_cast<T>(b > 0.0 ? 0.5 * a / (b * p) : -0.5 / p);
}

template <typename T>
__global__ __launch_bounds__(256) void ovl_fill_kernel(const OvlShape sh, const void *__restrict__ sidxA,
                                                       const void *__restrict__ sidxB, int itype, int64_t B,
                                                       int nblk_perm, int nblk_tile, const int *__restrict__ rank,
                                                       const int *__restrict__ blk_off, const int *__restrict__ toff,
                                                       int *__restrict__ perm, int *__restrict__ tile_key,
                                                       int maxtiles, const double *__restrict__ coef64,
                                                       const double *__restrict__ alpha64,
                                                       const int *__restrict__ nlist, const int *__restrict__ lidx,
                                                       typename RecOf<T>::type *__restrict__ rec,
                                                       PairPb<T> *__restrict__ rec2)
{
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ, NK = NZ * NZ;
    const int bid = blockIdx.x;
    if (bid < nblk_perm) {
        const int64_t i = static_cast<int64_t>(bid) * 256 + threadIdx.x;
        if (i >= B) return;
        int sa = load_sigma(sidxA, itype, i), sb = load_sigma(sidxB, itype, i);
        sa = (sa < 0 || sa >= NZ) ? 0 : sa;
        sb = (sb < 0 || sb >= NZ) ? 0 : sb;
        const int key = sa * NZ + sb;
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
    const int64_t total = static_cast<int64_t>(NK) * ns * ns * K * K;
    if (gid >= total) return;
    const int j = static_cast<int>(gid % K);
    const int i = static_cast<int>((gid / K) % K);
    const int64_t g = gid / (static_cast<int64_t>(K) * K);
    const int sa = static_cast<int>(g % ns);
    const int sb = static_cast<int>((g / ns) % ns);
    const int key = static_cast<int>(g / (static_cast<int64_t>(ns) * ns));
    const int sgA = key / NZ, sgB = key - sgA * NZ;
    const int la_idx = sgA * ns + sa, lb_idx = sgB * ns + sb;
    if (i >= nlist[la_idx] || j >= nlist[lb_idx]) return;
    const int ka = lidx[static_cast<int64_t>(la_idx) * K + i];
    const int kb = lidx[static_cast<int64_t>(lb_idx) * K + j];
    const int64_t qa = sa + static_cast<int64_t>(ns) * (ka + static_cast<int64_t>(K) * sgA);
    const int64_t qb = sb + static_cast<int64_t>(ns) * (kb + static_cast<int64_t>(K) * sgB);
    typename RecOf<T>::type r;
    make_record<T>(alpha64[qa], alpha64[qb], coef64[qa], coef64[qb], r);
    rec[gid] = r;
    if (rec2) {
        PairPb<T> r2;
        make_pb_record<T>(alpha64[qa], alpha64[qb], coef64[qa], coef64[qb], sh.ls[sa] == 0 && sh.ls[sb] == 0, r2);
        rec2[gid] = r2;
    }
}

// ---------------------------------------------------------------------------------
// Per-lane primitive-pair math
// ---------------------------------------------------------------------------------
__device__ __forceinline__ float pair_exp(const PairRecF &r, double R2)
{
    const double y = r.mu2 * R2;  // mu |AB|^2 log2(e), in FP64
    const double n = rint(y);
    const float f = static_cast<float>(y - n);
    return ldexpf(__builtin_amdgcn_exp2f(-f), -static_cast<int>(n));
}
__device__ __forceinline__ double pair_exp(const PairRecD &r, double R2) { return exp(-r.mu2 * R2); }
// Pullback: the argument mu |AB|^2 log2(e) formed in FP64 and rounded once to FP32, then
// v_exp_f32 (3 instructions instead of the forward's 7-instruction range reduction).  The
// rounding perturbs a term e^{-y} by about y 2^-24 ln 2 of itself; pullback terms are
// summed over the batch and judged against the largest cotangent, so the large-y terms
// (whose per-block accuracy needs the forward's exact reduction) are negligible here.
// C1 (B = 65536): pullback 110 -> 107 us; FP32 cotangents against the FP64 oracle on
// 32768 pairs 4.7e-7 (dzeta) and 1.9e-7 (dD) of the largest entry, against 4.5e-7 and
// 1.6e-7 with the exact reduction (with the product also in FP32: 106 us, 7.8e-7).
__device__ __forceinline__ float pair_exp_pb(const PairRecF &r, double R2, float)
{
    return __builtin_amdgcn_exp2f(-static_cast<float>(r.mu2 * R2));
}
__device__ __forceinline__ double pair_exp_pb(const PairRecD &r, double R2, double) { return pair_exp(r, R2); }

// 1-D Obara-Saika overlaps s[i][j], i <= IA, j <= IB, normalised to s[0][0] = 1.
template <typename T, int IA, int IB>
__device__ __forceinline__ void os1d(T PA, T PB, T i2p, T (&s)[IA + 1][IB + 1])
{
    s[0][0] = T(1);
#pragma unroll
    for (int i = 1; i <= IA; ++i) s[i][0] = PA * s[i - 1][0] + (i > 1 ? T(i - 1) * i2p * s[i > 1 ? i - 2 : 0][0] : T(0));
#pragma unroll
    for (int j = 1; j <= IB; ++j)
#pragma unroll
        for (int i = 0; i <= IA; ++i) {
            T v = PB * s[i][j - 1];
            if (i > 0) v += T(i) * i2p * s[i > 0 ? i - 1 : 0][j - 1];
            if (j > 1) v += T(j - 1) * i2p * s[i][j > 1 ? j - 2 : 0];
            s[i][j] = v;
        }
}

// Unroll factors of the primitive-pair loops (launch-shape tuning, fixed at build time).
#ifndef AOK_FWD_UNROLL
#define AOK_FWD_UNROLL 4
#endif
#ifndef AOK_PB_UNROLL
#define AOK_PB_UNROLL 1
#endif
#define AOK_PRAGMA_(x) _Pragma(#x)
#define AOK_PRAGMA(x) AOK_PRAGMA_(x)

template <int L>
struct ShellC {
    static constexpr int nbf = (L + 1) * (L + 2) / 2;
};

// Forward of one shell pair for the lane's b: accumulate over the pair list and write
// the aliased nbf_a x nbf_b block into the LDS stripe (column stride nrows).
template <typename T, int LA, int LB>
__device__ __forceinline__ void pair_fwd(const typename RecOf<T>::type *__restrict__ recs, int nra, int nrb, int K,
                                         double R2, const T (&ab)[3], T *dst, int nrows)
{
    constexpr int NBA = ShellC<LA>::nbf, NBB = ShellC<LB>::nbf, N1 = 2 * LA + 1;
    constexpr int NLIN = (NBA - 1) + N1 * (NBB - 1) + 1;
    T acc[NLIN];
#pragma unroll
    for (int q = 0; q < NLIN; ++q) acc[q] = T(0);
    for (int i = 0; i < nra; ++i) {
        const typename RecOf<T>::type *row = recs + i * K;
        AOK_PRAGMA(unroll AOK_FWD_UNROLL)
        for (int j = 0; j < nrb; ++j) {
            const typename RecOf<T>::type r = row[j];
            const T G = r.Kab * pair_exp(r, R2);
            if constexpr (LA + LB == 0) {
                acc[0] += G;
            } else {
                T sx[LA + 1][LB + 1], sy[LA + 1][LB + 1], sz[LA + 1][LB + 1];
                os1d<T, LA, LB>(-r.bop * ab[0], r.aop * ab[0], r.i2p, sx);
                os1d<T, LA, LB>(-r.bop * ab[1], r.aop * ab[1], r.i2p, sy);
                os1d<T, LA, LB>(-r.bop * ab[2], r.aop * ab[2], r.i2p, sz);
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
                                acc[index1 + N1 * index2] += G * sx[ll1][ll2] * sy[m1][m2] * sz[n1][n2];
                                ++index2;
                            }
                        ++index1;
                    }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NBB; ++j)
#pragma unroll
        for (int i = 0; i < NBA; ++i) {
            const int q = i + NBA * j;
            dst[i + nrows * j] = q < NLIN ? acc[q < NLIN ? q : 0] : T(0);
        }
}

// Contraction of one primitive pair against the aliased cotangent block g (LDS stripe
// slice of the lane, column stride nrows).  H = sum g s_x s_y s_z; with RAISED also the
// exponent sums Ha, Hb of the bra and the ket.  d/dalpha x^i e^{-alpha x^2} = -x^{i+2} e^{..}
// raises a 1-D factor by two; the sums here take each raised factor minus (2i+1)/(4 alpha)
// times itself, u(i, j) = s(i+2, j) - (2i+1)/(4 alpha) s(i, j) per axis, which over the
// three axes removes n_a H with n_a = (2 l_a + 3)/(4 alpha) (the logarithmic derivative of
// a primitive's L2 norm).  By the vertical recursion
//   u(i, j) = PA s(i+1, j) + (alpha - (2i+1) beta)/(4 alpha p) s(i, j) + j/(2p) s(i+1, j-1),
// so the n_a H that the raised moment nearly cancels (diffuse primitives: the raised moment
// and n_a H agree to 5-15x in C1's p shells) never appears as a difference of rounded
// values; the reduction adds n_a coef_a dcoef_a back in FP64 (ovl2c_bwd_reduce).  The ket
// side likewise with v(i, j) = s(i, j+2) - (2j+1)/(4 beta) s(i, j).
// Cotangent accessors for pair_contract: the lane's slice of the LDS stripe (column stride
// nrows), or the shell pair's block held in registers (column stride NBA).
template <typename T>
struct GStripe {
    const T *g;
    int nrows;
    __device__ __forceinline__ T operator()(int i, int j) const { return g[i + nrows * j]; }
};
template <typename T, int NBA, int NBB>
struct GRegs {
    T v[NBA * NBB];
    __device__ __forceinline__ T operator()(int i, int j) const { return v[i + NBA * j]; }
};

// Shell pairs with an l >= AOK_FACTOR_L shell use the factored contraction below (C2,
// B = 262144, FP32: pullback 6.58 -> 6.28 ms with every pair factored; the d-and-lighter
// pairs of C1 gain no time from it, and their FP32 cotangent error against FP64 moved from
// 4.9e-7 to 8.0e-7 of the largest entry, so they keep the per-element form).
#ifndef AOK_FACTOR_L
#define AOK_FACTOR_L 3
#endif
// Exponent tables of one axis from the base table s(i, j), i <= LA, j <= LB, extended
// here by the one row s(LA+1, j) (bra) or column s(i, LB+1) (ket) they need: the bra's
// u(i, j) (KET = false; c[i] = (alpha - (2i+1) beta)/(4 alpha p)) or the ket's v(i, j)
// (KET = true; c[j] = (beta - (2j+1) alpha)/(4 beta p)), i <= LA, j <= LB.
template <typename T, int LA, int LB, bool KET>
__device__ __forceinline__ void os1d_exp(T PA, T PB, T i2p, const T (&c)[(KET ? LB : LA) + 1],
                                         const T (&s)[LA + 1][LB + 1], T (&u)[LA + 1][LB + 1])
{
    if constexpr (!KET) {
        T t[LB + 1];  // s(LA+1, j)
        t[0] = PA * s[LA][0];
        if (LA > 0) t[0] = fma(T(LA) * i2p, s[LA > 0 ? LA - 1 : 0][0], t[0]);
#pragma unroll
        for (int j = 1; j <= LB; ++j) {
            T v = fma(PB, t[j - 1], T(LA + 1) * i2p * s[LA][j - 1]);
            if (j > 1) v = fma(T(j - 1) * i2p, t[j > 1 ? j - 2 : 0], v);
            t[j] = v;
        }
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
            for (int j = 0; j <= LB; ++j) {
                const T up = i < LA ? s[i < LA ? i + 1 : 0][j] : t[j];
                T x = fma(PA, up, c[i] * s[i][j]);
                if (j > 0) x = fma(T(j) * i2p, i < LA ? s[i < LA ? i + 1 : 0][j > 0 ? j - 1 : 0] : t[j > 0 ? j - 1 : 0], x);
                u[i][j] = x;
            }
    } else {
        T t[LA + 1];  // s(i, LB+1)
#pragma unroll
        for (int i = 0; i <= LA; ++i) {
            T v = PB * s[i][LB];
            if (i > 0) v = fma(T(i) * i2p, s[i > 0 ? i - 1 : 0][LB], v);
            if (LB > 0) v = fma(T(LB) * i2p, s[i][LB > 0 ? LB - 1 : 0], v);
            t[i] = v;
        }
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
            for (int j = 0; j <= LB; ++j) {
                const T up = j < LB ? s[i][j < LB ? j + 1 : 0] : t[i];
                T x = fma(PB, up, c[j] * s[i][j]);
                if (i > 0) x = fma(T(i) * i2p, j < LB ? s[i > 0 ? i - 1 : 0][j < LB ? j + 1 : 0] : t[i > 0 ? i - 1 : 0], x);
                u[i][j] = x;
            }
    }
}
// The exponent coefficients c[i] = c0 - i s (PairPb).
template <typename T, int L>
__device__ __forceinline__ void exp_coefs(T c0, T st, T (&c)[L + 1])
{
    c[0] = c0;
#pragma unroll
    for (int i = 1; i <= L; ++i) c[i] = fma(T(-i), st, c0);
}

template <typename T, int LA, int LB, bool RAISED, typename G>
__device__ __forceinline__ void pair_contract(const typename RecOf<T>::type &r, const PairPb<T> &q, const T (&ab)[3],
                                              T r2, const G &g, T &H, T &Ha, T &Hb)
{
    if constexpr (LA == 0 && LB == 0) {
        // s-s: s_x s_y s_z = 1 and the exponent sums close over the axes,
        // sum_ax u_ax(0, 0) = sum_ax (PA_ax^2 + (alpha - beta)/(4 alpha p))
        //                   = (beta/p)^2 |AB|^2 + 3 (alpha - beta)/(4 alpha p), likewise the ket
        // (with the companion's s-s constants, Ha and Hb come out multiplied by Kab)
        const T gv = g(0, 0);
        H = gv;
        Ha = RAISED ? gv * fma(q.ca, r2, q.cb) : T(0);
        Hb = RAISED ? gv * fma(q.sa, r2, q.sb) : T(0);
        return;
    }
    constexpr int NBA = ShellC<LA>::nbf, N1 = 2 * LA + 1;
    T sx[LA + 1][LB + 1], sy[LA + 1][LB + 1], sz[LA + 1][LB + 1];
    const T PA[3] = {-r.bop * ab[0], -r.bop * ab[1], -r.bop * ab[2]};
    const T PB[3] = {r.aop * ab[0], r.aop * ab[1], r.aop * ab[2]};
    os1d<T, LA, LB>(PA[0], PB[0], r.i2p, sx);
    os1d<T, LA, LB>(PA[1], PB[1], r.i2p, sy);
    os1d<T, LA, LB>(PA[2], PB[2], r.i2p, sz);
    // exponent tables (RAISED only; 1 x 1 and unused otherwise)
    constexpr int UA = RAISED ? LA + 1 : 1, UB = RAISED ? LB + 1 : 1;
    H = T(0);
    Ha = T(0);
    Hb = T(0);
    if constexpr (LA >= AOK_FACTOR_L || LB >= AOK_FACTOR_L) {
        T ux[UA][UB], uy[UA][UB], uz[UA][UB], vx[UA][UB], vy[UA][UB], vz[UA][UB];
        if constexpr (RAISED) {
            T ca[LA + 1], cb[LB + 1];
            exp_coefs<T, LA>(q.ca, q.sa, ca);
            exp_coefs<T, LB>(q.cb, q.sb, cb);
            os1d_exp<T, LA, LB, false>(PA[0], PB[0], r.i2p, ca, sx, ux);
            os1d_exp<T, LA, LB, false>(PA[1], PB[1], r.i2p, ca, sy, uy);
            os1d_exp<T, LA, LB, false>(PA[2], PB[2], r.i2p, ca, sz, uz);
            os1d_exp<T, LA, LB, true>(PA[0], PB[0], r.i2p, cb, sx, vx);
            os1d_exp<T, LA, LB, true>(PA[1], PB[1], r.i2p, cb, sy, vy);
            os1d_exp<T, LA, LB, true>(PA[2], PB[2], r.i2p, cb, sz, vz);
        }
        // Factored over the x exponents: for each (lx_a, lx_b) the y-z sums
        //   t = sum g s_y s_z,  ty = sum g u_y s_z,  tz = sum g s_y u_z   (u: exponent table)
        // run over the elements sharing those x exponents, then
        //   H += s_x t,  Ha += u_x t + s_x (ty + tz)  (bra), Hb likewise with v (ket):
        // 2 mul + 5 fma per element instead of about 20 operations.
        int base1 = 0;
#pragma unroll
        for (int ll1 = LA; ll1 >= 0; --ll1) {
            int base2 = 0;
#pragma unroll
            for (int ll2 = LB; ll2 >= 0; --ll2) {
                T t = T(0), tya = T(0), tza = T(0), tyb = T(0), tzb = T(0);
#pragma unroll
                for (int n1 = 0; n1 <= LA - ll1; ++n1) {
                    const int m1 = LA - ll1 - n1;
#pragma unroll
                    for (int n2 = 0; n2 <= LB - ll2; ++n2) {
                        const int m2 = LB - ll2 - n2;
                        const int lin = (base1 + n1) + N1 * (base2 + n2);
                        const T gv = g(lin % NBA, lin / NBA);
                        const T u = gv * sz[n1][n2];
                        t = fma(u, sy[m1][m2], t);
                        if constexpr (RAISED) {
                            const T v = gv * sy[m1][m2];
                            tya = fma(u, uy[m1][m2], tya);
                            tyb = fma(u, vy[m1][m2], tyb);
                            tza = fma(v, uz[n1][n2], tza);
                            tzb = fma(v, vz[n1][n2], tzb);
                        }
                    }
                }
                const T x0 = sx[ll1][ll2];
                H = fma(x0, t, H);
                if constexpr (RAISED) {
                    Ha = fma(ux[ll1][ll2], t, fma(x0, tya + tza, Ha));
                    Hb = fma(vx[ll1][ll2], t, fma(x0, tyb + tzb, Hb));
                }
                base2 += LB - ll2 + 1;
            }
            base1 += LA - ll1 + 1;
        }
    } else {
        // two passes over the elements, the bra's exponent tables in the first and the
        // ket's in the second, so only one side's tables are live at a time
#pragma unroll
        for (int side = 0; side < (RAISED ? 2 : 1); ++side) {
            T ux[UA][UB], uy[UA][UB], uz[UA][UB];
            if constexpr (RAISED) {
                if (side == 0) {
                    T ca[LA + 1];
                    exp_coefs<T, LA>(q.ca, q.sa, ca);
                    os1d_exp<T, LA, LB, false>(PA[0], PB[0], r.i2p, ca, sx, ux);
                    os1d_exp<T, LA, LB, false>(PA[1], PB[1], r.i2p, ca, sy, uy);
                    os1d_exp<T, LA, LB, false>(PA[2], PB[2], r.i2p, ca, sz, uz);
                } else {
                    T cb[LB + 1];
                    exp_coefs<T, LB>(q.cb, q.sb, cb);
                    os1d_exp<T, LA, LB, true>(PA[0], PB[0], r.i2p, cb, sx, ux);
                    os1d_exp<T, LA, LB, true>(PA[1], PB[1], r.i2p, cb, sy, uy);
                    os1d_exp<T, LA, LB, true>(PA[2], PB[2], r.i2p, cb, sz, uz);
                }
            }
            T E = T(0);
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
                            const int lin = index1 + N1 * index2;
                            const T gv = g(lin % NBA, lin / NBA);
                            const T x0 = sx[ll1][ll2], y0 = sy[m1][m2], z0 = sz[n1][n2];
                            if (side == 0) H += gv * (x0 * y0 * z0);
                            if constexpr (RAISED)
                                E += gv * (ux[ll1][ll2] * y0 * z0 + x0 * uy[m1][m2] * z0 + x0 * y0 * uz[n1][n2]);
                            ++index2;
                        }
                    ++index1;
                }
            if (side == 0)
                Ha = E;
            else
                Hb = E;
        }
    }
}

// Pullback of one shell pair.  Every per-pair contribution is first summed per lane in T
// into the lane's own LDS rows (stride LROW, odd, so the transposed read below is free of
// bank conflicts): bra rows 2*i + c for list entry i (c = 0 coefficient, 1 exponent), ket
// rows 2*nla + 2*j + c.  Bra sums run over j in registers, ket sums over i in the rows.
// After the shell pair, lane v sums row v over the 64 lanes in FP64 (a transposed
// reduction: one pass for all rows instead of one cross-lane reduction per row) and adds
// it to the chunk's FP64 sums: the bra rows are stored to the shell's slice of the
// workgroup's partial, accA[2*i + c] (global; each bra shell of a chunk is visited once),
// the ket rows added to accB[2*j + c] (LDS, summed over the chunk's bra shells).
constexpr int LROW = OVL_TILE + 1;

template <typename T, int LA, int LB, typename G>
__device__ __forceinline__ void pair_bwd(const typename RecOf<T>::type *__restrict__ recs,
                                         const PairPb<T> *__restrict__ recs2, int nra, int nla, int nrb,
                                         int nlb, int K, double R2, const T (&ab)[3], const G &g, double *accA,
                                         double *accB, T *lac, int lane)
{
    using R = typename RecOf<T>::type;
    constexpr bool SS = LA + LB == 0;  // s-s exponent sums arrive multiplied by Kab
    const T r2f = static_cast<T>(R2);
    const int nrow = 2 * (nla + nlb);
    T *ket = lac + 2 * nla * LROW;
    // bra rows are stored whole below; only the ket rows (accumulated over i) and the
    // exponent rows of padded bra slots (never written) start at zero
    for (int v = 2 * nla; v < nrow; ++v) lac[v * LROW + lane] = T(0);
    for (int i = nra; i < nla; ++i) lac[(2 * i + 1) * LROW + lane] = T(0);
    // real x real: all four cotangents
    int i0 = 0;
    // s-s and s-p: two bra entries per pass, one read-modify-write of each ket row per two
    // primitive pairs (C1 pullback 110 -> 106 us; three or four per pass for s-s, or two
    // for p-p, s-d as well, measured no faster)
    if constexpr (LA + LB <= 1) {
        constexpr int NJ = 2;
        for (; i0 + NJ - 1 < nra; i0 += NJ) {
            T cA[NJ], aA[NJ];
#pragma unroll
            for (int u = 0; u < NJ; ++u) cA[u] = aA[u] = T(0);
            for (int j = 0; j < nrb; ++j) {
                T kc = T(0), ka = T(0);
#pragma unroll
                for (int u = 0; u < NJ; ++u) {
                    const R r = recs[(i0 + u) * K + j];
                    const T e = pair_exp_pb(r, R2, r2f);
                    T H, Ha, Hb;
                    pair_contract<T, LA, LB, true>(r, recs2[(i0 + u) * K + j], ab, r2f, g, H, Ha, Hb);
                    const T eH = e * H;
                    cA[u] += r.w1 * eH;
                    aA[u] -= (SS ? e : r.Kab * e) * Ha;
                    kc += r.w2 * eH;
                    ka += (SS ? e : r.Kab * e) * Hb;
                }
                ket[2 * j * LROW + lane] += kc;
                ket[(2 * j + 1) * LROW + lane] -= ka;
            }
#pragma unroll
            for (int u = 0; u < NJ; ++u) {
                lac[2 * (i0 + u) * LROW + lane] = cA[u];
                lac[(2 * (i0 + u) + 1) * LROW + lane] = aA[u];
            }
        }
    }
    for (int i = i0; i < nra; ++i) {
        T cA = T(0), aA = T(0);
        AOK_PRAGMA(unroll AOK_PB_UNROLL)
        for (int j = 0; j < nrb; ++j) {
            const R r = recs[i * K + j];
            const T e = pair_exp_pb(r, R2, r2f);
            T H, Ha, Hb;
            pair_contract<T, LA, LB, true>(r, recs2[i * K + j], ab, r2f, g, H, Ha, Hb);
            const T eH = e * H;
            cA += r.w1 * eH;
            aA -= (SS ? e : r.Kab * e) * Ha;
            ket[2 * j * LROW + lane] += r.w2 * eH;
            ket[(2 * j + 1) * LROW + lane] -= (SS ? e : r.Kab * e) * Hb;
        }
        lac[2 * i * LROW + lane] = cA;
        lac[(2 * i + 1) * LROW + lane] = aA;
    }
    // padded bra representative x real ket: coefficient cotangent of the bra slot only
    for (int i = nra; i < nla; ++i) {
        T cA = T(0);
        for (int j = 0; j < nrb; ++j) {
            const R r = recs[i * K + j];
            T H, Ha, Hb;
            pair_contract<T, LA, LB, false>(r, PairPb<T>{}, ab, r2f, g, H, Ha, Hb);
            cA += r.w1 * pair_exp_pb(r, R2, r2f) * H;
        }
        lac[2 * i * LROW + lane] = cA;
    }
    // real bra x padded ket representative: coefficient cotangent of the ket slot only
    for (int j = nrb; j < nlb; ++j) {
        T cB = T(0);
        for (int i = 0; i < nra; ++i) {
            const R r = recs[i * K + j];
            T H, Ha, Hb;
            pair_contract<T, LA, LB, false>(r, PairPb<T>{}, ab, r2f, g, H, Ha, Hb);
            cB += r.w2 * pair_exp_pb(r, R2, r2f) * H;
        }
        ket[2 * j * LROW + lane] = cB;
    }
    __syncthreads();
    for (int v = lane; v < nrow; v += OVL_TILE) {
        const T *row = lac + v * LROW;
        double sum = 0.0;
        if constexpr (std::is_same<T, float>::value) {
            // lanes summed in FP32 quads, then in FP64: 5 instead of 8 VALU per 4 lanes (C1
            // pullback 106.5 -> 104 us; FP32 cotangent errors 2.3e-7 / 5.7e-7 of the largest
            // entry against FP64 and the oracle sample, from 4.9e-7 / 4.7e-7)
#pragma unroll 4
            for (int l = 0; l < OVL_TILE; l += 4) sum += static_cast<double>((row[l] + row[l + 1]) + (row[l + 2] + row[l + 3]));
        } else {
#pragma unroll 8
            for (int l = 0; l < OVL_TILE; ++l) sum += static_cast<double>(row[l]);
        }
        if (v < 2 * nla)
            accA[v] = sum;
        else
            accB[v - 2 * nla] += sum;
    }
    __syncthreads();
}

// Position derivative of one shell pair for the lane's b, accumulated in FP64:
// acc += sum over real primitive pairs of G * sum_lin g[lin] * d/dA (s_x s_y s_z), with
// d/dA_x [x_A^i e^{-alpha x_A^2}] = 2 alpha x_A^{i+1} e - i x_A^{i-1} e, i.e. the 1-D
// tables raised by one on the bra: d_x(i, j) = 2 alpha s_x(i+1, j) - i s_x(i-1, j).
template <typename T, int LA, int LB, typename GAcc>
__device__ __forceinline__ void pair_dx(const typename RecOf<T>::type *__restrict__ recs, int nra, int nrb, int K,
                                        double R2, const T (&ab)[3], const GAcc &g, double (&acc)[3])
{
    constexpr int NBA = ShellC<LA>::nbf, N1 = 2 * LA + 1;
    for (int i = 0; i < nra; ++i)
        for (int j = 0; j < nrb; ++j) {
            const typename RecOf<T>::type r = recs[i * K + j];
            const T G = r.Kab * pair_exp(r, R2);
            T sx[LA + 2][LB + 1], sy[LA + 2][LB + 1], sz[LA + 2][LB + 1];
            os1d<T, LA + 1, LB>(-r.bop * ab[0], r.aop * ab[0], r.i2p, sx);
            os1d<T, LA + 1, LB>(-r.bop * ab[1], r.aop * ab[1], r.i2p, sy);
            os1d<T, LA + 1, LB>(-r.bop * ab[2], r.aop * ab[2], r.i2p, sz);
            const T a2 = r.aop / r.i2p;  // 2 alpha = (alpha / p) / (1 / 2p)
            T hx = T(0), hy = T(0), hz = T(0);
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
                            const int lin = index1 + N1 * index2;
                            const T gv = g(lin % NBA, lin / NBA);
                            const T x0 = sx[ll1][ll2], y0 = sy[m1][m2], z0 = sz[n1][n2];
                            const T dxv = a2 * sx[ll1 + 1][ll2] - (ll1 > 0 ? T(ll1) * sx[ll1 > 0 ? ll1 - 1 : 0][ll2] : T(0));
                            const T dyv = a2 * sy[m1 + 1][m2] - (m1 > 0 ? T(m1) * sy[m1 > 0 ? m1 - 1 : 0][m2] : T(0));
                            const T dzv = a2 * sz[n1 + 1][n2] - (n1 > 0 ? T(n1) * sz[n1 > 0 ? n1 - 1 : 0][n2] : T(0));
                            hx += gv * (dxv * y0 * z0);
                            hy += gv * (x0 * dyv * z0);
                            hz += gv * (x0 * y0 * dzv);
                            ++index2;
                        }
                    ++index1;
                }
            acc[0] += static_cast<double>(G * hx);
            acc[1] += static_cast<double>(G * hy);
            acc[2] += static_cast<double>(G * hz);
        }
}

template <typename T, int LMAXB>
__host__ __device__ constexpr bool bwd_direct()
{
    return LMAXB <= 2 && std::is_same<T, float>::value;
}


// (la, lb) dispatch over the compile-time cases allowed by LMAXB
#define AOK_LL_CASES(X)                                                                                       \
    X(0, 0) X(0, 1) X(0, 2) X(0, 3) X(0, 4) X(1, 0) X(1, 1) X(1, 2) X(1, 3) X(1, 4) X(2, 0) X(2, 1) X(2, 2) \
        X(2, 3) X(2, 4) X(3, 0) X(3, 1) X(3, 2) X(3, 3) X(3, 4) X(4, 0) X(4, 1) X(4, 2) X(4, 3) X(4, 4)

// Walks the elements e = lane, lane + 64, ... of a tile's chunk stripe (64 pairs x len
// elements, len = nrows * nbb): e = bl * len + j * nrows + i, advanced by counters instead
// of divisions.
struct StripeWalk {
    int bl, j, i, dbl, dj, di;
    __device__ __forceinline__ StripeWalk(int lane, int nrows, int nbb)
    {
        const int len = nrows * nbb;
        bl = lane / len;
        const int q = lane - bl * len;
        j = q / nrows;
        i = q - j * nrows;
        dbl = OVL_TILE / len;
        const int dq = OVL_TILE - dbl * len;
        dj = dq / nrows;
        di = dq - dj * nrows;
    }
    __device__ __forceinline__ void next(int nrows, int nbb)
    {
        i += di;
        j += dj;
        bl += dbl;
        if (i >= nrows) {
            i -= nrows;
            ++j;
        }
        if (j >= nbb) {
            j -= nbb;
            ++bl;
        }
    }
};

// Stage the lanes' cotangent stripe (rows row0.., the chunk's nbb ket columns) into LDS
// (zeros for inactive lanes), STAGE_BATCH elements per lane in flight at a time: one load
// at a time paid the full memory latency per element (C2, B = 2^18: the l <= 3 pullback
// instance spent 3.7 of its 5.0 ms staging; batches of 8: pullback 6.12 -> 4.48 ms,
// batches of 16: 4.64 ms).
constexpr int STAGE_BATCH = 8;
template <typename T>
__device__ __forceinline__ void stage_stripe(T *stripe, int str, const T *__restrict__ dS, const int *bl_b,
                                             int64_t NN, int64_t colbase, int N, int nrows, int nbb, int nact,
                                             int lane)
{
    const int total = OVL_TILE * nrows * nbb;
    StripeWalk w(lane, nrows, nbb);
    constexpr int SB = STAGE_BATCH;
    for (int e0 = lane; e0 < total; e0 += OVL_TILE * SB) {
        T v[SB];
        int dst[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const bool in = e0 + OVL_TILE * u < total;
            dst[u] = w.bl * str + w.j * nrows + w.i;
            v[u] = (in && w.bl < nact) ? dS[static_cast<int64_t>(bl_b[w.bl]) * NN + colbase + N * w.j + w.i] : T(0);
            w.next(nrows, nbb);
        }
#pragma unroll
        for (int u = 0; u < SB; ++u)
            if (e0 + OVL_TILE * u < total) stripe[dst[u]] = v[u];
    }
}

// ---------------------------------------------------------------------------------
// Tile bookkeeping shared by the fast kernels
// ---------------------------------------------------------------------------------
// Workgroup -> (tile, chunk), grid (8 G x tiles/8, chunk groups): the 8 XCDs receive
// workgroups round-robin by linear id, so tile 8 q + x % 8 always runs on XCD x % 8, and
// within a group of G consecutive chunks its workgroups follow one another.  G = 1: the
// workgroups resident on a CU run one chunk's shell-pair code and records (small
// blocks); G = all chunks: a tile's chunks back to back share the XCD's L2 (the pairs'
// cotangent lines, the output block's partial lines) -- large blocks.  Returns false for
// the padding workgroups of the last group.
__device__ __forceinline__ bool tile_chunk(int chunk0, int nch, int group, int &tile, int &chunk)
{
    const int r = static_cast<int>(blockIdx.x >> 3);  // q * group + chunk in group
    const int q = r / group, gi = r - q * group;
    tile = q * 8 + static_cast<int>(blockIdx.x & 7u);
    const int c = static_cast<int>(blockIdx.y) * group + gi;
    chunk = c + chunk0;
    return c < nch;
}
struct TileInfo {
    int key, first, nact;
};

__device__ __forceinline__ bool tile_info(int t, int NK, const int *__restrict__ tile_key,
                                          const int *__restrict__ boff, const int *__restrict__ bcnt,
                                          const int *__restrict__ toff, TileInfo &ti)
{
    if (t >= toff[NK]) return false;
    ti.key = tile_key[t];
    ti.first = boff[ti.key] + (t - toff[ti.key]) * OVL_TILE;
    ti.nact = min(OVL_TILE, boff[ti.key] + bcnt[ti.key] - ti.first);
    return true;
}

// ---------------------------------------------------------------------------------
// Fast forward
// ---------------------------------------------------------------------------------
template <typename T, int LMAXB>
__global__ __launch_bounds__(64) void ovl2c_fwd_fast(const OvlShape sh, const T *__restrict__ posA,
                                                     const T *__restrict__ posB, T scale,
                                                     const int *__restrict__ perm, const int *__restrict__ tile_key,
                                                     const int *__restrict__ boff, const int *__restrict__ bcnt,
                                                     const int *__restrict__ toff,
                                                     const typename RecOf<T>::type *__restrict__ rec,
                                                     const int *__restrict__ nreal, T *__restrict__ out, int chunk0,
                                                     int nch, int group, int narrow)
{
    extern __shared__ unsigned char smem_raw[];
    // LDS: the lanes' block offsets b * N * N, the column-segment offsets of the chunk's
    // stripe elements, then the stripe itself (fwd_lds_bytes).  With `narrow` (the output
    // spans less than 4 GiB) both offsets are 32-bit byte offsets.
    int64_t *bl_off = reinterpret_cast<int64_t *>(smem_raw);
    uint32_t *bl_off32 = reinterpret_cast<uint32_t *>(smem_raw);
    int *q_off = reinterpret_cast<int *>(smem_raw + 8 * OVL_TILE);
    T *stripe = reinterpret_cast<T *>(smem_raw + 8 * OVL_TILE + 4 * ((sh.str + 3) & ~3));
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ, N = sh.N;
    TileInfo ti;
    int tile, chunk;
    if (!tile_chunk(chunk0, nch, group, tile, chunk)) return;
    if (!tile_info(tile, NZ * NZ, tile_key, boff, bcnt, toff, ti)) return;
    const int lane = threadIdx.x;
    const int sgA = ti.key / NZ, sgB = ti.key - sgA * NZ;
    const int64_t b = perm[ti.first + (lane < ti.nact ? lane : 0)];
    if (narrow)
        bl_off32[lane] = static_cast<uint32_t>(b * N * N * static_cast<int64_t>(sizeof(T)));
    else
        bl_off[lane] = static_cast<int64_t>(b) * N * N;
    const T Ax = scale * posA[3 * b], Ay = scale * posA[3 * b + 1], Az = scale * posA[3 * b + 2];
    const T Bx = scale * posB[3 * b], By = scale * posB[3 * b + 1], Bz = scale * posB[3 * b + 2];
    const double dx = static_cast<double>(Ax) - static_cast<double>(Bx);
    const double dy = static_cast<double>(Ay) - static_cast<double>(By);
    const double dz = static_cast<double>(Az) - static_cast<double>(Bz);
    const double R2 = dx * dx + dy * dy + dz * dz;
    const T ab[3] = {static_cast<T>(dx), static_cast<T>(dy), static_cast<T>(dz)};
    T *mine = stripe + lane * sh.str;
    {
        const OvlChunk ch = unpack_chunk(sh.ch[chunk]);
        const int sb = ch.sb, lb = sh.ls[sb];
        const int nrb = nreal[sgB * ns + sb];
        const int nbb = nbf_of(lb), nrows = ch.nrows;
        const int len = nrows * nbb;
        const int qscale = narrow ? static_cast<int>(sizeof(T)) : 1;
        for (int q = lane; q < len; q += OVL_TILE) q_off[q] = (N * (q / nrows) + q % nrows) * qscale;
        for (int sa = ch.sa0; sa < ch.sa1; ++sa) {
            const int la = sh.ls[sa];
            const int64_t g = (static_cast<int64_t>(ti.key) * ns + sb) * ns + sa;
            const typename RecOf<T>::type *recs = rec + g * K * K;
            const int nra = nreal[sgA * ns + sa];
            T *dst = mine + (sh.off[sa] - ch.row0);
#define AOK_FWD_CASE(A, B_)                                                          \
    if constexpr (A <= LMAXB && B_ <= LMAXB)                                         \
        if (la == A && lb == B_) {                                                   \
            pair_fwd<T, A, B_>(recs, nra, nrb, K, R2, ab, dst, ch.nrows);            \
            continue;                                                                \
        }
            AOK_LL_CASES(AOK_FWD_CASE)
#undef AOK_FWD_CASE
        }
        __syncthreads();
        // coalesced write-out: element e = bl * len + q of the stripe goes to pair bl's block
        // at column-segment offset q_off[q]; (bl, q) advance by 64 elements per step
        const int total = ti.nact * len;
        T *obase = out + static_cast<int64_t>(N) * sh.off[sb] + ch.row0;
        const int dbl = OVL_TILE / len, dq = OVL_TILE - dbl * len;
        int bl = lane / len, q = lane - bl * len;
        if (narrow) {
            // 32-bit byte offsets from the uniform base (one address add per element) and
            // four elements per pass (C1 forward 53.6 -> 52.6 us)
            char *ob = reinterpret_cast<char *>(obase);
            int si = bl * sh.str + q;
            const int dsi = dbl * sh.str + dq, wrap = sh.str - len;
            auto next = [&]() {
                q += dq;
                bl += dbl;
                si += dsi;
                if (q >= len) {
                    q -= len;
                    ++bl;
                    si += wrap;
                }
            };
            int e = lane;
            for (; e + 3 * OVL_TILE < total; e += 4 * OVL_TILE) {
                uint32_t o[4];
                T v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    o[u] = bl_off32[bl] + static_cast<uint32_t>(q_off[q]);
                    v[u] = stripe[si];
                    next();
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) *reinterpret_cast<T *>(ob + o[u]) = v[u];
            }
            for (; e < total; e += OVL_TILE) {
                *reinterpret_cast<T *>(ob + (bl_off32[bl] + static_cast<uint32_t>(q_off[q]))) = stripe[si];
                next();
            }
        } else {
            for (int e = lane; e < total; e += OVL_TILE) {
                obase[bl_off[bl] + q_off[q]] = stripe[bl * sh.str + q];
                q += dq;
                bl += dbl;
                if (q >= len) {
                    q -= len;
                    ++bl;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// Fast pullback
// ---------------------------------------------------------------------------------
// FP32 instances that stage the cotangent stripe in LDS copy each shell pair's block of at
// least BWD_GREG_MIN values into registers before its primitive loop (instead of LDS reads
// per primitive pair), and the l <= 3 instance is compiled for one wave per SIMD (the
// registers to hold an f-f block): C2 pullback 4.48 -> 4.22 ms.  Thresholds 20-60 and two
// waves per SIMD measured within 1%.
constexpr int BWD_GREG_MIN = 30;
template <typename T, int LMAXB>
constexpr int bwd_max_waves()
{
    return std::is_same<T, float>::value && LMAXB == 3 ? 1 : 8;
}
template <typename T, int LMAXB>
constexpr int bwd_min_waves()
{
    return std::is_same<T, float>::value && LMAXB == 2 ? 4 : 1;
}

template <typename T, int LMAXB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(bwd_min_waves<T, LMAXB>(), bwd_max_waves<T, LMAXB>())))
void ovl2c_bwd_fast(const OvlShape sh, const T *__restrict__ dS,
                                                     const T *__restrict__ posA, const T *__restrict__ posB, T scale,
                                                     const int *__restrict__ perm, const int *__restrict__ tile_key,
                                                     const int *__restrict__ boff, const int *__restrict__ bcnt,
                                                     const int *__restrict__ toff,
                                                     const typename RecOf<T>::type *__restrict__ rec,
                                                     const PairPb<T> *__restrict__ rec2,
                                                     const int *__restrict__ nreal, const int *__restrict__ nlist,
                                                     double *__restrict__ partial, int chunk0, int nch, int group)
{
    extern __shared__ unsigned char smem_raw[];
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ, N = sh.N;
    const int pstride = (ns + 1) * K * 2;
    int *bl_b = reinterpret_cast<int *>(smem_raw);
    // FP32 with bases up to d shells: each shell pair's cotangent block is read straight
    // into registers (no LDS stripe, so more waves fit per CU); larger blocks, and FP64
    // (twice the registers), stage it through the stripe.
    constexpr bool DIRECT = bwd_direct<T, LMAXB>();
    // LDS: [pair index per lane (stripe staging only)][ket sums K x 2 FP64][stripe][lane rows];
    // the bra sums go straight to the partial, which keeps the FP32 C1 instance at 9.5 KB
    // (16 workgroups per CU, the VGPR limit; with the bra sums in LDS, 10.6 KB: 15)
    double *accB = reinterpret_cast<double *>(smem_raw + (DIRECT ? 0 : 256));
    T *stripe = reinterpret_cast<T *>(accB + 2 * K);
    T *lac = DIRECT ? stripe : stripe + OVL_TILE * sh.str;  // per-lane sums [4 * K][LROW]
    TileInfo ti;
    int tile, chunk;
    if (!tile_chunk(chunk0, nch, group, tile, chunk)) return;
    if (!tile_info(tile, NZ * NZ, tile_key, boff, bcnt, toff, ti)) return;
    const int lane = threadIdx.x;
    for (int q = lane; q < 2 * K; q += OVL_TILE) accB[q] = 0.0;
    const int sgA = ti.key / NZ, sgB = ti.key - sgA * NZ;
    const int64_t b = perm[ti.first + (lane < ti.nact ? lane : 0)];
    if constexpr (!DIRECT) bl_b[lane] = static_cast<int>(b);  // b < 2^30 (the batch cap)
    const T Ax = scale * posA[3 * b], Ay = scale * posA[3 * b + 1], Az = scale * posA[3 * b + 2];
    const T Bx = scale * posB[3 * b], By = scale * posB[3 * b + 1], Bz = scale * posB[3 * b + 2];
    const double dx = static_cast<double>(Ax) - static_cast<double>(Bx);
    const double dy = static_cast<double>(Ay) - static_cast<double>(By);
    const double dz = static_cast<double>(Az) - static_cast<double>(Bz);
    const double R2 = dx * dx + dy * dy + dz * dz;
    const T ab[3] = {static_cast<T>(dx), static_cast<T>(dy), static_cast<T>(dz)};
    T *mine = stripe + lane * sh.str;
    const int64_t NN = static_cast<int64_t>(N) * N;
    __syncthreads();
    const OvlChunk ch = unpack_chunk(sh.ch[chunk]);
    const int sb = ch.sb, lb = sh.ls[sb];
    const int nbb = nbf_of(lb), nrows = ch.nrows;
    const int64_t colbase = static_cast<int64_t>(N) * sh.off[sb] + ch.row0;
    // stage the cotangent stripe (zeros for inactive lanes)
    if constexpr (!DIRECT) {
        stage_stripe(stripe, sh.str, dS, bl_b, NN, colbase, N, nrows, nbb, ti.nact, lane);
    }
    __syncthreads();
    const bool active = lane < ti.nact;
    const T *gblk = dS + static_cast<int64_t>(b) * NN + colbase - ch.row0;  // + off[sa] + i + N * j
    const int nrb = nreal[sgB * ns + sb], nlb = nlist[sgB * ns + sb];
    double *pt = partial + (static_cast<int64_t>(tile) * sh.nchunk + chunk) * pstride;
    for (int sa = ch.sa0; sa < ch.sa1; ++sa) {
        const int la = sh.ls[sa];
        const int64_t g = (static_cast<int64_t>(ti.key) * ns + sb) * ns + sa;
        const typename RecOf<T>::type *recs = rec + g * K * K;
        const PairPb<T> *recs2 = rec2 + g * K * K;
        const int nra = nreal[sgA * ns + sa], nla = nlist[sgA * ns + sa];
        const T *gsl = mine + (sh.off[sa] - ch.row0);
        double *aA = pt + 2 * (sa - ch.sa0) * K;
#define AOK_BWD_CASE(A, B_)                                                                        \
    if constexpr (A <= LMAXB && B_ <= LMAXB)                                                       \
        if (la == A && lb == B_) {                                                                 \
            if constexpr (DIRECT) {                                                                \
                GRegs<T, ShellC<A>::nbf, ShellC<B_>::nbf> gr;                                      \
                const T *src = gblk + sh.off[sa];                                                  \
                _Pragma("unroll") for (int j = 0; j < ShellC<B_>::nbf; ++j)                        \
                    _Pragma("unroll") for (int i = 0; i < ShellC<A>::nbf; ++i)                     \
                        gr.v[i + ShellC<A>::nbf * j] = active ? src[i + static_cast<int64_t>(N) * j] : T(0); \
                pair_bwd<T, A, B_>(recs, recs2, nra, nla, nrb, nlb, K, R2, ab, gr, aA, accB, lac, lane);    \
            } else if constexpr (std::is_same<T, float>::value &&                                  \
                                 ShellC<A>::nbf * ShellC<B_>::nbf >= BWD_GREG_MIN) {                \
                GRegs<T, ShellC<A>::nbf, ShellC<B_>::nbf> gr;                                      \
                _Pragma("unroll") for (int j = 0; j < ShellC<B_>::nbf; ++j)                        \
                    _Pragma("unroll") for (int i = 0; i < ShellC<A>::nbf; ++i)                     \
                        gr.v[i + ShellC<A>::nbf * j] = gsl[i + nrows * j];                         \
                pair_bwd<T, A, B_>(recs, recs2, nra, nla, nrb, nlb, K, R2, ab, gr, aA, accB, lac, lane);    \
            } else {                                                                               \
                pair_bwd<T, A, B_>(recs, recs2, nra, nla, nrb, nlb, K, R2, ab, GStripe<T>{gsl, nrows}, aA, accB, \
                                   lac, lane);                                                     \
            }                                                                                      \
            continue;                                                                              \
        }
        AOK_LL_CASES(AOK_BWD_CASE)
#undef AOK_BWD_CASE
    }
    __syncthreads();
    for (int q = lane; q < 2 * K; q += OVL_TILE) pt[ns * K * 2 + q] = accB[q];
}

// ---------------------------------------------------------------------------------
// Position cotangents: grid (tile, chunk) as the pullback; each lane sums its pair's
// A-derivative over the chunk's shell pairs and writes one FP64 partial per (chunk,
// bucket position): partial[(chunk * B + pos) * 3 + c], pos = the lane's place in the
// species-pair bucket order (contiguous per tile, so the stores coalesce).
// ---------------------------------------------------------------------------------
template <typename T, int LMAXB>
__global__ __launch_bounds__(64) void ovl2c_dx_fast(const OvlShape sh, int64_t B, const T *__restrict__ dS,
                                                    const T *__restrict__ posA, const T *__restrict__ posB, T scale,
                                                    const int *__restrict__ perm, const int *__restrict__ tile_key,
                                                    const int *__restrict__ boff, const int *__restrict__ bcnt,
                                                    const int *__restrict__ toff,
                                                    const typename RecOf<T>::type *__restrict__ rec,
                                                    const int *__restrict__ nreal, double *__restrict__ partial,
                                                    int chunk0, int nch, int group)
{
    extern __shared__ unsigned char smem_raw[];
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ, N = sh.N;
    int *bl_b = reinterpret_cast<int *>(smem_raw);
    T *stripe = reinterpret_cast<T *>(smem_raw + 256);
    // as the pullback: FP32 with shells up to d reads each shell pair's block into registers
    constexpr bool DIRECT = bwd_direct<T, LMAXB>();
    TileInfo ti;
    int tile, chunk;
    if (!tile_chunk(chunk0, nch, group, tile, chunk)) return;
    if (!tile_info(tile, NZ * NZ, tile_key, boff, bcnt, toff, ti)) return;
    const int lane = threadIdx.x;
    const int sgA = ti.key / NZ, sgB = ti.key - sgA * NZ;
    const int64_t b = perm[ti.first + (lane < ti.nact ? lane : 0)];
    bl_b[lane] = static_cast<int>(b);  // b < 2^30 (the batch cap)
    const T Ax = scale * posA[3 * b], Ay = scale * posA[3 * b + 1], Az = scale * posA[3 * b + 2];
    const T Bx = scale * posB[3 * b], By = scale * posB[3 * b + 1], Bz = scale * posB[3 * b + 2];
    const double dx = static_cast<double>(Ax) - static_cast<double>(Bx);
    const double dy = static_cast<double>(Ay) - static_cast<double>(By);
    const double dz = static_cast<double>(Az) - static_cast<double>(Bz);
    const double R2 = dx * dx + dy * dy + dz * dz;
    const T ab[3] = {static_cast<T>(dx), static_cast<T>(dy), static_cast<T>(dz)};
    T *mine = stripe + lane * sh.str;
    const int64_t NN = static_cast<int64_t>(N) * N;
    __syncthreads();
    const OvlChunk ch = unpack_chunk(sh.ch[chunk]);
    const int sb = ch.sb, lb = sh.ls[sb];
    const int nbb = nbf_of(lb), nrows = ch.nrows;
    const int64_t colbase = static_cast<int64_t>(N) * sh.off[sb] + ch.row0;
    if constexpr (!DIRECT) {
        stage_stripe(stripe, sh.str, dS, bl_b, NN, colbase, N, nrows, nbb, ti.nact, lane);
    }
    __syncthreads();
    const bool active = lane < ti.nact;
    const T *gblk = dS + static_cast<int64_t>(b) * NN + colbase - ch.row0;  // + off[sa] + i + N * j
    const int nrb = nreal[sgB * ns + sb];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int sa = ch.sa0; sa < ch.sa1; ++sa) {
        const int la = sh.ls[sa];
        const int64_t g = (static_cast<int64_t>(ti.key) * ns + sb) * ns + sa;
        const typename RecOf<T>::type *recs = rec + g * K * K;
        const int nra = nreal[sgA * ns + sa];
        const T *gsl = mine + (sh.off[sa] - ch.row0);
#define AOK_DX_CASE(A, B_)                                                                          \
    if constexpr (A <= LMAXB && B_ <= LMAXB)                                                        \
        if (la == A && lb == B_) {                                                                  \
            if constexpr (DIRECT) {                                                                 \
                GRegs<T, ShellC<A>::nbf, ShellC<B_>::nbf> gr;                                       \
                const T *src = gblk + sh.off[sa];                                                   \
                _Pragma("unroll") for (int j = 0; j < ShellC<B_>::nbf; ++j)                         \
                    _Pragma("unroll") for (int i = 0; i < ShellC<A>::nbf; ++i)                      \
                        gr.v[i + ShellC<A>::nbf * j] = active ? src[i + static_cast<int64_t>(N) * j] : T(0); \
                pair_dx<T, A, B_>(recs, nra, nrb, K, R2, ab, gr, acc);                              \
            } else {                                                                                \
                pair_dx<T, A, B_>(recs, nra, nrb, K, R2, ab, GStripe<T>{gsl, nrows}, acc);          \
            }                                                                                       \
            continue;                                                                               \
        }
        AOK_LL_CASES(AOK_DX_CASE)
#undef AOK_DX_CASE
    }
    if (lane < ti.nact) {
        double *pp = partial + (static_cast<int64_t>(chunk) * B + ti.first + lane) * 3;
        pp[0] = acc[0];
        pp[1] = acc[1];
        pp[2] = acc[2];
    }
}

// Sum of the chunk partials of each pair in chunk order (deterministic), chain rule to
// input length units, dB = -dA.
template <typename T>
__global__ __launch_bounds__(256) void ovl2c_dx_reduce(int64_t B, int nchunk, const int *__restrict__ perm,
                                                       const double *__restrict__ partial, double scale,
                                                       T *__restrict__ dA, T *__restrict__ dB)
{
    const int64_t pos = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (pos >= B) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < nchunk; ++c) {
        const double *pp = partial + (static_cast<int64_t>(c) * B + pos) * 3;
        s[0] += pp[0];
        s[1] += pp[1];
        s[2] += pp[2];
    }
    const int64_t b = perm[pos];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        dA[3 * b + k] = static_cast<T>(scale * s[k]);
        if (dB) dB[3 * b + k] = static_cast<T>(-scale * s[k]);
    }
}

// ---------------------------------------------------------------------------------
// Pullback reduction, stage 1: block (slot, side, partner species).  A primitive slot
// (species sg, shell s, k) receives bra-side sums from the tiles of species pairs
// (sg, partner) in the chunks whose row range holds s, and ket-side sums from the tiles
// of (partner, sg) in the chunks of column shell s.  Tiles of one species pair are
// contiguous, so each block reads only its own rows; the block sums them in a fixed
// order and writes part[(side * NZ + partner) * npar + q] (q the Julia [ns x K x NZ]
// index), dcoef first, then dalpha at + 2 * NZ * npar.  Stage 2 (the sum over the
// 2 * NZ parts) runs in the output stage.
// ---------------------------------------------------------------------------------
// The exponent rows of the partials hold the exponent cotangent at fixed primitive norm,
// da' = da + n coef dcoef with n = (2l + 3)/(4 alpha) (pair_contract); the slot's block
// subtracts n coef dcoef in FP64 after its sum (real slots; padded ones have coef = 0).
__global__ __launch_bounds__(256) void ovl2c_bwd_reduce(const OvlShape sh, const int *__restrict__ toff,
                                                        const int *__restrict__ pcls, const int *__restrict__ nreal,
                                                        const double *__restrict__ coef64,
                                                        const double *__restrict__ alpha64,
                                                        const double *__restrict__ partial, double *__restrict__ part)
{
    __shared__ int rel_ci[OVL_MAXCH], rel_col[OVL_MAXCH];
    __shared__ int wcnt[4];
    __shared__ double red[8];
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ, nchunk = sh.nchunk;
    const int pstride = (ns + 1) * K * 2;
    const int64_t npar = static_cast<int64_t>(NZ) * ns * K;
    const int q = blockIdx.x;
    const int sg = q / (ns * K), rem = q - sg * ns * K, s = rem / K, k = rem - s * K;
    const int li = pcls[(sg * ns + s) * K + k];
    const int side = blockIdx.y / NZ, partner = blockIdx.y - side * NZ;
    const int key = side == 0 ? sg * NZ + partner : partner * NZ + sg;
    // loaded before the barriers below (the compiler keeps global loads on their side), so
    // their latencies overlap the chunk compaction
    const int t0 = toff[key], t1 = toff[key + 1];
    const int64_t dst = s + static_cast<int64_t>(ns) * (k + static_cast<int64_t>(K) * sg);
    const double al = alpha64[dst], cf = coef64[dst];
    const int nr = nreal[sg * ns + s];
    // relevant chunks, compacted in chunk order with ballots
    const int ci = threadIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bool use = false;
    int col = 0;
    if (ci < nchunk) {
        const OvlChunk ch = unpack_chunk(sh.ch[ci]);
        if (side == 0) {
            use = s >= ch.sa0 && s < ch.sa1;
            col = 2 * ((s - ch.sa0) * K + li);
        } else {
            use = ch.sb == s;
            col = 2 * (ns * K + li);
        }
    }
    const unsigned long long m = __ballot(use);
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    int base = 0, nrel = 0;
    for (int w = 0; w < 4; ++w) {
        if (w < wave) base += wcnt[w];
        nrel += wcnt[w];
    }
    if (use) {
        const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
        rel_ci[pos] = ci;
        rel_col[pos] = col;
    }
    __syncthreads();
    const int nw = (t1 - t0) * nrel;
    double c = 0.0, a = 0.0;
    for (int w = threadIdx.x; w < nw; w += 256) {
        const int t = w / nrel, r = w - t * nrel;
        const double *pt = partial + (static_cast<int64_t>(t0 + t) * nchunk + rel_ci[r]) * pstride + rel_col[r];
        c += pt[0];
        a += pt[1];
    }
    block_sum256x2(c, a, red);
    if (threadIdx.x == 0) {
        if (li < nr && al > 0.0) a -= (2 * sh.ls[s] + 3) / (4.0 * al) * cf * c;
        part[blockIdx.y * npar + dst] = c;
        part[(2 * NZ + blockIdx.y) * npar + dst] = a;
    }
}

// ---------------------------------------------------------------------------------
// Output stages: cast, or normalisation VJP (overlap_2c_grad.jl:251-304) in FP64.
// ---------------------------------------------------------------------------------
// Parameter cotangent q summed over its nparts partial sums (stride pstride), in order;
// eight loads in flight at a time (one at a time, each part cost a memory latency).
__device__ __forceinline__ double sum_parts(const double *__restrict__ p, int nparts, int64_t pstride, int64_t q)
{
    double v = 0.0;
    if (p)
        for (int i0 = 0; i0 < nparts; i0 += 8) {
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = i0 + u < nparts ? p[(i0 + u) * pstride + q] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (i0 + u < nparts) v += x[u];
        }
    return v;
}

template <typename T>
__global__ void ovl2c_cast_out(int n, const double *__restrict__ pc, const double *__restrict__ pa, int nparts,
                               int64_t pstride, T *__restrict__ oa, T *__restrict__ ob)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (oa) oa[q] = static_cast<T>(sum_parts(pc, nparts, pstride, q));
    if (ob) ob[q] = static_cast<T>(sum_parts(pa, nparts, pstride, q));
}

// dD, dzeta from (dcoef, dalpha) in FP64: one block per (species, shell).
//   dD_m = dc_m / n - 2 c_l g_m s,  dzeta_m = da_m + 2 c_l (l + 3/2) D_m h_m s,
//   s = (sum_i dc_i D_i) / (2 n^3)
template <typename T>
__global__ __launch_bounds__(256) void ovl2c_vjp_out(const OvlShape sh, const double *__restrict__ pc,
                                                     const double *__restrict__ pa, int nparts, int64_t pstride,
                                                     const T *__restrict__ zetaT, const T *__restrict__ DT,
                                                     T *__restrict__ dzeta, T *__restrict__ dD)
{
    __shared__ ShellNorm S;
    __shared__ double dcs[OVL_MAXK], das[OVL_MAXK];
    __shared__ double s2n3;
    const int ns = sh.ns, K = sh.K;
    const int sg = blockIdx.x / ns, s = blockIdx.x - sg * ns, l = sh.ls[s];
    const int64_t q0 = s + static_cast<int64_t>(ns) * K * sg;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const int64_t q = q0 + static_cast<int64_t>(ns) * k;
        dcs[k] = sum_parts(pc, nparts, pstride, q);
        das[k] = sum_parts(pa, nparts, pstride, q);
    }
    shell_norm(zetaT, DT, q0, ns, K, l, true, S);
    const double nsq = S.nsq;
    if (!(nsq > 0.0)) {
        for (int k = threadIdx.x; k < K; k += blockDim.x) {
            const int64_t q = q0 + static_cast<int64_t>(ns) * k;
            if (dD) dD[q] = T(0);
            if (dzeta) dzeta[q] = static_cast<T>(das[k]);
        }
        return;
    }
    const double n = sqrt(nsq), c = cart_cl(l), e = l + 1.5;
    if (threadIdx.x == 0) {
        double dot = 0.0;
        for (int k = 0; k < K; ++k) dot += dcs[k] * S.d[k];
        s2n3 = dot / (2.0 * n * n * n);
    }
    __syncthreads();
    for (int m = threadIdx.x; m < K; m += blockDim.x) {
        const int64_t q = q0 + static_cast<int64_t>(ns) * m;
        const double gm = 2.0 * c * S.g[m];
        const double hm = -2.0 * c * e * S.d[m] * S.h[m];
        if (dD) dD[q] = static_cast<T>(dcs[m] / n - gm * s2n3);
        if (dzeta) dzeta[q] = static_cast<T>(das[m] - hm * s2n3);
    }
}

// coef = D / sqrt(normsq) per (species, shell) in FP64 (cartesian_gto_basis.jl:110-120),
// one block per (species, shell)
template <typename T>
__global__ __launch_bounds__(256) void compile_coef_kernel(const OvlShape sh, const T *__restrict__ zetaT,
                                                           const T *__restrict__ DT, T *__restrict__ coef)
{
    __shared__ ShellNorm S;
    const int ns = sh.ns, K = sh.K;
    const int sg = blockIdx.x / ns, s = blockIdx.x - sg * ns;
    const int64_t q0 = s + static_cast<int64_t>(ns) * K * sg;
    shell_norm(zetaT, DT, q0, ns, K, sh.ls[s], false, S);
    const double nsq = S.nsq;
    for (int k = threadIdx.x; k < K; k += blockDim.x)
        coef[q0 + static_cast<int64_t>(ns) * k] = nsq > 0.0 ? static_cast<T>(S.d[k] / sqrt(nsq)) : T(0);
}

// ---------------------------------------------------------------------------------
// Plain path: the reference decomposition (one item per (b, s_a, s_b)).
// ---------------------------------------------------------------------------------
template <typename T, int LMAXP, bool RAISE>
__device__ void plain_etab(T (&E)[2 * (LMAXP + (RAISE ? 2 : 0)) + 2][LMAXP + (RAISE ? 2 : 0) + 1]
                                  [LMAXP + (RAISE ? 2 : 0) + 1][3],
                           int la1, int lb1, T PA, T PB, T oo2p, T E00, int ax)
{
    // MD recursion of overlap_2c.jl:222-253 (0-based: t, i, j)
    constexpr int LR = LMAXP + (RAISE ? 2 : 0);
    constexpr int Tmax = 2 * LR + 2;
    for (int jb = 0; jb < lb1; ++jb)
        for (int ia = 0; ia < la1; ++ia) {
            const int ab = ia + jb + 1;  // first index beyond t = ia + jb
            if (ab < Tmax) E[ab][ia][jb][ax] = T(0);
            if (ab + 1 < Tmax) E[ab + 1][ia][jb][ax] = T(0);
        }
    E[0][0][0][ax] = E00;
    for (int i = 1; i < la1; ++i) {
        E[0][i][0][ax] = PA * E[0][i - 1][0][ax] + E[1][i - 1][0][ax];
        for (int t = 1; t <= i; ++t)
            E[t][i][0][ax] = PA * E[t][i - 1][0][ax] + T(t + 1) * E[t + 1][i - 1][0][ax] + oo2p * E[t - 1][i - 1][0][ax];
    }
    for (int j = 1; j < lb1; ++j) {
        E[0][0][j][ax] = PB * E[0][0][j - 1][ax] + E[1][0][j - 1][ax];
        for (int t = 1; t <= j; ++t)
            E[t][0][j][ax] = PB * E[t][0][j - 1][ax] + T(t + 1) * E[t + 1][0][j - 1][ax] + oo2p * E[t - 1][0][j - 1][ax];
        for (int i = 1; i < la1; ++i) {
            E[0][i][j][ax] = PA * E[0][i - 1][j][ax] + E[1][i - 1][j][ax];
            for (int t = 1; t <= i + j; ++t)
                E[t][i][j][ax] =
                    PA * E[t][i - 1][j][ax] + T(t + 1) * E[t + 1][i - 1][j][ax] + oo2p * E[t - 1][i - 1][j][ax];
        }
    }
}

template <typename T, int LMAXP>
__global__ __launch_bounds__(256) void ovl2c_fwd_plain(const OvlShape sh, int64_t B, const T *__restrict__ coef,
                                                       const T *__restrict__ alpha, const void *__restrict__ sidxA,
                                                       const void *__restrict__ sidxB, int itype,
                                                       const T *__restrict__ posA, const T *__restrict__ posB,
                                                       T scale, T *__restrict__ out, unsigned *__restrict__ err)
{
    constexpr int NBF = (LMAXP + 1) * (LMAXP + 2) / 2;
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ, N = sh.N;
    const int64_t item = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (item >= B * ns * ns) return;
    const int64_t b = item % B;
    const int sa = static_cast<int>((item / B) % ns), sb = static_cast<int>(item / (B * ns));
    int sgA = load_sigma(sidxA, itype, b), sgB = load_sigma(sidxB, itype, b);
    if (sgA < 0 || sgA >= NZ || sgB < 0 || sgB >= NZ) {
        *err = ERR_SPECIES;
        return;
    }
    const int la = sh.ls[sa], lb = sh.ls[sb];
    const int nba = nbf_of(la), nbb = nbf_of(lb), N1 = 2 * la + 1;
    T E[2 * LMAXP + 2][LMAXP + 1][LMAXP + 1][3];
    T blk[NBF * NBF];
    for (int k = 0; k < nba * nbb; ++k) blk[k] = T(0);
    const T A[3] = {scale * posA[3 * b], scale * posA[3 * b + 1], scale * posA[3 * b + 2]};
    const T Bv[3] = {scale * posB[3 * b], scale * posB[3 * b + 1], scale * posB[3 * b + 2]};
    const T pi32 = T(5.568327996831707845);
    for (int ip = 0; ip < K; ++ip)
        for (int jp = 0; jp < K; ++jp) {
            const int64_t qa = sa + static_cast<int64_t>(ns) * (ip + static_cast<int64_t>(K) * sgA);
            const int64_t qb = sb + static_cast<int64_t>(ns) * (jp + static_cast<int64_t>(K) * sgB);
            const T ca = coef[qa], cb = coef[qb];
            if (ca == T(0) || cb == T(0)) continue;
            const T aa = alpha[qa], ab = alpha[qb];
            const T p = aa + ab, mu = aa * ab / p, oo2p = T(0.5) / p;
            const T prefac = ca * cb * pi32 / (p * sqrt(p));
            for (int ax = 0; ax < 3; ++ax) {
                const T P = (aa * A[ax] + ab * Bv[ax]) / p;
                const T AB = A[ax] - Bv[ax];
                plain_etab<T, LMAXP, false>(E, la + 1, lb + 1, P - A[ax], P - Bv[ax], oo2p, exp(-mu * AB * AB), ax);
            }
            int index1 = 0;
            for (int ll1 = la; ll1 >= 0; --ll1)
                for (int n1 = 0; n1 <= la - ll1; ++n1) {
                    const int m1 = la - ll1 - n1;
                    int index2 = 0;
                    for (int ll2 = lb; ll2 >= 0; --ll2)
                        for (int n2 = 0; n2 <= lb - ll2; ++n2) {
                            const int m2 = lb - ll2 - n2;
                            blk[index1 + N1 * index2] += E[0][ll1][ll2][0] * E[0][m1][m2][1] * E[0][n1][n2][2] * prefac;
                            ++index2;
                        }
                    ++index1;
                }
        }
    const int64_t base = static_cast<int64_t>(N) * N * b;
    for (int j = 0; j < nbb; ++j)
        for (int i = 0; i < nba; ++i)
            out[base + (sh.off[sa] + i) + static_cast<int64_t>(N) * (sh.off[sb] + j)] = blk[i + nba * j];
}

template <typename T, int LMAXP>
__global__ __launch_bounds__(256) void ovl2c_bwd_plain(const OvlShape sh, int64_t B, const T *__restrict__ dS,
                                                       const T *__restrict__ coef, const T *__restrict__ alpha,
                                                       const void *__restrict__ sidxA, const void *__restrict__ sidxB,
                                                       int itype, const T *__restrict__ posA,
                                                       const T *__restrict__ posB, T scale, T *__restrict__ dcoef,
                                                       T *__restrict__ dalpha, unsigned *__restrict__ err)
{
    constexpr int LR = LMAXP + 2;
    const int ns = sh.ns, K = sh.K, NZ = sh.NZ, N = sh.N;
    const int64_t item = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (item >= B * ns * ns) return;
    const int64_t b = item % B;
    const int sa = static_cast<int>((item / B) % ns), sb = static_cast<int>(item / (B * ns));
    int sgA = load_sigma(sidxA, itype, b), sgB = load_sigma(sidxB, itype, b);
    if (sgA < 0 || sgA >= NZ || sgB < 0 || sgB >= NZ) {
        *err = ERR_SPECIES;
        return;
    }
    const int la = sh.ls[sa], lb = sh.ls[sb];
    const int nba = nbf_of(la), N1 = 2 * la + 1;
    T E[2 * LR + 2][LR + 1][LR + 1][3];
    const T A[3] = {scale * posA[3 * b], scale * posA[3 * b + 1], scale * posA[3 * b + 2]};
    const T Bv[3] = {scale * posB[3 * b], scale * posB[3 * b + 1], scale * posB[3 * b + 2]};
    const T pi32 = T(5.568327996831707845);
    const int64_t base = static_cast<int64_t>(N) * N * b;
    for (int ip = 0; ip < K; ++ip)
        for (int jp = 0; jp < K; ++jp) {
            const int64_t qa = sa + static_cast<int64_t>(ns) * (ip + static_cast<int64_t>(K) * sgA);
            const int64_t qb = sb + static_cast<int64_t>(ns) * (jp + static_cast<int64_t>(K) * sgB);
            const T ca = coef[qa], cb = coef[qb];
            const T aa = alpha[qa], ab = alpha[qb];
            const T p = aa + ab, mu = aa * ab / p, oo2p = T(0.5) / p;
            const T pref0 = pi32 / (p * sqrt(p));
            for (int ax = 0; ax < 3; ++ax) {
                const T P = (aa * A[ax] + ab * Bv[ax]) / p;
                const T AB = A[ax] - Bv[ax];
                plain_etab<T, LMAXP, true>(E, la + 3, lb + 3, P - A[ax], P - Bv[ax], oo2p, exp(-mu * AB * AB), ax);
            }
            T Bsum = T(0), Aa = T(0), Ab = T(0);
            int index1 = 0;
            for (int ll1 = la; ll1 >= 0; --ll1)
                for (int n1 = 0; n1 <= la - ll1; ++n1) {
                    const int m1 = la - ll1 - n1;
                    int index2 = 0;
                    for (int ll2 = lb; ll2 >= 0; --ll2)
                        for (int n2 = 0; n2 <= lb - ll2; ++n2) {
                            const int m2 = lb - ll2 - n2;
                            const int lin = index1 + N1 * index2;
                            const int i = lin % nba, j = lin / nba;
                            const T g = dS[base + (sh.off[sa] + i) + static_cast<int64_t>(N) * (sh.off[sb] + j)];
                            const T Ex = E[0][ll1][ll2][0], Ey = E[0][m1][m2][1], Ez = E[0][n1][n2][2];
                            const T dGa = -(E[0][ll1 + 2][ll2][0] * Ey * Ez + Ex * E[0][m1 + 2][m2][1] * Ez +
                                            Ex * Ey * E[0][n1 + 2][n2][2]);
                            const T dGb = -(E[0][ll1][ll2 + 2][0] * Ey * Ez + Ex * E[0][m1][m2 + 2][1] * Ez +
                                            Ex * Ey * E[0][n1][n2 + 2][2]);
                            Bsum += g * (Ex * Ey * Ez);
                            Aa += g * dGa;
                            Ab += g * dGb;
                            ++index2;
                        }
                    ++index1;
                }
            atomicAdd(&dcoef[qa], cb * pref0 * Bsum);
            atomicAdd(&dcoef[qb], ca * pref0 * Bsum);
            atomicAdd(&dalpha[qa], ca * cb * pref0 * Aa);
            atomicAdd(&dalpha[qb], ca * cb * pref0 * Ab);
        }
}

template <typename T>
__global__ void widen_kernel(int64_t n, const T *__restrict__ a, const T *__restrict__ b, double *__restrict__ oa,
                             double *__restrict__ ob)
{
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    oa[q] = static_cast<double>(a[q]);
    ob[q] = static_cast<double>(b[q]);
}

template <typename T>
__global__ void fill_zero(int64_t n, T *__restrict__ a)
{
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < n) a[q] = T(0);
}

// ---------------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------------


static int lmax_of(const int64_t *ls, int64_t ns)
{
    int m = 0;
    for (int64_t q = 0; q < ns; ++q) m = std::max<int>(m, static_cast<int>(ls[q]));
    return m;
}

// Default stripe capacities (floats per lane) when the ovl_cap knob is 0: the forward's
// chunks, and the pullback's and position cotangents' (the FP32 l <= 2 instances of those
// hold no stripe, so there the capacity only sets how many bra shells a workgroup walks).
// C1 (B = 65536): pullback 99.6 -> 93.9 us at 28 against the forward's 40; the forward at
// 28: 46.7 -> 53.9 us.
constexpr int OVL_STRIPE_CAP = 40;
constexpr int OVL_STRIPE_CAP_PB = 28;

template <typename T>
static aok_status fast_prep(aok_ctx ctx, const Ovl2cCall &c, const OvlShape &sh, const Ovl2cWs &w, char *ws)
{
    using R = typename RecOf<T>::type;
    const int nblk_h = w.nblk_h;
    hipLaunchKernelGGL(ovl_hist_kernel, dim3(nblk_h), dim3(256), 0, ctx->stream, c.sidxA, c.sidxB, c.itype, c.B,
                       sh.NZ, reinterpret_cast<int *>(ws + w.rank), reinterpret_cast<int *>(ws + w.blk_hist),
                       ctx->d_err);
    AOK_HIP(ctx, hipGetLastError());
    // 1024 threads: block 0's bucket offsets give each key's block histogram column to one of
    // 16 waves (36 keys: 3 passes of dependent loads per wave instead of 9)
    hipLaunchKernelGGL((ovl_plan_kernel<T>), dim3(1 + sh.NZ * sh.ns), dim3(1024), 0, ctx->stream, sh, c.mode,
                       static_cast<const T *>(c.coef), static_cast<const T *>(c.alpha), static_cast<const T *>(c.D),
                       reinterpret_cast<double *>(ws + w.coef64), reinterpret_cast<double *>(ws + w.alpha64),
                       reinterpret_cast<int *>(ws + w.nreal), reinterpret_cast<int *>(ws + w.nlist),
                       reinterpret_cast<int *>(ws + w.lidx), reinterpret_cast<int *>(ws + w.pcls), nblk_h,
                       reinterpret_cast<const int *>(ws + w.blk_hist), reinterpret_cast<int *>(ws + w.blk_off),
                       reinterpret_cast<int *>(ws + w.bcnt), reinterpret_cast<int *>(ws + w.boff),
                       reinterpret_cast<int *>(ws + w.toff));
    AOK_HIP(ctx, hipGetLastError());
    const int nblk_tile = (w.maxtiles + 255) / 256;
    const int64_t nrec = static_cast<int64_t>(w.NK) * sh.ns * sh.ns * sh.K * sh.K;
    const int nblk_rec = static_cast<int>((nrec + 255) / 256);
    hipLaunchKernelGGL((ovl_fill_kernel<T>), dim3(nblk_h + nblk_tile + nblk_rec), dim3(256), 0, ctx->stream, sh,
                       c.sidxA, c.sidxB, c.itype, c.B, nblk_h, nblk_tile, reinterpret_cast<const int *>(ws + w.rank),
                       reinterpret_cast<const int *>(ws + w.blk_off), reinterpret_cast<const int *>(ws + w.toff),
                       reinterpret_cast<int *>(ws + w.perm), reinterpret_cast<int *>(ws + w.tile_key), w.maxtiles,
                       reinterpret_cast<const double *>(ws + w.coef64), reinterpret_cast<const double *>(ws + w.alpha64),
                       reinterpret_cast<const int *>(ws + w.nlist), reinterpret_cast<const int *>(ws + w.lidx),
                       reinterpret_cast<R *>(ws + w.rec),
                       w.rec2 ? reinterpret_cast<PairPb<T> *>(ws + w.rec2) : nullptr);
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

// Chunk group of the fast kernels' workgroup order (tile_chunk): the ovl_order knob (1:
// one chunk, 2: all chunks of a launch, k >= 3: k - 1), else all chunks for output blocks
// above 32 x 32 (measured: C2, N = 50, forward 3.0 -> 1.8 ms and pullback 7.6 -> 6.6 ms;
// C1, N = 15, is faster one chunk at a time).  Grids too large for groups use G = 1.
static int grid_group(aok_ctx ctx, const OvlShape &sh, const Ovl2cWs &w, int nch)
{
    const int o = ctx->tune.ovl_order > 0 ? ctx->tune.ovl_order : (sh.N > 32 ? 2 : 1);
    const int g = o == 1 ? 1 : (o == 2 ? nch : std::min(nch, o - 1));
    const int64_t t8 = (w.maxtiles + 7) / 8;
    return 8 * static_cast<int64_t>(g) * t8 <= INT32_MAX ? g : 1;
}

static dim3 launch_grid(const Ovl2cWs &w, int nch, int g)
{
    const unsigned t8 = static_cast<unsigned>((w.maxtiles + 7) / 8);
    return dim3(8u * static_cast<unsigned>(g) * t8, static_cast<unsigned>((nch + g - 1) / g));
}

template <typename T, int LMAXB>
static aok_status fast_fwd_launch(aok_ctx ctx, const Ovl2cCall &c, const OvlShape &sh, const Ovl2cWs &w, char *ws,
                                  int chunk0, int nch)
{
    if (nch <= 0) return AOK_OK;
    using R = typename RecOf<T>::type;
    const int grp = grid_group(ctx, sh, w, nch);
    const size_t shm = 8 * OVL_TILE + 4 * ((sh.str + 3) & ~3) + sizeof(T) * OVL_TILE * sh.str;
    auto kern = &ovl2c_fwd_fast<T, LMAXB>;
    if (shm > 64 * 1024)
        AOK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm)));
    hipLaunchKernelGGL(kern, launch_grid(w, nch, grp), dim3(OVL_TILE), shm, ctx->stream, sh,
                       static_cast<const T *>(c.posA), static_cast<const T *>(c.posB), static_cast<T>(c.scale),
                       reinterpret_cast<const int *>(ws + w.perm), reinterpret_cast<const int *>(ws + w.tile_key),
                       reinterpret_cast<const int *>(ws + w.boff), reinterpret_cast<const int *>(ws + w.bcnt),
                       reinterpret_cast<const int *>(ws + w.toff), reinterpret_cast<const R *>(ws + w.rec),
                       reinterpret_cast<const int *>(ws + w.nreal), static_cast<T *>(c.out), chunk0, nch,
                       grp, c.B * sh.N * sh.N * static_cast<int64_t>(sizeof(T)) < (int64_t(1) << 32) ? 1 : 0);
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

template <typename T, int LMAXB>
static aok_status fast_bwd_launch(aok_ctx ctx, const Ovl2cCall &c, const OvlShape &sh, const Ovl2cWs &w, char *ws,
                                  int chunk0, int nch)
{
    if (nch <= 0) return AOK_OK;
    using R = typename RecOf<T>::type;
    const int grp = grid_group(ctx, sh, w, nch);
    constexpr bool DIRECT = bwd_direct<T, LMAXB>();
    const size_t shm = (DIRECT ? 0 : 256) + sizeof(double) * sh.K * 2 +
                       sizeof(T) * ((DIRECT ? 0 : OVL_TILE * sh.str) + 4 * sh.K * LROW);
    auto kern = &ovl2c_bwd_fast<T, LMAXB>;
    if (shm > 64 * 1024)
        AOK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm)));
    hipLaunchKernelGGL(kern, launch_grid(w, nch, grp), dim3(OVL_TILE), shm, ctx->stream, sh,
                       static_cast<const T *>(c.dS),
                       static_cast<const T *>(c.posA), static_cast<const T *>(c.posB), static_cast<T>(c.scale),
                       reinterpret_cast<const int *>(ws + w.perm), reinterpret_cast<const int *>(ws + w.tile_key),
                       reinterpret_cast<const int *>(ws + w.boff), reinterpret_cast<const int *>(ws + w.bcnt),
                       reinterpret_cast<const int *>(ws + w.toff), reinterpret_cast<const R *>(ws + w.rec),
                       reinterpret_cast<const PairPb<T> *>(ws + w.rec2),
                       reinterpret_cast<const int *>(ws + w.nreal), reinterpret_cast<const int *>(ws + w.nlist),
                       reinterpret_cast<double *>(ws + w.partial), chunk0, nch, grp);
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

static aok_status bwd_reduce_launch(aok_ctx ctx, const OvlShape &sh, const Ovl2cWs &w, char *ws)
{
    const int nslot = sh.NZ * sh.ns * sh.K;
    hipLaunchKernelGGL(ovl2c_bwd_reduce, dim3(nslot, 2 * sh.NZ), dim3(256), 0, ctx->stream, sh,
                       reinterpret_cast<const int *>(ws + w.toff), reinterpret_cast<const int *>(ws + w.pcls),
                       reinterpret_cast<const int *>(ws + w.nreal), reinterpret_cast<const double *>(ws + w.coef64),
                       reinterpret_cast<const double *>(ws + w.alpha64),
                       reinterpret_cast<const double *>(ws + w.partial), reinterpret_cast<double *>(ws + w.part2));
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

// (dcoef, dalpha) = sums of nparts partial arrays (stride pstride) -> the caller's outputs:
// cast, or the normalisation VJP (overlap_2c_grad.jl:251-304) in FP64.
template <typename T>
static aok_status out_stage(aok_ctx ctx, const Ovl2cCall &c, const OvlShape &sh, const double *pc, const double *pa,
                            int nparts, int64_t pstride)
{
    const int npar = sh.NZ * sh.ns * sh.K;
    if (c.vjp) {
        hipLaunchKernelGGL((ovl2c_vjp_out<T>), dim3(sh.NZ * sh.ns), dim3(256), 0, ctx->stream, sh, pc, pa, nparts,
                           pstride, static_cast<const T *>(c.alpha), static_cast<const T *>(c.D),
                           static_cast<T *>(c.o1), static_cast<T *>(c.o2));
    } else {
        hipLaunchKernelGGL((ovl2c_cast_out<T>), dim3((npar + 255) / 256), dim3(256), 0, ctx->stream, npar, pc, pa,
                           nparts, pstride, static_cast<T *>(c.o1), static_cast<T *>(c.o2));
    }
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

template <typename T, int LMAXP>
static aok_status plain_launch(aok_ctx ctx, const Ovl2cCall &c, const OvlShape &sh, bool pullback)
{
    const int64_t items = c.B * sh.ns * sh.ns;
    const int grid = static_cast<int>((items + 255) / 256);
    const int64_t npar = static_cast<int64_t>(sh.NZ) * sh.ns * sh.K;
    // workspace: [compiled coef (T)] [dcoef, dalpha (T)] [dc64, da64 (double)], sized once
    const size_t sT = align256(sizeof(T) * npar), sD = align256(sizeof(double) * npar);
    aok_status st = ensure_ws(ctx, sT + 2 * sT + 2 * sD);
    if (st != AOK_OK) return st;
    char *ws = static_cast<char *>(ctx->ws);
    const T *coef = static_cast<const T *>(c.coef);
    if (c.mode == 1) {  // compile coef in FP64 into the workspace first
        T *cw = reinterpret_cast<T *>(ws);
        hipLaunchKernelGGL((compile_coef_kernel<T>), dim3(sh.NZ * sh.ns), dim3(256), 0, ctx->stream, sh,
                           static_cast<const T *>(c.alpha), static_cast<const T *>(c.D), cw);
        AOK_HIP(ctx, hipGetLastError());
        coef = cw;
    }
    if (!pullback) {
        if (grid == 0) return AOK_OK;
        time_begin(ctx, "ovl2c_fwd_plain");
        hipLaunchKernelGGL((ovl2c_fwd_plain<T, LMAXP>), dim3(grid), dim3(256), 0, ctx->stream, sh, c.B, coef,
                           static_cast<const T *>(c.alpha), c.sidxA, c.sidxB, c.itype, static_cast<const T *>(c.posA),
                           static_cast<const T *>(c.posB), static_cast<T>(c.scale), static_cast<T *>(c.out),
                           ctx->d_err);
        time_end(ctx);
        AOK_HIP(ctx, hipGetLastError());
        return AOK_OK;
    }
    // pullback: T-precision atomics like the reference, into (dcoef, dalpha) buffers
    T *dcoef = c.vjp ? reinterpret_cast<T *>(ws + sT) : static_cast<T *>(c.o1);
    T *dalpha = c.vjp ? reinterpret_cast<T *>(ws + 2 * sT) : static_cast<T *>(c.o2);
    double *dc64 = reinterpret_cast<double *>(ws + 3 * sT);
    double *da64 = reinterpret_cast<double *>(ws + 3 * sT + sD);
    hipLaunchKernelGGL((fill_zero<T>), dim3((npar + 255) / 256), dim3(256), 0, ctx->stream, npar, dcoef);
    hipLaunchKernelGGL((fill_zero<T>), dim3((npar + 255) / 256), dim3(256), 0, ctx->stream, npar, dalpha);
    if (grid > 0) {
        time_begin(ctx, "ovl2c_bwd_plain");
        hipLaunchKernelGGL((ovl2c_bwd_plain<T, LMAXP>), dim3(grid), dim3(256), 0, ctx->stream, sh, c.B,
                           static_cast<const T *>(c.dS), coef, static_cast<const T *>(c.alpha), c.sidxA, c.sidxB,
                           c.itype, static_cast<const T *>(c.posA), static_cast<const T *>(c.posB),
                           static_cast<T>(c.scale), dcoef, dalpha, ctx->d_err);
        time_end(ctx);
    }
    AOK_HIP(ctx, hipGetLastError());
    if (c.vjp) {
        // widen to FP64, then the FP64 normalisation VJP
        hipLaunchKernelGGL((widen_kernel<T>), dim3((npar + 255) / 256), dim3(256), 0, ctx->stream, npar, dcoef,
                           dalpha, dc64, da64);
        AOK_HIP(ctx, hipGetLastError());
        return out_stage<T>(ctx, c, sh, dc64, da64, 1, npar);
    }
    return AOK_OK;
}

// Shape of a call: validation and the stripe plan.
// Stripe chunks of one call.  The pullback and the position cotangents run their light
// chunks in the l <= OVL_LIGHT_L instance (their wide instances hold one or two waves per
// SIMD); the forward keeps the fewer, longer chunks of the unsplit plan in one instance.
static aok_status shape_of(aok_ctx ctx, const Ovl2cCall &c, OvlShape &sh, bool split_light)
{
    const int knob = split_light && ctx->tune.ovl_cap_pb > 0 ? ctx->tune.ovl_cap_pb : ctx->tune.ovl_cap;
    const int cap = knob > 0 ? knob : (split_light ? OVL_STRIPE_CAP_PB : OVL_STRIPE_CAP);
    const std::string msg =
        build_shape(c.ns, c.K, c.NZ, c.ls, AOK_OVL_LMAX, cap, sh, split_light && ctx->tune.ovl_split != 2);
    if (!msg.empty()) return fail(ctx, msg.find("shell l") != std::string::npos ? AOK_ERR_LMAX : AOK_ERR_ARG, msg);
    if (c.B < 0) return fail(ctx, AOK_ERR_ARG, "B must be >= 0");
    if (c.B > (int64_t(1) << 30)) return fail(ctx, AOK_ERR_ARG, "B too large");
    return AOK_OK;
}

template <typename T, int LMAXB>
static aok_status fast_dx_launch(aok_ctx ctx, const Ovl2cCall &c, const OvlShape &sh, const Ovl2cWs &w, char *ws,
                                 double *partial, int chunk0, int nch)
{
    if (nch <= 0) return AOK_OK;
    using R = typename RecOf<T>::type;
    const int grp = grid_group(ctx, sh, w, nch);
    const size_t shm = 256 + (bwd_direct<T, LMAXB>() ? 0 : sizeof(T) * OVL_TILE * sh.str);
    auto kern = &ovl2c_dx_fast<T, LMAXB>;
    if (shm > 64 * 1024)
        AOK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm)));
    hipLaunchKernelGGL(kern, launch_grid(w, nch, grp), dim3(OVL_TILE), shm, ctx->stream, sh, c.B,
                       static_cast<const T *>(c.dS), static_cast<const T *>(c.posA), static_cast<const T *>(c.posB),
                       static_cast<T>(c.scale), reinterpret_cast<const int *>(ws + w.perm),
                       reinterpret_cast<const int *>(ws + w.tile_key), reinterpret_cast<const int *>(ws + w.boff),
                       reinterpret_cast<const int *>(ws + w.bcnt), reinterpret_cast<const int *>(ws + w.toff),
                       reinterpret_cast<const R *>(ws + w.rec), reinterpret_cast<const int *>(ws + w.nreal),
                       partial, chunk0, nch, grp);
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

// The light chunks in the l <= OVL_LIGHT_L instance (with their own stripe length), the
// rest in the instance for the basis' largest l.  LAUNCH<T, L>::run(ctx, c, shape, w, ws,
// x..., chunk0, nch) launches one instance over chunks chunk0 .. chunk0 + nch - 1.
template <typename T, template <typename, int> class LAUNCH, typename... X>
static aok_status launch_split(aok_ctx ctx, const Ovl2cCall &c, const OvlShape &sh, const Ovl2cWs &w, char *ws,
                               X... x)
{
    const int lmax = lmax_of(c.ls, c.ns);
    if (lmax == 0) return LAUNCH<T, 0>::run(ctx, c, sh, w, ws, x..., 0, sh.nchunk);
    if (lmax == 1) return LAUNCH<T, 1>::run(ctx, c, sh, w, ws, x..., 0, sh.nchunk);
    if (lmax == 2) return LAUNCH<T, 2>::run(ctx, c, sh, w, ws, x..., 0, sh.nchunk);
    if (sh.nlight > 0) {
        OvlShape shl = sh;
        shl.str = sh.strl;
        const aok_status st = LAUNCH<T, OVL_LIGHT_L>::run(ctx, c, shl, w, ws, x..., 0, sh.nlight);
        if (st != AOK_OK) return st;
    }
    const int nh = sh.nchunk - sh.nlight;
    if (lmax == 3) return LAUNCH<T, 3>::run(ctx, c, sh, w, ws, x..., sh.nlight, nh);
    return LAUNCH<T, 4>::run(ctx, c, sh, w, ws, x..., sh.nlight, nh);
}

template <typename T, int L>
struct FwdL {
    template <typename... A>
    static aok_status run(A &&...a) { return fast_fwd_launch<T, L>(a...); }
};
template <typename T, int L>
struct BwdL {
    template <typename... A>
    static aok_status run(A &&...a) { return fast_bwd_launch<T, L>(a...); }
};
template <typename T, int L>
struct DxL {
    template <typename... A>
    static aok_status run(A &&...a) { return fast_dx_launch<T, L>(a...); }
};

// Position cotangents on a prepared workspace; `partial` holds nchunk * B * 3 doubles.
template <typename T>
static aok_status fast_dx(aok_ctx ctx, const Ovl2cCall &c, const OvlShape &sh, const Ovl2cWs &w, char *ws,
                          double *partial)
{
    if (c.B == 0) return AOK_OK;
    time_begin(ctx, "ovl2c_dx_fast");
    aok_status st = launch_split<T, DxL>(ctx, c, sh, w, ws, partial);
    time_end(ctx);
    if (st != AOK_OK) return st;
    hipLaunchKernelGGL((ovl2c_dx_reduce<T>), dim3(static_cast<unsigned>((c.B + 255) / 256)), dim3(256), 0,
                       ctx->stream, c.B, sh.nchunk, reinterpret_cast<const int *>(ws + w.perm), partial, c.scale,
                       static_cast<T *>(c.dA), static_cast<T *>(c.dB));
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

static size_t dx_partial_bytes(const OvlShape &sh, int64_t B)
{
    return align256(sizeof(double) * 3 * static_cast<size_t>(sh.nchunk) * static_cast<size_t>(B > 0 ? B : 1));
}

template <typename T>
static aok_status fast_forward(aok_ctx ctx, const Ovl2cCall &c, const OvlShape &sh, const Ovl2cWs &w, char *ws)
{
    if (c.B == 0) return AOK_OK;
    time_begin(ctx, "ovl2c_fwd_fast");
    const aok_status st = launch_split<T, FwdL>(ctx, c, sh, w, ws);
    time_end(ctx);
    return st;
}

template <typename T>
static aok_status fast_pullback(aok_ctx ctx, const Ovl2cCall &c, const OvlShape &sh, const Ovl2cWs &w, char *ws)
{
    aok_status st = AOK_OK;
    const int64_t npar = static_cast<int64_t>(sh.NZ) * sh.ns * sh.K;
    if (c.B == 0)  // no pairs: zero cotangents (sums over no parts)
        return out_stage<T>(ctx, c, sh, nullptr, nullptr, 0, npar);
    time_begin(ctx, "ovl2c_bwd_fast");
    st = launch_split<T, BwdL>(ctx, c, sh, w, ws);
    time_end(ctx);
    if (st != AOK_OK) return st;
    if ((st = bwd_reduce_launch(ctx, sh, w, ws)) != AOK_OK) return st;
    const double *part = reinterpret_cast<const double *>(ws + w.part2);
    return out_stage<T>(ctx, c, sh, part, part + 2 * sh.NZ * npar, 2 * sh.NZ, npar);
}

template <typename T>
static aok_status plain_run(aok_ctx ctx, const Ovl2cCall &c, const OvlShape &sh, bool pullback)
{
    switch (lmax_of(c.ls, c.ns)) {
        case 0: return plain_launch<T, 0>(ctx, c, sh, pullback);
        case 1: return plain_launch<T, 1>(ctx, c, sh, pullback);
        case 2: return plain_launch<T, 2>(ctx, c, sh, pullback);
        case 3: return plain_launch<T, 3>(ctx, c, sh, pullback);
        default: return plain_launch<T, 4>(ctx, c, sh, pullback);
    }
}

// Entry used by the C ABI: validation, workspace, dispatch on the basis' largest l.
template <typename T>
aok_status ovl2c_run(aok_ctx ctx, const Ovl2cCall &c, bool pullback)
{
    OvlShape sh;
    aok_status st = shape_of(ctx, c, sh, pullback);
    if (st != AOK_OK) return st;
    if (c.flags & AOK_PLAIN) return plain_run<T>(ctx, c, sh, pullback);
    const Ovl2cWs w = ovl2c_ws_map(c.B, sh.ns, sh.K, sh.NZ, sh.nchunk, sizeof(typename RecOf<T>::type),
                                   sizeof(PairPb<T>), pullback);
    st = ensure_ws(ctx, w.total);
    if (st != AOK_OK) return st;
    char *ws = static_cast<char *>(ctx->ws);
    if (c.B > 0 && (st = fast_prep<T>(ctx, c, sh, w, ws)) != AOK_OK) return st;
    return pullback ? fast_pullback<T>(ctx, c, sh, w, ws) : fast_forward<T>(ctx, c, sh, w, ws);
}

// ---------------------------------------------------------------------------------
// Allocation-free forward (aok_ovl2c_noalloc): the same prep and forward kernels, with
// the bucketing tables and primitive-pair records in a fixed per-device array instead of
// the context workspace.  Host code never passes `scratch` itself (its host-side name is
// not a device address): each device's copy is looked up once with hipGetSymbolAddress,
// with that device current, and kept in scratch_ptr.
// ---------------------------------------------------------------------------------
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[AOK_NOALLOC_SCRATCH_BYTES];
static std::mutex scratch_mu;
static void *scratch_ptr[NOALLOC_MAX_DEVICES];

aok_status ovl2c_noalloc_setup(aok_ctx ctx)
{
    const int d = ctx->device;
    if (d < 0 || d >= NOALLOC_MAX_DEVICES) return fail(ctx, AOK_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_ptr[d]) return AOK_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        return fail(ctx, AOK_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    scratch_ptr[d] = p;
    return AOK_OK;
}

template <typename T>
aok_status ovl2c_run_noalloc(aok_ctx ctx, const Ovl2cCall &c)
{
    if (c.flags & AOK_PLAIN) return fail(ctx, AOK_ERR_ARG, "aok_ovl2c_noalloc runs the fast kernels only (flags 0)");
    OvlShape sh;
    aok_status st = shape_of(ctx, c, sh, false);
    if (st != AOK_OK) return st;
    const Ovl2cWs w = ovl2c_ws_map(c.B, sh.ns, sh.K, sh.NZ, sh.nchunk, sizeof(typename RecOf<T>::type),
                                   sizeof(PairPb<T>), false);
    if (w.total > AOK_NOALLOC_SCRATCH_BYTES)
        return fail(ctx, AOK_ERR_ARG,
                    "aok_ovl2c_noalloc: this batch needs " + std::to_string(w.total) +
                        " bytes of scratch, the fixed maximum is " + std::to_string(size_t(AOK_NOALLOC_SCRATCH_BYTES)) +
                        ": split the batch");
    if ((st = ovl2c_noalloc_setup(ctx)) != AOK_OK) return st;
    char *ws = static_cast<char *>(scratch_ptr[ctx->device]);
    if (c.B == 0) return AOK_OK;
    // untimed: timing would draw events from the context pool, which may create some
    const bool timing = ctx->timing;
    ctx->timing = false;
    st = fast_prep<T>(ctx, c, sh, w, ws);
    if (st == AOK_OK) st = fast_forward<T>(ctx, c, sh, w, ws);
    ctx->timing = timing;
    return st;
}

// ---------------------------------------------------------------------------------
// Plans: the species-pair bucketing and the primitive-pair records of one (sidx, params)
// set, built once and reused by forward and pullback calls (the rrule's forward and
// its pullback see the same batch and parameters).
// ---------------------------------------------------------------------------------
template <typename T>
aok_status ovl2c_plan_build(aok_ctx ctx, const Ovl2cCall &c, Ovl2cPlanState &p)
{
    aok_status st = shape_of(ctx, c, p.sh, true);
    if (st == AOK_OK) st = shape_of(ctx, c, p.shf, false);
    if (st != AOK_OK) return st;
    p.call = c;
    if (c.flags & AOK_PLAIN) return AOK_OK;
    p.w = ovl2c_ws_map(c.B, p.sh.ns, p.sh.K, p.sh.NZ, p.sh.nchunk, sizeof(typename RecOf<T>::type),
                          sizeof(PairPb<T>), true);
    if (p.w.total > p.bytes) {
        if (in_capture(ctx))
            return fail(ctx, AOK_ERR_ARG,
                        "a plan workspace of " + std::to_string(p.w.total) +
                            " bytes is needed inside a graph capture: build the plan before capturing");
        if (p.buf) {
            void *old = p.buf;
            p.buf = nullptr;
            p.bytes = 0;
            st = release_buffer(ctx, old);
            if (st != AOK_OK) return st;
        }
        if (hipMalloc(&p.buf, p.w.total) != hipSuccess) {
            p.buf = nullptr;
            return fail(ctx, AOK_ERR_NOMEM, "plan workspace allocation failed");
        }
        p.bytes = p.w.total;
    }
    if (c.B > 0) return fast_prep<T>(ctx, c, p.sh, p.w, static_cast<char *>(p.buf));
    return AOK_OK;
}

// Rebuild a plan's bucketing and records from the species indices and parameters it was
// created with, as they are now in device memory (same sizes and pointers).
template <typename T>
aok_status ovl2c_plan_refresh(aok_ctx ctx, Ovl2cPlanState &p)
{
    if ((p.call.flags & AOK_PLAIN) || p.call.B == 0) return AOK_OK;
    return fast_prep<T>(ctx, p.call, p.sh, p.w, static_cast<char *>(p.buf));
}

template <typename T>
aok_status ovl2c_plan_exec(aok_ctx ctx, const Ovl2cPlanState &p, const Ovl2cCall &io, bool pullback)
{
    Ovl2cCall c = p.call;
    c.posA = io.posA;
    c.posB = io.posB;
    c.scale = io.scale;
    c.out = io.out;
    c.dS = io.dS;
    c.o1 = io.o1;
    c.o2 = io.o2;
    if (c.flags & AOK_PLAIN) return plain_run<T>(ctx, c, p.sh, pullback);
    char *ws = static_cast<char *>(p.buf);
    return pullback ? fast_pullback<T>(ctx, c, p.sh, p.w, ws) : fast_forward<T>(ctx, c, p.shf, p.w, ws);
}

template aok_status ovl2c_plan_build<float>(aok_ctx, const Ovl2cCall &, Ovl2cPlanState &);
template aok_status ovl2c_plan_build<double>(aok_ctx, const Ovl2cCall &, Ovl2cPlanState &);
template aok_status ovl2c_plan_refresh<float>(aok_ctx, Ovl2cPlanState &);
template aok_status ovl2c_plan_refresh<double>(aok_ctx, Ovl2cPlanState &);
template aok_status ovl2c_plan_exec<float>(aok_ctx, const Ovl2cPlanState &, const Ovl2cCall &, bool);
template aok_status ovl2c_plan_exec<double>(aok_ctx, const Ovl2cPlanState &, const Ovl2cCall &, bool);

template <typename T>
aok_status compile_coef_run(aok_ctx ctx, const void *zeta, const void *D, int64_t ns, int64_t K, int64_t NZ,
                            const int64_t *ls, void *coef)
{
    OvlShape sh;
    const std::string msg = build_shape(ns, K, NZ, ls, 64, 1 << 30, sh);
    if (!msg.empty()) return fail(ctx, AOK_ERR_ARG, msg);
    hipLaunchKernelGGL((compile_coef_kernel<T>), dim3(sh.NZ * sh.ns), dim3(256), 0, ctx->stream, sh,
                       static_cast<const T *>(zeta), static_cast<const T *>(D), static_cast<T *>(coef));
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

template <typename T>
aok_status compile_coef_pb_run(aok_ctx ctx, const void *dcoef, const void *zeta, const void *D, int64_t ns,
                               int64_t K, int64_t NZ, const int64_t *ls, void *dD, void *dzeta_norm)
{
    OvlShape sh;
    const std::string msg = build_shape(ns, K, NZ, ls, 64, 1 << 30, sh);
    if (!msg.empty()) return fail(ctx, AOK_ERR_ARG, msg);
    const int64_t npar = static_cast<int64_t>(NZ) * ns * K;
    aok_status st = ensure_ws(ctx, align256(sizeof(double) * npar));
    if (st != AOK_OK) return st;
    double *dc64 = static_cast<double *>(ctx->ws);
    hipLaunchKernelGGL((widen_kernel<T>), dim3((npar + 255) / 256), dim3(256), 0, ctx->stream, npar,
                       static_cast<const T *>(dcoef), static_cast<const T *>(dcoef), dc64, dc64);
    hipLaunchKernelGGL((ovl2c_vjp_out<T>), dim3(sh.NZ * sh.ns), dim3(256), 0, ctx->stream, sh, dc64,
                       static_cast<const double *>(nullptr), 1, npar, static_cast<const T *>(zeta),
                       static_cast<const T *>(D), static_cast<T *>(dzeta_norm), static_cast<T *>(dD));
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

// Position cotangents, one-shot: the fast path's prep, then the partials after it in the
// context workspace.
template <typename T>
aok_status ovl2c_dx_run(aok_ctx ctx, const Ovl2cCall &c)
{
    OvlShape sh;
    aok_status st = shape_of(ctx, c, sh, true);
    if (st != AOK_OK) return st;
    if (c.flags & AOK_PLAIN) return fail(ctx, AOK_ERR_ARG, "position cotangents are computed by the fast path only");
    const Ovl2cWs w = ovl2c_ws_map(c.B, sh.ns, sh.K, sh.NZ, sh.nchunk, sizeof(typename RecOf<T>::type),
                                   sizeof(PairPb<T>), false);
    st = ensure_ws(ctx, w.total + dx_partial_bytes(sh, c.B));
    if (st != AOK_OK) return st;
    char *ws = static_cast<char *>(ctx->ws);
    if (c.B == 0) return AOK_OK;
    if ((st = fast_prep<T>(ctx, c, sh, w, ws)) != AOK_OK) return st;
    return fast_dx<T>(ctx, c, sh, w, ws, reinterpret_cast<double *>(ws + w.total));
}

// Position cotangents through a plan: its prep, the partials in the context workspace.
template <typename T>
aok_status ovl2c_plan_dx(aok_ctx ctx, const Ovl2cPlanState &p, const Ovl2cCall &io)
{
    if (p.call.flags & AOK_PLAIN)
        return fail(ctx, AOK_ERR_ARG, "position cotangents are computed by the fast path only");
    if (p.call.B == 0) return AOK_OK;
    aok_status st = ensure_ws(ctx, dx_partial_bytes(p.sh, p.call.B));
    if (st != AOK_OK) return st;
    Ovl2cCall c = p.call;
    c.posA = io.posA;
    c.posB = io.posB;
    c.scale = io.scale;
    c.dS = io.dS;
    c.dA = io.dA;
    c.dB = io.dB;
    return fast_dx<T>(ctx, c, p.sh, p.w, static_cast<char *>(p.buf), static_cast<double *>(ctx->ws));
}

template aok_status ovl2c_dx_run<float>(aok_ctx, const Ovl2cCall &);
template aok_status ovl2c_dx_run<double>(aok_ctx, const Ovl2cCall &);
template aok_status ovl2c_plan_dx<float>(aok_ctx, const Ovl2cPlanState &, const Ovl2cCall &);
template aok_status ovl2c_plan_dx<double>(aok_ctx, const Ovl2cPlanState &, const Ovl2cCall &);
template aok_status ovl2c_run<float>(aok_ctx, const Ovl2cCall &, bool);
template aok_status ovl2c_run_noalloc<float>(aok_ctx, const Ovl2cCall &);
template aok_status ovl2c_run_noalloc<double>(aok_ctx, const Ovl2cCall &);
template aok_status ovl2c_run<double>(aok_ctx, const Ovl2cCall &, bool);
template aok_status compile_coef_run<float>(aok_ctx, const void *, const void *, int64_t, int64_t, int64_t,
                                            const int64_t *, void *);
template aok_status compile_coef_run<double>(aok_ctx, const void *, const void *, int64_t, int64_t, int64_t,
                                             const int64_t *, void *);
template aok_status compile_coef_pb_run<float>(aok_ctx, const void *, const void *, const void *, int64_t, int64_t,
                                               int64_t, const int64_t *, void *, void *);
template aok_status compile_coef_pb_run<double>(aok_ctx, const void *, const void *, const void *, int64_t, int64_t,
                                                int64_t, const int64_t *, void *, void *);

}  // namespace aok
