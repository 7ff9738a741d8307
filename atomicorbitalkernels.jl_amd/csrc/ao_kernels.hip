// ao_kernels.hip -- H1: batched atomic-orbital evaluation, gradients and pullback.
//
// One lane per point.  Radials R_k = r^p sum_m D exp(-zeta f(r)) and the solid
// harmonics stay in registers; the orbital loop walks a wave-uniform run table
// (one run per radial shell for the canonical AtomicOrbitals index maps) so every
// P[i + nX*o] store is a coalesced column segment.  The per-species (zeta, D)
// slice is staged in LDS once per block.
#include "aok_common.hpp"
#include "ao_plan.hpp"
#include "ylm.hpp"

namespace aok {

// One primitive of the LDS parameter table: a = zeta*log2(e) for float (the exponential
// runs on v_exp_f32 in base 2), a = zeta for double; d = D; ad = a * d (the radial's
// derivative sum, so each primitive costs two FMAs after its exponential).
template <typename T>
struct ParamSet {
    T a, d, ad, pad;
};

template <typename T>
__device__ __forceinline__ T exp_param(T a, T f);
template <>
__device__ __forceinline__ float exp_param<float>(float a, float f)
{
    return __builtin_amdgcn_exp2f(-a * f);
}
template <>
__device__ __forceinline__ double exp_param<double>(double a, double f)
{
    return exp(-a * f);
}
// factor turning sum(a * d * e) into sum(zeta * d * e)
template <typename T>
__device__ __forceinline__ T a_to_zeta()
{
    return std::is_same<T, float>::value ? T(0.69314718055994530942) : T(1);
}

// (zeta*log2e or zeta, D) per (species, radial, primitive) at q = (sg*nRad + k)*K + m,
// built once per call (FP64 products rounded once) and copied into each block's LDS.
template <typename T>
__global__ void ao_param_kernel(const T *__restrict__ zeta, const T *__restrict__ D, int nRad, int K, int NZ,
                                ParamSet<T> *__restrict__ tab)
{
    const int n = NZ * nRad * K;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int sg = q / (nRad * K);
    const int rem = q - sg * nRad * K;
    const int k = rem / K;
    const int m = rem - k * K;
    const int64_t src = k + static_cast<int64_t>(nRad) * (m + static_cast<int64_t>(K) * sg);
    const double z = static_cast<double>(zeta[src]);
    ParamSet<T> pp;
    const double a = std::is_same<T, float>::value ? z * 1.4426950408889634 : z;
    pp.a = T(a);
    pp.d = D[src];
    pp.ad = T(a * static_cast<double>(D[src]));
    pp.pad = T(0);
    tab[q] = pp;
}

template <typename T>
__device__ __forceinline__ void load_param_table(ParamSet<T> *tab, const ParamSet<T> *__restrict__ gtab, int n)
{
    for (int q = threadIdx.x; q < n; q += blockDim.x) tab[q] = gtab[q];
}

template <typename T>
__device__ __forceinline__ T int_pow(T r, int p)
{
    T v = T(1);
    for (int q = 0; q < p; ++q) v *= r;
    return v;
}

// Radial value R and G = (dR/dr)/r for radial k of species sg at decay argument f.
// Gaussian with p = 0 uses the closed form G = -2 sum zeta D e (finite at r = 0);
// otherwise the reference's dR/dr * (1/r) (NaN at r = 0 like _aorb_ed_ka!).
template <typename T, bool GRAD>
__device__ __forceinline__ void radial(const ParamSet<T> *row, int K, T f, T r, int p, int decay, T &R,
                                       T &G)
{
    T s = T(0), ds = T(0);
    for (int m = 0; m < K; ++m) {
        const ParamSet<T> pp = row[m];
        const T ex = exp_param<T>(pp.a, f);
        s = fma(ex, pp.d, s);
        if (GRAD) ds = fma(ex, pp.ad, ds);
    }
    if (p == 0) {
        R = s;
        if (GRAD) G = (decay == AOK_DECAY_GAUSSIAN) ? T(-2) * a_to_zeta<T>() * ds
                                                    : -a_to_zeta<T>() * ds / r;
    } else {
        const T rp = int_pow(r, p);
        R = rp * s;
        if (GRAD) {
            const T drp = T(p) * int_pow(r, p - 1);
            const T dfdr = (decay == AOK_DECAY_GAUSSIAN) ? T(2) * r : T(1);
            G = (drp * s - rp * dfdr * a_to_zeta<T>() * ds) / r;
        }
    }
}

template <int LL, int L, bool GRAD, typename T>
__device__ __forceinline__ void store_shell(const T *Z, const T *Zx, const T *Zy, const T *Zz, T R, T G,
                                           T x, T y, T z, T scale, int64_t i, int64_t nX, int o0,
                                           T *__restrict__ P, T *__restrict__ dP)
{
    if constexpr (LL <= L) {
#pragma unroll
        for (int j = 0; j <= 2 * LL; ++j) {
            const int yy = LL * LL + j;
            const int64_t col = i + nX * (o0 + j);
            store_nt(P + col, R * Z[yy]);
            if constexpr (GRAD) {
                const T gz = G * Z[yy];
                T *q = dP + 3 * col;
                store_nt(q + 0, scale * (gz * x + R * Zx[yy]));
                store_nt(q + 1, scale * (gz * y + R * Zy[yy]));
                store_nt(q + 2, scale * (gz * z + R * Zz[yy]));
            }
        }
    }
}

// Runtime harmonic index y (non-canonical maps) -> register value, by a uniform
// select chain.
template <int NY, typename T>
__device__ __forceinline__ T pick(const T *v, int y)
{
    T r = v[0];
#pragma unroll
    for (int q = 1; q < NY; ++q) r = (y == q) ? v[q] : r;
    return r;
}

template <typename T, int L, bool GRAD, bool GENERIC>
__global__ __launch_bounds__(256) void ao_eval_kernel(const T *__restrict__ X, const void *__restrict__ sidx,
                                                      int itype, int64_t nX, const ParamSet<T> *__restrict__ gtab,
                                                       int nRad, int K, int NZ, T scale,
                                                      int decay, AoRunTable runs, T *__restrict__ P,
                                                      T *__restrict__ dP, unsigned *__restrict__ err)
{
    extern __shared__ unsigned char smem_raw[];
    ParamSet<T> *tab = reinterpret_cast<ParamSet<T> *>(smem_raw);
    load_param_table(tab, gtab, NZ * nRad * K);
    __syncthreads();
    constexpr int NY = (L + 1) * (L + 1);
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nX;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const T x = scale * X[3 * i], y = scale * X[3 * i + 1], z = scale * X[3 * i + 2];
        int sg = load_sigma(sidx, itype, i);
        if (sg < 0 || sg >= NZ) {
            *err = ERR_SPECIES;
            sg = 0;
        }
        const T r2 = x * x + y * y + z * z;
        const T r = sqrt(r2);
        const T f = (decay == AOK_DECAY_GAUSSIAN) ? r2 : r;
        T Z[NY], Zx[GRAD ? NY : 1], Zy[GRAD ? NY : 1], Zz[GRAD ? NY : 1];
        solid_harmonics<L, GRAD, T>(x, y, z, YlmOut<T>{Z, Zx, Zy, Zz});
        for (int q = 0; q < runs.nrun; ++q) {
            const uint32_t rw = runs.run[q];
            const int k = run_k(rw), lk = run_l(rw), o0 = run_o0(rw);
            T R, G = T(0);
            radial<T, GRAD>(tab + (sg * nRad + k) * K, K, f, r, run_poly(runs, k), decay, R, G);
            switch (lk) {
                case 0: store_shell<0, L, GRAD>(Z, Zx, Zy, Zz, R, G, x, y, z, scale, i, nX, o0, P, dP); break;
                case 1: store_shell<1, L, GRAD>(Z, Zx, Zy, Zz, R, G, x, y, z, scale, i, nX, o0, P, dP); break;
                case 2: store_shell<2, L, GRAD>(Z, Zx, Zy, Zz, R, G, x, y, z, scale, i, nX, o0, P, dP); break;
                case 3: store_shell<3, L, GRAD>(Z, Zx, Zy, Zz, R, G, x, y, z, scale, i, nX, o0, P, dP); break;
                case 4: store_shell<4, L, GRAD>(Z, Zx, Zy, Zz, R, G, x, y, z, scale, i, nX, o0, P, dP); break;
                default:
                    if constexpr (GENERIC) {  // single orbital with harmonic index -1-lk
                    const int yy = -1 - lk;
                    const int64_t col = i + nX * o0;
                    const T zy = pick<NY>(Z, yy);
                    store_nt(P + col, R * zy);
                    if constexpr (GRAD) {
                        const T gz = G * zy;
                        T *qd = dP + 3 * col;
                        store_nt(qd + 0, scale * (gz * x + R * pick<NY>(Zx, yy)));
                        store_nt(qd + 1, scale * (gz * y + R * pick<NY>(Zy, yy)));
                        store_nt(qd + 2, scale * (gz * z + R * pick<NY>(Zz, yy)));
                    }
                    }
            }
        }
    }
}

// Values-only evaluation with PPL consecutive points per lane: every orbital column of
// a wave is one contiguous 256*PPL-byte segment written by vector stores.
template <int PPL, typename T>
struct VecOf;
template <>
struct VecOf<2, float> {
    typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct VecOf<4, float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct VecOf<2, double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct VecOf<4, double> {
    typedef double type __attribute__((ext_vector_type(4)));
};

template <int LL, int L, int PPL, typename T>
__device__ __forceinline__ void store_shell_vec(const T (*Z)[(L + 1) * (L + 1)], const T *R, int64_t i0, int64_t nX,
                                               int o0, T *__restrict__ P)
{
    if constexpr (LL <= L) {
        using V = typename VecOf<PPL, T>::type;
#pragma unroll
        for (int j = 0; j <= 2 * LL; ++j) {
            const int yy = LL * LL + j;
            V v;
#pragma unroll
            for (int p = 0; p < PPL; ++p) v[p] = R[p] * Z[p][yy];
            __builtin_nontemporal_store(v, reinterpret_cast<V *>(P + i0 + nX * (o0 + j)));
        }
    }
}

template <typename T, int L, int PPL>
__global__ __launch_bounds__(256) void ao_eval_vec_kernel(const T *__restrict__ X, const void *__restrict__ sidx,
                                                          int itype, int64_t nX, const ParamSet<T> *__restrict__ gtab,
                                                           int nRad, int K, int NZ, T scale,
                                                          int decay, AoRunTable runs, T *__restrict__ P,
                                                          unsigned *__restrict__ err)
{
    extern __shared__ unsigned char smem_raw[];
    ParamSet<T> *tab = reinterpret_cast<ParamSet<T> *>(smem_raw);
    load_param_table(tab, gtab, NZ * nRad * K);
    __syncthreads();
    constexpr int NY = (L + 1) * (L + 1);
    const int64_t ntile = (nX + 256 * PPL - 1) / (256 * PPL);
    for (int64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
        const int64_t i0 = (t * 256 + threadIdx.x) * PPL;
        if (i0 >= nX) continue;  // nX is a multiple of PPL (host-checked)
        T Z[PPL][NY], f[PPL], r[PPL];
        int sg[PPL];
#pragma unroll
        for (int p = 0; p < PPL; ++p) {
            const int64_t i = i0 + p;
            const T x = scale * X[3 * i], y = scale * X[3 * i + 1], z = scale * X[3 * i + 2];
            int s = load_sigma(sidx, itype, i);
            if (s < 0 || s >= NZ) {
                *err = ERR_SPECIES;
                s = 0;
            }
            sg[p] = s;
            const T r2 = x * x + y * y + z * z;
            r[p] = sqrt(r2);
            f[p] = (decay == AOK_DECAY_GAUSSIAN) ? r2 : r[p];
            solid_harmonics<L, false, T>(x, y, z, YlmOut<T>{Z[p], nullptr, nullptr, nullptr});
        }
        for (int q = 0; q < runs.nrun; ++q) {
            const uint32_t rw = runs.run[q];
            const int k = run_k(rw), lk = run_l(rw), o0 = run_o0(rw);
            T R[PPL];
#pragma unroll
            for (int p = 0; p < PPL; ++p) {
                T G;
                radial<T, false>(tab + (sg[p] * nRad + k) * K, K, f[p], r[p], run_poly(runs, k), decay, R[p], G);
            }
            switch (lk) {
                case 0: store_shell_vec<0, L, PPL>(Z, R, i0, nX, o0, P); break;
                case 1: store_shell_vec<1, L, PPL>(Z, R, i0, nX, o0, P); break;
                case 2: store_shell_vec<2, L, PPL>(Z, R, i0, nX, o0, P); break;
                case 3: store_shell_vec<3, L, PPL>(Z, R, i0, nX, o0, P); break;
                case 4: store_shell_vec<4, L, PPL>(Z, R, i0, nX, o0, P); break;
                default: break;
            }
        }
    }
}

// Pullback contraction of one full shell run: dr += sum_j w_j Z_j and W_y += w_j R for
// its harmonics y (the spatial gradient is contracted once per point from W).
// The point's cotangent of column o is src[cs * o] (src = dPin + i, cs = nX).
template <int LL, int L, typename T>
__device__ __forceinline__ void contract_shell(const T *Z, T *W, T R, const T *src, int64_t cs, int o0, T &dr)
{
    if constexpr (LL <= L) {
#pragma unroll
        for (int j = 0; j <= 2 * LL; ++j) {
            const int yy = LL * LL + j;
            const T w = src[cs * (o0 + j)];
            dr += w * Z[yy];
            W[yy] += w * R;
        }
    }
}

// Prefetched shell runs of the pullback: the cotangents of a group of runs are loaded
// into registers (load_run) before their radials are formed, and contracted after
// (use_run), so the loads overlap the exponentials.  Measured on nX = 2^24, 10 radials
// x 4 primitives, L = 3 (FP32): one run per group at 6 waves/SIMD 0.89 ms; at 7 waves
// the compiler serialises the loads (0.92 ms); two or three runs per group at 4-6 waves
// 0.94-1.67 ms.
template <int LL, int MJ, typename T>
__device__ __forceinline__ void load_run(T (&v)[MJ], const T *src, int64_t cs, int o0)
{
    if constexpr (2 * LL + 1 <= MJ) {
#pragma unroll
        for (int j = 0; j <= 2 * LL; ++j) v[j] = src[cs * (o0 + j)];
    }
}
template <int LL, int L, int MJ, typename T>
__device__ __forceinline__ void use_run(const T *Z, T *W, T R, const T (&v)[MJ], T &dr)
{
    if constexpr (LL <= L && 2 * LL + 1 <= MJ) {
#pragma unroll
        for (int j = 0; j <= 2 * LL; ++j) {
            dr += v[j] * Z[LL * LL + j];
            W[LL * LL + j] += v[j] * R;
        }
    }
}
constexpr int AO_PB_RUN_GROUP = 1;

// ---------------------------------------------------------------------------------
// Pullback, one 256-point tile per block iteration.
//  Phase A (one lane per point): dX = s * sum_o dP_io grad P_io, and dR_k =
//    sum_{o in k} dP_io Z_io into an LDS row; rows are ordered by species (the point's
//    rank among the tile's points of its species, by wave ballots).
//  Phase B (one lane per parameter item): item = (sub, species, k, m); its lane walks
//    its share of that species' rows and accumulates
//      dD[k,m,sg] += dR_k r^p e,   dzeta[k,m,sg] -= dR_k r^p D f e,   e = exp(-zeta f)
//    in FP32 over the tile, then adds the tile sum into its own FP64 LDS accumulator.
//  No cross-lane reductions and no atomics: every sum has a fixed order.  The block
//  writes one FP64 partial per item and ao_pullback_finalize sums them (block per slot,
//  fixed LDS tree).
// ---------------------------------------------------------------------------------
constexpr int AO_PB_THREADS = 256;

struct AoPbLayout {
    int nslot, nsub, nitem, nRadP;
    size_t tab, dr, f, cnt, rng, dec, acc, total;
};

// LDS plan of one pullback block.
template <typename T>
__host__ __device__ inline AoPbLayout ao_pb_layout(int nRad, int K, int NZ)
{
    AoPbLayout y;
    y.nslot = NZ * nRad * K;
    y.nsub = y.nslot >= AO_PB_THREADS ? 1 : AO_PB_THREADS / y.nslot;
    y.nitem = y.nslot * y.nsub;
    y.nRadP = nRad | 1;  // odd row stride: rows of different points start in different banks
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = (o + bytes + 15) & ~size_t(15);
        return at;
    };
    y.tab = take(sizeof(ParamSet<T>) * y.nslot);
    y.dr = take(sizeof(T) * AO_PB_THREADS * y.nRadP);
    y.f = take(sizeof(T) * AO_PB_THREADS);
    y.cnt = take(sizeof(int) * (4 * NZ + NZ + 1));
    y.rng = take(sizeof(int) * NZ * (y.nsub + 1));
    y.dec = take(sizeof(int2) * y.nitem);
    y.acc = take(sizeof(double) * 2 * y.nitem);
    y.total = o;
    return y;
}

// Waves per SIMD the FP32 pullback is compiled for: at most 84 VGPRs for L <= 3
// (measured on nX = 2^24, 10 radials x 4 primitives, L = 3: 4 waves/SIMD at the
// compiler's own 99 VGPRs 1.06-1.11 ms, 7 waves 0.89-0.92 ms, 6 waves with the run's
// loads issued before its radial 0.89 ms, the grid matched to the resident blocks).
// FP64 and arbitrary orbital maps keep the compiler's choice.
template <typename T, int L, bool GENERIC>
constexpr int ao_pb_waves()
{
    return std::is_same<T, float>::value && L <= 3 && !GENERIC ? 6 : 1;
}

template <typename T, int L, bool GENERIC>
__global__ __launch_bounds__(AO_PB_THREADS) __attribute__((amdgpu_waves_per_eu(ao_pb_waves<T, L, GENERIC>(), 8)))
void ao_pullback_kernel(
    const T *__restrict__ X, const void *__restrict__ sidx, int itype, int64_t nX, const ParamSet<T> *__restrict__ gtab,
    int nRad, int K, int NZ, T scale, int decay, AoRunTable runs, const T *__restrict__ dPin, T *__restrict__ dX,
    double *__restrict__ partial, int want_params, unsigned *__restrict__ err)
{
    constexpr int NY = (L + 1) * (L + 1);
    constexpr int TILE = AO_PB_THREADS;
    extern __shared__ unsigned char smem_raw[];
    const AoPbLayout y = ao_pb_layout<T>(nRad, K, NZ);
    ParamSet<T> *tab = reinterpret_cast<ParamSet<T> *>(smem_raw + y.tab);
    T *dru = reinterpret_cast<T *>(smem_raw + y.dr);  // [row][nRadP], rows in species order
    T *fu = reinterpret_cast<T *>(smem_raw + y.f);
    int *whist = reinterpret_cast<int *>(smem_raw + y.cnt);  // [4 waves][NZ]
    int *soff = whist + 4 * NZ;                              // [NZ + 1]
    int *rng = reinterpret_cast<int *>(smem_raw + y.rng);    // [NZ][nsub + 1] row bounds of the subsets
    int2 *idec = reinterpret_cast<int2 *>(smem_raw + y.dec);  // item -> (u | p << 8 | g << 16, k)
    double *acc = reinterpret_cast<double *>(smem_raw + y.acc);

    load_param_table(tab, gtab, y.nslot);
    // parameter items decoded once per block: item = (sub u, slot), slot = (g * nRad + k) * K + m
    for (int q = threadIdx.x; q < y.nitem; q += AO_PB_THREADS) {
        acc[2 * q] = 0.0;
        acc[2 * q + 1] = 0.0;
        const int u = q / y.nslot, slot = q - u * y.nslot;
        const int g = slot / (nRad * K), k = (slot - g * nRad * K) / K;
        idec[q] = make_int2(u | (run_poly(runs, k) << 8) | (g << 16), k);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

    const int64_t ntiles = (nX + TILE - 1) / TILE;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t i = t * TILE + threadIdx.x;
        int sg = -1;
        if (i < nX) {
            sg = load_sigma(sidx, itype, i);
            if (sg < 0 || sg >= NZ) {
                *err = ERR_SPECIES;
                sg = 0;
            }
        }
        // row of this point: its position in the tile's species-ordered point list
        // (rank among the wave's points of the same species by ballots, plus the counts
        // of earlier species and waves), so phase B reads each species' rows contiguously
        int pos = threadIdx.x;
        if (want_params) {
            for (int q = threadIdx.x; q < 4 * NZ; q += AO_PB_THREADS) whist[q] = 0;
            __syncthreads();
            unsigned long long todo = __ballot(sg >= 0);
            int rk = 0;
            while (todo) {
                const int leader = __ffsll(static_cast<long long>(todo)) - 1;
                const int ls = __shfl(sg, leader, 64);
                const unsigned long long m = __ballot(sg == ls);
                if (sg == ls) rk = __popcll(m & ((1ull << lane) - 1ull));
                if (lane == leader) whist[wave * NZ + ls] = __popcll(m);
                todo &= ~m;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                int o = 0;
                for (int g = 0; g < NZ; ++g) {
                    soff[g] = o;
                    o += whist[g] + whist[NZ + g] + whist[2 * NZ + g] + whist[3 * NZ + g];
                }
                soff[NZ] = o;
            }
            __syncthreads();
            if (sg >= 0) {
                pos = soff[sg] + rk;
                for (int w = 0; w < wave; ++w) pos += whist[w * NZ + sg];
            }
            // bounds of the nsub row subsets of every species
            const int nb = y.nsub + 1;
            for (int q = threadIdx.x; q < NZ * nb; q += AO_PB_THREADS) {
                const int g = q / nb, u = q - g * nb;
                const int n0 = soff[g], cnt = soff[g + 1] - n0;
                rng[q] = n0 + (cnt * u) / y.nsub;
            }
        }
        // ---- phase A ----
        if (i < nX) {
            const T *src = dPin + i;
            const int64_t cs = nX;
            T *myrow = dru + pos * y.nRadP;
            for (int k = 0; k < nRad; ++k) myrow[k] = T(0);
            const T x = scale * X[3 * i], y3 = scale * X[3 * i + 1], z = scale * X[3 * i + 2];
            const T r2 = x * x + y3 * y3 + z * z;
            const T r = sqrt(r2);
            const T f = (decay == AOK_DECAY_GAUSSIAN) ? r2 : r;
            // dX = s (sum_k G_k dR_k) x + s sum_y W_y grad Z_y,  W_y = sum_{o: iY[o]=y} dP_io R_iR[o]
            T Z[NY], W[NY];
            solid_harmonics<L, false, T>(x, y3, z, YlmOut<T>{Z, nullptr, nullptr, nullptr});
#pragma unroll
            for (int q = 0; q < NY; ++q) W[q] = T(0);
            T sgd = T(0);
            if constexpr (!GENERIC) {
                constexpr int MJ = 2 * L + 1, RG = AO_PB_RUN_GROUP;
                for (int q = 0; q < runs.nrun; q += RG) {
                    T v[RG][MJ];
#pragma unroll
                    for (int g = 0; g < RG; ++g) {
                        if (q + g >= runs.nrun) break;
                        const uint32_t rw = runs.run[q + g];
                        const int o0 = run_o0(rw);
                        switch (run_l(rw)) {
                            case 0: load_run<0>(v[g], src, cs, o0); break;
                            case 1: load_run<1>(v[g], src, cs, o0); break;
                            case 2: load_run<2>(v[g], src, cs, o0); break;
                            case 3: load_run<3>(v[g], src, cs, o0); break;
                            case 4: load_run<4>(v[g], src, cs, o0); break;
                        }
                    }
#pragma unroll
                    for (int g = 0; g < RG; ++g) {
                        if (q + g >= runs.nrun) break;
                        const uint32_t rw = runs.run[q + g];
                        const int k = run_k(rw);
                        T R, G;
                        radial<T, true>(tab + (sg * nRad + k) * K, K, f, r, run_poly(runs, k), decay, R, G);
                        T dr = T(0);
                        switch (run_l(rw)) {
                            case 0: use_run<0, L>(Z, W, R, v[g], dr); break;
                            case 1: use_run<1, L>(Z, W, R, v[g], dr); break;
                            case 2: use_run<2, L>(Z, W, R, v[g], dr); break;
                            case 3: use_run<3, L>(Z, W, R, v[g], dr); break;
                            case 4: use_run<4, L>(Z, W, R, v[g], dr); break;
                        }
                        sgd += G * dr;
                        myrow[k] += dr;
                    }
                }
            }
            for (int q = 0; GENERIC && q < runs.nrun; ++q) {
                const uint32_t rw = runs.run[q];
                const int k = run_k(rw), lk = run_l(rw), o0 = run_o0(rw);
                T R, G;
                radial<T, true>(tab + (sg * nRad + k) * K, K, f, r, run_poly(runs, k), decay, R, G);
                T dr = T(0);
                switch (lk) {
                    case 0: contract_shell<0, L>(Z, W, R, src, cs, o0, dr); break;
                    case 1: contract_shell<1, L>(Z, W, R, src, cs, o0, dr); break;
                    case 2: contract_shell<2, L>(Z, W, R, src, cs, o0, dr); break;
                    case 3: contract_shell<3, L>(Z, W, R, src, cs, o0, dr); break;
                    case 4: contract_shell<4, L>(Z, W, R, src, cs, o0, dr); break;
                    default:
                        if constexpr (GENERIC) {
                            const int yy = -1 - lk;
                            const T w = src[cs * o0];
                            dr += w * pick<NY>(Z, yy);
#pragma unroll
                            for (int q2 = 0; q2 < NY; ++q2) W[q2] += (q2 == yy) ? w * R : T(0);
                        }
                }
                sgd += G * dr;
                myrow[k] += dr;
            }
            if (dX != nullptr) {
                T gx = sgd * x, gy = sgd * y3, gz = sgd * z;
                solid_harmonics_grad_dot<L, T>(x, y3, z, W, gx, gy, gz);
                dX[3 * i] = scale * gx;
                dX[3 * i + 1] = scale * gy;
                dX[3 * i + 2] = scale * gz;
            }
            fu[pos] = f;
        }
        if (!want_params) continue;
        __syncthreads();
        // ---- phase B ----
        for (int it = threadIdx.x; it < y.nitem; it += AO_PB_THREADS) {
            const int2 dc = idec[it];
            const int u = dc.x & 0xff, pk = (dc.x >> 8) & 0xff, g = dc.x >> 16, k = dc.y;
            const ParamSet<T> pp = tab[it - u * y.nslot];
            const int j0 = rng[g * (y.nsub + 1) + u], j1 = rng[g * (y.nsub + 1) + u + 1];
            const T *col = dru + k;
            T sD = T(0), sZ = T(0);
            auto term = [&](int j) {
                const T f = fu[j];
                T dr = col[j * y.nRadP];
                if (pk != 0) dr *= int_pow((decay == AOK_DECAY_GAUSSIAN) ? sqrt(f) : f, pk);
                const T w = dr * exp_param<T>(pp.a, f);
                sD += w;
                sZ = fma(w, f, sZ);
            };
            int j = j0;
            for (; j + 4 <= j1; j += 4) {
                term(j);
                term(j + 1);
                term(j + 2);
                term(j + 3);
            }
            for (; j < j1; ++j) term(j);
            acc[2 * it] += static_cast<double>(sD);
            acc[2 * it + 1] -= static_cast<double>(pp.d) * static_cast<double>(sZ);  // dzeta = -D sum w f
        }
        __syncthreads();  // LDS rows are rewritten by the next tile
    }
    if (!want_params) return;
    // partial[b][item][2] (items of this block; the accumulators are owned per thread)
    for (int q = threadIdx.x; q < 2 * y.nitem; q += AO_PB_THREADS)
        partial[static_cast<int64_t>(blockIdx.x) * 2 * y.nitem + q] = acc[q];
}

// Per slot: sum its nsub items over all blocks in a fixed order (block per slot, thread
// stride over blocks, LDS tree), then scatter to the Julia [nRad x K x NZ] layout.
template <typename T>
__global__ __launch_bounds__(256) void ao_pullback_finalize(const double *__restrict__ partial, int nblk, int nRad,
                                                            int K, int NZ, int nsub, T *__restrict__ dzeta,
                                                            T *__restrict__ dD)
{
    __shared__ double rd[256], rz[256];
    const int nslot = NZ * nRad * K, nitem = nslot * nsub;
    const int q = blockIdx.x;
    double sD = 0.0, sZ = 0.0;
    if (partial)
        for (int b = threadIdx.x; b < nblk; b += 256)
            for (int u = 0; u < nsub; ++u) {
                const double *pb = partial + (static_cast<int64_t>(b) * nitem + u * nslot + q) * 2;
                sD += pb[0];
                sZ += pb[1];
            }
    rd[threadIdx.x] = sD;
    rz[threadIdx.x] = sZ;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) {
            rd[threadIdx.x] += rd[threadIdx.x + h];
            rz[threadIdx.x] += rz[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const int sg = q / (nRad * K);
    const int rem = q - sg * nRad * K;
    const int k = rem / K, m = rem - k * K;
    const int64_t dst = k + static_cast<int64_t>(nRad) * (m + static_cast<int64_t>(K) * sg);
    if (dD) dD[dst] = static_cast<T>(rd[0]);
    if (dzeta) dzeta[dst] = static_cast<T>(rz[0]);
}

// ---------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------
// The per-call parameter table, at byte offset `at` of the context workspace.
template <typename T>
static aok_status param_table(aok_ctx ctx, const AoArgs &a, size_t at, const ParamSet<T> *&gtab)
{
    const int n = a.NZ * a.nRad * a.K;
    aok_status st = ensure_ws(ctx, at + sizeof(ParamSet<T>) * n);
    if (st != AOK_OK) return st;
    ParamSet<T> *t = reinterpret_cast<ParamSet<T> *>(static_cast<char *>(ctx->ws) + at);
    hipLaunchKernelGGL((ao_param_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<const T *>(a.zeta), static_cast<const T *>(a.D), a.nRad, a.K, a.NZ, t);
    AOK_HIP(ctx, hipGetLastError());
    gtab = t;
    return AOK_OK;
}

template <typename T, int L, bool GENERIC>
static aok_status launch_eval_L(aok_ctx ctx, const AoArgs &a, const AoRunTable &runs)
{
    const size_t shm = sizeof(ParamSet<T>) * a.NZ * a.nRad * a.K;
    if (a.nX <= 0) return AOK_OK;
    const ParamSet<T> *gtab = nullptr;
    aok_status st = param_table<T>(ctx, a, 0, gtab);
    if (st != AOK_OK) return st;
    // launch shape: points per lane (values-only, canonical maps) and the grid cap.  The
    // defaults (one point per lane, one 256-point tile per block) measured best on MI355X.
    int ppl = ctx->tune.ao_ppl > 0 ? ctx->tune.ao_ppl : 1;
    if (a.dP || GENERIC || a.nX % ppl != 0 || (ppl != 2 && ppl != 4)) ppl = 1;
    const int64_t ntile = (a.nX + 256 * ppl - 1) / (256 * ppl);
    const int cap = ctx->tune.ao_grid > 0 ? ctx->tune.ao_grid : (1 << 20);
    const int grid = static_cast<int>(std::min<int64_t>(ntile, cap));
    time_begin(ctx, a.dP ? "ao_eval_ed" : "ao_eval");
    if (a.dP)
        hipLaunchKernelGGL((ao_eval_kernel<T, L, true, GENERIC>), dim3(grid), dim3(256), shm, ctx->stream,
                           static_cast<const T *>(a.X), a.sidx, a.itype, a.nX, gtab,
                            a.nRad, a.K, a.NZ, static_cast<T>(a.scale), a.decay, runs,
                           static_cast<T *>(a.P), static_cast<T *>(a.dP), ctx->d_err);
    else if (ppl == 4)
        hipLaunchKernelGGL((ao_eval_vec_kernel<T, L, 4>), dim3(grid), dim3(256), shm, ctx->stream,
                           static_cast<const T *>(a.X), a.sidx, a.itype, a.nX, gtab,
                            a.nRad, a.K, a.NZ, static_cast<T>(a.scale), a.decay, runs,
                           static_cast<T *>(a.P), ctx->d_err);
    else if (ppl == 2)
        hipLaunchKernelGGL((ao_eval_vec_kernel<T, L, 2>), dim3(grid), dim3(256), shm, ctx->stream,
                           static_cast<const T *>(a.X), a.sidx, a.itype, a.nX, gtab,
                            a.nRad, a.K, a.NZ, static_cast<T>(a.scale), a.decay, runs,
                           static_cast<T *>(a.P), ctx->d_err);
    else
        hipLaunchKernelGGL((ao_eval_kernel<T, L, false, GENERIC>), dim3(grid), dim3(256), shm, ctx->stream,
                           static_cast<const T *>(a.X), a.sidx, a.itype, a.nX, gtab,
                            a.nRad, a.K, a.NZ, static_cast<T>(a.scale), a.decay, runs,
                           static_cast<T *>(a.P), nullptr, ctx->d_err);
    time_end(ctx);
    AOK_HIP(ctx, hipGetLastError());
    return AOK_OK;
}

static bool runs_canonical(const AoRunTable &runs)
{
    for (int q = 0; q < runs.nrun; ++q)
        if (run_l(runs.run[q]) < 0) return false;
    return true;
}

template <typename T, bool G>
static aok_status ao_eval_launch_g(aok_ctx ctx, const AoArgs &a, const AoRunTable &runs)
{
    switch (a.L) {
        case 0: return launch_eval_L<T, 0, G>(ctx, a, runs);
        case 1: return launch_eval_L<T, 1, G>(ctx, a, runs);
        case 2: return launch_eval_L<T, 2, G>(ctx, a, runs);
        case 3: return launch_eval_L<T, 3, G>(ctx, a, runs);
        case 4: return launch_eval_L<T, 4, G>(ctx, a, runs);
    }
    return fail(ctx, AOK_ERR_LMAX, "L out of range");
}

template <typename T>
aok_status ao_eval_launch(aok_ctx ctx, const AoArgs &a, const AoRunTable &runs)
{
    return runs_canonical(runs) ? ao_eval_launch_g<T, false>(ctx, a, runs) : ao_eval_launch_g<T, true>(ctx, a, runs);
}


template <typename T, int L, bool GENERIC>
static aok_status launch_pb_L(aok_ctx ctx, const AoArgs &a, const AoRunTable &runs, void *dX, void *dzeta,
                              void *dD)
{
    const int64_t ntiles = (a.nX + AO_PB_THREADS - 1) / AO_PB_THREADS;
    const AoPbLayout y = ao_pb_layout<T>(a.nRad, a.K, a.NZ);
    const size_t shm = y.total;
    auto kern = &ao_pullback_kernel<T, L, GENERIC>;
    // persistent grid: every block resident at once (blocks per CU from the kernel's
    // registers and LDS x CUs), or the ao_grid knob
    int cap = ctx->tune.ao_grid;
    if (cap <= 0 && ntiles > 0) {
        if (ctx->num_cu <= 0)
            AOK_HIP(ctx, hipDeviceGetAttribute(&ctx->num_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
        if (shm > 64 * 1024 && shm <= 160 * 1024)
            AOK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm)));
        int nb = 0;
        AOK_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void *>(kern),
                                                                  AO_PB_THREADS, shm));
        cap = std::max(1, nb) * std::max(1, ctx->num_cu);
    }
    const int grid = static_cast<int>(std::min<int64_t>(ntiles, cap));
    const bool want = (dzeta != nullptr) || (dD != nullptr);
    if (grid <= 0) {
        if (want) {
            // no points: zero cotangents (the reference returns zero-filled arrays)
            hipLaunchKernelGGL((ao_pullback_finalize<T>), dim3(y.nslot), dim3(256), 0, ctx->stream,
                               static_cast<const double *>(nullptr), 0, a.nRad, a.K, a.NZ, y.nsub,
                               static_cast<T *>(dzeta), static_cast<T *>(dD));
            AOK_HIP(ctx, hipGetLastError());
        }
        return AOK_OK;
    }
    if (shm > 160 * 1024)
        return fail(ctx, AOK_ERR_UNSUPPORTED, "ao_pullback: basis too large for the LDS tile");
    const size_t part_bytes = align256(sizeof(double) * 2 * static_cast<size_t>(y.nitem) * grid);
    const ParamSet<T> *gtab = nullptr;
    aok_status st = param_table<T>(ctx, a, part_bytes, gtab);
    if (st != AOK_OK) return st;
    double *partial = static_cast<double *>(ctx->ws);
    if (shm > 64 * 1024)
        AOK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm)));
    time_begin(ctx, "ao_pullback");
    hipLaunchKernelGGL(kern, dim3(grid), dim3(AO_PB_THREADS), shm, ctx->stream, static_cast<const T *>(a.X),
                       a.sidx, a.itype, a.nX, gtab, a.nRad, a.K, a.NZ, static_cast<T>(a.scale), a.decay, runs,
                       static_cast<const T *>(a.dPin), static_cast<T *>(dX), partial, want ? 1 : 0, ctx->d_err);
    time_end(ctx);
    AOK_HIP(ctx, hipGetLastError());
    if (want) {
        hipLaunchKernelGGL((ao_pullback_finalize<T>), dim3(y.nslot), dim3(256), 0, ctx->stream,
                           static_cast<const double *>(partial), grid, a.nRad, a.K, a.NZ, y.nsub,
                           static_cast<T *>(dzeta), static_cast<T *>(dD));
        AOK_HIP(ctx, hipGetLastError());
    }
    return AOK_OK;
}

template <typename T, bool G>
static aok_status ao_pullback_launch_g(aok_ctx ctx, const AoArgs &a, const AoRunTable &runs, void *dX, void *dzeta,
                                       void *dD)
{
    switch (a.L) {
        case 0: return launch_pb_L<T, 0, G>(ctx, a, runs, dX, dzeta, dD);
        case 1: return launch_pb_L<T, 1, G>(ctx, a, runs, dX, dzeta, dD);
        case 2: return launch_pb_L<T, 2, G>(ctx, a, runs, dX, dzeta, dD);
        case 3: return launch_pb_L<T, 3, G>(ctx, a, runs, dX, dzeta, dD);
        case 4: return launch_pb_L<T, 4, G>(ctx, a, runs, dX, dzeta, dD);
    }
    return fail(ctx, AOK_ERR_LMAX, "L out of range");
}

template <typename T>
aok_status ao_pullback_launch(aok_ctx ctx, const AoArgs &a, const AoRunTable &runs, void *dX, void *dzeta,
                              void *dD)
{
    return runs_canonical(runs) ? ao_pullback_launch_g<T, false>(ctx, a, runs, dX, dzeta, dD)
                                : ao_pullback_launch_g<T, true>(ctx, a, runs, dX, dzeta, dD);
}

template aok_status ao_eval_launch<float>(aok_ctx, const AoArgs &, const AoRunTable &);
template aok_status ao_eval_launch<double>(aok_ctx, const AoArgs &, const AoRunTable &);
template aok_status ao_pullback_launch<float>(aok_ctx, const AoArgs &, const AoRunTable &, void *, void *, void *);
template aok_status ao_pullback_launch<double>(aok_ctx, const AoArgs &, const AoRunTable &, void *, void *,
                                               void *);

}  // namespace aok
