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

// (a, d) pair of the LDS parameter table: a = zeta*log2(e) for float (the exponential
// runs on v_exp_f32 in base 2), a = zeta for double.
template <typename T>
struct ParamPair {
    T a, d;
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
                                ParamPair<T> *__restrict__ tab)
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
    ParamPair<T> pp;
    pp.a = std::is_same<T, float>::value ? T(z * 1.4426950408889634) : T(z);
    pp.d = D[src];
    tab[q] = pp;
}

template <typename T>
__device__ __forceinline__ void load_param_table(ParamPair<T> *tab, const ParamPair<T> *__restrict__ gtab, int n)
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
__device__ __forceinline__ void radial(const ParamPair<T> *row, int K, T f, T r, int p, int decay, T &R,
                                       T &G)
{
    T s = T(0), ds = T(0);
    for (int m = 0; m < K; ++m) {
        const ParamPair<T> pp = row[m];
        const T e = exp_param<T>(pp.a, f) * pp.d;
        s += e;
        if (GRAD) ds += pp.a * e;
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
                                                      int itype, int64_t nX, const ParamPair<T> *__restrict__ gtab,
                                                       int nRad, int K, int NZ, T scale,
                                                      int decay, AoRunTable runs, T *__restrict__ P,
                                                      T *__restrict__ dP, unsigned *__restrict__ err)
{
    extern __shared__ unsigned char smem_raw[];
    ParamPair<T> *tab = reinterpret_cast<ParamPair<T> *>(smem_raw);
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
            const int k = runs.k[q], lk = runs.l[q], o0 = runs.o0[q];
            T R, G = T(0);
            radial<T, GRAD>(tab + (sg * nRad + k) * K, K, f, r, runs.poly[q], decay, R, G);
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
                                                          int itype, int64_t nX, const ParamPair<T> *__restrict__ gtab,
                                                           int nRad, int K, int NZ, T scale,
                                                          int decay, AoRunTable runs, T *__restrict__ P,
                                                          unsigned *__restrict__ err)
{
    extern __shared__ unsigned char smem_raw[];
    ParamPair<T> *tab = reinterpret_cast<ParamPair<T> *>(smem_raw);
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
            const int k = runs.k[q], lk = runs.l[q], o0 = runs.o0[q];
            T R[PPL];
#pragma unroll
            for (int p = 0; p < PPL; ++p) {
                T G;
                radial<T, false>(tab + (sg[p] * nRad + k) * K, K, f[p], r[p], runs.poly[q], decay, R[p], G);
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

// Pullback contraction of one full shell run: dr += sum_j w_j Z_j, g += w_j grad P_j.
template <int LL, int L, typename T>
__device__ __forceinline__ void contract_shell(const T *Z, const T *Zx, const T *Zy, const T *Zz, T R, T G, T x,
                                               T y, T z, int64_t i, int64_t nX, int o0,
                                               const T *__restrict__ dPin, T &dr, T &gx, T &gy, T &gz)
{
    if constexpr (LL <= L) {
#pragma unroll
        for (int j = 0; j <= 2 * LL; ++j) {
            const int yy = LL * LL + j;
            const T w = dPin[i + nX * (o0 + j)];
            dr += w * Z[yy];
            const T wg = w * G * Z[yy];
            const T wr = w * R;
            gx += wg * x + wr * Zx[yy];
            gy += wg * y + wr * Zy[yy];
            gz += wg * z + wr * Zz[yy];
        }
    }
}

// ---------------------------------------------------------------------------------
// Pullback.  Phase A (one lane per point): dX = s * sum_o dP_io grad P_io and
// dR_k = sum_{o in k} dP_io Z_io, staged per lane in LDS.  Phase B: for every radial k
// and primitive m, the per-species sums dD[k,m,sg] += dR_k r^p e and
// dzeta[k,m,sg] -= dR_k r^p D f e are accumulated in registers with a species mask,
// reduced over the wave and added into per-wave FP64 LDS accumulators; the block
// writes one FP64 partial per slot and a second kernel sums the partials in a fixed
// order (deterministic, no float atomics).
// ---------------------------------------------------------------------------------
constexpr int AO_PB_THREADS = 256;
constexpr int AO_PB_PPT = 2;      // points per thread per tile
constexpr int AO_PB_KC = 8;       // primitive chunk held in registers
constexpr int AO_PB_NZG = 4;      // species handled per pass

template <typename T, int L, bool GENERIC>
__global__ __launch_bounds__(AO_PB_THREADS) void ao_pullback_kernel(
    const T *__restrict__ X, const void *__restrict__ sidx, int itype, int64_t nX, const ParamPair<T> *__restrict__ gtab,
     int nRad, int K, int NZ, T scale, int decay, AoRunTable runs,
    const T *__restrict__ dPin, T *__restrict__ dX, double *__restrict__ partial, int want_params,
    unsigned *__restrict__ err)
{
    constexpr int TILE = AO_PB_THREADS * AO_PB_PPT;
    constexpr int NY = (L + 1) * (L + 1);
    extern __shared__ unsigned char smem_raw[];
    // LDS layout: param table | wave accumulators (double) | dR[nRad][TILE] | f[TILE] | sg[TILE]
    ParamPair<T> *tab = reinterpret_cast<ParamPair<T> *>(smem_raw);
    const int ntab = NZ * nRad * K;
    const int nslot = ntab * 2;
    size_t off = (sizeof(ParamPair<T>) * ntab + 15) & ~size_t(15);
    double *wacc = reinterpret_cast<double *>(smem_raw + off);
    off += sizeof(double) * nslot * (AO_PB_THREADS / 64);
    T *dRb = reinterpret_cast<T *>(smem_raw + off);
    off += sizeof(T) * nRad * TILE;
    T *fb = reinterpret_cast<T *>(smem_raw + off);
    off += sizeof(T) * TILE;
    int *sgb = reinterpret_cast<int *>(smem_raw + off);

    load_param_table(tab, gtab, NZ * nRad * K);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    double *myacc = wacc + wave * nslot;
    for (int q = lane; q < nslot; q += 64) myacc[q] = 0.0;
    __syncthreads();

    const int64_t ntiles = (nX + TILE - 1) / TILE;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        // ---- phase A ----
#pragma unroll
        for (int p = 0; p < AO_PB_PPT; ++p) {
            const int slot = p * AO_PB_THREADS + threadIdx.x;
            const int64_t i = t * TILE + slot;
            for (int k = 0; k < nRad; ++k) dRb[k * TILE + slot] = T(0);
            if (i >= nX) {
                sgb[slot] = -1;
                fb[slot] = T(0);
                continue;
            }
            const T x = scale * X[3 * i], y = scale * X[3 * i + 1], z = scale * X[3 * i + 2];
            int sg = load_sigma(sidx, itype, i);
            if (sg < 0 || sg >= NZ) {
                *err = ERR_SPECIES;
                sg = 0;
            }
            const T r2 = x * x + y * y + z * z;
            const T r = sqrt(r2);
            const T f = (decay == AOK_DECAY_GAUSSIAN) ? r2 : r;
            T Z[NY], Zx[NY], Zy[NY], Zz[NY];
            solid_harmonics<L, true, T>(x, y, z, YlmOut<T>{Z, Zx, Zy, Zz});
            T gx = T(0), gy = T(0), gz = T(0);
            for (int q = 0; q < runs.nrun; ++q) {
                const int k = runs.k[q], lk = runs.l[q], o0 = runs.o0[q];
                T R, G;
                radial<T, true>(tab + (sg * nRad + k) * K, K, f, r, runs.poly[q], decay, R, G);
                T dr = T(0);
                switch (lk) {
                    case 0: contract_shell<0, L>(Z, Zx, Zy, Zz, R, G, x, y, z, i, nX, o0, dPin, dr, gx, gy, gz); break;
                    case 1: contract_shell<1, L>(Z, Zx, Zy, Zz, R, G, x, y, z, i, nX, o0, dPin, dr, gx, gy, gz); break;
                    case 2: contract_shell<2, L>(Z, Zx, Zy, Zz, R, G, x, y, z, i, nX, o0, dPin, dr, gx, gy, gz); break;
                    case 3: contract_shell<3, L>(Z, Zx, Zy, Zz, R, G, x, y, z, i, nX, o0, dPin, dr, gx, gy, gz); break;
                    case 4: contract_shell<4, L>(Z, Zx, Zy, Zz, R, G, x, y, z, i, nX, o0, dPin, dr, gx, gy, gz); break;
                    default:
                        if constexpr (GENERIC) {
                        const int yy = -1 - lk;
                        const T w = dPin[i + nX * o0];
                        const T zv = pick<NY>(Z, yy);
                        dr += w * zv;
                        const T wg = w * G * zv;
                        gx += wg * x + w * R * pick<NY>(Zx, yy);
                        gy += wg * y + w * R * pick<NY>(Zy, yy);
                        gz += wg * z + w * R * pick<NY>(Zz, yy);
                    }
                }
                dRb[k * TILE + slot] += dr;
            }
            if (dX != nullptr) {
                dX[3 * i] = scale * gx;
                dX[3 * i + 1] = scale * gy;
                dX[3 * i + 2] = scale * gz;
            }
            sgb[slot] = sg;
            fb[slot] = f;
        }
        if (!want_params) continue;
        // ---- phase B ---- (each thread reads only its own LDS slots: no barrier needed)
        for (int g0 = 0; g0 < NZ; g0 += AO_PB_NZG) {
            const int nzg = min(AO_PB_NZG, NZ - g0);
            for (int k = 0; k < nRad; ++k) {
                const int pk = runs.kpoly[k];
                for (int m0 = 0; m0 < K; m0 += AO_PB_KC) {
                    T accD[AO_PB_NZG][AO_PB_KC], accZ[AO_PB_NZG][AO_PB_KC];
#pragma unroll
                    for (int q = 0; q < AO_PB_NZG; ++q)
#pragma unroll
                        for (int m = 0; m < AO_PB_KC; ++m) accD[q][m] = accZ[q][m] = T(0);
#pragma unroll
                    for (int p = 0; p < AO_PB_PPT; ++p) {
                        const int slot = p * AO_PB_THREADS + threadIdx.x;
                        const int sg = sgb[slot];
                        const T f = fb[slot];
                        T dr = dRb[k * TILE + slot];
                        if (pk != 0) {
                            const T r = (decay == AOK_DECAY_GAUSSIAN) ? sqrt(f) : f;
                            dr *= int_pow(r, pk);
                        }
                        const int sgs = sg < 0 ? 0 : sg;
                        const ParamPair<T> *row = tab + (sgs * nRad + k) * K;
                        T mask[AO_PB_NZG];
#pragma unroll
                        for (int q = 0; q < AO_PB_NZG; ++q) mask[q] = (sg == g0 + q) ? T(1) : T(0);
#pragma unroll
                        for (int m = 0; m < AO_PB_KC; ++m) {
                            if (m0 + m < K) {
                                const ParamPair<T> pp = row[m0 + m];
                                const T w = dr * exp_param<T>(pp.a, f);
                                const T wz = -w * pp.d * f;
#pragma unroll
                                for (int q = 0; q < AO_PB_NZG; ++q) {
                                    accD[q][m] += mask[q] * w;
                                    accZ[q][m] += mask[q] * wz;
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int q = 0; q < AO_PB_NZG; ++q) {
                        if (q >= nzg) break;
#pragma unroll
                        for (int m = 0; m < AO_PB_KC; ++m) {
                            if (m0 + m >= K) break;
                            const double sD = wave_sum(static_cast<double>(accD[q][m]));
                            const double sZ = wave_sum(static_cast<double>(accZ[q][m]));
                            const int base = 2 * (((g0 + q) * nRad + k) * K + m0 + m);
                            if (lane == 0) {
                                myacc[base] += sD;
                                myacc[base + 1] += sZ;
                            }
                        }
                    }
                }
            }
        }
    }
    if (!want_params) return;
    __syncthreads();
    const int nw = AO_PB_THREADS / 64;
    for (int q = threadIdx.x; q < nslot; q += AO_PB_THREADS) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += wacc[w * nslot + q];
        partial[static_cast<int64_t>(q) * gridDim.x + blockIdx.x] = s;
    }
}

// Sum the per-block partials of every (sg, k, m) slot in block order and scatter to
// the Julia [nRad x K x NZ] layout.
template <typename T>
__global__ void ao_pullback_finalize(const double *__restrict__ partial, int nblk, int nRad, int K, int NZ,
                                     T *__restrict__ dzeta, T *__restrict__ dD)
{
    const int ntab = NZ * nRad * K;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ntab) return;
    double sD = 0.0, sZ = 0.0;
    for (int b = 0; b < nblk; ++b) {
        sD += partial[static_cast<int64_t>(2 * q) * nblk + b];
        sZ += partial[static_cast<int64_t>(2 * q + 1) * nblk + b];
    }
    const int sg = q / (nRad * K);
    const int rem = q - sg * nRad * K;
    const int k = rem / K, m = rem - k * K;
    const int64_t dst = k + static_cast<int64_t>(nRad) * (m + static_cast<int64_t>(K) * sg);
    if (dD) dD[dst] = static_cast<T>(sD);
    if (dzeta) dzeta[dst] = static_cast<T>(sZ);
}

// ---------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------
// The per-call parameter table, at byte offset `at` of the context workspace.
template <typename T>
static aok_status param_table(aok_ctx ctx, const AoArgs &a, size_t at, const ParamPair<T> *&gtab)
{
    const int n = a.NZ * a.nRad * a.K;
    aok_status st = ensure_ws(ctx, at + sizeof(ParamPair<T>) * n);
    if (st != AOK_OK) return st;
    ParamPair<T> *t = reinterpret_cast<ParamPair<T> *>(static_cast<char *>(ctx->ws) + at);
    hipLaunchKernelGGL((ao_param_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<const T *>(a.zeta), static_cast<const T *>(a.D), a.nRad, a.K, a.NZ, t);
    AOK_HIP(ctx, hipGetLastError());
    gtab = t;
    return AOK_OK;
}

template <typename T, int L, bool GENERIC>
static aok_status launch_eval_L(aok_ctx ctx, const AoArgs &a, const AoRunTable &runs)
{
    const size_t shm = sizeof(ParamPair<T>) * a.NZ * a.nRad * a.K;
    if (a.nX <= 0) return AOK_OK;
    const ParamPair<T> *gtab = nullptr;
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
        if (runs.l[q] < 0) return false;
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


template <typename T>
static size_t ao_pb_shmem(const AoArgs &a)
{
    constexpr int TILE = AO_PB_THREADS * AO_PB_PPT;
    const int ntab = a.NZ * a.nRad * a.K;
    size_t off = (sizeof(ParamPair<T>) * ntab + 15) & ~size_t(15);
    off += sizeof(double) * ntab * 2 * (AO_PB_THREADS / 64);
    off += sizeof(T) * a.nRad * TILE + sizeof(T) * TILE + sizeof(int) * TILE;
    return off;
}

template <typename T, int L, bool GENERIC>
static aok_status launch_pb_L(aok_ctx ctx, const AoArgs &a, const AoRunTable &runs, void *dX, void *dzeta,
                              void *dD)
{
    constexpr int TILE = AO_PB_THREADS * AO_PB_PPT;
    const int64_t ntiles = (a.nX + TILE - 1) / TILE;
    const int grid = static_cast<int>(std::min<int64_t>(ntiles, 2048));
    const int ntab = a.NZ * a.nRad * a.K;
    const bool want = (dzeta != nullptr) || (dD != nullptr);
    if (grid <= 0) {
        if (want) {
            // no points: zero cotangents (the reference returns zero-filled arrays)
            hipLaunchKernelGGL((ao_pullback_finalize<T>), dim3((ntab + 255) / 256), dim3(256), 0, ctx->stream,
                               static_cast<const double *>(nullptr), 0, a.nRad, a.K, a.NZ, static_cast<T *>(dzeta),
                               static_cast<T *>(dD));
            AOK_HIP(ctx, hipGetLastError());
        }
        return AOK_OK;
    }
    const size_t shm = ao_pb_shmem<T>(a);
    if (shm > 160 * 1024)
        return fail(ctx, AOK_ERR_UNSUPPORTED, "ao_pullback: basis too large for the LDS tile");
    const size_t part_bytes = align256(sizeof(double) * static_cast<size_t>(ntab) * 2 * grid);
    const ParamPair<T> *gtab = nullptr;
    aok_status st = param_table<T>(ctx, a, part_bytes, gtab);
    if (st != AOK_OK) return st;
    double *partial = static_cast<double *>(ctx->ws);
    if (shm > 64 * 1024)
        AOK_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(&ao_pullback_kernel<T, L, GENERIC>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm)));
    time_begin(ctx, "ao_pullback");
    hipLaunchKernelGGL((ao_pullback_kernel<T, L, GENERIC>), dim3(grid), dim3(AO_PB_THREADS), shm, ctx->stream,
                       static_cast<const T *>(a.X), a.sidx, a.itype, a.nX, gtab,
                        a.nRad, a.K, a.NZ, static_cast<T>(a.scale), a.decay, runs,
                       static_cast<const T *>(a.dPin), static_cast<T *>(dX), partial, want ? 1 : 0, ctx->d_err);
    time_end(ctx);
    AOK_HIP(ctx, hipGetLastError());
    if (want) {
        hipLaunchKernelGGL((ao_pullback_finalize<T>), dim3((ntab + 255) / 256), dim3(256), 0, ctx->stream,
                           static_cast<const double *>(partial), grid, a.nRad, a.K, a.NZ, static_cast<T *>(dzeta),
                           static_cast<T *>(dD));
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
