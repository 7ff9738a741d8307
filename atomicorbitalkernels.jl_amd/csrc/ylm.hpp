// ylm.hpp -- real solid harmonics Z_lm(x) (and gradients) fully unrolled in registers.
//
// Convention of SpheriCart.jl's SolidHarmonics (the reference calls it at
// src/orbitals/atomicorbitals.jl:155,200,239,259), L2-normalised, stored at index
// l*l + l + m:
//   Z_l^m = F_l^|m| Q_l^|m|(z, r^2) * { c_m (m>0), 1/sqrt2 (m=0), s_|m| (m<0) },
//   c_m + i s_m = (x + i y)^m,  F_l^m = (-1)^m sqrt((2l+1)/(2pi) (l-m)!/(l+m)!),
//   Q_l^l = (-1)^l (2l-1)!!,  (l-m) Q_l^m = (2l-1) z Q_{l-1}^m - (l+m-1) r^2 Q_{l-2}^m.
// Here the recursion runs directly on Phat_l^m = F_l^m Q_l^m (1/sqrt2 folded into
// m = 0): Phat_l^m = A_lm z Phat_{l-1}^m - B_lm r^2 Phat_{l-2}^m, with A, B folded to
// compile-time constants.  Gradients carry d/dz and d/d(r^2) of Phat through the same
// recursion and use dc_m/dx = m c_{m-1}, dc_m/dy = -m s_{m-1}, ds_m/dx = m s_{m-1},
// ds_m/dy = m c_{m-1}.
#pragma once

#include <hip/hip_runtime.h>

namespace aok {

constexpr double kPi = 3.14159265358979323846;

constexpr double c_sqrt(double x)
{
    if (x <= 0.0) return 0.0;
    double g = x > 1.0 ? x : 1.0;
    for (int i = 0; i < 200; ++i) {
        const double n = 0.5 * (g + x / g);
        if (n == g) break;
        g = n;
    }
    return g;
}

// F_l^m
constexpr double ylm_F(int l, int m)
{
    double ratio = 1.0;
    for (int k = l - m + 1; k <= l + m; ++k) ratio /= static_cast<double>(k);
    const double F = c_sqrt((2.0 * l + 1.0) / (2.0 * kPi) * ratio);
    return (m & 1) ? -F : F;
}

// Phat_m^m (1/sqrt2 folded for m = 0)
constexpr double ylm_diag(int m)
{
    double q = 1.0;  // Q_m^m = (-1)^m (2m-1)!!
    for (int l = 1; l <= m; ++l) q *= -(2.0 * l - 1.0);
    const double v = ylm_F(m, m) * q;
    return m == 0 ? v / c_sqrt(2.0) : v;
}

constexpr double ylm_A(int l, int m)
{
    return (2.0 * l - 1.0) / static_cast<double>(l - m) * ylm_F(l, m) / ylm_F(l - 1, m);
}

constexpr double ylm_B(int l, int m)
{
    return (l - 2 < m) ? 0.0
                       : static_cast<double>(l + m - 1) / static_cast<double>(l - m) * ylm_F(l, m) /
                             ylm_F(l - 2, m);
}

template <typename T>
struct YlmOut {
    T *Z;
    T *Zx;
    T *Zy;
    T *Zz;
};

// Sinks of the recursion: store every Z (and gradient), or contract the gradients with
// weights W as they are produced: g = sum_y W_y grad Z_y (no gradient arrays kept).
template <typename T>
struct YlmStore {
    YlmOut<T> o;
    __device__ __forceinline__ void val(int idx, T v) { o.Z[idx] = v; }
    __device__ __forceinline__ void grad(int idx, T gx, T gy, T gz)
    {
        o.Zx[idx] = gx;
        o.Zy[idx] = gy;
        o.Zz[idx] = gz;
    }
};

template <typename T>
struct YlmGradDot {
    const T *W;
    T gx, gy, gz;
    __device__ __forceinline__ void val(int, T) {}
    __device__ __forceinline__ void grad(int idx, T ax, T ay, T az)
    {
        gx += W[idx] * ax;
        gy += W[idx] * ay;
        gz += W[idx] * az;
    }
};

template <int L, bool GRAD, typename T, int M, int LL, typename S>
__device__ __forceinline__ void ylm_column(const T x, const T y, const T z, const T r2, const T *c,
                                           const T *s, T Pm1, T Pzm1, T Prm1, T Pm2, T Pzm2, T Prm2,
                                           S &o)
{
    if constexpr (LL <= L) {
        T P, Pz = T(0), Pr = T(0);
        if constexpr (LL == M) {
            constexpr T d = T(ylm_diag(M));  // constant-evaluated: no run-time square roots
            P = d;
        } else {
            constexpr T a = T(ylm_A(LL, M));
            constexpr T b = T(ylm_B(LL, M));
            if constexpr (LL - 2 >= M) {
                P = a * z * Pm1 - b * r2 * Pm2;
                if constexpr (GRAD) {
                    Pz = a * (Pm1 + z * Pzm1) - b * r2 * Pzm2;
                    Pr = a * z * Prm1 - b * (Pm2 + r2 * Prm2);
                }
            } else {
                P = a * z * Pm1;
                if constexpr (GRAD) {
                    Pz = a * (Pm1 + z * Pzm1);
                    Pr = a * z * Prm1;
                }
            }
        }
        constexpr int base = LL * LL + LL;
        if constexpr (M == 0) {
            o.val(base, P);
            if constexpr (GRAD) o.grad(base, T(2) * x * Pr, T(2) * y * Pr, Pz + T(2) * z * Pr);
        } else {
            o.val(base + M, P * c[M]);
            o.val(base - M, P * s[M]);
            if constexpr (GRAD) {
                const T pzz = Pz + T(2) * z * Pr;
                const T tx = T(2) * x * Pr, ty = T(2) * y * Pr;
                const T mP = T(M) * P;
                o.grad(base + M, tx * c[M] + mP * c[M - 1], ty * c[M] - mP * s[M - 1], pzz * c[M]);
                o.grad(base - M, tx * s[M] + mP * s[M - 1], ty * s[M] + mP * c[M - 1], pzz * s[M]);
            }
        }
        ylm_column<L, GRAD, T, M, LL + 1>(x, y, z, r2, c, s, P, Pz, Pr, Pm1, Pzm1, Prm1, o);
    }
}

template <int L, bool GRAD, typename T, int M, typename S>
__device__ __forceinline__ void ylm_columns(const T x, const T y, const T z, const T r2, const T *c,
                                            const T *s, S &o)
{
    if constexpr (M <= L) {
        ylm_column<L, GRAD, T, M, M>(x, y, z, r2, c, s, T(0), T(0), T(0), T(0), T(0), T(0), o);
        ylm_columns<L, GRAD, T, M + 1>(x, y, z, r2, c, s, o);
    }
}

template <int L, bool GRAD, typename T, typename S>
__device__ __forceinline__ void ylm_run(const T x, const T y, const T z, S &o)
{
    T c[L + 1], s[L + 1];
    c[0] = T(1);
    s[0] = T(0);
#pragma unroll
    for (int m = 1; m <= L; ++m) {
        c[m] = x * c[m - 1] - y * s[m - 1];
        s[m] = x * s[m - 1] + y * c[m - 1];
    }
    const T r2 = x * x + y * y + z * z;
    ylm_columns<L, GRAD, T, 0>(x, y, z, r2, c, s, o);
}

// Z[(L+1)^2] (and, with GRAD, Zx/Zy/Zz) at (x, y, z).
template <int L, bool GRAD, typename T>
__device__ __forceinline__ void solid_harmonics(const T x, const T y, const T z, YlmOut<T> o)
{
    YlmStore<T> st{o};
    ylm_run<L, GRAD, T>(x, y, z, st);
}

// sum_y W[y] grad Z_y at (x, y, z), accumulated into (gx, gy, gz).
template <int L, typename T>
__device__ __forceinline__ void solid_harmonics_grad_dot(const T x, const T y, const T z, const T *W, T &gx, T &gy,
                                                         T &gz)
{
    YlmGradDot<T> d{W, gx, gy, gz};
    ylm_run<L, true, T>(x, y, z, d);
    gx = d.gx;
    gy = d.gy;
    gz = d.gz;
}

}  // namespace aok
