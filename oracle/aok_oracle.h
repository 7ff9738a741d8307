/*
 * aok_oracle.h -- FP64 CPU restatement of AtomicOrbitalKernels.jl's two hot paths.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle: it is linked
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always
 * as the checker (or the timed CPU baseline), never as a product code path.  The
 * product (libaok.so) never loads it.
 *
 * Every array uses the reference's Julia column-major layouts, with 0-based C
 * indexing; species indices are 1-based int64 as Julia passes them.
 *   X, dX      : double[3*nX]            xyz interleaved (Vector{SVector{3}})
 *   zeta, D    : double[nRad*K*NZ]       [k + nRad*(m + K*sigma)]
 *   P          : double[nX*nOrb]         [i + nX*o]
 *   dP         : double[3*nX*nOrb]       [3*(i + nX*o) + c]
 *   pos A/B/C  : double[3*B]
 *   out 2c     : double[N*N*B]           [r + N*(c + N*b)]
 *   out 3c     : double[N*N*N*B]
 *
 * Pins (tests/test_oracle.py): Gauss-Hermite quadrature golden vectors for every 2c
 * shell pair up to (4,4) and 3c triples, composed with the reference's alias map;
 * closed forms (s-s same centre, l <= 2 harmonics, self-overlap 1); 3c -> 2c reduction;
 * finite differences of every gradient and pullback; addition theorem, homogeneity and
 * orthonormality of the harmonics.
 * Parity unpinned: orc_ylm's assignment of m to cos/sin and its sign convention within
 * each l (SpheriCart.jl 0.3.1, not vendored).  The tests above fix everything except
 * a signed permutation of the 2l+1 functions of each l; no reference test or fixture
 * distinguishes those.
 */
#ifndef AOK_ORACLE_H
#define AOK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Real solid harmonics Z_lm(x) for l <= L, stored at index l*l + l + m
 * (SpheriCart SolidHarmonics convention, L2-normalised).  dZ (may be NULL)
 * receives the gradient, dZ[3*idx + c]. */
void orc_ylm(int L, double x, double y, double z, double *Z, double *dZ);

/* evaluate / evaluate_ed of AtomicOrbitals (atomicorbitals.jl:148-162,192-210).
 * dP may be NULL (values only).  decay: 0 = Gaussian (r^2), 1 = Slater (r). */
int orc_ao_eval(const double *X, const int64_t *sidx, int64_t nX,
                const double *zeta, const double *D, int64_t nRad, int64_t K, int64_t NZ,
                const int64_t *poly, const int64_t *iR, const int64_t *iY, int64_t nOrb,
                int L, double lengthscale, int decay, double *P, double *dP);

/* rrule pullback (atomicorbitals.jl:254-267,283-299, gtostoradials.jl:232-267):
 * dX = sum_o dPin * grad P (NULL to skip); dzeta, dD (NULL to skip). */
int orc_ao_pullback(const double *X, const int64_t *sidx, int64_t nX,
                    const double *zeta, const double *D, int64_t nRad, int64_t K, int64_t NZ,
                    const int64_t *poly, const int64_t *iR, const int64_t *iY, int64_t nOrb,
                    int L, double lengthscale, int decay, const double *dPin,
                    double *dX, double *dzeta, double *dD);

/* compile_basis normalisation (cartesian_gto_basis.jl:69-120) and its VJP
 * (overlap_2c_grad.jl:251-286).  Arrays [ns*K*NZ]. */
double orc_cart_normsq(const double *D, const double *zeta, int64_t K, int64_t stride, int l);
void orc_compile_coef(const double *zeta, const double *D, int64_t ns, int64_t K, int64_t NZ,
                      const int64_t *ls, double *coef);
void orc_compile_coef_pb(const double *dcoef, const double *zeta, const double *D,
                         int64_t ns, int64_t K, int64_t NZ, const int64_t *ls,
                         double *dD, double *dzeta_norm);

/* batch_S_orbital_kernel! (overlap_2c.jl:15-131): out[N,N,B], positions already
 * in Bohr (the caller applies lengthscale).  Returns N. */
int64_t orc_ovl2c(const double *posA, const double *posB, const int64_t *sidxA,
                  const int64_t *sidxB, int64_t B, int64_t ns, int64_t K, int64_t NZ,
                  const int64_t *ls, const double *coef, const double *alpha, double *out);

/* batch_S_orbital_pb_kernel! (overlap_2c_grad.jl:71-214): dcoef, dalpha [ns*K*NZ]. */
void orc_ovl2c_pb(const double *dS, const double *posA, const double *posB,
                  const int64_t *sidxA, const int64_t *sidxB, int64_t B, int64_t ns,
                  int64_t K, int64_t NZ, const int64_t *ls, const double *coef,
                  const double *alpha, double *dcoef, double *dalpha);
/* The same split over OpenMP threads (per-thread partials summed in thread order). */
void orc_ovl2c_pb_omp(const double *dS, const double *posA, const double *posB,
                      const int64_t *sidxA, const int64_t *sidxB, int64_t B, int64_t ns,
                      int64_t K, int64_t NZ, const int64_t *ls, const double *coef,
                      const double *alpha, double *dcoef, double *dalpha);

/* Position cotangents of the 2-centre overlap (no reference counterpart): dA, dB
 * [3*B] in Bohr, dB = -dA; dB may be NULL. */
void orc_ovl2c_dx(const double *dS, const double *posA, const double *posB,
                  const int64_t *sidxA, const int64_t *sidxB, int64_t B, int64_t ns,
                  int64_t K, int64_t NZ, const int64_t *ls, const double *coef,
                  const double *alpha, double *dA, double *dB);

/* Species-aware 3-centre overlap following batch_V_kernel! (kernels_3c.jl:12-162)
 * with the per-species primitive slices of the orbital path; sidx* may be NULL
 * (species 1).  Returns N. */
int64_t orc_ovl3c(const double *posA, const double *posB, const double *posC,
                  const int64_t *sidxA, const int64_t *sidxB, const int64_t *sidxC,
                  int64_t B, int64_t ns, int64_t K, int64_t NZ, const int64_t *ls,
                  const double *coef, const double *alpha, double *out);

/* Parameter cotangents of the 3-centre overlap (no reference counterpart): dcoef,
 * dalpha [ns*K*NZ] from dT [N*N*N*B]; every primitive slot, padded ones included. */
void orc_ovl3c_pb(const double *dT, const double *posA, const double *posB, const double *posC,
                  const int64_t *sidxA, const int64_t *sidxB, const int64_t *sidxC, int64_t B,
                  int64_t ns, int64_t K, int64_t NZ, const int64_t *ls, const double *coef,
                  const double *alpha, double *dcoef, double *dalpha);

/* OpenMP variants used only for the CPU baseline timing (same arithmetic). */
int64_t orc_ovl2c_omp(const double *posA, const double *posB, const int64_t *sidxA,
                      const int64_t *sidxB, int64_t B, int64_t ns, int64_t K, int64_t NZ,
                      const int64_t *ls, const double *coef, const double *alpha, double *out);
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
