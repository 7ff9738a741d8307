/*
 * aok_oracle.c -- FP64 CPU restatement of AtomicOrbitalKernels.jl's hot paths.
 *
 * TEST INFRASTRUCTURE ONLY (the correctness oracle and the timed CPU baseline);
 * the product library never links or loads it.  Each function names the reference
 * file:line whose semantics it restates.  Scalar loops in plain C; the only
 * parallel code is orc_ovl2c_omp / orc_ovl2c_pb_omp, the OpenMP split of the same work
 * items that stands in for KernelAbstractions' CPU backend in the CPU-baseline timing.
 */
#include "aok_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_LMAX 8                 /* largest shell l the oracle accepts         */
#define ORC_LE (ORC_LMAX + 2)      /* E tables run to l+2 for the pullback       */
#define ORC_T (2 * ORC_LE + 3)     /* Hermite index range (with zero margin)     */
#define ORC_T3 (3 * ORC_LMAX + 4)

static const double ORC_PI = 3.14159265358979323846;

/* ------------------------------------------------------------------------- */
/* Real solid harmonics.  SpheriCart.jl (SolidHarmonics, compat 0.3.1) is not  */
/* vendored; this follows its published recursion: Z_l^m = F_l^|m| Q_l^|m|     */
/* times {c_m (m>0), 1/sqrt2 (m=0), s_|m| (m<0)}, c_m + i s_m = (x+iy)^m,      */
/* F_l^m = (-1)^m sqrt((2l+1)/(2pi) (l-m)!/(l+m)!), Q_l^l = (-1)^l (2l-1)!!,   */
/* (l-m) Q_l^m = (2l-1) z Q_{l-1}^m - (l+m-1) r^2 Q_{l-2}^m, index l^2+l+m.    */
/* Called at atomicorbitals.jl:155,200,239,259.                               */
/* ------------------------------------------------------------------------- */
void orc_ylm(int L, double x, double y, double z, double *Z, double *dZ)
{
    double c[ORC_LMAX + 2], s[ORC_LMAX + 2];
    double Q[ORC_LMAX + 1][ORC_LMAX + 1], Qz[ORC_LMAX + 1][ORC_LMAX + 1],
        Qr[ORC_LMAX + 1][ORC_LMAX + 1];
    const double r2 = x * x + y * y + z * z;
    memset(Q, 0, sizeof Q);
    memset(Qz, 0, sizeof Qz);
    memset(Qr, 0, sizeof Qr);
    c[0] = 1.0;
    s[0] = 0.0;
    for (int m = 1; m <= L; ++m) {
        c[m] = x * c[m - 1] - y * s[m - 1];
        s[m] = x * s[m - 1] + y * c[m - 1];
    }
    Q[0][0] = 1.0;
    for (int l = 1; l <= L; ++l) {
        Q[l][l] = -(2.0 * l - 1.0) * Q[l - 1][l - 1];
        for (int m = 0; m < l; ++m) {
            const double a = 2.0 * l - 1.0, bq = (double)(l + m - 1), inv = 1.0 / (double)(l - m);
            const double q2 = (m <= l - 2) ? Q[l - 2][m] : 0.0;
            const double q2z = (m <= l - 2) ? Qz[l - 2][m] : 0.0;
            const double q2r = (m <= l - 2) ? Qr[l - 2][m] : 0.0;
            Q[l][m] = (a * z * Q[l - 1][m] - bq * r2 * q2) * inv;
            Qz[l][m] = (a * (Q[l - 1][m] + z * Qz[l - 1][m]) - bq * r2 * q2z) * inv;
            Qr[l][m] = (a * z * Qr[l - 1][m] - bq * (q2 + r2 * q2r)) * inv;
        }
    }
    for (int l = 0; l <= L; ++l) {
        for (int m = 0; m <= l; ++m) {
            /* F_l^m with the factorial ratio (l-m)!/(l+m)! */
            double ratio = 1.0;
            for (int k = l - m + 1; k <= l + m; ++k) ratio /= (double)k;
            double F = sqrt((2.0 * l + 1.0) / (2.0 * ORC_PI) * ratio);
            if (m & 1) F = -F;
            const double fq = F * Q[l][m];
            const double fqz = F * (Qz[l][m] + 2.0 * z * Qr[l][m]); /* d/dz (through z and r^2) */
            const double fqr = F * Qr[l][m];
            if (m == 0) {
                const double w = 1.0 / sqrt(2.0);
                Z[l * l + l] = w * fq;
                if (dZ) {
                    double *g = dZ + 3 * (l * l + l);
                    g[0] = w * fqr * 2.0 * x;
                    g[1] = w * fqr * 2.0 * y;
                    g[2] = w * fqz;
                }
            } else {
                Z[l * l + l + m] = fq * c[m];
                Z[l * l + l - m] = fq * s[m];
                if (dZ) {
                    double *gp = dZ + 3 * (l * l + l + m);
                    double *gm = dZ + 3 * (l * l + l - m);
                    gp[0] = fqr * 2.0 * x * c[m] + fq * m * c[m - 1];
                    gp[1] = fqr * 2.0 * y * c[m] - fq * m * s[m - 1];
                    gp[2] = fqz * c[m];
                    gm[0] = fqr * 2.0 * x * s[m] + fq * m * s[m - 1];
                    gm[1] = fqr * 2.0 * y * s[m] + fq * m * c[m - 1];
                    gm[2] = fqz * s[m];
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Atomic orbitals                                                             */
/* ------------------------------------------------------------------------- */
static inline double ipow(double r, int64_t p)
{
    double v = 1.0;
    for (int64_t q = 0; q < p; ++q) v *= r;
    return v;
}

static int check_species(const int64_t *sidx, int64_t n, int64_t NZ)
{
    for (int64_t i = 0; i < n; ++i)
        if (sidx && (sidx[i] < 1 || sidx[i] > NZ)) return -1;
    return 0;
}

/* evaluate(::AtomicOrbitals, Rs, sidx, ps, st)  atomicorbitals.jl:148-162 with
 * _rnl_val_ka! gtostoradials.jl:117-130 and _aorb_val_ka! atomicorbitals.jl:110-113;
 * evaluate_ed  atomicorbitals.jl:192-210 with _rnl_ed_ka! gtostoradials.jl:132-153
 * and _aorb_ed_ka! atomicorbitals.jl:115-125 (including the X/r factor, so r = 0
 * yields NaN exactly as the reference does), then dRnlm .*= s (:208). */
int orc_ao_eval(const double *X, const int64_t *sidx, int64_t nX, const double *zeta,
                const double *D, int64_t nRad, int64_t K, int64_t NZ, const int64_t *poly,
                const int64_t *iR, const int64_t *iY, int64_t nOrb, int L, double lengthscale,
                int decay, double *P, double *dP)
{
    if (L < 0 || L > ORC_LMAX) return -2;
    if (check_species(sidx, nX, NZ)) return -1;
    const int nY = (L + 1) * (L + 1);
    double *Rn = (double *)malloc(sizeof(double) * (size_t)nRad * 2);
    double Z[(ORC_LMAX + 1) * (ORC_LMAX + 1)], dZ[3 * (ORC_LMAX + 1) * (ORC_LMAX + 1)];
    (void)nY;
    for (int64_t i = 0; i < nX; ++i) {
        const double x = lengthscale * X[3 * i], y = lengthscale * X[3 * i + 1],
                     z = lengthscale * X[3 * i + 2];
        const double r = sqrt(x * x + y * y + z * z);
        const int64_t sg = sidx ? sidx[i] - 1 : 0;
        const double fx = decay ? r : r * r, dfx = decay ? 1.0 : 2.0 * r;
        for (int64_t k = 0; k < nRad; ++k) {
            double sacc = 0.0, dsacc = 0.0;
            for (int64_t m = 0; m < K; ++m) {
                const int64_t q = k + nRad * (m + K * sg);
                const double a = D[q] * exp(-zeta[q] * fx);
                sacc += a;
                dsacc += -zeta[q] * dfx * a;
            }
            const int64_t p = poly ? poly[k] : 0;
            const double rp = ipow(r, p);
            const double drp = (p == 0) ? 0.0 : (double)p * ipow(r, p - 1);
            Rn[2 * k] = rp * sacc;
            Rn[2 * k + 1] = drp * sacc + rp * dsacc;
        }
        orc_ylm(L, x, y, z, Z, dP ? dZ : NULL);
        for (int64_t o = 0; o < nOrb; ++o) {
            const int64_t k = iR[o] - 1, yy = iY[o] - 1;
            const double rn = Rn[2 * k], yl = Z[yy];
            P[i + nX * o] = rn * yl;
            if (dP) {
                double *g = dP + 3 * (i + nX * o);
                const double xr[3] = {x / r, y / r, z / r};
                for (int c = 0; c < 3; ++c)
                    g[c] = lengthscale * (Rn[2 * k + 1] * xr[c] * yl + rn * dZ[3 * yy + c]);
            }
        }
    }
    free(Rn);
    return 0;
}

/* rrule(evaluate, ...) pullback  atomicorbitals.jl:307-316:
 *   dX: _aorb_pullback_x (:283-299) over the evaluate_ed Jacobian;
 *   (dzeta, dD): pullback_ps (:254-267) -> _aorb_pbrad_ka! scatter (:248-252)
 *   -> radial pullback _gtostoradials_pb_ka! (gtostoradials.jl:232-253). */
int orc_ao_pullback(const double *X, const int64_t *sidx, int64_t nX, const double *zeta,
                    const double *D, int64_t nRad, int64_t K, int64_t NZ, const int64_t *poly,
                    const int64_t *iR, const int64_t *iY, int64_t nOrb, int L,
                    double lengthscale, int decay, const double *dPin, double *dX,
                    double *dzeta, double *dD)
{
    if (L < 0 || L > ORC_LMAX) return -2;
    if (check_species(sidx, nX, NZ)) return -1;
    const int64_t npar = nRad * K * NZ;
    if (dzeta) memset(dzeta, 0, sizeof(double) * (size_t)npar);
    if (dD) memset(dD, 0, sizeof(double) * (size_t)npar);
    double *dR = (double *)malloc(sizeof(double) * (size_t)nRad);
    double *Rn = (double *)malloc(sizeof(double) * (size_t)nRad * 2);
    double Z[(ORC_LMAX + 1) * (ORC_LMAX + 1)], dZ[3 * (ORC_LMAX + 1) * (ORC_LMAX + 1)];
    for (int64_t i = 0; i < nX; ++i) {
        const double x = lengthscale * X[3 * i], y = lengthscale * X[3 * i + 1],
                     z = lengthscale * X[3 * i + 2];
        const double r = sqrt(x * x + y * y + z * z);
        const int64_t sg = sidx ? sidx[i] - 1 : 0;
        const double fx = decay ? r : r * r, dfx = decay ? 1.0 : 2.0 * r;
        for (int64_t k = 0; k < nRad; ++k) {
            double sacc = 0.0, dsacc = 0.0;
            for (int64_t m = 0; m < K; ++m) {
                const int64_t q = k + nRad * (m + K * sg);
                const double a = D[q] * exp(-zeta[q] * fx);
                sacc += a;
                dsacc += -zeta[q] * dfx * a;
            }
            const int64_t p = poly ? poly[k] : 0;
            const double rp = ipow(r, p);
            const double drp = (p == 0) ? 0.0 : (double)p * ipow(r, p - 1);
            Rn[2 * k] = rp * sacc;
            Rn[2 * k + 1] = drp * sacc + rp * dsacc;
            dR[k] = 0.0;
        }
        orc_ylm(L, x, y, z, Z, dZ);
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (int64_t o = 0; o < nOrb; ++o) {
            const int64_t k = iR[o] - 1, yy = iY[o] - 1;
            const double w = dPin[i + nX * o];
            dR[k] += w * Z[yy];
            const double rn = Rn[2 * k], drn = Rn[2 * k + 1], yl = Z[yy];
            gx += w * lengthscale * (drn * (x / r) * yl + rn * dZ[3 * yy]);
            gy += w * lengthscale * (drn * (y / r) * yl + rn * dZ[3 * yy + 1]);
            gz += w * lengthscale * (drn * (z / r) * yl + rn * dZ[3 * yy + 2]);
        }
        if (dX) {
            dX[3 * i] = gx;
            dX[3 * i + 1] = gy;
            dX[3 * i + 2] = gz;
        }
        for (int64_t k = 0; k < nRad; ++k) {
            const int64_t p = poly ? poly[k] : 0;
            const double gg = dR[k] * ipow(r, p);
            for (int64_t m = 0; m < K; ++m) {
                const int64_t q = k + nRad * (m + K * sg);
                const double a = exp(-zeta[q] * fx);
                if (dD) dD[q] += gg * a;
                if (dzeta) dzeta[q] += gg * D[q] * a * (-fx);
            }
        }
    }
    free(dR);
    free(Rn);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Cartesian normalisation  cartesian_gto_basis.jl:69-120                      */
/* ------------------------------------------------------------------------- */
static double cart_dfac(int n) /* n!! for odd n, 1 for n <= 1  (:69) */
{
    double p = 1.0;
    for (int k = n; k > 1; k -= 2) p *= (double)k;
    return p;
}

static double cart_cl(int l) /* pi^{3/2} (2l-1)!! / 2^l */
{
    return pow(ORC_PI, 1.5) * (l == 0 ? 1.0 : cart_dfac(2 * l - 1)) / ldexp(1.0, l);
}

double orc_cart_normsq(const double *D, const double *zeta, int64_t K, int64_t stride, int l)
{
    const double e = (double)l + 1.5;
    double s = 0.0;
    for (int64_t i = 0; i < K; ++i)
        for (int64_t j = 0; j < K; ++j)
            s += D[i * stride] * D[j * stride] / pow(zeta[i * stride] + zeta[j * stride], e);
    return cart_cl(l) * s;
}

void orc_compile_coef(const double *zeta, const double *D, int64_t ns, int64_t K, int64_t NZ,
                      const int64_t *ls, double *coef)
{
    for (int64_t sg = 0; sg < NZ; ++sg)
        for (int64_t k = 0; k < ns; ++k) {
            const int64_t base = k + ns * K * sg;
            const double nsq = orc_cart_normsq(D + base, zeta + base, K, ns, (int)ls[k]);
            for (int64_t m = 0; m < K; ++m)
                coef[base + ns * m] = nsq > 0.0 ? D[base + ns * m] / sqrt(nsq) : 0.0;
        }
}

/* _compile_coef_pb  overlap_2c_grad.jl:251-286 */
void orc_compile_coef_pb(const double *dcoef, const double *zeta, const double *D, int64_t ns,
                         int64_t K, int64_t NZ, const int64_t *ls, double *dD,
                         double *dzeta_norm)
{
    for (int64_t sg = 0; sg < NZ; ++sg)
        for (int64_t k = 0; k < ns; ++k) {
            const int64_t base = k + ns * K * sg;
            const int l = (int)ls[k];
            for (int64_t m = 0; m < K; ++m) {
                dD[base + ns * m] = 0.0;
                dzeta_norm[base + ns * m] = 0.0;
            }
            const double nsq = orc_cart_normsq(D + base, zeta + base, K, ns, l);
            if (!(nsq > 0.0)) continue;
            const double n = sqrt(nsq), c = cart_cl(l), e = (double)l + 1.5;
            double dot = 0.0;
            for (int64_t i = 0; i < K; ++i) dot += dcoef[base + ns * i] * D[base + ns * i];
            const double s2n3 = dot / (2.0 * n * n * n);
            for (int64_t m = 0; m < K; ++m) {
                double gm = 0.0, hm = 0.0;
                for (int64_t j = 0; j < K; ++j) {
                    const double den = zeta[base + ns * m] + zeta[base + ns * j];
                    gm += D[base + ns * j] / pow(den, e);
                    hm += D[base + ns * j] / pow(den, e + 1.0);
                }
                gm *= 2.0 * c;
                hm *= -2.0 * c * e * D[base + ns * m];
                dD[base + ns * m] = dcoef[base + ns * m] / n - gm * s2n3;
                dzeta_norm[base + ns * m] = -hm * s2n3;
            }
        }
}

/* ------------------------------------------------------------------------- */
/* McMurchie-Davidson E coefficients (reference/ecoeffs.jl:42-68; the same     */
/* recursion as overlap_2c.jl:87-106): E[t][i][j] for i<=la, j<=lb, zero      */
/* beyond t = i+j.                                                            */
/* ------------------------------------------------------------------------- */
/* Flat table E[t + T*(i + (la+1)*j)], T = la+lb+2 (one zero margin in t). */
#define EIDX(t, i, j) ((t) + T * ((i) + (la + 1) * (j)))

static void md_etab(double *E, int la, int lb, double PA, double PB, double oo2p, double E00)
{
    const int T = la + lb + 2;
    memset(E, 0, sizeof(double) * (size_t)(T * (la + 1) * (lb + 1)));
    E[0] = E00;
    for (int j = 0; j <= lb; ++j)
        for (int i = 0; i <= la; ++i) {
            if (i == 0 && j == 0) continue;
            const int pi = i > 0 ? i - 1 : i, pj = i > 0 ? j : j - 1;
            const double X = i > 0 ? PA : PB;
            for (int t = 0; t <= i + j; ++t) {
                const double em = t > 0 ? E[EIDX(t - 1, pi, pj)] : 0.0;
                E[EIDX(t, i, j)] = X * E[EIDX(t, pi, pj)] + (double)(t + 1) * E[EIDX(t + 1, pi, pj)] + oo2p * em;
            }
        }
}

/* E^{ij}_0 read-back from a table built for (LA, LB). */
static inline double e0(const double *E, int LA, int LB, int i, int j)
{
    const int T = LA + LB + 2;
    return E[T * (i + (LA + 1) * j)];
}

static int64_t shell_layout(int64_t ns, const int64_t *ls, int64_t *nbf, int64_t *off)
{
    off[0] = 0;
    for (int64_t s = 0; s < ns; ++s) {
        nbf[s] = (ls[s] + 1) * (ls[s] + 2) / 2;
        off[s + 1] = off[s] + nbf[s];
    }
    return off[ns];
}

/* One (b, s_a, s_b) work item of batch_S_orbital_kernel!  overlap_2c.jl:165-281. */
static void ovl2c_item(const double *posA, const double *posB, int64_t sgA, int64_t sgB,
                       int64_t sa, int64_t sb, int64_t ns, int64_t K, const int64_t *ls,
                       const int64_t *nbf, const int64_t *off, int64_t N, const double *coef,
                       const double *alpha, double *outb)
{
    const int la = (int)ls[sa], lb = (int)ls[sb];
    const int64_t nba = nbf[sa], nbb = nbf[sb];
    const int N1 = 2 * la + 1;
    double blk[(ORC_LMAX + 1) * (ORC_LMAX + 2) / 2 * (ORC_LMAX + 1) * (ORC_LMAX + 2) / 2];
    double Ex[ORC_T * (ORC_LE + 1) * (ORC_LE + 1)], Ey[ORC_T * (ORC_LE + 1) * (ORC_LE + 1)],
        Ez[ORC_T * (ORC_LE + 1) * (ORC_LE + 1)];
    for (int64_t q = 0; q < nba * nbb; ++q) blk[q] = 0.0;
    const double A[3] = {posA[0], posA[1], posA[2]}, Bv[3] = {posB[0], posB[1], posB[2]};
    for (int64_t ip = 0; ip < K; ++ip)
        for (int64_t jp = 0; jp < K; ++jp) {
            const double ca = coef[sa + ns * (ip + K * sgA)];
            const double cb = coef[sb + ns * (jp + K * sgB)];
            if (ca == 0.0 || cb == 0.0) continue; /* overlap_2c.jl:202 */
            const double aa = alpha[sa + ns * (ip + K * sgA)];
            const double ab = alpha[sb + ns * (jp + K * sgB)];
            const double p = aa + ab, mu = aa * ab / p, oo2p = 0.5 / p;
            const double prefac = ca * cb * pow(ORC_PI / p, 1.5);
            double *Es[3] = {Ex, Ey, Ez};
            for (int ax = 0; ax < 3; ++ax) {
                const double P = (aa * A[ax] + ab * Bv[ax]) / p;
                const double AB = A[ax] - Bv[ax];
                md_etab(Es[ax], la, lb, P - A[ax], P - Bv[ax], oo2p, exp(-mu * AB * AB));
            }
            int index1 = 0;
            for (int ll1 = la; ll1 >= 0; --ll1)
                for (int n1 = 0; n1 <= la - ll1; ++n1) {
                    const int m1 = la - ll1 - n1;
                    int index2 = 0;
                    for (int ll2 = lb; ll2 >= 0; --ll2)
                        for (int n2 = 0; n2 <= lb - ll2; ++n2) {
                            const int m2 = lb - ll2 - n2;
                            const int lin = index1 + N1 * index2;
                            blk[lin] += e0(Ex, la, lb, ll1, ll2) * e0(Ey, la, lb, m1, m2) *
                                        e0(Ez, la, lb, n1, n2) * prefac;
                            ++index2;
                        }
                    ++index1;
                }
        }
    for (int64_t j = 0; j < nbb; ++j)
        for (int64_t i = 0; i < nba; ++i) outb[(off[sa] + i) + N * (off[sb] + j)] = blk[i + nba * j];
}

int64_t orc_ovl2c(const double *posA, const double *posB, const int64_t *sidxA,
                  const int64_t *sidxB, int64_t B, int64_t ns, int64_t K, int64_t NZ,
                  const int64_t *ls, const double *coef, const double *alpha, double *out)
{
    int64_t *nbf = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * ns + 1));
    int64_t *off = nbf + ns;
    const int64_t N = shell_layout(ns, ls, nbf, off);
    for (int64_t s = 0; s < ns; ++s)
        if (ls[s] < 0 || ls[s] > ORC_LMAX) { free(nbf); return -2; }
    if (check_species(sidxA, B, NZ) || check_species(sidxB, B, NZ)) { free(nbf); return -1; }
    for (int64_t b = 0; b < B; ++b)
        for (int64_t sb = 0; sb < ns; ++sb)
            for (int64_t sa = 0; sa < ns; ++sa)
                ovl2c_item(posA + 3 * b, posB + 3 * b, sidxA ? sidxA[b] - 1 : 0,
                           sidxB ? sidxB[b] - 1 : 0, sa, sb, ns, K, ls, nbf, off, N, coef, alpha,
                           out + N * N * b);
    free(nbf);
    return N;
}

int64_t orc_ovl2c_omp(const double *posA, const double *posB, const int64_t *sidxA,
                      const int64_t *sidxB, int64_t B, int64_t ns, int64_t K, int64_t NZ,
                      const int64_t *ls, const double *coef, const double *alpha, double *out)
{
    int64_t *nbf = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * ns + 1));
    int64_t *off = nbf + ns;
    const int64_t N = shell_layout(ns, ls, nbf, off);
    (void)NZ;
    /* the linearised (b, s_a, s_b) ndrange split statically across threads */
    const int64_t items = B * ns * ns;
#pragma omp parallel for schedule(static)
    for (int64_t it = 0; it < items; ++it) {
        const int64_t b = it % B, sa = (it / B) % ns, sb = it / (B * ns);
        ovl2c_item(posA + 3 * b, posB + 3 * b, sidxA ? sidxA[b] - 1 : 0, sidxB ? sidxB[b] - 1 : 0,
                   sa, sb, ns, K, ls, nbf, off, N, coef, alpha, out + N * N * b);
    }
    free(nbf);
    return N;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* batch_S_orbital_pb_kernel!  overlap_2c_grad.jl:71-214, one work item (b, s_a, s_b):
 * every K x K primitive pair (padded ones included), E tables to l+2, contraction
 * against the aliased read-back position of dS, four accumulations per pair (into the
 * given arrays; the reference's atomics). */
static void ovl2c_pb_item(const double *dS, const double *posA, const double *posB,
                          const int64_t *sidxA, const int64_t *sidxB, int64_t b, int64_t sa, int64_t sb,
                          int64_t ns, int64_t K, const int64_t *ls, const int64_t *nbf, const int64_t *off,
                          int64_t N, const double *coef, const double *alpha, double *Ex,
                          double *dcoef, double *dalpha)
{
    double *Ey = Ex + ORC_T * (ORC_LE + 1) * (ORC_LE + 1), *Ez = Ey + ORC_T * (ORC_LE + 1) * (ORC_LE + 1);
    const int64_t sgA = sidxA ? sidxA[b] - 1 : 0, sgB = sidxB ? sidxB[b] - 1 : 0;
    const double *A = posA + 3 * b, *Bv = posB + 3 * b;
    const double *g_b = dS + N * N * b;
    const int la = (int)ls[sa], lb = (int)ls[sb];
    const int64_t nba = nbf[sa];
    const int N1 = 2 * la + 1;
    for (int64_t ip = 0; ip < K; ++ip)
        for (int64_t jp = 0; jp < K; ++jp) {
            const int64_t qa = sa + ns * (ip + K * sgA), qb = sb + ns * (jp + K * sgB);
            const double ca = coef[qa], cb = coef[qb];
            const double aa = alpha[qa], ab = alpha[qb];
            const double p = aa + ab, mu = aa * ab / p, oo2p = 0.5 / p;
            const double pref0 = pow(ORC_PI / p, 1.5);
            double *Es[3] = {Ex, Ey, Ez};
            for (int ax = 0; ax < 3; ++ax) {
                const double P = (aa * A[ax] + ab * Bv[ax]) / p;
                const double AB = A[ax] - Bv[ax];
                md_etab(Es[ax], la + 2, lb + 2, P - A[ax], P - Bv[ax], oo2p, exp(-mu * AB * AB));
            }
            double Bsum = 0.0, Aa = 0.0, Ab = 0.0;
            int index1 = 0;
            for (int ll1 = la; ll1 >= 0; --ll1)
                for (int n1 = 0; n1 <= la - ll1; ++n1) {
                    const int m1 = la - ll1 - n1;
                    int index2 = 0;
                    for (int ll2 = lb; ll2 >= 0; --ll2)
                        for (int n2 = 0; n2 <= lb - ll2; ++n2) {
                            const int m2 = lb - ll2 - n2;
                            const int lin = index1 + N1 * index2;
                            const int64_t i = lin % nba, j = lin / nba;
                            const double g = g_b[(off[sa] + i) + N * (off[sb] + j)];
                            const int LA = la + 2, LB = lb + 2;
                            const double ex = e0(Ex, LA, LB, ll1, ll2), ey = e0(Ey, LA, LB, m1, m2),
                                         ez = e0(Ez, LA, LB, n1, n2);
                            const double eprod = ex * ey * ez;
                            const double dGa = -(e0(Ex, LA, LB, ll1 + 2, ll2) * ey * ez +
                                                 ex * e0(Ey, LA, LB, m1 + 2, m2) * ez +
                                                 ex * ey * e0(Ez, LA, LB, n1 + 2, n2));
                            const double dGb = -(e0(Ex, LA, LB, ll1, ll2 + 2) * ey * ez +
                                                 ex * e0(Ey, LA, LB, m1, m2 + 2) * ez +
                                                 ex * ey * e0(Ez, LA, LB, n1, n2 + 2));
                            Bsum += g * eprod;
                            Aa += g * dGa;
                            Ab += g * dGb;
                            ++index2;
                        }
                    ++index1;
                }
            dcoef[qa] += cb * pref0 * Bsum;
            dcoef[qb] += ca * pref0 * Bsum;
            dalpha[qa] += ca * cb * pref0 * Aa;
            dalpha[qb] += ca * cb * pref0 * Ab;
        }
}

#define ORC_ETAB_DOUBLES (3 * ORC_T * (ORC_LE + 1) * (ORC_LE + 1))

void orc_ovl2c_pb(const double *dS, const double *posA, const double *posB,
                  const int64_t *sidxA, const int64_t *sidxB, int64_t B, int64_t ns, int64_t K,
                  int64_t NZ, const int64_t *ls, const double *coef, const double *alpha,
                  double *dcoef, double *dalpha)
{
    int64_t *nbf = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * ns + 1));
    int64_t *off = nbf + ns;
    const int64_t N = shell_layout(ns, ls, nbf, off);
    memset(dcoef, 0, sizeof(double) * (size_t)(ns * K * NZ));
    memset(dalpha, 0, sizeof(double) * (size_t)(ns * K * NZ));
    double *Ex = (double *)malloc(sizeof(double) * ORC_ETAB_DOUBLES);
    for (int64_t b = 0; b < B; ++b)
        for (int64_t sb = 0; sb < ns; ++sb)
            for (int64_t sa = 0; sa < ns; ++sa)
                ovl2c_pb_item(dS, posA, posB, sidxA, sidxB, b, sa, sb, ns, K, ls, nbf, off, N, coef, alpha, Ex,
                              dcoef, dalpha);
    free(Ex);
    free(nbf);
}

/* The same work items split statically across OpenMP threads over the linearised
 * (b, s_a, s_b) ndrange (KernelAbstractions' CPU backend splits the ndrange across
 * Julia threads); the reference's atomics become per-thread partial arrays, summed in
 * thread order at the end.  CPU-baseline timing only. */
void orc_ovl2c_pb_omp(const double *dS, const double *posA, const double *posB,
                      const int64_t *sidxA, const int64_t *sidxB, int64_t B, int64_t ns, int64_t K,
                      int64_t NZ, const int64_t *ls, const double *coef, const double *alpha,
                      double *dcoef, double *dalpha)
{
    int64_t *nbf = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * ns + 1));
    int64_t *off = nbf + ns;
    const int64_t N = shell_layout(ns, ls, nbf, off);
    const int64_t npar = ns * K * NZ;
    const int nt = orc_num_threads();
    double *part = (double *)calloc((size_t)(2 * npar * nt), sizeof(double));
    const int64_t items = B * ns * ns;
#pragma omp parallel
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num();
#else
        const int t = 0;
#endif
        double *Ex = (double *)malloc(sizeof(double) * ORC_ETAB_DOUBLES);
        double *pc = part + 2 * npar * t, *pa = pc + npar;
#pragma omp for schedule(static)
        for (int64_t it = 0; it < items; ++it) {
            const int64_t b = it % B, sa = (it / B) % ns, sb = it / (B * ns);
            ovl2c_pb_item(dS, posA, posB, sidxA, sidxB, b, sa, sb, ns, K, ls, nbf, off, N, coef, alpha, Ex, pc,
                          pa);
        }
        free(Ex);
    }
    for (int64_t q = 0; q < npar; ++q) {
        double c = 0.0, a = 0.0;
        for (int t = 0; t < nt; ++t) {
            c += part[2 * npar * t + q];
            a += part[2 * npar * t + npar + q];
        }
        dcoef[q] = c;
        dalpha[q] = a;
    }
    free(part);
    free(nbf);
}

/* Position cotangents of the 2-centre overlap.  No reference counterpart (the
 * reference's rrule returns NoTangent for the positions, overlap_2c_grad.jl:309-318).
 * Differentiates the forward exactly as computed (same alias map, same skip of zero
 * coefficients): d/dA_x [x_A^i e^{-a x_A^2}] = 2a x_A^{i+1} e - i x_A^{i-1} e, so each
 * term's A-derivative is prefac (2a E^{i+1,j}_0 - i E^{i-1,j}_0) E^y E^z (and likewise
 * for y, z).  dA[3b+c] = sum_ij dS[i,j,b] dS_ij/dA_c; dB = -dA (S depends on A - B). */
void orc_ovl2c_dx(const double *dS, const double *posA, const double *posB,
                  const int64_t *sidxA, const int64_t *sidxB, int64_t B, int64_t ns, int64_t K,
                  int64_t NZ, const int64_t *ls, const double *coef, const double *alpha,
                  double *dA, double *dB)
{
    int64_t *nbf = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * ns + 1));
    int64_t *off = nbf + ns;
    const int64_t N = shell_layout(ns, ls, nbf, off);
    (void)NZ;
    double *Ex = (double *)malloc(sizeof(double) * 3 * ORC_T * (ORC_LE + 1) * (ORC_LE + 1));
    double *Ey = Ex + ORC_T * (ORC_LE + 1) * (ORC_LE + 1), *Ez = Ey + ORC_T * (ORC_LE + 1) * (ORC_LE + 1);
    for (int64_t b = 0; b < B; ++b) {
        const int64_t sgA = sidxA ? sidxA[b] - 1 : 0, sgB = sidxB ? sidxB[b] - 1 : 0;
        const double *A = posA + 3 * b, *Bv = posB + 3 * b;
        const double *g_b = dS + N * N * b;
        double acc[3] = {0.0, 0.0, 0.0};
        for (int64_t sb = 0; sb < ns; ++sb)
            for (int64_t sa = 0; sa < ns; ++sa) {
                const int la = (int)ls[sa], lb = (int)ls[sb];
                const int64_t nba = nbf[sa];
                const int N1 = 2 * la + 1, LA = la + 1, LB = lb;
                for (int64_t ip = 0; ip < K; ++ip)
                    for (int64_t jp = 0; jp < K; ++jp) {
                        const int64_t qa = sa + ns * (ip + K * sgA), qb = sb + ns * (jp + K * sgB);
                        const double ca = coef[qa], cb = coef[qb];
                        if (ca == 0.0 || cb == 0.0) continue;
                        const double aa = alpha[qa], ab = alpha[qb];
                        const double p = aa + ab, mu = aa * ab / p, oo2p = 0.5 / p;
                        const double prefac = ca * cb * pow(ORC_PI / p, 1.5);
                        double *Es[3] = {Ex, Ey, Ez};
                        for (int ax = 0; ax < 3; ++ax) {
                            const double P = (aa * A[ax] + ab * Bv[ax]) / p;
                            const double AB = A[ax] - Bv[ax];
                            md_etab(Es[ax], LA, LB, P - A[ax], P - Bv[ax], oo2p, exp(-mu * AB * AB));
                        }
#define ORC_DE(E, i, j) (2.0 * aa * e0((E), LA, LB, (i) + 1, (j)) - ((i) > 0 ? (i) * e0((E), LA, LB, (i) - 1, (j)) : 0.0))
                        double hx = 0.0, hy = 0.0, hz = 0.0;
                        int index1 = 0;
                        for (int ll1 = la; ll1 >= 0; --ll1)
                            for (int n1 = 0; n1 <= la - ll1; ++n1) {
                                const int m1 = la - ll1 - n1;
                                int index2 = 0;
                                for (int ll2 = lb; ll2 >= 0; --ll2)
                                    for (int n2 = 0; n2 <= lb - ll2; ++n2) {
                                        const int m2 = lb - ll2 - n2;
                                        const int lin = index1 + N1 * index2;
                                        const int64_t i = lin % nba, j = lin / nba;
                                        const double g = g_b[(off[sa] + i) + N * (off[sb] + j)];
                                        const double ex = e0(Ex, LA, LB, ll1, ll2), ey = e0(Ey, LA, LB, m1, m2),
                                                     ez = e0(Ez, LA, LB, n1, n2);
                                        hx += g * ORC_DE(Ex, ll1, ll2) * ey * ez;
                                        hy += g * ex * ORC_DE(Ey, m1, m2) * ez;
                                        hz += g * ex * ey * ORC_DE(Ez, n1, n2);
                                        ++index2;
                                    }
                                ++index1;
                            }
#undef ORC_DE
                        acc[0] += prefac * hx;
                        acc[1] += prefac * hy;
                        acc[2] += prefac * hz;
                    }
            }
        for (int c = 0; c < 3; ++c) {
            dA[3 * b + c] = acc[c];
            if (dB) dB[3 * b + c] = -acc[c];
        }
    }
    free(Ex);
    free(nbf);
}

/* ------------------------------------------------------------------------- */
/* 3-centre overlap: E3 recursion about Q (reference/ecoeffs.jl:70-107, as in  */
/* kernels_3c.jl:54-131), contraction with the N1, N1*N2 strides and the       */
/* nbf_a, nbf_a*nbf_b read-back (kernels_3c.jl:137-161).                       */
/* ------------------------------------------------------------------------- */
/* Flat table E3[t + T*(i + (la+1)*(j + (lb+1)*k))], T = la+lb+lc+2. */
static void md_e3tab(double *E, int la, int lb, int lc, double QA, double QB, double QC,
                     double oo2q, double E000)
{
    const int T = la + lb + lc + 2;
#define E3IDX(t, i, j, k) ((t) + T * ((i) + (la + 1) * ((j) + (lb + 1) * (k))))
    memset(E, 0, sizeof(double) * (size_t)(T * (la + 1) * (lb + 1) * (lc + 1)));
    E[0] = E000;
    for (int k = 0; k <= lc; ++k)
        for (int j = 0; j <= lb; ++j)
            for (int i = 0; i <= la; ++i) {
                if (i == 0 && j == 0 && k == 0) continue;
                int pi = i, pj = j, pk = k;
                double X;
                if (i > 0) { pi = i - 1; X = QA; }
                else if (j > 0) { pj = j - 1; X = QB; }
                else { pk = k - 1; X = QC; }
                for (int t = 0; t <= i + j + k; ++t) {
                    const double em = t > 0 ? E[E3IDX(t - 1, pi, pj, pk)] : 0.0;
                    E[E3IDX(t, i, j, k)] = X * E[E3IDX(t, pi, pj, pk)] +
                                           (double)(t + 1) * E[E3IDX(t + 1, pi, pj, pk)] + oo2q * em;
                }
            }
#undef E3IDX
}

static inline double e30(const double *E, int la, int lb, int lc, int i, int j, int k)
{
    const int T = la + lb + lc + 2;
    return E[T * (i + (la + 1) * (j + (lb + 1) * k))];
}

int64_t orc_ovl3c(const double *posA, const double *posB, const double *posC,
                  const int64_t *sidxA, const int64_t *sidxB, const int64_t *sidxC, int64_t B,
                  int64_t ns, int64_t K, int64_t NZ, const int64_t *ls, const double *coef,
                  const double *alpha, double *out)
{
    int64_t *nbf = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * ns + 1));
    int64_t *off = nbf + ns;
    const int64_t N = shell_layout(ns, ls, nbf, off);
    for (int64_t s = 0; s < ns; ++s)
        if (ls[s] < 0 || ls[s] > ORC_LMAX) { free(nbf); return -2; }
    if (check_species(sidxA, B, NZ) || check_species(sidxB, B, NZ) ||
        check_species(sidxC, B, NZ)) { free(nbf); return -1; }
    const size_t e3n = (size_t)ORC_T3 * (ORC_LMAX + 1) * (ORC_LMAX + 1) * (ORC_LMAX + 1);
    double *E3 = (double *)malloc(sizeof(double) * 3 * e3n);
    const int64_t NB = (ORC_LMAX + 1) * (ORC_LMAX + 2) / 2;
    double *blk = (double *)malloc(sizeof(double) * (size_t)(NB * NB * NB));
    for (int64_t b = 0; b < B; ++b) {
        const int64_t sA = sidxA ? sidxA[b] - 1 : 0, sB = sidxB ? sidxB[b] - 1 : 0,
                      sC = sidxC ? sidxC[b] - 1 : 0;
        const double *A = posA + 3 * b, *Bv = posB + 3 * b, *C = posC + 3 * b;
        double *ob = out + N * N * N * b;
        for (int64_t sc = 0; sc < ns; ++sc)
            for (int64_t sb = 0; sb < ns; ++sb)
                for (int64_t sa = 0; sa < ns; ++sa) {
                    const int la = (int)ls[sa], lb = (int)ls[sb], lc = (int)ls[sc];
                    const int64_t na = nbf[sa], nb = nbf[sb], nc = nbf[sc];
                    const int N1 = 2 * la + 1, N2 = 2 * lb + 1;
                    for (int64_t q = 0; q < na * nb * nc; ++q) blk[q] = 0.0;
                    for (int64_t ip = 0; ip < K; ++ip)
                        for (int64_t jp = 0; jp < K; ++jp)
                            for (int64_t kp = 0; kp < K; ++kp) {
                                const int64_t qa = sa + ns * (ip + K * sA), qb = sb + ns * (jp + K * sB),
                                              qc = sc + ns * (kp + K * sC);
                                const double ca = coef[qa], cb = coef[qb], cc = coef[qc];
                                if (ca == 0.0 || cb == 0.0 || cc == 0.0) continue;
                                const double aa = alpha[qa], ab = alpha[qb], ac = alpha[qc];
                                const double p = aa + ab, q = p + ac;
                                const double mu = aa * ab / p, nu = p * ac / q, oo2q = 0.5 / q;
                                const double prefac = ca * cb * cc * pow(ORC_PI / q, 1.5);
                                for (int ax = 0; ax < 3; ++ax) {
                                    const double P = (aa * A[ax] + ab * Bv[ax]) / p;
                                    const double Qc = (p * P + ac * C[ax]) / q;
                                    const double AB = A[ax] - Bv[ax], PC = P - C[ax];
                                    md_e3tab(E3 + e3n * (size_t)ax, la, lb, lc, Qc - A[ax], Qc - Bv[ax], Qc - C[ax],
                                             oo2q, exp(-mu * AB * AB - nu * PC * PC));
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
                                                        const int lin = index1 + N1 * index2 + N1 * N2 * index3;
                                                        blk[lin] += e30(E3, la, lb, lc, ll1, ll2, ll3) *
                                                                    e30(E3 + e3n, la, lb, lc, m1, m2, m3) *
                                                                    e30(E3 + 2 * e3n, la, lb, lc, n1, n2, n3) *
                                                                    prefac;
                                                        ++index3;
                                                    }
                                                ++index2;
                                            }
                                        ++index1;
                                    }
                            }
                    for (int64_t kk = 0; kk < nc; ++kk)
                        for (int64_t jj = 0; jj < nb; ++jj)
                            for (int64_t ii = 0; ii < na; ++ii)
                                ob[(off[sa] + ii) + N * ((off[sb] + jj) + N * (off[sc] + kk))] =
                                    blk[ii + na * (jj + nb * kk)];
                }
    }
    free(blk);
    free(E3);
    free(nbf);
    return N;
}

/* Parameter cotangents of the 3-centre overlap.  No reference counterpart (the reference
 * has no 3-centre pullback).  As the 2-centre pullback (overlap_2c_grad.jl:71-214): every
 * K^3 primitive triple, padded slots included (a padded slot's coefficient cotangent is
 * the derivative at coef = 0); E3 tables to l+2 on every centre; the cotangent read back
 * through the 3-centre alias map.  d/dalpha x^l e^{-alpha x^2} = -x^{l+2} e^{-alpha x^2},
 * so dV/dalpha_a is minus the A-raised integral summed over the three axes. */
void orc_ovl3c_pb(const double *dT, const double *posA, const double *posB, const double *posC,
                  const int64_t *sidxA, const int64_t *sidxB, const int64_t *sidxC, int64_t B,
                  int64_t ns, int64_t K, int64_t NZ, const int64_t *ls, const double *coef,
                  const double *alpha, double *dcoef, double *dalpha)
{
    int64_t *nbf = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * ns + 1));
    int64_t *off = nbf + ns;
    const int64_t N = shell_layout(ns, ls, nbf, off);
    memset(dcoef, 0, sizeof(double) * (size_t)(ns * K * NZ));
    memset(dalpha, 0, sizeof(double) * (size_t)(ns * K * NZ));
    const size_t e3n = (size_t)(3 * (ORC_LMAX + 2) + 2) * (ORC_LMAX + 3) * (ORC_LMAX + 3) * (ORC_LMAX + 3);
    double *E3 = (double *)malloc(sizeof(double) * 3 * e3n);
    for (int64_t b = 0; b < B; ++b) {
        const int64_t sA = sidxA ? sidxA[b] - 1 : 0, sB = sidxB ? sidxB[b] - 1 : 0,
                      sC = sidxC ? sidxC[b] - 1 : 0;
        const double *A = posA + 3 * b, *Bv = posB + 3 * b, *C = posC + 3 * b;
        const double *g_b = dT + N * N * N * b;
        for (int64_t sc = 0; sc < ns; ++sc)
            for (int64_t sb = 0; sb < ns; ++sb)
                for (int64_t sa = 0; sa < ns; ++sa) {
                    const int la = (int)ls[sa], lb = (int)ls[sb], lc = (int)ls[sc];
                    const int LA = la + 2, LB = lb + 2, LC = lc + 2;
                    const int64_t na = nbf[sa], nb = nbf[sb];
                    const int N1 = 2 * la + 1, N2 = 2 * lb + 1;
                    for (int64_t ip = 0; ip < K; ++ip)
                        for (int64_t jp = 0; jp < K; ++jp)
                            for (int64_t kp = 0; kp < K; ++kp) {
                                const int64_t qa = sa + ns * (ip + K * sA), qb = sb + ns * (jp + K * sB),
                                              qc = sc + ns * (kp + K * sC);
                                const double ca = coef[qa], cb = coef[qb], cc = coef[qc];
                                /* two zero coefficients make every cotangent of the triple 0 */
                                if ((ca == 0.0) + (cb == 0.0) + (cc == 0.0) > 1) continue;
                                const double aa = alpha[qa], ab = alpha[qb], ac = alpha[qc];
                                const double p = aa + ab, q = p + ac;
                                const double mu = aa * ab / p, nu = p * ac / q, oo2q = 0.5 / q;
                                const double pref0 = pow(ORC_PI / q, 1.5);
                                for (int ax = 0; ax < 3; ++ax) {
                                    const double P = (aa * A[ax] + ab * Bv[ax]) / p;
                                    const double Qc = (p * P + ac * C[ax]) / q;
                                    const double AB = A[ax] - Bv[ax], PC = P - C[ax];
                                    md_e3tab(E3 + e3n * (size_t)ax, LA, LB, LC, Qc - A[ax], Qc - Bv[ax], Qc - C[ax],
                                             oo2q, exp(-mu * AB * AB - nu * PC * PC));
                                }
                                const double *Ex = E3, *Ey = E3 + e3n, *Ez = E3 + 2 * e3n;
#define ORC_E(E, i, j, k) e30((E), LA, LB, LC, (i), (j), (k))
                                double H = 0.0, Ha = 0.0, Hb = 0.0, Hc = 0.0;
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
                                                        const int64_t lin = index1 + N1 * index2 + N1 * N2 * index3;
                                                        const int64_t ii = lin % na, jj = (lin / na) % nb,
                                                                      kk = lin / (na * nb);
                                                        const double g =
                                                            g_b[(off[sa] + ii) + N * ((off[sb] + jj) + N * (off[sc] + kk))];
                                                        const double x = ORC_E(Ex, ll1, ll2, ll3),
                                                                     y = ORC_E(Ey, m1, m2, m3),
                                                                     z = ORC_E(Ez, n1, n2, n3);
                                                        H += g * x * y * z;
                                                        Ha -= g * (ORC_E(Ex, ll1 + 2, ll2, ll3) * y * z +
                                                                   x * ORC_E(Ey, m1 + 2, m2, m3) * z +
                                                                   x * y * ORC_E(Ez, n1 + 2, n2, n3));
                                                        Hb -= g * (ORC_E(Ex, ll1, ll2 + 2, ll3) * y * z +
                                                                   x * ORC_E(Ey, m1, m2 + 2, m3) * z +
                                                                   x * y * ORC_E(Ez, n1, n2 + 2, n3));
                                                        Hc -= g * (ORC_E(Ex, ll1, ll2, ll3 + 2) * y * z +
                                                                   x * ORC_E(Ey, m1, m2, m3 + 2) * z +
                                                                   x * y * ORC_E(Ez, n1, n2, n3 + 2));
                                                        ++index3;
                                                    }
                                                ++index2;
                                            }
                                        ++index1;
                                    }
#undef ORC_E
                                dcoef[qa] += cb * cc * pref0 * H;
                                dcoef[qb] += ca * cc * pref0 * H;
                                dcoef[qc] += ca * cb * pref0 * H;
                                dalpha[qa] += ca * cb * cc * pref0 * Ha;
                                dalpha[qb] += ca * cb * cc * pref0 * Hb;
                                dalpha[qc] += ca * cb * cc * pref0 * Hc;
                            }
                }
    }
    free(E3);
    free(nbf);
}
