"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by
``bench.py``'s ``cpu_baseline`` leg, always as the checker (or the timed CPU
baseline), never as a product code path.

Every function takes numpy arrays in the reference's Julia column-major layouts
(see ``aok_oracle.h``) and returns numpy arrays in the same layouts.  Species
indices are 1-based like Julia's.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int64)


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_ylm.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D]
        L.orc_ao_eval.argtypes = [_D, _I, ctypes.c_int64, _D, _D, ctypes.c_int64, ctypes.c_int64,
                                  ctypes.c_int64, _I, _I, _I, ctypes.c_int64, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_int, _D, _D]
        L.orc_ao_pullback.argtypes = [_D, _I, ctypes.c_int64, _D, _D, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_int64, _I, _I, _I, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_int, _D, _D, _D, _D]
        L.orc_cart_normsq.argtypes = [_D, _D, ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
        L.orc_cart_normsq.restype = ctypes.c_double
        L.orc_compile_coef.argtypes = [_D, _D, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _I, _D]
        L.orc_compile_coef_pb.argtypes = [_D, _D, _D, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                          _I, _D, _D]
        for name in ("orc_ovl2c", "orc_ovl2c_omp"):
            f = getattr(L, name)
            f.argtypes = [_D, _D, _I, _I, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                          ctypes.c_int64, _I, _D, _D, _D]
            f.restype = ctypes.c_int64
        for name in ("orc_ovl2c_pb", "orc_ovl2c_pb_omp"):
            getattr(L, name).argtypes = [_D, _D, _D, _I, _I, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.c_int64, _I, _D, _D, _D, _D]
        L.orc_ovl2c_dx.argtypes = [_D, _D, _D, _I, _I, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_int64, _I, _D, _D, _D, _D]
        L.orc_ovl3c.argtypes = [_D, _D, _D, _I, _I, _I, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                ctypes.c_int64, _I, _D, _D, _D]
        L.orc_ovl3c.restype = ctypes.c_int64
        L.orc_ovl3c_pb.argtypes = [_D, _D, _D, _D, _I, _I, _I, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_int64, _I, _D, _D, _D, _D]
        L.orc_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def num_threads() -> int:
    """OpenMP threads the oracle's parallel forms use (OMP_NUM_THREADS or the machine's)."""
    return int(lib().orc_num_threads())


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


def _i(a):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(_I)


def _fortran_flat(a):
    """Column-major flattening of a numpy array (Julia memory order)."""
    return np.asarray(a, dtype=np.float64).reshape(-1, order="F")


def ylm(L: int, xyz):
    """Z[(L+1)^2] and dZ[(L+1)^2, 3] at one point."""
    n = (L + 1) ** 2
    Z = np.zeros(n)
    dZ = np.zeros(3 * n)
    lib().orc_ylm(L, float(xyz[0]), float(xyz[1]), float(xyz[2]), Z.ctypes.data_as(_D), dZ.ctypes.data_as(_D))
    return Z, dZ.reshape(n, 3)


def ao_eval(X, sidx, zeta, D, poly, iR, iY, L, lengthscale=1.0, decay=0, grad=False):
    """P (nX, nOrb) and, with grad, dP (nX, nOrb, 3).  X is (nX, 3); zeta/D are
    (nRad, K, NZ) numpy arrays; iR/iY/sidx 1-based."""
    X = np.asarray(X, dtype=np.float64)
    nX = X.shape[0]
    nRad, K, NZ = zeta.shape
    nOrb = len(iR)
    Xa, Xp = _d(X.reshape(-1))
    sa, sp = _i(sidx)
    za, zp = _d(_fortran_flat(zeta))
    Da, Dp = _d(_fortran_flat(D))
    pa, pp = _i(poly if poly is not None else np.zeros(nRad, np.int64))
    ra, rp = _i(iR)
    ya, yp = _i(iY)
    P = np.zeros(nX * nOrb)
    dP = np.zeros(3 * nX * nOrb) if grad else None
    rc = lib().orc_ao_eval(Xp, sp, nX, zp, Dp, nRad, K, NZ, pp, rp, yp, nOrb, L, float(lengthscale), decay,
                           P.ctypes.data_as(_D), dP.ctypes.data_as(_D) if grad else None)
    if rc != 0:
        raise ValueError(f"oracle ao_eval rc={rc}")
    P = P.reshape(nOrb, nX).T
    if grad:
        return P, dP.reshape(nOrb, nX, 3).transpose(1, 0, 2)
    return P


def ao_pullback(X, sidx, zeta, D, poly, iR, iY, L, dP, lengthscale=1.0, decay=0):
    """(dX (nX,3), dzeta (nRad,K,NZ), dD (nRad,K,NZ)) for cotangent dP (nX, nOrb)."""
    X = np.asarray(X, dtype=np.float64)
    nX = X.shape[0]
    nRad, K, NZ = zeta.shape
    nOrb = len(iR)
    Xa, Xp = _d(X.reshape(-1))
    sa, sp = _i(sidx)
    za, zp = _d(_fortran_flat(zeta))
    Da, Dp = _d(_fortran_flat(D))
    pa, pp = _i(poly if poly is not None else np.zeros(nRad, np.int64))
    ra, rp = _i(iR)
    ya, yp = _i(iY)
    ga, gp = _d(_fortran_flat(dP))
    dX = np.zeros(3 * nX)
    dz = np.zeros(nRad * K * NZ)
    dd = np.zeros(nRad * K * NZ)
    rc = lib().orc_ao_pullback(Xp, sp, nX, zp, Dp, nRad, K, NZ, pp, rp, yp, nOrb, L, float(lengthscale),
                               decay, gp, dX.ctypes.data_as(_D), dz.ctypes.data_as(_D), dd.ctypes.data_as(_D))
    if rc != 0:
        raise ValueError(f"oracle ao_pullback rc={rc}")
    shp = (nRad, K, NZ)
    return dX.reshape(nX, 3), dz.reshape(shp, order="F"), dd.reshape(shp, order="F")


def cart_normsq(Dk, zk, l):
    Da, Dp = _d(Dk)
    za, zp = _d(zk)
    return lib().orc_cart_normsq(Dp, zp, len(Da), 1, int(l))


def compile_coef(zeta, D, ls):
    ns, K, NZ = zeta.shape
    za, zp = _d(_fortran_flat(zeta))
    Da, Dp = _d(_fortran_flat(D))
    la, lp = _i(ls)
    coef = np.zeros(ns * K * NZ)
    lib().orc_compile_coef(zp, Dp, ns, K, NZ, lp, coef.ctypes.data_as(_D))
    return coef.reshape((ns, K, NZ), order="F")


def compile_coef_pb(dcoef, zeta, D, ls):
    ns, K, NZ = zeta.shape
    ga, gp = _d(_fortran_flat(dcoef))
    za, zp = _d(_fortran_flat(zeta))
    Da, Dp = _d(_fortran_flat(D))
    la, lp = _i(ls)
    dD = np.zeros(ns * K * NZ)
    dz = np.zeros(ns * K * NZ)
    lib().orc_compile_coef_pb(gp, zp, Dp, ns, K, NZ, lp, dD.ctypes.data_as(_D), dz.ctypes.data_as(_D))
    shp = (ns, K, NZ)
    return dD.reshape(shp, order="F"), dz.reshape(shp, order="F")


def nbf_total(ls):
    return int(sum((l + 1) * (l + 2) // 2 for l in ls))


def ovl2c(posA, posB, sidxA, sidxB, ls, coef, alpha, omp=False):
    """out (N, N, B) in Bohr positions posA/posB (B, 3)."""
    posA = np.asarray(posA, np.float64)
    B = posA.shape[0]
    ns, K, NZ = coef.shape
    N = nbf_total(ls)
    Aa, Ap = _d(posA.reshape(-1))
    Ba, Bp = _d(np.asarray(posB, np.float64).reshape(-1))
    sa, sap = _i(sidxA)
    sb, sbp = _i(sidxB)
    la, lp = _i(ls)
    ca, cp = _d(_fortran_flat(coef))
    aa, ap = _d(_fortran_flat(alpha))
    out = np.zeros(N * N * B)
    f = lib().orc_ovl2c_omp if omp else lib().orc_ovl2c
    rc = f(Ap, Bp, sap, sbp, B, ns, K, NZ, lp, cp, ap, out.ctypes.data_as(_D))
    if rc < 0:
        raise ValueError(f"oracle ovl2c rc={rc}")
    return out.reshape((N, N, B), order="F")


def ovl2c_pb(dS, posA, posB, sidxA, sidxB, ls, coef, alpha, omp=False):
    posA = np.asarray(posA, np.float64)
    B = posA.shape[0]
    ns, K, NZ = coef.shape
    ga, gp = _d(_fortran_flat(dS))
    Aa, Ap = _d(posA.reshape(-1))
    Ba, Bp = _d(np.asarray(posB, np.float64).reshape(-1))
    sa, sap = _i(sidxA)
    sb, sbp = _i(sidxB)
    la, lp = _i(ls)
    ca, cp = _d(_fortran_flat(coef))
    aa, ap = _d(_fortran_flat(alpha))
    dc = np.zeros(ns * K * NZ)
    da = np.zeros(ns * K * NZ)
    f = lib().orc_ovl2c_pb_omp if omp else lib().orc_ovl2c_pb
    f(gp, Ap, Bp, sap, sbp, B, ns, K, NZ, lp, cp, ap, dc.ctypes.data_as(_D),
                       da.ctypes.data_as(_D))
    shp = (ns, K, NZ)
    return dc.reshape(shp, order="F"), da.reshape(shp, order="F")


def ovl2c_dx(dS, posA, posB, sidxA, sidxB, ls, coef, alpha):
    """Position cotangents (dA, dB), each (B, 3), of the 2-centre overlap (Bohr)."""
    posA = np.asarray(posA, np.float64)
    B = posA.shape[0]
    ns, K, NZ = coef.shape
    ga, gp = _d(_fortran_flat(dS))
    Aa, Ap = _d(posA.reshape(-1))
    Ba, Bp = _d(np.asarray(posB, np.float64).reshape(-1))
    sa, sap = _i(sidxA)
    sb, sbp = _i(sidxB)
    la, lp = _i(ls)
    ca, cp = _d(_fortran_flat(coef))
    aa, ap = _d(_fortran_flat(alpha))
    dA = np.zeros(3 * B)
    dB = np.zeros(3 * B)
    lib().orc_ovl2c_dx(gp, Ap, Bp, sap, sbp, B, ns, K, NZ, lp, cp, ap, dA.ctypes.data_as(_D),
                       dB.ctypes.data_as(_D))
    return dA.reshape(B, 3), dB.reshape(B, 3)


def ovl3c(posA, posB, posC, sidxA, sidxB, sidxC, ls, coef, alpha):
    posA = np.asarray(posA, np.float64)
    B = posA.shape[0]
    ns, K, NZ = coef.shape
    N = nbf_total(ls)
    Aa, Ap = _d(posA.reshape(-1))
    Ba, Bp = _d(np.asarray(posB, np.float64).reshape(-1))
    Ca, Cp = _d(np.asarray(posC, np.float64).reshape(-1))
    sa, sap = _i(sidxA)
    sb, sbp = _i(sidxB)
    sc, scp = _i(sidxC)
    la, lp = _i(ls)
    ca, cp = _d(_fortran_flat(coef))
    aa, ap = _d(_fortran_flat(alpha))
    out = np.zeros(N * N * N * B)
    rc = lib().orc_ovl3c(Ap, Bp, Cp, sap, sbp, scp, B, ns, K, NZ, lp, cp, ap, out.ctypes.data_as(_D))
    if rc < 0:
        raise ValueError(f"oracle ovl3c rc={rc}")
    return out.reshape((N, N, N, B), order="F")


def ovl3c_pb(dT, posA, posB, posC, sidxA, sidxB, sidxC, ls, coef, alpha):
    """(dcoef, dalpha), each (ns, K, NZ), of the 3-centre overlap (Bohr positions)."""
    posA = np.asarray(posA, np.float64)
    B = posA.shape[0]
    ns, K, NZ = coef.shape
    ga, gp = _d(_fortran_flat(dT))
    Aa, Ap = _d(posA.reshape(-1))
    Ba, Bp = _d(np.asarray(posB, np.float64).reshape(-1))
    Ca, Cp = _d(np.asarray(posC, np.float64).reshape(-1))
    sa, sap = _i(sidxA)
    sb, sbp = _i(sidxB)
    sc, scp = _i(sidxC)
    la, lp = _i(ls)
    ca, cp = _d(_fortran_flat(coef))
    aa, ap = _d(_fortran_flat(alpha))
    dc = np.zeros(ns * K * NZ)
    da = np.zeros(ns * K * NZ)
    lib().orc_ovl3c_pb(gp, Ap, Bp, Cp, sap, sbp, scp, B, ns, K, NZ, lp, cp, ap, dc.ctypes.data_as(_D),
                       da.ctypes.data_as(_D))
    shp = (ns, K, NZ)
    return dc.reshape(shp, order="F"), da.reshape(shp, order="F")


def pullback_ps_3c(dT, posA, posB, posC, sidxA, sidxB, sidxC, ls, zeta, D):
    """(dzeta, dD) of the 3-centre overlap of an orbital basis (coef compiled from (zeta, D))."""
    coef = compile_coef(zeta, D, ls)
    dc, da = ovl3c_pb(dT, posA, posB, posC, sidxA, sidxB, sidxC, ls, coef, zeta)
    dD, dzn = compile_coef_pb(dc, zeta, D, ls)
    return da + dzn, dD


def pullback_ps_2c(dS, posA, posB, sidxA, sidxB, ls, zeta, D, omp=False):
    """Full (dzeta, dD) of the Lux 2c overlap (overlap_2c_grad.jl:296-304)."""
    coef = compile_coef(zeta, D, ls)
    dcoef, dalpha = ovl2c_pb(dS, posA, posB, sidxA, sidxB, ls, coef, zeta, omp=omp)
    dD, dz_norm = compile_coef_pb(dcoef, zeta, D, ls)
    return dalpha + dz_norm, dD
