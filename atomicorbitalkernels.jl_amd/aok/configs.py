"""Seeded synthetic workloads of the benchmark configurations.

All generators are numpy ``default_rng`` based and return FP64 host arrays in the
reference's logical shapes (``zeta, D`` as ``(n, K, NZ)``, positions ``(n, 3)``,
1-based species indices).  FP32 runs round these to FP32 and the FP64 oracle is fed
the same rounded values, so measured errors are arithmetic only.

Padding follows the reference's union-spec convention (src/orbitals/gaussianbasis.jl:69-71):
unused primitive slots carry zeta = 1, D = 0.  D is normalised per contracted function so
that the (L2-normalised solid-harmonic) orbital has unit norm:
    sum_ij D_i D_j 1/2 Gamma(l + 3/2) (zeta_i + zeta_j)^-(l + 3/2) = 1.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from math import gamma


def normalise(D, zeta, l):
    e = l + 1.5
    s = 0.0
    for i in range(len(D)):
        for j in range(len(D)):
            s += D[i] * D[j] * 0.5 * gamma(e) * (zeta[i] + zeta[j]) ** (-e)
    return D / np.sqrt(s) if s > 0 else D


def _logu(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


@dataclass
class ShellBasis:
    """Species-aware shell table: ls (per shell), zeta/D (ns, K, NZ), nprim (ns, NZ)."""

    ls: list
    zeta: np.ndarray
    D: np.ndarray
    nprim: np.ndarray

    @property
    def ns(self):
        return len(self.ls)

    @property
    def K(self):
        return self.zeta.shape[1]

    @property
    def NZ(self):
        return self.zeta.shape[2]

    @property
    def N(self):
        return int(sum((l + 1) * (l + 2) // 2 for l in self.ls))

    @property
    def nOrb(self):
        return int(sum(2 * l + 1 for l in self.ls))

    def ao_maps(self):
        """spec (n, l), iR, iY (1-based) of the AtomicOrbitals built on these radials."""
        spec, iR, iY, count = [], [], [], {}
        for k, l in enumerate(self.ls):
            count[l] = count.get(l, 0) + 1
            spec.append((count[l], l))
            for m in range(-l, l + 1):
                iR.append(k + 1)
                iY.append(l * l + l + m + 1)
        return spec, iR, iY


def _pad_basis(ls, shells_per_species, K):
    """shells_per_species[s][k] = (zeta list, D list) or None (absent shell)."""
    NZ = len(shells_per_species)
    ns = len(ls)
    zeta = np.ones((ns, K, NZ))
    D = np.zeros((ns, K, NZ))
    nprim = np.zeros((ns, NZ), dtype=np.int64)
    for s, shells in enumerate(shells_per_species):
        for k, sh in enumerate(shells):
            if sh is None:
                continue
            z, d = sh
            zeta[k, :len(z), s] = z
            D[k, :len(d), s] = normalise(np.asarray(d, float), np.asarray(z, float), ls[k])
            nprim[k, s] = len(z)
    return ShellBasis(list(ls), zeta, D, nprim)


# ---- C0: AO evaluation, FP64, single species, 3s2p1d, K = 4 --------------------------------
def c0_basis(seed=0):
    rng = np.random.default_rng(seed)
    ls = [0, 0, 0, 1, 1, 2]
    shells = []
    for l in ls:
        z = np.sort(_logu(rng, 0.1, 50.0, 4))[::-1]
        shells.append((z, rng.standard_normal(4)))
    return _pad_basis(ls, [shells], 4), rng


def c0(seed=0, nX=10_000):
    basis, rng = c0_basis(seed)
    X = rng.normal(0.0, 1.5, size=(nX, 3))
    return {"basis": basis, "X": X, "sidx": None, "L": 2}


# ---- C1 species: three double-zeta-shaped sets, 9s4p1d -> 3s2p1d, K = 9 -----------------------
C1_LS = [0, 0, 0, 1, 1, 2]


def c1_basis(seed=1):
    rng = np.random.default_rng(seed)
    species = []
    for _ in range(3):
        zs = np.sort(_logu(rng, 0.1, 7000.0, 9))[::-1]
        zp = np.sort(_logu(rng, 0.1, 7000.0, 4))[::-1]
        shells = [
            (zs, rng.standard_normal(9)),                       # s1: general contraction
            (zs, rng.standard_normal(9)),                       # s2: same exponents, mixed signs
            (_logu(rng, 0.1, 7000.0, 1), np.ones(1)),           # s3
            (zp, rng.standard_normal(4)),                       # p1
            (_logu(rng, 0.1, 7000.0, 1), np.ones(1)),           # p2
            (_logu(rng, 0.1, 7000.0, 1), np.ones(1)),           # d
        ]
        species.append(shells)
    return _pad_basis(C1_LS, species, 9)


def _pairs(rng, B, rmin, rmax, sigma_A=0.4):
    A = rng.normal(0.0, sigma_A, size=(B, 3))
    u = rng.standard_normal((B, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    R = rng.uniform(rmin, rmax, size=(B, 1))
    return A, A + R * u


def c1(seed=1, B=65536):
    basis = c1_basis(1)
    rng = np.random.default_rng(seed + 1000)
    posA, posB = _pairs(rng, B, 0.5, 6.0)
    sA = rng.integers(1, 4, size=B)
    sB = rng.integers(1, 4, size=B)
    return {"basis": basis, "posA": posA, "posB": posB, "sidxA": sA, "sidxB": sB}


# ---- C2: uneven batching -----------------------------------------------------------------------
def c2_basis(seed=2):
    rng = np.random.default_rng(seed)
    per_species = []
    while len(per_species) < 6:
        lmax = int(rng.integers(0, 5))
        counts = {l: int(rng.integers(1, 4)) for l in range(lmax + 1)}
        nfun = sum(c * (l + 1) * (l + 2) // 2 for l, c in counts.items())
        if not 5 <= nfun <= 55:
            continue
        shells = {}
        for l, c in counts.items():
            for n in range(c):
                depth = int(rng.integers(1, 11))
                z = np.sort(_logu(rng, 0.1, 7000.0, depth))[::-1]
                shells[(l, n)] = (z, rng.standard_normal(depth))
        per_species.append(shells)
    union = sorted({key for sp in per_species for key in sp})
    ls = [l for (l, _) in union]
    table = [[sp.get(key) for key in union] for sp in per_species]
    return _pad_basis(ls, table, 10)


def c2(seed=2, B=2**18):
    basis = c2_basis(seed)
    rng = np.random.default_rng(seed + 1000)
    pairs = [(a, b) for a in range(1, 7) for b in range(1, 7)]
    order = rng.permutation(len(pairs))
    w = np.arange(1, 37, dtype=float) ** -1.2
    w /= w.sum()
    pick = rng.choice(36, size=B, p=w)
    chosen = np.array([pairs[order[k]] for k in pick])
    posA, posB = _pairs(rng, B, 0.3, 8.0)
    return {"basis": basis, "posA": posA, "posB": posB, "sidxA": chosen[:, 0], "sidxB": chosen[:, 1]}


# ---- C3: differentiation -----------------------------------------------------------------------
def c3_overlap(seed=3, B=65536):
    d = c1(1, B)
    rng = np.random.default_rng(seed)
    basis = d["basis"]
    zt = basis.zeta * np.exp(0.05 * rng.standard_normal(basis.zeta.shape))
    Dt = basis.D * np.exp(0.05 * rng.standard_normal(basis.D.shape))
    zt = np.where(basis.D != 0, zt, basis.zeta)  # padded slots stay (1, 0)
    d["target"] = (zt, Dt)
    return d


def c3_ao(seed=3, nX=2**20):
    basis = c1_basis(1)
    rng = np.random.default_rng(seed + 2000)
    X = rng.normal(0.0, 2.0, size=(nX, 3))
    sidx = rng.integers(1, 4, size=nX)
    dP = rng.standard_normal((nX, basis.nOrb))
    return {"basis": basis, "X": X, "sidx": sidx, "dP": dP, "L": 2}


# ---- C4 -----------------------------------------------------------------------------------------
def c4a(seed=4, B=2**14):
    basis = c1_basis(1)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-2.0, 2.0, size=(3, B, 3))
    sp = rng.integers(1, 4, size=(3, B))
    return {"basis": basis, "posA": pos[0], "posB": pos[1], "posC": pos[2], "sidxA": sp[0], "sidxB": sp[1],
            "sidxC": sp[2]}


C4B_LS = [0, 0, 0, 0, 1, 1, 1, 2, 2, 3]  # 4s3p2d1f -> 30 orbitals


def c4b_basis(seed=4):
    rng = np.random.default_rng(seed + 3000)
    species = []
    for _ in range(3):
        shells = []
        for l in C4B_LS:
            z = np.sort(_logu(rng, 0.1, 50.0, 4))[::-1]
            shells.append((z, rng.standard_normal(4)))
        species.append(shells)
    return _pad_basis(C4B_LS, species, 4)


def c4b(seed=4, nX=2**24):
    basis = c4b_basis(seed)
    rng = np.random.default_rng(seed + 4000)
    X = rng.normal(0.0, 3.0, size=(nX, 3))
    sidx = rng.integers(1, 4, size=nX)
    return {"basis": basis, "X": X, "sidx": sidx, "L": 3}


# ---- g shells: an AO / overlap basis with every l up to 4 -------------------------------------
G_LS = [0, 0, 1, 1, 2, 3, 4]  # 2s2p1d1f1g -> 29 orbitals, 39 Cartesian functions


def g_basis(seed=5, NZ=3, K=4, zmax=50.0):
    """Species with s..g shells, K primitives each (depths 1..K drawn per shell, so the
    padded (zeta = 1, D = 0) slots of the union layout are exercised), zeta ~ logU[0.1, zmax]."""
    rng = np.random.default_rng(seed + 5000)
    species = []
    for _ in range(NZ):
        shells = []
        for l in G_LS:
            depth = int(rng.integers(1, K + 1))
            z = np.sort(_logu(rng, 0.1, zmax, depth))[::-1]
            shells.append((z, rng.standard_normal(depth)))
        species.append(shells)
    return _pad_basis(G_LS, species, K)
