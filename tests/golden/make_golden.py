"""Write tests/golden/golden.json: known-answer vectors from mathematics that is
independent of both the oracle and the kernels.

* Z_lm for l <= 2 from the closed forms of the SpheriCart convention (the reference
  stores harmonic (l, m) at l^2 + l + m + 1; test/orbitals/test_gaussianbasis.jl:40).
* Z_lm for l <= 4 from the textbook definition of the real solid harmonics (associated
  Legendre functions without the Condon-Shortley phase, sin |m| phi for m < 0, cos m phi for
  m > 0, r^l Y_lm, L2-normalised on the sphere), cross-checked against the published
  polynomial tables for l = 3 and l = 4 before anything is written.  This is the convention
  SpheriCart documents (the reference's harmonics, called at src/orbitals/atomicorbitals.jl:155,200).
* 2-centre contracted Cartesian overlaps by Gauss-Hermite quadrature of the true
  primitive overlaps, composed with the reference's contraction/read-back alias map
  (blk[mu + (2 l_a + 1) nu] += S_true[mu, nu]; out[i, j] = blk[i + nbf_a j],
  src/integrals/orbital/overlap_2c.jl:256-281) -- shells up to l = 4.
* 3-centre blocks the same way with the N1, N1*N2 strides (src/integrals/kernels_3c.jl:134-161).
* The reference's own known answers: (pi/(a+b))^{3/2} for two same-centre s
  primitives (test/integrals/test_reference.jl:10-20), ang2bohr.

Run:  python tests/golden/make_golden.py
"""
import json
import os
from math import factorial, pi, sqrt

import numpy as np
from numpy.polynomial import legendre as _leg

HERE = os.path.dirname(os.path.abspath(__file__))
GH_T, GH_W = np.polynomial.hermite.hermgauss(24)


def comps(l):
    """Cartesian component order: for ll = l..0, n = 0..l-ll: (x, y, z) = (ll, l-ll-n, n)."""
    return [(ll, l - ll - n, n) for ll in range(l, -1, -1) for n in range(0, l - ll + 1)]


def s1d(i, j, a, b, A, B):
    """int (x-A)^i (x-B)^j exp(-a (x-A)^2 - b (x-B)^2) dx by Gauss-Hermite."""
    p = a + b
    P = (a * A + b * B) / p
    mu = a * b / p
    x = P + GH_T / np.sqrt(p)
    return np.exp(-mu * (A - B) ** 2) / np.sqrt(p) * np.sum(GH_W * (x - A) ** i * (x - B) ** j)


def s1d3(i, j, k, a, b, c, A, B, C):
    q = a + b + c
    Q = (a * A + b * B + c * C) / q
    pre = np.exp(-(a * b * (A - B) ** 2 + a * c * (A - C) ** 2 + b * c * (B - C) ** 2) / q)
    x = Q + GH_T / np.sqrt(q)
    return pre / np.sqrt(q) * np.sum(GH_W * (x - A) ** i * (x - B) ** j * (x - C) ** k)


def true_block(la, lb, za, ca, zb, cb, A, B):
    ca_, cb_ = comps(la), comps(lb)
    S = np.zeros((len(ca_), len(cb_)))
    for (alpha, c1) in zip(za, ca):
        for (beta, c2) in zip(zb, cb):
            for mu, pa in enumerate(ca_):
                for nu, pb in enumerate(cb_):
                    v = 1.0
                    for ax in range(3):
                        v *= s1d(pa[ax], pb[ax], alpha, beta, A[ax], B[ax])
                    S[mu, nu] += c1 * c2 * v
    return S


def alias_2c(S, la):
    nba, nbb = S.shape
    blk = np.zeros(nba * nbb)
    for mu in range(nba):
        for nu in range(nbb):
            blk[mu + (2 * la + 1) * nu] += S[mu, nu]
    return blk.reshape((nba, nbb), order="F")


def true_block3(la, lb, lc, pa, pb, pc, A, B, C):
    c_a, c_b, c_c = comps(la), comps(lb), comps(lc)
    V = np.zeros((len(c_a), len(c_b), len(c_c)))
    for (za, ca) in zip(*pa):
        for (zb, cb) in zip(*pb):
            for (zc, cc) in zip(*pc):
                for i, qa in enumerate(c_a):
                    for j, qb in enumerate(c_b):
                        for k, qc in enumerate(c_c):
                            v = 1.0
                            for ax in range(3):
                                v *= s1d3(qa[ax], qb[ax], qc[ax], za, zb, zc, A[ax], B[ax], C[ax])
                            V[i, j, k] += ca * cb * cc * v
    return V


def alias_3c(V, la, lb):
    na, nb, nc = V.shape
    blk = np.zeros(na * nb * nc)
    N1, N2 = 2 * la + 1, 2 * lb + 1
    for i in range(na):
        for j in range(nb):
            for k in range(nc):
                blk[i + N1 * j + N1 * N2 * k] += V[i, j, k]
    return blk.reshape((na, nb, nc), order="F")


def ylm_closed(x, y, z):
    r2 = x * x + y * y + z * z
    return [0.28209479177387814, 0.4886025119029199 * y, 0.4886025119029199 * z, 0.4886025119029199 * x,
            1.0925484305920792 * x * y, 1.0925484305920792 * y * z, 0.31539156525252005 * (3 * z * z - r2),
            1.0925484305920792 * x * z, 0.5462742152960396 * (x * x - y * y)]


def zlm_definition(l, m, p):
    """r^l Y_lm from the definition: N_l|m| r^(l-|m|) P_l^(|m|)(z/r) times 1 (m = 0),
    sqrt(2) Re (x + iy)^|m| (m > 0) or sqrt(2) Im (x + iy)^|m| (m < 0), where P_l^(|m|) is
    the |m|-th derivative of the Legendre polynomial and N_lk = sqrt((2l+1)/(4 pi) (l-k)!/(l+k)!)."""
    x, y, z = p
    r = np.sqrt(x * x + y * y + z * z)
    am = abs(m)
    norm = sqrt((2 * l + 1) / (4 * pi) * factorial(l - am) / factorial(l + am))
    dleg = _leg.Legendre.basis(l).deriv(am)(z / r)
    c = complex(x, y) ** am
    ang = 1.0 if m == 0 else sqrt(2.0) * (c.real if m > 0 else c.imag)
    return norm * r ** (l - am) * dleg * ang


def zlm_table_l34(x, y, z):
    """Published polynomial forms of the real solid harmonics, l = 3 (m = -3..3) then
    l = 4 (m = -4..4)."""
    r2 = x * x + y * y + z * z
    l3 = [0.25 * sqrt(35 / (2 * pi)) * y * (3 * x * x - y * y), 0.5 * sqrt(105 / pi) * x * y * z,
          0.25 * sqrt(21 / (2 * pi)) * y * (5 * z * z - r2), 0.25 * sqrt(7 / pi) * z * (5 * z * z - 3 * r2),
          0.25 * sqrt(21 / (2 * pi)) * x * (5 * z * z - r2), 0.25 * sqrt(105 / pi) * (x * x - y * y) * z,
          0.25 * sqrt(35 / (2 * pi)) * x * (x * x - 3 * y * y)]
    l4 = [0.75 * sqrt(35 / pi) * x * y * (x * x - y * y), 0.75 * sqrt(35 / (2 * pi)) * y * z * (3 * x * x - y * y),
          0.75 * sqrt(5 / pi) * x * y * (7 * z * z - r2), 0.75 * sqrt(5 / (2 * pi)) * y * z * (7 * z * z - 3 * r2),
          3 / 16 * sqrt(1 / pi) * (35 * z ** 4 - 30 * z * z * r2 + 3 * r2 * r2),
          0.75 * sqrt(5 / (2 * pi)) * x * z * (7 * z * z - 3 * r2), 3 / 8 * sqrt(5 / pi) * (x * x - y * y) * (7 * z * z - r2),
          0.75 * sqrt(35 / (2 * pi)) * x * z * (x * x - 3 * y * y),
          3 / 16 * sqrt(35 / pi) * (x * x * (x * x - 3 * y * y) - y * y * (3 * x * x - y * y))]
    return l3 + l4


def ylm_l4_fixture(seed=20261204, npts=8):
    rng = np.random.default_rng(seed)
    pts = rng.normal(0, 1.3, size=(npts, 3))
    vals = []
    for p in pts:
        v = [zlm_definition(l, m, p) for l in range(5) for m in range(-l, l + 1)]
        tab = zlm_table_l34(*p)
        scale = max(1.0, float(np.max(np.abs(v))))
        assert np.max(np.abs(np.array(v[9:]) - tab)) < 1e-13 * scale, "l = 3, 4 tables disagree with the definition"
        assert np.max(np.abs(np.array(v[:9]) - ylm_closed(*p))) < 1e-13 * scale
        vals.append(v)
    return {"points": pts.tolist(), "values": vals}


def main():
    rng = np.random.default_rng(20261203)
    out = {"ang2bohr": 1.8897261246257702}
    pts = rng.normal(0, 1.2, size=(6, 3))
    out["ylm_l2"] = {"points": pts.tolist(), "values": [ylm_closed(*p) for p in pts]}
    # same-centre s/s closed form (reference test_reference.jl:10-20)
    out["ss_same_centre"] = {"alpha": 1.3, "beta": 0.7, "value": (np.pi / 2.0) ** 1.5}
    # 2c shell-pair blocks (primitive data given directly, positions in Bohr)
    blocks = []
    for (la, lb) in [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (2, 1), (1, 2), (2, 2), (3, 2), (2, 3), (3, 3), (4, 1),
                     (4, 4)]:
        ka, kb = int(rng.integers(1, 4)), int(rng.integers(1, 4))
        za = np.exp(rng.uniform(np.log(0.2), np.log(30), ka))
        zb = np.exp(rng.uniform(np.log(0.2), np.log(30), kb))
        ca = rng.standard_normal(ka)
        cb = rng.standard_normal(kb)
        A = rng.normal(0, 0.5, 3)
        B = A + rng.normal(0, 0.8, 3)
        S = true_block(la, lb, za, ca, zb, cb, A, B)
        blocks.append({"la": la, "lb": lb, "za": za.tolist(), "ca": ca.tolist(), "zb": zb.tolist(),
                       "cb": cb.tolist(), "A": A.tolist(), "B": B.tolist(),
                       "out": alias_2c(S, la).reshape(-1, order="F").tolist()})
    out["blocks_2c"] = blocks
    # p-shell transpose symmetry fixture (reference test_reference.jl:24-41)
    b3 = []
    for (la, lb, lc) in [(0, 0, 0), (1, 0, 1), (2, 1, 0), (1, 2, 2), (2, 2, 2)]:
        prims = []
        for l in (la, lb, lc):
            k = int(rng.integers(1, 3))
            prims.append((np.exp(rng.uniform(np.log(0.3), np.log(10), k)), rng.standard_normal(k)))
        A = rng.uniform(-1, 1, 3)
        B = rng.uniform(-1, 1, 3)
        C = rng.uniform(-1, 1, 3)
        V = true_block3(la, lb, lc, prims[0], prims[1], prims[2], A, B, C)
        b3.append({"la": la, "lb": lb, "lc": lc, "prims": [[p[0].tolist(), p[1].tolist()] for p in prims],
                   "A": A.tolist(), "B": B.tolist(), "C": C.tolist(),
                   "out": alias_3c(V, la, lb).reshape(-1, order="F").tolist()})
    out["blocks_3c"] = b3
    out["ylm_l4"] = ylm_l4_fixture()
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
