"""CPU tests that pin the FP64 oracle (oracle/aok_oracle.c) against mathematics it does
not share code with: golden quadrature vectors, closed forms, invariants and central
finite differences.  A plausible slip in the restatement (a dropped recursion term, a
wrong sign, a transposed operand, a wrong alias stride) fails at least one of these."""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def _two_shell(la, lb, za, ca, zb, cb):
    ka, kb = len(za), len(zb)
    K = max(ka, kb)
    zeta = np.ones((2, K, 1))
    coef = np.zeros((2, K, 1))
    zeta[0, :ka, 0], coef[0, :ka, 0] = za, ca
    zeta[1, :kb, 0], coef[1, :kb, 0] = zb, cb
    return [la, lb], coef, zeta


# ---------------------------------------------------------------- overlaps
@pytest.mark.parametrize("blk", GOLD["blocks_2c"], ids=lambda b: f"l{b['la']}{b['lb']}")
def test_ovl2c_golden_quadrature(blk):
    """MD recursion + alias map vs Gauss-Hermite quadrature of the true overlaps."""
    la, lb = blk["la"], blk["lb"]
    ls, coef, zeta = _two_shell(la, lb, blk["za"], blk["ca"], blk["zb"], blk["cb"])
    out = O.ovl2c(np.array([blk["A"]]), np.array([blk["B"]]), None, None, ls, coef, zeta)
    nba, nbb = (la + 1) * (la + 2) // 2, (lb + 1) * (lb + 2) // 2
    ref = np.array(blk["out"]).reshape((nba, nbb), order="F")
    got = out[:nba, nba:nba + nbb, 0]
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("blk", GOLD["blocks_3c"], ids=lambda b: f"l{b['la']}{b['lb']}{b['lc']}")
def test_ovl3c_golden_quadrature(blk):
    la, lb, lc = blk["la"], blk["lb"], blk["lc"]
    pr = blk["prims"]
    K = max(len(p[0]) for p in pr)
    zeta = np.ones((3, K, 1))
    coef = np.zeros((3, K, 1))
    for s, (z, c) in enumerate(pr):
        zeta[s, :len(z), 0], coef[s, :len(c), 0] = z, c
    out = O.ovl3c(np.array([blk["A"]]), np.array([blk["B"]]), np.array([blk["C"]]), None, None, None,
                  [la, lb, lc], coef, zeta)
    na, nb, nc = [(l + 1) * (l + 2) // 2 for l in (la, lb, lc)]
    got = out[:na, na:na + nb, na + nb:na + nb + nc, 0]
    ref = np.array(blk["out"]).reshape((na, nb, nc), order="F")
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_ss_same_centre_closed_form():
    """test/integrals/test_reference.jl:10-20: (pi/(a+b))^{3/2}."""
    g = GOLD["ss_same_centre"]
    ls, coef, zeta = _two_shell(0, 0, [g["alpha"]], [1.0], [g["beta"]], [1.0])
    out = O.ovl2c(np.zeros((1, 3)), np.zeros((1, 3)), None, None, ls, coef, zeta)
    assert abs(out[0, 1, 0] - g["value"]) < 1e-12
    assert abs(g["value"] - (np.pi / 2.0) ** 1.5) < 1e-15


def test_p_shell_transpose_symmetry():
    """test/integrals/test_reference.jl:24-41."""
    ls, coef, zeta = _two_shell(1, 1, [1.2, 0.4], [0.7, 0.3], [1.2, 0.4], [0.7, 0.3])
    A, B = np.array([[0.0, 0.0, 0.0]]), np.array([[0.4, 0.7, -0.2]])
    sab = O.ovl2c(A, B, None, None, ls, coef, zeta)[0:3, 3:6, 0]
    sba = O.ovl2c(B, A, None, None, ls, coef, zeta)[0:3, 3:6, 0]
    assert np.max(np.abs(sab - sba.T)) < 1e-12


def test_3c_reduces_to_2c():
    """V(alpha_c = 0, l_c = 0, c_c = 1) = S_2c (q = p, Q = P, nu = 0)."""
    rng = np.random.default_rng(3)
    ls = [0, 1, 2, 0]
    K = 3
    zeta = np.exp(rng.uniform(-1, 2, size=(4, K, 1)))
    coef = rng.standard_normal((4, K, 1))
    zeta[3], coef[3] = 1.0, 0.0
    zeta[3, 0, 0], coef[3, 0, 0] = 0.0, 1.0
    A, B, C = rng.normal(size=(2, 3)), rng.normal(size=(2, 3)), rng.normal(size=(2, 3))
    V = O.ovl3c(A, B, C, None, None, None, ls, coef, zeta)
    S = O.ovl2c(A, B, None, None, ls[:3], coef[:3], zeta[:3])
    N = 1 + 3 + 6
    assert np.max(np.abs(V[:N, :N, N, :] - S)) < 1e-12


def test_self_overlap_normalised():
    """compile_coef normalises x^l: out[off, off] = 1 at A = B for every shell."""
    rng = np.random.default_rng(5)
    ls = [0, 1, 2, 3, 4]
    K = 4
    zeta = np.exp(rng.uniform(-1, 3, size=(5, K, 2)))
    D = rng.standard_normal((5, K, 2))
    coef = O.compile_coef(zeta, D, ls)
    X = np.array([[0.3, -0.2, 0.1]])
    for sg in (1, 2):
        out = O.ovl2c(X, X, [sg], [sg], ls, coef, zeta)
        off = 0
        for l in ls:
            assert abs(out[off, off, 0] - 1.0) < 1e-12
            off += (l + 1) * (l + 2) // 2


def test_forward_skips_zero_coef_but_pullback_does_not():
    """Padded slots: no value contribution, non-zero coefficient cotangent, zero exponent
    cotangent (overlap_2c.jl:202, overlap_2c_grad.jl:104-110)."""
    ls, coef, zeta = _two_shell(1, 0, [0.8, 2.0], [1.0, 0.0], [0.5], [1.0])
    A, B = np.array([[0.0, 0.0, 0.0]]), np.array([[0.3, 0.2, 0.9]])
    S = O.ovl2c(A, B, None, None, ls, coef, zeta)
    ls2, coef2, zeta2 = _two_shell(1, 0, [0.8], [1.0], [0.5], [1.0])
    S2 = O.ovl2c(A, B, None, None, ls2, coef2, zeta2)
    assert np.allclose(S, S2, atol=1e-15)
    dS = np.ones_like(S)
    dc, da = O.ovl2c_pb(dS, A, B, None, None, ls, coef, zeta)
    assert dc[0, 1, 0] != 0.0 and da[0, 1, 0] == 0.0


# ---------------------------------------------------------------- overlap derivatives (FD)
def _small_orbital_overlap_case(seed, NZ=2, K=3, B=5):
    rng = np.random.default_rng(seed)
    ls = [0, 0, 1, 2]
    zeta = np.exp(rng.uniform(-1, 2, size=(len(ls), K, NZ)))
    D = rng.standard_normal((len(ls), K, NZ))
    if K > 2:
        D[3, 2, :] = 0.0   # a padded slot
        zeta[3, 2, :] = 1.0
    posA = rng.normal(0, 0.5, size=(B, 3))
    posB = posA + rng.normal(0, 1.0, size=(B, 3))
    sA = rng.integers(1, NZ + 1, B)
    sB = rng.integers(1, NZ + 1, B)
    return ls, zeta, D, posA, posB, sA, sB


def _lux_overlap(ls, zeta, D, posA, posB, sA, sB):
    return O.ovl2c(posA, posB, sA, sB, ls, O.compile_coef(zeta, D, ls), zeta)


def test_overlap_param_pullback_vs_fd():
    ls, zeta, D, posA, posB, sA, sB = _small_orbital_overlap_case(7)
    S = _lux_overlap(ls, zeta, D, posA, posB, sA, sB)
    rng = np.random.default_rng(8)
    dS = rng.standard_normal(S.shape)
    dz, dD = O.pullback_ps_2c(dS, posA, posB, sA, sB, ls, zeta, D)
    V, W = rng.standard_normal(zeta.shape), rng.standard_normal(D.shape)
    h = 1e-6
    f = lambda z, d: np.sum(dS * _lux_overlap(ls, z, d, posA, posB, sA, sB))
    fd_z = (f(zeta + h * V, D) - f(zeta - h * V, D)) / (2 * h)
    fd_D = (f(zeta, D + h * W) - f(zeta, D - h * W)) / (2 * h)
    assert abs(np.sum(dz * V) - fd_z) <= 1e-6 * max(1.0, abs(fd_z))
    assert abs(np.sum(dD * W) - fd_D) <= 1e-6 * max(1.0, abs(fd_D))


def test_position_cotangent_vs_fd():
    """orc_ovl2c_dx: central differences of <dS, S> in every coordinate of A and of B,
    shells up to l = 4 (the alias map included), two species."""
    rng = np.random.default_rng(31)
    ls = [0, 1, 2, 3, 4, 2]
    K, NZ, B = 3, 2, 4
    zeta = np.exp(rng.uniform(-1, 1.5, size=(len(ls), K, NZ)))
    coef = rng.standard_normal((len(ls), K, NZ))
    coef[2, 2, 0] = 0.0  # a zero coefficient (skipped in the forward)
    posA = rng.normal(0, 0.5, size=(B, 3))
    posB = posA + rng.normal(0, 0.9, size=(B, 3))
    sA, sB = rng.integers(1, NZ + 1, B), rng.integers(1, NZ + 1, B)
    S = O.ovl2c(posA, posB, sA, sB, ls, coef, zeta)
    dS = rng.standard_normal(S.shape)
    dA, dB = O.ovl2c_dx(dS, posA, posB, sA, sB, ls, coef, zeta)
    h = 1e-6
    for b in range(B):
        for c in range(3):
            for which, d in ((0, dA), (1, dB)):
                P = [posA.copy(), posB.copy()]
                P[which][b, c] += h
                fp = np.sum(dS * O.ovl2c(P[0], P[1], sA, sB, ls, coef, zeta))
                P[which][b, c] -= 2 * h
                fm = np.sum(dS * O.ovl2c(P[0], P[1], sA, sB, ls, coef, zeta))
                fd = (fp - fm) / (2 * h)
                assert abs(d[b, c] - fd) <= 1e-7 * max(1.0, np.abs(d).max()), (b, c, which, d[b, c], fd)
    assert np.array_equal(dB, -dA)


def test_position_cotangent_ss_closed_form():
    """s-s primitives: S = c_a c_b (pi/p)^(3/2) exp(-mu R^2), dS/dA = -2 mu (A - B) S."""
    ls, coef, zeta = _two_shell(0, 0, [1.3], [0.7], [0.45], [1.1])
    A, B = np.array([[0.2, -0.4, 0.9]]), np.array([[-0.5, 0.3, 0.1]])
    S = O.ovl2c(A, B, None, None, ls, coef, zeta)
    g = np.zeros_like(S)
    g[0, 1, 0] = 1.0  # the (a, b) entry only
    dA, _ = O.ovl2c_dx(g, A, B, None, None, ls, coef, zeta)
    mu = 1.3 * 0.45 / 1.75
    np.testing.assert_allclose(dA[0], -2 * mu * (A[0] - B[0]) * S[0, 1, 0], rtol=1e-13)


def _three_centre_case(seed, ls=(0, 1, 2, 1), K=2, NZ=2, B=3):
    rng = np.random.default_rng(seed)
    ls = list(ls)
    zeta = np.exp(rng.uniform(-1, 1.2, size=(len(ls), K, NZ)))
    D = rng.standard_normal((len(ls), K, NZ))
    D[3, 1, :] = 0.0  # a padded slot (zeta 1)
    zeta[3, 1, :] = 1.0
    P = [rng.normal(0, 0.6, size=(B, 3)) for _ in range(3)]
    S = [rng.integers(1, NZ + 1, B) for _ in range(3)]
    return ls, zeta, D, P, S


def test_3c_param_pullback_vs_fd():
    """orc_ovl3c_pb (coef, alpha) and its (zeta, D) chain vs central differences of
    <dT, T>, padded slot included, shells up to l = 2, two species per centre."""
    ls, zeta, D, P, S = _three_centre_case(3)
    coef = O.compile_coef(zeta, D, ls)
    T = O.ovl3c(*P, *S, ls, coef, zeta)
    rng = np.random.default_rng(4)
    dT = rng.standard_normal(T.shape)
    dc, da = O.ovl3c_pb(dT, *P, *S, ls, coef, zeta)
    h = 1e-6
    f = lambda c, a: np.sum(dT * O.ovl3c(*P, *S, ls, c, a))
    V = rng.standard_normal(coef.shape)
    fd = (f(coef + h * V, zeta) - f(coef - h * V, zeta)) / (2 * h)
    assert abs(np.sum(dc * V) - fd) <= 1e-7 * max(1.0, abs(fd))
    W = rng.standard_normal(zeta.shape) * 0.1
    fd = (f(coef, zeta + h * W) - f(coef, zeta - h * W)) / (2 * h)
    assert abs(np.sum(da * W) - fd) <= 1e-6 * max(1.0, abs(fd))
    # the padded slot (coef 0) has a coefficient cotangent and no exponent cotangent
    assert dc[3, 1, 0] != 0.0 and da[3, 1, 0] == 0.0
    # (zeta, D) through the normalisation
    dz, dD = O.pullback_ps_3c(dT, *P, *S, ls, zeta, D)
    g = lambda z, d: np.sum(dT * O.ovl3c(*P, *S, ls, O.compile_coef(z, d, ls), z))
    fd = (g(zeta + h * W, D) - g(zeta - h * W, D)) / (2 * h)
    assert abs(np.sum(dz * W) - fd) <= 1e-6 * max(1.0, abs(fd))
    fd = (g(zeta, D + h * V) - g(zeta, D - h * V)) / (2 * h)
    assert abs(np.sum(dD * V) - fd) <= 1e-6 * max(1.0, abs(fd))


def test_3c_pullback_reduces_to_2c():
    """A third centre with an s shell of exponent ~0 (1e-14: exactly 0 would make p = 0 in
    the s3 x s3 pairs, whose cotangent is zero anyway) and coefficient 1 leaves the
    2-centre overlap (q -> p, Q -> P): the 3c (coef, alpha) cotangents of the A and B shells
    equal the 2-centre pullback's (orc_ovl2c_pb, an independent code path)."""
    rng = np.random.default_rng(12)
    ls2 = [0, 1, 2]
    K, B = 2, 3
    zeta2 = np.exp(rng.uniform(-0.5, 1.0, size=(3, K, 1)))
    coef2 = rng.standard_normal((3, K, 1))
    A, Bp = rng.normal(0, 0.5, (B, 3)), rng.normal(0, 0.5, (B, 3))
    S = O.ovl2c(A, Bp, None, None, ls2, coef2, zeta2)
    dS = rng.standard_normal(S.shape)
    dc2, da2 = O.ovl2c_pb(dS, A, Bp, None, None, ls2, coef2, zeta2)
    ls3 = ls2 + [0]
    zeta3 = np.concatenate([zeta2, np.array([[[1e-14], [1.0]]])], axis=0)
    coef3 = np.concatenate([coef2, np.array([[[1.0], [0.0]]])], axis=0)
    N2, N3 = S.shape[0], S.shape[0] + 1
    dT = np.zeros((N3, N3, N3, B))
    dT[:N2, :N2, N2, :] = dS
    C = rng.normal(0, 0.5, (B, 3))
    dc3, da3 = O.ovl3c_pb(dT, A, Bp, C, None, None, None, ls3, coef3, zeta3)
    # A-side and B-side contributions both land on the shared (ns, K) slots: compare sums
    np.testing.assert_allclose(dc3[:3], dc2, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(da3[:3], da2, rtol=1e-10, atol=1e-12)


def test_normalisation_vjp_vs_fd():
    rng = np.random.default_rng(2)
    ls = [0, 1, 2]
    zeta = np.exp(rng.uniform(-1, 2, size=(3, 4, 2)))
    D = rng.standard_normal((3, 4, 2))
    dc = rng.standard_normal((3, 4, 2))
    dD, dz = O.compile_coef_pb(dc, zeta, D, ls)
    V, W = rng.standard_normal(zeta.shape), rng.standard_normal(D.shape)
    h = 1e-6
    f = lambda z, d: np.sum(dc * O.compile_coef(z, d, ls))
    assert abs(np.sum(dD * W) - (f(zeta, D + h * W) - f(zeta, D - h * W)) / (2 * h)) < 1e-6
    assert abs(np.sum(dz * V) - (f(zeta + h * V, D) - f(zeta - h * V, D)) / (2 * h)) < 1e-6


def test_single_primitive_has_no_D_gradient():
    """K = 1: coef = D / sqrt(normsq) is scale invariant, so dS/dD = 0 (AGENTS.md:111-114)."""
    ls, zeta, D, posA, posB, sA, sB = _small_orbital_overlap_case(9, K=1)
    D[3, 0, :] = 1.0
    zeta[3, 0, :] = 1.3
    dS = np.random.default_rng(1).standard_normal((1 + 1 + 3 + 6,) * 2 + (len(sA),))
    _, dD = O.pullback_ps_2c(dS, posA, posB, sA, sB, ls, zeta, D)
    assert np.max(np.abs(dD)) < 1e-12


def test_absent_species_zero_gradient():
    ls, zeta, D, posA, posB, sA, sB = _small_orbital_overlap_case(4, NZ=3, B=6)
    sA[:] = np.where(sA == 3, 1, sA)
    sB[:] = np.where(sB == 3, 2, sB)
    S = _lux_overlap(ls, zeta, D, posA, posB, sA, sB)
    dz, dD = O.pullback_ps_2c(np.ones_like(S), posA, posB, sA, sB, ls, zeta, D)
    assert np.all(dz[:, :, 2] == 0.0) and np.all(dD[:, :, 2] == 0.0)


def test_oracle_rejects_unknown_species():
    ls, coef, zeta = _two_shell(0, 0, [1.0], [1.0], [1.0], [1.0])
    with pytest.raises(ValueError):
        O.ovl2c(np.zeros((1, 3)), np.zeros((1, 3)), [2], [1], ls, coef, zeta)


# ---------------------------------------------------------------- harmonics
def test_ylm_closed_forms_l2():
    g = GOLD["ylm_l2"]
    for p, vals in zip(g["points"], g["values"]):
        Z, _ = O.ylm(2, p)
        assert np.max(np.abs(Z - np.array(vals))) < 1e-14


def test_ylm_l3_l4_published_forms():
    """m-order and sign of every harmonic up to l = 4 against the textbook definition
    (associated Legendre without the Condon-Shortley phase, sin for m < 0, cos for m > 0),
    itself checked against the published l = 3, 4 polynomial tables when the fixture is made.
    A permuted m, a flipped sign or a cos/sin swap at any (l, m) fails here."""
    g = GOLD["ylm_l4"]
    for p, vals in zip(g["points"], g["values"]):
        Z, _ = O.ylm(4, p)
        ref = np.array(vals)
        assert Z.shape == ref.shape == (25,)
        assert np.max(np.abs(Z - ref)) < 1e-13 * max(1.0, np.max(np.abs(ref)))


def test_ylm_orthonormal_on_sphere():
    """L2 normalisation and orthogonality for l <= 4 by product quadrature."""
    L = 4
    xg, wg = np.polynomial.legendre.leggauss(24)
    nphi = 48
    phis = 2 * np.pi * np.arange(nphi) / nphi
    G = np.zeros(((L + 1) ** 2, (L + 1) ** 2))
    for ct, w in zip(xg, wg):
        st = np.sqrt(1 - ct * ct)
        for ph in phis:
            Z, _ = O.ylm(L, (st * np.cos(ph), st * np.sin(ph), ct))
            G += w * (2 * np.pi / nphi) * np.outer(Z, Z)
    assert np.max(np.abs(G - np.eye(G.shape[0]))) < 1e-12


def test_ylm_addition_theorem_and_homogeneity():
    rng = np.random.default_rng(1)
    x = rng.normal(size=3)
    r2 = x @ x
    Z, _ = O.ylm(4, x)
    Z2, _ = O.ylm(4, 2.0 * x)
    for l in range(5):
        sl = slice(l * l, (l + 1) ** 2)
        assert abs(np.sum(Z[sl] ** 2) - (2 * l + 1) / (4 * np.pi) * r2 ** l) < 1e-12 * max(1, r2 ** l)
        assert np.allclose(Z2[sl], 2.0 ** l * Z[sl], rtol=1e-13, atol=1e-15)


def test_ylm_gradient_fd():
    rng = np.random.default_rng(2)
    x = rng.normal(size=3)
    _, dZ = O.ylm(4, x)
    h = 1e-6
    for c in range(3):
        e = np.zeros(3)
        e[c] = h
        fd = (O.ylm(4, x + e)[0] - O.ylm(4, x - e)[0]) / (2 * h)
        assert np.max(np.abs(fd - dZ[:, c])) < 1e-7


# ---------------------------------------------------------------- atomic orbitals
def _ao_case(seed, NZ=2, K=3, nX=7, L=3, scale=1.0, decay=0):
    rng = np.random.default_rng(seed)
    ls = [0, 0, 1, 2, 3][: 2 + L]
    nRad = len(ls)
    zeta = np.exp(rng.uniform(-1, 1.5, size=(nRad, K, NZ)))
    D = rng.standard_normal((nRad, K, NZ))
    iR, iY = [], []
    for k, l in enumerate(ls):
        for m in range(-l, l + 1):
            iR.append(k + 1)
            iY.append(l * l + l + m + 1)
    X = rng.normal(0, 1.2, size=(nX, 3))
    sidx = rng.integers(1, NZ + 1, nX)
    poly = np.zeros(nRad, np.int64) if decay == 0 else np.array([l for l in ls], np.int64)
    return dict(X=X, sidx=sidx, zeta=zeta, D=D, poly=poly, iR=np.array(iR), iY=np.array(iY), L=L,
                lengthscale=scale, decay=decay)


def test_ao_values_closed_form():
    c = _ao_case(1, L=2)
    P = O.ao_eval(**c)
    for i in range(len(c["X"])):
        x = c["X"][i]
        r2 = x @ x
        Zc = [0.28209479177387814, 0.4886025119029199 * x[1], 0.4886025119029199 * x[2],
              0.4886025119029199 * x[0], 1.0925484305920792 * x[0] * x[1], 1.0925484305920792 * x[1] * x[2],
              0.31539156525252005 * (3 * x[2] ** 2 - r2), 1.0925484305920792 * x[0] * x[2],
              0.5462742152960396 * (x[0] ** 2 - x[1] ** 2)]
        sg = c["sidx"][i] - 1
        for o, (k, y) in enumerate(zip(c["iR"], c["iY"])):
            R = np.sum(c["D"][k - 1, :, sg] * np.exp(-c["zeta"][k - 1, :, sg] * r2))
            assert abs(P[i, o] - R * Zc[y - 1]) < 1e-13


@pytest.mark.parametrize("decay,scale", [(0, 1.0), (0, 1.8897261246257702), (1, 1.0)])
def test_ao_gradient_fd(decay, scale):
    c = _ao_case(2, decay=decay, scale=scale)
    _, dP = O.ao_eval(grad=True, **c)
    h = 1e-6
    for comp in range(3):
        Xp, Xm = c["X"].copy(), c["X"].copy()
        Xp[:, comp] += h
        Xm[:, comp] -= h
        fd = (O.ao_eval(**{**c, "X": Xp}) - O.ao_eval(**{**c, "X": Xm})) / (2 * h)
        assert np.max(np.abs(fd - dP[:, :, comp])) < 1e-6


def test_ao_length_unit_equivalence():
    """test/orbitals/test_length_unit.jl:35-58: values equal at the same physical point,
    gradients carry the lengthscale factor."""
    a2b = 1.8897261246257702
    c = _ao_case(3)
    PA, dA = O.ao_eval(grad=True, **{**c, "lengthscale": a2b})
    PB, dB = O.ao_eval(grad=True, **{**c, "X": a2b * c["X"]})
    assert np.max(np.abs(PA - PB)) < 1e-12
    assert np.max(np.abs(dA - a2b * dB)) < 1e-12


def test_ao_pullback_vs_fd():
    c = _ao_case(4)
    rng = np.random.default_rng(5)
    nX, nOrb = len(c["X"]), len(c["iR"])
    dP = rng.standard_normal((nX, nOrb))
    dX, dz, dD = O.ao_pullback(dP=dP, **c)
    h = 1e-6
    f = lambda **kw: np.sum(dP * O.ao_eval(**{**c, **kw}))
    V, W, U = rng.standard_normal(c["zeta"].shape), rng.standard_normal(c["D"].shape), rng.standard_normal((nX, 3))
    fz = (f(zeta=c["zeta"] + h * V) - f(zeta=c["zeta"] - h * V)) / (2 * h)
    fD = (f(D=c["D"] + h * W) - f(D=c["D"] - h * W)) / (2 * h)
    fX = (f(X=c["X"] + h * U) - f(X=c["X"] - h * U)) / (2 * h)
    assert abs(np.sum(dz * V) - fz) < 1e-6 * max(1, abs(fz))
    assert abs(np.sum(dD * W) - fD) < 1e-6 * max(1, abs(fD))
    assert abs(np.sum(dX * U) - fX) < 1e-6 * max(1, abs(fX))


def test_ao_absent_species_zero_gradient():
    c = _ao_case(6, NZ=3)
    c["sidx"] = np.where(c["sidx"] == 3, 1, c["sidx"])
    dP = np.ones((len(c["X"]), len(c["iR"])))
    _, dz, dD = O.ao_pullback(dP=dP, **c)
    assert np.all(dz[:, :, 2] == 0) and np.all(dD[:, :, 2] == 0)


def test_ovl2c_pb_openmp_matches_serial():
    """The OpenMP split of the 2c pullback (per-thread partials; the CPU-baseline form)
    gives the serial oracle's cotangents."""
    from aok import configs

    d = configs.c1(1, B=96)
    cb = d["basis"]
    dS = np.random.default_rng(4).standard_normal((cb.N, cb.N, 96))
    coef = O.compile_coef(cb.zeta, cb.D, cb.ls)
    a = O.ovl2c_pb(dS, d["posA"], d["posB"], d["sidxA"], d["sidxB"], cb.ls, coef, cb.zeta)
    b = O.ovl2c_pb(dS, d["posA"], d["posB"], d["sidxA"], d["sidxB"], cb.ls, coef, cb.zeta, omp=True)
    for x, y in zip(a, b):
        assert np.max(np.abs(x - y)) <= 1e-12 * np.max(np.abs(x))
