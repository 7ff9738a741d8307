"""CPU checks of the algebra behind the 2-centre pullback's exponent tables.

The pullback kernel (csrc/ovl2c_kernels.hip, pair_contract / os1d_exp) never forms the
raised 1-D overlap s(i+2, j) of d/dalpha x^i e^{-alpha x^2} = -x^{i+2} e^{-alpha x^2}.  It forms
    u(i, j) = s(i+2, j) - (2i+1)/(4 alpha) s(i, j)
by its own recursion
    u(i, j) = PA s(i+1, j) + (alpha - (2i+1) beta)/(4 alpha p) s(i, j) + j/(2p) s(i+1, j-1),
with s(LA+1, j) and s(i, LB+1) extended from the base table by the vertical and horizontal
Obara-Saika steps, and the ket side likewise.  These tests restate those recursions in
numpy and check them against the definition (raised tables minus the norm term) in FP64,
the s-s closed form, the sum over the three axes (n_a = (2l + 3)/(4 alpha)), and that in
FP32 the recursion loses less than the difference form where the two terms nearly cancel.
"""
import numpy as np
import pytest


def os1d(PA, PB, i2p, IA, IB, dt=np.float64):
    """1-D Obara-Saika overlaps s[i][j] (s[0][0] = 1): vertical then horizontal steps."""
    s = np.zeros((IA + 1, IB + 1), dtype=dt)
    s[0, 0] = 1
    for i in range(1, IA + 1):
        s[i, 0] = PA * s[i - 1, 0] + (dt(i - 1) * i2p * s[i - 2, 0] if i > 1 else 0)
    for j in range(1, IB + 1):
        for i in range(IA + 1):
            v = PB * s[i, j - 1]
            if i > 0:
                v += dt(i) * i2p * s[i - 1, j - 1]
            if j > 1:
                v += dt(j - 1) * i2p * s[i, j - 2]
            s[i, j] = v
    return s


def exp_tables(PA, PB, i2p, ca, sa, cb, sb, s, dt=np.float64):
    """The kernel's u (bra) and v (ket) tables from the base table s (LA+1 x LB+1)."""
    LA, LB = s.shape[0] - 1, s.shape[1] - 1
    c = [ca - dt(i) * sa for i in range(LA + 1)]
    d = [cb - dt(j) * sb for j in range(LB + 1)]
    t = np.zeros(LB + 1, dtype=dt)  # s(LA+1, j)
    t[0] = PA * s[LA, 0] + (dt(LA) * i2p * s[LA - 1, 0] if LA > 0 else 0)
    for j in range(1, LB + 1):
        t[j] = PB * t[j - 1] + dt(LA + 1) * i2p * s[LA, j - 1] + (dt(j - 1) * i2p * t[j - 2] if j > 1 else 0)
    u = np.zeros_like(s)
    for i in range(LA + 1):
        for j in range(LB + 1):
            up = s[i + 1, j] if i < LA else t[j]
            x = PA * up + c[i] * s[i, j]
            if j > 0:
                x += dt(j) * i2p * (s[i + 1, j - 1] if i < LA else t[j - 1])
            u[i, j] = x
    w = np.zeros(LA + 1, dtype=dt)  # s(i, LB+1)
    for i in range(LA + 1):
        w[i] = PB * s[i, LB] + (dt(i) * i2p * s[i - 1, LB] if i > 0 else 0) + (dt(LB) * i2p * s[i, LB - 1] if LB > 0 else 0)
    v = np.zeros_like(s)
    for i in range(LA + 1):
        for j in range(LB + 1):
            up = s[i, j + 1] if j < LB else w[i]
            x = PB * up + d[j] * s[i, j]
            if i > 0:
                x += dt(i) * i2p * (s[i - 1, j + 1] if j < LB else w[i - 1])
            v[i, j] = x
    return u, v


def record(a, b, dt=np.float64):
    """The pair constants the kernel reads (ovl_fill_kernel: make_record / make_pb_record);
    a side with a non-positive exponent keeps the plain raised table (c0 = 1/(2p), s = -1/(2p))."""
    p = a + b
    ca = 0.25 * (a - b) / (a * p) if a > 0 else 0.5 / p
    cb = -0.25 * (a - b) / (b * p) if b > 0 else 0.5 / p
    sa = 0.5 * b / (a * p) if a > 0 else -0.5 / p
    sb = 0.5 * a / (b * p) if b > 0 else -0.5 / p
    return dict(aop=dt(a / p), bop=dt(b / p), i2p=dt(0.5 / p), ca=dt(ca), cb=dt(cb), sa=dt(sa), sb=dt(sb))


@pytest.mark.parametrize("LA,LB", [(la, lb) for la in range(5) for lb in range(5)])
def test_tables_match_raised_minus_norm_term(LA, LB):
    rng = np.random.default_rng(10 * LA + LB)
    for _ in range(20):
        a, b = np.exp(rng.uniform(np.log(0.1), np.log(7000.0), 2))
        ab = rng.uniform(-4, 4)
        r = record(a, b)
        PA, PB = -r["bop"] * ab, r["aop"] * ab
        big = os1d(PA, PB, r["i2p"], LA + 2, LB + 2)
        s = big[:LA + 1, :LB + 1]
        u, v = exp_tables(PA, PB, r["i2p"], r["ca"], r["sa"], r["cb"], r["sb"], s)
        for i in range(LA + 1):
            for j in range(LB + 1):
                ru = big[i + 2, j] - (2 * i + 1) / (4 * a) * big[i, j]
                rv = big[i, j + 2] - (2 * j + 1) / (4 * b) * big[i, j]
                scale_u = abs(big[i + 2, j]) + abs((2 * i + 1) / (4 * a) * big[i, j]) + 1e-300
                scale_v = abs(big[i, j + 2]) + abs((2 * j + 1) / (4 * b) * big[i, j]) + 1e-300
                assert abs(u[i, j] - ru) <= 1e-12 * scale_u, (i, j, u[i, j], ru)
                assert abs(v[i, j] - rv) <= 1e-12 * scale_v, (i, j, v[i, j], rv)


def test_ss_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(50):
        a, b = np.exp(rng.uniform(np.log(0.1), np.log(7000.0), 2))
        AB = rng.uniform(-4, 4, 3)
        r = record(a, b)
        tot = 0.0
        for ax in range(3):
            s = os1d(-r["bop"] * AB[ax], r["aop"] * AB[ax], r["i2p"], 0, 0)
            u, _ = exp_tables(-r["bop"] * AB[ax], r["aop"] * AB[ax], r["i2p"], r["ca"], r["sa"], r["cb"], r["sb"], s)
            tot += u[0, 0]
        closed = r["bop"] ** 2 * AB @ AB + 3 * r["ca"]
        assert abs(tot - closed) <= 1e-13 * (abs(closed) + r["bop"] ** 2 * AB @ AB + 3 * abs(r["ca"]))


def test_axis_sum_removes_the_primitive_norm_term():
    """Per Cartesian component (ix, iy, iz), l = ix + iy + iz: the exponent sum over the axes
    equals the raised sum minus (2l + 3)/(4 alpha) times the plain product."""
    rng = np.random.default_rng(4)
    la, lb = 2, 1
    a, b = 0.3, 2.5
    AB = rng.uniform(-2, 2, 3)
    r = record(a, b)
    S, U, BIG = [], [], []
    for ax in range(3):
        PA, PB = -r["bop"] * AB[ax], r["aop"] * AB[ax]
        big = os1d(PA, PB, r["i2p"], la + 2, lb)
        s = os1d(PA, PB, r["i2p"], la, lb)
        u, _ = exp_tables(PA, PB, r["i2p"], r["ca"], r["sa"], r["cb"], r["sb"], s)
        S.append(s), U.append(u), BIG.append(big)
    for ix in range(la + 1):
        for iy in range(la + 1 - ix):
            iz = la - ix - iy
            for jx in range(lb + 1):
                for jy in range(lb + 1 - jx):
                    jz = lb - jx - jy
                    x, y, z = S[0][ix, jx], S[1][iy, jy], S[2][iz, jz]
                    e = U[0][ix, jx] * y * z + x * U[1][iy, jy] * z + x * y * U[2][iz, jz]
                    raised = BIG[0][ix + 2, jx] * y * z + x * BIG[1][iy + 2, jy] * z + x * y * BIG[2][iz + 2, jz]
                    n = (2 * la + 3) / (4 * a)
                    assert abs(e - (raised - n * x * y * z)) <= 1e-12 * (abs(raised) + abs(n * x * y * z))


def test_fp32_recursion_keeps_the_cancelling_digits():
    """Equal diffuse exponents: s(2,0) = PA^2 + 1/(2p) and the norm term 1/(4 alpha) = 1/(2p)
    cancel exactly, leaving PA^2.  In FP32 the recursion gives PA^2 to rounding of PA^2; the
    difference form loses the digits of 1/(2p)."""
    a = b = np.float32(0.13)
    AB = np.float32(0.05)
    r = record(float(a), float(b), np.float32)
    PA, PB = np.float32(-r["bop"] * AB), np.float32(r["aop"] * AB)
    exact = float(PA) ** 2
    s = os1d(PA, PB, r["i2p"], 0, 0, np.float32)
    u, _ = exp_tables(PA, PB, r["i2p"], r["ca"], r["sa"], r["cb"], r["sb"], s, np.float32)
    big = os1d(PA, PB, r["i2p"], 2, 0, np.float32)
    diff = np.float32(big[2, 0] - np.float32(1 / (4 * a)) * big[0, 0])
    err_rec = abs(float(u[0, 0]) - exact) / exact
    err_diff = abs(float(diff) - exact) / exact
    assert err_rec < 1e-6
    assert err_diff > 100 * err_rec


@pytest.mark.parametrize("LA,LB", [(0, 0), (1, 2), (2, 1), (3, 3)])
def test_zero_exponent_keeps_the_raised_table(LA, LB):
    """alpha = 0 (no L2 norm to take out): the bra table is the plain raised table s(i+2, j)."""
    b = 1.7
    AB = 0.9
    r = record(0.0, b)
    PA, PB = -r["bop"] * AB, r["aop"] * AB
    big = os1d(PA, PB, r["i2p"], LA + 2, LB + 2)
    s = big[:LA + 1, :LB + 1]
    u, v = exp_tables(PA, PB, r["i2p"], r["ca"], r["sa"], r["cb"], r["sb"], s)
    for i in range(LA + 1):
        for j in range(LB + 1):
            assert abs(u[i, j] - big[i + 2, j]) <= 1e-12 * (abs(big[i + 2, j]) + 1e-300)
            rv = big[i, j + 2] - (2 * j + 1) / (4 * b) * big[i, j]
            assert abs(v[i, j] - rv) <= 1e-12 * (abs(big[i, j + 2]) + abs((2 * j + 1) / (4 * b) * big[i, j]))
