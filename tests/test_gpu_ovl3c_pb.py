"""GPU parity tests of the 3-centre parameter pullback (aok_ovl3c_pullback and the Python
rrule_batch_overlap_3c; no reference counterpart, pinned through the oracle's
orc_ovl3c_pb, itself pinned by finite differences and the 3c -> 2c reduction in
tests/test_oracle.py).

Tolerances (relative to the largest |cotangent|): FP64 1e-10; FP32 vs the FP64 oracle on
the same FP32-rounded inputs 1e-5 (FP32 terms, FP64 sums over the batch)."""
import numpy as np
import pytest
import torch

import aok
import oracle as O
from aok import configs
from helpers import dev_f, dev_rows, host, log, rel_max, rounded

pytestmark = pytest.mark.gpu


def _basis(ls, zeta, coef, dtype, dev):
    nbf = [(l + 1) * (l + 2) // 2 for l in ls]
    off = [0]
    for n in nbf:
        off.append(off[-1] + n)
    return aok.CartesianGTOBasis(max(ls), len(ls), zeta.shape[1], list(ls), nbf, off, off[-1],
                                 dev_f(zeta, dtype, dev), dev_f(coef, dtype, dev),
                                 tuple(range(1, zeta.shape[2] + 1)), 1.0)


def _inputs(d, dtype, dev):
    return [(dev_rows(d[f"pos{c}"], dtype, dev), torch.as_tensor(d[f"sidx{c}"], device=dev)) for c in "ABC"]


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_c4a_coef_alpha_pullback_vs_oracle(dev, dtype, tol):
    B = 60  # the FP64 oracle walks every primitive triple: seconds per 10 elements
    d = configs.c4a(B=B)
    cb = d["basis"]
    log(f"3c pullback C4a B={B} dtype={dtype}")
    coef = rounded(O.compile_coef(rounded(cb.zeta, dtype), rounded(cb.D, dtype), cb.ls), dtype)
    basis = _basis(cb.ls, cb.zeta, coef, dtype, dev)
    rng = np.random.default_rng(9)
    dT = rng.standard_normal((cb.N, cb.N, cb.N, B))
    dc, da = aok.overlap_3c_pb_coef_alpha(dev_f(dT, dtype, dev), basis, *_inputs(d, dtype, dev))
    rc, ra = O.ovl3c_pb(rounded(dT, dtype), *(rounded(d[f"pos{c}"], dtype) for c in "ABC"), d["sidxA"], d["sidxB"],
                        d["sidxC"], cb.ls, coef, rounded(cb.zeta, dtype))
    assert rel_max(host(dc), rc) < tol
    assert rel_max(host(da), ra) < tol
    # padded slots: a coefficient cotangent, no exponent cotangent (coef = 0)
    pad = cb.D == 0
    assert np.all(host(da)[pad] == 0.0)
    assert np.any(host(dc)[pad] != 0.0)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 1e-6)])
def test_parts_agree(dev, dtype, tol):
    """Every split of the shell loops (ovl3_parts 1 none, 2 by third-centre shell, 3 by
    (third, second) shell pair, 0 automatic) sums the same terms in another grouping:
    equal to rounding.  B = 333 leaves a ragged last batch tile."""
    B = 333
    d = configs.c4a(B=B)
    cb = d["basis"]
    coef = rounded(O.compile_coef(rounded(cb.zeta, dtype), rounded(cb.D, dtype), cb.ls), dtype)
    basis = _basis(cb.ls, cb.zeta, coef, dtype, dev)
    rng = np.random.default_rng(11)
    dT = dev_f(rng.standard_normal((cb.N, cb.N, cb.N, B)), dtype, dev)
    ctx = aok.Context(dev.index or 0)
    X = _inputs(d, dtype, dev)
    got = {}
    try:
        for parts in (1, 2, 3, 0):
            log(f"3c pullback C4a B={B} dtype={dtype} ovl3_parts={parts}")
            ctx.tune("ovl3_parts", parts)
            dc, da = aok.overlap_3c_pb_coef_alpha(dT, basis, *X, ctx=ctx)
            got[parts] = (host(dc), host(da))
    finally:
        ctx.tune("ovl3_parts", 0)
    for parts in (2, 3, 0):
        assert rel_max(got[parts][0], got[1][0]) < tol
        assert rel_max(got[parts][1], got[1][1]) < tol
    assert np.array_equal(got[0][0], got[3][0]) and np.array_equal(got[0][1], got[3][1])


def test_chunked_batch(dev):
    """B = 30000 C4a triples: the (third, second)-shell split rows (41 KB per triple) exceed
    the 1 GiB budget, so the batch runs in two chunks folded into one set of rows; equal
    to rounding to the unsplit form (ovl3_parts 1), FP32."""
    B = 30000
    d = configs.c4a(B=B)
    cb = d["basis"]
    dtype = torch.float32
    coef = rounded(O.compile_coef(rounded(cb.zeta, dtype), rounded(cb.D, dtype), cb.ls), dtype)
    basis = _basis(cb.ls, cb.zeta, coef, dtype, dev)
    g = torch.Generator(device=dev).manual_seed(12)
    dT = torch.randn((B, cb.N, cb.N, cb.N), generator=g, device=dev, dtype=dtype).permute(3, 2, 1, 0)
    ctx = aok.Context(dev.index or 0)
    X = _inputs(d, dtype, dev)
    got = {}
    try:
        for parts in (0, 1):
            log(f"3c pullback C4a B={B} FP32 ovl3_parts={parts}")
            ctx.tune("ovl3_parts", parts)
            dc, da = aok.overlap_3c_pb_coef_alpha(dT, basis, *X, ctx=ctx)
            got[parts] = (host(dc), host(da))
    finally:
        ctx.tune("ovl3_parts", 0)
    assert rel_max(got[0][0], got[1][0]) < 1e-6
    assert rel_max(got[0][1], got[1][1]) < 1e-6


@pytest.mark.parametrize("ls", [[0, 0], [0, 1, 0]])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_small_l_instances_vs_oracle(dev, ls, dtype, tol):
    """Bases without d shells take the single rows-kernel instance compiled for l <= 1 or
    l = 0 (C4a's d shells take the per-class ones); one padded slot."""
    B, K, NZ = 50, 3, 2
    ns = len(ls)
    rng = np.random.default_rng(21 + ns)
    zeta = np.exp(rng.uniform(np.log(0.2), np.log(20.0), size=(ns, K, NZ)))
    D = rng.uniform(-1.0, 1.0, size=(ns, K, NZ))
    D[0, 2, 1] = 0.0
    log(f"3c pullback ls={ls} B={B} dtype={dtype}")
    coef = rounded(O.compile_coef(rounded(zeta, dtype), rounded(D, dtype), ls), dtype)
    basis = _basis(ls, zeta, coef, dtype, dev)
    pos = rng.uniform(-1.5, 1.5, size=(3, B, 3))
    sp = rng.integers(1, NZ + 1, size=(3, B))
    d = {"posA": pos[0], "posB": pos[1], "posC": pos[2], "sidxA": sp[0], "sidxB": sp[1], "sidxC": sp[2]}
    N = sum((l + 1) * (l + 2) // 2 for l in ls)
    dT = rng.standard_normal((N, N, N, B))
    dc, da = aok.overlap_3c_pb_coef_alpha(dev_f(dT, dtype, dev), basis, *_inputs(d, dtype, dev))
    rc, ra = O.ovl3c_pb(rounded(dT, dtype), *(rounded(d[f"pos{c}"], dtype) for c in "ABC"), d["sidxA"], d["sidxB"],
                        d["sidxC"], ls, coef, rounded(zeta, dtype))
    assert rel_max(host(dc), rc) < tol
    assert rel_max(host(da), ra) < tol
    assert host(da)[0, 2, 1] == 0.0 and host(dc)[0, 2, 1] != 0.0


def test_orbital_rrule_vs_oracle(dev):
    """(zeta, D) cotangents through the FP64 normalisation VJP, FP64."""
    B = 40
    d = configs.c4a(B=B)
    cb = d["basis"]
    log(f"3c rrule (zeta, D) C4a B={B}")
    spec, _, _ = cb.ao_maps()
    R = aok.GaussianTypeRadials(dev_f(cb.zeta, torch.float64, dev), dev_f(cb.D, torch.float64, dev), spec, None,
                                tuple(range(1, cb.NZ + 1)))
    orb = aok.AtomicOrbitals(R, max(cb.ls), length_unit="bohr")
    X = _inputs(d, torch.float64, dev)
    T, pb = aok.rrule_batch_overlap_3c(orb, *X, orb.static_params())
    Tr = O.ovl3c(*(d[f"pos{c}"] for c in "ABC"), d["sidxA"], d["sidxB"], d["sidxC"], cb.ls,
                 O.compile_coef(cb.zeta, cb.D, cb.ls), cb.zeta)
    assert rel_max(host(T), Tr) < 1e-12
    rng = np.random.default_rng(10)
    dT = rng.standard_normal(Tr.shape)
    dps = pb(dev_f(dT, torch.float64, dev))[5]
    rz, rD = O.pullback_ps_3c(dT, *(d[f"pos{c}"] for c in "ABC"), d["sidxA"], d["sidxB"], d["sidxC"], cb.ls,
                              cb.zeta, cb.D)
    assert rel_max(host(dps["Rnl"]["zeta"]), rz) < 1e-10
    assert rel_max(host(dps["Rnl"]["D"]), rD) < 1e-10


def test_edges(dev):
    d = configs.c4a(B=4)
    cb = d["basis"]
    coef = O.compile_coef(cb.zeta, cb.D, cb.ls)
    basis = _basis(cb.ls, cb.zeta, coef, torch.float64, dev)
    log("3c pullback edges")
    z = (dev_rows(np.zeros((0, 3)), torch.float64, dev), None)
    dT0 = torch.zeros((cb.N, cb.N, cb.N, 0), dtype=torch.float64, device=dev)
    dc, da = aok.overlap_3c_pb_coef_alpha(dT0, basis, z, z, z)
    assert torch.count_nonzero(dc) == 0 and torch.count_nonzero(da) == 0
    # shells beyond d are not supported by the 3-centre pullback
    ls = [0, 3]
    zeta = np.ones((2, 1, 1))
    c1 = np.ones((2, 1, 1))
    b3 = _basis(ls, zeta, c1, torch.float64, dev)
    N = 1 + 10
    P = (dev_rows(np.zeros((1, 3)), torch.float64, dev), None)
    with pytest.raises(aok.AOKLmaxError):
        aok.overlap_3c_pb_coef_alpha(torch.zeros((N, N, N, 1), dtype=torch.float64, device=dev), b3, P, P, P)


def test_pullback_cotangent_beyond_4gib(dev):
    """A 4.32 GB FP32 cotangent (C4a basis, B = 320,000 triples) nonzero only on the last 40
    triples and 40 scattered ones past the 4 GiB mark (the oracle takes ~0.1 s per triple): the (coef, alpha) cotangents are those
    triples' alone, checked against the oracle on that subset (64-bit addressing of dT)."""
    dtype, B = torch.float32, 320_000
    d = configs.c4a(B=B)
    cb = d["basis"]
    N = cb.N
    first_high = -(-(1 << 32) // (N ** 3 * 4))
    rng = np.random.default_rng(13)
    idx = np.concatenate([np.sort(rng.choice(np.arange(first_high, B - 40), 40, replace=False)),
                          np.arange(B - 40, B)])
    log(f"3c pullback cotangent beyond 4 GiB B={B} first_high={first_high}")
    coef = rounded(O.compile_coef(rounded(cb.zeta, dtype), rounded(cb.D, dtype), cb.ls), dtype)
    basis = _basis(cb.ls, cb.zeta, coef, dtype, dev)
    vals = rng.standard_normal((N, N, N, len(idx)))
    dT = torch.zeros((B, N, N, N), dtype=dtype, device=dev).permute(3, 2, 1, 0)
    dT[:, :, :, torch.as_tensor(idx, device=dev)] = torch.as_tensor(vals, dtype=dtype, device=dev)
    dc, da = aok.overlap_3c_pb_coef_alpha(dT, basis, *_inputs(d, dtype, dev))
    dc, da = host(dc), host(da)
    del dT
    torch.cuda.empty_cache()
    rc, ra = O.ovl3c_pb(rounded(vals, dtype), *(rounded(d[f"pos{c}"][idx], dtype) for c in "ABC"), d["sidxA"][idx],
                        d["sidxB"][idx], d["sidxC"][idx], cb.ls, coef, rounded(cb.zeta, dtype))
    print(f"3c pullback beyond 4 GiB: dcoef {rel_max(dc, rc):.3e} dalpha {rel_max(da, ra):.3e}")
    assert rel_max(dc, rc) < 1e-5
    assert rel_max(da, ra) < 1e-5
