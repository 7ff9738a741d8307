"""GPU parity of H1 at L = 3 and L = 4 through the C ABI against the FP64 oracle: the
4s3p2d1f basis of the HBM-bound AO workload (3 species, 30 orbitals, L = 3), a g-shell
basis (L = 4, ragged contraction depths) and the reference's own device test case, the
default ``gaussian_orbitals`` (L = 3; test/orbitals/test_gpu.jl:13-97).

Tolerances (relative to the largest |value| of each output, same FP32-rounded inputs fed to
the oracle): FP64 1e-12 values and gradients, 1e-11 pullbacks; FP32 2e-6 values, 1e-5
gradients, 2e-5 pullbacks.  The reference's own device tests allow 1e-3 / 1e-2 in a norm."""
import numpy as np
import pytest
import torch

import aok
import oracle as O
from aok import configs
from helpers import dev_f, dev_rows, host, log, rel_max, rounded

pytestmark = pytest.mark.gpu

BASES = {"c4b_L3": lambda: configs.c4b_basis(4), "g_L4": lambda: configs.g_basis(5)}
TOL = {torch.float64: (1e-12, 1e-12, 1e-11), torch.float32: (2e-6, 1e-5, 2e-5)}


def _orb(cb, dtype, dev):
    spec, _, _ = cb.ao_maps()
    R = aok.GaussianTypeRadials(dev_f(cb.zeta, dtype, dev), dev_f(cb.D, dtype, dev), spec, None,
                                tuple(range(1, cb.NZ + 1)))
    return aok.AtomicOrbitals(R, max(cb.ls), length_unit="bohr")


def _case(cb, nX, seed, sd=3.0):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, sd, size=(nX, 3))
    sp = rng.integers(1, cb.NZ + 1, nX)
    dP = rng.standard_normal((nX, cb.nOrb))
    return X, sp, dP


def _oracle(cb, X, sp, dtype, grad=False):
    _, iR, iY = cb.ao_maps()
    return O.ao_eval(rounded(X, dtype), sp, rounded(cb.zeta, dtype), rounded(cb.D, dtype),
                     np.zeros(cb.ns, np.int64), np.array(iR), np.array(iY), max(cb.ls), grad=grad)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("name", list(BASES))
def test_values_and_gradients(dev, name, dtype):
    cb = BASES[name]()
    nX = 20000 + 37
    X, sp, _ = _case(cb, nX, 41)
    log(f"ao high-l {name} dtype={dtype} nX={nX} L={max(cb.ls)}")
    orb = _orb(cb, dtype, dev)
    Xd = dev_rows(X, dtype, dev)
    sidx = torch.as_tensor(sp, dtype=torch.int64, device=dev)
    P = aok.evaluate(orb, Xd, sidx=sidx)
    P2, dP = aok.evaluate_ed(orb, Xd, sidx=sidx)
    Pr, dPr = _oracle(cb, X, sp, dtype, grad=True)
    tv, tg, _ = TOL[dtype]
    assert rel_max(host(P), Pr) < tv
    assert rel_max(host(P2), Pr) < tv
    for c in range(3):  # each gradient component against its own scale
        assert rel_max(host(dP[:, :, c]), dPr[:, :, c]) < tg


@pytest.mark.parametrize("ppl", [2, 4])
@pytest.mark.parametrize("name", list(BASES))
def test_values_vector_stores(dev, name, ppl):
    """The values-only kernel with 2 or 4 points per lane (the ao_ppl knob) at L = 3, 4."""
    cb = BASES[name]()
    nX = 8192 + 4 * 37  # a multiple of ppl, not of the tile
    X, sp, _ = _case(cb, nX, 43)
    log(f"ao high-l vec {name} ppl={ppl} nX={nX}")
    orb = _orb(cb, torch.float32, dev)
    ctx = aok.Context(dev.index or 0)
    ctx.tune("ao_ppl", ppl)
    P = aok.evaluate(orb, dev_rows(X, torch.float32, dev), sidx=torch.as_tensor(sp, device=dev), ctx=ctx)
    ctx.close()
    Pr = _oracle(cb, X, sp, torch.float32)
    assert rel_max(host(P), Pr) < 2e-6


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("name", list(BASES))
def test_pullback(dev, name, dtype):
    cb = BASES[name]()
    nX = 20000 + 19
    X, sp, dPin = _case(cb, nX, 47)
    log(f"ao high-l pullback {name} dtype={dtype} nX={nX}")
    orb = _orb(cb, dtype, dev)
    sidx = torch.as_tensor(sp, dtype=torch.int64, device=dev)
    _, pb = aok.rrule_evaluate(orb, dev_rows(X, dtype, dev), sidx=sidx)
    _, _, dX, dps, _ = pb(dev_f(dPin, dtype, dev))
    _, iR, iY = cb.ao_maps()
    dXr, dzr, dDr = O.ao_pullback(rounded(X, dtype), sp, rounded(cb.zeta, dtype), rounded(cb.D, dtype),
                                  np.zeros(cb.ns, np.int64), np.array(iR), np.array(iY), max(cb.ls),
                                  rounded(dPin, dtype))
    tp = TOL[dtype][2]
    assert rel_max(host(dX), dXr) < tp
    assert rel_max(host(dps["Rnl"]["zeta"]), dzr) < tp
    assert rel_max(host(dps["Rnl"]["D"]), dDr) < tp


def _ref_norm_checks(P, Pc, dP, dPc):
    """The reference's own device criteria (test/orbitals/test_gpu.jl:31,35-37)."""
    assert np.linalg.norm(P - Pc) / np.linalg.norm(Pc) < 1e-3
    num = np.max(np.linalg.norm(dP - dPc, axis=2))
    den = np.max(np.linalg.norm(dPc, axis=2))
    assert num / den < 1e-2


def _state_arrays(orb):
    st = orb.static_state()
    return (host(orb.Rnl.zeta), host(orb.Rnl.D), np.array(st["Rnl"]["poly"]), np.array(st["iR"]),
            np.array(st["iY"]), orb.L)


def test_reference_default_basis_fp32(dev):
    """gaussian_orbitals(; length_unit = :bohr) -- N1 = 4, N2 = 3, K = 1, L = 3 -- in FP32
    on the device vs FP64 (test/orbitals/test_gpu.jl:13-58)."""
    rng = np.random.default_rng(1234)
    Xh = rng.standard_normal((64, 3))
    orb64 = aok.gaussian_orbitals(length_unit="bohr", device=dev)
    orb32 = aok.gaussian_orbitals(length_unit="bohr", dtype=torch.float32, device=dev)
    assert orb32.L == 3 and len(orb32) == 4 * 3 * 16
    log("ao reference default basis fp32")
    z, D, poly, iR, iY, L = _state_arrays(orb64)
    X32 = rounded(Xh, torch.float32)
    Pc, dPc = O.ao_eval(X32, None, z, D, poly, iR, iY, L, grad=True)
    Xg = dev_rows(Xh, torch.float32, dev)
    P = host(aok.evaluate(orb32, Xg))
    _, dP = aok.evaluate_ed(orb32, Xg)
    dP = host(dP)
    _ref_norm_checks(P, Pc, dP, dPc)
    assert rel_max(P, Pc) < 2e-6 and rel_max(dP, dPc) < 1e-5
    dPin = rng.standard_normal((64, len(orb32)))
    _, pb = aok.rrule_evaluate(orb32, Xg)
    _, _, dX, dps, _ = pb(dev_f(dPin, torch.float32, dev))
    dXc, dzc, dDc = O.ao_pullback(X32, None, z, D, poly, iR, iY, L, rounded(dPin, torch.float32))
    for got, ref in ((host(dps["Rnl"]["zeta"]), dzc), (host(dps["Rnl"]["D"]), dDc)):
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-2
        assert rel_max(got, ref) < 2e-5
    dX = host(dX)
    assert np.max(np.linalg.norm(dX - dXc, axis=1)) / np.max(np.linalg.norm(dXc, axis=1)) < 1e-2
    assert rel_max(dX, dXc) < 2e-5


def test_reference_multi_species_fp32(dev):
    """gaussian_orbitals(4, 3; nspecies = 2, zlist = (1, 6)) with species indices
    (test/orbitals/test_gpu.jl:60-97)."""
    rng = np.random.default_rng(99)
    nX = 64
    Xh = rng.standard_normal((nX, 3))
    sp = rng.integers(1, 3, nX)
    orb64 = aok.gaussian_orbitals(4, 3, length_unit="bohr", nspecies=2, zlist=(1, 6), device=dev)
    orb32 = aok.gaussian_orbitals(4, 3, length_unit="bohr", nspecies=2, zlist=(1, 6), dtype=torch.float32,
                                  device=dev)
    log("ao reference multi-species fp32")
    z, D, poly, iR, iY, L = _state_arrays(orb64)
    X32 = rounded(Xh, torch.float32)
    Pc, dPc = O.ao_eval(X32, sp, z, D, poly, iR, iY, L, grad=True)
    Xg = dev_rows(Xh, torch.float32, dev)
    sg = torch.as_tensor(sp, device=dev)
    P = host(aok.evaluate(orb32, Xg, sidx=sg))
    _, dP = aok.evaluate_ed(orb32, Xg, sidx=sg)
    _ref_norm_checks(P, Pc, host(dP), dPc)
    assert rel_max(P, Pc) < 2e-6
    dPin = rng.standard_normal((nX, len(orb32)))
    dps = aok.pullback_ps(dev_f(dPin, torch.float32, dev), orb32, Xg, None, None, sidx=sg)
    _, dzc, dDc = O.ao_pullback(X32, sp, z, D, poly, iR, iY, L, rounded(dPin, torch.float32))
    for got, ref in ((host(dps["Rnl"]["zeta"]), dzc), (host(dps["Rnl"]["D"]), dDc)):
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-2
        assert rel_max(got, ref) < 2e-5


def test_outputs_beyond_4gib(dev):
    """P of 4.32 GB and dP of 13 GB (the 4s3p2d1f basis, 30 orbitals, nX = 36M, FP32): values and
    gradients of the first, last and 300 scattered points against the oracle, and the rrule
    pullback of a 4.32 GB cotangent that is nonzero only at those points (64-bit column
    addressing of P, dP and the cotangent)."""
    cb = configs.c4b_basis(4)
    nX = 36_000_000
    assert nX * cb.nOrb * 4 >= 1 << 32
    rng = np.random.default_rng(61)
    X = rng.normal(0.0, 3.0, size=(nX, 3)).astype(np.float32)
    sp = rng.integers(1, cb.NZ + 1, nX).astype(np.int32)
    idx = np.concatenate([np.arange(300), np.arange(nX - 300, nX),
                          np.sort(rng.choice(np.arange(300, nX - 300), 300, replace=False))])
    log(f"ao beyond 4 GiB nX={nX}")
    orb = _orb(cb, torch.float32, dev)
    Xd = torch.as_tensor(X, device=dev)
    sidx = torch.as_tensor(sp, device=dev)
    ti = torch.as_tensor(idx, device=dev)
    P = aok.evaluate(orb, Xd, sidx=sidx)
    Pg = host(P[ti])
    del P
    P2, dP = aok.evaluate_ed(orb, Xd, sidx=sidx)
    P2g, dPg = host(P2[ti]), host(dP[ti])
    del P2, dP
    dPin = torch.zeros((nX, cb.nOrb), dtype=torch.float32, device=dev)
    vals = rng.standard_normal((len(idx), cb.nOrb))
    dPin[ti] = torch.as_tensor(vals, dtype=torch.float32, device=dev)
    _, pb = aok.rrule_evaluate(orb, Xd, sidx=sidx)
    _, _, dX, dps, _ = pb(dPin)
    dXg, dzg, dDg = host(dX[ti]), host(dps["Rnl"]["zeta"]), host(dps["Rnl"]["D"])
    del dPin, dX, Xd
    torch.cuda.empty_cache()
    Pr, dPr = _oracle(cb, X[idx].astype(np.float64), sp[idx], torch.float32, grad=True)
    _, iR, iY = cb.ao_maps()
    dXr, dzr, dDr = O.ao_pullback(X[idx].astype(np.float64), sp[idx], rounded(cb.zeta, torch.float32),
                                  rounded(cb.D, torch.float32), np.zeros(cb.ns, np.int64), np.array(iR),
                                  np.array(iY), max(cb.ls), rounded(vals, torch.float32))
    tv, tg, tp = TOL[torch.float32]
    assert rel_max(Pg, Pr) < tv and rel_max(P2g, Pr) < tv
    for c in range(3):
        assert rel_max(dPg[:, :, c], dPr[:, :, c]) < tg
    assert rel_max(dXg, dXr) < tp
    assert rel_max(dzg, dzr) < tp and rel_max(dDg, dDr) < tp
