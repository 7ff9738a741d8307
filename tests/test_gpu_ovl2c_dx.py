"""GPU parity tests of the 2-centre position cotangents (aok_ovl2c_pullback_x and
aok_ovl2c_plan_pullback_x; no reference counterpart, pinned through the oracle's
orc_ovl2c_dx, itself pinned by finite differences and the s-s closed form in
tests/test_oracle.py).

Tolerances (relative to the largest |dA| of the batch): FP64 1e-10 (sums of a few
hundred primitive-pair terms per pair); FP32 vs the FP64 oracle on the same FP32-rounded
inputs 1e-5 (FP32 terms, FP64 per-pair sums)."""
import numpy as np
import pytest
import torch

import aok
import oracle as O
from aok import configs
from helpers import dev_f, dev_rows, host, log, rel_max, rounded

pytestmark = pytest.mark.gpu


def _orb(cb, dtype, dev, unit="bohr"):
    spec, _, _ = cb.ao_maps()
    R = aok.GaussianTypeRadials(dev_f(cb.zeta, dtype, dev), dev_f(cb.D, dtype, dev), spec, None,
                                tuple(range(1, cb.NZ + 1)))
    return aok.AtomicOrbitals(R, max(cb.ls), length_unit=unit)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_c1_position_cotangents_vs_oracle(dev, dtype, tol):
    B = 1500  # not a multiple of the 64-pair tiles
    d = configs.c1(1, B=B)
    cb = d["basis"]
    rng = np.random.default_rng(23)
    dS = rng.standard_normal((cb.N, cb.N, B))
    log(f"2c position cotangents C1 B={B} dtype={dtype}")
    orb = _orb(cb, dtype, dev)
    ps = orb.static_params()
    XA = (dev_rows(d["posA"], dtype, dev), torch.as_tensor(d["sidxA"], device=dev))
    XB = (dev_rows(d["posB"], dtype, dev), torch.as_tensor(d["sidxB"], device=dev))
    S, pb = aok.rrule_batch_overlap(orb, XA, XB, ps, None, wrt_positions=True)
    _, _, dXA, dXB, dps, _ = pb(dev_f(dS, dtype, dev))
    coef = O.compile_coef(rounded(cb.zeta, dtype), rounded(cb.D, dtype), cb.ls)
    rA, rB = O.ovl2c_dx(rounded(dS, dtype), rounded(d["posA"], dtype), rounded(d["posB"], dtype), d["sidxA"],
                        d["sidxB"], cb.ls, coef, rounded(cb.zeta, dtype))
    assert rel_max(host(dXA), rA) < tol
    assert torch.equal(dXB, -dXA)
    # the parameter cotangents are unchanged by asking for the positions as well
    _, pb2 = aok.rrule_batch_overlap(orb, XA, XB, ps, None)
    dps2 = pb2(dev_f(dS, dtype, dev))[4]
    assert torch.equal(dps["Rnl"]["zeta"], dps2["Rnl"]["zeta"]) and torch.equal(dps["Rnl"]["D"], dps2["Rnl"]["D"])
    # the one-shot coefficient form agrees (its coefficients are compiled separately, so
    # they may differ from the plan's in the last bit)
    basis = aok.compile_basis(orb)
    oA, oB = aok.overlap_pullback_x(dev_f(dS, dtype, dev), basis, XA, XB)
    assert rel_max(host(oA), host(dXA)) < (1e-13 if dtype == torch.float64 else 1e-6)
    assert torch.equal(oB, -oA)


def test_high_l_and_length_unit(dev):
    """Shells up to l = 4 (the alias map for l >= 2), two species, a zero coefficient;
    an angstrom basis scales the cotangent by ang2bohr (d/dx_input = ang2bohr d/dx_bohr)."""
    rng = np.random.default_rng(41)
    ls = [0, 1, 2, 3, 4, 2]
    K, NZ, B = 3, 2, 700
    zeta = np.exp(rng.uniform(-1, 1.5, size=(len(ls), K, NZ)))
    coef = rng.standard_normal((len(ls), K, NZ))
    coef[2, 2, 0] = 0.0
    posA = rng.normal(0, 0.5, size=(B, 3))
    posB = posA + rng.normal(0, 0.9, size=(B, 3))
    sA, sB = rng.integers(1, NZ + 1, B), rng.integers(1, NZ + 1, B)
    N = sum((l + 1) * (l + 2) // 2 for l in ls)
    dS = rng.standard_normal((N, N, B))
    rA, _ = O.ovl2c_dx(dS, posA, posB, sA, sB, ls, coef, zeta)
    nbf = [(l + 1) * (l + 2) // 2 for l in ls]
    off = list(np.concatenate([[0], np.cumsum(nbf)]))
    for unit, scale in (("bohr", 1.0), ("angstrom", aok.ANG2BOHR)):
        log(f"2c position cotangents l<=4 unit={unit}")
        basis = aok.CartesianGTOBasis(max(ls), len(ls), K, list(ls), nbf, off, N,
                                      dev_f(zeta, torch.float64, dev), dev_f(coef, torch.float64, dev),
                                      tuple(range(1, NZ + 1)), aok.length_factor(unit))
        XA = (dev_rows(posA / scale, torch.float64, dev), torch.as_tensor(sA, device=dev))
        XB = (dev_rows(posB / scale, torch.float64, dev), torch.as_tensor(sB, device=dev))
        dA, dB = aok.overlap_pullback_x(dev_f(dS, torch.float64, dev), basis, XA, XB)
        assert rel_max(host(dA), rA * scale) < 1e-10, unit
        assert torch.equal(dB, -dA)


def test_empty_batch(dev):
    d = configs.c1(1, B=4)
    cb = d["basis"]
    orb = _orb(cb, torch.float64, dev)
    log("2c position cotangents B=0")
    z = dev_rows(np.zeros((0, 3)), torch.float64, dev)
    S, pb = aok.rrule_batch_overlap(orb, (z, None), (z, None), orb.static_params(), None, wrt_positions=True)
    _, _, dXA, dXB, _, _ = pb(torch.zeros_like(S))
    assert dXA.shape == (0, 3) and dXB.shape == (0, 3)
