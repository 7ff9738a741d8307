"""Differentiation at full size (the C3 workloads) on the GPU.

Overlap loss: C1 basis, B = 65536 pairs, target parameters (zeta, D) perturbed by
exp(0.05 eps); L = sum_b ||S_b - S_b^tgt||^2, so dS = 2 (S - S^tgt).  The rrule's
(zeta, D) cotangents are checked against central finite differences of L computed with
the FP64 kernels over the whole batch (h = 1e-6 in log zeta and in D), and the FP32
cotangents against the FP64 ones on the same FP32-rounded inputs.

AO pullback: nX = 2^20 points of the C1 species as atomic orbitals, a seeded cotangent;
(dX, dzeta, dD) against the FP64 oracle on the same inputs.

Tolerances are written per assertion."""
import numpy as np
import pytest
import torch

import aok
import oracle as O
from aok import configs
from helpers import dev_f, dev_rows, host, log, rel_max, rounded

pytestmark = pytest.mark.gpu


def _orb(cb, zeta, D, dtype, dev):
    spec, _, _ = cb.ao_maps()
    R = aok.GaussianTypeRadials(dev_f(zeta, dtype, dev), dev_f(D, dtype, dev), spec, None,
                                tuple(range(1, cb.NZ + 1)))
    return aok.AtomicOrbitals(R, max(cb.ls), length_unit="bohr")


def _loss_setup(dev, dtype):
    d = configs.c3_overlap()
    cb = d["basis"]
    zt, Dt = d["target"]
    XA = (dev_rows(d["posA"], dtype, dev), torch.as_tensor(d["sidxA"], device=dev))
    XB = (dev_rows(d["posB"], dtype, dev), torch.as_tensor(d["sidxB"], device=dev))
    tgt = _orb(cb, zt, Dt, dtype, dev)
    S_tgt = aok.batch_overlap(tgt, XA, XB, tgt.static_params(), None)
    return d, cb, XA, XB, S_tgt


def test_c3_overlap_loss_gradient_vs_fd(dev):
    """FP64, full batch: rrule cotangents of L vs central differences of L.
    Bound: |fd - grad| <= 1e-6 * max|grad| (FP64 roundoff of L over 14.7e6 entries
    divided by 2h is ~1e-10 relative; truncation is O(h^2))."""
    d, cb, XA, XB, S_tgt = _loss_setup(dev, torch.float64)
    log(f"C3 overlap loss FD, B={len(d['posA'])}")
    orb = _orb(cb, cb.zeta, cb.D, torch.float64, dev)
    ps = orb.static_params()
    S, pb = aok.rrule_batch_overlap(orb, XA, XB, ps, None)
    dps = pb(2.0 * (S - S_tgt))[4]
    gz, gD = host(dps["Rnl"]["zeta"]), host(dps["Rnl"]["D"])

    def loss(zeta, D):
        o = _orb(cb, zeta, D, torch.float64, dev)
        Sx = aok.batch_overlap(o, XA, XB, o.static_params(), None)
        return float(((Sx - S_tgt) ** 2).sum())

    rng = np.random.default_rng(5)
    real = np.argwhere(cb.D != 0)
    picks = real[rng.choice(len(real), size=6, replace=False)]
    h = 1e-6
    scale_z = np.abs(gz * cb.zeta).max()
    scale_D = np.abs(gD).max()
    for (s, k, g) in picks:
        zp, zm = cb.zeta.copy(), cb.zeta.copy()
        zp[s, k, g] *= np.exp(h)
        zm[s, k, g] *= np.exp(-h)
        fd = (loss(zp, cb.D) - loss(zm, cb.D)) / (2 * h)  # d L / d log zeta
        an = gz[s, k, g] * cb.zeta[s, k, g]
        assert abs(fd - an) <= 1e-6 * scale_z, (s, k, g, fd, an)
        Dp, Dm = cb.D.copy(), cb.D.copy()
        Dp[s, k, g] += h
        Dm[s, k, g] -= h
        fd = (loss(cb.zeta, Dp) - loss(cb.zeta, Dm)) / (2 * h)
        assert abs(fd - gD[s, k, g]) <= 1e-6 * scale_D, (s, k, g, fd, gD[s, k, g])


def test_c3_fp32_gradients_vs_fp64(dev):
    """FP32 rrule cotangents vs FP64 ones on the same FP32-rounded inputs, full batch.
    Bound: 1e-6 of the largest |cotangent| (FP32 terms, FP64 reductions; measured 4.9e-7)."""
    d, cb, XA, XB, S_tgt = _loss_setup(dev, torch.float32)
    log("C3 FP32 vs FP64 cotangents")
    o32 = _orb(cb, cb.zeta, cb.D, torch.float32, dev)
    S32, pb32 = aok.rrule_batch_overlap(o32, XA, XB, o32.static_params(), None)
    dS = 2.0 * (S32 - S_tgt)
    g32 = pb32(dS)[4]["Rnl"]
    o64 = _orb(cb, rounded(cb.zeta, torch.float32), rounded(cb.D, torch.float32), torch.float64, dev)
    X64A = (XA[0].double().contiguous(), XA[1])
    X64B = (XB[0].double().contiguous(), XB[1])
    _, pb64 = aok.rrule_batch_overlap(o64, X64A, X64B, o64.static_params(), None)
    g64 = pb64(aok.as_fortran(dS.double()))[4]["Rnl"]
    assert rel_max(host(g32["zeta"]), host(g64["zeta"])) < 1e-6
    assert rel_max(host(g32["D"]), host(g64["D"])) < 1e-6


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_c3_ao_pullback_full_size(dev, dtype, tol):
    """nX = 2^20 AO pullback (dX, dzeta, dD) vs the oracle.  Bounds: FP64 1e-10, FP32 1e-5
    relative to the largest |value| (sums over 2^20 points)."""
    d = configs.c3_ao()
    b = d["basis"]
    log(f"C3 AO pullback nX={len(d['X'])} dtype={dtype}")
    orb = _orb(b, b.zeta, b.D, dtype, dev)
    X = dev_rows(d["X"], dtype, dev)
    sidx = torch.as_tensor(d["sidx"], device=dev)
    dP = dev_f(d["dP"], dtype, dev)
    _, pb = aok.rrule_evaluate(orb, X, sidx=sidx)
    _, _, dX, dps, _ = pb(dP)
    spec, iR, iY = b.ao_maps()
    dXr, dzr, dDr = O.ao_pullback(rounded(d["X"], dtype), d["sidx"], rounded(b.zeta, dtype), rounded(b.D, dtype),
                                  np.zeros(b.ns, np.int64), np.array(iR), np.array(iY), d["L"],
                                  rounded(d["dP"], dtype))
    assert rel_max(host(dX), dXr) < tol
    assert rel_max(host(dps["Rnl"]["zeta"]), dzr) < tol
    assert rel_max(host(dps["Rnl"]["D"]), dDr) < tol
