"""Randomised parity sweep of the 2-centre fast path: bases of random shape (1-7 shells,
l <= 4, 1-8 primitives per shell with absent and padded slots, 1-4 species), random batch
sizes (ragged tiles, a single pair) and stripe capacities (shared and pullback-only), through the forward, the
(zeta, D) pullback and the position cotangents of the rrule, FP64 and FP32, against the
oracle; then the 3-centre and AO paths likewise.

Tolerances: FP64 1e-12 blockwise (values) / 1e-10 (cotangents).  FP32 is held to the
conditioning of each case.  A block's values are checked to 3e-6 * max(1, kappa_b) of its
largest entry, kappa_b = max|S_b| contracted with |coefficients| / max|S_b| (mixed-sign
contractions of near-equal diffuse exponents cancel: kappa reaches 99 in this sweep, and the
FP32 error stays below 1.2e-6 * kappa_b over 40 seeds); blocks whose FP64 values lie below
FP32's normal range (1e-30) are checked to stay there instead.  Cotangents are checked to
1e-5 * max(1, kappa / 5) of the largest entry, kappa the largest kappa_b of the case.  The
D-cotangent is measured against the larger of its own scale and that of the coefficient
cotangent before the normalisation projection: a species whose shells have one primitive each
has dD = 0 exactly (D is normalised away), so the result is rounding of the projection.
"""
import numpy as np
import pytest
import torch

import aok
import oracle as O
from aok import configs
from helpers import dev_f, dev_rows, host, log, rel_max, rounded

pytestmark = pytest.mark.gpu
# Results near the bottom of a format's range carry no relative precision (their terms and
# intermediate products underflow first): a case whose values all lie below the floor (far-apart
# centres, tight exponents) is judged against the floor instead of its own tiny maximum.
F32_FLOOR = 1e-25
F64_FLOOR = 1e-290


def _rel(got, ref, f64, scale=0.0):
    """max|got - ref| over the largest of max|ref|, `scale` and the format's floor."""
    d = np.max(np.abs(got - ref)) if got.size else 0.0
    return d / max(np.max(np.abs(ref)) if ref.size else 0.0, scale, F64_FLOOR if f64 else F32_FLOOR)


def _random_case(seed):
    rng = np.random.default_rng(1000 + seed)
    ns = int(rng.integers(1, 8))
    ls = sorted(int(x) for x in rng.integers(0, 5, size=ns))
    K = int(rng.integers(1, 9))
    NZ = int(rng.integers(1, 5))
    species = []
    for _ in range(NZ):
        shells = []
        for l in ls:
            if rng.random() < 0.15:
                shells.append(None)  # absent shell: all slots padded
                continue
            depth = int(rng.integers(1, K + 1))
            z = np.sort(np.exp(rng.uniform(np.log(0.1), np.log(3000.0), depth)))[::-1]
            shells.append((z, rng.standard_normal(depth)))
        species.append(shells)
    cb = configs._pad_basis(ls, species, K)
    B = int(rng.choice([1, 63, 64, 65, int(rng.integers(100, 700))]))
    A = rng.normal(0.0, 0.6, size=(B, 3))
    d = rng.normal(size=(B, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    Bp = A + d * rng.uniform(0.3, 6.0, size=(B, 1))
    sA = rng.integers(1, NZ + 1, size=B)
    sB = rng.integers(1, NZ + 1, size=B)
    cap = int(rng.choice([0, 16, 24, 112]))
    cap_pb = int(rng.choice([0, 16, 40]))  # the pullback's own capacity (0: cap)
    return cb, A, Bp, sA, sB, cap, cap_pb


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("seed", list(range(32)))
def test_random_basis_forward_and_pullback(dev, seed, dtype):
    cb, A, Bp, sA, sB, cap, cap_pb = _random_case(seed)
    B = len(sA)
    log(f"2c random basis seed={seed} ls={cb.ls} K={cb.K} NZ={cb.NZ} B={B} cap={cap}/{cap_pb} dtype={dtype}")
    ctx = aok.context(0)
    ctx.tune("ovl_cap", cap)
    ctx.tune("ovl_cap_pb", cap_pb)
    assert ctx.tune_get("ovl_cap") == cap and ctx.tune_get("ovl_cap_pb") == cap_pb
    try:
        spec, _, _ = cb.ao_maps()
        R = aok.GaussianTypeRadials(dev_f(cb.zeta, dtype, dev), dev_f(cb.D, dtype, dev), spec, None,
                                    tuple(range(1, cb.NZ + 1)))
        orb = aok.AtomicOrbitals(R, max(cb.ls), length_unit="bohr")
        XA = (dev_rows(A, dtype, dev), torch.as_tensor(sA, device=dev))
        XB = (dev_rows(Bp, dtype, dev), torch.as_tensor(sB, device=dev))
        S, pb = aok.rrule_batch_overlap(orb, XA, XB, orb.static_params(), None, ctx=ctx, wrt_positions=True)
        dS = np.random.default_rng(seed).standard_normal((cb.N, cb.N, B))
        _, _, dXA, dXB, dps, _ = pb(dev_f(dS, dtype, dev))
    finally:
        ctx.tune("ovl_cap", 0)
        ctx.tune("ovl_cap_pb", 0)
    z, D = rounded(cb.zeta, dtype), rounded(cb.D, dtype)
    rA, rB = rounded(A, dtype), rounded(Bp, dtype)
    coef = O.compile_coef(z, D, cb.ls)
    Sr = O.ovl2c(rA, rB, sA, sB, cb.ls, coef, z)
    Sabs = O.ovl2c(rA, rB, sA, sB, cb.ls, np.abs(coef), z)
    dc, da = O.ovl2c_pb(rounded(dS, dtype), rA, rB, sA, sB, cb.ls, coef, z)
    dDr, dzn = O.compile_coef_pb(dc, z, D, cb.ls)
    dzr = da + dzn
    f64 = dtype == torch.float64
    Sg, dz, dD = host(S), host(dps["Rnl"]["zeta"]), host(dps["Rnl"]["D"])
    assert np.all(np.isfinite(Sg)) and np.all(np.isfinite(dz)) and np.all(np.isfinite(dD))
    bmax = np.max(np.abs(Sr), axis=(0, 1))
    kappa_b = np.max(np.abs(Sabs), axis=(0, 1)) / np.maximum(bmax, 1e-300)
    normal = bmax >= (F64_FLOOR if f64 else 1e-30)
    err_b = np.max(np.abs(Sg - Sr), axis=(0, 1)) / np.maximum(bmax, 1e-300)
    kappa = float(kappa_b[normal].max()) if normal.any() else 1.0
    es = float((err_b[normal] / np.maximum(kappa_b[normal], 1.0)).max()) if normal.any() else 0.0
    tiny = float(np.abs(Sg[:, :, ~normal]).max()) if (~normal).any() else 0.0
    live = D != 0
    raw_D = np.abs(dc[live] * coef[live] / D[live]).max() if live.any() else 0.0
    scale_D = max(np.abs(dDr).max(), raw_D, F64_FLOOR if f64 else F32_FLOOR)
    ez = _rel(dz, dzr, f64)
    eD = float(np.abs(dD - dDr).max() / scale_D)
    rXA, _ = O.ovl2c_dx(rounded(dS, dtype), rA, rB, sA, sB, cb.ls, coef, z)
    ex = _rel(host(dXA), rXA, f64)
    print(f"seed {seed} {dtype}: S {es:.2e} (per kappa; kappa {kappa:.1f}, {(~normal).sum()} blocks"
          f" below FP32 range, max {tiny:.1e}) dzeta {ez:.2e} dD {eD:.2e} dA {ex:.2e}")
    assert torch.equal(dXB, -dXA)
    assert es < (1e-12 if f64 else 3e-6), es
    assert tiny < 1e-29, tiny
    tol = 1e-10 if f64 else 1e-5 * max(1.0, kappa / 5)
    assert ez < tol, (ez, tol)
    assert eD < tol, (eD, tol)
    assert ex < tol, (ex, tol)


def _random_species(rng, ls, K, NZ, zlo, zhi):
    species = []
    for _ in range(NZ):
        shells = []
        for l in ls:
            if rng.random() < 0.15:
                shells.append(None)
                continue
            depth = int(rng.integers(1, K + 1))
            z = np.sort(np.exp(rng.uniform(np.log(zlo), np.log(zhi), depth)))[::-1]
            shells.append((z, rng.standard_normal(depth)))
        species.append(shells)
    return configs._pad_basis(ls, species, K)


def _orb(cb, dtype, dev):
    spec, _, _ = cb.ao_maps()
    R = aok.GaussianTypeRadials(dev_f(cb.zeta, dtype, dev), dev_f(cb.D, dtype, dev), spec, None,
                                tuple(range(1, cb.NZ + 1)))
    return aok.AtomicOrbitals(R, max(cb.ls), length_unit="bohr")


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("seed", list(range(32)))
def test_random_basis_3c_forward_and_pullback(dev, seed, dtype):
    """3-centre fast path (shells up to d): random bases of 1-5 shells, 1-6 primitive slots,
    1-3 species, batches of 1 to 130 triples (ragged tiles); the rrule forward and its
    (zeta, D) pullback.  FP32 bounds and metrics as in the 2-centre sweep: 3e-6 x max(1,
    kappa_b) for values, 1e-5 x max(1, kappa / 5) for cotangents, dD against the larger of its
    own scale and the coefficient cotangent's before the normalisation projection; dzeta
    against the larger of its own scale and a tenth of its terms' magnitudes summed (so at
    most 1e-6 of those: about 17 FP32 ulp)."""
    rng = np.random.default_rng(2000 + seed)
    ns = int(rng.integers(1, 6))
    ls = sorted(int(x) for x in rng.integers(0, 3, size=ns))
    K, NZ = int(rng.integers(1, 7)), int(rng.integers(1, 4))
    cb = _random_species(rng, ls, K, NZ, 0.1, 3000.0)
    B = int(rng.choice([1, 7, 33, 64, 65, 130]))
    P3 = rng.normal(0.0, 1.2, size=(3, B, 3))
    S3 = rng.integers(1, NZ + 1, size=(3, B))
    log(f"3c random basis seed={seed} ls={cb.ls} K={cb.K} NZ={cb.NZ} B={B} dtype={dtype}")
    orb = _orb(cb, dtype, dev)
    X = [(dev_rows(P3[c], dtype, dev), torch.as_tensor(S3[c], device=dev)) for c in range(3)]
    T, pb = aok.rrule_batch_overlap_3c(orb, *X, orb.static_params())
    dT = np.random.default_rng(seed).standard_normal((cb.N, cb.N, cb.N, B))
    dps = pb(dev_f(dT, dtype, dev))[5]
    z, D = rounded(cb.zeta, dtype), rounded(cb.D, dtype)
    coef = O.compile_coef(z, D, cb.ls)
    pos = [rounded(P3[c], dtype) for c in range(3)]
    Tr = O.ovl3c(*pos, *S3, cb.ls, coef, z)
    Tabs = O.ovl3c(*pos, *S3, cb.ls, np.abs(coef), z)
    dc, da = O.ovl3c_pb(rounded(dT, dtype), *pos, *S3, cb.ls, coef, z)
    dDr, dzn = O.compile_coef_pb(dc, z, D, cb.ls)
    dzr = da + dzn
    # the exponent cotangent's own conditioning: its terms' magnitudes summed (|dT|, |coef|)
    # against the result, which cancels over the elements of a block and then against the
    # normalisation VJP (up to ~500x for a single triple or a 2-primitive d shell)
    _, da_abs = O.ovl3c_pb(np.abs(rounded(dT, dtype)), *pos, *S3, cb.ls, np.abs(coef), z)
    f64 = dtype == torch.float64
    Tg, dz, dD = host(T), host(dps["Rnl"]["zeta"]), host(dps["Rnl"]["D"])
    assert np.all(np.isfinite(Tg)) and np.all(np.isfinite(dz)) and np.all(np.isfinite(dD))
    bmax = np.max(np.abs(Tr), axis=(0, 1, 2))
    kappa_b = np.max(np.abs(Tabs), axis=(0, 1, 2)) / np.maximum(bmax, 1e-300)
    normal = bmax >= (F64_FLOOR if f64 else 1e-30)
    err_b = np.max(np.abs(Tg - Tr), axis=(0, 1, 2)) / np.maximum(bmax, 1e-300)
    kappa = float(kappa_b[normal].max()) if normal.any() else 1.0
    es = float((err_b[normal] / np.maximum(kappa_b[normal], 1.0)).max()) if normal.any() else 0.0
    tiny = float(np.abs(Tg[..., ~normal]).max()) if (~normal).any() else 0.0
    live = D != 0
    raw_D = np.abs(dc[live] * coef[live] / D[live]).max() if live.any() else 0.0
    scale_D = max(np.abs(dDr).max(), raw_D, F64_FLOOR if f64 else F32_FLOOR)
    ez = _rel(dz, dzr, f64, scale=0.1 * np.abs(da_abs).max())
    eD = float(np.abs(dD - dDr).max() / scale_D)
    print(f"3c seed {seed} {dtype}: T {es:.2e} (per kappa; kappa {kappa:.1f}) dzeta {ez:.2e} dD {eD:.2e}"
          f" (|dD| {np.abs(dDr).max():.2e} of the unprojected {raw_D:.2e})")
    assert es < (1e-12 if f64 else 3e-6), es
    assert tiny < 1e-29, tiny
    tol = 1e-10 if f64 else 1e-5 * max(1.0, kappa / 5)
    assert ez < tol, (ez, tol)
    assert eD < tol, (eD, tol)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("seed", list(range(32)))
def test_random_basis_ao(dev, seed, dtype):
    """AO path: random bases of 1-6 shells up to l = 4, 1-6 primitive slots, 1-4 species,
    ragged point counts (1 to 20000); evaluate, evaluate_ed and the rrule pullback
    against the oracle.  Tolerances as the fixed-basis AO tests (relative to each output's
    largest entry): FP64 1e-12 values and gradients, 1e-11 pullbacks; FP32 2e-6, 1e-5, 2e-5,
    each times max(1, kappa, kappa_e / 20), kappa = max|P| with |D| / max|P| (mixed-sign
    radials), kappa_e = the relative change of P per relative change of the exponents."""
    rng = np.random.default_rng(3000 + seed)
    ns = int(rng.integers(1, 7))
    ls = sorted(int(x) for x in rng.integers(0, 5, size=ns))
    K, NZ = int(rng.integers(1, 7)), int(rng.integers(1, 5))
    cb = _random_species(rng, ls, K, NZ, 0.1, 50.0)
    nX = int(rng.choice([1, 63, 257, int(rng.integers(1000, 20000))]))
    X = rng.normal(0.0, 1.5, size=(nX, 3))
    sp = rng.integers(1, NZ + 1, nX)
    dPin = rng.standard_normal((nX, cb.nOrb))
    log(f"ao random basis seed={seed} ls={cb.ls} K={cb.K} NZ={cb.NZ} nX={nX} dtype={dtype}")
    orb = _orb(cb, dtype, dev)
    Xd = dev_rows(X, dtype, dev)
    sidx = torch.as_tensor(sp, dtype=torch.int64, device=dev)
    P = host(aok.evaluate(orb, Xd, sidx=sidx))
    P2, dP = (host(t) for t in aok.evaluate_ed(orb, Xd, sidx=sidx))
    _, pb = aok.rrule_evaluate(orb, Xd, sidx=sidx)
    _, _, dX, dps, _ = pb(dev_f(dPin, dtype, dev))
    _, iR, iY = cb.ao_maps()
    args = (rounded(X, dtype), sp, rounded(cb.zeta, dtype))
    maps = (np.zeros(cb.ns, np.int64), np.array(iR), np.array(iY), max(cb.ls))
    Pr, dPr = O.ao_eval(*args, rounded(cb.D, dtype), *maps, grad=True)
    Pa = O.ao_eval(*args, np.abs(rounded(cb.D, dtype)), *maps)
    dXr, dzr, dDr = O.ao_pullback(*args, rounded(cb.D, dtype), *maps, rounded(dPin, dtype))
    kappa = max(1.0, float(np.abs(Pa).max() / max(np.abs(Pr).max(), 1e-300)))
    # exponent conditioning: e^{-zeta r^2} moves by zeta r^2 times a relative change of its
    # argument, which FP32 rounds (a point far from a tight primitive: zeta r^2 ~ 50 gives
    # 3e-6 alone); measured as the relative change of P under zeta (1 + 1e-6)
    Pz = O.ao_eval(args[0], args[1], args[2] * (1 + 1e-6), rounded(cb.D, dtype), *maps)
    kappa_e = rel_max(Pz, Pr) / 1e-6
    kappa = max(kappa, kappa_e / 20)
    tv, tg, tp = (1e-12, 1e-12, 1e-11) if dtype == torch.float64 else (2e-6 * kappa, 1e-5 * kappa, 2e-5 * kappa)
    e = {"P": rel_max(P, Pr), "P_ed": rel_max(P2, Pr),
         "dP": max(rel_max(dP[:, :, c], dPr[:, :, c]) for c in range(3)),
         "dX": rel_max(host(dX), dXr), "dzeta": rel_max(host(dps["Rnl"]["zeta"]), dzr),
         "dD": rel_max(host(dps["Rnl"]["D"]), dDr)}
    print(f"ao seed {seed} {dtype}: kappa {kappa:.1f} (exponent {kappa_e:.1f}) " + " ".join(f"{k} {v:.2e}" for k, v in e.items()))
    assert e["P"] < tv and e["P_ed"] < tv, e
    assert e["dP"] < tg, e
    assert e["dX"] < tp and e["dzeta"] < tp and e["dD"] < tp, e
