"""GPU parity tests of H1 (atomic orbitals) through the C ABI against the FP64 oracle.

Tolerances: FP64 kernels 1e-12 (relative to the largest |value| of the output);
FP32 kernels 2e-6 on values, 1e-5 on gradients and pullbacks (same FP32-rounded inputs
fed to the oracle).  The reference's own device tests use 1e-3 / 1e-2
(test/orbitals/test_gpu.jl:31,37,50-55)."""
import numpy as np
import pytest
import torch

import aok
import oracle as O
from aok import configs
from helpers import dev_f, dev_rows, host, log, rel_max, rounded

pytestmark = pytest.mark.gpu


def _basis_from(cfg_basis, dtype, device, zlist=None, length_unit="bohr", decay="gaussian"):
    spec, iR, iY = cfg_basis.ao_maps()
    zl = zlist or tuple(range(1, cfg_basis.NZ + 1))
    TR = aok.GaussianTypeRadials if decay == "gaussian" else aok.SlaterTypeRadials
    R = TR(dev_f(cfg_basis.zeta, dtype, device), dev_f(cfg_basis.D, dtype, device), spec, None, zl)
    return aok.AtomicOrbitals(R, max(cfg_basis.ls), length_unit=length_unit)


def _oracle_ao(cfg_basis, X, sidx, dtype, grad=False, scale=1.0, decay=0, L=None):
    spec, iR, iY = cfg_basis.ao_maps()
    return O.ao_eval(rounded(X, dtype), sidx, rounded(cfg_basis.zeta, dtype), rounded(cfg_basis.D, dtype),
                     np.zeros(cfg_basis.ns, np.int64) if decay == 0 else None, np.array(iR), np.array(iY),
                     L if L is not None else max(cfg_basis.ls), lengthscale=scale, decay=decay, grad=grad)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_c0_values_and_gradients(dev, dtype, tol):
    cfg = configs.c0()
    log(f"ao C0 dtype={dtype} nX={len(cfg['X'])}")
    basis = _basis_from(cfg["basis"], dtype, dev)
    X = dev_rows(cfg["X"], dtype, dev)
    P, dP = aok.evaluate_ed(basis, X)
    P0 = aok.evaluate(basis, X)
    Pr, dPr = _oracle_ao(cfg["basis"], cfg["X"], None, dtype, grad=True)
    assert rel_max(host(P), Pr) < tol
    assert rel_max(host(P0), Pr) < tol
    assert rel_max(host(dP), dPr) < (tol if dtype == torch.float64 else 1e-5)


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_multi_species_pstates_and_sidx(dev, dtype, itype):
    cb = configs.c1_basis(1)
    rng = np.random.default_rng(11)
    nX = 1000 + 37  # not a multiple of the block size
    Xh = rng.normal(0, 2.0, size=(nX, 3))
    sp = rng.integers(1, 4, nX)
    log(f"ao multi-species dtype={dtype} itype={itype} nX={nX}")
    zlist = (6, 1, 8)
    basis = _basis_from(cb, dtype, dev, zlist=zlist)
    X = dev_rows(Xh, dtype, dev)
    labels = [zlist[s - 1] for s in sp]
    P1 = aok.evaluate(basis, aok.PStates(X, labels))
    P2 = aok.evaluate(basis, X, sidx=torch.as_tensor(sp, dtype=itype, device=dev))
    Pr = _oracle_ao(cb, Xh, sp, dtype)
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    assert rel_max(host(P1), Pr) < tol
    assert np.array_equal(host(P1), host(P2))


def test_length_unit_angstrom(dev):
    cb = configs.c0_basis(0)[0]
    rng = np.random.default_rng(3)
    Xh = rng.normal(0, 1.0, size=(300, 3))
    log("ao length unit angstrom")
    ba = _basis_from(cb, torch.float64, dev, length_unit="angstrom")
    bb = _basis_from(cb, torch.float64, dev, length_unit="bohr")
    PA, dA = aok.evaluate_ed(ba, dev_rows(Xh, torch.float64, dev))
    PB, dB = aok.evaluate_ed(bb, dev_rows(aok.ANG2BOHR * Xh, torch.float64, dev))
    assert rel_max(host(PA), host(PB)) < 1e-12
    assert rel_max(host(dA), aok.ANG2BOHR * host(dB)) < 1e-12


def test_non_canonical_orbital_map(dev):
    """Arbitrary iR/iY maps run through the generic kernel variant."""
    cb = configs.c0_basis(0)[0]
    spec, iR, iY = cb.ao_maps()
    perm = np.random.default_rng(5).permutation(len(iR))
    iRp, iYp = [iR[p] for p in perm], [iY[p] for p in perm]
    rng = np.random.default_rng(6)
    Xh = rng.normal(0, 1.5, size=(500, 3))
    log("ao non-canonical map")
    basis = _basis_from(cb, torch.float64, dev)
    st = basis.static_state()
    st["iR"], st["iY"] = iRp, iYp
    P, dP = aok.evaluate_ed(basis, dev_rows(Xh, torch.float64, dev), None, st)
    Pr, dPr = O.ao_eval(Xh, None, cb.zeta, cb.D, None, np.array(iRp), np.array(iYp), 2, grad=True)
    assert rel_max(host(P), Pr) < 1e-12
    assert rel_max(host(dP), dPr) < 1e-12
    dPin = np.random.default_rng(7).standard_normal(Pr.shape)
    P_, pb = aok.rrule_evaluate(basis, dev_rows(Xh, torch.float64, dev), basis.static_params(), st)
    _, _, dX, dps, _ = pb(dev_f(dPin, torch.float64, dev))
    dXr, dzr, dDr = O.ao_pullback(Xh, None, cb.zeta, cb.D, None, np.array(iRp), np.array(iYp), 2, dPin)
    assert rel_max(host(dX), dXr) < 1e-11
    assert rel_max(host(dps["Rnl"]["zeta"]), dzr) < 1e-11
    assert rel_max(host(dps["Rnl"]["D"]), dDr) < 1e-11


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 2e-5)])
def test_pullback_multi_species(dev, dtype, tol):
    cb = configs.c1_basis(1)
    rng = np.random.default_rng(21)
    nX = 5000 + 13
    Xh = rng.normal(0, 2.0, size=(nX, 3))
    sp = rng.integers(1, 4, nX)
    dPin = rng.standard_normal((nX, cb.nOrb))
    log(f"ao pullback multi-species dtype={dtype} nX={nX}")
    basis = _basis_from(cb, dtype, dev)
    X = dev_rows(Xh, dtype, dev)
    sidx = torch.as_tensor(sp, dtype=torch.int64, device=dev)
    P, pb = aok.rrule_evaluate(basis, X, sidx=sidx)
    _, _, dX, dps, _ = pb(dev_f(dPin, dtype, dev))
    spec, iR, iY = cb.ao_maps()
    dXr, dzr, dDr = O.ao_pullback(rounded(Xh, dtype), sp, rounded(cb.zeta, dtype), rounded(cb.D, dtype), None,
                                  np.array(iR), np.array(iY), 2, rounded(dPin, dtype))
    assert rel_max(host(dX), dXr) < tol
    assert rel_max(host(dps["Rnl"]["zeta"]), dzr) < tol
    assert rel_max(host(dps["Rnl"]["D"]), dDr) < tol
    # deterministic: a second call is bitwise identical
    _, _, dX2, dps2, _ = pb(dev_f(dPin, dtype, dev))
    assert torch.equal(dps["Rnl"]["zeta"], dps2["Rnl"]["zeta"]) and torch.equal(dX, dX2)


def test_absent_species_zero_and_slater(dev):
    cb = configs.c1_basis(1)
    rng = np.random.default_rng(31)
    Xh = rng.normal(0, 1.5, size=(400, 3))
    sp = rng.integers(1, 3, 400)  # species 3 absent
    log("ao absent species")
    basis = _basis_from(cb, torch.float64, dev)
    dps = aok.pullback_ps(dev_f(np.ones((400, cb.nOrb)), torch.float64, dev), basis,
                          dev_rows(Xh, torch.float64, dev), None, None,
                          sidx=torch.as_tensor(sp, device=dev))
    assert torch.all(dps["Rnl"]["zeta"][:, :, 2] == 0) and torch.all(dps["Rnl"]["D"][:, :, 2] == 0)
    # Slater radials (decay r, r^{n1-1}) -- the "next" family, same kernels
    log("ao slater")
    sl = aok.slater_orbitals(3, 2, length_unit="bohr", K=4, device=dev)
    Xs = rng.normal(0, 1.5, size=(257, 3))
    P, dP = aok.evaluate_ed(sl, dev_rows(Xs, torch.float64, dev))
    st = sl.static_state()
    Pr, dPr = O.ao_eval(Xs, None, host(sl.Rnl.zeta), host(sl.Rnl.D), np.array(st["Rnl"]["poly"]),
                        np.array(st["iR"]), np.array(st["iY"]), sl.L, decay=1, grad=True)
    assert rel_max(host(P), Pr) < 1e-12
    assert rel_max(host(dP), dPr) < 1e-11


def test_errors_and_edges(dev):
    cb = configs.c0_basis(0)[0]
    basis = _basis_from(cb, torch.float64, dev, zlist=(6,))
    X = dev_rows(np.zeros((5, 3)), torch.float64, dev)
    log("ao errors")
    with pytest.raises(aok.AOKSpeciesError):
        aok.evaluate(basis, aok.PStates(X, [6, 6, 99, 6, 6]))
    with pytest.raises(aok.AOKSpeciesError):  # out-of-range index caught by the kernel
        aok.evaluate(basis, X, sidx=torch.tensor([1, 1, 2, 1, 1], device=dev))
    st = basis.static_state()
    st["Ylm"]["L"] = 7
    with pytest.raises(aok.AOKLmaxError):
        aok.evaluate(basis, X, None, st)
    # r = 0: the Gaussian gradient is finite (closed form), values exact
    P, dP = aok.evaluate_ed(basis, X)
    assert torch.all(torch.isfinite(dP))
    # empty batch
    P0 = aok.evaluate(basis, dev_rows(np.zeros((0, 3)), torch.float64, dev))
    assert P0.shape == (0, len(basis))
