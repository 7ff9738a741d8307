"""GPU tests of AtomicOrbitals built from basis-set data (test_gaussianbasis.jl
pattern): values against an independent per-point reference built from the same shell
data, chi = [sum_k coef_k exp(-exp_k |R|^2)] * Z_lm(R) with the oracle's solid
harmonics, per-species slices, padded columns exactly zero, FP32 conversion.

Tolerances: FP64 atol 1e-10 (the reference's own, test_gaussianbasis.jl:81-83), FP32
atol 1e-4 (:88)."""
import numpy as np
import pytest
import torch

import aok
import oracle as O
from helpers import dev_rows, host, log

pytestmark = pytest.mark.gpu


def _gb_ref(orb, bs, X, Zs):
    """test_gaussianbasis.jl:26-43 restated: orbital (n, l, m) of point i uses element
    Zs[i]'s n-th shell of angular momentum l (zero if it has none)."""
    P = np.zeros((len(X), len(orb)))
    for i, R in enumerate(X):
        sh = [s for s in bs.basis if s.Z == Zs[i]]
        Zh, _ = O.ylm(orb.L, R)
        r2 = float(R @ R)
        for col, (n, l, m) in enumerate(orb.spec):
            sl = [s for s in sh if s.l == l]
            rad = float(sl[n - 1].coef @ np.exp(-sl[n - 1].exp * r2)) if n <= len(sl) else 0.0
            P[i, col] = rad * Zh[l * l + l + m]
    return P


@pytest.mark.parametrize("dtype,atol", [(torch.float64, 1e-10), (torch.float32, 1e-4)])
def test_multi_element_values(dev, dtype, atol):
    bs = aok.basis_set("sto-3g", ["H", "C", "N", "O"])
    orb = aok.gaussian_orbitals(bs, length_unit="bohr", dtype=dtype, device=dev)
    rng = np.random.default_rng(21)
    n = 400
    X = rng.normal(0, 1.2, (n, 3))
    Zs = rng.choice([1, 6, 7, 8], n)
    log(f"basis-set AO values H/C/N/O sto-3g n={n} dtype={dtype}")
    Xs = aok.PStates(dev_rows(X, dtype, dev), [int(z) for z in Zs])
    P = host(aok.evaluate(orb, Xs))
    ref = _gb_ref(orb, bs, X, Zs)
    assert np.max(np.abs(P - ref)) < atol
    Pe, _ = aok.evaluate_ed(orb, Xs)
    assert np.max(np.abs(host(Pe) - ref)) < atol
    # hydrogen points: every column but 1s is a padded slot and exactly zero
    h = Zs == 1
    c1s = orb.spec.index((1, 0, 0))
    others = [c for c in range(len(orb)) if c != c1s]
    assert np.all(P[np.ix_(h, others)] == 0.0)
    assert np.all(P[h, c1s] != 0.0)


def test_pairs_form_matches_single_element(dev):
    """C-tagged points of the C + H mix reproduce the C-only orbitals (same spec)."""
    mix = aok.gaussian_orbitals([("C", "sto-3g"), ("H", "sto-3g")], length_unit="angstrom", device=dev)
    onlyC = aok.gaussian_orbitals("sto-3g", ["C"], length_unit="angstrom", device=dev)
    assert mix.spec == onlyC.spec
    X = dev_rows(np.random.default_rng(2).normal(0, 1.0, (128, 3)), torch.float64, dev)
    log("basis-set pairs form vs single element")
    Pm = aok.evaluate(mix, aok.PStates(X, [6] * 128))
    Pc = aok.evaluate(onlyC, X)
    assert torch.max(torch.abs(Pm - Pc)).item() < 1e-12
