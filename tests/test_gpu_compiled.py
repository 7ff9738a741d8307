"""GPU tests of the geometry-in overlaps on a CompiledBasis (kernels_2c.jl:158-184,
kernels_3c.jl:183-222): unit-carrying positions, one basis, no species axis.

Tolerances (written per assertion):
  FP64 vs the quadrature golden vectors and vs the FP64 oracle: 1e-12 relative to the
  block max;
  FP32 fast kernel vs the FP64 oracle on the same FP32-rounded inputs: blockwise 1e-6
  (the target metric; the measured value is printed);
  same geometry in angstrom and in nm: 1e-13 (the positions differ by one rounding);
  closed forms (normalised s and p self-overlap = 1): 1e-13."""
import json
import os

import numpy as np
import pytest
import torch

import aok
import oracle as O
from helpers import blockwise_rel, host, log, rel_max

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")))


def _bohr(rows):
    """(B, 3) Bohr values as unit-carrying (3, B) positions."""
    return aok.Length(np.asarray(rows, np.float64).T.copy(), "bohr")


def _random_basis(seed, lmax=3, kmax=5, nshell=6):
    rng = np.random.default_rng(seed)
    shells = []
    for i in range(nshell):
        l = i % (lmax + 1)
        k = int(rng.integers(1, kmax + 1))
        shells.append((l, np.exp(rng.uniform(-1.5, 2.5, k)), rng.uniform(-1.0, 1.0, k)))
    return aok.compile_shells(shells)


def _oracle_2c(cb, A, B, dtype):
    c, a = cb.padded()
    c, a = c.cpu().numpy(), a.cpu().numpy()
    if dtype == torch.float32:
        c, a = c.astype(np.float32).astype(np.float64), a.astype(np.float32).astype(np.float64)
    ones = np.ones(len(A), np.int64)
    return O.ovl2c(A, B, ones, ones, list(cb.ls), c, a)


@pytest.mark.parametrize("mode", ["fast", "plain"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_golden_blocks_compiled(dev, dtype, mode):
    """Quadrature golden vectors through compile_shells + batch_overlap (Bohr positions)."""
    for blk in GOLD["blocks_2c"]:
        la, lb = blk["la"], blk["lb"]
        log(f"compiled 2c golden l=({la},{lb}) dtype={dtype} mode={mode}")
        cb = aok.adapt_basis(aok.compile_shells([(la, blk["za"], blk["ca"]), (lb, blk["zb"], blk["cb"])]), dev, dtype)
        out = aok.batch_overlap(cb, _bohr([blk["A"]]), _bohr([blk["B"]]), dtype=dtype, mode=mode)
        assert out.device == dev and out.dtype == dtype and tuple(out.shape) == (cb.nbf_total, cb.nbf_total, 1)
        nba, nbb = (la + 1) * (la + 2) // 2, (lb + 1) * (lb + 2) // 2
        got = host(out)[:nba, nba:nba + nbb, 0]
        ref = np.array(blk["out"]).reshape((nba, nbb), order="F")
        # FP32: the golden inputs are FP64, so the bound also covers input rounding
        t = 1e-12 if dtype == torch.float64 else (2e-5 if mode == "fast" else 1e-4)
        assert rel_max(got, ref) < t, (la, lb)


@pytest.mark.parametrize("mode,dtype", [("fast", torch.float64), ("plain", torch.float64), ("fast", torch.float32)])
def test_random_basis_vs_oracle(dev, mode, dtype):
    cb = aok.adapt_basis(_random_basis(11), dev, dtype)
    rng = np.random.default_rng(4)
    Bn = 777
    ang_A = rng.uniform(-1.5, 1.5, (3, Bn))
    ang_B = ang_A + rng.normal(0.0, 0.8, (3, Bn))
    pA, pB = aok.Length(ang_A, "angstrom"), aok.Length(ang_B, "angstrom")
    log(f"compiled 2c random basis B={Bn} mode={mode} dtype={dtype}")
    S = aok.batch_overlap(cb, pA, pB, dtype=dtype, mode=mode)
    A = host(aok.to_bohr(pA, dtype))
    B = host(aok.to_bohr(pB, dtype))
    Sr = _oracle_2c(cb, A, B, dtype)
    err = blockwise_rel(host(S), Sr)
    print(f"compiled 2c {mode} {dtype}: blockwise {err:.3e}")
    assert err < (1e-12 if dtype == torch.float64 else 1e-6), err
    # the mutating form into a preallocated column-major output gives the same values
    out = torch.empty((Bn, cb.nbf_total, cb.nbf_total), dtype=dtype, device=dev).permute(2, 1, 0)
    aok.batch_overlap_(out, cb, pA, pB, mode=mode)
    assert torch.equal(out, S)


def test_units_angstrom_vs_nm(dev):
    cb = aok.adapt_basis(_random_basis(12, lmax=2), dev)
    rng = np.random.default_rng(8)
    ang_A, ang_B = rng.uniform(-2, 2, (3, 300)), rng.uniform(-2, 2, (3, 300))
    log("compiled 2c angstrom vs nm")
    S1 = aok.batch_overlap(cb, aok.Length(ang_A, "angstrom"), aok.Length(ang_B, "angstrom"))
    S2 = aok.batch_overlap(cb, aok.Length(ang_A / 10, "nm"), aok.Length(ang_B / 10, "nm"))
    assert blockwise_rel(host(S2), host(S1)) < 1e-13


def _normalised(l, exps, d):
    """Primitive-normalised, contraction-normalised coefficients of the x^l component
    of an s (l = 0) or p (l = 1) shell: <g|g> = 1 on one centre (closed form)."""
    exps, d = np.asarray(exps, np.float64), np.asarray(d, np.float64)
    c = d * (2 * exps / np.pi) ** 0.75 * (4 * exps) ** (l / 2)
    S = np.sum(np.outer(c, c) * (np.pi / np.add.outer(exps, exps)) ** 1.5 / (2 * np.add.outer(exps, exps)) ** l)
    return c / np.sqrt(S)


def test_self_overlap_and_symmetry(dev):
    """STO-3G hydrogen s and a 2-primitive p shell, normalised in closed form: the
    same-centre diagonal is 1; swapping the centres transposes every block."""
    sto = ([3.42525091, 0.62391373, 0.16885540], [0.15432897, 0.53532814, 0.44463454])
    pz = ([1.3, 0.35], [0.4, 0.7])
    cb = aok.adapt_basis(aok.compile_shells([(0, sto[0], _normalised(0, *sto)), (1, pz[0], _normalised(1, *pz))]),
                         dev)
    rng = np.random.default_rng(9)
    X = rng.uniform(-1, 1, (3, 64))
    Y = X + rng.normal(0, 0.7, (3, 64))
    log("compiled 2c self overlap / symmetry")
    S0 = host(aok.batch_overlap(cb, aok.Length(X, "angstrom"), aok.Length(X, "angstrom")))
    for i in range(cb.nbf_total):
        np.testing.assert_allclose(S0[i, i, :], 1.0, rtol=1e-13)
    Sab = host(aok.batch_overlap(cb, aok.Length(X, "angstrom"), aok.Length(Y, "angstrom")))
    Sba = host(aok.batch_overlap(cb, aok.Length(Y, "angstrom"), aok.Length(X, "angstrom")))
    assert rel_max(Sba.transpose(1, 0, 2), Sab) < 1e-13


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_golden_triples_compiled(dev, dtype):
    for blk in GOLD["blocks_3c"]:
        ls = (blk["la"], blk["lb"], blk["lc"])
        log(f"compiled 3c golden l={ls} dtype={dtype}")
        cb = aok.adapt_basis(aok.compile_shells([(l, z, c) for l, (z, c) in zip(ls, blk["prims"])]), dev, dtype)
        out = host(aok.batch_overlap_3c(cb, _bohr([blk["A"]]), _bohr([blk["B"]]), _bohr([blk["C"]]), dtype=dtype))
        nb = [(l + 1) * (l + 2) // 2 for l in ls]
        o = np.cumsum([0] + nb)
        got = out[o[0]:o[1], o[1]:o[2], o[2]:o[3], 0]
        ref = np.array(blk["out"]).reshape(tuple(nb), order="F")
        assert rel_max(got, ref) < (1e-12 if dtype == torch.float64 else 2e-5), ls


def test_random_basis_3c_vs_oracle(dev):
    cb = aok.adapt_basis(_random_basis(13, lmax=2, kmax=3, nshell=4), dev)
    rng = np.random.default_rng(6)
    Bn = 150
    P = [aok.Length(rng.uniform(-1.5, 1.5, (3, Bn)), "angstrom") for _ in range(3)]
    log(f"compiled 3c random basis B={Bn}")
    T = aok.batch_overlap_3c(cb, *P)
    c, a = cb.padded()
    ones = np.ones(Bn, np.int64)
    Pb = [host(aok.to_bohr(p)) for p in P]
    Tr = O.ovl3c(*Pb, ones, ones, ones, list(cb.ls), c.cpu().numpy(), a.cpu().numpy())
    assert rel_max(host(T), Tr) < 1e-12
