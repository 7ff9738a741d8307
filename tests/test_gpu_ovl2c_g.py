"""GPU parity of the 2-centre overlap and its (zeta, D) pullback with g shells (l = 4):
a batched, three-species basis with s..g shells of ragged contraction depth, B = 300
pairs (not a multiple of the 64-pair tiles), so the species-pair bucketing, the padded
primitive slots and the light (s/p/d) / heavy (f/g) chunk split all run in the fast
kernels (src/integrals/orbital/overlap_2c.jl:15-131, overlap_2c_grad.jl:71-242, E to
l + 2 = 6 in the pullback).

Tolerances: FP64 1e-12 values / 1e-10 gradients against the FP64 oracle; FP32 against the
oracle on the same FP32-rounded inputs, blockwise 3e-6 (mixed-sign contractions over
zeta in [0.1, 7000], the bound of the uneven six-species batch) and 1e-5 on gradients."""
import numpy as np
import pytest
import torch

import aok
import oracle as O
from aok import configs
from helpers import blockwise_rel, dev_f, dev_rows, host, log, rel_max, rounded

pytestmark = pytest.mark.gpu


def _case(B=300, seed=6):
    cb = configs.g_basis(seed, NZ=3, K=6, zmax=7000.0)
    rng = np.random.default_rng(seed + 100)
    A = rng.normal(0.0, 0.4, size=(B, 3))
    u = rng.standard_normal((B, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    Bp = A + rng.uniform(0.3, 6.0, size=(B, 1)) * u
    return cb, A, Bp, rng.integers(1, 4, B), rng.integers(1, 4, B)


def _orb(cb, dtype, dev):
    spec, _, _ = cb.ao_maps()
    R = aok.GaussianTypeRadials(dev_f(cb.zeta, dtype, dev), dev_f(cb.D, dtype, dev), spec, None,
                                tuple(range(1, cb.NZ + 1)))
    return aok.AtomicOrbitals(R, max(cb.ls), length_unit="bohr")


def _dev_inputs(A, Bp, sA, sB, dtype, dev):
    return ((dev_rows(A, dtype, dev), torch.as_tensor(sA, device=dev)),
            (dev_rows(Bp, dtype, dev), torch.as_tensor(sB, device=dev)))


@pytest.mark.parametrize("mode,dtype", [("fast", torch.float64), ("fast", torch.float32), ("plain", torch.float64)],
                         ids=["fast-f64", "fast-f32", "plain-f64"])
def test_g_shell_forward(dev, mode, dtype):
    cb, A, Bp, sA, sB = _case()
    assert max(cb.ls) == 4
    log(f"2c g shells forward B={len(sA)} mode={mode} dtype={dtype}")
    orb = _orb(cb, dtype, dev)
    XA, XB = _dev_inputs(A, Bp, sA, sB, dtype, dev)
    S = aok.batch_overlap(orb, XA, XB, orb.static_params(), None, mode=mode)
    z, D = rounded(cb.zeta, dtype), rounded(cb.D, dtype)
    coef = O.compile_coef(z, D, cb.ls)
    Sr = O.ovl2c(rounded(A, dtype), rounded(Bp, dtype), sA, sB, cb.ls, coef, z)
    err = blockwise_rel(host(S), Sr)
    print(f"g shells {mode} {dtype}: blockwise {err:.3e}")
    assert err < (1e-12 if dtype == torch.float64 else 3e-6), err


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_g_shell_pullback(dev, dtype):
    cb, A, Bp, sA, sB = _case()
    B = len(sA)
    dS = np.random.default_rng(8).standard_normal((cb.N, cb.N, B))
    log(f"2c g shells pullback B={B} dtype={dtype}")
    orb = _orb(cb, dtype, dev)
    XA, XB = _dev_inputs(A, Bp, sA, sB, dtype, dev)
    _, pb = aok.rrule_batch_overlap(orb, XA, XB, orb.static_params(), None)
    dps = pb(dev_f(dS, dtype, dev))[4]
    z, D = rounded(cb.zeta, dtype), rounded(cb.D, dtype)
    rA, rB, rdS = rounded(A, dtype), rounded(Bp, dtype), rounded(dS, dtype)
    dzr, dDr = O.pullback_ps_2c(rdS, rA, rB, sA, sB, cb.ls, z, D)
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    assert rel_max(host(dps["Rnl"]["zeta"]), dzr) < tol
    assert rel_max(host(dps["Rnl"]["D"]), dDr) < tol
    # kernel level: (dcoef, dalpha) of batch_S_orbital_pb_kernel!
    basis = aok.compile_basis(orb)
    dc, da = aok.overlap_pb_coef_alpha(dev_f(dS, dtype, dev), basis, XA, XB)
    coef = O.compile_coef(z, D, cb.ls)
    dcr, dar = O.ovl2c_pb(rdS, rA, rB, sA, sB, cb.ls, coef, z)
    assert rel_max(host(dc), dcr) < tol
    assert rel_max(host(da), dar) < tol
