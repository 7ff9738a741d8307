"""GPU parity tests of the 3-centre overlap through the C ABI.

Tolerances: FP64 (fast and plain) 1e-12 relative to the block max; FP32 fast 2e-6
blockwise against the FP64 oracle on the same rounded inputs."""
import json
import os

import numpy as np
import pytest
import torch

import aok
import oracle as O
from aok import configs
from helpers import dev_f, dev_rows, host, log, rel_max, rounded

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")))


def _basis(ls, zeta, coef, dtype, dev, zlist=None):
    nbf = [(l + 1) * (l + 2) // 2 for l in ls]
    off = [0]
    for n in nbf:
        off.append(off[-1] + n)
    return aok.CartesianGTOBasis(max(ls), len(ls), zeta.shape[1], list(ls), nbf, off, off[-1],
                                 dev_f(zeta, dtype, dev), dev_f(coef, dtype, dev),
                                 tuple(zlist or range(1, zeta.shape[2] + 1)), 1.0)


def _blockwise(got, ref):
    num = np.max(np.abs(got - ref), axis=(0, 1, 2))
    den = np.maximum(np.max(np.abs(ref), axis=(0, 1, 2)), 1e-300)
    return float(np.max(num / den))


@pytest.mark.parametrize("mode", ["fast", "plain"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_golden_triples(dev, dtype, mode):
    for blk in GOLD["blocks_3c"]:
        la, lb, lc = blk["la"], blk["lb"], blk["lc"]
        pr = blk["prims"]
        K = max(len(p[0]) for p in pr)
        zeta = np.ones((3, K, 1))
        coef = np.zeros((3, K, 1))
        for s, (z, c) in enumerate(pr):
            zeta[s, :len(z), 0], coef[s, :len(c), 0] = z, c
        log(f"3c golden l=({la},{lb},{lc}) dtype={dtype} mode={mode}")
        basis = _basis([la, lb, lc], zeta, coef, dtype, dev)
        P = [dev_rows(np.array([blk[k]]), dtype, dev) for k in ("A", "B", "C")]
        out = aok.batch_overlap_3c(basis, (P[0], None), (P[1], None), (P[2], None), mode=mode)
        na, nb, nc = [(l + 1) * (l + 2) // 2 for l in (la, lb, lc)]
        got = host(out)[:na, na:na + nb, na + nb:na + nb + nc, 0]
        ref = np.array(blk["out"]).reshape((na, nb, nc), order="F")
        assert rel_max(got, ref) < (1e-12 if dtype == torch.float64 else 2e-5), (la, lb, lc)


@pytest.mark.parametrize("mode,dtype", [("fast", torch.float64), ("plain", torch.float64), ("fast", torch.float32)])
def test_c4a_species_triples_vs_oracle(dev, mode, dtype):
    d = configs.c4a(B=300)
    cb = d["basis"]
    log(f"3c C4a B=300 mode={mode} dtype={dtype}")
    coef = O.compile_coef(rounded(cb.zeta, dtype), rounded(cb.D, dtype), cb.ls)
    coef_r = rounded(coef, dtype)
    basis = _basis(cb.ls, cb.zeta, coef, dtype, dev)
    X = [(dev_rows(d[f"pos{c}"], dtype, dev), torch.as_tensor(d[f"sidx{c}"], device=dev)) for c in "ABC"]
    out = aok.batch_overlap_3c(basis, *X, mode=mode)
    ref = O.ovl3c(*(rounded(d[f"pos{c}"], dtype) for c in "ABC"), d["sidxA"], d["sidxB"], d["sidxC"], cb.ls,
                  coef_r, rounded(cb.zeta, dtype))
    err = _blockwise(host(out), ref)
    print(f"3c {mode} {dtype}: blockwise {err:.3e}")
    assert err < (1e-12 if dtype == torch.float64 else 2e-6)


def test_single_species_reduces_to_reference_form(dev):
    """sigma_A = sigma_B = sigma_C: the species-aware kernel is the reference's single-basis
    batch_V_kernel! (NULL species arrays mean species 1)."""
    d = configs.c4a(B=70)
    cb = d["basis"]
    coef = O.compile_coef(cb.zeta, cb.D, cb.ls)
    log("3c single species")
    basis1 = _basis(cb.ls, cb.zeta[:, :, :1], coef[:, :, :1], torch.float64, dev)
    X = [(dev_rows(d[f"pos{c}"], torch.float64, dev), None) for c in "ABC"]
    out1 = host(aok.batch_overlap_3c(basis1, *X))
    basis3 = _basis(cb.ls, cb.zeta, coef, torch.float64, dev)
    ones = torch.ones(70, dtype=torch.int64, device=dev)
    out3 = host(aok.batch_overlap_3c(basis3, *[(x[0], ones) for x in X]))
    assert np.array_equal(out1, out3)
    ref = O.ovl3c(*(d[f"pos{c}"] for c in "ABC"), None, None, None, cb.ls, coef[:, :, :1], cb.zeta[:, :, :1])
    assert _blockwise(out1, ref) < 1e-12


def test_forward_output_beyond_4gib(dev):
    """A 4.32 GB FP32 output (C4a basis, N = 15, B = 320,000 triples): the first, last (ragged
    tile) and 100 scattered triples' blocks against the oracle (64-bit block addressing)."""
    B = 320_000
    d = configs.c4a(B=B)
    cb = d["basis"]
    assert B * cb.N ** 3 * 4 >= 1 << 32
    log(f"3c forward beyond 4 GiB B={B}")
    dtype = torch.float32
    coef = O.compile_coef(rounded(cb.zeta, dtype), rounded(cb.D, dtype), cb.ls)
    basis = _basis(cb.ls, cb.zeta, coef, dtype, dev)
    X = [(dev_rows(d[f"pos{c}"], dtype, dev), torch.as_tensor(d[f"sidx{c}"], device=dev)) for c in "ABC"]
    out = aok.batch_overlap_3c(basis, *X)
    idx = np.concatenate([np.arange(100), np.arange(B - 100, B),
                          np.sort(np.random.default_rng(3).choice(np.arange(100, B - 100), 100, replace=False))])
    got = host(out[:, :, :, torch.as_tensor(idx, device=dev)])
    del out
    torch.cuda.empty_cache()
    ref = O.ovl3c(*(rounded(d[f"pos{c}"][idx], dtype) for c in "ABC"), d["sidxA"][idx], d["sidxB"][idx],
                  d["sidxC"][idx], cb.ls, rounded(coef, dtype), rounded(cb.zeta, dtype))
    err = _blockwise(got, ref)
    print(f"3c beyond 4 GiB: blockwise {err:.3e} over {len(idx)} triples")
    assert err < 2e-6, err
