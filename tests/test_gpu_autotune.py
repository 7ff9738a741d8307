"""The autotuner's search on the GPU (aok.autotune; no reference counterpart): it sets the
fastest measured knobs on the context and records them.  The 2-centre knobs leave S bit
for bit unchanged (each element is computed by one lane in a fixed order) and move the
FP32 pullback by at most 1e-5 of its largest entry (regrouped FP64 partial sums, FP32
output); AO values agree to 1e-6 of the largest |P| (another kernel, same formulas)."""
import numpy as np
import pytest
import torch

import aok
from aok import autotune, configs
from helpers import dev_f, dev_rows, log, rel_max

pytestmark = pytest.mark.gpu


def test_tune_overlap_uneven_basis(dev, tmp_path):
    d = configs.c2(2, B=4096)
    cb = d["basis"]
    spec, _, _ = cb.ao_maps()
    R = aok.GaussianTypeRadials(dev_f(cb.zeta, torch.float32, dev), dev_f(cb.D, torch.float32, dev), spec, None,
                                tuple(range(1, cb.NZ + 1)))
    orb = aok.AtomicOrbitals(R, max(cb.ls), length_unit="bohr")
    ps = orb.static_params()
    XA = (dev_rows(d["posA"], torch.float32, dev), torch.as_tensor(d["sidxA"], device=dev))
    XB = (dev_rows(d["posB"], torch.float32, dev), torch.as_tensor(d["sidxB"], device=dev))
    ctx = aok.Context(dev.index or 0)
    S0, pb0 = aok.rrule_batch_overlap(orb, XA, XB, ps, ctx=ctx)
    g = torch.Generator(device=dev).manual_seed(3)
    dS = torch.randn(tuple(reversed(S0.shape)), generator=g, device=dev).permute(2, 1, 0)
    d0 = pb0(dS)[4]
    table = autotune.TuneTable(str(tmp_path / "t.json"))
    log("autotune C2 B=4096")
    knobs, ms, results = autotune.tune_overlap(orb, XA, XB, ps, dS, ctx=ctx, table=table,
                                               space={"ovl_cap": [32, 40, 80], "ovl_split": [1, 2]})
    assert len(results) == 6 and ms == min(r[1] for r in results) and ms > 0
    assert {k: ctx.tune_get(k) for k in knobs} == knobs
    assert ctx.blocking
    table.save()
    assert autotune.TuneTable(table.path).get(next(iter(table.entries))) == knobs
    S1, pb1 = aok.rrule_batch_overlap(orb, XA, XB, ps, ctx=ctx)
    d1 = pb1(dS)[4]
    assert torch.equal(S0, S1)
    for k in ("zeta", "D"):
        assert rel_max(d1["Rnl"][k].cpu().numpy(), d0["Rnl"][k].cpu().numpy()) < 1e-5
    ctx.close()


def test_tune_ao(dev, tmp_path):
    rng = np.random.default_rng(4)
    orb = aok.gaussian_orbitals(3, 2, length_unit="bohr", K=3, dtype=torch.float32, nspecies=2, device=dev)
    X = torch.as_tensor(rng.normal(0, 2, size=(50000, 3)), dtype=torch.float32, device=dev)
    sidx = torch.as_tensor(rng.integers(1, 3, 50000), device=dev)
    ctx = aok.Context(dev.index or 0)
    P0 = aok.evaluate(orb, X, sidx=sidx, ctx=ctx)
    log("autotune AO nX=50000")
    knobs, ms, results = autotune.tune_ao(orb, X, sidx=sidx, ctx=ctx, save=False,
                                          space={"ao_ppl": [1, 4], "ao_grid": [0, 1024]})
    assert len(results) == 4 and ms > 0
    P1 = aok.evaluate(orb, X, sidx=sidx, ctx=ctx)
    assert rel_max(P1.cpu().numpy(), P0.cpu().numpy()) < 1e-6
    ctx.close()
