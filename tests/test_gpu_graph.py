"""HIP-graph capture of the 2-centre rrule step (aok_graph_*, aok_ovl2c_plan_update and the
Python OverlapStepGraph; no reference counterpart).

Every replay is checked bit for bit against the eager plan path on the same inputs, with
the parameters and the cotangent changed in place between replays and eager work and
synchronisations in between (a replay that reused stale records, a stale workspace or a
recorded clear would differ)."""
import numpy as np
import pytest
import torch

import aok
from aok import configs
from helpers import dev_f, dev_rows, log

pytestmark = pytest.mark.gpu


def _orb(cb, dtype, dev):
    spec, _, _ = cb.ao_maps()
    R = aok.GaussianTypeRadials(dev_f(cb.zeta, dtype, dev), dev_f(cb.D, dtype, dev), spec, None,
                                tuple(range(1, cb.NZ + 1)))
    return aok.AtomicOrbitals(R, max(cb.ls), length_unit="bohr")


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_step_graph_replays_match_eager(dev, dtype):
    B = 3000
    d = configs.c1(1, B=B)
    cb = d["basis"]
    orb = _orb(cb, dtype, dev)
    XA = (dev_rows(d["posA"], dtype, dev), torch.as_tensor(d["sidxA"], device=dev))
    XB = (dev_rows(d["posB"], dtype, dev), torch.as_tensor(d["sidxB"], device=dev))
    log(f"graph step C1 B={B} dtype={dtype}")
    g = aok.OverlapStepGraph(orb, XA, XB, orb.static_params())
    gen = torch.Generator(device=dev).manual_seed(5)
    z0 = g.params["Rnl"]["zeta"].clone()
    for it in range(4):
        # in-place parameter step and a new cotangent
        g.params["Rnl"]["zeta"].copy_(z0 * (1.0 + 0.02 * it))
        g.params["Rnl"]["D"].mul_(1.0 - 0.01 * it)
        g.dS.copy_(torch.randn(g.dS.shape, generator=gen, device=dev, dtype=dtype))
        torch.cuda.synchronize()
        S, dps = g.replay()
        torch.cuda.synchronize()
        ps = {"Rnl": {"zeta": g.params["Rnl"]["zeta"].clone(), "D": g.params["Rnl"]["D"].clone()}, "Ylm": {}}
        Se, pb = aok.rrule_batch_overlap(orb, XA, XB, ps)
        de = pb(g.dS.clone())[4]
        log(f"  replay {it}")
        assert torch.equal(S, Se)
        assert torch.equal(dps["Rnl"]["zeta"], de["Rnl"]["zeta"])
        assert torch.equal(dps["Rnl"]["D"], de["Rnl"]["D"])
    g.close()


def test_generic_capture_and_errors(dev):
    """Any calls of a context can be recorded; a call that would grow the workspace
    inside the capture fails it cleanly and the context keeps working."""
    d = configs.c1(1, B=500)
    cb = d["basis"]
    orb = _orb(cb, torch.float64, dev)
    ctx = aok.Context.side_stream(dev.index or 0)
    XA = (dev_rows(d["posA"], torch.float64, dev), torch.as_tensor(d["sidxA"], device=dev))
    XB = (dev_rows(d["posB"], torch.float64, dev), torch.as_tensor(d["sidxB"], device=dev))
    ps = orb.static_params()
    log("graph generic capture")
    basis = aok.compile_basis(orb)
    out = torch.empty((500, cb.N, cb.N), dtype=torch.float64, device=dev).permute(2, 1, 0)
    aok.batch_overlap_(out, basis, XA, XB, ctx=ctx)  # eager: sizes the workspace
    ref = out.clone()
    out.zero_()
    torch.cuda.synchronize()
    with ctx.capture() as gr:
        aok.batch_overlap_(out, basis, XA, XB, ctx=ctx)
    gr.keep(out, basis, XA, XB)
    assert torch.count_nonzero(out) == 0  # recorded, not run
    for _ in range(2):
        out.zero_()
        torch.cuda.synchronize()
        gr.replay()
        assert torch.equal(out, ref)
    gr.close()
    # a larger batch needs a larger workspace: refused inside the capture
    d2 = configs.c1(1, B=200000)
    XA2 = (dev_rows(d2["posA"], torch.float64, dev), torch.as_tensor(d2["sidxA"], device=dev))
    XB2 = (dev_rows(d2["posB"], torch.float64, dev), torch.as_tensor(d2["sidxB"], device=dev))
    out2 = torch.empty((200000, cb.N, cb.N), dtype=torch.float64, device=dev).permute(2, 1, 0)
    with pytest.raises(aok.AOKArgumentError, match="capture"):
        with ctx.capture():
            aok.batch_overlap_(out2, basis, XA2, XB2, ctx=ctx)
    # still usable eagerly
    aok.batch_overlap_(out, basis, XA, XB, ctx=ctx)
    assert torch.equal(out, ref)
    # destroying the context first detaches its graphs: replay refuses, destroy works
    with ctx.capture() as gr2:
        aok.batch_overlap_(out, basis, XA, XB, ctx=ctx)
    gr2.keep(out, basis, XA, XB)
    ctx.close()
    with pytest.raises(aok.AOKArgumentError):
        gr2.replay()
    gr2.close()
    del ps
