"""GPU tests of the stream-ordered execution mode (aok_set_blocking(ctx, 0), aok_sync,
aok_kernel_times): calls return once enqueued on the context stream (torch's default
stream), torch work queued after a call sees its outputs, results are bit-identical to
blocking mode, device-detected errors are latched until aok_sync, and the timed-launch
record returns every dominant kernel in launch order."""
import numpy as np
import pytest
import torch

import aok
from aok import configs
from helpers import dev_f, dev_rows, log

pytestmark = pytest.mark.gpu


def _orb(cb, dtype, dev, zlist=None):
    spec, _, _ = cb.ao_maps()
    R = aok.GaussianTypeRadials(dev_f(cb.zeta, dtype, dev), dev_f(cb.D, dtype, dev), spec, None,
                                tuple(zlist or range(1, cb.NZ + 1)))
    return aok.AtomicOrbitals(R, max(cb.ls), length_unit="bohr")


def _c1(dev, B, dtype):
    d = configs.c1(1, B=B)
    cb = d["basis"]
    XA = (dev_rows(d["posA"], dtype, dev), torch.as_tensor(d["sidxA"], device=dev))
    XB = (dev_rows(d["posB"], dtype, dev), torch.as_tensor(d["sidxB"], device=dev))
    return cb, XA, XB


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_stream_ordered_equals_blocking(dev, dtype):
    cb, XA, XB = _c1(dev, 3000, dtype)
    orb = _orb(cb, dtype, dev)
    ps = orb.static_params()
    g = torch.Generator(device=dev).manual_seed(3)
    dS = torch.randn((3000, cb.N, cb.N), generator=g, device=dev, dtype=dtype).permute(2, 1, 0)
    X = dev_rows(np.random.default_rng(1).normal(0, 1.5, (5000, 3)), dtype, dev)
    sidx = torch.as_tensor(np.random.default_rng(2).integers(1, cb.NZ + 1, 5000), device=dev)
    log(f"stream-ordered vs blocking dtype={dtype}")
    ctx = aok.Context(0)
    assert ctx.blocking

    def run():
        S, pb = aok.rrule_batch_overlap(orb, XA, XB, ps, None, ctx=ctx)
        dps = pb(dS)[4]["Rnl"]
        P, pbe = aok.rrule_evaluate(orb, X, sidx=sidx, ctx=ctx)
        _, _, dX, dpa, _ = pbe(P * 0.5 + 1.0)  # torch work on the outputs, queued in between
        return S, dps["zeta"], dps["D"], P, dX, dpa["Rnl"]["zeta"], dpa["Rnl"]["D"]

    ref = [t.clone() for t in run()]
    ctx.set_blocking(False)
    assert not ctx.blocking
    outs = [run() for _ in range(3)]  # back-to-back, no host waits
    ctx.sync()
    ctx.set_blocking(True)
    for o in outs:
        for a, b in zip(o, ref):
            assert torch.equal(a, b)
    ctx.close()


def test_latched_species_error(dev):
    cb, XA, XB = _c1(dev, 64, torch.float64)
    orb = _orb(cb, torch.float64, dev)
    log("stream-ordered latched species error")
    ctx = aok.Context(0)
    ctx.set_blocking(False)
    bad = (XA[0][:2], torch.tensor([1, 4], device=dev))
    aok.batch_overlap(orb, bad, (XB[0][:2], torch.tensor([1, 1], device=dev)), ctx=ctx)  # enqueued, no raise
    S = aok.batch_overlap(orb, XA, XB, ctx=ctx)  # a good call after it
    with pytest.raises(aok.AOKSpeciesError):
        ctx.sync()
    ctx.sync()  # the latch was cleared
    S_ref = aok.batch_overlap(orb, XA, XB)
    assert torch.equal(S, S_ref)
    # switching back to blocking reports what was latched
    aok.batch_overlap(orb, bad, (XB[0][:2], torch.tensor([1, 1], device=dev)), ctx=ctx)
    with pytest.raises(aok.AOKSpeciesError):
        ctx.set_blocking(True)
    assert ctx.blocking
    with pytest.raises(aok.AOKSpeciesError):  # blocking mode raises at the call again
        aok.batch_overlap(orb, bad, (XB[0][:2], torch.tensor([1, 1], device=dev)), ctx=ctx)
    ctx.close()


def test_kernel_times_record(dev):
    cb, XA, XB = _c1(dev, 2000, torch.float32)
    orb = _orb(cb, torch.float32, dev)
    ps = orb.static_params()
    log("stream-ordered kernel-time record")
    ctx = aok.Context(0)
    ctx.set_timing(True)
    ctx.set_blocking(False)
    for _ in range(4):
        S, pb = aok.rrule_batch_overlap(orb, XA, XB, ps, None, ctx=ctx)
        pb(torch.ones_like(S))
    assert ctx.last_kernel_name() == "ovl2c_bwd_fast" and ctx.last_kernel_ms() > 0
    rec = ctx.kernel_times()
    assert [n for n, _ in rec] == ["ovl2c_fwd_fast", "ovl2c_bwd_fast"] * 4
    assert all(ms > 0 for _, ms in rec)
    assert ctx.kernel_times() == []
    ctx.sync()
    ctx.set_blocking(True)
    ctx.close()
    y = torch.ones(10, device=dev)
    torch.cuda.synchronize()
    assert float(y.sum()) == 10.0
