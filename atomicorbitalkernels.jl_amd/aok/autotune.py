"""Autotuner for libaok's launch-shape knobs (aok_tune; no reference counterpart: the
reference has no tunables).

It times every candidate setting on the caller's own workload -- stream-ordered steps on
the context stream, wall clock between two synchronisations, best of ``trials`` runs of
``steps`` steps -- keeps the fastest, leaves it set on the context, and records it in a
JSON table keyed by a workload signature, so that a later process applies the choice
without timing again (``TuneTable.apply``).

Knobs (include/aok.h, aok_tune):
  * 2-centre overlaps: ``ovl_cap`` (LDS stripe capacity per lane, i.e. how many bra
    shells share one (tile, chunk) workgroup), ``ovl_split`` (light chunks of bases with
    f/g shells in the l <= 2 kernel instance or not), ``ovl_order`` (tile-major or
    XCD-grouped workgroup order);
  * AO evaluation: ``ao_ppl`` (consecutive points per lane), ``ao_grid`` (persistent grid).

The 2-centre knobs never change S (each element is computed by one lane in a fixed
order); ``ovl_cap`` regroups the pullback's FP64 partial sums, which can move its last
bits.  ``ao_ppl`` > 1 runs the same formulas in another kernel, so AO values may differ
in the last bits.
"""
from __future__ import annotations

import itertools
import json
import math
import os
import time

import torch

from . import _lib

OVL_SPACE = {"ovl_cap": [24, 32, 40, 56, 80, 112], "ovl_split": [1, 2], "ovl_order": [1, 2]}
AO_SPACE = {"ao_ppl": [1, 2, 4], "ao_grid": [0, 1024, 2048, 4096]}


def default_path() -> str:
    return os.environ.get("AOK_AUTOTUNE_PATH",
                          os.path.join(os.path.expanduser("~"), ".cache", "aok", "autotune.json"))


class TuneTable:
    """Tuned knob settings per workload signature, persisted as JSON:
    ``{signature: {"knobs": {key: value}, "ms": best step ms, "candidates": n}}``."""

    def __init__(self, path: str | None = None):
        self.path = path or default_path()
        self.entries: dict = {}
        if os.path.exists(self.path):
            with open(self.path) as f:
                self.entries = json.load(f)

    def get(self, sig: str):
        e = self.entries.get(sig)
        return dict(e["knobs"]) if e else None

    def put(self, sig: str, knobs: dict, ms: float, candidates: int):
        self.entries[sig] = {"knobs": {k: int(v) for k, v in knobs.items()}, "ms": float(ms),
                             "candidates": int(candidates)}

    def save(self):
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.entries, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def apply(self, ctx, sig: str) -> bool:
        """Set the recorded knobs of ``sig`` on ``ctx``; False if there is no record."""
        knobs = self.get(sig)
        if knobs is None:
            return False
        for k, v in knobs.items():
            ctx.tune(k, v)
        return True


def _device_name(device) -> str:
    try:
        return torch.cuda.get_device_properties(device).gcnArchName.split(":")[0]
    except Exception:  # pragma: no cover - property missing on some builds
        return torch.cuda.get_device_name(device)


def _bucket(n: int) -> int:
    return 0 if n <= 0 else int(round(math.log2(n)))


def overlap_signature(ls, K: int, NZ: int, B: int, dtype, device) -> str:
    """Signature of a 2-centre workload: shell l's, primitive slots, species, batch size
    to the nearest power of two, precision, GPU."""
    return (f"ovl2c|{str(dtype).replace('torch.', '')}|ls={','.join(str(int(l)) for l in ls)}|K={K}|NZ={NZ}"
            f"|B~2^{_bucket(B)}|{_device_name(device)}")


def ao_signature(nRad: int, K: int, NZ: int, nOrb: int, L: int, nX: int, dtype, grad: bool, device) -> str:
    return (f"ao|{str(dtype).replace('torch.', '')}|nRad={nRad}|K={K}|NZ={NZ}|nOrb={nOrb}|L={L}|grad={int(grad)}"
            f"|nX~2^{_bucket(nX)}|{_device_name(device)}")


def _time(ctx, fn, steps: int, trials: int) -> float:
    fn()  # warm-up: plans, workspaces
    ctx.sync()
    best = math.inf
    for _ in range(trials):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.sync()
        torch.cuda.synchronize()
        best = min(best, (time.perf_counter() - ts) * 1e3 / steps)
    return best


def _search(ctx, fn, space: dict, steps: int, trials: int):
    keys = list(space)
    saved = {k: ctx.tune_get(k) for k in keys}
    was_blocking = ctx.blocking
    results = []
    try:
        ctx.set_blocking(False)
        for vals in itertools.product(*(space[k] for k in keys)):
            knobs = dict(zip(keys, vals))
            for k, v in knobs.items():
                ctx.tune(k, v)
            results.append((knobs, _time(ctx, fn, steps, trials)))
    except Exception:
        for k, v in saved.items():
            ctx.tune(k, v)
        raise
    finally:
        ctx.set_blocking(was_blocking)
    best = min(results, key=lambda r: r[1])
    for k, v in best[0].items():
        ctx.tune(k, v)
    return best, results


def tune_overlap(orb, XA, XB, ps, dS=None, *, ctx=None, space=None, steps: int = 3, trials: int = 2,
                 table: TuneTable | None = None, save: bool = True):
    """Time the rrule(batch_overlap) step (forward S and the (zeta, D) pullback of ``dS``,
    random if None) of this batch for every combination in ``space`` (default OVL_SPACE),
    set the fastest on ``ctx`` and record it.  Returns (best knobs, best ms, [(knobs, ms)])."""
    from .basis import compile_struct
    from .overlap import _pair_inputs, rrule_batch_overlap

    ls, _, off = compile_struct(orb)
    pA, _, _, _ = _pair_inputs(XA, XB, orb.Rnl.zlist)
    zeta = ps["Rnl"]["zeta"]
    dt = torch.promote_types(pA.dtype, torch.promote_types(zeta.dtype, ps["Rnl"]["D"].dtype))
    ctx = ctx or _lib.context(pA.device.index or 0)
    B, N = pA.shape[0], off[-1]
    if dS is None:
        g = torch.Generator(device=pA.device).manual_seed(0)
        dS = torch.randn((B, N, N), generator=g, device=pA.device, dtype=dt).permute(2, 1, 0)

    def step():
        _, pb = rrule_batch_overlap(orb, XA, XB, ps, ctx=ctx)
        pb(dS)

    (knobs, ms), results = _search(ctx, step, dict(space or OVL_SPACE), steps, trials)
    sig = overlap_signature(ls, zeta.shape[1], zeta.shape[2], B, dt, pA.device)
    if save:
        table = table or TuneTable()
        table.put(sig, knobs, ms, len(results))
        table.save()
    elif table is not None:
        table.put(sig, knobs, ms, len(results))
    return knobs, ms, results


def tune_ao(orb, X, ps=None, st=None, *, sidx=None, grad: bool = False, ctx=None, space=None, steps: int = 3,
            trials: int = 2, table: TuneTable | None = None, save: bool = True):
    """Time evaluate (or evaluate_ed with ``grad``) on X (species ``sidx`` as evaluate
    takes them) for every combination in ``space`` (default AO_SPACE), set the fastest on
    ``ctx`` and record it.  Returns (best knobs, best ms, [(knobs, ms)])."""
    from .orbitals import _prep, _resolve, evaluate, positions

    ps_, st_ = _resolve(orb, ps, st)
    Rs = positions(X)
    _, zeta, _, dt, nRad, K, NZ, _, _, _, nOrb, L = _prep(orb, Rs, ps_, st_)
    ctx = ctx or _lib.context(Rs.device.index or 0)

    def step():
        evaluate(orb, X, ps, st, sidx=sidx, grad=grad, ctx=ctx)

    sp = dict(space or AO_SPACE)
    if grad:
        sp.pop("ao_ppl", None)  # points per lane applies to the values-only kernel
    (knobs, ms), results = _search(ctx, step, sp, steps, trials)
    sig = ao_signature(nRad, K, NZ, nOrb, L, Rs.shape[0], dt, grad, Rs.device)
    if save:
        table = table or TuneTable()
        table.put(sig, knobs, ms, len(results))
        table.save()
    elif table is not None:
        table.put(sig, knobs, ms, len(results))
    return knobs, ms, results
