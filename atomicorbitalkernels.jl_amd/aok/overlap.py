"""H2 host layer: batched 2-centre overlaps and their (zeta, D) pullback.

Mirrors src/integrals/orbital/overlap_2c.jl:171-235 and overlap_2c_grad.jl:22-318:
  batch_overlap(basis::CartesianGTOBasis, XA, XB)            -> out (N, N, B)
  batch_overlap_(out, basis, XA, XB)                          (the mutating ``!`` form)
  batch_overlap(orb::AtomicOrbitals, XA, XB[, ps, st])        (compile-on-call / Lux form)
  pullback_ps(dS, orb, XA, XB, ps, st)                        -> {'Rnl': {'zeta', 'D'}, 'Ylm': {}}
  rrule_batch_overlap(orb, XA, XB, ps, st)                    -> (S, pullback)
  overlap_pb_coef_alpha(dS, basis, XA, XB)                    -> (dcoef, dalpha)
``XA``/``XB`` are ``PStates`` (positions in the basis' length unit + species labels),
or ``(positions, sidx)`` tuples for the kernel-facing layer.  ``mode`` selects the
fast kernels (default) or the reference-decomposition ``"plain"`` kernels.
"""
from __future__ import annotations

import torch

from . import _lib
from .basis import (AtomicOrbitals, CartesianGTOBasis, _dtype_code, as_fortran, compile_basis, compile_coef,
                    compile_coef_pb, compile_struct, fortran_empty)
from . import compiled as _cb
from .compiled import CompiledBasis
from .orbitals import NO_TANGENT, PStates


def _flags(mode: str) -> int:
    if mode == "fast":
        return _lib.AOK_FAST
    if mode == "plain":
        return _lib.AOK_PLAIN
    raise _lib.AOKArgumentError(f"unknown mode {mode!r}")


def _unpack(X, zlist, what):
    """_cgto_posmat + _cgto_sidx (overlap_2c.jl:139-158)."""
    if isinstance(X, PStates):
        lut = {z: i + 1 for i, z in enumerate(zlist)}
        idx = []
        for s in X.species:
            s = s.item() if hasattr(s, "item") else s
            if s not in lut:
                raise _lib.AOKSpeciesError(f"species {s} not in basis species list {zlist}")
            idx.append(lut[s])
        pos = X.positions.contiguous()
        return pos, torch.tensor(idx, dtype=torch.int64, device=pos.device)
    if isinstance(X, tuple) and len(X) == 2:
        pos, sidx = X
        return pos.contiguous(), (sidx.contiguous() if sidx is not None else None)
    raise _lib.AOKArgumentError(f"{what} must be PStates or a (positions, sidx) tuple")


def _sidx_ptr(sidx):
    if sidx is None:
        return None, _lib.AOK_I64
    if sidx.dtype == torch.int64:
        return sidx.data_ptr(), _lib.AOK_I64
    if sidx.dtype == torch.int32:
        return sidx.data_ptr(), _lib.AOK_I32
    raise _lib.AOKArgumentError("species indices must be int64 or int32")


def _pair_inputs(XA, XB, zlist):
    pA, sA = _unpack(XA, zlist, "XA")
    pB, sB = _unpack(XB, zlist, "XB")
    if pA.shape[0] != pB.shape[0]:  # overlap_2c.jl:177-178
        raise _lib.AOKArgumentError(f"XA and XB must have equal length; got {pA.shape[0]} and {pB.shape[0]}")
    if pA.shape[1:] != (3,) or pB.shape[1:] != (3,):
        raise _lib.AOKArgumentError("positions must have shape (B, 3)")
    return pA, sA, pB, sB


def batch_overlap_(out, basis, XA, XB, *, mode="fast", ctx=None):
    """batch_overlap!(out, basis, XA, XB) (overlap_2c.jl:171-197); ``out`` is (N, N, B),
    column-major, and sets the kernel precision.  A CompiledBasis takes the geometry-in
    form (kernels_2c.jl:158-168) with unit-carrying (3, B) positions."""
    if isinstance(basis, CompiledBasis):
        return _cb.batch_overlap_(out, basis, XA, XB, mode=mode, ctx=ctx)
    pA, sA, pB, sB = _pair_inputs(XA, XB, basis.zlist)
    B = pA.shape[0]
    N = basis.nbf_total
    if tuple(out.shape) != (N, N, B):
        raise _lib.AOKArgumentError(f"out must be {(N, N, B)}, got {tuple(out.shape)}")
    if not out.permute(2, 1, 0).is_contiguous():
        raise _lib.AOKArgumentError("out must be column-major contiguous")
    dt = out.dtype
    pA, pB = pA.to(dt), pB.to(dt)
    coef = as_fortran(basis.coef.to(dt))
    zeta = as_fortran(basis.zeta.to(dt))
    _lib.require_device(out, pA, pB, sA, sB, coef, zeta)
    ctx = ctx or _lib.context(out.device.index or 0)
    a, it = _sidx_ptr(sA)
    b, _ = _sidx_ptr(sB)
    ns, K, NZ = zeta.shape
    ctx.call("aok_ovl2c", _dtype_code(dt), _flags(mode), pA.data_ptr(), pB.data_ptr(), a, b, it, B, ns, K, NZ,
             _lib.i64_array(basis.ls), coef.data_ptr(), zeta.data_ptr(), float(basis.lengthscale), out.data_ptr())
    return out


def _out_tensor(N, B, dtype, device):
    return torch.empty((B, N, N), dtype=dtype, device=device).permute(2, 1, 0)


def batch_overlap(basis, XA, XB, ps=None, st=None, *, mode="fast", dtype=None, ctx=None):
    """batch_overlap(basis::CartesianGTOBasis, XA, XB) or batch_overlap(orb, XA, XB[, ps, st]).
    The output precision follows the positions (promoted with (zeta, D) in the Lux form).
    A CompiledBasis takes the geometry-in form (kernels_2c.jl:170-184), FP64 by default."""
    if isinstance(basis, CompiledBasis):
        return _cb.batch_overlap(basis, XA, XB, dtype=dtype or torch.float64, mode=mode, ctx=ctx)
    if isinstance(basis, CartesianGTOBasis):
        pA = XA.positions if isinstance(XA, PStates) else XA[0]
        dt = dtype or pA.dtype
        N = basis.nbf_total
        out = _out_tensor(N, pA.shape[0], dt, pA.device)
        return batch_overlap_(out, basis, XA, XB, mode=mode, ctx=ctx)
    if not isinstance(basis, AtomicOrbitals):
        raise _lib.AOKArgumentError("basis must be a CartesianGTOBasis or an AtomicOrbitals")
    orb = basis
    if ps is None:  # batch_overlap(orb, XA, XB) = batch_overlap(compile_basis(orb), XA, XB)
        return batch_overlap(compile_basis(orb), XA, XB, mode=mode, dtype=dtype, ctx=ctx)
    # Lux form: coef recompiled from ps on the device (overlap_2c_grad.jl:41-50)
    if orb.Rnl.decay != _lib.AOK_DECAY_GAUSSIAN:
        raise _lib.AOKArgumentError("the overlap kernels integrate plain Gaussians")
    ls, nbf, off = compile_struct(orb)
    pA, sA, pB, sB = _pair_inputs(XA, XB, orb.Rnl.zlist)
    zeta = ps["Rnl"]["zeta"]
    D = ps["Rnl"]["D"]
    dt = dtype or torch.promote_types(pA.dtype, torch.promote_types(zeta.dtype, D.dtype))
    zeta, D = as_fortran(zeta.to(dt)), as_fortran(D.to(dt))
    pA, pB = pA.to(dt), pB.to(dt)
    B = pA.shape[0]
    _lib.require_device(pA, pB, sA, sB, zeta, D)
    out = _out_tensor(off[-1], B, dt, pA.device)
    ctx = ctx or _lib.context(pA.device.index or 0)
    a, it = _sidx_ptr(sA)
    b, _ = _sidx_ptr(sB)
    ns, K, NZ = zeta.shape
    ctx.call("aok_ovl2c_ps", _dtype_code(dt), _flags(mode), pA.data_ptr(), pB.data_ptr(), a, b, it, B, ns, K, NZ,
             _lib.i64_array(ls), zeta.data_ptr(), D.data_ptr(), float(orb.lengthscale), out.data_ptr())
    return out


def overlap_pb_coef_alpha(dS, basis: CartesianGTOBasis, XA, XB, *, mode="fast", ctx=None):
    """_overlap_pb_coef_alpha (overlap_2c_grad.jl:218-242) -> (dcoef, dalpha), (ns, K, NZ)."""
    pA, sA, pB, sB = _pair_inputs(XA, XB, basis.zlist)
    dt = torch.promote_types(dS.dtype, basis.coef.dtype)
    dS = as_fortran(dS.to(dt))
    B = pA.shape[0]
    N = basis.nbf_total
    if tuple(dS.shape) != (N, N, B):
        raise _lib.AOKArgumentError(f"dS must be {(N, N, B)}")
    pA, pB = pA.to(dt), pB.to(dt)
    coef = as_fortran(basis.coef.to(dt))
    zeta = as_fortran(basis.zeta.to(dt))
    ns, K, NZ = zeta.shape
    dc = fortran_empty(zeta.shape, dt, dS.device)
    _lib.require_device(dS, pA, pB, sA, sB, coef, zeta)
    da = fortran_empty(zeta.shape, dt, dS.device)
    ctx = ctx or _lib.context(dS.device.index or 0)
    a, it = _sidx_ptr(sA)
    b, _ = _sidx_ptr(sB)
    ctx.call("aok_ovl2c_pullback", _dtype_code(dt), _flags(mode), dS.data_ptr(), pA.data_ptr(), pB.data_ptr(), a, b,
             it, B, ns, K, NZ, _lib.i64_array(basis.ls), coef.data_ptr(), zeta.data_ptr(), float(basis.lengthscale),
             dc.data_ptr(), da.data_ptr())
    return dc, da


def overlap_pullback_x(dS, basis: CartesianGTOBasis, XA, XB, *, ctx=None):
    """Position cotangents of batch_overlap(basis, XA, XB) (aok_ovl2c_pullback_x; no
    reference counterpart) -> (dXA, dXB), each (B, 3), in the basis' length unit;
    dXB = -dXA."""
    pA, sA, pB, sB = _pair_inputs(XA, XB, basis.zlist)
    dt = torch.promote_types(dS.dtype, basis.coef.dtype)
    dS = as_fortran(dS.to(dt))
    B = pA.shape[0]
    N = basis.nbf_total
    if tuple(dS.shape) != (N, N, B):
        raise _lib.AOKArgumentError(f"dS must be {(N, N, B)}")
    pA, pB = pA.to(dt), pB.to(dt)
    coef = as_fortran(basis.coef.to(dt))
    zeta = as_fortran(basis.zeta.to(dt))
    ns, K, NZ = zeta.shape
    _lib.require_device(dS, pA, pB, sA, sB, coef, zeta)
    dA = torch.empty((B, 3), dtype=dt, device=dS.device)
    dB = torch.empty((B, 3), dtype=dt, device=dS.device)
    ctx = ctx or _lib.context(dS.device.index or 0)
    a, it = _sidx_ptr(sA)
    b, _ = _sidx_ptr(sB)
    ctx.call("aok_ovl2c_pullback_x", _dtype_code(dt), _lib.AOK_FAST, dS.data_ptr(), pA.data_ptr(), pB.data_ptr(), a,
             b, it, B, ns, K, NZ, _lib.i64_array(basis.ls), coef.data_ptr(), zeta.data_ptr(),
             float(basis.lengthscale), dA.data_ptr(), dB.data_ptr())
    return dA, dB


def pullback_ps(dS, orb: AtomicOrbitals, XA, XB, ps, st=None, *, mode="fast", ctx=None):
    """pullback_ps(dS, orb, XA, XB, ps, st) (overlap_2c_grad.jl:296-304)."""
    ls, _, off = compile_struct(orb)
    pA, sA, pB, sB = _pair_inputs(XA, XB, orb.Rnl.zlist)
    zeta, D = ps["Rnl"]["zeta"], ps["Rnl"]["D"]
    dt = torch.promote_types(dS.dtype, torch.promote_types(zeta.dtype, D.dtype))
    zeta, D = as_fortran(zeta.to(dt)), as_fortran(D.to(dt))
    dS = as_fortran(dS.to(dt))
    B = pA.shape[0]
    if tuple(dS.shape) != (off[-1], off[-1], B):
        raise _lib.AOKArgumentError(f"dS must be {(off[-1], off[-1], B)}")
    pA, pB = pA.to(dt), pB.to(dt)
    ns, K, NZ = zeta.shape
    dz = fortran_empty(zeta.shape, dt, dS.device)
    _lib.require_device(dS, pA, pB, sA, sB, zeta, D)
    dD = fortran_empty(zeta.shape, dt, dS.device)
    ctx = ctx or _lib.context(dS.device.index or 0)
    a, it = _sidx_ptr(sA)
    b, _ = _sidx_ptr(sB)
    ctx.call("aok_ovl2c_pullback_ps", _dtype_code(dt), _flags(mode), dS.data_ptr(), pA.data_ptr(), pB.data_ptr(), a,
             b, it, B, ns, K, NZ, _lib.i64_array(ls), zeta.data_ptr(), D.data_ptr(), float(orb.lengthscale),
             dz.data_ptr(), dD.data_ptr())
    return {"Rnl": {"zeta": dz, "D": dD}, "Ylm": {}}


class Ovl2cPlan:
    """A 2-centre plan (aok_ovl2c_plan_*): the species-pair bucketing of one batch and the
    primitive-pair records of one parameter set, shared by a forward and its pullback.
    Holds the species-index and parameter tensors it was built from."""

    def __init__(self, ls, sA, sB, B, p1, p2, params, lengthscale, dtype, mode="fast", ctx=None):
        _lib.require_device(sA, sB, p1, p2)
        self.ctx = ctx or _lib.context(p1.device.index or 0)
        self._keep = (sA, sB, p1, p2)
        self.dtype = dtype
        self.B = B
        self.shape = tuple(p1.shape)
        a, it = _sidx_ptr(sA)
        b, _ = _sidx_ptr(sB)
        ns, K, NZ = p1.shape
        h = _lib.c_void_p()
        _lib._order_inputs(self.ctx)
        rc = self.ctx.lib.aok_ovl2c_plan_create(self.ctx.handle, _dtype_code(dtype), _flags(mode), a, b, it, B, ns, K,
                                                NZ, _lib.i64_array(ls), params, p1.data_ptr(), p2.data_ptr(),
                                                float(lengthscale), _lib.byref(h))
        self.ctx.check(rc, "aok_ovl2c_plan_create")
        self.handle = h

    def forward(self, pA, pB, out):
        _lib.require_device(pA, pB, out)
        _lib._order_inputs(self.ctx)
        rc = self.ctx.lib.aok_ovl2c_plan_forward(self.handle, pA.data_ptr(), pB.data_ptr(), out.data_ptr())
        self.ctx.check(rc, "aok_ovl2c_plan_forward")
        return out

    def update(self):
        """Rebuild the bucketing and primitive-pair records from the species-index and
        parameter tensors this plan holds, as they are now (aok_ovl2c_plan_update; no
        reference counterpart)."""
        _lib._order_inputs(self.ctx)
        self.ctx.check(self.ctx.lib.aok_ovl2c_plan_update(self.handle), "aok_ovl2c_plan_update")

    def pullback(self, dS, pA, pB, out=None):
        _lib.require_device(dS, pA, pB)
        if out is None:
            d1 = fortran_empty(self.shape, self.dtype, dS.device)
            d2 = fortran_empty(self.shape, self.dtype, dS.device)
        else:
            d1, d2 = out
        _lib._order_inputs(self.ctx)
        rc = self.ctx.lib.aok_ovl2c_plan_pullback(self.handle, dS.data_ptr(), pA.data_ptr(), pB.data_ptr(),
                                                  d1.data_ptr(), d2.data_ptr())
        self.ctx.check(rc, "aok_ovl2c_plan_pullback")
        return d1, d2

    def pullback_x(self, dS, pA, pB):
        """Position cotangents (dA, dB), each (B, 3), of dS (aok_ovl2c_plan_pullback_x)."""
        _lib.require_device(dS, pA, pB)
        dA = torch.empty((self.B, 3), dtype=self.dtype, device=dS.device)
        dB = torch.empty((self.B, 3), dtype=self.dtype, device=dS.device)
        _lib._order_inputs(self.ctx)
        rc = self.ctx.lib.aok_ovl2c_plan_pullback_x(self.handle, dS.data_ptr(), pA.data_ptr(), pB.data_ptr(),
                                                    dA.data_ptr(), dB.data_ptr())
        self.ctx.check(rc, "aok_ovl2c_plan_pullback_x")
        return dA, dB

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.aok_ovl2c_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


def rrule_batch_overlap(orb: AtomicOrbitals, XA, XB, ps, st=None, *, mode="fast", dtype=None, ctx=None,
                        wrt_positions=False):
    """rrule(batch_overlap, orb, XA, XB, ps, st) (overlap_2c_grad.jl:309-318) -> (S, pullback);
    pullback(dS) -> (NoTangent, NoTangent, NoTangent, NoTangent, dps, NoTangent).  Forward
    and pullback share one plan (coefficient compilation, bucketing, pair records).
    With ``wrt_positions`` (no reference counterpart: its rrule has NoTangent there) the
    third and fourth slots hold the position cotangents dXA, dXB as (B, 3) tensors in the
    input length unit (aok_ovl2c_plan_pullback_x)."""
    if orb.Rnl.decay != _lib.AOK_DECAY_GAUSSIAN:
        raise _lib.AOKArgumentError("the overlap kernels integrate plain Gaussians")
    ls, _, off = compile_struct(orb)
    pA, sA, pB, sB = _pair_inputs(XA, XB, orb.Rnl.zlist)
    zeta, D = ps["Rnl"]["zeta"], ps["Rnl"]["D"]
    dt = dtype or torch.promote_types(pA.dtype, torch.promote_types(zeta.dtype, D.dtype))
    zeta, D = as_fortran(zeta.to(dt)), as_fortran(D.to(dt))
    pA, pB = pA.to(dt), pB.to(dt)
    B = pA.shape[0]
    N = off[-1]
    plan = Ovl2cPlan(ls, sA, sB, B, zeta, D, _lib.AOK_PARAMS_PS, orb.lengthscale, dt, mode=mode,
                     ctx=ctx or _lib.context(pA.device.index or 0))
    S = plan.forward(pA, pB, _out_tensor(N, B, dt, pA.device))

    def pullback(dS):
        if tuple(dS.shape) != (N, N, B):
            raise _lib.AOKArgumentError(f"dS must be {(N, N, B)}")
        g = as_fortran(dS.to(dt))
        dz, dD = plan.pullback(g, pA, pB)
        dps = {"Rnl": {"zeta": dz, "D": dD}, "Ylm": {}}
        if wrt_positions:
            dXA, dXB = plan.pullback_x(g, pA, pB)
            return NO_TANGENT, NO_TANGENT, dXA, dXB, dps, NO_TANGENT
        return NO_TANGENT, NO_TANGENT, NO_TANGENT, NO_TANGENT, dps, NO_TANGENT

    return S, pullback


class OverlapStepGraph:
    """The rrule(batch_overlap) step of one fixed batch -- plan update (bucketing and
    primitive-pair records from the current parameters), forward S and (zeta, D) pullback
    of the cotangent ``dS`` -- recorded once into a HIP graph and replayed as one launch
    (no reference counterpart: the reference relaunches its kernels per call).

    Replays read ``params["Rnl"]["zeta"]``, ``params["Rnl"]["D"]``, ``posA``, ``posB`` and
    ``dS`` (the tensors held here) as they are at replay time and write ``S``, ``dzeta``
    and ``dD``: update them in place between replays (an optimiser step, the next
    cotangent).  The graph runs on its own context on a torch side stream unless ``ctx``
    is given."""

    def __init__(self, orb: AtomicOrbitals, XA, XB, ps, *, dtype=None, ctx=None):
        if orb.Rnl.decay != _lib.AOK_DECAY_GAUSSIAN:
            raise _lib.AOKArgumentError("the overlap kernels integrate plain Gaussians")
        ls, _, off = compile_struct(orb)
        pA, sA, pB, sB = _pair_inputs(XA, XB, orb.Rnl.zlist)
        zeta, D = ps["Rnl"]["zeta"], ps["Rnl"]["D"]
        dt = dtype or torch.promote_types(pA.dtype, torch.promote_types(zeta.dtype, D.dtype))
        dev = pA.device
        self.ctx = ctx or _lib.Context.side_stream(dev.index or 0)
        self.params = {"Rnl": {"zeta": as_fortran(zeta.to(dt)).clone(memory_format=torch.preserve_format),
                               "D": as_fortran(D.to(dt)).clone(memory_format=torch.preserve_format)}, "Ylm": {}}
        self.posA, self.posB = pA.to(dt).clone(), pB.to(dt).clone()
        B, N = pA.shape[0], off[-1]
        self.B, self.N = B, N
        z, d = self.params["Rnl"]["zeta"], self.params["Rnl"]["D"]
        self.plan = Ovl2cPlan(ls, sA, sB, B, z, d, _lib.AOK_PARAMS_PS, orb.lengthscale, dt, ctx=self.ctx)
        self.S = _out_tensor(N, B, dt, dev)
        self.dS = _out_tensor(N, B, dt, dev).zero_()
        self.dzeta = fortran_empty(tuple(z.shape), dt, dev)
        self.dD = fortran_empty(tuple(z.shape), dt, dev)
        torch.cuda.synchronize(dev)
        self._record()  # eager pass: sizes every workspace before the capture
        with self.ctx.capture() as g:
            self._record()
        g.keep(self.plan, self.params, self.posA, self.posB, self.S, self.dS, self.dzeta, self.dD, sA, sB)
        self.graph = g

    def _record(self):
        self.plan.update()
        self.plan.forward(self.posA, self.posB, self.S)
        self.plan.pullback(self.dS, self.posA, self.posB, out=(self.dzeta, self.dD))

    def replay(self):
        """One step: returns (S, dps) -- the graph's own output tensors, overwritten by the
        next replay."""
        self.graph.replay()
        return self.S, {"Rnl": {"zeta": self.dzeta, "D": self.dD}, "Ylm": {}}

    def close(self):
        self.graph.close()
        self.plan.close()


def batch_overlap_3c_(out, basis, XA, XB, XC, *, mode="fast", ctx=None):
    """batch_overlap_3c!(out, basis, posA, posB, posC) (src/integrals/kernels_3c.jl:183-200),
    species-aware: XA/XB/XC are PStates or (positions, sidx) tuples, positions in the
    basis' length unit.  ``out`` is (N, N, N, B), column-major.  A CompiledBasis takes the
    reference's own single-basis form with unit-carrying (3, B) positions."""
    if isinstance(basis, CompiledBasis):
        return _cb.batch_overlap_3c_(out, basis, XA, XB, XC, mode=mode, ctx=ctx)
    pA, sA = _unpack(XA, basis.zlist, "XA")
    pB, sB = _unpack(XB, basis.zlist, "XB")
    pC, sC = _unpack(XC, basis.zlist, "XC")
    B = pA.shape[0]
    if pB.shape[0] != B or pC.shape[0] != B:
        raise _lib.AOKArgumentError("XA, XB and XC must have equal length")
    N = basis.nbf_total
    if tuple(out.shape) != (N, N, N, B) or not out.permute(3, 2, 1, 0).is_contiguous():
        raise _lib.AOKArgumentError(f"out must be a column-major {(N, N, N, B)} tensor")
    dt = out.dtype
    pA, pB, pC = pA.to(dt), pB.to(dt), pC.to(dt)
    coef = as_fortran(basis.coef.to(dt))
    zeta = as_fortran(basis.zeta.to(dt))
    ns, K, NZ = zeta.shape
    _lib.require_device(out, pA, pB, pC, sA, sB, sC, coef, zeta)
    ctx = ctx or _lib.context(out.device.index or 0)
    a, it = _sidx_ptr(sA)
    b, _ = _sidx_ptr(sB)
    c, _ = _sidx_ptr(sC)
    ctx.call("aok_ovl3c", _dtype_code(dt), _flags(mode), pA.data_ptr(), pB.data_ptr(), pC.data_ptr(), a, b, c, it,
             B, ns, K, NZ, _lib.i64_array(basis.ls), coef.data_ptr(), zeta.data_ptr(), float(basis.lengthscale),
             out.data_ptr())
    return out


def overlap_3c_pb_coef_alpha(dT, basis: CartesianGTOBasis, XA, XB, XC, *, ctx=None):
    """(dcoef, dalpha), each (ns, K, NZ), of batch_overlap_3c(basis, XA, XB, XC) for the
    cotangent dT (N, N, N, B) (aok_ovl3c_pullback; no reference counterpart).  Every
    primitive slot receives its coefficient cotangent, padded slots included, as in the
    2-centre pullback; shells up to l = 2."""
    pA, sA = _unpack(XA, basis.zlist, "XA")
    pB, sB = _unpack(XB, basis.zlist, "XB")
    pC, sC = _unpack(XC, basis.zlist, "XC")
    B = pA.shape[0]
    if pB.shape[0] != B or pC.shape[0] != B:
        raise _lib.AOKArgumentError("XA, XB and XC must have equal length")
    N = basis.nbf_total
    dt = torch.promote_types(dT.dtype, basis.coef.dtype)
    dT = as_fortran(dT.to(dt))
    if tuple(dT.shape) != (N, N, N, B):
        raise _lib.AOKArgumentError(f"dT must be {(N, N, N, B)}")
    pA, pB, pC = pA.to(dt), pB.to(dt), pC.to(dt)
    coef = as_fortran(basis.coef.to(dt))
    zeta = as_fortran(basis.zeta.to(dt))
    ns, K, NZ = zeta.shape
    _lib.require_device(dT, pA, pB, pC, sA, sB, sC, coef, zeta)
    dc = fortran_empty(zeta.shape, dt, dT.device)
    da = fortran_empty(zeta.shape, dt, dT.device)
    ctx = ctx or _lib.context(dT.device.index or 0)
    a, it = _sidx_ptr(sA)
    b, _ = _sidx_ptr(sB)
    c, _ = _sidx_ptr(sC)
    ctx.call("aok_ovl3c_pullback", _dtype_code(dt), _lib.AOK_FAST, dT.data_ptr(), pA.data_ptr(), pB.data_ptr(),
             pC.data_ptr(), a, b, c, it, B, ns, K, NZ, _lib.i64_array(basis.ls), coef.data_ptr(), zeta.data_ptr(),
             float(basis.lengthscale), dc.data_ptr(), da.data_ptr())
    return dc, da


def rrule_batch_overlap_3c(orb: AtomicOrbitals, XA, XB, XC, ps, st=None, *, dtype=None, ctx=None):
    """3-centre overlap of an orbital basis with (zeta, D) from ps, and its pullback
    (no reference counterpart) -> (T, pullback); pullback(dT) -> (NoTangent, NoTangent,
    NoTangent, NoTangent, NoTangent, dps, NoTangent) with dps = (Rnl = (zeta, D), Ylm = ()):
    dzeta = dalpha + dzeta_norm, dD through the FP64 normalisation VJP."""
    if orb.Rnl.decay != _lib.AOK_DECAY_GAUSSIAN:
        raise _lib.AOKArgumentError("the overlap kernels integrate plain Gaussians")
    ls, nbf, off = compile_struct(orb)
    zeta, D = ps["Rnl"]["zeta"], ps["Rnl"]["D"]
    pA = XA.positions if isinstance(XA, PStates) else XA[0]
    dt = dtype or torch.promote_types(pA.dtype, torch.promote_types(zeta.dtype, D.dtype))
    zeta, D = as_fortran(zeta.to(dt)), as_fortran(D.to(dt))
    coef = compile_coef(zeta, D, ls, ctx=ctx)
    basis = CartesianGTOBasis(max(ls), len(ls), zeta.shape[1], list(ls), nbf, off, off[-1], zeta, coef,
                              orb.Rnl.zlist, orb.lengthscale)
    T = batch_overlap_3c(basis, XA, XB, XC, dtype=dt, ctx=ctx)

    def pullback(dT):
        dc, da = overlap_3c_pb_coef_alpha(dT, basis, XA, XB, XC, ctx=ctx)
        dD, dzn = compile_coef_pb(dc, zeta, D, ls, ctx=ctx)
        dps = {"Rnl": {"zeta": da + dzn, "D": dD}, "Ylm": {}}
        return NO_TANGENT, NO_TANGENT, NO_TANGENT, NO_TANGENT, NO_TANGENT, dps, NO_TANGENT

    return T, pullback


def batch_overlap_3c(basis, XA, XB, XC, *, mode="fast", dtype=None, ctx=None):
    """batch_overlap_3c(basis, posA, posB, posC) -> (N, N, N, B) (kernels_3c.jl:213-222)."""
    if isinstance(basis, CompiledBasis):
        return _cb.batch_overlap_3c(basis, XA, XB, XC, dtype=dtype or torch.float64, mode=mode, ctx=ctx)
    pA = XA.positions if isinstance(XA, PStates) else XA[0]
    dt = dtype or pA.dtype
    N = basis.nbf_total
    out = torch.empty((pA.shape[0], N, N, N), dtype=dt, device=pA.device).permute(3, 2, 1, 0)
    return batch_overlap_3c_(out, basis, XA, XB, XC, mode=mode, ctx=ctx)
