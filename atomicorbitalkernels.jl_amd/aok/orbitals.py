"""H1 host layer: evaluate / evaluate_ed / pullback_ps / rrule of ``AtomicOrbitals``.

Mirrors the reference's three call layers (src/orbitals/atomicorbitals.jl:148-328):
public ``evaluate(basis, X[, ps, st])`` -> input extraction (positions, species
indices) -> kernel-facing layer, here one ``aok_ao_eval`` / ``aok_ao_pullback`` call.

Inputs: ``X`` is either a ``(nX, 3)`` tensor of positions (every point species 1, like
the reference's plain-coordinate input) or a ``PStates`` batch (positions + species
labels, the DecoratedParticles ``PState`` vector).  Outputs keep the reference's
column-major shapes: ``P`` is ``(nX, nOrb)``, ``dP`` is ``(nX, nOrb, 3)``
(``Matrix{SVector{3}}``), parameter cotangents ``(nRad, K, NZ)``.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _lib
from .basis import AtomicOrbitals, _dtype_code, as_fortran, fortran_empty


class NoTangent:
    """ChainRulesCore.NoTangent stand-in."""

    def __repr__(self):
        return "NoTangent()"


NO_TANGENT = NoTangent()


@dataclass
class PStates:
    """A batch of DecoratedParticles ``PState``s: positions (n, 3) and species labels."""

    positions: torch.Tensor
    species: list

    def __len__(self):
        return self.positions.shape[0]


def positions(X) -> torch.Tensor:
    """_positions (atomicorbitals.jl:130-131)."""
    P = X.positions if isinstance(X, PStates) else X
    if P.dim() != 2 or P.shape[1] != 3:
        raise _lib.AOKArgumentError("positions must have shape (n, 3)")
    return P.contiguous()


def species_indices(basis, X):
    """_species_indices (gtostoradials.jl:172-178): PState labels -> 1-based int64 device
    tensor via the zlist lookup; plain coordinates -> None (species 1)."""
    if not isinstance(X, PStates):
        return None
    rad = basis.Rnl if isinstance(basis, AtomicOrbitals) else basis
    lut = {z: i + 1 for i, z in enumerate(rad.zlist)}
    idx = []
    for s in X.species:
        s = s.item() if hasattr(s, "item") else s
        if s not in lut:
            raise _lib.AOKSpeciesError(f"species {s} not in basis species list {rad.zlist}")
        idx.append(lut[s])
    return torch.tensor(idx, dtype=torch.int64, device=X.positions.device)


def _resolve(basis: AtomicOrbitals, ps, st):
    ps = ps if ps is not None else basis.static_params()
    st = st if st is not None else basis.static_state()
    return ps, st


def _sidx_args(sidx):
    if sidx is None:
        return None, _lib.AOK_I64
    if sidx.dtype == torch.int64:
        return sidx.contiguous().data_ptr(), _lib.AOK_I64
    if sidx.dtype == torch.int32:
        return sidx.contiguous().data_ptr(), _lib.AOK_I32
    raise _lib.AOKArgumentError("species indices must be int64 or int32")


def _prep(basis, Rs, ps, st):
    zeta = as_fortran(ps["Rnl"]["zeta"])
    D = as_fortran(ps["Rnl"]["D"])
    dt = torch.promote_types(Rs.dtype, torch.promote_types(zeta.dtype, D.dtype))
    if Rs.dtype != dt:
        Rs = Rs.to(dt)
    if zeta.dtype != dt:
        zeta = as_fortran(zeta.to(dt))
    if D.dtype != dt:
        D = as_fortran(D.to(dt))
    for t, what in ((zeta, "zeta"), (D, "D")):
        if t.device != Rs.device:
            raise _lib.AOKArgumentError(f"{what} must live on the positions' device")
    nRad, K, NZ = zeta.shape
    poly = _lib.i64_array(st["Rnl"]["poly"])
    iR = _lib.i64_array(st["iR"])
    iY = _lib.i64_array(st["iY"])
    return Rs, zeta, D, dt, nRad, K, NZ, poly, iR, iY, len(st["iR"]), int(st["Ylm"]["L"])


def evaluate(basis: AtomicOrbitals, X, ps=None, st=None, *, sidx=None, grad=False, ctx=None):
    """evaluate(basis, X[, ps, st]) -> P (nX, nOrb); with ``grad`` also dP (nX, nOrb, 3)
    (evaluate_ed).  ``sidx`` (1-based species indices) selects the kernel-facing layer
    evaluate(basis, Rs, sidx, ps, st) directly."""
    ps, st = _resolve(basis, ps, st)
    Rs = positions(X)
    if sidx is None:
        sidx = species_indices(basis, X)
    Rs, zeta, D, dt, nRad, K, NZ, poly, iR, iY, nOrb, L = _prep(basis, Rs, ps, st)
    _lib.require_device(Rs, zeta, D, sidx if isinstance(sidx, torch.Tensor) else None)
    ctx = ctx or _lib.context(Rs.device.index or 0)
    nX = Rs.shape[0]
    P = fortran_empty((nX, nOrb), dt, Rs.device)
    dP = torch.empty((nOrb, nX, 3), dtype=dt, device=Rs.device).permute(1, 0, 2) if grad else None
    sp, it = _sidx_args(sidx)
    ctx.call("aok_ao_eval", _dtype_code(dt), Rs.data_ptr(), sp, it, nX, zeta.data_ptr(), D.data_ptr(), nRad, K, NZ,
             poly, iR, iY, nOrb, L, float(basis.lengthscale), basis.Rnl.decay, P.data_ptr(),
             dP.data_ptr() if grad else None)
    return (P, dP) if grad else P


def evaluate_ed(basis: AtomicOrbitals, X, ps=None, st=None, *, sidx=None, ctx=None):
    """evaluate_ed(basis, X[, ps, st]) -> (P, dP), dP w.r.t. input-unit coordinates."""
    return evaluate(basis, X, ps, st, sidx=sidx, grad=True, ctx=ctx)


def _pullback(dP, basis, X, ps, st, sidx, want_x, want_ps, ctx=None):
    ps, st = _resolve(basis, ps, st)
    Rs = positions(X)
    if sidx is None:
        sidx = species_indices(basis, X)
    Rs, zeta, D, dt, nRad, K, NZ, poly, iR, iY, nOrb, L = _prep(basis, Rs, ps, st)
    nX = Rs.shape[0]
    if tuple(dP.shape) != (nX, nOrb):
        raise _lib.AOKArgumentError(f"cotangent shape {tuple(dP.shape)} != {(nX, nOrb)}")
    dP = as_fortran(dP.to(dt))
    _lib.require_device(Rs, zeta, D, dP, sidx if isinstance(sidx, torch.Tensor) else None)
    ctx = ctx or _lib.context(Rs.device.index or 0)
    dX = torch.empty((nX, 3), dtype=dt, device=Rs.device) if want_x else None
    dz = fortran_empty((nRad, K, NZ), dt, Rs.device) if want_ps else None
    dD = fortran_empty((nRad, K, NZ), dt, Rs.device) if want_ps else None
    sp, it = _sidx_args(sidx)
    ctx.call("aok_ao_pullback", _dtype_code(dt), Rs.data_ptr(), sp, it, nX, zeta.data_ptr(), D.data_ptr(), nRad, K,
             NZ, poly, iR, iY, nOrb, L, float(basis.lengthscale), basis.Rnl.decay, dP.data_ptr(),
             dX.data_ptr() if want_x else None, dz.data_ptr() if want_ps else None,
             dD.data_ptr() if want_ps else None)
    return dX, dz, dD


def pullback_ps(dP, basis: AtomicOrbitals, X, ps, st, *, sidx=None, ctx=None):
    """pullback_ps(dP, basis, X, ps, st) (atomicorbitals.jl:254-270) ->
    {'Rnl': {'zeta': dzeta, 'D': dD}, 'Ylm': {}}."""
    _, dz, dD = _pullback(dP, basis, X, ps, st, sidx, False, True, ctx)
    return {"Rnl": {"zeta": dz, "D": dD}, "Ylm": {}}


def pullback_x(dP, basis: AtomicOrbitals, X, ps=None, st=None, *, sidx=None, ctx=None):
    """_aorb_pullback_x (atomicorbitals.jl:283-299): dX[i] = sum_o dP[i,o] grad P[i,o]."""
    dX, _, _ = _pullback(dP, basis, X, ps, st, sidx, True, False, ctx)
    return dX


def rrule_evaluate(basis: AtomicOrbitals, X, ps=None, st=None, *, sidx=None, ctx=None):
    """rrule(evaluate, basis, X, ps, st) (atomicorbitals.jl:307-328).  The forward returns
    P; the pullback recomputes the spatial Jacobian in registers instead of keeping
    dP alive, and returns (NoTangent, NoTangent, dX, dps, NoTangent); dX is the
    position cotangent (for PStates input, a PStates-shaped cotangent of the positions)."""
    P = evaluate(basis, X, ps, st, sidx=sidx, ctx=ctx)

    def pullback(dP):
        dX, dz, dD = _pullback(dP, basis, X, ps, st, sidx, True, True, ctx)
        return NO_TANGENT, NO_TANGENT, dX, {"Rnl": {"zeta": dz, "D": dD}, "Ylm": {}}, NO_TANGENT

    return P, pullback
