"""Geometry-in overlaps on a compiled Gaussian basis (no species axis).

Mirrors src/integrals/compiled_basis.jl:36-99 (CompiledBasis, compile_basis),
src/integrals/adapt.jl:13-25 (adapt_basis), src/integrals/units.jl:13-38 (to_bohr and
the explicit-unit rule) and the wrappers of src/integrals/kernels_2c.jl:139-184 and
src/integrals/kernels_3c.jl:183-222 (batch_overlap[!], batch_overlap_3c[!] on a
CompiledBasis).

Positions carry explicit length units, as the reference requires Unitful lengths:
``Length(values, "angstrom")`` with ``values`` of Julia shape (3, B) (a (3, B) tensor or
array), or (B, 3) when ``rows=True``.  Plain tensors are rejected with the reference's
hint.  A CompiledBasis runs on the same kernels as the orbital path: one species, the
per-shell primitive lists zero-padded to the longest contraction (a zero coefficient
adds an exact zero to every block, so the sums are the reference's).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .basis import ANG2BOHR, _dtype_code

UNITFUL_HINT = (
    "positions must carry explicit length units, e.g. Length(pos, \"angstrom\") or "
    "Length(pos, \"nm\") (\"bohr\", \"pm\" and \"m\" are accepted too).  Plain numeric "
    "matrices are not accepted: the unit choice must be explicit.")

# length of one unit in angstrom
_UNIT_TO_ANG = {"angstrom": 1.0, "nm": 10.0, "pm": 0.01, "m": 1e10, "bohr": 1.0 / ANG2BOHR}


@dataclass
class Length:
    """A matrix of positions with a length unit (the Unitful.Length matrices the reference
    accepts).  ``values``: Julia shape (3, B); with ``rows=True``, (B, 3) instead."""

    values: object
    unit: str
    rows: bool = False

    def __post_init__(self):
        if self.unit not in _UNIT_TO_ANG:
            raise _lib.AOKArgumentError(f"unknown length unit {self.unit!r}; one of {sorted(_UNIT_TO_ANG)}")


def to_bohr(pos, dtype=torch.float64, device=None) -> torch.Tensor:
    """to_bohr(pos, FT) (units.jl:28-38): positions in Bohr as a (B, 3) contiguous tensor,
    i.e. the memory of the reference's (3, B) matrix.  Plain arrays raise."""
    if not isinstance(pos, Length):
        raise _lib.AOKArgumentError(UNITFUL_HINT)
    v = pos.values
    t = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v, dtype=np.float64))
    if device is not None:
        t = t.to(device)
    if not pos.rows:
        if t.dim() != 2 or t.shape[0] != 3:
            raise _lib.AOKArgumentError("positions must have shape (3, B)")
        t = t.t()
    elif t.dim() != 2 or t.shape[1] != 3:
        raise _lib.AOKArgumentError("positions must have shape (B, 3) with rows=True")
    # the same two roundings as the reference: value in angstrom (FT), times ang2bohr (FT)
    ang = (t.to(torch.float64) * _UNIT_TO_ANG[pos.unit]).to(dtype)
    return (ang * torch.tensor(ANG2BOHR, dtype=dtype)).contiguous()


@dataclass
class CompiledBasis:
    """CompiledBasis (compiled_basis.jl:36-47): shells as struct-of-arrays.  Structural
    arrays are host int64; coef / alpha are 1-D tensors (host or device)."""

    Lmax: int
    nshells: int
    ls: np.ndarray
    nprim: np.ndarray
    prim_offset: np.ndarray
    coef: torch.Tensor
    alpha: torch.Tensor
    nbf: np.ndarray
    basis_offset: np.ndarray
    nbf_total: int
    _padded: dict = field(default_factory=dict, repr=False, compare=False)

    @property
    def K(self) -> int:
        return int(self.nprim.max()) if self.nshells else 1

    def padded(self):
        """(coef, alpha) as (nshells, K, 1) column-major tensors: slot k < nprim holds the
        shell's k-th primitive, the rest coefficient 0, exponent 1 (cached)."""
        key = (self.coef.dtype, self.coef.device)
        if key not in self._padded:
            ns, K = self.nshells, self.K
            c = torch.zeros((1, K, ns), dtype=self.coef.dtype, device=self.coef.device)
            a = torch.ones((1, K, ns), dtype=self.alpha.dtype, device=self.alpha.device)
            for s in range(ns):
                o, n = int(self.prim_offset[s]), int(self.nprim[s])
                c[0, :n, s] = self.coef[o:o + n]
                a[0, :n, s] = self.alpha[o:o + n]
            self._padded[key] = (c.permute(2, 1, 0), a.permute(2, 1, 0))
        return self._padded[key]


def compile_shells(shells, dtype=torch.float64, device="cpu") -> CompiledBasis:
    """compile_basis(BS) (compiled_basis.jl:60-99) for a basis given as a list of shells
    (l, exponents, coefficients), coefficients taken verbatim (the reference copies
    s.coef).  Cartesian shells only."""
    ls, nprim, coef, alpha = [], [], [], []
    for l, exps, cs in shells:
        exps, cs = np.asarray(exps, np.float64).ravel(), np.asarray(cs, np.float64).ravel()
        if l < 0 or len(exps) != len(cs) or len(exps) == 0:
            raise _lib.AOKArgumentError("each shell needs l >= 0 and equal, non-empty exponent/coefficient lists")
        ls.append(int(l))
        nprim.append(len(exps))
        coef.extend(cs.tolist())
        alpha.extend(exps.tolist())
    ls = np.array(ls, np.int64)
    nprim = np.array(nprim, np.int64)
    prim_offset = np.concatenate([[0], np.cumsum(nprim)]).astype(np.int64)
    nbf = (ls + 1) * (ls + 2) // 2
    basis_offset = np.concatenate([[0], np.cumsum(nbf)]).astype(np.int64)
    return CompiledBasis(int(ls.max()) if len(ls) else 0, len(ls), ls, nprim, prim_offset,
                         torch.tensor(coef, dtype=dtype, device=device),
                         torch.tensor(alpha, dtype=dtype, device=device), nbf, basis_offset,
                         int(basis_offset[-1]))


def adapt_basis(basis: CompiledBasis, device, dtype=None) -> CompiledBasis:
    """adapt_basis(basis, ArrayCtor, FT) (adapt.jl:13-25): the same basis with coef / alpha
    on ``device`` in ``dtype``."""
    dt = dtype or basis.coef.dtype
    return CompiledBasis(basis.Lmax, basis.nshells, basis.ls.copy(), basis.nprim.copy(), basis.prim_offset.copy(),
                         basis.coef.to(device=device, dtype=dt), basis.alpha.to(device=device, dtype=dt),
                         basis.nbf.copy(), basis.basis_offset.copy(), basis.nbf_total)


def _padded_on(basis: CompiledBasis, device, dtype):
    """Padded (coef, alpha) on ``device`` in ``dtype``, column-major (nshells, K, 1)."""
    def fort(t):
        return t.to(device=device, dtype=dtype).permute(2, 1, 0).contiguous().permute(2, 1, 0)

    c, a = basis.padded()
    return fort(c), fort(a)


def _flags(mode):
    if mode == "fast":
        return _lib.AOK_FAST
    if mode == "plain":
        return _lib.AOK_PLAIN
    raise _lib.AOKArgumentError(f"unknown mode {mode!r}")


def batch_overlap_(out, basis: CompiledBasis, posA, posB, *, mode="fast", ctx=None):
    """batch_overlap!(out, basis::CompiledBasis, posA, posB) (kernels_2c.jl:158-168):
    ``out`` (N, N, B) column-major; its dtype sets the kernel precision."""
    dt = out.dtype
    A = to_bohr(posA, dt, out.device)
    Bm = to_bohr(posB, dt, out.device)
    if A.shape[0] != Bm.shape[0]:
        raise _lib.AOKArgumentError(f"posA and posB must have equal length; got {A.shape[0]} and {Bm.shape[0]}")
    B, N = A.shape[0], basis.nbf_total
    if tuple(out.shape) != (N, N, B) or not out.permute(2, 1, 0).is_contiguous():
        raise _lib.AOKArgumentError(f"out must be a column-major {(N, N, B)} tensor")
    _lib.require_device(out, A, Bm)
    coef, alpha = _padded_on(basis, out.device, dt)
    ctx = ctx or _lib.context(out.device.index or 0)
    ctx.call("aok_ovl2c", _dtype_code(dt), _flags(mode), A.data_ptr(), Bm.data_ptr(), None, None, _lib.AOK_I64, B,
             basis.nshells, basis.K, 1, _lib.i64_array(basis.ls), coef.data_ptr(), alpha.data_ptr(), 1.0,
             out.data_ptr())
    return out


def batch_overlap(basis: CompiledBasis, posA, posB, *, dtype=torch.float64, device=None, mode="fast", ctx=None):
    """batch_overlap(basis::CompiledBasis, posA, posB; FT) (kernels_2c.jl:170-184)."""
    dev = device or basis.coef.device
    B = to_bohr(posA, dtype).shape[0]
    N = basis.nbf_total
    out = torch.empty((B, N, N), dtype=dtype, device=dev).permute(2, 1, 0)
    return batch_overlap_(out, basis, posA, posB, mode=mode, ctx=ctx)


def batch_overlap_3c_(out, basis: CompiledBasis, posA, posB, posC, *, mode="fast", ctx=None):
    """batch_overlap_3c!(out, basis::CompiledBasis, posA, posB, posC) (kernels_3c.jl:183-205):
    ``out`` (N, N, N, B) column-major."""
    dt = out.dtype
    P = [to_bohr(p, dt, out.device) for p in (posA, posB, posC)]
    B = P[0].shape[0]
    if P[1].shape[0] != B or P[2].shape[0] != B:
        raise _lib.AOKArgumentError("posA, posB and posC must have equal length")
    N = basis.nbf_total
    if tuple(out.shape) != (N, N, N, B) or not out.permute(3, 2, 1, 0).is_contiguous():
        raise _lib.AOKArgumentError(f"out must be a column-major {(N, N, N, B)} tensor")
    _lib.require_device(out, *P)
    coef, alpha = _padded_on(basis, out.device, dt)
    ctx = ctx or _lib.context(out.device.index or 0)
    ctx.call("aok_ovl3c", _dtype_code(dt), _flags(mode), P[0].data_ptr(), P[1].data_ptr(), P[2].data_ptr(), None,
             None, None, _lib.AOK_I64, B, basis.nshells, basis.K, 1, _lib.i64_array(basis.ls), coef.data_ptr(),
             alpha.data_ptr(), 1.0, out.data_ptr())
    return out


def batch_overlap_3c(basis: CompiledBasis, posA, posB, posC, *, dtype=torch.float64, device=None, mode="fast",
                     ctx=None):
    """batch_overlap_3c(basis::CompiledBasis, posA, posB, posC; FT) (kernels_3c.jl:213-222)."""
    dev = device or basis.coef.device
    B = to_bohr(posA, dtype).shape[0]
    N = basis.nbf_total
    out = torch.empty((B, N, N, N), dtype=dtype, device=dev).permute(3, 2, 1, 0)
    return batch_overlap_3c_(out, basis, posA, posB, posC, mode=mode, ctx=ctx)
