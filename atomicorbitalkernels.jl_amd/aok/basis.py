"""Basis types mirroring AtomicOrbitalKernels.jl's orbital-side layer.

* ``GaussianTypeRadials`` / ``SlaterTypeRadials``  (src/orbitals/gtostoradials.jl:14-46):
  learnable ``zeta, D`` of shape ``[nRad x K x NZ]``, radial spec ``(n, l)``, provenance
  ``(n1, n2, l)``, species labels ``zlist``.
* ``AtomicOrbitals``  (src/orbitals/atomicorbitals.jl:31-98): product basis R_nl * Z_lm with
  orbital spec ``(n, l, m)``, m = -l..l per radial, ``radidx``/``ylmidx`` (1-based,
  ``ylmidx = l^2 + l + m + 1``) and ``lengthscale`` (:bohr -> 1, :angstrom -> ang2bohr).
* ``CartesianGTOBasis`` + ``compile_basis``  (src/integrals/orbital/cartesian_gto_basis.jl:48-150).

Parameter tensors keep the reference's column-major memory order: a ``[n x K x NZ]``
array is a torch view of shape ``(n, K, NZ)`` with strides ``(1, n, n*K)`` so its
``data_ptr()`` is exactly the Julia layout the C ABI consumes.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib

ANG2BOHR = 1.8897261246257702  # src/AtomicOrbitalKernels.jl:56

_LENGTH_UNITS = {
    "bohr": 1.0,
    "angstrom": ANG2BOHR,
    "Å": ANG2BOHR,
    "nm": 10.0 * ANG2BOHR,
    "pm": 0.01 * ANG2BOHR,
}


def length_factor(u) -> float:
    """_length_factor (atomicorbitals.jl:50-54): native Bohr per input length unit."""
    if isinstance(u, (int, float)) and not isinstance(u, bool):
        return float(u)
    if u in _LENGTH_UNITS:
        return _LENGTH_UNITS[u]
    raise ValueError(f"unknown length unit :{u} — use :bohr, :angstrom/:Å, or a Unitful length")


def fortran_zeros(shape, dtype, device):
    """Zero tensor of logical ``shape`` stored column-major (Julia order)."""
    t = torch.zeros(tuple(reversed(shape)), dtype=dtype, device=device)
    return t.permute(*reversed(range(len(shape))))


def fortran_empty(shape, dtype, device):
    """Uninitialised column-major tensor (for outputs the kernels fill completely)."""
    t = torch.empty(tuple(reversed(shape)), dtype=dtype, device=device)
    return t.permute(*reversed(range(len(shape))))


def as_fortran(t: torch.Tensor) -> torch.Tensor:
    """Column-major copy of ``t`` unless it already is column-major contiguous."""
    if t.dim() <= 1:
        return t.contiguous()
    if t.permute(*reversed(range(t.dim()))).is_contiguous():
        return t
    out = fortran_zeros(t.shape, t.dtype, t.device)
    out.copy_(t)
    return out


def fortran_from_numpy(a, dtype, device):
    a = np.asarray(a)
    t = torch.as_tensor(np.ascontiguousarray(a.transpose(*reversed(range(a.ndim)))), dtype=dtype, device=device)
    return t.permute(*reversed(range(a.ndim)))


def _check_fortran(t: torch.Tensor, what: str):
    if t.dim() > 1 and not t.permute(*reversed(range(t.dim()))).is_contiguous():
        raise _lib.AOKArgumentError(f"{what} must be column-major contiguous")


@dataclass
class _Radials:
    zeta: torch.Tensor
    D: torch.Tensor
    spec: list  # [(n, l)]
    nnspec: list  # [(n1, n2, l)]
    zlist: tuple
    decay: int = _lib.AOK_DECAY_GAUSSIAN

    def __post_init__(self):
        nRad, K, NZ = self.zeta.shape
        assert tuple(self.zeta.shape) == tuple(self.D.shape), "size(zeta) == size(D)"
        assert nRad == len(self.spec) == len(self.nnspec), "nRad == length(spec) == length(nnspec)"
        assert NZ == len(self.zlist), "NZ == length(zlist)"
        self.zeta = as_fortran(self.zeta)
        self.D = as_fortran(self.D)

    def __len__(self):
        return len(self.spec)

    @property
    def nspecies(self):
        return len(self.zlist)

    def powers(self):
        """_powers (gtostoradials.jl:212-214): 0 for GTO, n1-1 for STO."""
        if self.decay == _lib.AOK_DECAY_GAUSSIAN:
            return [0] * len(self.spec)
        return [nn[0] - 1 for nn in self.nnspec]

    def z2i(self, s) -> int:
        """_z2i (gtostoradials.jl:164-168): species label -> 1-based species index."""
        for i, z in enumerate(self.zlist):
            if z == s:
                return i + 1
        raise _lib.AOKSpeciesError(f"species {s} not in basis species list {self.zlist}")


class GaussianTypeRadials(_Radials):
    def __init__(self, zeta, D, spec, nnspec=None, zlist=(1,)):
        super().__init__(zeta, D, list(spec), list(nnspec or [(n, 1, l) for (n, l) in spec]), tuple(zlist),
                         _lib.AOK_DECAY_GAUSSIAN)


class SlaterTypeRadials(_Radials):
    def __init__(self, zeta, D, spec, nnspec=None, zlist=(1,)):
        super().__init__(zeta, D, list(spec), list(nnspec or [(n, 1, l) for (n, l) in spec]), tuple(zlist),
                         _lib.AOK_DECAY_SLATER)


class AtomicOrbitals:
    """Product basis phi_nlm = R_nl(r) Z_lm(x); ``L`` is the SolidHarmonics degree."""

    def __init__(self, Rnl: _Radials, L: int | None = None, *, length_unit):
        self.Rnl = Rnl
        self.L = max(l for (_, l) in Rnl.spec) if L is None else int(L)
        self.spec, self.radidx, self.ylmidx = [], [], []
        for k, (n, l) in enumerate(Rnl.spec):  # atomicorbitals.jl:62-68
            for m in range(-l, l + 1):
                self.spec.append((n, l, m))
                self.radidx.append(k + 1)
                self.ylmidx.append(l * l + l + m + 1)
        self.lengthscale = length_factor(length_unit)

    def __len__(self):
        return len(self.spec)

    @property
    def dtype(self):
        return self.Rnl.zeta.dtype

    def static_params(self):
        """_static_params (atomicorbitals.jl:90-91): ps = (Rnl = (zeta, D), Ylm = (;))."""
        return {"Rnl": {"zeta": self.Rnl.zeta, "D": self.Rnl.D}, "Ylm": {}}

    def static_state(self):
        """_static_state (atomicorbitals.jl:93-95): st = (Rnl = (poly,), Ylm = (L,), iR, iY)."""
        return {"Rnl": {"poly": self.Rnl.powers()}, "Ylm": {"L": self.L}, "iR": list(self.radidx),
                "iY": list(self.ylmidx)}

    # LuxCore.setup equivalents
    def initialparameters(self):
        ps = self.static_params()
        return {"Rnl": {"zeta": ps["Rnl"]["zeta"].clone(), "D": ps["Rnl"]["D"].clone()}, "Ylm": {}}

    def initialstates(self):
        return self.static_state()

    def to(self, device=None, dtype=None):
        z = self.Rnl.zeta.to(device=device, dtype=dtype)
        D = self.Rnl.D.to(device=device, dtype=dtype)
        R = type(self.Rnl)(as_fortran(z), as_fortran(D), self.Rnl.spec, self.Rnl.nnspec, self.Rnl.zlist)
        out = AtomicOrbitals.__new__(AtomicOrbitals)
        out.__dict__.update(self.__dict__)
        out.Rnl = R
        return out


def _example_radial(TR, N1, N2, K=1, dtype=torch.float64, zlist=(1,), device="cpu"):
    """Deterministic toy radial basis (src/orbitals/utils.jl:10-37)."""
    spec, nnspec, zrows, drows = [], [], [], []
    for l in range(N1):
        n = 0
        for n1 in range(1, N1 + 1):
            for n2 in range(1, N2 + 1):
                n += 1
                spec.append((n, l))
                nnspec.append((n1, n2, l))
                zrows.append([0.5 * (n2 + j) + l for j in range(1, K + 1)])
                drows.append([1.0] * K if TR is GaussianTypeRadials else [1.0 / j for j in range(1, K + 1)])
    z2 = np.array(zrows)
    d2 = np.array(drows)
    NZ = len(zlist)
    zeta = np.zeros((len(spec), K, NZ))
    D = np.zeros((len(spec), K, NZ))
    for s in range(NZ):
        zeta[:, :, s] = z2 + 0.1 * s
        D[:, :, s] = d2 * (1 + 0.05 * s)
    return TR(fortran_from_numpy(zeta, dtype, device), fortran_from_numpy(D, dtype, device), spec, nnspec, zlist)


def gaussian_orbitals(N1=4, N2=3, *, length_unit, K=1, dtype=torch.float64, nspecies=1, zlist=None, device="cpu",
                      path=None):
    """gaussian_orbitals (src/orbitals/utils.jl:47-51): the synthetic (N1, N2) basis.
    Also the basis-set forms of src/orbitals/gaussianbasis.jl: ``gaussian_orbitals(bs)``
    for a BasisSet, ``gaussian_orbitals(name, elements)`` and
    ``gaussian_orbitals([(element, name), ...])``."""
    from . import basisset as _bs

    if isinstance(N1, _bs.BasisSet):
        return _bs.orbitals_from_basis_set(N1, length_unit=length_unit, dtype=dtype, device=device)
    if isinstance(N1, str):
        return _bs.orbitals_from_name(N1, N2, length_unit=length_unit, dtype=dtype, device=device, path=path)
    if isinstance(N1, (list, tuple)):
        return _bs.orbitals_from_pairs(N1, length_unit=length_unit, dtype=dtype, device=device, path=path)
    zl = tuple(zlist) if zlist is not None else tuple(range(1, nspecies + 1))
    return AtomicOrbitals(_example_radial(GaussianTypeRadials, N1, N2, K, dtype, zl, device), N1 - 1,
                          length_unit=length_unit)


def slater_orbitals(N1=4, N2=3, *, length_unit, K=1, dtype=torch.float64, nspecies=1, zlist=None, device="cpu"):
    """slater_orbitals (src/orbitals/utils.jl:60-64)."""
    zl = tuple(zlist) if zlist is not None else tuple(range(1, nspecies + 1))
    return AtomicOrbitals(_example_radial(SlaterTypeRadials, N1, N2, K, dtype, zl, device), N1 - 1,
                          length_unit=length_unit)


@dataclass
class CartesianGTOBasis:
    """Species-aware Cartesian SoA basis (cartesian_gto_basis.jl:48-61)."""

    Lmax: int
    nshells: int
    K: int
    ls: list
    nbf: list
    basis_offset: list
    nbf_total: int
    zeta: torch.Tensor  # [ns x K x NZ], column-major
    coef: torch.Tensor  # [ns x K x NZ], column-major
    zlist: tuple
    lengthscale: float
    _ls_c: object = field(default=None, repr=False)

    def __post_init__(self):
        self._ls_c = _lib.i64_array(self.ls)

    @property
    def nspecies(self):
        return len(self.zlist)

    def z2i(self, s) -> int:
        for i, z in enumerate(self.zlist):
            if z == s:
                return i + 1
        raise _lib.AOKSpeciesError(f"species {s} not in basis species list {self.zlist}")


def compile_struct(orb: AtomicOrbitals):
    """_compile_struct (cartesian_gto_basis.jl:88-104)."""
    ls = [l for (_, l) in orb.Rnl.spec]
    nbf = [(l + 1) * (l + 2) // 2 for l in ls]
    off = [0]
    for n in nbf:
        off.append(off[-1] + n)
    return ls, nbf, off


def compile_coef(zeta: torch.Tensor, D: torch.Tensor, ls, ctx=None) -> torch.Tensor:
    """_compile_coef (cartesian_gto_basis.jl:110-120) on the device, FP64 inside."""
    ctx = ctx or _lib.context(zeta.device.index or 0)
    zeta, D = as_fortran(zeta), as_fortran(D)
    ns, K, NZ = zeta.shape
    coef = fortran_zeros(zeta.shape, zeta.dtype, zeta.device)
    ls_c = _lib.i64_array(ls)
    ctx.call("aok_compile_coef", _dtype_code(zeta.dtype), zeta.data_ptr(), D.data_ptr(), ns, K, NZ, ls_c,
             coef.data_ptr())
    return coef


def compile_coef_pb(dcoef, zeta, D, ls, ctx=None):
    """_compile_coef_pb (overlap_2c_grad.jl:251-286) on the device, FP64 inside.  Returns (dD, dzeta_norm)."""
    ctx = ctx or _lib.context(zeta.device.index or 0)
    dcoef, zeta, D = as_fortran(dcoef), as_fortran(zeta), as_fortran(D)
    ns, K, NZ = zeta.shape
    dD = fortran_zeros(zeta.shape, zeta.dtype, zeta.device)
    dz = fortran_zeros(zeta.shape, zeta.dtype, zeta.device)
    ctx.call("aok_compile_coef_pb", _dtype_code(zeta.dtype), dcoef.data_ptr(), zeta.data_ptr(), D.data_ptr(), ns, K,
             NZ, _lib.i64_array(ls), dD.data_ptr(), dz.data_ptr())
    return dD, dz


def compile_basis(orb: AtomicOrbitals) -> CartesianGTOBasis:
    """compile_basis(::GaussianTypeOrbitals) (cartesian_gto_basis.jl:138-146); Slater -> error (:148-150)."""
    if orb.Rnl.decay != _lib.AOK_DECAY_GAUSSIAN:
        raise _lib.AOKArgumentError("the overlap kernels integrate plain Gaussians; a Slater-type orbital basis "
                                    "carries a nonzero radial power and is unsupported")
    ls, nbf, off = compile_struct(orb)
    coef = compile_coef(orb.Rnl.zeta, orb.Rnl.D, ls)
    return CartesianGTOBasis(max(ls), len(ls), orb.Rnl.zeta.shape[1], ls, nbf, off, off[-1], orb.Rnl.zeta, coef,
                             orb.Rnl.zlist, orb.lengthscale)


def _dtype_code(dt: torch.dtype) -> int:
    if dt == torch.float32:
        return _lib.AOK_F32
    if dt == torch.float64:
        return _lib.AOK_F64
    raise _lib.AOKArgumentError(f"unsupported dtype {dt}; use float32 or float64")
