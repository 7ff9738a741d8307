"""Gaussian basis-set ingestion: element-keyed shell data -> AtomicOrbitals.

Mirrors src/orbitals/gaussianbasis.jl:39-114 (gaussian_orbitals on a BasisSet, on a
named set for a list of elements, and on element => name pairs).  GaussianBasis.jl is
not available here, so the BasisSet side is restated: ``read_gbs`` parses Gaussian94
``.gbs`` text (the on-disk format of basis-set libraries), ``basis_set`` loads a named
set for a list of elements from ``<name>.gbs`` files (``path=``, then ``AOK_BASIS_PATH``,
then the bundled ``basis_library/``), and every shell is stored L2-normalised for real
solid harmonics:

    D_k = c * d_k * N_k,   N_k = [G_l (2 zeta_k)^-(l+3/2)]^-1/2,   G_l = Gamma(l+3/2)/2,
    c chosen so that  sum_ij D_i D_j G_l (zeta_i + zeta_j)^-(l+3/2) = 1,

with d_k the file's coefficients (which refer to normalised primitives).  This is the
convention the reference's tests fix (test/orbitals/test_gaussianbasis.jl:48-66,94-97:
the analytic single-atom overlap of the stored (zeta, D) equals libcint's overlap of the
same set).  The conversion to AtomicOrbitals follows gaussianbasis.jl:39-92: one species
per element, a shared (n, l) spec whose count per l is the maximum over the elements,
padding zeta = 1, D = 0, K = the longest contraction, L = the largest l.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .basis import AtomicOrbitals, GaussianTypeRadials, fortran_from_numpy

_SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br "
            "Kr").split()
_LETTERS = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4, "H": 5, "I": 6}
LIBRARY_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_library")


def atomic_number(el) -> int:
    """An element label (atomic number, symbol string) -> atomic number; the labels
    ChemicalSpecies accepts (gaussianbasis.jl:85-90)."""
    if isinstance(el, (int, np.integer)) and not isinstance(el, bool):
        if 1 <= int(el) <= len(_SYMBOLS):
            return int(el)
        raise _lib.AOKArgumentError(f"atomic number {el} out of range 1..{len(_SYMBOLS)}")
    if isinstance(el, str):
        s = el.strip()
        s = s[:1].upper() + s[1:].lower()
        if s in _SYMBOLS:
            return _SYMBOLS.index(s) + 1
    raise _lib.AOKArgumentError(f"unknown element {el!r}")


def symbol(Z: int) -> str:
    return _SYMBOLS[Z - 1]


def _gamma_half(l: int) -> float:
    """Gamma(l + 3/2) / 2 = int_0^inf r^(2l+2) exp(-r^2) dr."""
    return 0.5 * math.gamma(l + 1.5)


def normalise_shell(l: int, exps, coefs) -> np.ndarray:
    """File coefficients (of normalised primitives) -> D of an L2-normalised radial
    R(r) = sum_k D_k exp(-zeta_k r^2) paired with L2-normalised solid harmonics."""
    z = np.asarray(exps, np.float64)
    d = np.asarray(coefs, np.float64)
    g = _gamma_half(l)
    D = d * (g * (2.0 * z) ** (-(l + 1.5))) ** -0.5
    s = float(D @ (g * np.add.outer(z, z) ** (-(l + 1.5))) @ D)
    if not s > 0:
        raise _lib.AOKArgumentError(f"shell l={l} has a non-positive norm")
    return D / math.sqrt(s)


@dataclass
class Shell:
    """One spherical shell on element ``Z``: exponents and (normalised) coefficients."""

    Z: int
    l: int
    exp: np.ndarray
    coef: np.ndarray
    spherical: bool = True

    def __post_init__(self):
        self.exp = np.asarray(self.exp, np.float64).ravel()
        self.coef = np.asarray(self.coef, np.float64).ravel()
        if self.l < 0 or self.exp.size == 0 or self.exp.size != self.coef.size:
            raise _lib.AOKArgumentError("a shell needs l >= 0 and equal, non-empty exponent/coefficient lists")


@dataclass
class BasisSet:
    """The part of GaussianBasis.BasisSet the conversion reads: the atoms' elements and
    the shells (positions are irrelevant: the orbitals are centre-free)."""

    name: str
    atoms: list
    basis: list = field(default_factory=list)


def read_gbs(text: str) -> dict:
    """Parse Gaussian94 basis text -> {Z: [(l, exps, file_coefs), ...]} in file order.
    ``SP`` shells give an s and a p shell on the same exponents; exponents are scaled by
    the shell's scale factor squared; Fortran ``D`` exponents are accepted."""
    out = {}
    lines = [ln.split("!")[0].strip() for ln in text.splitlines()]
    lines = [ln for ln in lines if ln]
    i = 0

    def num(t):
        return float(t.replace("D", "E").replace("d", "e"))

    while i < len(lines):
        if lines[i].startswith("****"):
            i += 1
            continue
        head = lines[i].split()
        Z = atomic_number(head[0])
        if Z in out:
            raise _lib.AOKArgumentError(f"element {head[0]} appears twice in the basis file")
        shells = out[Z] = []
        i += 1
        while i < len(lines) and not lines[i].startswith("****"):
            t = lines[i].split()
            kind, n = t[0].upper(), int(t[1])
            sf = num(t[2]) if len(t) > 2 else 1.0
            rows = [[num(x) for x in lines[i + 1 + j].split()] for j in range(n)]
            i += 1 + n
            ex = np.array([r[0] for r in rows]) * sf * sf
            if kind == "SP":
                shells.append((0, ex, np.array([r[1] for r in rows])))
                shells.append((1, ex, np.array([r[2] for r in rows])))
            elif kind in _LETTERS:
                shells.append((_LETTERS[kind], ex, np.array([r[1] for r in rows])))
            else:
                raise _lib.AOKArgumentError(f"unsupported shell type {kind!r}")
    return out


def _find(name: str, path=None) -> str:
    dirs = []
    if path:
        dirs += [path] if isinstance(path, str) else list(path)
    dirs += [d for d in os.environ.get("AOK_BASIS_PATH", "").split(os.pathsep) if d]
    dirs.append(LIBRARY_DIR)
    fn = name.lower() + ".gbs"
    for d in dirs:
        p = os.path.join(d, fn)
        if os.path.isfile(p):
            return p
    raise _lib.AOKArgumentError(f"basis set {name!r} not found (looked for {fn} in {dirs})")


def basis_set(name: str, elements, path=None) -> BasisSet:
    """BasisSet(name, geometry) for one atom of each element (positions ignored), with
    every shell normalised as above."""
    Zs = [atomic_number(e) for e in elements]
    with open(_find(name, path)) as f:
        data = read_gbs(f.read())
    shells = []
    for Z in Zs:
        if Z not in data:
            raise _lib.AOKArgumentError(f"basis set {name!r} has no entry for {symbol(Z)}")
        shells += [Shell(Z, l, ex, normalise_shell(l, ex, d)) for (l, ex, d) in data[Z]]
    return BasisSet(name, Zs, shells)


def orbitals_from_basis_set(bs: BasisSet, *, length_unit, dtype=torch.float64, device="cpu") -> AtomicOrbitals:
    """gaussian_orbitals(bs::BasisSet; length_unit, T) (gaussianbasis.jl:39-83)."""
    shells = bs.basis
    if not shells:
        raise _lib.AOKArgumentError("the basis set has no shells")
    if not all(s.spherical for s in shells):
        raise _lib.AOKArgumentError("only spherical basis sets are supported")
    Zs = [int(z) for z in bs.atoms]
    if len(set(Zs)) != len(Zs):
        raise _lib.AOKArgumentError(f"each chemical element may appear at most once in the basis set; "
                                    f"got atomic numbers {Zs}")
    NZ = len(Zs)
    Lmax = max(s.l for s in shells)
    count = np.zeros(Lmax + 1, np.int64)  # per l: max shell count over the elements
    for Z in Zs:
        c = np.zeros(Lmax + 1, np.int64)
        for s in shells:
            if s.Z == Z:
                c[s.l] += 1
        count = np.maximum(count, c)
    spec = [(n, l) for l in range(Lmax + 1) for n in range(1, count[l] + 1)]
    nnspec = [(n, 1, l) for (n, l) in spec]
    K = max(s.exp.size for s in shells)
    zeta = np.ones((len(spec), K, NZ))
    D = np.zeros((len(spec), K, NZ))
    for sg, Z in enumerate(Zs):
        seen = np.zeros(Lmax + 1, np.int64)
        for s in shells:
            if s.Z != Z:
                continue
            seen[s.l] += 1
            k = spec.index((int(seen[s.l]), s.l))
            zeta[k, :s.exp.size, sg] = s.exp
            D[k, :s.exp.size, sg] = s.coef
    R = GaussianTypeRadials(fortran_from_numpy(zeta, dtype, device), fortran_from_numpy(D, dtype, device), spec,
                            nnspec, tuple(Zs))
    return AtomicOrbitals(R, Lmax, length_unit=length_unit)


def orbitals_from_name(name: str, elements, *, length_unit, dtype=torch.float64, device="cpu",
                       path=None) -> AtomicOrbitals:
    """gaussian_orbitals(basisname, elements; length_unit, T) (gaussianbasis.jl:95-102)."""
    elements = list(elements)
    if not elements:
        raise _lib.AOKArgumentError("`elements` is empty")
    return orbitals_from_basis_set(basis_set(name, elements, path), length_unit=length_unit, dtype=dtype,
                                   device=device)


def orbitals_from_pairs(pairs, *, length_unit, dtype=torch.float64, device="cpu", path=None) -> AtomicOrbitals:
    """gaussian_orbitals([el => name, ...]; length_unit, T) (gaussianbasis.jl:104-114)."""
    pairs = list(pairs)
    if not pairs:
        raise _lib.AOKArgumentError("the `element => basis` list is empty")
    subs = [basis_set(str(name), [el], path) for (el, name) in pairs]
    merged = BasisSet("mixed", [z for s in subs for z in s.atoms], [sh for s in subs for sh in s.basis])
    return orbitals_from_basis_set(merged, length_unit=length_unit, dtype=dtype, device=device)
