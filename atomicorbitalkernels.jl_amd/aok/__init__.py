"""aok -- MI355X (gfx950) implementation of AtomicOrbitalKernels.jl's two hot paths.

H1  batched atomic-orbital evaluation (values, spatial gradients, (zeta, D, x) pullback);
H2  batched 2-/3-centre Cartesian-Gaussian overlaps and the 2-centre (zeta, D) pullback.

All compute runs in libaok.so (hand-written HIP kernels); this package is the host
mirror of the reference's API.  Importing it does not load the library; the first
call does, and raises ``AOKLibraryError`` if it is not built.
"""
from ._lib import (AOKArgumentError, AOKDeviceError, AOKError, AOKLibraryError, AOKLmaxError, AOKSpeciesError,
                   Context, Graph, context)
from . import autotune
from .basis import (ANG2BOHR, AtomicOrbitals, CartesianGTOBasis, GaussianTypeRadials, SlaterTypeRadials, as_fortran,
                    compile_basis, compile_coef, compile_coef_pb, fortran_from_numpy, fortran_zeros,
                    gaussian_orbitals, length_factor, slater_orbitals)
from .basisset import BasisSet, Shell, atomic_number, basis_set, normalise_shell, read_gbs
from .compiled import UNITFUL_HINT, CompiledBasis, Length, adapt_basis, compile_shells, to_bohr
from .orbitals import (NO_TANGENT, NoTangent, PStates, evaluate, evaluate_ed, pullback_ps, pullback_x,
                       rrule_evaluate)
from .overlap import (OverlapStepGraph, Ovl2cPlan, batch_overlap, batch_overlap_, batch_overlap_3c,
                      batch_overlap_3c_, overlap_3c_pb_coef_alpha, overlap_pb_coef_alpha, overlap_pullback_x,
                      rrule_batch_overlap, rrule_batch_overlap_3c)
from .overlap import pullback_ps as overlap_pullback_ps

__all__ = [
    "ANG2BOHR", "AOKArgumentError", "AOKDeviceError", "AOKError", "AOKLibraryError", "AOKLmaxError",
    "AOKSpeciesError", "AtomicOrbitals", "BasisSet", "CartesianGTOBasis", "CompiledBasis", "Context",
    "GaussianTypeRadials", "Graph", "Length", "NO_TANGENT", "NoTangent", "OverlapStepGraph", "Ovl2cPlan", "PStates",
    "Shell", "SlaterTypeRadials", "UNITFUL_HINT", "adapt_basis", "as_fortran", "atomic_number", "autotune",
    "basis_set", "batch_overlap", "batch_overlap_", "batch_overlap_3c", "batch_overlap_3c_", "compile_basis",
    "compile_coef", "compile_coef_pb", "compile_shells", "context", "evaluate", "evaluate_ed", "fortran_from_numpy",
    "fortran_zeros", "gaussian_orbitals", "length_factor", "normalise_shell", "overlap_3c_pb_coef_alpha",
    "overlap_pb_coef_alpha", "overlap_pullback_ps", "overlap_pullback_x", "pullback_ps", "pullback_x", "read_gbs",
    "rrule_batch_overlap", "rrule_batch_overlap_3c", "rrule_evaluate", "slater_orbitals", "to_bohr",
]
