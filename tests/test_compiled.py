"""CPU tests of the geometry-in host layer: explicit length units (to_bohr), the
CompiledBasis built from shells, adapt_basis, and the rejections the reference makes
before any kernel runs (plain numeric positions, host tensors handed to the library)."""
import json
import os

import numpy as np
import pytest
import torch

import aok
from aok import _lib

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")))

STO3G_H = (0, [3.42525091, 0.62391373, 0.16885540], [0.15432897, 0.53532814, 0.44463454])


def test_to_bohr_angstrom_is_ang2bohr():
    p = aok.to_bohr(aok.Length(np.array([[1.0], [-2.0], [0.5]]), "angstrom"))
    assert p.shape == (1, 3) and p.is_contiguous() and p.dtype == torch.float64
    np.testing.assert_array_equal(p.numpy()[0], np.array([1.0, -2.0, 0.5]) * GOLD["ang2bohr"])


@pytest.mark.parametrize("unit,scale", [("nm", 10.0), ("pm", 0.01), ("m", 1e10)])
def test_to_bohr_other_units(unit, scale):
    """Same positions in another unit give the same Bohr values (one rounding apart)."""
    rng = np.random.default_rng(3)
    ang = rng.standard_normal((3, 50))
    ref = aok.to_bohr(aok.Length(ang, "angstrom")).numpy()
    got = aok.to_bohr(aok.Length(ang / scale, unit)).numpy()
    np.testing.assert_allclose(got, ref, rtol=4e-16, atol=0)


def test_to_bohr_bohr_round_trip():
    x = np.random.default_rng(1).standard_normal((3, 20))
    np.testing.assert_allclose(aok.to_bohr(aok.Length(x, "bohr")).numpy(), x.T, rtol=5e-16, atol=0)


def test_to_bohr_fp32_two_roundings():
    """FT(angstrom value) * FT(ang2bohr), as units.jl:28-38 does for FT = Float32."""
    x = np.array([[0.1, 1.7], [2.3, -0.9], [3.3, 0.0]])
    got = aok.to_bohr(aok.Length(x, "angstrom"), torch.float32)
    assert got.dtype == torch.float32
    ref = (x.T.astype(np.float32) * np.float32(GOLD["ang2bohr"])).astype(np.float32)
    np.testing.assert_array_equal(got.numpy(), ref)


def test_to_bohr_rows_layout():
    x = np.random.default_rng(2).standard_normal((3, 7))
    a = aok.to_bohr(aok.Length(x, "nm"))
    b = aok.to_bohr(aok.Length(torch.tensor(x.T.copy()), "nm", rows=True))
    assert torch.equal(a, b)


def test_plain_positions_rejected_with_hint():
    for bad in (np.zeros((3, 4)), torch.zeros(3, 4)):
        with pytest.raises(aok.AOKArgumentError, match="explicit length units"):
            aok.to_bohr(bad)
    cb = aok.compile_shells([STO3G_H])
    with pytest.raises(aok.AOKArgumentError, match="explicit length units"):
        aok.batch_overlap(cb, np.zeros((3, 2)), np.zeros((3, 2)))
    with pytest.raises(aok.AOKArgumentError, match="explicit length units"):
        aok.batch_overlap_3c(cb, np.zeros((3, 2)), np.zeros((3, 2)), np.zeros((3, 2)))


def test_bad_shapes_and_units_rejected():
    with pytest.raises(aok.AOKArgumentError, match="unknown length unit"):
        aok.Length(np.zeros((3, 1)), "furlong")
    with pytest.raises(aok.AOKArgumentError, match=r"\(3, B\)"):
        aok.to_bohr(aok.Length(np.zeros((4, 2)), "angstrom"))
    with pytest.raises(aok.AOKArgumentError, match=r"\(B, 3\)"):
        aok.to_bohr(aok.Length(np.zeros((3, 2)), "angstrom", rows=True))


def test_compile_shells_layout():
    """The struct-of-arrays fields of compiled_basis.jl:60-99 (offsets 0-based here)."""
    shells = [STO3G_H, (1, [1.0, 0.3], [0.5, 0.6]), (2, [0.8], [1.0]), (3, [2.0, 0.9, 0.4, 0.1], [1, 2, 3, 4])]
    cb = aok.compile_shells(shells)
    assert cb.Lmax == 3 and cb.nshells == 4 and cb.K == 4
    np.testing.assert_array_equal(cb.ls, [0, 1, 2, 3])
    np.testing.assert_array_equal(cb.nprim, [3, 2, 1, 4])
    np.testing.assert_array_equal(cb.prim_offset, [0, 3, 5, 6, 10])
    np.testing.assert_array_equal(cb.nbf, [1, 3, 6, 10])
    np.testing.assert_array_equal(cb.basis_offset, [0, 1, 4, 10, 20])
    assert cb.nbf_total == 20
    np.testing.assert_array_equal(cb.coef.numpy()[5:], [1.0, 1, 2, 3, 4])
    c, a = cb.padded()
    assert c.shape == (4, 4, 1) and c.permute(2, 1, 0).is_contiguous()
    # slot k >= nprim: coefficient 0 (adds an exact zero), exponent 1 (finite)
    np.testing.assert_array_equal(c[1, :, 0].numpy(), [0.5, 0.6, 0, 0])
    np.testing.assert_array_equal(a[1, :, 0].numpy(), [1.0, 0.3, 1, 1])
    np.testing.assert_array_equal(c[0, :3, 0].numpy(), STO3G_H[2])


def test_compile_shells_rejects_bad_shells():
    with pytest.raises(aok.AOKArgumentError):
        aok.compile_shells([(0, [1.0, 2.0], [1.0])])
    with pytest.raises(aok.AOKArgumentError):
        aok.compile_shells([(-1, [1.0], [1.0])])
    with pytest.raises(aok.AOKArgumentError):
        aok.compile_shells([(0, [], [])])


def test_adapt_basis_converts_and_copies():
    cb = aok.compile_shells([STO3G_H, (1, [1.0, 0.3], [0.5, 0.6])])
    c32 = aok.adapt_basis(cb, "cpu", torch.float32)
    assert c32.coef.dtype == torch.float32 and c32.alpha.dtype == torch.float32
    np.testing.assert_array_equal(c32.coef.numpy(), cb.coef.numpy().astype(np.float32))
    assert c32.nbf_total == cb.nbf_total and c32.ls is not cb.ls
    c64 = aok.adapt_basis(cb, "cpu")
    assert c64.coef.dtype == torch.float64


def test_host_tensors_never_reach_the_library():
    """No CPU path: host tensors raise before any library call (the check runs first)."""
    cb = aok.compile_shells([STO3G_H])
    pos = aok.Length(np.zeros((3, 2)), "angstrom")
    with pytest.raises(aok.AOKArgumentError, match="no CPU path"):
        aok.batch_overlap(cb, pos, pos)
    with pytest.raises(aok.AOKArgumentError, match="no CPU path"):
        aok.batch_overlap_3c(cb, pos, pos, pos)
    with pytest.raises(aok.AOKArgumentError, match="no CPU path"):
        _lib.require_device(torch.zeros(2))
