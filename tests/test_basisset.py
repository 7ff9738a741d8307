"""CPU tests of basis-set ingestion (gaussianbasis.jl:39-114): the Gaussian94 reader,
the L2 normalisation of spherical shells, the element -> species conversion with a
shared zero-padded (n, l) spec, and the errors the reference raises.

The normalisation is pinned three independent ways: the reference's analytic
single-atom overlap of the stored (zeta, D) (test_gaussianbasis.jl:48-66) has a unit
diagonal; it equals the textbook overlap of contractions over normalised primitives,
<g_a|g_b> = (2 sqrt(zeta_a zeta_b) / (zeta_a + zeta_b))^(l+3/2), which is what libcint
computes from the file's coefficients (the reference's eigenvalue check, :94-97); and it
equals a numerical radial quadrature of R_i R_j r^(2l+2)."""
import math

import numpy as np
import pytest
import torch
from scipy.integrate import quad

import aok
from aok import basisset as bsmod

STO3G_C_FILE = {  # the published STO-3G carbon rows (exponent, s coef, p coef)
    "s1": ([71.6168370, 13.0450960, 3.5305122], [0.15432897, 0.53532814, 0.44463454]),
    "sp": ([2.9412494, 0.6834831, 0.2222899], [-0.09996723, 0.39951283, 0.70011547],
           [0.15591627, 0.60768372, 0.39195739]),
}


def _analytic_overlap(orb, sigma=0):
    """The reference's _analytic_overlap (test_gaussianbasis.jl:50-66) for species sigma."""
    z, D = orb.Rnl.zeta.numpy(), orb.Rnl.D.numpy()
    nb = len(orb)
    S = np.zeros((nb, nb))
    for i in range(nb):
        for j in range(nb):
            (_, li, mi), (_, lj, mj) = orb.spec[i], orb.spec[j]
            if li != lj or mi != mj:
                continue
            ki, kj = orb.radidx[i] - 1, orb.radidx[j] - 1
            g = 0.5 * math.gamma(li + 1.5)
            S[i, j] = g * np.sum(np.outer(D[ki, :, sigma], D[kj, :, sigma])
                                 * np.add.outer(z[ki, :, sigma], z[kj, :, sigma]) ** (-(li + 1.5)))
    return S


def _textbook_overlap(l, ea, da, eb, db):
    """Contracted overlap over normalised primitives, each contraction normalised."""
    def prim(x, y):
        return (2 * np.sqrt(np.outer(x, y)) / np.add.outer(x, y)) ** (l + 1.5)

    na = da @ prim(ea, ea) @ da
    nb = db @ prim(eb, eb) @ db
    return (da @ prim(ea, eb) @ db) / math.sqrt(na * nb)


def test_read_gbs_bundled_sto3g():
    with open(bsmod._find("sto-3g")) as f:
        data = aok.read_gbs(f.read())
    assert sorted(data) == [1, 2, 6, 7, 8]
    C = data[6]
    assert [sh[0] for sh in C] == [0, 0, 1]  # SP -> s then p on the same exponents
    np.testing.assert_array_equal(C[1][1], STO3G_C_FILE["sp"][0])
    np.testing.assert_array_equal(C[2][1], STO3G_C_FILE["sp"][0])
    np.testing.assert_array_equal(C[1][2], STO3G_C_FILE["sp"][1])
    np.testing.assert_array_equal(C[2][2], STO3G_C_FILE["sp"][2])
    np.testing.assert_array_equal(C[0][2], STO3G_C_FILE["s1"][1])


def test_read_gbs_scale_factor_and_fortran_exponents():
    """Scale factor 1.24 on the universal STO-3G 1s exponents gives hydrogen's."""
    txt = """! comment
****
H 0
S 3 1.24
 2.227660584D+00 0.154328967D+00
 0.405771156D+00 0.535328142D+00
 0.109817510D+00 0.444634542D+00
****
"""
    (l, ex, d), = aok.read_gbs(txt)[1]
    assert l == 0
    np.testing.assert_allclose(ex, [3.42525091, 0.62391373, 0.16885540], rtol=2e-8)
    np.testing.assert_allclose(d, [0.154328967, 0.535328142, 0.444634542], rtol=0)


def test_normalisation_three_ways():
    for l in (0, 1, 2, 3):
        rng = np.random.default_rng(l)
        ex = np.exp(rng.uniform(-2, 3, 4))
        d = rng.uniform(-1, 1, 4)
        D = aok.normalise_shell(l, ex, d)
        g = 0.5 * math.gamma(l + 1.5)
        assert abs(D @ (g * np.add.outer(ex, ex) ** (-(l + 1.5))) @ D - 1.0) < 1e-13
        # relative weights: D_k proportional to d_k times the primitive norm
        val, _ = quad(lambda r: (np.exp(-np.outer([r * r], ex)) @ D)[0] ** 2 * r ** (2 * l + 2), 0, np.inf,
                      epsabs=1e-14, epsrel=1e-12, limit=200)
        assert abs(val - 1.0) < 1e-9
    # two shells of one l: cross overlap = textbook normalised-primitive overlap
    (_, e1, d1), (_, e2, d2) = (0, *STO3G_C_FILE["s1"]), (0, STO3G_C_FILE["sp"][0], STO3G_C_FILE["sp"][1])
    D1, D2 = aok.normalise_shell(0, e1, d1), aok.normalise_shell(0, e2, d2)
    g = 0.5 * math.gamma(1.5)
    cross = g * D1 @ (np.add.outer(np.array(e1), np.array(e2)) ** -1.5) @ D2
    assert abs(cross - _textbook_overlap(0, np.array(e1), np.array(d1), np.array(e2), np.array(d2))) < 1e-13
    quad_cross, _ = quad(lambda r: (np.exp(-np.outer([r * r], e1)) @ D1)[0] * (np.exp(-np.outer([r * r], e2)) @ D2)[0]
                         * r * r, 0, np.inf, epsabs=1e-14, epsrel=1e-12, limit=200)
    assert abs(cross - quad_cross) < 1e-9


@pytest.mark.parametrize("el", ["H", "C", "N", "O"])
def test_single_element_overlap_matches_textbook(el):
    orb = aok.gaussian_orbitals("sto-3g", [el], length_unit="bohr")
    assert orb.Rnl.nspecies == 1
    S = _analytic_overlap(orb)
    np.testing.assert_allclose(np.diag(S), 1.0, rtol=1e-13)
    with open(bsmod._find("sto-3g")) as f:
        raw = aok.read_gbs(f.read())[aok.atomic_number(el)]
    ref = np.zeros_like(S)
    for i, (ni, li, mi) in enumerate(orb.spec):
        for j, (nj, lj, mj) in enumerate(orb.spec):
            if li == lj and mi == mj:
                si = [r for r in raw if r[0] == li][ni - 1]
                sj = [r for r in raw if r[0] == lj][nj - 1]
                ref[i, j] = _textbook_overlap(li, si[1], si[2], sj[1], sj[2])
    np.testing.assert_allclose(S, ref, atol=1e-13)


def test_union_spec_and_padding():
    """H (1s) + C (1s 2s 2p) -> shared spec [(1,0), (2,0), (1,1)]; H pads 2s and 2p."""
    orb = aok.gaussian_orbitals([("H", "sto-3g"), (6, "sto-3g")], length_unit="bohr")
    assert orb.Rnl.zlist == (1, 6)
    assert orb.Rnl.spec == [(1, 0), (2, 0), (1, 1)]
    assert orb.Rnl.nnspec == [(1, 1, 0), (2, 1, 0), (1, 1, 1)]
    assert len(orb) == 5 and orb.L == 1
    z, D = orb.Rnl.zeta.numpy(), orb.Rnl.D.numpy()
    assert z.shape == (3, 3, 2)
    np.testing.assert_array_equal(z[1:, :, 0], 1.0)
    np.testing.assert_array_equal(D[1:, :, 0], 0.0)
    np.testing.assert_array_equal(z[2, :, 1], STO3G_C_FILE["sp"][0])
    # carbon's slice equals the carbon-only conversion
    oc = aok.gaussian_orbitals("sto-3g", ["C"], length_unit="bohr")
    np.testing.assert_array_equal(D[:, :, 1], oc.Rnl.D.numpy()[:, :, 0])


def test_forms_agree_and_labels():
    a = aok.gaussian_orbitals("sto-3g", ["C", "N", "O"], length_unit="bohr")
    b = aok.gaussian_orbitals(aok.basis_set("sto-3g", [6, 7, 8]), length_unit="bohr")
    c = aok.gaussian_orbitals([("c", "sto-3g"), ("N", "STO-3G"), (8, "sto-3g")], length_unit="bohr")
    for o in (b, c):
        assert o.Rnl.zlist == a.Rnl.zlist == (6, 7, 8)
        assert o.Rnl.spec == a.Rnl.spec
        assert torch.equal(o.Rnl.zeta, a.Rnl.zeta) and torch.equal(o.Rnl.D, a.Rnl.D)
    o32 = aok.gaussian_orbitals("sto-3g", ["O"], length_unit="angstrom", dtype=torch.float32)
    assert o32.Rnl.D.dtype == torch.float32 and o32.lengthscale == aok.ANG2BOHR


def test_errors():
    with pytest.raises(aok.AOKArgumentError, match="at most once"):
        aok.gaussian_orbitals("sto-3g", ["H", "H"], length_unit="bohr")
    with pytest.raises(aok.AOKArgumentError, match="empty"):
        aok.gaussian_orbitals("sto-3g", [], length_unit="bohr")
    with pytest.raises(aok.AOKArgumentError, match="empty"):
        aok.gaussian_orbitals([], length_unit="bohr")
    with pytest.raises(aok.AOKArgumentError, match="no entry"):
        aok.gaussian_orbitals("sto-3g", ["Ne"], length_unit="bohr")
    with pytest.raises(aok.AOKArgumentError, match="not found"):
        aok.gaussian_orbitals("cc-pvqz-nonexistent", ["H"], length_unit="bohr")
    with pytest.raises(aok.AOKArgumentError, match="unknown element"):
        aok.atomic_number("Xx")
    with pytest.raises(aok.AOKArgumentError, match="no shells"):
        aok.gaussian_orbitals(aok.BasisSet("x", [1], []), length_unit="bohr")
    with pytest.raises(aok.AOKArgumentError, match="spherical"):
        aok.gaussian_orbitals(aok.BasisSet("x", [1], [aok.Shell(1, 0, [1.0], [1.0], spherical=False)]),
                              length_unit="bohr")
    with pytest.raises(TypeError):
        aok.gaussian_orbitals("sto-3g", ["H"])  # length_unit is required


def test_user_library_path(tmp_path, monkeypatch):
    (tmp_path / "mini.gbs").write_text("****\nLi 0\nS 1 1.00\n 0.5 1.0\nP 1 1.00\n 0.3 1.0\n****\n")
    orb = aok.gaussian_orbitals("mini", ["Li"], length_unit="bohr", path=str(tmp_path))
    assert orb.Rnl.spec == [(1, 0), (1, 1)] and orb.Rnl.zlist == (3,)
    monkeypatch.setenv("AOK_BASIS_PATH", str(tmp_path))
    orb2 = aok.gaussian_orbitals("mini", ["Li"], length_unit="bohr")
    assert torch.equal(orb2.Rnl.D, orb.Rnl.D)
    # single primitive: D = primitive radial norm, so the stored radial is normalised
    g = 0.5 * math.gamma(1.5)
    assert abs(orb.Rnl.D[0, 0, 0].item() ** 2 * g * (2 * 0.5) ** -1.5 - 1.0) < 1e-14
