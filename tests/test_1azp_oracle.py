"""The oracle on the 1AZP complex of the reference's own unit tests.

Golden values and tolerances are the reference's (pydock.rs:554-593,
dna.rs:1112-1180, vdw.rs:351-378); the per-atom arrays come from
tests/golden/1azp.npz (the reference's test PDB files typed with its amber.rs /
dna.rs tables, see tests/golden/make_1azp.py).  These pin the oracle's
formulas, constants, far-field build and sample, and fold against absolute
numbers the reference's tests hold, on a real protein-DNA complex.
"""
import numpy as np
import pytest

from oracle import lkl_oracle as O

import azp_fixture as A


@pytest.fixture(scope="module")
def pydock_cx():
    cx = A.load("pydock")
    cx.set_field(*O.field_build(cx.rec.xyz, cx.rec.ele))
    return cx, O.Oracle(cx)


@pytest.fixture(scope="module")
def dna_cx():
    cx = A.load("dna")
    cx.set_field(*O.field_build(cx.rec.xyz, cx.rec.ele))
    return cx, O.Oracle(cx)


def test_fixture_shapes():
    z = np.load(A.PATH)
    assert z["rec_xyz"].shape == (1094, 3) and z["lig_xyz"].shape == (506, 3)
    # the DNA ligand: 16 nucleotides, each with a phosphate (16 P atoms), and
    # a net charge of -16 e under both tables
    for fam in ("pydock", "dna"):
        assert abs(z["lig_%s_ele" % fam].sum() + 16.0) < 1e-9


def test_pydock_exact_kat(pydock_cx):
    """pydock.rs:578-582: energy_exact at the identity pose."""
    cx, orc = pydock_cx
    e = orc.pydock_exact(A.pose7((0.0, 0.0, 0.0)))
    assert abs(e - A.PYDOCK_EXACT_IDENTITY) < 1e-8, e


@pytest.mark.parametrize("t", A.PYDOCK_POSES)
def test_pydock_grid_vs_exact(pydock_cx, t):
    """pydock.rs:583-592: the grid path within 5 % of the exact 30 A sum."""
    cx, orc = pydock_cx
    p = A.pose7(t)
    parts = orc.parts("pydock", p)
    g = O.fold("pydock", parts.e_raw, parts.v)
    e = orc.pydock_exact(p)
    r = abs(g - e) / abs(e) if abs(e) > 1.0 else abs(g - e)
    assert r < 0.05, (t, g, e)


def test_dna_energy_kat(dna_cx):
    """dna.rs:1143-1151: the grid energy at the identity pose.  The reference
    allows 2 %; the oracle reproduces the reference's value to its last digit,
    which is the tolerance asserted here (1e-12 relative)."""
    cx, orc = dna_cx
    parts = orc.parts("dna", A.pose7((0.0, 0.0, 0.0)))
    g = O.fold("dna", parts.e_raw, parts.v)
    assert abs(g - A.DNA_ENERGY_IDENTITY) <= 1e-12 * abs(A.DNA_ENERGY_IDENTITY), repr(g)


@pytest.mark.parametrize("t", A.DNA_POSES)
def test_dna_grid_vs_exact(dna_cx, t):
    """dna.rs:1168-1178: grid within 2 % of exact (relative, floor 1)."""
    cx, orc = dna_cx
    p = A.pose7(t)
    parts = orc.parts("dna", p)
    g = O.fold("dna", parts.e_raw, parts.v)
    e = orc.dna_exact(p)
    assert abs(e - g) / max(abs(e), 1.0) < 0.02, (t, g, e)


@pytest.mark.parametrize("t", A.PYDOCK_POSES)
def test_vdw_grid_equals_exact(pydock_cx, t):
    """vdw.rs:369-377: on pydock's model the cell path and the all-pairs scan
    agree to f64 summation order (1e-9)."""
    cx, orc = pydock_cx
    p = A.pose7(t)
    g = O.fold("vdw", 0.0, orc.parts("vdw", p).v)
    e = O.fold("vdw", 0.0, orc.bruteforce("vdw", p).v)
    assert abs(g - e) < 1e-9, (t, g, e)


@pytest.fixture(scope="module")
def cpydock_cx():
    cx = A.load("cpydock", sasa=O.lr_sasa)
    cx.set_field(*O.field_build(cx.rec.xyz, cx.rec.ele))
    return cx, O.Oracle(cx)


def test_cpydock_areas(cpydock_cx):
    """Every heavy atom has an area in [0, 4 pi (r + 1.4)^2] and the monomers
    are not fully buried; hydrogens keep -1 (cpydock.rs:279-332)."""
    cx, _ = cpydock_cx
    z = np.load(A.PATH)
    for m, k in ((cx.rec, "rec"), (cx.lig, "lig")):
        heavy = m.hyd == 1
        full = 4 * np.pi * (z[k + "_cpy_desrad"] + 1.4) ** 2
        assert (m.asa[~heavy] == -1.0).all()
        assert (m.asa[heavy] >= 0).all() and (m.asa[heavy] <= full[heavy] * (1 + 1e-12)).all()
        assert (m.asa[heavy] > 0).mean() > 0.5


@pytest.mark.parametrize("t", A.CPYDOCK_POSES)
def test_cpydock_grid_vs_exact(cpydock_cx, t):
    """cpydock.rs:861-875: grid and exact within 12 (absolute) at each pose."""
    cx, orc = cpydock_cx
    p = A.pose7(t)
    parts = orc.parts("cpydock", p)
    g = O.fold("cpydock", parts.e_raw, parts.v, parts.s)
    e = orc.cpydock_exact(p)
    assert abs(g - e) < A.CPYDOCK_GRID_EXACT, (t, g, e)


def test_f32_inputs_stay_within_budget(dna_cx):
    """The device is handed f32 coordinates and parameters (the reference's
    GPU bridge casts them the same way); on this complex that rounding moves
    the dna grid energy by far less than the 1e-5 parity budget."""
    cx, orc = dna_cx
    cx32 = A.load("dna", f32=True)
    cx32.set_field(*O.field_build(cx32.rec.xyz, cx32.rec.ele))
    o32 = O.Oracle(cx32)
    for t in A.DNA_POSES:
        p = A.pose7(t)
        a, b = orc.parts("dna", p), o32.parts("dna", p)
        g, g32 = O.fold("dna", a.e_raw, a.v), O.fold("dna", b.e_raw, b.v)
        assert abs(g - g32) <= 1e-6 * max(abs(g), 1.0), (t, g, g32)
