"""lkl_lr_sasa (the gfx950 Lee & Richards SASA) against the oracle's
restatement of lr_sasa.rs.

Bar: per-atom areas within 1e-10 of the atom's full sphere area 4 pi (r +
probe)^2 (f64 throughout; only the device acos/atan2 differ from the host
libm, by an ulp or two); a lone sphere exact to rounding; bad arguments
refused with an error, not a fault.
"""
import math
import sys

import numpy as np
import pytest

import lklight
from lklight import synth
from oracle import lkl_oracle as O

import azp_fixture as A

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _check(xyz, radii, probe=1.4, n_slices=20):
    g = lklight.lr_sasa(xyz, radii, probe, n_slices)
    o = O.lr_sasa(xyz, radii, probe, n_slices)
    full = 4 * math.pi * (np.asarray(radii) + probe) ** 2
    err = np.abs(g - o) / full
    say("n %d slices %d: max |gpu - oracle| / full area %.3e, bitwise equal %.4f"
        % (len(radii), n_slices, float(err.max()), float(np.mean(g == o))))
    assert (err <= 1e-10).all(), float(err.max())
    return g


def test_lone_sphere_and_pairs(gpu):
    g = lklight.lr_sasa([[1.0, 2.0, 3.0]], [1.6])
    assert abs(g[0] - 4 * math.pi * 3.0 ** 2) <= 1e-12 * 113.1
    _check([[0, 0, 0], [2.4, 1.8, 0.0]], [1.9, 1.6], n_slices=5000)
    _check([[0, 0, 0], [0.4, 0.2, -0.3]], [3.0, 0.8])  # one sphere inside the other


def test_1azp_monomers(gpu):
    """The heavy atoms of each 1AZP monomer with cpydock's desolvation radii,
    the areas cpydock.rs:287-332 computes."""
    z = np.load(A.PATH)
    for k in ("rec", "lig"):
        heavy = z[k + "_cpy_hyd"] == 1
        _check(z[k + "_xyz"][heavy], z[k + "_cpy_desrad"][heavy])


def test_receptor_of_8000_atoms(gpu):
    """An 8,000-atom receptor at the benchmark density, radii U(1.5, 2.0)."""
    xyz, _ = synth.rsa(8000, synth.scaled_axes(8000, (1.0, 1.0, 1.0)), 42)
    radii = np.random.default_rng(4).uniform(1.5, 2.0, 8000)
    g = _check(xyz, radii)
    assert (g > 0).mean() > 0.3


def test_wave_path_equals_lane_path(gpu):
    """The wavefront-per-atom areas (arcs formed in parallel, ranked in LDS)
    equal, bit for bit, those of the one-lane-per-atom path (arcs insertion-
    sorted in scratch), which atoms with more than 128 neighbours take."""
    import os
    xyz, _ = synth.rsa(8000, synth.scaled_axes(8000, (1.0, 1.0, 1.0)), 42)
    radii = np.random.default_rng(4).uniform(1.5, 2.0, 8000)
    z = np.load(A.PATH)
    heavy = z["rec_cpy_hyd"] == 1
    cases = [(xyz, radii, 1.4), (z["rec_xyz"][heavy], z["rec_cpy_desrad"][heavy], 1.4),
             (xyz[:3000], radii[:3000] + 1.0, 3.0)]  # large spheres: many neighbours
    for x, r, probe in cases:
        wave = lklight.lr_sasa(x, r, probe)
        os.environ["LKL_SASA_LANE"] = "1"
        try:
            lane = lklight.lr_sasa(x, r, probe)
        finally:
            del os.environ["LKL_SASA_LANE"]
        say("wave vs lane path: n %d probe %.1f, bitwise equal %.4f" % (len(r), probe, float(np.mean(wave == lane))))
        assert np.array_equal(wave, lane)


def test_bad_arguments_are_refused(gpu):
    assert lklight.lr_sasa(np.zeros((0, 3)), np.zeros(0)).shape == (0,)
    with pytest.raises(lklight.GpuError):
        lklight.lr_sasa([[0, 0, 0]], [1.5], n_slices=0)
    with pytest.raises(lklight.GpuError):
        lklight.lr_sasa([[0, 0, np.nan]], [1.5])
    with pytest.raises(lklight.GpuError):
        lklight.lr_sasa([[0, 0, 0]], [0.0], probe=0.0)
