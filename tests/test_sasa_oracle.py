"""Pins of the oracle's Lee & Richards SASA (lr_sasa.rs:1-291) by what the
geometry fixes, so that a plausible restatement mistake (slice placement, the
probe, the wrap of an arc through angle 0, a missed fully-buried slice, the
union of overlapping arcs) fails one of them.
"""
import math

import numpy as np
import pytest

from oracle import lkl_oracle as O

PROBE = 1.4


def test_lone_sphere_is_exact():
    """One sphere: every slice is a full circle, the sum is 4 pi (r + probe)^2."""
    for r in (1.5, 1.7, 2.0):
        a = O.lr_sasa([[0.3, -1.2, 7.1]], [r])
        R = r + PROBE
        assert abs(a[0] - 4 * math.pi * R * R) <= 1e-12 * R * R


def test_sphere_inside_a_larger_one():
    """A sphere inside another has no exposed area; the outer one keeps 4 pi R^2."""
    a = O.lr_sasa([[0, 0, 0], [0.4, 0.2, -0.3]], [3.0, 0.8])
    assert a[1] == 0.0
    assert abs(a[0] - 4 * math.pi * (3.0 + PROBE) ** 2) <= 1e-12 * 100


@pytest.mark.parametrize("d", [3.1, 4.0, 5.5])
def test_two_spheres_cap_closed_form(d):
    """Two overlapping spheres: the exposed area of sphere i is
    2 pi R_i (R_i + a_i), a_i = (d^2 + R_i^2 - R_j^2) / (2 d); the slice sum
    converges to it (20,000 slices: within 1e-6)."""
    r1, r2 = 1.9, 1.6
    c2 = np.array([d * 0.6, d * 0.64, d * 0.48])  # |c2| = d
    a = O.lr_sasa([[0, 0, 0], c2], [r1, r2], n_slices=20000)
    R1, R2 = r1 + PROBE, r2 + PROBE
    ex = [2 * math.pi * R * (R + (d * d + R * R - Ro * Ro) / (2 * d)) for R, Ro in ((R1, R2), (R2, R1))]
    assert np.allclose(a, ex, rtol=1e-6, atol=0), (a, ex)


def dense_slices(xyz, radii, probe, n_slices, n_ang):
    """Independent restatement of the slice sum: slice k of atom i is the plane
    z = z_i - R_i + (k - 1/2) delta (delta = 2 R_i / n_slices); its exposed arc
    is measured by sampling n_ang points of the slice circle and testing each
    against every other sphere; area = sum of delta * R_i * arc."""
    xyz = np.asarray(xyz, float)
    R = np.asarray(radii, float) + probe
    th = (np.arange(n_ang) + 0.5) * (2 * math.pi / n_ang)
    c, s = np.cos(th), np.sin(th)
    out = np.zeros(len(R))
    for i in range(len(R)):
        others = np.array([j for j in range(len(R)) if j != i])
        delta = 2 * R[i] / n_slices
        for k in range(1, n_slices + 1):
            z = xyz[i, 2] - R[i] + (k - 0.5) * delta
            rp2 = R[i] ** 2 - (xyz[i, 2] - z) ** 2
            if rp2 <= 0:
                continue
            rp = math.sqrt(rp2)
            px, py = xyz[i, 0] + rp * c, xyz[i, 1] + rp * s
            cov = np.zeros(n_ang, bool)
            for j in others:
                cov |= (px - xyz[j, 0]) ** 2 + (py - xyz[j, 1]) ** 2 + (z - xyz[j, 2]) ** 2 < R[j] ** 2
            out[i] += delta * R[i] * 2 * math.pi * (1.0 - cov.mean())
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_twenty_slices_match_dense_sampling(seed):
    """The oracle's 20-slice areas of a random cluster equal the dense-angle
    restatement within its sampling error (20,000 angles per slice)."""
    rng = np.random.default_rng(seed)
    n = 18
    xyz = rng.uniform(-5, 5, (n, 3))
    radii = rng.uniform(1.5, 2.0, n)
    a = O.lr_sasa(xyz, radii)
    b = dense_slices(xyz, radii, PROBE, 20, 20000)
    full = 4 * math.pi * (radii + PROBE) ** 2
    assert (a < full * (1 + 1e-12)).all() and (a >= 0).all()
    assert np.abs(a - b).max() <= 2e-3 * full.max(), np.abs(a - b).max()
    # a cluster this dense buries some atoms more than half
    assert (a < 0.5 * full).any()


def test_contact_counts_and_order_free():
    """Reordering the atoms permutes the areas and changes nothing else (the
    arcs are sorted before the exposed length is summed)."""
    rng = np.random.default_rng(9)
    xyz = rng.uniform(-6, 6, (40, 3))
    radii = rng.uniform(1.5, 2.0, 40)
    a = O.lr_sasa(xyz, radii)
    p = rng.permutation(40)
    b = O.lr_sasa(xyz[p], radii[p])
    assert np.array_equal(a[p], b)
