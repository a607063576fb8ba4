"""The 1AZP complex of the reference's own unit tests, from tests/golden/1azp.npz
(written by tests/golden/make_1azp.py), as a synth.Complex of one family's
parameter set.

`f32=True` rounds coordinates and parameters to f32, which is what the GPU
path is handed (the reference's gpu_score.rs:73-145 / gpu_family.rs:82-153 cast
the f64 model arrays to f32 the same way); the oracle is then run on the same
rounded values so the two see identical inputs.
"""
import os

import numpy as np

from lklight import synth

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "1azp.npz")

# Poses of the reference's 1AZP tests (identity rotation, these translations).
PYDOCK_POSES = [(0.0, 0.0, 0.0), (2.0, 0.0, 0.0), (0.0, 3.0, 0.0), (-4.0, 2.0, 5.0),
                (1.0, -2.0, 6.0)]                                   # pydock.rs:563-592
DNA_POSES = [(0.0, 0.0, 0.0), (1.0, 2.0, 3.0), (-2.0, 1.0, -4.0), (5.0, -3.0, 2.0)]  # dna.rs:1170

CPYDOCK_POSES = [(0.0, 0.0, 0.0), (2.0, 3.0, 1.0), (-4.0, 5.0, 2.0), (0.0, -6.0, 8.0),
                 (3.0, 3.0, -3.0)]                                  # cpydock.rs:861-875

PYDOCK_EXACT_IDENTITY = -364.88126358158974   # pydock.rs:581-582, |d| < 1e-8
DNA_ENERGY_IDENTITY = -1211.6718762645426     # dna.rs:1149-1151, relative 2 %
CPYDOCK_GRID_EXACT = 12.0                     # cpydock.rs:866-874, |grid - exact| < 12


class _Mol(synth.Mol):
    """synth.Mol whose epsilon is the model's own value (vdw.rs:225 takes
    sqrt(eps_i * eps_j) of it) rather than sqrt_eps squared."""

    @property
    def eps(self):
        return self._eps


def pose7(t):
    return np.array([1.0, 0.0, 0.0, 0.0, t[0], t[1], t[2]])


def load(family, f32=False, sasa=None):
    """family: "pydock" (amber.rs tables; vdw.rs's tests use this model too),
    "dna" (dna.rs tables) or "cpydock" (amber.rs tables plus the hydrogen
    flags, desolvation coefficients and Lee & Richards areas of
    cpydock.rs:172-332).  For cpydock, `sasa(xyz, radii)` gives the areas of
    each monomer's heavy atoms from the f64 coordinates and the desolvation
    radii (hydrogens get -1, as cpydock.rs:279 leaves them)."""
    z = np.load(PATH)
    cpy = family == "cpydock"
    if cpy:
        if sasa is None:
            raise ValueError("cpydock needs a SASA function for the reference areas")
        family = "pydock"

    def r32(a):
        return np.asarray(a, np.float32).astype(np.float64) if f32 else np.asarray(a, np.float64)

    def mol(k):
        xyz = r32(z[k + "_xyz"])
        n = xyz.shape[0]
        eps = np.asarray(z["%s_%s_eps" % (k, family)], np.float64)
        m = _Mol(xyz=xyz, ele=r32(z["%s_%s_ele" % (k, family)]),
                 rad=r32(z["%s_%s_rad" % (k, family)]), seps=r32(np.sqrt(eps)),
                 heavy=np.asarray(z[k + "_heavy"], np.uint8),
                 hyd=np.zeros(n, np.int32), des=np.zeros(n), asa=np.zeros(n))
        # f32 inputs: the device forms sqrt(eps_i) * sqrt(eps_j) from the f32
        # sqrt values, so the oracle is given eps = sqrt_eps^2 of those.
        m._eps = m.seps * m.seps if f32 else eps
        if cpy:
            hyd = np.asarray(z[k + "_cpy_hyd"], np.int32)
            heavy = hyd == 1
            asa = np.full(n, -1.0)
            asa[heavy] = sasa(z[k + "_xyz"][heavy], z[k + "_cpy_desrad"][heavy])
            m.hyd, m.des, m.asa = hyd, r32(z[k + "_cpy_des"]), r32(asa)
        return m

    return synth.Complex(rec=mol("rec"), lig=mol("lig"),
                         meta={"name": "1azp", "family": "cpydock" if cpy else family, "f32": f32})
