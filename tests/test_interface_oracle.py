"""Interface flags (restraints / membrane) in the oracle: closed forms at the
3.9 A cutoff, all-pairs brute force, and the score fold of dna.rs:989-995."""
import numpy as np

from lklight import synth
from oracle import lkl_oracle as O


def _mol(xyz):
    n = len(xyz)
    return synth.Mol(xyz=np.asarray(xyz, np.float64), ele=np.zeros(n), rad=np.full(n, 1.8),
                     seps=np.full(n, 0.3), heavy=np.ones(n, np.uint8), hyd=np.ones(n, np.int32),
                     des=np.zeros(n), asa=np.zeros(n))


def test_interface_closed_form():
    # receptor atoms at x = 0, 20, 40; the ligand (identity pose) puts one atom
    # exactly 3.9 A from receptor atom 0 (inside: d2 == 3.9*3.9), one 3.95 A
    # from receptor atom 1 (outside) and one 3.0 A from receptor atom 2.
    rec = _mol([[0, 0, 0], [20, 0, 0], [40, 0, 0]])
    lig = _mol([[3.9, 0, 0], [20, 3.95, 0], [40, 0, 3.0]])
    cx = synth.Complex(rec=rec, lig=lig)
    orc = O.Oracle(cx)
    ir, il = orc.interface(np.array([1.0, 0, 0, 0, 0, 0, 0]))
    assert ir.tolist() == [1, 0, 1] and il.tolist() == [1, 0, 1]
    # a translation of 0.001 A moves the first pair outside
    ir, il = orc.interface(np.array([1.0, 0, 0, 0, 0.001, 0, 0]))
    assert ir.tolist() == [0, 0, 1] and il.tolist() == [0, 0, 1]
    # fractions and fold: residues {0}, {1, 2} -> 2/2; ligand residue {1} -> 0;
    # membrane {0, 1} -> 1/2 of beads at the interface (first pose)
    s = orc.restraint_score(-10.0, np.array([1.0, 0, 0, 0, 0, 0, 0]), [[0], [1, 2]], [[1]], [0, 1])
    assert s == -10.0 + (1.0 * -10.0 + 0.0 * -10.0 - 999.0 * 0.5)
    s = orc.restraint_score(-10.0, np.array([1.0, 0, 0, 0, 0, 0, 0]), [], [], [])
    assert s == -10.0


def test_interface_against_all_pairs():
    cx = synth.make_config("small")
    orc = O.Oracle(cx)
    r = cx.rec.xyz
    for pose in synth.make_poses(12, cx.meta["rec_axes"], 3):
        ir, il = orc.interface(pose)
        lc = orc.transform(pose)
        dx = lc[:, None, 0] - r[None, :, 0]
        dy = lc[:, None, 1] - r[None, :, 1]
        dz = lc[:, None, 2] - r[None, :, 2]
        m = dx * dx + dy * dy + dz * dz <= 3.9 * 3.9
        assert np.array_equal(ir, m.any(0).astype(np.uint8))
        assert np.array_equal(il, m.any(1).astype(np.uint8))
