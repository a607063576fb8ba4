"""Writes tests/golden/tiny_scores.json: oracle outputs (f64 E_raw, V, S and the
folded score per family) for 24 poses of the seeded 300-atom / 60-atom
complex, with the oracle's own far field.  Inputs are regenerated from the
seeds by lklight.synth; the file stores the seeds, sizes and a checksum of the
coordinates so a generator change is detected.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "lklight-gpu_amd"))
sys.path.insert(0, ROOT)

from lklight import synth  # noqa: E402
from oracle import lkl_oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiny_scores.json")


def inputs():
    cx = synth.make_config("tiny")
    phi, o, d = O.field_build(cx.rec.xyz, cx.rec.ele)
    cx.set_field(phi, o, d)
    poses = synth.make_poses(24, cx.meta["rec_axes"], 1234, shell=6.0)
    return cx, poses


def digest(cx, poses):
    h = hashlib.sha256()
    for a in (cx.rec.xyz, cx.rec.ele, cx.lig.xyz, cx.lig.ele, poses, cx.phi):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    cx, poses = inputs()
    orc = O.Oracle(cx)
    out = {"complex": "tiny", "seed": cx.meta["seed"], "n_rec": cx.rec.n, "n_lig": cx.lig.n,
           "pose_seed": 1234, "n_poses": int(poses.shape[0]), "inputs_sha256": digest(cx, poses),
           "families": {}}
    for m in ("dna", "vdw", "pydock", "cpydock"):
        r = orc.batch(m, poses)
        out["families"][m] = {"e": r["e"].tolist(), "v": r["v"].tolist(),
                              "s": r["s"].tolist(),
                              "score": [O.fold(m, float(e), float(v), float(t))
                                        for e, v, t in zip(r["e"], r["v"], r["s"])]}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
