"""The dataset directory of the reference's DNA::dump_metal_dataset
(dna.rs:473-663): a deterministic batch of dna grid-scoring inputs with the
CPU grid path's per-pose energies, written so that an external kernel can be
validated and timed on byte-identical data.

Files (little-endian, packed, no padding):
  meta.json  nx ny nz ox oy oz spacing phi_len nr nl n_pose ncx ncy ncz
             c_ox c_oy c_oz c_sp cell_start_len cpu_grid_ms cpu_exact_ms
  field.bin  [nx ny nz: i32][ox oy oz: f64][spacing: f64][phi: f32 ...]
  rec.bin    [nr: i32][xyz f32 nr*3][ele f32 nr][svdw f32 nr][vdwr f32 nr]
             [heavy u8 nr][ncx ncy ncz: i32][c_ox c_oy c_oz c_sp: f32]
             [cell_start i32 (ncx+1)(ncy+1)(ncz+1)+1][cell_atoms i32 nr]
  lig.bin    [nl: i32][ele f32 nl][svdw f32 nl][vdwr f32 nl][heavy u8 nl]
  poses.bin  [n_pose: i32][nl: i32][placed ligand xyz f32 n_pose*nl*3, pose-major]
  ref.bin    [n_pose: i32][energy f64 n_pose]

poses.bin holds the placed coordinates (f64 rigid transform, then f32), not the
rotations, so `Dataset.score` scores them with lkl_ctx_score_coords.  The
energies are dna scores -(E * 332/4 + V) (dna.rs:976-995 without restraints).
The writer takes them from the caller: it computes no energies itself.
"""
import ctypes as C
import json
import os
from dataclasses import dataclass, field

import numpy as np

from . import _capi, synth

CELL = 10.0  # CLOSE_DIST, the reference's cell edge (gpu_score.rs:75)


def cell_list(xyz):
    """gpu_score.rs:73-145: 10 A cells over the receptor (lo = floor(min -
    0.01), n = floor((hi - lo) / 10) + 1), each atom in its clamped cell,
    counting sort keeping input order.  Returns (cell_start, cell_atoms,
    (ncx, ncy, ncz), (c_ox, c_oy, c_oz, c_sp)) with the f32 origin and edge."""
    xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
    lo = np.floor(xyz.min(0) - 0.01)
    hi = xyz.max(0)
    n = np.maximum(np.floor((hi - lo) / CELL).astype(np.int64) + 1, 1)
    idx = np.clip(np.floor((xyz - lo) / CELL).astype(np.int64), 0, n - 1)
    cell = (idx[:, 2] * n[1] + idx[:, 1]) * n[0] + idx[:, 0]
    ncells = int((n[0] + 1) * (n[1] + 1) * (n[2] + 1))
    counts = np.bincount(cell, minlength=ncells)
    start = np.zeros(ncells + 1, np.int32)
    start[1:] = np.cumsum(counts)
    atoms = np.argsort(cell, kind="stable").astype(np.int32)
    return start, atoms, tuple(int(v) for v in n), tuple(float(np.float32(v)) for v in (*lo, CELL))


def placed(lig_xyz, poses):
    """The ligand placed by each pose (w, x, y, z, tx, ty, tz): qt.rs:67-74's
    matrix, qt.rs:122-128's product and the translation in f64, in the
    reference's operation order, then f32 (dna.rs:546-556).  -> (N, nl, 3)."""
    v = np.asarray(lig_xyz, np.float64).reshape(-1, 3)
    P = np.asarray(poses, np.float64).reshape(-1, 7)
    out = np.empty((P.shape[0], v.shape[0], 3), np.float32)
    for k, (w, x, y, z, tx, ty, tz) in enumerate(P):
        m = ((1. - 2. * (y * y + z * z), 2. * (x * y - w * z), 2. * (x * z + w * y)),
             (2. * (x * y + w * z), 1. - 2. * (x * x + z * z), 2. * (y * z - w * x)),
             (2. * (x * z - w * y), 2. * (y * z + w * x), 1. - 2. * (x * x + y * y)))
        for a, t in enumerate((tx, ty, tz)):
            out[k, :, a] = (m[a][0] * v[:, 0] + m[a][1] * v[:, 1] + m[a][2] * v[:, 2]) + t
    return out


def seeded_poses(seed, n):
    """The reference's pose set for `seed` (dna.rs:497-519, lkl_dataset_poses)."""
    out = np.zeros((n, 7), np.float64)
    if _capi.lib().lkl_dataset_poses(int(seed) & 0xFFFFFFFFFFFFFFFF, int(n),
                                     _capi.ptr(out, C.c_double)) != 0:
        raise ValueError("lkl_dataset_poses: bad arguments")
    return out


@dataclass
class Dataset:
    meta: dict
    dims: tuple
    origin: tuple
    spacing: float
    phi: np.ndarray
    rec: dict              # xyz (nr, 3) f32, ele, svdw, vdwr f32, heavy u8
    cells: tuple           # cell_start, cell_atoms, (ncx, ncy, ncz), (c_ox, c_oy, c_oz, c_sp)
    lig: dict              # ele, svdw, vdwr f32, heavy u8
    coords: np.ndarray     # (n_pose, nl, 3) f32
    ref: np.ndarray        # (n_pose,) f64
    extra: dict = field(default_factory=dict)

    @property
    def n_pose(self):
        return self.coords.shape[0]

    def complex(self):
        """A synth.Complex of the dataset (the ligand at its first placement;
        the scores use the per-pose placements)."""
        def mol(xyz, p):
            n = xyz.shape[0]
            f = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731
            return synth.Mol(xyz=f(xyz), ele=f(p["ele"]), rad=f(p["vdwr"]), seps=f(p["svdw"]),
                             heavy=np.asarray(p["heavy"], np.uint8), hyd=np.ones(n, np.int32),
                             des=np.zeros(n), asa=np.zeros(n))
        cx = synth.Complex(rec=mol(self.rec["xyz"], self.rec), lig=mol(self.coords[0], self.lig),
                           meta={"dataset": self.meta})
        cx.set_field(self.phi, self.origin, self.dims)
        return cx

    def score(self, ctx=None):
        """dna scores of every placement on the GPU (lkl_ctx_score_coords)."""
        from . import Context, family_flags, fold_scores
        if self.spacing != 0.5:
            raise ValueError("the far field must have 0.5 A spacing (grid_dna.rs:29)")
        own = ctx is None
        ctx = ctx or Context(self.complex())
        try:
            e, v, _ = ctx.score_coords(self.coords, family_flags("dna"))
        finally:
            if own:
                ctx.close()
        return fold_scores("dna", e, v)


def write(path, cx, poses, ref, cpu_grid_ms=None, cpu_exact_ms=None):
    """Write the dataset of complex `cx` (with its far field) for `poses`
    (N x 7) and per-pose energies `ref` in the reference's layout."""
    os.makedirs(path, exist_ok=True)
    r, l = cx.rec, cx.lig
    coords = placed(l.xyz, poses)
    n = coords.shape[0]
    ref = np.asarray(ref, np.float64).reshape(-1)
    if ref.shape[0] != n:
        raise ValueError("one energy per pose")
    i32 = lambda *v: np.asarray(v, "<i4").tobytes()  # noqa: E731
    f32 = lambda a: np.asarray(a, "<f4").tobytes()  # noqa: E731
    f64 = lambda a: np.asarray(a, "<f8").tobytes()  # noqa: E731
    u8 = lambda a: np.asarray(a, np.uint8).tobytes()  # noqa: E731
    dims, origin = cx.dims, cx.origin
    with open(os.path.join(path, "field.bin"), "wb") as f:
        f.write(i32(*dims) + f64(origin) + f64([0.5]) + f32(cx.phi))
    start, atoms, nc, co = cell_list(r.xyz)
    with open(os.path.join(path, "rec.bin"), "wb") as f:
        f.write(i32(r.n) + f32(r.xyz) + f32(r.ele) + f32(r.seps) + f32(r.rad) + u8(r.heavy) + i32(*nc)
                + f32(co) + np.asarray(start, "<i4").tobytes() + np.asarray(atoms, "<i4").tobytes())
    with open(os.path.join(path, "lig.bin"), "wb") as f:
        f.write(i32(l.n) + f32(l.ele) + f32(l.seps) + f32(l.rad) + u8(l.heavy))
    with open(os.path.join(path, "poses.bin"), "wb") as f:
        f.write(i32(n, l.n) + f32(coords))
    with open(os.path.join(path, "ref.bin"), "wb") as f:
        f.write(i32(n) + f64(ref))
    meta = {"nx": dims[0], "ny": dims[1], "nz": dims[2], "ox": origin[0], "oy": origin[1],
            "oz": origin[2], "spacing": 0.5, "phi_len": int(np.asarray(cx.phi).size), "nr": r.n,
            "nl": l.n, "n_pose": n, "ncx": nc[0], "ncy": nc[1], "ncz": nc[2], "c_ox": co[0],
            "c_oy": co[1], "c_oz": co[2], "c_sp": co[3], "cell_start_len": int(start.size),
            "cpu_grid_ms": cpu_grid_ms, "cpu_exact_ms": cpu_exact_ms}
    with open(os.path.join(path, "meta.json"), "w") as f:
        json.dump(meta, f)
    return meta


class _Reader:
    def __init__(self, path):
        with open(path, "rb") as f:
            self.b = f.read()
        self.o = 0
        self.path = path

    def take(self, dtype, n):
        dt = np.dtype(dtype)
        end = self.o + dt.itemsize * n
        if end > len(self.b):
            raise ValueError("%s: truncated (need %d bytes, have %d)" % (self.path, end, len(self.b)))
        a = np.frombuffer(self.b, dt, n, self.o).copy()
        self.o = end
        return a

    def done(self):
        if self.o != len(self.b):
            raise ValueError("%s: %d trailing bytes" % (self.path, len(self.b) - self.o))


def read(path):
    """Read a dataset directory written by the reference (or by `write`)."""
    with open(os.path.join(path, "meta.json")) as f:
        meta = json.load(f)
    fb = _Reader(os.path.join(path, "field.bin"))
    dims = tuple(int(v) for v in fb.take("<i4", 3))
    origin = tuple(float(v) for v in fb.take("<f8", 3))
    spacing = float(fb.take("<f8", 1)[0])
    phi = fb.take("<f4", int(np.prod(dims)))
    fb.done()
    rb = _Reader(os.path.join(path, "rec.bin"))
    nr = int(rb.take("<i4", 1)[0])
    rec = {"xyz": rb.take("<f4", 3 * nr).reshape(nr, 3), "ele": rb.take("<f4", nr),
           "svdw": rb.take("<f4", nr), "vdwr": rb.take("<f4", nr), "heavy": rb.take("u1", nr)}
    nc = tuple(int(v) for v in rb.take("<i4", 3))
    co = tuple(float(v) for v in rb.take("<f4", 4))
    start = rb.take("<i4", (nc[0] + 1) * (nc[1] + 1) * (nc[2] + 1) + 1)
    atoms = rb.take("<i4", nr)
    rb.done()
    lb = _Reader(os.path.join(path, "lig.bin"))
    nl = int(lb.take("<i4", 1)[0])
    lig = {"ele": lb.take("<f4", nl), "svdw": lb.take("<f4", nl), "vdwr": lb.take("<f4", nl),
           "heavy": lb.take("u1", nl)}
    lb.done()
    pb = _Reader(os.path.join(path, "poses.bin"))
    n, nl2 = (int(v) for v in pb.take("<i4", 2))
    if nl2 != nl:
        raise ValueError("poses.bin holds %d ligand atoms, lig.bin %d" % (nl2, nl))
    coords = pb.take("<f4", n * nl * 3).reshape(n, nl, 3)
    pb.done()
    eb = _Reader(os.path.join(path, "ref.bin"))
    if int(eb.take("<i4", 1)[0]) != n:
        raise ValueError("ref.bin and poses.bin disagree on the pose count")
    ref = eb.take("<f8", n)
    eb.done()
    for k, v in (("nr", nr), ("nl", nl), ("n_pose", n), ("cell_start_len", start.size),
                 ("phi_len", phi.size)):
        if k in meta and int(meta[k]) != v:
            raise ValueError("meta.json %s = %s, files hold %d" % (k, meta[k], v))
    return Dataset(meta=meta, dims=dims, origin=origin, spacing=spacing, phi=phi, rec=rec,
                   cells=(start, atoms, nc, co), lig=lig, coords=coords, ref=ref)
