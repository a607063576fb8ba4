"""lklight — MI355X (gfx950) batched GSO pose scoring for LKlight's all-atom families.

Python view of the C ABI in include/lklight_gpu.h.  Every scoring call runs the
HIP kernels of liblklight_gpu.so; if the library is missing, or no gfx950
device is present, calls raise instead of computing anything on the CPU.
"""
import ctypes as C

import numpy as np

from . import _capi
from ._capi import (F_CLASH, F_DESOLV, F_ELEC, F_FAR, FLAGS_DNA, FLAGS_PYDOCK, FLAGS_VDW,
                    GpuError, LibraryMissing, lib, last_error)

FACTOR, EPSILON = 332.0, 4.0

__all__ = ["Context", "Multi", "device_available", "lib", "last_error", "GpuError", "LibraryMissing",
           "family_flags", "fold_scores", "tuned_tiles", "fold_restraints", "F_FAR", "F_ELEC", "F_CLASH", "F_DESOLV",
           "FLAGS_DNA", "FLAGS_PYDOCK", "FLAGS_VDW"]


def tuned_tiles(key: str):
    """Stored autotuner choice for a workload key ("c2-dna", "c3-cpydock", "c4",
    ...) from tiles.json beside this module, or None."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiles.json")) as f:
        return json.load(f).get(key)


def device_available() -> bool:
    """lkl_device_available (replaces gpu_score.rs:224-232 cuda_available)."""
    return bool(lib().lkl_device_available())


def family_flags(method: str) -> int:
    """gpu_family.rs:187-194: dna -> 7, pydock/cpydock -> 3, vdw -> 0, other -> 7."""
    return {"dna": FLAGS_DNA, "pydock": FLAGS_PYDOCK, "cpydock": FLAGS_PYDOCK,
            "vdw": FLAGS_VDW}.get(method, FLAGS_DNA)


MEMBRANE_PENALTY = 999.0  # constants.rs:21


def _csr(residues):
    """List of atom-index lists -> (offsets, atoms) int32 arrays."""
    off = np.zeros(len(residues) + 1, np.int32)
    off[1:] = np.cumsum([len(r) for r in residues])
    atoms = np.asarray([a for r in residues for a in r], np.int32)
    return off, atoms


def fold_restraints(scores, counts, n_rec_res, n_lig_res, n_membrane):
    """dna.rs:989-995 on folded scores: score += perc_r*score + perc_l*score -
    penalty, with perc = satisfied residues / residues (0 without residues)
    and penalty = 999 * beads at the interface / beads when > 0
    (scoring.rs:83-110).  counts: N x 3 from Context.interface()."""
    s = np.asarray(scores, np.float64)
    c = np.asarray(counts).reshape(-1, 3)
    perc_r = c[:, 0] / n_rec_res if n_rec_res else np.zeros(len(s))
    perc_l = c[:, 1] / n_lig_res if n_lig_res else np.zeros(len(s))
    inter = c[:, 2] / n_membrane if n_membrane else np.zeros(len(s))
    penalty = np.where(inter > 0.0, MEMBRANE_PENALTY * inter, 0.0)
    return s + ((perc_r * s + perc_l * s) - penalty)


def lr_sasa(xyz, radii, probe=1.4, n_slices=20):
    """Lee & Richards solvent-accessible area per atom (A^2) on the GPU
    (lkl_lr_sasa; the reference's lr_sasa.rs:277-291 with its probe 1.4 and 20
    slices by default).  xyz: (n, 3), radii without the probe."""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    radii = np.ascontiguousarray(radii, np.float64)
    if radii.shape != (xyz.shape[0],):
        raise ValueError("radii must hold one value per atom")
    out = np.zeros(xyz.shape[0])
    _capi.check(lib().lkl_lr_sasa(_capi.ptr(xyz, C.c_double), _capi.ptr(radii, C.c_double),
                                  xyz.shape[0], float(probe), int(n_slices),
                                  _capi.ptr(out, C.c_double)), "lkl_lr_sasa")
    return out


def fold_scores(method: str, e_raw, v, s=None):
    """Per-pose scores from kernel parts: -(E*332/4 + V) (gpu_score.rs:326-331,
    gpu_family.rs:283-288); cpydock -(E*332/4 + 0.1 V) + S (cpydock.rs:789-791)."""
    e = np.asarray(e_raw) * FACTOR / EPSILON
    if method == "cpydock":
        return -(e + 0.1 * np.asarray(v)) + np.asarray(s)
    return -(e + np.asarray(v))


class Context:
    """One uploaded complex (lkl_ctx).  Host arrays are copied at creation."""

    def __init__(self, cx, with_field=True, desolv=False, cell_size=0.0, device_field=False):
        L = lib()
        self._keep = []
        r, l = cx.rec, cx.lig

        def f(a):
            a = np.ascontiguousarray(a, np.float32)
            self._keep.append(a)
            return _capi.ptr(a, C.c_float)

        def u8(a):
            a = np.ascontiguousarray(a, np.uint8)
            self._keep.append(a)
            return _capi.ptr(a, C.c_ubyte)

        d = _capi.ComplexDesc()
        if with_field:
            if cx.phi is None:
                raise ValueError("complex has no far field; call build_field() or set_field()")
            d.phi = f(cx.phi)
            d.nx, d.ny, d.nz = cx.dims
            d.ox, d.oy, d.oz = cx.origin
            d.sp = 0.5
        d.nr = r.n
        d.r_xyz, d.r_ele, d.r_svdw, d.r_vdwr = f(r.xyz), f(r.ele), f(r.seps), f(r.rad)
        d.r_heavy = u8(r.heavy)
        d.nl = l.n
        d.l_xyz, d.l_ele, d.l_svdw, d.l_vdwr = f(l.xyz), f(l.ele), f(l.seps), f(l.rad)
        d.l_heavy = u8(l.heavy)
        if desolv:
            d.r_flag, d.r_des, d.r_asa = u8(r.flag), f(r.des), f(r.asa)
            d.l_flag, d.l_des, d.l_asa = u8(l.flag), f(l.des), f(l.asa)
        d.cell_size = float(cell_size)
        h = L.lkl_ctx_create(C.byref(d))
        self._keep = []
        if not h:
            raise GpuError(f"lkl_ctx_create failed: {last_error()}")
        self.h = h
        self.nl, self.nr = l.n, r.n
        if device_field:
            self.build_field()

    def build_field(self):
        """lkl_ctx_build_field: phi built on the device, no host copy."""
        _capi.check(lib().lkl_ctx_build_field(self.h), "lkl_ctx_build_field")
        o = np.zeros(3)
        d = np.zeros(3, np.int32)
        lib().lkl_ctx_field_info(self.h, _capi.ptr(o, C.c_double), _capi.ptr(d, C.c_int))
        self.field_origin, self.field_dims = tuple(o), tuple(int(v) for v in d)

    def close(self):
        if getattr(self, "h", None):
            lib().lkl_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def score(self, poses, flags):
        """Blocking: returns (E_raw, V, S) arrays (S None without F_DESOLV)."""
        p = _capi.as_f64(poses).reshape(-1, 7)
        n = p.shape[0]
        out = np.zeros((n, 2), np.float64)
        s = np.zeros(n, np.float64) if flags & F_DESOLV else None
        _capi.check(lib().lkl_ctx_score(self.h, _capi.ptr(p, C.c_double), n, flags,
                                        _capi.ptr(out, C.c_double), _capi.ptr(s, C.c_double)),
                    "lkl_ctx_score")
        return out[:, 0].copy(), out[:, 1].copy(), s

    def set_timing(self, on):
        """lkl_ctx_set_timing: HIP events around every launch's kernels."""
        _capi.check(lib().lkl_ctx_set_timing(self.h, int(bool(on))), "lkl_ctx_set_timing")

    def kernel_ms(self):
        """(pre-pass ms, scoring-kernel ms) of the last launch (timing on)."""
        a, b = C.c_double(), C.c_double()
        _capi.check(lib().lkl_ctx_kernel_ms(self.h, C.byref(a), C.byref(b)), "lkl_ctx_kernel_ms")
        return a.value, b.value

    def set_pose_order(self, on):
        """lkl_ctx_set_pose_order: longest-first dispatch of large launches."""
        _capi.check(lib().lkl_ctx_set_pose_order(self.h, int(bool(on))), "lkl_ctx_set_pose_order")

    def score_coords(self, coords, flags):
        """lkl_ctx_score_coords: ligands already placed, coords (N, nl, 3) in
        the descriptor's ligand order -> (E_raw, V, S)."""
        xyz = _capi.as_f32(coords)
        n = xyz.size // (3 * self.nl) if self.nl else 0
        if xyz.size != n * 3 * self.nl or n == 0:
            raise ValueError("coords must hold N x nl x 3 values (nl = %d)" % self.nl)
        out = np.zeros((n, 2), np.float64)
        s = np.zeros(n, np.float64) if flags & F_DESOLV else None
        _capi.check(lib().lkl_ctx_score_coords(self.h, _capi.ptr(xyz, C.c_float), n, flags,
                                               _capi.ptr(out, C.c_double), _capi.ptr(s, C.c_double)),
                    "lkl_ctx_score_coords")
        return out[:, 0].copy(), out[:, 1].copy(), s

    def score_async(self, d_poses, n, flags, d_out, d_outS=None, stream=None):
        """Device pointers (ints), enqueued on `stream` (hipStream_t as int)."""
        _capi.check(lib().lkl_ctx_score_async(self.h, C.c_void_p(d_poses), n, flags,
                                              C.c_void_p(d_out),
                                              C.c_void_p(d_outS) if d_outS else None,
                                              C.c_void_p(stream) if stream else None),
                    "lkl_ctx_score_async")

    def transform(self, poses):
        p = _capi.as_f64(poses).reshape(-1, 7)
        out = np.zeros((p.shape[0], self.nl, 3), np.float64)
        _capi.check(lib().lkl_ctx_transform(self.h, _capi.ptr(p, C.c_double), p.shape[0],
                                            _capi.ptr(out, C.c_double)), "lkl_ctx_transform")
        return out

    def autotune(self, poses, flags):
        """lkl_ctx_autotune -> (cell, cell_x, block, chunk,
        table[(cell, cell_x, block, chunk, ms)])."""
        p = _capi.as_f64(poses).reshape(-1, 7)
        cell, cx, block, chunk, n = C.c_float(0), C.c_float(0), C.c_int(0), C.c_int(0), C.c_int(0)
        table = np.zeros((320, 5), np.float32)
        _capi.check(lib().lkl_ctx_autotune(self.h, _capi.ptr(p, C.c_double), p.shape[0], flags,
                                           C.byref(cell), C.byref(cx), C.byref(block), C.byref(chunk),
                                           _capi.ptr(table, C.c_float), C.byref(n)),
                    "lkl_ctx_autotune")
        return cell.value, cx.value, block.value, chunk.value, table[:n.value].tolist()

    def set_work_sort(self, on):
        """lkl_ctx_set_work_sort: order each tile's near field by expected work."""
        _capi.check(lib().lkl_ctx_set_work_sort(self.h, int(bool(on))), "lkl_ctx_set_work_sort")

    def set_list_spacing(self, spacing):
        """lkl_ctx_set_list_spacing: neighbour-list node spacing (A); 0 = cell-row walk."""
        _capi.check(lib().lkl_ctx_set_list_spacing(self.h, float(spacing)), "lkl_ctx_set_list_spacing")

    def set_interface(self, rec_residues=(), lig_residues=(), membrane=()):
        """lkl_ctx_set_interface: restraint residues (lists of atom indices, reference
        order) of each side and the receptor membrane beads."""
        ro, ra = _csr(list(rec_residues))
        lo, la = _csr(list(lig_residues))
        m = np.asarray(list(membrane), np.int32)
        self._iface = (ro, ra, lo, la, m)
        _capi.check(lib().lkl_ctx_set_interface(self.h, _capi.ptr(ro, C.c_int), _capi.ptr(ra, C.c_int),
                                                len(ro) - 1, _capi.ptr(lo, C.c_int), _capi.ptr(la, C.c_int),
                                                len(lo) - 1, _capi.ptr(m, C.c_int), len(m)),
                    "lkl_ctx_set_interface")
        self.iface_sizes = (len(ro) - 1, len(lo) - 1, len(m))

    def interface(self, poses):
        """lkl_ctx_interface -> counts (N x 3 int32): satisfied receptor residues,
        satisfied ligand residues, membrane beads at the interface."""
        p = _capi.as_f64(poses).reshape(-1, 7)
        out = np.zeros((p.shape[0], 3), np.int32)
        _capi.check(lib().lkl_ctx_interface(self.h, _capi.ptr(p, C.c_double), p.shape[0],
                                            _capi.ptr(out, C.c_int)), "lkl_ctx_interface")
        return out

    def set_phi_packed(self, on):
        """lkl_ctx_set_phi_packed: keep the corner-packed far field (8 corners per cell)."""
        _capi.check(lib().lkl_ctx_set_phi_packed(self.h, int(bool(on))), "lkl_ctx_set_phi_packed")

    def phi_packed(self):
        return bool(lib().lkl_ctx_phi_packed(self.h))

    def set_param_types(self, level):
        """lkl_ctx_set_param_types: LDS type tables of the list walk (0 none,
        1 the receptor's, 2 pair types when the ligand has types too)."""
        _capi.check(lib().lkl_ctx_set_param_types(self.h, int(level)), "lkl_ctx_set_param_types")

    def param_types(self):
        """Receptor parameter types in use by the list walk (0 = none)."""
        return int(lib().lkl_ctx_param_types(self.h))

    def list_info(self):
        """(list entries, node grid dims); entries 0 = lists off."""
        d = np.zeros(3, np.int32)
        n = lib().lkl_ctx_list_info(self.h, _capi.ptr(d, C.c_int))
        return int(n), tuple(int(v) for v in d)

    def set_tiles(self, cell_size, block, chunk=0, cell_x=0.0):
        """Cell edge (y, z; and x unless cell_x > 0), workgroup size, ligand atoms per tile."""
        _capi.check(lib().lkl_ctx_set_cell_x(self.h, float(cell_x)), "lkl_ctx_set_cell_x")
        _capi.check(lib().lkl_ctx_set_tiles(self.h, float(cell_size), int(block), int(chunk)),
                    "lkl_ctx_set_tiles")


class Multi:
    """Many complexes scored by one kernel launch per step (lkl_multi_*)."""

    def __init__(self, ctxs, flags, n_poses):
        K = len(ctxs)
        hs = (C.c_void_p * K)(*[c.h for c in ctxs])
        fl = (C.c_uint * K)(*[int(f) for f in flags])
        npz = np.ascontiguousarray(n_poses, np.int32)
        h = lib().lkl_multi_create(hs, fl, _capi.ptr(npz, C.c_int), K)
        if not h:
            raise GpuError(f"lkl_multi_create failed: {last_error()}")
        self.h, self.ctxs, self.P = h, list(ctxs), int(npz.sum())

    def close(self):
        if getattr(self, "h", None):
            lib().lkl_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def score(self, poses):
        p = _capi.as_f64(poses).reshape(-1, 7)
        assert p.shape[0] == self.P
        out = np.zeros((self.P, 2))
        _capi.check(lib().lkl_multi_score(self.h, _capi.ptr(p, C.c_double), _capi.ptr(out, C.c_double)),
                    "lkl_multi_score")
        return out

    def score_async(self, d_poses, d_out, stream=None):
        _capi.check(lib().lkl_multi_score_async(self.h, C.c_void_p(d_poses), C.c_void_p(d_out),
                                                C.c_void_p(stream) if stream else None),
                    "lkl_multi_score_async")

    def use_graph(self, on):
        lib().lkl_multi_use_graph(self.h, int(bool(on)))

    def set_trace(self, on):
        """lkl_multi_set_trace: record per-tile start/end, CU and problem of each launch."""
        _capi.check(lib().lkl_multi_set_trace(self.h, int(bool(on))), "lkl_multi_set_trace")

    def trace(self):
        """The last traced launch: (records uint64 [tiles, 4] = start, end, CU id, problem;
        wall-clock rate in kHz)."""
        khz = C.c_int()
        rec = np.zeros((1, 4), np.uint64)
        n = lib().lkl_multi_trace(self.h, _capi.ptr(rec, C.c_uint64), 0, C.byref(khz))  # the tile count
        if n >= 0:
            rec = np.zeros((max(n, 1), 4), np.uint64)
            n = lib().lkl_multi_trace(self.h, _capi.ptr(rec, C.c_uint64), n, C.byref(khz))
        if n < 0:
            raise _capi.GpuError("lkl_multi_trace failed: " + _capi.last_error())
        return rec[:n], khz.value

    def set_tiles(self, block, chunk_mult):
        _capi.check(lib().lkl_multi_set_tiles(self.h, int(block), int(chunk_mult)),
                    "lkl_multi_set_tiles")

    def autotune(self, poses):
        p = _capi.as_f64(poses).reshape(-1, 7)
        b, m, n = C.c_int(), C.c_int(), C.c_int()
        table = np.zeros((64, 3), np.float32)
        _capi.check(lib().lkl_multi_autotune(self.h, _capi.ptr(p, C.c_double), C.byref(b), C.byref(m),
                                             _capi.ptr(table, C.c_float), C.byref(n)),
                    "lkl_multi_autotune")
        return b.value, m.value, table[:n.value].tolist()


def load_balance(rec, clock_khz):
    """Load balance of one traced batch launch (Multi.trace()): per CU the union
    of its tiles' [start, end) intervals is its busy time.  Returns the span
    (first start to last end), the busy fraction (sum of CU busy times over
    CUs x span), the mean number of tiles in flight per CU, and the spread of
    the CUs' last finish times, all times in us."""
    r = np.asarray(rec, dtype=np.int64)
    if r.shape[0] == 0:
        return None
    us = 1000.0 / float(clock_khz)
    start, end, cu = r[:, 0], r[:, 1], r[:, 2]
    t0, t1 = int(start.min()), int(end.max())
    span = float(t1 - t0)
    busy, fin = [], []
    for c in np.unique(cu):
        m = cu == c
        iv = sorted(zip(start[m].tolist(), end[m].tolist()))
        tot, cs, ce = 0, iv[0][0], iv[0][1]
        for a, b in iv[1:]:
            if a > ce:
                tot += ce - cs
                cs, ce = a, b
            else:
                ce = max(ce, b)
        tot += ce - cs
        busy.append(tot)
        fin.append(int(end[m].max()) - t0)
    busy, fin = np.array(busy, float), np.array(fin, float)
    ncu = len(busy)
    return {"tiles": int(r.shape[0]), "cus": ncu, "span_us": span * us,
            "busy_frac": float(busy.sum() / (ncu * span)),
            "tiles_in_flight_per_cu": float((end - start).sum() / (ncu * span)),
            "cu_finish_us": {"min": float(fin.min()) * us, "median": float(np.median(fin)) * us,
                             "max": float(fin.max()) * us},
            "tile_us": {"median": float(np.median(end - start)) * us, "max": float((end - start).max()) * us}}
