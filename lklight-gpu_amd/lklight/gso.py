"""GSO driver over the GPU scorer (lkl_gso_* of include/lklight_gpu.h)."""
import ctypes as C

import numpy as np

from . import _capi
from ._capi import GpuError, lib

FAMILY = {"dna": 0, "vdw": 1, "pydock": 2, "cpydock": 3}


class GSO:
    """n_swarms independent swarms (swarm_sizes), one StdRng(seed) per swarm.

    positions: rows (tx, ty, tz, w, x, y, z).  ctx: lklight.Context (GPU
    scoring) or None.  fallback(poses[n, 7]) -> scores[n]: the caller's CPU
    path, used only when the GPU call fails or ctx is None."""

    def __init__(self, ctx, method, positions, seed, swarm_sizes=None, fallback=None):
        pos = _capi.as_f64(positions).reshape(-1, 7)
        sizes = np.asarray(swarm_sizes if swarm_sizes is not None else [pos.shape[0]], np.int32)
        self._cb = None
        if fallback is not None:
            def _call(user, poses, n, scores):
                try:
                    p = np.ctypeslib.as_array(poses, shape=(n, 7)).copy()
                    out = np.ctypeslib.as_array(scores, shape=(n,))
                    out[:] = fallback(p)
                    return 0
                except Exception:  # reported as a failed scoring call
                    return -1
            self._cb = _capi.SCORE_FALLBACK(_call)
        h = lib().lkl_gso_create(ctx.h if ctx is not None else None, FAMILY[method], len(sizes),
                                 _capi.ptr(sizes, C.c_int), _capi.ptr(pos, C.c_double),
                                 int(seed), self._cb if self._cb else _capi.SCORE_FALLBACK(0), None)
        if not h:
            raise GpuError("lkl_gso_create failed (bad arguments or no scorer)")
        self.h, self.ctx, self.n = h, ctx, int(pos.shape[0])

    def close(self):
        if getattr(self, "h", None):
            lib().lkl_gso_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update_luciferin(self):
        _capi.check(lib().lkl_gso_update_luciferin(self.h), "lkl_gso_update_luciferin")

    def movement_phase(self):
        _capi.check(lib().lkl_gso_movement_phase(self.h), "lkl_gso_movement_phase")

    def run(self, steps):
        _capi.check(lib().lkl_gso_run(self.h, int(steps)), "lkl_gso_run")

    def device_run(self, steps, stream=None):
        """lkl_gso_device_run: `steps` full steps with the state on the device."""
        _capi.check(lib().lkl_gso_device_run(self.h, int(steps), C.c_void_p(stream) if stream else None),
                    "lkl_gso_device_run")

    def state(self):
        n = self.n
        pose = np.zeros((n, 7))
        luc, sc, vis = np.zeros(n), np.zeros(n), np.zeros(n)
        nnb = np.zeros(n, np.int32)
        mv = np.zeros(n, np.uint8)
        _capi.check(lib().lkl_gso_state(self.h, _capi.ptr(pose, C.c_double), _capi.ptr(luc, C.c_double),
                                        _capi.ptr(sc, C.c_double), _capi.ptr(vis, C.c_double),
                                        _capi.ptr(nnb, C.c_int), _capi.ptr(mv, C.c_ubyte)),
                    "lkl_gso_state")
        return {"pose": pose, "luciferin": luc, "scoring": sc, "vision": vis, "neighbors": nnb,
                "moved": mv}

    def set_poses(self, pose7, mask=None):
        p = _capi.as_f64(pose7).reshape(-1, 7)
        m = None if mask is None else _capi.as_u8(mask)
        _capi.check(lib().lkl_gso_set_poses(self.h, _capi.ptr(p, C.c_double),
                                            _capi.ptr(m, C.c_ubyte)), "lkl_gso_set_poses")

    def plan(self):
        """Plan the step's GSO moves (same draws as movement_phase) without
        applying them; returns the planned poses (n, 7) as (w, x, y, z, tx, ty, tz)."""
        out = np.zeros((self.n, 7))
        _capi.check(lib().lkl_gso_plan(self.h, _capi.ptr(out, C.c_double)), "lkl_gso_plan")
        return out

    def commit(self, pose7, scores, accept):
        p = _capi.as_f64(pose7).reshape(-1, 7)
        s = _capi.as_f64(scores)
        a = _capi.as_u8(accept)
        _capi.check(lib().lkl_gso_commit(self.h, _capi.ptr(p, C.c_double), _capi.ptr(s, C.c_double),
                                         _capi.ptr(a, C.c_ubyte)), "lkl_gso_commit")

    def stats(self):
        a, b, c = C.c_longlong(), C.c_longlong(), C.c_longlong()
        lib().lkl_gso_stats(self.h, C.byref(a), C.byref(b), C.byref(c))
        return {"gpu_launches": a.value, "fallback_calls": b.value, "gpu_scored": c.value}

    def save(self, swarm, step, directory):
        _capi.check(lib().lkl_gso_save(self.h, int(swarm), int(step), directory.encode()),
                    "lkl_gso_save")


def rng_f64s(seed, n):
    out = np.zeros(n)
    lib().lkl_rng_f64s(int(seed), int(n), _capi.ptr(out, C.c_double))
    return out


def slerp(a, b, t):
    a, b = _capi.as_f64(a), _capi.as_f64(b)
    out = np.zeros(4)
    lib().lkl_quat_slerp(_capi.ptr(a, C.c_double), _capi.ptr(b, C.c_double), float(t),
                         _capi.ptr(out, C.c_double))
    return out
