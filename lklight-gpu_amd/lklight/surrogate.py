"""Learned surrogate of the exact pose score.

Pair-histogram features (16 receptor type groups x 16 ligand type groups x 24
distance bins over 0-30 A, log1p, bf16), a bf16 MLP 6144 -> 512 -> 512 -> 512
-> 1 trained with Adam, all computed by liblklight_gpu.so (lkl_feat_*,
lkl_mlp_*).  There is no counterpart in the reference: the surrogate only ranks
candidate GSO moves; every accepted move is exact-scored by the scoring kernel.
"""
import ctypes as C

import numpy as np

from . import _capi

NF, HIDDEN = 6144, 512


class Features:
    """Feature extractor of one complex (receptor + ligand type groups).

    grid=None: exact pair histograms.  grid=h (A): the receptor's counts are
    tabulated on a node grid of spacing h and each ligand atom takes its
    nearest node's counts (lkl_feat_create_grid)."""

    def __init__(self, cx, grid=None):
        L = _capi.lib()
        self._keep = [cx.rec.f32("xyz"), np.ascontiguousarray(cx.rec.group, np.uint8),
                      cx.lig.f32("xyz"), np.ascontiguousarray(cx.lig.group, np.uint8)]
        rx, rg, lx, lg = self._keep
        args = (_capi.ptr(rx, C.c_float), _capi.ptr(rg, C.c_ubyte), rx.shape[0],
                _capi.ptr(lx, C.c_float), _capi.ptr(lg, C.c_ubyte), lx.shape[0])
        self.grid = grid
        if grid is None:
            self.h = L.lkl_feat_create(*args)
        else:
            self.h = L.lkl_feat_create_grid(*args, float(grid))
        if not self.h:
            raise _capi.GpuError(f"lkl_feat_create failed: {_capi.last_error()}")

    def grid_info(self):
        o = np.zeros(3)
        d = np.zeros(3, np.int32)
        sp = np.zeros(1)
        n = _capi.lib().lkl_feat_grid_info(self.h, _capi.ptr(o, C.c_double), _capi.ptr(d, C.c_int),
                                           _capi.ptr(sp, C.c_double))
        return o, tuple(int(v) for v in d), float(sp[0]), int(n)

    def grid_nodes(self, first, count):
        out = np.empty((count, 384), np.uint8)
        _capi.check(_capi.lib().lkl_feat_grid_read(self.h, int(first), int(count),
                                                   _capi.ptr(out, C.c_ubyte)), "lkl_feat_grid_read")
        return out

    def _host(self, poses, want_bits, want_counts):
        P = _capi.as_f64(poses).reshape(-1, 7)
        n = P.shape[0]
        bits = np.empty((n, NF), np.uint16) if want_bits else None
        cnt = np.empty((n, NF), np.uint32) if want_counts else None
        _capi.check(_capi.lib().lkl_feat_compute_host(
            self.h, _capi.ptr(P, C.c_double), n, _capi.ptr(bits, C.c_ushort),
            _capi.ptr(cnt, C.c_uint)), "lkl_feat_compute_host")
        return bits, cnt

    def counts(self, poses):
        return self._host(poses, False, True)[1]

    def bits(self, poses):
        return self._host(poses, True, False)[0]

    def compute_device(self, d_poses, d_out, stream=0):
        """d_poses: device (N, 7) float64; d_out: device (N, 6144) 16-bit."""
        n = int(d_poses.shape[0])
        _capi.check(_capi.lib().lkl_feat_compute(self.h, d_poses.data_ptr(), n, d_out.data_ptr(),
                                                 None, stream or None), "lkl_feat_compute")

    def close(self):
        if self.h:
            _capi.lib().lkl_feat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MLP:
    """The surrogate network on the GPU (fp32 master weights + Adam state)."""

    def __init__(self, seed=0):
        L = _capi.lib()
        self.h = L.lkl_mlp_create(int(seed) & 0xFFFFFFFFFFFFFFFF)
        if not self.h:
            raise _capi.GpuError(f"lkl_mlp_create failed: {_capi.last_error()}")
        self.nparam = int(L.lkl_mlp_param_count())

    @property
    def params(self):
        out = np.empty(self.nparam, np.float32)
        _capi.check(_capi.lib().lkl_mlp_get_params(self.h, _capi.ptr(out, C.c_float)), "get_params")
        return out

    @params.setter
    def params(self, p):
        p = _capi.as_f32(p)
        assert p.shape == (self.nparam,)
        _capi.check(_capi.lib().lkl_mlp_set_params(self.h, _capi.ptr(p, C.c_float)), "set_params")

    def grads(self):
        out = np.empty(self.nparam, np.float32)
        _capi.check(_capi.lib().lkl_mlp_get_grads(self.h, _capi.ptr(out, C.c_float)), "get_grads")
        return out

    def forward(self, xbits):
        x = np.ascontiguousarray(xbits, np.uint16)
        y = np.empty(x.shape[0], np.float32)
        _capi.check(_capi.lib().lkl_mlp_forward_host(self.h, _capi.ptr(x, C.c_ushort), x.shape[0],
                                                     _capi.ptr(y, C.c_float)), "lkl_mlp_forward_host")
        return y

    def grad(self, xbits, t):
        x = np.ascontiguousarray(xbits, np.uint16)
        t = _capi.as_f32(t)
        loss = np.zeros(1, np.float32)
        _capi.check(_capi.lib().lkl_mlp_grad_host(self.h, _capi.ptr(x, C.c_ushort),
                                                  _capi.ptr(t, C.c_float), x.shape[0],
                                                  _capi.ptr(loss, C.c_float)), "lkl_mlp_grad_host")
        return float(loss[0])

    def forward_device(self, d_x, d_y, stream=0):
        _capi.check(_capi.lib().lkl_mlp_forward(self.h, d_x.data_ptr(), int(d_x.shape[0]),
                                                d_y.data_ptr(), stream or None), "lkl_mlp_forward")

    def grad_device(self, d_x, d_t, want_loss=True):
        loss = np.zeros(1, np.float32)
        _capi.check(_capi.lib().lkl_mlp_grad(self.h, d_x.data_ptr(), d_t.data_ptr(),
                                             int(d_x.shape[0]),
                                             _capi.ptr(loss, C.c_float) if want_loss else None),
                    "lkl_mlp_grad")
        return float(loss[0]) if want_loss else None

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        _capi.check(_capi.lib().lkl_mlp_adam(self.h, lr, beta1, beta2, eps), "lkl_mlp_adam")

    def sync(self):
        _capi.check(_capi.lib().lkl_mlp_sync(self.h), "lkl_mlp_sync")

    def close(self):
        if self.h:
            _capi.lib().lkl_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gemm_bf16(a_bits, sam, sak, b_bits, sbk, sbn, M, N, K):
    """C = sum_k A(m, k) B(k, n) on the matrix cores (bf16 operand bits)."""
    A = np.ascontiguousarray(a_bits, np.uint16).ravel()
    B = np.ascontiguousarray(b_bits, np.uint16).ravel()
    out = np.empty((M, N), np.float32)
    _capi.check(_capi.lib().lkl_gemm_bf16_host(_capi.ptr(A, C.c_ushort), A.size, sam, sak,
                                               _capi.ptr(B, C.c_ushort), B.size, sbk, sbn, M, N, K,
                                               _capi.ptr(out, C.c_float)), "lkl_gemm_bf16_host")
    return out
