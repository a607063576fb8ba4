"""ctypes binding of the CPU oracle (oracle/_build/liblkl_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker of the HIP path.  Imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liblkl_oracle.so")

DNA, VDW, PYDOCK, CPYDOCK = 0, 1, 2, 3
FAMILY = {"dna": DNA, "vdw": VDW, "pydock": PYDOCK, "cpydock": CPYDOCK}

_d = C.POINTER(C.c_double)


class Mol(C.Structure):
    _fields_ = [("n", C.c_int), ("xyz", _d), ("ele", _d), ("eps", _d), ("seps", _d),
                ("rad", _d), ("heavy", C.POINTER(C.c_ubyte)), ("hyd", C.POINTER(C.c_int)),
                ("des", _d), ("asa", _d)]


class Field(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("n", C.c_int64 * 3), ("spacing", C.c_double),
                ("phi", C.POINTER(C.c_float))]


class Parts(C.Structure):
    _fields_ = [("e_raw", C.c_double), ("v", C.c_double), ("s", C.c_double),
                ("n_near", C.c_int64), ("n_far", C.c_int64), ("n_cand", C.c_int64),
                ("e_abs", C.c_double), ("f_abs", C.c_double), ("v_abs", C.c_double),
                ("c_abs", C.c_double), ("s_abs", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}


def build_native(out_dir):
    """Compile the same oracle with -march=native into out_dir (the host it
    runs on decides the ISA, so this is done on the machine that times it);
    returns the library path."""
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "liblkl_oracle_native.so")
    src = [os.path.join(HERE, f) for f in ("lkl_oracle.c", "lkl_oracle_gso.c", "lkl_oracle_sur.c",
                                           "lkl_oracle_sasa.c")]
    subprocess.run(["gcc", "-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
                    "-std=gnu11", "-shared", "-o", out] + src + ["-lm", "-lpthread"], check=True)
    return out


def lib(path=None):
    """The oracle library (default: oracle/_build/liblkl_oracle.so, -march=x86-64)."""
    path = path or LIB
    if path not in _libs:
        if path == LIB and not os.path.exists(LIB):
            build()
        L = C.CDLL(path)
        L.orc_field_dims.restype = C.c_int64
        L.orc_field_dims.argtypes = [_d, C.c_int, _d, C.POINTER(C.c_int64)]
        L.orc_field_fill.restype = None
        L.orc_field_fill.argtypes = [_d, _d, C.c_int, _d, C.POINTER(C.c_int64),
                                     C.POINTER(C.c_float)]
        L.orc_field_sample.restype = C.c_double
        L.orc_field_sample.argtypes = [C.POINTER(Field), C.c_double, C.c_double, C.c_double]
        L.orc_crec_build.restype = C.c_int
        L.orc_crec_build.argtypes = [_d, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_transform.restype = None
        L.orc_transform.argtypes = [_d, C.c_int, _d, _d]
        for name in ("orc_dna_energy_grid", "orc_pydock_energy_grid", "orc_cpydock_energy_grid"):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [C.POINTER(Mol), C.POINTER(Mol), C.POINTER(Field), _d, C.POINTER(Parts)]
        L.orc_vdw_energy_grid.restype = None
        L.orc_vdw_energy_grid.argtypes = [C.POINTER(Mol), C.POINTER(Mol), _d, C.POINTER(Parts)]
        L.orc_fold_score.restype = C.c_double
        L.orc_fold_score.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_batch_energy_grid.restype = None
        L.orc_batch_energy_grid.argtypes = [C.c_int, C.POINTER(Mol), C.POINTER(Mol),
                                            C.POINTER(Field), _d, C.c_int, _d, _d, _d,
                                            C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int]
        L.orc_batch_energy_grid_mag.restype = None
        L.orc_batch_energy_grid_mag.argtypes = [C.c_int, C.POINTER(Mol), C.POINTER(Mol),
                                                C.POINTER(Field), _d, C.c_int, _d, _d, _d,
                                                C.POINTER(C.c_int64), C.POINTER(C.c_int64), _d, C.c_int]
        L.orc_bruteforce_parts.restype = None
        L.orc_bruteforce_parts.argtypes = [C.c_int, C.POINTER(Mol), C.POINTER(Mol),
                                           C.POINTER(Field), _d, C.POINTER(Parts)]
        L.orc_dna_energy_exact.restype = C.c_double
        L.orc_dna_energy_exact.argtypes = [C.POINTER(Mol), C.POINTER(Mol), _d]
        L.orc_pydock_energy_exact.restype = C.c_double
        L.orc_pydock_energy_exact.argtypes = [C.POINTER(Mol), C.POINTER(Mol), _d]
        L.orc_cpydock_energy_exact.restype = C.c_double
        L.orc_cpydock_energy_exact.argtypes = [C.POINTER(Mol), C.POINTER(Mol), _d]
        L.orc_lr_sasa.restype = None
        L.orc_lr_sasa.argtypes = [_d, _d, C.c_int, C.c_double, C.c_int, _d]
        L.orc_rng_f64s.restype = None
        L.orc_rng_f64s.argtypes = [C.c_uint64, C.c_int, _d]
        L.orc_quat_random.restype = None
        L.orc_quat_random.argtypes = [C.c_uint64, C.c_int, _d]
        L.orc_slerp.restype = None
        L.orc_slerp.argtypes = [_d, _d, C.c_double, _d]
        L.orc_gso_run.restype = None
        L.orc_gso_run.argtypes = [C.c_int, C.POINTER(Mol), C.POINTER(Mol), C.POINTER(Field), _d,
                                  C.c_int, C.c_uint64, C.c_int, C.c_int, _d, _d, _d, _d,
                                  C.POINTER(C.c_int)]
        L.orc_feat_counts.restype = None
        L.orc_feat_counts.argtypes = [_d, C.POINTER(C.c_ubyte), C.c_int, _d, C.POINTER(C.c_ubyte),
                                      C.c_int, _d, C.c_int, C.POINTER(C.c_uint32), C.c_int]
        L.orc_feat_grid_dims.restype = C.c_int64
        L.orc_feat_grid_dims.argtypes = [_d, C.c_int, C.c_double, _d, C.POINTER(C.c_int)]
        L.orc_feat_grid_node.restype = None
        L.orc_feat_grid_node.argtypes = [_d, C.POINTER(C.c_ubyte), C.c_int, _d, C.c_double, C.c_int,
                                         C.c_int, C.c_int, C.POINTER(C.c_uint32)]
        L.orc_feat_grid_counts.restype = None
        L.orc_feat_grid_counts.argtypes = [_d, C.POINTER(C.c_ubyte), C.c_int, _d, C.POINTER(C.c_ubyte),
                                           C.c_int, _d, C.c_int, C.c_double, C.POINTER(C.c_uint32),
                                           C.c_int]
        L.orc_nearcells_pair_count.restype = C.c_int64
        L.orc_nearcells_pair_count.argtypes = [_d, C.c_int, C.c_double]
        L.orc_interface_flags.restype = None
        L.orc_interface_flags.argtypes = [C.POINTER(Mol), C.POINTER(Mol), _d, C.POINTER(C.c_ubyte),
                                          C.POINTER(C.c_ubyte)]
        L.orc_satisfied_restraints.restype = C.c_double
        L.orc_satisfied_restraints.argtypes = [C.POINTER(C.c_ubyte), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                               C.c_int]
        L.orc_membrane_intersection.restype = C.c_double
        L.orc_membrane_intersection.argtypes = [C.POINTER(C.c_ubyte), C.POINTER(C.c_int), C.c_int]
        L.orc_restraint_fold.restype = C.c_double
        L.orc_restraint_fold.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
        _libs[path] = L
    return _libs[path]


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class OMol:
    """Oracle view of a molecule (keeps the f64 arrays alive)."""

    def __init__(self, m):
        self.arrs = dict(
            xyz=np.ascontiguousarray(m.xyz, np.float64), ele=np.ascontiguousarray(m.ele, np.float64),
            eps=np.ascontiguousarray(m.eps, np.float64), seps=np.ascontiguousarray(m.seps, np.float64),
            rad=np.ascontiguousarray(m.rad, np.float64), heavy=np.ascontiguousarray(m.heavy, np.uint8),
            hyd=np.ascontiguousarray(m.hyd, np.int32), des=np.ascontiguousarray(m.des, np.float64),
            asa=np.ascontiguousarray(m.asa, np.float64))
        a = self.arrs
        self.s = Mol(m.xyz.shape[0], _p(a["xyz"]), _p(a["ele"]), _p(a["eps"]), _p(a["seps"]),
                     _p(a["rad"]), _p(a["heavy"], C.c_ubyte), _p(a["hyd"], C.c_int),
                     _p(a["des"]), _p(a["asa"]))


def field_build(xyz, q):
    """grid_dna.rs:44-149 -> (phi f32 flat, origin, dims)."""
    L = lib()
    xyz = np.ascontiguousarray(xyz, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    o = np.zeros(3, np.float64)
    d = np.zeros(3, np.int64)
    n = L.orc_field_dims(_p(xyz), xyz.shape[0], _p(o), _p(d, C.c_int64))
    phi = np.empty(int(n), np.float32)
    L.orc_field_fill(_p(xyz), _p(q), xyz.shape[0], _p(o), _p(d, C.c_int64), _p(phi, C.c_float))
    return phi, tuple(o), tuple(int(v) for v in d)


class OField:
    def __init__(self, phi, origin, dims):
        self.phi = np.ascontiguousarray(phi, np.float32)
        self.s = Field((C.c_double * 3)(*origin), (C.c_int64 * 3)(*dims), 0.5,
                       _p(self.phi, C.c_float))

    def sample(self, x, y, z):
        return lib().orc_field_sample(C.byref(self.s), x, y, z)


class Oracle:
    """CPU grid path of one complex."""

    def __init__(self, cx, lib_path=None):
        self.lib_path = lib_path
        self.rec, self.lig = OMol(cx.rec), OMol(cx.lig)
        self.field = OField(cx.phi, cx.origin, cx.dims) if cx.phi is not None else None

    def _fp(self):
        return C.byref(self.field.s) if self.field is not None else None

    def parts(self, method, pose):
        """One pose of `method`'s energy_grid -> Parts (e_raw, v, s, n_near, n_far)."""
        L = lib()
        p = np.ascontiguousarray(pose, np.float64)
        out = Parts()
        if method == "vdw":
            L.orc_vdw_energy_grid(C.byref(self.rec.s), C.byref(self.lig.s), _p(p), C.byref(out))
        else:
            fn = {"dna": L.orc_dna_energy_grid, "pydock": L.orc_pydock_energy_grid,
                  "cpydock": L.orc_cpydock_energy_grid}[method]
            fn(C.byref(self.rec.s), C.byref(self.lig.s), self._fp(), _p(p), C.byref(out))
        return out

    def batch(self, method, poses, nthreads=1, counts=False, bound=False):
        """Batched energy_grid -> dict of arrays e, v, s (+ n_near, n_far with
        counts; + "bound", the FP32 rounding bound of each folded score
        (fp32_bound), with bound)."""
        L = lib(self.lib_path)
        p = np.ascontiguousarray(poses, np.float64).reshape(-1, 7)
        n = p.shape[0]
        e, v, s = np.zeros(n), np.zeros(n), np.zeros(n)
        nn = np.zeros(n, np.int64)
        nf = np.zeros(n, np.int64)
        mag = np.zeros((n, 5))
        counts = counts or bound
        L.orc_batch_energy_grid_mag(FAMILY[method], C.byref(self.rec.s), C.byref(self.lig.s),
                                    self._fp() if method != "vdw" else None, _p(p), n, _p(e), _p(v),
                                    _p(s), _p(nn, C.c_int64) if counts else None,
                                    _p(nf, C.c_int64) if counts else None,
                                    _p(mag) if bound else None, int(nthreads))
        out = {"e": e, "v": v, "s": s}
        if counts:
            out["n_near"], out["n_far"] = nn, nf
        if bound:
            out["mag"] = mag
            out["bound"] = fp32_bound(method, mag, nn, nf, self.rec.s.n, self.lig.s.n)
        return out

    def bruteforce(self, method, pose):
        out = Parts()
        p = np.ascontiguousarray(pose, np.float64)
        lib().orc_bruteforce_parts(FAMILY[method], C.byref(self.rec.s), C.byref(self.lig.s),
                                   self._fp() if method != "vdw" else None, _p(p), C.byref(out))
        return out

    def gso(self, method, positions, seed, steps, nthreads=4):
        """lib.rs:47-78 with CPU-grid scoring; per-step arrays after each move."""
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 7)
        n = pos.shape[0]
        pose = np.zeros((steps, n, 7))
        luc, sc, vis = np.zeros((steps, n)), np.zeros((steps, n)), np.zeros((steps, n))
        nnb = np.zeros((steps, n), np.int32)
        lib().orc_gso_run(FAMILY[method], C.byref(self.rec.s), C.byref(self.lig.s),
                          self._fp() if method != "vdw" else None, _p(pos), n, int(seed),
                          int(steps), int(nthreads), _p(pose), _p(luc), _p(sc), _p(vis),
                          _p(nnb, C.c_int))
        return {"pose": pose, "luciferin": luc, "scoring": sc, "vision": vis, "neighbors": nnb}

    def interface(self, pose):
        """dna.rs:885-995 interface flags of one pose -> (iface_r, iface_l) u8 arrays."""
        p = np.ascontiguousarray(pose, np.float64)
        ir = np.zeros(self.rec.arrs["xyz"].shape[0], np.uint8)
        il = np.zeros(self.lig.arrs["xyz"].shape[0], np.uint8)
        lib().orc_interface_flags(C.byref(self.rec.s), C.byref(self.lig.s), _p(p), _p(ir, C.c_ubyte),
                                  _p(il, C.c_ubyte))
        return ir, il

    def restraint_score(self, score, pose, rec_residues, lig_residues, membrane):
        """The folded score of one pose with restraints / membrane (dna.rs:985-995)."""
        L = lib()
        ir, il = self.interface(pose)

        def csr(res):
            off = np.zeros(len(res) + 1, np.int32)
            off[1:] = np.cumsum([len(r) for r in res])
            atoms = np.asarray([a for r in res for a in r] or [0], np.int32)
            return off, atoms
        ro, ra = csr(rec_residues)
        lo, la = csr(lig_residues)
        m = np.asarray(list(membrane) or [0], np.int32)
        pr = L.orc_satisfied_restraints(_p(ir, C.c_ubyte), _p(ro, C.c_int), _p(ra, C.c_int), len(rec_residues))
        pl = L.orc_satisfied_restraints(_p(il, C.c_ubyte), _p(lo, C.c_int), _p(la, C.c_int), len(lig_residues))
        mi = L.orc_membrane_intersection(_p(ir, C.c_ubyte), _p(m, C.c_int), len(membrane))
        return L.orc_restraint_fold(float(score), pr, pl, mi)

    def dna_exact(self, pose):
        p = np.ascontiguousarray(pose, np.float64)
        return lib().orc_dna_energy_exact(C.byref(self.rec.s), C.byref(self.lig.s), _p(p))

    def pydock_exact(self, pose):
        p = np.ascontiguousarray(pose, np.float64)
        return lib().orc_pydock_energy_exact(C.byref(self.rec.s), C.byref(self.lig.s), _p(p))

    def cpydock_exact(self, pose):
        p = np.ascontiguousarray(pose, np.float64)
        return lib().orc_cpydock_energy_exact(C.byref(self.rec.s), C.byref(self.lig.s), _p(p))

    def transform(self, pose):
        p = np.ascontiguousarray(pose, np.float64)
        xyz = self.lig.arrs["xyz"]
        out = np.zeros_like(xyz)
        lib().orc_transform(_p(xyz), xyz.shape[0], _p(p), _p(out))
        return out


# ── FP32 rounding bound of the device's evaluation of the same sum ───────────
U32 = 2.0 ** -24  # unit roundoff of f32


def fp32_bound(method, mag, n_near, n_far, nr, nl):
    """Per-pose bound on |device score - oracle score| from the device's
    operation sequence (first order in u = 2^-24; the f64 oracle's own
    rounding is below 1e-15 relative and neglected), with the magnitudes the
    oracle accumulates (orc_parts):
      electrostatics, per pair: d = (x_hi - r) + x_lo (2 roundings), d2 by a
        product and two fmas (3), hardware reciprocal (1 ulp = 2u), two
        products  ->  |dE_ij| <= 11u |E_ij|, plus 2^-38 per pair for the
        fixed-point rounding (2^37 units);
      far field: f64 from the same f32 grid values as the CPU: 2^-45 of
        sum |q phi| for the evaluation order, 2^-38 per sampled atom;
      LJ, per pair: sqrt(eps_i) sqrt(eps_j) from f32 inputs against the
        CPU's epsilon (3u), r_i + r_j (u), t = (r_i + r_j)^2 / d2 (13u),
        p6 = t^3 (41u), p6^2 - 2 p6 and the product  ->  |dV_ij| <= 89u
        |eps_ij| (p6^2 + 2 p6) over the pairs not clearly capped (the cap is
        exact), plus 2^-27 per pair (2^26 units);
      clash: sqrt (1 ulp) of d2 and 0.75 (r_i + r_j)  ->  10u of the clash
        weight times the clash distance, over the pairs at or inside it;
      desolvation: receptor minima kept as f32 keys (u of d2, so 5u sqrt(d2)
        <= 32u of 10 sqrt(d2) in the window) -> 32u sum |des|, plus 2^-38 per
        atom.
    The folds weight E by 332 / 4 and V by 1 (cpydock 0.1)."""
    mag = np.asarray(mag, np.float64).reshape(-1, 5)
    n_near = np.asarray(n_near, np.float64)
    n_far = np.asarray(n_far, np.float64)
    u = U32
    b_e = 11 * u * mag[:, 0] + 2.0 ** -45 * mag[:, 1] + 2.0 ** -38 * (n_near + n_far)
    b_v = 89 * u * mag[:, 2] + 10 * u * mag[:, 3] + 2.0 ** -27 * n_near
    b_s = 32 * u * mag[:, 4] + 2.0 ** -38 * (nr + nl)
    if method == "vdw":
        return b_v
    if method == "cpydock":
        return 83.0 * b_e + 0.1 * b_v + b_s
    return 83.0 * b_e + b_v


def fold(method, e, v, s=0.0):
    return lib().orc_fold_score(FAMILY[method], e, v, s)


def rng_f64s(seed, n):
    out = np.zeros(n)
    lib().orc_rng_f64s(int(seed), int(n), _p(out))
    return out


def quat_random(seed, n):
    out = np.zeros((n, 4))
    lib().orc_quat_random(int(seed), int(n), _p(out))
    return out


def slerp(a, b, t):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    out = np.zeros(4)
    lib().orc_slerp(_p(a), _p(b), float(t), _p(out))
    return out


def nearcells_pair_count(xyz, cell):
    xyz = np.ascontiguousarray(xyz, np.float64)
    return lib().orc_nearcells_pair_count(_p(xyz), xyz.shape[0], float(cell))


def lr_sasa(xyz, radii, probe=1.4, n_slices=20):
    """lr_sasa.rs:277-291: Lee & Richards area per atom (radii without probe)."""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    radii = np.ascontiguousarray(radii, np.float64)
    out = np.zeros(xyz.shape[0])
    lib().orc_lr_sasa(_p(xyz), _p(radii), xyz.shape[0], float(probe), int(n_slices), _p(out))
    return out


def crec_build(xyz):
    L = lib()
    xyz = np.ascontiguousarray(xyz, np.float64)
    nc = (C.c_int * 3)()
    co = (C.c_float * 3)()
    ln = L.orc_crec_build(_p(xyz), xyz.shape[0], nc, co, None, None)
    cs = np.zeros(ln, np.int32)
    ca = np.zeros(max(1, xyz.shape[0]), np.int32)
    L.orc_crec_build(_p(xyz), xyz.shape[0], nc, co, _p(cs, C.c_int), _p(ca, C.c_int))
    return cs, ca[:xyz.shape[0]], tuple(nc), tuple(co)


def feat_counts(rec_xyz, rec_group, lig_xyz, lig_group, poses, nthreads=8):
    """Surrogate pair-histogram counts (lkl_oracle_sur.c), n x 6144 u32."""
    rx = np.ascontiguousarray(rec_xyz, np.float64)
    lx = np.ascontiguousarray(lig_xyz, np.float64)
    rg = np.ascontiguousarray(rec_group, np.uint8)
    lg = np.ascontiguousarray(lig_group, np.uint8)
    P = np.ascontiguousarray(poses, np.float64).reshape(-1, 7)
    out = np.zeros((P.shape[0], 6144), np.uint32)
    lib().orc_feat_counts(_p(rx), _p(rg, C.c_ubyte), rx.shape[0], _p(lx), _p(lg, C.c_ubyte),
                          lx.shape[0], _p(P), P.shape[0], _p(out, C.c_uint32), int(nthreads))
    return out


def feat_grid_dims(rec_xyz, h):
    rx = np.ascontiguousarray(rec_xyz, np.float64)
    o = np.zeros(3)
    d = np.zeros(3, np.int32)
    n = lib().orc_feat_grid_dims(_p(rx), rx.shape[0], float(h), _p(o), _p(d, C.c_int))
    return o, tuple(int(v) for v in d), int(n)


def feat_grid_node(rec_xyz, rec_group, h, ix, iy, iz):
    """Counts (16 x 24) of one node of the grid-mode table."""
    rx = np.ascontiguousarray(rec_xyz, np.float64)
    rg = np.ascontiguousarray(rec_group, np.uint8)
    o, _, _ = feat_grid_dims(rx, h)
    out = np.zeros(384, np.uint32)
    lib().orc_feat_grid_node(_p(rx), _p(rg, C.c_ubyte), rx.shape[0], _p(o), float(h), int(ix), int(iy),
                             int(iz), _p(out, C.c_uint32))
    return out


def feat_grid_counts(rec_xyz, rec_group, lig_xyz, lig_group, poses, h, nthreads=8):
    """Grid-mode features (nearest-node receptor counts), n x 6144 u32."""
    rx = np.ascontiguousarray(rec_xyz, np.float64)
    lx = np.ascontiguousarray(lig_xyz, np.float64)
    rg = np.ascontiguousarray(rec_group, np.uint8)
    lg = np.ascontiguousarray(lig_group, np.uint8)
    P = np.ascontiguousarray(poses, np.float64).reshape(-1, 7)
    out = np.zeros((P.shape[0], 6144), np.uint32)
    lib().orc_feat_grid_counts(_p(rx), _p(rg, C.c_ubyte), rx.shape[0], _p(lx), _p(lg, C.c_ubyte),
                               lx.shape[0], _p(P), P.shape[0], float(h), _p(out, C.c_uint32),
                               int(nthreads))
    return out


# ── screen candidates (own definition; no reference counterpart) ─────────────
# A restatement of lkl_screen_candidates: candidate i = g * C + c of glowworm
# g; c = 0 copies the base pose, c > 0 adds trans_radius * U(-1, 1) to each
# translation component and rot_sigma * U(-1, 1) to each quaternion component,
# then renormalises the quaternion.  U = (splitmix64 output >> 11) * 2^-53 with
# key = mix(seed ^ mix(8 i + d)) for draw d of candidate i, mix being
# splitmix64's output function applied to z + 0x9E3779B97F4A7C15 (so mix(0) is
# the first output of splitmix64 seeded with 0, 0xE220A8397B1DCDAF).

def splitmix64_mix(z):
    """numpy uint64 array -> splitmix64 output of state z (wrapping)."""
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def screen_candidates(base, C, trans_radius, rot_sigma, seed):
    base = np.asarray(base, np.float64).reshape(-1, 7)
    n = base.shape[0]
    i = np.arange(n * C, dtype=np.uint64)
    out = np.repeat(base, C, axis=0).copy()
    with np.errstate(over="ignore"):
        u = np.stack([(splitmix64_mix(np.uint64(seed) ^ splitmix64_mix(i * np.uint64(8) + np.uint64(d)))
                       >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) for d in range(7)], 1)
    q = out[:, :4] + rot_sigma * (2.0 * u[:, :4] - 1.0)
    nrm = np.sqrt(q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2] + q[:, 3] * q[:, 3])
    t = out[:, 4:] + trans_radius * (2.0 * u[:, 4:] - 1.0)
    pert = (i % np.uint64(C)) != 0
    out[pert, :4] = q[pert] / nrm[pert, None]
    out[pert, 4:] = t[pert]
    return out
