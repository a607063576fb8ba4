"""Seeded synthetic complexes of the benchmark configurations.

Receptor and ligand atoms are placed by random sequential addition (RSA, 3.0 A
minimum separation) inside an ellipsoid / sphere.  Atom types come from a
20-entry Amber-like table (LJ radius 1.7-2.1 A, epsilon 0.05-0.2); charges are
N(0, 0.3) shifted additively so the receptor carries +10 e and the ligand -10 e.
All coordinates and parameters are f32-representable, so the f32 arrays given to
the GPU and the f64 arrays of the CPU grid path hold identical values.

Containers that the stated atom counts cannot fill at 3.0 A separation are
scaled isotropically to the 12,625-atom / (60, 55, 50) A receptor's density
(0.01827 atoms/A^3); the actual semi-axes are recorded in `meta`.
"""
from dataclasses import dataclass, field
import math

import numpy as np

from . import _capi

RHO = 12625 / (4.0 / 3.0 * math.pi * 60 * 55 * 50)  # atoms / A^3 of the 12.6k receptor

# 20-entry type table: LJ radius (sigma) and sqrt(epsilon), both f32 values;
# epsilon = sqrt_eps^2 is then exact in f64.
_K = np.arange(20)
TYPE_RADIUS = np.float32(1.7 + 0.4 * _K / 19.0)
TYPE_SQRT_EPS = np.float32(np.sqrt(0.05 + 0.15 * ((_K * 7) % 20) / 19.0))
# ASP desolvation coefficients (cpydock.rs:50-53, entries 1..10), cycled over types.
ASP = np.float32([0.01918, 0.1108, -0.0391, -0.12604, -0.06256,
                  -0.04255, -0.03128, -0.06877, 0.02576, 0.00506])


@dataclass
class Mol:
    xyz: np.ndarray    # (n, 3) f64, f32-exact
    ele: np.ndarray    # f64, f32-exact
    rad: np.ndarray    # vdw radius
    seps: np.ndarray   # sqrt(epsilon)
    heavy: np.ndarray  # u8
    hyd: np.ndarray    # i32, 1 = heavy (cpydock model field)
    des: np.ndarray
    asa: np.ndarray
    typ: np.ndarray = None  # type index (0..19) of the 20-entry table

    @property
    def group(self):
        """Type group (0..15) of the surrogate's pair histograms: 4 size
        classes of the 20-entry type table (types 0-4, 5-9, 10-14, 15-19) x 4
        charge classes (q < -0.2, -0.2 <= q < 0, 0 <= q < 0.2, q >= 0.2).  The
        synthetic charges are drawn per atom, not per type, so the charge class
        is what lets the histograms see electrostatics."""
        t = np.asarray(self.typ if self.typ is not None else np.zeros(self.n, np.int64))
        qc = np.searchsorted(np.array([-0.2, 0.0, 0.2]), np.asarray(self.ele), side="right")
        return ((t * 4) // 20 * 4 + qc).astype(np.uint8)

    @property
    def n(self):
        return self.xyz.shape[0]

    @property
    def eps(self):
        return self.seps * self.seps

    @property
    def flag(self):
        """cpydock.rs:808-819 exclusion mask: (i even) and hydrogens[i // 2] != 0."""
        i = np.arange(self.n)
        f = np.zeros(self.n, np.uint8)
        even = i % 2 == 0
        f[even] = (self.hyd[i[even] // 2] != 0).astype(np.uint8)
        return f

    def f32(self, name):
        return np.ascontiguousarray(getattr(self, name), dtype=np.float32)


@dataclass
class Complex:
    rec: Mol
    lig: Mol
    meta: dict = field(default_factory=dict)
    phi: np.ndarray = None      # f32 far field, flat, x fastest
    origin: tuple = None
    dims: tuple = None

    def build_field(self):
        """Build phi on the GPU (lkl_field_build)."""
        L = _capi.lib()
        rx = self.rec.f32("xyz")
        re = self.rec.f32("ele")
        o = np.zeros(3, np.float64)
        d = np.zeros(3, np.int32)
        n = L.lkl_field_dims(_capi.ptr(rx, _capi.C.c_float), self.rec.n,
                             _capi.ptr(o, _capi.C.c_double), _capi.ptr(d, _capi.C.c_int))
        phi = np.empty(int(n), np.float32)
        _capi.check(L.lkl_field_build(_capi.ptr(rx, _capi.C.c_float),
                                      _capi.ptr(re, _capi.C.c_float), self.rec.n,
                                      _capi.ptr(phi, _capi.C.c_float)), "lkl_field_build")
        self.phi, self.origin, self.dims = phi, tuple(o), tuple(int(v) for v in d)
        return phi

    def set_field(self, phi, origin, dims):
        self.phi = np.ascontiguousarray(phi, np.float32)
        self.origin = tuple(float(v) for v in origin)
        self.dims = tuple(int(v) for v in dims)


def rsa(n, axes, seed, dmin=3.0, max_attempts=None):
    """Random sequential addition of n points in an ellipsoid (semi-axes)."""
    L = _capi.lib()
    xyz = np.zeros((n, 3), np.float64)
    att = _capi.C.c_int64(0)
    cap = max_attempts or 400 * n + 100000
    got = L.lkl_synth_rsa(n, float(axes[0]), float(axes[1]), float(axes[2]), float(dmin),
                          int(seed) & 0xFFFFFFFFFFFFFFFF, int(cap),
                          _capi.ptr(xyz, _capi.C.c_double), _capi.C.byref(att))
    if got != n:
        raise ValueError(f"RSA placed {got}/{n} atoms in {axes} after {att.value} attempts")
    return xyz, att.value


def scaled_axes(n, axes):
    """Axes scaled isotropically (never shrunk) so n atoms sit at density RHO."""
    vol = 4.0 / 3.0 * math.pi * axes[0] * axes[1] * axes[2]
    s = max(1.0, (n / (RHO * vol)) ** (1.0 / 3.0))
    return tuple(float(a * s) for a in axes)


def _f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def make_mol(n, axes, seed, net_charge, radii=None, center=False):
    rng = np.random.default_rng(seed + 7919)
    xyz, attempts = rsa(n, axes, seed)
    if center:
        xyz = np.round((xyz - xyz.mean(axis=0)) * 1024.0) / 1024.0
    t = rng.integers(0, 20, n)
    q = rng.normal(0.0, 0.3, n)
    q = q + (net_charge - q.sum()) / n
    if radii is None:
        rad = TYPE_RADIUS[t].astype(np.float64)
    else:
        rad = _f32(rng.uniform(radii[0], radii[1], n))
    # exposure proxy for the reference SASA (non-negative), from the neighbour
    # count; make_config replaces it with Lee & Richards areas for "c3"
    from scipy.spatial import cKDTree
    nb = np.array([len(v) - 1 for v in cKDTree(xyz).query_ball_point(xyz, 8.0)])
    area = 4.0 * math.pi * (rad + 1.4) ** 2 * np.clip(1.0 - nb / 30.0, 0.0, 1.0)
    mol = Mol(xyz=xyz, ele=_f32(q), rad=rad, seps=TYPE_SQRT_EPS[t].astype(np.float64),
              heavy=np.ones(n, np.uint8), hyd=np.ones(n, np.int32),
              des=ASP[t % 10].astype(np.float64), asa=_f32(area), typ=t.astype(np.int32))
    return mol, attempts


def random_quaternions(n, rng):
    """Uniform unit quaternions (w, x, y, z), the construction of qt.rs:106-116."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
    return q


def make_poses(n, rec_axes, seed, shell=15.0):
    """n poses (w, x, y, z, tx, ty, tz), f32 values: uniform rotation; translation =
    receptor-surface point along a uniform direction + U(0, shell) outward."""
    rng = np.random.default_rng(seed)
    q = random_quaternions(n, rng)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    a, b, c = rec_axes
    r_surf = 1.0 / np.sqrt((u[:, 0] / a) ** 2 + (u[:, 1] / b) ** 2 + (u[:, 2] / c) ** 2)
    t = u * (r_surf + rng.uniform(0.0, shell, n))[:, None]
    return _f32(np.concatenate([q, t], 1))


def make_swarm_positions(n_swarms, n_glow, rec_axes, seed, lig_radius=0.0, spread=3.0):
    """Initial GSO positions, rows (tx, ty, tz, w, x, y, z), f32 values: each
    swarm centre sits on the receptor surface (uniform direction) pushed out by
    the ligand radius; glowworms are uniform in a ball of `spread` A around it
    with uniform random rotations."""
    rng = np.random.default_rng(seed)
    rows = []
    a, b, c = rec_axes
    for _ in range(n_swarms):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        r_surf = 1.0 / math.sqrt((u[0] / a) ** 2 + (u[1] / b) ** 2 + (u[2] / c) ** 2)
        centre = u * (r_surf + lig_radius)
        d = rng.normal(size=(n_glow, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        t = centre + d * (spread * rng.random(n_glow) ** (1.0 / 3.0))[:, None]
        q = random_quaternions(n_glow, rng)
        rows.append(np.concatenate([t, q], 1))
    return _f32(np.concatenate(rows, 0))


CONFIGS = {
    # name: (n_rec, rec axes as stated, n_lig, lig radius as stated, seed)
    "c1": (1500, (22.0, 18.0, 16.0), 800, 10.0, 42),
    "c2": (12625, (60.0, 55.0, 50.0), 1500, None, 42),
    "c3": (8000, None, 3000, None, 42),
    "tiny": (300, (12.0, 11.0, 10.0), 60, None, 5),
    "small": (1200, None, 200, None, 11),
}


def make_config(name, seed=None, radii=None):
    nr, raxes, nl, lrad, s0 = CONFIGS[name]
    seed = s0 if seed is None else seed
    if raxes is None:
        r = (3 * nr / (4 * math.pi * RHO)) ** (1 / 3)
        raxes = (r, r, r)
    raxes = scaled_axes(nr, raxes)
    lr = lrad if lrad is not None else (3 * nl / (4 * math.pi * RHO)) ** (1 / 3)
    laxes = scaled_axes(nl, (lr, lr, lr))
    if name == "c3" and radii is None:
        radii = (1.5, 2.0)
    rec, ra = make_mol(nr, raxes, seed, +10.0, radii=radii)
    lig, la = make_mol(nl, laxes, seed + 1, -10.0, radii=radii, center=True)
    meta = {"config": name, "seed": seed, "rec_axes": raxes, "lig_axes": laxes,
            "rec_rsa_attempts": ra, "lig_rsa_attempts": la, "density": RHO,
            "asa": "neighbour-count proxy"}
    if name == "c3":
        # the desolvation workload's reference areas: Lee & Richards per
        # monomer (probe 1.4 A, 20 slices, the atoms' radii), on the GPU
        from . import lr_sasa
        for m in (rec, lig):
            m.asa = _f32(lr_sasa(m.xyz, m.rad))
        meta["asa"] = "Lee & Richards, probe 1.4 A, 20 slices (lkl_lr_sasa)"
    return Complex(rec=rec, lig=lig, meta=meta)


def make_config4(n_complex=64, seed0=0, scale=1.0):
    """64 independent complexes: receptor atoms log-uniform 2,000-20,000, ligand
    atoms log-uniform 500-5,000 (spheres at the benchmark density), glowworm
    counts uniform 100-500, families drawn from dna/pydock/vdw; complex k uses
    seed seed0 + k.  `scale` shrinks every size (tests)."""
    out = []
    for k in range(n_complex):
        rng = np.random.default_rng(seed0 + k)
        nr = max(50, int(round(math.exp(rng.uniform(math.log(2000), math.log(20000))) * scale)))
        nl = max(20, int(round(math.exp(rng.uniform(math.log(500), math.log(5000))) * scale)))
        ng = max(4, int(round(rng.integers(100, 501) * scale)))
        fam = ["dna", "pydock", "vdw"][int(rng.integers(0, 3))]
        rr = (3 * nr / (4 * math.pi * RHO)) ** (1 / 3)
        lr = (3 * nl / (4 * math.pi * RHO)) ** (1 / 3)
        rec, _ = make_mol(nr, (rr, rr, rr), seed0 + k, +10.0)
        lig, _ = make_mol(nl, (lr, lr, lr), seed0 + k + 1000, -10.0, center=True)
        cx = Complex(rec=rec, lig=lig, meta={"config": "c4", "index": k, "family": fam,
                                             "glowworms": ng, "rec_axes": (rr, rr, rr),
                                             "lig_axes": (lr, lr, lr), "seed": seed0 + k})
        out.append(cx)
    return out
