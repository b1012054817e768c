"""CPU oracle for the XPBD substep — TEST INFRASTRUCTURE ONLY.

Python driver around ``oracle/_build/liboracle.so`` (the C restatement of the
reference kernels, oracle/xpbd_oracle.c).  ``OracleSim`` exposes the same
methods as ``rlmusclev2_amd.engine.XpbdEngine`` so the same sim program runs on
either; only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline
leg use it, as the checker and the reported CPU baseline.

Program order follows MultiGridSim.build_program (src/muscle/mg/mg.py:1322-1378):
integrate -> pin -> iters x (clear, dispatch, apply_dP, pin) -> bone push ->
update_velocities.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

ORC_TETARAPSVD = -92199132
ORC_TETVOLUME = -215389979
ORC_TETSNH = -503462311
ORC_TETARAPANISO = -92199133
ORC_ATTACH = 1650556
ORC_PINTOTARGET = 771624311

_P = C.c_void_p


class OrcCons(C.Structure):
    _fields_ = [("n", C.c_int), ("type", _P), ("pts", _P), ("stiffness", _P),
                ("compressionstiffness", _P), ("dampingratio", _P), ("tetid", _P), ("L", _P),
                ("restlength", _P), ("restvector", _P), ("restdir", _P), ("reaction_accum", _P)]


class OrcState(C.Structure):
    _fields_ = [("nv", C.c_int), ("nt", C.c_int), ("pos", _P), ("pprev", _P), ("vel", _P),
                ("force", _P), ("mass", _P), ("stopped", _P), ("dP", _P), ("dPw", _P),
                ("rest_matrix", _P), ("v_fiber_dir", _P), ("activation", _P), ("tendonmask", _P)]


class OrcParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("gravity", C.c_float * 3), ("veldamping", C.c_float),
                ("quasi_static", C.c_int), ("contraction_ratio", C.c_float),
                ("fiber_stiffness_scale", C.c_float), ("has_compress", C.c_int),
                ("use_jacobi", C.c_int), ("exp_damp", C.c_float), ("const_damp", C.c_float)]


_lib = None


def build():
    """Compile the C oracle (gcc, no reference build involved)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        lib.orc_svd3.argtypes = [_P, _P, _P, _P]
        lib.orc_svd3_double.argtypes = [_P, _P, _P, _P]
        lib.orc_precompute_rest.argtypes = [C.c_int, _P, _P, C.c_float, _P, _P, _P, _P]
        lib.orc_integrate.argtypes = [C.POINTER(OrcState), C.POINTER(OrcParams)]
        lib.orc_clear.argtypes = [C.POINTER(OrcState), C.POINTER(OrcCons)]
        lib.orc_dispatch.argtypes = [C.POINTER(OrcCons), C.POINTER(OrcState), C.POINTER(OrcParams), C.c_int]
        lib.orc_dispatch.restype = C.c_int
        lib.orc_apply_dP.argtypes = [C.POINTER(OrcState), C.c_int]
        lib.orc_pintotarget_masked.argtypes = [C.c_int, _P, _P, _P, _P, _P]
        lib.orc_pintotarget_cons.argtypes = [C.POINTER(OrcCons), _P]
        lib.orc_update_velocities.argtypes = [C.POINTER(OrcState), C.POINTER(OrcParams)]
        lib.orc_sdf_grid_push.argtypes = [C.c_int, _P, _P, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P,
                                          C.c_float, C.c_float, _P]
        lib.orc_sdf_grid_push.restype = C.c_int
        lib.orc_mesh_normals.argtypes = [C.c_int, _P, C.c_int, _P, _P, _P]
        lib.orc_mesh_query.argtypes = [C.c_int, _P, C.c_int, _P, C.c_int, _P, C.c_int, _P, _P, _P, _P, _P, _P,
                                       C.c_float, _P, _P, _P, _P, C.c_int]
        lib.orc_mesh_push.argtypes = [C.c_int, _P, _P, _P, _P, _P, C.c_float, C.c_float]
        lib.orc_mesh_push.restype = C.c_int
        lib.orc_mesh_bind.argtypes = [C.c_int, _P, _P, C.c_int, _P, C.c_float, _P, _P, C.c_int]
        lib.orc_residuals.argtypes = [C.c_int, _P, _P, _P, _P, _P, _P, _P]
        lib.orc_count_inverted.argtypes = [C.c_int, _P, _P, _P]
        lib.orc_count_inverted.restype = C.c_int
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def svd3(F):
    """Sign-fixed SVD of a 3x3 float32 matrix (U, sigma, V) — ssvd_fn semantics."""
    lib = load()
    F = np.ascontiguousarray(F, np.float32).reshape(3, 3)
    U = np.empty((3, 3), np.float32)
    V = np.empty((3, 3), np.float32)
    s = np.empty(3, np.float32)
    lib.orc_svd3(_p(F), _p(U), _p(s), _p(V))
    return U, s, V


def precompute_rest(pos0, tets, density):
    lib = load()
    pos0 = np.ascontiguousarray(pos0, np.float32)
    tets = np.ascontiguousarray(tets, np.int32)
    nt, nv = tets.shape[0], pos0.shape[0]
    rv = np.zeros(nt, np.float32)
    rm = np.zeros((nt, 3, 3), np.float32)
    m = np.zeros(nv, np.float32)
    tot = np.zeros(1, np.float32)
    lib.orc_precompute_rest(nt, _p(pos0), _p(tets), C.c_float(density), _p(rv), _p(rm), _p(m), _p(tot))
    return rv, rm, m


def mesh_adjacency(tris) -> np.ndarray:
    """(ntris, 3) neighbour across edge ab, bc, ca: the first other triangle (in
    triangle order) on the same undirected edge, -1 on a boundary edge."""
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    nt = len(t)
    u = np.concatenate([t[:, 0], t[:, 1], t[:, 2]])
    v = np.concatenate([t[:, 1], t[:, 2], t[:, 0]])
    f = np.tile(np.arange(nt), 3)
    e = np.repeat(np.arange(3), nt)
    key = np.minimum(u, v) * (int(t.max()) + 1) + np.maximum(u, v)
    order = np.lexsort((e, f, key))
    k, fs = key[order], f[order]
    n = len(k)
    start = np.r_[0, np.nonzero(np.diff(k))[0] + 1]
    grp = np.repeat(np.arange(len(start)), np.diff(np.r_[start, n]))
    first = fs[start][grp]
    cand = np.where(fs != first, np.arange(n), n)
    sec_pos = np.minimum.reduceat(cand, start)
    sec = np.where(sec_pos < n, fs[np.minimum(sec_pos, n - 1)], -1)[grp]
    nb = np.where(fs != first, first, sec)
    adj = np.full((nt, 3), -1, np.int32)
    adj[f[order], e[order]] = nb
    return adj


class MeshQuery:
    """Brute-force closest point + pseudonormal sign over a bone triangle mesh
    (orc_mesh_query): the checker of the library's BVH query."""

    def __init__(self, verts, tris, tri_off=None):
        self.lib = load()
        self.tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
        nt = self.tris.shape[0]
        self.tri_off = np.ascontiguousarray([0, nt] if tri_off is None else tri_off, np.int32)
        self.adj = np.ascontiguousarray(mesh_adjacency(self.tris), np.int32)
        self.set_points(verts)

    def set_points(self, verts):
        self.verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        nt = self.tris.shape[0]
        self.fn = np.zeros((nt, 3), np.float64)
        self.vn = np.zeros((self.verts.shape[0], 3), np.float64)
        self.lib.orc_mesh_normals(self.verts.shape[0], _p(self.verts), nt, _p(self.tris), _p(self.fn), _p(self.vn))

    def query(self, q, max_dist, vgroup=None, goff=None, nthreads=8):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        n = q.shape[0]
        sd = np.zeros(n, np.float64)
        cp = np.zeros((n, 3), np.float64)
        face = np.zeros(n, np.int32)
        region = np.zeros(n, np.int32)
        vg = None if vgroup is None else np.ascontiguousarray(vgroup, np.int32)
        go = None if goff is None else np.ascontiguousarray(goff, np.int32)
        self.lib.orc_mesh_query(n, _p(q), self.verts.shape[0], _p(self.verts), self.tris.shape[0], _p(self.tris),
                                len(self.tri_off) - 1, _p(self.tri_off), _p(self.fn), _p(self.vn), _p(self.adj),
                                _p(vg), _p(go), C.c_float(max_dist), _p(sd), _p(cp), _p(face), _p(region),
                                int(nthreads))
        return sd, cp, face, region

    def push(self, pos, max_dist, contact_offset, max_push, coll_w=None, vgroup=None, goff=None, nthreads=8):
        """solve_mesh_contact_pos_kernel on pos (float32, in place); returns the number pushed"""
        sd, cp, face, _ = self.query(pos, max_dist, vgroup, goff, nthreads)
        cw = None if coll_w is None else np.ascontiguousarray(coll_w, np.float32)
        return self.lib.orc_mesh_push(pos.shape[0], _p(pos), _p(cw), _p(sd), _p(cp), _p(face),
                                      C.c_float(contact_offset), C.c_float(max_push))

    def depth(self, pos, max_dist, meter_w=None, vgroup=None, goff=None, nthreads=8):
        """_mark_inside_kernel: > 0 inside, 0 outside or unmetered"""
        sd, _, face, _ = self.query(pos, max_dist, vgroup, goff, nthreads)
        d = np.where((face >= 0) & (sd < 0.0), -sd, 0.0)
        if meter_w is not None:
            d[np.asarray(meter_w) == 0] = 0.0
        return d


def fascia_bind(surf_pos, surf_tris, points, max_dist, nthreads=8):
    """_build_muscle_bind (fascia.py:443-490) by brute force: per point the nearest
    surface triangle among those nearer than max_dist (lowest index at near-ties),
    barycentric weights of the closest point and the normal gap (x - closest)·n.
    Returns face (-1 free), weights (n, 3) and gap (n,), in float64."""
    lib = load()
    sp = np.ascontiguousarray(surf_pos, np.float32).reshape(-1, 3)
    tr = np.ascontiguousarray(surf_tris, np.int32).reshape(-1, 3)
    q = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    n = len(q)
    face = np.zeros(n, np.int32)
    cp = np.zeros((n, 3), np.float64)
    lib.orc_mesh_bind(n, _p(q), _p(sp), len(tr), _p(tr), C.c_float(max_dist), _p(face), _p(cp), int(nthreads))
    w = np.zeros((n, 3))
    gap = np.zeros(n)
    ok = face >= 0
    A, B, Cc = (sp[tr[face[ok], k]].astype(np.float64) for k in range(3))
    v0, v1, v2 = B - A, Cc - A, cp[ok] - A
    d00, d01, d11 = (v0 * v0).sum(1), (v0 * v1).sum(1), (v1 * v1).sum(1)
    d20, d21 = (v2 * v0).sum(1), (v2 * v1).sum(1)
    den = d00 * d11 - d01 * d01
    safe = den > 0
    wb = np.where(safe, (d11 * d20 - d01 * d21) / np.where(safe, den, 1.0), 0.0).clip(0.0)
    wc = np.where(safe, (d00 * d21 - d01 * d20) / np.where(safe, den, 1.0), 0.0).clip(0.0)
    over = wb + wc > 1.0
    ssum = np.where(over, wb + wc, 1.0)
    wb, wc = wb / ssum, wc / ssum
    w[ok] = np.c_[1.0 - wb - wc, wb, wc]
    nrm = np.cross(v0, v1)
    nl = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = np.where(nl > 0, nrm / np.where(nl > 0, nl, 1.0), 0.0)
    gap[ok] = ((q[ok].astype(np.float64) - cp[ok]) * nrm).sum(1)
    return face, w, gap


def fascia_ride_targets(x, surf_tris, face, w, gap):
    """fascia_ride_kernel (fascia.py:159-176) in FP32: barycentric point of the current
    triangle plus gap along its unit normal (rows with face >= 0)"""
    tr = np.asarray(surf_tris, np.int64).reshape(-1, 3)[face]
    x = np.asarray(x, np.float32)
    a, b, c = x[tr[:, 0]], x[tr[:, 1]], x[tr[:, 2]]
    w = np.asarray(w, np.float32)
    cp = (w[:, 0:1] * a + w[:, 1:2] * b) + w[:, 2:3] * c
    n = np.cross(b - a, c - a).astype(np.float32)
    ln = np.sqrt((n * n).sum(1, keepdims=True))
    u = np.where(ln > 0, n / np.where(ln > 0, ln, 1), 0).astype(np.float32)
    return (cp + np.asarray(gap, np.float32)[:, None] * u).astype(np.float32)


def glue_step(xf, xm, wj, restlen, alpha, omega):
    """solve_glue_kernel (fascia.py:81-125) in FP32 with the multiplier at zero: the
    kinematic fascia point xf pulls xm along the pair by omega * C * wj / (wj + alpha),
    C = |xf - xm| - restlen; pairs with |xf - xm| < 1e-9 or wj + alpha < 1e-12 stay.
    Returns the new xm."""
    xf = np.asarray(xf, np.float32)
    xm = np.asarray(xm, np.float32)
    n = xf - xm
    d = np.sqrt((n * n).sum(1)).astype(np.float32)
    denom = (wj + alpha).astype(np.float32)
    ok = (d >= np.float32(1e-9)) & (denom >= np.float32(1e-12))
    dirn = n / np.where(ok, d, np.float32(1.0))[:, None]
    C = d - np.asarray(restlen, np.float32)
    dl = -C / np.where(ok, denom, np.float32(1.0))
    m = ((-(np.float32(omega) * dl) * wj)[:, None] * dirn).astype(np.float32)
    return np.where(ok[:, None], xm + m, xm).astype(np.float32)


class OracleSim:
    """Same interface as XpbdEngine, computed by the C oracle on the CPU."""

    def __init__(self, scene, nthreads: int = 1):
        self.lib = load()
        self.scene = scene
        self.nthreads = int(nthreads)
        s = scene
        self.NV, self.NT = s.nv, s.nt
        self.pos0 = np.ascontiguousarray(s.pos0, np.float32).reshape(-1, 3)
        self.tets = np.ascontiguousarray(s.tets, np.int32).reshape(-1, 4)
        self.rest_volume, self.rest_matrix, mass = precompute_rest(self.pos0, self.tets, s.density)
        rs = float(np.float32(getattr(s, "rest_scale", 1.0)))
        if abs(rs - 1.0) >= 1e-9:         # Fascia._apply_rest_shrink (fascia.py:341-352)
            self.rest_matrix = (self.rest_matrix.astype(np.float64) * (1.0 / rs)).astype(np.float32)
        self.mass = np.ascontiguousarray(s.mass if s.mass is not None else mass, np.float32)
        self.pos = self.pos0.copy()
        self.pprev = self.pos0.copy()
        self.vel = np.zeros_like(self.pos)
        self.force = np.zeros_like(self.pos)
        self.stopped = np.ascontiguousarray(s.stopped if s.stopped is not None else np.zeros(self.NV), np.int32)
        self.dP = np.zeros_like(self.pos)
        self.dPw = np.zeros(self.NV, np.float32)
        self.vf = np.ascontiguousarray(s.v_fiber_dir if s.v_fiber_dir is not None
                                       else np.zeros((self.NV, 3)), np.float32)
        self.tendonmask = np.ascontiguousarray(s.tendonmask if s.tendonmask is not None
                                               else np.zeros(self.NT), np.float32)
        self.activation = np.zeros(self.NT, np.float32)
        t = s.cons.contiguous()
        self.cons = t
        self.L = np.zeros((t.n, 3), np.float32)
        self.reaction = np.zeros((t.n, 3), np.float32)
        self.has_pt_cons = bool(np.any(t.type == ORC_PINTOTARGET))
        self._c = OrcCons(t.n, _p(t.type), _p(t.pts), _p(t.stiffness), _p(t.compressionstiffness),
                          _p(t.dampingratio), _p(t.tetid), _p(self.L), _p(t.restlength),
                          _p(t.restvector), _p(t.restdir), _p(self.reaction))
        self._s = OrcState(self.NV, self.NT, _p(self.pos), _p(self.pprev), _p(self.vel), _p(self.force),
                           _p(self.mass), _p(self.stopped), _p(self.dP), _p(self.dPw), _p(self.rest_matrix),
                           _p(self.vf), _p(self.activation), _p(self.tendonmask))
        self.set_params(s.params)
        self.pin_vids = None if s.pin_vids is None else np.ascontiguousarray(s.pin_vids, np.int32)
        self.pin_active = None
        if self.pin_vids is not None:
            self.pin_active = np.ascontiguousarray(s.pin_active if s.pin_active is not None
                                                   else np.ones(len(self.pin_vids)), np.int32)
            self.pin_target = self.pos0[self.pin_vids].copy()
            self.pin_zero = np.zeros_like(self.pin_target)
        self.sdf = s.sdf
        if self.sdf is not None:
            sd = self.sdf
            self.grids = np.ascontiguousarray(sd.grids, np.float32)
            self.origins = np.ascontiguousarray(sd.origins, np.float32)
            self.spacing = np.ascontiguousarray(sd.spacing, np.float32)
            self.coll_w = np.ascontiguousarray(sd.coll_w if sd.coll_w is not None else np.ones(self.NV), np.float32)
            self.xf = np.tile(np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32), (sd.nbones, 1))
            self.vgroup = None if sd.vertex_group is None else np.ascontiguousarray(sd.vertex_group, np.int32)
            self.goff = None if sd.group_bone_off is None else np.ascontiguousarray(sd.group_bone_off, np.int32)
            self.cells = np.full((self.NV, 2), -1, np.int32)

        self.bm = None
        if s.bone_mesh is not None:
            self.set_bone_mesh(s.bone_mesh)

    # exact-mesh bone collider -------------------------------------------------
    def set_bone_mesh(self, bm):
        self.bm = bm
        if bm is None:
            self.bmq = None
            return
        self.bmq = MeshQuery(bm.verts, bm.tris, bm.tri_off)
        self.bm_resolve = bool(bm.resolve)

    def set_bone_mesh_points(self, verts):
        self.bmq.set_points(verts)

    def set_bone_mesh_resolve(self, on):
        self.bm_resolve = bool(on)

    def bone_mesh_push(self):
        bm = self.bm
        return self.bmq.push(self.pos, bm.max_dist, bm.contact_offset, bm.max_push, bm.coll_w, bm.vertex_group,
                             bm.group_bone_off, self.nthreads)

    def bone_penetration(self, collide_eps=1e-6, want_depth=True):
        bm = self.bm
        d = self.bmq.depth(self.pos, bm.max_dist, bm.meter_w, bm.vertex_group, bm.group_bone_off, self.nthreads)
        return (d.astype(np.float32) if want_depth else None), int((d > collide_eps).sum()), float(d.max(initial=0.0))

    def set_params(self, p):
        q = OrcParams()
        q.dt = p.dt
        for i in range(3):
            q.gravity[i] = p.gravity[i]
        q.veldamping = p.veldamping
        q.quasi_static = int(bool(p.quasi_static))
        q.contraction_ratio = p.contraction_ratio
        q.fiber_stiffness_scale = p.fiber_stiffness_scale
        q.has_compress = int(bool(p.has_compress))
        q.use_jacobi = 1
        q.exp_damp = p.exp_damp
        q.const_damp = p.const_damp
        self._prm = q
        self.params = p

    # inputs ---------------------------------------------------------------
    def set_pin_targets(self, xyz):
        self.pin_target = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)

    def set_bone_transforms(self, xf):
        self.xf = np.ascontiguousarray(xf, np.float32).reshape(-1, 12)

    def set_fiber_activation(self, act):
        self.activation[:] = np.asarray(act, np.float32)

    def set_activation_groups(self, tet_group):
        self._act_group = np.asarray(tet_group, np.int64)

    def set_group_activation(self, act):
        self.activation[:] = np.asarray(act, np.float32)[self._act_group]

    def set_attach_targets(self, xyz):
        att = np.nonzero(self.cons.type == ORC_ATTACH)[0]
        self.cons.restvector[att, :3] = np.asarray(xyz, np.float32).reshape(-1, 3)

    # program pieces ----------------------------------------------------------
    def apply_pintotarget(self):
        if self.pin_vids is not None:
            self.lib.orc_pintotarget_masked(len(self.pin_vids), _p(self.pin_vids), _p(self.pin_target),
                                            _p(self.pin_zero), _p(self.pin_active), _p(self.pos))
        if self.has_pt_cons:
            self.lib.orc_pintotarget_cons(C.byref(self._c), _p(self.pos))

    def integrate(self):
        self.lib.orc_integrate(C.byref(self._s), C.byref(self._prm))
        self.apply_pintotarget()

    def clear(self):
        self.lib.orc_clear(C.byref(self._s), C.byref(self._c))

    def solve_constraints(self, run_extras=True):
        rc = self.lib.orc_dispatch(C.byref(self._c), C.byref(self._s), C.byref(self._prm), self.nthreads)
        if rc != 0:
            raise RuntimeError("oracle: unsupported constraint type in table")

    def apply_dP(self):
        self.lib.orc_apply_dP(C.byref(self._s), self.nthreads)

    def iterate(self, n):
        for _ in range(int(n)):
            self.clear()
            self.solve_constraints()
            self.apply_dP()
            self.apply_pintotarget()

    def collide(self):
        if self.bm is not None and self.bm_resolve:
            return self.bone_mesh_push()
        if self.sdf is None:
            return 0
        sd = self.sdf
        return self.lib.orc_sdf_grid_push(self.NV, _p(self.pos), _p(self.coll_w), sd.nbones, sd.res,
                                          _p(self.grids), _p(self.origins), _p(self.spacing), _p(self.xf),
                                          _p(self.vgroup), _p(self.goff), C.c_float(sd.contact_offset),
                                          C.c_float(sd.max_push), _p(self.cells))

    def update_velocities(self):
        self.lib.orc_update_velocities(C.byref(self._s), C.byref(self._prm))

    # two-level correction, restated in NumPy from mg.py:1064-1079, mg/kernel.py:9-40, 138-180
    # (least-squares restriction per muscle block, as built by rlmusclev2_amd/multigrid.py)
    def set_coarse_level(self, coarse, level, kinematic=True):
        self.coarse = coarse
        self.mg = level
        self._mg_shift_pprev = bool(kinematic)

    def vcycle(self):
        cc, lv = getattr(self, "coarse", None), getattr(self, "mg", None)
        if cc is None or lv is None or lv.iters <= 0:
            return
        if getattr(self, "_mg_restrict", None) is None:
            self._mg_restrict = CageRestriction(lv)
        cc.pos[:] = self._mg_restrict(self.pos, self.pos0).astype(np.float32)
        qprev = cc.pos.astype(np.float64).copy()
        if cc.pin_vids is not None:
            if getattr(lv, "pin_rule", 0) == 1:
                # the reference's rule (_coarse_kinematic_pin then stopped nodes held,
                # mg.py:999-1062): pins start at the restricted bone goal
                goal = getattr(self, "_mg_pin_goal", None)
                goal = lv.pos0 if goal is None else goal
                cc.set_pin_targets(goal[cc.pin_vids])
                cc.pos[cc.pin_vids] = goal[cc.pin_vids]
            else:                          # coarse pins keep their restricted positions
                cc.set_pin_targets(cc.pos[cc.pin_vids])
        cc.iterate(lv.iters)
        d = cc.pos.astype(np.float64) - qprev
        pidx = np.asarray(lv.p_idx, np.int64)
        dx = np.einsum("ik,ikc->ic", lv.p_w.astype(np.float64), d[pidx])
        if self.pin_vids is not None:      # pinned vertices are kinematic: no correction
            dx[self.pin_vids[self.pin_active != 0]] = 0.0
        self.pos[:] = (self.pos.astype(np.float64) + dx).astype(np.float32)
        if getattr(self, "_mg_shift_pprev", False):
            self.pprev[:] = (self.pprev.astype(np.float64) + dx).astype(np.float32)
        self.apply_pintotarget()

    def set_coarse_pin_targets(self, xyz):
        self._mg_pin_goal = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)

    # fascia shell (Fascia, fascia.py:245-572): another OracleSim, in ride or squeeze mode
    def set_fascia(self, fascia, link):
        if fascia is None:
            self._fascia = None
            return
        iters = int(link["iters"])
        if int(link.get("mode", 0)) == 1:
            self._fascia = dict(mode=1, f=fascia, iters=iters,
                                prev=np.asarray(link["follow0"], np.float32).reshape(-1, 3).copy(),
                                gf=np.asarray(link["glue_fascia"], np.int64), gm=np.asarray(link["glue_muscle"], np.int64),
                                rl=np.asarray(link["glue_restlen"], np.float32),
                                cm=np.asarray(link["glue_compliance"], np.float32), omega=np.float32(link["omega"]),
                                anc=np.asarray(link["anchor"], np.int64),
                                off=np.asarray(link["anchor_offset"], np.float32).reshape(-1, 3))
            return
        surf_tris = link["surf_tris"]
        face, w, gap = fascia_bind(self.pos0, surf_tris, fascia.pos0, link["bind_max_dist"], max(self.nthreads, 8))
        bound = np.nonzero(face >= 0)[0].astype(np.int32)
        tris = np.asarray(surf_tris, np.int32).reshape(-1, 3)
        fascia.pin_vids = bound
        fascia.pin_active = np.ones(len(bound), np.int32)
        fascia.pin_target = fascia_ride_targets(self.pos0, tris, face[bound], w[bound], gap[bound])
        fascia.pin_zero = np.zeros_like(fascia.pin_target)
        self._fascia = dict(mode=0, f=fascia, tris=tris, bound=bound, face=face[bound], w=w[bound], gap=gap[bound],
                            iters=iters)
        self.fascia_bind = (face, w, gap)

    def set_fascia_follow(self, xyz):
        """squeeze: advect_bone (fascia.py:540-545)"""
        fz = self._fascia
        cur = np.asarray(xyz, np.float32).reshape(-1, 3)
        fz["f"].pos[:] = fz["f"].pos + (cur - fz["prev"])
        fz["prev"] = cur.copy()

    def _inv_mass(self):
        w = np.where(self.mass > 0, (np.float32(1.0) / np.where(self.mass > 0, self.mass, 1)).astype(np.float32),
                     np.float32(0.0))
        return np.where(self.stopped != 0, np.float32(0.0), w).astype(np.float32)

    def fascia_substep(self):
        if getattr(self, "_fascia", None) is not None:
            self._fascia_substep()

    def _fascia_substep(self):
        fz = self._fascia
        f = fz["f"]
        if fz["mode"] == 0:
            if len(fz["bound"]):
                g = fascia_ride_targets(self.pos, fz["tris"], fz["face"], fz["w"], fz["gap"])
                f.pos[fz["bound"]] = g          # the ride snaps before the fascia's integrate
                f.pin_target = g
            f.native_substep(fz["iters"])
            return
        # squeeze (Fascia.build_program, fascia.py:547-565)
        f.pos[fz["anc"]] = fz["prev"][fz["anc"]] + fz["off"]
        f.integrate()
        wj = self._inv_mass()[fz["gm"]]
        inv_dt2 = np.float32(1.0) / np.float32(np.float32(f.params.dt) * np.float32(f.params.dt))
        alpha = fz["cm"] * inv_dt2
        for _ in range(fz["iters"]):
            f.iterate(1)
            self.pos[fz["gm"]] = glue_step(f.pos[fz["gf"]], self.pos[fz["gm"]], wj, fz["rl"], alpha, fz["omega"])

    def native_substep(self, iters):
        self.integrate()
        self.vcycle()
        self.iterate(iters)
        if getattr(self, "_fascia", None) is not None:
            self._fascia_substep()
        self.collide()
        self.update_velocities()

    def native_frame(self, nsub, iters, use_graph=True):
        for _ in range(int(nsub)):
            self.native_substep(iters)

    def synchronize(self):
        pass

    # state -----------------------------------------------------------------
    def positions_np(self):
        return self.pos.copy()

    def velocities_np(self):
        return self.vel.copy()

    def get_state(self):
        return self.pos.copy(), self.pprev.copy(), self.vel.copy()

    def set_state(self, pos=None, pprev=None, vel=None):
        if pos is not None:
            self.pos[:] = pos
        if pprev is not None:
            self.pprev[:] = pprev
        if vel is not None:
            self.vel[:] = vel

    def get_rest(self):
        return self.rest_volume.copy(), self.rest_matrix.copy(), self.mass.copy()

    def sdf_cells(self):
        return self.cells.copy()

    def count_inverted(self):
        return int(self.lib.orc_count_inverted(self.NT, _p(self.tets), _p(self.pos), _p(self.rest_matrix)))

    def residuals(self) -> dict:
        return residuals_numpy(self.cons, self.pos, self.rest_matrix, self.vf, self.tendonmask,
                               self.activation, self.params)


class CageRestriction:
    """Regularised least-squares restriction of the fine state onto the coarse
    lattice cage, per muscle block b (the Galerkin restrict of _restrict_cage,
    src/muscle/mg/mg.py:985-1011: restrict_galerkin_scatter_kernel +
    block_solve_kernel, src/muscle/mg/kernel.py:138-180, with a ridge):

        q_b = argmin_q ||P_b q - x_b||^2 + eps_b ||q - q_ref,b||^2

    P_b: the block's rows of the trilinear prolongation (p_idx, p_w);
    q_ref,b = A q0 + t, (A, t) the least-squares affine map of the block's rest
    vertices onto their current positions.  Assembled from p_idx / p_w and solved
    here in float64 (Cholesky per block), independent of the precomputed
    inverses the GPU path uploads.  ``eps_scale`` multiplies eps_b (0 gives the
    plain least-squares restriction, where P_b has full column rank).  A block
    with eps_b = 0 (the reference's voxel cage) is the Galerkin projection
    pinv(P^T P) P^T x with rcond 1e-8 (tools/gen_coarse_cage.py)."""

    def __init__(self, level, eps_scale: float = 1.0):
        from scipy.linalg import cho_factor
        self.lv = level
        self.blocks = []
        p_idx = np.asarray(level.p_idx, np.int64)
        p_w = np.asarray(level.p_w, np.float64)
        for b in range(len(level.blk_off) - 1):
            s, e = int(level.blk_off[b]), int(level.blk_off[b + 1])
            vids = np.asarray(level.fit_fine[level.fit_off[b]:level.fit_off[b + 1]], np.int64)
            n = e - s
            rows = np.repeat(np.arange(len(vids)), 8)
            cols = p_idx[vids].ravel() - s
            if cols.size and (cols.min() < 0 or cols.max() >= n):
                raise ValueError("a fine vertex of the block interpolates from another block")
            PtP = np.zeros((n, n))
            wv = p_w[vids]
            for k in range(8):
                for m in range(8):
                    np.add.at(PtP, (p_idx[vids, k] - s, p_idx[vids, m] - s), wv[:, k] * wv[:, m])
            eps = float(level.eps[b]) * float(eps_scale)
            if eps > 0.0:
                fac = ("cho", cho_factor(PtP + eps * np.eye(n)))
            else:   # Galerkin projection pinv(P^T P) P^T (tools/gen_coarse_cage.py, rcond 1e-8)
                fac = ("pinv", np.linalg.pinv(PtP, rcond=1e-8))
            self.blocks.append((s, e, vids, rows, cols, wv.ravel(), eps, fac))

    def __call__(self, x, rest) -> np.ndarray:
        from scipy.linalg import cho_solve
        x = np.asarray(x, np.float64).reshape(-1, 3)
        rest = np.asarray(rest, np.float64).reshape(-1, 3)
        q0 = np.asarray(self.lv.pos0, np.float64)
        q = np.zeros_like(q0)
        for s, e, vids, rows, cols, w, eps, fac in self.blocks:
            rhs = np.zeros((e - s, 3))
            for a in range(3):
                np.add.at(rhs[:, a], cols, w * x[vids][rows, a])
            if eps > 0.0:
                rh = np.concatenate([rest[vids], np.ones((len(vids), 1))], 1)
                T = np.linalg.lstsq(rh, x[vids], rcond=None)[0]          # (4, 3): x ~ [r 1] T
                rhs += eps * (np.concatenate([q0[s:e], np.ones((e - s, 1))], 1) @ T)
            q[s:e] = cho_solve(fac[1], rhs) if fac[0] == "cho" else fac[1] @ rhs
        return q


def residuals_numpy(cons, pos, rest_matrix, v_fiber, tendonmask, activation, params) -> dict:
    """Per-type residual RMS computed independently with NumPy (float64 SVD):
    ARAP ||sigma-1|| (arap_residual_accum_kernel, src/muscle/mg/kernel.py:87-118),
    volume V - V0, SNH (C_D, C_H), aniso 1/2 (|F f0| - fs)^2."""
    out = {}
    pos = pos.astype(np.float64)

    def F_of(sel):
        pts = cons.pts[sel]
        p3 = pos[pts[:, 3]]
        Ds = np.stack([pos[pts[:, 0]] - p3, pos[pts[:, 1]] - p3, pos[pts[:, 2]] - p3], axis=2)
        return Ds @ rest_matrix[cons.tetid[sel]].astype(np.float64), Ds

    t = cons.type
    sel = np.nonzero((t == ORC_TETARAPSVD) & (cons.stiffness > 0))[0]
    if sel.size:
        F, _ = F_of(sel)
        U, S, Vt = np.linalg.svd(F)
        sgn = np.sign(np.linalg.det(U)) * np.sign(np.linalg.det(Vt))
        S[:, 2] *= sgn
        out["arap"] = float(np.sqrt(np.mean(np.sum((S - 1.0) ** 2, axis=1))))
    sel = np.nonzero((t == ORC_TETVOLUME) & (cons.stiffness > 0))[0]
    if sel.size:
        _, Ds = F_of(sel)
        V = np.linalg.det(Ds) / 6.0
        out["volume"] = float(np.sqrt(np.mean((V - cons.restlength[sel]) ** 2)))
    sel = np.nonzero((t == ORC_TETSNH) & (cons.stiffness > 0))[0]
    if sel.size:
        F, _ = F_of(sel)
        mu = cons.stiffness[sel].astype(np.float64)
        lam = cons.restdir[sel, 0].astype(np.float64)
        cd = np.sqrt(np.maximum(np.sum(F * F, axis=(1, 2)), 1e-12)) - np.sqrt(3.0)
        ch = np.linalg.det(F) - (1.0 + mu / lam)
        out["snh"] = float(np.sqrt(np.mean(cd * cd + ch * ch)))
    sel = np.nonzero((t == ORC_TETARAPANISO) & (cons.stiffness > 0))[0]
    if sel.size:
        pts = cons.pts[sel]
        f0 = (v_fiber[pts[:, 0]] + v_fiber[pts[:, 1]] + v_fiber[pts[:, 2]] + v_fiber[pts[:, 3]]) / 4.0
        fn = np.linalg.norm(f0, axis=1)
        ok = fn >= 1e-9
        if ok.any():
            F, _ = F_of(sel[ok])
            f0 = f0[ok] / fn[ok, None]
            d = np.einsum("nij,nj->ni", F, f0)
            tid = cons.tetid[sel[ok]]
            fs = 1.0 - (1.0 - tendonmask[tid]) * activation[tid] * params.contraction_ratio
            C_ = 0.5 * (np.linalg.norm(d, axis=1) - fs) ** 2
            out["aniso"] = float(np.sqrt(np.mean(C_ * C_)))
    return out


# ---------------------------------------------------------------------------
# tri-tri DCD (src/muscle/collision/udcf/kernel.py)
# ---------------------------------------------------------------------------
class OrcContact(C.Structure):
    _fields_ = [("cur", C.c_int * 3), ("closest", C.c_int * 3), ("C", C.c_float), ("C1", C.c_float),
                ("C2", C.c_float), ("n", C.c_float * 3), ("b1", C.c_float * 3), ("b2", C.c_float * 3),
                ("valid", C.c_int)]


def _dcd_lib():
    lib = load()
    if not getattr(lib, "_dcd_bound", False):
        lib.orc_dcd_tri_data.argtypes = [C.c_int, _P, _P, _P, _P, _P]
        lib.orc_dcd_pairs_bruteforce.argtypes = [C.c_int, _P, _P, _P, _P, _P, C.c_float, C.c_int, C.c_float,
                                                 C.c_long, _P, _P]
        lib.orc_dcd_pairs_bruteforce.restype = C.c_long
        lib.orc_dcd_narrow.argtypes = [C.c_long, _P, _P, _P, _P, C.c_float, _P]
        lib.orc_dcd_respond.argtypes = [C.c_long, _P, _P, _P, _P, C.c_float, C.c_float, C.c_float, _P, _P]
        lib.orc_dcd_apply.argtypes = [C.c_int, _P, _P, _P, C.c_int]
        lib._dcd_bound = True
    return lib


def dcd_tri_data(tris, x):
    lib = _dcd_lib()
    tris = np.ascontiguousarray(tris, np.int32)
    x = np.ascontiguousarray(x, np.float32)
    m = tris.shape[0]
    c = np.zeros((m, 3), np.float32)
    r = np.zeros(m, np.float32)
    n = np.zeros((m, 3), np.float32)
    lib.orc_dcd_tri_data(m, _p(tris), _p(x), _p(c), _p(r), _p(n))
    return c, r, n


def dcd_pairs(tris, x, group=None, allow_self=True, thickness=0.0, normal_cull=1.0, max_pairs=1 << 22):
    """Sorted (i<j) candidate pairs by exhaustive enumeration of the emit rules."""
    lib = _dcd_lib()
    tris = np.ascontiguousarray(tris, np.int32)
    c, r, n = dcd_tri_data(tris, x)
    g = None if group is None else np.ascontiguousarray(group, np.int32)
    pa = np.zeros(max_pairs, np.int32)
    pb = np.zeros(max_pairs, np.int32)
    k = lib.orc_dcd_pairs_bruteforce(tris.shape[0], _p(tris), _p(c), _p(r), _p(n), _p(g),
                                     C.c_float(normal_cull), int(allow_self), C.c_float(thickness),
                                     max_pairs, _p(pa), _p(pb))
    k = min(int(k), max_pairs)
    return np.stack([pa[:k], pb[:k]], axis=1)


def dcd_narrow(pairs, tris, x, offset_eps=1e-6):
    lib = _dcd_lib()
    pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    pa = np.ascontiguousarray(pairs[:, 0])
    pb = np.ascontiguousarray(pairs[:, 1])
    out = (OrcContact * max(1, len(pairs)))()
    tris = np.ascontiguousarray(tris, np.int32)
    x = np.ascontiguousarray(x, np.float32)
    lib.orc_dcd_narrow(len(pairs), _p(pa), _p(pb), _p(tris), _p(x), C.c_float(offset_eps), out)
    return out, len(pairs)


def dcd_respond_apply(contacts, ncontacts, x, x_prev, inv_mass, multiplier, mu_s, mu_d, use_avg=True):
    lib = _dcd_lib()
    x = np.ascontiguousarray(x, np.float32).copy()
    xp = np.ascontiguousarray(x_prev, np.float32)
    im = np.ascontiguousarray(inv_mass, np.float32)
    dP = np.zeros_like(x)
    dPw = np.zeros(x.shape[0], np.float32)
    lib.orc_dcd_respond(ncontacts, contacts, _p(x), _p(xp), _p(im), C.c_float(multiplier), C.c_float(mu_s),
                        C.c_float(mu_d), _p(dP), _p(dPw))
    lib.orc_dcd_apply(x.shape[0], _p(x), _p(dP), _p(dPw), int(use_avg))
    return x


def make_contact(cur, closest, Cval, n, b1, b2):
    c = (OrcContact * 1)()
    c[0].cur[:] = list(cur)
    c[0].closest[:] = list(closest)
    c[0].C = Cval
    c[0].n[:] = list(n)
    c[0].b1[:] = list(b1)
    c[0].b2[:] = list(b2)
    c[0].valid = 1
    return c
