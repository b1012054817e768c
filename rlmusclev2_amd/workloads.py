"""Seeded synthetic workloads (the reference ships no asset-free scenes).

* ``cantilever``: a box split into hex cells, each cut into 6 tets with the
  corner pattern of tools/gen_coarse_cage.py:43-44, rest positions jittered.
* ``humanv2_scale``: fusiform spindle muscles packed around capsule bones,
  tet-meshed by lattice clipping, plus a fascia shell (tube) per bone; tendon
  tips pinned and driven by rigid bone motion; capsule SDFs baked on grids.
* ``batch``: many independent spindle instances of uneven size, each with its
  own capsule bone, packed into one mesh.

All randomness comes from ``numpy.random.Generator(PCG64(seed))``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# 6-tet split of a unit cube; corner b has offset bits (b&1, (b>>1)&1, (b>>2)&1)
CUBE6 = np.array([(0, 1, 3, 7), (0, 3, 2, 7), (0, 2, 6, 7),
                  (0, 6, 4, 7), (0, 4, 5, 7), (0, 5, 1, 7)], np.int64)
_CORNER = np.array([(b & 1, (b >> 1) & 1, (b >> 2) & 1) for b in range(8)], np.int64)


def _oriented_cube6():
    """CUBE6 with p0/p1 swapped where needed so every tet has det(Dm) > 0 (p3 reference)."""
    out = CUBE6.copy()
    for t in range(6):
        p = _CORNER[out[t]].astype(np.float64)
        d = np.linalg.det(np.stack([p[0] - p[3], p[1] - p[3], p[2] - p[3]], axis=1))
        if d < 0:
            out[t, [0, 1]] = out[t, [1, 0]]
    return out


CUBE6_POS = _oriented_cube6()


def lattice_mesh(keep: np.ndarray, origin, h: float):
    """Tet mesh of the kept cells of an (nx,ny,nz) lattice: returns pos (nv,3) f64,
    tets (nt,4) i32, lattice index (nv,3) of each vertex."""
    nx, ny, nz = keep.shape
    ci = np.argwhere(keep)                                   # (ncell, 3) i,j,k
    if ci.shape[0] == 0:
        return np.zeros((0, 3)), np.zeros((0, 4), np.int32), np.zeros((0, 3), np.int64)
    sx, sy = nx + 1, ny + 1
    corner_ids = []
    for b in range(8):
        q = ci + _CORNER[b]
        corner_ids.append(q[:, 0] + sx * (q[:, 1] + sy * q[:, 2]))
    corner_ids = np.stack(corner_ids, axis=1)                # (ncell, 8) dense ids
    used, inv = np.unique(corner_ids, return_inverse=True)
    inv = inv.reshape(corner_ids.shape)
    tets = inv[:, CUBE6_POS].reshape(-1, 4).astype(np.int32)
    li = np.stack([used % sx, (used // sx) % sy, used // (sx * sy)], axis=1)
    pos = np.asarray(origin, np.float64) + li * h
    return pos, tets, li


@dataclass
class Workload:
    name: str
    pos0: np.ndarray
    tets: np.ndarray
    density: float = 1060.0
    v_fiber_dir: np.ndarray | None = None
    v_tendonmask: np.ndarray | None = None
    vertex_group: np.ndarray | None = None     # bone / instance per vertex
    tet_instance: np.ndarray | None = None
    is_fascia: np.ndarray | None = None
    pin_vids: np.ndarray | None = None
    pin_group: np.ndarray | None = None
    stopped: np.ndarray | None = None
    bones: list = field(default_factory=list)  # dicts: a, b, radius, pivot, axis, amp, phase
    info: dict = field(default_factory=dict)

    @property
    def nv(self):
        return int(self.pos0.shape[0])

    @property
    def nt(self):
        return int(self.tets.shape[0])


def cantilever(seed: int = 0, n=(16, 4, 4), size=(0.4, 0.1, 0.1), jitter=1e-4) -> Workload:
    """16x4x4 hex lattice, 6 tets per cube (384 tets / 425 verts), x=0 face pinned."""
    rng = np.random.Generator(np.random.PCG64(seed))
    h = size[0] / n[0]
    keep = np.ones(n, bool)
    pos, tets, li = lattice_mesh(keep, (0.0, 0.0, 0.0), h)
    pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    pinned = np.nonzero(li[:, 0] == 0)[0].astype(np.int32)
    stopped = np.zeros(pos.shape[0], np.int32)
    stopped[pinned] = 1
    return Workload("cantilever", pos.astype(np.float32), tets, density=1060.0,
                    pin_vids=pinned, stopped=stopped, info={"h": h})


# ---------------------------------------------------------------------------
# spindles and capsules
# ---------------------------------------------------------------------------
def _rotation_to(axis):
    """Proper rotation taking +z to unit ``axis``."""
    z = np.array([0.0, 0.0, 1.0])
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    v = np.cross(z, a)
    s = np.linalg.norm(v)
    c = float(np.dot(z, a))
    if s < 1e-12:
        return np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
    K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]]) / s
    ang = np.arctan2(s, c)
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)


def spindle_profile(s):
    """Fusiform radius factor at normalised axial coordinate s in [-1, 1]."""
    return np.clip(1.0 - s * s, 0.0, None) ** 0.75


def spindle_mesh(length, radius, h, center, axis):
    """Lattice-clipped spindle (tapered ellipsoid along local z). Returns pos,
    tets, s (axial coordinate in [-1,1] per vertex)."""
    nxy = int(np.ceil(2 * radius / h)) + 1
    nz = int(np.ceil(length / h)) + 1
    org = np.array([-nxy * h / 2, -nxy * h / 2, -nz * h / 2])
    ii, jj, kk = np.meshgrid(np.arange(nxy), np.arange(nxy), np.arange(nz), indexing="ij")
    cx = org[0] + (ii + 0.5) * h
    cy = org[1] + (jj + 0.5) * h
    cz = org[2] + (kk + 0.5) * h
    s = 2.0 * cz / length
    keep = (np.abs(s) < 1.0) & (np.hypot(cx, cy) < radius * spindle_profile(s))
    pos, tets, _ = lattice_mesh(keep, org, h)
    sv = np.clip(2.0 * pos[:, 2] / length, -1.0, 1.0)
    R = _rotation_to(axis)
    pos = pos @ R.T + np.asarray(center, np.float64)
    return pos, tets, sv


def tube_mesh(r_in, thickness, length, h, center, axis):
    n = int(np.ceil(2 * (r_in + thickness) / h)) + 2
    nz = int(np.ceil(length / h)) + 1
    org = np.array([-n * h / 2, -n * h / 2, -nz * h / 2])
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(nz), indexing="ij")
    cx = org[0] + (ii + 0.5) * h
    cy = org[1] + (jj + 0.5) * h
    cz = org[2] + (kk + 0.5) * h
    r = np.hypot(cx, cy)
    keep = (r >= r_in) & (r < r_in + thickness) & (np.abs(cz) < length / 2)
    pos, tets, _ = lattice_mesh(keep, org, h)
    sv = np.clip(2.0 * pos[:, 2] / length, -1.0, 1.0)
    R = _rotation_to(axis)
    return pos @ R.T + np.asarray(center, np.float64), tets, sv


def capsule_sdf_grid(a, b, radius, res, margin):
    """Capsule SDF on a res^3 node grid in the bone's local (= rest world) frame.
    Returns grid[z, y, x] f32, origin (3,), spacing h."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    lo = np.minimum(a, b) - radius - margin
    hi = np.maximum(a, b) + radius + margin
    h = float(np.max(hi - lo) / (res - 1))
    origin = 0.5 * (lo + hi) - 0.5 * h * (res - 1)
    t = np.arange(res) * h
    # node coordinates relative to a, broadcast over (z, y, x)
    dz = (origin[2] + t - a[2])[:, None, None]
    dy = (origin[1] + t - a[1])[None, :, None]
    dx = (origin[0] + t - a[0])[None, None, :]
    ab = b - a
    u = np.clip((dx * ab[0] + dy * ab[1] + dz * ab[2]) / float(ab @ ab), 0.0, 1.0)
    phi = np.sqrt((dx - u * ab[0]) ** 2 + (dy - u * ab[1]) ** 2 + (dz - u * ab[2]) ** 2) - radius
    return phi.astype(np.float32), origin.astype(np.float32), np.float32(h)


def capsule_phi(p, a, b, radius):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    ab = b - a
    u = np.clip(((np.asarray(p, np.float64) - a) @ ab) / float(ab @ ab), 0.0, 1.0)
    q = a + u[..., None] * ab
    return np.linalg.norm(p - q, axis=-1) - radius


def capsule_mesh(a, b, radius, n_around=96, n_cap=16, n_len=96):
    """Closed triangle mesh of the capsule (segment a-b, radius), outward winding:
    (p1 - p0) x (p2 - p0) points out.  Rings of n_around vertices: n_cap latitudes
    on each cap (the equators included), n_len - 1 rings along the cylinder, and
    the two poles.  Returns (verts (n, 3) float64, tris (m, 3) int32)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    u = b - a
    u /= np.linalg.norm(u)
    e1 = np.cross(u, [1.0, 0.0, 0.0] if abs(u[0]) < 0.9 else [0.0, 1.0, 0.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(u, e1)                      # e1 x e2 = u
    th = 2.0 * np.pi * np.arange(n_around) / n_around
    ring_dir = np.cos(th)[:, None] * e1 + np.sin(th)[:, None] * e2
    rings = []
    for i in range(1, n_cap + 1):             # bottom cap, from the pole side up to a's equator
        ph = -0.5 * np.pi + 0.5 * np.pi * i / n_cap
        rings.append(a + radius * (np.cos(ph) * ring_dir + np.sin(ph) * u))
    for j in range(1, n_len):
        rings.append(a + (j / n_len) * (b - a) + radius * ring_dir)
    for i in range(n_cap):                    # top cap, from b's equator towards the pole
        ph = 0.5 * np.pi * i / n_cap
        rings.append(b + radius * (np.cos(ph) * ring_dir + np.sin(ph) * u))
    nr = len(rings)
    verts = np.concatenate([np.stack(rings).reshape(-1, 3), [a - radius * u], [b + radius * u]])
    bot, top = nr * n_around, nr * n_around + 1
    j = np.arange(n_around)
    jn = (j + 1) % n_around
    tris = []
    for k in range(nr - 1):
        p00, p01 = k * n_around + j, k * n_around + jn
        p10, p11 = (k + 1) * n_around + j, (k + 1) * n_around + jn
        tris.append(np.stack([p00, p01, p11], 1))
        tris.append(np.stack([p00, p11, p10], 1))
    tris.append(np.stack([np.full(n_around, bot), jn, j], 1))
    last = (nr - 1) * n_around
    tris.append(np.stack([last + j, last + jn, np.full(n_around, top)], 1))
    return verts, np.concatenate(tris).astype(np.int32)


def bone_meshes(w, n_around=96, n_cap=16, n_len=96):
    """Triangle meshes of the workload's capsule bones in the rest pose, one
    sub-mesh per bone: (verts, tris, tri_off, vertex bone)."""
    key = ("_bone_mesh", n_around, n_cap, n_len)
    if key not in w.info:
        vs, ts, toff, vb = [], [], [0], []
        off = 0
        for i, bn in enumerate(w.bones):
            v, t = capsule_mesh(bn["a"], bn["b"], bn["radius"], n_around, n_cap, n_len)
            vs.append(v)
            ts.append(t + off)
            toff.append(toff[-1] + t.shape[0])
            vb.append(np.full(v.shape[0], i, np.int32))
            off += v.shape[0]
        w.info[key] = (np.concatenate(vs).astype(np.float32), np.concatenate(ts),
                       np.asarray(toff, np.int32), np.concatenate(vb))
    return w.info[key]


def bone_mesh_points(w, frame: int, **res) -> np.ndarray:
    """Bone mesh vertices at ``frame``: each bone's rigid motion on its rest vertices."""
    verts, _, _, vb = bone_meshes(w, **res)
    out = np.empty_like(verts)
    for i, bn in enumerate(w.bones):
        R, t = bone_transform(bn, frame)
        sel = vb == i
        out[sel] = (verts[sel].astype(np.float64) @ R.T + t).astype(np.float32)
    return out


def rotation_about(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def bone_transform(bone: dict, frame: int, fps: float = 24.0):
    """Rigid swing about the bone's pivot, zero at t=0:
    theta = amp*(sin(2*pi*freq*t + phase) - sin(phase)), |dtheta/dt| <= amp*2*pi*freq.
    Returns (R, t) with world = R @ local + t."""
    tt = frame / fps
    w = 2 * np.pi * bone["freq"]
    ang = bone["amp"] * (np.sin(w * tt + bone["phase"]) - np.sin(bone["phase"]))
    R = rotation_about(bone["axis"], ang)
    piv = np.asarray(bone["pivot"], np.float64)
    return R, piv - R @ piv


def tip_pins(s, frac=0.05):
    """Indices of the lowest-s and highest-s ``frac`` of vertices (rank order)."""
    n = s.shape[0]
    k = max(1, int(round(frac * n)))
    order = np.argsort(s, kind="stable")
    return np.concatenate([order[:k], order[-k:]])


def humanv2_scale(seed: int = 0, n_bones=4, per_bone=10, target_tets=468_000,
                  fascia_tets=67_000, sdf_res=128) -> Workload:
    """40 spindles around 4 capsule bones (~468k tets) + 4 fascia tubes (~67k tets)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rb = 0.025
    bone_len = 0.5
    bones = []
    muscles = []
    for b in range(n_bones):
        x0 = b * 0.9
        bones.append(dict(a=np.array([x0, 0.0, -bone_len / 2]), b=np.array([x0, 0.0, bone_len / 2]),
                          radius=rb, pivot=np.array([x0, 0.0, bone_len / 2]),
                          axis=np.array([1.0, 0.0, 0.0]) if b % 2 == 0 else np.array([0.0, 1.0, 0.0]),
                          amp=np.deg2rad(9.0), freq=0.5, phase=float(rng.uniform(0, 2 * np.pi))))
        for k in range(per_bone):
            L = float(rng.uniform(0.15, 0.45))
            R = float(rng.uniform(0.015, 0.05))
            ring = k % 2
            ang = 2 * np.pi * (k // 2) / (per_bone // 2) + ring * np.pi / (per_bone // 2)
            d = rb + 0.98 * R
            zc = (-0.12 if ring == 0 else 0.12) + float(rng.uniform(-0.03, 0.03))
            tilt = rng.normal(0, 0.08, size=2)
            axis = np.array([tilt[0], tilt[1], 1.0])
            muscles.append(dict(L=L, R=R, center=np.array([x0 + d * np.cos(ang), d * np.sin(ang), zc]),
                                axis=axis / np.linalg.norm(axis), bone=b))
    vol = sum(np.pi * m["R"] ** 2 * m["L"] * (3 * np.pi / 16) for m in muscles)
    h = (6.0 * vol / target_tets) ** (1.0 / 3.0)
    for _ in range(3):
        nt = sum(spindle_mesh(m["L"], m["R"], h, m["center"], m["axis"])[1].shape[0] for m in muscles)
        h *= (nt / target_tets) ** (1.0 / 3.0)
    parts_p, parts_t, s_all, grp, mid, fas, fib = [], [], [], [], [], [], []
    pins, pin_grp = [], []
    off = 0
    for i, m in enumerate(muscles):
        p, t, s = spindle_mesh(m["L"], m["R"], h, m["center"], m["axis"])
        parts_p.append(p)
        parts_t.append(t + off)
        s_all.append(s)
        grp.append(np.full(p.shape[0], m["bone"], np.int32))
        mid.append(np.full(p.shape[0], i, np.int32))
        fas.append(np.zeros(p.shape[0], bool))
        fib.append(np.tile(m["axis"], (p.shape[0], 1)))
        pv = tip_pins(s) + off
        pins.append(pv)
        pin_grp.append(np.full(pv.shape[0], m["bone"], np.int32))
        off += p.shape[0]
    # fascia tubes around each bone's muscle ring
    tubes = []
    for b in range(n_bones):
        ms = [m for m in muscles if m["bone"] == b]
        x0 = bones[b]["a"][0]
        r_in = max(np.hypot(m["center"][0] - x0, m["center"][1]) + m["R"] for m in ms) + 0.005
        tubes.append(r_in)
    # 2-cell thick shell: tets ~ 6 * 2*pi*r*len*2h / h^3
    tl = 0.55
    hf = np.sqrt(sum(6 * 2 * np.pi * r * tl * 2 for r in tubes) / fascia_tets)
    for _ in range(3):
        nt = sum(tube_mesh(r, 2 * hf, tl, hf, (bones[b]["a"][0], 0, 0), (0, 0, 1))[1].shape[0]
                 for b, r in enumerate(tubes))
        hf *= (nt / fascia_tets) ** 0.5
    for b, r in enumerate(tubes):
        p, t, s = tube_mesh(r, 2 * hf, tl, hf, (bones[b]["a"][0], 0, 0), (0, 0, 1))
        parts_p.append(p)
        parts_t.append(t + off)
        s_all.append(s)
        grp.append(np.full(p.shape[0], b, np.int32))
        mid.append(np.full(p.shape[0], len(muscles) + b, np.int32))
        fas.append(np.ones(p.shape[0], bool))
        fib.append(np.zeros((p.shape[0], 3)))
        pv = tip_pins(s) + off
        pins.append(pv)
        pin_grp.append(np.full(pv.shape[0], b, np.int32))
        off += p.shape[0]
    pos = np.concatenate(parts_p).astype(np.float32)
    tets = np.concatenate(parts_t).astype(np.int32)
    s_all = np.concatenate(s_all)
    is_f = np.concatenate(fas)
    tm = np.clip((np.abs(s_all) - 0.7) / 0.2, 0.0, 1.0).astype(np.float32)
    tm[is_f] = 0.0
    pin_vids = np.concatenate(pins).astype(np.int32)
    return Workload(
        "humanv2_scale", pos, tets, density=1060.0,
        v_fiber_dir=np.concatenate(fib).astype(np.float32), v_tendonmask=tm,
        vertex_group=np.concatenate(grp), is_fascia=is_f, pin_vids=pin_vids,
        pin_group=np.concatenate(pin_grp), bones=bones,
        info=dict(h_muscle=h, h_fascia=hf, muscle_id=np.concatenate(mid), s=s_all,
                  n_muscle_tets=int(sum(t.shape[0] for t in parts_t[:len(muscles)])),
                  sdf_res=sdf_res))


def split_fascia(w: Workload):
    """The humanV2-scale workload with its fascia tubes taken out as their own mesh
    (the reference's fascia is a separate sub-solver with its own tet mesh,
    fascia.py:45-60). Returns (muscle workload, fascia positions, fascia tets)."""
    fas = np.asarray(w.is_fascia, bool)
    keep = ~fas
    tf = fas[w.tets].all(1)
    if (fas[w.tets].any(1) & ~tf).any():
        raise ValueError("fascia and muscle tets share vertices")
    new = np.full(w.nv, -1, np.int64)
    new[keep] = np.arange(int(keep.sum()))
    fnew = np.full(w.nv, -1, np.int64)
    fnew[fas] = np.arange(int(fas.sum()))
    pin_keep = keep[w.pin_vids] if w.pin_vids is not None else None
    info = dict(w.info)
    info.pop("_pin_groups", None)
    if info.get("muscle_id") is not None:
        info["muscle_id"] = np.asarray(info["muscle_id"])[keep]
    if info.get("s") is not None:
        info["s"] = np.asarray(info["s"])[keep]
    m = Workload(
        w.name, w.pos0[keep], new[w.tets[~tf]].astype(np.int32), density=w.density,
        v_fiber_dir=None if w.v_fiber_dir is None else w.v_fiber_dir[keep],
        v_tendonmask=None if w.v_tendonmask is None else w.v_tendonmask[keep],
        vertex_group=None if w.vertex_group is None else w.vertex_group[keep],
        tet_instance=None if w.tet_instance is None else w.tet_instance[~tf],
        is_fascia=np.zeros(int(keep.sum()), bool),
        pin_vids=None if w.pin_vids is None else new[w.pin_vids[pin_keep]].astype(np.int32),
        pin_group=None if w.pin_group is None else w.pin_group[pin_keep],
        stopped=None if w.stopped is None else w.stopped[keep], bones=w.bones, info=info)
    return m, w.pos0[fas].astype(np.float32), fnew[w.tets[tf]].astype(np.int32)


def bake_bone_sdfs(w: Workload, res: int | None = None, margin: float = 0.03):
    """Capsule SDF per bone on a res^3 grid (own spacing per bone)."""
    res = int(res or w.info.get("sdf_res", 64))
    grids, origins, hs = [], [], []
    for bn in w.bones:
        g, o, h = capsule_sdf_grid(bn["a"], bn["b"], bn["radius"], res, margin)
        grids.append(g)
        origins.append(o)
        hs.append(h)
    return np.stack(grids), np.stack(origins), np.asarray(hs, np.float32)


def bone_xforms(w: Workload, frame: int) -> np.ndarray:
    out = np.zeros((len(w.bones), 12), np.float32)
    for i, bn in enumerate(w.bones):
        R, t = bone_transform(bn, frame)
        out[i] = np.concatenate([R, t[:, None]], axis=1).reshape(-1)
    return out


def _pin_groups(w: Workload):
    """(rest positions of the pins, per-bone pin index lists), built once per workload."""
    cache = w.info.get("_pin_groups")
    if cache is None or cache[0] is not w.pin_vids:
        x0 = w.pos0[w.pin_vids].astype(np.float64)
        order = np.argsort(w.pin_group, kind="stable")
        cuts = np.searchsorted(w.pin_group[order], np.arange(len(w.bones) + 1))
        cache = (w.pin_vids, x0, [order[cuts[b]:cuts[b + 1]] for b in range(len(w.bones))])
        w.info["_pin_groups"] = cache
    return cache[1], cache[2]


def pin_targets(w: Workload, frame: int) -> np.ndarray:
    """Rigid bone motion applied to the rest position of each pinned vertex."""
    x0, groups = _pin_groups(w)
    out = np.empty_like(x0)
    for b, bn in enumerate(w.bones):
        sel = groups[b]
        if sel.size:
            R, t = bone_transform(bn, frame)
            out[sel] = x0[sel] @ R.T + t
    return out.astype(np.float32)


def follow_targets(w: Workload, frame: int, vids: np.ndarray) -> np.ndarray:
    """Rest positions of ``vids`` moved rigidly by the bone of each vertex
    (``vertex_group``): the belly-attach goal target_pos + rest_offset of
    MultiGridSim._update_belly_attach_targets (mg.py:725-734) for a synthetic
    scene whose targets ride the bones."""
    key = ("_follow", id(vids))
    cache = w.info.get(key)
    if cache is None or cache[0] is not vids:
        v = np.asarray(vids, np.int64)
        grp = np.asarray(w.vertex_group, np.int64)[v]
        order = np.argsort(grp, kind="stable")
        cuts = np.searchsorted(grp[order], np.arange(len(w.bones) + 1))
        cache = (vids, w.pos0[v].astype(np.float64), [order[cuts[b]:cuts[b + 1]] for b in range(len(w.bones))])
        w.info[key] = cache
    x0, groups = cache[1], cache[2]
    out = x0.copy()
    for b, bn in enumerate(w.bones):
        sel = groups[b]
        if sel.size:
            R, t = bone_transform(bn, frame)
            out[sel] = x0[sel] @ R.T + t
    return out.astype(np.float32)


def batch(seed: int = 0, n_inst: int = 512, lo: int = 2000, hi: int = 40000, h: float = 0.004,
          sdf_res: int = 64) -> Workload:
    """Independent spindle instances, tet count log-uniform in [lo, hi], each with
    its own capsule bone; instances laid out on a grid so they never touch."""
    rng = np.random.Generator(np.random.PCG64(seed))
    targets = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_inst))
    parts_p, parts_t, grp, fib, tm_all, pins, pin_grp, tet_inst = [], [], [], [], [], [], [], []
    bones = []
    side = int(np.ceil(n_inst ** (1 / 3)))
    pitch = 0.6
    off = 0
    for i in range(n_inst):
        aspect = float(rng.uniform(5.0, 9.0))        # L / R
        # tets ~ 6 * V / h^3, V = pi R^2 L * 3pi/16
        V = targets[i] * h ** 3 / 6.0
        R = (V / (np.pi * aspect * 3 * np.pi / 16)) ** (1 / 3)
        L = aspect * R
        gi = np.array([i % side, (i // side) % side, i // (side * side)], np.float64)
        c = gi * pitch
        rb = 0.5 * R
        bone_c = c + np.array([R + 0.9 * rb, 0.0, 0.0])
        p, t, s = spindle_mesh(L, R, h, c, (0, 0, 1))
        parts_p.append(p)
        parts_t.append(t + off)
        grp.append(np.full(p.shape[0], i, np.int32))
        tet_inst.append(np.full(t.shape[0], i, np.int32))
        fib.append(np.tile([0.0, 0.0, 1.0], (p.shape[0], 1)))
        tm_all.append(np.clip((np.abs(s) - 0.7) / 0.2, 0, 1))
        pv = tip_pins(s) + off
        pins.append(pv)
        pin_grp.append(np.full(pv.shape[0], i, np.int32))
        bones.append(dict(a=bone_c - [0, 0, 0.4 * L], b=bone_c + [0, 0, 0.4 * L], radius=rb,
                          pivot=bone_c + [0, 0, 0.4 * L], axis=np.array([1.0, 0.0, 0.0]),
                          amp=np.deg2rad(9.0), freq=0.5, phase=float(rng.uniform(0, 2 * np.pi))))
        off += p.shape[0]
    pos = np.concatenate(parts_p).astype(np.float32)
    tets = np.concatenate(parts_t).astype(np.int32)
    return Workload(
        "batch", pos, tets, density=1060.0, v_fiber_dir=np.concatenate(fib).astype(np.float32),
        v_tendonmask=np.concatenate(tm_all).astype(np.float32), vertex_group=np.concatenate(grp),
        tet_instance=np.concatenate(tet_inst), is_fascia=np.zeros(pos.shape[0], bool),
        pin_vids=np.concatenate(pins).astype(np.int32), pin_group=np.concatenate(pin_grp),
        bones=bones, info=dict(n_inst=n_inst, sdf_res=sdf_res, target_tets=targets, h=h))
