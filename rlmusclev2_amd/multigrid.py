"""Two-level hierarchical correction for MultiGridSim: a coarse hexahedral
lattice (spacing ``factor`` x the fine lattice spacing) around each muscle
(connected component of the fine mesh), split into tets, with trilinear
prolongation P.

Per substep, after integrate (MultiGridSim._hpbd_vcycle, src/muscle/mg/mg.py:1064-1079):

1. restrict the fine state onto the coarse nodes by regularised least squares,
   ``q = (P^T P + eps I)^-1 (P^T x + eps q_ref)``, one dense block per muscle
   (``_restrict_cage`` with ``mg_sync_restrict = "galerkin"``: the scatter
   ``restrict_galerkin_scatter_kernel`` and the per-muscle ``block_solve_kernel``,
   src/muscle/mg/kernel.py:138-180).  ``q_ref`` is the coarse rest lattice moved
   by the muscle's least-squares affine fit of its fine vertices, so nodes the
   fine vertices barely constrain follow the muscle instead of collapsing (the
   reference takes a pseudo-inverse, which sends such directions to zero);
2. save ``q_prev = q``; run ``iters`` Jacobi ARAP-SVD sweeps on the coarse
   cage with the coarse pins held (``_coarse_solve`` / ``_coarse_arap_sweep``,
   mg.py:939-952, 1047-1062; lambda = 0, no coarse attach, beta = 0 is the
   reference default);
3. add the coarse displacement back, ``x_i += sum_j P_ij (q_j - q_prev_j)``
   (``prolong_delta_kernel``, mg/kernel.py:9-40), then re-pin.

The reference embeds the fine vertices in an enclosing tetrahedral cage
(barycentric weights).  Here the cage is a regular lattice and P is
trilinear, so P reproduces affine fields; together with the affine ``q_ref``
the restriction returns the lattice rest pose for the fine rest pose and the
moved lattice for any affine (in particular rigid) motion of a muscle.  A
coarse node is a hard pin when at least ``pin_frac`` of its P-weight comes
from pinned fine vertices
(``mg_coarse_pin_frac``, mg.py:217-222).  Fascia components (a thin shell,
not a muscle) get no cage: their P weights are zero.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import builder

# Kuhn split of the unit cube into 6 tets around the 000-111 diagonal; conforming
# between neighbouring cells because every cell uses the same diagonal.
_KUHN = []
for _a, _b, _c in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
    e = np.eye(3, dtype=np.int64)
    _KUHN.append((np.zeros(3, np.int64), e[_a], e[_a] + e[_b], e[_a] + e[_b] + e[_c]))


# limits of the one-workgroup-per-block coarse solve (k_mg_coarse: 1,024 threads)
MAX_BLOCK_NODES = 1024
MAX_BLOCK_TETS = 2048


@dataclass
class CoarseLevel:
    pos0: np.ndarray        # (nc, 3) coarse rest positions (lattice nodes)
    tets: np.ndarray        # (nct, 4)
    mass: np.ndarray        # (nc,) sum_i P_ij m_i
    stopped: np.ndarray     # (nc,) 1 = coarse pin (kept at its restricted position by every sweep)
    p_idx: np.ndarray       # (nv, 8) coarse node of each trilinear corner
    p_w: np.ndarray         # (nv, 8) trilinear weights (sum to 1; 0 for vertices without a cage)
    blk_off: np.ndarray     # (nblk+1,) coarse node range of each muscle block
    minv: np.ndarray        # concatenated row-major (P^T P + eps I)^-1 blocks
    eps: np.ndarray         # (nblk,) ridge of each block
    fit_off: np.ndarray     # (nblk+1,) range of each block in fit_fine
    fit_fine: np.ndarray    # fine vertices of each block (affine fit)
    fit_center: np.ndarray  # (nblk, 3) rest centroid of each block
    fit_sinv: np.ndarray    # (nblk, 4, 4) inverse of sum [x-c;1][x-c;1]^T over the block at rest
    iters: int
    stiffness: float
    spacing: np.ndarray     # per block
    n_components: int
    # pin rule: 0 = coarse pins take part in the sweeps and keep their restricted
    # positions; 1 = the reference's (mg.py:1013-1062): pins are moved to the
    # restricted bone goal and held with zero inverse mass
    pin_rule: int = 0
    blk_vgroup: np.ndarray | None = None   # (nblk,) the fine vertex group (bone) of each block


def components(nv: int, tets: np.ndarray) -> np.ndarray:
    """Connected components of the tet mesh (vertices without tets get their own)."""
    parent = np.arange(nv)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for t in np.asarray(tets, np.int64):
        r0 = find(t[0])
        for k in (1, 2, 3):
            rk = find(t[k])
            if rk != r0:
                parent[rk] = r0
    roots = np.array([find(i) for i in range(nv)])
    _, comp = np.unique(roots, return_inverse=True)
    return comp.astype(np.int32)


def build_coarse_level(pos0, tets, mass, pin_mask=None, factor: float = 4.0, iters: int = 15,
                       stiffness: float = 1.0e6, pin_frac: float = 0.5, comp=None, skip=None,
                       ridge: float = 5e-2, mass_floor: float = 0.05) -> CoarseLevel:
    """``comp``: per-vertex component id (default: connected components);
    ``skip``: per-vertex mask of vertices that get no cage (e.g. fascia)."""
    pos0 = np.asarray(pos0, np.float64).reshape(-1, 3)
    tets = np.asarray(tets, np.int64).reshape(-1, 4)
    nv = pos0.shape[0]
    comp = components(nv, tets) if comp is None else np.asarray(comp, np.int64)
    skip = np.zeros(nv, bool) if skip is None else np.asarray(skip, bool)
    vols = np.abs(np.einsum("ij,ij->i", np.cross(pos0[tets[:, 1]] - pos0[tets[:, 0]],
                                                 pos0[tets[:, 2]] - pos0[tets[:, 0]]),
                            pos0[tets[:, 3]] - pos0[tets[:, 0]])) / 6.0
    tcomp = comp[tets[:, 0]]
    p_idx = np.zeros((nv, 8), np.int64)
    p_w = np.zeros((nv, 8), np.float64)
    node_pos, all_tets, spacing, blk_off, minv = [], [], [], [0], []
    eps_l, fit_off, fit_fine, fit_c, fit_s = [], [0], [], [], []
    base = 0
    corner_off = np.array([[i >> 2 & 1, i >> 1 & 1, i & 1] for i in range(8)])
    for c in (np.unique(comp[~skip]) if nv else []):
        vids = np.nonzero((comp == c) & ~skip)[0]
        tv = vols[tcomp == c]
        # fine lattice spacing: a cube of side h holds 6 tets of volume h^3/6
        h = (6.0 * tv.mean()) ** (1.0 / 3.0) if tv.size else 1.0
        H = factor * h
        x = pos0[vids]
        origin = x.min(0) - 0.125 * H
        rel = (x - origin) / H
        cell = np.floor(rel).astype(np.int64)
        uvw = rel - cell
        used = np.unique(cell, axis=0)
        nodes = np.unique((used[:, None, :] + corner_off[None]).reshape(-1, 3), axis=0)
        dims = nodes.max(0) + 2
        key = lambda ijk: (ijk[..., 0] * dims[1] + ijk[..., 1]) * dims[2] + ijk[..., 2]
        nkeys = key(nodes)
        order = np.argsort(nkeys)
        skeys = nkeys[order]

        def nid(ijk):
            return order[np.searchsorted(skeys, key(ijk))]

        n = len(nodes)
        # the one-workgroup coarse solve holds a block's nodes one per thread and its
        # tets two per thread
        if n > MAX_BLOCK_NODES or 6 * len(used) > MAX_BLOCK_TETS:
            raise ValueError(f"coarse cage of component {c} has {n} nodes / {6 * len(used)} tets; the "
                             f"coarse solve takes at most {MAX_BLOCK_NODES} nodes and {MAX_BLOCK_TETS} "
                             f"tets per block (raise mg_coarse_factor)")
        Pl = np.zeros((len(vids), n))
        for k in range(8):
            off = corner_off[k]
            j = nid(cell + off)
            wk = np.ones(len(vids))
            for a in range(3):
                wk = wk * (uvw[:, a] if off[a] else 1.0 - uvw[:, a])
            p_idx[vids, k] = base + j
            p_w[vids, k] = wk
            Pl[np.arange(len(vids)), j] += wk
        for t in _KUHN:
            all_tets.append(np.stack([base + nid(used + o) for o in t], 1))
        node_pos.append(origin + nodes * H)
        PtP = Pl.T @ Pl
        eps = ridge * float(np.mean(np.diag(PtP)))
        minv.append(np.linalg.inv(PtP + eps * np.eye(n)).ravel())
        eps_l.append(eps)
        ctr = x.mean(0)
        xh = np.concatenate([x - ctr, np.ones((len(vids), 1))], 1)
        fit_s.append(np.linalg.inv(xh.T @ xh))
        fit_c.append(ctr)
        fit_fine.append(vids)
        fit_off.append(fit_off[-1] + len(vids))
        spacing.append(H)
        base += n
        blk_off.append(base)
    nc = base
    cpos = np.concatenate(node_pos) if node_pos else np.zeros((0, 3))
    ctets = np.concatenate(all_tets).astype(np.int32) if all_tets else np.zeros((0, 4), np.int32)
    wsum = np.zeros(nc)
    np.add.at(wsum, p_idx.ravel(), p_w.ravel())
    cmass = np.zeros(nc)
    np.add.at(cmass, p_idx.ravel(), (p_w * np.asarray(mass, np.float64)[:, None]).ravel())
    # nodes at the edge of a muscle get little P-weight; a floor keeps the coarse
    # inverse masses within a bounded ratio so the coarse Jacobi sweep stays stable
    for k in range(len(blk_off) - 1):
        sl = slice(blk_off[k], blk_off[k + 1])
        cmass[sl] = np.maximum(cmass[sl], mass_floor * cmass[sl].mean())
    stopped = np.zeros(nc, np.int32)
    if pin_mask is not None:
        pw = np.zeros(nc)
        np.add.at(pw, p_idx.ravel(), (p_w * np.asarray(pin_mask, np.float64)[:, None]).ravel())
        stopped = ((wsum > 0) & (pw >= pin_frac * wsum)).astype(np.int32)
    return CoarseLevel(pos0=cpos.astype(np.float32), tets=ctets, mass=cmass.astype(np.float32),
                       stopped=stopped, p_idx=p_idx.astype(np.int32), p_w=p_w.astype(np.float32),
                       blk_off=np.asarray(blk_off, np.int32),
                       minv=(np.concatenate(minv) if minv else np.zeros(0)).astype(np.float32),
                       eps=np.asarray(eps_l, np.float32), fit_off=np.asarray(fit_off, np.int32),
                       fit_fine=(np.concatenate(fit_fine) if fit_fine else np.zeros(0)).astype(np.int32),
                       fit_center=np.asarray(fit_c, np.float32).reshape(-1, 3),
                       fit_sinv=np.asarray(fit_s, np.float32).reshape(-1, 4, 4),
                       iters=int(iters), stiffness=float(stiffness), spacing=np.array(spacing),
                       n_components=len(spacing))


# 6-tet split of a unit cube (corner b at offset (b & 1, b >> 1 & 1, b >> 2 & 1)), the
# split of the reference's voxel cage (tools/gen_coarse_cage.py CUBE)
_VOXEL_TETS = np.array([(0, 1, 3, 7), (0, 3, 2, 7), (0, 2, 6, 7), (0, 6, 4, 7), (0, 4, 5, 7), (0, 5, 1, 7)],
                       np.int64)
_VOXEL_CORNER = np.array([(b & 1, (b >> 1) & 1, (b >> 2) & 1) for b in range(8)], np.int64)


def _voxel_barycentric(u):
    """(tet index within the voxel, 4 barycentric weights) of local points u in
    [0, 1]^3: the voxel tet with the largest smallest weight (the containing one)."""
    best_t = np.zeros(len(u), np.int64)
    best_w = np.zeros((len(u), 4))
    best_s = np.full(len(u), -np.inf)
    for t, tet in enumerate(_VOXEL_TETS):
        p = _VOXEL_CORNER[tet].astype(np.float64)
        T = np.stack([p[0] - p[3], p[1] - p[3], p[2] - p[3]], axis=1)
        lam = np.linalg.solve(T, (u - p[3]).T).T
        w = np.concatenate([lam, 1.0 - lam.sum(1, keepdims=True)], axis=1)
        m = w.min(1)
        upd = m > best_s
        best_s[upd], best_t[upd], best_w[upd] = m[upd], t, w[upd]
    return best_t, best_w


def build_voxel_level(pos0, tets, total_mass: float, pin_mask=None, voxel: float | None = None,
                      factor: float = 4.0, iters: int = 15, stiffness: float = 1.0e6, pin_frac: float = 0.5,
                      comp=None, skip=None, vgroup=None) -> CoarseLevel:
    """The reference's two-level cage (tools/gen_coarse_cage.py --method voxel,
    MultiGridSim._init_coarse, mg.py:744-866): per muscle, the voxels (side
    ``voxel``, default ``factor`` x the fine lattice spacing) that hold a fine
    vertex, each split into 6 tets; a fine vertex interpolates from the voxel
    tet that contains it (barycentric weights in [0, 1]); the restriction is the
    Galerkin least-squares projection pinv(P^T P) P^T per muscle (rcond 1e-8, so
    a corner that no fine vertex references restricts to the origin, as in the
    reference); the coarse rest is the restriction of the fine rest (FAS); the
    coarse masses spread ``total_mass`` over the cage by tet volume
    (_node_mass_redistribute, mg.py:56-68); a node is a pin when the pinned
    share of its |weight| is at least ``pin_frac`` (mg.py:835-853).  The fit
    tables are those of the lattice level; with eps = 0 they do not enter the
    restriction."""
    pos0 = np.asarray(pos0, np.float64).reshape(-1, 3)
    tets = np.asarray(tets, np.int64).reshape(-1, 4)
    nv = pos0.shape[0]
    comp = components(nv, tets) if comp is None else np.asarray(comp, np.int64)
    skip = np.zeros(nv, bool) if skip is None else np.asarray(skip, bool)
    vols = np.abs(np.einsum("ij,ij->i", np.cross(pos0[tets[:, 1]] - pos0[tets[:, 0]],
                                                 pos0[tets[:, 2]] - pos0[tets[:, 0]]),
                            pos0[tets[:, 3]] - pos0[tets[:, 0]])) / 6.0
    tcomp = comp[tets[:, 0]]
    p_idx = np.zeros((nv, 8), np.int64)
    p_w = np.zeros((nv, 8), np.float64)
    cage_pos, all_tets, q0_all, spacing, blk_off, minv = [], [], [], [], [0], []
    fit_off, fit_fine, fit_c, fit_s, bgroup = [0], [], [], [], []
    base = 0
    for c in (np.unique(comp[~skip]) if nv else []):
        vids = np.nonzero((comp == c) & ~skip)[0]
        tv = vols[tcomp == c]
        h = voxel if voxel else factor * ((6.0 * tv.mean()) ** (1.0 / 3.0) if tv.size else 1.0)
        x = pos0[vids]
        mn = x.min(0)
        rel = (x - mn) / h
        ijk = np.floor(rel).astype(np.int64)
        occ, vox_of = np.unique(ijk, axis=0, return_inverse=True)
        vox_of = vox_of.reshape(-1)
        corners = (occ[:, None, :] + _VOXEL_CORNER[None]).reshape(-1, 3)
        nodes, node_of = np.unique(corners, axis=0, return_inverse=True)
        node_of = node_of.reshape(len(occ), 8)
        n = len(nodes)
        if n > MAX_BLOCK_NODES or 6 * len(occ) > MAX_BLOCK_TETS:
            raise ValueError(f"voxel cage of component {c} has {n} nodes / {6 * len(occ)} tets; the coarse "
                             f"solve takes at most {MAX_BLOCK_NODES} nodes and {MAX_BLOCK_TETS} tets per block "
                             f"(use a larger voxel)")
        ctet = node_of[:, _VOXEL_TETS].reshape(-1, 4)                # voxel-major, 6 per voxel
        tloc, w4 = _voxel_barycentric(rel - ijk)
        w4 = np.clip(w4, 0.0, 1.0)                                   # rounding at a tet face
        cn = ctet[6 * vox_of + tloc]                                  # (nv_b, 4) block-local nodes
        p_idx[vids, :4] = base + cn
        p_idx[vids, 4:] = (base + cn[:, :1])        # unused slots: weight 0, same block
        p_w[vids, :4] = w4
        P = np.zeros((len(vids), n))
        np.add.at(P, (np.repeat(np.arange(len(vids)), 4), cn.ravel()), w4.ravel())
        M = np.linalg.pinv(P.T @ P, rcond=1e-8)
        minv.append(M.ravel())
        cage_pos.append(mn + nodes * h)
        q0_all.append(M @ (P.T @ x))                                  # FAS rest of the block
        all_tets.append(base + ctet)
        ctr = x.mean(0)
        xh = np.concatenate([x - ctr, np.ones((len(vids), 1))], 1)
        fit_s.append(np.linalg.inv(xh.T @ xh))
        fit_c.append(ctr)
        fit_fine.append(vids)
        fit_off.append(fit_off[-1] + len(vids))
        bgroup.append(int(vgroup[vids[0]]) if vgroup is not None else -1)
        spacing.append(h)
        base += n
        blk_off.append(base)
    nc = base
    cpos = np.concatenate(cage_pos) if cage_pos else np.zeros((0, 3))
    q0 = np.concatenate(q0_all) if q0_all else np.zeros((0, 3))
    ctets = np.concatenate(all_tets).astype(np.int32) if all_tets else np.zeros((0, 4), np.int32)
    # masses: total_mass spread by cage tet volume (the baked cage rest, as the reference)
    r = cpos[ctets]
    vol = np.abs(np.linalg.det(np.stack([r[:, 1] - r[:, 0], r[:, 2] - r[:, 0], r[:, 3] - r[:, 0]], axis=-1))) / 6.0
    cmass = np.zeros(nc)
    np.add.at(cmass, ctets.reshape(-1), np.repeat(vol / 4.0, 4))
    if cmass.sum() > 0:
        cmass *= float(total_mass) / cmass.sum()
    stopped = np.zeros(nc, np.int32)
    if pin_mask is not None:
        wacc = np.zeros(nc)
        np.add.at(wacc, p_idx.ravel(), np.abs(p_w).ravel())
        pacc = np.zeros(nc)
        np.add.at(pacc, p_idx.ravel(), (np.abs(p_w) * np.asarray(pin_mask, np.float64)[:, None]).ravel())
        with np.errstate(divide="ignore", invalid="ignore"):
            frac = np.where(wacc > 1e-9, pacc / wacc, 0.0)
        stopped = ((frac >= pin_frac) & (wacc > 1e-9)).astype(np.int32)
    return CoarseLevel(pos0=q0.astype(np.float32), tets=ctets, mass=cmass.astype(np.float32), stopped=stopped,
                       p_idx=p_idx.astype(np.int32), p_w=p_w.astype(np.float32),
                       blk_off=np.asarray(blk_off, np.int32),
                       minv=(np.concatenate(minv) if minv else np.zeros(0)).astype(np.float32),
                       eps=np.zeros(len(spacing), np.float32), fit_off=np.asarray(fit_off, np.int32),
                       fit_fine=(np.concatenate(fit_fine) if fit_fine else np.zeros(0)).astype(np.int32),
                       fit_center=np.asarray(fit_c, np.float32).reshape(-1, 3),
                       fit_sinv=np.asarray(fit_s, np.float32).reshape(-1, 4, 4),
                       iters=int(iters), stiffness=float(stiffness), spacing=np.array(spacing),
                       n_components=len(spacing), pin_rule=1, blk_vgroup=np.asarray(bgroup, np.int32))


def coarse_pin_targets(level: CoarseLevel, w, frame: int) -> np.ndarray:
    """Restricted bone goal of every coarse node (_refresh_coarse_target,
    mg.py:1013-1030): each muscle's fine goals follow its bone rigidly, the
    restriction is linear and reproduces constants, so the restricted goal is
    the coarse rest moved by that bone."""
    from . import workloads
    q0 = level.pos0.astype(np.float64)
    out = q0.copy()
    for b in range(len(level.blk_off) - 1):
        g = int(level.blk_vgroup[b]) if level.blk_vgroup is not None else -1
        if 0 <= g < len(w.bones):
            s, e = int(level.blk_off[b]), int(level.blk_off[b + 1])
            R, t = workloads.bone_transform(w.bones[g], frame)
            out[s:e] = q0[s:e] @ R.T + t
    return out.astype(np.float32)


def restrict_numpy(level: CoarseLevel, x, rest) -> np.ndarray:
    """q = (P^T P + eps I)^-1 (P^T x + eps q_ref) per block, float64 (the oracle's
    restriction); ``rest`` are the fine rest positions."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    rest = np.asarray(rest, np.float64).reshape(-1, 3)
    nc = level.pos0.shape[0]
    b = np.zeros((nc, 3))
    for a in range(3):
        np.add.at(b[:, a], level.p_idx.ravel(), (level.p_w.astype(np.float64) * x[:, a][:, None]).ravel())
    q = np.zeros((nc, 3))
    o = 0
    for k in range(len(level.blk_off) - 1):
        s, e = int(level.blk_off[k]), int(level.blk_off[k + 1])
        n = e - s
        vids = level.fit_fine[level.fit_off[k]:level.fit_off[k + 1]]
        ctr = level.fit_center[k].astype(np.float64)
        xh = np.concatenate([rest[vids] - ctr, np.ones((len(vids), 1))], 1)
        T = (x[vids].T @ xh) @ level.fit_sinv[k].astype(np.float64)          # [A t], (3, 4)
        qr = (level.pos0[s:e].astype(np.float64) - ctr) @ T[:, :3].T + T[:, 3]
        M = level.minv[o:o + n * n].astype(np.float64).reshape(n, n)
        q[s:e] = M @ (b[s:e] + float(level.eps[k]) * qr)
        o += n * n
    return q


def coarse_scene(level: CoarseLevel, params):
    """Scene of the coarse cage: ARAP-SVD only, no gravity, masses from P.

    Coarse pins follow the fine level's pin rule (MultiGridSim keeps its pinned
    vertices in the Jacobi sums with their inverse mass and snaps them back
    after every iteration, mg.py:1337,1355): a pinned node takes part in every
    projection and keeps its restricted position.  Held with zero inverse mass
    instead (``coarse_stopped``, mg.py:860-866), a tet whose other corners are
    held moves one corner along the gradient of C = ||sigma - 1||; that step
    overshoots the minimum of C along its line by 1/cos^2 of the angle between
    F - R and the one-corner direction, and the sweeps amplify a perturbation
    20-60x each."""
    from .scene import Scene, SimParams
    cons = builder.build_from_configs(level.pos0, level.tets,
                                      [dict(type="tetarapsvd", name="arap", stiffness=level.stiffness,
                                            dampingratio=0.0)])
    p = SimParams(dt=params.dt, gravity=(0.0, 0.0, 0.0), veldamping=0.0, quasi_static=False,
                  contraction_ratio=0.0, fiber_stiffness_scale=1.0)
    pins = np.nonzero(level.stopped)[0].astype(np.int32)
    held = getattr(level, "pin_rule", 0) == 1      # the reference's rule: pins stopped at their goal
    return Scene(pos0=level.pos0, tets=level.tets, density=1.0, cons=cons, params=p, mass=level.mass,
                 stopped=level.stopped if held else None, pin_vids=pins if pins.size else None,
                 pin_active=np.ones(pins.size, np.int32) if pins.size else None,
                 meta=dict(workload="coarse cage"))
