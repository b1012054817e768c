"""Sim programs on top of the HIP XpbdEngine: MultiGridSim and MuscleSim.

``MultiGridSim`` follows the JSON-driven pin-to-target program of the
reference (src/muscle/mg/mg.py): config semantics (:95-400), constraint set
(:379-399), density from total mass (:736-740), pins (:486-535), substep order
(``build_program`` :1322-1378) and ``step_frame`` (:1409-1418).  Meshes come
from a ``synthetic`` config block (the reference's .geo/USD assets are not
available) or from arrays passed in.

``MuscleSim.from_procedural`` follows src/muscle/musclesim.py:181-287: one
constraint iteration per substep, pins as stopped vertices.
"""
from __future__ import annotations

import numpy as np

from . import builder, capi, collision, multigrid, workloads
from .builder import ConsTable
from .capi import XpbdError
from .config import find_constraint
from .engine import XpbdEngine
from .scene import BoneMeshCollider, Scene, SdfColliders, SimParams


def mg_constraint_configs(cfg: dict) -> list:
    """The fine-level constraint set of MultiGridSim (mg.py:103-131, 379-399)."""
    arap = find_constraint(cfg, "arap")
    if arap is None:
        raise ValueError("constraints[] must define an 'arap' entry (the muscle ARAP-SVD shape match)")
    mu = float(arap["stiffness"])
    arap_damp = float(arap.get("dampingratio", 0.0))
    snh = find_constraint(cfg, "snh")
    vol = find_constraint(cfg, "volume")
    fib = find_constraint(cfg, "fiber")
    out = []
    if snh is not None:
        out.append(dict(type="snh", name="snh", mu=float(snh.get("mu", snh.get("stiffness", mu))),
                        lam=float(snh.get("lam", 1.0e8)),
                        dampingratio=float(snh.get("dampingratio", arap_damp))))
    else:
        out.append(dict(type="tetarapsvd", name="arap", stiffness=mu, dampingratio=arap_damp))
    if vol is not None and snh is None:
        out.append(dict(type="volume", name="volume", stiffness=float(vol["stiffness"]),
                        dampingratio=float(vol.get("dampingratio", 0.0))))
    if fib is not None:
        model = str(fib.get("fiber_model", "aniso_arap"))
        if model != "aniso_arap":
            raise NotImplementedError(f"fiber_model={model!r}: only aniso_arap is on the HIP path")
        out.append(dict(type="aniso_arap", name="fiber", stiffness=float(fib.get("stiffness", 1.0e8)),
                        dampingratio=float(fib.get("dampingratio", 0.0))))
    return out


def fascia_scene(cfg: dict, pos, tets, dt: float, ccfg=None) -> Scene:
    """The fascia shell (Fascia.__init__, fascia.py:245-330): in ride mode Stable
    Neo-Hookean from the 'fascia_snh' entry (or the given constraint set), density from
    fascia_total_mass (else the muscle's total_mass) over the shell's rest volume,
    quasi-static, no gravity."""
    if ccfg is None:
        snh = find_constraint(cfg, "fascia_snh") or {}
        ccfg = [dict(type="snh", name="fascia_snh", mu=float(snh.get("mu", 1.0e3)),
                     lam=float(snh.get("lam", 1.0e8)), dampingratio=float(snh.get("dampingratio", 0.0)))]
    pos = np.asarray(pos, np.float32)
    tets = np.asarray(tets, np.int32)
    cons = builder.build_from_configs(pos, tets, ccfg)
    volumes, valid = builder.tet_rest_data(pos, tets)
    total = float(cfg.get("fascia_total_mass", cfg.get("total_mass", 1.0)))
    density = total / float(max(volumes[valid].sum(), 1e-30))
    params = SimParams(dt=dt, gravity=(0.0, 0.0, 0.0), quasi_static=True)
    return Scene(pos0=pos, tets=tets, density=density, cons=cons, params=params,
                 meta=dict(workload="fascia", constraint_configs=ccfg))


def fascia_scene_squeeze(cfg: dict, pos, tets, dt: float) -> Scene:
    """The fascia shell in squeeze mode (Fascia.__init__ and _apply_rest_shrink,
    fascia.py:245-352): ARAP-SVD ('fascia_arap': stiffness, dampingratio, rest_scale)
    plus volume ('fascia_volume'), the ARAP rest shape scaled by rest_scale."""
    arap = find_constraint(cfg, "fascia_arap") or {}
    vol = find_constraint(cfg, "fascia_volume") or {}
    k = float(arap.get("stiffness", -1.0))
    if k < 0:
        raise ValueError("fascia_arap.stiffness < 0 asks for the fascia geometry's solidshapestiffness attribute, "
                         "which a synthetic shell does not carry: give the stiffness")
    ccfg = [dict(type="tetarapsvd", name="fascia_arap", stiffness=k, dampingratio=float(arap.get("dampingratio", 0.0))),
            dict(type="volume", name="fascia_volume", stiffness=float(vol.get("stiffness", 1.0e10)),
                 dampingratio=float(vol.get("dampingratio", 0.0)))]
    sc = fascia_scene(cfg, pos, tets, dt, ccfg)
    sc.rest_scale = float(arap.get("rest_scale", 0.95))
    return sc


class FasciaShell:
    """Fascia sub-solver (Fascia, fascia.py:217-572): its own engine on its own mesh;
    its substep runs inside the muscle's.  "ride": bound to the muscle surface, Stable
    Neo-Hookean; "squeeze": follows the skeleton (advected per frame, anchors snapped per
    substep), ARAP + volume with a shrunk rest shape, glued one way to the muscle surface."""

    def __init__(self, parent, engine_cls, device, tuning):
        cfg = parent.cfg
        self.mode = str(cfg.get("fascia_mode", "squeeze"))
        if self.mode not in ("ride", "squeeze"):
            raise ValueError(f"fascia_mode={self.mode!r}: 'ride' or 'squeeze'")
        w = parent.workload
        if w is None or "fascia_mesh" not in w.info:
            raise ValueError("use_fascia needs a fascia mesh on the workload")
        fpos, ftets = w.info["fascia_mesh"]
        self.scene = (fascia_scene(cfg, fpos, ftets, parent.dt) if self.mode == "ride"
                      else fascia_scene_squeeze(cfg, fpos, ftets, parent.dt))
        self.NV, self.NT = self.scene.nv, self.scene.nt
        self.engine = (self._shell_engine(device, tuning) if engine_cls is XpbdEngine
                       else engine_cls(self.scene))
        glue = find_constraint(cfg, "fascia_glue") or {}
        self.iters = int(glue.get("iterations", parent.fine_iters))
        self.surf_tris = collision.surface_tris(parent.scene.tets)
        self.bind_max_dist = float(cfg.get("fascia_bind_max_dist", 0.05))
        if self.mode == "ride":
            link = dict(mode=0, iters=self.iters, surf_tris=self.surf_tris, bind_max_dist=self.bind_max_dist)
        else:
            link = self._squeeze_link(parent, np.asarray(fpos, np.float32), glue)
        parent.engine.set_fascia(self.engine, link)

    def _shell_engine(self, device, tuning):
        """The shell with the parent's launch shape; with 256-thread blocks its plain
        iterations run resident (every iteration of a shell substep in one launch,
        the smallest parts that fit) unless ``resident_blocks`` was given: a shell
        is small, and small meshes gain most from one launch per substep."""
        t = dict(tuning)
        if (t.get("strategy") == 2 and t.get("block_tets") == t.get("block_threads") and
                int(t.get("block_threads", 0)) <= 256 and not t.get("resident_blocks")):
            try:
                return XpbdEngine(self.scene, device=device, **dict(t, resident_blocks=-1))
            except XpbdError as e:
                if e.code != capi.XPBD_ELAYOUT:
                    raise
        return XpbdEngine(self.scene, device=device, **t)

    def _squeeze_link(self, parent, fpos, glue):
        """_build_coupling / _pair_surface (fascia.py:354-441): every muscle surface vertex
        glued to its nearest fascia vertex at their rest distance; every fascia vertex
        follows the bone-driven position of its nearest muscle vertex; anchors: a uniform
        subsample of fascia_anchor.count vertices (no 'rigid' group on a synthetic shell)."""
        from scipy.spatial import cKDTree
        cfg = parent.cfg
        rest_m = np.asarray(parent.scene.pos0, np.float32)
        surf = np.unique(self.surf_tris).astype(np.int64)
        d, floc = cKDTree(fpos).query(rest_m[surf], k=1)
        if glue.get("stiffness") is None:
            raise ValueError("fascia_glue.stiffness is needed: a synthetic shell has no tissuetomusclestiffness "
                             "attribute")
        stiff = np.full(len(surf), float(glue["stiffness"]) * float(glue.get("stiffness_scale", 0.001)), np.float32)
        _, fm = cKDTree(rest_m).query(fpos, k=1)
        self.follow_vids = fm.astype(np.int32)
        follow0 = rest_m[self.follow_vids]
        n_anc = max(1, int((find_constraint(cfg, "fascia_anchor") or {}).get("count", 256)))
        rid = np.arange(0, len(fpos), max(1, len(fpos) // n_anc), dtype=np.int32)
        self.anchor = rid
        return dict(mode=1, iters=self.iters, follow0=follow0, glue_fascia=floc.astype(np.int32),
                    glue_muscle=surf.astype(np.int32), glue_restlen=d.astype(np.float32),
                    glue_compliance=(1.0 / np.maximum(stiff, 1e-30)).astype(np.float32),
                    omega=float(cfg.get("omega", 1.0)), anchor=rid,
                    anchor_offset=(fpos[rid] - follow0[rid]).astype(np.float32))

    def follow(self, w, frame):
        """squeeze: the bone-driven positions of the followed muscle vertices at frame"""
        return workloads.follow_targets(w, frame, self.follow_vids)

    def positions_np(self):
        return self.engine.positions_np()


def multigrid_volume(scene) -> float:
    volumes, valid = builder.tet_rest_data(scene.pos0, scene.tets)
    return float(np.abs(volumes[valid]).sum())


def build_workload(cfg: dict, base=None) -> workloads.Workload:
    """A synthetic workload from the config's ``synthetic`` block, or the
    reference's asset-driven inputs (``rest_path`` / ``geo_dir``, ``cons_dir``,
    ``bone_dir``; ``assets.from_assets``, paths relative to ``base``)."""
    syn = cfg.get("synthetic")
    if not syn and ("rest_path" in cfg or "geo_dir" in cfg):
        from . import assets
        return assets.from_assets(cfg, base)
    if not syn:
        raise ValueError("config has no 'synthetic' block, no rest_path / geo_dir and no mesh arrays were given")
    name = syn["workload"]
    seed = int(syn.get("seed", 0))
    if name == "cantilever":
        return workloads.cantilever(seed)
    if name == "humanv2_scale":
        w = workloads.humanv2_scale(seed, target_tets=int(syn.get("muscle_tets", 468_000)),
                                    fascia_tets=int(syn.get("fascia_tets", 67_000)),
                                    sdf_res=int(syn.get("sdf_res", 128)))
        if cfg.get("use_fascia", False):
            # the fascia tubes become the fascia sub-solver's own mesh
            w, fpos, ftets = workloads.split_fascia(w)
            w.info["fascia_mesh"] = (fpos, ftets)
        return w
    if name == "batch":
        return workloads.batch(seed, n_inst=int(syn.get("instances", 512)), lo=int(syn.get("min_tets", 2000)),
                               hi=int(syn.get("max_tets", 40000)), sdf_res=int(syn.get("sdf_res", 64)))
    raise ValueError(f"unknown synthetic workload {name!r}")


def _mesh_res(coll: dict) -> dict:
    """tessellation of the synthetic capsule bones (around, per cap, along)"""
    r = coll.get("mesh_res", [96, 16, 96])
    return dict(n_around=int(r[0]), n_cap=int(r[1]), n_len=int(r[2]))


def scene_from_config(cfg: dict, w: workloads.Workload) -> Scene:
    """Scene for the MultiGridSim program on workload ``w``."""
    pos0 = w.pos0.astype(np.float32)
    tets = w.tets.astype(np.int32)
    ccfg = mg_constraint_configs(cfg)
    fib = find_constraint(cfg, "fiber")
    cons = builder.build_from_configs(pos0, tets, ccfg)
    volumes, valid = builder.tet_rest_data(pos0, tets)
    if cfg.get("density", -1.0) and float(cfg.get("density", -1.0)) > 0:
        density = float(cfg["density"])
    else:  # _rest_density (mg.py:736-740)
        density = float(cfg["total_mass"]) / float(max(volumes[valid].sum(), 1e-30))
    dt = float(cfg["delta_t"]) / int(cfg["nsubsteps"])
    params = SimParams(dt=dt, gravity=tuple(float(g) for g in cfg.get("gravity", [0.0, 0.0, 0.0])),
                       veldamping=float(cfg.get("veldamping", 0.0)),
                       quasi_static=bool(cfg.get("quasi_static", False)),
                       contraction_ratio=float((fib or {}).get("contraction_ratio", 0.4)),
                       fiber_stiffness_scale=1.0)
    tm = None
    if w.v_tendonmask is not None:
        tm = builder.tet_tendonmask(tets, w.v_tendonmask)
    vf = w.v_fiber_dir if fib is not None else None
    # kinematic pins (muscleToBonePin); synthetic meshes carry the tendon-tip set
    pin_vids = pin_active = None
    pin = find_constraint(cfg, "muscleToBonePin")
    asset = w.info.get("assets")
    if asset is not None:
        # the constrained vertices of the asset's targets, hard-pinned by the region rule
        # (mg.py:486-536); the others ride the belly attach
        from . import assets
        pin_vids = np.asarray(w.pin_vids, np.int32)
        pin_active = assets.pin_active(pin, w)
    elif pin is not None and w.pin_vids is not None and w.stopped is None:
        pin_vids = w.pin_vids
        region = str(pin.get("region", "tips")).lower()
        pin_active = np.ones(len(pin_vids), np.int32) if region != "none" else np.zeros(len(pin_vids), np.int32)
    # belly attach (mg.py:466-469, 700-734): every muscle vertex that is not an
    # active hard pin gets an ATTACH toward its per-frame bone goal (fascia is a
    # separate sub-solver in the reference and gets none)
    attach_vids = None
    belly = find_constraint(cfg, "belly_attach")
    if belly is not None and asset is not None:
        # the constrained vertices that are not hard-pinned, in constraint order
        # (_build_fine_belly_attach, mg.py:700-734)
        attach_vids = pin_vids[pin_active == 0].astype(np.int32)
        att = builder.create_attach(attach_vids, pos0[attach_vids], float(belly["stiffness"]),
                                    float(belly.get("dampingratio", 0.0)))
        cons = builder.ConsTable.concat([cons, att]).sorted_by_type().contiguous()
    elif belly is not None:
        free = np.ones(w.nv, bool)
        if w.is_fascia is not None:
            free &= ~np.asarray(w.is_fascia, bool)
        if pin_vids is not None:
            free[pin_vids[pin_active != 0]] = False
        if w.stopped is not None:
            free &= np.asarray(w.stopped) == 0
        if w.vertex_group is None and w.bones:
            raise ValueError("belly_attach needs a bone per vertex (workload vertex_group)")
        attach_vids = np.nonzero(free)[0].astype(np.int32)
        att = builder.create_attach(attach_vids, pos0[attach_vids], float(belly["stiffness"]),
                                    float(belly.get("dampingratio", 0.0)))
        cons = builder.ConsTable.concat([cons, att]).sorted_by_type().contiguous()
    sdf = bone_mesh = None
    coll = cfg.get("demo", {}).get("collision", {})
    kind = str(coll.get("collider", "grid")).lower()
    if kind not in ("grid", "mesh"):
        raise ValueError(f"demo.collision.collider={kind!r}: 'grid' (bone SDF grids) or 'mesh' (bone triangle mesh)")
    bsrc = asset.get("bone_source") if asset is not None else None
    if bsrc is not None and (coll.get("enabled", False) or kind == "mesh"):
        # the asset's skeleton (.geo bone sequence) as the exact-mesh collider: Houdini winding,
        # flipped to outward normals as AnimatedBoneMeshCollider does (humanV2.py:319-384)
        m = np.ones(w.nv, np.float32)
        bone_mesh = BoneMeshCollider(verts=np.asarray(bsrc.positions(asset["start_frame"]), np.float32),
                                     tris=np.asarray(bsrc.tris, np.int32).reshape(-1, 3)[:, [0, 2, 1]],
                                     tri_off=None, max_dist=float(coll.get("max_dist", 0.2)),
                                     contact_offset=float(coll.get("contact_offset", 1e-5)),
                                     max_push=float(coll.get("max_push", -1.0)), coll_w=m, meter_w=m.copy(),
                                     vertex_group=None, group_bone_off=None,
                                     resolve=bool(coll.get("enabled", False)), coherent=int(coll.get("coherent", 2)))
    elif w.bones and (coll.get("enabled", False) or kind == "mesh"):
        m = np.ones(w.nv, np.float32)
        if w.is_fascia is not None:
            m[w.is_fascia] = 0.0   # fascia excluded from bone collision (humanV2.py:361-366)
        vg = goff = None
        if w.vertex_group is not None and w.name == "batch":
            vg = w.vertex_group.astype(np.int32)
            goff = np.arange(len(w.bones) + 1, dtype=np.int32)
        if kind == "grid":
            grids, origins, spacing = workloads.bake_bone_sdfs(w, int(cfg.get("synthetic", {}).get("sdf_res", 64)))
            sdf = SdfColliders(grids=grids, origins=origins, spacing=spacing,
                               contact_offset=float(coll.get("contact_offset", 1e-5)),
                               max_push=float(coll.get("max_push", -1.0)), coll_w=m,
                               vertex_group=vg, group_bone_off=goff)
        else:
            # AnimatedBoneMeshCollider (humanV2.py:319-384): the skeleton as one triangle mesh,
            # pushing when collision is enabled, the penetration meter always
            verts, tris, toff, _ = workloads.bone_meshes(w, **_mesh_res(coll))
            bone_mesh = BoneMeshCollider(verts=workloads.bone_mesh_points(w, int(cfg.get("start_frame", 1)),
                                                                          **_mesh_res(coll)),
                                         tris=tris, tri_off=toff, max_dist=float(coll.get("max_dist", 0.2)),
                                         contact_offset=float(coll.get("contact_offset", 1e-5)),
                                         max_push=float(coll.get("max_push", -1.0)), coll_w=m, meter_w=m.copy(),
                                         vertex_group=vg, group_bone_off=goff,
                                         resolve=bool(coll.get("enabled", False)),
                                         coherent=int(coll.get("coherent", 2)))
    dcd = None
    sc = cfg.get("demo", {}).get("self_collision", {})
    if sc.get("enabled", False):
        method = str(sc.get("method", "dcd")).lower()
        if method != "dcd":
            raise NotImplementedError(f"self_collision.method={method!r}: only the DCD pass is on the HIP path")
        # surface triangles of the selected tets (humanV2.py:820-866 passes sim.surf_flat)
        which = str(sc.get("surfaces", "all")).lower()
        sel = np.ones(len(tets), bool)
        if which in ("fascia", "muscle"):
            if w.is_fascia is None:
                raise ValueError(f"self_collision.surfaces={which!r} needs a fascia mask on the workload")
            fas = w.is_fascia[tets].all(1)
            sel = fas if which == "fascia" else ~fas
        p = collision.DcdParams(thickness=float(sc.get("thickness", 1e-3)), multiplier=float(sc.get("multiplier", 0.3)),
                                friction_static=float(sc.get("mu_s", 0.5)), friction_dynamic=float(sc.get("mu_d", 0.4)),
                                normal_cull=float(sc.get("normal_cull", 1.0)), cell_size=float(sc.get("cell_size", 0.0)),
                                iters=int(sc.get("iters", 2)), rebuild_substeps=bool(sc.get("rebuild_substeps", False)),
                                max_pairs=int(sc.get("max_pairs", 1 << 22)))
        dcd = collision.DcdSpec(tris=collision.surface_tris(tets[sel]), group=None, params=p)
    if attach_vids is not None:
        ccfg = ccfg + [dict(type="attach", name="belly_attach", stiffness=float(belly["stiffness"]),
                            dampingratio=float(belly.get("dampingratio", 0.0)))]
    return Scene(pos0=pos0, tets=tets, density=density, cons=cons, params=params, v_fiber_dir=vf,
                 tendonmask=tm, stopped=w.stopped, pin_vids=pin_vids, pin_active=pin_active, sdf=sdf,
                 bone_mesh=bone_mesh,
                 dcd=dcd, attach_vids=attach_vids, meta=dict(workload=w.name, constraint_configs=ccfg))


class MultiGridSim:
    """Pin-to-target muscle sim on the HIP substep (mg.py:85-1434, single level)."""

    def __init__(self, cfg: dict, workload: workloads.Workload | None = None, device: int = 0,
                 patch_verts: int = 256, engine_cls=XpbdEngine, strategy: int = 0, block_tets: int = 1024,
                 block_threads: int = 256, resident_blocks: int = 0):
        if not cfg.get("use_jacobi", True):
            raise NotImplementedError("use_jacobi=false: Gauss-Seidel is not on the GPU path")
        self.cfg = cfg
        self.workload = workload if workload is not None else build_workload(cfg)
        self.scene = scene_from_config(cfg, self.workload)
        self.nsubsteps = int(cfg["nsubsteps"])
        self.fine_iters = int(cfg["fine_iterations"])
        self.dt = self.scene.params.dt
        self.start_frame = int(cfg.get("start_frame", 1))
        self.engine = engine_cls(self.scene, device=device, patch_verts=patch_verts, strategy=strategy,
                                 block_tets=block_tets, block_threads=block_threads,
                                 resident_blocks=resident_blocks) \
            if engine_cls is XpbdEngine else engine_cls(self.scene)
        self.NV = self.scene.nv
        self.NT = self.scene.nt
        self.coarse = None
        if cfg.get("use_multigrid", False):
            self._build_coarse(engine_cls, device)
        self.fascia = None
        if cfg.get("use_fascia", False):
            self.fascia = FasciaShell(self, engine_cls, device,
                                      dict(strategy=strategy, patch_verts=patch_verts, block_tets=block_tets,
                                           block_threads=block_threads, resident_blocks=resident_blocks))
        self.use_graph = True
        self._native_collider = None
        self.collision_solvers = []
        self.collision_solvers_iter = []
        self._frame = None

    def _build_coarse(self, engine_cls, device):
        """Two-level hierarchical correction on a factor-x coarser lattice cage
        (use_multigrid with mg_mode "hpbd", mg.py:192-245, 1064-1079)."""
        cfg = self.cfg
        mode = str(cfg.get("mg_mode", "hpbd")).lower()
        if mode != "hpbd":
            raise NotImplementedError(f"mg_mode={mode!r}: only the hpbd V-cycle is on the HIP path")
        w = self.workload
        arap = find_constraint(cfg, "arap")
        pin_mask = None
        if self.scene.pin_vids is not None:
            pin_mask = np.zeros(self.scene.nv, np.float32)
            pin_mask[self.scene.pin_vids] = 1.0
        mass = self.engine.get_rest()[2] if hasattr(self.engine, "get_rest") else None
        comp = None
        if w is not None and w.info.get("muscle_id") is not None:
            comp = w.info["muscle_id"]
        cage = str(cfg.get("mg_cage", "lattice")).lower()
        common = dict(pin_mask=pin_mask, iters=int(cfg.get("mg_coarse_iters", 15)),
                      stiffness=float(arap["stiffness"]), pin_frac=float(cfg.get("mg_coarse_pin_frac", 0.5)),
                      comp=comp, skip=w.is_fascia if w is not None else None)
        if cage == "lattice":
            self.coarse = multigrid.build_coarse_level(self.scene.pos0, self.scene.tets, mass,
                                                       factor=float(cfg.get("mg_coarse_factor", 4.0)), **common)
        elif cage == "voxel":
            # the reference's scheme (tools/gen_coarse_cage.py --method voxel, mg.py:744-1079)
            total = float(cfg["total_mass"]) if float(cfg.get("total_mass", -1.0)) > 0 else float(
                self.scene.density * multigrid_volume(self.scene))
            vox = float(cfg.get("mg_coarse_voxel", 0.0)) or None
            self.coarse = multigrid.build_voxel_level(
                self.scene.pos0, self.scene.tets, total, voxel=vox, factor=float(cfg.get("mg_coarse_factor", 4.0)),
                vgroup=w.vertex_group if w is not None else None, **common)
        else:
            raise NotImplementedError(f"mg_cage={cage!r}: 'lattice' or 'voxel'")
        cs = multigrid.coarse_scene(self.coarse, self.scene.params)
        self.coarse_engine = (engine_cls(cs, device=device, strategy=1) if engine_cls is XpbdEngine
                              else engine_cls(cs))
        # the reference's dynamic form by default: the prolonged shift enters the
        # velocity (prolong_delta_kernel, mg/kernel.py:9-40); mg_kinematic_correction
        # also shifts the previous positions, so the shift does not
        self.engine.set_coarse_level(self.coarse_engine, self.coarse,
                                     kinematic=bool(cfg.get("mg_kinematic_correction", False)))

    # kinematic drive ---------------------------------------------------------
    def load_frame_targets(self, frame: int) -> None:
        """Per-frame bone pose -> pin targets, belly-attach targets and SDF
        transforms (mg.py:1233-1249)."""
        w = self.workload
        asset = w.info.get("assets")
        if asset is not None:
            # the asset's per-frame goals: cons_<frame>.geo target_pos + rest_offset, or the
            # bone-surface binding riding bone_<frame>.geo (mg.py:1233-1249)
            goals = asset["targets"].goals(frame)
            if self.scene.pin_vids is not None:
                self.engine.set_pin_targets(goals)
            if self.scene.attach_vids is not None:
                self.engine.set_attach_targets(goals[self.scene.pin_active == 0])
            if self.scene.bone_mesh is not None:
                self.engine.set_bone_mesh_points(asset["bone_source"].positions(frame))
            self._frame = frame
            return
        if self.scene.pin_vids is not None and w.bones:
            self.engine.set_pin_targets(workloads.pin_targets(w, frame))
        if self.scene.attach_vids is not None and w.bones:
            self.engine.set_attach_targets(workloads.follow_targets(w, frame, self.scene.attach_vids))
        if self.scene.sdf is not None:
            self.engine.set_bone_transforms(workloads.bone_xforms(w, frame))
        if self.scene.bone_mesh is not None and w.bones:
            coll = self.cfg.get("demo", {}).get("collision", {})
            self.engine.set_bone_mesh_points(workloads.bone_mesh_points(w, frame, **_mesh_res(coll)))
        if self.coarse is not None and self.coarse.pin_rule == 1 and w.bones:
            self.engine.set_coarse_pin_targets(multigrid.coarse_pin_targets(self.coarse, w, frame))
        if self.fascia is not None and self.fascia.mode == "squeeze" and w.bones:
            # Fascia.advect_bone once per frame, before the substeps (mg.py:1413-1416)
            self.engine.set_fascia_follow(self.fascia.follow(w, frame))
        self._frame = frame

    def set_fiber_activation(self, acti) -> None:
        if find_constraint(self.cfg, "fiber") is None:
            raise RuntimeError("set_fiber_activation requires a 'fiber' constraint")
        self.engine.set_fiber_activation(acti)

    def set_activation_groups(self, tet_group) -> None:
        """one activation per group of tets from now on (batched instances)"""
        if find_constraint(self.cfg, "fiber") is None:
            raise RuntimeError("activation groups require a 'fiber' constraint")
        self.engine.set_activation_groups(tet_group)

    def set_group_activation(self, act) -> None:
        self.engine.set_group_activation(act)

    def apply_pintotarget(self) -> None:
        self.engine.apply_pintotarget()

    # collision_solvers: Python post-phase hooks, plus at most one native bone
    # collider (collider.AnimatedBoneMeshCollider), which runs inside the native
    # substep (and its graph) as the last word of the collision stage, as the
    # reference's bone collider does (humanV2.py:616-622)
    @property
    def collision_solvers(self):
        return self._collision_solvers

    @collision_solvers.setter
    def collision_solvers(self, fns):
        fns = list(fns)
        nat = [f for f in fns if getattr(f, "native_stage", False)]
        if len(nat) > 1:
            raise ValueError("at most one native bone collider per sim")
        if nat and nat[0].sim is not self:
            raise ValueError("the bone collider belongs to another sim")
        if nat or self._native_collider is not None:
            self.engine.set_bone_mesh_resolve(bool(nat))
        self._native_collider = nat[0] if nat else None
        self._collision_solvers = fns
        self._py_solvers = [f for f in fns if not getattr(f, "native_stage", False)]

    # program -------------------------------------------------------------------
    def substep(self) -> None:
        """integrate -> pin -> fine_iters x (clear, dispatch, apply_dP, pin) ->
        collision_solvers -> update_velocities (mg.py:1322-1384)."""
        if self._py_solvers or self.collision_solvers_iter:
            e = self.engine
            e.integrate()
            if self.coarse is not None:
                e.vcycle()
            for it in range(self.fine_iters):
                e.iterate(1)
                for fn in self.collision_solvers_iter:
                    fn(it)
            if self.fascia is not None:
                e.fascia_substep()     # Fascia.substep first in post_phase (mg.py:1365-1366)
            for fn in self._py_solvers:
                fn()
            e.collide()            # bone push last (humanV2.py:616-622)
            e.update_velocities()
        else:
            self.engine.native_substep(self.fine_iters)

    def step_frame(self, frame: int) -> None:
        """Kinematic step once, then nsubsteps substeps (mg.py:1409-1418)."""
        self.load_frame_targets(frame)
        self.engine.apply_pintotarget()
        if self._py_solvers or self.collision_solvers_iter:
            for _ in range(self.nsubsteps):
                self.substep()
        else:
            self.engine.native_frame(self.nsubsteps, self.fine_iters, self.use_graph)

    def positions_np(self) -> np.ndarray:
        return self.engine.positions_np()

    def count_inverted(self) -> int:
        return self.engine.count_inverted()

    def residuals(self) -> dict:
        return self.engine.residuals()


def _procedural_attach(v, t, bone, attrs, params):
    """create_attach_constraints (constraint/builder.py:347-398) for a procedural
    mesh: mask -> sources, nearest bone target (scipy cKDTree, as the reference),
    rest length = distance, tetid = the first tet holding the source."""
    import warnings
    from scipy.spatial import cKDTree
    name = params.get("mask_name")
    mask = attrs.get(name) if name is not None else None
    if mask is None:
        warnings.warn(f"mask '{name}' not found in geometry.")
        return None
    src = np.nonzero(np.asarray(mask, np.float32) > float(params.get("mask_threshold", 0.75)))[0].astype(np.int32)
    if src.size == 0:
        warnings.warn(f"No vertices with mask > {params.get('mask_threshold', 0.75)}")
        return None
    if params.get("target_group") is not None or params.get("source_nearest_group") is not None:
        # bone vertex groups come from a loaded bone geometry; a procedural mesh has none
        warnings.warn(f"bone target groups not found: {params.get('target_group') or params.get('source_nearest_group')}")
        return None
    if bone.shape[0] == 0:
        return None
    dist, tgt = cKDTree(bone).query(v[src], k=1)
    first = np.full(v.shape[0], t.shape[0], np.int64)
    np.minimum.at(first, t.ravel(), np.repeat(np.arange(t.shape[0]), 4))
    tetid = np.where(first[src] < t.shape[0], first[src], -1)
    return builder.create_attach(src, bone[tgt], float(params.get("stiffness", 1e10)),
                                 float(params.get("dampingratio", 0.0)), restlength=dist.astype(np.float32),
                                 target_ids=tgt, tetid=tetid)


class MuscleSim:
    """MuscleSim program on the HIP substep: 1 Jacobi iteration per substep."""

    def __init__(self, scene: Scene, num_substeps: int, device: int = 0, engine_cls=None):
        self.scene = scene
        self.num_substeps = int(num_substeps)
        self.engine = XpbdEngine(scene, device=device) if engine_cls is None else engine_cls(scene)
        self.n_verts = scene.nv
        self.dt = scene.params.dt
        self.step_cnt = 0
        self.post_smooth_iters = 0
        self.program = None

    @classmethod
    def from_procedural(cls, vertices, tets, fiber_dirs_per_tet, bone_targets=None, *,
                        constraint_configs, dts, num_substeps=1, device=0, density=1060.0,
                        veldamping=0.02, gravity=0.0, sigma0=0.0, lambda_opt=1.0, contraction_ratio=0.0,
                        fiber_stiffness_scale=1.0, v_max_norm=0.0, d_damp=0.0, stopped=None, engine_cls=None,
                        vertex_attrs=None):
        """musclesim.py:181-287 (gravity is the y component, as the reference engine).

        ``attach`` entries follow create_attach_constraints (constraint/builder.py:347-398):
        the vertices whose ``mask_name`` attribute exceeds ``mask_threshold`` attach to their
        nearest point of ``bone_targets`` (rest length = that distance).  The reference reads
        the mask from the loaded geometry, which a procedural mesh does not have, so it warns
        and adds none; ``vertex_attrs`` (name -> per-vertex array, no reference counterpart)
        supplies such attributes here.  ``sigma0 > 0`` (the explicit active fibre force,
        musclesim.py:400-412) is not on the HIP path and raises; ``lambda_opt``,
        ``v_max_norm`` and ``d_damp`` only act through it."""
        if float(sigma0) > 0.0:
            raise NotImplementedError("sigma0 > 0: the explicit active fibre force (musclesim.py:400-412) "
                                      "is not on the HIP substep path")
        v = np.asarray(vertices, np.float32)
        t = np.asarray(tets, np.int32)
        vf = builder.vertex_fiber_from_tets(v.shape[0], t, fiber_dirs_per_tet)
        kinds = [str(c["type"]).lower() for c in constraint_configs]
        parts = [builder.build_from_configs(v, t, [c for c, k in zip(constraint_configs, kinds) if k != "attach"])]
        bone = np.zeros((0, 3), np.float32) if bone_targets is None else np.asarray(bone_targets, np.float32)
        for c, k in zip(constraint_configs, kinds):
            if k == "attach" and c.get("enable", True):
                parts.append(_procedural_attach(v, t, bone, vertex_attrs or {}, c))
        cons = builder.ConsTable.concat(parts).sorted_by_type().contiguous()
        params = SimParams(dt=float(dts) / num_substeps, gravity=(0.0, float(gravity), 0.0),
                           veldamping=float(veldamping), contraction_ratio=float(contraction_ratio),
                           fiber_stiffness_scale=float(fiber_stiffness_scale))
        scene = Scene(pos0=v, tets=t, density=float(density), cons=cons, params=params,
                      v_fiber_dir=vf, tendonmask=np.zeros(t.shape[0], np.float32),
                      stopped=None if stopped is None else np.asarray(stopped, np.int32))
        sim = cls(scene, num_substeps, device=device, engine_cls=engine_cls)
        sim.bone_pos = bone
        return sim

    # extension hooks (musclesim.py:107-142, engine.py:1039-1067): Python extras run in the
    # solve phase, and a custom program replaces the phase list; either takes the phase-by-phase
    # path (one C-ABI call per phase), the default program runs as one native substep
    @property
    def constraint_solvers(self):
        return self.engine.constraint_solvers

    @constraint_solvers.setter
    def constraint_solvers(self, fns):
        self.engine.constraint_solvers = list(fns)

    def build_program(self):
        """The default program as phase closures (build_program, musclesim.py:107-130)."""
        e = self.engine

        def integrate_phase():
            e.integrate()
            e.apply_pintotarget()
            e.clear()

        def solve_phase():
            e.solve_constraints(run_extras=False)
            e.apply_dP()
            e.solve_extras()

        def finalize_phase():
            e.post_smooth(self.post_smooth_iters)
            e.apply_pintotarget()
            e.update_velocities()

        self.program = [integrate_phase, solve_phase, finalize_phase]
        self._default_program = list(self.program)
        return self.program

    def run_program(self):
        for _ in range(self.num_substeps):
            for phase in self.program:
                phase()

    def step(self):
        """integrate_phase, solve_phase, finalize_phase per substep (musclesim.py:107-142):
        one native substep each, or the phase program when extras or a custom program
        are installed."""
        custom = getattr(self, "program", None) is not None and self.program != getattr(self, "_default_program", None)
        if custom or getattr(self.engine, "constraint_solvers", None):
            if getattr(self, "program", None) is None:
                self.build_program()
            self.run_program()
        else:
            for _ in range(self.num_substeps):
                self.engine.native_substep(1)
                if self.post_smooth_iters:
                    self.engine.post_smooth(self.post_smooth_iters)
        self.step_cnt += 1

    def reset(self):
        p0 = self.scene.pos0
        self.engine.set_state(p0, p0, np.zeros_like(p0))

    def positions_np(self):
        return self.engine.positions_np()
