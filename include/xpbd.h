/*
 * xpbd.h — C ABI of the MI355X-native XPBD muscle / soft-tissue substep.
 *
 * Drop-in for the substep of XpbdEngine under MuscleSim / MultiGridSim in
 * chunleili/rlmuscleV2.  Every entry point is plain C: opaque context, plain
 * pointers and sizes, int32 status (0 = ok, see XPBD_E*), the message of the
 * last failure from xpbd_last_error().  No exceptions cross the boundary.
 *
 * All host arrays are in the caller's vertex / tet / constraint order; the
 * library keeps its own device layout (one fused element per tet holding all
 * of its constraint types, elements in Morton order and cut into blocks of
 * consecutive elements for the block strategy, the autotuner's pick;
 * vertices in Morton order, renumbered block by block for the block
 * strategy) and converts on upload / download.
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef XPBD_H
#define XPBD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XPBD_OK           0
#define XPBD_EINVAL       1   /* bad argument / size / state order            */
#define XPBD_EUNSUPPORTED 2   /* constraint type or mode not on this path      */
#define XPBD_EOOM         3   /* device allocation failed                      */
#define XPBD_EHIP         4   /* HIP runtime error                             */
#define XPBD_ELAYOUT      5   /* mesh does not fit the patch layout limits     */
#define XPBD_ENOTREADY    6   /* xpbd_finalize() not called yet                */

/* constraint type tags — src/muscle/defs.py:20-36 (same integer values) */
#define XPBD_CT_ATTACH       1650556
#define XPBD_CT_PINTOTARGET  771624311
#define XPBD_CT_TETVOLUME    (-215389979)
#define XPBD_CT_TETARAPSVD   (-92199132)
#define XPBD_CT_TETARAPANISO (-92199133)
#define XPBD_CT_TETSNH       (-503462311)

typedef struct xpbd_ctx xpbd_ctx;

/* Engine scalars.  Reference: XpbdEngine.integrate / update_velocities
 * (src/muscle/engine.py:798-829), MuscleSim.__init__ (src/muscle/musclesim.py:72-82),
 * MultiGridSim.__init__ (src/muscle/mg/mg.py:95-187). */
typedef struct {
    float dt;                     /* substep dt = delta_t / nsubsteps            */
    float gravity[3];             /* _gravity_vec, or (0, cfg.gravity, 0)        */
    float veldamping;             /* cfg.veldamping                              */
    int32_t quasi_static;         /* cfg.quasi_static                            */
    float contraction_ratio;      /* aniso_arap active shortening ratio          */
    float fiber_stiffness_scale;  /* aniso_arap stiffness scale                  */
    int32_t has_compress;         /* HAS_compressstiffness                       */
    float exp_damp, const_damp;   /* Gaia velocity damping (1, 0 = off)          */
} xpbd_params;

/* Layout / launch tuning, filled by the autotuner or by hand.  0 = default. */
typedef struct {
    int32_t patch_verts;          /* patch strategy: owned vertices per patch = threads per block (64..256) */
    int32_t strategy;             /* 0 = patch (LDS-staged, one launch per iteration), 1 = flat
                                     (tet pass + vertex pass, two launches per iteration), 2 = block
                                     (tets in blocks reduced in LDS + shared-vertex pass)        */
    int32_t block_tets;           /* block strategy: tets per block (64..8192, default 1024)     */
    int32_t block_threads;        /* block strategy: threads per block (64..512, default 256)    */
    int32_t resident_blocks;      /* block strategy: 0 = one launch per iteration (plus the
                                     shared-vertex pass); R in 1..4 = resident mode, runs of R
                                     blocks form a part and one launch runs all plain iterations
                                     of a substep with each part's vertices in LDS (needs
                                     block_tets == block_threads <= 256 and every part resident
                                     at once, else XPBD_ELAYOUT); -1 = resident with the smallest
                                     R that fits; -2 = resident with as many parts as fit, the
                                     blocks spread evenly over them.  Same positions as 0, bit
                                     for bit.  The parts
                                     wait for each other, so no other long-running kernel may
                                     hold the device during a resident launch; a wait that runs
                                     out sets a status that xpbd_synchronize / xpbd_get_state
                                     return as XPBD_EHIP.  No counterpart in the reference (its
                                     Warp launches are one per constraint type and iteration). */
    int32_t reserved[3];
} xpbd_tuning;

typedef struct {
    int32_t nv, nt, n_elems, n_patches;  /* n_patches: patches, or blocks (block strategy) */
    int32_t n_patch_tets;         /* sum over patches of replicated tets          */
    int32_t max_table;            /* largest patch vertex table (owned + halo)    */
    int32_t max_patch_tets;
    int32_t feature_mask;         /* bit0 arap, bit1 volume, bit2 snh, bit3 aniso, bit4 damping, bit5 attach, bit6 pins */
    double  bytes_per_iter_alg;   /* algorithmic bytes of one Jacobi iteration   */
    double  bytes_per_iter_impl;  /* bytes the projection kernel actually moves  */
    int32_t resident_blocks;      /* resident mode: blocks per part (0 = off)     */
    int32_t resident_parts;       /* resident mode: parts = workgroups per launch  */
    int32_t max_part_verts;       /* resident mode: largest part vertex table      */
    int32_t max_part_neighbours;  /* resident mode: most parts one part waits for  */
    int32_t cross_vertex_entries; /* resident mode: part-local entries of vertices other parts touch too */
    int32_t resident_lds;         /* resident mode: LDS bytes per workgroup        */
} xpbd_stats;

/* ---- lifetime --------------------------------------------------------- */
/* replaces wp.init()/device selection (src/muscle/musclesim.py:44-86) */
int32_t xpbd_create(int32_t device, xpbd_ctx **out);
void    xpbd_destroy(xpbd_ctx *ctx);
const char *xpbd_last_error(const xpbd_ctx *ctx);
/* stream: a hipStream_t (NULL = the context's own stream) */
int32_t xpbd_set_stream(xpbd_ctx *ctx, void *stream);

/* ---- problem definition ---------------------------------------------- */
/* MuscleSim._allocate_fields/_init_fields + precompute_rest_kernel
 * (src/muscle/musclesim.py:331-352,490-543; src/muscle/engine.py:199-224):
 * rest positions and tets; computes B = inv(Dm), |V|, lumped mass with density. */
int32_t xpbd_set_mesh(xpbd_ctx *ctx, int32_t nv, const float *pos0, int32_t nt,
                      const int32_t *tets, float density);
/* read back what xpbd_set_mesh computed (any pointer may be NULL) */
int32_t xpbd_get_rest(xpbd_ctx *ctx, float *rest_volume, float *rest_matrix, float *mass);
/* MuscleSim mass / stopped fields (engine.py:431-464 reads them) */
int32_t xpbd_set_mass(xpbd_ctx *ctx, const float *mass);
int32_t xpbd_set_stopped(xpbd_ctx *ctx, const int32_t *stopped);
int32_t xpbd_set_force(xpbd_ctx *ctx, const float *force);
/* v_fiber_dir (nv*3) and per-tet tendonmask (nt); musclesim.py:490-512 */
int32_t xpbd_set_fiber(xpbd_ctx *ctx, const float *v_fiber_dir, const float *tendonmask);
/* per-tet activation (nt); MultiGridSim.set_fiber_activation (mg.py:1098-1111) */
int32_t xpbd_set_activation(xpbd_ctx *ctx, const float *activation);
/* activation by group, for batched instances driven by one activation each: tet_group (nt ids in
 * [0, ngroups)) once, then per frame xpbd_set_group_activation with ngroups values, expanded to
 * the tets on the device (ngroups floats cross PCIe instead of nt).  Same result as
 * xpbd_set_activation with activation[t] = act[tet_group[t]].  No reference counterpart (the
 * reference sets activation per tet, mg.py:1098-1111). */
int32_t xpbd_set_activation_groups(xpbd_ctx *ctx, int32_t ngroups, const int32_t *tet_group);
int32_t xpbd_set_group_activation(xpbd_ctx *ctx, int32_t ngroups, const float *act);
/* the SoA constraint table of XpbdEngine._upload_cons (engine.py:1179-1211);
 * pts n*4, restvector n*4, restdir n*3, others n. */
int32_t xpbd_upload_constraints(xpbd_ctx *ctx, int32_t n, const int32_t *type,
                                const int32_t *pts, const float *stiffness,
                                const float *compressionstiffness, const float *dampingratio,
                                const int32_t *tetid, const float *restlength,
                                const float *restvector, const float *restdir);
/* ATTACH targets refreshed per frame, in table order of the ATTACH range
 * (MultiGridSim._update_belly_attach_targets, mg.py:725-734); xyz n*3 */
int32_t xpbd_set_attach_targets(xpbd_ctx *ctx, int32_t n, const float *xyz);
int32_t xpbd_set_params(xpbd_ctx *ctx, const xpbd_params *p);
/* kinematic pins: vertex ids + active mask (MultiGridSim pts / pin_active,
 * mg.py:486-535; pintotarget_overwrite_masked_kernel mg/kernel.py:212-226) */
int32_t xpbd_set_pins(xpbd_ctx *ctx, int32_t n, const int32_t *vids, const int32_t *active);
/* per-pin target = target_pos + rest_offset (n*3), per frame (mg.py:1233-1261) */
int32_t xpbd_set_pin_targets(xpbd_ctx *ctx, int32_t n, const float *xyz);
/* bone push-out on trilinear SDF grids; the push rule of
 * solve_mesh_contact_pos_kernel (src/muscle/collision/sdf/kernel.py:123-170).
 * grids: nbones*res^3 f32, x fastest; origins nbones*3 (bone-local); spacing nbones;
 * coll_w: per-vertex collision inverse mass (0 = skip), humanV2.py:361-366 */
int32_t xpbd_set_sdf(xpbd_ctx *ctx, int32_t nbones, int32_t res, const float *grids,
                     const float *origins, const float *spacing, float contact_offset,
                     float max_push, const float *coll_w);
/* which bones each vertex tests: vertex_group[nv] selects the bone range
 * [group_bone_off[g], group_bone_off[g+1]) (default: one group, all bones).
 * Batched instances each test only their own bone (no reference counterpart:
 * the reference has one skeleton collider, humanV2.py:319-418). */
int32_t xpbd_set_sdf_groups(xpbd_ctx *ctx, int32_t ngroups, const int32_t *group_bone_off,
                            const int32_t *vertex_group);
/* per-bone rigid transforms, row-major [R|t] 3x4 each (AnimatedBoneMeshCollider.set_frame,
 * src/example/humanV2.py:386-398) */
int32_t xpbd_set_bone_transforms(xpbd_ctx *ctx, int32_t nbones, const float *xforms);
/* build the device layout; call after the definition calls above */
int32_t xpbd_finalize(xpbd_ctx *ctx, const xpbd_tuning *tuning);
int32_t xpbd_get_stats(xpbd_ctx *ctx, xpbd_stats *out);

/* Closest point of each query point on a triangle mesh among the faces nearer than max_dist, on the
 * host (no context, no GPU): per point the face (-1: none), the barycentric weights (3 floats) of the
 * closest point on (tri[0], tri[1], tri[2]) and the signed distance along the face normal
 * (B - A) x (C - A); near-ties go to the lowest face index.  Replaces the bind queries of
 * bone_bind_kernel (src/muscle/mg/kernel.py:288-312: a muscle vertex bound to its bone surface
 * for target_mode "mesh") and fascia_bind_kernel (src/muscle/fascia/fascia.py:135-157), which
 * call Warp's mesh_query_point. */
int32_t xpbd_mesh_bind(int32_t nv, const float *verts, int32_t ntris, const int32_t *tris, int32_t npts,
                       const float *pts, float max_dist, int32_t *face, float *bary, float *gap);

/* Two-level hierarchical correction once per substep, after integrate and before the fine
 * iterations (MultiGridSim._hpbd_vcycle, src/muscle/mg/mg.py:1064-1079; coarse sweeps
 * _coarse_arap_sweep mg.py:939-952; restriction restrict_galerkin_scatter_kernel +
 * block_solve_kernel, src/muscle/mg/kernel.py:138-180; prolong_delta_kernel kernel.py:9-40).
 * The coarse cage is another finalized context (flat strategy): ARAP-SVD constraints, masses,
 * and its active pins are the coarse pins: they take part in every coarse projection with their
 * inverse mass and keep their restricted positions (the fine level's pin rule, mg.py:1337,1355);
 * stopped nodes are held with zero inverse mass.  Coarse nodes are grouped in blocks (one per
 * muscle, contiguous in the coarse context's caller order; at most 1024 nodes and 2048 tets per
 * block).  Per substep, per block b:
 *   [A t]_b = (sum_{i in fit_b} x_i [r_i - c_b; 1]^T) S_b^-1      (r = fine rest positions)
 *   q_j     = sum_l G_b[j, l] (sum_i P_il x_i + eps_b (A_b (q0_l - c_b) + t_b))
 * then iters coarse Jacobi sweeps, then x_i += sum_j P_ij (q_j - q_prev_j) for every vertex
 * that is not an active pin (also added to the previous positions when kinematic != 0, so the
 * shift does not enter the velocity), then the fine pins are re-applied. */
typedef struct {
    const int32_t *p_idx;       /* nv*8 coarse node per trilinear corner (coarse caller order) */
    const float *p_w;           /* nv*8 weights in [0, 1]                                       */
    int32_t nblk;
    const int32_t *blk_off;     /* nblk+1 coarse node ranges                                    */
    const float *minv;          /* sum n_b^2: G_b = (P^T P + eps_b I)^-1, row-major per block   */
    const float *eps;           /* nblk                                                         */
    const int32_t *fit_off;     /* nblk+1 ranges into fit_fine                                  */
    const int32_t *fit_fine;    /* fine vertices of each block's affine fit                     */
    const float *fit_center;    /* nblk*3 c_b                                                   */
    const float *fit_sinv;      /* nblk*16 S_b^-1, S_b = sum [r_i - c_b; 1][r_i - c_b; 1]^T     */
    int32_t iters;              /* coarse sweeps per substep; <= 0 turns the level off          */
    int32_t kinematic;
    int32_t pin_rule;           /* 0: coarse pins take part in the sweeps and keep their restricted
                                   positions; 1: the reference's rule, pins (stopped in the coarse
                                   context) start at their goal (xpbd_set_coarse_pin_targets) and are
                                   held (_coarse_solve, mg.py:1047-1062)                            */
} xpbd_coarse_desc;
/* coarse = NULL or desc = NULL turns it off.  The coarse context must outlive its use;
 * xpbd_finalize drops the link. */
int32_t xpbd_set_coarse_level(xpbd_ctx *ctx, xpbd_ctx *coarse, const xpbd_coarse_desc *desc);
/* pin rule 1: the restricted bone goal of every coarse node (n = coarse nodes, xyz n*3, coarse
 * caller order; only the pins' entries are read), per frame (_refresh_coarse_target,
 * mg.py:1013-1030) */
int32_t xpbd_set_coarse_pin_targets(xpbd_ctx *ctx, int32_t n, const float *xyz);
/* ---- fascia shell (Fascia, src/muscle/fascia/fascia.py:45-572) ----------
 * The fascia is another finalized context (its own tet mesh and constraints, quasi-static
 * parameters) that runs one substep inside every substep of this context, after this context's
 * iterations and before its DCD and bone push (MultiGridSim post_phase, mg.py:1362-1366), on this
 * context's stream.  Finalizing either context drops the link; NULL fascia or desc removes it.
 *
 * mode 0, ride (fascia_mode "ride", the humanV2 default; _build_muscle_bind fascia.py:443-490,
 * build_program_ride :518-538): xpbd_set_fascia binds each fascia rest vertex to the closest point
 * of this context's surface triangles among the faces nearer than bind_max_dist (barycentric
 * weights and the normal gap (x - closest)·n, n the triangle's unit normal; at near-ties the lowest
 * triangle index) and makes the bound vertices the fascia's active pins.  Each substep moves them
 * to their barycentric point on the current surface plus gap along the current normal, then runs
 * the fascia's integrate and iters iterations, each re-applying the pins.  The fascia never moves
 * this context.
 *
 * mode 1, squeeze (_build_coupling :354-424, build_program :547-565, advect_bone :540-545): the
 * fascia follows the skeleton and a one-way glue pulls this context's surface vertices towards it.
 * follow0 gives, per fascia vertex, the bone-driven world position of its follow point at rest
 * (the reference's target_pos + rest_offset of the nearest muscle vertex); xpbd_set_fascia_follow
 * gives it per frame and adds its change to every fascia vertex (advect_bone).  Each substep snaps
 * the anchors to their follow point plus anchor_offset, integrates, then runs iters x (fascia
 * iteration, glue): per pair, d = |x_f - x_m|, C = d - restlen, dl = -C / (w_m + compliance/dt^2)
 * (the glue multiplier restarts at zero each iteration), x_m += omega * (-dl * w_m) (x_f - x_m)/d,
 * skipped for d < 1e-9 or a denominator below 1e-12; w_m is this context's inverse mass. */
typedef struct {
    int32_t nsurf;
    const int32_t *surf_tris;   /* ride: nsurf*3 vertex ids of this context's surface (caller order) */
    float bind_max_dist;        /* ride: fascia_bind_max_dist                                       */
    int32_t iters;              /* fascia iterations per substep (fascia_glue.iterations)           */
    int32_t mode;               /* 0 ride, 1 squeeze                                                */
    const float *follow0;       /* squeeze: fascia nv*3 follow positions at rest (fascia caller order) */
    int32_t nglue;
    const int32_t *glue_fascia; /* squeeze: nglue fascia vertices (fascia caller order)              */
    const int32_t *glue_muscle; /* squeeze: nglue vertices of this context (caller order), distinct  */
    const float *glue_restlen;  /* squeeze: nglue rest distances                                    */
    const float *glue_compliance; /* squeeze: nglue 1/stiffness                                     */
    float omega;                /* squeeze: the glue's relaxation factor (the parent's omega)       */
    int32_t nanchor;
    const int32_t *anchor;      /* squeeze: fascia vertices snapped each substep (caller order)    */
    const float *anchor_offset; /* squeeze: nanchor*3, fascia rest minus its follow point at rest  */
} xpbd_fascia_desc;
int32_t xpbd_set_fascia(xpbd_ctx *ctx, xpbd_ctx *fascia, const xpbd_fascia_desc *desc);
/* the rest matrices B = Dm^-1 times 1/scale (Fascia._apply_rest_shrink, fascia.py:341-352): the
 * shape constraints see the rest shape scaled by `scale`; rest volumes and masses are kept.  After
 * xpbd_set_mesh, before xpbd_finalize. */
int32_t xpbd_scale_rest_shape(xpbd_ctx *ctx, float scale);
/* squeeze: this frame's follow positions (fascia nv*3, fascia caller order); enqueues the
 * advection of every fascia vertex by their change (advect_bone, fascia.py:540-545; per frame,
 * before the substeps) */
int32_t xpbd_set_fascia_follow(xpbd_ctx *ctx, const float *xyz);
/* the fascia's substep alone on the current positions (Fascia.substep, fascia.py:567-572), for
 * callers that drive the phases one by one */
int32_t xpbd_fascia_substep(xpbd_ctx *ctx);
/* the bind, per fascia vertex (caller order; any pointer may be NULL): surface triangle (-1:
 * free), 3 barycentric weights, gap; and the number bound */
int32_t xpbd_get_fascia_bind(xpbd_ctx *ctx, int32_t *face, float *w, float *gap, int32_t *nbound);
/* the V-cycle alone, for callers that drive the phases one by one */
int32_t xpbd_vcycle(xpbd_ctx *ctx);

/* ---- state ----------------------------------------------------------- */
int32_t xpbd_set_state(xpbd_ctx *ctx, const float *pos, const float *pprev, const float *vel);
int32_t xpbd_get_state(xpbd_ctx *ctx, float *pos, float *pprev, float *vel);

/* ---- phases (XpbdEngine methods, engine.py:798-1134) ------------------ */
int32_t xpbd_integrate(xpbd_ctx *ctx);                /* integrate() + apply_pintotarget()  */
int32_t xpbd_clear(xpbd_ctx *ctx);                    /* clear()                             */
int32_t xpbd_solve_constraints(xpbd_ctx *ctx);        /* solve_constraints(run_extras=False) */
int32_t xpbd_apply_dP(xpbd_ctx *ctx);                 /* apply_dP()                          */
int32_t xpbd_apply_pins(xpbd_ctx *ctx);               /* apply_pintotarget()                 */
int32_t xpbd_iterate(xpbd_ctx *ctx, int32_t n);       /* n x (clear, dispatch, apply_dP, pin), fused */
int32_t xpbd_collide(xpbd_ctx *ctx);                  /* collision_solvers: bone push (grid SDF or bone mesh) */
int32_t xpbd_update_velocities(xpbd_ctx *ctx);        /* update_velocities()                 */
/* MultiGridSim.substep (mg.py:1322-1384): integrate, pin, iters x iterate, collide, velocity */
int32_t xpbd_substep(xpbd_ctx *ctx, int32_t iters);
/* MultiGridSim.step_frame substep loop (mg.py:1409-1418); use_graph=1 replays a
 * captured hipGraph of the whole frame (re-captured when iters/nsub change) */
int32_t xpbd_step_frame(xpbd_ctx *ctx, int32_t nsub, int32_t iters, int32_t use_graph);
int32_t xpbd_synchronize(xpbd_ctx *ctx);

/* ---- exact-mesh bone collider ------------------------------------------
 * AnimatedBoneMeshCollider (src/example/humanV2.py:319-418): one triangle mesh over the
 * skeleton (wp.Mesh, built once; points refreshed and refit per frame), the push of
 * solve_mesh_contact_pos_kernel (src/muscle/collision/sdf/kernel.py:123-170) and the
 * penetration meter _mark_inside_kernel (humanV2.py:67-87) with CollisionMeter.measure
 * (humanV2.py:163-167).  Query: the closest point among the faces nearer than max_dist, signed
 * by the angle-weighted pseudonormal of the closest feature (face, edge or vertex), as Warp's
 * mesh_query_point_sign_normal; push: x -= n err with err = d - contact_offset < 0,
 * n = (x - closest) / |x - closest| * sign, skipped for |x - closest| < 1e-9 and, when
 * max_push > 0, for -err > max_push.  Vertex groups restrict which sub-meshes ("bones") a
 * vertex tests (the closest point over the union of those sub-meshes). */
typedef struct {
    int32_t nverts;             /* bone mesh vertices                                          */
    const float *verts;         /* nverts*3 positions of the current frame                     */
    int32_t ntris;
    const int32_t *tris;        /* ntris*3 vertex ids, (b - a) x (c - a) pointing out of the bone
                                   (the reference flips the SNU bones' winding to get this)     */
    int32_t nbones;             /* sub-meshes: triangles tri_off[b] .. tri_off[b + 1] - 1        */
    const int32_t *tri_off;     /* nbones+1; NULL with nbones <= 1: one sub-mesh                 */
    int32_t ngroups;            /* 0: every vertex tests every sub-mesh                          */
    const int32_t *goff;        /* ngroups+1: group g tests sub-meshes goff[g] .. goff[g + 1] - 1 */
    const int32_t *vgroup;      /* nv group ids, caller's vertex order                           */
    const float *coll_w;        /* nv, 0 = not pushed (coll_inv_mass, humanV2.py:361-366); NULL: all */
    const float *meter_w;       /* nv, 0 = not metered (meter_mask, humanV2.py:367-375); NULL: all  */
    float max_dist, contact_offset, max_push;
    int32_t resolve;            /* 1: the substep's collision stage pushes with this mesh (the grid
                                   SDF is then not used); 0: meter only                           */
    int32_t coherent;           /* 1: skip the query of a vertex that its last query and the motion
                                   since (its own, and the largest bone-vertex displacement of each
                                   upload) prove to be outside by more than the contact offset; the
                                   results equal querying every vertex when each sub-mesh bounds a
                                   solid and the sub-meshes a vertex tests do not overlap; needs
                                   closed sub-meshes (EINVAL otherwise).  0: query every vertex.
                                   2: 1 if the mesh is closed, else 0                              */
} xpbd_bone_mesh_desc;
/* desc = NULL removes the collider.  Allowed before or after xpbd_finalize. */
int32_t xpbd_set_bone_mesh(xpbd_ctx *ctx, const xpbd_bone_mesh_desc *desc);
/* per frame: new bone vertex positions (nverts*3, same topology), then the refit
 * (AnimatedBoneMeshCollider.set_frame, humanV2.py:386-398); staged like the other setters */
int32_t xpbd_set_bone_mesh_points(xpbd_ctx *ctx, const float *verts);
/* turn the push in the collision stage on or off (the sim's collision_solvers list) */
int32_t xpbd_set_bone_mesh_resolve(xpbd_ctx *ctx, int32_t resolve);
/* one push of the current positions (AnimatedBoneMeshCollider.__call__, humanV2.py:400-406) */
int32_t xpbd_bone_mesh_push(xpbd_ctx *ctx);
/* queries made since the collider was set (push and meter; with coherent queries the vertices
 * that were not certified outside), synchronises.  No reference counterpart. */
int32_t xpbd_bone_mesh_queries(xpbd_ctx *ctx, int64_t *queries);
/* the meter on the current positions: depth (nv floats, caller order, > 0 inside; may be NULL),
 * the number of metered vertices deeper than collide_eps and the largest depth; synchronises */
int32_t xpbd_bone_penetration(xpbd_ctx *ctx, float collide_eps, float *depth, int32_t *count,
                              float *max_depth);

/* ---- tri-tri DCD (UdcfDcdCollider, src/muscle/collision/udcf/solver.py:146-457) --- */
typedef struct {
    float thickness;          /* broadphase sphere slack (UdcfParams.thickness)        */
    float multiplier;         /* collisionConstraintMultiplier                          */
    float friction_static, friction_dynamic;
    float offset;             /* narrow-phase accept threshold (1e-6)                   */
    float normal_cull;        /* >= 1 disables the co-oriented self-pair cull           */
    float cell_size;          /* base cell of the grid hierarchy; 0 = 1.01 (2 r50 + thickness), r50 the
                                 median triangle radius at each rebuild */
    int32_t allow_self;       /* movable-vs-movable pairs                               */
    int32_t iters;            /* DCD passes per substep (humanV2 "iters", default 2)    */
    int32_t rebuild_substeps; /* 1 = broadphase every substep, 0 = once per frame       */
    int64_t max_pairs;        /* pair / contact capacity (default 1 << 22, below 2^31)  */
} xpbd_dcd_params;
/* surface triangles (ntris*3 vertex ids) and optional group (0 movable, 1 static);
 * enables the DCD post pass before the bone push (humanV2.py:820-866) */
int32_t xpbd_set_dcd(xpbd_ctx *ctx, int32_t ntris, const int32_t *tris, const int32_t *group,
                     const xpbd_dcd_params *p);
/* one DCD pass on the current positions (UdcfDcdCollider.solve, solver.py:365-440) */
int32_t xpbd_dcd_solve(xpbd_ctx *ctx, int32_t rebuild);
/* the (i<j)-sorted candidate pair list of the last broadphase (int32 pairs, 2 per entry).
 * The broadphase keeps its pairs in append order; this call sorts a copy (segments by first
 * index, each by the second).  n_total = pairs found; beyond max_pairs some were dropped. */
int32_t xpbd_dcd_pairs(xpbd_ctx *ctx, int32_t *out, int64_t cap, int64_t *n_total);
/* UdcfDcdCollider.stats (solver.py:442-457): raw pair count, contacts of the last pass, cell */
int32_t xpbd_dcd_stats(xpbd_ctx *ctx, int64_t *n_pairs, int64_t *n_contacts, float *cell);
/* broadphase limits hit at the last rebuild (0 and 0 on a valid mesh): triangles left out
 * (bounding sphere above the top grid level, or a non-finite or out-of-range centroid) and
 * truncated work (searches stopped at the visit limit, pairs dropped beyond max_pairs, and
 * pair segments too long to sort in the last pair-list query).  No reference counterpart. */
int32_t xpbd_dcd_overflow(xpbd_ctx *ctx, int64_t *left_out, int64_t *truncated);

/* ---- learned warm start: per-tet MLP (no reference counterpart) ------- */
/* 24 features (22 used) -> 64 -> 64 -> 4 outputs (3 used), ReLU hidden layers; layers 1-2
 * on BF16 MFMA with FP32 accumulate.  All pointers are device pointers. */
#define XPBD_MLP_IN 24
#define XPBD_MLP_HIDDEN 64
#define XPBD_MLP_OUT 4
typedef struct {
    const uint16_t *W1;   /* bf16 bits [HIDDEN][IN]     */
    const float *b1;      /* [HIDDEN]                   */
    const uint16_t *W2;   /* bf16 bits [HIDDEN][HIDDEN] */
    const float *b2;      /* [HIDDEN]                   */
    const float *W3;      /* fp32 [OUT][HIDDEN]         */
    const float *b3;      /* [OUT]                      */
} xpbd_mlp_dev;
/* Y (n x OUT); H1, H2 (n x HIDDEN, post-ReLU) written when non-NULL; stream = hipStream_t */
int32_t xpbd_mlp_forward(const xpbd_mlp_dev *net, const float *X, int64_t n, float *Y, float *H1,
                         float *H2, void *stream);
/* number of floats of the gradient vector: W1, b1, W2, b2, W3, b3 flattened in that order */
int32_t xpbd_mlp_grad_size(void);
/* d loss / d params from dY (n x OUT); work: device workspace of >= ceil(n/32) * grad_size floats */
int32_t xpbd_mlp_backward(const xpbd_mlp_dev *net, const float *X, const float *H1, const float *H2,
                          const float *dY, int64_t n, float *grad, float *work, int64_t work_floats,
                          void *stream);
/* xpbd_mlp_backward without the workspace argument; allocates nothing (no hipMalloc / hipFree,
 * no allocating library call), for callers that cannot allocate device memory: code under HIP
 * graph capture, training loops that must not allocate per step.  No reference counterpart.
 * The workspace is a static device array sized for n <= XPBD_MLP_NOALLOC_MAX_ROWS rows (12.3 MB,
 * held on each GPU the library runs on until the process ends); larger n returns XPBD_EINVAL
 * before anything is launched.  Runs on the current device (stream must belong to it).  All calls
 * on one GPU share that array: calls that overlap in time on one GPU race. */
#define XPBD_MLP_NOALLOC_MAX_ROWS 16384
int32_t xpbd_mlp_backward_noalloc(const xpbd_mlp_dev *net, const float *X, const float *H1,
                                  const float *H2, const float *dY, int64_t n, float *grad,
                                  void *stream);

/* multiplier modes (flat and block strategies; no reference counterpart):
 * 0 = plain zero-lambda Jacobi (the reference GPU path, solve_kernels.py:271-281 with clear()
 *     every iteration, mg.py:1348);
 * 1 = lambda carried across the iterations of a substep from 0 (a reference solve: at the end
 *     of the substep its multipliers become the base of the next warm start);
 * 2 = learned warm start: the network predicts, per tet and constraint type, a gain on the base,
 *     lambda0 = base * clamp(1 + out_scale * y, 0, 2), which also becomes the next base; the
 *     first iteration applies lambda0 as an impulse along the constraint gradients, the rest
 *     are plain zero-lambda iterations;
 * 3 = the same with lambda0 = base (no network);
 * 4 = frames with a schedule (xpbd_set_warm_schedule): reference substeps run mode 1 with the
 *     reference iteration count, the others mode 2 (network) or mode 3 (base only) with the
 *     frame's iteration count; xpbd_substep and the phase calls use mode 2 or 3. */
int32_t xpbd_set_lambda_mode(xpbd_ctx *ctx, int32_t mode);
/* network weights (host fp32; W1, W2 stored as bf16), feature standardisation (IN each) and
 * per-output scales (OUT) */
int32_t xpbd_set_mlp(xpbd_ctx *ctx, const float *W1, const float *b1, const float *W2, const float *b2,
                     const float *W3, const float *b3, const float *feat_mean, const float *feat_invstd,
                     const float *out_scale);
/* start-of-substep multiplier set-up of the current mode (xpbd_substep calls it after integrate) */
int32_t xpbd_warm_start(xpbd_ctx *ctx);
/* per-element features (n_elems x IN, device pointer; normalized = standardised) */
int32_t xpbd_features(xpbd_ctx *ctx, float *dev_out, int32_t normalized);
/* per-element multipliers (n_elems x 4 floats, device pointers) */
int32_t xpbd_get_lambda(xpbd_ctx *ctx, float *dev_out);
int32_t xpbd_set_lambda(xpbd_ctx *ctx, const float *dev_in);   /* sets the carried multipliers and the base */
int32_t xpbd_set_reference_iters(xpbd_ctx *ctx, int32_t n);
/* mode 4 schedule: reference substeps (lambda carried from zero, ref_iters iterations) are
 * those of index % ref_every == 0 within xpbd_step_frame (ref_every 0: the first substep of
 * each frame only, the default); between them the base of the last reference is applied as
 * an impulse, scaled by the network when use_network is 1.  No reference counterpart. */
int32_t xpbd_set_warm_schedule(xpbd_ctx *ctx, int32_t ref_iters, int32_t ref_every, int32_t use_network);
/* save (1) / restore (0) positions, previous positions and velocities on the device */
int32_t xpbd_snapshot(xpbd_ctx *ctx, int32_t save);

/* ---- diagnostics ------------------------------------------------------ */
/* per-type residual sums: out[0] sum ||sigma-1||^2 over ARAP tets, out[1] count,
 * out[2] sum (V-V0)^2 over volume tets, out[3] count, out[4] sum C_D^2 + C_H^2 (SNH),
 * out[5] count, out[6] sum C^2 (aniso), out[7] count.
 * arap_residual_accum_kernel (src/muscle/mg/kernel.py:87-118) generalised per type. */
int32_t xpbd_residuals(xpbd_ctx *ctx, double *out8);
/* count_inverted_tets_kernel (src/muscle/engine.py:410-428) */
int32_t xpbd_count_inverted(xpbd_ctx *ctx, int32_t *out);
/* SDF cells of the last collide pass (nv*2: bone, linear cell; -1 = none) */
int32_t xpbd_get_sdf_cells(xpbd_ctx *ctx, int32_t *cells);
/* event timing of projection launches: enable=1 starts accumulating */
int32_t xpbd_kernel_timing(xpbd_ctx *ctx, int32_t enable);
int32_t xpbd_get_kernel_timing(xpbd_ctx *ctx, double *total_ms, int32_t *launches);
/* average duration of one projection iteration (the kernels of xpbd_iterate,
 * lambda mode 0): n * reps (<= 4096, n even) iterations captured in one HIP graph,
 * launched once to warm up and once between two events on the context stream;
 * X is restored afterwards.  Measures the kernels without host dispatch gaps.  No reference counterpart
 * (the reference brackets sections with wp.synchronize() and perf_counter,
 * tools/profile_humanV2.py:56-131). */
int32_t xpbd_time_iterations(xpbd_ctx *ctx, int32_t n, int32_t reps, double *us_per_iter);

#ifdef __cplusplus
}
#endif
#endif
