/*
 * CPU oracle for the XPBD muscle substep — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference (chunleili/rlmuscleV2) Warp kernels on
 * the XPBD hot path, written from the reference source text.  It is the checker
 * the HIP library is compared against and the reported CPU baseline; nothing in
 * the product library links or calls it.
 *
 * Data model mirrors the reference's struct-of-arrays constraint table
 * (src/muscle/engine.py:1179-1211) and per-vertex / per-tet fields
 * (src/muscle/musclesim.py:331-352).  All arithmetic is FP32 as in the
 * reference, except the 3x3 SVD, which is a double-precision one-sided Jacobi
 * SVD (a textbook method) with the reference's ssvd sign fix applied.
 */
#ifndef XPBD_ORACLE_H
#define XPBD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* constraint type tags: src/muscle/defs.py:20-36 */
#define ORC_PIN             157323
#define ORC_ATTACH          1650556
#define ORC_PINTOTARGET     771624311
#define ORC_TETVOLUME       (-215389979)
#define ORC_TETARAPSVD      (-92199132)
#define ORC_TETARAPANISO    (-92199133)
#define ORC_TETSNH          (-503462311)

typedef struct {
    int n;
    const int   *type;                 /* n      */
    const int   *pts;                  /* n*4    */
    const float *stiffness;            /* n      */
    const float *compressionstiffness; /* n      */
    const float *dampingratio;         /* n      */
    const int   *tetid;                /* n      */
    float       *L;                    /* n*3    */
    const float *restlength;           /* n      */
    const float *restvector;           /* n*4    */
    const float *restdir;              /* n*3    */
    float       *reaction_accum;       /* n*3 (may be NULL) */
} orc_cons_t;

typedef struct {
    int nv, nt;
    float *pos, *pprev, *vel, *force, *mass;   /* nv*3 / nv */
    int   *stopped;                            /* nv */
    float *dP, *dPw;                           /* nv*3, nv */
    const float *rest_matrix;                  /* nt*9 row-major B = inv(Dm) */
    const float *v_fiber_dir;                  /* nv*3 */
    const float *activation;                   /* nt */
    const float *tendonmask;                   /* nt */
} orc_state_t;

typedef struct {
    float dt;
    float gravity[3];
    float veldamping;
    int   quasi_static;
    float contraction_ratio;
    float fiber_stiffness_scale;
    int   has_compress;
    int   use_jacobi;
    float exp_damp, const_damp;
} orc_params_t;

/* status codes */
#define ORC_OK 0
#define ORC_EUNSUPPORTED 2

void orc_svd3(const float *F, float *U, float *sigma, float *V);
void orc_svd3_double(const double *F, double *U, double *sigma, double *V);

void orc_precompute_rest(int nt, const float *pos0, const int *tets, float density,
                         float *rest_volume, float *rest_matrix, float *mass,
                         float *total_rest_volume);
void orc_integrate(orc_state_t *s, const orc_params_t *p);
void orc_clear(orc_state_t *s, const orc_cons_t *c);
int  orc_dispatch(const orc_cons_t *c, orc_state_t *s, const orc_params_t *p, int nthreads);
void orc_apply_dP(orc_state_t *s, int nthreads);
void orc_pintotarget_masked(int n, const int *pts, const float *target, const float *rest_offset,
                            const int *active, float *pos);
void orc_pintotarget_cons(const orc_cons_t *c, float *pos);
void orc_update_velocities(orc_state_t *s, const orc_params_t *p);
int  orc_sdf_grid_push(int nv, float *pos, const float *coll_w, int nbones, int res,
                       const float *grids, const float *origin, const float *spacing,
                       const float *xforms, const int *vgroup, const int *goff,
                       float contact_offset, float max_push, int *cell_out);
void orc_residuals(int nt, const int *tets, const float *pos, const float *rest_matrix,
                   const float *rest_volume, double *out_sumsq_arap, double *out_sumsq_vol,
                   int *out_inverted);
int  orc_count_inverted(int nt, const int *tets, const float *pos, const float *rest_matrix);

#ifdef __cplusplus
}
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif
/* ---- tri-tri DCD (src/muscle/collision/udcf/kernel.py) ---------------- */
typedef struct {
    int cur[3], closest[3];
    float C, C1, C2;
    float n[3], b1[3], b2[3];
    int valid;
} orc_contact_t;

void orc_dcd_tri_data(int m, const int *tris, const float *x, float *centroids, float *radii,
                      float *normals);
/* all pairs i<j passing the broadphase emit rules, in (i, j) order; returns the
 * number found (entries beyond max_pairs are counted but not stored) */
long orc_dcd_pairs_bruteforce(int m, const int *tris, const float *centroids, const float *radii,
                              const float *normals, const int *group, float normal_cull,
                              int allow_self, float thickness, long max_pairs, int *pair_a,
                              int *pair_b);
void orc_dcd_narrow(long npairs, const int *pair_a, const int *pair_b, const int *tris,
                    const float *x, float offset_eps, orc_contact_t *out);
void orc_dcd_respond(long ncontacts, const orc_contact_t *c, const float *x, const float *x_prev,
                     const float *inv_mass, float multiplier, float mu_s, float mu_d, float *dP,
                     float *dPw);
void orc_dcd_apply(int n, float *pos, const float *dP, const float *dPw, int use_avg);
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
/* ---- exact-mesh bone collider (src/muscle/collision/sdf/kernel.py:123-170) -- */
void orc_mesh_normals(int nbv, const float *bv, int ntris, const int *tris, double *fn, double *vn);
void orc_mesh_query(int nq, const float *q, int nbv, const float *bv, int ntris, const int *tris,
                    int nbones, const int *tri_off, const double *fn, const double *vn, const int *adj,
                    const int *vgroup, const int *goff, float max_dist, double *sd, double *cp, int *face,
                    int *region, int nthreads);
int  orc_mesh_push(int nv, float *pos, const float *coll_w, const double *sd, const double *cp,
                   const int *face, float contact_offset, float max_push);
void orc_mesh_bind(int nq, const float *q, const float *bv, int ntris, const int *tris, float max_dist,
                   int *face, double *cp, int nthreads);
#ifdef __cplusplus
}
#endif
