/*
 * CPU oracle for the XPBD muscle substep — TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference Warp kernels (chunleili/rlmuscleV2).
 * Each function cites the reference file:line it follows.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this code, as
 * the checker / reported CPU baseline; the product library never does.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -shared -fPIC (see oracle/Makefile).
 * -ffp-contract=off keeps every FP32 expression rounded as written, so the
 * integer results that are derived from floats (SDF cell indices) are exact.
 */
#include "xpbd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* small vector helpers (FP32, left-to-right evaluation)               */
/* ------------------------------------------------------------------ */
typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 ld3(const float *a, int i) { return mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
static inline void st3(float *a, int i, v3 v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline v3 scl(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 dvs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float length(v3 a) { return sqrtf(dot(a, a)); }

/* Ds = [c0 c1 c2] (columns); F = Ds * B, both row-major 3x3 */
static inline void ds_times_b(v3 c0, v3 c1, v3 c2, const float *B, float *F) {
    const float Ds[9] = {c0.x, c1.x, c2.x, c0.y, c1.y, c2.y, c0.z, c1.z, c2.z};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            F[3 * i + j] = Ds[3 * i + 0] * B[0 * 3 + j] + Ds[3 * i + 1] * B[1 * 3 + j] +
                           Ds[3 * i + 2] * B[2 * 3 + j];
}

static inline float det3(const float *m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* get_inv_mass_fn: src/muscle/constraint/solve_kernels.py:71-80 */
static inline float inv_mass(const orc_state_t *s, int i) {
    if (s->stopped[i] != 0) return 0.0f;
    float m = s->mass[i];
    return m > 0.0f ? 1.0f / m : 0.0f;
}

/* update_dP: solve_kernels.py:52-56 (serial accumulation, or atomics when threaded) */
static inline void update_dP(orc_state_t *s, v3 dp, int pt, int threaded) {
    if (threaded) {
#pragma omp atomic
        s->dP[3 * pt] += dp.x;
#pragma omp atomic
        s->dP[3 * pt + 1] += dp.y;
#pragma omp atomic
        s->dP[3 * pt + 2] += dp.z;
#pragma omp atomic
        s->dPw[pt] += 1.0f;
    } else {
        s->dP[3 * pt] += dp.x;
        s->dP[3 * pt + 1] += dp.y;
        s->dP[3 * pt + 2] += dp.z;
        s->dPw[pt] += 1.0f;
    }
}

static inline void gs_move(orc_state_t *s, int pt, v3 dp) {
    st3(s->pos, pt, add(ld3(s->pos, pt), dp));
}

/* ------------------------------------------------------------------ */
/* 3x3 SVD: one-sided Jacobi in double (textbook, Demmel-Veselic),      */
/* sigma sorted descending, then the ssvd_fn sign fix                   */
/* (solve_kernels.py:85-112): flip U[:,2] and sigma[2] if det U < 0,    */
/* same for V.  Result: U, V proper rotations, sigma[2] carries sign.   */
/* ------------------------------------------------------------------ */
static double det3d(const double *m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

static void complete_basis(double *U, int known) {
    /* columns 0..known-1 orthonormal; fill the rest */
    if (known == 0) { memset(U, 0, 9 * sizeof(double)); U[0] = U[4] = U[8] = 1.0; return; }
    if (known == 1) {
        double a[3] = {U[0], U[3], U[6]};
        /* pick the axis least aligned with a */
        double e[3] = {0, 0, 0};
        int k = 0;
        double m = fabs(a[0]);
        if (fabs(a[1]) < m) { m = fabs(a[1]); k = 1; }
        if (fabs(a[2]) < m) { k = 2; }
        e[k] = 1.0;
        double b[3] = {a[1] * e[2] - a[2] * e[1], a[2] * e[0] - a[0] * e[2], a[0] * e[1] - a[1] * e[0]};
        double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        for (int i = 0; i < 3; ++i) U[3 * i + 1] = b[i] / nb;
    }
    double a[3] = {U[0], U[3], U[6]}, b[3] = {U[1], U[4], U[7]};
    double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double nc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int i = 0; i < 3; ++i) U[3 * i + 2] = c[i] / nc;
}

void orc_svd3_double(const double *F, double *U, double *sigma, double *V) {
    double A[9], W[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(A, F, sizeof(A));
    static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 60; ++sweep) {
        int rotated = 0;
        for (int r = 0; r < 3; ++r) {
            int p = P[r], q = Q[r];
            double al = 0, be = 0, ga = 0;
            for (int i = 0; i < 3; ++i) {
                al += A[3 * i + p] * A[3 * i + p];
                be += A[3 * i + q] * A[3 * i + q];
                ga += A[3 * i + p] * A[3 * i + q];
            }
            if (ga == 0.0 || fabs(ga) <= 1e-17 * sqrt(al * be)) continue;
            rotated = 1;
            double zeta = (be - al) / (2.0 * ga);
            double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
            for (int i = 0; i < 3; ++i) {
                double ap = A[3 * i + p], aq = A[3 * i + q];
                A[3 * i + p] = c * ap - s * aq;
                A[3 * i + q] = s * ap + c * aq;
                double vp = W[3 * i + p], vq = W[3 * i + q];
                W[3 * i + p] = c * vp - s * vq;
                W[3 * i + q] = s * vp + c * vq;
            }
        }
        if (!rotated) break;
    }
    double sg[3];
    for (int j = 0; j < 3; ++j)
        sg[j] = sqrt(A[j] * A[j] + A[3 + j] * A[3 + j] + A[6 + j] * A[6 + j]);
    /* sort descending */
    int ord[3] = {0, 1, 2};
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b)
            if (sg[ord[b]] > sg[ord[a]]) { int t = ord[a]; ord[a] = ord[b]; ord[b] = t; }
    double smax = sg[ord[0]];
    int known = 0;
    for (int j = 0; j < 3; ++j) {
        int o = ord[j];
        sigma[j] = sg[o];
        for (int i = 0; i < 3; ++i) V[3 * i + j] = W[3 * i + o];
        if (sg[o] > 1e-300 && sg[o] > 1e-14 * smax) {
            for (int i = 0; i < 3; ++i) U[3 * i + j] = A[3 * i + o] / sg[o];
            known = j + 1;
        }
    }
    if (known < 3) complete_basis(U, known);
    if (det3d(U) < 0.0) {
        for (int i = 0; i < 3; ++i) U[3 * i + 2] = -U[3 * i + 2];
        sigma[2] = -sigma[2];
    }
    if (det3d(V) < 0.0) {
        for (int i = 0; i < 3; ++i) V[3 * i + 2] = -V[3 * i + 2];
        sigma[2] = -sigma[2];
    }
}

void orc_svd3(const float *F, float *U, float *sigma, float *V) {
    double Fd[9], Ud[9], sd[3], Vd[9];
    for (int i = 0; i < 9; ++i) Fd[i] = F[i];
    orc_svd3_double(Fd, Ud, sd, Vd);
    for (int i = 0; i < 9; ++i) { U[i] = (float)Ud[i]; V[i] = (float)Vd[i]; }
    for (int i = 0; i < 3; ++i) sigma[i] = (float)sd[i];
}

/* ------------------------------------------------------------------ */
/* rest precompute: precompute_rest_kernel, src/muscle/engine.py:199-224 */
/* (serial mass accumulation instead of atomics)                         */
/* ------------------------------------------------------------------ */
static void inverse3(const float *m, float *out) {
    float d = det3(m);
    if (d == 0.0f) { memset(out, 0, 9 * sizeof(float)); return; }
    float inv = 1.0f / d;
    out[0] = (m[4] * m[8] - m[5] * m[7]) * inv;
    out[1] = (m[2] * m[7] - m[1] * m[8]) * inv;
    out[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
    out[3] = (m[5] * m[6] - m[3] * m[8]) * inv;
    out[4] = (m[0] * m[8] - m[2] * m[6]) * inv;
    out[5] = (m[2] * m[3] - m[0] * m[5]) * inv;
    out[6] = (m[3] * m[7] - m[4] * m[6]) * inv;
    out[7] = (m[1] * m[6] - m[0] * m[7]) * inv;
    out[8] = (m[0] * m[4] - m[1] * m[3]) * inv;
}

void orc_precompute_rest(int nt, const float *pos0, const int *tets, float density,
                         float *rest_volume, float *rest_matrix, float *mass,
                         float *total_rest_volume) {
    float tot = 0.0f;
    for (int c = 0; c < nt; ++c) {
        const int *pts = tets + 4 * c;
        v3 p3 = ld3(pos0, pts[3]);
        v3 c0 = sub(ld3(pos0, pts[0]), p3), c1 = sub(ld3(pos0, pts[1]), p3),
           c2 = sub(ld3(pos0, pts[2]), p3);
        float Dm[9] = {c0.x, c1.x, c2.x, c0.y, c1.y, c2.y, c0.z, c1.z, c2.z};
        float vol = fabsf(det3(Dm)) / 6.0f;
        rest_volume[c] = vol;
        tot += vol;
        float mc = vol * density / 4.0f;
        for (int i = 0; i < 4; ++i) mass[pts[i]] += mc;
        inverse3(Dm, rest_matrix + 9 * c);
    }
    if (total_rest_volume) *total_rest_volume = tot;
}

/* integrate_kernel: src/muscle/engine.py:431-464 */
void orc_integrate(orc_state_t *s, const orc_params_t *p) {
    const v3 g = mk(p->gravity[0], p->gravity[1], p->gravity[2]);
    for (int i = 0; i < s->nv; ++i) {
        if (s->stopped[i] != 0) {
            st3(s->pprev, i, ld3(s->pos, i));
            st3(s->vel, i, mk(0, 0, 0));
            continue;
        }
        v3 extacc = g;
        float m = s->mass[i];
        if (m > 1.0e-12f) extacc = add(extacc, dvs(ld3(s->force, i), m));
        v3 x = ld3(s->pos, i);
        st3(s->pprev, i, x);
        if (p->quasi_static) {
            st3(s->pos, i, add(x, scl(p->dt * p->dt, extacc)));
            st3(s->vel, i, mk(0, 0, 0));
        } else {
            v3 v = scl(1.0f - p->veldamping, ld3(s->vel, i));
            v = add(v, scl(p->dt, extacc));
            st3(s->vel, i, v);
            st3(s->pos, i, add(x, scl(p->dt, v)));
        }
    }
}

/* clear: src/muscle/engine.py:858-863 (clear_dP_kernel, clear_cons_L_kernel :664-678) */
void orc_clear(orc_state_t *s, const orc_cons_t *c) {
    memset(s->dP, 0, sizeof(float) * 3 * s->nv);
    memset(s->dPw, 0, sizeof(float) * s->nv);
    if (c && c->n > 0) memset(c->L, 0, sizeof(float) * 3 * c->n);
}

/* ------------------------------------------------------------------ */
/* constraint projections                                               */
/* ------------------------------------------------------------------ */

/* gradient rows: grad_k[a] = sum_j dCdF[a][j] * B[k][j]  (solve_kernels.py:1269-1281) */
static inline v3 grad_row(const float *G, const float *B, int k) {
    return mk(G[0] * B[3 * k] + G[1] * B[3 * k + 1] + G[2] * B[3 * k + 2],
              G[3] * B[3 * k] + G[4] * B[3 * k + 1] + G[5] * B[3 * k + 2],
              G[6] * B[3 * k] + G[7] * B[3 * k + 1] + G[8] * B[3 * k + 2]);
}

/* solve_tetarap_svd_kernel: solve_kernels.py:1194-1307 */
static void solve_arap_svd(const orc_cons_t *cs, int c, orc_state_t *s, const orc_params_t *p,
                           int thr) {
    float k = cs->stiffness[c];
    if (k <= 0.0f) return;
    const int *pts = cs->pts + 4 * c;
    int tetid = cs->tetid[c];
    v3 P0 = ld3(s->pos, pts[0]), P1 = ld3(s->pos, pts[1]), P2 = ld3(s->pos, pts[2]),
       P3 = ld3(s->pos, pts[3]);
    float w[4];
    for (int i = 0; i < 4; ++i) w[i] = inv_mass(s, pts[i]);
    const float *B = s->rest_matrix + 9 * tetid;
    float F[9], U[9], sg[3], V[9];
    ds_times_b(sub(P0, P3), sub(P1, P3), sub(P2, P3), B, F);
    orc_svd3(F, U, sg, V);
    float sm[3] = {sg[0] - 1.0f, sg[1] - 1.0f, sg[2] - 1.0f};
    float sum_sq = sm[0] * sm[0] + sm[1] * sm[1] + sm[2] * sm[2];
    if (sum_sq < 1.0e-12f) return;
    float C = sqrtf(sum_sq), inv_C = 1.0f / C;
    float d0 = inv_C * sm[0], d1 = inv_C * sm[1], d2 = inv_C * sm[2];
    float G[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            G[3 * i + j] = U[3 * i + 0] * V[3 * j + 0] * d0 + U[3 * i + 1] * V[3 * j + 1] * d1 +
                           U[3 * i + 2] * V[3 * j + 2] * d2;
    v3 g[4];
    g[0] = grad_row(G, B, 0);
    g[1] = grad_row(G, B, 1);
    g[2] = grad_row(G, B, 2);
    g[3] = sub(sub(neg(g[0]), g[1]), g[2]);
    float wsum = w[0] * dot(g[0], g[0]) + w[1] * dot(g[1], g[1]) + w[2] * dot(g[2], g[2]) +
                 w[3] * dot(g[3], g[3]);
    if (wsum == 0.0f) return;
    float alpha_tilde = (1.0f / k) / (p->dt * p->dt);
    float L0 = cs->L[3 * c];
    float dL = -(C + alpha_tilde * L0) / (wsum + alpha_tilde);
    if (p->use_jacobi) {
        for (int i = 0; i < 4; ++i) update_dP(s, scl(dL * w[i], g[i]), pts[i], thr);
    } else {
        for (int i = 0; i < 4; ++i) gs_move(s, pts[i], scl(dL * w[i], g[i]));
        cs->L[3 * c] = L0 + dL;
    }
}

/* solve_tetvolume_kernel + tet_volume_update_xpbd_fn: solve_kernels.py:553-585, 187-281 */
static void solve_volume(const orc_cons_t *cs, int c, orc_state_t *s, const orc_params_t *p,
                         int thr) {
    float cst = cs->stiffness[c];
    if (cst <= 0.0f) return;
    float kcomp = cs->compressionstiffness[c];
    kcomp = kcomp >= 0.0f ? kcomp : cst;
    const int *pts = cs->pts + 4 * c;
    float w[4];
    for (int i = 0; i < 4; ++i) w[i] = inv_mass(s, pts[i]);
    v3 P[4];
    for (int i = 0; i < 4; ++i) P[i] = ld3(s->pos, pts[i]);
    v3 d0 = sub(P[0], P[3]), d1 = sub(P[1], P[3]), d2 = sub(P[2], P[3]);
    v3 g[4];
    g[0] = dvs(cross(d1, d2), 6.0f);
    g[1] = dvs(cross(d2, d0), 6.0f);
    g[2] = dvs(cross(d0, d1), 6.0f);
    g[3] = neg(add(add(g[0], g[1]), g[2]));
    float w_sum = w[0] * dot(g[0], g[0]) + w[1] * dot(g[1], g[1]) + w[2] * dot(g[2], g[2]) +
                  w[3] * dot(g[3], g[3]);
    if (!(w_sum > 1e-9f)) return;
    float restlength = cs->restlength[c];
    float volume = dot(cross(d0, d1), d2) / 6.0f;
    int comp = 0;
    if (p->has_compress && volume < restlength) comp = 1;
    float kstiff = comp ? kcomp : cst;
    int loff = comp < 2 ? comp : 2;
    if (kstiff == 0.0f) return;
    float l = cs->L[3 * c + loff];
    float alpha = 1.0f / kstiff;
    alpha = alpha / (p->dt * p->dt);
    float C = volume - restlength;
    float dsum = 0.0f, gamma = 1.0f;
    float kd = cs->dampingratio[c];
    if (kd > 0.0f) {
        float beta = kstiff * kd * p->dt * p->dt;
        gamma = alpha * beta / p->dt;
        dsum = 0.0f;
        for (int i = 0; i < 4; ++i) dsum += dot(g[i], sub(P[i], ld3(s->pprev, pts[i])));
        dsum = dsum * gamma;
        gamma = gamma + 1.0f;
    }
    float dlambda = (-C - alpha * l - dsum) / (gamma * w_sum + alpha);
    if (p->use_jacobi) {
        for (int i = 0; i < 4; ++i) update_dP(s, scl(dlambda * w[i], g[i]), pts[i], thr);
    } else {
        for (int i = 0; i < 4; ++i) gs_move(s, pts[i], scl(dlambda * w[i], g[i]));
        cs->L[3 * c + loff] = l + dlambda;
    }
}

/* solve_tetsnh_kernel + tet_snh_update_xpbd_fn: solve_kernels.py:1587-1617, 1429-1584 */
static void solve_snh(const orc_cons_t *cs, int c, orc_state_t *s, const orc_params_t *p, int thr) {
    float mu = cs->stiffness[c];
    if (mu <= 0.0f) return;
    const int *pts = cs->pts + 4 * c;
    int tetid = cs->tetid[c];
    float lam = cs->restdir[3 * c];
    float restvolume = cs->restlength[c];
    float kd = cs->dampingratio[c];
    float dt = p->dt;
    const float *B = s->rest_matrix + 9 * tetid;
    v3 P[4];
    for (int i = 0; i < 4; ++i) P[i] = ld3(s->pos, pts[i]);
    float w[4];
    for (int i = 0; i < 4; ++i) w[i] = inv_mass(s, pts[i]);
    v3 c0 = sub(P[0], P[3]), c1 = sub(P[1], P[3]), c2 = sub(P[2], P[3]);
    float F[9];
    ds_times_b(c0, c1, c2, B, F);
    float Ic = 0.0f;
    for (int i = 0; i < 9; ++i) Ic += F[i] * F[i];
    float r_s = sqrtf(Ic > 1.0e-12f ? Ic : 1.0e-12f);
    float dC = r_s - 1.7320508075688772f;
    if (r_s < 1.0e-6f) return;
    float inv_rs = 1.0f / r_s;
    float Fn[9];
    for (int i = 0; i < 9; ++i) Fn[i] = F[i] * inv_rs;
    /* Ht = B * Fn^T ; dgrad_k = row k of Ht */
    v3 dg[4];
    for (int k = 0; k < 3; ++k) {
        float r[3];
        for (int j = 0; j < 3; ++j)
            r[j] = B[3 * k + 0] * Fn[3 * j + 0] + B[3 * k + 1] * Fn[3 * j + 1] +
                   B[3 * k + 2] * Fn[3 * j + 2];
        dg[k] = mk(r[0], r[1], r[2]);
    }
    dg[3] = sub(sub(neg(dg[0]), dg[1]), dg[2]);
    float dsum_w = w[0] * dot(dg[0], dg[0]) + w[1] * dot(dg[1], dg[1]) + w[2] * dot(dg[2], dg[2]) +
                   w[3] * dot(dg[3], dg[3]);
    if (dsum_w == 0.0f) return;
    float alpha_snh = 1.0f + mu / lam;
    float J = det3(F);
    float hC = J - alpha_snh;
    float adm = fabsf(restvolume) * 6.0f;
    float inv_detDm = 1.0f / (adm > 1.0e-20f ? adm : 1.0e-20f);
    v3 hg[4];
    hg[0] = scl(inv_detDm, cross(c1, c2));
    hg[1] = scl(inv_detDm, cross(c2, c0));
    hg[2] = scl(inv_detDm, cross(c0, c1));
    hg[3] = sub(sub(neg(hg[0]), hg[1]), hg[2]);
    float hsum_w = w[0] * dot(hg[0], hg[0]) + w[1] * dot(hg[1], hg[1]) + w[2] * dot(hg[2], hg[2]) +
                   w[3] * dot(hg[3], hg[3]);
    if (hsum_w == 0.0f) return;
    float dhsum = w[0] * dot(dg[0], hg[0]) + w[1] * dot(dg[1], hg[1]) + w[2] * dot(dg[2], hg[2]) +
                  w[3] * dot(dg[3], hg[3]);
    float av = fabsf(restvolume);
    float abs_vol = av > 1.0e-20f ? av : 1.0e-20f;
    float dalpha = 1.0f / (mu * abs_vol * dt * dt);
    float halpha = 1.0f / (lam * abs_vol * dt * dt);
    float ddamp = 0.0f, dgamma = 1.0f, hdamp = 0.0f, hgamma = 1.0f;
    if (kd > 0.0f) {
        float dbeta = mu * abs_vol * kd * dt * dt;
        float hbeta = lam * abs_vol * kd * dt * dt;
        dgamma = dalpha * dbeta / dt;
        hgamma = halpha * hbeta / dt;
        ddamp = 0.0f;
        hdamp = 0.0f;
        for (int i = 0; i < 4; ++i) {
            v3 dv = sub(P[i], ld3(s->pprev, pts[i]));
            ddamp += dot(dg[i], dv);
        }
        for (int i = 0; i < 4; ++i) {
            v3 dv = sub(P[i], ld3(s->pprev, pts[i]));
            hdamp += dot(hg[i], dv);
        }
        ddamp = ddamp * dgamma;
        hdamp = hdamp * hgamma;
        dgamma = dgamma + 1.0f;
        hgamma = hgamma + 1.0f;
    }
    float dL = cs->L[3 * c], hL = cs->L[3 * c + 2];
    float Axx = dgamma * dsum_w + dalpha;
    float Ayy = hgamma * hsum_w + halpha;
    float det_A = Axx * Ayy - dhsum * dhsum;
    if (det_A < 1.0e-8f) return;
    float inv_det = 1.0f / det_A;
    float db = -dC - dalpha * dL - ddamp;
    float hb = -hC - halpha * hL - hdamp;
    float ddL = inv_det * (Ayy * db - dhsum * hb);
    float hdL = inv_det * (Axx * hb - dhsum * db);
    for (int i = 0; i < 4; ++i) {
        v3 dp = scl(w[i], add(scl(ddL, dg[i]), scl(hdL, hg[i])));
        if (p->use_jacobi) update_dP(s, dp, pts[i], thr);
        else st3(s->pos, pts[i], add(P[i], dp));
    }
    if (!p->use_jacobi) {
        cs->L[3 * c] = dL + ddL;
        cs->L[3 * c + 2] = hL + hdL;
    }
}

/* solve_tet_aniso_arap_kernel: solve_kernels.py:1310-1426 */
static void solve_aniso(const orc_cons_t *cs, int c, orc_state_t *s, const orc_params_t *p,
                        int thr) {
    float cst = cs->stiffness[c];
    if (cst <= 0.0f) return;
    const int *pts = cs->pts + 4 * c;
    int tetid = cs->tetid[c];
    v3 f0 = dvs(add(add(add(ld3(s->v_fiber_dir, pts[0]), ld3(s->v_fiber_dir, pts[1])),
                        ld3(s->v_fiber_dir, pts[2])),
                    ld3(s->v_fiber_dir, pts[3])),
                4.0f);
    float fn = length(f0);
    if (fn < 1.0e-9f) return;
    f0 = dvs(f0, fn);
    v3 P3 = ld3(s->pos, pts[3]);
    v3 c0 = sub(ld3(s->pos, pts[0]), P3), c1 = sub(ld3(s->pos, pts[1]), P3),
       c2 = sub(ld3(s->pos, pts[2]), P3);
    const float *B = s->rest_matrix + 9 * tetid;
    float F[9];
    ds_times_b(c0, c1, c2, B, F);
    v3 d = mk(F[0] * f0.x + F[1] * f0.y + F[2] * f0.z, F[3] * f0.x + F[4] * f0.y + F[5] * f0.z,
              F[6] * f0.x + F[7] * f0.y + F[8] * f0.z);
    float sl = length(d);
    if (sl < 1.0e-9f) return;
    float detF = det3(F);
    if (detF <= 0.0f || sl > 1.5f) return;
    float SI4 = detF < 0.0f ? -1.0f : 1.0f;
    float belly = 1.0f - s->tendonmask[tetid];
    float fiberscale = 1.0f - belly * s->activation[tetid] * p->contraction_ratio;
    float diff = sl - SI4 * fiberscale;
    float Cval = 0.5f * diff * diff;
    float gradscale = diff / sl;
    float G[9];
    const float dv[3] = {d.x, d.y, d.z}, fv[3] = {f0.x, f0.y, f0.z};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) G[3 * i + j] = gradscale * (dv[i] * fv[j]);
    v3 g[4];
    g[0] = grad_row(G, B, 0);
    g[1] = grad_row(G, B, 1);
    g[2] = grad_row(G, B, 2);
    g[3] = sub(sub(neg(g[0]), g[1]), g[2]);
    float w[4];
    for (int i = 0; i < 4; ++i) w[i] = inv_mass(s, pts[i]);
    float wsum = w[0] * dot(g[0], g[0]) + w[1] * dot(g[1], g[1]) + w[2] * dot(g[2], g[2]) +
                 w[3] * dot(g[3], g[3]);
    if (wsum < 1.0e-20f) return;
    float stiffness_val = cst * p->fiber_stiffness_scale;
    float alpha_tilde = (1.0f / stiffness_val) / (p->dt * p->dt);
    float L0 = cs->L[3 * c];
    float dL = -(Cval + alpha_tilde * L0) / (wsum + alpha_tilde);
    if (p->use_jacobi) {
        for (int i = 0; i < 4; ++i) update_dP(s, scl(dL * w[i], g[i]), pts[i], thr);
    } else {
        for (int i = 0; i < 4; ++i) gs_move(s, pts[i], scl(dL * w[i], g[i]));
        cs->L[3 * c] = L0 + dL;
    }
}

/* solve_attach_kernel + attach_bilateral_update_fn: solve_kernels.py:892-927, 462-526 */
static void solve_attach(const orc_cons_t *cs, int c, orc_state_t *s, const orc_params_t *p,
                         int thr) {
    float cst = cs->stiffness[c];
    if (cst <= 0.0f) return;
    float kcomp = cs->compressionstiffness[c];
    kcomp = kcomp >= 0.0f ? kcomp : cst;
    int pt1 = cs->pts[4 * c];
    v3 p0 = mk(cs->restvector[4 * c], cs->restvector[4 * c + 1], cs->restvector[4 * c + 2]);
    float invmass1 = inv_mass(s, pt1);
    float wsum = invmass1;
    if (wsum == 0.0f) return;
    v3 p1 = ld3(s->pos, pt1);
    v3 n = sub(p1, p0);
    float d = length(n);
    if (!(d >= 1e-6f)) return;
    float restlength = cs->restlength[c];
    int comp = 0;
    if (p->has_compress && d < restlength) comp = 1;
    float kv = comp ? kcomp : cst;
    int loff = comp < 2 ? comp : 2;
    if (kv == 0.0f) return;
    float l = cs->L[3 * c + loff];
    float alpha = 1.0f / kv;
    alpha = alpha / (p->dt * p->dt);
    float C = d - restlength;
    n = dvs(n, d);
    float dsum = 0.0f, gamma = 1.0f;
    float kd = cs->dampingratio[c];
    if (kd > 0.0f) {
        v3 prev1 = ld3(s->pprev, pt1);
        float beta = kv * kd * p->dt * p->dt;
        gamma = alpha * beta / p->dt;
        dsum = gamma * dot(n, sub(p1, prev1));
        gamma = gamma + 1.0f;
    }
    float dL = (-C - alpha * l - dsum) / (gamma * wsum + alpha);
    v3 dp = mk(n.x * (-dL), n.y * (-dL), n.z * (-dL));
    if (p->use_jacobi) {
        update_dP(s, scl(-invmass1, dp), pt1, thr);
    } else {
        st3(s->pos, pt1, sub(ld3(s->pos, pt1), scl(invmass1, dp)));
        cs->L[3 * c + loff] = l + dL;
    }
    if (cs->reaction_accum) {
        v3 r = scl(C, n);
        if (thr) {
#pragma omp atomic
            cs->reaction_accum[3 * c] += r.x;
#pragma omp atomic
            cs->reaction_accum[3 * c + 1] += r.y;
#pragma omp atomic
            cs->reaction_accum[3 * c + 2] += r.z;
        } else {
            cs->reaction_accum[3 * c] += r.x;
            cs->reaction_accum[3 * c + 1] += r.y;
            cs->reaction_accum[3 * c + 2] += r.z;
        }
    }
}

/* _dispatch_constraints: src/muscle/engine.py:890-1027.  The table is sorted by
 * type (musclesim.py:369-389), so walking it in order is the reference's
 * per-type-range dispatch order.  PINTOTARGET is a post-pass (apply_pintotarget). */
int orc_dispatch(const orc_cons_t *cs, orc_state_t *s, const orc_params_t *p, int nthreads) {
    int n = cs->n;
    for (int c = 0; c < n; ++c) {
        int t = cs->type[c];
        if (t != ORC_TETARAPSVD && t != ORC_TETVOLUME && t != ORC_TETSNH && t != ORC_TETARAPANISO &&
            t != ORC_ATTACH && t != ORC_PINTOTARGET)
            return ORC_EUNSUPPORTED;
    }
    int thr = (nthreads > 1 && p->use_jacobi) ? 1 : 0;
    int start = 0;
    while (start < n) {
        int t = cs->type[start];
        int end = start;
        while (end < n && cs->type[end] == t) ++end;
        if (thr) {
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
            for (int c = start; c < end; ++c) {
                switch (t) {
                    case ORC_TETARAPSVD: solve_arap_svd(cs, c, s, p, 1); break;
                    case ORC_TETVOLUME: solve_volume(cs, c, s, p, 1); break;
                    case ORC_TETSNH: solve_snh(cs, c, s, p, 1); break;
                    case ORC_TETARAPANISO: solve_aniso(cs, c, s, p, 1); break;
                    case ORC_ATTACH: solve_attach(cs, c, s, p, 1); break;
                    default: break;
                }
            }
        } else {
            for (int c = start; c < end; ++c) {
                switch (t) {
                    case ORC_TETARAPSVD: solve_arap_svd(cs, c, s, p, 0); break;
                    case ORC_TETVOLUME: solve_volume(cs, c, s, p, 0); break;
                    case ORC_TETSNH: solve_snh(cs, c, s, p, 0); break;
                    case ORC_TETARAPANISO: solve_aniso(cs, c, s, p, 0); break;
                    case ORC_ATTACH: solve_attach(cs, c, s, p, 0); break;
                    default: break;
                }
            }
        }
        start = end;
    }
    return ORC_OK;
}

/* apply_dP_kernel: src/muscle/engine.py:689-698 */
void orc_apply_dP(orc_state_t *s, int nthreads) {
    (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
#endif
    for (int i = 0; i < s->nv; ++i) {
        float w = s->dPw[i];
        if (w > 1e-9f) st3(s->pos, i, add(ld3(s->pos, i), dvs(ld3(s->dP, i), w)));
    }
}

/* pintotarget_overwrite_masked_kernel: src/muscle/mg/kernel.py:212-226 */
void orc_pintotarget_masked(int n, const int *pts, const float *target, const float *rest_offset,
                            const int *active, float *pos) {
    for (int i = 0; i < n; ++i) {
        if (active[i] == 0) continue;
        st3(pos, pts[i], add(ld3(target, i), ld3(rest_offset, i)));
    }
}

/* solve_pintotarget_kernel: solve_kernels.py:969-991 */
void orc_pintotarget_cons(const orc_cons_t *cs, float *pos) {
    for (int c = 0; c < cs->n; ++c) {
        if (cs->type[c] != ORC_PINTOTARGET) continue;
        int src = cs->pts[4 * c];
        if (src < 0) continue;
        const float *rv = cs->restvector + 4 * c, *rd = cs->restdir + 3 * c;
        st3(pos, src, mk(rv[0] + rd[0], rv[1] + rd[1], rv[2] + rd[2]));
    }
}

/* update_velocities_kernel: src/muscle/engine.py:582-605; skipped when quasi_static (:811-829) */
void orc_update_velocities(orc_state_t *s, const orc_params_t *p) {
    if (p->quasi_static) return;
    for (int i = 0; i < s->nv; ++i) {
        v3 v = dvs(sub(ld3(s->pos, i), ld3(s->pprev, i)), p->dt);
        if (p->const_damp != 0.0f || p->exp_damp != 1.0f) {
            float vmag = length(v);
            if (vmag > 1.0e-6f) {
                float vnew = vmag * p->exp_damp - p->const_damp;
                if (vnew < 1.0e-6f) vnew = 0.0f;
                st3(s->vel, i, scl(vnew / vmag, v));
            } else {
                st3(s->vel, i, mk(0, 0, 0));
            }
        } else {
            st3(s->vel, i, v);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Bone push-out on trilinear SDF grids.                                */
/* Push semantics follow solve_mesh_contact_pos_kernel                  */
/* (src/muscle/collision/sdf/kernel.py:123-170): skip verts with zero   */
/* collision weight; phi<offset -> x -= n*(phi-offset) with the outward */
/* normal; max_push>0 rejects deeper contacts.  The grid lookup is new: */
/* each bone b has a res^3 node grid (origin, spacing h) in its local   */
/* frame (spacing h_b) and a rigid transform xform[b] = [R | t] (3x4,  */
/* world = R*local + t).  The bone with the smallest phi wins (the      */
/* union of bones, as the reference queries one skeleton mesh).         */
/* cell = floor((R^T (x - t) - origin) * (1/h)), clamped to [0,res-2];  */
/* vgroup/goff (optional) restrict vertex i to bones [goff[g],goff[g+1]). */
/* points outside the node box are not in contact.                      */
/* ------------------------------------------------------------------ */
typedef struct { float phi; v3 n; int cell; int ok; } sdf_hit;

static sdf_hit sdf_grid_eval(const float *grid, int res, const float *origin, float inv_h,
                             const float *X, v3 x) {
    sdf_hit r;
    r.ok = 0; r.phi = 0; r.n = mk(0, 0, 0); r.cell = -1;
    v3 q = sub(x, mk(X[3], X[7], X[11]));
    /* local = R^T q */
    v3 y = mk(X[0] * q.x + X[4] * q.y + X[8] * q.z, X[1] * q.x + X[5] * q.y + X[9] * q.z,
              X[2] * q.x + X[6] * q.y + X[10] * q.z);
    float gx = (y.x - origin[0]) * inv_h, gy = (y.y - origin[1]) * inv_h,
          gz = (y.z - origin[2]) * inv_h;
    float lim = (float)(res - 1);
    if (!(gx >= 0.0f && gy >= 0.0f && gz >= 0.0f && gx <= lim && gy <= lim && gz <= lim)) return r;
    int ix = (int)floorf(gx), iy = (int)floorf(gy), iz = (int)floorf(gz);
    if (ix > res - 2) ix = res - 2;
    if (iy > res - 2) iy = res - 2;
    if (iz > res - 2) iz = res - 2;
    float fx = gx - (float)ix, fy = gy - (float)iy, fz = gz - (float)iz;
    const float *g0 = grid + ((size_t)iz * res + iy) * res + ix;
    size_t sy = (size_t)res, sz = (size_t)res * res;
    float c000 = g0[0], c100 = g0[1], c010 = g0[sy], c110 = g0[sy + 1];
    float c001 = g0[sz], c101 = g0[sz + 1], c011 = g0[sz + sy], c111 = g0[sz + sy + 1];
    float ux = 1.0f - fx, uy = 1.0f - fy, uz = 1.0f - fz;
    float c00 = c000 * ux + c100 * fx, c10 = c010 * ux + c110 * fx;
    float c01 = c001 * ux + c101 * fx, c11 = c011 * ux + c111 * fx;
    float c0 = c00 * uy + c10 * fy, c1 = c01 * uy + c11 * fy;
    r.phi = c0 * uz + c1 * fz;
    /* analytic gradient of the trilinear interpolant (grid units) */
    float dx = ((c100 - c000) * uy + (c110 - c010) * fy) * uz + ((c101 - c001) * uy + (c111 - c011) * fy) * fz;
    float dy = ((c010 - c000) * ux + (c110 - c100) * fx) * uz + ((c011 - c001) * ux + (c111 - c101) * fx) * fz;
    float dz = c1 - c0;
    /* world gradient direction = R * grad_local */
    v3 gw = mk(X[0] * dx + X[1] * dy + X[2] * dz, X[4] * dx + X[5] * dy + X[6] * dz,
               X[8] * dx + X[9] * dy + X[10] * dz);
    float gl = length(gw);
    if (!(gl > 1e-20f)) return r;
    r.n = dvs(gw, gl);
    r.cell = (iz * res + iy) * res + ix;
    r.ok = 1;
    return r;
}

int orc_sdf_grid_push(int nv, float *pos, const float *coll_w, int nbones, int res,
                      const float *grids, const float *origin, const float *spacing,
                      const float *xforms, const int *vgroup, const int *goff,
                      float contact_offset, float max_push, int *cell_out) {
    int pushed = 0;
    for (int i = 0; i < nv; ++i) {
        if (cell_out) { cell_out[2 * i] = -1; cell_out[2 * i + 1] = -1; }
        if (coll_w[i] == 0.0f) continue;
        v3 x = ld3(pos, i);
        sdf_hit best;
        memset(&best, 0, sizeof(best));
        int bb = -1;
        int b0 = 0, b1 = nbones;
        if (vgroup && goff) { b0 = goff[vgroup[i]]; b1 = goff[vgroup[i] + 1]; }
        for (int b = b0; b < b1; ++b) {
            float inv_h = 1.0f / spacing[b];
            sdf_hit hb = sdf_grid_eval(grids + (size_t)b * res * res * res, res, origin + 3 * b, inv_h,
                                       xforms + 12 * b, x);
            if (hb.ok && (!best.ok || hb.phi < best.phi)) { best = hb; bb = b; }
        }
        if (!best.ok) continue;
        if (cell_out) { cell_out[2 * i] = bb; cell_out[2 * i + 1] = best.cell; }
        float err = best.phi - contact_offset;
        if (err < 0.0f && (max_push <= 0.0f || -err <= max_push)) {
            st3(pos, i, sub(x, scl(err, best.n)));
            ++pushed;
        }
    }
    return pushed;
}

/* ------------------------------------------------------------------ */
/* residual metrics: arap_residual_accum_kernel (src/muscle/mg/kernel.py:87-118)
 * = sum over tets of ||sigma(F)-1||^2; plus the volume constraint sum of
 * (V - V0)^2 and the inverted count (count_inverted_tets_kernel,
 * src/muscle/engine.py:410-428).  Double accumulation (order-free).       */
/* ------------------------------------------------------------------ */
void orc_residuals(int nt, const int *tets, const float *pos, const float *rest_matrix,
                   const float *rest_volume, double *out_sumsq_arap, double *out_sumsq_vol,
                   int *out_inverted) {
    double sa = 0, sv = 0;
    int inv = 0;
    for (int c = 0; c < nt; ++c) {
        const int *pts = tets + 4 * c;
        v3 P3 = ld3(pos, pts[3]);
        v3 c0 = sub(ld3(pos, pts[0]), P3), c1 = sub(ld3(pos, pts[1]), P3), c2 = sub(ld3(pos, pts[2]), P3);
        float F[9], U[9], sg[3], V[9];
        ds_times_b(c0, c1, c2, rest_matrix + 9 * c, F);
        if (det3(F) <= 0.0f) ++inv;
        orc_svd3(F, U, sg, V);
        float a = (sg[0] - 1.0f) * (sg[0] - 1.0f) + (sg[1] - 1.0f) * (sg[1] - 1.0f) +
                  (sg[2] - 1.0f) * (sg[2] - 1.0f);
        sa += a;
        if (rest_volume) {
            float vol = dot(cross(c0, c1), c2) / 6.0f;
            double dv = (double)vol - (double)rest_volume[c];
            sv += dv * dv;
        }
    }
    if (out_sumsq_arap) *out_sumsq_arap = sa;
    if (out_sumsq_vol) *out_sumsq_vol = sv;
    if (out_inverted) *out_inverted = inv;
}

int orc_count_inverted(int nt, const int *tets, const float *pos, const float *rest_matrix) {
    int cnt = 0;
    for (int c = 0; c < nt; ++c) {
        const int *pts = tets + 4 * c;
        v3 P3 = ld3(pos, pts[3]);
        float F[9];
        ds_times_b(sub(ld3(pos, pts[0]), P3), sub(ld3(pos, pts[1]), P3), sub(ld3(pos, pts[2]), P3),
                   rest_matrix + 9 * c, F);
        if (det3(F) <= 0.0f) ++cnt;
    }
    return cnt;
}

/* ================================================================== */
/* tri-tri DCD: src/muscle/collision/udcf/kernel.py                    */
/* ================================================================== */

/* centroid_radius_kernel :47-63, face_normal_kernel :66-84 */
void orc_dcd_tri_data(int m, const int *tris, const float *x, float *centroids, float *radii,
                      float *normals) {
    for (int i = 0; i < m; ++i) {
        v3 a = ld3(x, tris[3 * i]), b = ld3(x, tris[3 * i + 1]), c = ld3(x, tris[3 * i + 2]);
        v3 ce = dvs(add(add(a, b), c), 3.0f);
        float r0 = length(sub(a, ce)), r1 = length(sub(b, ce)), r2 = length(sub(c, ce));
        float r = r0 > r1 ? r0 : r1;
        r = r > r2 ? r : r2;
        st3(centroids, i, ce);
        radii[i] = r;
        if (normals) {
            v3 n = cross(sub(b, a), sub(c, a));
            float ln = length(n);
            st3(normals, i, ln > 0.0f ? dvs(n, ln) : mk(0, 0, 0));
        }
    }
}

/* broadphase_find_pairs_kernel :851-915 emit rules, enumerated exhaustively */
long orc_dcd_pairs_bruteforce(int m, const int *tris, const float *centroids, const float *radii,
                              const float *normals, const int *group, float normal_cull,
                              int allow_self, float thickness, long max_pairs, int *pair_a,
                              int *pair_b) {
    long cnt = 0;
    for (int i = 0; i < m; ++i) {
        v3 ci = ld3(centroids, i);
        float ri = radii[i];
        int gi = group ? group[i] : 0;
        for (int j = i + 1; j < m; ++j) {
            int gj = group ? group[j] : 0;
            if (gi == 1 && gj == 1) continue;
            if (gi == 0 && gj == 0) {
                if (!allow_self) continue;
                const int *a = tris + 3 * i, *b = tris + 3 * j;
                int share = 0;
                for (int u = 0; u < 3; ++u)
                    for (int v = 0; v < 3; ++v) share |= (a[u] == b[v]);
                if (share) continue;
                if (normals && dot(ld3(normals, i), ld3(normals, j)) > normal_cull) continue;
            }
            if (length(sub(ci, ld3(centroids, j))) > ri + radii[j] + thickness) continue;
            if (cnt < max_pairs) { pair_a[cnt] = i; pair_b[cnt] = j; }
            ++cnt;
        }
    }
    return cnt;
}

/* geometry helpers of narrowPhase (kernel.py:165-309) */
static v3 pick3(v3 a, v3 b, v3 c, int i) { return i == 0 ? a : (i == 1 ? b : c); }

static int inside_tri(v3 p1, v3 p2, v3 p3, v3 p) {
    v3 c1 = cross(sub(p2, p1), sub(p, p2));
    v3 c2 = cross(sub(p3, p2), sub(p, p3));
    v3 c3 = cross(sub(p1, p3), sub(p, p1));
    return (dot(c1, c2) > 0.0f && dot(c1, c3) > 0.0f) ? 1 : 0;
}

static v3 bary(v3 A, v3 B, v3 C, v3 P) {
    v3 v0 = sub(B, A), v1 = sub(C, A), v2 = sub(P, A);
    float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    float denom = d00 * d11 - d01 * d01;
    if (denom == 0.0f) return mk(1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f);
    float v = (d11 * d20 - d01 * d21) / denom;
    float w = (d00 * d21 - d01 * d20) / denom;
    return mk(1.0f - v - w, v, w);
}

static int line_isect(v3 p1, v3 p2, v3 p3, v3 p4, v3 v1, v3 v2, v3 *out) {
    v3 d1 = sub(p2, p1), d2 = sub(p4, p3);
    v3 nrm = cross(d1, d2);
    if (length(nrm) < 1.0e-6f) return 0;
    float denom = dot(nrm, nrm);
    if (denom < 1.0e-6f) return 0;
    v3 v = sub(p3, p1);
    float t1 = dot(cross(v, d2), nrm) / denom;
    float t2 = dot(cross(v, d1), nrm) / denom;
    if (t2 >= 0.0f && t2 <= 1.0f && t1 >= 0.0f && t1 <= 1.0f) {
        *out = add(v1, scl(t1, sub(v2, v1)));
        return 1;
    }
    return 0;
}

static int find_isect(v3 N, v3 p0, v3 p, v3 dir, v3 *out) {
    float denom = dot(N, dir);
    if (denom == 0.0f) return 0;
    float t = dot(N, sub(p0, p)) / denom;
    *out = add(p, scl(t, dir));
    return 1;
}

#define UDCF_THRESH 1.0e-5f

static int gen_t(v3 N, v3 v0, v3 v1, v3 v2, float d, v3 D, float *lo, float *hi) {
    float d0 = dot(N, v0) + d, d1 = dot(N, v1) + d, d2 = dot(N, v2) + d;
    if (d0 <= UDCF_THRESH && d1 <= UDCF_THRESH && d2 <= UDCF_THRESH) return 0;
    if (d0 >= -UDCF_THRESH && d1 >= -UDCF_THRESH && d2 >= -UDCF_THRESH) return 0;
    v3 s0 = v0, s1 = v1, s2 = v2;
    float e0 = d0, e1 = d1, e2 = d2;
    if ((d0 <= 0.0f && d1 >= 0.0f && d2 >= 0.0f) || (d0 >= 0.0f && d1 <= 0.0f && d2 <= 0.0f)) {
        s0 = v1; s1 = v0; e0 = d1; e1 = d0;
    } else if ((d0 >= 0.0f && d1 >= 0.0f && d2 <= 0.0f) || (d0 <= 0.0f && d1 <= 0.0f && d2 >= 0.0f)) {
        s1 = v2; s2 = v1; e1 = d2; e2 = d1;
    }
    float p0 = dot(D, s0), p1 = dot(D, s1), p2 = dot(D, s2);
    float tA = p0 + (p1 - p0) * fabsf(e0 / (e0 - e1));
    float tB = p2 + (p1 - p2) * fabsf(e2 / (e2 - e1));
    if (tA > tB) { *lo = tB; *hi = tA; } else { *lo = tA; *hi = tB; }
    return 1;
}

/* narrow_kernel :317-474 */
void orc_dcd_narrow(long npairs, const int *pair_a, const int *pair_b, const int *tris,
                    const float *x, float offset_eps, orc_contact_t *out) {
    for (long k = 0; k < npairs; ++k) {
        orc_contact_t *o = out + k;
        memset(o, 0, sizeof(*o));
        int ta = pair_a[k], tb = pair_b[k];
        v3 q00 = ld3(x, tris[3 * ta]), q01 = ld3(x, tris[3 * ta + 1]), q02 = ld3(x, tris[3 * ta + 2]);
        v3 q10 = ld3(x, tris[3 * tb]), q11 = ld3(x, tris[3 * tb + 1]), q12 = ld3(x, tris[3 * tb + 2]);
        v3 n1f = cross(sub(q01, q00), sub(q02, q00)), n2f = cross(sub(q11, q10), sub(q12, q10));
        float l1 = length(n1f), l2 = length(n2f);
        if (l1 == 0.0f || l2 == 0.0f) continue;
        v3 N1 = dvs(n1f, l1), N2 = dvs(n2f, l2);
        float d1 = -dot(N1, q00), d2 = -dot(N2, q10);
        v3 cp = cross(N1, N2);
        float lcp = length(cp);
        if (lcp < 1.0e-8f) continue;
        v3 D = dvs(cp, lcp);
        float t1, t2, t3, t4;
        if (!gen_t(N2, q00, q01, q02, d2, D, &t1, &t2)) continue;
        if (!gen_t(N1, q10, q11, q12, d1, D, &t3, &t4)) continue;
        if (!(t2 >= t3 && t4 >= t1)) continue;
        float C1 = 0.0f, C2 = 0.0f;
        v3 b1 = mk(0, 0, 0), b2 = b1, b3 = b1, b4 = b1;
        for (int j = 0; j < 3; ++j) {
            v3 q0j = pick3(q00, q01, q02, j), q1j = pick3(q10, q11, q12, j);
            v3 P0j = sub(q0j, scl(dot(N2, sub(q0j, q10)), N2));
            if (inside_tri(q10, q11, q12, P0j)) {
                float cand = dot(N2, q0j) + d2;
                if (C1 < -cand) {
                    C1 = -cand;
                    b1 = bary(q00, q01, q02, q0j);
                    b2 = bary(q10, q11, q12, sub(q0j, scl(cand, N2)));
                }
            }
            v3 Q0j;
            if (find_isect(N1, q00, q1j, N2, &Q0j) && inside_tri(q00, q01, q02, Q0j)) {
                float cand = dot(N2, Q0j) + d2;
                if (C1 < -cand) {
                    C1 = -cand;
                    b1 = bary(q00, q01, q02, Q0j);
                    b2 = bary(q10, q11, q12, sub(Q0j, scl(cand, N2)));
                }
            }
            v3 P1j = sub(q1j, scl(dot(N1, sub(q1j, q00)), N1));
            if (inside_tri(q00, q01, q02, P1j)) {
                float cand = dot(N1, q1j) + d1;
                if (C2 < -cand) {
                    C2 = -cand;
                    b3 = bary(q10, q11, q12, q1j);
                    b4 = bary(q00, q01, q02, sub(q1j, scl(cand, N1)));
                }
            }
            v3 Q1j;
            if (find_isect(N2, q10, q0j, N1, &Q1j) && inside_tri(q10, q11, q12, Q1j)) {
                float cand = dot(N1, Q1j) + d1;
                if (C2 < -cand) {
                    C2 = -cand;
                    b3 = bary(q10, q11, q12, Q1j);
                    b4 = bary(q00, q01, q02, sub(Q1j, scl(cand, N1)));
                }
            }
            int jn = (j + 1) % 3;
            v3 q0jn = pick3(q00, q01, q02, jn), q1jn = pick3(q10, q11, q12, jn);
            v3 P0jn = sub(q0jn, scl(dot(N2, sub(q0jn, q10)), N2));
            v3 P1jn = sub(q1jn, scl(dot(N1, sub(q1jn, q00)), N1));
            for (int kk = 0; kk < 3; ++kk) {
                int kn = (kk + 1) % 3;
                v3 q1k = pick3(q10, q11, q12, kk), q1kn = pick3(q10, q11, q12, kn);
                v3 hit;
                if (line_isect(P0j, P0jn, q1k, q1kn, q0j, q0jn, &hit)) {
                    float cand = dot(N2, hit) + d2;
                    if (C1 < -cand) {
                        C1 = -cand;
                        b1 = bary(q00, q01, q02, hit);
                        b2 = bary(q10, q11, q12, sub(hit, scl(cand, N2)));
                    }
                }
                v3 q0k = pick3(q00, q01, q02, kk), q0kn = pick3(q00, q01, q02, kn);
                if (line_isect(P1j, P1jn, q0k, q0kn, q1j, q1jn, &hit)) {
                    float cand = dot(N1, hit) + d1;
                    if (C2 < -cand) {
                        C2 = -cand;
                        b3 = bary(q10, q11, q12, hit);
                        b4 = bary(q00, q01, q02, sub(hit, scl(cand, N1)));
                    }
                }
            }
        }
        if (C1 <= offset_eps || C2 <= offset_eps) continue;
        o->valid = 1;
        o->C1 = C1;
        o->C2 = C2;
        if (C1 < C2) {
            for (int j = 0; j < 3; ++j) { o->cur[j] = tris[3 * ta + j]; o->closest[j] = tris[3 * tb + j]; }
            o->C = C1;
            o->n[0] = N2.x; o->n[1] = N2.y; o->n[2] = N2.z;
            o->b1[0] = b1.x; o->b1[1] = b1.y; o->b1[2] = b1.z;
            o->b2[0] = b2.x; o->b2[1] = b2.y; o->b2[2] = b2.z;
        } else {
            for (int j = 0; j < 3; ++j) { o->cur[j] = tris[3 * tb + j]; o->closest[j] = tris[3 * ta + j]; }
            o->C = C2;
            o->n[0] = N1.x; o->n[1] = N1.y; o->n[2] = N1.z;
            o->b1[0] = b3.x; o->b1[1] = b3.y; o->b1[2] = b3.z;
            o->b2[0] = b4.x; o->b2[1] = b4.y; o->b2[2] = b4.z;
        }
    }
}

/* respond_kernel :714-799 (serial accumulation, contacts in pair order) */
void orc_dcd_respond(long ncontacts, const orc_contact_t *cs, const float *x, const float *x_prev,
                     const float *inv_mass, float multiplier, float mu_s, float mu_d, float *dP,
                     float *dPw) {
    for (long k = 0; k < ncontacts; ++k) {
        const orc_contact_t *c = cs + k;
        if (!c->valid) continue;
        float pen = c->C;
        float C = pen * multiplier;
        if (C == 0.0f) continue;
        v3 nrm = mk(c->n[0], c->n[1], c->n[2]);
        const int *cu = c->cur, *cl = c->closest;
        int solid_cur = (inv_mass[cu[0]] == 0.0f || inv_mass[cu[1]] == 0.0f || inv_mass[cu[2]] == 0.0f);
        int solid_clo = (inv_mass[cl[0]] == 0.0f || inv_mass[cl[1]] == 0.0f || inv_mass[cl[2]] == 0.0f);
        if (!solid_cur && !solid_clo) C = C * 0.5f;
        v3 push_clo = solid_clo ? mk(0, 0, 0) : scl(-C, nrm);
        v3 push_cur = solid_cur ? mk(0, 0, 0) : scl(C, nrm);
        v3 v0 = sub(add(ld3(x, cl[0]), push_clo), ld3(x_prev, cl[0]));
        v3 v1 = sub(add(ld3(x, cl[1]), push_clo), ld3(x_prev, cl[1]));
        v3 v2 = sub(add(ld3(x, cl[2]), push_clo), ld3(x_prev, cl[2]));
        v3 v3_ = sub(add(ld3(x, cu[0]), push_cur), ld3(x_prev, cu[0]));
        v3 v4 = sub(add(ld3(x, cu[1]), push_cur), ld3(x_prev, cu[1]));
        v3 v5 = sub(add(ld3(x, cu[2]), push_cur), ld3(x_prev, cu[2]));
        v3 g = dvs(nrm, length(nrm));
        v3 vclo = add(add(scl(c->b1[0], v0), scl(c->b1[1], v1)), scl(c->b1[2], v2));
        v3 perp_clo = sub(vclo, scl(dot(vclo, g), g));
        v3 vcur = add(add(scl(c->b2[0], v3_), scl(c->b2[1], v4)), scl(c->b2[2], v5));
        v3 perp_cur = sub(vcur, scl(dot(vcur, g), g));
        v3 diff = sub(perp_cur, perp_clo);
        float dn = length(diff);
        float static_norm = pen * mu_s;
        if (dn > static_norm && dn > 0.0f) diff = scl(pen * mu_d / dn, diff);
        v3 diffV = scl(0.5f, diff);
        v3 diffT = sub(diff, diffV);
        v3 dcur = sub(push_cur, diffV), dclo = add(push_clo, diffT);
        for (int j = 0; j < 3; ++j) {
            if (inv_mass[cu[j]] > 0.0f) {
                dP[3 * cu[j]] += dcur.x; dP[3 * cu[j] + 1] += dcur.y; dP[3 * cu[j] + 2] += dcur.z;
                dPw[cu[j]] += 1.0f;
            }
        }
        for (int j = 0; j < 3; ++j) {
            if (inv_mass[cl[j]] > 0.0f) {
                dP[3 * cl[j]] += dclo.x; dP[3 * cl[j] + 1] += dclo.y; dP[3 * cl[j] + 2] += dclo.z;
                dPw[cl[j]] += 1.0f;
            }
        }
    }
}

/* apply_kernel :802-815 */
void orc_dcd_apply(int n, float *pos, const float *dP, const float *dPw, int use_avg) {
    for (int i = 0; i < n; ++i) {
        float w = dPw[i];
        if (w > 0.0f) {
            v3 d = ld3(dP, i);
            st3(pos, i, add(ld3(pos, i), use_avg ? dvs(d, w) : d));
        }
    }
}

/* ------------------------------------------------------------------ */
/* exact-mesh bone collider                                            */
/*                                                                     */
/* solve_mesh_contact_pos_kernel (src/muscle/collision/sdf/kernel.py:  */
/* 123-170) and the penetration meter _mark_inside_kernel              */
/* (src/example/humanV2.py:67-87) query the bone mesh through Warp     */
/* 1.13.0's mesh_query_point_sign_normal (warp-lang, uv.lock:831-833), */
/* which is not vendored: its documented behaviour is restated here.   */
/* The closest point on the mesh among the faces nearer than max_dist  */
/* (none: no result), found by brute force over every triangle, in     */
/* double precision; the inside/outside sign is that of the            */
/* angle-weighted pseudonormal of the closest feature (Baerentzen and  */
/* Aanaes, "Signed distance computation using the angle weighted       */
/* pseudonormal", IEEE TVCG 11(3), 2005): the face normal inside a     */
/* face, the sum of the two face normals on an edge, the sum of the    */
/* incident face normals weighted by their corner angles at a vertex.  */
/* Closest point on a triangle: Ericson, Real-Time Collision Detection */
/* (2005), section 5.1.5 (Voronoi regions of the vertices, edges and   */
/* face).                                                              */
/* ------------------------------------------------------------------ */
typedef struct { double x, y, z; } d3;
static inline d3 dmk(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 dld(const float *a, int i) { return dmk(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
static inline d3 dsub(d3 a, d3 b) { return dmk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 dadd(d3 a, d3 b) { return dmk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 dscl(double s, d3 a) { return dmk(s * a.x, s * a.y, s * a.z); }
static inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 dcross(d3 a, d3 b) {
    return dmk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* closest point of segment a-b to p; *t = parameter */
static d3 seg_closest(d3 p, d3 a, d3 b, double *t) {
    d3 ab = dsub(b, a);
    double l2 = ddot(ab, ab), s = 0.0;
    if (l2 > 0.0) {
        s = ddot(dsub(p, a), ab) / l2;
        s = s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
    }
    *t = s;
    return dadd(a, dscl(s, ab));
}

/* region: 0 face, 1 2 3 vertex a b c, 4 edge ab, 5 edge bc, 6 edge ca */
static d3 tri_closest(d3 p, d3 a, d3 b, d3 c, int *region) {
    d3 ab = dsub(b, a), ac = dsub(c, a), ap = dsub(p, a);
    double d1 = ddot(ab, ap), d2 = ddot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { *region = 1; return a; }
    d3 bp = dsub(p, b);
    double d3_ = ddot(ab, bp), d4 = ddot(ac, bp);
    if (d3_ >= 0.0 && d4 <= d3_) { *region = 2; return b; }
    double vc = d1 * d4 - d3_ * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3_ <= 0.0) { *region = 4; return dadd(a, dscl(d1 / (d1 - d3_), ab)); }
    d3 cp = dsub(p, c);
    double d5 = ddot(ab, cp), d6 = ddot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) { *region = 3; return c; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) { *region = 6; return dadd(a, dscl(d2 / (d2 - d6), ac)); }
    double va = d3_ * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3_) >= 0.0 && (d5 - d6) >= 0.0) {
        *region = 5;
        return dadd(b, dscl((d4 - d3_) / ((d4 - d3_) + (d5 - d6)), dsub(c, b)));
    }
    double den = va + vb + vc;
    if (!(den > 0.0)) {          /* zero-area triangle: nearest of its three edges */
        double t;
        d3 q[3];
        q[0] = seg_closest(p, a, b, &t);
        q[1] = seg_closest(p, b, c, &t);
        q[2] = seg_closest(p, c, a, &t);
        int k = 0;
        double best = ddot(dsub(p, q[0]), dsub(p, q[0]));
        for (int j = 1; j < 3; ++j) {
            double e = ddot(dsub(p, q[j]), dsub(p, q[j]));
            if (e < best) { best = e; k = j; }
        }
        *region = 4 + k;
        return q[k];
    }
    *region = 0;
    return dadd(a, dadd(dscl(vb / den, ab), dscl(vc / den, ac)));
}

/* unit face normals (ntris x 3) and angle-weighted vertex pseudonormals (nbv x 3) */
void orc_mesh_normals(int nbv, const float *bv, int ntris, const int *tris, double *fn, double *vn) {
    memset(vn, 0, sizeof(double) * 3 * (size_t)nbv);
    for (int f = 0; f < ntris; ++f) {
        const int *t = tris + 3 * f;
        d3 P[3] = {dld(bv, t[0]), dld(bv, t[1]), dld(bv, t[2])};
        d3 n = dcross(dsub(P[1], P[0]), dsub(P[2], P[0]));
        double l = sqrt(ddot(n, n));
        n = l > 0.0 ? dscl(1.0 / l, n) : dmk(0, 0, 0);
        fn[3 * f] = n.x; fn[3 * f + 1] = n.y; fn[3 * f + 2] = n.z;
        for (int k = 0; k < 3; ++k) {
            d3 e1 = dsub(P[(k + 1) % 3], P[k]), e2 = dsub(P[(k + 2) % 3], P[k]);
            double l1 = sqrt(ddot(e1, e1)), l2 = sqrt(ddot(e2, e2));
            double cang = (l1 > 0.0 && l2 > 0.0) ? ddot(e1, e2) / (l1 * l2) : 1.0;
            cang = cang > 1.0 ? 1.0 : (cang < -1.0 ? -1.0 : cang);
            double ang = acos(cang);
            double *v = vn + 3 * t[k];
            v[0] += ang * n.x; v[1] += ang * n.y; v[2] += ang * n.z;
        }
    }
}

/* closest point over triangles [t0, t1) nearer than sqrt(*best_d2); returns the face or -1 */
static int mesh_scan(d3 p, const float *bv, const int *tris, int t0, int t1, double *best_d2, d3 *best_cp,
                     int *best_region) {
    int face = -1;
    for (int f = t0; f < t1; ++f) {
        const int *t = tris + 3 * f;
        int region;
        d3 q = tri_closest(p, dld(bv, t[0]), dld(bv, t[1]), dld(bv, t[2]), &region);
        d3 d = dsub(p, q);
        double e = ddot(d, d);
        if (e < *best_d2) { *best_d2 = e; *best_cp = q; *best_region = region; face = f; }
    }
    return face;
}

static double feature_sign(d3 p, d3 cp, int face, int region, const int *tris, const double *fn,
                           const double *vn, const int *adj) {
    const double *nf = fn + 3 * face;
    d3 N = dmk(nf[0], nf[1], nf[2]);
    if (region >= 1 && region <= 3) {
        const double *v = vn + 3 * tris[3 * face + region - 1];
        N = dmk(v[0], v[1], v[2]);
    } else if (region >= 4) {
        /* edge ab = 0, bc = 1, ca = 2 */
        int g = adj ? adj[3 * face + (region - 4)] : -1;
        if (g >= 0) N = dadd(N, dmk(fn[3 * g], fn[3 * g + 1], fn[3 * g + 2]));
    }
    return ddot(dsub(p, cp), N) < 0.0 ? -1.0 : 1.0;
}

/* query nq points; bones b0..b1 of the point's group (vgroup/goff NULL: every bone);
 * out: sd (signed distance, < 0 inside; 0 without a result), cp (3 per point), face
 * (-1: no face nearer than max_dist), region */
void orc_mesh_query(int nq, const float *q, int nbv, const float *bv, int ntris, const int *tris,
                    int nbones, const int *tri_off, const double *fn, const double *vn, const int *adj,
                    const int *vgroup, const int *goff, float max_dist, double *sd, double *cp, int *face,
                    int *region, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < nq; ++i) {
        d3 p = dld(q, i);
        double best = (double)max_dist * (double)max_dist;
        d3 bcp = dmk(0, 0, 0);
        int breg = -1, bf = -1;
        int b0 = 0, b1 = nbones;
        if (vgroup && goff) { b0 = goff[vgroup[i]]; b1 = goff[vgroup[i] + 1]; }
        for (int b = b0; b < b1; ++b) {
            int f = mesh_scan(p, bv, tris, tri_off[b], tri_off[b + 1], &best, &bcp, &breg);
            if (f >= 0) bf = f;
        }
        face[i] = bf;
        region[i] = breg;
        if (bf < 0) {
            sd[i] = 0.0;
            cp[3 * i] = cp[3 * i + 1] = cp[3 * i + 2] = 0.0;
            continue;
        }
        cp[3 * i] = bcp.x; cp[3 * i + 1] = bcp.y; cp[3 * i + 2] = bcp.z;
        sd[i] = sqrt(best) * feature_sign(p, bcp, bf, breg, tris, fn, vn, adj);
    }
}

/* the push of solve_mesh_contact_pos_kernel (kernel.py:155-170) from the query above:
 * vertices with coll_w == 0 are skipped; returns the number pushed */
int orc_mesh_push(int nv, float *pos, const float *coll_w, const double *sd, const double *cp, const int *face,
                  float contact_offset, float max_push) {
    int pushed = 0;
    for (int i = 0; i < nv; ++i) {
        if (coll_w && coll_w[i] == 0.0f) continue;
        if (face[i] < 0) continue;
        d3 x = dld(pos, i);
        d3 delta = dsub(x, dmk(cp[3 * i], cp[3 * i + 1], cp[3 * i + 2]));
        double dlen = sqrt(ddot(delta, delta));
        if (dlen < 1.0e-9) continue;
        double sgn = sd[i] < 0.0 ? -1.0 : 1.0;
        double err = dlen * sgn - (double)contact_offset;
        if (err < 0.0 && (max_push <= 0.0f || -err <= (double)max_push)) {
            d3 n = dscl(sgn / dlen, delta);
            d3 y = dsub(x, dscl(err, n));
            pos[3 * i] = (float)y.x; pos[3 * i + 1] = (float)y.y; pos[3 * i + 2] = (float)y.z;
            ++pushed;
        }
    }
    return pushed;
}

/* the fascia bind's closest point (fascia_bind_kernel, src/muscle/fascia/fascia.py:135-157):
 * for each point the face of the triangle mesh nearest to it among the faces nearer than
 * max_dist (brute force, double); at near-ties (within 1e-6 relative of the nearest distance)
 * the lowest triangle index, the rule the library's bind uses as well; face -1 = none */
void orc_mesh_bind(int nq, const float *q, const float *bv, int ntris, const int *tris, float max_dist,
                   int *face, double *cp, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < nq; ++i) {
        d3 p = dld(q, i);
        double md2 = (double)max_dist * (double)max_dist, best = INFINITY;
        for (int f = 0; f < ntris; ++f) {
            int reg;
            const int *t = tris + 3 * f;
            d3 c = tri_closest(p, dld(bv, t[0]), dld(bv, t[1]), dld(bv, t[2]), &reg);
            d3 d = dsub(p, c);
            double e = ddot(d, d);
            if (e < md2 && e < best) best = e;
        }
        face[i] = -1;
        if (!(best < INFINITY)) continue;
        for (int f = 0; f < ntris; ++f) {
            int reg;
            const int *t = tris + 3 * f;
            d3 c = tri_closest(p, dld(bv, t[0]), dld(bv, t[1]), dld(bv, t[2]), &reg);
            d3 d = dsub(p, c);
            double e = ddot(d, d);
            if (e < md2 && e <= best * (1.0 + 2e-6) + 1e-14) {
                face[i] = f;
                cp[3 * i] = c.x; cp[3 * i + 1] = c.y; cp[3 * i + 2] = c.z;
                break;
            }
        }
    }
}
