// MI355X (gfx950) XPBD muscle substep: kernels + C ABI (include/xpbd.h).
//
// Hot loop (block strategy, the autotuner's pick on every workload here): per
// Jacobi iteration, k_block_tets gives each workgroup a block of consecutive
// Morton-ordered tet elements (ARAP-SVD / volume / SNH / aniso fused per tet),
// stages the vertices they touch in LDS, projects one element per thread and
// parks the four corner corrections in LDS; every local vertex then sums its
// corners in a fixed (element, corner) order.  A vertex touched by this block
// only is finished in place (dP/dPw, attach, pins); a shared one gets one
// partial per block, which k_block_shared adds in ascending block order.  No
// float atomics, so a run repeats bit for bit.  The patch (k_iterate) and flat
// (k_flat_tets + k_flat_verts) strategies are the alternatives the autotuner
// weighs.  Same semantics as the reference per-type kernels + update_dP atomics
// + apply_dP + pin overwrite (src/muscle/constraint/solve_kernels.py,
// src/muscle/engine.py:689-698, src/muscle/mg/kernel.py:212-226).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xpbd.h"
#include "xpbd_dcd.h"
#include "xpbd_kernels.h"
#include "xpbd_layout.h"

using namespace xpbd;

// ---------------------------------------------------------------------------
// small FP32 vector helpers
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 mk3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 xyz(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float det3(const float *m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}
// gradient of corner k: G * B[k,:]^T  (solve_kernels.py:1269-1281)
__device__ __forceinline__ V3 grad_row(const float *G, const float *B, int k) {
    return V3{G[0] * B[3 * k] + G[1] * B[3 * k + 1] + G[2] * B[3 * k + 2],
              G[3] * B[3 * k] + G[4] * B[3 * k + 1] + G[5] * B[3 * k + 2],
              G[6] * B[3 * k] + G[7] * B[3 * k + 1] + G[8] * B[3 * k + 2]};
}

// ---------------------------------------------------------------------------
// Closest proper rotation R of F (the rotation of the sign-fixed SVD,
// R = U V^T, solve_kernels.py:85-132).  ARAP-SVD's C = ||sigma - 1|| equals
// ||F - R||_F and dC/dF = (F - R) / C, so only R is needed.
//
// det F > 0: near rest the Bjorck-Bowie (Newton-Schulz) iteration below,
// otherwise Higham's scaled Newton iteration X <- (g X + X^{-T}/g)/2 with
// Frobenius-norm scaling (N. J. Higham, "Computing the polar decomposition —
// with applications", SIAM J. Sci. Stat. Comput. 7, 1986); both quadratically
// convergent.  Otherwise (inverted / degenerate, or no convergence) a
// one-sided Jacobi SVD gives R = U diag(1,1,det U det V) V^T.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool polar_newton(const float *F, float *X) {
#pragma unroll
    for (int i = 0; i < 9; ++i) X[i] = F[i];
    bool scale = true;
    for (int it = 0; it < 20; ++it) {
        const float C0 = X[4] * X[8] - X[5] * X[7], C1 = X[5] * X[6] - X[3] * X[8],
                    C2 = X[3] * X[7] - X[4] * X[6];
        const float C3 = X[2] * X[7] - X[1] * X[8], C4 = X[0] * X[8] - X[2] * X[6],
                    C5 = X[1] * X[6] - X[0] * X[7];
        const float C6 = X[1] * X[5] - X[2] * X[4], C7 = X[2] * X[3] - X[0] * X[5],
                    C8 = X[0] * X[4] - X[1] * X[3];
        const float det = X[0] * C0 + X[1] * C1 + X[2] * C2;
        if (!(det > 1.0e-30f)) return false;
        const float idet = __builtin_amdgcn_rcpf(det);
        float g = 1.0f, ig = 1.0f;
        if (scale) {
            float nx = 0.f;
#pragma unroll
            for (int i = 0; i < 9; ++i) nx += X[i] * X[i];
            const float nc = C0 * C0 + C1 * C1 + C2 * C2 + C3 * C3 + C4 * C4 + C5 * C5 + C6 * C6 +
                             C7 * C7 + C8 * C8;
            // g = (||X^-1|| / ||X||)^(1/2) = (nc / nx)^(1/4) / sqrt(det)
            g = __builtin_amdgcn_sqrtf(__builtin_amdgcn_sqrtf(nc * __builtin_amdgcn_rcpf(nx)) * idet);
            ig = __builtin_amdgcn_rcpf(g);
        }
        const float a = 0.5f * g, b = 0.5f * idet * ig;
        const float C[9] = {C0, C1, C2, C3, C4, C5, C6, C7, C8};
        float delta = 0.f;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const float xn = a * X[i] + b * C[i];
            const float d = xn - X[i];
            delta += d * d;
            X[i] = xn;
        }
        if (delta < 1.0e-4f) scale = false;
        // near convergence the next error is about delta / 2 (quadratic
        // convergence, delta = |X_k+1 - X_k|^2 ~ error_k^2): stop once that is
        // below FP32 resolution of a rotation
        if (delta < 1.0e-7f) return true;
    }
    return false;
}

// Near rest, Bjorck and Bowie's orthogonalisation (the Newton-Schulz iteration
// for the polar factor; A. Bjorck and C. Bowie, "An iterative algorithm for
// computing the best estimate of an orthogonal matrix", SIAM J. Numer. Anal.
// 8, 1971; N. J. Higham, "Functions of Matrices", SIAM 2008, sec. 8.3):
// X <- X (3 I - X^T X) / 2 = X (I + E / 2) with E = I - X^T X.  It converges
// quadratically (E' = 3/4 E^2 + 1/4 E^3) to the polar factor when |E| < 1 and
// needs no inverse, division or square root: about 54 FP32 operations a step
// against about 100 for a scaled Newton step.  Taken when |E0|_F < 0.5 (all
// singular values of F within (0.7, 1.23) for det F > 0); it stops once
// 3/4 |E|^2, the error after the step, is below FP32 resolution.
__device__ __forceinline__ bool polar_schulz(const float *F, float *X) {
#pragma unroll
    for (int i = 0; i < 9; ++i) X[i] = F[i];
    for (int it = 0; it < 8; ++it) {
        // S = X^T X (symmetric), E = I - S
        const float s00 = X[0] * X[0] + X[3] * X[3] + X[6] * X[6];
        const float s11 = X[1] * X[1] + X[4] * X[4] + X[7] * X[7];
        const float s22 = X[2] * X[2] + X[5] * X[5] + X[8] * X[8];
        const float s01 = X[0] * X[1] + X[3] * X[4] + X[6] * X[7];
        const float s02 = X[0] * X[2] + X[3] * X[5] + X[6] * X[8];
        const float s12 = X[1] * X[2] + X[4] * X[5] + X[7] * X[8];
        const float e00 = 1.0f - s00, e11 = 1.0f - s11, e22 = 1.0f - s22;
        const float en = e00 * e00 + e11 * e11 + e22 * e22 + 2.0f * (s01 * s01 + s02 * s02 + s12 * s12);
        if (it == 0 && !(en < 0.25f)) return false;
        // M = I + E / 2
        const float m00 = 1.0f + 0.5f * e00, m11 = 1.0f + 0.5f * e11, m22 = 1.0f + 0.5f * e22;
        const float m01 = -0.5f * s01, m02 = -0.5f * s02, m12 = -0.5f * s12;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const float x0 = X[3 * r], x1 = X[3 * r + 1], x2 = X[3 * r + 2];
            X[3 * r] = x0 * m00 + x1 * m01 + x2 * m02;
            X[3 * r + 1] = x0 * m01 + x1 * m11 + x2 * m12;
            X[3 * r + 2] = x0 * m02 + x1 * m12 + x2 * m22;
        }
        if (en < 8.0e-8f) return true;
    }
    return false;
}

struct M9 {
    float m[9];
};

// by-value in/out keeps the caller's matrices in registers (no address taken)
__device__ __noinline__ M9 rotation_from_svd(M9 Fm) {
    const float *F = Fm.m;
    M9 Rm;
    float *R = Rm.m;
    float A[9], W[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 9; ++i) A[i] = F[i];
    const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 12; ++sweep) {
        bool rotated = false;
        for (int r = 0; r < 3; ++r) {
            const int p = P[r], q = Q[r];
            float al = 0, be = 0, ga = 0;
            for (int i = 0; i < 3; ++i) {
                al += A[3 * i + p] * A[3 * i + p];
                be += A[3 * i + q] * A[3 * i + q];
                ga += A[3 * i + p] * A[3 * i + q];
            }
            if (ga == 0.0f || fabsf(ga) <= 1.0e-7f * sqrtf(al * be)) continue;
            rotated = true;
            const float zeta = (be - al) / (2.0f * ga);
            const float t = (zeta >= 0.f ? 1.0f : -1.0f) / (fabsf(zeta) + sqrtf(1.0f + zeta * zeta));
            const float c = 1.0f / sqrtf(1.0f + t * t), s = c * t;
            for (int i = 0; i < 3; ++i) {
                const float ap = A[3 * i + p], aq = A[3 * i + q];
                A[3 * i + p] = c * ap - s * aq;
                A[3 * i + q] = s * ap + c * aq;
                const float vp = W[3 * i + p], vq = W[3 * i + q];
                W[3 * i + p] = c * vp - s * vq;
                W[3 * i + q] = s * vp + c * vq;
            }
        }
        if (!rotated) break;
    }
    float sg[3];
    for (int j = 0; j < 3; ++j) sg[j] = sqrtf(A[j] * A[j] + A[3 + j] * A[3 + j] + A[6 + j] * A[6 + j]);
    int o[3] = {0, 1, 2};
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b)
            if (sg[o[b]] > sg[o[a]]) { int t = o[a]; o[a] = o[b]; o[b] = t; }
    float U[9], V[9];
    const float smax = sg[o[0]];
    int known = 0;
    for (int j = 0; j < 3; ++j) {
        for (int i = 0; i < 3; ++i) V[3 * i + j] = W[3 * i + o[j]];
        if (sg[o[j]] > 1.0e-30f && sg[o[j]] > 1.0e-6f * smax) {
            for (int i = 0; i < 3; ++i) U[3 * i + j] = A[3 * i + o[j]] / sg[o[j]];
            known = j + 1;
        }
    }
    if (known == 0) { for (int i = 0; i < 9; ++i) U[i] = (i % 4 == 0) ? 1.f : 0.f; known = 1; }
    if (known == 1) {
        const float a0 = U[0], a1 = U[3], a2 = U[6];
        float e0 = 0, e1 = 0, e2 = 0;
        if (fabsf(a0) <= fabsf(a1) && fabsf(a0) <= fabsf(a2)) e0 = 1;
        else if (fabsf(a1) <= fabsf(a2)) e1 = 1;
        else e2 = 1;
        float b0 = a1 * e2 - a2 * e1, b1 = a2 * e0 - a0 * e2, b2 = a0 * e1 - a1 * e0;
        const float nb = sqrtf(b0 * b0 + b1 * b1 + b2 * b2);
        U[1] = b0 / nb; U[4] = b1 / nb; U[7] = b2 / nb;
        known = 2;
    }
    if (known == 2) {
        U[2] = U[3] * U[7] - U[6] * U[4];
        U[5] = U[6] * U[1] - U[0] * U[7];
        U[8] = U[0] * U[4] - U[3] * U[1];
    }
    const float s3 = (det3(U) * det3(V) < 0.f) ? -1.f : 1.f;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            R[3 * i + j] = U[3 * i] * V[3 * j] + U[3 * i + 1] * V[3 * j + 1] + s3 * U[3 * i + 2] * V[3 * j + 2];
    return Rm;
}

__device__ __forceinline__ void closest_rotation(const float *F, float *R) {
    if (!(det3(F) > 1.0e-20f) || !(polar_schulz(F, R) || polar_newton(F, R))) {
        M9 f;
#pragma unroll
        for (int i = 0; i < 9; ++i) f.m[i] = F[i];
        const M9 r = rotation_from_svd(f);
#pragma unroll
        for (int i = 0; i < 9; ++i) R[i] = r.m[i];
    }
}

// ARAP-SVD constraint of one tet (solve_tetarap_svd_kernel, solve_kernels.py:1194-1307):
// C = |sigma - 1| = |F - R|_F with R the closest rotation, dC/dF = (F - R) / C,
// corner gradients g_k = dC/dF B[k,:]^T, g_3 = -sum.  Returns C, or 0 when the
// constraint is satisfied (sum (sigma - 1)^2 < 1e-12).
__device__ __forceinline__ float arap_grad(const float *F, const float *B, const float w[4], V3 gr[4], float &wsum) {
    float R[9];
    closest_rotation(F, R);
    float G[9], c2s = 0.f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        G[i] = F[i] - R[i];
        c2s += G[i] * G[i];
    }
    wsum = 0.f;
    if (c2s < 1.0e-12f) return 0.f;
    const float C = __builtin_amdgcn_sqrtf(c2s), invC = __builtin_amdgcn_rcpf(C);
#pragma unroll
    for (int i = 0; i < 9; ++i) G[i] *= invC;
    gr[0] = grad_row(G, B, 0);
    gr[1] = grad_row(G, B, 1);
    gr[2] = grad_row(G, B, 2);
    gr[3] = -gr[0] - gr[1] - gr[2];
    wsum = w[0] * dot(gr[0], gr[0]) + w[1] * dot(gr[1], gr[1]) + w[2] * dot(gr[2], gr[2]) + w[3] * dot(gr[3], gr[3]);
    return C;
}

// ---------------------------------------------------------------------------
// per-element projection (all constraint types of one tet, Jacobi, lambda = 0)
// ---------------------------------------------------------------------------
struct Corr {
    V3 d[4];
    float cnt;
};

// an element's own data: everything the projection reads that does not depend
// on the positions, loaded apart so that a caller can issue the loads before
// it stages the vertices (every per-element array is sized for all elements)
template <int FEAT, int LAM>
struct ElemIn {
    float4 lam, fib;
    uint32_t meta;
    float rvol, rsnh;
    float B[9];
};

template <int FEAT, int LAM>
__device__ __forceinline__ void load_elem(const IterArgs &a, int64_t g, ElemIn<FEAT, LAM> &e) {
    e.lam = LAM ? a.lam[g] : make_float4(0.f, 0.f, 0.f, 0.f);
    e.meta = a.meta[g];
    e.rvol = (FEAT & F_VOL) ? a.rvol[g] : 0.f;
    e.rsnh = (FEAT & F_SNH) ? a.rsnh[g] : 0.f;
    e.fib = (FEAT & F_ANISO) ? a.fib[g] : make_float4(0.f, 0.f, 0.f, 0.f);
    constexpr bool NEED_F = (FEAT & (F_ARAP | F_SNH | F_ANISO)) != 0;
#pragma unroll
    for (int q = 0; q < 9; ++q) e.B[q] = NEED_F ? a.Bp[(int64_t)q * a.nptet + g] : 0.f;
}

// Block layouts hold fewer than 2^26 elements and vertices (checked when the
// layout is built), so every byte offset into a per-element array, the nine B
// planes included, fits 32 bits: an element load is a uniform base (scalar
// registers) plus one 32-bit lane offset shared by all arrays, and the B planes
// are buffer loads whose plane offset is a scalar operand, with no 64-bit
// address arithmetic per lane.  (The vertex gathers keep 64-bit indexing: in
// the block kernel the 32-bit form measured slower.)
constexpr int64_t BLOCK_MAX_ELEMS = int64_t(1) << 26;
constexpr int BUF_RSRC_W3 = 0x00020000;         // buffer descriptor word 3 (32-bit data format)

template <typename T>
__device__ __forceinline__ T ld_off(const T *base, uint32_t byte_off) {
    return *(const T *)((const char *)base + byte_off);
}

template <int FEAT, int LAM>
__device__ __forceinline__ void load_elem(const IterArgs &a, uint32_t g, ElemIn<FEAT, LAM> &e) {
    const uint32_t o4 = g * 4u, o16 = g * 16u;
    e.lam = LAM ? ld_off(a.lam, o16) : make_float4(0.f, 0.f, 0.f, 0.f);
    e.meta = ld_off(a.meta, o4);
    e.rvol = (FEAT & F_VOL) ? ld_off(a.rvol, o4) : 0.f;
    e.rsnh = (FEAT & F_SNH) ? ld_off(a.rsnh, o4) : 0.f;
    e.fib = (FEAT & F_ANISO) ? ld_off(a.fib, o16) : make_float4(0.f, 0.f, 0.f, 0.f);
    constexpr bool NEED_F = (FEAT & (F_ARAP | F_SNH | F_ANISO)) != 0;
    if (NEED_F) {
        const uint32_t plane = (uint32_t)a.nptet * 4u;
        const __amdgpu_buffer_rsrc_t br = __builtin_amdgcn_make_buffer_rsrc((void *)a.Bp, 0, (int)(9u * plane), BUF_RSRC_W3);
#pragma unroll
        for (int q = 0; q < 9; ++q)
            e.B[q] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(br, (int)o4, (int)(q * plane), 0));
    } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) e.B[q] = 0.f;
    }
}

// LAM = 0: plain zero-lambda Jacobi (the reference GPU path); 1: lambda carried
// across the iterations of a substep (a.lam read and updated in place, one
// thread per element); 2: warm-start impulse, the step applies dlambda = lambda0
// from a.lam along each constraint gradient and keeps lambda = lambda0.
// cls: the material-class table (global, or a copy in LDS)
// The corner positions (and the class parameters) are read where a constraint
// uses them, again after the polar iteration, instead of being held in
// registers across it: ARAP-SVD goes first with only F, B and the polar
// iterate live, and the volume constraint, which needs the positions and the
// previous positions, goes last.  A compiler fence before each re-read keeps
// the compiler from reusing the first read.  (Fewer live registers: more
// waves per SIMD to hide the vertex-table and LDS latencies.)
__device__ __forceinline__ void reg_fence() { asm volatile("" ::: "memory"); }

template <int FEAT, int DAMP, int LAM = 0>
__device__ __forceinline__ void project_elem(const IterArgs &a, int64_t g, const ElemIn<FEAT, LAM> &e,
                                             const float4 *sx, const float4 *spp, const int lv[4], Corr &out,
                                             const DevClass *cls) {
    const float4 lam = e.lam;
    float4 nl = lam;
    const uint32_t meta = e.meta;
    const uint32_t mask = (meta >> 16) & 15u;
    const uint32_t cls_off = (meta & 0xffffu) * (uint32_t)sizeof(DevClass);
    const float dt = a.sp.dt;
    for (int i = 0; i < 4; ++i) out.d[i] = mk3(0.f, 0.f, 0.f);
    out.cnt = 0.f;

    // F is formed for every element when a deformation-gradient type is
    // compiled in (finite for finite positions, which the selected fibre
    // update relies on)
    float B[9], F[9];
    constexpr bool NEED_F = (FEAT & (F_ARAP | F_SNH | F_ANISO)) != 0;
    if (NEED_F) {
        const float4 X0 = sx[lv[0]], X1 = sx[lv[1]], X2 = sx[lv[2]], X3 = sx[lv[3]];
        const V3 c0 = xyz(X0) - xyz(X3), c1 = xyz(X1) - xyz(X3), c2 = xyz(X2) - xyz(X3);
#pragma unroll
        for (int q = 0; q < 9; ++q) B[q] = e.B[q];
        const float Ds[9] = {c0.x, c1.x, c2.x, c0.y, c1.y, c2.y, c0.z, c1.z, c2.z};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                F[3 * i + j] = Ds[3 * i] * B[j] + Ds[3 * i + 1] * B[3 + j] + Ds[3 * i + 2] * B[6 + j];
    }

    // ---- ARAP-SVD: solve_tetarap_svd_kernel, solve_kernels.py:1194-1307 ----
    // The polar iteration keeps its branches; the update is selected instead of
    // branched (a satisfied constraint gets C = 0, zero gradients and dL = 0),
    // so the wave issues no exec-mask bookkeeping around it.
    if ((FEAT & F_ARAP) && (mask & M_ARAP)) {
        float R[9];
        closest_rotation(F, R);
        reg_fence();
        const float w[4] = {sx[lv[0]].w, sx[lv[1]].w, sx[lv[2]].w, sx[lv[3]].w};
        const float at_arap = ld_off(cls, cls_off).at_arap;
        float G[9], c2s = 0.f;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            G[i] = F[i] - R[i];
            c2s += G[i] * G[i];
        }
        const bool sat = c2s < 1.0e-12f;
        const float C = sat ? 0.0f : __builtin_amdgcn_sqrtf(c2s);
        const float invC = sat ? 0.0f : __builtin_amdgcn_rcpf(C);
#pragma unroll
        for (int i = 0; i < 9; ++i) G[i] *= invC;
        V3 gr[4];
        gr[0] = grad_row(G, B, 0);
        gr[1] = grad_row(G, B, 1);
        gr[2] = grad_row(G, B, 2);
        gr[3] = -gr[0] - gr[1] - gr[2];
        const float wsum =
            w[0] * dot(gr[0], gr[0]) + w[1] * dot(gr[1], gr[1]) + w[2] * dot(gr[2], gr[2]) + w[3] * dot(gr[3], gr[3]);
        const bool on = C > 0.0f && wsum != 0.0f;
        float dL = LAM == 2 ? lam.x : -(C + at_arap * lam.x) * __builtin_amdgcn_rcpf(wsum + at_arap);
        dL = on ? dL : 0.0f;
        if (LAM == 1 && on) nl.x = lam.x + dL;
#pragma unroll
        for (int i = 0; i < 4; ++i) out.d[i] = (dL * w[i]) * gr[i];
        out.cnt = on ? 1.0f : 0.0f;
    }

    // ---- SNH: tet_snh_update_xpbd_fn, solve_kernels.py:1429-1584 ----
    // (square roots and reciprocals in the hardware's single instructions,
    // about 1 ulp, instead of the correctly rounded sequences, as in ARAP)
    if ((FEAT & F_SNH) && (mask & M_SNH)) {
        reg_fence();
        const float4 X0 = sx[lv[0]], X1 = sx[lv[1]], X2 = sx[lv[2]], X3 = sx[lv[3]];
        const V3 P[4] = {xyz(X0), xyz(X1), xyz(X2), xyz(X3)};
        const float w[4] = {X0.w, X1.w, X2.w, X3.w};
        const V3 c0 = P[0] - P[3], c1 = P[1] - P[3], c2 = P[2] - P[3];
        const ClassParams cp = ld_off(cls, cls_off).p;
        float Ic = 0.f;
#pragma unroll
        for (int i = 0; i < 9; ++i) Ic += F[i] * F[i];
        const float r_s = __builtin_amdgcn_sqrtf(fmaxf(Ic, 1.0e-12f));
        if (!(r_s < 1.0e-6f)) {
            const float dC = r_s - 1.7320508075688772f;
            const float inv_rs = __builtin_amdgcn_rcpf(r_s);
            V3 dg[4];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                dg[k] = V3{B[3 * k] * (F[0] * inv_rs) + B[3 * k + 1] * (F[1] * inv_rs) + B[3 * k + 2] * (F[2] * inv_rs),
                           B[3 * k] * (F[3] * inv_rs) + B[3 * k + 1] * (F[4] * inv_rs) + B[3 * k + 2] * (F[5] * inv_rs),
                           B[3 * k] * (F[6] * inv_rs) + B[3 * k + 1] * (F[7] * inv_rs) + B[3 * k + 2] * (F[8] * inv_rs)};
            }
            dg[3] = -dg[0] - dg[1] - dg[2];
            const float dsw = w[0] * dot(dg[0], dg[0]) + w[1] * dot(dg[1], dg[1]) +
                              w[2] * dot(dg[2], dg[2]) + w[3] * dot(dg[3], dg[3]);
            if (dsw != 0.0f) {
                const float mu = cp.mu, lamS = cp.lam;
                const float hC = det3(F) - (1.0f + mu * __builtin_amdgcn_rcpf(lamS));
                const float restv = e.rsnh;
                const float inv_detDm = __builtin_amdgcn_rcpf(fmaxf(fabsf(restv) * 6.0f, 1.0e-20f));
                V3 hg[4];
                hg[0] = inv_detDm * cross(c1, c2);
                hg[1] = inv_detDm * cross(c2, c0);
                hg[2] = inv_detDm * cross(c0, c1);
                hg[3] = -hg[0] - hg[1] - hg[2];
                const float hsw = w[0] * dot(hg[0], hg[0]) + w[1] * dot(hg[1], hg[1]) +
                                  w[2] * dot(hg[2], hg[2]) + w[3] * dot(hg[3], hg[3]);
                if (hsw != 0.0f) {
                    const float dh = w[0] * dot(dg[0], hg[0]) + w[1] * dot(dg[1], hg[1]) +
                                     w[2] * dot(dg[2], hg[2]) + w[3] * dot(dg[3], hg[3]);
                    const float abs_vol = fmaxf(fabsf(restv), 1.0e-20f);
                    const float dalpha = __builtin_amdgcn_rcpf(mu * abs_vol * dt * dt);
                    const float halpha = __builtin_amdgcn_rcpf(lamS * abs_vol * dt * dt);
                    float ddamp = 0.f, dgam = 1.f, hdamp = 0.f, hgam = 1.f;
                    if (DAMP && cp.damp_snh > 0.0f) {
                        const float kd = cp.damp_snh;
                        const float dbeta = mu * abs_vol * kd * dt * dt;
                        const float hbeta = lamS * abs_vol * kd * dt * dt;
                        dgam = dalpha * dbeta * __builtin_amdgcn_rcpf(dt);
                        hgam = halpha * hbeta * __builtin_amdgcn_rcpf(dt);
                        for (int i = 0; i < 4; ++i) {
                            const V3 dv = P[i] - xyz(spp[lv[i]]);
                            ddamp += dot(dg[i], dv);
                            hdamp += dot(hg[i], dv);
                        }
                        ddamp = ddamp * dgam;
                        hdamp = hdamp * hgam;
                        dgam = dgam + 1.0f;
                        hgam = hgam + 1.0f;
                    }
                    const float Axx = dgam * dsw + dalpha;
                    const float Ayy = hgam * hsw + halpha;
                    const float detA = Axx * Ayy - dh * dh;
                    if (!(detA < 1.0e-8f)) {
                        const float inv_det = __builtin_amdgcn_rcpf(detA);
                        const float db = -dC - dalpha * lam.x - ddamp;
                        const float hb = -hC - halpha * lam.y - hdamp;
                        const float ddL = LAM == 2 ? lam.x : inv_det * (Ayy * db - dh * hb);
                        const float hdL = LAM == 2 ? lam.y : inv_det * (Axx * hb - dh * db);
                        if (LAM == 1) { nl.x = lam.x + ddL; nl.y = lam.y + hdL; }
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            out.d[i] = out.d[i] + w[i] * (ddL * dg[i] + hdL * hg[i]);
                        out.cnt += 1.0f;
                    }
                }
            }
        }
    }

    // ---- aniso_arap: solve_tet_aniso_arap_kernel, solve_kernels.py:1310-1426 ----
    // (selected like the volume constraint: a skipped fibre gets gs = 0, so
    // zero gradients, and dL = 0)
    if (FEAT & F_ANISO) {
        reg_fence();
        const float w[4] = {sx[lv[0]].w, sx[lv[1]].w, sx[lv[2]].w, sx[lv[3]].w};
        const float at_aniso = ld_off(cls, cls_off).at_aniso;
        const float4 fb = e.fib;  // f0 (normalised at build), fiberscale
        const V3 f0 = xyz(fb);
        const V3 d = V3{F[0] * f0.x + F[1] * f0.y + F[2] * f0.z, F[3] * f0.x + F[4] * f0.y + F[5] * f0.z,
                        F[6] * f0.x + F[7] * f0.y + F[8] * f0.z};
        const float s = __builtin_amdgcn_sqrtf(dot(d, d));
        const float detF = det3(F);
        const bool act = (mask & M_ANISO) && !(s < 1.0e-9f) && !(detF <= 0.0f || s > 1.5f);
        const float diff = s - fb.w;  // SI4 = +1 since det F > 0
        const float C = 0.5f * diff * diff;
        const float gs = act ? diff * __builtin_amdgcn_rcpf(s) : 0.0f;
        float G[9];
        const float dv[3] = {d.x, d.y, d.z}, fv[3] = {f0.x, f0.y, f0.z};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) G[3 * i + j] = gs * (dv[i] * fv[j]);
        V3 gr[4];
        gr[0] = grad_row(G, B, 0);
        gr[1] = grad_row(G, B, 1);
        gr[2] = grad_row(G, B, 2);
        gr[3] = -gr[0] - gr[1] - gr[2];
        const float wsum = w[0] * dot(gr[0], gr[0]) + w[1] * dot(gr[1], gr[1]) +
                           w[2] * dot(gr[2], gr[2]) + w[3] * dot(gr[3], gr[3]);
        const bool on = act && !(wsum < 1.0e-20f);
        float dL = LAM == 2 ? lam.z : -(C + at_aniso * lam.z) * __builtin_amdgcn_rcpf(wsum + at_aniso);
        dL = on ? dL : 0.0f;
        if (LAM == 1 && on) nl.z = lam.z + dL;
#pragma unroll
        for (int i = 0; i < 4; ++i) out.d[i] = out.d[i] + (dL * w[i]) * gr[i];
        out.cnt += on ? 1.0f : 0.0f;
    }
    // ---- volume: tet_volume_update_xpbd_fn, solve_kernels.py:187-281 ----
    // (last: it needs the positions and the previous positions)
    //
    // Computed without branches: every quantity is finite for finite positions,
    // and a constraint that the reference skips (wsum <= 1e-9, stiffness 0, type
    // absent) gets dl = 0 by a select, so the corrections equal the branched
    // form's (up to the sign of a zero) while the wave issues no exec-mask
    // bookkeeping and keeps no zero copies of the accumulators live.
    if (FEAT & F_VOL) {
        reg_fence();
        const float4 X0 = sx[lv[0]], X1 = sx[lv[1]], X2 = sx[lv[2]], X3 = sx[lv[3]];
        const V3 P[4] = {xyz(X0), xyz(X1), xyz(X2), xyz(X3)};
        const float w[4] = {X0.w, X1.w, X2.w, X3.w};
        const V3 c0 = P[0] - P[3], c1 = P[1] - P[3], c2 = P[2] - P[3];
        const DevClass dc = ld_off(cls, cls_off);
        const ClassParams &cp = dc.p;
        constexpr float SIXTH = 1.0f / 6.0f;
        V3 gr[4];
        gr[0] = SIXTH * cross(c1, c2);
        gr[1] = SIXTH * cross(c2, c0);
        gr[2] = SIXTH * cross(c0, c1);
        gr[3] = -(gr[0] + gr[1] + gr[2]);
        const float wsum = w[0] * dot(gr[0], gr[0]) + w[1] * dot(gr[1], gr[1]) +
                           w[2] * dot(gr[2], gr[2]) + w[3] * dot(gr[3], gr[3]);
        const float vol = SIXTH * dot(cross(c0, c1), c2);
        const float rest = e.rvol;
        const bool comp = a.sp.has_compress && (vol < rest);
        const float k = comp ? cp.kcomp_vol : cp.k_vol;
        const float alpha = comp ? dc.a_comp : dc.a_vol;
        const float C = vol - rest;
        float dsum = 0.f, gamma = 1.f;
        if (DAMP && cp.damp_vol > 0.0f) {
            gamma = comp ? dc.g_comp : dc.g_vol;
            for (int i = 0; i < 4; ++i) {
                const float4 pp = spp[lv[i]];
                dsum += dot(gr[i], P[i] - xyz(pp));
            }
            dsum = dsum * gamma;
            gamma = gamma + 1.0f;
        }
        const bool on = (mask & M_VOL) && wsum > 1e-9f && k != 0.0f;
        float dl = LAM == 2 ? lam.y : (-C - alpha * lam.y - dsum) * __builtin_amdgcn_rcpf(gamma * wsum + alpha);
        dl = on ? dl : 0.0f;
        if (LAM == 1 && on) nl.y = lam.y + dl;
#pragma unroll
        for (int i = 0; i < 4; ++i) out.d[i] = out.d[i] + (dl * w[i]) * gr[i];
        out.cnt += on ? 1.0f : 0.0f;
    }

    if (LAM == 1) a.lam[g] = nl;
}

// ---------------------------------------------------------------------------
// attach (per owned vertex): attach_bilateral_update_fn, solve_kernels.py:462-526
// ---------------------------------------------------------------------------
__device__ __forceinline__ void attach_vertex(const IterArgs &a, int v, float4 x, float4 pp,
                                              float &ax, float &ay, float &az, float &aw) {
    const float4 a0 = a.att0[v];
    if (!(a0.w > 0.0f)) return;
    const float wv = x.w;
    if (wv == 0.0f) return;
    const float4 a1 = a.att1[v];
    V3 n = xyz(x) - xyz(a0);
    const float d = sqrtf(dot(n, n));
    if (!(d >= 1e-6f)) return;
    const bool comp = a.sp.has_compress && (d < a1.x);
    const float kv = comp ? a1.y : a0.w;
    if (kv == 0.0f) return;
    const float dt = a.sp.dt;
    const float alpha = (1.0f / kv) / (dt * dt);
    const float C = d - a1.x;
    n = n / d;
    float dsum = 0.f, gamma = 1.f;
    if (a1.z > 0.0f) {
        const float beta = kv * a1.z * dt * dt;
        gamma = alpha * beta / dt;
        dsum = gamma * dot(n, xyz(x) - xyz(pp));
        gamma = gamma + 1.0f;
    }
    const float dL = (-C - dsum) / (gamma * wv + alpha);
    const V3 dp = V3{n.x * (-dL), n.y * (-dL), n.z * (-dL)};
    const V3 c = (-wv) * dp;
    ax += c.x;
    ay += c.y;
    az += c.z;
    aw += 1.0f;
}

// ---------------------------------------------------------------------------
// one Jacobi iteration over all patches
// ---------------------------------------------------------------------------
template <int BS, int FEAT, int DAMP, int MODE>
__global__ __launch_bounds__(BS) void k_iterate(IterArgs a, const float4 *__restrict__ xin,
                                                float4 *__restrict__ xout,
                                                const float4 *__restrict__ pprev) {
    extern __shared__ float4 smem[];
    float4 *sx = smem;
    float4 *spp = smem + a.max_table;
    float4 *sc = smem + (DAMP ? 2 : 1) * a.max_table;
    const PatchDesc P = a.patch[blockIdx.x];
    const int tid = threadIdx.x;
    for (int i = tid; i < P.nown; i += BS) {
        sx[i] = xin[P.vbeg + i];
        if (DAMP) spp[i] = pprev[P.vbeg + i];
    }
    for (int i = tid; i < P.nhalo; i += BS) {
        const int gv = a.halo[P.hbeg + i];
        sx[P.nown + i] = xin[gv];
        if (DAMP) spp[P.nown + i] = pprev[gv];
    }
    __syncthreads();
    float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
    int cur = 0, end = 0;
    if (tid < P.nown) {
        cur = a.vslot_off[P.vbeg + tid];
        end = a.vslot_off[P.vbeg + tid + 1];
    }
    for (int base = 0; base < P.ntet; base += BS) {
        const int t = base + tid;
        if (t < P.ntet) {
            Corr c;
            const ushort4 li = a.lidx[(int64_t)P.tbeg + t];
            const int lv[4] = {li.x, li.y, li.z, li.w};
            ElemIn<FEAT, 0> e;
            load_elem(a, (int64_t)P.tbeg + t, e);
            project_elem<FEAT, DAMP>(a, (int64_t)P.tbeg + t, e, sx, spp, lv, c, a.cls);
#pragma unroll
            for (int k = 0; k < 4; ++k) sc[tid * 4 + k] = make_float4(c.d[k].x, c.d[k].y, c.d[k].z, c.cnt);
        }
        __syncthreads();
        if (tid < P.nown) {
            const int lim = (base + BS) * 4;
            while (cur < end) {
                const int s = a.vslot[cur];
                if (s >= lim) break;
                const float4 q = sc[s - base * 4];
                ax += q.x;
                ay += q.y;
                az += q.z;
                aw += q.w;
                ++cur;
            }
        }
        __syncthreads();
    }
    if (tid < P.nown) {
        const int v = P.vbeg + tid;
        float4 x = sx[tid];
        if (a.att0) attach_vertex(a, v, x, DAMP ? spp[tid] : make_float4(0, 0, 0, 0), ax, ay, az, aw);
        if (MODE == MODE_ACCUM) {
            float4 o = a.dpw[v];
            o.x += ax;
            o.y += ay;
            o.z += az;
            o.w += aw;
            a.dpw[v] = o;
        } else {
            if (aw > 1e-9f) {  // apply_dP_kernel, engine.py:689-698 (one reciprocal, as finish_value)
                const float inv = 1.0f / aw;
                x.x = x.x + ax * inv;
                x.y = x.y + ay * inv;
                x.z = x.z + az * inv;
            }
            if (a.pin) {  // pintotarget overwrite, mg/kernel.py:212-226
                const float4 p = a.pin[v];
                if (p.w != 0.0f) { x.x = p.x; x.y = p.y; x.z = p.z; }
            }
            xout[v] = x;
        }
    }
}


// ---------------------------------------------------------------------------
// "flat" strategy: one thread per tet element (each element once, Morton
// order) writes its four corner corrections straight into a vertex-major
// corner array; one thread per vertex then sums its contiguous run of corners
// (ascending element order, so the sum order is fixed), applies dP/dPw,
// attach and pins.
// ---------------------------------------------------------------------------
template <int FEAT, int DAMP, int LAM = 0>
__global__ __launch_bounds__(256) void k_flat_tets(IterArgs a, const float4 *__restrict__ xin,
                                                   const float4 *__restrict__ pprev,
                                                   float4 *__restrict__ corr) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= a.nptet) return;
    const int4 gi = a.gidx[g];
    const int lv[4] = {gi.x, gi.y, gi.z, gi.w};
    Corr c;
    ElemIn<FEAT, LAM> e;
    load_elem(a, g, e);
    project_elem<FEAT, DAMP, LAM>(a, g, e, xin, pprev, lv, c, a.cls);
    const int4 so = a.slot_of[g];
    corr[so.x] = make_float4(c.d[0].x, c.d[0].y, c.d[0].z, c.cnt);
    corr[so.y] = make_float4(c.d[1].x, c.d[1].y, c.d[1].z, c.cnt);
    corr[so.z] = make_float4(c.d[2].x, c.d[2].y, c.d[2].z, c.cnt);
    corr[so.w] = make_float4(c.d[3].x, c.d[3].y, c.d[3].z, c.cnt);
}

// finish one vertex: attach, then either accumulate (MODE_ACCUM) or apply
// dP/dPw (apply_dP_kernel, engine.py:689-698) and the pin overwrite
// (pintotarget_overwrite_masked_kernel, mg/kernel.py:212-226)
// finished position of a vertex after an iteration: attach, apply dP/dPw
// (apply_dP_kernel, engine.py:689-698) and the pin overwrite
// (pintotarget_overwrite_masked_kernel, mg/kernel.py:212-226)
// pv: the vertex's pin record (pin_of), loaded by the caller ahead of time
__device__ __forceinline__ float4 finish_value(const IterArgs &a, int v, float4 x, float4 pp, float ax, float ay,
                                               float az, float aw, float4 pv) {
    if (a.att0) attach_vertex(a, v, x, pp, ax, ay, az, aw);
    if (aw > 1e-9f) {
        // one correctly rounded reciprocal instead of three divisions (the
        // correction differs from dP / dPw by at most an ulp or two of the
        // correction, far below the solver's own tolerance)
        const float inv = 1.0f / aw;
        x.x = x.x + ax * inv;
        x.y = x.y + ay * inv;
        x.z = x.z + az * inv;
    }
    if (a.pin && pv.w != 0.0f) { x.x = pv.x; x.y = pv.y; x.z = pv.z; }
    return x;
}

__device__ __forceinline__ float4 pin_of(const IterArgs &a, int v) {
    return a.pin ? a.pin[v] : make_float4(0.f, 0.f, 0.f, 0.f);
}

// finish one vertex: either accumulate (MODE_ACCUM, attach included) or write
// the finished position
template <int DAMP, int MODE>
__device__ __forceinline__ void finish_vertex(const IterArgs &a, int v, float4 x, float4 pp, float ax, float ay,
                                              float az, float aw, float4 pv, float4 *__restrict__ xout) {
    if (MODE == MODE_ACCUM) {
        if (a.att0) attach_vertex(a, v, x, pp, ax, ay, az, aw);
        float4 o = a.dpw[v];
        o.x += ax;
        o.y += ay;
        o.z += az;
        o.w += aw;
        a.dpw[v] = o;
        return;
    }
    xout[v] = finish_value(a, v, x, pp, ax, ay, az, aw, pv);
}

// Vertex pass of the flat strategy: four lanes per vertex read its contiguous
// run of corner corrections (lane j takes corners j, j+4, ...; a wavefront's
// loads then touch 16 runs rather than 64), then two cross-lane adds combine
// them in a fixed order and the first lane finishes the vertex.
template <int DAMP, int MODE>
__global__ __launch_bounds__(256) void k_flat_verts(IterArgs a, const float4 *__restrict__ xin,
                                                    float4 *__restrict__ xout,
                                                    const float4 *__restrict__ pprev,
                                                    const float4 *__restrict__ corr, int nv) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int v = t >> 2, j = t & 3;
    float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
    if (v < nv) {
        const int e = a.vslot_off[v + 1];
        for (int q = a.vslot_off[v] + j; q < e; q += 4) {
            const float4 c = corr[q];
            ax += c.x;
            ay += c.y;
            az += c.z;
            aw += c.w;
        }
    }
    // lanes of a vertex are 4 consecutive lanes of one wavefront
    ax += __shfl_xor(ax, 1);
    ay += __shfl_xor(ay, 1);
    az += __shfl_xor(az, 1);
    aw += __shfl_xor(aw, 1);
    ax += __shfl_xor(ax, 2);
    ay += __shfl_xor(ay, 2);
    az += __shfl_xor(az, 2);
    aw += __shfl_xor(aw, 2);
    if (v >= nv || j != 0) return;
    finish_vertex<DAMP, MODE>(a, v, xin[v], DAMP ? pprev[v] : make_float4(0, 0, 0, 0), ax, ay, az, aw,
                              MODE == MODE_FUSED ? pin_of(a, v) : make_float4(0, 0, 0, 0), xout);
}

// ---------------------------------------------------------------------------
// "block" strategy: one workgroup per block of consecutive elements.  The
// block's vertex table is staged in LDS, every element is projected once,
// the four corner corrections go to LDS, and each local vertex sums its
// corners in ascending slot order.  Vertices owned by this block alone are
// finished here; the others get this block's partial sum, finished by
// k_block_shared, which adds a vertex's partials in ascending block order.
// The sum order is fixed by the layout, so results are run-to-run identical.
// ---------------------------------------------------------------------------
struct BlockArgs {
    const BlockDesc *blk;
    const int32_t *bvt, *bps;
    const uint32_t *brun;
    const ushort4 *blidx;
    const ushort4 *bcpos;        // per element: LDS positions of its four corners (vertex-major)
    const int32_t *svid, *soff;
    float4 *part;
    int max_nloc, block_tets;
};

// XCD-aware block order.  The dispatcher deals workgroups to the eight XCDs of
// the chip in turn (workgroup i to XCD i mod 8), and each XCD has its own L2.
// Workgroup i therefore takes the j-th block (j = i / 8) of XCD (i mod 8)'s
// contiguous range of blocks: neighbouring blocks, which share vertices and
// the cache lines of their positions, pins and partials, run on one XCD and
// meet in its L2.  A bijection on [0, n), so the result does not depend on it.
constexpr uint32_t N_XCD = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t i, uint32_t n) {
    const uint32_t x = i % N_XCD, j = i / N_XCD, q = n / N_XCD, r = n % N_XCD;
    return x * q + min(x, r) + j;
}

// LDS: vertex table (and previous positions when damping), then the corner
// corrections of all block elements in vertex-major order: each element stores
// its four corners at their places in the block's local CSR (bcpos), so after
// one barrier every local entry adds a contiguous run, in ascending (element,
// corner) order.
template <int FEAT, int DAMP, int LAM, int MODE>
__global__ __launch_bounds__(512) void k_block_tets(IterArgs a, BlockArgs b, const float4 *__restrict__ xin,
                                                    float4 *__restrict__ xout,
                                                    const float4 *__restrict__ pprev) {
    extern __shared__ float4 smem[];
    const int M = b.max_nloc, TB = b.block_tets;
    float4 *sx = smem;
    float4 *spp = smem + M;
    float4 *sc = smem + (DAMP ? 2 : 1) * M;
    const int tid = threadIdx.x, nth = blockDim.x;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    // blocks are consecutive runs of TB elements, so the element loads of the
    // one-element-per-thread path are issued before the block record arrives
    const uint32_t tb0 = blk * (uint32_t)TB;
    const int nte = (int)min((int64_t)TB, a.nptet - (int64_t)tb0);
    const bool has_t = nte <= nth && tid < nte;
    const uint32_t g = tb0 + tid;
    ElemIn<FEAT, LAM> e;
    ushort4 li = make_ushort4(0, 0, 0, 0), cp = li;
    if (has_t) {
        load_elem(a, g, e);
        li = ld_off(b.blidx, g * 8u);
        cp = ld_off(b.bcpos, g * 8u);
    }
    const BlockDesc D = b.blk[blk];
    if (D.ntet <= nth && D.nloc <= nth) {
        // one element and at most one local vertex per thread: the element's
        // data and the vertex's tables are loaded before the staging barrier,
        // so their latency overlaps the vertex gather
        const bool has_v = tid < D.nloc;
        int gv = 0, ps = -1;
        uint32_t run = 0;
        float4 pv = make_float4(0.f, 0.f, 0.f, 0.f);
        if (has_v) {
            gv = b.bvt[D.vbeg + tid];
            run = b.brun[D.vbeg + tid];
            ps = b.bps[D.vbeg + tid];
            sx[tid] = xin[gv];
            if (DAMP) spp[tid] = pprev[gv];
            // the pin record of a vertex this block finishes, loaded with the gather
            if (MODE == MODE_FUSED && ps < 0 && a.pin) pv = ld_off(a.pin, (uint32_t)gv * 16u);
        }
        __syncthreads();
        if (has_t) {
            const int lv[4] = {li.x, li.y, li.z, li.w};
            Corr c;
            project_elem<FEAT, DAMP, LAM>(a, g, e, sx, spp, lv, c, a.cls);
            sc[cp.x] = make_float4(c.d[0].x, c.d[0].y, c.d[0].z, c.cnt);
            sc[cp.y] = make_float4(c.d[1].x, c.d[1].y, c.d[1].z, c.cnt);
            sc[cp.z] = make_float4(c.d[2].x, c.d[2].y, c.d[2].z, c.cnt);
            sc[cp.w] = make_float4(c.d[3].x, c.d[3].y, c.d[3].z, c.cnt);
        }
        __syncthreads();
        if (!has_v) return;
        const int s0 = (int)(run & 0xffffu), n = (int)(run >> 16);
        float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
        for (int j = 0; j < n; ++j) {
            const float4 q = sc[s0 + j];
            ax += q.x;
            ay += q.y;
            az += q.z;
            aw += q.w;
        }
        if (ps >= 0) {
            b.part[ps] = make_float4(ax, ay, az, aw);
        } else {
            finish_vertex<DAMP, MODE>(a, gv, sx[tid], DAMP ? spp[tid] : make_float4(0, 0, 0, 0), ax, ay, az, aw, pv,
                                      xout);
        }
        return;
    }
    for (int i = tid; i < D.nloc; i += nth) {
        const int gv = b.bvt[D.vbeg + i];
        sx[i] = xin[gv];
        if (DAMP) spp[i] = pprev[gv];
    }
    __syncthreads();
    for (int t = tid; t < D.ntet; t += nth) {
        const int64_t g = (int64_t)D.tbeg + t;
        const ushort4 li = b.blidx[g];
        const int lv[4] = {li.x, li.y, li.z, li.w};
        Corr c;
        ElemIn<FEAT, LAM> e;
        load_elem(a, g, e);
        project_elem<FEAT, DAMP, LAM>(a, g, e, sx, spp, lv, c, a.cls);
        const ushort4 cp = b.bcpos[g];
        sc[cp.x] = make_float4(c.d[0].x, c.d[0].y, c.d[0].z, c.cnt);
        sc[cp.y] = make_float4(c.d[1].x, c.d[1].y, c.d[1].z, c.cnt);
        sc[cp.z] = make_float4(c.d[2].x, c.d[2].y, c.d[2].z, c.cnt);
        sc[cp.w] = make_float4(c.d[3].x, c.d[3].y, c.d[3].z, c.cnt);
    }
    __syncthreads();
    for (int i = tid; i < D.nloc; i += nth) {
        const uint32_t run = b.brun[D.vbeg + i];
        const int s0 = (int)(run & 0xffffu), n = (int)(run >> 16);
        float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
        for (int j = 0; j < n; ++j) {
            const float4 q = sc[s0 + j];
            ax += q.x;
            ay += q.y;
            az += q.z;
            aw += q.w;
        }
        const int ps = b.bps[D.vbeg + i];
        if (ps >= 0) {
            b.part[ps] = make_float4(ax, ay, az, aw);
        } else {
            const int gv = b.bvt[D.vbeg + i];
            finish_vertex<DAMP, MODE>(a, gv, sx[i], DAMP ? spp[i] : make_float4(0, 0, 0, 0), ax, ay, az, aw,
                                      MODE == MODE_FUSED ? pin_of(a, gv) : make_float4(0, 0, 0, 0), xout);
        }
    }
}

template <int DAMP, int MODE>
__global__ __launch_bounds__(256) void k_block_shared(IterArgs a, BlockArgs b, const float4 *__restrict__ xin,
                                                      float4 *__restrict__ xout,
                                                      const float4 *__restrict__ pprev, int ns) {
    const int i = (int)xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (i >= ns) return;
    const int v = b.svid[i];
    const int s0 = b.soff[i], e = b.soff[i + 1];
    // a shared vertex has a few partials: the first four are loaded together,
    // with the vertex's own position, and added in slot order
    const int n = e - s0;
    float4 q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = k < n ? b.part[s0 + k] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 x = xin[v];
    // the previous position is read only by the attach update
    const float4 pp = (DAMP && a.att0) ? pprev[v] : make_float4(0, 0, 0, 0);
    const float4 pv = MODE == MODE_FUSED ? pin_of(a, v) : make_float4(0, 0, 0, 0);
    float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k < n) {
            ax += q[k].x;
            ay += q[k].y;
            az += q[k].z;
            aw += q[k].w;
        }
    for (int s = s0 + 4; s < e; ++s) {
        const float4 r = b.part[s];
        ax += r.x;
        ay += r.y;
        az += r.z;
        aw += r.w;
    }
    finish_vertex<DAMP, MODE>(a, v, x, pp, ax, ay, az, aw, pv, xout);
}

// ---------------------------------------------------------------------------
// "resident" mode of the block strategy: one launch runs all K plain Jacobi
// iterations of a substep.  Runs of R consecutive blocks form a part and one
// workgroup per part keeps the part's vertices (positions, previous positions)
// and the class table in LDS for the whole launch, and loads each block's
// elements one block ahead.  Per iteration a
// part projects its blocks one after the other (one element per thread) and
// sums every block-local vertex's corners as k_block_tets does; the sum of a
// vertex that only this part touches goes into an LDS accumulator (ascending
// block order), the block partial of a cross vertex (touched by blocks of other
// parts too) goes to its slot of the block strategy's partial array.  After the
// part has finished its own vertices it publishes its iteration count, waits
// for the parts it shares vertices with, and adds the partials of each cross
// vertex in slot order, exactly as k_block_shared does: every sum has the
// block strategy's operands in its order, so the positions equal those of the
// per-iteration launches bit for bit.
//
// Partials cross workgroups (and XCDs) as 16-byte agent-scope (sc1) stores and
// loads, iteration counts as agent-scope atomics; a part publishes its count
// only after its partial stores have completed.  The partial array is double-buffered by iteration
// parity, so a part that runs ahead never overwrites a partial that a slower
// neighbour still has to read (it waits for that neighbour's next count
// first).  Every part of a launch must be resident at once: the host checks
// the occupancy, and a wait that runs out of polls sets a status word instead
// of hanging (the iterations that follow skip their waits, later launches
// return at once, and the host reports the error).
// ---------------------------------------------------------------------------
struct ResArgs {
    const PartDesc *parts;
    const int4 *prec;          // per part-local entry: vertex, svid index or -1, write-back flag, 0
    const uint32_t *bmap;      // per block-local entry: part-local index | cross << 31
    const ushort4 *plidx;      // per element: part-local corner ids
    const int32_t *nbr;
    float4 *xp;                // block partials of cross vertices: two arrays of np, by iteration parity
    int np;
    uint32_t *count;           // per part: iterations completed since the counts were reset
    uint32_t *status;          // [0] != 0: a wait ran out
    int max_nloc, spin_limit, nclass;
};

// 16-byte agent-scope stores and loads of the partials: buffer accesses with
// the sc1 cache policy (the bits an agent-scope atomic access carries), one
// memory transaction per partial instead of four dword atomics
typedef float res_v4f __attribute__((ext_vector_type(4)));
constexpr int RES_SC1 = 16;                     // cache policy: sc1
constexpr int RES_RSRC_W3 = BUF_RSRC_W3;

__device__ __forceinline__ void res_store(__amdgpu_buffer_rsrc_t r, int slot, float4 v) {
    const res_v4f x = {v.x, v.y, v.z, v.w};
    __builtin_amdgcn_raw_buffer_store_b128(x, r, slot * 16, 0, RES_SC1);
}

__device__ __forceinline__ float4 res_load(__amdgpu_buffer_rsrc_t r, int slot) {
    const res_v4f x = __builtin_amdgcn_raw_buffer_load_b128(r, slot * 16, 0, RES_SC1);
    return make_float4(x.x, x.y, x.z, x.w);
}

// one element per thread per block (block_tets == blockDim.x), at most VPT
// part-local vertices per thread.  LDS: positions, accumulators, previous
// positions (damping), the corner corrections of one block, the class table.
template <int FEAT, int DAMP, int VPT>
__global__ __launch_bounds__(256) void k_resident(IterArgs a, BlockArgs b, ResArgs rs, const float4 *__restrict__ xin,
                                                  float4 *__restrict__ xout, const float4 *__restrict__ pprev,
                                                  int iters) {
    extern __shared__ float4 smem[];
    __shared__ uint32_t s_base, s_fail;
    const int M = rs.max_nloc, TB = b.block_tets;
    float4 *sx = smem, *sacc = smem + M;
    float4 *spp = smem + 2 * M;
    float4 *sc = smem + (DAMP ? 3 : 2) * M;
    DevClass *scls = (DevClass *)(sc + 4 * TB);
    const int tid = threadIdx.x, nth = blockDim.x;
    const int p = (int)xcd_block(blockIdx.x, gridDim.x);
    const PartDesc P = rs.parts[p];
    if (tid == 0) {
        s_base = __hip_atomic_load(rs.count + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_fail = __hip_atomic_load(rs.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    for (int i = tid; i < rs.nclass; i += nth) scls[i] = a.cls[i];
    // the elements of a round (one block) are loaded one round ahead
    ElemIn<FEAT, 0> e;
    ushort4 pl = make_ushort4(0, 0, 0, 0), cp = pl;
    {
        const uint32_t g = (uint32_t)P.b0 * TB + tid;
        if ((int64_t)g < a.nptet) {
            load_elem(a, g, e);
            pl = ld_off(rs.plidx, g * 8u);
            cp = ld_off(b.bcpos, g * 8u);
        }
    }
    // the part's vertices: state into LDS; their table entries, partial
    // ranges and pin records in registers
    int4 pr[VPT];
    int2 so[VPT];
    float4 pv[VPT];
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
        const int j = tid + k * nth;
        pr[k] = make_int4(0, -1, 0, 0);
        so[k] = make_int2(0, 0);
        pv[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j < P.nloc) {
            pr[k] = rs.prec[P.vbeg + j];
            sx[j] = xin[pr[k].x];
            if (DAMP) spp[j] = pprev[pr[k].x];
            sacc[j] = make_float4(0.f, 0.f, 0.f, 0.f);
            pv[k] = pin_of(a, pr[k].x);
            if (pr[k].y >= 0) so[k] = make_int2(b.soff[pr[k].y], b.soff[pr[k].y + 1]);
        }
    }
    __syncthreads();
    if (s_fail) return;
    const uint32_t base = s_base;
    for (int it = 0; it < iters; ++it) {
        const __amdgpu_buffer_rsrc_t xp = __builtin_amdgcn_make_buffer_rsrc(
            rs.xp + (((base + it) & 1u) ? rs.np : 0), 0, rs.np * 16, RES_RSRC_W3);
#pragma unroll 1
        for (int rr = 0; rr < P.nblk; ++rr) {
            const BlockDesc D = b.blk[P.b0 + rr];
            // this block's local entries, used after the projection
            uint32_t run = 0, bm = 0;
            int bps = -1;
            if (tid < D.nloc) {
                run = b.brun[D.vbeg + tid];
                bm = rs.bmap[D.vbeg + tid];
                bps = b.bps[D.vbeg + tid];
            }
            // the next round's elements (the first block's again for the next iteration)
            const int rn = (rr + 1 < P.nblk) ? rr + 1 : 0;
            const bool more = rr + 1 < P.nblk || it + 1 < iters;
            const uint32_t gn = (uint32_t)(P.b0 + rn) * TB + tid;
            ElemIn<FEAT, 0> en;
            ushort4 pln = make_ushort4(0, 0, 0, 0), cpn = pln;
            if (more && (int64_t)gn < a.nptet) {
                load_elem(a, gn, en);
                pln = ld_off(rs.plidx, gn * 8u);
                cpn = ld_off(b.bcpos, gn * 8u);
            }
            if (tid < D.ntet) {
                const int64_t g = (int64_t)D.tbeg + tid;
                const int lv[4] = {pl.x, pl.y, pl.z, pl.w};
                Corr c;
                project_elem<FEAT, DAMP, 0>(a, g, e, sx, spp, lv, c, scls);
                sc[cp.x] = make_float4(c.d[0].x, c.d[0].y, c.d[0].z, c.cnt);
                sc[cp.y] = make_float4(c.d[1].x, c.d[1].y, c.d[1].z, c.cnt);
                sc[cp.z] = make_float4(c.d[2].x, c.d[2].y, c.d[2].z, c.cnt);
                sc[cp.w] = make_float4(c.d[3].x, c.d[3].y, c.d[3].z, c.cnt);
            }
            e = en;
            pl = pln;
            cp = cpn;
            __syncthreads();
            if (tid < D.nloc) {
                const int s0 = (int)(run & 0xffffu), n = (int)(run >> 16);
                float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
                for (int j = 0; j < n; ++j) {
                    const float4 q = sc[s0 + j];
                    ax += q.x;
                    ay += q.y;
                    az += q.z;
                    aw += q.w;
                }
                if (bm >> 31) {
                    res_store(xp, bps, make_float4(ax, ay, az, aw));
                } else {
                    float4 &acc = sacc[bm & 0x7fffffffu];
                    acc.x += ax;
                    acc.y += ay;
                    acc.z += az;
                    acc.w += aw;
                }
            }
            __syncthreads();
        }
        // the vertices only this part touches: every projection of this
        // iteration has read its positions, so each is finished in place
#pragma unroll
        for (int k = 0; k < VPT; ++k) {
            const int j = tid + k * nth;
            if (j < P.nloc && pr[k].y < 0) {
                const float4 acc = sacc[j];
                sx[j] = finish_value(a, pr[k].x, sx[j], DAMP ? spp[j] : make_float4(0, 0, 0, 0), acc.x, acc.y,
                                     acc.z, acc.w, pv[k]);
                sacc[j] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        if (P.nnb > 0) {
            // publish: once every thread's partial stores of iteration it have
            // completed (agent-scope stores: complete = at the agent's point of
            // coherence), the part's count says so.  Nothing else crosses
            // workgroups, so no cache write-back or invalidate is needed.
            const uint32_t want = base + (uint32_t)it + 1u;
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (tid == 0) __hip_atomic_store(rs.count + p, want, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // wait for the neighbours' partials of iteration it: one neighbour
            // per thread, a bounded number of polls
            if (!s_fail) {
                for (int q = tid; q < P.nnb; q += nth) {
                    const int nb = rs.nbr[P.nbeg + q];
                    int polls = 0;
                    while ((int)(__hip_atomic_load(rs.count + nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - want) < 0) {
                        if (++polls > rs.spin_limit ||
                            __hip_atomic_load(rs.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                            __hip_atomic_store(rs.status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            s_fail = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                }
            }
            // the partial loads below are agent-scope loads issued after the
            // counts were seen (the barrier waits for the polls' results)
            __syncthreads();
            // cross vertices: all block partials in slot order (k_block_shared)
#pragma unroll
            for (int k = 0; k < VPT; ++k) {
                const int j = tid + k * nth;
                if (j < P.nloc && pr[k].y >= 0) {
                    float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
                    for (int s = so[k].x; s < so[k].y; ++s) {
                        const float4 q = res_load(xp, s);
                        ax += q.x;
                        ay += q.y;
                        az += q.z;
                        aw += q.w;
                    }
                    sx[j] = finish_value(a, pr[k].x, sx[j], DAMP ? spp[j] : make_float4(0, 0, 0, 0), ax, ay, az,
                                         aw, pv[k]);
                }
            }
        } else if (tid == 0) {
            __hip_atomic_store(rs.count + p, base + (uint32_t)it + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
    // write back the vertices this part owns
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
        const int j = tid + k * nth;
        if (j < P.nloc && pr[k].z) xout[pr[k].x] = sx[j];
    }
}


// per-element input features of the warm-start network (22 used of 24):
// F - I (9), fibre direction (3), fibre target length (1), previous-substep
// multipliers (3), log10 rest volume, det F - 1,
// |F|^2 - 3, log10 mean inverse mass, log10 ARAP and volume compliances.
// the four features that do not change during a run (rest volume, inverse
// masses, compliances), per element, as elem_features computes them
__device__ __forceinline__ float4 elem_static_features(const IterArgs &a, int64_t g, const float4 *__restrict__ X) {
    const int4 gi = a.gidx[g];
    const uint32_t meta = a.meta[g];
    const DevClass dc = a.cls[meta & 0xffffu];
    float B[9];
    for (int q = 0; q < 9; ++q) B[q] = a.Bp[(int64_t)q * a.nptet + g];
    const float detB = det3(B);
    const float wsum = X[gi.x].w + X[gi.y].w + X[gi.z].w + X[gi.w].w;
    return make_float4(log10f(1.0f / (6.0f * fabsf(detB) + 1e-30f) + 1e-30f), log10f(0.25f * wsum + 1e-30f),
                       log10f(dc.at_arap + 1e-30f), log10f(dc.a_vol + 1e-30f));
}

__global__ void k_static_features(IterArgs a, const float4 *__restrict__ X, float4 *__restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < a.nptet) out[g] = elem_static_features(a, g, X);
}

// sfeat: the element's static features (k_static_features)
__device__ __forceinline__ void elem_features(const IterArgs &a, int64_t g, const float4 *__restrict__ X,
                                              const float4 *__restrict__ base, const float4 *__restrict__ sfeat,
                                              float f[24]) {
    const int4 gi = a.gidx[g];
    const uint32_t meta = a.meta[g];
    const uint32_t mask = (meta >> 16) & 15u;
    const float4 X0 = X[gi.x], X1 = X[gi.y], X2 = X[gi.z], X3 = X[gi.w];
    const V3 c0 = xyz(X0) - xyz(X3), c1 = xyz(X1) - xyz(X3), c2 = xyz(X2) - xyz(X3);
    float B[9], F[9];
    for (int q = 0; q < 9; ++q) B[q] = a.Bp[(int64_t)q * a.nptet + g];
    const float Ds[9] = {c0.x, c1.x, c2.x, c0.y, c1.y, c2.y, c0.z, c1.z, c2.z};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F[3 * i + j] = Ds[3 * i] * B[j] + Ds[3 * i + 1] * B[3 + j] + Ds[3 * i + 2] * B[6 + j];
    for (int i = 0; i < 9; ++i) f[i] = F[i] - ((i % 4 == 0) ? 1.0f : 0.0f);
    const float4 fb = (mask & M_ANISO) ? a.fib[g] : make_float4(0.f, 0.f, 0.f, 1.f);
    f[9] = fb.x; f[10] = fb.y; f[11] = fb.z; f[12] = fb.w;
    const float4 lp = base[g];
    f[13] = lp.x; f[14] = lp.y; f[15] = lp.z;
    const float4 sf = sfeat[g];
    f[16] = sf.x;
    float n2 = 0.f;
    for (int i = 0; i < 9; ++i) n2 += F[i] * F[i];
    f[17] = det3(F) - 1.0f;
    f[18] = n2 - 3.0f;
    f[19] = sf.y;
    f[20] = sf.z;
    f[21] = sf.w;
    f[22] = 0.f;
    f[23] = 0.f;
}

__global__ void k_features(IterArgs a, const float4 *__restrict__ X, const float4 *__restrict__ base,
                           const float4 *__restrict__ sfeat, float *__restrict__ out,
                           const float *__restrict__ fmean, const float *__restrict__ finv, int normalize) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.nptet) return;
    float f[24];
    elem_features(a, g, X, base, sfeat, f);
    float *o = out + g * 24;
    for (int i = 0; i < 24; ++i) o[i] = normalize ? (f[i] - fmean[i]) * finv[i] : f[i];
}

// ---------------------------------------------------------------------------
// fused warm start: features -> 24x64 -> 64x64 (BF16 MFMA 32x32x16, FP32
// accumulate, ReLU) -> 64x4 (VALU) -> lambda0 = lambda_prev * clamp(1 + s y).
// One wavefront per 32 elements; the operand layout is the one of k_mlp_fwd
// (row l%32, k-half l/32; D column l%32, rows (i/4)*8 + 4h + i%4).  Weights,
// biases and the standardisation are staged in LDS once per workgroup, which
// then walks its tiles grid-stride.
// ---------------------------------------------------------------------------
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 ws_bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float ws_f32x16;

struct WarmNet {
    const uint16_t *W1, *W2;      // bf16 bits, row-major (out, in)
    const float *b1, *b2, *W3, *b3, *fmean, *finv, *oscale;
};

// LDS written by one lane of a wavefront and read by another lane of the same
// wavefront: complete the writes and keep the compiler from moving the reads
// above them (no workgroup barrier: waves work on their own LDS rows)
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// base: lambda_prev in, lambda0 out (the base of the next substep); lam: lambda0 out.
// Each wavefront runs its own 32-element tiles start to finish (its own LDS
// rows, wavefront-level synchronisation only), so the waves of a workgroup
// overlap one another's gathers with their matrix work.
// write_base = 0 keeps the base (the schedule's reference multipliers) for the
// next warm substeps instead of replacing it with lambda0
__global__ __launch_bounds__(256) void k_warm_start(IterArgs a, WarmNet net, const float4 *__restrict__ X,
                                                    float4 *base, const float4 *__restrict__ sfeat,
                                                    float4 *__restrict__ lam, int write_base) {
    constexpr int NI = XPBD_MLP_IN, NH = XPBD_MLP_HIDDEN, NO = XPBD_MLP_OUT;
    __shared__ uint16_t sW1[NH * NI], sW2[NH * NH];
    __shared__ __align__(16) float sW3[NO * NH];
    __shared__ float sb1[NH], sb2[NH], sb3[NO], smu[NI], sinv[NI], sos[NO];
    // activation rows padded to 68 floats: 16-byte aligned rows for the layer-3
    // float4 reads, rows of the two k-halves 16 banks apart for the writes
    constexpr int NHP = NH + 4;
    __shared__ __align__(16) float sH[4][32 * NHP];
    const int tid = threadIdx.x;
    for (int i = tid; i < NH * NI; i += 256) sW1[i] = net.W1[i];
    for (int i = tid; i < NH * NH; i += 256) sW2[i] = net.W2[i];
    for (int i = tid; i < NO * NH; i += 256) sW3[i] = net.W3[i];
    if (tid < NH) { sb1[tid] = net.b1[tid]; sb2[tid] = net.b2[tid]; }
    if (tid < NI) { smu[tid] = net.fmean[tid]; sinv[tid] = net.finv[tid]; }
    if (tid < NO) { sb3[tid] = net.b3[tid]; sos[tid] = net.oscale[tid]; }
    __syncthreads();
    const int wave = tid >> 6, l = tid & 63, col = l & 31, h = l >> 5;
    float *H = sH[wave];
    const int64_t nwt = (a.nptet + 31) / 32;                    // 32-element tiles
    const int64_t wstride = (int64_t)gridDim.x * 4;
    for (int64_t wt = (int64_t)blockIdx.x * 4 + wave; wt < nwt; wt += wstride) {
        const int64_t row = wt * 32 + col;
        const bool rv = row < a.nptet;
        float f[24];
        if (rv) {
            elem_features(a, row, X, base, sfeat, f);
        } else {
#pragma unroll
            for (int i = 0; i < 24; ++i) f[i] = 0.f;
        }
        const float4 lp = rv ? base[row] : make_float4(0.f, 0.f, 0.f, 0.f);
        // layer 1: K = 24 padded to 32, two k-steps of 16; this lane holds k = 16s + 8h + i
        ws_f32x16 acc[2];
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[t][i] = 0.f;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            ws_bf16x8 av;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = 16 * s + 8 * h + i;
                // f[k] with k = 16s + 8h + i; only s = 1, h = 1 falls in the padding
                const float fk = s == 0 ? (h ? f[8 + i] : f[i]) : (h ? 0.f : f[16 + i]);
                av[i] = (__bf16)((rv && k < NI) ? (fk - smu[k]) * sinv[k] : 0.f);
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                ws_bf16x8 bv;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int k = 16 * s + 8 * h + i;
                    bv[i] = k < NI ? __builtin_bit_cast(__bf16, sW1[(32 * t + col) * NI + k]) : (__bf16)0.f;
                }
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[t], 0, 0, 0);
            }
        }
        wave_lds_sync();   // the previous tile's layer-3 reads of H are done
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int c = 32 * t + col;
                H[((i >> 2) * 8 + 4 * h + (i & 3)) * NHP + c] = fmaxf(acc[t][i] + sb1[c], 0.f);
            }
        wave_lds_sync();
        // layer 2: K = 64, four k-steps
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[t][i] = 0.f;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            ws_bf16x8 av;
#pragma unroll
            for (int i = 0; i < 8; ++i) av[i] = (__bf16)H[col * NHP + 16 * s + 8 * h + i];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                ws_bf16x8 bv;
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    bv[i] = __builtin_bit_cast(__bf16, sW2[(32 * t + col) * NH + 16 * s + 8 * h + i]);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[t], 0, 0, 0);
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int c = 32 * t + col;
                H[((i >> 2) * 8 + 4 * h + (i & 3)) * NHP + c] = fmaxf(acc[t][i] + sb2[c], 0.f);
            }
        wave_lds_sync();
        // layer 3 on the vector ALUs: this lane computes outputs 2h and 2h+1 of
        // its row, reading the row and the two W3 rows four floats at a time
        if (rv) {
            const float4 *hr = reinterpret_cast<const float4 *>(H + col * NHP);
            const float4 *w0 = reinterpret_cast<const float4 *>(sW3 + (2 * h) * NH);
            const float4 *w1 = reinterpret_cast<const float4 *>(sW3 + (2 * h + 1) * NH);
            float y0 = sb3[2 * h], y1 = sb3[2 * h + 1];
#pragma unroll
            for (int q = 0; q < NH / 4; ++q) {
                const float4 hv = hr[q], a0 = w0[q], a1 = w1[q];
                y0 += hv.x * a0.x;
                y0 += hv.y * a0.y;
                y0 += hv.z * a0.z;
                y0 += hv.w * a0.w;
                y1 += hv.x * a1.x;
                y1 += hv.y * a1.y;
                y1 += hv.z * a1.z;
                y1 += hv.w * a1.w;
            }
            float2 v;
            if (h == 0) {
                const float s0 = fminf(fmaxf(1.0f + y0 * sos[0], 0.0f), 2.0f);
                const float s1 = fminf(fmaxf(1.0f + y1 * sos[1], 0.0f), 2.0f);
                v = make_float2(lp.x * s0, lp.y * s1);
            } else {
                const float s2 = fminf(fmaxf(1.0f + y0 * sos[2], 0.0f), 2.0f);
                v = make_float2(lp.z * s2, 0.0f);
            }
            reinterpret_cast<float2 *>(lam + row)[h] = v;
            if (write_base) reinterpret_cast<float2 *>(base + row)[h] = v;
        }
    }
}


// ---------------------------------------------------------------------------
// per-vertex kernels
// ---------------------------------------------------------------------------
// integrate_kernel (engine.py:431-464) + apply_pintotarget (mg.py:1337)
__device__ __forceinline__ void integrate_vertex(int i, float4 &x, float4 &pp, float4 &v, const float4 *FRC,
                                                 const uint32_t *flags, const float4 *pin, const SimParams &sp) {
    if (flags[i] & 1u) {
        pp.x = x.x; pp.y = x.y; pp.z = x.z;
        v.x = v.y = v.z = 0.f;
    } else {
        V3 ex = V3{sp.gx, sp.gy, sp.gz};
        const float m = v.w;
        if (FRC && m > 1.0e-12f) ex = ex + xyz(FRC[i]) / m;
        pp.x = x.x; pp.y = x.y; pp.z = x.z;
        if (sp.quasi_static) {
            const float dt2 = sp.dt * sp.dt;
            x.x = x.x + dt2 * ex.x; x.y = x.y + dt2 * ex.y; x.z = x.z + dt2 * ex.z;
            v.x = v.y = v.z = 0.f;
        } else {
            const float dmp = 1.0f - sp.veldamping;
            V3 vv = dmp * xyz(v);
            vv = vv + sp.dt * ex;
            v.x = vv.x; v.y = vv.y; v.z = vv.z;
            x.x = x.x + sp.dt * vv.x; x.y = x.y + sp.dt * vv.y; x.z = x.z + sp.dt * vv.z;
        }
    }
    if (pin) {
        const float4 p = pin[i];
        if (p.w != 0.0f) { x.x = p.x; x.y = p.y; x.z = p.z; }
    }
}

__global__ void k_integrate(int nv, float4 *X, float4 *PP, float4 *VEL, const float4 *FRC,
                            const uint32_t *flags, const float4 *pin, SimParams sp) {
    const int i = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    float4 x = X[i], pp = PP[i], v = VEL[i];
    integrate_vertex(i, x, pp, v, FRC, flags, pin, sp);
    X[i] = x;
    PP[i] = pp;
    VEL[i] = v;
}

// SDF lookup on one bone grid; every FP32 step rounded as written so the
// cell index and the push match the CPU oracle bit for bit.
struct SdfHit {
    float phi;
    V3 n;
    int cell;
    int ok;
};

__device__ SdfHit sdf_eval(const float *grid, int res, const float *bn, V3 x) {
#pragma clang fp contract(off)
    const float inv_h = bn[15];
    SdfHit r;
    r.ok = 0; r.phi = 0.f; r.n = mk3(0, 0, 0); r.cell = -1;
    const float *X = bn;            // 3x4 row-major [R|t]
    const float *origin = bn + 12;
    const V3 q = x - V3{X[3], X[7], X[11]};
    const V3 y = V3{X[0] * q.x + X[4] * q.y + X[8] * q.z, X[1] * q.x + X[5] * q.y + X[9] * q.z,
                    X[2] * q.x + X[6] * q.y + X[10] * q.z};
    const float gx = (y.x - origin[0]) * inv_h, gy = (y.y - origin[1]) * inv_h,
                gz = (y.z - origin[2]) * inv_h;
    const float lim = (float)(res - 1);
    if (!(gx >= 0.0f && gy >= 0.0f && gz >= 0.0f && gx <= lim && gy <= lim && gz <= lim)) return r;
    int ix = (int)floorf(gx), iy = (int)floorf(gy), iz = (int)floorf(gz);
    ix = min(ix, res - 2); iy = min(iy, res - 2); iz = min(iz, res - 2);
    const float fx = gx - (float)ix, fy = gy - (float)iy, fz = gz - (float)iz;
    const float *g0 = grid + ((size_t)iz * res + iy) * res + ix;
    const size_t sy = (size_t)res, sz = (size_t)res * res;
    const float c000 = g0[0], c100 = g0[1], c010 = g0[sy], c110 = g0[sy + 1];
    const float c001 = g0[sz], c101 = g0[sz + 1], c011 = g0[sz + sy], c111 = g0[sz + sy + 1];
    const float ux = 1.0f - fx, uy = 1.0f - fy, uz = 1.0f - fz;
    const float c00 = c000 * ux + c100 * fx, c10 = c010 * ux + c110 * fx;
    const float c01 = c001 * ux + c101 * fx, c11 = c011 * ux + c111 * fx;
    const float c0 = c00 * uy + c10 * fy, c1 = c01 * uy + c11 * fy;
    r.phi = c0 * uz + c1 * fz;
    const float dx = ((c100 - c000) * uy + (c110 - c010) * fy) * uz + ((c101 - c001) * uy + (c111 - c011) * fy) * fz;
    const float dy = ((c010 - c000) * ux + (c110 - c100) * fx) * uz + ((c011 - c001) * ux + (c111 - c101) * fx) * fz;
    const float dz = c1 - c0;
    const V3 gw = V3{X[0] * dx + X[1] * dy + X[2] * dz, X[4] * dx + X[5] * dy + X[6] * dz,
                     X[8] * dx + X[9] * dy + X[10] * dz};
    const float gl = sqrtf(gw.x * gw.x + gw.y * gw.y + gw.z * gw.z);
    if (!(gl > 1e-20f)) return r;
    r.n = V3{gw.x / gl, gw.y / gl, gw.z / gl};
    r.cell = (iz * res + iy) * res + ix;
    r.ok = 1;
    return r;
}

// post phase: bone SDF push (collision_solvers, mg.py:1370-1371) then
// update_velocities_kernel (engine.py:582-605); result lands in X (buffer A).
// INTEG = 1 goes on with the next substep's integrate of the same vertex
// (nothing runs between the two inside a frame), one launch instead of two.
// dacc != null: the substep's last DCD pass is applied here first, with the
// arithmetic of k_dcd_apply (averaged 2^-32 m fixed-point sums, accumulators
// and the contact counter reset for the next pass)
template <int INTEG>
__global__ void k_post(int nv, const float4 *Xc, float4 *X, float4 *PP, float4 *VEL,
                       SdfParams sd, int32_t *cells, SimParams sp, int do_sdf, int do_vel, const float4 *FRC,
                       const uint32_t *flags, const float4 *pin, unsigned long long *dacc, uint32_t *dacc_n,
                       uint32_t *dnc) {
    const int i = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    float4 x = Xc[i];
    if (dacc) {
        if (i == 0) {
            dnc[1] = dnc[0];
            dnc[0] = 0u;
        }
        const uint32_t n = dacc_n[i];
        if (n > 0) {
            const float w = (float)n;
            const float inv = 1.0f / DCD_FIX;
            x.x = x.x + ((float)(long long)dacc[3 * i + 0] * inv) / w;
            x.y = x.y + ((float)(long long)dacc[3 * i + 1] * inv) / w;
            x.z = x.z + ((float)(long long)dacc[3 * i + 2] * inv) / w;
            dacc[3 * i + 0] = dacc[3 * i + 1] = dacc[3 * i + 2] = 0ull;
            dacc_n[i] = 0u;
        }
    }
    const float4 pp = PP[i];
    float4 v_new = make_float4(0.f, 0.f, 0.f, 0.f);
    bool vel_set = false;
    if (do_sdf) {
#pragma clang fp contract(off)
        int bb = -1, bc = -1;
        if (pp.w != 0.0f) {
            const V3 p = xyz(x);
            SdfHit best;
            best.ok = 0; best.phi = 0.f; best.n = mk3(0, 0, 0); best.cell = -1;
            const size_t gsz = (size_t)sd.res * sd.res * sd.res;
            const int grp = sd.vgroup[i];
            const int b1 = sd.goff[grp + 1];
            for (int b = sd.goff[grp]; b < b1; ++b) {
                const SdfHit h = sdf_eval(sd.grids + (size_t)b * gsz, sd.res, sd.bone + 16 * b, p);
                if (h.ok && (!best.ok || h.phi < best.phi)) { best = h; bb = b; }
            }
            if (best.ok) {
                bc = best.cell;
                const float err = best.phi - sd.contact_offset;
                if (err < 0.0f && (sd.max_push <= 0.0f || -err <= sd.max_push)) {
                    x.x = p.x - err * best.n.x;
                    x.y = p.y - err * best.n.y;
                    x.z = p.z - err * best.n.z;
                }
            } else {
                bb = -1;
            }
        }
        if (cells) { cells[2 * i] = bb; cells[2 * i + 1] = bc; }
    }
    if (do_vel && !sp.quasi_static) {
        float4 v = VEL[i];
        V3 vv = V3{(x.x - pp.x) / sp.dt, (x.y - pp.y) / sp.dt, (x.z - pp.z) / sp.dt};
        if (sp.const_damp != 0.0f || sp.exp_damp != 1.0f) {
            const float vm = sqrtf(dot(vv, vv));
            if (vm > 1.0e-6f) {
                float vn = vm * sp.exp_damp - sp.const_damp;
                if (vn < 1.0e-6f) vn = 0.0f;
                vv = (vn / vm) * vv;
            } else {
                vv = mk3(0, 0, 0);
            }
        }
        v.x = vv.x; v.y = vv.y; v.z = vv.z;
        v_new = v;
        vel_set = true;
    }
    if (INTEG) {
        float4 v = vel_set ? v_new : VEL[i];
        float4 p2 = pp;
        integrate_vertex(i, x, p2, v, FRC, flags, pin, sp);
        PP[i] = p2;
        VEL[i] = v;
    } else if (vel_set) {
        VEL[i] = v_new;
    }
    X[i] = x;
}

#include "xpbd_bonemesh.h"

__global__ void k_apply_dpw(int nv, float4 *X, const float4 *dpw, const float4 *pin, int do_pin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    float4 x = X[i];
    const float4 d = dpw[i];
    if (d.w > 1e-9f) {   // as finish_value: one reciprocal
        const float inv = 1.0f / d.w;
        x.x = x.x + d.x * inv;
        x.y = x.y + d.y * inv;
        x.z = x.z + d.z * inv;
    }
    if (do_pin && pin) {
        const float4 p = pin[i];
        if (p.w != 0.0f) { x.x = p.x; x.y = p.y; x.z = p.z; }
    }
    X[i] = x;
}

// fascia ride (fascia_ride_kernel, src/muscle/fascia/fascia.py:159-176): each bound fascia
// vertex goes to its barycentric point on the current surface triangle plus the rest normal gap
// along the triangle's unit normal; written to the fascia positions (before its integrate, as the
// reference snaps before it) and to its pin targets (re-applied after every fascia iteration)
__global__ void k_fascia_ride(int n, const float4 *Xm, const int32_t *vid, const int4 *tri, const float4 *w,
                              float4 *Xf, float4 *PINf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 t = tri[i];
    const float4 ww = w[i];
    const V3 a = xyz(Xm[t.x]), b = xyz(Xm[t.y]), c = xyz(Xm[t.z]);
    const V3 cp = ww.x * a + ww.y * b + ww.z * c;
    const V3 nn = cross(b - a, c - a);
    const float l = sqrtf(dot(nn, nn));
    const V3 u = l > 0.f ? nn / l : mk3(0.f, 0.f, 0.f);
    const V3 g = cp + ww.w * u;
    const int k = vid[i];
    const float4 x = Xf[k];
    Xf[k] = make_float4(g.x, g.y, g.z, x.w);
    PINf[k] = make_float4(g.x, g.y, g.z, 1.0f);
}

// fascia squeeze: advect_bone (fascia.py:45-61), once per frame: every fascia vertex moves by
// the change of its follow point since the last frame
__global__ void k_fascia_advect(int n, const float4 *ff, float4 *prev, float4 *Xf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = ff[i], p = prev[i];
    float4 x = Xf[i];
    x.x = x.x + (c.x - p.x);
    x.y = x.y + (c.y - p.y);
    x.z = x.z + (c.z - p.z);
    Xf[i] = x;
    prev[i] = c;
}

// fascia squeeze: fascia_anchor_bone_kernel (fascia.py:64-78), the anchors snapped to their
// follow point plus the rest offset at the start of the fascia substep
__global__ void k_fascia_anchor(int n, const int32_t *anc, const float4 *off, const float4 *ff, float4 *Xf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = anc[i];
    const float4 c = ff[k], o = off[i];
    const float4 x = Xf[k];
    Xf[k] = make_float4(c.x + o.x, c.y + o.y, c.z + o.z, x.w);
}

// fascia squeeze: solve_glue_kernel (fascia.py:81-125) with the multiplier at zero (the reference
// zeroes it before every launch); the fascia side is kinematic, the muscle vertex moves by
// omega * dl * w_m along the pair; every muscle vertex is in one pair at most
__global__ void k_fascia_glue(int n, const int4 *pairs, const float4 *Xf, float4 *Xm, float inv_dt2, float omega) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int4 q = pairs[c];
    const float4 xf = Xf[q.x];
    float4 xm = Xm[q.y];
    const V3 nn = V3{xf.x - xm.x, xf.y - xm.y, xf.z - xm.z};
    const float d = sqrtf(dot(nn, nn));
    if (d < 1.0e-9f) return;
    const V3 dirn = nn / d;
    const float C = d - __int_as_float(q.z);
    const float wj = xm.w;
    const float alpha = __int_as_float(q.w) * inv_dt2;
    const float denom = wj + alpha;
    if (denom < 1.0e-12f) return;
    const float dl = -C / denom;
    const float f = omega * dl;
    const V3 m = (-f * wj) * dirn;
    xm.x = xm.x + m.x;
    xm.y = xm.y + m.y;
    xm.z = xm.z + m.z;
    Xm[q.y] = xm;
}

__global__ void k_pins(int nv, float4 *X, const float4 *pin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const float4 p = pin[i];
    if (p.w != 0.0f) { float4 x = X[i]; x.x = p.x; x.y = p.y; x.z = p.z; X[i] = x; }
}

// two-level correction (MultiGridSim._hpbd_vcycle, mg.py:1064-1079).
// Restriction, per muscle block b with coarse nodes j (block-contiguous order):
//   [A t]_b = (sum_i x_i [r_i - c_b; 1]^T) S_b^-1        affine fit of the block
//   rhs_j   = sum_i P_ij x_i + eps_b (A_b (q0_j - c_b) + t_b)
//   q_j     = sum_l G_b[j, l] rhs_l,  G_b = (P^T P + eps_b I)^-1
// (restrict_galerkin_scatter_kernel + block_solve_kernel, mg/kernel.py:138-180,
// with the ridge anchored at the block's affine motion).  Then the coarse
// sweeps, and the delta prolongation (prolong_delta_kernel, mg/kernel.py:9-40).
struct MgArgs {
    const int32_t *fit_off, *fit_fine;   // per block: fine vertices (device ids)
    const float4 *fit_rest;              // per fit entry: rest position - block centre
    const float *fit_sinv;               // per block: 4x4
    float *fitT;                         // per block: 3x4 result
    const int32_t *ptoff, *ptfv;         // per coarse node (block order): P^T entries
    const float *ptw;
    const int32_t *nblk_of;              // per coarse node (block order): block id
    const float4 *q0c;                   // per coarse node (block order): rest - block centre
    const float *eps;                    // per block
    float4 *rhs;                         // per coarse node (block order)
    const int32_t *boff;                 // nblk + 1
    const int64_t *goff;                 // per block: offset of G_b
    const float *G;
    const int32_t *dev_of;               // block order -> coarse device id
    float4 *q;                           // coarse positions, block (caller) order
};

// the whole coarse solve of one cage block in one workgroup: nodes and the
// corner corrections of its tets live in LDS for all sweeps (_coarse_arap_sweep,
// mg.py:939-952: clear, ARAP-SVD Jacobi with lambda = 0, apply dP / dPw)
struct MgCoarse {
    const int32_t *tet_off;      // nblk + 1
    const ushort4 *tlidx;        // per tet: block-local node ids
    const float *tB;             // per tet: rest inverse B, 9 floats row-major
    const float *tat;            // per tet: 1 / (k dt^2)
    const float *node_w;         // per node: inverse mass (0 = held)
    const uint8_t *node_pin;     // per node: 1 = coarse pin (never moved by a sweep)
    const float4 *pin_tgt;       // pin rule 1: per node (block order) the restricted bone goal of a pin
    int pin_rule;                // 0: pins keep their restricted positions; 1: they start at pin_tgt
    const int32_t *slot_off;     // per node + 1: its run in the block's vertex-major corner list
    const ushort4 *tcpos;        // per tet: its corners' places in that list (block-local)
    int max_nodes, max_tets;
};

// sum over a wavefront in a fixed tree order
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// per-muscle affine fit: 1,024 threads per muscle, two vertices per thread and
// round with both gathers issued first; wavefront sums, then the 16 wavefront
// partials added in order (a fixed reduction order, so the fit is the same on
// every run)
constexpr int MGF_T = 1024;
__global__ __launch_bounds__(MGF_T) void k_mg_fit(MgArgs m, const float4 *__restrict__ xf) {
    __shared__ float red[12][MGF_T / 64];
    const int b = blockIdx.x, tid = threadIdx.x;
    float acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = 0.f;
    const int e0 = m.fit_off[b], e1 = m.fit_off[b + 1];
    for (int e = e0 + tid; e < e1; e += 2 * MGF_T) {
        const bool two = e + MGF_T < e1;
        const float4 xa = xf[m.fit_fine[e]];
        const float4 ra = m.fit_rest[e];
        const float4 xb = two ? xf[m.fit_fine[e + MGF_T]] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 rb = two ? m.fit_rest[e + MGF_T] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float ha[4] = {ra.x, ra.y, ra.z, 1.0f};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            acc[k] += xa.x * ha[k];
            acc[4 + k] += xa.y * ha[k];
            acc[8 + k] += xa.z * ha[k];
        }
        if (two) {
            const float hb[4] = {rb.x, rb.y, rb.z, 1.0f};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                acc[k] += xb.x * hb[k];
                acc[4 + k] += xb.y * hb[k];
                acc[8 + k] += xb.z * hb[k];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = wave_sum(acc[k]);
    if ((tid & 63) == 0)
#pragma unroll
        for (int k = 0; k < 12; ++k) red[k][tid >> 6] = acc[k];
    __syncthreads();
    if (tid < 12) {
        // this entry's row of sums over the 16 wavefronts, in order
        float sk[4];
        const int row = tid / 4, col = tid % 4;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float v = 0.f;
            for (int w = 0; w < MGF_T / 64; ++w) v += red[4 * row + k][w];
            sk[k] = v;
        }
        const float *S = m.fit_sinv + 16 * b;
        float v = 0.f;
        for (int k = 0; k < 4; ++k) v += sk[k] * S[4 * k + col];
        m.fitT[12 * b + tid] = v;
    }
}

// one wavefront per coarse node: the lanes split its P^T entries
__global__ __launch_bounds__(256) void k_mg_rhs(MgArgs m, int nc, const float4 *__restrict__ xf) {
    const int j = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    const int l = threadIdx.x & 63;
    if (j >= nc) return;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int e = m.ptoff[j] + l; e < m.ptoff[j + 1]; e += 64) {
        const float w = m.ptw[e];
        const float4 x = xf[m.ptfv[e]];
        sx += w * x.x;
        sy += w * x.y;
        sz += w * x.z;
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    sz = wave_sum(sz);
    if (l != 0) return;
    const int b = m.nblk_of[j];
    const float *T = m.fitT + 12 * b;
    const float4 q = m.q0c[j];
    const float ep = m.eps[b];
    sx += ep * (T[0] * q.x + T[1] * q.y + T[2] * q.z + T[3]);
    sy += ep * (T[4] * q.x + T[5] * q.y + T[6] * q.z + T[7]);
    sz += ep * (T[8] * q.x + T[9] * q.y + T[10] * q.z + T[11]);
    m.rhs[j] = make_float4(sx, sy, sz, 0.f);
}

// one wavefront per coarse node: coalesced reads of its row of G_b
__global__ __launch_bounds__(256) void k_mg_solve(MgArgs m, int nc, float4 *__restrict__ qprev) {
    const int j = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    const int l = threadIdx.x & 63;
    if (j >= nc) return;
    const int b = m.nblk_of[j];
    const int s = m.boff[b], n = m.boff[b + 1] - s;
    const float *row = m.G + m.goff[b] + (int64_t)(j - s) * n;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int k = l; k < n; k += 64) {
        const float g = row[k];
        const float4 r = m.rhs[s + k];
        sx += g * r.x;
        sy += g * r.y;
        sz += g * r.z;
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    sz = wave_sum(sz);
    if (l != 0) return;
    const float4 q = make_float4(sx, sy, sz, 0.f);
    m.q[j] = q;
    qprev[j] = q;
}

constexpr int MGC_T = 1024;  // threads per coarse block workgroup

// projection of one coarse tet from the block's LDS node table; its four
// corner corrections go to their places in the block's vertex-major list
__device__ __forceinline__ void mg_coarse_tet(const float4 *sq, float4 *sc, ushort4 li, ushort4 cp, const float *B,
                                              float at) {
    const float4 X0 = sq[li.x], X1 = sq[li.y], X2 = sq[li.z], X3 = sq[li.w];
    const V3 c0 = xyz(X0) - xyz(X3), c1 = xyz(X1) - xyz(X3), c2 = xyz(X2) - xyz(X3);
    const float w[4] = {X0.w, X1.w, X2.w, X3.w};
    float F[9];
    const float Ds[9] = {c0.x, c1.x, c2.x, c0.y, c1.y, c2.y, c0.z, c1.z, c2.z};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            F[3 * i + j] = Ds[3 * i] * B[j] + Ds[3 * i + 1] * B[3 + j] + Ds[3 * i + 2] * B[6 + j];
    V3 gr[4];
    float wsum;
    const float C = arap_grad(F, B, w, gr, wsum);
    float dL = 0.f, cnt = 0.f;
    if (C > 0.0f && wsum != 0.0f) {
        dL = -C * __builtin_amdgcn_rcpf(wsum + at);
        cnt = 1.0f;
    } else {
        gr[0] = gr[1] = gr[2] = gr[3] = mk3(0.f, 0.f, 0.f);
    }
    const V3 d0 = (dL * w[0]) * gr[0], d1 = (dL * w[1]) * gr[1], d2 = (dL * w[2]) * gr[2], d3 = (dL * w[3]) * gr[3];
    sc[cp.x] = make_float4(d0.x, d0.y, d0.z, cnt);
    sc[cp.y] = make_float4(d1.x, d1.y, d1.z, cnt);
    sc[cp.z] = make_float4(d2.x, d2.y, d2.z, cnt);
    sc[cp.w] = make_float4(d3.x, d3.y, d3.z, cnt);
}

// all sweeps of one muscle's cage in one workgroup.  A thread keeps its (at
// most two) tets' data in registers and its node's corner range across the
// sweeps, so a sweep touches only LDS: project every tet, barrier, each node
// adds its contiguous corner run (ascending (tet, corner) order) and applies
// dP / dPw, barrier.  A coarse pin takes part in the projections with its
// inverse mass and keeps its restricted position (the fine level's pin rule:
// apply, then snap back to the target, mg.py:1337,1355).
__global__ __launch_bounds__(MGC_T) void k_mg_coarse(MgArgs m, MgCoarse mc, int iters, float4 *__restrict__ xc) {
    extern __shared__ float4 smem[];
    float4 *sq = smem;                         // xyz, inverse mass
    float4 *sc = smem + mc.max_nodes;          // corner corrections, vertex-major
    const int b = blockIdx.x, tid = threadIdx.x;
    const int n0 = m.boff[b], nn = m.boff[b + 1] - n0;
    const int t0 = mc.tet_off[b], nt = mc.tet_off[b + 1] - t0;
    const bool has0 = tid < nt, has1 = tid + MGC_T < nt;
    ushort4 li0 = make_ushort4(0, 0, 0, 0), li1 = li0, cp0 = li0, cp1 = li0;
    float B0[9], B1[9], at0 = 0.f, at1 = 0.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) B0[q] = B1[q] = 0.f;
    if (has0) {
        const int t = t0 + tid;
        li0 = mc.tlidx[t];
        cp0 = mc.tcpos[t];
        at0 = mc.tat[t];
#pragma unroll
        for (int q = 0; q < 9; ++q) B0[q] = mc.tB[9 * (int64_t)t + q];
    }
    if (has1) {
        const int t = t0 + tid + MGC_T;
        li1 = mc.tlidx[t];
        cp1 = mc.tcpos[t];
        at1 = mc.tat[t];
#pragma unroll
        for (int q = 0; q < 9; ++q) B1[q] = mc.tB[9 * (int64_t)t + q];
    }
    if (tid < nn) {
        float4 q = m.q[n0 + tid];
        // the reference's rule: a pin moves to its restricted bone goal, then is held
        // (_coarse_kinematic_pin, mg.py:999-1011; zero inverse mass via coarse_stopped)
        if (mc.pin_rule && mc.node_pin[n0 + tid]) q = mc.pin_tgt[n0 + tid];
        sq[tid] = make_float4(q.x, q.y, q.z, mc.node_w[n0 + tid]);
    }
    // a node's corner run is summed by two adjacent lanes (even and odd
    // places), which then combine, when the cage has at most MGC_T / 2 nodes
    const int lpn = (2 * nn <= MGC_T) ? 2 : 1;
    const int node = lpn == 2 ? (tid >> 1) : tid, part = lpn == 2 ? (tid & 1) : 0;
    const bool hasn = node < nn;
    int s0 = 0, ns = 0;
    bool pinned = false;
    if (hasn) {
        pinned = mc.node_pin[n0 + node] != 0;
        const int base = mc.slot_off[n0];
        s0 = mc.slot_off[n0 + node] - base;
        ns = mc.slot_off[n0 + node + 1] - mc.slot_off[n0 + node];
    }
    for (int it = 0; it < iters; ++it) {
        __syncthreads();
        if (has0) mg_coarse_tet(sq, sc, li0, cp0, B0, at0);
        if (has1) mg_coarse_tet(sq, sc, li1, cp1, B1, at1);
        __syncthreads();
        float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
        if (hasn)
            for (int j = part; j < ns; j += lpn) {
                const float4 c = sc[s0 + j];
                ax += c.x;
                ay += c.y;
                az += c.z;
                aw += c.w;
            }
        if (lpn == 2) {   // the node's two lanes are adjacent lanes of one wavefront
            ax += __shfl_xor(ax, 1);
            ay += __shfl_xor(ay, 1);
            az += __shfl_xor(az, 1);
            aw += __shfl_xor(aw, 1);
        }
        if (hasn && part == 0 && !pinned && aw > 1e-9f) {   // apply_dP_kernel, engine.py:689-698
            float4 x = sq[node];
            x.x = x.x + ax / aw;
            x.y = x.y + ay / aw;
            x.z = x.z + az / aw;
            sq[node] = x;
        }
    }
    __syncthreads();
    if (tid < nn) {
        const float4 x = sq[tid];
        m.q[n0 + tid] = make_float4(x.x, x.y, x.z, 0.f);
        float4 *o = xc + m.dev_of[n0 + tid];   // the coarse context's own state, for readback
        *o = make_float4(x.x, x.y, x.z, o->w);
    }
}

__global__ void k_mg_prolong(int nv, const int4 *__restrict__ pidx, const float4 *__restrict__ pw,
                             const float4 *__restrict__ xc, const float4 *__restrict__ qprev, float4 *__restrict__ xf,
                             float4 *__restrict__ pp, const float4 *__restrict__ pin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    if (pin && pin[i].w != 0.0f) return;   // pinned vertices are kinematic (snapped right after)
    const int4 ia = pidx[2 * i], ib = pidx[2 * i + 1];
    const float4 wa = pw[2 * i], wb = pw[2 * i + 1];
    const int id[8] = {ia.x, ia.y, ia.z, ia.w, ib.x, ib.y, ib.z, ib.w};
    const float w[8] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
    float dx = 0.f, dy = 0.f, dz = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float4 a = xc[id[k]], b = qprev[id[k]];
        dx += w[k] * (a.x - b.x);
        dy += w[k] * (a.y - b.y);
        dz += w[k] * (a.z - b.z);
    }
    float4 x = xf[i];
    x.x += dx;
    x.y += dy;
    x.z += dz;
    xf[i] = x;
    if (pp) {  // kinematic correction: the shift does not enter the velocity
        float4 q = pp[i];
        q.x += dx;
        q.y += dy;
        q.z += dz;
        pp[i] = q;
    }
}

__global__ void k_zero_i32(int n, int32_t *p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_copy4(int64_t n, float4 *__restrict__ dst, const float4 *__restrict__ src) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_zero4(int64_t n, float4 *p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// fiberscale = 1 - belly * activation[tetid] * contraction_ratio (solve_kernels.py:1388-1389)
__global__ void k_refresh_fiber(int64_t nptet, const float4 *fstat, const int32_t *tetid,
                                const float *act, float cr, float4 *fib) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nptet) return;
    float4 f = fstat[i];
    f.w = 1.0f - f.w * act[tetid[i]] * cr;
    fib[i] = f;
}

// the same from per-group activations: the element's tet's group selects the value
__global__ void k_refresh_fiber_groups(int64_t nptet, const float4 *fstat, const int32_t *egroup,
                                       const float *gact, float cr, float4 *fib) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nptet) return;
    float4 f = fstat[i];
    f.w = 1.0f - f.w * gact[egroup[i]] * cr;
    fib[i] = f;
}

// residual sums per type over primary elements: out per block 8 doubles
template <int BS>
__global__ __launch_bounds__(BS) void k_residual(IterArgs a, const float4 *__restrict__ X,
                                                 const int32_t *__restrict__ pvert,
                                                 double *partial) {
    __shared__ double red[BS];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t g = (int64_t)blockIdx.x * BS + threadIdx.x; g < a.nptet; g += (int64_t)gridDim.x * BS) {
        const uint32_t meta = a.meta[g];
        if (!((meta >> 20) & 1u)) continue;
        const uint32_t mask = (meta >> 16) & 15u;
        const ClassParams cp = a.cls[meta & 0xffffu].p;
        const int32_t *pv = pvert + 4 * g;
        const V3 P0 = xyz(X[pv[0]]), P1 = xyz(X[pv[1]]), P2 = xyz(X[pv[2]]), P3 = xyz(X[pv[3]]);
        const V3 c0 = P0 - P3, c1 = P1 - P3, c2 = P2 - P3;
        float B[9], F[9];
        for (int q = 0; q < 9; ++q) B[q] = a.Bp[(int64_t)q * a.nptet + g];
        const float Ds[9] = {c0.x, c1.x, c2.x, c0.y, c1.y, c2.y, c0.z, c1.z, c2.z};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                F[3 * i + j] = Ds[3 * i] * B[j] + Ds[3 * i + 1] * B[3 + j] + Ds[3 * i + 2] * B[6 + j];
        if (mask & M_ARAP) {
            float R[9];
            closest_rotation(F, R);
            float s = 0.f;
            for (int i = 0; i < 9; ++i) s += (F[i] - R[i]) * (F[i] - R[i]);
            acc[0] += s;
            acc[1] += 1.0;
        }
        if (mask & M_VOL) {
            const float vol = dot(cross(c0, c1), c2) / 6.0f;
            const double dv = (double)vol - (double)a.rvol[g];
            acc[2] += dv * dv;
            acc[3] += 1.0;
        }
        if (mask & M_SNH) {
            float Ic = 0.f;
            for (int i = 0; i < 9; ++i) Ic += F[i] * F[i];
            const double cd = (double)sqrtf(fmaxf(Ic, 1e-12f)) - 1.7320508075688772;
            const double ch = (double)det3(F) - (1.0 + (double)cp.mu / (double)cp.lam);
            acc[4] += cd * cd + ch * ch;
            acc[5] += 1.0;
        }
        if (mask & M_ANISO) {
            const float4 fb = a.fib[g];
            const V3 d = V3{F[0] * fb.x + F[1] * fb.y + F[2] * fb.z, F[3] * fb.x + F[4] * fb.y + F[5] * fb.z,
                            F[6] * fb.x + F[7] * fb.y + F[8] * fb.z};
            const double s = sqrt((double)dot(d, d)) - (double)fb.w;
            const double C = 0.5 * s * s;
            acc[6] += C * C;
            acc[7] += 1.0;
        }
    }
    for (int k = 0; k < 8; ++k) {
        red[threadIdx.x] = acc[k];
        __syncthreads();
        for (int s = BS / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(size_t)blockIdx.x * 8 + k] = red[0];
        __syncthreads();
    }
}

// without contraction, as the oracle: the integer count is bit-exact also for
// tets whose determinant is zero up to rounding
__global__ void k_count_inverted(int nt, const int4 *tets, const float *B, const float4 *X,
                                 int *count) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    int inv = 0;
    if (c < nt) {
        const int4 t = tets[c];
        const V3 P3 = xyz(X[t.w]);
        const V3 c0 = xyz(X[t.x]) - P3, c1 = xyz(X[t.y]) - P3, c2 = xyz(X[t.z]) - P3;
        const float *b = B + 9 * (size_t)c;
        const float Ds[9] = {c0.x, c1.x, c2.x, c0.y, c1.y, c2.y, c0.z, c1.z, c2.z};
        float F[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                F[3 * i + j] = Ds[3 * i] * b[j] + Ds[3 * i + 1] * b[3 + j] + Ds[3 * i + 2] * b[6 + j];
        const float det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                          F[2] * (F[3] * F[7] - F[4] * F[6]);
        inv = det <= 0.0f ? 1 : 0;
    }
    // integer reduction is order-free
    if (inv) atomicAdd(count, 1);
}


// ---------------------------------------------------------------------------
// tri-tri DCD kernels (semantics: src/muscle/collision/udcf/kernel.py)
// ---------------------------------------------------------------------------
#define DCD_THRESH 1.0e-5f

__device__ __forceinline__ uint32_t dcd_hash(int x, int y, int z, int level, int bits) {
    const uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u) ^
                       ((uint32_t)level * 2654435761u);
    return h & ((1u << bits) - 1u);
}

// centroid_radius_kernel (:47-63) + face_normal_kernel (:66-84) + a log2
// histogram of the radii, in place of max_radius_reduce_kernel (:87-100): the
// grid hierarchy takes its base cell from the radius distribution, not from
// the maximum.  The histogram is reduced in the workgroup first: one
// device-scope atomic per non-empty bin rather than per triangle.
__global__ __launch_bounds__(256) void k_dcd_tridata(DcdDev d, const float4 *X) {
#pragma clang fp contract(off)
    __shared__ uint32_t sh[DCD_HIST];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x < DCD_HIST) sh[threadIdx.x] = 0u;
    __syncthreads();
    if (i < d.ntris) {
    const int4 t = d.tris[i];
    const V3 a = xyz(X[t.x]), b = xyz(X[t.y]), c = xyz(X[t.z]);
    const V3 s = (a + b) + c;
    const V3 ce = V3{s.x / 3.0f, s.y / 3.0f, s.z / 3.0f};
    const V3 e0 = a - ce, e1 = b - ce, e2 = c - ce;
    const float r0 = sqrtf(dot(e0, e0)), r1 = sqrtf(dot(e1, e1)), r2 = sqrtf(dot(e2, e2));
    float r = fmaxf(fmaxf(r0, r1), r2);
    if (!(isfinite(r) && isfinite(ce.x) && isfinite(ce.y) && isfinite(ce.z))) r = INFINITY;
    d.cr[i] = make_float4(ce.x, ce.y, ce.z, r);
    const V3 n = cross(b - a, c - a);
    const float ln = sqrtf(dot(n, n));
    d.nrm[i] = ln > 0.0f ? make_float4(n.x / ln, n.y / ln, n.z / ln, 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
    if (isfinite(r)) {
        // bin k holds r in [2^((k-80)/4), 2^((k-79)/4)); the ends are open
        const int k = r > 0.0f ? (int)floorf(4.0f * log2f(r)) + 80 : 0;
        atomicAdd(&sh[min(max(k, 0), DCD_HIST - 1)], 1u);
    }
    }
    __syncthreads();
    if (threadIdx.x < DCD_HIST && sh[threadIdx.x]) atomicAdd(&d.hist[threadIdx.x], sh[threadIdx.x]);
}

// base cell c0 = 1.01 (2 r50 + thickness), r50 the upper edge of the histogram
// bin holding the median radius (or the caller's fixed cell): triangles up to
// the median size live at level 0, larger ones one or more levels up, which
// keeps the candidates per search low when the sizes are spread out
__global__ void k_dcd_cellsize(DcdDev d) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float c0 = d.cell_fixed;
    if (!(c0 > 0.0f)) {
        uint32_t tot = 0;
        for (int k = 0; k < DCD_HIST; ++k) tot += d.hist[k];
        const uint32_t want = (uint32_t)(((uint64_t)tot + 1u) / 2u);
        uint32_t acc = 0;
        int k = 0;
        for (; k < DCD_HIST - 1; ++k) {
            acc += d.hist[k];
            if (acc >= want) break;
        }
        const float r50 = exp2f((float)(k - 79) * 0.25f);
        c0 = 1.01f * fmaxf(2.0f * r50 + d.thickness, 1.0e-4f);
    }
    d.scal[0] = c0;
    d.scal[1] = 1.0f / c0;
}

// level b cell and its inverse: exact power-of-two scalings of the base
__device__ __forceinline__ float dcd_cell(const DcdDev &d, int b) { return ldexpf(d.scal[0], b); }
__device__ __forceinline__ float dcd_inv(const DcdDev &d, int b) { return ldexpf(d.scal[1], -b); }

// level, cell and hash key of one triangle; counts per level into the workgroup
__device__ __forceinline__ void dcd_key_one(DcdDev &d, int i, uint32_t *sl) {
    const float4 c = d.cr[i];
    const float need = 2.0f * c.w + d.thickness;
    int lv = 0;
    while (lv < DCD_LEVELS && !(need <= dcd_cell(d, lv) * (1.0f / 1.01f))) ++lv;
    if (lv == DCD_LEVELS) {   // non-finite or above the top level
        d.tlvl[i] = -1;
        d.cell[i] = make_int4(0, 0, 0, -1);
        atomicAdd(&d.overflow[0], 1u);
        return;
    }
    const float inv = dcd_inv(d, lv);
    if (!(fabsf(c.x) * inv < 1.0e9f && fabsf(c.y) * inv < 1.0e9f && fabsf(c.z) * inv < 1.0e9f)) {
        d.tlvl[i] = -1;   // cell coordinates would leave the int range
        d.cell[i] = make_int4(0, 0, 0, -1);
        atomicAdd(&d.overflow[0], 1u);
        return;
    }
    const int x = (int)floorf(c.x * inv), y = (int)floorf(c.y * inv), z = (int)floorf(c.z * inv);
    const uint32_t h = dcd_hash(x, y, z, lv, d.tbits);
    d.tlvl[i] = lv;
    d.cell[i] = make_int4(x, y, z, (int)h);
    atomicAdd(&d.bucket_cnt[h], 1u);
    atomicAdd(&sl[lv], 1u);
}

__global__ __launch_bounds__(256) void k_dcd_keys(DcdDev d) {
    __shared__ uint32_t sl[DCD_LEVELS];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x < DCD_LEVELS) sl[threadIdx.x] = 0u;
    __syncthreads();
    if (i < d.ntris) dcd_key_one(d, i, sl);
    __syncthreads();
    // the search only asks whether a level holds triangles: a flag, set with a
    // plain store (every writer stores 1) once this workgroup's copy reads 0,
    // instead of an atomic count on one of a dozen addresses per workgroup
    if (threadIdx.x < DCD_LEVELS && sl[threadIdx.x] && d.lvl_cnt[threadIdx.x] == 0u) d.lvl_cnt[threadIdx.x] = 1u;
}


__global__ void k_dcd_scatter(DcdDev d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.ntris) return;
    if (d.tlvl[i] < 0) return;
    const uint32_t h = (uint32_t)d.cell[i].w;
    const uint32_t q = d.bucket_off[h] + atomicAdd(&d.bucket_fill[h], 1u);
    const int4 c = d.cell[i];
    const int4 t = d.tris[i];
    d.rec[q] = d.cr[i];
    d.key[q] = make_int4(i | (d.tlvl[i] << DCD_LVL_SHIFT), c.x, c.y, c.z);
    d.tv[q] = make_int4(t.x, t.y, t.z, d.group ? d.group[i] : 0);
}

// emit rules of broadphase_find_pairs_kernel (:851-915) for a candidate j of i
// that passed the sphere test; tj = j's vertex ids, w = its group
__device__ __forceinline__ bool dcd_emit(const DcdDev &d, int j, const int4 ti, float4 ni, int gi, int4 tj) {
#pragma clang fp contract(off)
    const int gj = tj.w;
    if (gi == 1 && gj == 1) return false;
    if (gi == 0 && gj == 0) {
        if (!d.allow_self) return false;
        if (ti.x == tj.x || ti.x == tj.y || ti.x == tj.z || ti.y == tj.x || ti.y == tj.y || ti.y == tj.z ||
            ti.z == tj.x || ti.z == tj.y || ti.z == tj.z)
            return false;
        if (d.normal_cull < 1.0f) {
            const float4 nj = d.nrm[j];
            if (ni.x * nj.x + ni.y * nj.y + ni.z * nj.z > d.normal_cull) return false;
        }
    }
    return true;
}

// Broadphase search, one pass.  DCD_LANES consecutive lanes of a wavefront
// share one triangle (taken in bucket order, so a wavefront searches a few
// neighbouring cells).  Per grid level the group first reads the bucket ranges
// of all cells of its search box at once (lane L takes cells L and L + 16, in
// z, y, x order) and scans their sizes; the cells up to the visit limit are
// kept (whole cells, as a cell-by-cell walk would keep them) and their
// candidates form one flattened range that the lanes split evenly,
// DCD_UNROLL candidates per lane per round with all loads issued first (a box
// wider than three cells, from rounding at a level boundary, is walked cell by
// cell).  Emitted pairs (min, max) are staged in LDS and appended to the pair
// list with one cursor atomic per workgroup (past a full stage, one cursor
// atomic per wavefront and emit point).
constexpr int DCD_LANES = 16;
constexpr int DCD_BOX = 27;        // cells of one search box (at most 3 per axis)
constexpr int DCD_UNROLL = 2;

// candidate j (bucket slot q, bucket-ordered record cj / key kj) of triangle i
// searching cell (cx, cy, cz) of level b: sphere test, own-cell rule, index
// rule, emit rules; an emitted pair goes to the workgroup's LDS stage
__device__ __forceinline__ void dcd_try(const DcdDev &d, int i, int la, int b, int4 ti, float4 ni, float4 ci,
                                        float4 cj, int4 kj, int cx, int cy, int cz, uint32_t q, int2 *stage,
                                        uint32_t *s_n) {
#pragma clang fp contract(off)
    const V3 dd = V3{ci.x - cj.x, ci.y - cj.y, ci.z - cj.z};
    if (sqrtf(dot(dd, dd)) > ci.w + cj.w + d.thickness) return;
    const int j = kj.x & ((1 << DCD_LVL_SHIFT) - 1);
    // visit j only from its own cell and level
    if ((kj.x >> DCD_LVL_SHIFT) != b || kj.y != cx || kj.z != cy || kj.w != cz) return;
    if (j == i || (b == la && j < i)) return;
    if (!dcd_emit(d, j, ti, ni, ti.w, d.tv[q])) return;
    const int2 pr = make_int2(min(i, j), max(i, j));
    // the lanes of the wavefront that emit here take consecutive stage places
    // with one LDS atomic; lanes past the stage's end append with one cursor
    // atomic for all of them
    const int lane = (int)__lane_id();
    const uint64_t m = __ballot(1);
    const int lead = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == lead) base = atomicAdd(s_n, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, lead);
    const uint32_t p = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (p < (uint32_t)DCD_STAGE) {
        stage[p] = pr;
    } else {
        const uint64_t mo = __ballot(1);
        const int lo = __ffsll((unsigned long long)mo) - 1;
        unsigned long long g = 0ull;
        if (lane == lo) g = atomicAdd(d.npairs, (unsigned long long)__popcll(mo));
        const uint32_t glo = (uint32_t)__shfl((int)(uint32_t)g, lo), ghi = (uint32_t)__shfl((int)(uint32_t)(g >> 32), lo);
        const unsigned long long o = (((unsigned long long)ghi << 32) | glo) +
                                     __builtin_amdgcn_mbcnt_hi((uint32_t)(mo >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mo, 0u));
        if (o < (unsigned long long)d.max_pairs) d.pairs[o] = pr;
    }
}

__device__ __forceinline__ void dcd_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(256) void k_dcd_search(DcdDev d) {
#pragma clang fp contract(off)
    constexpr int NG = 256 / DCD_LANES;
    __shared__ int2 stage[DCD_STAGE];
    __shared__ int4 s_cell[NG][DCD_BOX];        // cell x, y, z, first slot - flattened start
    __shared__ uint32_t s_incl[NG][DCD_BOX];    // inclusive end of the cell in the flattened range
    __shared__ uint32_t s_n;
    __shared__ unsigned long long s_base;
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    const int lane = (int)(threadIdx.x & (DCD_LANES - 1));
    const int grp = (int)(threadIdx.x / DCD_LANES);
    // XCD-aware order: neighbouring triangles (bucket order) search on one XCD
    const int64_t q0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * NG + grp;
    const bool live = q0 < (int64_t)d.bucket_off[(size_t)1 << d.tbits];
    bool stop = false;
    if (live) {
        const int4 ki = d.key[q0];
        const int i = ki.x & ((1 << DCD_LVL_SHIFT) - 1);
        const int la = ki.x >> DCD_LVL_SHIFT;
        const int4 ti = d.tv[q0];
        const float4 ci = d.rec[q0];
        const float4 ni = d.normal_cull < 1.0f ? d.nrm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        uint32_t visits = 0;
        for (int b = la; b < DCD_LEVELS && !stop; ++b) {
            if (d.lvl_cnt[b] == 0) continue;
            const float cb = dcd_cell(d, b), ib = dcd_inv(d, b);
            // reach of a partner at level b: r_j <= (cell_b / 1.01 - thickness) / 2
            const float R = 1.01f * (ci.w + 0.5f * (cb * (1.0f / 1.01f) - d.thickness) + d.thickness);
            const int x0 = (int)floorf((ci.x - R) * ib), x1 = (int)floorf((ci.x + R) * ib);
            const int y0 = (int)floorf((ci.y - R) * ib), y1 = (int)floorf((ci.y + R) * ib);
            const int z0 = (int)floorf((ci.z - R) * ib), z1 = (int)floorf((ci.z + R) * ib);
            if (x1 - x0 > 2 || y1 - y0 > 2 || z1 - z0 > 2) {
                // a box wider than 3 cells (rounding at a level boundary): walk it cell by cell
                for (int z = z0; z <= z1 && !stop; ++z)
                    for (int y = y0; y <= y1 && !stop; ++y)
                        for (int x = x0; x <= x1 && !stop; ++x) {
                            const uint32_t h = dcd_hash(x, y, z, b, d.tbits);
                            const uint32_t e0 = d.bucket_off[h], e = d.bucket_off[h + 1];
                            visits += e - e0;
                            if (visits > DCD_MAX_VISIT) {
                                stop = true;
                                break;
                            }
                            for (uint32_t q = e0 + lane; q < e; q += DCD_LANES)
                                dcd_try(d, i, la, b, ti, ni, ci, d.rec[q], d.key[q], x, y, z, q, stage, &s_n);
                        }
                continue;
            }
            const int nx = x1 - x0 + 1, ny = y1 - y0 + 1, nz = z1 - z0 + 1;
            const int ncell = nx * ny * nz;
            // bucket ranges of cells lane and lane + 16
            uint32_t cnt[2] = {0u, 0u}, s0[2] = {0u, 0u};
            int cx[2] = {0, 0}, cy[2] = {0, 0}, cz[2] = {0, 0};
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const int c = lane + hh * DCD_LANES;
                if (c < ncell) {
                    cx[hh] = x0 + c % nx;
                    cy[hh] = y0 + (c / nx) % ny;
                    cz[hh] = z0 + c / (nx * ny);
                    const uint32_t h = dcd_hash(cx[hh], cy[hh], cz[hh], b, d.tbits);
                    s0[hh] = d.bucket_off[h];
                    cnt[hh] = d.bucket_off[h + 1] - s0[hh];
                }
            }
            // inclusive scan of the sizes over the 32 cells
            uint32_t inc[2] = {cnt[0], cnt[1]};
#pragma unroll
            for (int o = 1; o < DCD_LANES; o <<= 1) {
#pragma unroll
                for (int hh = 0; hh < 2; ++hh) {
                    const uint32_t v = (uint32_t)__shfl_up((int)inc[hh], o, DCD_LANES);
                    if (lane >= o) inc[hh] += v;
                }
            }
            inc[1] += (uint32_t)__shfl((int)inc[0], DCD_LANES - 1, DCD_LANES);
            // keep the cells whose running visit count stays within the limit
            uint32_t T = 0u;
            bool over = false;
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const int c = lane + hh * DCD_LANES;
                if (c < ncell) {
                    if (visits + inc[hh] <= DCD_MAX_VISIT) {
                        T = max(T, inc[hh]);
                        s_cell[grp][c] = make_int4(cx[hh], cy[hh], cz[hh], (int)(s0[hh] - (inc[hh] - cnt[hh])));
                        s_incl[grp][c] = inc[hh];
                    } else {
                        over = true;
                    }
                }
            }
#pragma unroll
            for (int o = 1; o < DCD_LANES; o <<= 1) {
                T = max(T, (uint32_t)__shfl_xor((int)T, o, DCD_LANES));
                over = over || (__shfl_xor((int)over, o, DCD_LANES) != 0);
            }
            dcd_wave_sync();
            // the flattened candidates: f = lane, lane + 16, ...
            int c = 0;
            for (uint32_t f0 = (uint32_t)lane; f0 < T; f0 += DCD_UNROLL * DCD_LANES) {
                float4 cj[DCD_UNROLL];
                int4 kj[DCD_UNROLL];
                int cc[DCD_UNROLL];
#pragma unroll
                for (int u = 0; u < DCD_UNROLL; ++u) {
                    const uint32_t f = f0 + (uint32_t)(u * DCD_LANES);
                    cc[u] = -1;
                    if (f < T) {
                        while (s_incl[grp][c] <= f) ++c;
                        cc[u] = c;
                        const uint32_t q = (uint32_t)s_cell[grp][c].w + f;
                        cj[u] = d.rec[q];
                        kj[u] = d.key[q];
                    }
                }
#pragma unroll
                for (int u = 0; u < DCD_UNROLL; ++u) {
                    if (cc[u] < 0) continue;
                    const int4 cl = s_cell[grp][cc[u]];
                    dcd_try(d, i, la, b, ti, ni, ci, cj[u], kj[u], cl.x, cl.y, cl.z,
                            (uint32_t)cl.w + f0 + (uint32_t)(u * DCD_LANES), stage, &s_n);
                }
            }
            dcd_wave_sync();
            visits += T;
            stop = over;
        }
        if (stop && lane == 0) atomicAdd(&d.overflow[1], 1u);
    }
    __syncthreads();
    const uint32_t n = min(s_n, (uint32_t)DCD_STAGE);
    if (threadIdx.x == 0) s_base = n ? atomicAdd(d.npairs, (unsigned long long)n) : 0ull;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
        const unsigned long long o = s_base + t;
        if (o < (unsigned long long)d.max_pairs) d.pairs[o] = stage[t];
    }
}

// the (i, j)-sorted pair list, built on request: count per first index,
// scan, scatter into segments, then sort each segment by the second index
__global__ void k_dcd_segcount(DcdDev d, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) atomicAdd(&d.pair_cnt[d.pairs[k].x], 1u);
}

__global__ void k_dcd_segscatter(DcdDev d, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int2 pr = d.pairs[k];
    d.pairs_seg[d.pair_off[pr.x] + atomicAdd(&d.pair_fill[pr.x], 1u)] = pr;
}

// sort each triangle's segment by partner index.  DCD_LANES lanes share a
// segment; up to 4 DCD_LANES entries each lane ranks its entries against the
// whole segment (partners are distinct) and stores them at their ranks once
// every lane of the group has read (the group lies in one wavefront, whose
// loads complete before the dependent stores); longer segments are
// insertion-sorted by the group's first lane.
__global__ __launch_bounds__(256) void k_dcd_sortseg(DcdDev d) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = (int)(g & (DCD_LANES - 1));
    const int64_t i = g / DCD_LANES;
    int64_t lo = 0, hi = 0;
    if (i < d.ntris) {
        lo = min((int64_t)d.pair_off[i], d.max_pairs);
        hi = min((int64_t)d.pair_off[i + 1], d.max_pairs);
    }
    const int64_t n = hi - lo;
    if (n <= 4 * DCD_LANES) {
        int y[4], rk[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t e = lane + (int64_t)k * DCD_LANES;
            y[k] = e < n ? d.pairs_seg[lo + e].y : 0;
            rk[k] = 0;
        }
        for (int64_t m = 0; m < n; ++m) {
            const int ym = d.pairs_seg[lo + m].y;
#pragma unroll
            for (int k = 0; k < 4; ++k) rk[k] += ym < y[k] ? 1 : 0;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (lane + (int64_t)k * DCD_LANES < n) d.pairs_seg[lo + rk[k]] = make_int2((int)i, y[k]);
        return;
    }
    if (lane != 0) return;
    if (n > DCD_SORT_MAX) {
        atomicAdd(&d.overflow[2], 1u);
        return;
    }
    for (int64_t a = lo + 1; a < hi; ++a) {
        const int2 v = d.pairs_seg[a];
        int64_t b = a - 1;
        while (b >= lo && d.pairs_seg[b].y > v.y) {
            d.pairs_seg[b + 1] = d.pairs_seg[b];
            --b;
        }
        d.pairs_seg[b + 1] = v;
    }
}

__device__ __forceinline__ V3 dcd_pick(V3 a, V3 b, V3 c, int i) { return i == 0 ? a : (i == 1 ? b : c); }

__device__ __forceinline__ int dcd_inside(V3 p1, V3 p2, V3 p3, V3 p) {
#pragma clang fp contract(off)
    const V3 c1 = cross(p2 - p1, p - p2), c2 = cross(p3 - p2, p - p3), c3 = cross(p1 - p3, p - p1);
    return (dot(c1, c2) > 0.0f && dot(c1, c3) > 0.0f) ? 1 : 0;
}

__device__ __forceinline__ V3 dcd_bary(V3 A, V3 B, V3 C, V3 P) {
#pragma clang fp contract(off)
    const V3 v0 = B - A, v1 = C - A, v2 = P - A;
    const float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    const float den = d00 * d11 - d01 * d01;
    if (den == 0.0f) return V3{1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f};
    const float v = (d11 * d20 - d01 * d21) / den;
    const float w = (d00 * d21 - d01 * d20) / den;
    return V3{1.0f - v - w, v, w};
}

__device__ __forceinline__ int dcd_line(V3 p1, V3 p2, V3 p3, V3 p4, V3 v1, V3 v2, V3 &out) {
#pragma clang fp contract(off)
    const V3 d1 = p2 - p1, d2 = p4 - p3;
    const V3 nr = cross(d1, d2);
    if (sqrtf(dot(nr, nr)) < 1.0e-6f) return 0;
    const float den = dot(nr, nr);
    if (den < 1.0e-6f) return 0;
    const V3 v = p3 - p1;
    const float t1 = dot(cross(v, d2), nr) / den;
    const float t2 = dot(cross(v, d1), nr) / den;
    if (t2 >= 0.0f && t2 <= 1.0f && t1 >= 0.0f && t1 <= 1.0f) {
        out = v1 + t1 * (v2 - v1);
        return 1;
    }
    return 0;
}

__device__ __forceinline__ int dcd_gen_t(V3 N, V3 v0, V3 v1, V3 v2, float d, V3 D, float &lo, float &hi) {
#pragma clang fp contract(off)
    const float d0 = dot(N, v0) + d, d1 = dot(N, v1) + d, d2 = dot(N, v2) + d;
    if (d0 <= DCD_THRESH && d1 <= DCD_THRESH && d2 <= DCD_THRESH) return 0;
    if (d0 >= -DCD_THRESH && d1 >= -DCD_THRESH && d2 >= -DCD_THRESH) return 0;
    V3 s0 = v0, s1 = v1, s2 = v2;
    float e0 = d0, e1 = d1, e2 = d2;
    if ((d0 <= 0.0f && d1 >= 0.0f && d2 >= 0.0f) || (d0 >= 0.0f && d1 <= 0.0f && d2 <= 0.0f)) {
        s0 = v1; s1 = v0; e0 = d1; e1 = d0;
    } else if ((d0 >= 0.0f && d1 >= 0.0f && d2 <= 0.0f) || (d0 <= 0.0f && d1 <= 0.0f && d2 >= 0.0f)) {
        s1 = v2; s2 = v1; e1 = d2; e2 = d1;
    }
    const float p0 = dot(D, s0), p1 = dot(D, s1), p2 = dot(D, s2);
    const float tA = p0 + (p1 - p0) * fabsf(e0 / (e0 - e1));
    const float tB = p2 + (p1 - p2) * fabsf(e2 / (e2 - e1));
    if (tA > tB) { lo = tB; hi = tA; } else { lo = tA; hi = tB; }
    return 1;
}

// narrow_kernel (:317-474) with respond_kernel (:714-799) fused in.  The
// straddle test (each triangle's plane cuts the other triangle, overlapping
// intervals on the intersection line) is all that most pairs reach; a pair
// that passes gets its deepest points (C1, C2) and, if the contact is valid,
// its response at once.  The response reads positions that no kernel of the
// pass writes before the apply and sums in fixed point, so fusing it changes
// nothing in the result.
struct DcdGeo {
    V3 q00, q01, q02, q10, q11, q12, N1, N2;
    float d1, d2;
};

__device__ __forceinline__ void dcd_load(const float4 *X, int4 A, int4 Bt, DcdGeo &g) {
    g.q00 = xyz(X[A.x]); g.q01 = xyz(X[A.y]); g.q02 = xyz(X[A.z]);
    g.q10 = xyz(X[Bt.x]); g.q11 = xyz(X[Bt.y]); g.q12 = xyz(X[Bt.z]);
}

// the straddle test on loaded vertices; fills in the unit normals and offsets
__device__ __forceinline__ bool dcd_straddle(DcdGeo &g) {
#pragma clang fp contract(off)
    const V3 n1f = cross(g.q01 - g.q00, g.q02 - g.q00), n2f = cross(g.q11 - g.q10, g.q12 - g.q10);
    const float l1 = sqrtf(dot(n1f, n1f)), l2 = sqrtf(dot(n2f, n2f));
    if (l1 == 0.0f || l2 == 0.0f) return false;
    g.N1 = n1f / l1;
    g.N2 = n2f / l2;
    g.d1 = -dot(g.N1, g.q00);
    g.d2 = -dot(g.N2, g.q10);
    const V3 cp = cross(g.N1, g.N2);
    const float lcp = sqrtf(dot(cp, cp));
    if (lcp < 1.0e-8f) return false;
    const V3 D = cp / lcp;
    float t1, t2, t3, t4;
    if (!dcd_gen_t(g.N2, g.q00, g.q01, g.q02, g.d2, D, t1, t2)) return false;
    if (!dcd_gen_t(g.N1, g.q10, g.q11, g.q12, g.d1, D, t3, t4)) return false;
    return t2 >= t3 && t4 >= t1;
}

__device__ __forceinline__ void dcd_add(const DcdDev &d, int v, V3 c) {
    atomicAdd(&d.acc[3 * v + 0], (unsigned long long)__float2ll_rn(c.x * DCD_FIX));
    atomicAdd(&d.acc[3 * v + 1], (unsigned long long)__float2ll_rn(c.y * DCD_FIX));
    atomicAdd(&d.acc[3 * v + 2], (unsigned long long)__float2ll_rn(c.z * DCD_FIX));
    atomicAdd(&d.acc_n[v], 1u);
}

// contact and response of one pair that passed the straddle test
// returns whether the pair made a valid contact (counted by the caller)
__device__ __forceinline__ bool dcd_pair_contact(const DcdDev &d, const float4 *X, int4 A, int4 Bt, const DcdGeo &g) {
#pragma clang fp contract(off)
    const V3 q00 = g.q00, q01 = g.q01, q02 = g.q02, q10 = g.q10, q11 = g.q11, q12 = g.q12;
    const V3 N1 = g.N1, N2 = g.N2;
    const float d1 = g.d1, d2 = g.d2;
    float C1 = 0.f, C2 = 0.f;
    V3 b1{0, 0, 0}, b2{0, 0, 0}, b3{0, 0, 0}, b4{0, 0, 0};
    for (int j = 0; j < 3; ++j) {
        const V3 q0j = dcd_pick(q00, q01, q02, j), q1j = dcd_pick(q10, q11, q12, j);
        const V3 P0j = q0j - dot(N2, q0j - q10) * N2;
        if (dcd_inside(q10, q11, q12, P0j)) {
            const float cand = dot(N2, q0j) + d2;
            if (C1 < -cand) { C1 = -cand; b1 = dcd_bary(q00, q01, q02, q0j); b2 = dcd_bary(q10, q11, q12, q0j - cand * N2); }
        }
        const float dq = dot(N1, N2);
        if (dq != 0.0f) {
            const V3 Q0j = q1j + (dot(N1, q00 - q1j) / dq) * N2;
            if (dcd_inside(q00, q01, q02, Q0j)) {
                const float cand = dot(N2, Q0j) + d2;
                if (C1 < -cand) { C1 = -cand; b1 = dcd_bary(q00, q01, q02, Q0j); b2 = dcd_bary(q10, q11, q12, Q0j - cand * N2); }
            }
        }
        const V3 P1j = q1j - dot(N1, q1j - q00) * N1;
        if (dcd_inside(q00, q01, q02, P1j)) {
            const float cand = dot(N1, q1j) + d1;
            if (C2 < -cand) { C2 = -cand; b3 = dcd_bary(q10, q11, q12, q1j); b4 = dcd_bary(q00, q01, q02, q1j - cand * N1); }
        }
        const float dq2 = dot(N2, N1);
        if (dq2 != 0.0f) {
            const V3 Q1j = q0j + (dot(N2, q10 - q0j) / dq2) * N1;
            if (dcd_inside(q10, q11, q12, Q1j)) {
                const float cand = dot(N1, Q1j) + d1;
                if (C2 < -cand) { C2 = -cand; b3 = dcd_bary(q10, q11, q12, Q1j); b4 = dcd_bary(q00, q01, q02, Q1j - cand * N1); }
            }
        }
        const int jn = (j + 1) % 3;
        const V3 q0jn = dcd_pick(q00, q01, q02, jn), q1jn = dcd_pick(q10, q11, q12, jn);
        const V3 P0jn = q0jn - dot(N2, q0jn - q10) * N2;
        const V3 P1jn = q1jn - dot(N1, q1jn - q00) * N1;
        for (int kk = 0; kk < 3; ++kk) {
            const int kn = (kk + 1) % 3;
            V3 hit;
            if (dcd_line(P0j, P0jn, dcd_pick(q10, q11, q12, kk), dcd_pick(q10, q11, q12, kn), q0j, q0jn, hit)) {
                const float cand = dot(N2, hit) + d2;
                if (C1 < -cand) { C1 = -cand; b1 = dcd_bary(q00, q01, q02, hit); b2 = dcd_bary(q10, q11, q12, hit - cand * N2); }
            }
            if (dcd_line(P1j, P1jn, dcd_pick(q00, q01, q02, kk), dcd_pick(q00, q01, q02, kn), q1j, q1jn, hit)) {
                const float cand = dot(N1, hit) + d1;
                if (C2 < -cand) { C2 = -cand; b3 = dcd_bary(q10, q11, q12, hit); b4 = dcd_bary(q00, q01, q02, hit - cand * N1); }
            }
        }
    }
    if (C1 <= d.offset_eps || C2 <= d.offset_eps) return false;
    const bool first = C1 < C2;
    const int4 cu4 = first ? A : Bt, cl4 = first ? Bt : A;
    const int cu[3] = {cu4.x, cu4.y, cu4.z}, cl[3] = {cl4.x, cl4.y, cl4.z};
    const V3 nrm = first ? N2 : N1;
    const V3 cb1 = first ? b1 : b3, cb2 = first ? b2 : b4;
    // respond_kernel (:714-799)
    const float pen = first ? C1 : C2;
    float C = pen * d.multiplier;
    if (C == 0.0f) return true;
    float wcu[3], wcl[3];
    for (int j = 0; j < 3; ++j) { wcu[j] = X[cu[j]].w; wcl[j] = X[cl[j]].w; }
    const bool solid_cur = wcu[0] == 0.0f || wcu[1] == 0.0f || wcu[2] == 0.0f;
    const bool solid_clo = wcl[0] == 0.0f || wcl[1] == 0.0f || wcl[2] == 0.0f;
    if (!solid_cur && !solid_clo) C = C * 0.5f;
    const V3 push_clo = solid_clo ? V3{0, 0, 0} : (-C) * nrm;
    const V3 push_cur = solid_cur ? V3{0, 0, 0} : C * nrm;
    V3 vcl[3], vcu[3];
    for (int j = 0; j < 3; ++j) {
        vcl[j] = (xyz(X[cl[j]]) + push_clo) - xyz(d.xprev[cl[j]]);
        vcu[j] = (xyz(X[cu[j]]) + push_cur) - xyz(d.xprev[cu[j]]);
    }
    const V3 gn = nrm / sqrtf(dot(nrm, nrm));
    const V3 vclo = (cb1.x * vcl[0] + cb1.y * vcl[1]) + cb1.z * vcl[2];
    const V3 perp_clo = vclo - dot(vclo, gn) * gn;
    const V3 vcur = (cb2.x * vcu[0] + cb2.y * vcu[1]) + cb2.z * vcu[2];
    const V3 perp_cur = vcur - dot(vcur, gn) * gn;
    V3 diff = perp_cur - perp_clo;
    const float dn = sqrtf(dot(diff, diff));
    if (dn > pen * d.mu_s && dn > 0.0f) diff = ((pen * d.mu_d) / dn) * diff;
    const V3 diffV = 0.5f * diff;
    const V3 diffT = diff - diffV;
    const V3 dcur = push_cur - diffV, dclo = push_clo + diffT;
    for (int j = 0; j < 3; ++j) {
        if (wcu[j] > 0.0f) dcd_add(d, cu[j], dcur);
        if (wcl[j] > 0.0f) dcd_add(d, cl[j], dclo);
    }
    return true;
}

// A workgroup takes DCD_NR rounds of 256 pairs, queues the pairs that pass
// the straddle test in LDS (a few per thousand) and then runs the contact
// stage on the queue with consecutive lanes, so one wavefront, not every
// wavefront that holds a passing pair, goes through the long contact code.
constexpr int DCD_NR = 4;
constexpr int DCD_QCAP = DCD_NR * 256;   // a round's pairs all fit: the queue never overflows

__global__ __launch_bounds__(256) void k_dcd_narrow(DcdDev d, const float4 *X) {
    // launched with 256 threads
    __shared__ int64_t s_q[DCD_QCAP];
    __shared__ uint32_t s_nq, s_nc;
    if (threadIdx.x == 0) s_nc = 0u;
    const int64_t total = (int64_t)min(*d.npairs, (unsigned long long)d.max_pairs);
    const int64_t span = (int64_t)DCD_NR * 256;
    // XCD-aware order over the chunks of span pairs: XCD x (the workgroups
    // i with i mod 8 = x) takes a contiguous range of chunks, so pairs that
    // the search emitted together (neighbouring triangles) meet in one L2
    const int64_t nchunk = (total + span - 1) / span;
    const uint32_t xcd = blockIdx.x % N_XCD, jx = blockIdx.x / N_XCD;
    const uint32_t gx = gridDim.x / N_XCD + (xcd < gridDim.x % N_XCD ? 1u : 0u);
    const int64_t cq = nchunk / N_XCD, cr = nchunk % N_XCD;
    const int64_t c0 = (int64_t)xcd * cq + min((int64_t)xcd, cr), cn = cq + ((int64_t)xcd < cr ? 1 : 0);
    for (int64_t cc = jx; cc < cn; cc += gx) {
        const int64_t base = (c0 + cc) * span;
        if (threadIdx.x == 0) s_nq = 0u;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < DCD_NR; ++r) {
            const int64_t k = base + (int64_t)r * 256 + threadIdx.x;
            if (k >= total) continue;
            const int2 pr = d.pairs[k];
            const int4 A = d.tris[pr.x], Bt = d.tris[pr.y];
            DcdGeo g;
            dcd_load(X, A, Bt, g);
            if (!dcd_straddle(g)) continue;
            s_q[atomicAdd(&s_nq, 1u)] = k;
        }
        __syncthreads();
        const uint32_t nq = s_nq;
        for (uint32_t t = threadIdx.x; t < nq; t += 256) {
            const int2 pr = d.pairs[s_q[t]];
            const int4 A = d.tris[pr.x], Bt = d.tris[pr.y];
            DcdGeo g;
            dcd_load(X, A, Bt, g);
            // the same test as above: always passes
            if (dcd_straddle(g) && dcd_pair_contact(d, X, A, Bt, g)) atomicAdd(&s_nc, 1u);
        }
        __syncthreads();
    }
    // contacts counted per workgroup, one device-scope atomic each
    if (threadIdx.x == 0 && s_nc) atomicAdd(&d.ncontact[0], s_nc);
}

// apply_kernel (:802-815), averaged; clears the accumulators and the contact
// counter for the next pass (the count is kept as the last pass's)
__global__ void k_dcd_apply(DcdDev d, float4 *X, int nv) {
    const int i = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i == 0) {
        d.ncontact[1] = d.ncontact[0];
        d.ncontact[0] = 0u;
    }
    if (i >= nv) return;
    const uint32_t n = d.acc_n[i];
    if (n > 0) {
        const float w = (float)n;
        float4 x = X[i];
        const float inv = 1.0f / DCD_FIX;
        x.x = x.x + ((float)(long long)d.acc[3 * i + 0] * inv) / w;
        x.y = x.y + ((float)(long long)d.acc[3 * i + 1] * inv) / w;
        x.z = x.z + ((float)(long long)d.acc[3 * i + 2] * inv) / w;
        X[i] = x;
        d.acc[3 * i + 0] = d.acc[3 * i + 1] = d.acc[3 * i + 2] = 0ull;
        d.acc_n[i] = 0u;
    }
}

// broadphase counters back to zero (one fill kernel: graph-safe, one launch)
__global__ void k_dcd_reset(DcdDev d, int64_t T) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < T) {
        d.bucket_cnt[i] = 0u;
        d.bucket_fill[i] = 0u;
    }
    if (i == 0) {
        *d.npairs = 0ull;
    }
    if (i < 3) d.overflow[i] = 0u;
    if (i < DCD_HIST) d.hist[i] = 0u;
    if (i < DCD_LEVELS) d.lvl_cnt[i] = 0u;
}

__global__ void k_dcd_snapshot(float4 *dst, const float4 *src, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_zero_u32(uint32_t *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// exclusive scan of n uint32 -> out[0..n], three phases (block sums, scan of sums, add)
constexpr int SCAN_BLOCK = 1024;
__global__ __launch_bounds__(256) void k_scan_blocks(const uint32_t *in, uint32_t *out, uint32_t *bsum, int64_t n) {
    __shared__ uint32_t s[256];
    const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK;
    uint32_t v[4];
    uint32_t loc = 0;
    for (int k = 0; k < 4; ++k) {
        const int64_t idx = base + threadIdx.x * 4 + k;
        v[k] = idx < n ? in[idx] : 0u;
        loc += v[k];
    }
    s[threadIdx.x] = loc;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - loc;
    for (int k = 0; k < 4; ++k) {
        const int64_t idx = base + threadIdx.x * 4 + k;
        if (idx < n) out[idx] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = s[255];
}

__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t *bsum, int nb, uint32_t *total_out) {
    __shared__ uint32_t s[1024];
    uint32_t carry = 0;
    for (int b0 = 0; b0 < nb; b0 += 1024) {
        const int i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? bsum[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + s[threadIdx.x] - v;
        const uint32_t tot = s[1023];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ void k_scan_add(uint32_t *out, const uint32_t *bsum, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += bsum[i / SCAN_BLOCK];
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
namespace {

struct DBuf {
    void *p = nullptr;
    size_t n = 0;
};

struct EvPair {
    hipEvent_t a, b;
};

}  // namespace

// Double-buffered pinned host staging for per-frame uploads: the host fills one
// slot while the stream may still be copying from the other, so per-frame
// setters never wait for the GPU (only for the copy issued two calls ago).
struct HostStage {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    size_t cap = 0;
    int next = 0;
};

// exact-mesh bone collider: the caller's mesh, the BVH built from it and its device copy
struct BoneMesh {
    bool set = false, resolve = false, built = false;
    int nbv = 0, ntris = 0, nbones = 0;
    std::vector<float> pts;                       // nbv*3, current frame
    std::vector<int32_t> tris, tri_off, goff, vgroup;
    std::vector<float> cw, mw;
    float max_dist = 0.2f, contact_offset = 0.f, max_push = -1.f;
    int coherent_req = 0;                          // as asked: 0 off, 1 on, 2 when closed
    double motion = 0.0;                           // running sum of the largest bone-vertex displacement per upload
    BmTables tab;                                 // the tree's tables (BVH order)
    std::vector<void *> bufs;
    BoneMeshDev dev{};
    bm_f4 *d_pts = nullptr;                        // nbv positions, then (M, 0, 0, 0)
    float *d_depth = nullptr;
    uint32_t *d_stat = nullptr;
    HostStage st;
};

struct xpbd_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr, stream = nullptr;
    std::string err;
    // host definition
    int nv = 0, nt = 0;
    std::vector<float> pos0, rest_volume, rest_matrix, mass, force, fiber, tendonmask, activation;
    // activation by group (xpbd_set_activation_groups): per-tet group, the groups' values, whether
    // the host copy of the per-tet activation is stale (the groups are expanded on the device)
    std::vector<int32_t> act_group;
    std::vector<float> group_act;
    int n_act_groups = 0;
    bool act_host_stale = false;
    int32_t *d_elem_group = nullptr;   // per element: its tet's group
    float *d_gact = nullptr;
    std::vector<int32_t> tets, stopped;
    bool has_force = false;
    // constraint table copy
    int ncons = 0;
    std::vector<int32_t> c_type, c_pts, c_tetid;
    std::vector<float> c_stiff, c_comp, c_damp, c_rest, c_rvec, c_rdir;
    SimParams sp{};
    bool params_set = false;
    // pins
    std::vector<int32_t> pin_vids, pin_active;
    std::vector<float> pin_targets;
    // per-frame target path: PIN already holds the full table for the current
    // pin set, only the targeted active pins' rows are rewritten on the device
    bool pins_full = false;
    int n_pin_dev = 0, pin_cap = 0;
    int32_t *d_pin_dev = nullptr;    // device vertex ids of targeted active pins (pin order)
    float4 *d_pin_tgt = nullptr;
    // per-frame attach targets: staged compactly (attach order) and scattered into ATT0
    int32_t *d_att_dev = nullptr;    // device vertex of attach constraint #i (table order)
    float4 *d_att_tgt = nullptr;
    HostStage st_pin, st_act, st_bone, st_att, st_gact;
    // sdf
    int nbones = 0, res = 0;
    std::vector<float> sdf_grids, sdf_origin, sdf_invh, coll_w, bone_xf;
    std::vector<int32_t> vgroup, goff;
    float sdf_offset = 0.f, sdf_maxpush = -1.f;
    // layout
    xpbd::Layout L;
    bool ready = false;
    int P = 256;
    bool flat = false;
    int block = 0;          // block strategy: tets per block (0 = off)
    int block_bs = 256;     // block strategy: threads per block
    int feature = 0;
    // device buffers
    std::vector<DBuf> bufs;
    float4 *X[2] = {nullptr, nullptr}, *PPv = nullptr, *VEL = nullptr, *FRC = nullptr, *PIN = nullptr,
           *ATT0 = nullptr, *ATT1 = nullptr, *DPW = nullptr, *FIB = nullptr, *FIBS = nullptr;
    uint32_t *FLAGS = nullptr;
    PatchDesc *d_patch = nullptr;
    int32_t *d_halo = nullptr, *d_vslot_off = nullptr, *d_ptet_tetid = nullptr, *d_pvert = nullptr;
    ushort4 *d_lidx = nullptr;
    uint32_t *d_meta = nullptr;
    uint16_t *d_vslot = nullptr;
    int4 *d_gidx = nullptr;
    int32_t *d_vslot32 = nullptr;
    float4 *CORR = nullptr;
    // block strategy tables
    BlockDesc *d_blk = nullptr;
    int32_t *d_bvt = nullptr, *d_bps = nullptr, *d_svid = nullptr, *d_soff = nullptr;
    uint32_t *d_brun = nullptr;
    ushort4 *d_blidx = nullptr;
    ushort4 *d_bcpos = nullptr;
    float4 *BPART = nullptr;
    // resident mode of the block strategy (all plain iterations of a substep in one launch)
    int res_R = 0, res_vpt = 0;
    size_t res_lds = 0;
    xpbd::Parts parts;
    PartDesc *d_parts = nullptr;
    int4 *d_prec = nullptr;
    uint32_t *d_bmap = nullptr;
    ushort4 *d_plidx = nullptr;
    int32_t *d_nbr = nullptr;
    float4 *RXP = nullptr;            // 2 x n_partial
    uint32_t *d_rcount = nullptr, *d_rstatus = nullptr;
    // two-level correction: a finalized coarse context (not owned) and transfer tables
    xpbd_ctx *coarse = nullptr;
    // fascia shell in ride mode: a finalized context (not owned) whose bound vertices follow this
    // context's surface; gen counts this context's finalizes, so that a stale link is detected
    uint64_t gen = 0;
    xpbd_ctx *fascia = nullptr;
    uint64_t fascia_gen = 0;
    int fascia_iters = 0, fascia_nbound = 0;
    int32_t *d_fb_vid = nullptr;      // fascia library index of each bound vertex
    int4 *d_fb_tri = nullptr;         // this context's library indices of its triangle
    float4 *d_fb_w = nullptr;         // barycentric weights of a, b, c and the normal gap
    std::vector<int32_t> fb_face, fb_ok;   // per fascia vertex (caller order): surface triangle, bound
    std::vector<float> fb_w, fb_gap;       // 3 weights, gap
    // fascia squeeze: follow positions (fascia library order), the previous frame's, glue and anchors
    int fascia_mode = 0, fascia_nglue = 0, fascia_nanchor = 0;
    float fascia_omega = 1.f;
    float4 *d_ff = nullptr, *d_ff_prev = nullptr;
    int4 *d_fglue = nullptr;               // fascia vertex, muscle vertex (library orders), restlen, compliance bits
    int32_t *d_fanc = nullptr;             // anchor: fascia library index
    float4 *d_fanc_off = nullptr;
    HostStage st_ff;
    int mg_iters = 0;
    bool mg_pidx_set = false, mg_kinematic = false;
    int mg_nblk = 0;
    float4 *d_mg_pin_tgt = nullptr;   // pin rule 1: restricted bone goal per coarse node (block order)
    HostStage st_mgpin;
    int32_t *d_mg_pidx = nullptr;
    float *d_mg_pw = nullptr;
    float4 *MG_QPREV = nullptr;
    MgArgs mg{};
    MgCoarse mgc{};
    size_t mgc_lds = 0;
    float *d_Bp = nullptr, *d_rvol = nullptr, *d_rsnh = nullptr, *d_act = nullptr;
    DevClass *d_cls = nullptr;
    int4 *d_tets = nullptr;
    float *d_B = nullptr;
    float *d_grids = nullptr, *d_bone = nullptr;
    int32_t *d_vgroup = nullptr, *d_goff = nullptr;
    int32_t *d_cells = nullptr, *d_count = nullptr;
    double *d_partial = nullptr;
    int n_partial = 0;
    bool pins_on = false;
    // learned warm start
    int lam_mode = 0;             // 0 plain, 1 carry, 2 network warm start, 3 previous-substep warm start
    bool impulse_pending = false;
    int eff_mode = 0;             // mode of the current substep (mode 4 schedules 1 and 2)
    int ref_iters = 10;           // mode 4: iterations of a reference substep
    int ref_every = 0;            // mode 4: reference substeps are those with index % ref_every == 0
                                  // (0: the first substep of a frame only)
    int warm_net = 1;             // mode 4: between references, network (mode 2) or base only (mode 3)
    float4 *LAMB = nullptr;       // carried multipliers of the current substep
    float4 *LAMB0 = nullptr;      // base of the next warm start (lambda0 of the last substep, or a reference solve)
    float4 *SFEAT = nullptr;      // static network features per element (k_static_features)
    bool sfeat_ok = false;
    uint16_t *mW1 = nullptr, *mW2 = nullptr;
    float *mb1 = nullptr, *mb2 = nullptr, *mW3 = nullptr, *mb3 = nullptr;
    float *fmean = nullptr, *finv = nullptr, *oscale = nullptr;
    bool mlp_ready = false;
    float4 *S_X = nullptr, *S_PP = nullptr, *S_VEL = nullptr;
    float4 *T_X = nullptr;   // scratch copy of X[0] while iterations are timed
    // tri-tri DCD
    bool dcd_on = false;
    int dcd_iters = 2, dcd_rebuild_substeps = 0;
    bool dcd_fresh = true;
    std::vector<int32_t> dcd_tris, dcd_group;
    xpbd_dcd_params dcd_p{};
    DcdDev dcd{};
    BoneMesh bm;
    uint32_t *d_scan_bsum = nullptr;
    int n_scan_bsum = 0;
    // captured frames, keyed by (substeps, iterations, lambda mode); a small
    // cache so that plain and warm-started frames can alternate
    struct GraphSlot {
        hipGraph_t graph = nullptr;
        hipGraphExec_t gexec = nullptr;
        int nsub = -1, iters = -1, lam_mode = -1, ref_iters = -1, ref_every = -1, warm_net = -1;
        uint64_t used = 0;
    };
    static constexpr int NGRAPH = 4;
    GraphSlot gslot[NGRAPH];
    uint64_t gclock = 0;
    // timing
    bool timing = false;
    std::vector<EvPair> evs;
    double timed_ms = 0.0;
    int timed_n = 0;
};

namespace {

int fail(xpbd_ctx *c, int code, const std::string &m) {
    if (c) c->err = m;
    return code;
}

#define HIPCHK(ctx, expr)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(ctx, XPBD_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

template <typename T>
int dalloc(xpbd_ctx *c, T **p, size_t count) {
    size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(c, XPBD_EOOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->bufs.push_back(DBuf{q, bytes});
    *p = (T *)q;
    return XPBD_OK;
}

// free one buffer of the context's list (the stream must not use it any more)
template <typename T>
void dfree(xpbd_ctx *c, T *&p) {
    if (!p) return;
    for (size_t i = 0; i < c->bufs.size(); ++i)
        if (c->bufs[i].p == (void *)p) {
            (void)hipFree(c->bufs[i].p);
            c->bufs.erase(c->bufs.begin() + i);
            break;
        }
    p = nullptr;
}

template <typename T>
int dupload(xpbd_ctx *c, T **p, const T *src, size_t count) {
    int r = dalloc(c, p, count);
    if (r) return r;
    if (count) HIPCHK(c, hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
    return XPBD_OK;
}

void drop_graph_slot(xpbd_ctx::GraphSlot &g) {
    if (g.gexec) { (void)hipGraphExecDestroy(g.gexec); g.gexec = nullptr; }
    if (g.graph) { (void)hipGraphDestroy(g.graph); g.graph = nullptr; }
    g.nsub = g.iters = g.lam_mode = g.ref_iters = g.ref_every = g.warm_net = -1;
    g.used = 0;
}

void bm_free(BoneMesh &b) {
    for (void *p : b.bufs) (void)hipFree(p);
    b.bufs.clear();
    b.built = false;
    b.dev = BoneMeshDev{};
    b.d_pts = nullptr;
    b.d_depth = nullptr;
    b.d_stat = nullptr;
}

void free_device(xpbd_ctx *c) {
    bm_free(c->bm);
    for (auto &g : c->gslot) drop_graph_slot(g);
    for (auto &b : c->bufs) (void)hipFree(b.p);
    c->bufs.clear();
    c->ready = false;
    c->pins_full = false;
    c->T_X = nullptr;
    c->d_pin_dev = nullptr;
    c->d_pin_tgt = nullptr;
    c->d_att_dev = nullptr;
    c->d_att_tgt = nullptr;
    c->d_elem_group = nullptr;
    c->d_gact = nullptr;
    c->d_mg_pin_tgt = nullptr;
    c->d_fb_vid = nullptr;
    c->d_fb_tri = nullptr;
    c->d_fb_w = nullptr;
    c->d_ff = c->d_ff_prev = c->d_fanc_off = nullptr;
    c->d_fglue = nullptr;
    c->d_fanc = nullptr;
    c->SFEAT = nullptr;
    c->sfeat_ok = false;
    c->pin_cap = 0;
    c->res_R = 0;
    c->d_rcount = c->d_rstatus = nullptr;
}

void free_stage(HostStage &s) {
    for (int i = 0; i < 2; ++i) {
        if (s.ev[i]) (void)hipEventDestroy(s.ev[i]);
        if (s.buf[i]) (void)hipHostFree(s.buf[i]);
        s.ev[i] = nullptr;
        s.buf[i] = nullptr;
        s.pending[i] = false;
    }
    s.cap = 0;
}

// a pinned slot of at least `bytes` that no enqueued copy still reads
int stage_acquire(xpbd_ctx *c, HostStage &s, size_t bytes, int &slot, void *&host) {
    slot = s.next;
    s.next ^= 1;
    if (s.pending[slot]) {
        HIPCHK(c, hipEventSynchronize(s.ev[slot]));
        s.pending[slot] = false;
    }
    if (bytes > s.cap) {
        for (int i = 0; i < 2; ++i)
            if (s.pending[i]) {
                HIPCHK(c, hipEventSynchronize(s.ev[i]));
                s.pending[i] = false;
            }
        for (int i = 0; i < 2; ++i) {
            if (s.buf[i]) (void)hipHostFree(s.buf[i]);
            s.buf[i] = nullptr;
        }
        s.cap = 0;
        for (int i = 0; i < 2; ++i) {
            hipError_t e = hipHostMalloc(&s.buf[i], bytes, hipHostMallocDefault);
            if (e != hipSuccess) return fail(c, XPBD_EOOM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
            if (!s.ev[i]) HIPCHK(c, hipEventCreateWithFlags(&s.ev[i], hipEventDisableTiming));
        }
        s.cap = bytes;
    }
    host = s.buf[slot];
    return XPBD_OK;
}

// enqueue the host-to-device copy of a filled slot and mark it busy
int stage_submit(xpbd_ctx *c, HostStage &s, int slot, void *dst, size_t bytes) {
    HIPCHK(c, hipMemcpyAsync(dst, s.buf[slot], bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipEventRecord(s.ev[slot], c->stream));
    s.pending[slot] = true;
    return XPBD_OK;
}

void invalidate_graph(xpbd_ctx *c) {
    for (auto &g : c->gslot) drop_graph_slot(g);
}

// ---- exact-mesh bone collider: BVH build (host, once per mesh) and refit / query launches
template <typename T>
int bm_alloc(xpbd_ctx *c, T **p, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count * sizeof(T), 16));
    if (e != hipSuccess) return fail(c, XPBD_EOOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->bm.bufs.push_back(q);
    *p = (T *)q;
    return XPBD_OK;
}

template <typename T>
int bm_upload(xpbd_ctx *c, T **p, const std::vector<T> &v) {
    int r = bm_alloc(c, p, v.size());
    if (r) return r;
    if (!v.empty()) HIPCHK(c, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return XPBD_OK;
}

inline dim3 vgrid_bm(int n) { return dim3((unsigned)((n + 255) / 256)); }

int bm_refit(xpbd_ctx *c) {
    BoneMesh &b = c->bm;
    auto g = [](int n) { return dim3((unsigned)((n + 255) / 256)); };
    hipLaunchKernelGGL(k_bm_soup, g(b.ntris), dim3(256), 0, c->stream, b.dev);
    hipLaunchKernelGGL(k_bm_vnorm, g(b.nbv), dim3(256), 0, c->stream, b.dev);
    hipLaunchKernelGGL(k_bm_refit_low, dim3((unsigned)b.tab.sub.size()), dim3(1 << BM_SUB_LOG), 0, c->stream, b.dev);
    hipLaunchKernelGGL(k_bm_refit_top, dim3((unsigned)b.nbones), dim3(1024), 0, c->stream, b.dev);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

int bm_push(xpbd_ctx *c, float4 *X) {
    hipLaunchKernelGGL(k_bm_query<0>, dim3((unsigned)((c->nv + 255) / 256)), dim3(256), 0, c->stream, c->nv, X,
                       c->bm.dev, nullptr, 0.0f, nullptr);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

int bm_upload_points(xpbd_ctx *c) {
    BoneMesh &b = c->bm;
    int slot;
    void *h;
    const size_t bytes = ((size_t)b.nbv + 1) * sizeof(float4);
    int r = stage_acquire(c, b.st, bytes, slot, h);
    if (r) return r;
    bm_f4 *q = (bm_f4 *)h;
    for (int i = 0; i < b.nbv; ++i) q[i] = bm_f4{b.pts[3 * i], b.pts[3 * i + 1], b.pts[3 * i + 2], 0.f};
    // M rounded up, so that a certificate never under-counts the bones' motion
    q[b.nbv] = bm_f4{std::nextafter((float)b.motion, INFINITY), 0.f, 0.f, 0.f};
    if ((r = stage_submit(c, b.st, slot, b.d_pts, bytes))) return r;
    return bm_refit(c);
}

// the tree's tables from the first pose (bm_build_tables), vertex-side tables
// in the library's vertex order, device copies, then the first refit
int bm_build(xpbd_ctx *c) {
    BoneMesh &b = c->bm;
    bm_free(b);
    if (!b.set) return XPBD_OK;
    const int nb = b.nbones, nt = b.ntris, nbv = b.nbv, nv = c->nv;
    BmTables &T = b.tab;
    std::string e = bm_build_tables(nbv, b.pts.data(), nt, b.tris.data(), nb, b.tri_off.data(), T);
    if (!e.empty()) return fail(c, XPBD_EINVAL, e);
    if (b.coherent_req == 1 && !T.closed)
        return fail(c, XPBD_EINVAL, "coherent queries need closed sub-meshes (every edge in two triangles, "
                                    "opposite directions)");
    std::vector<int32_t> vg(nv), goff = b.goff;
    std::vector<float> cw(nv), mw;
    if (goff.empty()) goff = {0, nb};
    if (!b.mw.empty()) mw.resize(nv);
    for (int n = 0; n < nv; ++n) {
        const int o = c->L.perm[n];
        vg[n] = b.vgroup.empty() ? 0 : b.vgroup[o];
        cw[n] = b.cw.empty() ? 1.0f : b.cw[o];
        if (!mw.empty()) mw[n] = b.mw[o];
    }
    BoneMeshDev &d = b.dev;
    d.nbones = nb;
    d.ntris = nt;
    d.nbv = nbv;
    d.max_dist = b.max_dist;
    d.contact_offset = b.contact_offset;
    d.max_push = b.max_push;
    d.coherent = (b.coherent_req == 1 || (b.coherent_req == 2 && T.closed)) ? 1 : 0;
    int r;
    bm_i4 *bone = nullptr, *tv = nullptr, *adj = nullptr;
    bm_i2 *sub = nullptr;
    int32_t *vt_off = nullptr, *vt = nullptr, *dvg = nullptr, *dgoff = nullptr;
    float *dcw = nullptr, *dmw = nullptr;
    if ((r = bm_upload(c, &bone, T.bone)) || (r = bm_upload(c, &tv, T.tv)) || (r = bm_upload(c, &adj, T.adj)) ||
        (r = bm_upload(c, &sub, T.sub)) || (r = bm_upload(c, &vt_off, T.vt_off)) || (r = bm_upload(c, &vt, T.vt)) ||
        (r = bm_upload(c, &dvg, vg)) || (r = bm_upload(c, &dgoff, goff)) || (r = bm_upload(c, &dcw, cw)))
        return r;
    if (!mw.empty() && (r = bm_upload(c, &dmw, mw))) return r;
    d.bone = bone; d.tv = tv; d.adj = adj; d.sub = sub; d.vt_off = vt_off; d.vt = vt;
    d.vgroup = dvg; d.goff = dgoff; d.cw = dcw; d.mw = dmw;
    bm_f4 *pts = nullptr;
    if ((r = bm_alloc(c, &d.soup, 3 * (size_t)nt)) || (r = bm_alloc(c, &d.fn, (size_t)nt)) ||
        (r = bm_alloc(c, &d.vn, (size_t)nbv)) || (r = bm_alloc(c, &d.node, 2 * (size_t)T.nnodes)) ||
        (r = bm_alloc(c, &pts, (size_t)nbv + 1)) || (r = bm_alloc(c, &b.d_depth, (size_t)nv)) ||
        (r = bm_alloc(c, &b.d_stat, 2)) || (r = bm_alloc(c, &d.cache, (size_t)nv)) ||
        (r = bm_alloc(c, &d.nquery, 1)))
        return r;
    HIPCHK(c, hipMemsetAsync(d.nquery, 0, sizeof(unsigned long long), c->stream));
    b.d_pts = pts;
    d.pts = pts;
    hipLaunchKernelGGL(k_bm_cache_reset, vgrid_bm(nv), dim3(256), 0, c->stream, d, nv);
    HIPCHK(c, hipGetLastError());
    b.built = true;
    invalidate_graph(c);
    return bm_upload_points(c);
}

inline float inv_mass_of(const xpbd_ctx *c, int old) {
    if (!c->stopped.empty() && c->stopped[old] != 0) return 0.0f;
    float m = c->mass[old];
    return m > 0.0f ? 1.0f / m : 0.0f;
}

// upload the per-vertex W (inverse mass) / mass / flags after mass or stopped change
int upload_vertex_static(xpbd_ctx *c) {
    c->sfeat_ok = false;   // the inverse masses enter the network's static features
    const int nv = c->nv;
    std::vector<float4> x(nv), v(nv), pp(nv);
    HIPCHK(c, hipMemcpyAsync(x.data(), c->X[0], nv * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(v.data(), c->VEL, nv * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(pp.data(), c->PPv, nv * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<uint32_t> fl(nv);
    for (int n = 0; n < nv; ++n) {
        const int o = c->L.perm[n];
        x[n].w = inv_mass_of(c, o);
        v[n].w = c->mass[o];
        fl[n] = (!c->stopped.empty() && c->stopped[o]) ? 1u : 0u;
        pp[n].w = c->coll_w.empty() ? 0.0f : c->coll_w[o];
    }
    HIPCHK(c, hipMemcpy(c->X[0], x.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->X[1], x.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->VEL, v.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->PPv, pp.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->FLAGS, fl.data(), nv * sizeof(uint32_t), hipMemcpyHostToDevice));
    return XPBD_OK;
}

int upload_pins(xpbd_ctx *c) {
    const int nv = c->nv;
    std::vector<float4> pin(nv, make_float4(0, 0, 0, 0));
    bool any = false;
    for (int n = 0; n < nv; ++n) {
        const float *q = &c->L.pin_from_cons[4 * (size_t)n];
        if (q[3] != 0.0f) { pin[n] = make_float4(q[0], q[1], q[2], 1.0f); any = true; }
    }
    const int np = (int)c->pin_vids.size();
    std::vector<int32_t> dev;
    for (int i = 0; i < np; ++i) {
        if (!c->pin_active[i]) continue;
        const int n = c->L.iperm[c->pin_vids[i]];
        const float *t = c->pin_targets.empty() ? nullptr : &c->pin_targets[3 * (size_t)i];
        if (!t) continue;
        pin[n] = make_float4(t[0], t[1], t[2], 1.0f);
        dev.push_back(n);
        any = true;
    }
    c->pins_on = any;
    c->pins_full = false;
    HIPCHK(c, hipMemcpyAsync(c->PIN, pin.data(), nv * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    if (!dev.empty()) {
        if ((int)dev.size() > c->pin_cap) {
            int r = dalloc(c, &c->d_pin_dev, (size_t)np);
            if (!r) r = dalloc(c, &c->d_pin_tgt, (size_t)np);
            if (r) return r;
            c->pin_cap = np;
        }
        HIPCHK(c, hipMemcpyAsync(c->d_pin_dev, dev.data(), dev.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                 c->stream));
        c->n_pin_dev = (int)dev.size();
        // a vertex pinned twice keeps the later target on the host; the device
        // scatter would race, so such a pin set always takes the full upload
        std::vector<int32_t> u(dev);
        std::sort(u.begin(), u.end());
        c->pins_full = std::adjacent_find(u.begin(), u.end()) == u.end();
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return XPBD_OK;
}

// per-frame pin targets: rewrite the targeted rows of the pin table in place
__global__ __launch_bounds__(256) void k_pin_scatter(int n, const int32_t *__restrict__ dev,
                                                     const float4 *__restrict__ tgt, float4 *__restrict__ pin) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) pin[dev[i]] = tgt[i];
}

// per-frame attach targets (_update_belly_attach_targets, mg.py:725-734): the
// target xyz of each attach record, its stiffness (w) kept
__global__ __launch_bounds__(256) void k_att_scatter(int n, const int32_t *__restrict__ dev,
                                                     const float4 *__restrict__ tgt, float4 *__restrict__ att0) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const float4 t = tgt[i];
        float4 *a = att0 + dev[i];
        *a = make_float4(t.x, t.y, t.z, a->w);
    }
}

std::vector<DevClass> dev_classes(const std::vector<ClassParams> &cls, const SimParams &sp) {
    std::vector<DevClass> out(cls.size());
    const float dt = sp.dt;
    for (size_t i = 0; i < cls.size(); ++i) {
        const ClassParams &p = cls[i];
        DevClass &d = out[i];
        d = DevClass{};
        d.p = p;
        d.at_arap = p.k_arap > 0.0f ? (1.0f / p.k_arap) / (dt * dt) : 0.0f;
        d.a_vol = p.k_vol != 0.0f ? (1.0f / p.k_vol) / (dt * dt) : 0.0f;
        d.a_comp = p.kcomp_vol != 0.0f ? (1.0f / p.kcomp_vol) / (dt * dt) : 0.0f;
        // tet_volume_update_xpbd_fn damping: beta = k*ratio*dt^2, gamma = alpha*beta/dt
        d.g_vol = d.a_vol * (p.k_vol * p.damp_vol * dt * dt) / dt;
        d.g_comp = d.a_comp * (p.kcomp_vol * p.damp_vol * dt * dt) / dt;
        const float sv = p.k_aniso * sp.fiber_scale;
        d.at_aniso = sv > 0.0f ? (1.0f / sv) / (dt * dt) : 0.0f;
    }
    return out;
}

IterArgs make_args(xpbd_ctx *c, float4 *dpw) {
    IterArgs a;
    a.patch = c->d_patch;
    a.halo = c->d_halo;
    a.lidx = c->d_lidx;
    a.meta = c->d_meta;
    a.Bp = c->d_Bp;
    a.rvol = c->d_rvol;
    a.rsnh = c->d_rsnh;
    a.fib = c->FIB;
    a.cls = c->d_cls;
    a.vslot_off = c->d_vslot_off;
    a.vslot = c->d_vslot;
    a.gidx = c->d_gidx;
    a.slot_of = (const int4 *)c->d_vslot32;
    a.pin = c->pins_on ? c->PIN : nullptr;
    a.att0 = (c->feature & F_ATTACH) ? c->ATT0 : nullptr;
    a.att1 = c->ATT1;
    a.dpw = dpw;
    a.lam = c->LAMB;
    a.nptet = (int64_t)c->L.meta.size();
    a.max_table = c->L.max_table;
    a.sp = c->sp;
    return a;
}

// template dispatch of the iteration kernel
typedef void (*IterKernel)(IterArgs, const float4 *, float4 *, const float4 *);

template <int BS, int FEAT>
IterKernel pick_damp(int damp, int mode) {
    if (mode == MODE_ACCUM) return damp ? k_iterate<BS, FEAT, 1, MODE_ACCUM> : k_iterate<BS, FEAT, 0, MODE_ACCUM>;
    return damp ? k_iterate<BS, FEAT, 1, MODE_FUSED> : k_iterate<BS, FEAT, 0, MODE_FUSED>;
}

template <int BS>
IterKernel pick_feat(int feat, int damp, int mode) {
    constexpr int AV = F_ARAP | F_VOL, AVA = F_ARAP | F_VOL | F_ANISO, SA = F_SNH | F_ANISO,
                  ALL = F_ARAP | F_VOL | F_SNH | F_ANISO;
    const int f = feat & ALL;
    if (mode == MODE_ACCUM) return pick_damp<BS, ALL>(damp, mode);
    if ((f & ~AV) == 0) return pick_damp<BS, AV>(damp, mode);
    if ((f & ~SA) == 0) return pick_damp<BS, SA>(damp, mode);
    if ((f & ~AVA) == 0) return pick_damp<BS, AVA>(damp, mode);
    return pick_damp<BS, ALL>(damp, mode);
}

IterKernel pick_iter(int P, int feat, int mode) {
    const int damp = (feat & F_DAMP) ? 1 : 0;
    switch (P) {
        case 64: return pick_feat<64>(feat, damp, mode);
        case 128: return pick_feat<128>(feat, damp, mode);
        default: return pick_feat<256>(feat, damp, mode);
    }
}

typedef void (*FlatTetKernel)(IterArgs, const float4 *, const float4 *, float4 *);
typedef void (*FlatVertKernel)(IterArgs, const float4 *, float4 *, const float4 *, const float4 *, int);

template <int FEAT, int DAMP>
FlatTetKernel pick_flat_lam(int lam) {
    return lam == 2 ? k_flat_tets<FEAT, DAMP, 2> : (lam == 1 ? k_flat_tets<FEAT, DAMP, 1> : k_flat_tets<FEAT, DAMP, 0>);
}

template <int FEAT>
FlatTetKernel pick_flat_damp(int damp, int lam) {
    return damp ? pick_flat_lam<FEAT, 1>(lam) : pick_flat_lam<FEAT, 0>(lam);
}

FlatTetKernel pick_flat_tets(int feat, int lam) {
    constexpr int AV = F_ARAP | F_VOL, AVA = F_ARAP | F_VOL | F_ANISO, SA = F_SNH | F_ANISO,
                  ALL = F_ARAP | F_VOL | F_SNH | F_ANISO;
    const int f = feat & ALL;
    const int damp = (feat & F_DAMP) ? 1 : 0;
    if ((f & ~AV) == 0) return pick_flat_damp<AV>(damp, lam);
    if ((f & ~SA) == 0) return pick_flat_damp<SA>(damp, lam);
    if ((f & ~AVA) == 0) return pick_flat_damp<AVA>(damp, lam);
    return pick_flat_damp<ALL>(damp, lam);
}

FlatVertKernel pick_flat_verts(int feat, int mode) {
    const int damp = (feat & F_DAMP) ? 1 : 0;
    if (mode == MODE_ACCUM) return damp ? k_flat_verts<1, MODE_ACCUM> : k_flat_verts<0, MODE_ACCUM>;
    return damp ? k_flat_verts<1, MODE_FUSED> : k_flat_verts<0, MODE_FUSED>;
}

typedef void (*BlockTetKernel)(IterArgs, BlockArgs, const float4 *, float4 *, const float4 *);
typedef void (*BlockVertKernel)(IterArgs, BlockArgs, const float4 *, float4 *, const float4 *, int);

template <int FEAT, int DAMP>
BlockTetKernel pick_block_lam(int lam) {
    return lam == 2 ? k_block_tets<FEAT, DAMP, 2, MODE_FUSED>
                    : (lam == 1 ? k_block_tets<FEAT, DAMP, 1, MODE_FUSED> : k_block_tets<FEAT, DAMP, 0, MODE_FUSED>);
}

template <int FEAT>
BlockTetKernel pick_block_damp(int damp, int lam) {
    return damp ? pick_block_lam<FEAT, 1>(lam) : pick_block_lam<FEAT, 0>(lam);
}

BlockTetKernel pick_block_tets(int feat, int lam, int mode) {
    constexpr int AV = F_ARAP | F_VOL, AVA = F_ARAP | F_VOL | F_ANISO, SA = F_SNH | F_ANISO,
                  ALL = F_ARAP | F_VOL | F_SNH | F_ANISO;
    const int f = feat & ALL;
    const int damp = (feat & F_DAMP) ? 1 : 0;
    if (mode == MODE_ACCUM) return damp ? k_block_tets<ALL, 1, 0, MODE_ACCUM> : k_block_tets<ALL, 0, 0, MODE_ACCUM>;
    if ((f & ~AV) == 0) return pick_block_damp<AV>(damp, lam);
    if ((f & ~SA) == 0) return pick_block_damp<SA>(damp, lam);
    if ((f & ~AVA) == 0) return pick_block_damp<AVA>(damp, lam);
    return pick_block_damp<ALL>(damp, lam);
}

BlockVertKernel pick_block_shared(int feat, int mode) {
    const int damp = (feat & F_DAMP) ? 1 : 0;
    if (mode == MODE_ACCUM) return damp ? k_block_shared<1, MODE_ACCUM> : k_block_shared<0, MODE_ACCUM>;
    return damp ? k_block_shared<1, MODE_FUSED> : k_block_shared<0, MODE_FUSED>;
}

typedef void (*ResKernel)(IterArgs, BlockArgs, ResArgs, const float4 *, float4 *, const float4 *, int);

template <int FEAT, int DAMP>
ResKernel pick_res_v(int vpt) {
    return vpt <= 1 ? k_resident<FEAT, DAMP, 1> : (vpt == 2 ? k_resident<FEAT, DAMP, 2> : k_resident<FEAT, DAMP, 3>);
}

// part-local vertices per thread: 1 to 3
ResKernel pick_resident(int feat, int vpt) {
    constexpr int AV = F_ARAP | F_VOL, SA = F_SNH | F_ANISO, ALL = F_ARAP | F_VOL | F_SNH | F_ANISO;
    const int f = feat & ALL;
    const int damp = (feat & F_DAMP) ? 1 : 0;
    if ((f & ~AV) == 0) return damp ? pick_res_v<AV, 1>(vpt) : pick_res_v<AV, 0>(vpt);
    if ((f & ~SA) == 0) return damp ? pick_res_v<SA, 1>(vpt) : pick_res_v<SA, 0>(vpt);
    return damp ? pick_res_v<ALL, 1>(vpt) : pick_res_v<ALL, 0>(vpt);
}

size_t resident_lds(int feature, int M, int TB, int nclass) {
    const size_t planes = (feature & F_DAMP) ? 3 : 2;
    return (planes * (size_t)M + 4 * (size_t)TB) * sizeof(float4) + (size_t)nclass * sizeof(DevClass);
}

size_t block_lds(const xpbd_ctx *c) {
    const size_t M = (size_t)c->L.max_nloc;
    const size_t damp = (c->feature & F_DAMP) ? 2 : 1;
    return (damp * M + 4 * (size_t)c->block) * sizeof(float4);
}

BlockArgs make_block_args(xpbd_ctx *c) {
    BlockArgs b;
    b.blk = c->d_blk;
    b.bvt = c->d_bvt;
    b.bps = c->d_bps;
    b.brun = c->d_brun;
    b.blidx = c->d_blidx;
    b.bcpos = c->d_bcpos;
    b.svid = c->d_svid;
    b.soff = c->d_soff;
    b.part = c->BPART;
    b.max_nloc = c->L.max_nloc;
    b.block_tets = c->block;
    return b;
}

size_t iter_lds(const xpbd_ctx *c) {
    const int damp = (c->feature & F_DAMP) ? 2 : 1;
    return ((size_t)damp * c->L.max_table + (size_t)c->P * 4) * sizeof(float4);
}

int launch_iter(xpbd_ctx *c, const float4 *xin, float4 *xout, int mode, int lam = 0) {
    IterArgs a = make_args(c, c->DPW);
    EvPair ev{};
    const bool tm = c->timing;
    if (tm) {
        HIPCHK(c, hipEventCreate(&ev.a));
        HIPCHK(c, hipEventCreate(&ev.b));
        HIPCHK(c, hipEventRecord(ev.a, c->stream));
    }
    const float4 *pp = c->PPv;
    if (c->block) {
        BlockTetKernel kt = pick_block_tets(c->feature, lam, mode);
        BlockVertKernel kv = pick_block_shared(c->feature, mode);
        BlockArgs b = make_block_args(c);
        const unsigned nb = (unsigned)c->L.blocks.size();
        void *at[] = {(void *)&a, (void *)&b, (void *)&xin, (void *)&xout, (void *)&pp};
        if (nb > 0)
            HIPCHK(c, hipLaunchKernel((const void *)kt, dim3(nb), dim3(c->block_bs), at, block_lds(c), c->stream));
        int ns = (int)c->L.svid.size();
        void *av[] = {(void *)&a, (void *)&b, (void *)&xin, (void *)&xout, (void *)&pp, (void *)&ns};
        if (ns > 0)
            HIPCHK(c, hipLaunchKernel((const void *)kv, dim3((unsigned)((ns + 255) / 256)), dim3(256), av, 0,
                                      c->stream));
    } else if (c->flat) {
        FlatTetKernel kt = pick_flat_tets(c->feature, lam);
        FlatVertKernel kv = pick_flat_verts(c->feature, mode);
        const int64_t ne = (int64_t)c->L.meta.size();
        float4 *corr = c->CORR;
        void *at[] = {(void *)&a, (void *)&xin, (void *)&pp, (void *)&corr};
        if (ne > 0)
            HIPCHK(c, hipLaunchKernel((const void *)kt, dim3((unsigned)((ne + 255) / 256)), dim3(256), at, 0,
                                      c->stream));
        const float4 *cc = corr;
        int nv = c->nv;
        void *av[] = {(void *)&a, (void *)&xin, (void *)&xout, (void *)&pp, (void *)&cc, (void *)&nv};
        HIPCHK(c, hipLaunchKernel((const void *)kv, dim3((unsigned)((4 * (int64_t)nv + 255) / 256)), dim3(256), av,
                                  0, c->stream));
    } else {
        IterKernel k = pick_iter(c->P, c->feature, mode);
        const size_t lds = iter_lds(c);
        void *args[] = {(void *)&a, (void *)&xin, (void *)&xout, (void *)&pp};
        HIPCHK(c, hipLaunchKernel((const void *)k, dim3((unsigned)c->L.patches.size()), dim3(c->P), args, lds,
                                  c->stream));
    }
    if (tm) {
        HIPCHK(c, hipEventRecord(ev.b, c->stream));
        c->evs.push_back(ev);
    }
    return XPBD_OK;
}

// all `iters` plain iterations of a substep in one launch of the resident
// kernel: positions from xin, the result in xout (xout may be xin: every part
// has read its vertices before any neighbour can finish an iteration)
int launch_resident(xpbd_ctx *c, const float4 *xin, float4 *xout, int iters) {
    if (iters <= 0) return XPBD_OK;
    IterArgs a = make_args(c, c->DPW);
    BlockArgs b = make_block_args(c);
    ResArgs rs;
    rs.parts = c->d_parts;
    rs.prec = c->d_prec;
    rs.bmap = c->d_bmap;
    rs.plidx = c->d_plidx;
    rs.nbr = c->d_nbr;
    const size_t np = std::max<size_t>((size_t)c->L.n_partial, 1);
    rs.xp = c->RXP;
    rs.np = (int)np;
    rs.count = c->d_rcount;
    rs.status = c->d_rstatus;
    rs.max_nloc = c->parts.max_nloc;
    rs.spin_limit = 1 << 20;
    rs.nclass = (int)c->L.classes.size();
    ResKernel k = pick_resident(c->feature, c->res_vpt);
    const float4 *pp = c->PPv;
    EvPair ev{};
    const bool tm = c->timing;
    if (tm) {
        HIPCHK(c, hipEventCreate(&ev.a));
        HIPCHK(c, hipEventCreate(&ev.b));
        HIPCHK(c, hipEventRecord(ev.a, c->stream));
    }
    void *args[] = {(void *)&a, (void *)&b, (void *)&rs, (void *)&xin, (void *)&xout, (void *)&pp, (void *)&iters};
    HIPCHK(c, hipLaunchKernel((const void *)k, dim3((unsigned)c->parts.parts.size()), dim3(c->block_bs), args,
                              c->res_lds, c->stream));
    if (tm) {
        HIPCHK(c, hipEventRecord(ev.b, c->stream));
        c->evs.push_back(ev);
    }
    return XPBD_OK;
}

// the resident kernel's status word: set when a wait for a neighbouring part ran out
int check_resident(xpbd_ctx *c) {
    if (!c->res_R || !c->d_rstatus) return XPBD_OK;
    uint32_t st = 0;
    HIPCHK(c, hipMemcpy(&st, c->d_rstatus, sizeof(st), hipMemcpyDeviceToHost));
    if (st)
        return fail(c, XPBD_EHIP, "resident iterations: a wait for a neighbouring part ran out (the parts were not "
                                  "all resident at once); positions are invalid until xpbd_finalize runs again");
    return XPBD_OK;
}

inline dim3 vgrid(int n) { return dim3((unsigned)((n + 255) / 256)); }

int launch_integrate(xpbd_ctx *c) {
    hipLaunchKernelGGL(k_integrate, vgrid(c->nv), dim3(256), 0, c->stream, c->nv, c->X[0], c->PPv, c->VEL,
                       c->has_force ? (const float4 *)c->FRC : nullptr, (const uint32_t *)c->FLAGS,
                       c->pins_on ? (const float4 *)c->PIN : nullptr, c->sp);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

// integ: also the next substep's integrate (k_post<1>); dcd_apply: also the
// apply of the substep's last DCD pass
int launch_post(xpbd_ctx *c, const float4 *xc, int do_sdf, int do_vel, bool integ = false, bool dcd_apply = false) {
    SdfParams sd{};
    sd.nbones = c->nbones;
    sd.res = c->res;
    sd.contact_offset = c->sdf_offset;
    sd.max_push = c->sdf_maxpush;
    sd.grids = c->d_grids;
    sd.bone = c->d_bone;
    sd.vgroup = c->d_vgroup;
    sd.goff = c->d_goff;
    // the bone mesh, when it resolves contacts, takes the grid SDF's place: the
    // DCD pass's apply first, then the mesh push in place, then velocities
    const bool mesh = do_sdf && c->bm.built && c->bm.resolve;
    if (mesh) {
        if (dcd_apply) {
            hipLaunchKernelGGL(k_dcd_apply, vgrid(c->nv), dim3(256), 0, c->stream, c->dcd, (float4 *)xc, c->nv);
            HIPCHK(c, hipGetLastError());
            dcd_apply = false;
        }
        int r = bm_push(c, (float4 *)xc);
        if (r) return r;
    }
    const int sdf_on = (!mesh && do_sdf && c->nbones > 0 && c->d_grids) ? 1 : 0;
    const float4 *frc = c->has_force ? (const float4 *)c->FRC : nullptr;
    const float4 *pin = c->pins_on ? (const float4 *)c->PIN : nullptr;
    unsigned long long *dacc = dcd_apply ? c->dcd.acc : nullptr;
    uint32_t *dacc_n = dcd_apply ? c->dcd.acc_n : nullptr, *dnc = dcd_apply ? c->dcd.ncontact : nullptr;
    if (integ)
        hipLaunchKernelGGL(k_post<1>, vgrid(c->nv), dim3(256), 0, c->stream, c->nv, xc, c->X[0], c->PPv, c->VEL, sd,
                           sdf_on ? c->d_cells : nullptr, c->sp, sdf_on, do_vel, frc, (const uint32_t *)c->FLAGS,
                           pin, dacc, dacc_n, dnc);
    else
        hipLaunchKernelGGL(k_post<0>, vgrid(c->nv), dim3(256), 0, c->stream, c->nv, xc, c->X[0], c->PPv, c->VEL, sd,
                           sdf_on ? c->d_cells : nullptr, c->sp, sdf_on, do_vel, frc, (const uint32_t *)c->FLAGS,
                           pin, dacc, dacc_n, dnc);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}


int launch_scan(xpbd_ctx *c, const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total) {
    const int nb = (int)((n + SCAN_BLOCK - 1) / SCAN_BLOCK);
    if (nb > c->n_scan_bsum) return fail(c, XPBD_EINVAL, "scan scratch too small");
    if (nb > 0) {
        hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, c->stream, in, out, c->d_scan_bsum, n);
        HIPCHK(c, hipGetLastError());
    }
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, c->stream, c->d_scan_bsum, nb, total);
    HIPCHK(c, hipGetLastError());
    if (nb > 0) {
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, out,
                           (const uint32_t *)c->d_scan_bsum, n);
        HIPCHK(c, hipGetLastError());
    }
    return XPBD_OK;
}

// one DCD pass on positions X (in place); rebuild = broadphase first;
// apply = false leaves the correction in the accumulators for k_post
int launch_dcd(xpbd_ctx *c, float4 *X, bool rebuild, bool apply = true) {
    DcdDev &d = c->dcd;
    const int m = d.ntris;
    const int64_t T = (int64_t)1 << d.tbits;
    int r;
    auto g1 = [](int64_t n) { return dim3((unsigned)((n + 255) / 256)); };
    if (rebuild) {
        hipLaunchKernelGGL(k_dcd_reset, g1(T), dim3(256), 0, c->stream, d, T);
        hipLaunchKernelGGL(k_dcd_snapshot, g1(c->nv), dim3(256), 0, c->stream, d.xprev, (const float4 *)X, c->nv);
        hipLaunchKernelGGL(k_dcd_tridata, g1(m), dim3(256), 0, c->stream, d, (const float4 *)X);
        hipLaunchKernelGGL(k_dcd_cellsize, dim3(1), dim3(64), 0, c->stream, d);
        hipLaunchKernelGGL(k_dcd_keys, g1(m), dim3(256), 0, c->stream, d);
        HIPCHK(c, hipGetLastError());
        if ((r = launch_scan(c, d.bucket_cnt, d.bucket_off, T, d.bucket_off + T))) return r;
        hipLaunchKernelGGL(k_dcd_scatter, g1(m), dim3(256), 0, c->stream, d);
        hipLaunchKernelGGL(k_dcd_search, g1((int64_t)DCD_LANES * m), dim3(256), 0, c->stream, d);
        HIPCHK(c, hipGetLastError());
    }
    hipLaunchKernelGGL(k_dcd_narrow, dim3(1024), dim3(256), 0, c->stream, d, (const float4 *)X);
    if (apply) hipLaunchKernelGGL(k_dcd_apply, g1(c->nv), dim3(256), 0, c->stream, d, X, c->nv);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

int alloc_dcd(xpbd_ctx *c) {
    DcdDev &d = c->dcd;
    d = DcdDev{};
    const int m = (int)(c->dcd_tris.size() / 3);
    d.ntris = m;
    int bits = 10;
    while (((int64_t)1 << bits) < 2 * (int64_t)m && bits < 30) ++bits;
    d.tbits = bits;
    const int64_t T = (int64_t)1 << bits;
    d.max_pairs = c->dcd_p.max_pairs > 0 ? c->dcd_p.max_pairs : ((int64_t)1 << 22);
    d.thickness = c->dcd_p.thickness;
    d.multiplier = c->dcd_p.multiplier;
    d.mu_s = c->dcd_p.friction_static;
    d.mu_d = c->dcd_p.friction_dynamic;
    d.offset_eps = c->dcd_p.offset;
    d.normal_cull = c->dcd_p.normal_cull;
    d.allow_self = c->dcd_p.allow_self;
    d.cell_fixed = c->dcd_p.cell_size > 0.0f ? c->dcd_p.cell_size : 0.0f;
    std::vector<int4> tt(m);
    for (int i = 0; i < m; ++i)
        tt[i] = make_int4(c->L.iperm[c->dcd_tris[3 * i]], c->L.iperm[c->dcd_tris[3 * i + 1]],
                          c->L.iperm[c->dcd_tris[3 * i + 2]], 0);
    int r;
    int4 *dt = nullptr;
    if ((r = dupload(c, &dt, tt.data(), tt.size()))) return r;
    d.tris = dt;
    if (!c->dcd_group.empty()) {
        int32_t *g = nullptr;
        if ((r = dupload(c, &g, c->dcd_group.data(), c->dcd_group.size()))) return r;
        d.group = g;
    }
    if ((r = dalloc(c, &d.cr, m)) || (r = dalloc(c, &d.nrm, m)) || (r = dalloc(c, &d.cell, m)) ||
        (r = dalloc(c, &d.bucket_cnt, T)) || (r = dalloc(c, &d.bucket_off, T + 1)) ||
        (r = dalloc(c, &d.bucket_fill, T)) || (r = dalloc(c, &d.key, m)) || (r = dalloc(c, &d.tv, m)) ||
        (r = dalloc(c, &d.npairs, 1)) || (r = dalloc(c, &d.pairs_seg, (size_t)d.max_pairs)) ||
        (r = dalloc(c, &d.pair_cnt, (size_t)m + 1)) || (r = dalloc(c, &d.pair_off, (size_t)m + 1)) ||
        (r = dalloc(c, &d.pairs, (size_t)d.max_pairs)) ||
        (r = dalloc(c, &d.acc, 3 * (size_t)c->nv)) || (r = dalloc(c, &d.acc_n, (size_t)c->nv)) ||
        (r = dalloc(c, &d.xprev, (size_t)c->nv)) || (r = dalloc(c, &d.scal, 4)) ||
        (r = dalloc(c, &d.ncontact, 2)) || (r = dalloc(c, &d.overflow, 3)) ||
        (r = dalloc(c, &d.tlvl, (size_t)std::max(m, 1))) || (r = dalloc(c, &d.rec, (size_t)std::max(m, 1))) ||
        (r = dalloc(c, &d.hist, (size_t)DCD_HIST)) || (r = dalloc(c, &d.lvl_cnt, (size_t)DCD_LEVELS)) ||
        (r = dalloc(c, &d.pair_fill, (size_t)m + 1)))
        return r;
    HIPCHK(c, hipMemset(d.overflow, 0, 3 * sizeof(uint32_t)));
    HIPCHK(c, hipMemset(d.acc, 0, 3 * (size_t)c->nv * sizeof(unsigned long long)));
    HIPCHK(c, hipMemset(d.acc_n, 0, (size_t)c->nv * sizeof(uint32_t)));
    HIPCHK(c, hipMemset(d.npairs, 0, sizeof(unsigned long long)));
    HIPCHK(c, hipMemset(d.ncontact, 0, 2 * sizeof(uint32_t)));
    const int64_t need = std::max<int64_t>(T, (int64_t)m + 1);
    const int nb = (int)((need + SCAN_BLOCK - 1) / SCAN_BLOCK) + 1;
    if (nb > c->n_scan_bsum) {
        if ((r = dalloc(c, &c->d_scan_bsum, (size_t)nb))) return r;
        c->n_scan_bsum = nb;
    }
    c->dcd_on = m > 0;
    c->dcd_fresh = true;
    return XPBD_OK;
}

// static network features, computed once per layout, masses and parameters
// (called outside graph capture, before the first use)
int ensure_static_features(xpbd_ctx *c) {
    if (c->sfeat_ok || !c->flat || !c->SFEAT) return XPBD_OK;
    const int64_t ne = (int64_t)c->L.meta.size();
    IterArgs a = make_args(c, c->DPW);
    hipLaunchKernelGGL(k_static_features, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, c->stream, a,
                       (const float4 *)c->X[0], c->SFEAT);
    HIPCHK(c, hipGetLastError());
    c->sfeat_ok = true;
    return XPBD_OK;
}

// start-of-substep multiplier set-up for the effective mode of this substep
int launch_warm_start(xpbd_ctx *c) {
    const int64_t ne = (int64_t)c->L.meta.size();
    auto g1 = [](int64_t n) { return dim3((unsigned)((n + 255) / 256)); };
    c->impulse_pending = false;
    if (c->eff_mode == 1) {
        hipLaunchKernelGGL(k_zero4, g1(ne), dim3(256), 0, c->stream, ne, c->LAMB);
    } else if (c->eff_mode == 2) {
        if (!c->mlp_ready) return fail(c, XPBD_EINVAL, "lambda mode 2 needs xpbd_set_mlp()");
        IterArgs a = make_args(c, c->DPW);
        WarmNet net{c->mW1, c->mW2, c->mb1, c->mb2, c->mW3, c->mb3, c->fmean, c->finv, c->oscale};
        // four 32-element tiles per workgroup pass; 2,048 workgroups walk the rest grid-stride
        const int64_t ntile = (ne + 127) / 128;
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ntile, 2048));
        int r = ensure_static_features(c);
        if (r) return r;
        // inside a schedule (mode 4) the base stays the last reference solve's multipliers
        hipLaunchKernelGGL(k_warm_start, dim3(grid), dim3(256), 0, c->stream, a, net, (const float4 *)c->X[0],
                           c->LAMB0, (const float4 *)c->SFEAT, c->LAMB, c->lam_mode == 4 ? 0 : 1);
        c->impulse_pending = true;
    } else if (c->eff_mode == 3) {
        hipLaunchKernelGGL(k_copy4, g1(ne), dim3(256), 0, c->stream, ne, c->LAMB, (const float4 *)c->LAMB0);
        c->impulse_pending = true;
    }
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

// end of a substep: a reference solve (carried from zero) becomes the base
int finish_lambda(xpbd_ctx *c) {
    if (c->eff_mode != 1) return XPBD_OK;
    const int64_t ne = (int64_t)c->L.meta.size();
    hipLaunchKernelGGL(k_copy4, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, c->stream, ne, c->LAMB0,
                       (const float4 *)c->LAMB);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

// lambda template of the next iteration: mode 1 carries lambda from zero; modes
// 2 and 3 apply lambda0 as an impulse in their first iteration and then run
// plain zero-lambda iterations, which keep lowering the plain solver's own
// residual (a carried lambda converges towards the compliant solution, whose
// residual is alpha * lambda, and overshoots the plain residual over a frame)
int lam_for_iteration(xpbd_ctx *c) {
    if (c->eff_mode == 0) return 0;
    if (c->impulse_pending) {
        c->impulse_pending = false;
        return 2;
    }
    return c->eff_mode == 1 ? 1 : 0;
}

// restrict -> coarse Jacobi ARAP sweeps (coarse pins projected with their
// inverse mass and kept at their restricted positions) -> delta prolongation
// -> pins, on the fine context's stream
int launch_vcycle(xpbd_ctx *c, bool pins = true) {
    const int nc = c->coarse->nv;
    hipLaunchKernelGGL(k_mg_fit, dim3((unsigned)c->mg_nblk), dim3(MGF_T), 0, c->stream, c->mg, (const float4 *)c->X[0]);
    const dim3 gw((unsigned)((64 * (int64_t)nc + 255) / 256));   // one wavefront per coarse node
    hipLaunchKernelGGL(k_mg_rhs, gw, dim3(256), 0, c->stream, c->mg, nc, (const float4 *)c->X[0]);
    hipLaunchKernelGGL(k_mg_solve, gw, dim3(256), 0, c->stream, c->mg, nc, c->MG_QPREV);
    hipLaunchKernelGGL(k_mg_coarse, dim3((unsigned)c->mg_nblk), dim3(MGC_T), c->mgc_lds, c->stream, c->mg, c->mgc,
                       c->mg_iters, c->coarse->X[0]);
    hipLaunchKernelGGL(k_mg_prolong, vgrid(c->nv), dim3(256), 0, c->stream, c->nv, (const int4 *)c->d_mg_pidx,
                       (const float4 *)c->d_mg_pw, (const float4 *)c->mg.q, (const float4 *)c->MG_QPREV, c->X[0],
                       c->mg_kinematic ? c->PPv : nullptr, c->pins_on ? (const float4 *)c->PIN : nullptr);
    HIPCHK(c, hipGetLastError());
    // pinned vertices: inside a substep the integrate that precedes the
    // correction put them on their targets and the prolongation leaves them
    // alone, so only a stand-alone V-cycle ends with the pin pass
    if (pins && c->pins_on) {
        hipLaunchKernelGGL(k_pins, vgrid(c->nv), dim3(256), 0, c->stream, c->nv, c->X[0], (const float4 *)c->PIN);
        HIPCHK(c, hipGetLastError());
    }
    return XPBD_OK;
}

int enqueue_substep(xpbd_ctx *c, int iters, bool dcd_rebuild, int sub = 0, bool integrated = false,
                    bool fuse_next = false);

// the fascia shell's substep (Fascia.build_program_ride, fascia.py:518-538): ride, integrate,
// iters x (projection, apply, ride) -- the ride after each iteration is the pin overwrite
int launch_fascia(xpbd_ctx *c, float4 *Xm) {
    xpbd_ctx *f = c->fascia;
    if (f->gen != c->fascia_gen || !f->ready)
        return fail(c, XPBD_ENOTREADY, "the fascia context was finalized again; call xpbd_set_fascia again");
    if (c->fascia_mode == 1) {
        // squeeze (Fascia.build_program, fascia.py:547-565): anchors, integrate, iters x (fascia
        // iteration, glue on this context's positions)
        auto g = [](int n) { return dim3((unsigned)((n + 255) / 256)); };
        if (c->fascia_nanchor > 0)
            hipLaunchKernelGGL(k_fascia_anchor, g(c->fascia_nanchor), dim3(256), 0, c->stream, c->fascia_nanchor,
                               (const int32_t *)c->d_fanc, (const float4 *)c->d_fanc_off, (const float4 *)c->d_ff,
                               f->X[0]);
        HIPCHK(c, hipGetLastError());
        int r = launch_integrate(f);
        if (r) return fail(c, r, std::string("fascia integrate: ") + f->err);
        const float inv_dt2 = 1.0f / (f->sp.dt * f->sp.dt);
        for (int it = 0; it < c->fascia_iters; ++it) {
            if ((r = launch_iter(f, f->X[it & 1], f->X[(it + 1) & 1], MODE_FUSED, 0)))
                return fail(c, r, std::string("fascia iteration: ") + f->err);
            if (c->fascia_nglue > 0)
                hipLaunchKernelGGL(k_fascia_glue, g(c->fascia_nglue), dim3(256), 0, c->stream, c->fascia_nglue,
                                   (const int4 *)c->d_fglue, (const float4 *)f->X[(it + 1) & 1], Xm, inv_dt2,
                                   c->fascia_omega);
            HIPCHK(c, hipGetLastError());
        }
        if (c->fascia_iters & 1)
            HIPCHK(c, hipMemcpyAsync(f->X[0], f->X[1], f->nv * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
        return XPBD_OK;
    }
    if (c->fascia_nbound > 0) {
        hipLaunchKernelGGL(k_fascia_ride, dim3((unsigned)((c->fascia_nbound + 255) / 256)), dim3(256), 0, c->stream,
                           c->fascia_nbound, Xm, (const int32_t *)c->d_fb_vid, (const int4 *)c->d_fb_tri,
                           (const float4 *)c->d_fb_w, f->X[0], f->PIN);
        HIPCHK(c, hipGetLastError());
    }
    const int r = enqueue_substep(f, c->fascia_iters, false, 0, false, false);
    if (r) return fail(c, r, std::string("fascia substep: ") + f->err);
    return XPBD_OK;
}

// one substep, positions start and end in X[0]; sub = index within the frame
// integrated: the previous substep's post already ran this substep's
// integrate; fuse_next: this substep's post runs the next one's integrate
int enqueue_substep(xpbd_ctx *c, int iters, bool dcd_rebuild, int sub, bool integrated, bool fuse_next) {
    int r = integrated ? XPBD_OK : launch_integrate(c);
    if (r) return r;
    if (c->coarse && c->mg_iters > 0 && (r = launch_vcycle(c, false))) return r;
    // mode 4: reference substeps (the first of a frame, or every ref_every-th)
    // are solves carried from zero with ref_iters iterations that refresh the
    // base; the others are warm starts (network or base) with iters iterations
    bool ref = false;
    if (c->lam_mode == 4) {
        ref = c->ref_every > 0 ? (sub % c->ref_every == 0) : (sub == 0);
        c->eff_mode = ref ? 1 : (c->warm_net ? 2 : 3);
        if (ref) iters = c->ref_iters;
    } else {
        c->eff_mode = c->lam_mode;
    }
    if (c->eff_mode && (r = launch_warm_start(c))) return r;
    if (c->res_R > 0 && c->eff_mode == 0) {
        if ((r = launch_resident(c, c->X[0], c->X[iters & 1], iters))) return r;
    } else {
        for (int it = 0; it < iters; ++it) {
            r = launch_iter(c, c->X[it & 1], c->X[(it + 1) & 1], MODE_FUSED, lam_for_iteration(c));
            if (r) return r;
        }
    }
    if ((r = finish_lambda(c))) return r;
    // post phase (mg.py:1362-1376): the fascia shell's substep, the self DCD passes, then the
    // bone push-out and velocities
    if (c->fascia && (r = launch_fascia(c, c->X[iters & 1]))) return r;
    // the last pass's apply runs inside the post kernel
    const bool dcd = c->dcd_on && c->dcd_iters > 0;
    if (dcd)
        for (int k = 0; k < c->dcd_iters; ++k)
            if ((r = launch_dcd(c, c->X[iters & 1], k == 0 && (dcd_rebuild || c->dcd_rebuild_substeps),
                                k + 1 < c->dcd_iters)))
                return r;
    return launch_post(c, c->X[iters & 1], 1, 1, fuse_next, dcd);
}

// resident mode: parts of R consecutive blocks, the smallest R in the allowed
// range whose parts all fit on the device at once (occupancy of the kernel at
// its LDS size times the compute units), then the part tables on the device
int setup_resident(xpbd_ctx *c, int want, int maxlds) {
    c->res_R = 0;
    if (want == 0) return XPBD_OK;
    if (want < -2 || want > 4) return fail(c, XPBD_EINVAL, "resident_blocks must be -2, -1, 0 or 1..4");
    if (c->block != c->block_bs || c->block_bs > 256)
        return fail(c, XPBD_ELAYOUT, "resident mode needs block_tets == block_threads <= 256");
    if (c->L.max_nloc > c->block_bs)
        return fail(c, XPBD_ELAYOUT, "resident mode: a block touches more vertices than it has threads");
    if ((int64_t)c->L.n_partial > (1 << 26))
        return fail(c, XPBD_ELAYOUT, "resident mode: more than 2^26 block partials");
    if (c->L.blocks.empty()) return fail(c, XPBD_ELAYOUT, "resident mode needs at least one element");
    int ncu = 0;
    HIPCHK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    const int r0 = want > 0 ? want : 1, r1 = want > 0 ? want : 4;
    std::string why = "no part size fits";
    // -2: as many parts as fit at once (blocks spread evenly over them), found by
    // shrinking the part count to the capacity its own part size allows
    const int64_t nb = (int64_t)c->L.blocks.size();
    int64_t target = want == -2 ? nb : 0;
    const int tries = want == -2 ? 6 : r1 - r0 + 1;
    for (int t = 0, R = r0; t < tries; ++t, ++R) {
        xpbd::Parts P;
        int code = 0;
        std::string m = xpbd::build_parts(c->L, c->nv, want == -2 ? 0 : R, P, &code, (int)target);
        if (code) { why = m; continue; }
        const int vpt = (P.max_nloc + c->block_bs - 1) / c->block_bs;
        if (vpt > 3) {
            why = "a part of " + std::to_string(R) + " blocks has " + std::to_string(P.max_nloc) +
                  " vertices, more than 3 per thread";
            continue;
        }
        const size_t lds = resident_lds(c->feature, P.max_nloc, c->block, (int)c->L.classes.size());
        if (lds > (size_t)maxlds) { why = "a part needs " + std::to_string(lds) + " B of LDS"; continue; }
        ResKernel k = pick_resident(c->feature, std::max(vpt, 1));
        HIPCHK(c, hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int occ = 0;
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)k, c->block_bs, lds));
        if ((int64_t)occ * ncu < (int64_t)P.parts.size()) {
            why = std::to_string(P.parts.size()) + " parts of up to " + std::to_string(P.R) + " blocks, but only " +
                  std::to_string(occ) + " x " + std::to_string(ncu) + " fit on the device at once";
            if (want == -2) target = std::min<int64_t>((int64_t)P.parts.size() - 1, (int64_t)occ * ncu);
            continue;
        }
        c->parts = std::move(P);
        c->res_R = c->parts.R;
        c->res_lds = lds;
        c->res_vpt = std::max(vpt, 1);
        break;
    }
    if (!c->res_R) return fail(c, XPBD_ELAYOUT, "resident mode: " + why);
    const xpbd::Parts &P = c->parts;
    int r;
    if ((r = dupload(c, &c->d_parts, P.parts.data(), P.parts.size()))) return r;
    if ((r = dupload(c, &c->d_prec, (const int4 *)P.prec.data(), P.prec.size() / 4))) return r;
    if ((r = dupload(c, &c->d_bmap, P.bmap.data(), P.bmap.size()))) return r;
    if ((r = dupload(c, &c->d_plidx, (const ushort4 *)P.plidx.data(), P.plidx.size() / 4))) return r;
    if ((r = dupload(c, &c->d_nbr, P.nbr.data(), P.nbr.size()))) return r;
    const size_t np = std::max<size_t>((size_t)c->L.n_partial, 1);
    if ((r = dalloc(c, &c->RXP, 2 * np))) return r;
    if ((r = dalloc(c, &c->d_rcount, P.parts.size()))) return r;
    if ((r = dalloc(c, &c->d_rstatus, 4))) return r;
    HIPCHK(c, hipMemset(c->d_rcount, 0, P.parts.size() * sizeof(uint32_t)));
    HIPCHK(c, hipMemset(c->d_rstatus, 0, 4 * sizeof(uint32_t)));
    return XPBD_OK;
}

int check_ready(xpbd_ctx *c) {
    if (!c) return XPBD_EINVAL;
    if (!c->ready) return fail(c, XPBD_ENOTREADY, "call xpbd_finalize() first");
    return XPBD_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int32_t xpbd_create(int32_t device, xpbd_ctx **out) {
    if (!out) return XPBD_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return XPBD_EHIP;
    if (device < 0 || device >= n) return XPBD_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return XPBD_EHIP;
    xpbd_ctx *c = new xpbd_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return XPBD_EHIP;
    }
    c->stream = c->own_stream;
    c->sp.dt = 1.0f / 240.0f;
    c->sp.exp_damp = 1.0f;
    c->sp.fiber_scale = 1.0f;
    *out = c;
    return XPBD_OK;
}

void xpbd_destroy(xpbd_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &e : c->evs) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    free_device(c);
    free_stage(c->st_pin);
    free_stage(c->st_act);
    free_stage(c->st_bone);
    free_stage(c->st_att);
    free_stage(c->st_gact);
    free_stage(c->st_mgpin);
    free_stage(c->bm.st);
    free_stage(c->st_ff);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
}

const char *xpbd_last_error(const xpbd_ctx *c) { return c ? c->err.c_str() : "null context"; }

int32_t xpbd_set_stream(xpbd_ctx *c, void *stream) {
    if (!c) return XPBD_EINVAL;
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    invalidate_graph(c);
    if (c->fascia) {                 // the fascia's substeps run inside this context's
        c->fascia->stream = c->stream;
        invalidate_graph(c->fascia);
    }
    return XPBD_OK;
}

int32_t xpbd_set_mesh(xpbd_ctx *c, int32_t nv, const float *pos0, int32_t nt, const int32_t *tets,
                      float density) {
    if (!c || nv <= 0 || nt < 0 || !pos0 || (nt > 0 && !tets)) return fail(c, XPBD_EINVAL, "bad mesh");
    for (int64_t i = 0; i < 4 * (int64_t)nt; ++i)
        if (tets[i] < 0 || tets[i] >= nv) return fail(c, XPBD_EINVAL, "tet index out of range");
    free_device(c);
    c->nv = nv;
    c->nt = nt;
    c->pos0.assign(pos0, pos0 + 3 * (size_t)nv);
    c->tets.assign(tets, tets + 4 * (size_t)nt);
    c->rest_volume.assign(nt, 0.f);
    c->rest_matrix.assign(9 * (size_t)nt, 0.f);
    c->mass.assign(nv, 0.f);
    xpbd::precompute_rest(nv, pos0, nt, tets, density, c->rest_volume.data(), c->rest_matrix.data(),
                          c->mass.data());
    c->stopped.assign(nv, 0);
    c->force.clear();
    c->has_force = false;
    c->fiber.clear();
    c->tendonmask.assign(nt, 0.f);
    c->activation.assign(nt, 0.f);
    c->coll_w.clear();
    return XPBD_OK;
}

int32_t xpbd_get_rest(xpbd_ctx *c, float *rv, float *rm, float *mass) {
    if (!c || c->nv == 0) return fail(c, XPBD_EINVAL, "no mesh");
    if (rv) std::memcpy(rv, c->rest_volume.data(), c->rest_volume.size() * sizeof(float));
    if (rm) std::memcpy(rm, c->rest_matrix.data(), c->rest_matrix.size() * sizeof(float));
    if (mass) std::memcpy(mass, c->mass.data(), c->mass.size() * sizeof(float));
    return XPBD_OK;
}

int32_t xpbd_scale_rest_shape(xpbd_ctx *c, float scale) {
    if (!c || c->nt == 0) return fail(c, XPBD_EINVAL, "no mesh");
    if (c->ready) return fail(c, XPBD_EINVAL, "scale the rest shape before xpbd_finalize");
    if (!(scale > 0.0f) || !std::isfinite(scale)) return fail(c, XPBD_EINVAL, "rest scale must be > 0");
    if (std::fabs((double)scale - 1.0) < 1.0e-9) return XPBD_OK;
    const double k = 1.0 / (double)scale;
    for (float &b : c->rest_matrix) b = (float)((double)b * k);
    return XPBD_OK;
}

int32_t xpbd_set_mass(xpbd_ctx *c, const float *mass) {
    if (!c || !mass || c->nv == 0) return fail(c, XPBD_EINVAL, "bad mass");
    c->mass.assign(mass, mass + c->nv);
    return c->ready ? upload_vertex_static(c) : XPBD_OK;
}

int32_t xpbd_set_stopped(xpbd_ctx *c, const int32_t *st) {
    if (!c || !st || c->nv == 0) return fail(c, XPBD_EINVAL, "bad stopped");
    c->stopped.assign(st, st + c->nv);
    return c->ready ? upload_vertex_static(c) : XPBD_OK;
}

int32_t xpbd_set_force(xpbd_ctx *c, const float *f) {
    if (!c || c->nv == 0) return fail(c, XPBD_EINVAL, "bad force");
    if (!f) { c->force.clear(); c->has_force = false; }
    else { c->force.assign(f, f + 3 * (size_t)c->nv); c->has_force = true; }
    if (c->ready) {
        invalidate_graph(c);
        std::vector<float4> fr(c->nv, make_float4(0, 0, 0, 0));
        if (c->has_force)
            for (int n = 0; n < c->nv; ++n) {
                const int o = c->L.perm[n];
                fr[n] = make_float4(c->force[3 * o], c->force[3 * o + 1], c->force[3 * o + 2], 0.f);
            }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(c->FRC, fr.data(), c->nv * sizeof(float4), hipMemcpyHostToDevice));
    }
    return XPBD_OK;
}

int32_t xpbd_set_fiber(xpbd_ctx *c, const float *vf, const float *tm) {
    if (!c || c->nv == 0) return fail(c, XPBD_EINVAL, "no mesh");
    if (c->ready) return fail(c, XPBD_EINVAL, "fiber data must be set before xpbd_finalize");
    if (vf) c->fiber.assign(vf, vf + 3 * (size_t)c->nv);
    if (tm) c->tendonmask.assign(tm, tm + c->nt);
    return XPBD_OK;
}

// per-tet activation on the host from the last group values (only when a path needs it)
void sync_host_activation(xpbd_ctx *c) {
    if (!c->act_host_stale) return;
    for (int t = 0; t < c->nt; ++t) c->activation[t] = c->group_act[c->act_group[t]];
    c->act_host_stale = false;
}

// device tables of the activation groups (after every finalize)
int upload_act_groups(xpbd_ctx *c) {
    if (c->act_group.empty() || !c->ready) return XPBD_OK;
    const size_t n = c->L.meta.size();
    std::vector<int32_t> eg(std::max<size_t>(n, 1), 0);
    for (size_t g = 0; g < n; ++g) {
        const int t = c->L.ptet_tetid[g];
        eg[g] = (t >= 0 && t < c->nt) ? c->act_group[t] : 0;
    }
    int r;
    if ((r = dupload(c, &c->d_elem_group, eg.data(), eg.size()))) return r;
    if ((r = dalloc(c, &c->d_gact, (size_t)std::max(c->n_act_groups, 1)))) return r;
    return XPBD_OK;
}

int32_t xpbd_set_activation_groups(xpbd_ctx *c, int32_t ngroups, const int32_t *tet_group) {
    if (!c || c->nt == 0) return fail(c, XPBD_EINVAL, "no mesh");
    if (ngroups <= 0 || !tet_group) return fail(c, XPBD_EINVAL, "bad activation groups");
    for (int t = 0; t < c->nt; ++t)
        if (tet_group[t] < 0 || tet_group[t] >= ngroups) return fail(c, XPBD_EINVAL, "activation group out of range");
    sync_host_activation(c);
    c->act_group.assign(tet_group, tet_group + c->nt);
    c->n_act_groups = ngroups;
    c->group_act.assign(ngroups, 0.f);
    return upload_act_groups(c);
}

int32_t xpbd_set_group_activation(xpbd_ctx *c, int32_t ngroups, const float *act) {
    if (!c || !act) return fail(c, XPBD_EINVAL, "bad activation");
    if (c->act_group.empty() || ngroups != c->n_act_groups)
        return fail(c, XPBD_EINVAL, "group count differs from xpbd_set_activation_groups");
    c->group_act.assign(act, act + ngroups);
    c->act_host_stale = true;
    if (c->ready && (c->feature & F_ANISO)) {
        int slot;
        void *h;
        const size_t bytes = (size_t)ngroups * sizeof(float);
        int r = stage_acquire(c, c->st_gact, bytes, slot, h);
        if (r) return r;
        std::memcpy(h, act, bytes);
        if ((r = stage_submit(c, c->st_gact, slot, c->d_gact, bytes))) return r;
        const int64_t n = (int64_t)c->L.meta.size();
        hipLaunchKernelGGL(k_refresh_fiber_groups, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                           (const float4 *)c->FIBS, (const int32_t *)c->d_elem_group, (const float *)c->d_gact,
                           c->sp.contraction_ratio, c->FIB);
        HIPCHK(c, hipGetLastError());
    }
    return XPBD_OK;
}

int32_t xpbd_set_activation(xpbd_ctx *c, const float *act) {
    if (!c || !act || c->nt == 0) return fail(c, XPBD_EINVAL, "bad activation");
    if (act != c->activation.data()) c->activation.assign(act, act + c->nt);
    c->act_host_stale = false;
    if (c->ready && (c->feature & F_ANISO)) {
        int slot;
        void *h;
        const size_t bytes = (size_t)c->nt * sizeof(float);
        int r = stage_acquire(c, c->st_act, bytes, slot, h);
        if (r) return r;
        std::memcpy(h, act, bytes);
        if ((r = stage_submit(c, c->st_act, slot, c->d_act, bytes))) return r;
        const int64_t n = (int64_t)c->L.meta.size();
        hipLaunchKernelGGL(k_refresh_fiber, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                           (const float4 *)c->FIBS, (const int32_t *)c->d_ptet_tetid, (const float *)c->d_act,
                           c->sp.contraction_ratio, c->FIB);
        HIPCHK(c, hipGetLastError());
    }
    return XPBD_OK;
}

int32_t xpbd_upload_constraints(xpbd_ctx *c, int32_t n, const int32_t *type, const int32_t *pts,
                                const float *stiffness, const float *comp, const float *damp,
                                const int32_t *tetid, const float *restlength, const float *restvector,
                                const float *restdir) {
    if (!c || n < 0) return fail(c, XPBD_EINVAL, "bad constraint table");
    if (n > 0 && (!type || !pts || !stiffness || !comp || !damp || !tetid || !restlength || !restvector ||
                  !restdir))
        return fail(c, XPBD_EINVAL, "null constraint field");
    if (c->ready) free_device(c);
    c->ncons = n;
    c->c_type.assign(type, type + n);
    c->c_pts.assign(pts, pts + 4 * (size_t)n);
    c->c_stiff.assign(stiffness, stiffness + n);
    c->c_comp.assign(comp, comp + n);
    c->c_damp.assign(damp, damp + n);
    c->c_tetid.assign(tetid, tetid + n);
    c->c_rest.assign(restlength, restlength + n);
    c->c_rvec.assign(restvector, restvector + 4 * (size_t)n);
    c->c_rdir.assign(restdir, restdir + 3 * (size_t)n);
    return XPBD_OK;
}

int32_t xpbd_set_attach_targets(xpbd_ctx *c, int32_t n, const float *xyz) {
    int r = check_ready(c);
    if (r) return r;
    if (n != (int)c->L.attach_vertex.size() || (n > 0 && !xyz))
        return fail(c, XPBD_EINVAL, "attach target count mismatch");
    if (n == 0) return XPBD_OK;
    for (int i = 0; i < n; ++i) {   // host copy: a later finalize uploads the current targets
        float *a0 = &c->L.att0[4 * (size_t)c->L.attach_vertex[i]];
        a0[0] = xyz[3 * i];
        a0[1] = xyz[3 * i + 1];
        a0[2] = xyz[3 * i + 2];
    }
    // staged in pinned memory and scattered on the device: no host wait
    int slot;
    void *h;
    if ((r = stage_acquire(c, c->st_att, (size_t)n * sizeof(float4), slot, h))) return r;
    float4 *t = (float4 *)h;
    for (int i = 0; i < n; ++i) t[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.0f);
    if ((r = stage_submit(c, c->st_att, slot, c->d_att_tgt, (size_t)n * sizeof(float4)))) return r;
    hipLaunchKernelGGL(k_att_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                       (const int32_t *)c->d_att_dev, (const float4 *)c->d_att_tgt, c->ATT0);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

int32_t xpbd_set_params(xpbd_ctx *c, const xpbd_params *p) {
    if (!c || !p) return fail(c, XPBD_EINVAL, "null params");
    if (!(p->dt > 0.0f)) return fail(c, XPBD_EINVAL, "dt must be > 0");
    SimParams s{};
    s.dt = p->dt;
    s.inv_dt2 = 1.0f / (p->dt * p->dt);
    s.gx = p->gravity[0];
    s.gy = p->gravity[1];
    s.gz = p->gravity[2];
    s.veldamping = p->veldamping;
    s.quasi_static = p->quasi_static;
    s.contraction_ratio = p->contraction_ratio;
    s.fiber_scale = p->fiber_stiffness_scale;
    s.has_compress = p->has_compress;
    s.exp_damp = p->exp_damp;
    s.const_damp = p->const_damp;
    const bool cr_changed = s.contraction_ratio != c->sp.contraction_ratio;
    c->sp = s;
    c->sfeat_ok = false;   // the compliances enter the network's static features
    c->params_set = true;
    invalidate_graph(c);
    if (c->ready && c->d_cls) {
        std::vector<DevClass> dcl = dev_classes(c->L.classes, c->sp);
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(c->d_cls, dcl.data(), dcl.size() * sizeof(DevClass), hipMemcpyHostToDevice));
    }
    if (c->ready && cr_changed && (c->feature & F_ANISO)) {
        sync_host_activation(c);
        return xpbd_set_activation(c, c->activation.data());
    }
    return XPBD_OK;
}

int32_t xpbd_set_pins(xpbd_ctx *c, int32_t n, const int32_t *vids, const int32_t *active) {
    if (!c || n < 0 || (n > 0 && (!vids || !active))) return fail(c, XPBD_EINVAL, "bad pins");
    for (int i = 0; i < n; ++i)
        if (vids[i] < 0 || vids[i] >= c->nv) return fail(c, XPBD_EINVAL, "pin vertex out of range");
    c->pin_vids.assign(vids, vids + n);
    c->pin_active.assign(active, active + n);
    c->pin_targets.clear();
    invalidate_graph(c);
    return c->ready ? upload_pins(c) : XPBD_OK;
}

int32_t xpbd_set_pin_targets(xpbd_ctx *c, int32_t n, const float *xyz) {
    if (!c || n != (int)c->pin_vids.size() || (n > 0 && !xyz)) return fail(c, XPBD_EINVAL, "pin target count");
    const bool was_on = c->pins_on;
    const bool had_targets = !c->pin_targets.empty();
    c->pin_targets.assign(xyz, xyz + 3 * (size_t)n);
    if (!c->ready) return XPBD_OK;
    if (had_targets && c->pins_full && c->n_pin_dev > 0) {
        // same pin set as the table on the device: stage the targets of the
        // active pins (pin order, as d_pin_dev) and scatter them, no host wait
        const int m = c->n_pin_dev;
        int slot;
        void *h;
        int r = stage_acquire(c, c->st_pin, (size_t)m * sizeof(float4), slot, h);
        if (r) return r;
        float4 *t = (float4 *)h;
        int k = 0;
        for (int i = 0; i < n && k < m; ++i)
            if (c->pin_active[i]) t[k++] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 1.0f);
        if (k != m) return fail(c, XPBD_EINVAL, "pin set changed without xpbd_set_pins");
        if ((r = stage_submit(c, c->st_pin, slot, c->d_pin_tgt, (size_t)m * sizeof(float4)))) return r;
        hipLaunchKernelGGL(k_pin_scatter, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream, m,
                           (const int32_t *)c->d_pin_dev, (const float4 *)c->d_pin_tgt, c->PIN);
        HIPCHK(c, hipGetLastError());
        return XPBD_OK;
    }
    int r = upload_pins(c);
    if (r) return r;
    if (was_on != c->pins_on) invalidate_graph(c);
    return XPBD_OK;
}

int32_t xpbd_set_sdf(xpbd_ctx *c, int32_t nbones, int32_t res, const float *grids, const float *origins,
                     const float *spacing, float contact_offset, float max_push, const float *coll_w) {
    if (!c || c->nv == 0) return fail(c, XPBD_EINVAL, "no mesh");
    if (nbones < 0 || (nbones > 0 && (res < 2 || !grids || !origins || !spacing)))
        return fail(c, XPBD_EINVAL, "bad sdf grid");
    c->sdf_invh.assign(nbones, 0.f);
    for (int b = 0; b < nbones; ++b) {
        if (!(spacing[b] > 0.0f)) return fail(c, XPBD_EINVAL, "sdf spacing must be > 0");
        c->sdf_invh[b] = 1.0f / spacing[b];
    }
    c->nbones = nbones;
    c->res = res;
    c->sdf_grids.assign(grids, grids + (size_t)nbones * res * res * res);
    c->sdf_origin.assign(origins, origins + 3 * (size_t)nbones);
    c->sdf_offset = contact_offset;
    c->sdf_maxpush = max_push;
    if (coll_w) c->coll_w.assign(coll_w, coll_w + c->nv);
    else c->coll_w.assign(c->nv, 1.0f);
    c->goff.clear();
    c->vgroup.clear();
    if ((int)c->bone_xf.size() != 12 * nbones) {
        c->bone_xf.assign(12 * (size_t)nbones, 0.f);
        for (int b = 0; b < nbones; ++b) {
            c->bone_xf[12 * b + 0] = c->bone_xf[12 * b + 5] = c->bone_xf[12 * b + 10] = 1.f;
        }
    }
    if (c->ready) free_device(c);
    return XPBD_OK;
}

int32_t xpbd_set_sdf_groups(xpbd_ctx *c, int32_t ngroups, const int32_t *goff, const int32_t *vgroup) {
    if (!c || c->nv == 0 || ngroups <= 0 || !goff || !vgroup) return fail(c, XPBD_EINVAL, "bad sdf groups");
    if (goff[0] != 0 || goff[ngroups] != c->nbones) return fail(c, XPBD_EINVAL, "group offsets must span [0, nbones]");
    for (int g = 0; g < ngroups; ++g)
        if (goff[g + 1] < goff[g]) return fail(c, XPBD_EINVAL, "group offsets must be non-decreasing");
    for (int i = 0; i < c->nv; ++i)
        if (vgroup[i] < 0 || vgroup[i] >= ngroups) return fail(c, XPBD_EINVAL, "vertex group out of range");
    c->goff.assign(goff, goff + ngroups + 1);
    c->vgroup.assign(vgroup, vgroup + c->nv);
    if (c->ready) free_device(c);
    return XPBD_OK;
}

int32_t xpbd_set_bone_transforms(xpbd_ctx *c, int32_t nbones, const float *xf) {
    if (!c || nbones != c->nbones || (nbones > 0 && !xf)) return fail(c, XPBD_EINVAL, "bone count mismatch");
    c->bone_xf.assign(xf, xf + 12 * (size_t)nbones);
    if (c->ready && nbones > 0) {
        int slot;
        void *h;
        const size_t bytes = 16 * (size_t)nbones * sizeof(float);
        int r = stage_acquire(c, c->st_bone, bytes, slot, h);
        if (r) return r;
        float *bn = (float *)h;
        for (int b = 0; b < nbones; ++b) {
            std::memcpy(&bn[16 * b], &c->bone_xf[12 * b], 12 * sizeof(float));
            std::memcpy(&bn[16 * b + 12], &c->sdf_origin[3 * b], 3 * sizeof(float));
            bn[16 * b + 15] = c->sdf_invh[b];
        }
        if ((r = stage_submit(c, c->st_bone, slot, c->d_bone, bytes))) return r;
    }
    return XPBD_OK;
}

int32_t xpbd_set_bone_mesh(xpbd_ctx *c, const xpbd_bone_mesh_desc *d) {
    if (!c || c->nv == 0) return fail(c, XPBD_EINVAL, "no mesh");
    HIPCHK(c, hipSetDevice(c->device));
    BoneMesh &b = c->bm;
    if (!d) {
        if (c->stream) HIPCHK(c, hipStreamSynchronize(c->stream));
        bm_free(b);
        free_stage(b.st);
        b = BoneMesh{};
        invalidate_graph(c);
        return XPBD_OK;
    }
    if (d->nverts < 1 || !d->verts || d->ntris < 1 || !d->tris) return fail(c, XPBD_EINVAL, "empty bone mesh");
    if (d->ntris >= (1 << 29)) return fail(c, XPBD_EINVAL, "bone mesh has too many triangles");
    for (int64_t i = 0; i < 3 * (int64_t)d->ntris; ++i)
        if (d->tris[i] < 0 || d->tris[i] >= d->nverts) return fail(c, XPBD_EINVAL, "bone triangle vertex out of range");
    for (int64_t i = 0; i < 3 * (int64_t)d->nverts; ++i)
        if (!std::isfinite(d->verts[i])) return fail(c, XPBD_EINVAL, "non-finite bone vertex");
    std::vector<int32_t> toff;
    int nb = d->nbones;
    if (!d->tri_off) {
        if (nb > 1) return fail(c, XPBD_EINVAL, "tri_off is needed with more than one sub-mesh");
        nb = 1;
        toff = {0, d->ntris};
    } else {
        if (nb < 1) return fail(c, XPBD_EINVAL, "nbones must be >= 1");
        toff.assign(d->tri_off, d->tri_off + nb + 1);
        if (toff[0] != 0 || toff[nb] != d->ntris) return fail(c, XPBD_EINVAL, "tri_off must span [0, ntris]");
        for (int k = 0; k < nb; ++k)
            if (toff[k + 1] < toff[k]) return fail(c, XPBD_EINVAL, "tri_off must be non-decreasing");
    }
    std::vector<int32_t> goff, vg;
    if (d->ngroups > 0) {
        if (!d->goff || !d->vgroup) return fail(c, XPBD_EINVAL, "groups need goff and vgroup");
        goff.assign(d->goff, d->goff + d->ngroups + 1);
        if (goff[0] != 0 || goff[d->ngroups] != nb) return fail(c, XPBD_EINVAL, "group offsets must span [0, nbones]");
        for (int g = 0; g < d->ngroups; ++g)
            if (goff[g + 1] < goff[g]) return fail(c, XPBD_EINVAL, "group offsets must be non-decreasing");
        vg.assign(d->vgroup, d->vgroup + c->nv);
        for (int i = 0; i < c->nv; ++i)
            if (vg[i] < 0 || vg[i] >= d->ngroups) return fail(c, XPBD_EINVAL, "vertex group out of range");
    }
    if (!(d->max_dist > 0.0f) || !std::isfinite(d->max_dist)) return fail(c, XPBD_EINVAL, "max_dist must be > 0");
    if (d->coherent < 0 || d->coherent > 2) return fail(c, XPBD_EINVAL, "coherent must be 0, 1 or 2");
    if (c->stream) HIPCHK(c, hipStreamSynchronize(c->stream));
    bm_free(b);
    b.set = true;
    b.resolve = d->resolve != 0;
    b.nbv = d->nverts;
    b.ntris = d->ntris;
    b.nbones = nb;
    b.pts.assign(d->verts, d->verts + 3 * (size_t)d->nverts);
    b.tris.assign(d->tris, d->tris + 3 * (size_t)d->ntris);
    b.tri_off = toff;
    b.goff = goff;
    b.vgroup = vg;
    if (d->coll_w) b.cw.assign(d->coll_w, d->coll_w + c->nv); else b.cw.clear();
    if (d->meter_w) b.mw.assign(d->meter_w, d->meter_w + c->nv); else b.mw.clear();
    b.max_dist = d->max_dist;
    b.contact_offset = d->contact_offset;
    b.max_push = d->max_push;
    b.coherent_req = d->coherent;
    b.motion = 0.0;
    invalidate_graph(c);
    return c->ready ? bm_build(c) : XPBD_OK;
}

int32_t xpbd_set_bone_mesh_points(xpbd_ctx *c, const float *verts) {
    if (!c || !c->bm.set) return fail(c, XPBD_EINVAL, "no bone mesh");
    if (!verts) return fail(c, XPBD_EINVAL, "null bone points");
    BoneMesh &b = c->bm;
    for (int64_t i = 0; i < 3 * (int64_t)b.nbv; ++i)
        if (!std::isfinite(verts[i])) return fail(c, XPBD_EINVAL, "non-finite bone vertex");
    double mx = 0.0;
    for (int i = 0; i < b.nbv; ++i) {
        const double dx = (double)verts[3 * i] - b.pts[3 * i], dy = (double)verts[3 * i + 1] - b.pts[3 * i + 1],
                     dz = (double)verts[3 * i + 2] - b.pts[3 * i + 2];
        mx = std::max(mx, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    b.motion += mx;
    b.pts.assign(verts, verts + 3 * (size_t)b.nbv);
    return b.built ? bm_upload_points(c) : XPBD_OK;
}

int32_t xpbd_set_bone_mesh_resolve(xpbd_ctx *c, int32_t resolve) {
    if (!c || !c->bm.set) return fail(c, XPBD_EINVAL, "no bone mesh");
    if ((resolve != 0) != c->bm.resolve) {
        c->bm.resolve = resolve != 0;
        invalidate_graph(c);
    }
    return XPBD_OK;
}

int32_t xpbd_bone_mesh_push(xpbd_ctx *c) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->bm.built) return fail(c, XPBD_EINVAL, "no bone mesh");
    return bm_push(c, c->X[0]);
}

int32_t xpbd_bone_mesh_queries(xpbd_ctx *c, int64_t *queries) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->bm.built || !queries) return fail(c, XPBD_EINVAL, "no bone mesh");
    unsigned long long n = 0;
    HIPCHK(c, hipMemcpyAsync(&n, c->bm.dev.nquery, sizeof(n), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *queries = (int64_t)n;
    return XPBD_OK;
}

int32_t xpbd_bone_penetration(xpbd_ctx *c, float collide_eps, float *depth, int32_t *count, float *max_depth) {
    int r = check_ready(c);
    if (r) return r;
    BoneMesh &b = c->bm;
    if (!b.built) return fail(c, XPBD_EINVAL, "no bone mesh");
    HIPCHK(c, hipMemsetAsync(b.d_stat, 0, 2 * sizeof(uint32_t), c->stream));
    hipLaunchKernelGGL(k_bm_query<1>, vgrid(c->nv), dim3(256), 0, c->stream, c->nv, c->X[0], b.dev, b.d_depth,
                       collide_eps, b.d_stat);
    HIPCHK(c, hipGetLastError());
    uint32_t st[2];
    HIPCHK(c, hipMemcpyAsync(st, b.d_stat, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    std::vector<float> dd;
    if (depth) {
        dd.resize(c->nv);
        HIPCHK(c, hipMemcpyAsync(dd.data(), b.d_depth, c->nv * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (depth)
        for (int n = 0; n < c->nv; ++n) depth[c->L.perm[n]] = dd[n];
    if (count) *count = (int32_t)st[0];
    if (max_depth) {
        float m;
        std::memcpy(&m, &st[1], sizeof(m));
        *max_depth = m;
    }
    return XPBD_OK;
}

int32_t xpbd_finalize(xpbd_ctx *c, const xpbd_tuning *t) {
    if (!c || c->nv == 0) return fail(c, XPBD_EINVAL, "no mesh");
    if (!c->params_set) return fail(c, XPBD_EINVAL, "call xpbd_set_params() first");
    HIPCHK(c, hipSetDevice(c->device));
    free_device(c);
    c->coarse = nullptr;
    c->fascia = nullptr;
    c->fascia_nbound = 0;
    ++c->gen;
    c->mg_iters = 0;
    c->mg_pidx_set = false;
    int P = (t && t->patch_verts > 0) ? t->patch_verts : 256;
    if (P != 64 && P != 128 && P != 256) return fail(c, XPBD_EINVAL, "patch_verts must be 64, 128 or 256");
    c->P = P;
    const int strategy = t ? t->strategy : 0;
    if (strategy < 0 || strategy > 2)
        return fail(c, XPBD_EINVAL, "strategy must be 0 (patch), 1 (flat) or 2 (block)");
    c->flat = strategy >= 1;
    c->block = 0;
    if (strategy == 2) {
        c->block = (t->block_tets > 0) ? t->block_tets : 1024;
        c->block_bs = (t->block_threads > 0) ? t->block_threads : 256;
        if (c->block < 64 || c->block > 8192) return fail(c, XPBD_EINVAL, "block_tets must be in [64, 8192]");
        if (c->block_bs < 64 || c->block_bs > 512 || (c->block_bs & 63))
            return fail(c, XPBD_EINVAL, "block_threads must be a multiple of 64 in [64, 512]");
    }
    xpbd::HostInput in;
    in.nv = c->nv;
    in.nt = c->nt;
    in.pos0 = c->pos0.data();
    in.tets = c->tets.data();
    in.rest_matrix = c->rest_matrix.data();
    in.v_fiber_dir = c->fiber.empty() ? nullptr : c->fiber.data();
    in.tendonmask = c->tendonmask.data();
    in.ncons = c->ncons;
    in.type = c->c_type.data();
    in.pts = c->c_pts.data();
    in.tetid = c->c_tetid.data();
    in.stiffness = c->c_stiff.data();
    in.compressionstiffness = c->c_comp.data();
    in.dampingratio = c->c_damp.data();
    in.restlength = c->c_rest.data();
    in.restvector = c->c_rvec.data();
    in.restdir = c->c_rdir.data();
    in.patch_verts = P;
    in.flat = c->flat;
    in.block_tets = c->block;
    c->L = xpbd::Layout();
    int code = 0;
    std::string m = xpbd::build_layout(in, c->L, &code);
    if (code) return fail(c, code, m);
    // the block kernels address per-element and per-vertex arrays with 32-bit
    // byte offsets
    if (c->block && ((int64_t)c->L.meta.size() >= BLOCK_MAX_ELEMS || (int64_t)c->nv >= BLOCK_MAX_ELEMS))
        return fail(c, XPBD_ELAYOUT, "the block strategy holds fewer than 2^26 elements and vertices; use strategy 1");
    c->feature = c->L.feature;
    const size_t lds = c->block ? block_lds(c) : (c->flat ? 0 : iter_lds(c));
    int maxlds = 0;
    HIPCHK(c, hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    if (lds > (size_t)maxlds)
        return fail(c, XPBD_ELAYOUT, std::string(c->block ? "block" : "patch") + " tables need " +
                                         std::to_string(lds) + " B of LDS; limit " + std::to_string(maxlds));
    if (c->block) {
        const int modes[2] = {MODE_FUSED, MODE_ACCUM};
        for (int mode : modes)
            for (int lam = 0; lam < 3; ++lam) {
                BlockTetKernel k = pick_block_tets(c->feature, lam, mode);
                HIPCHK(c, hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            }
    }
    if (!c->flat && lds > 65536) {
        IterKernel k = pick_iter(c->P, c->feature, MODE_FUSED);
        HIPCHK(c, hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        IterKernel k2 = pick_iter(c->P, c->feature, MODE_ACCUM);
        HIPCHK(c, hipFuncSetAttribute((const void *)k2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    const xpbd::Layout &L = c->L;
    const int nv = c->nv;
    const size_t nptet = L.meta.size();
    int r;
#define UP(dst, vec) do { r = dupload(c, &dst, (vec).data(), (vec).size()); if (r) return r; } while (0)
    for (int b = 0; b < 2; ++b) if ((r = dalloc(c, &c->X[b], nv))) return r;
    if ((r = dalloc(c, &c->PPv, nv))) return r;
    if ((r = dalloc(c, &c->VEL, nv))) return r;
    if ((r = dalloc(c, &c->FRC, nv))) return r;
    if ((r = dalloc(c, &c->PIN, nv))) return r;
    if ((r = dalloc(c, &c->DPW, nv))) return r;
    if ((r = dalloc(c, &c->FLAGS, nv))) return r;
    r = dupload(c, &c->ATT0, (const float4 *)L.att0.data(), (size_t)nv); if (r) return r;
    r = dupload(c, &c->ATT1, (const float4 *)L.att1.data(), (size_t)nv); if (r) return r;
    if (!L.attach_vertex.empty()) {
        UP(c->d_att_dev, L.attach_vertex);
        if ((r = dalloc(c, &c->d_att_tgt, L.attach_vertex.size()))) return r;
    }
    UP(c->d_patch, L.patches);
    UP(c->d_halo, L.halo);
    if (c->flat) {
        r = dupload(c, &c->d_gidx, (const int4 *)L.gidx.data(), nptet); if (r) return r;
        UP(c->d_vslot32, L.vslot32);
        if (c->block) {
            r = dupload(c, &c->d_blk, L.blocks.data(), L.blocks.size()); if (r) return r;
            UP(c->d_bvt, L.bvt);
            UP(c->d_bps, L.bps);
            UP(c->d_brun, L.brun);
            r = dupload(c, &c->d_blidx, (const ushort4 *)L.blidx.data(), nptet); if (r) return r;
            r = dupload(c, &c->d_bcpos, (const ushort4 *)L.bcpos.data(), nptet); if (r) return r;
            UP(c->d_svid, L.svid);
            UP(c->d_soff, L.soff);
            if ((r = dalloc(c, &c->BPART, std::max<size_t>((size_t)L.n_partial, 1)))) return r;
            if ((r = setup_resident(c, t ? t->resident_blocks : 0, maxlds))) return r;
        } else {
            if ((r = dalloc(c, &c->CORR, 4 * std::max<size_t>(nptet, 1)))) return r;
        }
        if ((r = dalloc(c, &c->LAMB, std::max<size_t>(nptet, 1)))) return r;
        HIPCHK(c, hipMemset(c->LAMB, 0, std::max<size_t>(nptet, 1) * sizeof(float4)));
        if ((r = dalloc(c, &c->LAMB0, std::max<size_t>(nptet, 1)))) return r;
        HIPCHK(c, hipMemset(c->LAMB0, 0, std::max<size_t>(nptet, 1) * sizeof(float4)));
        if ((r = dalloc(c, &c->SFEAT, std::max<size_t>(nptet, 1)))) return r;
        c->sfeat_ok = false;
        if ((r = dalloc(c, &c->mW1, (size_t)XPBD_MLP_HIDDEN * XPBD_MLP_IN)) ||
            (r = dalloc(c, &c->mW2, (size_t)XPBD_MLP_HIDDEN * XPBD_MLP_HIDDEN)) ||
            (r = dalloc(c, &c->mb1, XPBD_MLP_HIDDEN)) || (r = dalloc(c, &c->mb2, XPBD_MLP_HIDDEN)) ||
            (r = dalloc(c, &c->mW3, (size_t)XPBD_MLP_OUT * XPBD_MLP_HIDDEN)) || (r = dalloc(c, &c->mb3, XPBD_MLP_OUT)) ||
            (r = dalloc(c, &c->fmean, 24)) || (r = dalloc(c, &c->finv, 24)) || (r = dalloc(c, &c->oscale, 4)))
            return r;
        {
            std::vector<float> zero(24, 0.f), one(24, 1.f);
            HIPCHK(c, hipMemcpy(c->fmean, zero.data(), 24 * sizeof(float), hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(c->finv, one.data(), 24 * sizeof(float), hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(c->oscale, one.data(), 4 * sizeof(float), hipMemcpyHostToDevice));
        }
        c->mlp_ready = false;
    } else {
        r = dupload(c, &c->d_lidx, (const ushort4 *)L.lidx.data(), nptet); if (r) return r;
    }
    UP(c->d_meta, L.meta);
    UP(c->d_Bp, L.Bp);
    UP(c->d_rvol, L.rvol);
    UP(c->d_rsnh, L.rsnh);
    r = dupload(c, &c->FIBS, (const float4 *)L.fib.data(), nptet); if (r) return r;
    r = dupload(c, &c->FIB, (const float4 *)L.fib.data(), nptet); if (r) return r;
    UP(c->d_ptet_tetid, L.ptet_tetid);
    {
        std::vector<DevClass> dcl = dev_classes(c->L.classes, c->sp);
        UP(c->d_cls, dcl);
    }
    UP(c->d_vslot_off, L.vslot_off);
    UP(c->d_vslot, L.vslot);
    sync_host_activation(c);
    UP(c->d_act, c->activation);
    // per-ptet global vertex ids (residual kernel)
    if (c->flat) {
        UP(c->d_pvert, L.gidx);
    } else {
        std::vector<int32_t> pv(4 * nptet);
        for (const auto &P_ : L.patches)
            for (int j = 0; j < P_.ntet; ++j) {
                const size_t g = (size_t)P_.tbeg + j;
                for (int k = 0; k < 4; ++k) {
                    const int li = L.lidx[4 * g + k];
                    pv[4 * g + k] = li < P_.nown ? P_.vbeg + li : L.halo[P_.hbeg + li - P_.nown];
                }
            }
        UP(c->d_pvert, pv);
    }
    {
        std::vector<int4> tt(c->nt);
        for (int i = 0; i < c->nt; ++i)
            tt[i] = make_int4(L.iperm[c->tets[4 * i]], L.iperm[c->tets[4 * i + 1]], L.iperm[c->tets[4 * i + 2]],
                              L.iperm[c->tets[4 * i + 3]]);
        UP(c->d_tets, tt);
        UP(c->d_B, c->rest_matrix);
    }
    if (c->nbones > 0) {
        UP(c->d_grids, c->sdf_grids);
        if (c->goff.empty()) { c->goff = {0, c->nbones}; c->vgroup.assign(nv, 0); }
        std::vector<int32_t> vg(nv);
        for (int n = 0; n < nv; ++n) vg[n] = c->vgroup[L.perm[n]];
        UP(c->d_vgroup, vg);
        UP(c->d_goff, c->goff);
        if ((r = dalloc(c, &c->d_bone, 16 * (size_t)c->nbones))) return r;
    }
    if ((r = dalloc(c, &c->d_cells, 2 * (size_t)nv))) return r;
    if ((r = dalloc(c, &c->S_X, nv)) || (r = dalloc(c, &c->S_PP, nv)) || (r = dalloc(c, &c->S_VEL, nv))) return r;
    if ((r = dalloc(c, &c->d_count, 1))) return r;
    c->n_partial = 512;
    if ((r = dalloc(c, &c->d_partial, 8 * (size_t)c->n_partial))) return r;
#undef UP
    // vertex state from rest pose
    {
        std::vector<float4> x(nv), z(nv, make_float4(0, 0, 0, 0));
        for (int n = 0; n < nv; ++n) {
            const int o = L.perm[n];
            x[n] = make_float4(c->pos0[3 * o], c->pos0[3 * o + 1], c->pos0[3 * o + 2], 0.f);
        }
        HIPCHK(c, hipMemcpy(c->X[0], x.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->PPv, x.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->VEL, z.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->FRC, z.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->DPW, z.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
    }
    c->ready = true;
    if (!c->dcd_tris.empty() && (r = alloc_dcd(c))) return r;
    if (c->bm.set && (r = bm_build(c))) return r;
    if ((r = upload_vertex_static(c))) return r;
    if ((r = upload_pins(c))) return r;
    if (c->has_force && (r = xpbd_set_force(c, c->force.data()))) return r;
    if ((r = xpbd_set_bone_transforms(c, c->nbones, c->bone_xf.data()))) return r;
    if ((r = upload_act_groups(c))) return r;
    if (c->feature & F_ANISO)
        if ((r = xpbd_set_activation(c, c->activation.data()))) return r;
    return XPBD_OK;
}

int32_t xpbd_get_stats(xpbd_ctx *c, xpbd_stats *o) {
    int r = check_ready(c);
    if (r) return r;
    if (!o) return fail(c, XPBD_EINVAL, "null stats");
    o->nv = c->nv;
    o->nt = c->nt;
    o->n_elems = c->L.n_elems;
    o->n_patches = (int)(c->block ? c->L.blocks.size() : c->L.patches.size());
    o->n_patch_tets = (int)c->L.meta.size();
    o->max_table = c->block ? c->L.max_nloc : c->L.max_table;
    o->max_patch_tets = c->L.max_ptets;
    o->feature_mask = c->feature;
    o->bytes_per_iter_alg = c->L.bytes_alg;
    o->bytes_per_iter_impl = c->L.bytes_impl;
    o->resident_blocks = c->res_R;
    o->resident_parts = c->res_R ? (int)c->parts.parts.size() : 0;
    o->max_part_verts = c->res_R ? c->parts.max_nloc : 0;
    o->max_part_neighbours = c->res_R ? c->parts.max_nnb : 0;
    o->cross_vertex_entries = c->res_R ? c->parts.n_cross : 0;
    o->resident_lds = (int)c->res_lds;
    return XPBD_OK;
}

int32_t xpbd_set_state(xpbd_ctx *c, const float *pos, const float *pprev, const float *vel) {
    int r = check_ready(c);
    if (r) return r;
    const int nv = c->nv;
    std::vector<float4> x(nv), pp(nv), v(nv);
    HIPCHK(c, hipMemcpyAsync(x.data(), c->X[0], nv * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(pp.data(), c->PPv, nv * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(v.data(), c->VEL, nv * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((r = check_resident(c))) return r;
    for (int n = 0; n < nv; ++n) {
        const int o = c->L.perm[n];
        if (pos) { x[n].x = pos[3 * o]; x[n].y = pos[3 * o + 1]; x[n].z = pos[3 * o + 2]; }
        if (pprev) { pp[n].x = pprev[3 * o]; pp[n].y = pprev[3 * o + 1]; pp[n].z = pprev[3 * o + 2]; }
        if (vel) { v[n].x = vel[3 * o]; v[n].y = vel[3 * o + 1]; v[n].z = vel[3 * o + 2]; }
    }
    HIPCHK(c, hipMemcpy(c->X[0], x.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->PPv, pp.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->VEL, v.data(), nv * sizeof(float4), hipMemcpyHostToDevice));
    return XPBD_OK;
}

int32_t xpbd_get_state(xpbd_ctx *c, float *pos, float *pprev, float *vel) {
    int r = check_ready(c);
    if (r) return r;
    const int nv = c->nv;
    std::vector<float4> x(nv), pp(nv), v(nv);
    HIPCHK(c, hipMemcpyAsync(x.data(), c->X[0], nv * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(pp.data(), c->PPv, nv * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(v.data(), c->VEL, nv * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((r = check_resident(c))) return r;
    for (int n = 0; n < nv; ++n) {
        const int o = c->L.perm[n];
        if (pos) { pos[3 * o] = x[n].x; pos[3 * o + 1] = x[n].y; pos[3 * o + 2] = x[n].z; }
        if (pprev) { pprev[3 * o] = pp[n].x; pprev[3 * o + 1] = pp[n].y; pprev[3 * o + 2] = pp[n].z; }
        if (vel) { vel[3 * o] = v[n].x; vel[3 * o + 1] = v[n].y; vel[3 * o + 2] = v[n].z; }
    }
    return XPBD_OK;
}

int32_t xpbd_integrate(xpbd_ctx *c) {
    int r = check_ready(c);
    return r ? r : launch_integrate(c);
}

int32_t xpbd_clear(xpbd_ctx *c) {
    int r = check_ready(c);
    if (r) return r;
    hipLaunchKernelGGL(k_zero4, vgrid(c->nv), dim3(256), 0, c->stream, (int64_t)c->nv, c->DPW);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

int32_t xpbd_solve_constraints(xpbd_ctx *c) {
    int r = check_ready(c);
    return r ? r : launch_iter(c, c->X[0], c->X[1], MODE_ACCUM);
}

int32_t xpbd_apply_dP(xpbd_ctx *c) {
    int r = check_ready(c);
    if (r) return r;
    hipLaunchKernelGGL(k_apply_dpw, vgrid(c->nv), dim3(256), 0, c->stream, c->nv, c->X[0],
                       (const float4 *)c->DPW, (const float4 *)c->PIN, 0);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

int32_t xpbd_apply_pins(xpbd_ctx *c) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->pins_on) return XPBD_OK;
    hipLaunchKernelGGL(k_pins, vgrid(c->nv), dim3(256), 0, c->stream, c->nv, c->X[0], (const float4 *)c->PIN);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

int32_t xpbd_iterate(xpbd_ctx *c, int32_t n) {
    int r = check_ready(c);
    if (r) return r;
    if (n < 0) return fail(c, XPBD_EINVAL, "negative iteration count");
    c->eff_mode = c->lam_mode == 4 ? (c->warm_net ? 2 : 3) : c->lam_mode;
    if (c->res_R > 0 && c->eff_mode == 0) {
        if ((r = launch_resident(c, c->X[0], c->X[n & 1], n))) return r;
    } else {
        for (int it = 0; it < n; ++it)
            if ((r = launch_iter(c, c->X[it & 1], c->X[(it + 1) & 1], MODE_FUSED, lam_for_iteration(c)))) return r;
    }
    if (n & 1) {
        HIPCHK(c, hipMemcpyAsync(c->X[0], c->X[1], c->nv * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
    }
    return XPBD_OK;
}

int32_t xpbd_collide(xpbd_ctx *c) {
    int r = check_ready(c);
    return r ? r : launch_post(c, c->X[0], 1, 0);
}

int32_t xpbd_update_velocities(xpbd_ctx *c) {
    int r = check_ready(c);
    return r ? r : launch_post(c, c->X[0], 0, 1);
}

int32_t xpbd_substep(xpbd_ctx *c, int32_t iters) {
    int r = check_ready(c);
    if (r) return r;
    if (iters < 0) return fail(c, XPBD_EINVAL, "negative iteration count");
    r = enqueue_substep(c, iters, c->dcd_fresh, c->lam_mode == 4 ? 1 : 0);
    c->dcd_fresh = false;
    return r;
}

int32_t xpbd_step_frame(xpbd_ctx *c, int32_t nsub, int32_t iters, int32_t use_graph) {
    int r = check_ready(c);
    if (r) return r;
    if (nsub < 0 || iters < 0) return fail(c, XPBD_EINVAL, "negative count");
    if (c->lam_mode && (r = ensure_static_features(c))) return r;   // before any capture
    // the DCD broadphase is rebuilt once per frame (humanV2.py:853-865)
    c->dcd_fresh = false;
    if (!use_graph || c->timing) {
        for (int s = 0; s < nsub; ++s)
            if ((r = enqueue_substep(c, iters, s == 0, s, s > 0, s + 1 < nsub))) return r;
        return XPBD_OK;
    }
    xpbd_ctx::GraphSlot *hit = nullptr, *victim = &c->gslot[0];
    for (auto &g : c->gslot) {
        if (g.gexec && g.nsub == nsub && g.iters == iters && g.lam_mode == c->lam_mode && g.ref_iters == c->ref_iters &&
            g.ref_every == c->ref_every && g.warm_net == c->warm_net)
            hit = &g;
        if (g.used < victim->used) victim = &g;
    }
    if (!hit) {
        drop_graph_slot(*victim);
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < nsub && r == XPBD_OK; ++s) r = enqueue_substep(c, iters, s == 0, s, s > 0, s + 1 < nsub);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (r) { if (g) (void)hipGraphDestroy(g); return r; }
        if (e != hipSuccess) return fail(c, XPBD_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        victim->graph = g;
        HIPCHK(c, hipGraphInstantiate(&victim->gexec, g, nullptr, nullptr, 0));
        victim->nsub = nsub;
        victim->iters = iters;
        victim->lam_mode = c->lam_mode;
        victim->ref_iters = c->ref_iters;
        victim->ref_every = c->ref_every;
        victim->warm_net = c->warm_net;
        hit = victim;
    }
    hit->used = ++c->gclock;
    HIPCHK(c, hipGraphLaunch(hit->gexec, c->stream));
    return XPBD_OK;
}

int32_t xpbd_synchronize(xpbd_ctx *c) {
    if (!c) return XPBD_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return check_resident(c);
}

int32_t xpbd_residuals(xpbd_ctx *c, double *out8) {
    int r = check_ready(c);
    if (r) return r;
    if (!out8) return fail(c, XPBD_EINVAL, "null output");
    IterArgs a = make_args(c, c->DPW);
    hipLaunchKernelGGL(k_residual<256>, dim3(c->n_partial), dim3(256), 0, c->stream, a, (const float4 *)c->X[0],
                       (const int32_t *)c->d_pvert, c->d_partial);
    HIPCHK(c, hipGetLastError());
    std::vector<double> h(8 * (size_t)c->n_partial);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_partial, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 8; ++k) out8[k] = 0.0;
    for (int b = 0; b < c->n_partial; ++b)
        for (int k = 0; k < 8; ++k) out8[k] += h[8 * (size_t)b + k];
    return XPBD_OK;
}

int32_t xpbd_count_inverted(xpbd_ctx *c, int32_t *out) {
    int r = check_ready(c);
    if (r) return r;
    if (!out) return fail(c, XPBD_EINVAL, "null output");
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, c->stream, 1, c->d_count);
    HIPCHK(c, hipGetLastError());
    if (c->nt > 0) {
        hipLaunchKernelGGL(k_count_inverted, vgrid(c->nt), dim3(256), 0, c->stream, c->nt, (const int4 *)c->d_tets,
                           (const float *)c->d_B, (const float4 *)c->X[0], c->d_count);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipMemcpyAsync(out, c->d_count, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return XPBD_OK;
}

int32_t xpbd_get_sdf_cells(xpbd_ctx *c, int32_t *cells) {
    int r = check_ready(c);
    if (r) return r;
    if (!cells) return fail(c, XPBD_EINVAL, "null output");
    std::vector<int32_t> h(2 * (size_t)c->nv);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_cells, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int n = 0; n < c->nv; ++n) {
        const int o = c->L.perm[n];
        cells[2 * o] = h[2 * n];
        cells[2 * o + 1] = h[2 * n + 1];
    }
    return XPBD_OK;
}


int32_t xpbd_set_dcd(xpbd_ctx *c, int32_t ntris, const int32_t *tris, const int32_t *group,
                     const xpbd_dcd_params *p) {
    if (!c || c->nv == 0 || ntris < 0 || (ntris > 0 && (!tris || !p))) return fail(c, XPBD_EINVAL, "bad dcd");
    if (ntris >= (1 << DCD_LVL_SHIFT)) return fail(c, XPBD_EINVAL, "too many dcd triangles (limit 2^27)");
    if (p && p->max_pairs >= ((int64_t)1 << 31)) return fail(c, XPBD_EINVAL, "dcd max_pairs must be below 2^31");
    for (int64_t i = 0; i < 3 * (int64_t)ntris; ++i)
        if (tris[i] < 0 || tris[i] >= c->nv) return fail(c, XPBD_EINVAL, "dcd triangle vertex out of range");
    if (group)
        for (int i = 0; i < ntris; ++i)
            if (group[i] != 0 && group[i] != 1) return fail(c, XPBD_EINVAL, "dcd group must be 0 or 1");
    c->dcd_tris.assign(tris, tris + 3 * (size_t)ntris);
    if (group) c->dcd_group.assign(group, group + ntris);
    else c->dcd_group.clear();
    if (p) {
        c->dcd_p = *p;
        c->dcd_iters = std::max(1, (int)p->iters);   // K = max(1, iters) passes (humanV2.py _make_self_collider)
        c->dcd_rebuild_substeps = p->rebuild_substeps;
    }
    if (c->ready) free_device(c);
    return XPBD_OK;
}

int32_t xpbd_dcd_solve(xpbd_ctx *c, int32_t rebuild) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->dcd_on) return fail(c, XPBD_EINVAL, "no DCD surface set");
    return launch_dcd(c, c->X[0], rebuild != 0);
}

int32_t xpbd_dcd_pairs(xpbd_ctx *c, int32_t *out, int64_t cap, int64_t *n_total) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->dcd_on) return fail(c, XPBD_EINVAL, "no DCD surface set");
    DcdDev &d = c->dcd;
    unsigned long long tot = 0;
    HIPCHK(c, hipMemcpyAsync(&tot, d.npairs, sizeof(tot), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n_total) *n_total = (int64_t)tot;
    const int64_t have = std::min<int64_t>((int64_t)tot, d.max_pairs);
    const int64_t n = std::min<int64_t>(have, cap);
    if (!out || n <= 0) return XPBD_OK;
    // segments by first index, each sorted by the second
    const int m = d.ntris;
    auto g1 = [](int64_t k) { return dim3((unsigned)((k + 255) / 256)); };
    hipLaunchKernelGGL(k_zero_u32, g1(m + 1), dim3(256), 0, c->stream, d.pair_cnt, (int64_t)m + 1);
    hipLaunchKernelGGL(k_zero_u32, g1(m), dim3(256), 0, c->stream, d.pair_fill, (int64_t)m);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c->stream, d.overflow + 2, (int64_t)1);
    hipLaunchKernelGGL(k_dcd_segcount, g1(have), dim3(256), 0, c->stream, d, have);
    HIPCHK(c, hipGetLastError());
    if ((r = launch_scan(c, d.pair_cnt, d.pair_off, m, d.pair_off + m))) return r;
    hipLaunchKernelGGL(k_dcd_segscatter, g1(have), dim3(256), 0, c->stream, d, have);
    hipLaunchKernelGGL(k_dcd_sortseg, g1((int64_t)DCD_LANES * m), dim3(256), 0, c->stream, d);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out, d.pairs_seg, (size_t)n * sizeof(int2), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return XPBD_OK;
}

int32_t xpbd_dcd_overflow(xpbd_ctx *c, int64_t *left_out, int64_t *truncated) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->dcd_on) return fail(c, XPBD_EINVAL, "no DCD surface set");
    uint32_t o[3] = {0, 0, 0};
    unsigned long long tot = 0;
    HIPCHK(c, hipMemcpyAsync(o, c->dcd.overflow, sizeof(o), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&tot, c->dcd.npairs, sizeof(tot), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t dropped = std::max<int64_t>((int64_t)tot - c->dcd.max_pairs, 0);
    if (left_out) *left_out = o[0];
    // stopped searches + pairs beyond the capacity + segments left unsorted
    if (truncated) *truncated = (int64_t)o[1] + dropped + o[2];
    return XPBD_OK;
}

int32_t xpbd_dcd_stats(xpbd_ctx *c, int64_t *n_pairs, int64_t *n_contacts, float *cell) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->dcd_on) return fail(c, XPBD_EINVAL, "no DCD surface set");
    unsigned long long tot = 0;
    uint32_t nc = 0;
    float sc[4];
    HIPCHK(c, hipMemcpyAsync(&tot, c->dcd.npairs, sizeof(tot), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&nc, c->dcd.ncontact + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(sc, c->dcd.scal, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n_pairs) *n_pairs = (int64_t)tot;
    if (n_contacts) *n_contacts = nc;
    if (cell) *cell = sc[0];
    return XPBD_OK;
}


static uint16_t f2bf(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u = u + (((u >> 16) & 1u) + 0x7FFFu);
    return (uint16_t)(u >> 16);
}

int32_t xpbd_set_mlp(xpbd_ctx *c, const float *W1, const float *b1, const float *W2, const float *b2,
                     const float *W3, const float *b3, const float *feat_mean, const float *feat_invstd,
                     const float *out_scale) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->flat) return fail(c, XPBD_EUNSUPPORTED, "the learned warm start runs on the flat strategy");
    if (!W1 || !b1 || !W2 || !b2 || !W3 || !b3 || !feat_mean || !feat_invstd || !out_scale)
        return fail(c, XPBD_EINVAL, "null MLP parameter");
    const int NH = XPBD_MLP_HIDDEN, NI = XPBD_MLP_IN, NO = XPBD_MLP_OUT;
    std::vector<uint16_t> w1(NH * NI), w2(NH * NH);
    for (int i = 0; i < NH * NI; ++i) w1[i] = f2bf(W1[i]);
    for (int i = 0; i < NH * NH; ++i) w2[i] = f2bf(W2[i]);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->mW1, w1.data(), w1.size() * 2, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->mW2, w2.data(), w2.size() * 2, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->mb1, b1, NH * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->mb2, b2, NH * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->mW3, W3, NO * NH * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->mb3, b3, NO * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->fmean, feat_mean, 24 * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->finv, feat_invstd, 24 * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->oscale, out_scale, 4 * 4, hipMemcpyHostToDevice));
    c->mlp_ready = true;
    return XPBD_OK;
}

namespace {

// closest point of triangle a b c to p in double precision (Ericson, Real-Time
// Collision Detection, 5.1.5), for the one-time fascia bind
struct D3 { double x, y, z; };
D3 d3sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
double d3dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 d3axpy(D3 a, double s, D3 d) { return D3{a.x + s * d.x, a.y + s * d.y, a.z + s * d.z}; }
D3 d3cross(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

D3 tri_closest_d(D3 p, D3 a, D3 b, D3 c) {
    const D3 ab = d3sub(b, a), ac = d3sub(c, a), ap = d3sub(p, a);
    const double d1 = d3dot(ab, ap), d2 = d3dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return a;
    const D3 bp = d3sub(p, b);
    const double d3 = d3dot(ab, bp), d4 = d3dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return b;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return d3axpy(a, d1 / (d1 - d3), ab);
    const D3 cq = d3sub(p, c);
    const double d5 = d3dot(ab, cq), d6 = d3dot(ac, cq);
    if (d6 >= 0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return d3axpy(a, d2 / (d2 - d6), ac);
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) return d3axpy(b, (d4 - d3) / ((d4 - d3) + (d5 - d6)), d3sub(c, b));
    const double den = va + vb + vc;
    if (!(den > 0)) {            // zero-area triangle: the nearest point of its three edges
        auto seg = [&](D3 s0, D3 s1) {
            const D3 e = d3sub(s1, s0);
            const double l2 = d3dot(e, e);
            double t = l2 > 0 ? d3dot(d3sub(p, s0), e) / l2 : 0.0;
            t = t < 0 ? 0 : (t > 1 ? 1 : t);
            return d3axpy(s0, t, e);
        };
        const D3 q0 = seg(a, b), q1 = seg(b, c), q2 = seg(c, a);
        const double e0 = d3dot(d3sub(p, q0), d3sub(p, q0)), e1 = d3dot(d3sub(p, q1), d3sub(p, q1)),
                     e2 = d3dot(d3sub(p, q2), d3sub(p, q2));
        return e0 <= e1 && e0 <= e2 ? q0 : (e1 <= e2 ? q1 : q2);
    }
    return d3axpy(d3axpy(a, vb / den, ab), vc / den, ac);
}

// squeeze: follow positions, glue pairs and anchors to the device (library orders)
int set_fascia_squeeze(xpbd_ctx *c, xpbd_ctx *f, const xpbd_fascia_desc *d) {
    const int nf = f->nv;
    if (!d->follow0) return fail(c, XPBD_EINVAL, "squeeze needs follow0");
    if (d->nglue < 0 || d->nanchor < 0 || (d->nglue > 0 && (!d->glue_fascia || !d->glue_muscle ||
                                                             !d->glue_restlen || !d->glue_compliance)) ||
        (d->nanchor > 0 && (!d->anchor || !d->anchor_offset)))
        return fail(c, XPBD_EINVAL, "bad glue or anchor arrays");
    std::vector<char> seen(c->nv, 0);
    std::vector<int4> glue(d->nglue);
    for (int k = 0; k < d->nglue; ++k) {
        const int a = d->glue_fascia[k], m = d->glue_muscle[k];
        if (a < 0 || a >= nf || m < 0 || m >= c->nv) return fail(c, XPBD_EINVAL, "glue vertex out of range");
        if (seen[m]) return fail(c, XPBD_EINVAL, "a muscle vertex is in two glue pairs");
        seen[m] = 1;
        float rl = d->glue_restlen[k], cm = d->glue_compliance[k];
        int rli, cmi;
        std::memcpy(&rli, &rl, 4);
        std::memcpy(&cmi, &cm, 4);
        glue[k] = make_int4(f->L.iperm[a], c->L.iperm[m], rli, cmi);
    }
    std::vector<int32_t> anc(d->nanchor);
    std::vector<float4> aoff(d->nanchor);
    for (int k = 0; k < d->nanchor; ++k) {
        if (d->anchor[k] < 0 || d->anchor[k] >= nf) return fail(c, XPBD_EINVAL, "anchor vertex out of range");
        anc[k] = f->L.iperm[d->anchor[k]];
        aoff[k] = make_float4(d->anchor_offset[3 * k], d->anchor_offset[3 * k + 1], d->anchor_offset[3 * k + 2], 0.f);
    }
    std::vector<float4> ff(nf);
    for (int k = 0; k < nf; ++k) {
        const int o = f->L.perm[k];
        ff[k] = make_float4(d->follow0[3 * o], d->follow0[3 * o + 1], d->follow0[3 * o + 2], 0.f);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c, c->d_ff);
    dfree(c, c->d_ff_prev);
    dfree(c, c->d_fglue);
    dfree(c, c->d_fanc);
    dfree(c, c->d_fanc_off);
    int r;
    if ((r = dupload(c, &c->d_ff, ff.data(), ff.size())) || (r = dupload(c, &c->d_ff_prev, ff.data(), ff.size())) ||
        (r = dupload(c, &c->d_fglue, glue.data(), glue.size())) || (r = dupload(c, &c->d_fanc, anc.data(), anc.size())) ||
        (r = dupload(c, &c->d_fanc_off, aoff.data(), aoff.size())))
        return r;
    c->fascia = f;
    c->fascia_gen = f->gen;
    c->fascia_iters = d->iters;
    c->fascia_mode = 1;
    c->fascia_nbound = 0;
    c->fascia_nglue = d->nglue;
    c->fascia_nanchor = d->nanchor;
    c->fascia_omega = d->omega;
    c->fb_face.assign(nf, -1);
    c->fb_ok.assign(nf, 0);
    c->fb_w.assign(3 * (size_t)nf, 0.f);
    c->fb_gap.assign(nf, 0.f);
    f->stream = c->stream;
    invalidate_graph(f);
    invalidate_graph(c);
    return XPBD_OK;
}

// closest point of each query point on a triangle mesh among the faces nearer than max_dist,
// found with the bone-mesh tree on the host, then exact double distances over every face within
// the float result plus slack; at near-ties (faces within 1e-6 relative of the nearest distance)
// the lowest triangle index wins, so the choice does not depend on the tree.  Per point: the face
// (-1: none within max_dist), the barycentric weights of the closest point on (tri[0], tri[1],
// tri[2]) and the signed distance along the face normal (B - A) x (C - A).
std::string mesh_bind_host(int nv, const float *pos, int ntri, const int32_t *tris, int npts, const float *pts,
                           float max_dist, int32_t *face_out, float *w_out, float *gap_out) {
    if (nv < 3 || ntri < 1 || !pos || !tris) return "no mesh";
    if (!(max_dist > 0.0f)) return "bad max_dist";
    for (int64_t i = 0; i < 3 * (int64_t)ntri; ++i)
        if (tris[i] < 0 || tris[i] >= nv) return "triangle vertex out of range";
    BmTables T;
    const int32_t off[2] = {0, ntri};
    std::string e = bm_build_tables(nv, pos, ntri, tris, 1, off, T);
    if (!e.empty()) return e;
    std::vector<bm_f4> bp(nv + 1), soup(3 * (size_t)ntri), fn(ntri), vn(nv), node(2 * (size_t)T.nnodes);
    for (int i = 0; i < nv; ++i) bp[i] = bm_f4{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.f};
    const int32_t goff[2] = {0, 1}, vg0 = 0;
    BoneMeshDev m{};
    m.nbones = 1; m.ntris = ntri; m.nbv = nv;
    m.bone = T.bone.data(); m.tv = T.tv.data(); m.adj = T.adj.data(); m.sub = T.sub.data();
    m.vt_off = T.vt_off.data(); m.vt = T.vt.data(); m.pts = bp.data(); m.soup = soup.data(); m.fn = fn.data();
    m.vn = vn.data(); m.node = node.data(); m.vgroup = &vg0; m.goff = goff;
    m.max_dist = max_dist;
    bm_refit_host(m);
    auto P = [&](int v) { return D3{pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]}; };
    const double md2 = (double)max_dist * max_dist;
    std::vector<int> stack, depth;
    std::vector<std::pair<double, int>> cand;
    for (int i = 0; i < npts; ++i) {
        face_out[i] = -1;
        w_out[3 * i] = w_out[3 * i + 1] = w_out[3 * i + 2] = 0.f;
        gap_out[i] = 0.f;
        const D3 p{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        BmHit h;
        float sg;
        if (!bm_query(m, 0, bm_mk((float)p.x, (float)p.y, (float)p.z), h, sg)) continue;
        // every face whose box is within the float result plus slack, then exact distances
        const double r2 = std::min(md2, (double)h.d2 * (1.0 + 1e-3) + 1e-12);
        const bm_i4 bn = T.bone[0];
        const int D = bm_depth(bn.y);
        double best = INFINITY;
        cand.clear();
        stack.assign(1, 0);
        depth.assign(1, 0);
        while (!stack.empty()) {
            const int k = stack.back(), dk = depth.back();
            stack.pop_back();
            depth.pop_back();
            if ((double)bm_box_d2(node.data() + 2 * (size_t)bn.x, k, bm_mk((float)p.x, (float)p.y, (float)p.z)) >
                r2 * (1.0 + 1e-3) + 1e-12)
                continue;
            if (dk == D) {
                const int leaf = k - (bn.y - 1);
                for (int j = leaf * BM_LEAF; j < std::min(leaf * BM_LEAF + BM_LEAF, bn.w); ++j) {
                    const int ft = T.tv[bn.z + j].w;
                    const int32_t *tt = tris + 3 * (size_t)ft;
                    const D3 q = tri_closest_d(p, P(tt[0]), P(tt[1]), P(tt[2]));
                    const D3 dd = d3sub(p, q);
                    const double e2 = d3dot(dd, dd);
                    if (e2 < md2) {
                        cand.push_back({e2, ft});
                        best = std::min(best, e2);
                    }
                }
            } else {
                stack.push_back(2 * k + 1); depth.push_back(dk + 1);
                stack.push_back(2 * k + 2); depth.push_back(dk + 1);
            }
        }
        if (cand.empty()) continue;
        int ft = -1;
        for (auto &q : cand)
            if (q.first <= best * (1.0 + 2e-6) + 1e-14 && (ft < 0 || q.second < ft)) ft = q.second;
        const int32_t *tt = tris + 3 * (size_t)ft;
        const D3 A = P(tt[0]), B = P(tt[1]), C = P(tt[2]);
        const D3 q = tri_closest_d(p, A, B, C);
        // barycentric weights of q (Cramer's rule), clamped onto the triangle
        const D3 v0 = d3sub(B, A), v1 = d3sub(C, A), v2 = d3sub(q, A);
        const double d00 = d3dot(v0, v0), d01 = d3dot(v0, v1), d11 = d3dot(v1, v1), d20 = d3dot(v2, v0),
                     d21 = d3dot(v2, v1), den = d00 * d11 - d01 * d01;
        double wb = 0, wc = 0;
        if (den > 0) { wb = (d11 * d20 - d01 * d21) / den; wc = (d00 * d21 - d01 * d20) / den; }
        wb = std::max(0.0, wb);
        wc = std::max(0.0, wc);
        if (wb + wc > 1.0) { const double s = wb + wc; wb /= s; wc /= s; }
        const double wa = 1.0 - wb - wc;
        D3 nrm = d3cross(v0, v1);
        const double nl = std::sqrt(d3dot(nrm, nrm));
        nrm = nl > 0 ? D3{nrm.x / nl, nrm.y / nl, nrm.z / nl} : D3{0, 0, 0};
        face_out[i] = ft;
        w_out[3 * i] = (float)wa; w_out[3 * i + 1] = (float)wb; w_out[3 * i + 2] = (float)wc;
        gap_out[i] = (float)d3dot(d3sub(p, q), nrm);
    }
    return std::string();
}

}  // namespace

int32_t xpbd_mesh_bind(int32_t nv, const float *verts, int32_t ntris, const int32_t *tris, int32_t npts,
                       const float *pts, float max_dist, int32_t *face, float *bary, float *gap) {
    if (nv < 3 || ntris < 1 || npts < 0 || !verts || !tris || (npts > 0 && (!pts || !face || !bary || !gap)))
        return XPBD_EINVAL;
    const std::string e = mesh_bind_host(nv, verts, ntris, tris, npts, pts, max_dist, face, bary, gap);
    return e.empty() ? XPBD_OK : XPBD_EINVAL;
}

int32_t xpbd_set_fascia(xpbd_ctx *c, xpbd_ctx *f, const xpbd_fascia_desc *d) {
    if (!c) return XPBD_EINVAL;
    if (!f || !d) {
        c->fascia = nullptr;
        c->fascia_nbound = 0;
        invalidate_graph(c);
        return XPBD_OK;
    }
    c->fascia = nullptr;             // until the new link is complete
    int r = check_ready(c);
    if (r) return r;
    if (f == c) return fail(c, XPBD_EINVAL, "a context cannot be its own fascia");
    if (!f->ready) return fail(c, XPBD_ENOTREADY, "finalize the fascia context first");
    if (f->device != c->device) return fail(c, XPBD_EINVAL, "the fascia context is on another device");
    if (d->iters < 0) return fail(c, XPBD_EINVAL, "negative fascia iterations");
    if (d->mode == 1) return set_fascia_squeeze(c, f, d);
    if (d->mode != 0) return fail(c, XPBD_EINVAL, "fascia mode must be 0 (ride) or 1 (squeeze)");
    if (d->nsurf < 1 || !d->surf_tris) return fail(c, XPBD_EINVAL, "no surface triangles");
    if (!(d->bind_max_dist > 0.0f)) return fail(c, XPBD_EINVAL, "bad bind_max_dist");
    for (int64_t i = 0; i < 3 * (int64_t)d->nsurf; ++i)
        if (d->surf_tris[i] < 0 || d->surf_tris[i] >= c->nv)
            return fail(c, XPBD_EINVAL, "surface triangle vertex out of range");
    // the bind (fascia_bind_kernel, fascia.py:135-157): the closest point of the rest surface to
    // each fascia rest vertex among the faces nearer than bind_max_dist
    const int ns = d->nsurf, nfv = f->nv;
    c->fb_face.assign(nfv, -1);
    c->fb_ok.assign(nfv, 0);
    c->fb_w.assign(3 * (size_t)nfv, 0.f);
    c->fb_gap.assign(nfv, 0.f);
    {
        std::string e = mesh_bind_host(c->nv, c->pos0.data(), ns, d->surf_tris, nfv, f->pos0.data(),
                                       d->bind_max_dist, c->fb_face.data(), c->fb_w.data(), c->fb_gap.data());
        if (!e.empty()) return fail(c, XPBD_EINVAL, e);
    }
    auto P = [&](int v) { return D3{c->pos0[3 * v], c->pos0[3 * v + 1], c->pos0[3 * v + 2]}; };
    std::vector<int32_t> bvid;
    std::vector<int4> btri;
    std::vector<float4> bw;
    std::vector<float> tgt;
    for (int i = 0; i < nfv; ++i) {
        const int ft = c->fb_face[i];
        if (ft < 0) continue;
        const int32_t *tt = d->surf_tris + 3 * (size_t)ft;
        const double wa = c->fb_w[3 * i], wb = c->fb_w[3 * i + 1], wc = c->fb_w[3 * i + 2];
        const double gap = c->fb_gap[i];
        const D3 A = P(tt[0]), B = P(tt[1]), C = P(tt[2]);
        const D3 q{wa * A.x + wb * B.x + wc * C.x, wa * A.y + wb * B.y + wc * C.y, wa * A.z + wb * B.z + wc * C.z};
        D3 nrm = d3cross(d3sub(B, A), d3sub(C, A));
        const double nl = std::sqrt(d3dot(nrm, nrm));
        nrm = nl > 0 ? D3{nrm.x / nl, nrm.y / nl, nrm.z / nl} : D3{0, 0, 0};
        c->fb_ok[i] = 1;
        bvid.push_back(i);
        btri.push_back(make_int4(c->L.iperm[tt[0]], c->L.iperm[tt[1]], c->L.iperm[tt[2]], 0));
        bw.push_back(make_float4((float)wa, (float)wb, (float)wc, (float)gap));
        const D3 g = d3axpy(q, gap, nrm);
        tgt.push_back((float)g.x); tgt.push_back((float)g.y); tgt.push_back((float)g.z);
    }
    // the bound vertices become the fascia's active pins (their targets at rest: the bind)
    const int nb = (int)bvid.size();
    std::vector<int32_t> act(nb, 1), fvid(nb);
    if ((r = xpbd_set_pins(f, nb, bvid.data(), act.data())) || (r = xpbd_set_pin_targets(f, nb, tgt.data())))
        return fail(c, r, std::string("fascia pins: ") + f->err);
    for (int k = 0; k < nb; ++k) fvid[k] = f->L.iperm[bvid[k]];
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c, c->d_fb_vid);
    dfree(c, c->d_fb_tri);
    dfree(c, c->d_fb_w);
    if ((r = dupload(c, &c->d_fb_vid, fvid.data(), (size_t)nb)) ||
        (r = dupload(c, &c->d_fb_tri, btri.data(), (size_t)nb)) || (r = dupload(c, &c->d_fb_w, bw.data(), (size_t)nb)))
        return r;
    c->fascia = f;
    c->fascia_gen = f->gen;
    c->fascia_iters = d->iters;
    c->fascia_nbound = nb;
    c->fascia_mode = 0;
    f->stream = c->stream;
    invalidate_graph(f);
    invalidate_graph(c);
    return XPBD_OK;
}

int32_t xpbd_set_fascia_follow(xpbd_ctx *c, const float *xyz) {
    if (!c || !c->fascia || c->fascia_mode != 1) return fail(c, XPBD_EINVAL, "no fascia in squeeze mode");
    if (!xyz) return fail(c, XPBD_EINVAL, "null follow positions");
    xpbd_ctx *f = c->fascia;
    if (f->gen != c->fascia_gen || !f->ready)
        return fail(c, XPBD_ENOTREADY, "the fascia context was finalized again; call xpbd_set_fascia again");
    const int n = f->nv;
    int slot;
    void *h;
    int r = stage_acquire(c, c->st_ff, (size_t)n * sizeof(float4), slot, h);
    if (r) return r;
    float4 *q = (float4 *)h;
    for (int k = 0; k < n; ++k) {
        const int o = f->L.perm[k];
        q[k] = make_float4(xyz[3 * o], xyz[3 * o + 1], xyz[3 * o + 2], 0.f);
    }
    if ((r = stage_submit(c, c->st_ff, slot, c->d_ff, (size_t)n * sizeof(float4)))) return r;
    hipLaunchKernelGGL(k_fascia_advect, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                       (const float4 *)c->d_ff, c->d_ff_prev, f->X[0]);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

int32_t xpbd_fascia_substep(xpbd_ctx *c) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->fascia) return fail(c, XPBD_EINVAL, "no fascia");
    return launch_fascia(c, c->X[0]);
}

int32_t xpbd_get_fascia_bind(xpbd_ctx *c, int32_t *face, float *w, float *gap, int32_t *nbound) {
    if (!c || !c->fascia) return fail(c, XPBD_EINVAL, "no fascia");
    if (face) std::copy(c->fb_face.begin(), c->fb_face.end(), face);
    if (w) std::copy(c->fb_w.begin(), c->fb_w.end(), w);
    if (gap) std::copy(c->fb_gap.begin(), c->fb_gap.end(), gap);
    if (nbound) *nbound = c->fascia_nbound;
    return XPBD_OK;
}

int32_t xpbd_set_coarse_level(xpbd_ctx *c, xpbd_ctx *cc, const xpbd_coarse_desc *d) {
    int r = check_ready(c);
    if (r) return r;
    invalidate_graph(c);
    if (!cc || !d || d->iters <= 0) {
        c->coarse = nullptr;
        c->mg_iters = 0;
        return XPBD_OK;
    }
    if (cc == c) return fail(c, XPBD_EINVAL, "the coarse level must be another context");
    if (!cc->ready) return fail(c, XPBD_ENOTREADY, "coarse context is not finalized");
    if (cc->device != c->device) return fail(c, XPBD_EINVAL, "coarse context is on another device");
    if (!cc->flat || cc->block) return fail(c, XPBD_EINVAL, "the coarse context must use the flat strategy");
    if (c->mg_pidx_set) return fail(c, XPBD_EINVAL, "coarse level already set; finalize again to change it");
    const int nv = c->nv, nc = cc->nv, nb = d->nblk;
    if (!d->p_idx || !d->p_w || nb <= 0 || !d->blk_off || !d->minv || !d->eps || !d->fit_off || !d->fit_center ||
        !d->fit_sinv || (d->fit_off[nb] > 0 && !d->fit_fine))
        return fail(c, XPBD_EINVAL, "incomplete coarse level description");
    if (d->blk_off[0] != 0 || d->blk_off[nb] != nc) return fail(c, XPBD_EINVAL, "blocks must cover the coarse nodes");
    for (int b = 0; b < nb; ++b)
        if (d->blk_off[b + 1] <= d->blk_off[b] || d->fit_off[b + 1] < d->fit_off[b])
            return fail(c, XPBD_EINVAL, "block ranges must increase");
    // prolongation, fine device order, coarse device ids
    std::vector<int32_t> pidx(8 * (size_t)nv);
    std::vector<float> pw(8 * (size_t)nv);
    std::vector<std::vector<std::pair<int32_t, float>>> pt(nc);   // P^T per coarse node (block order)
    for (int i = 0; i < nv; ++i) {
        const int fi = c->L.iperm[i];
        for (int k = 0; k < 8; ++k) {
            const int j = d->p_idx[8 * (size_t)i + k];
            const float w = d->p_w[8 * (size_t)i + k];
            if (j < 0 || j >= nc) return fail(c, XPBD_EINVAL, "prolongation index out of range");
            if (!(w >= 0.0f) || !(w <= 1.0f)) return fail(c, XPBD_EINVAL, "prolongation weight outside [0, 1]");
            pidx[8 * (size_t)fi + k] = j;
            pw[8 * (size_t)fi + k] = w;
            if (w > 0.0f) pt[j].push_back({fi, w});
        }
    }
    std::vector<int32_t> ptoff(nc + 1, 0), ptfv, nblk_of(nc), dev_of(nc);
    std::vector<float> ptw;
    std::vector<float4> q0c(nc);
    for (int b = 0; b < nb; ++b)
        for (int j = d->blk_off[b]; j < d->blk_off[b + 1]; ++j) nblk_of[j] = b;
    for (int j = 0; j < nc; ++j) {
        std::sort(pt[j].begin(), pt[j].end());
        for (const auto &e : pt[j]) {
            ptfv.push_back(e.first);
            ptw.push_back(e.second);
        }
        ptoff[j + 1] = (int32_t)ptfv.size();
        dev_of[j] = cc->L.iperm[j];
        const float *ctr = d->fit_center + 3 * (size_t)nblk_of[j];
        q0c[j] = make_float4(cc->pos0[3 * (size_t)j] - ctr[0], cc->pos0[3 * (size_t)j + 1] - ctr[1],
                             cc->pos0[3 * (size_t)j + 2] - ctr[2], 0.f);
    }
    if (ptfv.empty()) { ptfv.push_back(0); ptw.push_back(0.0f); }
    // affine-fit entries: fine device ids and centred rest positions
    const int nfit = d->fit_off[nb];
    std::vector<int32_t> fitf(std::max(nfit, 1), 0);
    std::vector<float4> fitr(std::max(nfit, 1), make_float4(0, 0, 0, 0));
    for (int b = 0; b < nb; ++b)
        for (int e = d->fit_off[b]; e < d->fit_off[b + 1]; ++e) {
            const int i = d->fit_fine[e];
            if (i < 0 || i >= nv) return fail(c, XPBD_EINVAL, "fit vertex out of range");
            fitf[e] = c->L.iperm[i];
            const float *ctr = d->fit_center + 3 * (size_t)b;
            fitr[e] = make_float4(c->pos0[3 * (size_t)i] - ctr[0], c->pos0[3 * (size_t)i + 1] - ctr[1],
                                  c->pos0[3 * (size_t)i + 2] - ctr[2], 1.f);
        }
    std::vector<int64_t> goff(nb + 1, 0);
    for (int b = 0; b < nb; ++b) {
        const int64_t n = d->blk_off[b + 1] - d->blk_off[b];
        goff[b + 1] = goff[b] + n * n;
    }
    std::vector<int32_t> boff(d->blk_off, d->blk_off + nb + 1), foff(d->fit_off, d->fit_off + nb + 1);
    MgArgs m{};
    int32_t *p_fit_off, *p_fit_fine, *p_ptoff, *p_ptfv, *p_nblk, *p_boff, *p_dev;
    float4 *p_fit_rest, *p_q0c, *p_rhs;
    float *p_sinv, *p_T, *p_ptw, *p_eps, *p_G;
    int64_t *p_goff;
    if ((r = dupload(c, &c->d_mg_pidx, pidx.data(), pidx.size())) || (r = dupload(c, &c->d_mg_pw, pw.data(), pw.size())) ||
        (r = dupload(c, &p_fit_off, foff.data(), foff.size())) || (r = dupload(c, &p_fit_fine, fitf.data(), fitf.size())) ||
        (r = dupload(c, &p_fit_rest, fitr.data(), fitr.size())) || (r = dupload(c, &p_sinv, d->fit_sinv, 16 * (size_t)nb)) ||
        (r = dalloc(c, &p_T, 12 * (size_t)nb)) || (r = dupload(c, &p_ptoff, ptoff.data(), ptoff.size())) ||
        (r = dupload(c, &p_ptfv, ptfv.data(), ptfv.size())) || (r = dupload(c, &p_ptw, ptw.data(), ptw.size())) ||
        (r = dupload(c, &p_nblk, nblk_of.data(), nblk_of.size())) || (r = dupload(c, &p_q0c, q0c.data(), q0c.size())) ||
        (r = dupload(c, &p_eps, d->eps, (size_t)nb)) || (r = dalloc(c, &p_rhs, (size_t)nc)) ||
        (r = dupload(c, &p_boff, boff.data(), boff.size())) || (r = dupload(c, &p_goff, goff.data(), goff.size())) ||
        (r = dupload(c, &p_G, d->minv, (size_t)goff[nb])) || (r = dupload(c, &p_dev, dev_of.data(), dev_of.size())) ||
        (r = dalloc(c, &c->MG_QPREV, (size_t)std::max(nc, 1))))
        return r;
    m.fit_off = p_fit_off;
    m.fit_fine = p_fit_fine;
    m.fit_rest = p_fit_rest;
    m.fit_sinv = p_sinv;
    m.fitT = p_T;
    m.ptoff = p_ptoff;
    m.ptfv = p_ptfv;
    m.ptw = p_ptw;
    m.nblk_of = p_nblk;
    m.q0c = p_q0c;
    m.eps = p_eps;
    m.rhs = p_rhs;
    m.boff = p_boff;
    m.goff = p_goff;
    m.G = p_G;
    m.dev_of = p_dev;
    // coarse cage per block: tets, rest matrices, compliances, inverse masses,
    // per-node corner slots; all in the coarse context's caller order
    {
        const int nct = cc->nt;
        std::vector<int32_t> tblk(nct, -1), toff(nb + 1, 0);
        for (int t = 0; t < nct; ++t) {
            const int v0 = cc->tets[4 * (size_t)t];
            const int bb = (v0 >= 0 && v0 < nc) ? nblk_of[v0] : -1;
            for (int k = 1; k < 4; ++k) {
                const int vk = cc->tets[4 * (size_t)t + k];
                if (vk < 0 || vk >= nc || nblk_of[vk] != bb) return fail(c, XPBD_EINVAL, "a coarse tet spans two blocks");
            }
            tblk[t] = bb;
            toff[bb + 1]++;
        }
        for (int bb = 0; bb < nb; ++bb) toff[bb + 1] += toff[bb];
        std::vector<int32_t> fill(toff.begin(), toff.end() - 1), order(nct);
        for (int t = 0; t < nct; ++t) order[fill[tblk[t]]++] = t;
        std::vector<ushort4> tl(std::max(nct, 1));
        std::vector<float> tB(9 * (size_t)std::max(nct, 1)), tat(std::max(nct, 1));
        // compliance from the coarse ARAP table (one constraint per tet)
        std::vector<float> kt(nct, 0.f);
        for (int q = 0; q < cc->ncons; ++q)
            if (cc->c_type[q] == XPBD_CT_TETARAPSVD && cc->c_tetid[q] >= 0 && cc->c_tetid[q] < nct)
                kt[cc->c_tetid[q]] = cc->c_stiff[q];
        const float dt = cc->sp.dt;
        std::vector<std::vector<uint16_t>> nslots(nc);
        int max_nodes = 0, max_tets = 0;
        for (int bb = 0; bb < nb; ++bb) {
            max_nodes = std::max(max_nodes, d->blk_off[bb + 1] - d->blk_off[bb]);
            max_tets = std::max(max_tets, toff[bb + 1] - toff[bb]);
        }
        if (4 * max_tets > 65535) return fail(c, XPBD_ELAYOUT, "coarse cage block too large for 16-bit slots");
        if (max_tets > 2 * MGC_T) return fail(c, XPBD_ELAYOUT, "coarse cage block has more than 2048 tets");
        // k_mg_coarse loads, sums and writes back one node per thread
        if (max_nodes > MGC_T) return fail(c, XPBD_ELAYOUT, "coarse cage block has more than 1024 nodes");
        for (int e = 0; e < nct; ++e) {
            const int t = order[e];
            const int bb = tblk[t], n0 = d->blk_off[bb], lt = e - toff[bb];
            int lv[4];
            for (int k = 0; k < 4; ++k) {
                const int v = cc->tets[4 * (size_t)t + k];
                lv[k] = v - n0;
                nslots[v].push_back((uint16_t)(4 * lt + k));
            }
            tl[e] = make_ushort4((unsigned short)lv[0], (unsigned short)lv[1], (unsigned short)lv[2], (unsigned short)lv[3]);
            for (int q = 0; q < 9; ++q) tB[9 * (size_t)e + q] = cc->rest_matrix[9 * (size_t)t + q];
            tat[e] = kt[t] > 0.f ? (1.0f / kt[t]) / (dt * dt) : 0.f;
        }
        // vertex-major corner list per block: node by node, each node's corners in
        // ascending (tet, corner) order; a corner's place is relative to its block
        std::vector<int32_t> soff(nc + 1, 0);
        std::vector<ushort4> tcp(std::max(nct, 1), make_ushort4(0, 0, 0, 0));
        std::vector<float> nw(std::max(nc, 1), 0.f);
        std::vector<uint8_t> npin(std::max(nc, 1), 0);
        for (size_t i = 0; i < cc->pin_vids.size(); ++i)
            if (cc->pin_active[i]) npin[cc->pin_vids[i]] = 1;
        for (int v = 0; v < nc; ++v) {
            std::sort(nslots[v].begin(), nslots[v].end());
            const int bb = nblk_of[v], base = soff[d->blk_off[bb]];
            for (size_t e = 0; e < nslots[v].size(); ++e) {
                const int code = nslots[v][e], lt = code >> 2, k = code & 3;
                const unsigned short pos = (unsigned short)(soff[v] + (int)e - base);
                ushort4 &c4 = tcp[toff[bb] + lt];
                if (k == 0) c4.x = pos; else if (k == 1) c4.y = pos; else if (k == 2) c4.z = pos; else c4.w = pos;
            }
            soff[v + 1] = soff[v] + (int32_t)nslots[v].size();
            nw[v] = inv_mass_of(cc, v);
        }
        MgCoarse g{};
        int32_t *p_toff, *p_soff;
        ushort4 *p_tl;
        float *p_tB, *p_tat, *p_nw;
        uint8_t *p_npin;
        ushort4 *p_tcp;
        if ((r = dupload(c, &p_toff, toff.data(), toff.size())) || (r = dupload(c, &p_tl, tl.data(), tl.size())) ||
            (r = dupload(c, &p_tB, tB.data(), tB.size())) || (r = dupload(c, &p_tat, tat.data(), tat.size())) ||
            (r = dupload(c, &p_nw, nw.data(), nw.size())) || (r = dupload(c, &p_npin, npin.data(), npin.size())) ||
            (r = dupload(c, &p_soff, soff.data(), soff.size())) ||
            (r = dupload(c, &p_tcp, tcp.data(), tcp.size())) || (r = dalloc(c, &m.q, (size_t)std::max(nc, 1))))
            return r;
        g.tet_off = p_toff;
        g.tlidx = p_tl;
        g.tB = p_tB;
        g.tat = p_tat;
        g.node_w = p_nw;
        g.node_pin = p_npin;
        g.pin_rule = d->pin_rule ? 1 : 0;
        {
            // pin goals start at the coarse rest (until xpbd_set_coarse_pin_targets)
            std::vector<float4> pt(std::max(nc, 1), make_float4(0.f, 0.f, 0.f, 0.f));
            for (int v = 0; v < nc; ++v)
                pt[v] = make_float4(cc->pos0[3 * (size_t)v], cc->pos0[3 * (size_t)v + 1], cc->pos0[3 * (size_t)v + 2], 0.f);
            if ((r = dupload(c, &c->d_mg_pin_tgt, pt.data(), pt.size()))) return r;
        }
        g.pin_tgt = c->d_mg_pin_tgt;
        g.slot_off = p_soff;
        g.tcpos = p_tcp;
        g.max_nodes = max_nodes;
        g.max_tets = std::max(max_tets, 1);
        const size_t lds = ((size_t)max_nodes + 4 * (size_t)g.max_tets) * sizeof(float4);
        int maxlds = 0;
        HIPCHK(c, hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
        if (lds > (size_t)maxlds)
            return fail(c, XPBD_ELAYOUT, "coarse cage block needs " + std::to_string(lds) + " B of LDS; limit " +
                                             std::to_string(maxlds));
        HIPCHK(c, hipFuncSetAttribute((const void *)k_mg_coarse, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        c->mgc = g;
        c->mgc_lds = lds;
    }
    c->mg = m;
    c->mg_nblk = nb;
    c->coarse = cc;
    c->mg_iters = d->iters;
    c->mg_kinematic = d->kinematic != 0;
    c->mg_pidx_set = true;
    return XPBD_OK;
}

int32_t xpbd_set_coarse_pin_targets(xpbd_ctx *c, int32_t n, const float *xyz) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->coarse || !c->mg_pidx_set) return fail(c, XPBD_EINVAL, "no coarse level set");
    if (n != c->coarse->nv || !xyz) return fail(c, XPBD_EINVAL, "one coarse pin goal per coarse node expected");
    int slot;
    void *h;
    const size_t bytes = (size_t)n * sizeof(float4);
    if ((r = stage_acquire(c, c->st_mgpin, bytes, slot, h))) return r;
    float4 *t = (float4 *)h;
    for (int i = 0; i < n; ++i) t[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.f);
    return stage_submit(c, c->st_mgpin, slot, c->d_mg_pin_tgt, bytes);
}

int32_t xpbd_vcycle(xpbd_ctx *c) {
    int r = check_ready(c);
    if (r) return r;
    return (c->coarse && c->mg_iters > 0) ? launch_vcycle(c) : XPBD_OK;
}

int32_t xpbd_set_lambda_mode(xpbd_ctx *c, int32_t mode) {
    int r = check_ready(c);
    if (r) return r;
    if (mode < 0 || mode > 4) return fail(c, XPBD_EINVAL, "lambda mode must be 0..4");
    if (mode && !c->flat) return fail(c, XPBD_EUNSUPPORTED, "lambda modes run on the flat strategy");
    if (mode && (c->feature & F_ARAP) && (c->feature & F_SNH))
        return fail(c, XPBD_EUNSUPPORTED, "lambda modes need ARAP-SVD and SNH on separate meshes");
    if ((mode == 2 || (mode == 4 && c->warm_net)) && !c->mlp_ready)
        return fail(c, XPBD_EINVAL, "lambda mode 2 and network schedules need xpbd_set_mlp()");
    c->lam_mode = mode;
    c->impulse_pending = false;
    return XPBD_OK;
}

int32_t xpbd_warm_start(xpbd_ctx *c) {
    int r = check_ready(c);
    if (r) return r;
    c->eff_mode = c->lam_mode == 4 ? (c->warm_net ? 2 : 3) : c->lam_mode;
    return c->eff_mode ? launch_warm_start(c) : XPBD_OK;
}

int32_t xpbd_features(xpbd_ctx *c, float *dev_out, int32_t normalized) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->flat) return fail(c, XPBD_EUNSUPPORTED, "features run on the flat strategy");
    if (!dev_out) return fail(c, XPBD_EINVAL, "null output");
    const int64_t ne = (int64_t)c->L.meta.size();
    if ((r = ensure_static_features(c))) return r;
    IterArgs a = make_args(c, c->DPW);
    hipLaunchKernelGGL(k_features, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, c->stream, a,
                       (const float4 *)c->X[0], (const float4 *)c->LAMB0, (const float4 *)c->SFEAT, dev_out,
                       (const float *)c->fmean, (const float *)c->finv, normalized ? 1 : 0);
    HIPCHK(c, hipGetLastError());
    return XPBD_OK;
}

int32_t xpbd_get_lambda(xpbd_ctx *c, float *dev_out) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->flat || !dev_out) return fail(c, XPBD_EINVAL, "flat strategy and an output buffer required");
    HIPCHK(c, hipMemcpyAsync(dev_out, c->LAMB, c->L.meta.size() * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
    return XPBD_OK;
}

int32_t xpbd_set_lambda(xpbd_ctx *c, const float *dev_in) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->flat || !dev_in) return fail(c, XPBD_EINVAL, "flat strategy and an input buffer required");
    HIPCHK(c, hipMemcpyAsync(c->LAMB, dev_in, c->L.meta.size() * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->LAMB0, dev_in, c->L.meta.size() * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
    return XPBD_OK;
}

int32_t xpbd_set_reference_iters(xpbd_ctx *c, int32_t n) {
    int r = check_ready(c);
    if (r) return r;
    if (n < 1) return fail(c, XPBD_EINVAL, "reference iterations must be >= 1");
    c->ref_iters = n;
    return XPBD_OK;
}

int32_t xpbd_set_warm_schedule(xpbd_ctx *c, int32_t ref_iters, int32_t ref_every, int32_t use_network) {
    int r = check_ready(c);
    if (r) return r;
    if (ref_iters < 1 || ref_every < 0) return fail(c, XPBD_EINVAL, "reference iterations >= 1, ref_every >= 0");
    if (use_network && !c->mlp_ready) return fail(c, XPBD_EINVAL, "a network schedule needs xpbd_set_mlp()");
    c->ref_iters = ref_iters;
    c->ref_every = ref_every;
    c->warm_net = use_network ? 1 : 0;
    return XPBD_OK;
}

int32_t xpbd_snapshot(xpbd_ctx *c, int32_t save) {
    int r = check_ready(c);
    if (r) return r;
    const size_t b = (size_t)c->nv * sizeof(float4);
    if (save) {
        HIPCHK(c, hipMemcpyAsync(c->S_X, c->X[0], b, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->S_PP, c->PPv, b, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->S_VEL, c->VEL, b, hipMemcpyDeviceToDevice, c->stream));
    } else {
        HIPCHK(c, hipMemcpyAsync(c->X[0], c->S_X, b, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->PPv, c->S_PP, b, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->VEL, c->S_VEL, b, hipMemcpyDeviceToDevice, c->stream));
    }
    return XPBD_OK;
}

int32_t xpbd_kernel_timing(xpbd_ctx *c, int32_t enable) {
    if (!c) return XPBD_EINVAL;
    if (enable) {
        for (auto &e : c->evs) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
        c->evs.clear();
        c->timed_ms = 0.0;
        c->timed_n = 0;
    }
    c->timing = enable != 0;
    return XPBD_OK;
}

int32_t xpbd_time_iterations(xpbd_ctx *c, int32_t n, int32_t reps, double *us_per_iter) {
    int r = check_ready(c);
    if (r) return r;
    if (n < 2 || (n & 1) || reps < 1 || (int64_t)n * reps > 4096 || !us_per_iter)
        return fail(c, XPBD_EINVAL, "time_iterations needs an even n >= 2, reps >= 1, n * reps <= 4096 and an output");
    if (c->lam_mode != 0) return fail(c, XPBD_EUNSUPPORTED, "iteration timing runs the plain (lambda mode 0) kernels");
    const size_t b = (size_t)c->nv * sizeof(float4);
    if (!c->T_X && (r = dalloc(c, &c->T_X, (size_t)c->nv))) return r;
    HIPCHK(c, hipMemcpyAsync(c->T_X, c->X[0], b, hipMemcpyDeviceToDevice, c->stream));
    const bool tm = c->timing;
    c->timing = false;
    c->eff_mode = 0;
    // all n * reps iterations go into one graph, so the timed region holds a
    // single graph launch and no host gaps between replays
    const int total = n * reps;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    if (c->res_R > 0) {
        // resident mode: one launch per n iterations (n is even, so each ends in X[0])
        for (int k = 0; k < reps && r == XPBD_OK; ++k) r = launch_resident(c, c->X[0], c->X[0], n);
    } else {
        for (int it = 0; it < total && r == XPBD_OK; ++it)
            r = launch_iter(c, c->X[it & 1], c->X[(it + 1) & 1], MODE_FUSED, 0);
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->timing = tm;
    if (r) { if (g) (void)hipGraphDestroy(g); return r; }
    if (e != hipSuccess) return fail(c, XPBD_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t ge = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    hipError_t he = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he == hipSuccess) he = hipGraphLaunch(ge, c->stream);   // warm
    if (he == hipSuccess) he = hipEventRecord(e0, c->stream);
    if (he == hipSuccess) he = hipGraphLaunch(ge, c->stream);
    if (he == hipSuccess) he = hipEventRecord(e1, c->stream);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (ge) (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    if (he != hipSuccess) return fail(c, XPBD_EHIP, std::string("iteration timing: ") + hipGetErrorString(he));
    HIPCHK(c, hipMemcpyAsync(c->X[0], c->T_X, b, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((r = check_resident(c))) return r;
    *us_per_iter = 1e3 * (double)ms / (double)total;
    return XPBD_OK;
}

int32_t xpbd_get_kernel_timing(xpbd_ctx *c, double *total_ms, int32_t *launches) {
    if (!c) return XPBD_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto &e : c->evs) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, e.a, e.b));
        c->timed_ms += ms;
        c->timed_n += 1;
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    c->evs.clear();
    if (total_ms) *total_ms = c->timed_ms;
    if (launches) *launches = c->timed_n;
    return XPBD_OK;
}

}  // extern "C"
