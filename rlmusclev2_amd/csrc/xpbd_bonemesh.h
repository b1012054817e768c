// Exact-mesh bone collider (gfx950).  Included by xpbd.hip; the part outside
// __HIPCC__ (data layout, tree tables, query) also builds with a plain C++
// compiler for the host checks of the tree.
//
// Semantics: solve_mesh_contact_pos_kernel (src/muscle/collision/sdf/kernel.py:
// 123-170) and the penetration meter _mark_inside_kernel (src/example/
// humanV2.py:67-87), which query the skeleton through Warp's
// mesh_query_point_sign_normal: the closest point on the bone mesh among the
// faces nearer than max_dist, and the inside/outside sign of the angle-weighted
// pseudonormal of the closest feature (face: its normal; edge: the sum of its
// two face normals; vertex: the corner-angle-weighted sum of its face normals;
// Baerentzen and Aanaes, IEEE TVCG 11(3), 2005).  A penetrating vertex is moved
// along (p - closest) * sign to contact_offset outside the surface.
//
// Mechanism (not the reference's): per bone, a binary BVH in implicit heap
// order (children of node k: 2k + 1, 2k + 2) over leaves of BM_LEAF
// consecutive triangles, padded to a power of two, the triangles ordered by
// top-down median splits along the longest axis; the tree's shape is fixed by the first pose, only its boxes are refitted
// per frame (new bone vertex positions, same topology, like wp.Mesh.refit).
// Per frame: triangle soup and face normals (k_bm_soup), vertex pseudonormals
// from a fixed incidence list (k_bm_vnorm, no atomics), leaf boxes plus the
// lower BM_SUB_LOG levels of each subtree in LDS (k_bm_refit_low), the levels
// above in LDS again (k_bm_refit_top).  Per query: near-child-first traversal
// without a stack (the heap index of the current node and one bit per level
// saying whether it is the second child taken), every box pruned against the
// best squared distance so far.
//
// Coherent mode (closed meshes): a vertex whose last query found it outside at
// unsigned distance s, and which has since moved by |dx| while no bone vertex
// moved by more than dM in total, is still outside at distance >= s - |dx| - dM
// (distance to the surface is 1-Lipschitz in both).  When that bound exceeds
// the contact offset (the meter: zero) the query cannot change the outcome and
// is skipped; the results are those of querying every vertex.  This needs the
// pseudonormal sign to be the true inside test: every sub-mesh closed (each
// edge shared by two triangles in opposite directions, checked when the tree is
// built) and the sub-meshes one vertex tests not overlapping (the caller's
// assurance).  Per vertex cache: the last query's position and s + M, where M
// is the running sum of the largest bone-vertex displacement of each upload.
#pragma once

#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BM_HD __host__ __device__ inline
#else
#define BM_HD inline
#endif

namespace xpbd {

constexpr int BM_LEAF = 4;          // triangles per leaf
constexpr int BM_SUB_LOG = 8;       // leaves per k_bm_refit_low workgroup = 2^BM_SUB_LOG
constexpr int BM_TOP_MAX = 2048;    // subtree roots one k_bm_refit_top workgroup reduces in LDS
constexpr int BM_MAX_DEPTH = 30;    // leaf levels (the bit trail is 32 bits)

struct alignas(16) bm_f4 { float x, y, z, w; };
struct alignas(16) bm_i4 { int x, y, z, w; };
struct alignas(8) bm_i2 { int x, y; };
struct bm_v3 { float x, y, z; };

BM_HD bm_v3 bm_mk(float x, float y, float z) { return bm_v3{x, y, z}; }
BM_HD bm_v3 bm_xyz(bm_f4 a) { return bm_v3{a.x, a.y, a.z}; }
BM_HD bm_v3 operator+(bm_v3 a, bm_v3 b) { return bm_v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
BM_HD bm_v3 operator-(bm_v3 a, bm_v3 b) { return bm_v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
BM_HD bm_v3 operator*(float s, bm_v3 a) { return bm_v3{s * a.x, s * a.y, s * a.z}; }
BM_HD float bm_dot(bm_v3 a, bm_v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BM_HD bm_v3 bm_cross(bm_v3 a, bm_v3 b) {
    return bm_v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct BoneMeshDev {
    int nbones, ntris, nbv;
    const bm_i4 *bone;         // per bone: x first node, y leaves P (power of two), z first sorted triangle, w count
    const bm_i4 *tv;           // sorted triangle: bone-mesh vertex ids a, b, c; w = caller's triangle id
    const bm_i4 *adj;          // sorted triangle: neighbour (sorted id) across ab, bc, ca; -1 none
    const bm_i2 *sub;          // k_bm_refit_low workgroups: (bone, first leaf)
    const int32_t *vt_off;     // nbv + 1: incidence list of each bone vertex
    const int32_t *vt;         // sorted triangle << 2 | corner
    const bm_f4 *pts;          // bone vertex positions (caller's bone vertex order)
    bm_f4 *soup;               // 3 per sorted triangle: a, b, c
    bm_f4 *fn;                 // sorted triangle: unit face normal
    bm_f4 *vn;                 // bone vertex: angle-weighted pseudonormal (not normalised)
    bm_f4 *node;               // 2 per node: lo.xyz, hi.xyz
    const int32_t *vgroup;     // per vertex (library order): group
    const int32_t *goff;       // ngroups + 1: bones goff[g] .. goff[g + 1] - 1
    const float *cw;           // per vertex (library order): 0 = not pushed
    const float *mw;           // per vertex (library order): 0 = not metered; null = all
    bm_f4 *cache;              // coherent mode, per vertex (library order): last query position, s + M (-inf: none)
    unsigned long long *nquery; // running count of the queries made (the vertices not certified)
    float max_dist, contact_offset, max_push;
    int coherent;              // 1: skip vertices certified outside (pts[nbv].x holds M)
};

// ------------------------------------------------------------ per-frame data
BM_HD void bm_soup_tri(const BoneMeshDev &m, int t) {
    const bm_i4 v = m.tv[t];
    const bm_f4 a = m.pts[v.x], b = m.pts[v.y], c = m.pts[v.z];
    m.soup[3 * t + 0] = bm_f4{a.x, a.y, a.z, 0.f};
    m.soup[3 * t + 1] = bm_f4{b.x, b.y, b.z, 0.f};
    m.soup[3 * t + 2] = bm_f4{c.x, c.y, c.z, 0.f};
    const bm_v3 n = bm_cross(bm_xyz(b) - bm_xyz(a), bm_xyz(c) - bm_xyz(a));
    const float l = sqrtf(bm_dot(n, n));
    m.fn[t] = l > 0.f ? bm_f4{n.x / l, n.y / l, n.z / l, 0.f} : bm_f4{0.f, 0.f, 0.f, 0.f};
}

BM_HD void bm_vertex_normal(const BoneMeshDev &m, int i) {
    bm_v3 s = bm_mk(0.f, 0.f, 0.f);
    for (int k = m.vt_off[i]; k < m.vt_off[i + 1]; ++k) {
        const int t = m.vt[k] >> 2, c = m.vt[k] & 3;
        const bm_v3 p0 = bm_xyz(m.soup[3 * t + c]), p1 = bm_xyz(m.soup[3 * t + (c + 1) % 3]),
                    p2 = bm_xyz(m.soup[3 * t + (c + 2) % 3]);
        const bm_v3 e1 = p1 - p0, e2 = p2 - p0;
        const float l1 = sqrtf(bm_dot(e1, e1)), l2 = sqrtf(bm_dot(e2, e2));
        float ca = (l1 > 0.f && l2 > 0.f) ? bm_dot(e1, e2) / (l1 * l2) : 1.f;
        ca = fminf(1.f, fmaxf(-1.f, ca));
        s = s + acosf(ca) * bm_xyz(m.fn[t]);
    }
    m.vn[i] = bm_f4{s.x, s.y, s.z, 0.f};
}

BM_HD int bm_depth(int P) {
    int D = 0;
    while ((1 << D) < P) ++D;
    return D;
}

BM_HD void bm_box_union(bm_f4 &lo, bm_f4 &hi, bm_f4 lo2, bm_f4 hi2) {
    lo.x = fminf(lo.x, lo2.x); lo.y = fminf(lo.y, lo2.y); lo.z = fminf(lo.z, lo2.z);
    hi.x = fmaxf(hi.x, hi2.x); hi.y = fmaxf(hi.y, hi2.y); hi.z = fmaxf(hi.z, hi2.z);
}

// box of leaf `leaf` of bone bn (empty: lo = +inf, hi = -inf)
BM_HD void bm_leaf_box(const BoneMeshDev &m, bm_i4 bn, int leaf, bm_f4 &lo, bm_f4 &hi) {
    lo = bm_f4{INFINITY, INFINITY, INFINITY, 0.f};
    hi = bm_f4{-INFINITY, -INFINITY, -INFINITY, 0.f};
    const int t0 = leaf * BM_LEAF, t1 = t0 + BM_LEAF < bn.w ? t0 + BM_LEAF : bn.w;
    for (int k = t0; k < t1; ++k)
        for (int c = 0; c < 3; ++c) {
            const bm_f4 p = m.soup[3 * (bn.z + k) + c];
            bm_box_union(lo, hi, p, p);
        }
}

// heap index (within the bone) of the node l levels above the leaves whose
// subtree starts at leaf first + (t << l)
BM_HD int bm_level_node(int D, int l, int first, int t) { return (1 << (D - l)) - 1 + (first >> l) + t; }

// ------------------------------------------------------------ query
BM_HD float bm_box_d2(const bm_f4 *node, int k, bm_v3 p) {
    const bm_f4 lo = node[2 * k], hi = node[2 * k + 1];
    const float dx = fmaxf(fmaxf(lo.x - p.x, 0.f), p.x - hi.x);
    const float dy = fmaxf(fmaxf(lo.y - p.y, 0.f), p.y - hi.y);
    const float dz = fmaxf(fmaxf(lo.z - p.z, 0.f), p.z - hi.z);
    return dx * dx + dy * dy + dz * dz;
}

BM_HD bm_v3 bm_seg(bm_v3 p, bm_v3 a, bm_v3 b) {
    const bm_v3 ab = b - a;
    const float l2 = bm_dot(ab, ab);
    float s = 0.f;
    if (l2 > 0.f) s = fminf(1.f, fmaxf(0.f, bm_dot(p - a, ab) / l2));
    return a + s * ab;
}

// closest point of triangle a b c to p by its Voronoi regions (Ericson,
// Real-Time Collision Detection, 2005, 5.1.5); region 0 face, 1 2 3 vertex
// a b c, 4 edge ab, 5 edge bc, 6 edge ca
BM_HD bm_v3 bm_tri_closest(bm_v3 p, bm_v3 a, bm_v3 b, bm_v3 c, int &region) {
    const bm_v3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = bm_dot(ab, ap), d2 = bm_dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) { region = 1; return a; }
    const bm_v3 bp = p - b;
    const float d3 = bm_dot(ab, bp), d4 = bm_dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) { region = 2; return b; }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) { region = 4; return a + (d1 / (d1 - d3)) * ab; }
    const bm_v3 cp = p - c;
    const float d5 = bm_dot(ab, cp), d6 = bm_dot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) { region = 3; return c; }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) { region = 6; return a + (d2 / (d2 - d6)) * ac; }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
        region = 5;
        return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
    }
    const float den = va + vb + vc;
    if (!(den > 0.f)) {          // zero-area triangle: nearest of its three edges
        const bm_v3 q0 = bm_seg(p, a, b), q1 = bm_seg(p, b, c), q2 = bm_seg(p, c, a);
        const float e0 = bm_dot(p - q0, p - q0), e1 = bm_dot(p - q1, p - q1), e2 = bm_dot(p - q2, p - q2);
        if (e0 <= e1 && e0 <= e2) { region = 4; return q0; }
        if (e1 <= e2) { region = 5; return q1; }
        region = 6;
        return q2;
    }
    region = 0;
    return a + (vb / den) * ab + (vc / den) * ac;
}

struct BmHit {
    float d2;     // squared distance to the closest point
    bm_v3 cp;
    int face;     // sorted triangle, -1 = none nearer than max_dist
    int region;
};

BM_HD void bm_traverse(const BoneMeshDev &m, int b, bm_v3 p, BmHit &h) {
    const bm_i4 bn = m.bone[b];
    const int P = bn.y, D = bm_depth(P);
    const bm_f4 *node = m.node + 2 * (size_t)bn.x;
    int k = 0, d = 0;
    uint32_t trail = 0u;          // bit d: the node at depth d is the second child taken
    bool go = bm_box_d2(node, 0, p) < h.d2;
    for (;;) {
        if (go) {
            if (d == D) {
                const int leaf = k - (P - 1);
                const int t0 = leaf * BM_LEAF, t1 = t0 + BM_LEAF < bn.w ? t0 + BM_LEAF : bn.w;
                for (int j = t0; j < t1; ++j) {
                    const int t = bn.z + j;
                    int reg;
                    const bm_v3 q = bm_tri_closest(p, bm_xyz(m.soup[3 * t]), bm_xyz(m.soup[3 * t + 1]),
                                                   bm_xyz(m.soup[3 * t + 2]), reg);
                    const bm_v3 dd = p - q;
                    const float e = bm_dot(dd, dd);
                    if (e < h.d2) { h.d2 = e; h.cp = q; h.face = t; h.region = reg; }
                }
            } else {
                const float dl = bm_box_d2(node, 2 * k + 1, p), dr = bm_box_d2(node, 2 * k + 2, p);
                ++d;
                trail &= ~(1u << d);
                k = dr < dl ? 2 * k + 2 : 2 * k + 1;     // nearer child first
                go = fminf(dl, dr) < h.d2;
                continue;
            }
        }
        // next: the other child of the deepest ancestor whose second child is not taken yet
        while (d > 0 && ((trail >> d) & 1u)) {
            k = (k - 1) >> 1;
            --d;
        }
        if (d == 0) break;
        trail |= 1u << d;
        k = (k & 1) ? k + 1 : k - 1;                     // the sibling
        go = bm_box_d2(node, k, p) < h.d2;
    }
}

// the sign of the closest feature's angle-weighted pseudonormal
BM_HD float bm_sign(const BoneMeshDev &m, bm_v3 p, const BmHit &h) {
    bm_v3 N = bm_xyz(m.fn[h.face]);
    if (h.region >= 1 && h.region <= 3) {
        const bm_i4 v = m.tv[h.face];
        N = bm_xyz(m.vn[h.region == 1 ? v.x : (h.region == 2 ? v.y : v.z)]);
    } else if (h.region >= 4) {
        const bm_i4 a = m.adj[h.face];
        const int g = h.region == 4 ? a.x : (h.region == 5 ? a.y : a.z);
        if (g >= 0) N = N + bm_xyz(m.fn[g]);
    }
    return bm_dot(p - h.cp, N) < 0.f ? -1.f : 1.f;
}

// closest point over the sub-meshes of group g; false without a face nearer than max_dist
BM_HD bool bm_query(const BoneMeshDev &m, int g, bm_v3 p, BmHit &h, float &sign) {
    h.d2 = m.max_dist * m.max_dist;
    h.face = -1;
    h.region = 0;
    h.cp = bm_mk(0.f, 0.f, 0.f);
    for (int b = m.goff[g]; b < m.goff[g + 1]; ++b) bm_traverse(m, b, p, h);
    if (h.face < 0) return false;
    sign = bm_sign(m, p, h);
    return true;
}

// the push of solve_mesh_contact_pos_kernel (kernel.py:155-170) from a query
// result; false = unchanged
BM_HD bool bm_push_from(const BoneMeshDev &m, const BmHit &h, float sg, bm_v3 &p) {
    const bm_v3 delta = p - h.cp;
    const float dlen = sqrtf(bm_dot(delta, delta));
    if (dlen < 1.0e-9f) return false;
    const float err = dlen * sg - m.contact_offset;
    if (!(err < 0.f && (m.max_push <= 0.f || -err <= m.max_push))) return false;
    const bm_v3 n = bm_mk(delta.x / dlen * sg, delta.y / dlen * sg, delta.z / dlen * sg);
    p = bm_mk(p.x - n.x * err, p.y - n.y * err, p.z - n.z * err);
    return true;
}

// _mark_inside_kernel (humanV2.py:76-87): > 0 inside, 0 outside
BM_HD float bm_depth_from(const BmHit &h, float sg, bm_v3 p) {
    const bm_v3 delta = p - h.cp;
    const float d = sqrtf(bm_dot(delta, delta)) * sg;
    return d < 0.f ? -d : 0.f;
}

BM_HD bool bm_push_point(const BoneMeshDev &m, int g, bm_v3 &p) {
    BmHit h;
    float sg;
    return bm_query(m, g, p, h, sg) && bm_push_from(m, h, sg, p);
}

BM_HD float bm_depth_point(const BoneMeshDev &m, int g, bm_v3 p) {
    BmHit h;
    float sg;
    return bm_query(m, g, p, h, sg) ? bm_depth_from(h, sg, p) : 0.f;
}

// coherent mode: true when the cached certificate shows vertex i outside by more
// than `clear` (the contact offset for the push, 0 for the meter)
BM_HD bool bm_certified_outside(const BoneMeshDev &m, int i, bm_v3 p, float clear) {
    const bm_f4 c = m.cache[i];
    const float M = m.pts[m.nbv].x;
    const bm_v3 d = p - bm_xyz(c);
    const float dx = sqrtf(bm_dot(d, d));
    // 10 um plus a relative term cover the rounding of s, M and |dx|
    return c.w - M - dx > fmaxf(clear, 0.f) + 1.0e-5f + 1.0e-6f * (fabsf(c.w) + M);
}

BM_HD void bm_certify(const BoneMeshDev &m, int i, bm_v3 p, bool found, const BmHit &h, float sg) {
    float s = -INFINITY;
    if (found && sg > 0.f) s = sqrtf(h.d2) + m.pts[m.nbv].x;
    m.cache[i] = bm_f4{p.x, p.y, p.z, s};
}

// ------------------------------------------------------------ host tables
struct BmTables {
    std::vector<bm_i4> bone, tv, adj;
    std::vector<bm_i2> sub;
    std::vector<int32_t> vt_off, vt;
    int nnodes = 0;
    bool closed = false;    // every edge shared by exactly two triangles, in opposite directions
};

// per sub-mesh: triangles ordered by median splits of their centroids in the
// given pose, leaves of BM_LEAF consecutive triangles padded to a power of two;
// edge neighbours: the first other triangle (caller's order) on
// the same undirected edge; vertex incidences by triangle then corner.
// Returns an error message, empty on success.  Inputs are already validated.
inline std::string bm_build_tables(int nbv, const float *pts, int nt, const int32_t *tris, int nb,
                                   const int32_t *tri_off, BmTables &o) {
    o.bone.assign(nb, bm_i4{0, 0, 0, 0});
    o.tv.assign(nt, bm_i4{0, 0, 0, 0});
    o.adj.assign(nt, bm_i4{-1, -1, -1, 0});
    o.sub.clear();
    std::vector<int32_t> sorted_of(nt);
    int64_t node = 0;
    for (int k = 0; k < nb; ++k) {
        const int t0 = tri_off[k], n = tri_off[k + 1] - t0;
        std::vector<float> cen(3 * (size_t)n);
        for (int j = 0; j < n; ++j)
            for (int a = 0; a < 3; ++a) {
                const int32_t *t = &tris[3 * (size_t)(t0 + j)];
                cen[3 * j + a] = (pts[3 * t[0] + a] + pts[3 * t[1] + a] + pts[3 * t[2] + a]) / 3.0f;
            }
        // top-down median splits along the longest axis of each node's centroid box, at the
        // leaf boundary that halves the node's leaves (equal-count splits keep the implicit
        // heap layout); ties by triangle index
        const int nleaf = std::max(1, (n + BM_LEAF - 1) / BM_LEAF);
        int Pk = 1;
        while (Pk < nleaf) Pk <<= 1;
        std::vector<int> perm(n);
        for (int j = 0; j < n; ++j) perm[j] = j;
        std::vector<std::pair<int, int>> todo{{0, Pk}};
        while (!todo.empty()) {
            const int l0 = todo.back().first, l1 = todo.back().second;
            todo.pop_back();
            if (l1 - l0 < 2) continue;
            const int a0 = std::min(l0 * BM_LEAF, n), a1 = std::min(l1 * BM_LEAF, n);
            const int lm = (l0 + l1) / 2, am = std::min(lm * BM_LEAF, n);
            if (am > a0 && am < a1) {
                float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
                for (int j = a0; j < a1; ++j)
                    for (int a = 0; a < 3; ++a) {
                        blo[a] = std::min(blo[a], cen[3 * perm[j] + a]);
                        bhi[a] = std::max(bhi[a], cen[3 * perm[j] + a]);
                    }
                int ax = 0;
                for (int a = 1; a < 3; ++a)
                    if (bhi[a] - blo[a] > bhi[ax] - blo[ax]) ax = a;
                std::nth_element(perm.begin() + a0, perm.begin() + am, perm.begin() + a1, [&](int x, int y) {
                    const float cx = cen[3 * x + ax], cy = cen[3 * y + ax];
                    return cx != cy ? cx < cy : x < y;
                });
            }
            todo.push_back({l0, lm});
            todo.push_back({lm, l1});
        }
        for (int j = 0; j < n; ++j) {
            const int f = t0 + perm[j], s = t0 + j;
            sorted_of[f] = s;
            o.tv[s] = bm_i4{tris[3 * (size_t)f], tris[3 * (size_t)f + 1], tris[3 * (size_t)f + 2], f};
        }
        const int leaves = std::max(1, (n + BM_LEAF - 1) / BM_LEAF);
        int P = 1, D = 0;
        while (P < leaves) { P <<= 1; ++D; }
        if (D > BM_MAX_DEPTH || (D > BM_SUB_LOG && (1 << (D - BM_SUB_LOG)) > BM_TOP_MAX))
            return "bone sub-mesh has too many triangles (at most 2^21)";
        o.bone[k] = bm_i4{(int)node, P, t0, n};
        node += 2 * (int64_t)P - 1;
        const int S = std::min(P, 1 << BM_SUB_LOG);
        for (int f = 0; f < P; f += S) o.sub.push_back(bm_i2{k, f});
        if (node > ((int64_t)1 << 30)) return "bone mesh too large";
    }
    o.nnodes = (int)node;
    struct E { int64_t key; int f, e; };
    std::vector<E> es(3 * (size_t)nt);
    for (int f = 0; f < nt; ++f)
        for (int e = 0; e < 3; ++e) {
            const int64_t u = tris[3 * (size_t)f + e], v = tris[3 * (size_t)f + (e + 1) % 3];
            es[3 * (size_t)f + e] = E{std::min(u, v) * nbv + std::max(u, v), f, e};
        }
    std::sort(es.begin(), es.end(), [](const E &x, const E &y) {
        return x.key != y.key ? x.key < y.key : (x.f != y.f ? x.f < y.f : x.e < y.e);
    });
    o.closed = true;
    for (size_t g0 = 0; g0 < es.size();) {
        size_t g1 = g0 + 1;
        while (g1 < es.size() && es[g1].key == es[g0].key) ++g1;
        if (g1 - g0 != 2 || tris[3 * (size_t)es[g0].f + es[g0].e] == tris[3 * (size_t)es[g0 + 1].f + es[g0 + 1].e])
            o.closed = false;
        for (size_t i = g0; i < g1; ++i) {
            int nbr = -1;
            for (size_t q = g0; q < g1; ++q)
                if (es[q].f != es[i].f) { nbr = es[q].f; break; }
            bm_i4 &a = o.adj[sorted_of[es[i].f]];
            const int s = nbr < 0 ? -1 : sorted_of[nbr];
            if (es[i].e == 0) a.x = s; else if (es[i].e == 1) a.y = s; else a.z = s;
        }
        g0 = g1;
    }
    o.vt_off.assign(nbv + 1, 0);
    for (int64_t i = 0; i < 3 * (int64_t)nt; ++i) ++o.vt_off[tris[i] + 1];
    for (int i = 0; i < nbv; ++i) o.vt_off[i + 1] += o.vt_off[i];
    o.vt.assign(3 * (size_t)nt, 0);
    std::vector<int32_t> fill(o.vt_off.begin(), o.vt_off.end() - 1);
    for (int f = 0; f < nt; ++f)
        for (int e = 0; e < 3; ++e) o.vt[fill[tris[3 * (size_t)f + e]]++] = (sorted_of[f] << 2) | e;
    return std::string();
}

// the per-frame refit in plain loops, with the kernels' per-thread steps
// (host checks of the tree; the library runs the kernels below)
inline void bm_refit_host(const BoneMeshDev &m) {
    for (int t = 0; t < m.ntris; ++t) bm_soup_tri(m, t);
    for (int i = 0; i < m.nbv; ++i) bm_vertex_normal(m, i);
    for (int b = 0; b < m.nbones; ++b) {
        const bm_i4 bn = m.bone[b];
        const int P = bn.y, D = bm_depth(P);
        for (int leaf = 0; leaf < P; ++leaf) {
            bm_f4 lo, hi;
            bm_leaf_box(m, bn, leaf, lo, hi);
            m.node[2 * (bn.x + P - 1 + leaf)] = lo;
            m.node[2 * (bn.x + P - 1 + leaf) + 1] = hi;
        }
        for (int l = 1; l <= D; ++l)
            for (int t = 0; t < (P >> l); ++t) {
                const int k = bm_level_node(D, l, 0, t);
                bm_f4 lo = m.node[2 * (bn.x + 2 * k + 1)], hi = m.node[2 * (bn.x + 2 * k + 1) + 1];
                bm_box_union(lo, hi, m.node[2 * (bn.x + 2 * k + 2)], m.node[2 * (bn.x + 2 * k + 2) + 1]);
                m.node[2 * (bn.x + k)] = lo;
                m.node[2 * (bn.x + k) + 1] = hi;
            }
    }
}

#if defined(__HIPCC__)
// ------------------------------------------------------------ kernels
__global__ void k_bm_soup(BoneMeshDev m) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m.ntris) bm_soup_tri(m, t);
}

__global__ void k_bm_vnorm(BoneMeshDev m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m.nbv) bm_vertex_normal(m, i);
}

// one workgroup of 2^BM_SUB_LOG threads per subtree of up to 2^BM_SUB_LOG leaves
__global__ __launch_bounds__(1 << BM_SUB_LOG) void k_bm_refit_low(BoneMeshDev m) {
    __shared__ bm_f4 slo[1 << BM_SUB_LOG], shi[1 << BM_SUB_LOG];
    const bm_i2 s = m.sub[blockIdx.x];
    const bm_i4 bn = m.bone[s.x];
    const int P = bn.y, D = bm_depth(P);
    const int sl = D < BM_SUB_LOG ? D : BM_SUB_LOG;     // levels handled here
    const int S = 1 << sl;
    const int t = threadIdx.x;
    if (t < S) {
        bm_f4 lo, hi;
        bm_leaf_box(m, bn, s.y + t, lo, hi);
        m.node[2 * (bn.x + P - 1 + s.y + t)] = lo;
        m.node[2 * (bn.x + P - 1 + s.y + t) + 1] = hi;
        slo[t] = lo;
        shi[t] = hi;
    }
    __syncthreads();
    for (int l = 1; l <= sl; ++l) {
        const int n = S >> l;
        bm_f4 a{}, b{};
        if (t < n) {
            a = slo[2 * t];
            b = shi[2 * t];
            bm_box_union(a, b, slo[2 * t + 1], shi[2 * t + 1]);
        }
        __syncthreads();
        if (t < n) {
            slo[t] = a;
            shi[t] = b;
            const int k = bm_level_node(D, l, s.y, t);
            m.node[2 * (bn.x + k)] = a;
            m.node[2 * (bn.x + k) + 1] = b;
        }
        __syncthreads();
    }
}

// one workgroup of 1024 threads per bone: the levels above the k_bm_refit_low
// subtree roots (at most BM_TOP_MAX of them, so at most 1024 nodes per level)
__global__ __launch_bounds__(1024) void k_bm_refit_top(BoneMeshDev m) {
    __shared__ bm_f4 slo[BM_TOP_MAX], shi[BM_TOP_MAX];
    const bm_i4 bn = m.bone[blockIdx.x];
    const int D = bm_depth(bn.y);
    if (D <= BM_SUB_LOG) return;                 // the whole tree was one subtree
    const int d0 = D - BM_SUB_LOG;               // depth of the subtree roots
    const int n0 = 1 << d0;
    for (int i = threadIdx.x; i < n0; i += blockDim.x) {
        const int k = n0 - 1 + i;
        slo[i] = m.node[2 * (bn.x + k)];
        shi[i] = m.node[2 * (bn.x + k) + 1];
    }
    __syncthreads();
    const int i = threadIdx.x;
    for (int d = d0 - 1; d >= 0; --d) {
        const int n = 1 << d;
        bm_f4 a{}, b{};
        if (i < n) {
            a = slo[2 * i];
            b = shi[2 * i];
            bm_box_union(a, b, slo[2 * i + 1], shi[2 * i + 1]);
            const int k = n - 1 + i;
            m.node[2 * (bn.x + k)] = a;
            m.node[2 * (bn.x + k) + 1] = b;
        }
        __syncthreads();
        if (i < n) {
            slo[i] = a;
            shi[i] = b;
        }
        __syncthreads();
    }
}

__global__ void k_bm_cache_reset(BoneMeshDev m, int nv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nv) m.cache[i] = bm_f4{0.f, 0.f, 0.f, -INFINITY};
}

// MODE 0: push X in place (collision stage); MODE 1: penetration depth per vertex
// into depth, and the count of depths > eps and the largest depth into stat
template <int MODE>
__global__ __launch_bounds__(256) void k_bm_query(int nv, float4 *X, BoneMeshDev m, float *depth, float eps,
                                                  uint32_t *stat) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float dep = 0.f;
    __shared__ uint32_t sq;
    if (threadIdx.x == 0) sq = 0u;
    __syncthreads();
    if (i < nv) {
        const bool on = MODE == 0 ? m.cw[i] != 0.f : (m.mw == nullptr || m.mw[i] != 0.f);
        if (on) {
            const float4 x = X[i];
            bm_v3 p = bm_mk(x.x, x.y, x.z);
            if (!(m.coherent && bm_certified_outside(m, i, p, MODE == 0 ? m.contact_offset : 0.f))) {
                atomicAdd(&sq, 1u);
                BmHit h;
                float sg = 1.f;
                const bool found = bm_query(m, m.vgroup[i], p, h, sg);
                if (m.coherent) bm_certify(m, i, p, found, h, sg);
                if (found) {
                    if (MODE == 0) {
                        if (bm_push_from(m, h, sg, p)) X[i] = make_float4(p.x, p.y, p.z, x.w);
                    } else {
                        dep = bm_depth_from(h, sg, p);
                    }
                }
            }
        }
        if (MODE == 1) depth[i] = dep;
    }
    __syncthreads();
    if (threadIdx.x == 0 && sq) atomicAdd(m.nquery, (unsigned long long)sq);
    if (MODE == 1) {
        // count and maximum over the workgroup, then one atomic each (order-free results)
        __shared__ uint32_t scnt, smax;
        if (threadIdx.x == 0) { scnt = 0u; smax = 0u; }
        __syncthreads();
        if (dep > eps) atomicAdd(&scnt, 1u);
        if (dep > 0.f) atomicMax(&smax, __float_as_uint(dep));
        __syncthreads();
        if (threadIdx.x == 0) {
            if (scnt) atomicAdd(&stat[0], scnt);
            if (smax) atomicMax(&stat[1], smax);
        }
    }
}
#endif  // __HIPCC__

}  // namespace xpbd
