// Host check of the bone-mesh BVH: builds the tables and refits the tree with
// the library's own per-thread steps (xpbd_bonemesh.h, plain C++), checks the
// box nesting, then answers closest-point queries and pushes for the Python test
// to compare with the oracle's brute force.
//
// in:  int32 nbv, nt, nb, nq; float32 verts[nbv*3]; int32 tris[nt*3];
//      int32 tri_off[nb+1]; float32 q[nq*3]; float32 max_dist, contact_offset, max_push
// out: float32 sd[nq]; float32 cp[nq*3]; int32 face[nq] (caller's id, -1 none);
//      int32 region[nq]; float32 pushed[nq*3]; float32 depth[nq]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "xpbd_bonemesh.h"

using namespace xpbd;

template <typename T>
static void rd(FILE *f, T *p, size_t n) {
    if (fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}

static bool inside(const bm_f4 &lo, const bm_f4 &hi, const bm_f4 &lo2, const bm_f4 &hi2) {
    if (lo2.x > hi2.x) return true;   // empty child
    return lo.x <= lo2.x && lo.y <= lo2.y && lo.z <= lo2.z && hi.x >= hi2.x && hi.y >= hi2.y && hi.z >= hi2.z;
}

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t h[4];
    rd(f, h, 4);
    const int nbv = h[0], nt = h[1], nb = h[2], nq = h[3];
    std::vector<float> verts(3 * (size_t)nbv), q(3 * (size_t)nq);
    std::vector<int32_t> tris(3 * (size_t)nt), toff(nb + 1);
    float par[3];
    rd(f, verts.data(), verts.size());
    rd(f, tris.data(), tris.size());
    rd(f, toff.data(), toff.size());
    rd(f, q.data(), q.size());
    rd(f, par, 3);
    fclose(f);

    BmTables T;
    std::string e = bm_build_tables(nbv, verts.data(), nt, tris.data(), nb, toff.data(), T);
    if (!e.empty()) { fprintf(stderr, "%s\n", e.c_str()); return 1; }
    std::vector<bm_f4> pts(nbv), soup(3 * (size_t)nt), fn(nt), vn(nbv), node(2 * (size_t)T.nnodes);
    for (int i = 0; i < nbv; ++i) pts[i] = bm_f4{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2], 0.f};
    std::vector<int32_t> goff = {0, nb}, vg(1, 0);
    BoneMeshDev m{};
    m.nbones = nb; m.ntris = nt; m.nbv = nbv;
    m.bone = T.bone.data(); m.tv = T.tv.data(); m.adj = T.adj.data(); m.sub = T.sub.data();
    m.vt_off = T.vt_off.data(); m.vt = T.vt.data(); m.pts = pts.data(); m.soup = soup.data();
    m.fn = fn.data(); m.vn = vn.data(); m.node = node.data(); m.vgroup = vg.data(); m.goff = goff.data();
    m.max_dist = par[0]; m.contact_offset = par[1]; m.max_push = par[2];
    bm_refit_host(m);

    // every triangle sits in exactly one leaf of its bone, inside the leaf box; boxes nest
    int bad = 0;
    std::vector<int> seen(nt, 0);
    for (int b = 0; b < nb; ++b) {
        const bm_i4 bn = T.bone[b];
        const int P = bn.y;
        if (P & (P - 1)) ++bad;
        for (int k = 0; k < P - 1; ++k)
            for (int ch = 1; ch <= 2; ++ch)
                if (!inside(node[2 * (bn.x + k)], node[2 * (bn.x + k) + 1], node[2 * (bn.x + 2 * k + ch)],
                            node[2 * (bn.x + 2 * k + ch) + 1])) ++bad;
        for (int leaf = 0; leaf < P; ++leaf) {
            const int t0 = leaf * BM_LEAF, t1 = std::min(t0 + BM_LEAF, bn.w);
            for (int j = t0; j < t1; ++j) {
                const int t = bn.z + j;
                ++seen[T.tv[t].w];
                for (int c = 0; c < 3; ++c) {
                    const bm_f4 p = soup[3 * t + c];
                    if (!inside(node[2 * (bn.x + P - 1 + leaf)], node[2 * (bn.x + P - 1 + leaf) + 1], p, p)) ++bad;
                }
                if (T.tv[t].w < toff[b] || T.tv[t].w >= toff[b + 1]) ++bad;
            }
        }
    }
    for (int t = 0; t < nt; ++t) bad += seen[t] != 1;
    if (bad) { fprintf(stderr, "tree invariants broken: %d\n", bad); return 1; }

    std::vector<float> sd(nq, 0.f), cp(3 * (size_t)nq, 0.f), pushed(q), depth(nq, 0.f);
    std::vector<int32_t> face(nq, -1), region(nq, -1);
    for (int i = 0; i < nq; ++i) {
        const bm_v3 p = bm_mk(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
        BmHit hh;
        float sg;
        if (bm_query(m, 0, p, hh, sg)) {
            const bm_v3 d = p - hh.cp;
            sd[i] = sqrtf(bm_dot(d, d)) * sg;
            cp[3 * i] = hh.cp.x; cp[3 * i + 1] = hh.cp.y; cp[3 * i + 2] = hh.cp.z;
            face[i] = T.tv[hh.face].w;
            region[i] = hh.region;
        }
        bm_v3 y = p;
        bm_push_point(m, 0, y);
        pushed[3 * i] = y.x; pushed[3 * i + 1] = y.y; pushed[3 * i + 2] = y.z;
        depth[i] = bm_depth_point(m, 0, p);
    }
    FILE *o = fopen(argv[2], "wb");
    if (!o) return 2;
    fwrite(sd.data(), 4, sd.size(), o);
    fwrite(cp.data(), 4, cp.size(), o);
    fwrite(face.data(), 4, face.size(), o);
    fwrite(region.data(), 4, region.size(), o);
    fwrite(pushed.data(), 4, pushed.size(), o);
    fwrite(depth.data(), 4, depth.size(), o);
    fclose(o);
    printf("ok %d nodes closed %d\n", T.nnodes, T.closed ? 1 : 0);
    return 0;
}
