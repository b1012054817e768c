// Host-only check of the resident parts (build_parts in rlmusclev2_amd/csrc/xpbd_layout.cpp) on a
// jittered lattice: the parts cover every block once, every vertex is written back by exactly one part,
// the neighbour lists are symmetric and hold every part that shares a vertex, the part-local corner ids
// name the element's vertices, and the resident scheme's sums (per-round LDS accumulation, block partials
// of cross vertices added in slot order) equal the block strategy's bit for bit in FP32.
#include "xpbd_layout.h"
#include "xpbd.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <set>
#include <vector>
#include <random>
using namespace xpbd;
int main(int argc, char **argv) {
    int n = argc > 1 ? atoi(argv[1]) : 12, TB = argc > 2 ? atoi(argv[2]) : 256, R = argc > 3 ? atoi(argv[3]) : 3;
    const int target = argc > 4 ? atoi(argv[4]) : 0;      // > 0: that many parts, blocks spread evenly
    int nx = n, ny = n / 2 + 1, nz = n / 3 + 1;
    auto vid = [&](int i, int j, int k) { return (i * (ny + 1) + j) * (nz + 1) + k; };
    std::vector<float> pos;
    std::mt19937 rng(2);
    std::uniform_real_distribution<float> U(-0.1f, 0.1f);
    for (int i = 0; i <= nx; ++i) for (int j = 0; j <= ny; ++j) for (int k = 0; k <= nz; ++k) {
        pos.push_back(i + U(rng)); pos.push_back(j + U(rng)); pos.push_back(k + U(rng)); }
    int nv = (int)pos.size() / 3;
    nv += 3;  // three isolated vertices (touched by no tet)
    for (int q = 0; q < 9; ++q) pos.push_back(100.f + q);
    std::vector<int> tets;
    const int T[6][4] = {{0,1,3,7},{0,1,5,7},{0,2,3,7},{0,2,6,7},{0,4,5,7},{0,4,6,7}};
    for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j) for (int k = 0; k < nz; ++k) {
        int c[8]; for (int b = 0; b < 8; ++b) c[b] = vid(i + (b >> 2 & 1), j + (b >> 1 & 1), k + (b & 1));
        for (auto &t : T) for (int q = 0; q < 4; ++q) tets.push_back(c[t[q]]); }
    int nt = (int)tets.size() / 4;
    std::vector<float> rest(9 * nt, 0.f), rv(nt), mass(nv);
    precompute_rest(nv, pos.data(), nt, tets.data(), 1000.f, rv.data(), rest.data(), mass.data());
    int nc = 2 * nt;
    std::vector<int> type(nc), pts(4 * nc), tetid(nc);
    std::vector<float> st(nc, 1e4f), cs(nc, -1.f), dr(nc, 0.f), rl(nc), rvec(4 * nc, 0.f), rdir(3 * nc, 0.f);
    for (int t = 0; t < nt; ++t) for (int h = 0; h < 2; ++h) {
        int c = 2 * t + h; type[c] = h ? XPBD_CT_TETVOLUME : XPBD_CT_TETARAPSVD; tetid[c] = t;
        for (int q = 0; q < 4; ++q) pts[4 * c + q] = tets[4 * t + q]; rl[c] = rv[t]; }
    HostInput in; in.nv = nv; in.nt = nt; in.pos0 = pos.data(); in.tets = tets.data(); in.rest_matrix = rest.data();
    in.ncons = nc; in.type = type.data(); in.pts = pts.data(); in.tetid = tetid.data(); in.stiffness = st.data();
    in.compressionstiffness = cs.data(); in.dampingratio = dr.data(); in.restlength = rl.data();
    in.restvector = rvec.data(); in.restdir = rdir.data(); in.flat = true; in.block_tets = TB;
    Layout L; int code = 0;
    std::string err = build_layout(in, L, &code);
    if (code) { printf("layout error %s\n", err.c_str()); return 1; }
    Parts P;
    err = build_parts(L, nv, R, P, &code, target);
    if (code) { printf("parts error %s\n", err.c_str()); return 1; }
    if (target > 0) {
        R = P.R;
        if ((int64_t)P.parts.size() != std::min<int64_t>(target, (int64_t)L.blocks.size())) { printf("part count\n"); return 1; }
    }
    const size_t ne = L.meta.size();
    const int G = (int)P.parts.size();
    int bad = 0;
    // every block in exactly one part, in order
    int nextb = 0;
    for (int p = 0; p < G; ++p) {
        const PartDesc &pd = P.parts[p];
        if (pd.b0 != nextb || pd.nblk < 1 || pd.nblk > R || pd.nloc > P.max_nloc) { printf("bad part %d\n", p); ++bad; }
        nextb = pd.b0 + pd.nblk;
    }
    if (nextb != (int)L.blocks.size()) { printf("parts cover %d of %zu blocks\n", nextb, L.blocks.size()); ++bad; }
    // corner ids, write-back flags, neighbour lists
    std::vector<int> written(nv, 0);
    std::vector<std::set<int>> touch(nv);
    for (int p = 0; p < G; ++p) {
        const PartDesc &pd = P.parts[p];
        for (int rr = 0; rr < pd.nblk; ++rr) {
            const BlockDesc &d = L.blocks[pd.b0 + rr];
            for (int g = d.tbeg; g < d.tbeg + d.ntet; ++g)
                for (int k = 0; k < 4; ++k) {
                    const int j = P.plidx[4 * (size_t)g + k];
                    if (j >= pd.nloc || P.prec[4 * (pd.vbeg + j)] != L.gidx[4 * (size_t)g + k]) { printf("plidx\n"); ++bad; }
                    touch[L.gidx[4 * (size_t)g + k]].insert(p);
                }
            for (int i = 0; i < d.nloc; ++i) {
                const uint32_t m = P.bmap[d.vbeg + i];
                const int j = (int)(m & 0x7fffffffu);
                if (j >= pd.nloc || P.prec[4 * (pd.vbeg + j)] != L.bvt[d.vbeg + i]) { printf("bmap\n"); ++bad; }
                if (((m >> 31) != 0) != (P.prec[4 * (pd.vbeg + j) + 1] >= 0)) { printf("bmap cross flag\n"); ++bad; }
            }
        }
        for (int j = 0; j < pd.nloc; ++j)
            if (P.prec[4 * (pd.vbeg + j) + 2]) written[P.prec[4 * (pd.vbeg + j)]]++;
    }
    for (int v = 0; v < nv; ++v)
        if (written[v] != 1) { printf("vertex %d written back %d times\n", v, written[v]); ++bad; }
    for (int p = 0; p < G; ++p) {
        const PartDesc &pd = P.parts[p];
        std::set<int> want;
        for (int j = 0; j < pd.nloc; ++j) {
            const int v = P.prec[4 * (pd.vbeg + j)];
            const bool cross = touch[v].size() > 1;
            if (cross != (P.prec[4 * (pd.vbeg + j) + 1] >= 0)) { printf("cross flag of %d\n", v); ++bad; }
            if (cross) for (int q : touch[v]) if (q != p) want.insert(q);
        }
        std::set<int> got(P.nbr.begin() + pd.nbeg, P.nbr.begin() + pd.nbeg + pd.nnb);
        if (got != want || (int)got.size() != pd.nnb) { printf("neighbours of part %d\n", p); ++bad; }
        for (int q : got) {
            const PartDesc &qd = P.parts[q];
            bool back = false;
            for (int k = 0; k < qd.nnb; ++k) back |= P.nbr[qd.nbeg + k] == p;
            if (!back) { printf("asymmetric neighbours %d %d\n", p, q); ++bad; }
        }
    }
    // FP32 sums: block strategy vs resident scheme
    std::vector<float> val(4 * ne);
    for (auto &x : val) x = std::uniform_real_distribution<float>(-1, 1)(rng);
    auto run_sum = [&](const BlockDesc &d, int i) {
        const uint32_t run = L.brun[d.vbeg + i];
        const int s0 = run & 0xffff, cnt = run >> 16;
        float s = 0.f;
        for (int j = 0; j < cnt; ++j) {
            const int c = L.bslot[4 * (size_t)d.tbeg + s0 + j];
            s += val[4 * ((size_t)d.tbeg + c % L.block_tets) + c / L.block_tets];
        }
        return s;
    };
    std::vector<float> blk(nv, 0.f), res(nv, 0.f), part(L.n_partial, 0.f), xpart(L.n_partial, 0.f);
    for (const BlockDesc &d : L.blocks)
        for (int i = 0; i < d.nloc; ++i) {
            const float s = run_sum(d, i);
            if (L.bps[d.vbeg + i] >= 0) part[L.bps[d.vbeg + i]] = s; else blk[L.bvt[d.vbeg + i]] = s;
        }
    for (size_t s = 0; s < L.svid.size(); ++s) {
        float a = 0.f;
        for (int k = L.soff[s]; k < L.soff[s + 1]; ++k) a += part[k];
        blk[L.svid[s]] = a;
    }
    std::vector<float> owner_val(nv, NAN);
    // exchange: every part writes the block partials of its cross vertices
    for (int p = 0; p < G; ++p) {
        const PartDesc &pd = P.parts[p];
        for (int rr = 0; rr < pd.nblk; ++rr) {
            const BlockDesc &d = L.blocks[pd.b0 + rr];
            for (int i = 0; i < d.nloc; ++i)
                if (P.bmap[d.vbeg + i] >> 31) xpart[L.bps[d.vbeg + i]] = run_sum(d, i);
        }
    }
    for (int p = 0; p < G; ++p) {
        const PartDesc &pd = P.parts[p];
        std::vector<float> acc(pd.nloc, 0.f);
        for (int rr = 0; rr < pd.nblk; ++rr) {
            const BlockDesc &d = L.blocks[pd.b0 + rr];
            for (int i = 0; i < d.nloc; ++i) {
                const uint32_t m = P.bmap[d.vbeg + i];
                if (!(m >> 31)) acc[m & 0x7fffffffu] += run_sum(d, i);
            }
        }
        for (int j = 0; j < pd.nloc; ++j) {
            const int *r = &P.prec[4 * (pd.vbeg + j)];
            float a = acc[j];
            if (r[1] >= 0) {
                a = 0.f;
                for (int k = L.soff[r[1]]; k < L.soff[r[1] + 1]; ++k) a += xpart[k];
                // every part that holds the vertex computes the same value
                if (!std::isnan(owner_val[r[0]]) && std::memcmp(&owner_val[r[0]], &a, 4)) { printf("sharers differ\n"); ++bad; }
                owner_val[r[0]] = a;
            }
            if (r[2]) res[r[0]] = a;
        }
    }
    int diff = 0;
    for (int v = 0; v < nv; ++v) diff += std::memcmp(&blk[v], &res[v], 4) != 0;
    if (diff) { printf("%d vertices differ from the block strategy\n", diff); ++bad; }
    printf("nv %d elems %zu blocks %zu parts %d R %d max_nloc %d max_nnb %d cross %d  bad %d\n", nv, ne,
           L.blocks.size(), G, R, P.max_nloc, P.max_nnb, P.n_cross, bad);
    return bad ? 1 : 0;
}
