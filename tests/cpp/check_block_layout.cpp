// Host-only check of the block layout (rlmusclev2_amd/csrc/xpbd_layout.cpp) on a jittered lattice:
// every element corner is reduced exactly once per block in ascending slot order, every partial slot
// is written once, and the emulated block pass + shared-vertex pass give the same per-vertex sums as
// the flat vertex-major corner order (double precision, so only the summation order differs).
#include "xpbd_layout.h"
#include "xpbd.h"
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <vector>
#include <random>
using namespace xpbd;
int main(int argc, char **argv) {
    int n = argc > 1 ? atoi(argv[1]) : 12, TB = argc > 2 ? atoi(argv[2]) : 256;
    int nx = n, ny = n / 2 + 1, nz = n / 3 + 1;
    auto vid = [&](int i, int j, int k) { return (i * (ny + 1) + j) * (nz + 1) + k; };
    std::vector<float> pos;
    std::mt19937 rng(1);
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
    const size_t ne = L.meta.size();
    int bad = 0;
    std::vector<double> val(4 * ne);
    for (auto &x : val) x = std::uniform_real_distribution<double>(-1, 1)(rng);
    // flat reference sums
    std::vector<double> ref(nv, 0.0), got(nv, 0.0);
    std::vector<int> written(nv, 0);
    for (size_t g = 0; g < ne; ++g) for (int k = 0; k < 4; ++k) ref[L.gidx[4 * g + k]] += val[4 * g + k];
    std::vector<double> part(L.n_partial, 0.0);
    std::vector<int> pw(L.n_partial, 0);
    for (size_t b = 0; b < L.blocks.size(); ++b) {
        const BlockDesc &D = L.blocks[b];
        if (D.ntet <= 0 || D.ntet > TB || D.nloc > L.max_nloc) { printf("bad block %zu\n", b); ++bad; }
        std::vector<int> seen(4 * D.ntet, 0);
        for (int i = 0; i < D.nloc; ++i) {
            uint32_t run = L.brun[D.vbeg + i]; int s0 = run & 0xffff, cnt = run >> 16;
            int v = L.bvt[D.vbeg + i];
            double sum = 0;
            int prev = -1;
            for (int j = 0; j < cnt; ++j) {
                if (s0 + j >= 4 * D.ntet) { printf("slot range\n"); ++bad; break; }
                const int code = L.bslot[4 * (size_t)D.tbeg + s0 + j];   // corner * block_tets + element
                const int el = code % L.block_tets, k = code / L.block_tets;
                const int sl = 4 * el + k;
                if (k > 3 || el >= D.ntet || sl <= prev) { printf("slot order/range\n"); ++bad; }
                prev = sl;
                seen[sl]++;
                int g = D.tbeg + el;
                // the corner's LDS position is its place in this vertex-major list
                if (L.bcpos[4 * (size_t)g + k] != (uint16_t)(s0 + j)) { printf("bcpos mismatch\n"); ++bad; }
                if (L.gidx[4 * g + k] != v || L.blidx[4 * g + k] != i) { printf("slot vertex mismatch\n"); ++bad; }
                sum += val[4 * g + k];
            }
            int ps = L.bps[D.vbeg + i];
            if (ps >= 0) { if (ps >= L.n_partial) { ++bad; continue; } part[ps] = sum; pw[ps]++; }
            else { got[v] = sum; written[v]++; }
        }
        for (int s = 0; s < 4 * D.ntet; ++s) if (seen[s] != 1) { printf("corner %d of block %zu seen %d\n", s, b, seen[s]); ++bad; }
    }
    for (int p = 0; p < L.n_partial; ++p) if (pw[p] != 1) { printf("partial %d written %d\n", p, pw[p]); ++bad; }
    for (size_t i = 0; i < L.svid.size(); ++i) {
        int v = L.svid[i]; double sum = 0;
        for (int s = L.soff[i]; s < L.soff[i + 1]; ++s) sum += part[s];
        got[v] = sum; written[v]++;
    }
    double maxerr = 0;
    for (int v = 0; v < nv; ++v) {
        if (written[v] != 1) { printf("vertex %d written %d\n", v, written[v]); ++bad; }
        maxerr = std::max(maxerr, std::fabs(got[v] - ref[v]));
    }
    printf("nv %d nt %d elems %zu blocks %zu max_nloc %d shared %zu partials %d  maxerr %.3g  bad %d\n", nv, nt, ne,
           L.blocks.size(), L.max_nloc, L.svid.size(), L.n_partial, maxerr, bad);
    return bad ? 1 : 0;
}
