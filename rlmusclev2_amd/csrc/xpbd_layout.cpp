// Host-side layout builder for the patch-owned Jacobi projection.
#include "xpbd_layout.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

#pragma clang fp contract(off)

namespace xpbd {

namespace {

constexpr int32_t CT_ATTACH = 1650556;
constexpr int32_t CT_PINTOTARGET = 771624311;
constexpr int32_t CT_TETVOLUME = -215389979;
constexpr int32_t CT_TETARAPSVD = -92199132;
constexpr int32_t CT_TETARAPANISO = -92199133;
constexpr int32_t CT_TETSNH = -503462311;

inline uint64_t spread21(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

struct Elem {
    int32_t p[4];
    int32_t tetid;
    uint32_t mask = 0;
    ClassParams cp{};
    float rvol = 0.f, rsnh = 0.f;
    float f0[3] = {0, 0, 0};
    float belly = 0.f;
};

struct KeyHash {
    size_t operator()(const std::array<int32_t, 5> &k) const {
        uint64_t h = 1469598103934665603ULL;
        for (int32_t v : k) { h ^= (uint32_t)v; h *= 1099511628211ULL; }
        return (size_t)h;
    }
};

}  // namespace

void precompute_rest(int nv, const float *pos0, int nt, const int32_t *tets, float density,
                     float *rest_volume, float *rest_matrix, float *mass) {
    std::fill(mass, mass + nv, 0.0f);
    for (int c = 0; c < nt; ++c) {
        const int32_t *t = tets + 4 * c;
        const float *p3 = pos0 + 3 * t[3];
        float col[3][3];
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) col[k][a] = pos0[3 * t[k] + a] - p3[a];
        // Dm row-major: Dm[a][k] = col[k][a]
        float m[9] = {col[0][0], col[1][0], col[2][0], col[0][1], col[1][1],
                      col[2][1], col[0][2], col[1][2], col[2][2]};
        float d = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                  m[2] * (m[3] * m[7] - m[4] * m[6]);
        float vol = std::fabs(d) / 6.0f;
        rest_volume[c] = vol;
        float mc = vol * density / 4.0f;
        for (int i = 0; i < 4; ++i) mass[t[i]] += mc;
        float *o = rest_matrix + 9 * c;
        if (d == 0.0f) {
            std::memset(o, 0, 9 * sizeof(float));
            continue;
        }
        float inv = 1.0f / d;
        o[0] = (m[4] * m[8] - m[5] * m[7]) * inv;
        o[1] = (m[2] * m[7] - m[1] * m[8]) * inv;
        o[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
        o[3] = (m[5] * m[6] - m[3] * m[8]) * inv;
        o[4] = (m[0] * m[8] - m[2] * m[6]) * inv;
        o[5] = (m[2] * m[3] - m[0] * m[5]) * inv;
        o[6] = (m[3] * m[7] - m[4] * m[6]) * inv;
        o[7] = (m[1] * m[6] - m[0] * m[7]) * inv;
        o[8] = (m[0] * m[4] - m[1] * m[3]) * inv;
    }
}


namespace {
// attach / pin records of the host arrays, renumbered to the new vertex order
void renumber_vertex_records(Layout &L, const std::vector<int32_t> &attach_old, int nv) {
    std::vector<float> a0(L.att0.size()), a1(L.att1.size()), pc(L.pin_from_cons.size());
    for (int v = 0; v < nv; ++v) {
        int n = L.iperm[v];
        std::memcpy(&a0[4 * (size_t)n], &L.att0[4 * (size_t)v], 16);
        std::memcpy(&a1[4 * (size_t)n], &L.att1[4 * (size_t)v], 16);
        std::memcpy(&pc[4 * (size_t)n], &L.pin_from_cons[4 * (size_t)v], 16);
    }
    L.att0.swap(a0);
    L.att1.swap(a1);
    L.pin_from_cons.swap(pc);
    L.attach_vertex.resize(attach_old.size());
    for (size_t i = 0; i < attach_old.size(); ++i) L.attach_vertex[i] = L.iperm[attach_old[i]];
}

double elem_static_bytes(int feature) {
    double b = 4.0;  // meta
    if (feature & (F_ARAP | F_SNH | F_ANISO)) b += 36.0;
    if (feature & F_VOL) b += 4.0;
    if (feature & F_SNH) b += 4.0;
    if (feature & F_ANISO) b += 16.0;
    return b;
}

double alg_bytes(int feature, double n_elems, double nv) {
    double btet = 16.0;
    if (feature & (F_ARAP | F_SNH | F_ANISO)) btet += 36.0;
    if (feature & (F_VOL | F_SNH)) btet += 4.0;
    if (feature & F_ANISO) btet += 20.0;
    return btet * n_elems + 28.0 * nv;
}
}  // namespace

std::string build_flat(const HostInput &in, Layout &L, const std::vector<Elem> &elems,
                       const std::vector<uint32_t> &ecls, const std::vector<int32_t> &eord, int feature,
                       int *code) {
    const int nv = in.nv;
    const size_t ne = eord.size();
    L.patches.clear();
    L.halo.clear();
    L.lidx.clear();
    L.gidx.assign(4 * ne, 0);
    L.meta.assign(ne, 0);
    L.Bp.assign(9 * ne, 0.f);
    L.rvol.assign(ne, 0.f);
    L.rsnh.assign(ne, 0.f);
    L.fib.assign(4 * ne, 0.f);
    L.ptet_tetid.assign(ne, 0);
    L.vslot_off.assign(nv + 1, 0);
    for (size_t g = 0; g < ne; ++g) {
        const Elem &e = elems[eord[g]];
        for (int k = 0; k < 4; ++k) {
            const int v = L.iperm[e.p[k]];
            L.gidx[4 * g + k] = v;
            L.vslot_off[v + 1]++;
        }
        L.meta[g] = ecls[eord[g]] | (e.mask << 16) | (1u << 20);
        const float *B = in.rest_matrix + 9 * (size_t)e.tetid;
        for (int q = 0; q < 9; ++q) L.Bp[(size_t)q * ne + g] = B[q];
        L.rvol[g] = e.rvol;
        L.rsnh[g] = e.rsnh;
        L.fib[4 * g + 0] = e.f0[0];
        L.fib[4 * g + 1] = e.f0[1];
        L.fib[4 * g + 2] = e.f0[2];
        L.fib[4 * g + 3] = e.belly;
        L.ptet_tetid[g] = e.tetid;
    }
    for (int v = 0; v < nv; ++v) L.vslot_off[v + 1] += L.vslot_off[v];
    // corner (g, k) is written straight to its place in the vertex-major corner array,
    // so each vertex later reads its corners as one contiguous run (ascending g)
    L.vslot32.assign(4 * ne, 0);
    std::vector<int32_t> fill(L.vslot_off.begin(), L.vslot_off.end() - 1);
    for (size_t g = 0; g < ne; ++g)
        for (int k = 0; k < 4; ++k) L.vslot32[4 * g + k] = fill[L.gidx[4 * g + k]]++;
    L.max_table = 0;
    L.max_ptets = (int)ne;
    L.bytes_alg = alg_bytes(feature, (double)ne, nv);
    const double damp = (feature & F_DAMP) ? 2.0 : 1.0;
    // tets: static data + index + 4 vertex gathers + 4 corner writes; vertices: offsets,
    // corner slots + gathers, position read and write
    L.bytes_impl = (elem_static_bytes(feature) + 32.0 + 64.0 * damp + 64.0) * (double)ne +
                   (4.0 + 16.0 * damp + 16.0) * nv + 16.0 * (double)L.vslot32.size();
    if (feature & F_PINS) L.bytes_impl += 16.0 * nv;
    if (feature & F_ATTACH) L.bytes_impl += 32.0 * nv;
    (void)code;
    return std::string();
}

std::string build_blocks(const HostInput &in, Layout &L, int feature, int *code) {
    const int nv = in.nv;
    const int TB = in.block_tets;
    if (TB < 64 || TB > 8192) { *code = 1; return "block_tets must be in [64, 8192]"; }
    const int64_t ne = (int64_t)L.meta.size();
    const int64_t nb = (ne + TB - 1) / TB;
    L.block_tets = TB;
    L.blocks.assign((size_t)nb, BlockDesc{});
    L.bvt.clear();
    L.bps.clear();
    L.brun.clear();
    L.blidx.assign(4 * (size_t)ne, 0);
    L.bslot.assign(4 * (size_t)ne, 0);
    L.bcpos.assign(4 * (size_t)ne, 0);
    std::vector<int32_t> nblk(nv, 0);
    std::vector<int32_t> loc(nv, -1);
    std::vector<int32_t> verts;
    int max_nloc = 0;
    // pass 1: local tables, local corner ids, per-vertex block counts
    for (int64_t b = 0; b < nb; ++b) {
        const int64_t t0 = b * TB, t1 = std::min<int64_t>(ne, t0 + TB);
        verts.clear();
        for (int64_t g = t0; g < t1; ++g)
            for (int k = 0; k < 4; ++k) verts.push_back(L.gidx[4 * g + k]);
        std::sort(verts.begin(), verts.end());
        verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
        if (verts.size() > 65535) { *code = 1; return "block vertex table exceeds 65535 entries"; }
        BlockDesc &d = L.blocks[(size_t)b];
        d.tbeg = (int32_t)t0;
        d.ntet = (int32_t)(t1 - t0);
        d.vbeg = (int32_t)L.bvt.size();
        d.nloc = (int32_t)verts.size();
        max_nloc = std::max(max_nloc, d.nloc);
        for (size_t i = 0; i < verts.size(); ++i) {
            loc[verts[i]] = (int32_t)i;
            nblk[verts[i]]++;
            L.bvt.push_back(verts[i]);
        }
        // corner slots of each local entry in ascending slot order (counting sort)
        std::vector<uint32_t> cnt(verts.size() + 1, 0);
        for (int64_t g = t0; g < t1; ++g)
            for (int k = 0; k < 4; ++k) {
                const int l = loc[L.gidx[4 * g + k]];
                L.blidx[4 * g + k] = (uint16_t)l;
                cnt[l + 1]++;
            }
        for (size_t i = 0; i < verts.size(); ++i) cnt[i + 1] += cnt[i];
        for (size_t i = 0; i < verts.size(); ++i) L.brun.push_back(cnt[i] | ((cnt[i + 1] - cnt[i]) << 16));
        std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
        for (int64_t g = t0; g < t1; ++g)
            for (int k = 0; k < 4; ++k) {
                const int l = L.blidx[4 * g + k];
                // listed in ascending (element, corner) order; the value is the
                // corner-major index k * TB + element; the corner's position in
                // this vertex-major list is where the kernel stores it in LDS
                const uint32_t pos = fill[l]++;
                L.bslot[4 * t0 + pos] = (uint16_t)(k * TB + (g - t0));
                L.bcpos[4 * g + k] = (uint16_t)pos;
            }
        for (int v : verts) loc[v] = -1;
    }
    L.max_nloc = max_nloc;
    // pass 2: vertices finished by the second pass (shared by several blocks or
    // touched by none), their partial ranges in vertex-major, block-ascending order
    L.svid.clear();
    L.soff.assign(1, 0);
    std::vector<int32_t> pbase(nv, -1);
    for (int v = 0; v < nv; ++v)
        if (nblk[v] != 1) {
            pbase[v] = L.soff.back();
            L.svid.push_back(v);
            L.soff.push_back(L.soff.back() + nblk[v]);
        }
    L.n_partial = L.soff.back();
    L.bps.assign(L.bvt.size(), -1);
    std::vector<int32_t> used(nv, 0);
    for (int64_t b = 0; b < nb; ++b) {
        const BlockDesc &d = L.blocks[(size_t)b];
        for (int i = 0; i < d.nloc; ++i) {
            const int v = L.bvt[d.vbeg + i];
            if (nblk[v] > 1) L.bps[d.vbeg + i] = pbase[v] + used[v]++;
        }
    }
    const double damp = (feature & F_DAMP) ? 2.0 : 1.0;
    // elements: static data + local corner ids + corner slots; local entries:
    // vertex id, partial slot, run, position gathers; partials written and read
    // back; per-vertex output
    L.bytes_impl = (elem_static_bytes(feature) + 8.0 + 8.0) * (double)ne +
                   (12.0 + 16.0 * damp) * (double)L.bvt.size() + 32.0 * L.n_partial +
                   (8.0 + 16.0 * damp) * (double)L.svid.size() + 16.0 * nv;
    if (feature & F_PINS) L.bytes_impl += 16.0 * nv;
    if (feature & F_ATTACH) L.bytes_impl += 32.0 * nv;
    return std::string();
}

std::string build_parts(const Layout &L, int nv, int R, Parts &P, int *code, int target_parts) {
    *code = 0;
    P = Parts();
    if (L.block_tets <= 0) { *code = 1; return "parts need the block strategy"; }
    const int64_t nb = (int64_t)L.blocks.size();
    if (target_parts > 0) {
        // as many parts as asked (at most one per block), the blocks spread evenly over them
        R = (int)std::max<int64_t>(1, (nb + target_parts - 1) / target_parts);
    }
    if (R < 1 || R > 64) { *code = 1; return "blocks per part must be in [1, 64]"; }
    const int64_t G = target_parts > 0 ? std::max<int64_t>(1, std::min<int64_t>(nb, target_parts)) : (nb + R - 1) / R;
    // part p owns blocks [first(p), first(p + 1))
    auto first = [&](int64_t p) { return target_parts > 0 ? p * nb / G : std::min<int64_t>(nb, p * R); };
    std::vector<int32_t> part_of((size_t)nb);
    for (int64_t p = 0; p < G; ++p)
        for (int64_t b = first(p); b < first(p + 1); ++b) part_of[(size_t)b] = (int32_t)p;
    P.R = R;
    P.parts.assign((size_t)G, PartDesc{});
    // the parts that touch each vertex, ascending (blocks are visited in order,
    // so a vertex's parts come in ascending order): counts, then a CSR
    std::vector<int32_t> npart(nv, 0), last(nv, -1);
    for (int64_t b = 0; b < nb; ++b) {
        const BlockDesc &d = L.blocks[(size_t)b];
        const int32_t p = part_of[(size_t)b];
        for (int i = 0; i < d.nloc; ++i) {
            const int v = L.bvt[d.vbeg + i];
            if (last[v] != p) { last[v] = p; ++npart[v]; }
        }
    }
    std::vector<int64_t> poff(nv + 1, 0);
    for (int v = 0; v < nv; ++v) poff[v + 1] = poff[v] + npart[v];
    std::vector<int32_t> plist((size_t)poff[nv]);
    {
        std::vector<int64_t> fill(poff.begin(), poff.end() - 1);
        std::fill(last.begin(), last.end(), -1);
        for (int64_t b = 0; b < nb; ++b) {
            const BlockDesc &d = L.blocks[(size_t)b];
            const int32_t p = part_of[(size_t)b];
            for (int i = 0; i < d.nloc; ++i) {
                const int v = L.bvt[d.vbeg + i];
                if (last[v] != p) { last[v] = p; plist[(size_t)fill[v]++] = p; }
            }
        }
    }
    std::vector<int32_t> sidx(nv, -1);
    for (size_t s = 0; s < L.svid.size(); ++s) sidx[L.svid[s]] = (int32_t)s;
    // vertices no element touches: finished every iteration like the others
    // (attach, pins), each by one part
    std::vector<std::vector<int32_t>> orphans((size_t)G);
    if (G > 0)
        for (int v = 0; v < nv; ++v)
            if (npart[v] == 0) orphans[(size_t)std::min<int64_t>(G - 1, (int64_t)v * G / nv)].push_back(v);
    P.bmap.assign(L.bvt.size(), 0u);
    P.plidx.assign(L.blidx.size(), 0);
    std::vector<int32_t> loc(nv, -1), verts, nb_list;
    for (int64_t p = 0; p < G; ++p) {
        PartDesc &pd = P.parts[(size_t)p];
        pd.b0 = (int32_t)first(p);
        pd.nblk = (int32_t)(first(p + 1) - first(p));
        verts.clear();
        for (int rr = 0; rr < pd.nblk; ++rr) {
            const BlockDesc &d = L.blocks[(size_t)(pd.b0 + rr)];
            for (int i = 0; i < d.nloc; ++i) verts.push_back(L.bvt[d.vbeg + i]);
        }
        std::sort(verts.begin(), verts.end());
        verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
        for (int v : orphans[(size_t)p]) verts.push_back(v);
        if (verts.size() > 65535) { *code = 5; return "part vertex table exceeds 65535 entries"; }
        pd.vbeg = (int32_t)(P.prec.size() / 4);
        pd.nloc = (int32_t)verts.size();
        P.max_nloc = std::max(P.max_nloc, pd.nloc);
        nb_list.clear();
        for (size_t j = 0; j < verts.size(); ++j) {
            const int v = verts[j];
            loc[v] = (int32_t)j;
            const bool cross = npart[v] > 1;
            P.prec.push_back(v);
            P.prec.push_back(cross ? sidx[v] : -1);
            P.prec.push_back(npart[v] == 0 || plist[(size_t)poff[v]] == (int32_t)p ? 1 : 0);
            P.prec.push_back(0);
            if (cross) {
                ++P.n_cross;
                for (int64_t q = poff[v]; q < poff[v + 1]; ++q)
                    if (plist[(size_t)q] != (int32_t)p) nb_list.push_back(plist[(size_t)q]);
            }
        }
        for (int rr = 0; rr < pd.nblk; ++rr) {
            const BlockDesc &d = L.blocks[(size_t)(pd.b0 + rr)];
            for (int i = 0; i < d.nloc; ++i) {
                const int v = L.bvt[d.vbeg + i];
                P.bmap[(size_t)d.vbeg + i] = (uint32_t)loc[v] | (npart[v] > 1 ? 0x80000000u : 0u);
            }
            for (int64_t g = d.tbeg; g < (int64_t)d.tbeg + d.ntet; ++g)
                for (int k = 0; k < 4; ++k) P.plidx[(size_t)(4 * g + k)] = (uint16_t)loc[L.gidx[(size_t)(4 * g + k)]];
        }
        std::sort(nb_list.begin(), nb_list.end());
        nb_list.erase(std::unique(nb_list.begin(), nb_list.end()), nb_list.end());
        pd.nbeg = (int32_t)P.nbr.size();
        pd.nnb = (int32_t)nb_list.size();
        P.max_nnb = std::max(P.max_nnb, pd.nnb);
        P.nbr.insert(P.nbr.end(), nb_list.begin(), nb_list.end());
        for (int v : verts) loc[v] = -1;
    }
    return std::string();
}

std::string build_layout(const HostInput &in, Layout &L, int *code) {
    *code = 0;
    const int nv = in.nv;
    const int P = in.patch_verts;
    if (P < 32 || P > 256 || (P & 31)) { *code = 1; return "patch_verts must be a multiple of 32 in [32,256]"; }

    // ---- 1. constraints -> fused tet elements ---------------------------
    std::vector<Elem> elems;
    std::unordered_map<std::array<int32_t, 5>, std::vector<int32_t>, KeyHash> by_key;
    L.att0.assign(4 * (size_t)nv, 0.f);
    L.att1.assign(4 * (size_t)nv, 0.f);
    L.pin_from_cons.assign(4 * (size_t)nv, 0.f);
    std::vector<int32_t> attach_old;  // attach #i -> old vertex
    std::vector<uint8_t> has_att(nv, 0);
    int feature = 0;
    for (int c = 0; c < in.ncons; ++c) {
        const int32_t t = in.type[c];
        const int32_t *pt = in.pts + 4 * c;
        const float k = in.stiffness[c];
        if (t == CT_ATTACH) {
            int v = pt[0];
            if (v < 0 || v >= nv) { *code = 1; return "ATTACH vertex out of range"; }
            attach_old.push_back(v);
            if (has_att[v]) { *code = 2; return "more than one ATTACH constraint on a vertex"; }
            has_att[v] = 1;
            float kc = in.compressionstiffness[c];
            kc = kc >= 0.0f ? kc : k;
            float *a0 = &L.att0[4 * (size_t)v], *a1 = &L.att1[4 * (size_t)v];
            a0[0] = in.restvector[4 * c];
            a0[1] = in.restvector[4 * c + 1];
            a0[2] = in.restvector[4 * c + 2];
            a0[3] = k;  // k <= 0 -> skipped in the kernel (solve_attach_kernel early return)
            a1[0] = in.restlength[c];
            a1[1] = kc;
            a1[2] = in.dampingratio[c];
            if (k > 0.0f) {
                feature |= F_ATTACH;
                if (in.dampingratio[c] > 0.0f) feature |= F_DAMP;
            }
            continue;
        }
        if (t == CT_PINTOTARGET) {
            int v = pt[0];
            if (v < 0) continue;
            if (v >= nv) { *code = 1; return "PINTOTARGET vertex out of range"; }
            float *q = &L.pin_from_cons[4 * (size_t)v];
            q[0] = in.restvector[4 * c] + in.restdir[3 * c];
            q[1] = in.restvector[4 * c + 1] + in.restdir[3 * c + 1];
            q[2] = in.restvector[4 * c + 2] + in.restdir[3 * c + 2];
            q[3] = 1.0f;
            feature |= F_PINS;
            continue;
        }
        uint32_t bit;
        if (t == CT_TETARAPSVD) bit = M_ARAP;
        else if (t == CT_TETVOLUME) bit = M_VOL;
        else if (t == CT_TETSNH) bit = M_SNH;
        else if (t == CT_TETARAPANISO) bit = M_ANISO;
        else { *code = 2; return "constraint type " + std::to_string(t) + " is not on the HIP path"; }
        for (int i = 0; i < 4; ++i)
            if (pt[i] < 0 || pt[i] >= nv) { *code = 1; return "tet constraint vertex out of range"; }
        int32_t tid = in.tetid[c];
        if ((bit & (M_ARAP | M_SNH | M_ANISO)) && (tid < 0 || tid >= in.nt)) {
            *code = 1;
            return "tet constraint tetid out of range";
        }
        // static early-outs of the reference become "constraint absent"
        if (k <= 0.0f) continue;  // every tet kernel returns when c_stiffness <= 0
        float f0[3] = {0, 0, 0};
        if (bit == M_ANISO) {
            if (!in.v_fiber_dir) { *code = 1; return "aniso_arap needs v_fiber_dir"; }
            // f0 = (sum of 4 vertex fiber dirs)/4, normalised (solve_kernels.py:1345-1350)
            float s[3];
            for (int a = 0; a < 3; ++a) {
                s[a] = in.v_fiber_dir[3 * pt[0] + a] + in.v_fiber_dir[3 * pt[1] + a];
                s[a] = s[a] + in.v_fiber_dir[3 * pt[2] + a];
                s[a] = s[a] + in.v_fiber_dir[3 * pt[3] + a];
                s[a] = s[a] / 4.0f;
            }
            float fn = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
            if (fn < 1.0e-9f) continue;
            for (int a = 0; a < 3; ++a) f0[a] = s[a] / fn;
        }
        // one element per vertex 4-tuple; volume does not read rest_matrix, so it
        // joins the element of any tetid
        int target = -1;
        auto place = [&](std::vector<int32_t> &lst) {
            for (int e : lst)
                if (!(elems[e].mask & bit) &&
                    (bit == M_VOL || elems[e].tetid < 0 || elems[e].tetid == tid)) {
                    target = e;
                    return;
                }
        };
        std::array<int32_t, 5> kk{pt[0], pt[1], pt[2], pt[3], 0};
        auto it = by_key.find(kk);
        if (it != by_key.end()) place(it->second);
        if (target < 0) {
            Elem e;
            for (int i = 0; i < 4; ++i) e.p[i] = pt[i];
            e.tetid = (bit == M_VOL) ? -1 : tid;
            target = (int)elems.size();
            elems.push_back(e);
            by_key[kk].push_back(target);
        }
        Elem &e = elems[target];
        e.mask |= bit;
        if (bit != M_VOL) e.tetid = tid;
        const float kd = in.dampingratio[c];
        if (bit == M_ARAP) {
            e.cp.k_arap = k;
        } else if (bit == M_VOL) {
            float kc = in.compressionstiffness[c];
            e.cp.k_vol = k;
            e.cp.kcomp_vol = kc >= 0.0f ? kc : k;
            e.cp.damp_vol = kd;
            e.rvol = in.restlength[c];
            if (kd > 0.0f) feature |= F_DAMP;
        } else if (bit == M_SNH) {
            e.cp.mu = k;
            e.cp.lam = in.restdir[3 * c];
            e.cp.damp_snh = kd;
            e.rsnh = in.restlength[c];
            if (kd > 0.0f) feature |= F_DAMP;
        } else {
            e.cp.k_aniso = k;
            for (int a = 0; a < 3; ++a) e.f0[a] = f0[a];
            e.belly = 1.0f - (in.tendonmask ? in.tendonmask[tid] : 0.0f);
        }
        feature |= (int)bit;
    }
    for (auto &e : elems)
        if (e.tetid < 0) e.tetid = 0;  // volume-only element: tetid unused
    L.feature = feature;
    L.n_elems = (int)elems.size();

    // ---- 2. material classes ----------------------------------------------
    std::vector<uint32_t> ecls(elems.size());
    {
        std::unordered_map<std::string, uint32_t> seen;
        for (size_t i = 0; i < elems.size(); ++i) {
            std::string key((const char *)&elems[i].cp, sizeof(ClassParams));
            auto f = seen.find(key);
            if (f == seen.end()) {
                uint32_t id = (uint32_t)L.classes.size();
                if (id > 0xffff) { *code = 5; return "too many distinct material classes"; }
                L.classes.push_back(elems[i].cp);
                seen.emplace(key, id);
                ecls[i] = id;
            } else {
                ecls[i] = f->second;
            }
        }
        if (L.classes.empty()) L.classes.push_back(ClassParams{});
    }

    // ---- 3. vertex renumbering: Morton order of rest positions -----------
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int v = 0; v < nv; ++v)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], in.pos0[3 * v + a]);
            hi[a] = std::max(hi[a], in.pos0[3 * v + a]);
        }
    double ext = 0;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, (double)hi[a] - lo[a]);
    if (!(ext > 0)) ext = 1;
    auto morton = [&](const float *x) {
        uint64_t c = 0;
        for (int a = 0; a < 3; ++a) {
            double q = ((double)x[a] - lo[a]) / ext * 2097151.0;
            uint64_t u = (uint64_t)std::min(std::max(q, 0.0), 2097151.0);
            c |= spread21(u) << a;
        }
        return c;
    };
    std::vector<uint64_t> vcode(nv);
    for (int v = 0; v < nv; ++v) vcode[v] = morton(in.pos0 + 3 * v);
    L.perm.resize(nv);
    std::iota(L.perm.begin(), L.perm.end(), 0);
    std::stable_sort(L.perm.begin(), L.perm.end(),
                     [&](int a, int b) { return vcode[a] < vcode[b]; });
    L.iperm.assign(nv, 0);
    for (int i = 0; i < nv; ++i) L.iperm[L.perm[i]] = i;
    const int npatch = (nv + P - 1) / P;

    // ---- 4. element order: Morton of centroid -----------------------------
    std::vector<int32_t> eord(elems.size());
    std::iota(eord.begin(), eord.end(), 0);
    {
        std::vector<uint64_t> ecode(elems.size());
        for (size_t i = 0; i < elems.size(); ++i) {
            float c[3] = {0, 0, 0};
            for (int k = 0; k < 4; ++k)
                for (int a = 0; a < 3; ++a) c[a] += 0.25f * in.pos0[3 * elems[i].p[k] + a];
            ecode[i] = morton(c);
        }
        std::stable_sort(eord.begin(), eord.end(),
                         [&](int a, int b) { return ecode[a] < ecode[b]; });
    }

    if (in.flat && in.block_tets >= 64) {
        // block strategy: vertices are renumbered block by block in element
        // order, each at the first block that touches it (Morton order among a
        // block's new vertices; untouched vertices last).  A block's own
        // vertices then form one contiguous range of its sorted vertex table,
        // so their staging gathers and finished-position stores are coalesced.
        // The sums do not depend on vertex numbers: every vertex still adds its
        // corners in (element, corner) order and its partials in block order.
        const size_t TB = (size_t)in.block_tets;
        std::vector<int32_t> nperm;
        nperm.reserve(nv);
        std::vector<char> seen(nv, 0);
        std::vector<int32_t> fresh;
        for (size_t t0 = 0; t0 < eord.size(); t0 += TB) {
            fresh.clear();
            const size_t t1 = std::min(eord.size(), t0 + TB);
            for (size_t g = t0; g < t1; ++g)
                for (int k = 0; k < 4; ++k) {
                    const int v = elems[eord[g]].p[k];
                    if (!seen[v]) {
                        seen[v] = 1;
                        fresh.push_back(v);
                    }
                }
            std::sort(fresh.begin(), fresh.end(), [&](int a, int b) { return L.iperm[a] < L.iperm[b]; });
            nperm.insert(nperm.end(), fresh.begin(), fresh.end());
        }
        for (int i = 0; i < nv; ++i)
            if (!seen[L.perm[i]]) nperm.push_back(L.perm[i]);
        L.perm.swap(nperm);
        for (int i = 0; i < nv; ++i) L.iperm[L.perm[i]] = i;
    }

    if (in.flat) {
        renumber_vertex_records(L, attach_old, nv);
        std::string err = build_flat(in, L, elems, ecls, eord, feature, code);
        if (err.empty() && in.block_tets > 0) err = build_blocks(in, L, feature, code);
        return err;
    }

    // ---- 5. (patch, element) entries ----------------------------------------
    std::vector<int32_t> pcount(npatch + 1, 0);
    auto patches_of = [&](const Elem &e, int *out) {
        int n = 0;
        for (int k = 0; k < 4; ++k) {
            int p = L.iperm[e.p[k]] / P;
            bool dup = false;
            for (int j = 0; j < n; ++j) dup |= (out[j] == p);
            if (!dup) out[n++] = p;
        }
        return n;
    };
    for (int e : eord) {
        int ps[4];
        int n = patches_of(elems[e], ps);
        for (int j = 0; j < n; ++j) pcount[ps[j] + 1]++;
    }
    for (int p = 0; p < npatch; ++p) pcount[p + 1] += pcount[p];
    const size_t nptet = (size_t)pcount[npatch];
    if (nptet > (size_t)INT32_MAX / 4) { *code = 5; return "too many replicated tets"; }
    std::vector<int32_t> pe(nptet);
    {
        std::vector<int32_t> fill(pcount.begin(), pcount.end() - 1);
        for (int e : eord) {
            int ps[4];
            int n = patches_of(elems[e], ps);
            for (int j = 0; j < n; ++j) pe[fill[ps[j]]++] = e;
        }
    }

    // ---- 6. per patch tables ------------------------------------------------
    L.patches.resize(npatch);
    L.lidx.assign(4 * nptet, 0);
    L.meta.assign(nptet, 0);
    L.Bp.assign(9 * nptet, 0.f);
    L.rvol.assign(nptet, 0.f);
    L.rsnh.assign(nptet, 0.f);
    L.fib.assign(4 * nptet, 0.f);
    L.ptet_tetid.assign(nptet, 0);
    L.vslot_off.assign(nv + 1, 0);
    L.halo.clear();
    std::vector<std::vector<uint16_t>> vs_tmp;  // reused per patch
    int max_table = 0, max_ptets = 0;
    std::vector<int32_t> hl;
    for (int p = 0; p < npatch; ++p) {
        const int vbeg = p * P, nown = std::min(P, nv - vbeg);
        const int tb = pcount[p], nt_p = pcount[p + 1] - tb;
        hl.clear();
        for (int j = tb; j < tb + nt_p; ++j)
            for (int k = 0; k < 4; ++k) {
                int v = L.iperm[elems[pe[j]].p[k]];
                if (v < vbeg || v >= vbeg + nown) hl.push_back(v);
            }
        std::sort(hl.begin(), hl.end());
        hl.erase(std::unique(hl.begin(), hl.end()), hl.end());
        const int table = nown + (int)hl.size();
        if (table > 65535) { *code = 5; return "patch vertex table exceeds 65535"; }
        if ((size_t)nt_p * 4 > 65535) { *code = 5; return "patch has more than 16383 tets"; }
        max_table = std::max(max_table, table);
        max_ptets = std::max(max_ptets, nt_p);
        L.patches[p] = PatchDesc{vbeg, nown, (int32_t)L.halo.size(), (int32_t)hl.size(), tb, nt_p, 0, 0};
        L.halo.insert(L.halo.end(), hl.begin(), hl.end());
        // slot counts per owned vertex
        std::vector<int32_t> cnt(nown, 0);
        for (int j = 0; j < nt_p; ++j) {
            const Elem &e = elems[pe[tb + j]];
            const size_t g = (size_t)tb + j;
            for (int k = 0; k < 4; ++k) {
                int v = L.iperm[e.p[k]];
                int li;
                if (v >= vbeg && v < vbeg + nown) {
                    li = v - vbeg;
                    cnt[li]++;
                } else {
                    li = nown + (int)(std::lower_bound(hl.begin(), hl.end(), v) - hl.begin());
                }
                L.lidx[4 * g + k] = (uint16_t)li;
            }
            bool primary = (L.iperm[e.p[0]] >= vbeg && L.iperm[e.p[0]] < vbeg + nown);
            L.meta[g] = ecls[pe[tb + j]] | (e.mask << 16) | ((primary ? 1u : 0u) << 20);
            const float *B = in.rest_matrix + 9 * (size_t)e.tetid;
            for (int q = 0; q < 9; ++q) L.Bp[(size_t)q * nptet + g] = B[q];
            L.rvol[g] = e.rvol;
            L.rsnh[g] = e.rsnh;
            L.fib[4 * g + 0] = e.f0[0];
            L.fib[4 * g + 1] = e.f0[1];
            L.fib[4 * g + 2] = e.f0[2];
            L.fib[4 * g + 3] = e.belly;
            L.ptet_tetid[g] = e.tetid;
        }
        for (int i = 0; i < nown; ++i) L.vslot_off[vbeg + i + 1] = cnt[i];
    }
    for (int v = 0; v < nv; ++v) L.vslot_off[v + 1] += L.vslot_off[v];
    L.vslot.assign((size_t)L.vslot_off[nv], 0);
    {
        std::vector<int32_t> fill(L.vslot_off.begin(), L.vslot_off.end() - 1);
        for (int p = 0; p < npatch; ++p) {
            const PatchDesc &d = L.patches[p];
            for (int j = 0; j < d.ntet; ++j) {
                const size_t g = (size_t)d.tbeg + j;
                for (int k = 0; k < 4; ++k) {
                    int li = L.lidx[4 * g + k];
                    if (li < d.nown) L.vslot[fill[d.vbeg + li]++] = (uint16_t)(j * 4 + k);
                }
            }
        }
    }
    L.max_table = max_table;
    L.max_ptets = max_ptets;

    renumber_vertex_records(L, attach_old, nv);

    // ---- 7. byte accounting ---------------------------------------------------
    // algorithmic: each tet element's static data read once, each vertex read+written once
    L.bytes_alg = alg_bytes(feature, (double)elems.size(), nv);
    double bpt = 8.0 + 4.0;  // lidx + meta
    if (feature & (F_ARAP | F_SNH | F_ANISO)) bpt += 36.0;
    if (feature & F_VOL) bpt += 4.0;
    if (feature & F_SNH) bpt += 4.0;
    if (feature & F_ANISO) bpt += 16.0;
    double tab = (double)nv + (double)L.halo.size();
    double per_table = (feature & F_DAMP) ? 32.0 : 16.0;
    L.bytes_impl = bpt * (double)nptet + per_table * tab + 4.0 * L.halo.size() +
                   2.0 * L.vslot.size() + 4.0 * nv /*vslot_off*/ + 16.0 * nv /*write*/ +
                   32.0 * npatch;
    if (feature & F_PINS) L.bytes_impl += 16.0 * nv;
    if (feature & F_ATTACH) L.bytes_impl += 32.0 * nv;
    return std::string();
}

}  // namespace xpbd
