// Host-side layout builder: constraint table -> fused tet elements -> patches.
//
// A patch owns a contiguous run of (renumbered) vertices and holds a copy of
// every tet element that touches one of them.  One workgroup processes one
// patch per Jacobi iteration: it stages the patch's vertex table (owned + halo)
// in LDS, evaluates its tets, and sums the corner corrections of each owned
// vertex in a fixed slot order — a deterministic segmented reduction that
// replaces the reference's float atomics (update_dP, solve_kernels.py:52-56).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace xpbd {

enum : uint32_t {
    M_ARAP = 1u,   // TETARAPSVD
    M_VOL = 2u,    // TETVOLUME
    M_SNH = 4u,    // TETSNH
    M_ANISO = 8u,  // TETARAPANISO
};
enum : int32_t {
    F_ARAP = 1, F_VOL = 2, F_SNH = 4, F_ANISO = 8, F_DAMP = 16, F_ATTACH = 32, F_PINS = 64,
};

struct PatchDesc {
    int32_t vbeg, nown, hbeg, nhalo, tbeg, ntet, pad0, pad1;
};

struct ClassParams {
    float k_arap, k_vol, kcomp_vol, damp_vol, mu, lam, damp_snh, k_aniso;
};

struct HostInput {
    int nv = 0, nt = 0;
    const float *pos0 = nullptr;       // nv*3
    const int32_t *tets = nullptr;     // nt*4
    const float *rest_matrix = nullptr;  // nt*9
    const float *v_fiber_dir = nullptr;  // nv*3 (may be null)
    const float *tendonmask = nullptr;   // nt   (may be null)
    // constraint table (SoA, reference engine.py:1179-1211)
    int ncons = 0;
    const int32_t *type = nullptr, *pts = nullptr, *tetid = nullptr;
    const float *stiffness = nullptr, *compressionstiffness = nullptr, *dampingratio = nullptr,
                *restlength = nullptr, *restvector = nullptr, *restdir = nullptr;
    int patch_verts = 256;
    bool flat = false;   // one copy per element + per-vertex corner CSR (no patches)
    int block_tets = 0;  // flat elements in blocks of this many, reduced in LDS (0 = off)
};

// "block" strategy: consecutive runs of elements (Morton order) form blocks.
// A block stages the vertices its elements touch in LDS, projects its elements
// once, and sums each local vertex's corner corrections in ascending slot
// order.  A vertex touched by one block only is finished by that block; a
// vertex touched by several blocks gets one partial sum per block, written to
// a vertex-major partial array and finished by a second, per-vertex pass that
// adds its partials in ascending block order.
struct BlockDesc {
    int32_t tbeg, ntet, vbeg, nloc;
};

// "resident" mode of the block strategy: runs of R consecutive blocks form a
// part, and one workgroup per part runs every Jacobi iteration of a substep in
// one launch.  The part's vertices stay in LDS across the iterations; its
// elements stay in registers.  A vertex that blocks of other parts also touch
// (a cross vertex) gets the same per-block partial slots as in the block
// strategy; each part writes its blocks' partials, waits for the parts it shares
// vertices with, then adds all partials of the vertex in ascending block order.
// Every other vertex is summed in LDS, also in ascending block order, so a
// substep gives bit for bit the positions of the block strategy's per-iteration
// launches.
struct PartDesc {
    int32_t b0, nblk, vbeg, nloc, nbeg, nnb, pad0, pad1;
};

struct Parts {
    int32_t R = 0;                   // blocks per part (the most, when the parts were spread evenly)
    std::vector<PartDesc> parts;
    // per part-local entry: vertex id (new order), index into svid / soff for a
    // cross vertex (else -1), write-back flag (the part of the vertex's first
    // block; vertices no element touches go to one part), unused
    std::vector<int32_t> prec;       // 4 per entry
    std::vector<uint32_t> bmap;      // per block-local entry: part-local index | cross << 31
    std::vector<uint16_t> plidx;     // 4 per element: part-local corner ids
    std::vector<int32_t> nbr;        // neighbour parts, ascending, CSR by PartDesc::nbeg / nnb
    int32_t max_nloc = 0, max_nnb = 0, n_cross = 0;
};

struct Layout {
    // vertex renumbering
    std::vector<int32_t> perm;   // new -> old
    std::vector<int32_t> iperm;  // old -> new
    std::vector<PatchDesc> patches;
    std::vector<int32_t> halo;               // new vertex ids
    // replicated tets ("ptets"), patch-major
    std::vector<uint16_t> lidx;              // 4 per ptet
    std::vector<uint32_t> meta;              // class | mask<<16 | primary<<20
    std::vector<float> Bp;                   // 9 planes of nptet
    std::vector<float> rvol, rsnh;           // per ptet
    std::vector<float> fib;                  // 4 per ptet: f0.xyz, belly
    std::vector<int32_t> ptet_tetid;         // per ptet (activation index)
    std::vector<int32_t> vslot_off;          // nv+1 (new order)
    std::vector<uint16_t> vslot;             // slot = local tet*4 + corner
    std::vector<int32_t> gidx;               // flat: 4 new vertex ids per element
    std::vector<int32_t> vslot32;            // flat: vertex-major position of corner (element, k)
    // block strategy
    std::vector<BlockDesc> blocks;
    std::vector<int32_t> bvt;                // per local entry: vertex id (new order)
    std::vector<int32_t> bps;                // per local entry: partial slot, -1 = finished in the block
    std::vector<uint32_t> brun;              // per local entry: first slot | slot count << 16
    std::vector<uint16_t> blidx;             // 4 per element: local entry of each corner
    std::vector<uint16_t> bcpos;             // 4 per element: the corner's position in the block's
                                             // vertex-major corner list (the inverse of bslot)
    std::vector<uint16_t> bslot;             // 4 per element, per block: corner slots grouped by local entry,
                                             // ascending (element, corner); value = corner * block_tets + element
    std::vector<int32_t> svid;               // vertices finished by the second pass
    std::vector<int32_t> soff;               // svid.size()+1 offsets into the partial array
    int32_t max_nloc = 0, block_tets = 0, n_partial = 0;
    std::vector<ClassParams> classes;
    // attach records in new vertex order
    std::vector<float> att0, att1;           // 4 per vertex
    std::vector<int32_t> attach_vertex;      // attach constraint #i (table order) -> new vertex
    std::vector<float> pin_from_cons;        // 4 per vertex (PINTOTARGET constraints)
    int32_t feature = 0;
    int32_t max_table = 0, max_ptets = 0, n_elems = 0;
    double bytes_alg = 0.0, bytes_impl = 0.0;
};

// returns empty string on success, else the error message; *code gets the status
std::string build_layout(const HostInput &in, Layout &out, int *code);

// parts of R consecutive blocks over a built block layout (L.blocks and its tables); with
// target_parts > 0, that many parts (at most one per block) of evenly spread block runs
std::string build_parts(const Layout &L, int nv, int R, Parts &out, int *code, int target_parts = 0);

// FP32 rest precompute in the reference's operation order (engine.py:199-224):
// Dm columns p_k - p3, |det|/6, mass += V*density/4 (tet order), B = inv(Dm).
void precompute_rest(int nv, const float *pos0, int nt, const int32_t *tets, float density,
                     float *rest_volume, float *rest_matrix, float *mass);

}  // namespace xpbd
