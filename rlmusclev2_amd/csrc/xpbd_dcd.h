// Tri-tri discrete collision detection and response on the GPU (gfx950).
// Included by xpbd.hip only.
//
// Semantics follow the reference UDCF collider (src/muscle/collision/udcf/):
// per-triangle centroid / bounding radius / normal (kernel.py:47-84), the
// broadphase emit rules (kernel.py:851-915), the narrow phase (kernel.py:317-474),
// respond with friction (kernel.py:714-799) and the averaged Jacobi apply
// (kernel.py:802-815).  Differences in mechanism, not in result:
//  * broadphase: uniform grid hashed into a power-of-two table, cell size
//    max(2 x mean rest edge, 2 x max radius + thickness) so every pair that
//    passes the sphere test lies in adjacent cells; each triangle is visited
//    only from its own cell, so the candidate set equals the exhaustive one;
//  * bounded work on broken geometry: the cell is capped at DCD_CELL_CAP x
//    (2 x mean rest edge); a triangle too large for the capped cell (or with a
//    non-finite centroid) takes no part and is counted, and a triangle keeps at
//    most DCD_MAX_PER_TRI candidates (the rest are counted).  Neither limit is
//    reached on a valid mesh, where the pair list equals the exhaustive one;
//  * pairs are written per triangle at scanned offsets and each triangle's
//    segment is sorted, giving the (i, j)-sorted pair list independent of
//    hashing and scheduling;
//  * contacts stay in their pair slot (no atomic compaction) and corrections
//    are summed in 2^-32 m fixed point with integer atomics, so the result is
//    the same on every run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xpbd {

struct DcdContact {
    int32_t cur[3], closest[3];
    float C, C1, C2;
    float n[3], b1[3], b2[3];
    int32_t valid;
};

constexpr float DCD_CELL_CAP = 4.0f;
constexpr uint32_t DCD_MAX_PER_TRI = 1024;

struct DcdDev {
    int ntris;
    int tbits;                  // hash table size = 1 << tbits
    int64_t max_pairs;
    const int4 *tris;           // internal vertex ids, w unused
    const int32_t *group;       // 0 movable / 1 static
    float4 *cr;                 // centroid.xyz, radius
    float4 *nrm;                // unit face normal
    int4 *cell;                 // cell coords, w = hash key
    uint32_t *bucket_cnt;       // T
    uint32_t *bucket_off;       // T + 1
    uint32_t *bucket_fill;      // T
    int32_t *sorted;            // ntris
    uint32_t *pair_cnt;         // ntris (+1 for the scan)
    uint32_t *pair_off;         // ntris + 1
    int2 *pairs;                // max_pairs
    DcdContact *contacts;       // max_pairs
    unsigned long long *acc;    // 3 per vertex, fixed point
    uint32_t *acc_n;            // per vertex
    float4 *xprev;              // snapshot at broadphase
    float *scal;                // [0] cell, [1] inv cell, [2] rmax bits (as float), [3] mean edge
    uint32_t *rmax_bits;
    uint32_t *ncontact;         // valid contacts of the last narrow pass
    uint32_t *overflow;         // [0] triangles left out (too large / non-finite), [1] truncated candidate lists
    float thickness, multiplier, mu_s, mu_d, offset_eps, normal_cull, mean_edge2;
    int allow_self;
};

}  // namespace xpbd
