// Tri-tri discrete collision detection and response on the GPU (gfx950).
// Included by xpbd.hip only.
//
// Semantics follow the reference UDCF collider (src/muscle/collision/udcf/):
// per-triangle centroid / bounding radius / normal (kernel.py:47-84), the
// broadphase emit rules (kernel.py:851-915), the narrow phase (kernel.py:317-474),
// respond with friction (kernel.py:714-799) and the averaged Jacobi apply
// (kernel.py:802-815).  Differences in mechanism, not in result:
//  * broadphase: a hierarchy of uniform grids hashed into one power-of-two
//    table (key = cell and level).  Level b has cell c0 * 2^b; a triangle
//    lives at the lowest level whose cell is at least 1.01 (2 r + thickness),
//    so one stretched triangle no longer inflates everybody's cell.  c0 comes
//    from the 90th percentile of the current radii (or the caller's
//    cell_size).  A triangle searches its own level for partners of higher
//    index and every higher level for any partner, over the cells that the
//    box of half-width 1.01 (r_i + r_max(level) + thickness) overlaps (at most
//    3 per axis) and takes each candidate only from its own cell, so every
//    pair passing the sphere test is found exactly once and the set equals
//    the exhaustive one;
//  * bounded work on broken geometry: a triangle above the top level (or with
//    a non-finite centroid) takes no part and is counted; a search stops
//    before the cell that would take it past DCD_MAX_VISIT candidates
//    (counted).  Neither limit is reached on a valid mesh;
//  * a pair (i, j), i < j, goes to i's segment whichever triangle found it
//    (slots by atomic counters), and each segment is then sorted by j, giving
//    the (i, j)-sorted pair list independent of hashing and scheduling;
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

constexpr int DCD_LEVELS = 12;                 // cells c0 .. 2048 c0
constexpr int DCD_HIST = 64;                   // radius histogram, 4 bins per octave
constexpr uint32_t DCD_MAX_VISIT = 1u << 15;   // candidates one search may visit
constexpr int DCD_SORT_MAX = 8192;             // longer segments stay unsorted (counted)

struct DcdDev {
    int ntris;
    int tbits;                  // hash table size = 1 << tbits
    int64_t max_pairs;
    const int4 *tris;           // internal vertex ids, w unused
    const int32_t *group;       // 0 movable / 1 static
    float4 *cr;                 // centroid.xyz, radius
    float4 *nrm;                // unit face normal
    int4 *cell;                 // cell coords at the triangle's level, w = hash key
    int32_t *tlvl;              // grid level per triangle, -1 = left out
    float4 *rec;                // centroid, radius in bucket order
    uint32_t *hist;             // DCD_HIST radius bins
    uint32_t *lvl_cnt;          // triangles per level
    uint32_t *bucket_cnt;       // T
    uint32_t *bucket_off;       // T + 1
    uint32_t *bucket_fill;      // T
    int32_t *sorted;            // ntris
    uint32_t *pair_cnt;         // ntris (+1 for the scan)
    uint32_t *pair_off;         // ntris + 1
    uint32_t *pair_fill;        // ntris
    int2 *pairs;                // max_pairs
    DcdContact *contacts;       // max_pairs
    unsigned long long *acc;    // 3 per vertex, fixed point
    uint32_t *acc_n;            // per vertex
    float4 *xprev;              // snapshot at broadphase
    float *scal;                // [0] base cell c0, [1] 1 / c0, [2] max radius, [3] unused
    uint32_t *rmax_bits;
    uint32_t *ncontact;         // valid contacts of the last narrow pass
    uint32_t *overflow;         // [0] triangles left out (too large / non-finite), [1] truncated searches,
                                // [2] segments too long to sort
    float thickness, multiplier, mu_s, mu_d, offset_eps, normal_cull, cell_fixed;
    int allow_self;
};

}  // namespace xpbd
