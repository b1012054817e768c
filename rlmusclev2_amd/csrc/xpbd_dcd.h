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
//    from the median of the current radii (or the caller's cell_size).  A triangle searches its own level for partners of higher
//    index and every higher level for any partner, over the cells that the
//    box of half-width 1.01 (r_i + r_max(level) + thickness) overlaps (at most
//    3 per axis) and takes each candidate only from its own cell, so every
//    pair passing the sphere test is found exactly once and the set equals
//    the exhaustive one;
//  * bounded work on broken geometry: a triangle above the top level (or with
//    a non-finite centroid) takes no part and is counted; a search stops
//    before the cell that would take it past DCD_MAX_VISIT candidates
//    (counted).  Neither limit is reached on a valid mesh;
//  * the bucket-ordered copy of each triangle's centroid/radius, key (index,
//    level, cell) and vertices is what the search reads, so a candidate costs
//    contiguous loads and no gather through the triangle index;
//  * a workgroup collects the pairs (min, max) its searches emit in LDS and
//    appends them with one atomic on the pair cursor, in no particular order;
//  * the response is fused into the narrow phase: a valid contact is applied
//    where it is found.  No order matters: corrections are summed in 2^-32 m
//    fixed point with integer atomics, so the result is the same on every run.  The
//    (i, j)-sorted list that the reference builds is made on request
//    (segments by first index, each sorted by the second) for the pair-list
//    query.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xpbd {

constexpr float DCD_FIX = 4294967296.0f;       // 2^32 fixed-point scale of the corrections (metres)
constexpr int DCD_LEVELS = 12;                 // cells c0 .. 2048 c0
constexpr int DCD_HIST = 64;                   // radius histogram, 4 bins per octave
constexpr uint32_t DCD_MAX_VISIT = 1u << 15;   // candidates one search may visit
constexpr int DCD_SORT_MAX = 8192;             // longer segments stay unsorted (counted)
constexpr int DCD_STAGE = 2048;                // pairs one workgroup stages in LDS
constexpr int DCD_LVL_SHIFT = 27;              // key.x = triangle | level << 27

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
    uint32_t *lvl_cnt;          // per level: 1 when a triangle lives there
    uint32_t *bucket_cnt;       // T
    uint32_t *bucket_off;       // T + 1
    uint32_t *bucket_fill;      // T
    int4 *key;                  // bucket order: triangle | level << 27, cell x, y, z
    int4 *tv;                   // bucket order: vertex ids, w = group
    unsigned long long *npairs; // pair cursor (pairs found, may exceed max_pairs)
    int2 *pairs;                // max_pairs, in append order
    uint32_t *pair_cnt;         // ntris (+1 for the scan)      } the sorted list,
    uint32_t *pair_off;         // ntris + 1                     } built on request
    uint32_t *pair_fill;        // ntris                         }
    int2 *pairs_seg;            // max_pairs                     }
    unsigned long long *acc;    // 3 per vertex, fixed point
    uint32_t *acc_n;            // per vertex
    float4 *xprev;              // snapshot at broadphase
    float *scal;                // [0] base cell c0, [1] 1 / c0, [2] [3] unused
    uint32_t *ncontact;         // [0] contacts of the running pass, [1] of the last finished pass
    uint32_t *overflow;         // [0] triangles left out (too large / non-finite), [1] truncated searches,
                                // [2] segments too long to sort
    float thickness, multiplier, mu_s, mu_d, offset_eps, normal_cull, cell_fixed;
    int allow_self;
};

}  // namespace xpbd
