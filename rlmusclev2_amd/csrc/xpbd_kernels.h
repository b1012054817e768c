// Device kernels of the XPBD substep (gfx950).  Included by xpbd.hip only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xpbd_layout.h"

namespace xpbd {

struct SimParams {
    float dt, inv_dt2;
    float gx, gy, gz;
    float veldamping;
    int quasi_static;
    float contraction_ratio, fiber_scale;
    int has_compress;
    float exp_damp, const_damp;
};

// per-class constants that depend on dt, folded on the host once per set_params
struct DevClass {
    ClassParams p;
    float at_arap;   // (1/k_arap)/dt^2
    float a_vol;     // (1/k_vol)/dt^2
    float a_comp;    // (1/kcomp)/dt^2
    float g_vol;     // damping gamma of the volume constraint (before +1)
    float g_comp;
    float at_aniso;  // (1/(k_aniso*fiber_scale))/dt^2
    float pad0, pad1;
};

struct SdfParams {
    int nbones, res;
    float contact_offset, max_push;
    const float *grids;      // nbones * res^3
    const float *bone;       // per bone: 12 xform + 3 origin + inv spacing = 16 floats
    const int32_t *vgroup;   // per vertex (new order): bone group
    const int32_t *goff;     // ngroups + 1 bone offsets
};

struct IterArgs {
    const PatchDesc *patch;
    const int32_t *halo;
    const ushort4 *lidx;
    const uint32_t *meta;
    const float *Bp;         // 9 planes of nptet
    const float *rvol, *rsnh;
    const float4 *fib;       // f0.xyz, fiberscale (per ptet, refreshed from activation)
    const DevClass *cls;
    const int32_t *vslot_off;
    const uint16_t *vslot;
    const int4 *gidx;        // flat strategy: element vertex ids
    const int4 *slot_of;     // flat strategy: vertex-major corner position of each element corner
    const float4 *pin;       // target.xyz, active
    const float4 *att0, *att1;
    float4 *dpw;             // accumulate mode output
    float4 *lam;             // per element multipliers (x: arap or snh-D, y: volume or snh-H, z: aniso)
    int64_t nptet;
    int max_table;
    SimParams sp;
};

enum { MODE_FUSED = 0, MODE_ACCUM = 1 };

}  // namespace xpbd
