// Per-tet MLP for warm-start Lagrange multipliers: 24 -> 64 -> 64 -> 4
// (22 features padded to 24, 3 multipliers padded to 4), ReLU hidden layers.
//
// Dense layers run on gfx950 matrix cores: v_mfma_f32_32x32x16_bf16 with BF16
// weights and activations and FP32 accumulation.  One wavefront owns 32 rows
// (tets).  Operand layout of the 32x32x16 MFMA used below (per lane l, row or
// column index l % 32, k-half h = l / 32):
//   A (32 x 16): lane holds row l%32, k = 8h .. 8h+7
//   B (16 x 32): lane holds column l%32, k = 8h .. 8h+7
//   D (32 x 32): lane holds column l%32, rows (i/4)*8 + 4h + (i%4), i = 0..15
// The layout is checked against NumPy by tests/test_gpu_mlp.py.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <mutex>

#include "../../include/xpbd.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int NIN = XPBD_MLP_IN;    // 24
constexpr int NH = XPBD_MLP_HIDDEN; // 64
constexpr int NOUT = XPBD_MLP_OUT;  // 4
constexpr int NHP = NH + 1;         // padded LDS row of one activation vector

__device__ __forceinline__ __bf16 tobf(float x) { return (__bf16)x; }
__device__ __forceinline__ __bf16 bfraw(uint16_t v) { return __builtin_bit_cast(__bf16, v); }

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int drow(int i, int h) { return (i >> 2) * 8 + 4 * h + (i & 3); }

// one wavefront: rows [r0, r0+32) of X (n x NIN fp32); writes Y (n x NOUT) and,
// when non-null, the post-ReLU activations H1, H2 (n x NH fp32) for backward.
// LDS: per wave 32 x 64 floats.
__global__ __launch_bounds__(256) void k_mlp_fwd(xpbd_mlp_dev net, const float *__restrict__ X, int64_t n,
                                                 float *__restrict__ Y, float *__restrict__ H1g,
                                                 float *__restrict__ H2g) {
    // activation rows are padded to NHP floats: a lane reads its own row, and an
    // unpadded stride of 64 floats would put all 32 rows in the same LDS bank
    __shared__ float sh[4][32 * NHP];
    __shared__ float w3[NOUT * NH], b3[NOUT];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, col = l & 31, h = l >> 5;
    for (int i = threadIdx.x; i < NOUT * NH; i += blockDim.x) w3[i] = net.W3[i];
    if (threadIdx.x < NOUT) b3[threadIdx.x] = net.b3[threadIdx.x];
    __syncthreads();
    const int64_t r0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + wave) * 32;
    float *H = sh[wave];  // rows past n are computed on zeros and never stored
    const int64_t row = r0 + col;
    const bool rv = row < n;
    // ---- layer 1: K = NIN padded to 32 (2 k-steps) ----
    f32x16 acc[2];
    for (int t = 0; t < 2; ++t)
        for (int i = 0; i < 16; ++i) acc[t][i] = 0.f;
    for (int s = 0; s < 2; ++s) {
        bf16x8 a;
        for (int i = 0; i < 8; ++i) {
            const int k = 16 * s + 8 * h + i;
            a[i] = tobf((rv && k < NIN) ? X[row * NIN + k] : 0.f);
        }
        for (int t = 0; t < 2; ++t) {
            bf16x8 b;
            for (int i = 0; i < 8; ++i) {
                const int k = 16 * s + 8 * h + i;
                b[i] = k < NIN ? bfraw(net.W1[(32 * t + col) * NIN + k]) : tobf(0.f);
            }
            acc[t] = mfma(a, b, acc[t]);
        }
    }
    for (int t = 0; t < 2; ++t)
        for (int i = 0; i < 16; ++i) {
            const int c = 32 * t + col;
            H[drow(i, h) * NHP + c] = fmaxf(acc[t][i] + net.b1[c], 0.f);
        }
    __syncthreads();
    if (H1g)
        for (int i = l; i < 32 * NH; i += 64) {
            const int64_t rr = r0 + i / NH;
            if (rr < n) H1g[rr * NH + (i % NH)] = H[(i / NH) * NHP + i % NH];
        }
    // ---- layer 2: K = 64 (4 k-steps) ----
    for (int t = 0; t < 2; ++t)
        for (int i = 0; i < 16; ++i) acc[t][i] = 0.f;
    for (int s = 0; s < 4; ++s) {
        bf16x8 a;
        for (int i = 0; i < 8; ++i) a[i] = tobf(H[col * NHP + 16 * s + 8 * h + i]);
        for (int t = 0; t < 2; ++t) {
            bf16x8 b;
            for (int i = 0; i < 8; ++i) b[i] = bfraw(net.W2[(32 * t + col) * NH + 16 * s + 8 * h + i]);
            acc[t] = mfma(a, b, acc[t]);
        }
    }
    __syncthreads();
    for (int t = 0; t < 2; ++t)
        for (int i = 0; i < 16; ++i) {
            const int c = 32 * t + col;
            H[drow(i, h) * NHP + c] = fmaxf(acc[t][i] + net.b2[c], 0.f);
        }
    __syncthreads();
    if (H2g)
        for (int i = l; i < 32 * NH; i += 64) {
            const int64_t rr = r0 + i / NH;
            if (rr < n) H2g[rr * NH + (i % NH)] = H[(i / NH) * NHP + i % NH];
        }
    // ---- layer 3 (64 -> 4) on the vector ALUs: lane (row col, outputs 2h, 2h+1) ----
    if (rv) {
        for (int oo = 0; oo < 2; ++oo) {
            const int o = 2 * h + oo;
            float y = b3[o];
            for (int k = 0; k < NH; ++k) y += H[col * NHP + k] * w3[o * NH + k];
            Y[row * NOUT + o] = y;
        }
    }
}

// backward, one wavefront per 32 rows.  dY (n x NOUT) -> per-block partial
// gradients (deterministic, reduced by k_mlp_reduce):
//   dW3 = dY^T H2, db3; dH2 = (dY W3) * [H2 > 0]; dW2 = dH2^T H1 (MFMA, K = rows);
//   dH1 = (dH2 W2) * [H1 > 0] (MFMA); dW1 = dH1^T X (MFMA, K = rows).
// Gradient vector layout: W1 (NH*NIN), b1 (NH), W2 (NH*NH), b2 (NH), W3 (NOUT*NH), b3 (NOUT).
__global__ __launch_bounds__(64) void k_mlp_bwd(xpbd_mlp_dev net, const float *__restrict__ X,
                                                const float *__restrict__ H1g, const float *__restrict__ H2g,
                                                const float *__restrict__ dY, int64_t n,
                                                float *__restrict__ partial) {
    __shared__ float dH2[32 * NH], dH1[32 * NH], H1s[32 * NH], Xs[32 * 32], dYs[32 * NOUT];
    const int l = threadIdx.x, col = l & 31, h = l >> 5;
    const int64_t r0 = (int64_t)blockIdx.x * 32;
    const int ng = NH * NIN + NH + NH * NH + NH + NOUT * NH + NOUT;
    float *out = partial + (int64_t)blockIdx.x * ng;
    float *gW1 = out, *gb1 = gW1 + NH * NIN, *gW2 = gb1 + NH, *gb2 = gW2 + NH * NH, *gW3 = gb2 + NH,
          *gb3 = gW3 + NOUT * NH;
    for (int i = l; i < 32 * NH; i += 64) {
        const int64_t rr = r0 + i / NH;
        H1s[i] = rr < n ? H1g[rr * NH + i % NH] : 0.f;
        dH2[i] = rr < n ? H2g[rr * NH + i % NH] : 0.f;  // H2 for now
    }
    for (int i = l; i < 32 * 32; i += 64) {
        const int64_t rr = r0 + i / 32;
        const int k = i % 32;
        Xs[i] = (rr < n && k < NIN) ? X[rr * NIN + k] : 0.f;
    }
    for (int i = l; i < 32 * NOUT; i += 64) {
        const int64_t rr = r0 + i / NOUT;
        dYs[i] = rr < n ? dY[rr * NOUT + i % NOUT] : 0.f;
    }
    __syncthreads();
    // layer 3 grads (fp32, fixed order)
    for (int i = l; i < NOUT * NH; i += 64) {
        const int o = i / NH, k = i % NH;
        float s = 0.f;
        for (int r = 0; r < 32; ++r) s += dYs[r * NOUT + o] * dH2[r * NH + k];
        gW3[i] = s;
    }
    if (l < NOUT) {
        float s = 0.f;
        for (int r = 0; r < 32; ++r) s += dYs[r * NOUT + l];
        gb3[l] = s;
    }
    __syncthreads();
    // dH2 = (dY W3) * [H2 > 0]
    for (int i = l; i < 32 * NH; i += 64) {
        const int r = i / NH, k = i % NH;
        float s = 0.f;
        for (int o = 0; o < NOUT; ++o) s += dYs[r * NOUT + o] * net.W3[o * NH + k];
        dH2[i] = dH2[i] > 0.f ? s : 0.f;
    }
    __syncthreads();
    // db2
    for (int j = l; j < NH; j += 64) {
        float s = 0.f;
        for (int r = 0; r < 32; ++r) s += dH2[r * NH + j];
        gb2[j] = s;
    }
    // dW2[j][k] = sum_r dH2[r][j] H1[r][k]: M = j (2 tiles), N = k (2 tiles), K = r (32: 2 steps)
    for (int tj = 0; tj < 2; ++tj)
        for (int tk = 0; tk < 2; ++tk) {
            f32x16 acc;
            for (int i = 0; i < 16; ++i) acc[i] = 0.f;
            for (int s = 0; s < 2; ++s) {
                bf16x8 a, b;
                for (int i = 0; i < 8; ++i) {
                    const int r = 16 * s + 8 * h + i;
                    a[i] = tobf(dH2[r * NH + 32 * tj + col]);
                    b[i] = tobf(H1s[r * NH + 32 * tk + col]);
                }
                acc = mfma(a, b, acc);
            }
            for (int i = 0; i < 16; ++i) gW2[(32 * tj + drow(i, h)) * NH + 32 * tk + col] = acc[i];
        }
    // dH1 = (dH2 W2) * [H1 > 0]: M = r (32), N = k (2 tiles), K = j (64: 4 steps)
    for (int tk = 0; tk < 2; ++tk) {
        f32x16 acc;
        for (int i = 0; i < 16; ++i) acc[i] = 0.f;
        for (int s = 0; s < 4; ++s) {
            bf16x8 a, b;
            for (int i = 0; i < 8; ++i) {
                const int j = 16 * s + 8 * h + i;
                a[i] = tobf(dH2[col * NH + j]);
                b[i] = bfraw(net.W2[j * NH + 32 * tk + col]);
            }
            acc = mfma(a, b, acc);
        }
        for (int i = 0; i < 16; ++i) {
            const int r = drow(i, h), k = 32 * tk + col;
            dH1[r * NH + k] = H1s[r * NH + k] > 0.f ? acc[i] : 0.f;
        }
    }
    __syncthreads();
    for (int j = l; j < NH; j += 64) {
        float s = 0.f;
        for (int r = 0; r < 32; ++r) s += dH1[r * NH + j];
        gb1[j] = s;
    }
    // dW1[j][k] = sum_r dH1[r][j] X[r][k]: M = j (2 tiles), N = k (1 tile of 32, k < NIN kept), K = r
    for (int tj = 0; tj < 2; ++tj) {
        f32x16 acc;
        for (int i = 0; i < 16; ++i) acc[i] = 0.f;
        for (int s = 0; s < 2; ++s) {
            bf16x8 a, b;
            for (int i = 0; i < 8; ++i) {
                const int r = 16 * s + 8 * h + i;
                a[i] = tobf(dH1[r * NH + 32 * tj + col]);
                b[i] = tobf(Xs[r * 32 + col]);
            }
            acc = mfma(a, b, acc);
        }
        if (col < NIN)
            for (int i = 0; i < 16; ++i) gW1[(32 * tj + drow(i, h)) * NIN + col] = acc[i];
    }
}

// sum of per-block partial gradients, fixed order (deterministic)
__global__ void k_mlp_reduce(const float *partial, int nblk, float *grad, int ng) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ng) return;
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += partial[(int64_t)b * ng + i];
    grad[i] = s;
}

constexpr int NG = NH * NIN + NH + NH * NH + NH + NOUT * NH + NOUT;   // gradient floats (6,020)

// Workspace of xpbd_mlp_backward_noalloc: the per-block partial gradients of
// up to XPBD_MLP_NOALLOC_MAX_ROWS rows (512 blocks x 6,020 floats, 12.3 MB on
// each GPU the library runs on, held until the process ends).  k_mlp_bwd
// writes every entry of the blocks it runs and k_mlp_reduce reads only those,
// so nothing has to start at zero.  Host code must use the per-device address
// from hipGetSymbolAddress, never this name.
constexpr int64_t NOALLOC_BLOCKS = XPBD_MLP_NOALLOC_MAX_ROWS / 32;
static __device__ __align__(256) float scratch[NOALLOC_BLOCKS * NG];

constexpr int MAX_DEVICES = 64;
struct ScratchSlot {
    std::once_flag once;
    float *ptr = nullptr;
    hipError_t err = hipSuccess;
};
ScratchSlot g_scratch[MAX_DEVICES];

// device address of the workspace on the current device, looked up once per
// device (each GPU holds its own copy of the array)
int32_t scratch_on_current_device(float **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return XPBD_EHIP;
    ScratchSlot &s = g_scratch[dev];
    std::call_once(s.once, [&s] { s.err = hipGetSymbolAddress((void **)&s.ptr, HIP_SYMBOL(scratch)); });
    if (s.err != hipSuccess || !s.ptr) return XPBD_EHIP;
    *out = s.ptr;
    return XPBD_OK;
}

int32_t launch_backward(const xpbd_mlp_dev *net, const float *X, const float *H1, const float *H2,
                        const float *dY, int64_t n, float *grad, float *partial, void *stream) {
    const int64_t nblk = (n + 31) / 32;
    hipLaunchKernelGGL(k_mlp_bwd, dim3((unsigned)nblk), dim3(64), 0, (hipStream_t)stream, *net, X, H1, H2, dY, n,
                       partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((unsigned)((NG + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const float *)partial, (int)nblk, grad, NG);
    return hipGetLastError() == hipSuccess ? XPBD_OK : XPBD_EHIP;
}

}  // namespace

extern "C" {

int32_t xpbd_mlp_forward(const xpbd_mlp_dev *net, const float *X, int64_t n, float *Y, float *H1, float *H2,
                         void *stream) {
    if (!net || !X || !Y || n < 0) return XPBD_EINVAL;
    if (n == 0) return XPBD_OK;
    const int64_t waves = (n + 31) / 32;
    const unsigned blocks = (unsigned)((waves + 3) / 4);
    hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks), dim3(256), 0, (hipStream_t)stream, *net, X, n, Y, H1, H2);
    return hipGetLastError() == hipSuccess ? XPBD_OK : XPBD_EHIP;
}

int32_t xpbd_mlp_grad_size(void) { return NG; }

int32_t xpbd_mlp_backward(const xpbd_mlp_dev *net, const float *X, const float *H1, const float *H2,
                          const float *dY, int64_t n, float *grad, float *work, int64_t work_floats,
                          void *stream) {
    if (!net || !X || !H1 || !H2 || !dY || !grad || !work || n <= 0) return XPBD_EINVAL;
    const int64_t nblk = (n + 31) / 32;
    if (nblk * NG > work_floats) return XPBD_EINVAL;
    return launch_backward(net, X, H1, H2, dY, n, grad, work, stream);
}

int32_t xpbd_mlp_backward_noalloc(const xpbd_mlp_dev *net, const float *X, const float *H1, const float *H2,
                                  const float *dY, int64_t n, float *grad, void *stream) {
    if (!net || !X || !H1 || !H2 || !dY || !grad || n <= 0) return XPBD_EINVAL;
    if (n > XPBD_MLP_NOALLOC_MAX_ROWS) return XPBD_EINVAL;   // beyond the static workspace
    float *ws = nullptr;
    const int32_t r = scratch_on_current_device(&ws);
    if (r) return r;
    return launch_backward(net, X, H1, H2, dY, n, grad, ws, stream);
}

}  // extern "C"
