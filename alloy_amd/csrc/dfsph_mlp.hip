// Learned warm start of the pressure solves: a per-particle MLP 13 -> 64 -> 64 -> 64 -> 2 (SiLU), bf16 weights and
// activations on the MFMA units with float32 accumulation (v_mfma_f32_16x16x16_bf16), forward for inference and
// forward + backward for training, with gradients reduced deterministically (fixed-order sums of per-wave
// partials, no float atomics) and an Adam update. The network predicts the initial kappa of the divergence and the
// density solve in place of solve_init's warm-start rule (pressure.py:254, 261); the solves still iterate to the
// same tolerance.
//
// Parameters are float32 masters in one buffer (layout MLP_* below, weights row-major [out][in]); every kernel rounds
// the weights to bf16 (round to nearest even) as it stages them in LDS.
//
// MFMA tile (16 x 16 x 16, bf16 in, f32 out), lane l of the 64-lane wavefront:
//   A (M x K): A[l % 16][4 (l / 16) + i], i = 0..3        B (K x N): B[4 (l / 16) + i][l % 16]
//   D (M x N): D[4 (l / 16) + i][l % 16]
// One wave owns 16 samples (rows); a 256-thread block owns 64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/dfsph_abi.h"

namespace {

constexpr int MLP_IN = 16;   // 13 features, zero padded
constexpr int MLP_IN_USED = 13;
constexpr int MLP_H = 64;
constexpr int MLP_OUT = 16;  // 2 outputs, zero padded
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_H * MLP_IN;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;
constexpr int OFF_W3 = OFF_B2 + MLP_H;
constexpr int OFF_B3 = OFF_W3 + MLP_H * MLP_H;
constexpr int OFF_W4 = OFF_B3 + MLP_H;
constexpr int OFF_B4 = OFF_W4 + MLP_OUT * MLP_H;
constexpr int N_PARAMS = OFF_B4 + MLP_OUT;  // 10448
constexpr int PART_STRIDE = N_PARAMS + 1;   // per-wave gradient partial + its squared-error sum
// LDS row strides (bf16 elements) of 16- and 64-wide rows, padded so that the 16 rows one MFMA fragment load
// touches start on distinct bank pairs (10 and 34 words): an unpadded 64-wide row (32 words) puts all 16 on one bank
constexpr int LS_IN = 20, LS_H = 68;
// bf16 weight staging in LDS: W1 [64][LS_IN] | W2 [64][LS_H] | W3 [64][LS_H] | W4 [16][LS_H]
constexpr int SW1 = 0, SW2 = SW1 + MLP_H * LS_IN, SW3 = SW2 + MLP_H * LS_H, SW4 = SW3 + MLP_H * LS_H;
constexpr int SW_TOTAL = SW4 + MLP_OUT * LS_H;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even: gfx950's v_cvt_pk_bf16_f32
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f)); }

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// sigmoid by the hardware exp and reciprocal (about 1 ulp each): the activations are rounded to bf16 right after
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
    float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

// stage the four weight matrices as bf16 in LDS (block-wide)
__device__ void stage_weights(const float* __restrict__ p, uint16_t* sW) {
    constexpr int N_W = MLP_H * MLP_IN + 2 * MLP_H * MLP_H + MLP_OUT * MLP_H;  // 10,240 weights
    static_assert(SW4 + (MLP_OUT - 1) * LS_H + MLP_H <= SW_TOTAL, "staged W4 fits");
    for (int t = threadIdx.x; t < N_W; t += blockDim.x) {
        // t runs over the weights in parameter order; dst is its padded LDS slot
        int dst;
        if (t < MLP_H * MLP_IN) dst = SW1 + (t / MLP_IN) * LS_IN + t % MLP_IN;
        else if (t < MLP_H * (MLP_IN + MLP_H)) {
            int u = t - MLP_H * MLP_IN;
            dst = SW2 + (u / MLP_H) * LS_H + u % MLP_H;
        } else if (t < MLP_H * (MLP_IN + 2 * MLP_H)) {
            int u = t - MLP_H * (MLP_IN + MLP_H);
            dst = SW3 + (u / MLP_H) * LS_H + u % MLP_H;
        } else {
            int u = t - MLP_H * (MLP_IN + 2 * MLP_H);
            dst = SW4 + (u / MLP_H) * LS_H + u % MLP_H;
        }
        int src = t < MLP_H * MLP_IN ? OFF_W1 + t
                  : t < MLP_H * (MLP_IN + MLP_H) ? OFF_W2 + (t - MLP_H * MLP_IN)
                  : t < MLP_H * (MLP_IN + 2 * MLP_H) ? OFF_W3 + (t - MLP_H * (MLP_IN + MLP_H))
                                                     : OFF_W4 + (t - MLP_H * (MLP_IN + 2 * MLP_H));
        sW[dst] = f2bf(p[src]);
    }
}

// A fragment: row-major [16][lda] bf16 tile, rows = samples, columns k
__device__ __forceinline__ s4 frag_rows(const uint16_t* A, int lda, int ks, int lane) {
    return *reinterpret_cast<const s4*>(A + (lane & 15) * lda + 16 * ks + 4 * (lane >> 4));
}
// B fragment of W^T for W row-major [N][ldw]: B[k][n] = W[n][k] (contiguous in W's row)
__device__ __forceinline__ s4 frag_wt(const uint16_t* W, int ldw, int ks, int nt, int lane) {
    return *reinterpret_cast<const s4*>(W + (16 * nt + (lane & 15)) * ldw + 16 * ks + 4 * (lane >> 4));
}
// B fragment of W for W row-major [K][ldw]: B[k][n] = W[k][n] (strided)
__device__ __forceinline__ s4 frag_w(const uint16_t* W, int ldw, int ks, int nt, int lane) {
    s4 r;
    const uint16_t* base = W + (16 * ks + 4 * (lane >> 4)) * ldw + 16 * nt + (lane & 15);
    r.x = short(base[0]);
    r.y = short(base[ldw]);
    r.z = short(base[2 * ldw]);
    r.w = short(base[3 * ldw]);
    return r;
}
// A fragment of M^T for M row-major [16 samples][ldm]: A[m][k] = M[k][16 mt + m] (strided)
__device__ __forceinline__ s4 frag_rows_t(const uint16_t* M, int ldm, int mt, int lane) {
    s4 r;
    const uint16_t* base = M + (4 * (lane >> 4)) * ldm + 16 * mt + (lane & 15);
    r.x = short(base[0]);
    r.y = short(base[ldm]);
    r.z = short(base[2 * ldm]);
    r.w = short(base[3 * ldm]);
    return r;
}
// B fragment of M for M row-major [16 samples][ldm]: B[k = sample][n] = M[k][16 nt + n] (strided)
__device__ __forceinline__ s4 frag_samples(const uint16_t* M, int ldm, int nt, int lane) {
    return frag_rows_t(M, ldm, nt, lane);
}

// D tile (4 floats per lane) -> row-major bf16 tile in LDS
__device__ __forceinline__ void store_d_bf16(uint16_t* T, int ld, int nt, int lane, f4 v) {
    int col = 16 * nt + (lane & 15), r0 = 4 * (lane >> 4);
    T[(r0 + 0) * ld + col] = f2bf(v.x);
    T[(r0 + 1) * ld + col] = f2bf(v.y);
    T[(r0 + 2) * ld + col] = f2bf(v.z);
    T[(r0 + 3) * ld + col] = f2bf(v.w);
}

struct WaveTiles {  // row strides LS_IN / LS_H
    uint16_t a0[16 * LS_IN];
    uint16_t a1[16 * LS_H];
    uint16_t a2[16 * LS_H];
    uint16_t a3[16 * LS_H];
    uint16_t g[16 * LS_H];   // backward: dZ of the current layer (bf16 operand)
    uint16_t gy[16 * LS_IN];
};

// forward of one wave's 16 samples: z1..z3 kept in registers (D layout), activations in LDS; returns y (D layout)
__device__ f4 forward_wave(const float* __restrict__ p, const uint16_t* sW, WaveTiles& T, int lane, f4 z1[4],
                           f4 z2[4], f4 z3[4]) {
    // layer 1: 16 -> 64
    for (int nt = 0; nt < 4; ++nt) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(frag_rows(T.a0, LS_IN, 0, lane), frag_wt(sW + SW1, LS_IN, 0, nt, lane), acc);
        float b = p[OFF_B1 + 16 * nt + (lane & 15)];
        acc += b;
        z1[nt] = acc;
        f4 a = {silu(acc.x), silu(acc.y), silu(acc.z), silu(acc.w)};
        store_d_bf16(T.a1, LS_H, nt, lane, a);
    }
    // layer 2
    for (int nt = 0; nt < 4; ++nt) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 4; ++ks) acc = mfma(frag_rows(T.a1, LS_H, ks, lane), frag_wt(sW + SW2, LS_H, ks, nt, lane), acc);
        acc += p[OFF_B2 + 16 * nt + (lane & 15)];
        z2[nt] = acc;
        f4 a = {silu(acc.x), silu(acc.y), silu(acc.z), silu(acc.w)};
        store_d_bf16(T.a2, LS_H, nt, lane, a);
    }
    // layer 3
    for (int nt = 0; nt < 4; ++nt) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 4; ++ks) acc = mfma(frag_rows(T.a2, LS_H, ks, lane), frag_wt(sW + SW3, LS_H, ks, nt, lane), acc);
        acc += p[OFF_B3 + 16 * nt + (lane & 15)];
        z3[nt] = acc;
        f4 a = {silu(acc.x), silu(acc.y), silu(acc.z), silu(acc.w)};
        store_d_bf16(T.a3, LS_H, nt, lane, a);
    }
    // layer 4: 64 -> 16 (2 used), linear
    f4 y = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 4; ++ks) y = mfma(frag_rows(T.a3, LS_H, ks, lane), frag_wt(sW + SW4, LS_H, ks, 0, lane), y);
    y += p[OFF_B4 + (lane & 15)];
    return y;
}

// stage a wave's input rows [16][MLP_IN] as bf16: the first MLP_IN_USED columns of X (rows past n and the padding
// columns, which carry other per-particle data, are zero)
__device__ void load_inputs(const float* __restrict__ X, int n, int row0, uint16_t* a0, int lane) {
    for (int t = lane; t < 16 * MLP_IN; t += 64) {
        int r = t / MLP_IN, k = t % MLP_IN;
        float v = (row0 + r < n && k < MLP_IN_USED) ? X[int64_t(row0 + r) * MLP_IN + k] : 0.0f;
        a0[r * LS_IN + k] = f2bf(v);
    }
}

// Y = MLP(X) for rows [0, n): the block stages the weights in LDS once and then walks 64-row tiles (16 per wave)
// with a grid stride, so the 42 KB of weights are read once per block rather than once per 64 rows. With tile_on, the
// tiles whose flag is 0 are skipped (their rows of Y are left as they are).
constexpr int FWD_MAX_BLOCKS = 2048;
__global__ void __launch_bounds__(256) k_mlp_forward(const float* __restrict__ p, const float* __restrict__ X, int n,
                                                     float* __restrict__ Y, const uint8_t* __restrict__ tile_on) {
    __shared__ uint16_t sW[SW_TOTAL];
    __shared__ WaveTiles tiles[4];
    stage_weights(p, sW);
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WaveTiles& T = tiles[w];
    const int n_tiles = (n + 63) / 64;
    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {  // block-uniform trip count
        if (tile_on && !tile_on[tile]) continue;  // block-uniform
        int row0 = tile * 64 + 16 * w;
        __syncthreads();  // weights staged (first pass); the previous tile's LDS reads done
        load_inputs(X, n, row0, T.a0, lane);
        __syncthreads();
        f4 z1[4], z2[4], z3[4];
        f4 y = forward_wave(p, sW, T, lane, z1, z2, z3);
        int col = lane & 15;
        if (col < 2) {
            float v[4] = {y.x, y.y, y.z, y.w};
            for (int i = 0; i < 4; ++i) {
                int r = row0 + 4 * (lane >> 4) + i;
                if (r < n) Y[int64_t(r) * 2 + col] = v[i];
            }
        }
    }
}
inline int fwd_blocks(int n) { return std::min((n + 63) / 64, FWD_MAX_BLOCKS); }

// sum over the 16 rows of a D tile, per column (lanes 0..15 hold column totals after the shuffles)
__device__ __forceinline__ float col_sum(f4 v) {
    float s = v.x + v.y + v.z + v.w;
    s += __shfl_xor(s, 16, 64);
    s += __shfl_xor(s, 32, 64);
    return s;
}

// dW[M = out][N = in] = dZ^T (16 samples) x A_prev: dZ tile g [16][ld_out], A_prev [16][ld_in]
__device__ void weight_grad(const uint16_t* g, int ld_out, int m_tiles, const uint16_t* aprev, int ld_in, int n_tiles,
                            float* out, int ldw, int lane) {
    for (int mt = 0; mt < m_tiles; ++mt)
        for (int nt = 0; nt < n_tiles; ++nt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(frag_rows_t(g, ld_out, mt, lane), frag_samples(aprev, ld_in, nt, lane), acc);
            int col = 16 * nt + (lane & 15), r0 = 16 * mt + 4 * (lane >> 4);
            out[(r0 + 0) * ldw + col] = acc.x;
            out[(r0 + 1) * ldw + col] = acc.y;
            out[(r0 + 2) * ldw + col] = acc.z;
            out[(r0 + 3) * ldw + col] = acc.w;
        }
}

// Forward + backward of the mean squared error over 2 outputs for rows [0, n): per-wave partial gradients
// part[slot][N_PARAMS] and the wave's squared-error sum part[slot][N_PARAMS]; inv_count = 1 / (2 n_total).
__global__ void __launch_bounds__(256) k_mlp_grad(const float* __restrict__ p, const float* __restrict__ X,
                                                  const float* __restrict__ Tgt, int n, float inv_count,
                                                  float* __restrict__ part) {
    __shared__ uint16_t sW[SW_TOTAL];
    __shared__ WaveTiles tiles[4];
    stage_weights(p, sW);
    __syncthreads();
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int row0 = blockIdx.x * 64 + 16 * w;
    WaveTiles& T = tiles[w];
    load_inputs(X, n, row0, T.a0, lane);
    __syncthreads();
    f4 z1[4], z2[4], z3[4];
    f4 y = forward_wave(p, sW, T, lane, z1, z2, z3);
    float* out = part + int64_t(blockIdx.x * 4 + w) * PART_STRIDE;
    // output error: dY = 2 (y - t) / (2 n_total) on the 2 real outputs of valid rows
    int col = lane & 15;
    float yv[4] = {y.x, y.y, y.z, y.w}, dy[4], se = 0.0f;
    for (int i = 0; i < 4; ++i) {
        int r = row0 + 4 * (lane >> 4) + i;
        float d = 0.0f;
        if (col < 2 && r < n) {
            d = yv[i] - Tgt[int64_t(r) * 2 + col];
            se += d * d;
        }
        dy[i] = 2.0f * d * inv_count;
    }
    for (int s = 32; s >= 1; s >>= 1) se += __shfl_xor(se, s, 64);
    if (lane == 0) out[N_PARAMS] = se;
    f4 gY = {dy[0], dy[1], dy[2], dy[3]};
    store_d_bf16(T.gy, LS_IN, 0, lane, gY);
    {
        float s = col_sum(gY);
        if (lane < 16) out[OFF_B4 + lane] = s;
    }
    __syncthreads();
    // layer 4: dW4 = dY^T a3, da3 = dY W4
    weight_grad(T.gy, LS_IN, 1, T.a3, LS_H, 4, out + OFF_W4, MLP_H, lane);
    f4 dz[4];
    for (int nt = 0; nt < 4; ++nt) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(frag_rows(T.gy, LS_IN, 0, lane), frag_w(sW + SW4, LS_H, 0, nt, lane), acc);
        f4 z = z3[nt];
        dz[nt] = f4{acc.x * silu_grad(z.x), acc.y * silu_grad(z.y), acc.z * silu_grad(z.z), acc.w * silu_grad(z.w)};
    }
    // layers 3, 2, 1
    const uint16_t* aprev[3] = {T.a2, T.a1, T.a0};
    const int woff[3] = {OFF_W3, OFF_W2, OFF_W1}, boff[3] = {OFF_B3, OFF_B2, OFF_B1}, sw[3] = {SW3, SW2, SW1};
    f4* zs[2] = {z2, z1};
    for (int L = 0; L < 3; ++L) {
        __syncthreads();
        for (int nt = 0; nt < 4; ++nt) {
            store_d_bf16(T.g, LS_H, nt, lane, dz[nt]);
            float s = col_sum(dz[nt]);
            if (lane < 16) out[boff[L] + 16 * nt + lane] = s;
        }
        __syncthreads();
        int in_dim = L == 2 ? MLP_IN : MLP_H, in_ls = L == 2 ? LS_IN : LS_H;
        weight_grad(T.g, LS_H, 4, aprev[L], in_ls, in_dim / 16, out + woff[L], in_dim, lane);
        if (L == 2) break;
        // da_prev = dZ W (W [64 out][64 in])
        for (int nt = 0; nt < 4; ++nt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int ks = 0; ks < 4; ++ks)
                acc = mfma(frag_rows(T.g, LS_H, ks, lane), frag_w(sW + sw[L], LS_H, ks, nt, lane), acc);
            f4 z = zs[L][nt];
            dz[nt] = f4{acc.x * silu_grad(z.x), acc.y * silu_grad(z.y), acc.z * silu_grad(z.z), acc.w * silu_grad(z.w)};
        }
    }
}

// grad[j] = sum of the slots' partials in slot order (deterministic); j = N_PARAMS gives the squared-error sum
__global__ void k_mlp_reduce(const float* __restrict__ part, int slots, float* __restrict__ grad) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > N_PARAMS) return;
    float s = 0.0f;
    for (int k = 0; k < slots; ++k) s += part[int64_t(k) * PART_STRIDE + j];
    grad[j] = s;
}

// Adam (Kingma & Ba 2015) on the float32 masters; padded weight rows/columns stay zero
__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N_PARAMS) return;
    bool pad = false;
    if (j < OFF_B1) pad = (j % MLP_IN) >= 13;                                    // W1 input columns 13..15
    else if (j >= OFF_W4 && j < OFF_B4) pad = (j - OFF_W4) / MLP_H >= 2;         // W4 rows 2..15
    else if (j >= OFF_B4) pad = (j - OFF_B4) >= 2;                               // b4 entries 2..15
    if (pad) {
        p[j] = 0.0f;
        return;
    }
    float gj = g[j];
    float mj = b1 * m[j] + (1.0f - b1) * gj;
    float vj = b2 * v[j] + (1.0f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    p[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
}

thread_local std::string g_mlp_err;

// The gradient partials of dfsph_mlp_grad_noalloc: a file-scope device array (one copy per GPU, held until the process
// ends) sized for DFSPH_MLP_NOALLOC_MAX_SAMPLES samples. Host code never passes the array's name to a kernel: it is a
// host-side stand-in, not the device address. Each device's address comes from hipGetSymbolAddress on that
// device's first call and is kept in g_noalloc_ptr.
constexpr int64_t NOALLOC_FLOATS = int64_t(4) * ((DFSPH_MLP_NOALLOC_MAX_SAMPLES + 63) / 64) * PART_STRIDE;
static __device__ __align__(256) float scratch[NOALLOC_FLOATS];
constexpr int MAX_DEVICES = 64;
float* g_noalloc_ptr[MAX_DEVICES] = {};
std::mutex g_noalloc_mu;

// the current device's address of `scratch` (looked up once per device, with that device current)
int noalloc_scratch(float** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return DFSPH_EDEVICE;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return DFSPH_EDEVICE;
        g_noalloc_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_noalloc_ptr[dev];
    return DFSPH_OK;
}

}  // namespace

// launched by the step (dfsph_kernels.hip) inside its captured segment
void dfsph_launch_mlp_forward(const float* params, const float* X, int n, float* Y, const uint8_t* tile_on,
                              hipStream_t st) {
    if (n > 0) k_mlp_forward<<<fwd_blocks(n), 256, 0, st>>>(params, X, n, Y, tile_on);
}

extern "C" {

int32_t dfsph_mlp_param_count(void) { return N_PARAMS; }

// Y[n][2] = MLP(X[n][16]) (device pointers)
int32_t dfsph_mlp_forward(const float* params, const float* X, int32_t n, float* Y, void* stream) {
    if (!params || !X || !Y || n < 0) return DFSPH_EINVAL;
    if (n == 0) return DFSPH_OK;
    k_mlp_forward<<<fwd_blocks(n), 256, 0, static_cast<hipStream_t>(stream)>>>(params, X, n, Y, nullptr);
    return hipGetLastError() == hipSuccess ? DFSPH_OK : DFSPH_EDEVICE;
}

// one partial per wave: 4 waves per 64-sample block, each N_PARAMS gradient floats + its squared-error sum
int64_t dfsph_mlp_grad_scratch_floats(int32_t n) {
    return n <= 0 ? 0 : int64_t(4) * ((int64_t(n) + 63) / 64) * PART_STRIDE;
}

// grad: float [N_PARAMS + 1] (last: squared-error sum)
int32_t dfsph_mlp_grad(const float* params, const float* X, const float* T, int32_t n, int32_t n_total, float* scratch,
                       int64_t scratch_floats, float* grad, void* stream) {
    if (!params || !X || !T || !scratch || !grad || n <= 0 || n_total <= 0) return DFSPH_EINVAL;
    if (scratch_floats < dfsph_mlp_grad_scratch_floats(n)) return DFSPH_EINVAL;
    hipStream_t st = static_cast<hipStream_t>(stream);
    int blocks = (n + 63) / 64;
    k_mlp_grad<<<blocks, 256, 0, st>>>(params, X, T, n, 1.0f / (2.0f * float(n_total)), scratch);
    k_mlp_reduce<<<(N_PARAMS + 1 + 255) / 256, 256, 0, st>>>(scratch, blocks * 4, grad);
    return hipGetLastError() == hipSuccess ? DFSPH_OK : DFSPH_EDEVICE;
}

// dfsph_mlp_grad with the partials in the library's static device array instead of a caller workspace: no
// allocation at all. n <= DFSPH_MLP_NOALLOC_MAX_SAMPLES. Every call on a device shares its array (calls that overlap
// on one device race); each launch writes every partial it then reduces, so nothing needs clearing between calls.
int32_t dfsph_mlp_grad_noalloc(const float* params, const float* X, const float* T, int32_t n, int32_t n_total,
                               float* grad, void* stream) {
    if (!params || !X || !T || !grad || n <= 0 || n_total <= 0) return DFSPH_EINVAL;
    if (n > DFSPH_MLP_NOALLOC_MAX_SAMPLES) return DFSPH_ECAPACITY;
    float* part = nullptr;
    int rc = noalloc_scratch(&part);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    int blocks = (n + 63) / 64;
    k_mlp_grad<<<blocks, 256, 0, st>>>(params, X, T, n, 1.0f / (2.0f * float(n_total)), part);
    k_mlp_reduce<<<(N_PARAMS + 1 + 255) / 256, 256, 0, st>>>(part, blocks * 4, grad);
    return hipGetLastError() == hipSuccess ? DFSPH_OK : DFSPH_EDEVICE;
}

int32_t dfsph_mlp_adam(float* params, const float* grad, float* m, float* v, double lr, double b1, double b2,
                       double eps, int32_t t, void* stream) {
    if (!params || !grad || !m || !v || t < 1) return DFSPH_EINVAL;
    double c1 = 1.0 - std::pow(b1, double(t)), c2 = 1.0 - std::pow(b2, double(t));
    k_mlp_adam<<<(N_PARAMS + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(
        params, grad, m, v, float(lr), float(b1), float(b2), float(eps), float(c1), float(c2));
    return hipGetLastError() == hipSuccess ? DFSPH_OK : DFSPH_EDEVICE;
}

}  // extern "C"
