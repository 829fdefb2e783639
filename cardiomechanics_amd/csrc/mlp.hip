// Learned Newton warm start (include/cmg.h, cmg_mlp_*): a per-node MLP
// 8 -> H -> H -> 3 (tanh hidden layers) that maps nodal features of the next
// Newmark step (the two previous increments, the pressure rate and the activation
// phase) to that step's displacement increment.  Hand-written gfx950 kernels for
// inference, the backward pass (activation gradients and weight-gradient
// reductions, two-stage and in a fixed order) and Adam.  The step only uses the
// proposed guess when its residual is below the predictor's (the predictor is both
// the fallback and the yardstick).
//
// Weight layout (FP32, flat): W1[8][H] b1[H] W2[H][H] b2[H] W3[H][3] b3[3]
// (W[k][c]: input k -> output c).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "cmg_ctx.h"

using namespace cmg;

namespace {

constexpr int kH = 128;         // hidden width of the kernels
constexpr int kIn = 8;
constexpr int kOut = 3;
constexpr int kTile = 64;       // rows per tile
constexpr int kThr = 256;       // threads per workgroup
constexpr int kPad = kH + 1;    // LDS row pitch (bank-conflict-free column walks)
constexpr int kGradBlocks = 128;

constexpr int64_t n_weights() { return kIn * kH + kH + kH * kH + kH + kH * kOut + kOut; }
constexpr int64_t off_b1 = kIn * kH;
constexpr int64_t off_W2 = off_b1 + kH;
constexpr int64_t off_b2 = off_W2 + kH * kH;
constexpr int64_t off_W3 = off_b2 + kH;
constexpr int64_t off_b3 = off_W3 + kH * kOut;

__device__ __forceinline__ float bf16_round(float x) {
    // round to nearest even at bfloat16 precision (finite inputs)
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

// ---------------------------------------------------------------------------
// forward (FP32): grid-stride over 64-row tiles; W2 stays in LDS for all tiles of a
// block.  Thread t owns rows 4*(t/16)..+4 and hidden columns 8*(t%16)..+8.
// store_act: write the hidden activations A1, A2 [rows][H] (training).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThr) void k_mlp_forward(const float* __restrict__ w, const float* __restrict__ X,
                                                      int64_t rows, float* __restrict__ Y, float* __restrict__ A1g,
                                                      float* __restrict__ A2g) {
    __shared__ float sW2[kH * kH];
    __shared__ float sW1[kIn * kH];
    __shared__ float sb1[kH], sb2[kH];
    __shared__ float sW3[kH * kOut];
    __shared__ float sb3[kOut];
    __shared__ float sA[kTile * kPad];
    const int t = threadIdx.x;
    for (int i = t; i < kH * kH; i += kThr) sW2[i] = w[off_W2 + i];
    for (int i = t; i < kIn * kH; i += kThr) sW1[i] = w[i];
    for (int i = t; i < kH * kOut; i += kThr) sW3[i] = w[off_W3 + i];
    if (t < kH) {
        sb1[t] = w[off_b1 + t];
        sb2[t] = w[off_b2 + t];
    }
    if (t < kOut) sb3[t] = w[off_b3 + t];
    __syncthreads();
    const int rg = t / 16, cg = t % 16;
    const int64_t ntiles = (rows + kTile - 1) / kTile;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t r0 = tile * kTile;
        // ---- layer 1 ----
        float acc[4][8];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t r = r0 + rg * 4 + i;
            float x[kIn];
#pragma unroll
            for (int k = 0; k < kIn; k++) x[k] = (r < rows) ? X[r * kIn + k] : 0.f;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int c = cg * 8 + j;
                float s = sb1[c];
#pragma unroll
                for (int k = 0; k < kIn; k++) s += x[k] * sW1[k * kH + c];
                acc[i][j] = tanhf(s);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t r = r0 + rg * 4 + i;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                sA[(rg * 4 + i) * kPad + cg * 8 + j] = acc[i][j];
                if (A1g && r < rows) A1g[r * kH + cg * 8 + j] = acc[i][j];
            }
        }
        __syncthreads();
        // ---- layer 2 ----
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 8; j++) acc[i][j] = sb2[cg * 8 + j];
        for (int k = 0; k < kH; k++) {
            float a[4], b[8];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = sA[(rg * 4 + i) * kPad + k];
#pragma unroll
            for (int j = 0; j < 8; j++) b[j] = sW2[k * kH + cg * 8 + j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 8; j++) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();   // every thread is done reading A1 from LDS
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t r = r0 + rg * 4 + i;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const float h = tanhf(acc[i][j]);
                sA[(rg * 4 + i) * kPad + cg * 8 + j] = h;
                if (A2g && r < rows) A2g[r * kH + cg * 8 + j] = h;
            }
        }
        __syncthreads();
        // ---- layer 3: 64 rows x 3 outputs on threads 0..191 ----
        if (t < kTile * kOut) {
            const int rr = t / kOut, o = t % kOut;
            const int64_t r = r0 + rr;
            float s = sb3[o];
            for (int k = 0; k < kH; k++) s += sA[rr * kPad + k] * sW3[k * kOut + o];
            if (r < rows) Y[r * kOut + o] = s;
        }
        __syncthreads();   // sA is rewritten by the next tile
    }
}

// ---------------------------------------------------------------------------
// bfloat16 inference: the 128 x 128 layer on the matrix cores
// (v_mfma_f32_16x16x16_bf16, FP32 accumulation), layers 1 and 3 in FP32 on the vector
// ALUs; inputs, weights and both hidden activations rounded to bfloat16.  One wave per
// 16 rows (four waves, 64 rows per tile).  Operand layout of the instruction (index
// m % 4 is the item within a lane's four values):
//   A[i][k] in lane i + 16 (k / 4),  B[k][j] in lane j + 16 (k / 4),  D[i][j] in lane j + 16 (i / 4).
// A lane therefore computes its own layer-1 values straight into its A operands
// (row i = lane % 16, columns 16 ks + 4 (lane / 16) + 0..3 for k-step ks), reads B from
// W2 stored transposed in LDS (four consecutive k in 8 bytes), and ends with the
// layer-2 outputs of rows 4 (lane / 16) + 0..3 and columns 16 nb + lane % 16, whose
// layer-3 products are summed over the 16 lanes of a row group.
// ---------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kW2tPitch = kH + 8;   // shorts per row of the transposed W2 (spreads the lanes over banks)

__device__ __forceinline__ short bf16_bits(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
    return (short)(u >> 16);
}

__global__ __launch_bounds__(kThr) void k_mlp_forward_mfma(const float* __restrict__ w, const float* __restrict__ X,
                                                           int64_t rows, float* __restrict__ Y) {
    __shared__ __align__(16) short sW2t[kH * kW2tPitch];   // sW2t[n * pitch + k] = bf16(W2[k][n])
    __shared__ float sW1[kIn * kH];
    __shared__ float sb1[kH], sb2[kH];
    __shared__ float sW3[kH * kOut];
    __shared__ float sb3[kOut];
    const int t = threadIdx.x;
    for (int i = t; i < kH * kH; i += kThr) sW2t[(i % kH) * kW2tPitch + i / kH] = bf16_bits(w[off_W2 + i]);
    for (int i = t; i < kIn * kH; i += kThr) sW1[i] = bf16_round(w[i]);
    for (int i = t; i < kH * kOut; i += kThr) sW3[i] = bf16_round(w[off_W3 + i]);
    if (t < kH) {
        sb1[t] = w[off_b1 + t];
        sb2[t] = w[off_b2 + t];
    }
    if (t < kOut) sb3[t] = w[off_b3 + t];
    __syncthreads();
    const int wv = t / 64, lane = t % 64;
    const int li = lane % 16, lk = lane / 16;
    const int64_t ntiles = (rows + kTile - 1) / kTile;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t r0 = tile * kTile + wv * 16;
        // ---- layer 1 into the A operands ----
        const int64_t ra = r0 + li;
        float x[kIn];
#pragma unroll
        for (int k = 0; k < kIn; k++) x[k] = (ra < rows) ? bf16_round(X[ra * kIn + k]) : 0.f;
        bf16x4 a[kH / 16];
#pragma unroll
        for (int ks = 0; ks < kH / 16; ks++)
#pragma unroll 1
            for (int j = 0; j < 4; j++) {
                const int c = 16 * ks + 4 * lk + j;
                float s = sb1[c];
#pragma unroll
                for (int k = 0; k < kIn; k++) s += x[k] * sW1[k * kH + c];
                a[ks][j] = bf16_bits(tanhf(s));
            }
        // ---- layers 2 (matrix cores) and 3, one 16-column block of layer 2 at a time ----
        float y[4][kOut];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int o = 0; o < kOut; o++) y[r][o] = 0.f;
#pragma unroll 1
        for (int nb = 0; nb < kH / 16; nb++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < kH / 16; ks++) {
                const bf16x4 b = *reinterpret_cast<const bf16x4*>(&sW2t[(16 * nb + li) * kW2tPitch + 16 * ks + 4 * lk]);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ks], b, acc, 0, 0, 0);
            }
            const int n = 16 * nb + li;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float h = bf16_round(tanhf(acc[r] + sb2[n]));
#pragma unroll
                for (int o = 0; o < kOut; o++) y[r][o] += h * sW3[n * kOut + o];
            }
        }
#pragma unroll
        for (int off = 8; off > 0; off >>= 1)
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int o = 0; o < kOut; o++) y[r][o] += __shfl_xor(y[r][o], off, 16);
        if (li == 0)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = r0 + 4 * lk + r;
                if (row < rows)
#pragma unroll
                    for (int o = 0; o < kOut; o++) Y[row * kOut + o] = y[r][o] + sb3[o];
            }
    }
}

// ---------------------------------------------------------------------------
// backward, activation part: dY = 2 (Y - T) / (3 rows); dZ2 = (dY W3^T) (1 - A2^2);
// dZ1 = (dZ2 W2^T) (1 - A1^2).  Per-block partial of the squared error.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThr) void k_mlp_backward_act(const float* __restrict__ w, int64_t rows,
                                                           const float* __restrict__ Y, const float* __restrict__ T,
                                                           const float* __restrict__ A1g,
                                                           const float* __restrict__ A2g, float* __restrict__ dYg,
                                                           float* __restrict__ dZ2g, float* __restrict__ dZ1g,
                                                           double* __restrict__ loss_part) {
    __shared__ float sW2[kH * kPad];   // padded: the W2^T product walks rows of W2
    __shared__ float sW3[kH * kOut];
    __shared__ float sdY[kTile * kOut];
    __shared__ float sZ[kTile * kPad];
    __shared__ double sl[kThr / 64];
    const int t = threadIdx.x;
    for (int i = t; i < kH * kH; i += kThr) sW2[(i / kH) * kPad + (i % kH)] = w[off_W2 + i];
    for (int i = t; i < kH * kOut; i += kThr) sW3[i] = w[off_W3 + i];
    __syncthreads();
    const int rg = t / 16, cg = t % 16;
    const int64_t ntiles = (rows + kTile - 1) / kTile;
    const float scale = 2.0f / (float)(3 * rows);
    double lsum = 0.0;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t r0 = tile * kTile;
        if (t < kTile * kOut) {
            const int64_t r = r0 + t / kOut;
            float d = 0.f;
            if (r < rows) {
                const float e = Y[r * kOut + t % kOut] - T[r * kOut + t % kOut];
                lsum += (double)e * (double)e;
                d = scale * e;
                dYg[r * kOut + t % kOut] = d;
            }
            sdY[t] = d;
        }
        __syncthreads();
        // dZ2 for rows rg*4..+4, columns cg*8..+8
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int rr = rg * 4 + i;
            const int64_t r = r0 + rr;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int c = cg * 8 + j;
                float da = 0.f;
#pragma unroll
                for (int o = 0; o < kOut; o++) da += sdY[rr * kOut + o] * sW3[c * kOut + o];
                const float a2 = (r < rows) ? A2g[r * kH + c] : 0.f;
                const float dz = da * (1.f - a2 * a2);
                sZ[rr * kPad + c] = dz;
                if (r < rows) dZ2g[r * kH + c] = dz;
            }
        }
        __syncthreads();
        // dZ1 = (dZ2 W2^T)(1 - A1^2): dA1[r][c] = sum_k dZ2[r][k] W2[c][k]
        float acc[4][8];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 8; j++) acc[i][j] = 0.f;
        for (int k = 0; k < kH; k++) {
            float a[4], b[8];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = sZ[(rg * 4 + i) * kPad + k];
#pragma unroll
            for (int j = 0; j < 8; j++) b[j] = sW2[(cg * 8 + j) * kPad + k];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 8; j++) acc[i][j] += a[i] * b[j];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t r = r0 + rg * 4 + i;
            if (r >= rows) continue;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int c = cg * 8 + j;
                const float a1 = A1g[r * kH + c];
                dZ1g[r * kH + c] = acc[i][j] * (1.f - a1 * a1);
            }
        }
        __syncthreads();   // sZ / sdY are rewritten by the next tile
    }
    // block partial of the squared error
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if (t % 64 == 0) sl[t / 64] = lsum;
    __syncthreads();
    if (t == 0) loss_part[blockIdx.x] = sl[0] + sl[1] + sl[2] + sl[3];
}

// ---------------------------------------------------------------------------
// weight gradient G[Ka][Kb] = A^T B and bias gradient gb[Kb] = sum_r B[r][:]
// over rows, in kGradBlocks contiguous row ranges (partials, fixed order).
// Thread t owns the entries t, t + 256, ... of G (and of gb).
// ---------------------------------------------------------------------------
template <int KA, int KB>
__global__ __launch_bounds__(kThr) void k_mlp_wgrad(const float* __restrict__ A, const float* __restrict__ B,
                                                    int64_t rows, float* __restrict__ part) {
    constexpr int NG = KA * KB;
    constexpr int PER = (NG + kThr - 1) / kThr;
    constexpr int CH = 16;   // rows staged per pass
    __shared__ float sa[CH * KA];
    __shared__ float sb[CH * KB];
    const int t = threadIdx.x;
    const int64_t per_block = (rows + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * per_block;
    const int64_t hi = std::min<int64_t>(rows, lo + per_block);
    float g[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) g[q] = 0.f;
    float gb = 0.f;
    for (int64_t r0 = lo; r0 < hi; r0 += CH) {
        const int nr = (int)std::min<int64_t>(CH, hi - r0);
        for (int i = t; i < CH * KA; i += kThr) sa[i] = (i / KA < nr) ? A[(r0 + i / KA) * KA + i % KA] : 0.f;
        for (int i = t; i < CH * KB; i += kThr) sb[i] = (i / KB < nr) ? B[(r0 + i / KB) * KB + i % KB] : 0.f;
        __syncthreads();
        for (int rr = 0; rr < CH; rr++) {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int e = t + q * kThr;
                if (e < NG) g[q] += sa[rr * KA + e / KB] * sb[rr * KB + e % KB];
            }
            if (t < KB) gb += sb[rr * KB + t];
        }
        __syncthreads();
    }
    float* out = part + (int64_t)blockIdx.x * (NG + KB);
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int e = t + q * kThr;
        if (e < NG) out[e] = g[q];
    }
    if (t < KB) out[NG + t] = gb;
}

// sum of the kGradBlocks partials into grad[off_w..] and grad[off_b..]
__global__ __launch_bounds__(kThr) void k_mlp_wgrad_reduce(const float* __restrict__ part, int nblk, int ng, int kb,
                                                           float* __restrict__ gw, float* __restrict__ gbias) {
    const int e = blockIdx.x * kThr + threadIdx.x;
    if (e >= ng + kb) return;
    float s = 0.f;
    for (int b = 0; b < nblk; b++) s += part[(int64_t)b * (ng + kb) + e];
    if (e < ng)
        gw[e] = s;
    else
        gbias[e - ng] = s;
}

__global__ __launch_bounds__(kThr) void k_adam(int64_t P, float* __restrict__ w, const float* __restrict__ g,
                                               float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                               float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (i >= P) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- features / targets / guesses on the context's state ----
__global__ __launch_bounds__(kThr) void k_mlp_features(int64_t N, const double* __restrict__ inc1,
                                                       const double* __restrict__ inc2,
                                                       const double* __restrict__ tact_node, double time,
                                                       double prate, double s_inc, double s_p, double s_t,
                                                       float* __restrict__ X) {
    const int64_t i = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (i >= N) return;
    float* x = X + i * kIn;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        x[c] = (float)(inc1[3 * i + c] * s_inc);
        x[3 + c] = (float)(inc2[3 * i + c] * s_inc);
    }
    x[6] = (float)(prate * s_p);
    x[7] = (float)((time - tact_node[i]) * s_t);
}

__global__ __launch_bounds__(kThr) void k_mlp_targets(int64_t n, const double* __restrict__ d, double s_inc,
                                                      float* __restrict__ T) {
    const int64_t i = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (i >= n) return;
    T[i] = (float)(d[i] * s_inc);
}

__global__ __launch_bounds__(kThr) void k_mlp_guess(int64_t N, const float* __restrict__ Y, double inv_s_inc,
                                                    const uint8_t* __restrict__ bcmask, double* __restrict__ g) {
    const int64_t i = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (i >= N) return;
    const uint8_t m = bcmask[i];
#pragma unroll
    for (int c = 0; c < 3; c++) g[3 * i + c] = ((m >> c) & 1) ? 0.0 : (double)Y[3 * i + c] * inv_s_inc;
}

// mean activation time of the elements incident to each node
__global__ __launch_bounds__(kThr) void k_node_activation(int64_t N, const int64_t* __restrict__ inc_ptr,
                                                          const int32_t* __restrict__ inc, int npe,
                                                          const double* __restrict__ tact, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (i >= N) return;
    double s = 0.0;
    const int64_t b = inc_ptr[i], e = inc_ptr[i + 1];
    for (int64_t k = b; k < e; k++) s += tact[inc[k] / npe];
    out[i] = (e > b) ? s / (double)(e - b) : 0.0;
}

unsigned grid_for(int64_t n) { return (unsigned)((n + kThr - 1) / kThr); }

}  // namespace

struct cmg_mlp {
    cmg_ctx* c = nullptr;
    int device = 0;
    DevBuf w, g, m, v, part, loss_part, tact_node;
    DevBuf A1, A2, dY, dZ1, dZ2, Yt;
    int64_t cap_rows = 0;
    int adam_t = 0;
    double s_inc = 1.0, s_p = 1.0, s_t = 1.0;
    double* h_pin = nullptr;
    ~cmg_mlp() {
        if (h_pin) (void)hipHostFree(h_pin);
    }
};

namespace {

int ensure_rows(cmg_mlp* m, int64_t rows) {
    if (rows <= m->cap_rows) return CMG_SUCCESS;
    const int64_t cap = std::max<int64_t>(rows, 2 * m->cap_rows);
    CMG_TRY(alloc(m->A1, (size_t)cap * kH * sizeof(float)));
    CMG_TRY(alloc(m->A2, (size_t)cap * kH * sizeof(float)));
    CMG_TRY(alloc(m->dZ1, (size_t)cap * kH * sizeof(float)));
    CMG_TRY(alloc(m->dZ2, (size_t)cap * kH * sizeof(float)));
    CMG_TRY(alloc(m->dY, (size_t)cap * kOut * sizeof(float)));
    CMG_TRY(alloc(m->Yt, (size_t)cap * kOut * sizeof(float)));
    m->cap_rows = cap;
    return CMG_SUCCESS;
}

unsigned fwd_blocks(int64_t rows) {
    const int64_t tiles = (rows + kTile - 1) / kTile;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, 1024));
}

}  // namespace

extern "C" {

int32_t cmg_mlp_num_weights(int32_t hidden) { return hidden == kH ? (int32_t)n_weights() : -1; }

int cmg_mlp_create(cmg_ctx* c, int32_t hidden, uint64_t seed, cmg_mlp** out) {
    if (!c || !out) return fail("cmg_mlp_create: null argument");
    if (hidden != kH) return fail("cmg_mlp_create: the kernels are built for 128 hidden units");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    std::unique_ptr<cmg_mlp> m(new cmg_mlp());
    m->c = c;
    m->device = c->device;
    const int64_t P = n_weights();
    for (DevBuf* b : {&m->w, &m->g, &m->m, &m->v}) CMG_TRY(alloc(*b, P * sizeof(float)));
    CMG_TRY(alloc(m->part, (size_t)kGradBlocks * (kH * kH + kH) * sizeof(float)));
    CMG_TRY(alloc(m->loss_part, 1024 * sizeof(double)));
    CMG_TRY(alloc(m->tact_node, c->N * sizeof(double)));
    CMG_HIP_CHECK(hipHostMalloc((void**)&m->h_pin, 1024 * sizeof(double), hipHostMallocDefault));
    // Glorot-uniform weights, zero biases (seeded host generator)
    std::vector<float> w(P, 0.f);
    std::mt19937_64 gen(seed);
    auto fill = [&](int64_t off, int fan_in, int fan_out) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        std::uniform_real_distribution<double> u(-a, a);
        for (int64_t i = 0; i < (int64_t)fan_in * fan_out; i++) w[off + i] = (float)u(gen);
    };
    fill(0, kIn, kH);
    fill(off_W2, kH, kH);
    fill(off_W3, kH, kOut);
    hipStream_t s = c->stream;
    CMG_HIP_CHECK(hipMemcpyAsync(m->w.p, w.data(), P * sizeof(float), hipMemcpyHostToDevice, s));
    CMG_HIP_CHECK(hipMemsetAsync(m->m.p, 0, P * sizeof(float), s));
    CMG_HIP_CHECK(hipMemsetAsync(m->v.p, 0, P * sizeof(float), s));
    CMG_HIP_CHECK(hipMemsetAsync(m->g.p, 0, P * sizeof(float), s));
    hipLaunchKernelGGL(k_node_activation, dim3(grid_for(c->N)), dim3(kThr), 0, s, c->N, c->inc_ptr.as<int64_t>(),
                       c->inc.as<int32_t>(), c->npe, c->tact.as<double>(), m->tact_node.as<double>());
    CMG_HIP_CHECK(hipGetLastError());
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    *out = m.release();
    return CMG_SUCCESS;
}

int cmg_mlp_destroy(cmg_mlp* m) {
    if (!m) return CMG_SUCCESS;
    (void)hipSetDevice(m->device);   // the context may already be gone: do not touch it
    (void)hipDeviceSynchronize();
    delete m;
    return CMG_SUCCESS;
}

int cmg_mlp_get_weights(cmg_mlp* m, float* w) {
    if (!m || !w) return fail("cmg_mlp_get_weights: null argument");
    CMG_HIP_CHECK(hipMemcpyAsync(w, m->w.p, n_weights() * sizeof(float), hipMemcpyDeviceToHost, m->c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(m->c->stream));
    return CMG_SUCCESS;
}

int cmg_mlp_set_weights(cmg_mlp* m, const float* w) {
    if (!m || !w) return fail("cmg_mlp_set_weights: null argument");
    CMG_HIP_CHECK(hipMemcpyAsync(m->w.p, w, n_weights() * sizeof(float), hipMemcpyHostToDevice, m->c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(m->c->stream));
    m->adam_t = 0;
    CMG_HIP_CHECK(hipMemsetAsync(m->m.p, 0, n_weights() * sizeof(float), m->c->stream));
    CMG_HIP_CHECK(hipMemsetAsync(m->v.p, 0, n_weights() * sizeof(float), m->c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(m->c->stream));
    return CMG_SUCCESS;
}

int cmg_mlp_set_scales(cmg_mlp* m, double inc_scale, double pressure_rate_scale, double phase_scale) {
    if (!m) return fail("null network");
    if (!(inc_scale > 0.0)) return fail("cmg_mlp_set_scales: inc_scale must be positive");
    m->s_inc = inc_scale;
    m->s_p = pressure_rate_scale;
    m->s_t = phase_scale;
    return CMG_SUCCESS;
}

int cmg_mlp_features(cmg_mlp* m, double time, double pressure_rate, float* X) {
    if (!m || !X) return fail("cmg_mlp_features: null argument");
    cmg_ctx* c = m->c;
    CMG_HIP_CHECK(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_mlp_features, dim3(grid_for(c->N)), dim3(kThr), 0, c->stream, c->N, c->inc1.as<double>(),
                       c->inc2.as<double>(), m->tact_node.as<double>(), time, pressure_rate, m->s_inc, m->s_p,
                       m->s_t, X);
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}

int cmg_mlp_targets(cmg_mlp* m, float* T) {
    if (!m || !T) return fail("cmg_mlp_targets: null argument");
    cmg_ctx* c = m->c;
    CMG_HIP_CHECK(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_mlp_targets, dim3(grid_for(c->n)), dim3(kThr), 0, c->stream, c->n, c->inc1.as<double>(),
                       m->s_inc, T);
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}

int cmg_mlp_forward(cmg_mlp* m, const float* X, int64_t rows, int32_t bf16, float* Y) {
    if (!m || !X || !Y) return fail("cmg_mlp_forward: null argument");
    if (rows <= 0) return CMG_SUCCESS;
    CMG_HIP_CHECK(hipSetDevice(m->c->device));
    hipStream_t s = m->c->stream;
    if (bf16)
        hipLaunchKernelGGL(k_mlp_forward_mfma, dim3(fwd_blocks(rows)), dim3(kThr), 0, s, m->w.as<float>(), X, rows, Y);
    else
        hipLaunchKernelGGL(k_mlp_forward, dim3(fwd_blocks(rows)), dim3(kThr), 0, s, m->w.as<float>(), X, rows, Y,
                           nullptr, nullptr);
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}

int cmg_mlp_loss_grad(cmg_mlp* m, const float* X, const float* T, int64_t rows, double* loss, float* grad_host) {
    if (!m || !X || !T) return fail("cmg_mlp_loss_grad: null argument");
    if (rows <= 0) return fail("cmg_mlp_loss_grad: no rows");
    cmg_ctx* c = m->c;
    CMG_HIP_CHECK(hipSetDevice(c->device));
    CMG_TRY(ensure_rows(m, rows));
    hipStream_t s = c->stream;
    const float* w = m->w.as<float>();
    float* g = m->g.as<float>();
    const unsigned nb = fwd_blocks(rows);
    hipLaunchKernelGGL(k_mlp_forward, dim3(nb), dim3(kThr), 0, s, w, X, rows, m->Yt.as<float>(),
                       m->A1.as<float>(), m->A2.as<float>());
    hipLaunchKernelGGL(k_mlp_backward_act, dim3(nb), dim3(kThr), 0, s, w, rows, m->Yt.as<float>(), T,
                       m->A1.as<float>(), m->A2.as<float>(), m->dY.as<float>(), m->dZ2.as<float>(),
                       m->dZ1.as<float>(), m->loss_part.as<double>());
    float* part = m->part.as<float>();
    // dW3 = A2^T dY, db3
    hipLaunchKernelGGL((k_mlp_wgrad<kH, kOut>), dim3(kGradBlocks), dim3(kThr), 0, s, m->A2.as<float>(),
                       m->dY.as<float>(), rows, part);
    hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3(grid_for(kH * kOut + kOut)), dim3(kThr), 0, s, part, kGradBlocks,
                       kH * kOut, kOut, g + off_W3, g + off_b3);
    // dW2 = A1^T dZ2, db2
    hipLaunchKernelGGL((k_mlp_wgrad<kH, kH>), dim3(kGradBlocks), dim3(kThr), 0, s, m->A1.as<float>(),
                       m->dZ2.as<float>(), rows, part);
    hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3(grid_for(kH * kH + kH)), dim3(kThr), 0, s, part, kGradBlocks, kH * kH,
                       kH, g + off_W2, g + off_b2);
    // dW1 = X^T dZ1, db1
    hipLaunchKernelGGL((k_mlp_wgrad<kIn, kH>), dim3(kGradBlocks), dim3(kThr), 0, s, X, m->dZ1.as<float>(), rows, part);
    hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3(grid_for(kIn * kH + kH)), dim3(kThr), 0, s, part, kGradBlocks,
                       kIn * kH, kH, g, g + off_b1);
    CMG_HIP_CHECK(hipGetLastError());
    CMG_HIP_CHECK(hipMemcpyAsync(m->h_pin, m->loss_part.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    if (grad_host) CMG_HIP_CHECK(hipMemcpyAsync(grad_host, g, n_weights() * sizeof(float), hipMemcpyDeviceToHost, s));
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    if (loss) {
        double l = 0.0;
        for (unsigned b = 0; b < nb; b++) l += m->h_pin[b];
        *loss = l / (3.0 * (double)rows);
    }
    return CMG_SUCCESS;
}

int cmg_mlp_adam_step(cmg_mlp* m, float lr) {
    if (!m) return fail("null network");
    CMG_HIP_CHECK(hipSetDevice(m->c->device));
    m->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
    const int64_t P = n_weights();
    hipLaunchKernelGGL(k_adam, dim3(grid_for(P)), dim3(kThr), 0, m->c->stream, P, m->w.as<float>(), m->g.as<float>(),
                       m->m.as<float>(), m->v.as<float>(), lr, b1, b2, eps, c1, c2);
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}

int cmg_mlp_propose_guess(cmg_mlp* m, const float* Y) {
    if (!m || !Y) return fail("cmg_mlp_propose_guess: null argument");
    cmg_ctx* c = m->c;
    CMG_HIP_CHECK(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_mlp_guess, dim3(grid_for(c->N)), dim3(kThr), 0, c->stream, c->N, Y, 1.0 / m->s_inc,
                       c->bcmask.as<uint8_t>(), c->override_guess.as<double>());
    CMG_HIP_CHECK(hipGetLastError());
    c->has_override = true;
    c->override_checked = true;
    return CMG_SUCCESS;
}

}  // extern "C"
