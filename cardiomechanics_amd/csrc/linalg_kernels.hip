// Sparse and vector kernels of the Krylov solve (gfx950): 3x3 block-CSR SpMV,
// nodal block Jacobi, fused multi-dot / multi-axpy for Gram-Schmidt, BLAS-1, and
// the Newmark predictor/corrector (CBSolverNewmarkBeta.cpp:845-852, 984-1015).
// All reductions are two-stage with a fixed order (no atomics): bitwise
// reproducible from run to run.
#include "cmg_internal.h"

namespace cmg {
namespace {

// y = A x, LANES lanes per block row (consecutive lanes read consecutive blocks)
template <typename VT, int LANES>
__global__ __launch_bounds__(256) void k_spmv(int64_t N, const int64_t* rowptr, const int32_t* colind,
                                              const VT* vals, const double* x, double* y) {
    constexpr int kSpmvGroup = LANES;
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = gt / kSpmvGroup;
    const int sub = (int)(gt % kSpmvGroup);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (r < N) {
        const int64_t end = rowptr[r + 1];
        for (int64_t k = rowptr[r] + sub; k < end; k += kSpmvGroup) {
            const int64_t c = colind[k];
            const VT* a = vals + 9 * k;
            const double x0 = x[3 * c], x1 = x[3 * c + 1], x2 = x[3 * c + 2];
            s0 += (double)a[0] * x0 + (double)a[1] * x1 + (double)a[2] * x2;
            s1 += (double)a[3] * x0 + (double)a[4] * x1 + (double)a[5] * x2;
            s2 += (double)a[6] * x0 + (double)a[7] * x1 + (double)a[8] * x2;
        }
    }
#pragma unroll
    for (int o = kSpmvGroup / 2; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, kSpmvGroup);
        s1 += __shfl_xor(s1, o, kSpmvGroup);
        s2 += __shfl_xor(s2, o, kSpmvGroup);
    }
    if (r < N && sub == 0) {
        y[3 * r] = s0;
        y[3 * r + 1] = s1;
        y[3 * r + 2] = s2;
    }
}


// y = A x reading the values in 144-byte chunks (2 FP64 or 4 FP32 blocks, nine
// 16-byte loads) and the column indices of a chunk in one 8- or 16-byte load: fewer,
// wider memory instructions than one block at a time.  The value and column arrays
// must be 16-byte aligned and padded by one chunk (blocks of other rows that share a
// chunk are skipped).
template <typename VT, int LANES>
__global__ __launch_bounds__(256) void k_spmv_chunk(int64_t N, const int64_t* rowptr, const int32_t* colind,
                                                    const VT* vals, const double* x, double* y) {
    constexpr int BPC = 16 / sizeof(VT) == 2 ? 2 : 4;   // blocks per 144-byte chunk
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = gt / LANES;
    const int sub = (int)(gt % LANES);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (r < N) {
        const int64_t k0 = rowptr[r], k1 = rowptr[r + 1];
        const uint4* V4 = reinterpret_cast<const uint4*>(vals);
        for (int64_t ch = k0 / BPC + sub; ch * BPC < k1; ch += LANES) {
            uint4 raw[9];
#pragma unroll
            for (int q = 0; q < 9; q++) raw[q] = V4[9 * ch + q];
            VT v[9 * BPC];
            __builtin_memcpy(v, raw, sizeof(v));
            int32_t cols[BPC];
            if constexpr (BPC == 2) {
                const int2 c2 = reinterpret_cast<const int2*>(colind)[ch];
                cols[0] = c2.x;
                cols[1] = c2.y;
            } else {
                const int4 c4 = reinterpret_cast<const int4*>(colind)[ch];
                cols[0] = c4.x;
                cols[1] = c4.y;
                cols[2] = c4.z;
                cols[3] = c4.w;
            }
#pragma unroll
            for (int b = 0; b < BPC; b++) {
                const int64_t k = ch * BPC + b;
                if (k < k0 || k >= k1) continue;
                const int64_t c = cols[b];
                const VT* a = v + 9 * b;
                const double x0 = x[3 * c], x1 = x[3 * c + 1], x2 = x[3 * c + 2];
                s0 += (double)a[0] * x0 + (double)a[1] * x1 + (double)a[2] * x2;
                s1 += (double)a[3] * x0 + (double)a[4] * x1 + (double)a[5] * x2;
                s2 += (double)a[6] * x0 + (double)a[7] * x1 + (double)a[8] * x2;
            }
        }
    }
#pragma unroll
    for (int o = LANES / 2; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, LANES);
        s1 += __shfl_xor(s1, o, LANES);
        s2 += __shfl_xor(s2, o, LANES);
    }
    if (r < N && sub == 0) {
        y[3 * r] = s0;
        y[3 * r + 1] = s1;
        y[3 * r + 2] = s2;
    }
}

// y = A x with the values read as one flat stream: G lanes share a block row and
// lane l takes the value words l, l + G, l + 2G, ... of the row's 9 nb values
// ([block][i][j] order), so every load instruction of a wave reads consecutive
// addresses (fully coalesced, unlike one 72-byte block per lane).  Each lane keeps
// three partial sums (its words' rows i) and the group reduces them with shuffles.
template <typename VT, int G>
__global__ __launch_bounds__(256) void k_spmv_flat(int64_t N, const int64_t* rowptr, const int32_t* colind,
                                                   const VT* vals, const double* x, double* y) {
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = gt / G;
    const int sub = (int)(gt % G);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (r < N) {
        const int64_t k0 = rowptr[r];
        const int cnt = (int)(9 * (rowptr[r + 1] - k0));
        const VT* v = vals + 9 * k0;
        const int32_t* ci = colind + k0;
        constexpr int GB = G / 9, GR = G % 9;
        int blk = sub / 9, rem = sub - 9 * (sub / 9);
        for (int idx = sub; idx < cnt; idx += G) {
            const int i = rem / 3, j = rem - 3 * (rem / 3);
            const double p = (double)v[idx] * x[3 * (int64_t)ci[blk] + j];
            s0 += (i == 0) ? p : 0.0;
            s1 += (i == 1) ? p : 0.0;
            s2 += (i == 2) ? p : 0.0;
            blk += GB;
            rem += GR;
            if (rem >= 9) {
                rem -= 9;
                blk++;
            }
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, G);
        s1 += __shfl_xor(s1, o, G);
        s2 += __shfl_xor(s2, o, G);
    }
    if (r < N && sub == 0) {
        y[3 * r] = s0;
        y[3 * r + 1] = s1;
        y[3 * r + 2] = s2;
    }
}

__global__ __launch_bounds__(256) void k_spmv_scalar_mass(int64_t N, const int64_t* rowptr, const int32_t* colind,
                                                          const double* mval, const uint8_t* bcmask,
                                                          const double* x, double* y) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
        const int64_t c = colind[k];
        s0 += mval[k] * x[3 * c];
        s1 += mval[k] * x[3 * c + 1];
        s2 += mval[k] * x[3 * c + 2];
    }
    const uint8_t m = bcmask ? bcmask[r] : 0;
    y[3 * r] = (m & 1) ? 0.0 : s0;
    y[3 * r + 1] = (m & 2) ? 0.0 : s1;
    y[3 * r + 2] = (m & 4) ? 0.0 : s2;
}

template <typename VT>
__global__ __launch_bounds__(256) void k_bj_setup(int64_t N, const int64_t* rowptr, const int32_t* colind,
                                                  const VT* vals, double* dinv, SetStrides b) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t set = blockIdx.y;
    if (r >= N || (b.active && !b.active[set])) return;
    vals += set * b.a;
    dinv += set * b.b;
    double a[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
        if (colind[k] == r) {
            for (int q = 0; q < 9; q++) a[q] = (double)vals[9 * k + q];
            break;
        }
    }
    const double d = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
    const double id = (d != 0.0) ? 1.0 / d : 0.0;
    double* o = dinv + 9 * r;
    o[0] = (a[4] * a[8] - a[5] * a[7]) * id;
    o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    o[3] = (a[5] * a[6] - a[3] * a[8]) * id;
    o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    o[6] = (a[3] * a[7] - a[4] * a[6]) * id;
    o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

__global__ __launch_bounds__(256) void k_bj_apply(int64_t N, const double* dinv, const double* r, double* z) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const double* a = dinv + 9 * n;
    const double r0 = r[3 * n], r1 = r[3 * n + 1], r2 = r[3 * n + 2];
    z[3 * n] = a[0] * r0 + a[1] * r1 + a[2] * r2;
    z[3 * n + 1] = a[3] * r0 + a[4] * r1 + a[5] * r2;
    z[3 * n + 2] = a[6] * r0 + a[7] * r1 + a[8] * r2;
}

// Multi-dot: partial[b][j] = sum over the b-th chunk of V_j . w, j < k.
constexpr int kDotThreads = 256;
constexpr int kDotPerThread = 8;
constexpr int kDotChunk = kDotThreads * kDotPerThread;
constexpr int kDotGroup = 8;

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kDotThreads / 64; i++) s += sh[i];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(kDotThreads) void k_multidot(int64_t n, int k, const double* V, int64_t ld,
                                                          const double* w, double* partial) {
    __shared__ double sh[kDotThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kDotChunk;
    double wv[kDotPerThread];
#pragma unroll
    for (int q = 0; q < kDotPerThread; q++) {
        const int64_t i = base + (int64_t)q * kDotThreads + threadIdx.x;
        wv[q] = (i < n) ? w[i] : 0.0;
    }
    for (int j0 = 0; j0 < k; j0 += kDotGroup) {
        double acc[kDotGroup];
#pragma unroll
        for (int g = 0; g < kDotGroup; g++) acc[g] = 0.0;
#pragma unroll
        for (int g = 0; g < kDotGroup; g++) {
            if (j0 + g < k) {
                const double* v = V + (int64_t)(j0 + g) * ld;
#pragma unroll
                for (int q = 0; q < kDotPerThread; q++) {
                    const int64_t i = base + (int64_t)q * kDotThreads + threadIdx.x;
                    if (i < n) acc[g] += v[i] * wv[q];
                }
            }
        }
        for (int g = 0; g < kDotGroup && j0 + g < k; g++) {
            const double s = block_sum(acc[g], sh);
            if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * k + j0 + g] = s;
        }
    }
}

// ---- fused BiCGStab vector updates (host-side scalars), one thread per node ----
// p = r + beta (p - omega v) (p = r on a (re)start), then ph = D^-1 p
__global__ __launch_bounds__(256) void k_bicg_p_bj(int64_t N, int fresh, double beta, double omega,
                                                   const double* r, const double* v, const double* dinv, double* p,
                                                   double* ph) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double q[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int64_t i = 3 * n + c;
        q[c] = fresh ? r[i] : r[i] + beta * (p[i] - omega * v[i]);
        p[i] = q[c];
    }
    if (!dinv) return;   // another preconditioner follows
    const double* a = dinv + 9 * n;
    ph[3 * n] = a[0] * q[0] + a[1] * q[1] + a[2] * q[2];
    ph[3 * n + 1] = a[3] * q[0] + a[4] * q[1] + a[5] * q[2];
    ph[3 * n + 2] = a[6] * q[0] + a[7] * q[1] + a[8] * q[2];
}

// s = r - alpha v, then sh = D^-1 s
__global__ __launch_bounds__(256) void k_bicg_s_bj(int64_t N, double alpha, const double* r, const double* v,
                                                   const double* dinv, double* sv, double* sh) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double q[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int64_t i = 3 * n + c;
        q[c] = r[i] - alpha * v[i];
        sv[i] = q[c];
    }
    if (!dinv) return;   // another preconditioner follows
    const double* a = dinv + 9 * n;
    sh[3 * n] = a[0] * q[0] + a[1] * q[1] + a[2] * q[2];
    sh[3 * n + 1] = a[3] * q[0] + a[4] * q[1] + a[5] * q[2];
    sh[3 * n + 2] = a[6] * q[0] + a[7] * q[1] + a[8] * q[2];
}

// x += alpha ph + omega sh;  r = s - omega t
__global__ __launch_bounds__(256) void k_bicg_xr(int64_t n, double alpha, double omega, const double* ph,
                                                 const double* sh, const double* sv, const double* t, double* x,
                                                 double* r) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = x[i] + alpha * ph[i] + omega * sh[i];
    r[i] = sv[i] - omega * t[i];
}

// partial[b][0..4] = (rh,t) (s,t) (t,t) (rh,s) (s,s) over the b-th chunk
__global__ __launch_bounds__(kDotThreads) void k_bicg_dots5(int64_t n, const double* rh, const double* sv,
                                                            const double* t, double* partial) {
    __shared__ double sh[kDotThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kDotChunk;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kDotPerThread; q++) {
        const int64_t i = base + (int64_t)q * kDotThreads + threadIdx.x;
        if (i < n) {
            const double a = rh[i], b = sv[i], c = t[i];
            acc[0] += a * c;
            acc[1] += b * c;
            acc[2] += c * c;
            acc[3] += a * b;
            acc[4] += b * b;
        }
    }
#pragma unroll
    for (int j = 0; j < 5; j++) {
        const double v = block_sum(acc[j], sh);
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * 5 + j] = v;
    }
}

__global__ __launch_bounds__(256) void k_reduce_partials(int nblk, int k, const double* partial, double* out) {
    __shared__ double sh[4];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += partial[(int64_t)b * k + j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[j] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ __launch_bounds__(256) void k_multiaxpy(int64_t n, int k, const double* V, int64_t ld, const double* h,
                                                   double* w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = w[i];
    for (int j = 0; j < k; j++) s -= h[j] * V[(int64_t)j * ld + i];
    w[i] = s;
}

__global__ __launch_bounds__(256) void k_axpby(int64_t n, double a, const double* x, double b, double* y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

__global__ __launch_bounds__(256) void k_waxpby(int64_t n, double a, const double* x, double b, const double* y,
                                                double* w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = a * x[i] + b * y[i];
}

__global__ __launch_bounds__(256) void k_scale_copy(int64_t n, double a, const double* x, double* y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

__global__ __launch_bounds__(256) void k_fill(int64_t n, double v, double* y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

__global__ __launch_bounds__(64) void k_fill_int(int64_t n, int v, int* y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

// CBSolverNewmarkBeta.cpp:845-852: d~ = dt v + dt^2/2 (1-2 beta) a ; v~ = v + (1-gamma) dt a
__global__ __launch_bounds__(256) void k_newmark_predict(int64_t n, double dt, double beta, double gamma,
                                                         const double* v, const double* a, double* dp, double* vp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dp[i] = dt * v[i] + dt * dt * 0.5 * (1 - 2 * beta) * a[i];
    vp[i] = 1.0 * v[i] + (1 - gamma) * dt * a[i];
}

// CBSolverNewmarkBeta.cpp:984-991: a = (d - d~)/(beta dt^2) ; v = v~ + gamma dt a
__global__ __launch_bounds__(256) void k_newmark_correct(int64_t n, double dt, double beta, double gamma,
                                                         const double* d, const double* dp, const double* vp,
                                                         double* v, double* a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ai = 1.0 / (beta * dt * dt) * d[i] + -1.0 / (beta * dt * dt) * dp[i];
    a[i] = ai;
    v[i] = vp[i] + gamma * dt * ai;
}

__global__ __launch_bounds__(256) void k_zero_bc(int64_t N, const uint8_t* bcmask, double* v) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const uint8_t m = bcmask[r];
    if (m & 1) v[3 * r] = 0.0;
    if (m & 2) v[3 * r + 1] = 0.0;
    if (m & 4) v[3 * r + 2] = 0.0;
}

inline dim3 grid1(int64_t n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

}  // namespace

void launch_spmv(int64_t N, const int64_t* rowptr, const int32_t* colind, const void* vals, bool fp32,
                 const double* x, double* y, hipStream_t s, int lanes, int layout) {
    if (N == 0) return;
    if (layout == kSpmvFlat) {
        if (lanes != 16 && lanes != 64) lanes = 32;
        const dim3 g = grid1(N * lanes);
#define CMG_SPMV(VT, L) \
    hipLaunchKernelGGL((k_spmv_flat<VT, L>), g, dim3(256), 0, s, N, rowptr, colind, (const VT*)vals, x, y)
        if (fp32) {
            if (lanes == 16) CMG_SPMV(float, 16);
            else if (lanes == 64) CMG_SPMV(float, 64);
            else CMG_SPMV(float, 32);
        } else {
            if (lanes == 16) CMG_SPMV(double, 16);
            else if (lanes == 64) CMG_SPMV(double, 64);
            else CMG_SPMV(double, 32);
        }
#undef CMG_SPMV
        return;
    }
    if (lanes != 4 && lanes != 16) lanes = 8;
    const dim3 g = grid1(N * lanes);
    if (layout == kSpmvChunk) {
#define CMG_SPMV(VT, L) \
    hipLaunchKernelGGL((k_spmv_chunk<VT, L>), g, dim3(256), 0, s, N, rowptr, colind, (const VT*)vals, x, y)
        if (fp32) {
            if (lanes == 4) CMG_SPMV(float, 4);
            else if (lanes == 16) CMG_SPMV(float, 16);
            else CMG_SPMV(float, 8);
        } else {
            if (lanes == 4) CMG_SPMV(double, 4);
            else if (lanes == 16) CMG_SPMV(double, 16);
            else CMG_SPMV(double, 8);
        }
#undef CMG_SPMV
        return;
    }
#define CMG_SPMV(VT, L) hipLaunchKernelGGL((k_spmv<VT, L>), g, dim3(256), 0, s, N, rowptr, colind, (const VT*)vals, x, y)
    if (fp32) {
        if (lanes == 4) CMG_SPMV(float, 4);
        else if (lanes == 16) CMG_SPMV(float, 16);
        else CMG_SPMV(float, 8);
    } else {
        if (lanes == 4) CMG_SPMV(double, 4);
        else if (lanes == 16) CMG_SPMV(double, 16);
        else CMG_SPMV(double, 8);
    }
#undef CMG_SPMV
}

void launch_spmv_scalar_mass(int64_t N, const int64_t* rowptr, const int32_t* colind, const double* mval,
                             const uint8_t* bcmask, const double* x, double* y, hipStream_t s) {
    if (N == 0) return;
    hipLaunchKernelGGL(k_spmv_scalar_mass, grid1(N), dim3(256), 0, s, N, rowptr, colind, mval, bcmask, x, y);
}

void launch_block_jacobi_setup(int64_t N, const int64_t* rowptr, const int32_t* colind, const void* vals, bool fp32,
                               double* dinv, hipStream_t s, const SetStrides& b) {
    if (N == 0) return;
    const dim3 grid((unsigned)((N + 255) / 256), b.n_sets);
    if (fp32)
        hipLaunchKernelGGL((k_bj_setup<float>), grid, dim3(256), 0, s, N, rowptr, colind, (const float*)vals,
                           dinv, b);
    else
        hipLaunchKernelGGL((k_bj_setup<double>), grid, dim3(256), 0, s, N, rowptr, colind, (const double*)vals,
                           dinv, b);
}

void launch_block_jacobi_apply(int64_t N, const double* dinv, const double* r, double* z, hipStream_t s) {
    if (N == 0) return;
    hipLaunchKernelGGL(k_bj_apply, grid1(N), dim3(256), 0, s, N, dinv, r, z);
}

int dot_blocks(int64_t n) { return (int)((n + kDotChunk - 1) / kDotChunk); }

void launch_multidot(int64_t n, int k, const double* V, int64_t ld, const double* w, double* partial, hipStream_t s) {
    const int nb = dot_blocks(n);
    if (nb == 0 || k == 0) return;
    hipLaunchKernelGGL(k_multidot, dim3(nb), dim3(kDotThreads), 0, s, n, k, V, ld, w, partial);
}

void launch_bicg_p_bj(int64_t N, bool fresh, double beta, double omega, const double* r, const double* v,
                      const double* dinv, double* p, double* ph, hipStream_t s) {
    if (N == 0) return;
    hipLaunchKernelGGL(k_bicg_p_bj, grid1(N), dim3(256), 0, s, N, fresh ? 1 : 0, beta, omega, r, v, dinv, p, ph);
}

void launch_bicg_s_bj(int64_t N, double alpha, const double* r, const double* v, const double* dinv, double* sv,
                      double* sh, hipStream_t s) {
    if (N == 0) return;
    hipLaunchKernelGGL(k_bicg_s_bj, grid1(N), dim3(256), 0, s, N, alpha, r, v, dinv, sv, sh);
}

void launch_bicg_xr(int64_t n, double alpha, double omega, const double* ph, const double* sh, const double* sv,
                    const double* t, double* x, double* r, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_bicg_xr, grid1(n), dim3(256), 0, s, n, alpha, omega, ph, sh, sv, t, x, r);
}

void launch_bicg_dots5(int64_t n, const double* rh, const double* sv, const double* t, double* partial,
                       hipStream_t s) {
    const int nb = dot_blocks(n);
    if (nb == 0) return;
    hipLaunchKernelGGL(k_bicg_dots5, dim3(nb), dim3(kDotThreads), 0, s, n, rh, sv, t, partial);
}

void launch_reduce_partials(int nblk, int k, const double* partial, double* out, hipStream_t s) {
    if (k == 0) return;
    hipLaunchKernelGGL(k_reduce_partials, dim3(k), dim3(256), 0, s, nblk, k, partial, out);
}

void launch_multiaxpy(int64_t n, int k, const double* V, int64_t ld, const double* h, double* w, hipStream_t s) {
    if (n == 0 || k == 0) return;
    hipLaunchKernelGGL(k_multiaxpy, grid1(n), dim3(256), 0, s, n, k, V, ld, h, w);
}

void launch_axpby(int64_t n, double a, const double* x, double b, double* y, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_axpby, grid1(n), dim3(256), 0, s, n, a, x, b, y);
}

void launch_waxpby(int64_t n, double a, const double* x, double b, const double* y, double* w, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_waxpby, grid1(n), dim3(256), 0, s, n, a, x, b, y, w);
}

void launch_scale_copy(int64_t n, double a, const double* x, double* y, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_scale_copy, grid1(n), dim3(256), 0, s, n, a, x, y);
}

void launch_fill(int64_t n, double v, double* y, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill, grid1(n), dim3(256), 0, s, n, v, y);
}

void launch_fill_int(int64_t n, int v, int* y, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_int, grid1(n, 64), dim3(64), 0, s, n, v, y);
}

void launch_zero_bc(int64_t N, const uint8_t* bcmask, double* v, hipStream_t s) {
    if (N) hipLaunchKernelGGL(k_zero_bc, grid1(N), dim3(256), 0, s, N, bcmask, v);
}

void launch_newmark_predict(int64_t n, double dt, double beta, double gamma, const double* v, const double* a,
                            double* dpred, double* vpred, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_newmark_predict, grid1(n), dim3(256), 0, s, n, dt, beta, gamma, v, a, dpred, vpred);
}

void launch_newmark_correct(int64_t n, double dt, double beta, double gamma, const double* d, const double* dpred,
                            const double* vpred, double* v, double* a, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_newmark_correct, grid1(n), dim3(256), 0, s, n, dt, beta, gamma, d, dpred, vpred, v, a);
}

}  // namespace cmg
