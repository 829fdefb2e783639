// Kernels of the parameter-set batch (gfx950): S independent systems on one mesh,
// laid out set-major ([S][n] vectors, [S][nnzb*9] block values, [S][N*9] block
// Jacobi inverses).  Every launch covers all sets (blockIdx.y = set); the scalar
// recurrences of the preconditioned conjugate gradient stay in device memory, so a
// run of iterations needs no host synchronisation and can be captured as one graph.
// Reductions are two-stage in a fixed order: results do not depend on timing.
#include "batch_kernels.h"

namespace cmg {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ double block_sum256(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
}

// q = A p for one set, LANES lanes per block row, plus the block's share of p . q
template <typename VT, int LANES>
__global__ __launch_bounds__(kThreads) void k_bcg_spmv_dot(int64_t N, const int64_t* rowptr, const int32_t* colind,
                                                           const VT* vals, int64_t set_vals, const double* p,
                                                           double* q, int64_t n, const int32_t* done,
                                                           double* partial, int nb) {
    __shared__ double sh[4];
    const int64_t set = blockIdx.y;
    if (done[set]) return;
    vals += set * set_vals;
    p += set * n;
    q += set * n;
    const int64_t gt = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t r = gt / LANES;
    const int sub = (int)(gt % LANES);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (r < N) {
        const int64_t end = rowptr[r + 1];
        for (int64_t k = rowptr[r] + sub; k < end; k += LANES) {
            const int64_t c = colind[k];
            const VT* a = vals + 9 * k;
            const double x0 = p[3 * c], x1 = p[3 * c + 1], x2 = p[3 * c + 2];
            s0 += (double)a[0] * x0 + (double)a[1] * x1 + (double)a[2] * x2;
            s1 += (double)a[3] * x0 + (double)a[4] * x1 + (double)a[5] * x2;
            s2 += (double)a[6] * x0 + (double)a[7] * x1 + (double)a[8] * x2;
        }
    }
#pragma unroll
    for (int o = LANES / 2; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, LANES);
        s1 += __shfl_xor(s1, o, LANES);
        s2 += __shfl_xor(s2, o, LANES);
    }
    double pq = 0.0;
    if (r < N && sub == 0) {
        q[3 * r] = s0;
        q[3 * r + 1] = s1;
        q[3 * r + 2] = s2;
        pq = p[3 * r] * s0 + p[3 * r + 1] * s1 + p[3 * r + 2] * s2;
    }
    const double tot = block_sum256(pq, sh);
    if (threadIdx.x == 0) partial[set * nb + blockIdx.x] = tot;
}

// sum of nb partials (stride `ld` between sets, `k` consecutive values per block)
__device__ __forceinline__ double reduce_partials(const double* part, int nb, int k, int j, double* sh) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += kThreads) v += part[(int64_t)b * k + j];
    return block_sum256(v, sh);
}

__global__ __launch_bounds__(kThreads) void k_bcg_alpha(BatchCG cg, int nb) {
    __shared__ double sh[4];
    const int set = blockIdx.x;
    if (cg.done[set]) return;
    const double pq = reduce_partials(cg.part + (int64_t)set * cg.part_ld, nb, 1, 0, sh);
    if (threadIdx.x == 0) {
        if (pq > 0.0 && isfinite(pq)) {
            cg.alpha[set] = cg.rz[set] / pq;
        } else {
            cg.alpha[set] = 0.0;
            cg.done[set] = 2;   // breakdown: the operator is not positive definite on p
        }
    }
}

// x += alpha p, r -= alpha q, z = D^-1 r; partial r.z and r.r (one node per thread)
__global__ __launch_bounds__(kThreads) void k_bcg_update(BatchCG cg, int64_t N, const double* dinv, int nb) {
    __shared__ double sh[4];
    const int64_t set = blockIdx.y;
    if (cg.done[set]) return;
    const int64_t n = 3 * N;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const double al = cg.alpha[set];
    double rz = 0.0, rr = 0.0;
    if (i < N) {
        double* x = cg.x + set * n + 3 * i;
        double* r = cg.r + set * n + 3 * i;
        double* z = cg.z + set * n + 3 * i;
        const double* p = cg.p + set * n + 3 * i;
        const double* q = cg.q + set * n + 3 * i;
        const double* a = dinv + set * 9 * N + 9 * i;
        double rv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            x[c] += al * p[c];
            rv[c] = r[c] - al * q[c];
            r[c] = rv[c];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double zc = a[3 * c] * rv[0] + a[3 * c + 1] * rv[1] + a[3 * c + 2] * rv[2];
            z[c] = zc;
            rz += rv[c] * zc;
            rr += rv[c] * rv[c];
        }
    }
    const double trz = block_sum256(rz, sh);
    const double trr = block_sum256(rr, sh);
    if (threadIdx.x == 0) {
        double* part = cg.part + set * cg.part_ld + 2 * blockIdx.x;
        part[0] = trz;
        part[1] = trr;
    }
}

__global__ __launch_bounds__(kThreads) void k_bcg_beta(BatchCG cg, int nb) {
    __shared__ double sh[4];
    const int set = blockIdx.x;
    if (cg.done[set]) return;
    const double* part = cg.part + (int64_t)set * cg.part_ld;
    const double rz = reduce_partials(part, nb, 2, 0, sh);
    const double rr = reduce_partials(part, nb, 2, 1, sh);
    if (threadIdx.x == 0) {
        cg.it[set] += 1;
        cg.rr[set] = rr;
        if (rr <= cg.tol2[set] || !isfinite(rr)) {
            cg.done[set] = 1;
        } else {
            cg.beta[set] = rz / cg.rz[set];
            cg.rz[set] = rz;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_bcg_p(BatchCG cg, int64_t n) {
    const int64_t set = blockIdx.y;
    if (cg.done[set]) return;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const double b = cg.beta[set];
    double* p = cg.p + set * n;
    p[i] = cg.z[set * n + i] + b * p[i];
}

// start: x = 0, r = b, z = D^-1 r, p = z; partial r.z and r.r
__global__ __launch_bounds__(kThreads) void k_bcg_init(BatchCG cg, int64_t N, const double* b, const double* dinv,
                                                       const int32_t* active) {
    __shared__ double sh[4];
    const int64_t set = blockIdx.y;
    if (!active[set]) return;
    const int64_t n = 3 * N;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (i < N) {
        const int64_t o = set * n + 3 * i;
        const double* a = dinv + set * 9 * N + 9 * i;
        const double rv[3] = {b[o], b[o + 1], b[o + 2]};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double zc = a[3 * c] * rv[0] + a[3 * c + 1] * rv[1] + a[3 * c + 2] * rv[2];
            cg.x[o + c] = 0.0;
            cg.r[o + c] = rv[c];
            cg.z[o + c] = zc;
            cg.p[o + c] = zc;
            rz += rv[c] * zc;
            rr += rv[c] * rv[c];
        }
    }
    const double trz = block_sum256(rz, sh);
    const double trr = block_sum256(rr, sh);
    if (threadIdx.x == 0) {
        double* part = cg.part + set * cg.part_ld + 2 * blockIdx.x;
        part[0] = trz;
        part[1] = trr;
    }
}

__global__ __launch_bounds__(kThreads) void k_bcg_init_reduce(BatchCG cg, int nb, const double* rtol,
                                                              const int32_t* active) {
    __shared__ double sh[4];
    const int set = blockIdx.x;
    if (!active[set]) {
        if (threadIdx.x == 0) {
            cg.done[set] = 1;
            cg.it[set] = 0;
        }
        return;
    }
    const double* part = cg.part + (int64_t)set * cg.part_ld;
    const double rz = reduce_partials(part, nb, 2, 0, sh);
    const double rr = reduce_partials(part, nb, 2, 1, sh);
    if (threadIdx.x == 0) {
        cg.rz[set] = rz;
        cg.rr[set] = rr;
        cg.rr0[set] = rr;
        cg.tol2[set] = rtol[set] * rtol[set] * rr;
        cg.it[set] = 0;
        cg.done[set] = (rr == 0.0 || !(rz > 0.0)) ? 1 : 0;
    }
}

// ---- per-set BLAS-1 and dots ----
__global__ __launch_bounds__(kThreads) void k_bdot(int64_t n, const double* x, const double* y, double* partial,
                                                   int nb, int64_t part_ld) {
    __shared__ double sh[4];
    const int64_t set = blockIdx.y;
    x += set * n;
    y += set * n;
    double v = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)nb * kThreads)
        v += x[i] * y[i];
    const double t = block_sum256(v, sh);
    if (threadIdx.x == 0) partial[set * part_ld + blockIdx.x] = t;
}

__global__ __launch_bounds__(kThreads) void k_breduce(const double* partial, int nb, int64_t part_ld, double* out) {
    __shared__ double sh[4];
    const int set = blockIdx.x;
    const double t = reduce_partials(partial + (int64_t)set * part_ld, nb, 1, 0, sh);
    if (threadIdx.x == 0) out[set] = t;
}

__global__ __launch_bounds__(kThreads) void k_baxpy(int64_t n, SetCoefs a, const double* x, double* y) {
    const int64_t set = blockIdx.y;
    const double c = a.c[set];
    if (c == 0.0) return;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    y[set * n + i] += c * x[set * n + i];
}

__global__ __launch_bounds__(kThreads) void k_bscale(int64_t n, SetCoefs a, double* y) {
    const int64_t set = blockIdx.y;
    const double c = a.c[set];
    if (c == 1.0) return;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    y[set * n + i] *= c;
}

}  // namespace

int batch_dot_blocks(int64_t n) {
    const int64_t b = (n + 8 * kThreads - 1) / (8 * kThreads);
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1024));
}

int batch_spmv_blocks(int64_t N, int lanes) { return (int)((N * lanes + kThreads - 1) / kThreads); }

int batch_node_blocks(int64_t N) { return (int)((N + kThreads - 1) / kThreads); }

void launch_bcg_spmv_dot(const BatchCG& cg, int lanes, int64_t N, const int64_t* rowptr, const int32_t* colind,
                         const void* vals, bool fp32, int64_t set_vals, hipStream_t s) {
    const int nb = batch_spmv_blocks(N, lanes);
    const dim3 grid(nb, cg.S);
    const int64_t n = 3 * N;
#define CMG_SPMV(VT, L)                                                                                          \
    hipLaunchKernelGGL((k_bcg_spmv_dot<VT, L>), grid, dim3(kThreads), 0, s, N, rowptr, colind, (const VT*)vals, \
                       set_vals, cg.p, cg.q, n, cg.done, cg.part, (int)cg.part_ld)
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

// The spmv writes one partial per block at stride part_ld: part_ld >= blocks (checked by the caller).
void launch_bcg_iteration(const BatchCG& cg, int lanes, int64_t N, const int64_t* rowptr, const int32_t* colind,
                          const void* vals, bool fp32, int64_t set_vals, const double* dinv, hipStream_t s) {
    launch_bcg_spmv_dot(cg, lanes, N, rowptr, colind, vals, fp32, set_vals, s);
    hipLaunchKernelGGL(k_bcg_alpha, dim3(cg.S), dim3(kThreads), 0, s, cg, batch_spmv_blocks(N, lanes));
    const int nbn = batch_node_blocks(N);
    hipLaunchKernelGGL(k_bcg_update, dim3(nbn, cg.S), dim3(kThreads), 0, s, cg, N, dinv, nbn);
    hipLaunchKernelGGL(k_bcg_beta, dim3(cg.S), dim3(kThreads), 0, s, cg, nbn);
    const int64_t n = 3 * N;
    hipLaunchKernelGGL(k_bcg_p, dim3((unsigned)((n + kThreads - 1) / kThreads), cg.S), dim3(kThreads), 0, s, cg, n);
}

void launch_bcg_init(const BatchCG& cg, int64_t N, const double* b, const double* dinv, const double* rtol,
                     const int32_t* active, hipStream_t s) {
    const int nbn = batch_node_blocks(N);
    hipLaunchKernelGGL(k_bcg_init, dim3(nbn, cg.S), dim3(kThreads), 0, s, cg, N, b, dinv, active);
    hipLaunchKernelGGL(k_bcg_init_reduce, dim3(cg.S), dim3(kThreads), 0, s, cg, nbn, rtol, active);
}

void launch_bdot(int S, int64_t n, const double* x, const double* y, double* partial, int64_t part_ld, double* out,
                 hipStream_t s) {
    const int nb = batch_dot_blocks(n);
    hipLaunchKernelGGL(k_bdot, dim3(nb, S), dim3(kThreads), 0, s, n, x, y, partial, nb, part_ld);
    hipLaunchKernelGGL(k_breduce, dim3(S), dim3(kThreads), 0, s, partial, nb, part_ld, out);
}

void launch_baxpy(int S, int64_t n, const SetCoefs& a, const double* x, double* y, hipStream_t s) {
    hipLaunchKernelGGL(k_baxpy, dim3((unsigned)((n + kThreads - 1) / kThreads), S), dim3(kThreads), 0, s, n, a, x, y);
}

void launch_bscale(int S, int64_t n, const SetCoefs& a, double* y, hipStream_t s) {
    hipLaunchKernelGGL(k_bscale, dim3((unsigned)((n + kThreads - 1) / kThreads), S), dim3(kThreads), 0, s, n, a, y);
}

}  // namespace cmg
