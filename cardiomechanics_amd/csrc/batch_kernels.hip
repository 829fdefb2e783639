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

// sum of nb partials (stride `ld` between sets, `k` consecutive values per block);
// eight independent loads in flight per thread, summed in a fixed order
__device__ __forceinline__ double reduce_partials(const double* part, int nb, int k, int j, double* sh) {
    double v = 0.0;
#pragma unroll 8
    for (int b = threadIdx.x; b < nb; b += kThreads) v += part[(int64_t)b * k + j];
    return block_sum256(v, sh);
}

// the K sums of nb partials of K consecutive values each, in one pass over contiguous
// memory (thread t takes blocks t, t + 256, ... in order, then the fixed block tree)
template <int K>
__device__ __forceinline__ void reduce_partials_k(const double* part, int nb, double* sh, double* out) {
    double v[K];
#pragma unroll
    for (int j = 0; j < K; j++) v[j] = 0.0;
#pragma unroll 4
    for (int b = threadIdx.x; b < nb; b += kThreads)
#pragma unroll
        for (int j = 0; j < K; j++) v[j] += part[(int64_t)b * K + j];
#pragma unroll
    for (int j = 0; j < K; j++) out[j] = block_sum256(v[j], sh);
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

// x += alpha p, r -= alpha q, z = D^-1 r; partial r.z and r.r (one node per thread).
// Without dinv only x and r (z then comes from the preconditioner, the partials from k_bcg_rz).
__global__ __launch_bounds__(kThreads) void k_bcg_update(BatchCG cg, int64_t N, const double* dinv, int nb) {
    __shared__ double sh[4];
    const int64_t set = blockIdx.y;
    if (cg.done[set]) return;
    const int64_t n = 3 * N;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const double al = cg.alpha[set];
    if (!dinv) {
        if (i < N) {
            double* x = cg.x + set * n + 3 * i;
            double* r = cg.r + set * n + 3 * i;
            const double* p = cg.p + set * n + 3 * i;
            const double* q = cg.q + set * n + 3 * i;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                x[c] += al * p[c];
                r[c] = r[c] - al * q[c];
            }
        }
        return;
    }
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

// partial r.z and r.r after the preconditioner wrote z (p = z as well when init)
__global__ __launch_bounds__(kThreads) void k_bcg_rz(BatchCG cg, int64_t N, const int32_t* active, bool init) {
    __shared__ double sh[4];
    const int64_t set = blockIdx.y;
    if (init ? !active[set] : cg.done[set] != 0) return;
    const int64_t n = 3 * N;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (i < N) {
        const int64_t o = set * n + 3 * i;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double rv = cg.r[o + c], zc = cg.z[o + c];
            if (init) cg.p[o + c] = zc;
            rz += rv * zc;
            rr += rv * rv;
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
    double t2[2];
    reduce_partials_k<2>(part, nb, sh, t2);
    const double rz = t2[0], rr = t2[1];
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

// start: x = 0, r = b, z = D^-1 r, p = z; partial r.z and r.r (without dinv: x and r only)
__global__ __launch_bounds__(kThreads) void k_bcg_init(BatchCG cg, int64_t N, const double* b, const double* dinv,
                                                       const int32_t* active) {
    __shared__ double sh[4];
    const int64_t set = blockIdx.y;
    if (!active[set]) return;
    const int64_t n = 3 * N;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (!dinv) {
        if (i < N)
            for (int c = 0; c < 3; c++) {
                cg.x[set * n + 3 * i + c] = 0.0;
                cg.r[set * n + 3 * i + c] = b[set * n + 3 * i + c];
            }
        return;
    }
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
    double t2[2];
    reduce_partials_k<2>(part, nb, sh, t2);
    const double rz = t2[0], rr = t2[1];
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


// ---------------------------------------------------------------------------
// BiCGStab with device scalars (BatchBiCG)
// ---------------------------------------------------------------------------
constexpr int kBicgDots = 5;

// y = A u for the rows of this block (LANES lanes per block row); DOTS = 1: partial
// rh.y; DOTS = 5: partials y.w, y.y, rh.y, rh.w, w.w (y = t, w = s)
// this lane's share of row r of A u: one block per step, or 144-byte chunks of
// 2 (FP64) / 4 (FP32) blocks read with 16-byte loads (values and columns padded)
template <typename VT, int LANES, bool CHUNKED>
__device__ __forceinline__ void row_part(int64_t r, int sub, const int64_t* rowptr, const int32_t* colind,
                                         const VT* vals, const double* u, double& s0, double& s1, double& s2) {
    const int64_t k0 = rowptr[r], k1 = rowptr[r + 1];
    if constexpr (!CHUNKED) {
        for (int64_t k = k0 + sub; k < k1; k += LANES) {
            const int64_t c = colind[k];
            const VT* a = vals + 9 * k;
            const double x0 = u[3 * c], x1 = u[3 * c + 1], x2 = u[3 * c + 2];
            s0 += (double)a[0] * x0 + (double)a[1] * x1 + (double)a[2] * x2;
            s1 += (double)a[3] * x0 + (double)a[4] * x1 + (double)a[5] * x2;
            s2 += (double)a[6] * x0 + (double)a[7] * x1 + (double)a[8] * x2;
        }
    } else {
        constexpr int BPC = sizeof(VT) == 8 ? 2 : 4;
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
                const double x0 = u[3 * c], x1 = u[3 * c + 1], x2 = u[3 * c + 2];
                s0 += (double)a[0] * x0 + (double)a[1] * x1 + (double)a[2] * x2;
                s1 += (double)a[3] * x0 + (double)a[4] * x1 + (double)a[5] * x2;
                s2 += (double)a[6] * x0 + (double)a[7] * x1 + (double)a[8] * x2;
            }
        }
    }
}

// y = A u for the rows of this block (LANES lanes per block row); DOTS = 1: partial
// rh.y; DOTS = 5: partials y.w, y.y, rh.y, rh.w, w.w (y = t, w = s)
template <typename VT, int LANES, int DOTS, bool CHUNKED>
__global__ __launch_bounds__(kThreads) void k_bicg_spmv(int64_t N, const int64_t* rowptr, const int32_t* colind,
                                                        const VT* vals, int64_t set_vals, const double* u, double* y,
                                                        const double* rh, const double* w, const int32_t* done,
                                                        double* part, int64_t part_ld) {
    __shared__ double sh[4];
    const int64_t set = blockIdx.y;
    if (done[set]) return;
    const int64_t n = 3 * N;
    vals += set * set_vals;
    u += set * n;
    y += set * n;
    rh += set * n;
    if (w) w += set * n;
    const int64_t gt = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t r = gt / LANES;
    const int sub = (int)(gt % LANES);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (r < N) row_part<VT, LANES, CHUNKED>(r, sub, rowptr, colind, vals, u, s0, s1, s2);
#pragma unroll
    for (int o = LANES / 2; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, LANES);
        s1 += __shfl_xor(s1, o, LANES);
        s2 += __shfl_xor(s2, o, LANES);
    }
    double d[DOTS];
#pragma unroll
    for (int j = 0; j < DOTS; j++) d[j] = 0.0;
    if (r < N && sub == 0) {
        y[3 * r] = s0;
        y[3 * r + 1] = s1;
        y[3 * r + 2] = s2;
        const double yv[3] = {s0, s1, s2};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double hr = rh[3 * r + c];
            if (DOTS == 1) {
                d[0] += hr * yv[c];
            } else {
                const double wc = w[3 * r + c];
                d[0] += yv[c] * wc;
                d[DOTS > 1 ? 1 : 0] += yv[c] * yv[c];
                d[DOTS > 2 ? 2 : 0] += hr * yv[c];
                d[DOTS > 3 ? 3 : 0] += hr * wc;
                d[DOTS > 4 ? 4 : 0] += wc * wc;
            }
        }
    }
    // one wave-level reduction per value, then a single exchange through LDS
    __shared__ double shm[DOTS][kThreads / 64];
#pragma unroll
    for (int j = 0; j < DOTS; j++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) d[j] += __shfl_xor(d[j], o, 64);
    if (threadIdx.x % 64 == 0)
#pragma unroll
        for (int j = 0; j < DOTS; j++) shm[j][threadIdx.x / 64] = d[j];
    __syncthreads();
    if (threadIdx.x < DOTS) {
        double* out = part + set * part_ld + (int64_t)DOTS * blockIdx.x;
        out[threadIdx.x] = shm[threadIdx.x][0] + shm[threadIdx.x][1] + shm[threadIdx.x][2] + shm[threadIdx.x][3];
    }
    (void)sh;
}

// z = D^-1 u (node per thread)
__device__ __forceinline__ void bj3(const double* a, const double* u, double* z) {
    z[0] = a[0] * u[0] + a[1] * u[1] + a[2] * u[2];
    z[1] = a[3] * u[0] + a[4] * u[1] + a[5] * u[2];
    z[2] = a[6] * u[0] + a[7] * u[1] + a[8] * u[2];
}

__global__ __launch_bounds__(kThreads) void k_bicg_precond_p(BatchBiCG k, int64_t N, const double* dinv) {
    const int64_t set = blockIdx.y;
    if (k.done[set]) return;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    const int64_t o = set * 3 * N + 3 * i;
    bj3(dinv + set * 9 * N + 9 * i, k.p + o, k.ph + o);
}

__global__ __launch_bounds__(kThreads) void k_bicg_alpha(BatchBiCG k, int nb) {
    __shared__ double sh[4];
    const int set = blockIdx.x;
    if (k.done[set]) return;
    const double rv = reduce_partials(k.part + (int64_t)set * k.part_ld, nb, 1, 0, sh);
    if (threadIdx.x == 0) {
        if (rv != 0.0 && isfinite(rv)) {
            k.alpha[set] = k.rho[set] / rv;
        } else {
            k.done[set] = 2;
        }
    }
}

// s = r - alpha v, sh = D^-1 s (s only without dinv)
__global__ __launch_bounds__(kThreads) void k_bicg_s(BatchBiCG k, int64_t N, const double* dinv) {
    const int64_t set = blockIdx.y;
    if (k.done[set]) return;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    const double al = k.alpha[set];
    const int64_t o = set * 3 * N + 3 * i;
    double sv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        sv[c] = k.r[o + c] - al * k.v[o + c];
        k.s[o + c] = sv[c];
    }
    if (dinv) bj3(dinv + set * 9 * N + 9 * i, sv, k.sh + o);
}

__global__ __launch_bounds__(kThreads) void k_bicg_omega(BatchBiCG k, int nb) {
    __shared__ double sh[4];
    const int set = blockIdx.x;
    if (k.done[set]) return;
    const double* part = k.part + (int64_t)set * k.part_ld;
    double d[kBicgDots];
    reduce_partials_k<kBicgDots>(part, nb, sh, d);
    if (threadIdx.x == 0) {
        const double ts = d[0], tt = d[1], rht = d[2], rhs = d[3], ss = d[4];
        const double om = (tt > 0.0) ? ts / tt : 0.0;
        const double rho_new = rhs - om * rht;
        const double rr = ss - 2.0 * om * ts + om * om * tt;
        k.omega[set] = om;
        k.rr[set] = rr;
        k.it[set] += 1;
        int fin = 0;
        if (rr <= k.tol2[set]) fin = 1;
        else if (om == 0.0 || rho_new == 0.0 || !isfinite(rho_new) || !isfinite(rr)) fin = 2;
        if (fin == 0) k.beta[set] = (rho_new / k.rho[set]) * (k.alpha[set] / om);
        k.rho[set] = rho_new;
        k.fin[set] = fin;
    }
}

// x += alpha ph + omega sh; r = s - omega t; p = r + beta (p - omega v) unless finishing
__global__ __launch_bounds__(kThreads) void k_bicg_update(BatchBiCG k, int64_t n) {
    const int64_t set = blockIdx.y;
    if (k.done[set]) return;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int64_t o = set * n + i;
    const double al = k.alpha[set], om = k.omega[set];
    k.x[o] += al * k.ph[o] + om * k.sh[o];
    const double rn = k.s[o] - om * k.t[o];
    k.r[o] = rn;
    if (k.fin[set] == 0) k.p[o] = rn + k.beta[set] * (k.p[o] - om * k.v[o]);
}

__global__ void k_bicg_commit(BatchBiCG k) {
    const int set = blockIdx.x * blockDim.x + threadIdx.x;
    if (set >= k.S || k.done[set]) return;
    if (k.fin[set]) k.done[set] = k.fin[set];
}

// after r = b - A x: r^ = r, p = r; partials r.r and b.b
__global__ __launch_bounds__(kThreads) void k_bicg_init(BatchBiCG k, int64_t n, const double* b,
                                                        const int32_t* active, int nb) {
    __shared__ double sh[4];
    const int64_t set = blockIdx.y;
    if (!active[set]) return;
    double rr = 0.0, bb = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)nb * kThreads) {
        const int64_t o = set * n + i;
        const double rv = k.r[o], bv = b[o];
        k.rh[o] = rv;
        k.p[o] = rv;
        rr += rv * rv;
        bb += bv * bv;
    }
    const double trr = block_sum256(rr, sh);
    const double tbb = block_sum256(bb, sh);
    if (threadIdx.x == 0) {
        double* part = k.part + set * k.part_ld + 2 * blockIdx.x;
        part[0] = trr;
        part[1] = tbb;
    }
}

__global__ __launch_bounds__(kThreads) void k_bicg_init_reduce(BatchBiCG k, int nb, const double* rtol,
                                                               const int32_t* active) {
    __shared__ double sh[4];
    const int set = blockIdx.x;
    if (!active[set]) {
        if (threadIdx.x == 0) {
            k.done[set] = 1;
            k.fin[set] = 0;
        }
        return;
    }
    const double* part = k.part + (int64_t)set * k.part_ld;
    double t2[2];
    reduce_partials_k<2>(part, nb, sh, t2);
    const double rr = t2[0], bb = t2[1];
    if (threadIdx.x == 0) {
        const double bt = (bb > 0.0) ? bb : 1.0;
        k.rho[set] = rr;
        k.rr[set] = rr;
        k.tol2[set] = rtol[set] * rtol[set] * bt;
        k.fin[set] = 0;
        k.done[set] = (rr <= k.tol2[set]) ? 1 : 0;
    }
}

}  // namespace

int64_t bicg_part_ld(int64_t N) {
    return std::max<int64_t>((int64_t)kBicgDots * batch_spmv_blocks(N, 16), 2 * (int64_t)batch_dot_blocks(3 * N)) + 1;
}

void launch_bicg_init(const BatchBiCG& k, int64_t N, const double* b, const double* rtol, const int32_t* active,
                      hipStream_t s) {
    const int nb = batch_dot_blocks(3 * N);
    hipLaunchKernelGGL(k_bicg_init, dim3(nb, k.S), dim3(kThreads), 0, s, k, 3 * N, b, active, nb);
    hipLaunchKernelGGL(k_bicg_init_reduce, dim3(k.S), dim3(kThreads), 0, s, k, nb, rtol, active);
}

void launch_bicg_iteration(const BatchBiCG& k, int lanes, bool chunked, int64_t N, const int64_t* rowptr,
                           const int32_t* colind, const void* vals, bool fp32, int64_t set_vals, const double* dinv,
                           hipStream_t s, PrecondApply apply, void* apply_ctx) {
    const int nbn = batch_node_blocks(N);
    const int nbs = batch_spmv_blocks(N, lanes);
    const int64_t n = 3 * N;
    if (apply)
        apply(apply_ctx, k.p, k.ph);
    else
        hipLaunchKernelGGL(k_bicg_precond_p, dim3(nbn, k.S), dim3(kThreads), 0, s, k, N, dinv);
#define CMG_BSPMV(VT, L, D, U, Y, W)                                                                              \
    if (chunked)                                                                                                  \
        hipLaunchKernelGGL((k_bicg_spmv<VT, L, D, true>), dim3(nbs, k.S), dim3(kThreads), 0, s, N, rowptr, colind, \
                           (const VT*)vals, set_vals, U, Y, k.rh, W, k.done, k.part, k.part_ld);                  \
    else                                                                                                          \
        hipLaunchKernelGGL((k_bicg_spmv<VT, L, D, false>), dim3(nbs, k.S), dim3(kThreads), 0, s, N, rowptr,       \
                           colind, (const VT*)vals, set_vals, U, Y, k.rh, W, k.done, k.part, k.part_ld)
#define CMG_BSPMV_L(D, U, Y, W)                                       \
    if (fp32) {                                                       \
        if (lanes == 4) CMG_BSPMV(float, 4, D, U, Y, W);              \
        else if (lanes == 16) CMG_BSPMV(float, 16, D, U, Y, W);       \
        else CMG_BSPMV(float, 8, D, U, Y, W);                         \
    } else {                                                          \
        if (lanes == 4) CMG_BSPMV(double, 4, D, U, Y, W);             \
        else if (lanes == 16) CMG_BSPMV(double, 16, D, U, Y, W);      \
        else CMG_BSPMV(double, 8, D, U, Y, W);                        \
    }
    CMG_BSPMV_L(1, k.ph, k.v, (const double*)nullptr)
    hipLaunchKernelGGL(k_bicg_alpha, dim3(k.S), dim3(kThreads), 0, s, k, nbs);
    hipLaunchKernelGGL(k_bicg_s, dim3(nbn, k.S), dim3(kThreads), 0, s, k, N, apply ? nullptr : dinv);
    if (apply) apply(apply_ctx, k.s, k.sh);
    CMG_BSPMV_L(kBicgDots, k.sh, k.t, k.s)
#undef CMG_BSPMV_L
#undef CMG_BSPMV
    hipLaunchKernelGGL(k_bicg_omega, dim3(k.S), dim3(kThreads), 0, s, k, nbs);
    hipLaunchKernelGGL(k_bicg_update, dim3((unsigned)((n + kThreads - 1) / kThreads), k.S), dim3(kThreads), 0, s, k,
                       n);
    hipLaunchKernelGGL(k_bicg_commit, dim3((k.S + 63) / 64), dim3(64), 0, s, k);
}

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
                          const void* vals, bool fp32, int64_t set_vals, const double* dinv, hipStream_t s,
                          PrecondApply apply, void* apply_ctx) {
    launch_bcg_spmv_dot(cg, lanes, N, rowptr, colind, vals, fp32, set_vals, s);
    hipLaunchKernelGGL(k_bcg_alpha, dim3(cg.S), dim3(kThreads), 0, s, cg, batch_spmv_blocks(N, lanes));
    const int nbn = batch_node_blocks(N);
    hipLaunchKernelGGL(k_bcg_update, dim3(nbn, cg.S), dim3(kThreads), 0, s, cg, N, apply ? nullptr : dinv, nbn);
    if (apply) {
        apply(apply_ctx, cg.r, cg.z);
        hipLaunchKernelGGL(k_bcg_rz, dim3(nbn, cg.S), dim3(kThreads), 0, s, cg, N, (const int32_t*)nullptr, false);
    }
    hipLaunchKernelGGL(k_bcg_beta, dim3(cg.S), dim3(kThreads), 0, s, cg, nbn);
    const int64_t n = 3 * N;
    hipLaunchKernelGGL(k_bcg_p, dim3((unsigned)((n + kThreads - 1) / kThreads), cg.S), dim3(kThreads), 0, s, cg, n);
}

void launch_bcg_init(const BatchCG& cg, int64_t N, const double* b, const double* dinv, const double* rtol,
                     const int32_t* active, hipStream_t s, PrecondApply apply, void* apply_ctx) {
    const int nbn = batch_node_blocks(N);
    hipLaunchKernelGGL(k_bcg_init, dim3(nbn, cg.S), dim3(kThreads), 0, s, cg, N, b, apply ? nullptr : dinv, active);
    if (apply) {
        apply(apply_ctx, cg.r, cg.z);
        hipLaunchKernelGGL(k_bcg_rz, dim3(nbn, cg.S), dim3(kThreads), 0, s, cg, N, active, true);
    }
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
