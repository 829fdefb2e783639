// Deterministic, atomic-free assembly for gfx950.
//
// The reference scatters every element vector/matrix with ADD_VALUES into a
// PETSc Vec/AIJ Mat (CBElementAdapter.cpp:138-163).  Here every output entry
// gathers its contributions instead, in a fixed order precomputed at context
// creation, so results are bitwise reproducible and no atomics are needed:
//   vector: node r sums fe over its incident (element, local node) pairs;
//   matrix: each 3x3 block sums the element/surface blocks that map to it, adds
//           cm * M (mass stored as one scalar per block) and puts the identity on
//           constrained DOFs (CBSolver.cpp:413-435, :323-324).
#include "cmg_internal.h"

namespace cmg {
namespace {

__global__ __launch_bounds__(256) void k_gather_vector(int64_t N, const int64_t* inc_ptr, const int32_t* inc,
                                                       const double* fe, const int64_t* sinc_ptr,
                                                       const int32_t* sinc, const double* fs,
                                                       const uint8_t* bcmask, double* f, SetStrides b) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t set = blockIdx.y;
    if (r >= N || (b.active && !b.active[set])) return;
    fe += set * b.a;
    if (fs) fs += set * b.b;
    f += set * b.c;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t k = inc_ptr[r]; k < inc_ptr[r + 1]; k++) {
        const double* p = fe + 3 * (int64_t)inc[k];
        s0 += p[0];
        s1 += p[1];
        s2 += p[2];
    }
    if (sinc_ptr) {
        for (int64_t k = sinc_ptr[r]; k < sinc_ptr[r + 1]; k++) {
            const double* p = fs + 3 * (int64_t)sinc[k];
            s0 += p[0];
            s1 += p[1];
            s2 += p[2];
        }
    }
    const uint8_t m = bcmask ? bcmask[r] : 0;
    f[3 * r] = (m & 1) ? 0.0 : s0;
    f[3 * r + 1] = (m & 2) ? 0.0 : s1;
    f[3 * r + 2] = (m & 4) ? 0.0 : s2;
}

__global__ __launch_bounds__(256) void k_add_mass_term(int64_t N, const int64_t* rowptr, const int32_t* colind,
                                                       const double* mval, double cm, const double* d,
                                                       const double* dp, const uint8_t* bcmask, double* f) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
        const int64_t c = colind[k];
        const double mv = mval[k];
        s0 += mv * (d[3 * c] - dp[3 * c]);
        s1 += mv * (d[3 * c + 1] - dp[3 * c + 1]);
        s2 += mv * (d[3 * c + 2] - dp[3 * c + 2]);
    }
    const uint8_t m = bcmask ? bcmask[r] : 0;
    if (!(m & 1)) f[3 * r] += cm * s0;
    if (!(m & 2)) f[3 * r + 1] += cm * s1;
    if (!(m & 4)) f[3 * r + 2] += cm * s2;
}

template <typename KT, typename VT>
__global__ __launch_bounds__(256) void k_gather_matrix(int64_t nnzb, const int64_t* cptr, const int64_t* contrib,
                                                       const KT* Ke, const double* Ks, int64_t n_solid,
                                                       const int32_t* blkrow, const int32_t* colind,
                                                       const double* mval, double cm, const uint8_t* bcmask,
                                                       VT* vals, SetStrides st) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t set = blockIdx.y;
    if (t >= 9 * nnzb || (st.active && !st.active[set])) return;
    Ke += set * st.a;
    if (Ks) Ks += set * st.b;
    vals += set * st.c;
    const int64_t slot = t / 9;
    const int k = (int)(t - 9 * slot);
    const int i = k / 3, j = k - 3 * (k / 3);
    double sum = 0.0;
    for (int64_t c = cptr[slot]; c < cptr[slot + 1]; c++) {
        const int64_t code = contrib[c];
        const bool trans = code < 0;
        const int64_t idx = trans ? ~code : code;
        const int ki = trans ? j : i, kj = trans ? i : j;
        if (idx < n_solid) {
            sum += (double)Ke[9 * idx + 3 * ki + kj];
        } else {
            // surface block: idx - n_solid = (s * nps + a) * nps + b, packed with nps in the low bits
            const int64_t sidx = idx - n_solid;
            const int nps = (int)(sidx & 7);
            const int64_t q = sidx >> 3;
            const int64_t s = q / (nps * nps);
            const int ab = (int)(q - s * nps * nps);
            const int a = ab / nps, b = ab - nps * (ab / nps);
            const int n = 3 * nps;
            sum += Ks[s * n * n + (int64_t)(3 * a + ki) * n + 3 * b + kj];
        }
    }
    const int64_t r = blkrow[slot], col = colind[slot];
    if (mval && i == j) sum += cm * mval[slot];
    const uint8_t mr = bcmask ? bcmask[r] : 0, mc = bcmask ? bcmask[col] : 0;
    if (((mr >> i) & 1) || ((mc >> j) & 1)) sum = (r == col && i == j) ? 1.0 : 0.0;
    vals[t] = (VT)sum;
}

// C = beta_R K + alpha M (Rayleigh damping) in place over K's gathered BSR values, with
// the rows and columns of constrained DOFs zeroed (their velocities are zero and their
// residual rows are cleared, so these entries never act); beta_R = 0 ignores vals' old
// contents
__global__ __launch_bounds__(256) void k_rayleigh_finish(int64_t nnzb, const int32_t* blkrow, const int32_t* colind,
                                                         const uint8_t* bcmask, const double* mval, double beta_r,
                                                         double alpha, double* vals) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 9 * nnzb) return;
    const int64_t slot = t / 9;
    const int k = (int)(t - 9 * slot);
    const int i = k / 3, j = k - 3 * (k / 3);
    double v = beta_r != 0.0 ? beta_r * vals[t] : 0.0;
    if (i == j) v += alpha * mval[slot];
    const uint8_t mr = bcmask ? bcmask[blkrow[slot]] : 0, mc = bcmask ? bcmask[colind[slot]] : 0;
    if (((mr >> i) & 1) || ((mc >> j) & 1)) v = 0.0;
    vals[t] = v;
}

// vals += a * C over the BSR values (the damping part of the Newmark tangent)
template <typename VT>
__global__ __launch_bounds__(256) void k_add_scaled_values(int64_t nv, double a, const double* C, VT* vals) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
    vals[t] = (VT)((double)vals[t] + a * C[t]);
}

}  // namespace

void launch_rayleigh_finish(int64_t nnzb, const int32_t* blkrow, const int32_t* colind, const uint8_t* bcmask,
                            const double* mval, double beta_r, double alpha, double* vals, hipStream_t s) {
    if (nnzb == 0) return;
    hipLaunchKernelGGL(k_rayleigh_finish, dim3((unsigned)((9 * nnzb + 255) / 256)), dim3(256), 0, s, nnzb, blkrow,
                       colind, bcmask, mval, beta_r, alpha, vals);
}

void launch_add_scaled_values(int64_t nnzb, double a, const double* C, void* vals, bool vals_fp32, hipStream_t s) {
    if (nnzb == 0) return;
    const int64_t nv = 9 * nnzb;
    const dim3 grid((unsigned)((nv + 255) / 256));
    if (vals_fp32)
        hipLaunchKernelGGL(k_add_scaled_values<float>, grid, dim3(256), 0, s, nv, a, C, (float*)vals);
    else
        hipLaunchKernelGGL(k_add_scaled_values<double>, grid, dim3(256), 0, s, nv, a, C, (double*)vals);
}

void launch_gather_vector(int64_t N, const int64_t* inc_ptr, const int32_t* inc, const double* fe, int,
                          const int64_t* sinc_ptr, const int32_t* sinc, const double* fs, int,
                          const uint8_t* bcmask, double* f, hipStream_t s, const SetStrides& b) {
    if (N == 0) return;
    hipLaunchKernelGGL(k_gather_vector, dim3((N + 255) / 256, b.n_sets), dim3(256), 0, s, N, inc_ptr, inc, fe,
                       sinc_ptr, sinc, fs, bcmask, f, b);
}

void launch_add_mass_term(int64_t N, const int64_t* rowptr, const int32_t* colind, const double* mval, double cm,
                          const double* d, const double* dpred, const uint8_t* bcmask, double* f, hipStream_t s) {
    if (N == 0) return;
    hipLaunchKernelGGL(k_add_mass_term, dim3((N + 255) / 256), dim3(256), 0, s, N, rowptr, colind, mval, cm, d, dpred,
                       bcmask, f);
}

void launch_gather_matrix(int64_t nnzb, const int64_t* cptr, const int64_t* contrib, const void* Ke, bool fp32,
                          const double* Ks, int64_t n_solid_blocks, const int32_t* blkrow, const int32_t* colind,
                          const double* mval, double cm, const uint8_t* bcmask, void* vals, bool vals_fp32,
                          hipStream_t s, const SetStrides& b) {
    if (nnzb == 0) return;
    const dim3 grid((unsigned)((9 * nnzb + 255) / 256), b.n_sets);
    if (fp32) {
        if (vals_fp32)
            hipLaunchKernelGGL((k_gather_matrix<float, float>), grid, dim3(256), 0, s, nnzb, cptr, contrib,
                               (const float*)Ke, Ks, n_solid_blocks, blkrow, colind, mval, cm, bcmask, (float*)vals, b);
        else
            hipLaunchKernelGGL((k_gather_matrix<float, double>), grid, dim3(256), 0, s, nnzb, cptr, contrib,
                               (const float*)Ke, Ks, n_solid_blocks, blkrow, colind, mval, cm, bcmask, (double*)vals, b);
    } else {
        if (vals_fp32)
            hipLaunchKernelGGL((k_gather_matrix<double, float>), grid, dim3(256), 0, s, nnzb, cptr, contrib,
                               (const double*)Ke, Ks, n_solid_blocks, blkrow, colind, mval, cm, bcmask, (float*)vals, b);
        else
            hipLaunchKernelGGL((k_gather_matrix<double, double>), grid, dim3(256), 0, s, nnzb, cptr, contrib,
                               (const double*)Ke, Ks, n_solid_blocks, blkrow, colind, mval, cm, bcmask,
                               (double*)vals, b);
    }
}

}  // namespace cmg
