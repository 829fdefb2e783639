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

constexpr int kMassLanes = 16;

__global__ __launch_bounds__(256) void k_add_mass_term(int64_t N, const int64_t* rowptr, const int32_t* colind,
                                                       const double* mval, double cm, const double* d,
                                                       const double* dp, const uint8_t* bcmask, double* f) {
    // kMassLanes lanes per block row (coalesced reads of the row's mass scalars and
    // columns), summed over the lanes in a fixed butterfly order
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = t / kMassLanes;
    const int l = (int)(t - r * kMassLanes);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (r < N)
        for (int64_t k = rowptr[r] + l; k < rowptr[r + 1]; k += kMassLanes) {
            const int64_t c = colind[k];
            const double mv = mval[k];
            s0 += mv * (d[3 * c] - dp[3 * c]);
            s1 += mv * (d[3 * c + 1] - dp[3 * c + 1]);
            s2 += mv * (d[3 * c + 2] - dp[3 * c + 2]);
        }
#pragma unroll
    for (int off = kMassLanes / 2; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off, kMassLanes);
        s1 += __shfl_xor(s1, off, kMassLanes);
        s2 += __shfl_xor(s2, off, kMassLanes);
    }
    if (r >= N || l != 0) return;
    const uint8_t m = bcmask ? bcmask[r] : 0;
    if (!(m & 1)) f[3 * r] += cm * s0;
    if (!(m & 2)) f[3 * r + 1] += cm * s1;
    if (!(m & 4)) f[3 * r + 2] += cm * s2;
}

// one 3x3 block of a contribution: solid element block idx (layout [e][a][b][3][3]) or a
// surface block (idx >= n_solid), accumulated into s[9] (transposed when trans)
template <typename KT>
__device__ __forceinline__ void add_block(double* s, int32_t code, const KT* Ke, const double* Ks, int32_t n_solid,
                                          int npe, bool flip) {
    bool trans = code < 0;
    int32_t idx = trans ? ~code : code;
    if (idx < n_solid) {
        if (flip) {
            // the same element's block (b,a) instead of (a,b), transposed: assembles K_e^T
            const int npe2 = npe * npe;
            const int32_t e = idx / npe2;
            const int32_t rem = idx - e * npe2;
            const int32_t p = rem / npe, q = rem - (rem / npe) * npe;
            idx = e * npe2 + q * npe + p;
            trans = !trans;
        }
        const KT* blk = Ke + 9 * (int64_t)idx;
        double v[9];
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = (double)blk[k];
        if (trans) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) s[3 * i + j] += v[3 * j + i];
        } else {
#pragma unroll
            for (int k = 0; k < 9; k++) s[k] += v[k];
        }
    } else {
        // surface block: idx - n_solid = (s * nps + a) * nps + b, packed with nps in the low bits
        const int32_t sidx = idx - n_solid;
        const int nps = (int)(sidx & 7);
        const int32_t q = sidx >> 3;
        const int32_t si = q / (nps * nps);
        const int ab = (int)(q - si * nps * nps);
        const int a = ab / nps, b = ab - nps * (ab / nps);
        const int n = 3 * nps;
        const double* base = Ks + (int64_t)si * n * n + (int64_t)(3 * a) * n + 3 * b;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) s[3 * i + j] += trans ? base[(int64_t)j * n + i] : base[(int64_t)i * n + j];
    }
}

// One thread per 3x3 block.  The first `light_blocks` workgroups take the slots in
// pattern order and skip the heavy ones (more than kHeavyContrib contributions: the
// diagonal blocks of element vertices, ~20-30 each); the remaining workgroups take the
// heavy slots from their own list, so no wave idles on one long contributor list.
template <typename KT, typename VT>
__global__ __launch_bounds__(256) void k_gather_matrix(int64_t nnzb, const int64_t* cptr, const int32_t* contrib,
                                                       const int32_t* heavy, int64_t n_heavy, int64_t light_blocks,
                                                       const KT* Ke, const double* Ks, int32_t n_solid, int npe,
                                                       const int32_t* blkrow, const int32_t* colind,
                                                       const double* mval, double cm, const uint8_t* bcmask,
                                                       VT* vals, bool flip, SetStrides st) {
    const int64_t set = blockIdx.y;
    if (st.active && !st.active[set]) return;
    // (workgroups in plain order: giving each XCD a contiguous range of slots measured
    // 19 % slower on the 200k-element LV, with the same HBM fetch)
    int64_t slot;
    if ((int64_t)blockIdx.x < light_blocks) {
        slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (slot >= nnzb) return;
        if (cptr[slot + 1] - cptr[slot] > kHeavyContrib) return;
    } else {
        const int64_t h = ((int64_t)blockIdx.x - light_blocks) * blockDim.x + threadIdx.x;
        if (h >= n_heavy) return;
        slot = heavy[h];
    }
    Ke += set * st.a;
    if (Ks) Ks += set * st.b;
    vals += set * st.c;
    double s[9];
#pragma unroll
    for (int k = 0; k < 9; k++) s[k] = 0.0;
    const int64_t c1 = cptr[slot + 1];
    for (int64_t c = cptr[slot]; c < c1; c++) add_block(s, contrib[c], Ke, Ks, n_solid, npe, flip);
    const int64_t r = blkrow[slot], col = colind[slot];
    if (mval) {
        const double m = cm * mval[slot];
#pragma unroll
        for (int i = 0; i < 3; i++) s[4 * i] += m;
    }
    const uint8_t mr = bcmask ? bcmask[r] : 0, mc = bcmask ? bcmask[col] : 0;
    if (mr | mc) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                if (((mr >> i) & 1) || ((mc >> j) & 1)) s[3 * i + j] = (r == col && i == j) ? 1.0 : 0.0;
    }
    VT* out = vals + 9 * slot;
#pragma unroll
    for (int k = 0; k < 9; k++) out[k] = (VT)s[k];
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
    hipLaunchKernelGGL(k_add_mass_term, dim3((unsigned)((N * kMassLanes + 255) / 256)), dim3(256), 0, s, N, rowptr,
                       colind, mval, cm, d, dpred,
                       bcmask, f);
}

void launch_gather_matrix(const GatherMaps& g, const void* Ke, bool fp32, const double* Ks, const double* mval,
                          double cm, const uint8_t* bcmask, void* vals, bool vals_fp32, bool flip, hipStream_t s,
                          const SetStrides& b) {
    if (g.nnzb == 0) return;
    const int64_t light_blocks = (g.nnzb + 255) / 256;
    const dim3 grid((unsigned)(light_blocks + (g.n_heavy + 255) / 256), b.n_sets);
#define CMG_GATHER(KT, VT)                                                                                       \
    hipLaunchKernelGGL((k_gather_matrix<KT, VT>), grid, dim3(256), 0, s, g.nnzb, g.cptr, g.contrib, g.heavy,      \
                       g.n_heavy, light_blocks, (const KT*)Ke, Ks, g.n_solid, g.npe, g.blkrow, g.colind, mval, cm, \
                       bcmask, (VT*)vals, flip, b)
    if (fp32) {
        if (vals_fp32)
            CMG_GATHER(float, float);
        else
            CMG_GATHER(float, double);
    } else {
        if (vals_fp32)
            CMG_GATHER(double, float);
        else
            CMG_GATHER(double, double);
    }
#undef CMG_GATHER
}

}  // namespace cmg
