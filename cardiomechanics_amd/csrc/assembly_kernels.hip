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
#pragma clang fp contract(off)
        const double m = cm * mval[slot];   // rounded, then added (as the row gather)
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

// Row gather: one wave per block row r, four rows per workgroup.  The row's blocks
// accumulate in LDS (row_cap blocks per wave).  For each solid incidence (e, a) of the
// row, in element order, K_e(a, 0..npe-1) is one contiguous run of 9 npe values that the
// first 9 npe / 2 lanes read two values at a time (coalesced, every fetched segment used
// by this wave) and add at the row positions rpos of the blocks; the surface incidences
// follow in surface order.  Within one incidence every (block, entry) has one lane, and
// a wave's LDS operations are carried out in issue order, so each block sums its
// contributions in the order of its contributor list (LDS adds round like v_add_f64):
// the values equal the slot gather's bit for bit.  The adds return nothing, so a wave
// issues them back to back.  The row's column indices, mass scalars and column
// constraint masks (cmask: the bcmask of each slot's column, for the context's bcmask)
// are loaded before the element values, so their latency overlaps.  Untransposed
// maps only (a transposed block of the run would be a 72-byte piece 9 npe values apart).
constexpr int kRowWaves = 4;
constexpr int kRowPrefetch = 8;   // incidences loaded before their values are added

template <typename KT, typename VT>
__global__ __launch_bounds__(64 * kRowWaves) void k_gather_rows(GatherMaps g, const KT* Ke, const double* Ks,
                                                               const double* mval, double cm,
                                                               const uint8_t* bcmask, VT* vals, SetStrides st) {
    extern __shared__ double row_acc[];   // [kRowWaves][row_cap * 9]
    const int64_t set = blockIdx.y;
    if (st.active && !st.active[set]) return;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int l = (int)(threadIdx.x & 63);
    const int64_t r = (int64_t)blockIdx.x * kRowWaves + w;
    if (r >= g.N) return;   // whole waves: nothing below waits on another wave
    Ke += set * st.a;
    if (Ks) Ks += set * st.b;
    vals += set * st.c;
    const int npe = g.npe;
    const int64_t rb = g.rowptr[r];
    const int nb = (int)(g.rowptr[r + 1] - rb);
    const int64_t k0 = g.inc_ptr[r], k1 = g.inc_ptr[r + 1];
    // the row's blocks l and l + 64 (lane per block in the final pass): columns, mass
    // scalars and column masks, requested before the element values
    int32_t colA = 0, colB = 0;
    double mA = 0.0, mB = 0.0;
    uint8_t cA = 0, cB = 0;
    if (l < nb) {
        colA = g.colind[rb + l];
        if (mval) mA = mval[rb + l];
        if (bcmask) cA = g.cmask[rb + l];
    }
    if (l + 64 < nb) {
        colB = g.colind[rb + l + 64];
        if (mval) mB = mval[rb + l + 64];
        if (bcmask) cB = g.cmask[rb + l + 64];
    }
    double* acc = row_acc + (int64_t)w * g.row_cap * 9;
    for (int t = l; t < 9 * nb; t += 64) acc[t] = 0.0;
    // solid incidences: lane l < 9 npe / 2 holds entries 2l and 2l + 1 of the run
    const int nv = 9 * npe;
    const bool act = 2 * l < nv;
    const int q0 = act ? 2 * l : 0, q1 = q0 + 1;
    const int b0 = q0 / 9, e0 = q0 - 9 * b0, b1 = q1 / 9, e1 = q1 - 9 * b1;
    for (int64_t kb = k0; kb < k1; kb += kRowPrefetch) {
        // unconditional loads (indices past the row's end repeat its last incidence), so
        // all of them are in flight together; only the adds below are conditional
        int64_t code[kRowPrefetch];
#pragma unroll
        for (int u = 0; u < kRowPrefetch; u++) code[u] = g.inc[kb + u < k1 ? kb + u : k1 - 1];
        double v0[kRowPrefetch], v1[kRowPrefetch];
        int p0[kRowPrefetch], p1[kRowPrefetch];
#pragma unroll
        for (int u = 0; u < kRowPrefetch; u++) {
            const int64_t k = kb + u < k1 ? kb + u : k1 - 1;
            const KT* run = Ke + code[u] * nv + q0;
            if constexpr (sizeof(KT) == 8) {
                const double2 v = *reinterpret_cast<const double2*>(run);   // runs are 16-byte aligned
                v0[u] = v.x;
                v1[u] = v.y;
            } else {
                const float2 v = *reinterpret_cast<const float2*>(run);     // 8-byte aligned
                v0[u] = (double)v.x;
                v1[u] = (double)v.y;
            }
            p0[u] = g.rpos[16 * k + b0];
            p1[u] = g.rpos[16 * k + b1];
        }
#pragma unroll
        for (int u = 0; u < kRowPrefetch; u++) {
            if (kb + u < k1 && act) {
                atomicAdd(&acc[9 * p0[u] + e0], v0[u]);
                atomicAdd(&acc[9 * p1[u] + e1], v1[u]);
            }
        }
    }
    // surface incidences: the rows 3a..3a+2 of the surface block matrix, 9 nps values
    // ([i][b][j]), one per lane
    if (g.sinc_ptr && Ks) {
        const int nps = g.nps, ns = 9 * nps;
        const bool sact = l < ns;
        const int i = sact ? l / (3 * nps) : 0, cc = l - i * 3 * nps;
        const int sb = cc / 3, sj = cc - 3 * sb;
        for (int64_t k = g.sinc_ptr[r]; k < g.sinc_ptr[r + 1]; k++) {
            if (sact) atomicAdd(&acc[9 * g.spos[8 * k + sb] + 3 * i + sj], Ks[(int64_t)g.sinc[k] * ns + l]);
        }
    }
    // + cm M on the diagonal, identity rows / columns of constrained DOFs; lane per block
    const uint8_t mr = bcmask ? bcmask[r] : 0;
    for (int kb = l; kb < nb; kb += 64) {
        const int64_t slot = rb + kb;
        int32_t col;
        double m;
        uint8_t mc;
        if (kb < 128) {
            col = kb < 64 ? colA : colB;
            m = kb < 64 ? mA : mB;
            mc = kb < 64 ? cA : cB;
        } else {
            col = g.colind[slot];
            m = mval ? mval[slot] : 0.0;
            mc = bcmask ? g.cmask[slot] : 0;
        }
        double v[9];
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = acc[9 * kb + k];
        if (mval) {
#pragma clang fp contract(off)
            const double cmm = cm * m;   // rounded, then added (as the slot gather)
#pragma unroll
            for (int i = 0; i < 3; i++) v[4 * i] += cmm;
        }
        if (mr | mc) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    if (((mr >> i) & 1) || ((mc >> j) & 1)) v[3 * i + j] = (r == col && i == j) ? 1.0 : 0.0;
        }
        VT* out = vals + 9 * slot;
#pragma unroll
        for (int k = 0; k < 9; k++) out[k] = (VT)v[k];
    }
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
    if (g.row_cap > 0 && !g.transposed && !flip && g.rpos && (g.npe == 10 || g.npe == 4)) {
        const dim3 grid((unsigned)((g.N + kRowWaves - 1) / kRowWaves), b.n_sets);
        const size_t lds = sizeof(double) * kRowWaves * 9 * (size_t)g.row_cap;
#define CMG_ROWS(KT, VT)                                                                                   \
    hipLaunchKernelGGL((k_gather_rows<KT, VT>), grid, dim3(64 * kRowWaves), lds, s, g, (const KT*)Ke, Ks, mval, \
                       cm, bcmask, (VT*)vals, b)
        if (fp32) {
            if (vals_fp32)
                CMG_ROWS(float, float);
            else
                CMG_ROWS(float, double);
        } else {
            if (vals_fp32)
                CMG_ROWS(double, float);
            else
                CMG_ROWS(double, double);
        }
#undef CMG_ROWS
        return;
    }
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
