// Internal declarations of the gfx950 mechanics library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "cmg.h"

namespace cmg {

#ifdef __HIPCC__
// Workgroups are dispatched round-robin over the 8 XCDs; this maps workgroup bid of
// nblk to a tile index so every XCD gets a contiguous range of tiles and neighbouring
// tiles share one L2.
__device__ __forceinline__ int xcd_swizzle(int bid, int nblk) {
    const int nx = 8;
    const int full = (nblk / nx) * nx;
    if (bid >= full) return bid;
    const int per = nblk / nx;
    return (bid % nx) * per + bid / nx;
}
#endif

// Device copy of one material (same meaning as cmg_material).
struct DevMaterial {
    int law;
    int ignore_corrupt;
    int tension_fn;
    int pad;
    double density;
    double p[12];
    double tension_max;
    double fn[12];
};

// Elements handled by one 64-lane workgroup in the tangent kernel: 9 lanes
// (one per spatial component pair i, j) per element, 7 elements per wave.
constexpr int kTanElemsPerWave = 7;
constexpr int kResidualBlock = 128;

// Per-element data, structure-of-arrays: entry k of element e at [k * E + e].
struct ElemArgs {
    int64_t E = 0;
    const int32_t* conn = nullptr;     // [npe][E]
    const double* dNdX = nullptr;      // [nqp*npe*3][E]  reference dN/dX at the quadrature points
    const double* bases = nullptr;     // [nqp*9][E]      fibre frame (columns f, s, n) per quadrature point
    const double* w = nullptr;         // [E]  detJ/24 (T10) or volume (T4)
    const int32_t* matidx = nullptr;   // [E] or nullptr
    const double* tact = nullptr;      // [E]
    const DevMaterial* mats = nullptr;
    const double* x = nullptr;         // current coordinates [3N]
    double time = 0.0;
    double* fe = nullptr;              // [E][3npe] element forces
    void* Ke = nullptr;                // [E][npe][npe][3][3] element tangents (double or float)
    int* corrupt = nullptr;            // smallest element id with det F <= 0 (INT_MAX if none)
    // Parameter-set batch: blockIdx.y selects the set.  Set s reads x + s*set_x and
    // materials mats + s*set_mats, writes fe + s*set_fe, Ke + s*set_Ke_bytes and
    // corrupt[s]; sets whose active flag is 0 are skipped.
    int n_sets = 1;
    int set_mats = 0;
    int64_t set_x = 0, set_fe = 0, set_Ke_bytes = 0;
    const int32_t* active = nullptr;
    // Land17 (materials with tension_fn CMG_FN_LAND17): parameter blocks [n_mats][40],
    // committed state [8][E], calcium transients [2][1001], (time - commit time) in ms
    const double* l17_params = nullptr;
    const double* l17_state = nullptr;
    const double* l17_cai = nullptr;
    double l17_dt_ms = 0.0;
};

struct SurfArgs {
    int64_t S = 0;
    int nps = 3;
    const int32_t* conn = nullptr;     // [S][nps]
    const int32_t* group = nullptr;    // [S]
    const double* pressure = nullptr;  // [n_groups]
    const uint8_t* bcmask = nullptr;   // [N]
    const double* x = nullptr;         // [3N]
    double* fs = nullptr;              // [S][3nps]
    double* Ks = nullptr;              // [S][3nps][3nps] or nullptr
    double* vol_partial = nullptr;     // [S] or nullptr
    // parameter-set batch (blockIdx.y = set), as in ElemArgs
    int n_sets = 1;
    int64_t set_x = 0, set_fs = 0, set_Ks = 0, set_vol = 0;
    const int32_t* active = nullptr;
};

// ---- launchers (element_kernels.hip) ----
// per-element export state: out[18][E] = F at the centroid (global frame), averaged passive PK2
void launch_element_export(int npe, const ElemArgs& a, const double* Xref, double* out, hipStream_t s);
void launch_element_residual(int npe, const ElemArgs& a, hipStream_t s);
// variant: 1 = per-lane quadrature points, 2 = shared stage (LDS), 3 = shared stage, two node passes,
// 4 = shared stage with two lanes (two waves) per element and spatial pair; 5 / 6 = variants 2 / 4
// with two 7-element groups per workgroup, so the quadrature stage fills 56 lanes of the first wave
void launch_element_residual_tangent(int npe, bool fp32, const ElemArgs& a, hipStream_t s, int variant);
void launch_surface(const SurfArgs& a, bool tangent, hipStream_t s);
void launch_t10_shape(int64_t E, const int32_t* conn_aos, const double* X, double* dNdX_soa, double* w, hipStream_t s);
// Land17 commit: each Land17 element advances its state [8][E] in place from the
// evaluation at its last quadrature point at coordinates a.x and time a.time
void launch_land17_commit(int npe, const ElemArgs& a, double* state, hipStream_t s);
constexpr int kLand17Params = 40;

// ---- launchers (assembly_kernels.hip) ----
// vector gather: f[dof] = sum over incidences; BC components zeroed
// Batched form (n_sets > 1): set k reads fe + k*set_fe, fs + k*set_fs and writes f + k*set_f.
struct SetStrides {
    int n_sets = 1;
    int64_t a = 0, b = 0, c = 0;   // per-set strides of the kernel's arrays, in elements
    const int32_t* active = nullptr;
};
void launch_gather_vector(int64_t N, const int64_t* inc_ptr, const int32_t* inc, const double* fe, int stride_e,
                          const int64_t* sinc_ptr, const int32_t* sinc, const double* fs, int stride_s,
                          const uint8_t* bcmask, double* f, hipStream_t s, const SetStrides& b = SetStrides());
// f += cm * M (d - dt) on free DOFs (Newmark inertia term); M stored as one scalar per block
void launch_add_mass_term(int64_t N, const int64_t* rowptr, const int32_t* colind, const double* mval, double cm,
                          const double* d, const double* dpred, const uint8_t* bcmask, double* f, hipStream_t s);
// Block-matrix assembly maps of a context.  Contributor codes (int32): a solid element
// block (e * npe + a) * npe + b of K_e ([E][npe][npe][3][3]) or n_solid + a packed surface
// block; ~code means the block enters transposed.  Slots with more than kHeavyContrib
// contributors are also listed in `heavy` and gathered by their own threads.
constexpr int kHeavyContrib = 8;
struct GatherMaps {
    int64_t nnzb = 0;
    const int64_t* cptr = nullptr;     // [nnzb + 1]
    const int32_t* contrib = nullptr;  // [cptr[nnzb]]
    const int32_t* heavy = nullptr;    // [n_heavy] slots
    int64_t n_heavy = 0;
    int32_t n_solid = 0;               // E * npe * npe
    int npe = 10;
    const int32_t* blkrow = nullptr;   // [nnzb] block row of each slot
    const int32_t* colind = nullptr;   // [nnzb]
    // Row gather (untransposed maps): the incidence lists of the residual gather give
    // each block row's (element, local node) pairs in element order and its (surface,
    // local node) pairs in surface order -- the order of every slot's contributor list
    // -- and rpos / spos the position within the row of the block of each local node.
    // row_cap = 0 when a row is longer than the gather's LDS holds.
    bool transposed = false;           // maps hold the transposed T10 convention
    int64_t N = 0;
    int nps = 3;
    int row_cap = 0;                   // longest block row, rounded up to 8
    const int64_t* rowptr = nullptr;   // [N + 1]
    const int64_t* inc_ptr = nullptr;  // [N + 1] solid incidences (e * npe + a)
    const int32_t* inc = nullptr;
    const int64_t* sinc_ptr = nullptr; // [N + 1] surface incidences (s * nps + a), or nullptr
    const int32_t* sinc = nullptr;
    const uint8_t* rpos = nullptr;     // [n_inc][16] row position of the block of local node b
    const uint8_t* spos = nullptr;     // [n_sinc][8]
    const uint8_t* cmask = nullptr;    // [nnzb] constraint mask of each slot's column node
};
constexpr int kRowGatherMaxCap = 224;   // 4 waves x 224 blocks x 72 B of LDS
// BSR values from element/surface contributions + cm*M + identity on constrained DOFs.
// flip = 1 assembles every solid element block as K_e^T instead (the reference's T10
// finite-difference layout when the maps hold the untransposed convention, and back).
void launch_gather_matrix(const GatherMaps& g, const void* Ke, bool fp32, const double* Ks, const double* mval,
                          double cm, const uint8_t* bcmask, void* vals, bool vals_fp32, bool flip, hipStream_t s,
                          const SetStrides& b = SetStrides());   // strides: Ke, Ks, vals

// Rayleigh damping C = beta_R K + alpha M in place over K's BSR values (constrained rows
// and columns zeroed); vals += a C
void launch_rayleigh_finish(int64_t nnzb, const int32_t* blkrow, const int32_t* colind, const uint8_t* bcmask,
                            const double* mval, double beta_r, double alpha, double* vals, hipStream_t s);
void launch_add_scaled_values(int64_t nnzb, double a, const double* C, void* vals, bool vals_fp32, hipStream_t s);

// ---- launchers (linalg_kernels.hip) ----
// SpMV layouts: kSpmvBlock one 3x3 block per lane step (4, 8 or 16 lanes per block
// row); kSpmvChunk 144-byte value chunks (needs 16-byte aligned values and columns
// padded by kSpmvPadBlocks blocks; 4, 8 or 16 lanes); kSpmvFlat the row's values as
// one coalesced word stream (16, 32 or 64 lanes)
constexpr int kSpmvPadBlocks = 4;
constexpr int kSpmvBlock = 0, kSpmvChunk = 1, kSpmvFlat = 2;
void launch_spmv(int64_t N, const int64_t* rowptr, const int32_t* colind, const void* vals, bool fp32,
                 const double* x, double* y, hipStream_t s, int lanes = 8, int layout = kSpmvBlock);
void launch_spmv_scalar_mass(int64_t N, const int64_t* rowptr, const int32_t* colind, const double* mval,
                             const uint8_t* bcmask, const double* x, double* y, hipStream_t s);
void launch_block_jacobi_setup(int64_t N, const int64_t* rowptr, const int32_t* colind, const void* vals,
                               bool fp32, double* dinv, hipStream_t s,
                               const SetStrides& b = SetStrides());   // strides: vals, dinv
void launch_block_jacobi_apply(int64_t N, const double* dinv, const double* r, double* z, hipStream_t s);
// partial dot products of k vectors V[j] (stride ld) with w: out[block][j]; returns #blocks
int dot_blocks(int64_t n);
void launch_multidot(int64_t n, int k, const double* V, int64_t ld, const double* w, double* partial, hipStream_t s);
void launch_reduce_partials(int nblk, int k, const double* partial, double* out, hipStream_t s);
void launch_multiaxpy(int64_t n, int k, const double* V, int64_t ld, const double* h, double* w, hipStream_t s);
// fused BiCGStab updates: p = r + beta (p - omega v) (p = r when fresh) and ph = D^-1 p;
// s = r - alpha v and sh = D^-1 s; x += alpha ph + omega sh and r = s - omega t;
// partial[b][5] = (rh,t) (s,t) (t,t) (rh,s) (s,s) per dot chunk (dot_blocks(n) chunks)
void launch_bicg_p_bj(int64_t N, bool fresh, double beta, double omega, const double* r, const double* v,
                      const double* dinv, double* p, double* ph, hipStream_t s);
void launch_bicg_s_bj(int64_t N, double alpha, const double* r, const double* v, const double* dinv, double* sv,
                      double* sh, hipStream_t s);
void launch_bicg_xr(int64_t n, double alpha, double omega, const double* ph, const double* sh, const double* sv,
                    const double* t, double* x, double* r, hipStream_t s);
void launch_bicg_dots5(int64_t n, const double* rh, const double* sv, const double* t, double* partial,
                       hipStream_t s);
void launch_axpby(int64_t n, double a, const double* x, double b, double* y, hipStream_t s);      // y = a x + b y
void launch_waxpby(int64_t n, double a, const double* x, double b, const double* y, double* w, hipStream_t s);
void launch_scale_copy(int64_t n, double a, const double* x, double* y, hipStream_t s);
void launch_fill(int64_t n, double v, double* y, hipStream_t s);
void launch_fill_int(int64_t n, int v, int* y, hipStream_t s);
void launch_zero_bc(int64_t N, const uint8_t* bcmask, double* v, hipStream_t s);
void launch_newmark_predict(int64_t n, double dt, double beta, double gamma, const double* v, const double* a,
                            double* dpred, double* vpred, hipStream_t s);
void launch_newmark_correct(int64_t n, double dt, double beta, double gamma, const double* d, const double* dpred,
                            const double* vpred, double* v, double* a, hipStream_t s);

// error helper
void set_error(const std::string& msg);

}  // namespace cmg

#define CMG_HIP_CHECK(expr)                                                                 \
    do {                                                                                    \
        hipError_t err__ = (expr);                                                          \
        if (err__ != hipSuccess) {                                                          \
            cmg::set_error(std::string(#expr) + ": " + hipGetErrorString(err__));           \
            return CMG_ERR_HIP;                                                             \
        }                                                                                   \
    } while (0)
