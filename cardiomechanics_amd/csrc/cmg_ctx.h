// Private layout of a mechanics context, shared by the translation units of the
// library (cmg_api.hip, batch.hip).  Not part of the ABI.
#pragma once

#include <string>
#include <vector>

#include "cmg_internal.h"

namespace cmg {

// owning device allocation
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    bool alloc(size_t b) {
        release();
        if (b == 0) return true;
        if (hipMalloc(&p, b) != hipSuccess) {
            p = nullptr;
            return false;
        }
        bytes = b;
        return true;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

int fail(const std::string& m, int code = CMG_ERR_INVALID);
int to_dev_material(const cmg_material& m, DevMaterial& d);   // validates the law
int alloc(DevBuf& b, size_t bytes);
double now_s();

template <typename T>
int upload(DevBuf& b, const T* src, size_t count, hipStream_t s) {
    if (!b.alloc(count * sizeof(T))) return fail("hipMalloc failed", CMG_ERR_NOMEM);
    if (count) CMG_HIP_CHECK(hipMemcpyAsync(b.p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    return CMG_SUCCESS;
}

}  // namespace cmg

#define CMG_TRY(expr)                         \
    do {                                      \
        int rc__ = (expr);                    \
        if (rc__ != CMG_SUCCESS) return rc__; \
    } while (0)

struct cmg_ctx {
    cmg_options opt;
    int device = 0;
    hipStream_t stream = nullptr;
    int npe = 10, nqp = 4, nps = 0;
    int64_t N = 0, E = 0, S = 0, n = 0;   // nodes, elements, surfaces, dofs
    int n_groups = 0;
    int n_mats = 0;
    int spmv_lanes = 0;         // lanes per block row of the SpMV, 0 until timed on the first solve
    int spmv_layout = 0;        // kSpmvBlock / kSpmvChunk / kSpmvFlat
    // the fastest block or chunk layout: the device-scalar BiCGStab fuses its SpMV with
    // dot products and exists in those two layouts only
    int bicg_lanes = 8;
    bool bicg_chunked = false;
    // device-scalar BiCGStab: workspace and the graph of `bicg_chunk` iterations
    cmg::DevBuf bicg_part, bicg_sc, bicg_flags;
    hipGraphExec_t bicg_graph = nullptr;
    const void* bicg_graph_x = nullptr;
    const void* bicg_graph_vals = nullptr;
    int bicg_chunk = 16;
    bool bicg_use_graph = true;
    double spmv_ms[9] = {};     // timed SpMV: block at 4/8/16 lanes, chunks at 4/8/16, flat at 16/32/64
    bool fp32 = false;
    bool trace = false;         // CMG_NEWTON_TRACE: one stderr line per Newton iteration
    // host copies needed for maps / exports
    std::vector<uint8_t> h_bc;
    std::vector<int64_t> h_rowptr;
    std::vector<int32_t> h_colind;
    std::vector<double> h_mval;
    std::vector<double> h_pressure;
    std::vector<int32_t> h_sgroup;   // surface groups, read back on the first volume query
    int64_t nnzb = 0;
    int64_t n_solid_blocks = 0;
    // element data (SoA)
    cmg::DevBuf conn, dNdX, bases, w, matidx, tact, mats, bcmask;
    cmg::DevBuf fe, Ke;
    // surfaces
    cmg::DevBuf sconn, sgroup, pressure, fs, Ks, vol_partial;
    // pattern and maps
    cmg::DevBuf rowptr, colind, blkrow, cptr, contrib, heavy, mval;
    int64_t n_heavy = 0;
    bool maps_transposed = false;   // T10 blocks enter the assembled tangent as K_e^T
    cmg::DevBuf inc_ptr, inc, sinc_ptr, sinc;
    cmg::DevBuf rpos, spos, cmask;  // row positions and column masks for the row gather (GatherMaps)
    int row_cap = 0;
    // state and work vectors
    cmg::DevBuf Xref, nodes, xeval, d, dpred, vpred, vel, acc, r, delta, dsave, tmpA, tmpB;
    cmg::DevBuf vals, dinv;
    // Rayleigh damping matrix C (FP64 BSR values, same pattern as vals), valid for the
    // configuration in nodes until nodes change
    cmg::DevBuf cvals;
    bool damping = false, damping_valid = false;
    // corrector energies of the last committed Newmark step
    double kinetic_energy = 0.0, damping_energy = 0.0;
    cmg::DevBuf gm_V, gm_partial, gm_h, gm_y;
    int gm_m = 0;
    double* h_pinned = nullptr;   // host staging for reductions
    double* h_part = nullptr;     // host staging for per-chunk dot partials (BiCGStab), 5 per chunk
    int64_t h_part_len = 0;
    cmg::DevBuf corrupt;
    // static solver bookkeeping (CBSolverEquilibrium::UpdateInitialGuess)
    double time_last = 0.0, step_last = 0.0;
    cmg::DevBuf guess;               // normalised previous increment
    bool has_override = false;  // cmg_set_initial_guess / cmg_mlp_propose_guess
    bool override_checked = false;   // compare with the predictor before use (Newmark)
    cmg::DevBuf inc1, inc2;     // increments of the last two Newmark steps
    // converged Newmark increment not yet committed (cmg_newmark_solve), with its time and step
    bool trial_pending = false;
    double trial_time = 0.0, trial_dt = 0.0;
    cmg::DevBuf override_guess;
    // Land17 tension model: parameter blocks, committed state [8][E], calcium tables and the
    // time of the last commit (the state's time)
    bool l17 = false;
    cmg::DevBuf l17_params, l17_state, l17_cai;
    std::vector<double> h_l17_params;
    std::vector<int32_t> h_matidx;
    double l17_t_commit = 0.0;
    // vertex-patch additive Schwarz preconditioner (schwarz.hip): patch node lists, slot
    // maps into the block pattern, inverse offsets, node -> patch-row positions, the
    // inverses and the per-patch results
    bool schwarz = false;
    int64_t sw_P = 0;
    int sw_nd_max = 0;
    cmg::DevBuf sw_pptr, sw_pnodes, sw_sptr, sw_slots, sw_iptr, sw_cptr, sw_cpos, sw_inv, sw_y;
    // patches by size class for the setup launches: sw_order[0, sw_n_small) have at most
    // sw_nd_small DOFs, the rest up to sw_nd_max
    cmg::DevBuf sw_order;
    bool precond_built = false;   // the preconditioner holds a factorisation of this Newton solve
    int64_t sw_n_small = 0;
    int sw_nd_small = 0;
    int64_t sw_inv_len = 0, sw_y_len = 0;   // inverse values and patch-vector rows of all patches
};

namespace cmg {
// kernel argument packs of a context for coordinates x_dev (device pointer)
ElemArgs elem_args(cmg_ctx* c, const double* x_dev, double time);
SurfArgs surf_args(cmg_ctx* c, const double* x_dev, bool with_volume);
GatherMaps gather_maps(cmg_ctx* c);
// vertex-patch additive Schwarz (schwarz.hip); patches of more than kMaxPatchDofs DOFs
// leave the context with block Jacobi
constexpr int kMaxPatchDofs = 72;
int build_vertex_patches(cmg_ctx* c, const int32_t* elements);   // sets c->schwarz
// S parameter sets over the context's patches (blockIdx.y = set): per-set strides of the
// block values, inverses and vectors; sets with flags[set] != 0 (setup: == 0) are skipped
struct PatchSets {
    int S = 1;
    int64_t set_vals = 0, set_inv = 0, set_vec = 0, set_y = 0;
    void* inv = nullptr;        // [S][sw_inv_len] (the context's sw_inv when null)
    double* y = nullptr;        // [S][sw_y_len] (the context's sw_y when null)
    const int32_t* flags = nullptr;
};
void launch_patch_setup(cmg_ctx* c, const void* vals, const PatchSets& ps = PatchSets());  // patch inverses
// z = sum_p R_p^T A_p^-1 R_p r
void launch_patch_apply(cmg_ctx* c, const double* r, double* z, const PatchSets& ps = PatchSets());
}  // namespace cmg
