// Parameter-set batch (include/cmg.h, cmg_batch_*): S quasi-static problems that
// share one mesh and differ in their material tables.  Each set follows the static
// step of one context (CBSolverEquilibrium::UpdateInitialGuess / SolverStep,
// CBSolverEquilibrium.cpp:234-363, with the Newton loop of cmg_static_step), but
// every kernel launch covers all sets at once: element kernels and assembly with
// blockIdx.y = set, and a block-Jacobi conjugate gradient whose scalar recurrences
// live on the device, so a run of `cg_chunk` iterations is one HIP graph replay
// followed by one small copy of the convergence flags.  The conjugate gradient is
// preconditioned by the context's vertex-patch additive Schwarz (one set of patch
// inverses per parameter set; symmetric like the tangents), or by 3x3 block Jacobi
// where the context has no patches.  An autotuner times the element-kernel variants
// and the SpMV lane count on the first tangent.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "batch_kernels.h"
#include "cmg_ctx.h"

using namespace cmg;

struct cmg_batch {
    cmg_ctx* c = nullptr;
    int device = 0;
    int S = 0;
    int64_t n = 0;
    DevBuf mats;                                   // [S][n_mats]
    DevBuf nodes, d, r, rhs, delta, xe, guess;     // [S][n]
    DevBuf fe, Ke, fs, Ks, vol;                    // per set element / surface outputs
    bool ptan = false;                             // surface-pressure tangent in the batched matrices
    DevBuf vals, dinv;                             // [S][nnzb*9], [S][N*9]
    DevBuf sw_inv, sw_y;                           // Schwarz: [S][sw_inv_len], [S][sw_y_len]
    bool schwarz = false;
    const int32_t* apply_skip = nullptr;           // sets the Schwarz application skips
    DevBuf cg_r, cg_z, cg_p, cg_q;                 // [S][n]
    DevBuf sc;                                     // CG scalars: 6 doubles per set
    DevBuf flags;                                  // done[S], it[S]
    DevBuf part, active, corrupt, rtol, dout;
    int64_t part_ld = 0;
    int32_t* h_int = nullptr;   // pinned: [0,S) flags to upload, [S,3S) flags read back
    double* h_dbl = nullptr;    // pinned: [0,S) upload, [S,3S) read back
    BatchCG cg;
    // step bookkeeping shared by the sets (they advance through the same times)
    double time_last = 0.0, step_last = 0.0;
    // tuning
    cmg_batch_info info;
    bool variant_fixed = false, lanes_fixed = false;
    hipGraphExec_t graph = nullptr;
    ~cmg_batch() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (h_int) (void)hipHostFree(h_int);
        if (h_dbl) (void)hipHostFree(h_dbl);
    }
};

namespace {

constexpr double kMinDomainLambda = 1.0 / (1 << 20);   // backtracking floor on inverted elements

int upload_flags(cmg_batch* b, const std::vector<int32_t>& f) {
    std::memcpy(b->h_int, f.data(), b->S * sizeof(int32_t));
    CMG_HIP_CHECK(hipMemcpyAsync(b->active.p, b->h_int, b->S * sizeof(int32_t), hipMemcpyHostToDevice, b->c->stream));
    return CMG_SUCCESS;
}

SetCoefs coefs(const std::vector<double>& v) {
    SetCoefs a;
    for (size_t i = 0; i < v.size(); i++) a.c[i] = v[i];
    return a;
}

bool any(const std::vector<int32_t>& f) {
    for (int32_t v : f)
        if (v) return true;
    return false;
}

// element (+ surface) kernels of the sets flagged in `act` at xe; corrupt[s] = first
// element with det F <= 0 or -1.  The flags must already be on the device.
int beval(cmg_batch* b, double time, bool tangent, std::vector<int32_t>& corrupt) {
    cmg_ctx* c = b->c;
    hipStream_t s = c->stream;
    const int S = b->S;
    launch_fill_int(S, INT_MAX, b->corrupt.as<int>(), s);
    ElemArgs a = elem_args(c, b->xe.as<double>(), time);
    a.n_sets = S;
    a.set_x = b->n;
    a.fe = b->fe.as<double>();
    a.set_fe = c->E * 3 * c->npe;
    a.Ke = b->Ke.p;
    a.set_Ke_bytes = c->E * 9 * c->npe * c->npe * (int64_t)(c->fp32 ? sizeof(float) : sizeof(double));
    a.mats = b->mats.as<DevMaterial>();
    a.set_mats = c->n_mats;
    a.corrupt = b->corrupt.as<int>();
    a.active = b->active.as<int32_t>();
    if (tangent) {
        if (!b->info.tuned && !b->variant_fixed) {
            // autotuner: one warm and one timed launch of each variant on this state
            hipEvent_t e0, e1;
            CMG_HIP_CHECK(hipEventCreate(&e0));
            CMG_HIP_CHECK(hipEventCreate(&e1));
            double best = 1e300;
            for (int v = 1; v <= 8; v++) {
                launch_element_residual_tangent(c->npe, c->fp32, a, s, v);
                CMG_HIP_CHECK(hipEventRecord(e0, s));
                launch_element_residual_tangent(c->npe, c->fp32, a, s, v);
                CMG_HIP_CHECK(hipEventRecord(e1, s));
                CMG_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.f;
                CMG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                b->info.element_ms[v - 1] = ms;
                if (ms < best) {
                    best = ms;
                    b->info.element_variant = v;
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            b->variant_fixed = true;
            launch_fill_int(S, INT_MAX, b->corrupt.as<int>(), s);
        }
        launch_element_residual_tangent(c->npe, c->fp32, a, s, b->info.element_variant);
    } else {
        launch_element_residual(c->npe, a, s);
    }
    if (c->S > 0) {
        SurfArgs sa = surf_args(c, b->xe.as<double>(), false);
        sa.n_sets = S;
        sa.set_x = b->n;
        sa.fs = b->fs.as<double>();
        sa.set_fs = c->S * 3 * c->nps;
        const bool st = tangent && b->ptan;
        sa.Ks = st ? b->Ks.as<double>() : nullptr;
        sa.set_Ks = c->S * 9 * c->nps * c->nps;
        sa.active = b->active.as<int32_t>();
        launch_surface(sa, st, s);
    }
    CMG_HIP_CHECK(hipGetLastError());
    CMG_HIP_CHECK(hipMemcpyAsync(b->h_int + S, b->corrupt.p, S * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    corrupt.assign(S, -1);
    for (int i = 0; i < S; i++)
        if (b->h_int[S + i] != INT_MAX) corrupt[i] = b->h_int[S + i];
    return CMG_SUCCESS;
}

void bgather_residual(cmg_batch* b) {
    cmg_ctx* c = b->c;
    const bool sp = c->S > 0;
    SetStrides st;
    st.n_sets = b->S;
    st.a = c->E * 3 * c->npe;
    st.b = c->S * 3 * c->nps;
    st.c = b->n;
    st.active = b->active.as<int32_t>();
    launch_gather_vector(c->N, c->inc_ptr.as<int64_t>(), c->inc.as<int32_t>(), b->fe.as<double>(), 3 * c->npe,
                         sp ? c->sinc_ptr.as<int64_t>() : nullptr, sp ? c->sinc.as<int32_t>() : nullptr,
                         sp ? b->fs.as<double>() : nullptr, 3 * c->nps, c->bcmask.as<uint8_t>(), b->r.as<double>(),
                         c->stream, st);
}

void bgather_tangent(cmg_batch* b) {
    cmg_ctx* c = b->c;
    hipStream_t s = c->stream;
    // without a pressure tangent the surface blocks of the pattern read the context's
    // Ks, cleared, for every set (stride 0)
    if (c->S > 0 && !b->ptan) launch_fill((int64_t)c->S * 9 * c->nps * c->nps, 0.0, c->Ks.as<double>(), s);
    SetStrides st;
    st.n_sets = b->S;
    st.a = c->E * 9 * c->npe * c->npe;
    st.b = b->ptan ? c->S * 9 * c->nps * c->nps : 0;
    st.c = c->nnzb * 9;
    st.active = b->active.as<int32_t>();
    launch_gather_matrix(gather_maps(c), b->Ke.p, c->fp32, b->ptan ? b->Ks.as<double>() : c->Ks.as<double>(), nullptr,
                         0.0, c->bcmask.as<uint8_t>(), b->vals.p, c->fp32, false, s, st);
    if (b->schwarz) {
        PatchSets ps;
        ps.S = b->S;
        ps.set_vals = c->nnzb * 9;
        ps.set_inv = c->sw_inv_len;
        ps.inv = b->sw_inv.p;
        ps.flags = b->active.as<int32_t>();
        launch_patch_setup(c, b->vals.p, ps);
        return;
    }
    SetStrides sj;
    sj.n_sets = b->S;
    sj.a = c->nnzb * 9;
    sj.b = c->N * 9;
    sj.active = b->active.as<int32_t>();
    launch_block_jacobi_setup(c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), b->vals.p, c->fp32,
                              b->dinv.as<double>(), s, sj);
}

// z = P^-1 u of every set not in b->apply_skip (the Schwarz application of the CG)
void batch_schwarz_apply(void* bb, const double* u, double* z) {
    cmg_batch* b = static_cast<cmg_batch*>(bb);
    PatchSets ps;
    ps.S = b->S;
    ps.set_inv = b->c->sw_inv_len;
    ps.set_vec = b->n;
    ps.set_y = b->c->sw_y_len;
    ps.inv = b->sw_inv.p;
    ps.y = b->sw_y.as<double>();
    ps.flags = b->apply_skip;
    launch_patch_apply(b->c, u, z, ps);
}

// out[k*S + s] = x_k,s . y_k,s for the given pairs; one synchronisation
int bdots(cmg_batch* b, const std::vector<std::pair<const double*, const double*>>& pairs, std::vector<double>& out) {
    cmg_ctx* c = b->c;
    const int S = b->S;
    for (size_t k = 0; k < pairs.size(); k++)
        launch_bdot(S, b->n, pairs[k].first, pairs[k].second, b->part.as<double>(), b->part_ld,
                    b->dout.as<double>() + k * S, c->stream);
    CMG_HIP_CHECK(hipMemcpyAsync(b->h_dbl + S, b->dout.p, pairs.size() * S * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    out.assign(b->h_dbl + S, b->h_dbl + S + pairs.size() * S);
    return CMG_SUCCESS;
}

int tune_spmv(cmg_batch* b) {
    cmg_ctx* c = b->c;
    hipStream_t s = c->stream;
    hipEvent_t e0, e1;
    CMG_HIP_CHECK(hipEventCreate(&e0));
    CMG_HIP_CHECK(hipEventCreate(&e1));
    const int lanes[3] = {4, 8, 16};
    double best = 1e300;
    for (int k = 0; k < 3; k++) {
        launch_bcg_spmv_dot(b->cg, lanes[k], c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), b->vals.p,
                            c->fp32, c->nnzb * 9, s);
        CMG_HIP_CHECK(hipEventRecord(e0, s));
        launch_bcg_spmv_dot(b->cg, lanes[k], c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), b->vals.p,
                            c->fp32, c->nnzb * 9, s);
        CMG_HIP_CHECK(hipEventRecord(e1, s));
        CMG_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CMG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        b->info.spmv_ms[k] = ms;
        if (ms < best) {
            best = ms;
            b->info.spmv_lanes = lanes[k];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->lanes_fixed = true;
    return CMG_SUCCESS;
}

void cg_chunk_eager(cmg_batch* b) {
    cmg_ctx* c = b->c;
    for (int k = 0; k < b->info.cg_chunk; k++)
        launch_bcg_iteration(b->cg, b->info.spmv_lanes, c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(),
                             b->vals.p, c->fp32, c->nnzb * 9, b->dinv.as<double>(), c->stream,
                             b->schwarz ? batch_schwarz_apply : nullptr, b);
}

int build_graph(cmg_batch* b) {
    hipStream_t s = b->c->stream;
    hipGraph_t g = nullptr;
    CMG_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    cg_chunk_eager(b);
    CMG_HIP_CHECK(hipStreamEndCapture(s, &g));
    hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    CMG_HIP_CHECK(e);
    return CMG_SUCCESS;
}

// CG on the sets flagged `solve` (already uploaded): delta = A^-1 rhs to rtol[s]
int bcg_solve(cmg_batch* b, const std::vector<double>& rtol, std::vector<int32_t>& its) {
    cmg_ctx* c = b->c;
    hipStream_t s = c->stream;
    const int S = b->S;
    std::memcpy(b->h_dbl, rtol.data(), S * sizeof(double));
    CMG_HIP_CHECK(hipMemcpyAsync(b->rtol.p, b->h_dbl, S * sizeof(double), hipMemcpyHostToDevice, s));
    b->apply_skip = nullptr;   // init: every set (the inactive ones' results are not used)
    launch_bcg_init(b->cg, c->N, b->rhs.as<double>(), b->dinv.as<double>(), b->rtol.as<double>(),
                    b->active.as<int32_t>(), s, b->schwarz ? batch_schwarz_apply : nullptr, b);
    b->apply_skip = b->cg.done;   // iterations: the sets still running
    if (!b->lanes_fixed) CMG_TRY(tune_spmv(b));
    if (b->info.use_graph && !b->graph) CMG_TRY(build_graph(b));
    b->info.tuned = 1;
    int total = 0;
    for (;;) {
        if (b->info.use_graph)
            CMG_HIP_CHECK(hipGraphLaunch(b->graph, s));
        else
            cg_chunk_eager(b);
        total += b->info.cg_chunk;
        CMG_HIP_CHECK(hipMemcpyAsync(b->h_int + S, b->cg.done, 2 * S * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        CMG_HIP_CHECK(hipStreamSynchronize(s));
        bool all = true;
        for (int i = 0; i < S; i++) all = all && b->h_int[S + i] != 0;
        if (all || total >= c->opt.krylov_maxit) break;
    }
    CMG_HIP_CHECK(hipGetLastError());
    its.assign(b->h_int + 2 * S, b->h_int + 3 * S);
    return CMG_SUCCESS;
}

// Newton of every set; the per-set logic is that of the single-context Newton
// (abstol / rtol / stol tests, Eisenstat-Walker forcing, backtracking on corrupt
// elements or on no decrease).
int bnewton(cmg_batch* b, double time, cmg_step_stats* st) {
    cmg_ctx* c = b->c;
    hipStream_t s = c->stream;
    const int S = b->S;
    const int64_t n = b->n;
    const cmg_options& o = c->opt;
    std::vector<int32_t> run(S, 1), its(S, 0), kry(S, 0), reason(S, 0), zero_tried(S, 0), stall(S, 0), corrupt;
    std::vector<double> lam(S, 1.0), rprev(S, 0.0), r0(S, 0.0), rn(S, 0.0), sn(S, 0.0), dn(S, 0.0), eta(S, 1e-1);
    std::vector<double> eta_k(S, 1e-1);
    const double t0 = now_s();
    double t_tan = 0.0, t_kry = 0.0;
    std::vector<double> dots;
    for (int guard = 0; guard < 100000 && any(run); guard++) {
        const double ta = now_s();
        launch_waxpby(S * n, 1.0, b->nodes.as<double>(), 1.0, b->d.as<double>(), b->xe.as<double>(), s);
        CMG_TRY(upload_flags(b, run));
        CMG_TRY(beval(b, time, true, corrupt));
        std::vector<double> back(S, 0.0), keep(S, 1.0);
        std::vector<int32_t> eval(S, 0);
        bool any_back = false, any_zero = false;
        for (int i = 0; i < S; i++) {
            if (!run[i]) continue;
            if (corrupt[i] >= 0) {
                st[i].corrupt_element = corrupt[i];
                if (its[i] > 0 && lam[i] > kMinDomainLambda) {
                    lam[i] *= 0.5;
                    back[i] = -lam[i];
                    any_back = true;
                } else if (its[i] == 0 && !zero_tried[i]) {
                    // extrapolated initial guess inverts an element: restart from a zero increment
                    zero_tried[i] = 1;
                    st[i].guess_source = 4;
                    keep[i] = 0.0;
                    any_zero = true;
                } else {
                    reason[i] = -6;
                    run[i] = 0;
                }
            } else {
                eval[i] = 1;
            }
        }
        if (any_zero) launch_bscale(S, n, coefs(keep), b->d.as<double>(), s);
        if (any(eval)) {
            CMG_TRY(upload_flags(b, eval));
            bgather_residual(b);
            CMG_TRY(bdots(b, {{b->r.as<double>(), b->r.as<double>()}}, dots));
        }
        t_tan += now_s() - ta;
        std::vector<int32_t> solve(S, 0);
        for (int i = 0; i < S; i++) {
            if (!eval[i]) continue;
            rn[i] = std::sqrt(std::max(dots[i], 0.0));
            if (!std::isfinite(dots[i])) {
                if (its[i] > 0 && lam[i] > kMinDomainLambda) {   // overflow in the laws: a domain error
                    lam[i] *= 0.5;
                    back[i] = -lam[i];
                    any_back = true;
                    continue;
                }
                reason[i] = -4;
                run[i] = 0;
                continue;
            }
            if (its[i] == 0) {
                r0[i] = rn[i];
                st[i].initial_residual_norm = rn[i];
            } else if (rn[i] > rprev[i] && lam[i] > 1.0 / 64 && rn[i] > o.newton_atol) {
                lam[i] *= 0.5;
                back[i] = -lam[i];
                any_back = true;
                continue;
            }
            // the single-set Newton's tests (cmg_api.hip newton()): divergence tolerance,
            // step-norm convergence only after a residual decrease
            if (rn[i] < o.newton_atol) reason[i] = 2;
            else if (its[i] > 0 && rn[i] > 1e4 * r0[i]) reason[i] = -9;
            else if (its[i] > 0 && rn[i] < o.newton_rtol * r0[i]) reason[i] = 3;
            else if (its[i] > 0 && rn[i] <= rprev[i] && sn[i] * lam[i] < o.newton_stol * dn[i]) reason[i] = 4;
            else if (its[i] >= o.newton_maxit) reason[i] = -5;
            if (reason[i] == 0 && its[i] > 0) {
                stall[i] = (rn[i] > 0.9 * rprev[i]) ? stall[i] + 1 : 0;   // as the single-set Newton
                if (o.newton_stall > 0 && stall[i] >= o.newton_stall) reason[i] = -7;
            }
            if (reason[i] != 0) {
                run[i] = 0;
                continue;
            }
            if (its[i] > 0)
                eta[i] = std::min(0.5, std::max(o.krylov_rtol, 0.9 * (rn[i] / rprev[i]) * (rn[i] / rprev[i])));
            // Kelley's safeguard, as in the single-set Newton
            eta_k[i] = std::min(0.5, std::max(eta[i], 0.5 * (o.newton_atol + o.newton_rtol * r0[i]) / rn[i]));
            rprev[i] = rn[i];
            solve[i] = 1;
        }
        if (any_back) launch_baxpy(S, n, coefs(back), b->delta.as<double>(), b->d.as<double>(), s);
        if (!any(solve)) continue;
        const double tb = now_s();
        CMG_TRY(upload_flags(b, solve));
        bgather_tangent(b);
        launch_scale_copy(S * n, -1.0, b->r.as<double>(), b->rhs.as<double>(), s);
        t_tan += now_s() - tb;
        const double tk = now_s();
        std::vector<int32_t> cits;
        CMG_TRY(bcg_solve(b, eta_k, cits));
        t_kry += now_s() - tk;
        std::vector<double> one(S, 0.0);
        for (int i = 0; i < S; i++)
            if (solve[i]) {
                one[i] = 1.0;
                kry[i] += cits[i];
                its[i]++;
                lam[i] = 1.0;
            }
        launch_baxpy(S, n, coefs(one), b->delta.as<double>(), b->d.as<double>(), s);
        CMG_TRY(bdots(b, {{b->delta.as<double>(), b->delta.as<double>()}, {b->d.as<double>(), b->d.as<double>()}},
                      dots));
        for (int i = 0; i < S; i++)
            if (solve[i]) {
                sn[i] = std::isfinite(dots[i]) ? std::sqrt(std::max(dots[i], 0.0)) : INFINITY;
                dn[i] = std::sqrt(std::max(dots[S + i], 0.0));
                if (!(sn[i] > 0.0) || !std::isfinite(sn[i])) {   // no usable update from the Krylov solve
                    reason[i] = -3;
                    run[i] = 0;
                }
            }
    }
    const double wall = now_s() - t0;
    bool ok = true;
    for (int i = 0; i < S; i++) {
        st[i].newton_iterations = its[i];
        st[i].krylov_iterations = kry[i];
        st[i].converged_reason = reason[i];
        st[i].residual_norm = rn[i];
        st[i].step_norm = sn[i];
        st[i].wall_seconds = wall;
        st[i].tangent_seconds = t_tan;
        st[i].krylov_seconds = t_kry;
        ok = ok && reason[i] > 0;
    }
    return ok ? CMG_SUCCESS : CMG_FAILED;
}

}  // namespace

extern "C" {

int cmg_batch_create(cmg_ctx* c, int32_t n_sets, const cmg_material* materials, cmg_batch** out) {
    if (!c || !materials || !out) return fail("cmg_batch_create: null argument");
    if (n_sets < 1 || n_sets > kMaxSets)
        return fail("cmg_batch_create: n_sets must be in [1, " + std::to_string(kMaxSets) + "]");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    std::unique_ptr<cmg_batch> b(new cmg_batch());
    b->c = c;
    b->device = c->device;
    b->S = n_sets;
    b->n = c->n;
    const int S = n_sets;
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    std::vector<DevMaterial> dm((size_t)S * c->n_mats);
    for (size_t i = 0; i < dm.size(); i++) {
        // the batched solve is a conjugate gradient: it needs symmetric tangents, which the
        // Holzapfel law as coded does not give (its isochoric factor is 1 and its switch
        // enters the stress without its derivative, CBConstitutiveModelHolzapfel.cpp:89,121,148-154)
        if (materials[i].tension_fn == CMG_FN_LAND17)
            return fail("cmg_batch_create: the Land17 tension model (per-element history) is not supported in "
                        "the parameter batch");
        if (materials[i].law != CMG_LAW_GUCCIONE)
            return fail("cmg_batch_create: the batched conjugate gradient needs symmetric tangents; only the "
                        "Guccione law (symmetric with FromFunction, TanH or no active tension) is accepted");
        CMG_TRY(to_dev_material(materials[i], dm[i]));
    }
    CMG_TRY(upload(b->mats, dm.data(), dm.size(), s));
    const size_t vb = (size_t)S * n * sizeof(double);
    for (DevBuf* v : {&b->nodes, &b->d, &b->r, &b->rhs, &b->delta, &b->xe, &b->guess, &b->cg_r, &b->cg_z, &b->cg_p,
                      &b->cg_q})
        CMG_TRY(alloc(*v, vb));
    const size_t kt = c->fp32 ? sizeof(float) : sizeof(double);
    CMG_TRY(alloc(b->fe, (size_t)S * c->E * 3 * c->npe * sizeof(double)));
    CMG_TRY(alloc(b->Ke, (size_t)S * c->E * 9 * c->npe * c->npe * kt));
    if (c->S > 0) {
        CMG_TRY(alloc(b->fs, (size_t)S * c->S * 3 * c->nps * sizeof(double)));
        CMG_TRY(alloc(b->vol, (size_t)S * c->S * sizeof(double)));
        b->ptan = c->opt.pressure_tangent != 0;
        if (b->ptan) CMG_TRY(alloc(b->Ks, (size_t)S * c->S * 9 * c->nps * c->nps * sizeof(double)));
    }
    CMG_TRY(alloc(b->vals, (size_t)S * c->nnzb * 9 * kt));
    CMG_TRY(alloc(b->dinv, (size_t)S * c->N * 9 * sizeof(double)));
    if (c->schwarz) {
        // S sets of patch inverses (S x 548 MB in FP64 on the 200k LV): where they do not
        // fit, the batch keeps block Jacobi
        b->schwarz = b->sw_inv.alloc((size_t)S * c->sw_inv_len * kt) &&
                     b->sw_y.alloc((size_t)S * c->sw_y_len * sizeof(double));
        if (!b->schwarz) {
            b->sw_inv.release();
            b->sw_y.release();
            (void)hipGetLastError();   // the failed allocation is not an error of the batch
        }
    }
    CMG_TRY(alloc(b->sc, (size_t)S * 6 * sizeof(double)));
    CMG_TRY(alloc(b->flags, (size_t)S * 2 * sizeof(int32_t)));
    b->part_ld = std::max<int64_t>({(int64_t)batch_spmv_blocks(c->N, 16), 2 * (int64_t)batch_node_blocks(c->N),
                                    (int64_t)batch_dot_blocks(n)}) + 1;
    CMG_TRY(alloc(b->part, (size_t)S * b->part_ld * sizeof(double)));
    CMG_TRY(alloc(b->active, S * sizeof(int32_t)));
    CMG_TRY(alloc(b->corrupt, S * sizeof(int32_t)));
    CMG_TRY(alloc(b->rtol, S * sizeof(double)));
    CMG_TRY(alloc(b->dout, 2 * S * sizeof(double)));
    CMG_HIP_CHECK(hipHostMalloc((void**)&b->h_int, 3 * S * sizeof(int32_t), hipHostMallocDefault));
    CMG_HIP_CHECK(hipHostMalloc((void**)&b->h_dbl, 3 * S * sizeof(double), hipHostMallocDefault));
    BatchCG& cg = b->cg;
    cg.S = S;
    cg.x = b->delta.as<double>();
    cg.r = b->cg_r.as<double>();
    cg.z = b->cg_z.as<double>();
    cg.p = b->cg_p.as<double>();
    cg.q = b->cg_q.as<double>();
    double* sc = b->sc.as<double>();
    cg.rz = sc;
    cg.rr = sc + S;
    cg.rr0 = sc + 2 * S;
    cg.alpha = sc + 3 * S;
    cg.beta = sc + 4 * S;
    cg.tol2 = sc + 5 * S;
    cg.done = b->flags.as<int32_t>();
    cg.it = b->flags.as<int32_t>() + S;
    cg.part = b->part.as<double>();
    cg.part_ld = b->part_ld;
    // every set starts from the context's current configuration
    for (int i = 0; i < S; i++)
        CMG_HIP_CHECK(hipMemcpyAsync(b->nodes.as<double>() + i * n, c->nodes.p, n * sizeof(double),
                                     hipMemcpyDeviceToDevice, s));
    launch_fill(S * n, 0.0, b->d.as<double>(), s);
    launch_fill(S * n, 0.0, b->guess.as<double>(), s);
    launch_fill(S * n, 0.0, b->delta.as<double>(), s);
    launch_fill(S * 6, 0.0, sc, s);
    launch_fill_int(2 * S, 1, cg.done, s);
    std::memset(&b->info, 0, sizeof(b->info));
    b->info.n_sets = S;
    b->info.preconditioner = b->schwarz ? 2 : 1;
    b->info.element_variant = c->opt.element_variant;
    b->info.spmv_lanes = 8;
    b->info.cg_chunk = 16;
    b->info.use_graph = 1;
    b->time_last = c->time_last;
    b->step_last = c->step_last;
    CMG_HIP_CHECK(hipGetLastError());
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    *out = b.release();
    return CMG_SUCCESS;
}

int cmg_batch_destroy(cmg_batch* b) {
    if (!b) return CMG_SUCCESS;
    (void)hipSetDevice(b->device);   // the context may already be gone: do not touch it
    (void)hipDeviceSynchronize();
    delete b;
    return CMG_SUCCESS;
}

int cmg_batch_configure(cmg_batch* b, int32_t use_graph, int32_t spmv_lanes, int32_t element_variant,
                        int32_t cg_chunk) {
    if (!b) return fail("null batch");
    if (spmv_lanes != 0 && spmv_lanes != 4 && spmv_lanes != 8 && spmv_lanes != 16)
        return fail("cmg_batch_configure: spmv_lanes must be 0, 4, 8 or 16");
    if (element_variant < 0 || element_variant > 8) return fail("cmg_batch_configure: element_variant in 0..8");
    if (cg_chunk < 1 || cg_chunk > 1024) return fail("cmg_batch_configure: cg_chunk in 1..1024");
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    b->info.use_graph = use_graph ? 1 : 0;
    b->info.cg_chunk = cg_chunk;
    b->lanes_fixed = spmv_lanes != 0;
    if (spmv_lanes) b->info.spmv_lanes = spmv_lanes;
    b->variant_fixed = element_variant != 0;
    if (element_variant) b->info.element_variant = element_variant;
    b->info.tuned = 0;
    return CMG_SUCCESS;
}

int cmg_batch_info_get(cmg_batch* b, cmg_batch_info* info) {
    if (!b || !info) return fail("cmg_batch_info_get: null argument");
    *info = b->info;
    return CMG_SUCCESS;
}

int cmg_batch_static_step(cmg_batch* b, double time, double dt, cmg_step_stats* stats) {
    if (!b) return fail("null batch");
    cmg_ctx* c = b->c;
    CMG_HIP_CHECK(hipSetDevice(c->device));
    const int S = b->S;
    const int64_t n = b->n;
    hipStream_t s = c->stream;
    std::vector<cmg_step_stats> local(S);
    cmg_step_stats* st = stats ? stats : local.data();
    std::memset(st, 0, S * sizeof(cmg_step_stats));
    for (int i = 0; i < S; i++) st[i].corrupt_element = -1;
    double* d = b->d.as<double>();
    // CBSolverEquilibrium::UpdateInitialGuess (:234-263), the same for every set
    const double t0 = now_s();
    if (time > b->time_last) {
        if (b->step_last > 0.0) {
            launch_scale_copy(S * n, 1. / b->step_last, d, b->guess.as<double>(), s);
            launch_scale_copy(S * n, 1. * dt / b->step_last, d, d, s);
        } else {
            launch_fill(S * n, 0.0, d, s);
        }
    } else if (time == b->time_last) {
        launch_fill(S * n, 0.0, d, s);
    } else {
        launch_scale_copy(S * n, dt, b->guess.as<double>(), d, s);
    }
    b->time_last = time;
    b->step_last = dt;
    for (int i = 0; i < S; i++) launch_zero_bc(c->N, c->bcmask.as<uint8_t>(), d + i * n, s);
    int rc = bnewton(b, time, st);
    // failed sets: increment reset to zero (:303-306); the others advance
    std::vector<double> keep(S, 1.0);
    for (int i = 0; i < S; i++)
        if (st[i].converged_reason <= 0) keep[i] = 0.0;
    launch_bscale(S, n, coefs(keep), d, s);
    launch_axpby(S * n, 1.0, d, 1.0, b->nodes.as<double>(), s);
    CMG_HIP_CHECK(hipGetLastError());
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    const double wall = now_s() - t0;
    for (int i = 0; i < S; i++) st[i].wall_seconds = wall;
    return rc;
}

int cmg_batch_get_nodes(cmg_batch* b, double* nodes) {
    if (!b || !nodes) return fail("cmg_batch_get_nodes: null argument");
    CMG_HIP_CHECK(hipMemcpyAsync(nodes, b->nodes.p, b->S * b->n * sizeof(double), hipMemcpyDeviceToHost,
                                 b->c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(b->c->stream));
    return CMG_SUCCESS;
}

int cmg_batch_set_nodes(cmg_batch* b, const double* nodes) {
    if (!b || !nodes) return fail("cmg_batch_set_nodes: null argument");
    CMG_HIP_CHECK(hipMemcpyAsync(b->nodes.p, nodes, b->S * b->n * sizeof(double), hipMemcpyHostToDevice,
                                 b->c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(b->c->stream));
    return CMG_SUCCESS;
}

int cmg_batch_residual(cmg_batch* b, const double* x, double time, double* f) {
    if (!b || !x || !f) return fail("cmg_batch_residual: null argument");
    CMG_HIP_CHECK(hipSetDevice(b->c->device));
    hipStream_t s = b->c->stream;
    const size_t vb = b->S * b->n * sizeof(double);
    CMG_HIP_CHECK(hipMemcpyAsync(b->xe.p, x, vb, hipMemcpyHostToDevice, s));
    CMG_TRY(upload_flags(b, std::vector<int32_t>(b->S, 1)));
    std::vector<int32_t> corrupt;
    CMG_TRY(beval(b, time, false, corrupt));
    for (int i = 0; i < b->S; i++)
        if (corrupt[i] >= 0) {
            set_error("set " + std::to_string(i) + ": element " + std::to_string(corrupt[i]) +
                      " is corrupt (det F <= 0)");
            return CMG_CORRUPT_ELEMENT;
        }
    bgather_residual(b);
    CMG_HIP_CHECK(hipMemcpyAsync(f, b->r.p, vb, hipMemcpyDeviceToHost, s));
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    return CMG_SUCCESS;
}

int cmg_batch_cavity_volumes(cmg_batch* b, double* volumes) {
    if (!b || !volumes) return fail("cmg_batch_cavity_volumes: null argument");
    cmg_ctx* c = b->c;
    CMG_HIP_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int S = b->S;
    std::vector<double> vp((size_t)S * c->S);
    if (c->S > 0) {
        CMG_TRY(upload_flags(b, std::vector<int32_t>(S, 1)));
        SurfArgs sa = surf_args(c, b->nodes.as<double>(), true);
        sa.n_sets = S;
        sa.set_x = b->n;
        sa.fs = b->fs.as<double>();
        sa.set_fs = c->S * 3 * c->nps;
        sa.Ks = nullptr;
        sa.vol_partial = b->vol.as<double>();
        sa.set_vol = c->S;
        sa.active = b->active.as<int32_t>();
        launch_surface(sa, false, s);
        CMG_HIP_CHECK(hipMemcpyAsync(vp.data(), b->vol.p, vp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    std::vector<int32_t> grp(c->S, 0);
    if (c->S > 0)
        CMG_HIP_CHECK(hipMemcpy(grp.data(), c->sgroup.p, c->S * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int i = 0; i < S; i++) {
        double* v = volumes + (int64_t)i * c->n_groups;
        for (int g = 0; g < c->n_groups; g++) v[g] = 0.0;
        for (int64_t k = 0; k < c->S; k++) v[grp[k]] += vp[(size_t)i * c->S + k];   // surface order
    }
    return CMG_SUCCESS;
}

}  // extern "C"
