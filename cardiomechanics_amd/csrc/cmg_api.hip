// Host side of the C ABI (include/cmg.h): context creation (mesh upload, block
// pattern, deterministic assembly maps), element-loop entry points, GMRES and the
// static / Newmark Newton drivers.  Runs one HIP stream per context.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "batch_kernels.h"
#include "cmg_ctx.h"
#include "land17_calcium.h"

namespace cmg {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace cmg

using namespace cmg;

namespace cmg {

int fail(const std::string& m, int code) {
    set_error(m);
    return code;
}

int alloc(DevBuf& b, size_t bytes) {
    if (!b.alloc(bytes)) return fail("hipMalloc of " + std::to_string(bytes) + " bytes failed", CMG_ERR_NOMEM);
    return CMG_SUCCESS;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int to_dev_material(const cmg_material& m, DevMaterial& d) {
    if (m.law != CMG_LAW_GUCCIONE && m.law != CMG_LAW_HOLZAPFEL) return fail("unknown material law");
    if (m.tension_fn < CMG_FN_NONE || m.tension_fn > CMG_FN_LAND17) return fail("unknown tension function");
    if (m.tension_fn == CMG_FN_TANH &&
        !(m.fn[0] > 0.0 && m.fn[4] >= 0.0 && m.fn[5] > 0.0 && m.fn[7] >= 0.0 && m.fn[8] > 0.0))
        return fail("TanH tension: cycleLength, ContractionTime and RelaxationTime must be positive, "
                    "Duration and LengthDependenceUpstrokeTime not negative");
    std::memset(&d, 0, sizeof(d));
    d.law = m.law;
    d.ignore_corrupt = m.ignore_corrupt;
    d.tension_fn = m.tension_fn;
    d.density = m.density;
    for (int k = 0; k < 12; k++) d.p[k] = m.params[k];
    d.tension_max = m.tension_max;
    for (int k = 0; k < 12; k++) d.fn[k] = m.fn[k];
    if (d.tension_fn == CMG_FN_DOUBLEHILL && !(d.fn[11] > 0.0)) {
        double mx = -INFINITY;   // CBAnalyticFunction.h:155-165
        for (int q = 0; q <= 1000; q++) {
            double t = q * d.fn[5] / 1000;
            double cc = std::pow(t / d.fn[6], d.fn[8]);
            cc = cc / (1.0 + cc);
            double rr = std::pow(t / d.fn[7], d.fn[9]);
            rr = 1.0 / (1.0 + rr);
            mx = std::max(mx, cc * rr);
        }
        d.fn[11] = mx;
    }
    return CMG_SUCCESS;
}

}  // namespace cmg

namespace {

// T4 shape derivatives, CBElementSolidT4::CalcShapeFunctionDerivatives (:73-121)
double t4_shape(const double* X, double* g) {
    double z43 = (X[11] - X[8]), z42 = (X[11] - X[5]), z41 = (X[11] - X[2]);
    double z32 = (X[8] - X[5]), z31 = (X[8] - X[2]), z21 = (X[5] - X[2]);
    double y43 = (X[10] - X[7]), y42 = (X[10] - X[4]), y41 = (X[10] - X[1]);
    double y32 = (X[7] - X[4]), y31 = (X[7] - X[1]), y21 = (X[4] - X[1]);
    double x41 = (X[9] - X[0]), x31 = (X[6] - X[0]), x21 = (X[3] - X[0]);
    double d = x21 * (y31 * z41 - y41 * z31) + y21 * (x41 * z31 - x31 * z41) + z21 * (x31 * y41 - x41 * y31);
    g[0] = 1.0 / d * (X[4] * z43 - X[7] * z42 + X[10] * z32);
    g[3] = 1.0 / d * (-X[1] * z43 + X[7] * z41 - X[10] * z31);
    g[6] = 1.0 / d * (X[1] * z42 - X[4] * z41 + X[10] * z21);
    g[9] = 1.0 / d * (-X[1] * z32 + X[4] * z31 - X[7] * z21);
    g[1] = 1.0 / d * (-X[3] * z43 + X[6] * z42 - X[9] * z32);
    g[4] = 1.0 / d * (X[0] * z43 - X[6] * z41 + X[9] * z31);
    g[7] = 1.0 / d * (-X[0] * z42 + X[3] * z41 - X[9] * z21);
    g[10] = 1.0 / d * (X[0] * z32 - X[3] * z31 + X[6] * z21);
    g[2] = 1.0 / d * (X[3] * y43 - X[6] * y42 + X[9] * y32);
    g[5] = 1.0 / d * (-X[0] * y43 + X[6] * y41 - X[9] * y31);
    g[8] = 1.0 / d * (X[0] * y42 - X[3] * y41 + X[9] * y21);
    g[11] = 1.0 / d * (-X[0] * y32 + X[3] * y31 - X[6] * y21);
    return d;
}

// T10 consistent mass pattern (CBElementSolidT10.cpp:371-380)
const double kR1[100] = {6,  1,  1,  1,  -4, -6, -4, -4, -6, -6, 1,  6,  1,  1,  -4, -4, -6, -6, -4, -6,
                         1,  1,  6,  1,  -6, -4, -4, -6, -6, -4, 1,  1,  1,  6,  -6, -6, -6, -4, -4, -4,
                         -4, -4, -6, -6, 32, 16, 16, 16, 16, 8,  -6, -4, -4, -6, 16, 32, 16, 8,  16, 16,
                         -4, -6, -4, -6, 16, 16, 32, 16, 8,  16, -4, -6, -6, -4, 16, 8,  16, 32, 16, 16,
                         -6, -4, -6, -4, 16, 16, 8,  16, 32, 16, -6, -6, -4, -4, 8,  16, 16, 16, 16, 32};

int64_t find_slot(const cmg_ctx* c, int64_t r, int32_t col) {
    const int32_t* b = c->h_colind.data() + c->h_rowptr[r];
    const int32_t* e = c->h_colind.data() + c->h_rowptr[r + 1];
    const int32_t* p = std::lower_bound(b, e, col);
    if (p == e || *p != col) return -1;
    return c->h_rowptr[r] + (p - b);
}

// Land17 defaults (InitParamsFromXml, CBTensionModelLand17.cpp:428-500): model constants,
// startTime, cycleLength, calcium transient (Coppini), rateDependancy (OFF), kff, and the
// initial state XS XW TRPN TmBlocked ZetaS ZetaW Cd
const double kLand17Defaults[kLand17Params] = {2.0,   0.1,   0.805, 1.0,    5.0,  0.35, 0.182, 0.012, 0.25, 0.5,
                                               0.0085, 0.615, 2.23, 25.0,  2.3,  -2.4, 120.0, 2.1,   7.0,  200.0,
                                               20.0,  1.0,   0.0,  1.0,    0.0,  0.0,  1.0,   0.0,   0.0,  0.0,
                                               1.0,   0.0,   0.0,  0.0,    0.0,  0.0,  0.0,   0.0,   0.0,  0.0};

// the committed state of every Land17 element from its material's initial values
// (lambda = 1, the undeformed configuration)
int land17_reset_state(cmg_ctx* c) {
    std::vector<double> st((size_t)8 * c->E, 0.0);
    for (int64_t e = 0; e < c->E; e++) {
        const double* p = c->h_l17_params.data() + (size_t)kLand17Params * c->h_matidx[e];
        for (int k = 0; k < 7; k++) st[(size_t)k * c->E + e] = p[27 + k];
        st[(size_t)7 * c->E + e] = 1.0;
    }
    CMG_HIP_CHECK(hipMemcpyAsync(c->l17_state.p, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

// Land17 commit after an accepted step at `time` with the nodes in x_dev
int land17_commit(cmg_ctx* c, const double* x_dev, double time) {
    if (!c->l17) return CMG_SUCCESS;
    launch_land17_commit(c->npe, elem_args(c, x_dev, time), c->l17_state.as<double>(), c->stream);
    CMG_HIP_CHECK(hipGetLastError());
    c->l17_t_commit = time;
    return CMG_SUCCESS;
}

// CSR incidence list node -> (item * width + local) in item order
void build_incidence(int64_t N, int64_t items, int width, const int32_t* conn, std::vector<int64_t>& ptr,
                     std::vector<int32_t>& inc) {
    ptr.assign(N + 1, 0);
    for (int64_t i = 0; i < items * width; i++) ptr[conn[i] + 1]++;
    for (int64_t r = 0; r < N; r++) ptr[r + 1] += ptr[r];
    inc.resize(ptr[N]);
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t i = 0; i < items * width; i++) inc[fill[conn[i]]++] = (int32_t)i;
}

}  // namespace

// ===========================================================================
// ABI
// ===========================================================================
extern "C" {

void cmg_default_options(cmg_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->precision = CMG_PREC_FP64;
    o->transpose_tangent = 0;
    o->pressure_tangent = 1;
    o->consistent_mass = 1;
    o->newton_maxit = 50;
    o->krylov_restart = 80;
    o->krylov_maxit = 4000;
    o->krylov_method = 2;   // BiCGStab, host-side scalars: the fastest on the 200k-element LV beat
    o->device = 0;
    o->element_variant = 0;   // by precision: 6 (FP64), 5 (mixed), the fastest of each on the LV
    o->newton_atol = 1e-8;
    o->newton_rtol = 1e-8;
    o->newton_stol = 1e-8;
    o->krylov_rtol = 1e-6;
    o->newmark_beta = 0.25;
    o->newmark_gamma = 0.5;
    o->fd_epsilon = 1e-12;
    o->preconditioner = 0;   // vertex-patch Schwarz for T10, block Jacobi otherwise
    o->krylov_eta_max = 0.5;
    o->precond_reuse = 1;   // as the reference lags it; 6 % (FP64) / 3 % (mixed) faster on the 200k beat window
    o->newton_stall = 0;    // the reference runs a stalled solve on to Solver.MaxIterations
}

const char* cmg_last_error(void) { return g_last_error.c_str(); }

int cmg_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void cmg_destroy(cmg_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->bicg_graph) (void)hipGraphExecDestroy(c->bicg_graph);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    if (c->h_part) (void)hipHostFree(c->h_part);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int cmg_create(const cmg_mesh* mesh, const cmg_material* materials, int32_t n_materials, const cmg_options* options,
               cmg_ctx** out) {
    if (!mesh || !materials || n_materials <= 0 || !out) return fail("cmg_create: null argument");
    *out = nullptr;
    const int npe = mesh->nodes_per_element;
    if (npe != 10 && npe != 4) return fail("cmg_create: nodes_per_element must be 10 or 4");
    if (mesh->n_nodes <= 0 || mesh->n_elements <= 0 || !mesh->ref_coords || !mesh->elements || !mesh->bases)
        return fail("cmg_create: empty mesh");
    if (mesh->n_nodes > INT_MAX - 1) return fail("cmg_create: too many nodes");
    if (mesh->n_surfaces > 0 && mesh->nodes_per_surface != 3 && mesh->nodes_per_surface != 6)
        return fail("cmg_create: nodes_per_surface must be 3 or 6");
    std::unique_ptr<cmg_ctx> c(new cmg_ctx());
    if (options)
        c->opt = *options;
    else
        cmg_default_options(&c->opt);
    if (c->opt.spmv_choice < 0 || c->opt.spmv_choice > 9) return fail("cmg_options.spmv_choice must be 0-9");
    if (c->opt.element_variant < 1 || c->opt.element_variant > 8)
        c->opt.element_variant = (c->opt.precision == CMG_PREC_MIXED) ? 5 : 6;
    if (!std::isfinite(c->opt.rayleigh_alpha) || !std::isfinite(c->opt.rayleigh_beta))
        return fail("cmg_create: Rayleigh damping coefficients must be finite");
    c->device = c->opt.device;
    CMG_HIP_CHECK(hipSetDevice(c->device));
    CMG_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CMG_HIP_CHECK(hipHostMalloc((void**)&c->h_pinned, 4096 * sizeof(double), hipHostMallocDefault));
    hipStream_t s = c->stream;
    c->npe = npe;
    c->nqp = (npe == 10) ? 4 : 1;
    c->N = mesh->n_nodes;
    c->E = mesh->n_elements;
    c->n = 3 * c->N;
    c->S = mesh->n_surfaces > 0 ? mesh->n_surfaces : 0;
    c->nps = c->S > 0 ? mesh->nodes_per_surface : 0;
    c->n_groups = std::max(1, mesh->n_groups);
    c->fp32 = c->opt.precision == CMG_PREC_MIXED;
    {
        const char* tr = std::getenv("CMG_NEWTON_TRACE");   // per-iteration Newton log on stderr
        c->trace = tr && *tr && *tr != '0';
    }
    c->damping = c->opt.rayleigh_alpha != 0.0 || c->opt.rayleigh_beta != 0.0;
    const int64_t N = c->N, E = c->E;
    // -- validate connectivity
    for (int64_t i = 0; i < E * npe; i++)
        if (mesh->elements[i] < 0 || mesh->elements[i] >= N) return fail("cmg_create: element node id out of range");
    for (int64_t i = 0; i < c->S * c->nps; i++)
        if (mesh->surfaces[i] < 0 || mesh->surfaces[i] >= N) return fail("cmg_create: surface node id out of range");
    if (mesh->material_index)
        for (int64_t e = 0; e < E; e++)
            if (mesh->material_index[e] < 0 || mesh->material_index[e] >= n_materials)
                return fail("cmg_create: material index out of range");
    if (c->S > 0 && mesh->surface_group)
        for (int64_t i = 0; i < c->S; i++)
            if (mesh->surface_group[i] < 0 || mesh->surface_group[i] >= c->n_groups)
                return fail("cmg_create: surface group out of range");
    if ((int64_t)E * npe * npe >= ((int64_t)1 << 60)) return fail("cmg_create: mesh too large");
    // -- materials
    std::vector<DevMaterial> dm(n_materials);
    for (int i = 0; i < n_materials; i++) CMG_TRY(to_dev_material(materials[i], dm[i]));
    c->n_mats = n_materials;
    CMG_TRY(upload(c->mats, dm.data(), dm.size(), s));
    c->h_matidx.assign(E, 0);
    if (mesh->material_index) std::copy(mesh->material_index, mesh->material_index + E, c->h_matidx.begin());
    for (int i = 0; i < n_materials; i++) c->l17 = c->l17 || materials[i].tension_fn == CMG_FN_LAND17;
    if (c->l17) {
        c->h_l17_params.assign((size_t)n_materials * kLand17Params, 0.0);
        for (int i = 0; i < n_materials; i++)
            std::copy(kLand17Defaults, kLand17Defaults + kLand17Params, c->h_l17_params.begin() + (size_t)i * kLand17Params);
        CMG_TRY(upload(c->l17_params, c->h_l17_params.data(), c->h_l17_params.size(), s));
        CMG_TRY(upload(c->l17_cai, &kLand17Cai[0][0], (size_t)2 * kLand17CaSamples, s));
        CMG_TRY(alloc(c->l17_state, (size_t)8 * E * sizeof(double)));
        CMG_TRY(land17_reset_state(c.get()));
    }
    // -- per-element SoA data
    std::vector<int32_t> conn_soa((size_t)npe * E);
    for (int64_t e = 0; e < E; e++)
        for (int k = 0; k < npe; k++) conn_soa[(size_t)k * E + e] = mesh->elements[e * npe + k];
    CMG_TRY(upload(c->conn, conn_soa.data(), conn_soa.size(), s));
    const int nb_in = (npe == 10) ? 5 : 1;
    std::vector<double> bases_soa((size_t)9 * c->nqp * E);
    for (int64_t e = 0; e < E; e++)
        for (int q = 0; q < c->nqp; q++) {
            const double* B = mesh->bases + ((size_t)e * nb_in + (npe == 10 ? q + 1 : 0)) * 9;
            for (int r = 0; r < 9; r++) bases_soa[(size_t)(q * 9 + r) * E + e] = B[r];
        }
    CMG_TRY(upload(c->bases, bases_soa.data(), bases_soa.size(), s));
    std::vector<double>().swap(bases_soa);
    std::vector<double> tact(E, 0.0);
    if (mesh->activation_time) std::copy(mesh->activation_time, mesh->activation_time + E, tact.begin());
    CMG_TRY(upload(c->tact, tact.data(), tact.size(), s));
    if (mesh->material_index) CMG_TRY(upload(c->matidx, mesh->material_index, E, s));
    c->h_bc.assign(N, 0);
    if (mesh->bc)
        for (int64_t i = 0; i < N; i++) c->h_bc[i] = mesh->bc[i] & 7;
    CMG_TRY(upload(c->bcmask, c->h_bc.data(), N, s));
    CMG_TRY(upload(c->Xref, mesh->ref_coords, 3 * N, s));
    CMG_TRY(alloc(c->w, E * sizeof(double)));
    CMG_TRY(alloc(c->dNdX, (size_t)c->nqp * npe * 3 * E * sizeof(double)));
    std::vector<double> wh(E);
    if (npe == 10) {
        DevBuf conn_aos;
        CMG_TRY(upload(conn_aos, mesh->elements, (size_t)E * npe, s));
        launch_t10_shape(E, conn_aos.as<int32_t>(), c->Xref.as<double>(), c->dNdX.as<double>(), c->w.as<double>(),
                         s);
        CMG_HIP_CHECK(hipGetLastError());
        CMG_HIP_CHECK(hipMemcpyAsync(wh.data(), c->w.p, E * sizeof(double), hipMemcpyDeviceToHost, s));
        CMG_HIP_CHECK(hipStreamSynchronize(s));
        for (auto& v : wh) v *= 24.0;   // detJ for the mass matrix
    } else {
        std::vector<double> g_soa((size_t)12 * E);
        std::vector<double> wv(E);
        for (int64_t e = 0; e < E; e++) {
            double X[12], g[12];
            for (int k = 0; k < 4; k++)
                for (int d = 0; d < 3; d++) X[3 * k + d] = mesh->ref_coords[3 * (int64_t)mesh->elements[e * 4 + k] + d];
            const double detJ = t4_shape(X, g);
            for (int r = 0; r < 12; r++) g_soa[(size_t)r * E + e] = g[r];
            wv[e] = detJ / 6.;   // initialVolume_ = GetVolume() = detJ/6
            wh[e] = detJ;
        }
        CMG_HIP_CHECK(hipMemcpyAsync(c->dNdX.p, g_soa.data(), g_soa.size() * sizeof(double), hipMemcpyHostToDevice, s));
        CMG_HIP_CHECK(hipMemcpyAsync(c->w.p, wv.data(), E * sizeof(double), hipMemcpyHostToDevice, s));
        CMG_HIP_CHECK(hipStreamSynchronize(s));
    }
    // -- incidence lists
    std::vector<int64_t> iptr;
    std::vector<int32_t> inc;
    build_incidence(N, E, npe, mesh->elements, iptr, inc);
    CMG_TRY(upload(c->inc_ptr, iptr.data(), iptr.size(), s));
    CMG_TRY(upload(c->inc, inc.data(), inc.size(), s));
    std::vector<int64_t> sptr;
    std::vector<int32_t> sinc;
    if (c->S > 0) {
        build_incidence(N, c->S, c->nps, mesh->surfaces, sptr, sinc);
        CMG_TRY(upload(c->sinc_ptr, sptr.data(), sptr.size(), s));
        CMG_TRY(upload(c->sinc, sinc.data(), sinc.size(), s));
    }
    // -- block pattern: node r couples to every node sharing an element (CBModel.cpp:43-83)
    {
        c->h_rowptr.assign(N + 1, 0);
        std::vector<std::vector<int32_t>> rows;   // built per node then flattened
        std::vector<int32_t> tmp;
        std::vector<int32_t> cols;
        cols.reserve((size_t)N * 30);
        for (int64_t r = 0; r < N; r++) {
            tmp.clear();
            for (int64_t k = iptr[r]; k < iptr[r + 1]; k++) {
                const int64_t e = inc[k] / npe;
                for (int b = 0; b < npe; b++) tmp.push_back(mesh->elements[e * npe + b]);
            }
            if (tmp.empty()) tmp.push_back((int32_t)r);   // isolated node: keep a diagonal block
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            cols.insert(cols.end(), tmp.begin(), tmp.end());
            c->h_rowptr[r + 1] = (int64_t)cols.size();
        }
        c->h_colind.swap(cols);
        c->nnzb = (int64_t)c->h_colind.size();
    }
    const int64_t nnzb = c->nnzb;
    std::vector<int32_t> blkrow(nnzb);
    for (int64_t r = 0; r < N; r++)
        for (int64_t k = c->h_rowptr[r]; k < c->h_rowptr[r + 1]; k++) blkrow[k] = (int32_t)r;
    CMG_TRY(upload(c->rowptr, c->h_rowptr.data(), N + 1, s));
    {
        // padded by kSpmvPadBlocks entries for the chunked SpMV (read, never used)
        std::vector<int32_t> colpad(c->h_colind);
        colpad.resize(nnzb + kSpmvPadBlocks, 0);
        CMG_TRY(upload(c->colind, colpad.data(), colpad.size(), s));
        CMG_HIP_CHECK(hipStreamSynchronize(s));
    }
    CMG_TRY(upload(c->blkrow, blkrow.data(), nnzb, s));
    // -- assembly contributor lists (deterministic order: solid elements, then surfaces)
    {
        const int64_t ns_blocks = E * npe * npe;
        c->n_solid_blocks = ns_blocks;
        const int nps = c->nps;
        std::vector<int64_t> slot_of((size_t)ns_blocks);
        std::vector<int64_t> cnt(nnzb + 1, 0);
        for (int64_t e = 0; e < E; e++)
            for (int a = 0; a < npe; a++)
                for (int b = 0; b < npe; b++) {
                    const int64_t sl = find_slot(c.get(), mesh->elements[e * npe + a], mesh->elements[e * npe + b]);
                    if (sl < 0) return fail("cmg_create: pattern construction failed");
                    slot_of[(e * npe + a) * npe + b] = sl;
                    cnt[sl + 1]++;
                }
        std::vector<int64_t> sslot((size_t)c->S * nps * nps);
        for (int64_t si = 0; si < c->S; si++)
            for (int a = 0; a < nps; a++)
                for (int b = 0; b < nps; b++) {
                    const int64_t sl = find_slot(c.get(), mesh->surfaces[si * nps + a], mesh->surfaces[si * nps + b]);
                    if (sl < 0) return fail("cmg_create: surface element nodes do not share a solid element");
                    sslot[(si * nps + a) * nps + b] = sl;
                    cnt[sl + 1]++;
                }
        for (int64_t k = 0; k < nnzb; k++) cnt[k + 1] += cnt[k];
        if (ns_blocks + ((c->S * nps * nps) << 3) + 8 >= (int64_t)INT_MAX)
            return fail("cmg_create: mesh too large for the int32 assembly maps");
        std::vector<int32_t> contrib(cnt[nnzb]);
        std::vector<int64_t> fillp(cnt.begin(), cnt.end() - 1);
        // T10 blocks may enter transposed (the reference convention, CBElementSolidT10.cpp:559-565);
        // the T4 finite-difference Jacobian is assembled as is (CBElementSolidT4.cpp:340-349)
        const bool trans = c->opt.transpose_tangent != 0 && npe == 10;
        c->maps_transposed = trans;
        for (int64_t e = 0; e < E; e++)
            for (int a = 0; a < npe; a++)
                for (int b = 0; b < npe; b++) {
                    const int64_t sl = slot_of[(e * npe + a) * npe + b];
                    // transposed: block (a,b) of the assembled matrix takes (K_e(b,a))^T
                    contrib[fillp[sl]++] = (int32_t)(trans ? ~((e * npe + b) * npe + a) : ((e * npe + a) * npe + b));
                }
        for (int64_t si = 0; si < c->S; si++)
            for (int a = 0; a < nps; a++)
                for (int b = 0; b < nps; b++) {
                    const int64_t sl = sslot[(si * nps + a) * nps + b];
                    contrib[fillp[sl]++] = (int32_t)(ns_blocks + ((((si * nps + a) * nps + b) << 3) | nps));
                }
        // row gather tables: position within its row of the block each incidence's local
        // nodes map to (element order and surface order, as the contributor lists)
        int64_t longest = 0;
        for (int64_t r = 0; r < N; r++) longest = std::max(longest, c->h_rowptr[r + 1] - c->h_rowptr[r]);
        const int64_t cap = (longest + 7) / 8 * 8;
        // (one surface value per lane: 9 nps <= 64)
        c->row_cap = (cap <= kRowGatherMaxCap && npe <= 16 && 9 * nps <= 64) ? (int)cap : 0;
        if (c->row_cap > 0) {
            std::vector<uint8_t> rpos(inc.size() * 16, 0);
            for (size_t k = 0; k < inc.size(); k++) {
                const int64_t r = mesh->elements[inc[k]];
                for (int b = 0; b < npe; b++)
                    rpos[16 * k + b] = (uint8_t)(slot_of[(int64_t)inc[k] * npe + b] - c->h_rowptr[r]);
            }
            CMG_TRY(upload(c->rpos, rpos.data(), rpos.size(), s));
            if (c->S > 0) {
                std::vector<uint8_t> spos(sinc.size() * 8, 0);
                for (size_t k = 0; k < sinc.size(); k++) {
                    const int64_t r = mesh->surfaces[sinc[k]];
                    for (int b = 0; b < nps; b++)
                        spos[8 * k + b] = (uint8_t)(sslot[(int64_t)sinc[k] * nps + b] - c->h_rowptr[r]);
                }
                CMG_TRY(upload(c->spos, spos.data(), spos.size(), s));
            }
            std::vector<uint8_t> cmask(nnzb);
            for (int64_t k = 0; k < nnzb; k++) cmask[k] = c->h_bc[c->h_colind[k]];
            CMG_TRY(upload(c->cmask, cmask.data(), cmask.size(), s));
        }
        std::vector<int32_t> heavy;
        for (int64_t k = 0; k < nnzb; k++)
            if (cnt[k + 1] - cnt[k] > kHeavyContrib) heavy.push_back((int32_t)k);
        c->n_heavy = (int64_t)heavy.size();
        CMG_TRY(upload(c->cptr, cnt.data(), cnt.size(), s));
        CMG_TRY(upload(c->contrib, contrib.data(), contrib.size(), s));
        CMG_TRY(upload(c->heavy, heavy.data(), heavy.size(), s));
        // -- mass matrix: one scalar per block (consistent CBElementSolidT10.cpp:352-394 /
        //    CBElementSolidT4.cpp:365-410, or lumped :396-422 / :432-456)
        c->h_mval.assign(nnzb, 0.0);
        for (int64_t e = 0; e < E; e++) {
            const int mi = mesh->material_index ? mesh->material_index[e] : 0;
            const double rho = dm[mi].density;
            for (int a = 0; a < npe; a++)
                for (int b = 0; b < npe; b++) {
                    double v;
                    if (npe == 10)
                        v = c->opt.consistent_mass ? kR1[10 * a + b] * ((rho * wh[e]) / 2520)
                                                   : ((a == b) ? (rho * (wh[e] / 6)) / 10 : 0.0);
                    else
                        v = c->opt.consistent_mass ? ((a == b) ? 2.0 : 1.0) * ((rho * (wh[e] / 6.)) / 20)
                                                   : ((a == b) ? (rho * (wh[e] / 6.)) / 4 : 0.0);
                    c->h_mval[slot_of[(e * npe + a) * npe + b]] += v;
                }
        }
        CMG_TRY(upload(c->mval, c->h_mval.data(), nnzb, s));
    }
    // -- preconditioner patches (need the block pattern)
    if (c->opt.preconditioner != 1) CMG_TRY(build_vertex_patches(c.get(), mesh->elements));
    // -- surfaces
    c->h_pressure.assign(c->n_groups, 0.0);
    CMG_TRY(upload(c->pressure, c->h_pressure.data(), c->n_groups, s));
    if (c->S > 0) {
        CMG_TRY(upload(c->sconn, mesh->surfaces, (size_t)c->S * c->nps, s));
        std::vector<int32_t> grp(c->S, 0);
        if (mesh->surface_group) std::copy(mesh->surface_group, mesh->surface_group + c->S, grp.begin());
        CMG_TRY(upload(c->sgroup, grp.data(), grp.size(), s));
        const int ns = 3 * c->nps;
        CMG_TRY(alloc(c->fs, (size_t)c->S * ns * sizeof(double)));
        CMG_TRY(alloc(c->Ks, (size_t)c->S * ns * ns * sizeof(double)));
        CMG_TRY(alloc(c->vol_partial, (size_t)c->S * sizeof(double)));
    }
    // -- work buffers
    const size_t vb = (size_t)c->n * sizeof(double);
    for (DevBuf* b : {&c->nodes, &c->xeval, &c->d, &c->dpred, &c->vpred, &c->vel, &c->acc, &c->r, &c->delta,
                      &c->dsave, &c->tmpA, &c->tmpB, &c->guess, &c->override_guess, &c->inc1, &c->inc2})
        CMG_TRY(alloc(*b, vb));
    CMG_HIP_CHECK(hipMemcpyAsync(c->nodes.p, mesh->ref_coords, vb, hipMemcpyHostToDevice, s));
    for (DevBuf* b : {&c->d, &c->vel, &c->acc, &c->guess, &c->dpred, &c->vpred, &c->inc1, &c->inc2})
        launch_fill(c->n, 0.0, b->as<double>(), s);
    CMG_TRY(alloc(c->fe, (size_t)E * 3 * npe * sizeof(double)));
    // FP64-sized in mixed precision when the Rayleigh damping matrix needs FP64 element tangents
    const bool ke64 = !c->fp32 || (c->damping && c->opt.rayleigh_beta != 0.0);
    CMG_TRY(alloc(c->Ke, (size_t)E * 9 * npe * npe * (ke64 ? sizeof(double) : sizeof(float))));
    CMG_TRY(alloc(c->vals, (size_t)(nnzb + kSpmvPadBlocks) * 9 * (c->fp32 ? sizeof(float) : sizeof(double))));
    launch_fill_int((nnzb + kSpmvPadBlocks) * 9 * (c->fp32 ? 1 : 2), 0, c->vals.as<int>(), s);   // defined padding
    CMG_TRY(alloc(c->dinv, (size_t)N * 9 * sizeof(double)));
    if (c->damping) {
        CMG_TRY(alloc(c->cvals, (size_t)(nnzb + kSpmvPadBlocks) * 9 * sizeof(double)));
        launch_fill((nnzb + kSpmvPadBlocks) * 9, 0.0, c->cvals.as<double>(), s);
    }
    CMG_TRY(alloc(c->corrupt, sizeof(int)));
    c->gm_m = std::max(8, std::min(c->opt.krylov_restart, 1000));
    CMG_TRY(alloc(c->gm_V, (size_t)(c->gm_m + 2) * vb));
    CMG_TRY(alloc(c->gm_partial, (size_t)dot_blocks(c->n) * (c->gm_m + 2) * sizeof(double) + 64));
    CMG_TRY(alloc(c->gm_h, (size_t)(c->gm_m + 2) * sizeof(double)));
    CMG_TRY(alloc(c->gm_y, (size_t)(c->gm_m + 2) * sizeof(double)));
    CMG_HIP_CHECK(hipGetLastError());
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    *out = c.release();
    return CMG_SUCCESS;
}

int cmg_sizes(cmg_ctx* c, int64_t* n_nodes, int64_t* n_elements, int64_t* n_dofs, int64_t* nnz_blocks) {
    if (!c) return fail("null context");
    if (n_nodes) *n_nodes = c->N;
    if (n_elements) *n_elements = c->E;
    if (n_dofs) *n_dofs = c->n;
    if (nnz_blocks) *nnz_blocks = c->nnzb;
    return CMG_SUCCESS;
}

int cmg_pattern(cmg_ctx* c, int64_t* rowptr, int32_t* colind) {
    if (!c) return fail("null context");
    if (rowptr) std::copy(c->h_rowptr.begin(), c->h_rowptr.end(), rowptr);
    if (colind) std::copy(c->h_colind.begin(), c->h_colind.end(), colind);
    return CMG_SUCCESS;
}

void* cmg_stream(cmg_ctx* c) { return c ? (void*)c->stream : nullptr; }

int cmg_synchronize(cmg_ctx* c) {
    if (!c) return fail("null context");
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

int cmg_set_pressures(cmg_ctx* c, const double* p) {
    if (!c || !p) return fail("cmg_set_pressures: null argument");
    std::copy(p, p + c->n_groups, c->h_pressure.begin());
    CMG_HIP_CHECK(hipMemcpyAsync(c->pressure.p, c->h_pressure.data(), c->n_groups * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

}  // extern "C"

// ===========================================================================
// internal evaluation helpers
// ===========================================================================
namespace cmg {

ElemArgs elem_args(cmg_ctx* c, const double* x_dev, double time) {
    ElemArgs a;
    a.E = c->E;
    a.conn = c->conn.as<int32_t>();
    a.dNdX = c->dNdX.as<double>();
    a.bases = c->bases.as<double>();
    a.w = c->w.as<double>();
    a.matidx = c->matidx.p ? c->matidx.as<int32_t>() : nullptr;
    a.tact = c->tact.as<double>();
    a.mats = c->mats.as<DevMaterial>();
    a.x = x_dev;
    a.time = time;
    a.fe = c->fe.as<double>();
    a.Ke = c->Ke.p;
    a.corrupt = c->corrupt.as<int>();
    if (c->l17) {
        a.l17_params = c->l17_params.as<double>();
        a.l17_state = c->l17_state.as<double>();
        a.l17_cai = c->l17_cai.as<double>();
        a.l17_dt_ms = (time - c->l17_t_commit) * 1000.0;
    }
    return a;
}

GatherMaps gather_maps(cmg_ctx* c) {
    GatherMaps g;
    g.nnzb = c->nnzb;
    g.cptr = c->cptr.as<int64_t>();
    g.contrib = c->contrib.as<int32_t>();
    g.heavy = c->heavy.as<int32_t>();
    g.n_heavy = c->n_heavy;
    g.n_solid = (int32_t)c->n_solid_blocks;
    g.npe = c->npe;
    g.blkrow = c->blkrow.as<int32_t>();
    g.colind = c->colind.as<int32_t>();
    g.transposed = c->maps_transposed;
    g.N = c->N;
    g.nps = c->nps;
    g.row_cap = c->row_cap;
    g.rowptr = c->rowptr.as<int64_t>();
    g.inc_ptr = c->inc_ptr.as<int64_t>();
    g.inc = c->inc.as<int32_t>();
    g.sinc_ptr = c->S > 0 ? c->sinc_ptr.as<int64_t>() : nullptr;
    g.sinc = c->S > 0 ? c->sinc.as<int32_t>() : nullptr;
    g.rpos = c->rpos.as<uint8_t>();
    g.spos = c->spos.as<uint8_t>();
    g.cmask = c->cmask.as<uint8_t>();
    return g;
}

SurfArgs surf_args(cmg_ctx* c, const double* x_dev, bool with_volume) {
    SurfArgs a;
    a.S = c->S;
    a.nps = c->nps;
    a.conn = c->sconn.as<int32_t>();
    a.group = c->sgroup.as<int32_t>();
    a.pressure = c->pressure.as<double>();
    a.bcmask = c->bcmask.as<uint8_t>();
    a.x = x_dev;
    a.fs = c->fs.as<double>();
    a.Ks = c->Ks.as<double>();
    a.vol_partial = with_volume ? c->vol_partial.as<double>() : nullptr;
    return a;
}

}  // namespace cmg

namespace {

// element kernels (+ surfaces) at x_dev; returns CMG_SUCCESS or CMG_CORRUPT_ELEMENT
int eval_elements(cmg_ctx* c, const double* x_dev, double time, bool tangent, bool pressure, int* corrupt_id,
                  bool fp64_tangent = false) {
    hipStream_t s = c->stream;
    launch_fill_int(1, INT_MAX, c->corrupt.as<int>(), s);
    ElemArgs a = elem_args(c, x_dev, time);
    if (tangent && fp64_tangent)
        launch_element_residual_tangent(c->npe, false, a, s, 6);
    else if (tangent)
        launch_element_residual_tangent(c->npe, c->fp32, a, s, c->opt.element_variant);
    else
        launch_element_residual(c->npe, a, s);
    if (pressure && c->S > 0) launch_surface(surf_args(c, x_dev, false), tangent && c->opt.pressure_tangent, s);
    CMG_HIP_CHECK(hipGetLastError());
    int flag = INT_MAX;
    CMG_HIP_CHECK(hipMemcpyAsync(c->h_pinned, c->corrupt.p, sizeof(int), hipMemcpyDeviceToHost, s));
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    std::memcpy(&flag, c->h_pinned, sizeof(int));
    if (corrupt_id) *corrupt_id = (flag == INT_MAX) ? -1 : flag;
    if (flag != INT_MAX) {
        set_error("element " + std::to_string(flag) + " is corrupt (det F <= 0)");
        return CMG_CORRUPT_ELEMENT;
    }
    return CMG_SUCCESS;
}

void gather_residual(cmg_ctx* c, bool pressure, double* f_dev) {
    const bool sp = pressure && c->S > 0;
    launch_gather_vector(c->N, c->inc_ptr.as<int64_t>(), c->inc.as<int32_t>(), c->fe.as<double>(), 3 * c->npe,
                         sp ? c->sinc_ptr.as<int64_t>() : nullptr, sp ? c->sinc.as<int32_t>() : nullptr,
                         sp ? c->fs.as<double>() : nullptr, 3 * c->nps, c->bcmask.as<uint8_t>(), f_dev, c->stream);
}

void gather_tangent(cmg_ctx* c, bool pressure, double cm, void* vals_dev) {
    const bool sp = pressure && c->S > 0 && c->opt.pressure_tangent;
    // without a pressure tangent the surface blocks read zeros: clear Ks
    if (c->S > 0 && !sp) launch_fill((int64_t)c->S * 9 * c->nps * c->nps, 0.0, c->Ks.as<double>(), c->stream);
    launch_gather_matrix(gather_maps(c), c->Ke.p, c->fp32, c->Ks.as<double>(), cm != 0.0 ? c->mval.as<double>() : nullptr,
                         cm, c->bcmask.as<uint8_t>(), vals_dev, c->fp32, false, c->stream);
}

// ||v||_2 (synchronising)
int norm2(cmg_ctx* c, const double* v, double* out) {
    hipStream_t s = c->stream;
    const int nb = dot_blocks(c->n);
    launch_multidot(c->n, 1, v, 0, v, c->gm_partial.as<double>(), s);
    launch_reduce_partials(nb, 1, c->gm_partial.as<double>(), c->gm_h.as<double>(), s);
    CMG_HIP_CHECK(hipMemcpyAsync(c->h_pinned, c->gm_h.p, sizeof(double), hipMemcpyDeviceToHost, s));
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    const double h = c->h_pinned[0];
    *out = std::isfinite(h) ? std::sqrt(std::max(0.0, h)) : INFINITY;   // NaN must not read as converged
    return CMG_SUCCESS;
}

// dots of V_0..V_{k-1} (stride ld) with w -> host array h
int multidot(cmg_ctx* c, int k, const double* V, int64_t ld, const double* w, double* h) {
    hipStream_t s = c->stream;
    const int nb = dot_blocks(c->n);
    launch_multidot(c->n, k, V, ld, w, c->gm_partial.as<double>(), s);
    launch_reduce_partials(nb, k, c->gm_partial.as<double>(), c->gm_h.as<double>(), s);
    CMG_HIP_CHECK(hipMemcpyAsync(c->h_pinned, c->gm_h.p, k * sizeof(double), hipMemcpyDeviceToHost, s));
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    std::memcpy(h, c->h_pinned, k * sizeof(double));
    return CMG_SUCCESS;
}

int to_device_coeffs(cmg_ctx* c, const double* h, int k) {
    std::memcpy(c->h_pinned + 2048, h, k * sizeof(double));
    CMG_HIP_CHECK(hipMemcpyAsync(c->gm_y.p, c->h_pinned + 2048, k * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // the pinned staging area is reused by the next call only after a later synchronisation
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

// the Krylov preconditioner: vertex-patch additive Schwarz when the context built its
// patches, 3x3 nodal block Jacobi otherwise
bool use_schwarz(const cmg_ctx* c) { return c->schwarz; }

void precond_setup(cmg_ctx* c, const void* vals) {
    // lagged (Schwarz only: a block-Jacobi setup is one pass over the diagonal blocks, and
    // reusing it cost the 200k beat window 50 % more Krylov iterations)
    if (c->opt.precond_reuse && c->precond_built && use_schwarz(c)) return;
    c->precond_built = true;
    if (use_schwarz(c))
        launch_patch_setup(c, vals);
    else
        launch_block_jacobi_setup(c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), vals, c->fp32,
                                  c->dinv.as<double>(), c->stream);
}

void precond_apply(cmg_ctx* c, const double* r, double* z) {
    if (use_schwarz(c))
        launch_patch_apply(c, r, z);
    else
        launch_block_jacobi_apply(c->N, c->dinv.as<double>(), r, z, c->stream);
}

// GMRES(m) with right preconditioning (block Jacobi or vertex-patch Schwarz): solves
// A x = b (x holds the initial guess).  Classical Gram-Schmidt with DGKS
// re-orthogonalisation.
int gmres(cmg_ctx* c, const void* vals, const double* b, double* x, double rtol, int maxit, cmg_solve_stats* st) {
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    const int m = c->gm_m;
    double* V = c->gm_V.as<double>();
    double* w = c->tmpA.as<double>();
    double* z = c->tmpB.as<double>();
    precond_setup(c, vals);
    double bnorm;
    CMG_TRY(norm2(c, b, &bnorm));
    // r = b - A x
    launch_spmv(c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), vals, c->fp32, x, w, s, c->spmv_lanes,
                c->spmv_layout);
    launch_waxpby(n, 1.0, b, -1.0, w, V, s);
    double beta;
    CMG_TRY(norm2(c, V, &beta));
    st->initial_residual = beta;
    st->iterations = 0;
    st->converged = 0;
    const double target = rtol * (bnorm > 0 ? bnorm : 1.0);
    if (beta <= target || beta == 0.0) {
        st->final_residual = beta;
        st->converged = 1;
        return CMG_SUCCESS;
    }
    std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), h(m + 2), h2(m + 2), y(m);
    int total = 0;
    while (true) {
        launch_scale_copy(n, 1.0 / beta, V, V, s);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        double resid = beta;
        for (; j < m && total < maxit;) {
            double* vj = V + (int64_t)j * n;
            double* vn = V + (int64_t)(j + 1) * n;
            precond_apply(c, vj, z);
            launch_spmv(c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), vals, c->fp32, z, vn, s, c->spmv_lanes,
                        c->spmv_layout);
            // pass 1: h_i = v_i . w (i <= j) and ||w||^2 in one multi-dot
            CMG_TRY(multidot(c, j + 2, V, n, vn, h.data()));
            const double wn2 = h[j + 1];
            launch_multiaxpy(n, j + 1, V, n, c->gm_h.as<double>(), vn, s);   // coefficients stay on device
            double hs = 0.0;
            for (int i = 0; i <= j; i++) hs += h[i] * h[i];
            double nrm2 = wn2 - hs;
            if (nrm2 < 0.5 * wn2) {
                // strong cancellation: second classical Gram-Schmidt pass
                CMG_TRY(multidot(c, j + 2, V, n, vn, h2.data()));
                launch_multiaxpy(n, j + 1, V, n, c->gm_h.as<double>(), vn, s);
                double hs2 = 0.0;
                for (int i = 0; i <= j; i++) {
                    h[i] += h2[i];
                    hs2 += h2[i] * h2[i];
                }
                nrm2 = h2[j + 1] - hs2;
            }
            const double hn = std::sqrt(std::max(nrm2, 0.0));
            for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = h[i];
            H[(size_t)(j + 1) * m + j] = hn;
            for (int i = 0; i < j; i++) {   // apply previous rotations
                const double a0 = H[(size_t)i * m + j], a1 = H[(size_t)(i + 1) * m + j];
                H[(size_t)i * m + j] = cs[i] * a0 + sn[i] * a1;
                H[(size_t)(i + 1) * m + j] = -sn[i] * a0 + cs[i] * a1;
            }
            const double a0 = H[(size_t)j * m + j], a1 = H[(size_t)(j + 1) * m + j];
            const double den = std::hypot(a0, a1);
            cs[j] = den > 0 ? a0 / den : 1.0;
            sn[j] = den > 0 ? a1 / den : 0.0;
            H[(size_t)j * m + j] = den;
            H[(size_t)(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            resid = std::fabs(g[j + 1]);
            total++;
            j++;
            if (hn > 0.0) launch_scale_copy(n, 1.0 / hn, vn, vn, s);
            if (resid <= target || hn == 0.0) break;
        }
        // y = H^-1 g, x += M^-1 (V y)
        for (int i = j - 1; i >= 0; i--) {
            double sum = g[i];
            for (int k = i + 1; k < j; k++) sum -= H[(size_t)i * m + k] * y[k];
            y[i] = sum / H[(size_t)i * m + i];
        }
        std::vector<double> ny(j);
        for (int i = 0; i < j; i++) ny[i] = -y[i];
        launch_fill(n, 0.0, w, s);
        CMG_TRY(to_device_coeffs(c, ny.data(), j));
        launch_multiaxpy(n, j, V, n, c->gm_y.as<double>(), w, s);   // w = V y
        precond_apply(c, w, z);
        launch_axpby(n, 1.0, z, 1.0, x, s);
        // true residual
        launch_spmv(c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), vals, c->fp32, x, w, s, c->spmv_lanes,
                c->spmv_layout);
        launch_waxpby(n, 1.0, b, -1.0, w, V, s);
        CMG_TRY(norm2(c, V, &beta));
        st->iterations = total;
        st->final_residual = beta;
        if (beta <= target) {
            st->converged = 1;
            break;
        }
        if (total >= maxit) break;
    }
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}


// BiCGStab with right preconditioning (van der Vorst 1992).  Two host
// synchronisations per iteration: (r^, v) for alpha, and one fused fetch of
// (r^,t) (s,t) (t,t) (r^,s) (s,s) from which omega, the next rho and the updated
// residual norm ||s - omega t|| follow without another pass.
int bicgstab(cmg_ctx* c, const void* vals, const double* b, double* x, double rtol, int maxit, cmg_solve_stats* st) {
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    const int nb = dot_blocks(n);
    double* W = c->gm_V.as<double>();   // (restart + 2) >= 10 vectors
    double* rh = W;
    double* sv = W + n;
    double* t = W + 2 * n;
    double* r = W + 3 * n;
    double* p = W + 4 * n;
    double* v = W + 5 * n;
    double* ph = W + 6 * n;
    double* sh = W + 7 * n;
    double* part = c->gm_partial.as<double>();
    const double* dinv = c->dinv.as<double>();
    const int64_t* rp = c->rowptr.as<int64_t>();
    const int32_t* ci = c->colind.as<int32_t>();
    if (c->h_part_len < 5 * (int64_t)nb) {
        if (c->h_part) CMG_HIP_CHECK(hipHostFree(c->h_part));
        c->h_part = nullptr;
        CMG_HIP_CHECK(hipHostMalloc((void**)&c->h_part, 5 * (size_t)nb * sizeof(double), hipHostMallocDefault));
        c->h_part_len = 5 * (int64_t)nb;
    }
    // dot products: per-chunk partials copied to the host and summed there in chunk order
    auto host_sums = [&](int k, double* out) -> int {
        CMG_HIP_CHECK(hipMemcpyAsync(c->h_part, part, (size_t)k * nb * sizeof(double), hipMemcpyDeviceToHost, s));
        CMG_HIP_CHECK(hipStreamSynchronize(s));
        for (int j = 0; j < k; j++) out[j] = 0.0;
        for (int blk = 0; blk < nb; blk++)
            for (int j = 0; j < k; j++) out[j] += c->h_part[(int64_t)blk * k + j];
        return CMG_SUCCESS;
    };
    const bool sw = use_schwarz(c);
    precond_setup(c, vals);
    double bnorm;
    CMG_TRY(norm2(c, b, &bnorm));
    launch_spmv(c->N, rp, ci, vals, c->fp32, x, v, s, c->spmv_lanes, c->spmv_layout);
    launch_waxpby(n, 1.0, b, -1.0, v, r, s);
    double rn;
    CMG_TRY(norm2(c, r, &rn));
    st->initial_residual = rn;
    st->iterations = 0;
    st->converged = 0;
    const double target = rtol * (bnorm > 0 ? bnorm : 1.0);
    if (rn <= target || rn == 0.0) {
        st->final_residual = rn;
        st->converged = 1;
        return CMG_SUCCESS;
    }
    int it = 0, restarts = 0;
    bool fresh = true;
    double rho = 1.0, alpha = 1.0, omega = 1.0, rho_new = rn * rn;
    while (it < maxit) {
        // p = r + beta (p - omega v) (p = r on a (re)start, with shadow residual r^ = r); ph = D^-1 p
        double beta = 0.0;
        if (fresh) {
            launch_scale_copy(n, 1.0, r, rh, s);
            rho_new = rn * rn;
        } else {
            beta = (rho_new / rho) * (alpha / omega);
        }
        launch_bicg_p_bj(c->N, fresh, beta, omega, r, v, sw ? nullptr : dinv, p, ph, s);
        if (sw) precond_apply(c, p, ph);
        fresh = false;
        rho = rho_new;
        launch_spmv(c->N, rp, ci, vals, c->fp32, ph, v, s, c->spmv_lanes, c->spmv_layout);
        launch_multidot(n, 1, rh, 0, v, part, s);
        double d0;
        CMG_TRY(host_sums(1, &d0));
        bool breakdown = !(std::fabs(d0) > 0.0) || !std::isfinite(d0);
        if (!breakdown) {
            alpha = rho / d0;
            launch_bicg_s_bj(c->N, alpha, r, v, sw ? nullptr : dinv, sv, sh, s);   // s = r - alpha v, sh = D^-1 s
            if (sw) precond_apply(c, sv, sh);
            launch_spmv(c->N, rp, ci, vals, c->fp32, sh, t, s, c->spmv_lanes, c->spmv_layout);
            launch_bicg_dots5(n, rh, sv, t, part, s);              // (rh,t) (s,t) (t,t) (rh,s) (s,s)
            double d5[5];
            CMG_TRY(host_sums(5, d5));
            const double rt = d5[0], st_ = d5[1], tt = d5[2], rs = d5[3], ss = d5[4];
            omega = tt > 0.0 ? st_ / tt : 0.0;
            launch_bicg_xr(n, alpha, omega, ph, sh, sv, t, x, r, s);   // x += alpha ph + omega sh, r = s - omega t
            it++;
            rn = std::sqrt(std::max(ss - 2.0 * omega * st_ + omega * omega * tt, 0.0));
            rho_new = rs - omega * rt;
            breakdown = omega == 0.0 || !(std::fabs(rho_new) > 0.0) || !std::isfinite(rn);
        }
        if (rn <= target || breakdown) {
            // the recurrence residual drifts from b - A x: confirm, restart if needed
            launch_spmv(c->N, rp, ci, vals, c->fp32, x, v, s, c->spmv_lanes, c->spmv_layout);
            launch_waxpby(n, 1.0, b, -1.0, v, r, s);
            CMG_TRY(norm2(c, r, &rn));
            if (rn <= target || ++restarts > 50 || !std::isfinite(rn)) break;
            fresh = true;
        }
    }
    st->iterations = it;
    st->final_residual = rn;
    st->converged = rn <= target;
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}

// SpMV layout: the fastest of 4 / 8 / 16 lanes per block row, one block or one
// 144-byte chunk per step, on this matrix (HIP events, one warm and two timed products
// each), chosen once per context
int tune_spmv(cmg_ctx* c, const void* vals, const double* b) {
    hipStream_t s = c->stream;
    hipEvent_t e0, e1;
    CMG_HIP_CHECK(hipEventCreate(&e0));
    CMG_HIP_CHECK(hipEventCreate(&e1));
    const int lanes[3][3] = {{4, 8, 16}, {4, 8, 16}, {16, 32, 64}};
    float best = 1e30f, best_blk = 1e30f;
    const int64_t* rp = c->rowptr.as<int64_t>();
    const int32_t* ci = c->colind.as<int32_t>();
    if (c->opt.spmv_choice >= 1 && c->opt.spmv_choice <= 9) {   // fixed by the caller, no timing
        const int layout = (c->opt.spmv_choice - 1) / 3, li = (c->opt.spmv_choice - 1) % 3;
        c->spmv_lanes = lanes[layout][li];
        c->spmv_layout = layout;
        c->bicg_lanes = layout == kSpmvFlat ? 16 : lanes[layout][li];
        c->bicg_chunked = layout == kSpmvChunk;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return CMG_SUCCESS;
    }
    for (int layout = kSpmvBlock; layout <= kSpmvFlat; layout++)
        for (int li = 0; li < 3; li++) {
            const int l = lanes[layout][li];
            launch_spmv(c->N, rp, ci, vals, c->fp32, b, c->tmpA.as<double>(), s, l, layout);
            CMG_HIP_CHECK(hipEventRecord(e0, s));
            for (int k = 0; k < 2; k++) launch_spmv(c->N, rp, ci, vals, c->fp32, b, c->tmpA.as<double>(), s, l, layout);
            CMG_HIP_CHECK(hipEventRecord(e1, s));
            CMG_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            CMG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            c->spmv_ms[layout * 3 + li] = 0.5 * ms;
            if (ms < best) {
                best = ms;
                c->spmv_lanes = l;
                c->spmv_layout = layout;
            }
            if (layout != kSpmvFlat && ms < best_blk) {
                best_blk = ms;
                c->bicg_lanes = l;
                c->bicg_chunked = layout == kSpmvChunk;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return CMG_SUCCESS;
}


// BiCGStab with the scalar recurrences on the device (BatchBiCG with one set): an
// iteration is eight launches without host synchronisation; `bicg_chunk` iterations
// are replayed as one HIP graph, after which the host reads the done flag.  A
// converged or broken-down recurrence is confirmed with the true residual b - A x
// and restarted from the current x when needed.
int bicgstab_device(cmg_ctx* c, const void* vals, const double* b, double* x, double rtol, int maxit,
                    cmg_solve_stats* st) {
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    const int64_t* rp = c->rowptr.as<int64_t>();
    const int32_t* ci = c->colind.as<int32_t>();
    precond_setup(c, vals);
    if (!c->bicg_part.p) {
        CMG_TRY(alloc(c->bicg_part, (size_t)bicg_part_ld(c->N) * sizeof(double)));
        CMG_TRY(alloc(c->bicg_sc, 16 * sizeof(double)));
        CMG_TRY(alloc(c->bicg_flags, 8 * sizeof(int32_t)));
    }
    double* W = c->gm_V.as<double>();   // >= 10 work vectors
    BatchBiCG k;
    k.S = 1;
    k.x = x;
    k.r = W;
    k.rh = W + n;
    k.p = W + 2 * n;
    k.v = W + 3 * n;
    k.s = W + 4 * n;
    k.t = W + 5 * n;
    k.ph = W + 6 * n;
    k.sh = W + 7 * n;
    double* sc = c->bicg_sc.as<double>();
    k.rho = sc;
    k.alpha = sc + 1;
    k.omega = sc + 2;
    k.beta = sc + 3;
    k.rr = sc + 4;
    k.tol2 = sc + 5;
    double* rtol_dev = sc + 6;
    int32_t* fl = c->bicg_flags.as<int32_t>();
    k.done = fl;
    k.it = fl + 1;
    k.fin = fl + 2;
    int32_t* active = fl + 3;
    k.part = c->bicg_part.as<double>();
    k.part_ld = bicg_part_ld(c->N);
    launch_fill(1, rtol, rtol_dev, s);
    launch_fill_int(4, 0, fl, s);
    launch_fill_int(1, 1, active, s);
    auto restart = [&]() {
        // r = b - A x, then r^ = r, p = r and the scalars
        launch_spmv(c->N, rp, ci, vals, c->fp32, x, k.r, s, c->spmv_lanes, c->spmv_layout);
        launch_waxpby(n, 1.0, b, -1.0, k.r, k.r, s);
        launch_bicg_init(k, c->N, b, rtol_dev, active, s);
    };
    restart();
    // Schwarz: the patch kernels skip, like the update kernels, once the recurrence has
    // finished within a replayed chunk (the done flag k.done)
    const PrecondApply apply = use_schwarz(c) ? [](void* cc, const double* u, double* z) {
        cmg_ctx* cx = static_cast<cmg_ctx*>(cc);
        PatchSets ps;
        ps.flags = cx->bicg_flags.as<int32_t>();
        launch_patch_apply(cx, u, z, ps);
    } : PrecondApply(nullptr);
    auto iterate = [&]() {
        launch_bicg_iteration(k, c->bicg_lanes, c->bicg_chunked, c->N, rp, ci, vals, c->fp32, 0,
                              c->dinv.as<double>(), s, apply, c);
    };
    const bool graph_ok = c->bicg_graph != nullptr && c->bicg_graph_x == x && c->bicg_graph_vals == vals;
    if (!graph_ok && c->bicg_use_graph) {
        if (c->bicg_graph) (void)hipGraphExecDestroy(c->bicg_graph);
        c->bicg_graph = nullptr;
        hipGraph_t g = nullptr;
        CMG_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < c->bicg_chunk; i++) iterate();
        CMG_HIP_CHECK(hipStreamEndCapture(s, &g));
        hipError_t e = hipGraphInstantiate(&c->bicg_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        CMG_HIP_CHECK(e);
        c->bicg_graph_x = x;
        c->bicg_graph_vals = vals;
    }
    double* hd = c->h_pinned + 3500;   // pinned staging beyond the GMRES areas
    int32_t* hi = reinterpret_cast<int32_t*>(c->h_pinned + 3600);
    CMG_HIP_CHECK(hipMemcpyAsync(hd, sc, 6 * sizeof(double), hipMemcpyDeviceToHost, s));
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    const double target = std::sqrt(std::max(hd[5], 0.0));
    double rn = std::sqrt(std::max(hd[4], 0.0));
    st->initial_residual = rn;
    st->iterations = 0;
    st->converged = 0;
    int restarts = 0;
    if (rn <= target) {
        st->final_residual = rn;
        st->converged = 1;
        return CMG_SUCCESS;
    }
    for (;;) {
        if (c->bicg_graph)
            CMG_HIP_CHECK(hipGraphLaunch(c->bicg_graph, s));
        else
            for (int i = 0; i < c->bicg_chunk; i++) iterate();
        CMG_HIP_CHECK(hipMemcpyAsync(hi, fl, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        CMG_HIP_CHECK(hipStreamSynchronize(s));
        const int done = hi[0], its = hi[1];
        st->iterations = its;
        if (done == 0) {
            if (its >= maxit) break;
            continue;
        }
        // confirm with the true residual (the recurrence drifts from b - A x)
        launch_spmv(c->N, rp, ci, vals, c->fp32, x, k.r, s, c->spmv_lanes, c->spmv_layout);
        launch_waxpby(n, 1.0, b, -1.0, k.r, k.r, s);
        CMG_TRY(norm2(c, k.r, &rn));
        if (rn <= target) {
            st->converged = 1;
            break;
        }
        if (++restarts > 50 || its >= maxit || !std::isfinite(rn)) break;
        launch_fill_int(1, 0, k.done, s);
        launch_bicg_init(k, c->N, b, rtol_dev, active, s);   // restart from the current x
    }
    if (!st->converged) {   // final residual of an unconverged solve
        launch_spmv(c->N, rp, ci, vals, c->fp32, x, k.r, s, c->spmv_lanes, c->spmv_layout);
        launch_waxpby(n, 1.0, b, -1.0, k.r, k.r, s);
        CMG_TRY(norm2(c, k.r, &rn));
    }
    st->final_residual = rn;
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}

int krylov(cmg_ctx* c, const void* vals, const double* b, double* x, double rtol, int maxit, cmg_solve_stats* st) {
    if (c->spmv_lanes == 0) CMG_TRY(tune_spmv(c, vals, b));
    if (c->opt.krylov_method == 1) return bicgstab_device(c, vals, b, x, rtol, maxit, st);
    if (c->opt.krylov_method == 2) return bicgstab(c, vals, b, x, rtol, maxit, st);
    return gmres(c, vals, b, x, rtol, maxit, st);
}

// Cavity volume per pressure group at the device coordinates xd: the surface kernel's
// per-element contributions, summed in surface order on the host like
// CBCirculationCavity::CalcVolume (CBCirculationCavity.cpp:77-94).
int group_volumes(cmg_ctx* c, const double* xd, double* volumes) {
    std::vector<double> vp(c->S);
    if (c->S > 0) {
        launch_surface(surf_args(c, xd, true), false, c->stream);
        CMG_HIP_CHECK(hipMemcpyAsync(vp.data(), c->vol_partial.p, c->S * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
    }
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (c->S > 0 && c->h_sgroup.empty()) {
        c->h_sgroup.resize(c->S);
        CMG_HIP_CHECK(hipMemcpy(c->h_sgroup.data(), c->sgroup.p, c->S * sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    for (int g = 0; g < c->n_groups; g++) volumes[g] = 0.0;
    for (int64_t i = 0; i < c->S; i++) volumes[c->h_sgroup[i]] += vp[i];
    return CMG_SUCCESS;
}

}  // namespace

// ===========================================================================
// element-loop entry points
// ===========================================================================
extern "C" {

int cmg_residual_device(cmg_ctx* c, const double* x_dev, double time, int32_t include_pressure, double* f_dev) {
    if (!c || !x_dev || !f_dev) return fail("cmg_residual_device: null argument");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    int rc = eval_elements(c, x_dev, time, false, include_pressure != 0, nullptr);
    if (rc != CMG_SUCCESS) return rc;
    gather_residual(c, include_pressure != 0, f_dev);
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}

int cmg_residual_tangent_device(cmg_ctx* c, const double* x_dev, double time, int32_t include_pressure, double* f_dev,
                                void* values_dev) {
    if (!c || !x_dev) return fail("cmg_residual_tangent_device: null argument");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    int rc = eval_elements(c, x_dev, time, true, include_pressure != 0, nullptr);
    if (rc != CMG_SUCCESS) return rc;
    if (f_dev) gather_residual(c, include_pressure != 0, f_dev);
    if (values_dev) gather_tangent(c, include_pressure != 0, 0.0, values_dev);
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}

int cmg_element_kernel_device(cmg_ctx* c, const double* x_dev, double time, int32_t with_tangent) {
    if (!c || !x_dev) return fail("cmg_element_kernel_device: null argument");
    ElemArgs a = elem_args(c, x_dev, time);
    if (with_tangent)
        launch_element_residual_tangent(c->npe, c->fp32, a, c->stream, c->opt.element_variant);
    else
        launch_element_residual(c->npe, a, c->stream);
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}

int cmg_copy_device(cmg_ctx* c, const double* src, double* dst, int64_t n) {
    if (!c || !src || !dst || n < 0) return fail("cmg_copy_device: bad argument");
    launch_scale_copy(n, 1.0, src, dst, c->stream);
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}

int cmg_residual(cmg_ctx* c, const double* x, double time, int32_t include_pressure, double* f) {
    if (!c || !x || !f) return fail("cmg_residual: null argument");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    const size_t vb = c->n * sizeof(double);
    CMG_HIP_CHECK(hipMemcpyAsync(c->xeval.p, x, vb, hipMemcpyHostToDevice, c->stream));
    int rc = cmg_residual_device(c, c->xeval.as<double>(), time, include_pressure, c->r.as<double>());
    if (rc != CMG_SUCCESS) return rc;
    CMG_HIP_CHECK(hipMemcpyAsync(f, c->r.p, vb, hipMemcpyDeviceToHost, c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

int cmg_tangent(cmg_ctx* c, const double* x, double time, int32_t include_pressure, void* values) {
    if (!c || !x || !values) return fail("cmg_tangent: null argument");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    CMG_HIP_CHECK(hipMemcpyAsync(c->xeval.p, x, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = cmg_residual_tangent_device(c, c->xeval.as<double>(), time, include_pressure, nullptr, c->vals.p);
    if (rc != CMG_SUCCESS) return rc;
    const size_t vbytes = (size_t)c->nnzb * 9 * (c->fp32 ? sizeof(float) : sizeof(double));
    CMG_HIP_CHECK(hipMemcpyAsync(values, c->vals.p, vbytes, hipMemcpyDeviceToHost, c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

int cmg_element_residuals(cmg_ctx* c, const double* x, double time, double* fe) {
    if (!c || !x || !fe) return fail("cmg_element_residuals: null argument");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    CMG_HIP_CHECK(hipMemcpyAsync(c->xeval.p, x, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = eval_elements(c, c->xeval.as<double>(), time, false, false, nullptr);
    if (rc != CMG_SUCCESS) return rc;
    CMG_HIP_CHECK(hipMemcpyAsync(fe, c->fe.p, (size_t)c->E * 3 * c->npe * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

int cmg_element_tangents(cmg_ctx* c, const double* x, double time, void* Ke) {
    if (!c || !x || !Ke) return fail("cmg_element_tangents: null argument");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    CMG_HIP_CHECK(hipMemcpyAsync(c->xeval.p, x, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = eval_elements(c, c->xeval.as<double>(), time, true, false, nullptr);
    if (rc != CMG_SUCCESS) return rc;
    const int npe = c->npe, nd = 3 * npe;
    const size_t es = c->fp32 ? sizeof(float) : sizeof(double);
    std::vector<char> blk((size_t)c->E * 9 * npe * npe * es);
    CMG_HIP_CHECK(hipMemcpyAsync(blk.data(), c->Ke.p, blk.size(), hipMemcpyDeviceToHost, c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    // block layout [e][a][b][i][j] -> Jacobian layout [e][3a+i][3b+j]
    for (int64_t e = 0; e < c->E; e++)
        for (int a = 0; a < npe; a++)
            for (int b = 0; b < npe; b++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        const size_t src = (size_t)e * 9 * npe * npe + (size_t)(a * npe + b) * 9 + 3 * i + j;
                        const size_t dst = (size_t)e * nd * nd + (size_t)(3 * a + i) * nd + 3 * b + j;
                        std::memcpy(static_cast<char*>(Ke) + dst * es, blk.data() + src * es, es);
                    }
    return CMG_SUCCESS;
}

int cmg_surface_forces(cmg_ctx* c, const double* x, const double* pressures, double* f) {
    if (!c || !x || !f) return fail("cmg_surface_forces: null argument");
    if (pressures) CMG_TRY(cmg_set_pressures(c, pressures));
    const size_t vb = c->n * sizeof(double);
    CMG_HIP_CHECK(hipMemcpyAsync(c->xeval.p, x, vb, hipMemcpyHostToDevice, c->stream));
    launch_fill((int64_t)c->E * 3 * c->npe, 0.0, c->fe.as<double>(), c->stream);
    if (c->S > 0) launch_surface(surf_args(c, c->xeval.as<double>(), false), false, c->stream);
    gather_residual(c, true, c->r.as<double>());
    CMG_HIP_CHECK(hipMemcpyAsync(f, c->r.p, vb, hipMemcpyDeviceToHost, c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

int cmg_cavity_volumes(cmg_ctx* c, const double* x, double* volumes) {
    if (!c || !volumes) return fail("cmg_cavity_volumes: null argument");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    const double* xd = c->nodes.as<double>();
    if (x) {
        CMG_HIP_CHECK(hipMemcpyAsync(c->xeval.p, x, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        xd = c->xeval.as<double>();
    }
    return group_volumes(c, xd, volumes);
}

int cmg_element_export(cmg_ctx* c, const double* x, double* out) {
    if (!c || !out) return fail("cmg_element_export: null argument");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    const double* xd = c->nodes.as<double>();
    if (x) {
        CMG_HIP_CHECK(hipMemcpyAsync(c->xeval.p, x, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        xd = c->xeval.as<double>();
    }
    DevBuf buf;
    CMG_TRY(alloc(buf, (size_t)18 * c->E * sizeof(double)));
    launch_element_export(c->npe, elem_args(c, xd, 0.0), c->Xref.as<double>(), buf.as<double>(), c->stream);
    CMG_HIP_CHECK(hipGetLastError());
    CMG_HIP_CHECK(hipMemcpyAsync(out, buf.p, (size_t)18 * c->E * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

int cmg_mass_matrix(cmg_ctx* c, double* values) {
    if (!c || !values) return fail("cmg_mass_matrix: null argument");
    // rows of constrained components are zero (CBElementSolidT10.cpp:386-389)
    for (int64_t r = 0; r < c->N; r++)
        for (int64_t k = c->h_rowptr[r]; k < c->h_rowptr[r + 1]; k++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    values[9 * k + 3 * i + j] = (i == j && !((c->h_bc[r] >> i) & 1)) ? c->h_mval[k] : 0.0;
    return CMG_SUCCESS;
}

int cmg_spmv_layout(cmg_ctx* c, const double* values, const double* x, double* y, int32_t lanes, int32_t layout) {
    if (!c || !values || !x || !y) return fail("cmg_spmv: null argument");
    if (layout < kSpmvBlock || layout > kSpmvFlat) return fail("cmg_spmv_layout: layout must be 0, 1 or 2");
    if (layout != kSpmvFlat && lanes != 4 && lanes != 8 && lanes != 16)
        return fail("cmg_spmv_layout: lanes must be 4, 8 or 16 for layouts 0 and 1");
    if (layout == kSpmvFlat && lanes != 16 && lanes != 32 && lanes != 64)
        return fail("cmg_spmv_layout: lanes must be 16, 32 or 64 for layout 2");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    const size_t vb = c->n * sizeof(double);
    DevBuf v64;   // padded for the chunked layout
    CMG_TRY(alloc(v64, (size_t)(c->nnzb + kSpmvPadBlocks) * 9 * sizeof(double)));
    launch_fill((c->nnzb + kSpmvPadBlocks) * 9, 0.0, v64.as<double>(), c->stream);
    CMG_HIP_CHECK(hipMemcpyAsync(v64.p, values, (size_t)c->nnzb * 9 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CMG_HIP_CHECK(hipMemcpyAsync(c->tmpA.p, x, vb, hipMemcpyHostToDevice, c->stream));
    launch_spmv(c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), v64.p, false, c->tmpA.as<double>(),
                c->tmpB.as<double>(), c->stream, lanes, layout);
    CMG_HIP_CHECK(hipMemcpyAsync(y, c->tmpB.p, vb, hipMemcpyDeviceToHost, c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

int cmg_spmv_tuning(cmg_ctx* c, double* ms, int32_t* layout, int32_t* lanes) {
    if (!c) return fail("null context");
    if (ms)
        for (int k = 0; k < 9; k++) ms[k] = c->spmv_ms[k];
    if (layout) *layout = c->spmv_lanes ? c->spmv_layout : 0;
    if (lanes) *lanes = c->spmv_lanes;
    return CMG_SUCCESS;
}

int cmg_spmv(cmg_ctx* c, const double* values, const double* x, double* y) {
    return cmg_spmv_layout(c, values, x, y, 8, 0);
}

int cmg_linear_solve(cmg_ctx* c, const double* values, const double* b, double* x, double rtol, cmg_solve_stats* st) {
    if (!c || !values || !b || !x) return fail("cmg_linear_solve: null argument");
    const size_t vb = c->n * sizeof(double);
    const size_t vbytes = (size_t)c->nnzb * 9 * (c->fp32 ? sizeof(float) : sizeof(double));
    CMG_HIP_CHECK(hipMemcpyAsync(c->vals.p, values, vbytes, hipMemcpyHostToDevice, c->stream));
    CMG_HIP_CHECK(hipMemcpyAsync(c->r.p, b, vb, hipMemcpyHostToDevice, c->stream));
    CMG_HIP_CHECK(hipMemcpyAsync(c->delta.p, x, vb, hipMemcpyHostToDevice, c->stream));
    cmg_solve_stats local;
    if (!st) st = &local;
    c->precond_built = false;   // a standalone solve always builds from its own matrix
    CMG_TRY(krylov(c, c->vals.p, c->r.as<double>(), c->delta.as<double>(), rtol, c->opt.krylov_maxit, st));
    CMG_HIP_CHECK(hipMemcpyAsync(x, c->delta.p, vb, hipMemcpyDeviceToHost, c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return st->converged ? CMG_SUCCESS : CMG_FAILED;
}

// ===========================================================================
// solver steps
// ===========================================================================
int cmg_set_state(cmg_ctx* c, const double* nodes, const double* vel, const double* acc) {
    if (!c) return fail("null context");
    c->trial_pending = false;
    const size_t vb = c->n * sizeof(double);
    if (nodes) {
        CMG_HIP_CHECK(hipMemcpyAsync(c->nodes.p, nodes, vb, hipMemcpyHostToDevice, c->stream));
        c->damping_valid = false;
    }
    if (vel) CMG_HIP_CHECK(hipMemcpyAsync(c->vel.p, vel, vb, hipMemcpyHostToDevice, c->stream));
    if (acc) CMG_HIP_CHECK(hipMemcpyAsync(c->acc.p, acc, vb, hipMemcpyHostToDevice, c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

int cmg_get_state(cmg_ctx* c, double* nodes, double* vel, double* acc, double* disp) {
    if (!c) return fail("null context");
    const size_t vb = c->n * sizeof(double);
    if (nodes) CMG_HIP_CHECK(hipMemcpyAsync(nodes, c->nodes.p, vb, hipMemcpyDeviceToHost, c->stream));
    if (vel) CMG_HIP_CHECK(hipMemcpyAsync(vel, c->vel.p, vb, hipMemcpyDeviceToHost, c->stream));
    if (acc) CMG_HIP_CHECK(hipMemcpyAsync(acc, c->acc.p, vb, hipMemcpyDeviceToHost, c->stream));
    if (disp) CMG_HIP_CHECK(hipMemcpyAsync(disp, c->d.p, vb, hipMemcpyDeviceToHost, c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return CMG_SUCCESS;
}

int cmg_set_initial_guess(cmg_ctx* c, const double* inc) {
    if (!c || !inc) return fail("cmg_set_initial_guess: null argument");
    CMG_HIP_CHECK(hipMemcpyAsync(c->override_guess.p, inc, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->has_override = true;
    c->override_checked = false;
    return CMG_SUCCESS;
}

}  // extern "C"

namespace {

// smallest backtracking factor when a Newton step inverts an element (the reference's
// line search goes down to 1e-12 on a domain error)
constexpr double kMinDomainLambda = 1.0 / (1 << 20);
// SNES divergence tolerance (PETSc's default divtol): ||r|| > 1e4 ||r_0|| ends the
// step as diverged (SNES_DIVERGED_DTOL)
constexpr double kNewtonDivTol = 1e4;
// a Newton iteration that keeps more than this fraction of the residual norm counts
// towards cmg_options.newton_stall
constexpr double kStallRatio = 0.9;

// Rayleigh damping matrix C = alpha M + beta_R K(x_n) for the configuration in nodes
// (CBSolverNewmarkBeta::CalcDampingMatrix, :657-702, called at the start of each new
// time step, :857-866): K is the element tangent without surface terms, evaluated at
// the step's time, in FP64 also in mixed precision (C w cancels strongly for the
// near-incompressible wall, so FP32 entries would cost ~1e-4 in the damping force);
// rebuilt only after nodes changed
int build_damping(cmg_ctx* c, double time) {
    if (!c->damping || c->damping_valid) return CMG_SUCCESS;
    hipStream_t s = c->stream;
    if (c->opt.rayleigh_beta != 0.0) {
        int corrupt = -1;
        const int rc = eval_elements(c, c->nodes.as<double>(), time, true, false, &corrupt, true);
        if (rc != CMG_SUCCESS) return rc;
        if (c->S > 0) launch_fill((int64_t)c->S * 9 * c->nps * c->nps, 0.0, c->Ks.as<double>(), s);
        // C is built from the tangent in the reference's layout: for T10 its finite-difference
        // Jacobian is assembled transposed (CBElementSolidT10.cpp:559-565), so the reference's
        // damping force is beta_R K^T w (CBSolverNewmarkBeta.cpp:657-668); Newton keeps K
        const bool flip = c->npe == 10 && !c->maps_transposed;
        launch_gather_matrix(gather_maps(c), c->Ke.p, false, c->Ks.as<double>(), nullptr, 0.0, c->bcmask.as<uint8_t>(),
                             c->cvals.p, false, flip, s);
    }
    launch_rayleigh_finish(c->nnzb, c->blkrow.as<int32_t>(), c->colind.as<int32_t>(), c->bcmask.as<uint8_t>(),
                           c->mval.as<double>(), c->opt.rayleigh_beta, c->opt.rayleigh_alpha, c->cvals.as<double>(), s);
    CMG_HIP_CHECK(hipGetLastError());
    c->damping_valid = true;
    return CMG_SUCCESS;
}

// r += C (v~ + cd (d - d~)) on free DOFs, cd = gamma / (beta dt) (:384-390)
void add_damping_force(cmg_ctx* c, double cd, double* r) {
    if (cd == 0.0 || !c->damping) return;
    hipStream_t s = c->stream;
    double* w = c->tmpB.as<double>();
    double* cw = c->tmpA.as<double>();
    launch_waxpby(c->n, 1.0, c->d.as<double>(), -1.0, c->dpred.as<double>(), w, s);
    launch_axpby(c->n, 1.0, c->vpred.as<double>(), cd, w, s);
    launch_spmv(c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), c->cvals.p, false, w, cw, s, 16, kSpmvBlock);
    launch_zero_bc(c->N, c->bcmask.as<uint8_t>(), cw, s);
    launch_axpby(c->n, 1.0, cw, 1.0, r, s);
}

// ||r(d)|| of the Newmark residual at the current increment d (infinite when an
// element inverts)
int residual_norm_at(cmg_ctx* c, double time, double cm, double cd, double* out) {
    hipStream_t s = c->stream;
    launch_waxpby(c->n, 1.0, c->nodes.as<double>(), 1.0, c->d.as<double>(), c->xeval.as<double>(), s);
    int corrupt = -1;
    if (eval_elements(c, c->xeval.as<double>(), time, false, true, &corrupt) != CMG_SUCCESS) {
        *out = INFINITY;
        return CMG_SUCCESS;
    }
    gather_residual(c, true, c->r.as<double>());
    if (cm != 0.0)
        launch_add_mass_term(c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), c->mval.as<double>(), cm,
                             c->d.as<double>(), c->dpred.as<double>(), c->bcmask.as<uint8_t>(), c->r.as<double>(), s);
    add_damping_force(c, cd, c->r.as<double>());
    return norm2(c, c->r.as<double>(), out);
}

// Newton on r(d) = f_int(nodes + d) + f_p(nodes + d) [+ cm M (d - d~)] = 0, the
// SNES newtonls loop of CBSolverNewmarkBeta.cpp:890 / CBSolverEquilibrium.cpp:287 with
// convergence on abstol / rtol / stol = Solver.Precision, a fresh exact tangent
// per iteration, an inexact Krylov solve (Eisenstat-Walker forcing) and
// backtracking on the residual norm.
int newton(cmg_ctx* c, double time, double cm, double cd, cmg_step_stats* st) {
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    double* d = c->d.as<double>();
    double* r = c->r.as<double>();
    double* delta = c->delta.as<double>();
    double* xe = c->xeval.as<double>();
    const double t0 = now_s();
    double t_tan = 0.0, t_kry = 0.0;
    double r0 = 0.0, rprev = 0.0, rnorm = 0.0, snorm = 0.0, dnorm = 0.0;
    int kry_total = 0;
    st->converged_reason = 0;
    st->corrupt_element = -1;
    // forcing: eta_0 = 0.1, eta_max = 0.5 (timed on whole beats against 1e-2 / 0.1:
    // fewer Krylov iterations for a few more Newton iterations)
    double eta = 1e-1;
    double lambda = 1.0;
    bool zero_tried = false;
    int stall = 0;
    c->precond_built = false;   // a new nonlinear solve builds its preconditioner afresh
    int it = 0;
    for (;; it++) {
        const double ta = now_s();
        launch_waxpby(n, 1.0, c->nodes.as<double>(), 1.0, d, xe, s);
        int corrupt = -1;
        // residual only: the tangent is evaluated below once the iterate is accepted and a
        // linear solve follows (not at the converged iterate, not on a backtracked one)
        int rc = eval_elements(c, xe, time, false, true, &corrupt);
        if (rc != CMG_SUCCESS) {
            st->corrupt_element = corrupt;
            if (c->trace) std::fprintf(stderr, "[newton]   element %d inverted at lambda=%.4g\n", corrupt, lambda);
            if (it > 0 && lambda > kMinDomainLambda) {
                // step into an inverted state: backtrack (SNESSetFunctionDomainError analogue)
                lambda *= 0.5;
                launch_axpby(n, -lambda, delta, 1.0, d, s);
                it--;
                continue;
            }
            if (it == 0 && !zero_tried) {
                // the extrapolated initial guess inverts an element: start from a zero increment
                // (the current configuration is valid)
                zero_tried = true;
                st->guess_source = 4;
                launch_fill(n, 0.0, d, s);
                it--;
                continue;
            }
            st->converged_reason = -6;
            st->newton_iterations = it;
            st->krylov_iterations = kry_total;
            return CMG_FAILED;
        }
        gather_residual(c, true, r);
        if (cm != 0.0)
            launch_add_mass_term(c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), c->mval.as<double>(), cm, d,
                                 c->dpred.as<double>(), c->bcmask.as<uint8_t>(), r, s);
        add_damping_force(c, cd, r);
        CMG_TRY(norm2(c, r, &rnorm));
        t_tan += now_s() - ta;
        if (!std::isfinite(rnorm)) {
            if (it > 0 && lambda > kMinDomainLambda) {   // overflow in the laws: a domain error too
                lambda *= 0.5;
                launch_axpby(n, -lambda, delta, 1.0, d, s);
                it--;
                continue;
            }
            st->converged_reason = -4;
            break;
        }
        if (it == 0) {
            r0 = rnorm;
            st->initial_residual_norm = r0;
        } else if (rnorm > rprev && lambda > 1.0 / 64 && rnorm > c->opt.newton_atol) {
            // no decrease: backtrack along the last Newton direction
            if (c->trace) std::fprintf(stderr, "[newton]   backtrack lambda=%.4g |r|=%.3e\n", lambda, rnorm);
            lambda *= 0.5;
            launch_axpby(n, -lambda, delta, 1.0, d, s);
            it--;
            continue;
        }
        if (rnorm < c->opt.newton_atol) {
            st->converged_reason = 2;
            break;
        }
        if (it > 0 && rnorm > kNewtonDivTol * r0) {
            st->converged_reason = -9;   // SNES_DIVERGED_DTOL
            break;
        }
        if (it > 0 && rnorm < c->opt.newton_rtol * r0) {
            st->converged_reason = 3;
            break;
        }
        // a short step ends Newton only after a residual decrease: the reference's
        // backtracking line search accepts no step that raises ||r||, where the
        // loop above gives up backtracking at 1/64 and carries on
        if (it > 0 && rnorm <= rprev && snorm * lambda < c->opt.newton_stol * dnorm) {
            st->converged_reason = 4;
            break;
        }
        if (it >= c->opt.newton_maxit) {
            st->converged_reason = -5;   // SNES_DIVERGED_MAX_IT
            break;
        }
        // stalled: less than 10 % reduction per iteration, newton_stall times in a row
        if (it > 0) stall = (rnorm > kStallRatio * rprev) ? stall + 1 : 0;
        if (c->opt.newton_stall > 0 && stall >= c->opt.newton_stall) {
            st->converged_reason = -7;
            break;
        }
        const double eta_max = c->opt.krylov_eta_max > 0.0 ? c->opt.krylov_eta_max : 0.5;
        if (it == 0) eta = std::min(eta, eta_max);
        if (it > 0) eta = std::min(eta_max, std::max(c->opt.krylov_rtol, 0.9 * (rnorm / rprev) * (rnorm / rprev)));
        // no oversolving: a linear residual below half the nonlinear tolerance is all the
        // next iterate needs (Kelley 1995, eq. 6.20: eta >= 0.5 tau_t / ||r||, tau_t =
        // atol + rtol ||r_0||)
        const double eta_k =
            std::min(eta_max, std::max(eta, 0.5 * (c->opt.newton_atol + c->opt.newton_rtol * r0) / rnorm));
        rprev = rnorm;
        const double tb = now_s();
        rc = eval_elements(c, xe, time, true, true, &corrupt);
        if (rc != CMG_SUCCESS) {   // the residual kernel accepted this state: not expected
            st->corrupt_element = corrupt;
            st->converged_reason = -6;
            break;
        }
        gather_tangent(c, true, cm, c->vals.p);
        if (cd != 0.0 && c->damping) launch_add_scaled_values(c->nnzb, cd, c->cvals.as<double>(), c->vals.p, c->fp32, s);
        t_tan += now_s() - tb;
        const double tk = now_s();
        launch_scale_copy(n, -1.0, r, c->dsave.as<double>(), s);
        launch_fill(n, 0.0, delta, s);
        cmg_solve_stats ls;
        CMG_TRY(krylov(c, c->vals.p, c->dsave.as<double>(), delta, eta_k, c->opt.krylov_maxit, &ls));
        kry_total += ls.iterations;
        t_kry += now_s() - tk;
        if (c->trace)
            std::fprintf(stderr, "[newton] t=%.6g it=%d |r|=%.3e eta=%.2e krylov=%d lin.res=%.2e/%.2e\n", time, it,
                         rnorm, eta_k, ls.iterations, ls.final_residual, ls.initial_residual);
        lambda = 1.0;
        launch_axpby(n, 1.0, delta, 1.0, d, s);
        CMG_TRY(norm2(c, delta, &snorm));
        CMG_TRY(norm2(c, d, &dnorm));
        if (!std::isfinite(ls.final_residual) || !(snorm > 0.0) || !std::isfinite(snorm)) {
            // the Krylov solve broke down without a usable direction (a zero or
            // non-finite update): SNES_DIVERGED_LINEAR_SOLVE, never a short step
            st->converged_reason = -3;
            break;
        }
    }
    st->newton_iterations = it;
    st->krylov_iterations = kry_total;
    st->residual_norm = rnorm;
    st->step_norm = snorm;
    st->tangent_seconds = t_tan;
    st->krylov_seconds = t_kry;
    st->wall_seconds = now_s() - t0;
    return st->converged_reason > 0 ? CMG_SUCCESS : CMG_FAILED;
}

}  // namespace

extern "C" {

int cmg_static_step(cmg_ctx* c, double time, double dt, cmg_step_stats* stats) {
    if (!c) return fail("null context");
    c->trial_pending = false;
    CMG_HIP_CHECK(hipSetDevice(c->device));
    cmg_step_stats local;
    cmg_step_stats* st = stats ? stats : &local;
    std::memset(st, 0, sizeof(*st));
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    double* d = c->d.as<double>();
    // CBSolverEquilibrium::UpdateInitialGuess (:234-263)
    if (time > c->time_last) {
        if (c->step_last > 0.0) {
            launch_scale_copy(n, 1. / c->step_last, d, c->guess.as<double>(), s);
            launch_scale_copy(n, 1. * dt / c->step_last, d, d, s);
        } else {
            launch_fill(n, 0.0, d, s);
        }
    } else if (time == c->time_last) {
        launch_fill(n, 0.0, d, s);
    } else {
        launch_scale_copy(n, dt, c->guess.as<double>(), d, s);
    }
    c->time_last = time;
    c->step_last = dt;
    if (c->has_override) {
        launch_scale_copy(n, 1.0, c->override_guess.as<double>(), d, s);
        c->has_override = false;
        st->guess_source = 1;
    }
    launch_zero_bc(c->N, c->bcmask.as<uint8_t>(), d, s);
    int rc = newton(c, time, 0.0, 0.0, st);
    if (rc != CMG_SUCCESS) {
        launch_fill(n, 0.0, d, s);   // :303-306
        CMG_HIP_CHECK(hipStreamSynchronize(s));
        return CMG_FAILED;
    }
    launch_axpby(n, 1.0, d, 1.0, c->nodes.as<double>(), s);   // :359
    c->damping_valid = false;
    CMG_TRY(land17_commit(c, c->nodes.as<double>(), time));
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    return CMG_SUCCESS;
}

int cmg_newmark_solve(cmg_ctx* c, double time, double dt, int32_t guess, cmg_step_stats* stats) {
    if (!c) return fail("null context");
    if (!(dt > 0.0)) return fail("cmg_newmark_solve: dt must be positive");
    if (guess != 0 && guess != 1) return fail("cmg_newmark_solve: guess must be 0 (predictor) or 1 (last trial)");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    cmg_step_stats local;
    cmg_step_stats* st = stats ? stats : &local;
    std::memset(st, 0, sizeof(*st));
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    const double beta = c->opt.newmark_beta, gamma = c->opt.newmark_gamma;
    // a pending trial of the same step size is a better start than the predictor when
    // only a load (the cavity pressure) changed since it was solved
    const bool from_trial = guess == 1 && c->trial_pending && c->trial_dt == dt;
    c->trial_pending = false;
    // predictor (:845-855): the initial guess is d~
    launch_newmark_predict(n, dt, beta, gamma, c->vel.as<double>(), c->acc.as<double>(), c->dpred.as<double>(),
                           c->vpred.as<double>(), s);
    const double cm = 1.0 / (beta * dt * dt);
    const double cd = gamma / (beta * dt);
    {
        const int rc = build_damping(c, time);
        if (rc != CMG_SUCCESS) {
            st->converged_reason = -6;
            return CMG_FAILED;
        }
    }
    double* d = c->d.as<double>();
    if (from_trial) {
        st->guess_source = 5;   // d still holds the last trial increment
    } else if (c->has_override && c->override_checked) {
        // proposed guess (learned warm start): kept only if its residual is below the predictor's
        double r_pred, r_prop;
        launch_scale_copy(n, 1.0, c->dpred.as<double>(), d, s);
        launch_zero_bc(c->N, c->bcmask.as<uint8_t>(), d, s);
        CMG_TRY(residual_norm_at(c, time, cm, cd, &r_pred));
        launch_scale_copy(n, 1.0, c->override_guess.as<double>(), d, s);
        launch_zero_bc(c->N, c->bcmask.as<uint8_t>(), d, s);
        CMG_TRY(residual_norm_at(c, time, cm, cd, &r_prop));
        st->guess_source = (r_prop < r_pred) ? 2 : 3;
        if (st->guess_source == 3) launch_scale_copy(n, 1.0, c->dpred.as<double>(), d, s);
        st->guess_residual_ratio = r_pred > 0.0 ? r_prop / r_pred : 0.0;
    } else if (c->has_override) {
        launch_scale_copy(n, 1.0, c->override_guess.as<double>(), d, s);
        st->guess_source = 1;
    } else {
        launch_scale_copy(n, 1.0, c->dpred.as<double>(), d, s);
    }
    if (!from_trial) {
        c->has_override = false;
        c->override_checked = false;
    }
    launch_zero_bc(c->N, c->bcmask.as<uint8_t>(), d, s);
    int rc = newton(c, time, cm, cd, st);
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    if (rc != CMG_SUCCESS) return CMG_FAILED;
    c->trial_pending = true;
    c->trial_time = time;
    c->trial_dt = dt;
    return CMG_SUCCESS;
}

int cmg_trial_volumes(cmg_ctx* c, double* volumes) {
    if (!c || !volumes) return fail("cmg_trial_volumes: null argument");
    if (!c->trial_pending) return fail("cmg_trial_volumes: no solved step is pending");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    launch_scale_copy(c->n, 1.0, c->nodes.as<double>(), c->xeval.as<double>(), c->stream);
    launch_axpby(c->n, 1.0, c->d.as<double>(), 1.0, c->xeval.as<double>(), c->stream);
    return group_volumes(c, c->xeval.as<double>(), volumes);
}

int cmg_commit_step(cmg_ctx* c) {
    if (!c) return fail("null context");
    if (!c->trial_pending) return fail("cmg_commit_step: no solved step is pending");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    const double dt = c->trial_dt, beta = c->opt.newmark_beta, gamma = c->opt.newmark_gamma;
    // corrector (:984-1015)
    launch_newmark_correct(n, dt, beta, gamma, c->d.as<double>(), c->dpred.as<double>(), c->vpred.as<double>(),
                           c->vel.as<double>(), c->acc.as<double>(), s);
    // kinetic energy v^T M v / 2 and damping dissipation (C v) . d (:990-1003), C
    // still that of the step's start configuration
    {
        double* mv = c->tmpA.as<double>();
        launch_fill(n, 0.0, mv, s);
        launch_fill(n, 0.0, c->tmpB.as<double>(), s);
        launch_add_mass_term(c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), c->mval.as<double>(), 1.0,
                             c->vel.as<double>(), c->tmpB.as<double>(), c->bcmask.as<uint8_t>(), mv, s);
        double h = 0.0;
        CMG_TRY(multidot(c, 1, mv, 0, c->vel.as<double>(), &h));
        c->kinetic_energy = 0.5 * h;
        if (c->damping) {
            launch_spmv(c->N, c->rowptr.as<int64_t>(), c->colind.as<int32_t>(), c->cvals.p, false,
                        c->vel.as<double>(), mv, s, 16, kSpmvBlock);
            CMG_TRY(multidot(c, 1, mv, 0, c->d.as<double>(), &h));
            c->damping_energy += h;
        }
    }
    launch_axpby(n, 1.0, c->d.as<double>(), 1.0, c->nodes.as<double>(), s);
    c->damping_valid = false;
    CMG_TRY(land17_commit(c, c->nodes.as<double>(), c->trial_time));
    // increment history (features of a learned warm start)
    launch_scale_copy(n, 1.0, c->inc1.as<double>(), c->inc2.as<double>(), s);
    launch_scale_copy(n, 1.0, c->d.as<double>(), c->inc1.as<double>(), s);
    c->trial_pending = false;
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    return CMG_SUCCESS;
}

int cmg_newmark_energies(cmg_ctx* c, double* kinetic, double* damping_dissipated) {
    if (!c) return fail("null context");
    if (kinetic) *kinetic = c->kinetic_energy;
    if (damping_dissipated) *damping_dissipated = c->damping_energy;
    return CMG_SUCCESS;
}

int cmg_newmark_step(cmg_ctx* c, double time, double dt, cmg_step_stats* stats) {
    int rc = cmg_newmark_solve(c, time, dt, 0, stats);
    if (rc != CMG_SUCCESS) return rc;
    return cmg_commit_step(c);
}

int cmg_set_land17_parameters(cmg_ctx* c, int32_t material, const double* params) {
    if (!c || !params) return fail("cmg_set_land17_parameters: null argument");
    if (!c->l17) return fail("cmg_set_land17_parameters: no material of this context uses the Land17 model");
    if (material < 0 || material >= c->n_mats) return fail("cmg_set_land17_parameters: material out of range");
    if (!(params[8] > 0.0 && params[8] < 1.0 && params[9] > 0.0 && params[9] < 1.0 && params[23] > 0.0 &&
          params[19] > 0.0 && params[20] > 0.0))
        return fail("cmg_set_land17_parameters: rs and rw must lie in (0, 1), cycleLength, eta_l and eta_s must "
                    "be positive");
    if (params[24] != 0.0 && params[24] != 1.0)
        return fail("cmg_set_land17_parameters: calcium transient must be 0 (Coppini) or 1 (Brixius)");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    std::copy(params, params + kLand17Params, c->h_l17_params.begin() + (size_t)material * kLand17Params);
    CMG_HIP_CHECK(hipMemcpyAsync(c->l17_params.p, c->h_l17_params.data(), c->h_l17_params.size() * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream));
    return land17_reset_state(c);
}

int cmg_get_land17_state(cmg_ctx* c, double* state, double* commit_time) {
    if (!c) return fail("null context");
    if (!c->l17) return fail("cmg_get_land17_state: no material of this context uses the Land17 model");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    if (state)
        CMG_HIP_CHECK(hipMemcpyAsync(state, c->l17_state.p, (size_t)8 * c->E * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (commit_time) *commit_time = c->l17_t_commit;
    return CMG_SUCCESS;
}

int cmg_set_land17_state(cmg_ctx* c, const double* state, double commit_time) {
    if (!c || !state) return fail("cmg_set_land17_state: null argument");
    if (!c->l17) return fail("cmg_set_land17_state: no material of this context uses the Land17 model");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    CMG_HIP_CHECK(hipMemcpyAsync(c->l17_state.p, state, (size_t)8 * c->E * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
    CMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->l17_t_commit = commit_time;
    return CMG_SUCCESS;
}

}  // extern "C"
