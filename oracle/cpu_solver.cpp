// CPU restatement of the reference's solver steps with its finite-difference element
// tangent -- TEST INFRASTRUCTURE ONLY (the CPU baseline of the benchmark and oracle
// pins; the product library never links or calls it).
//
//   orc_cpu_newmark_step   CBSolverNewmarkBeta::SolverStep (CBSolverNewmarkBeta.cpp:829-1019):
//                          predictor d~ = dt v + dt^2/2 (1 - 2 beta) a (:845-855), Newton on
//                          f_int(x + d) + f_p(x + d) + M (d - d~) / (beta dt^2) = 0 (:340-419),
//                          corrector (:978-1015)
//   orc_cpu_static_step    CBSolverEquilibrium::UpdateInitialGuess + SolverStep (:234-363)
//   tangent                every element's central-difference Jacobian exactly as
//                          CBElementSolidT10::CalcNodalForcesJacobian (:511-567; T4 :294-351)
//                          with h = CBSolver::CalcFiniteDifferencesEpsilon (CBSolver.cpp:1610-1625):
//                          (1 + ||a||_2) Solver.Epsilon, a = the Newton iterate's increment for
//                          static steps (CBSolverEquilibrium.cpp:114), a zero vector for Newmark
//                          (CBSolverNewmarkBeta.cpp:568-572); assembled in the reference's layout
//                          (T10 row = perturbed DOF, CBElementSolidT10.cpp:559-565), + M/(beta dt^2),
//                          identity on constrained DOFs (CBSolver.cpp:413-435); the T3 pressure
//                          Jacobian is zero as coded (CBElementSurfaceT3.cpp:268-269, 350-369)
//   linear solve           the reference factorises with MUMPS LU (CBSolverNewmarkBeta.cpp:176-192),
//                          which this image does not have: BiCGStab with 3x3 nodal block-Jacobi
//                          right preconditioning and Eisenstat-Walker forcing (OpenMP), the same
//                          Krylov method as the GPU path's block-Jacobi option.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct orc_cpu_solver {
    int npe = 10, nps = 0;
    int64_t E = 0, N = 0, S = 0, n = 0, nnzb = 0;
    std::vector<int32_t> conn, sconn, sgroup, matidx;
    std::vector<double> dNdX, bases, w, tact;
    std::vector<orc_material> mats;
    std::vector<uint8_t> bcmask;
    std::vector<int64_t> rowptr, inc_ptr;
    std::vector<int32_t> colind, inc, slot;   // slot[(e npe + a) npe + b] of block (node a, node b)
    std::vector<double> mval;                 // scalar mass per block (consistent)
    std::vector<double> nodes, vel, acc, d, dpred, pressures;
    std::vector<double> fe, Ke, vals, dinv;
    // static-step bookkeeping (UpdateInitialGuess)
    double time_last = 0.0, step_last = 0.0;
    std::vector<double> guess;
};

namespace {

void gather_elem(const orc_cpu_solver* s, int64_t e, const double* x, double* xe, uint8_t* bc) {
    for (int a = 0; a < s->npe; a++) {
        const int32_t nd = s->conn[e * s->npe + a];
        for (int c = 0; c < 3; c++) {
            xe[3 * a + c] = x[3 * (int64_t)nd + c];
            bc[3 * a + c] = (s->bcmask[nd] >> c) & 1;
        }
    }
}

// f = assembled internal + pressure forces at x, constrained components zero;
// ORC_CORRUPT_ELEMENT if an element's det F <= 0
int residual(orc_cpu_solver* s, const double* x, double time, double* f) {
    const int nd = 3 * s->npe;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t e = 0; e < s->E; e++) {
        double xe[30];
        uint8_t bc[30];
        gather_elem(s, e, x, xe, bc);
        const orc_material* m = &s->mats[s->matidx[e]];
        const int rc = s->npe == 10
                           ? orc_t10_residual(xe, &s->dNdX[150 * e], &s->bases[45 * e], s->w[e], bc, m, time,
                                              s->tact[e], &s->fe[nd * e])
                           : orc_t4_residual(xe, &s->dNdX[12 * e], &s->bases[9 * e], s->w[e], bc, m, time,
                                             s->tact[e], &s->fe[nd * e]);
        bad += rc != ORC_SUCCESS;
    }
    if (bad) return ORC_CORRUPT_ELEMENT;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < s->N; r++) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int64_t k = s->inc_ptr[r]; k < s->inc_ptr[r + 1]; k++) {
            const double* p = &s->fe[3 * (int64_t)s->inc[k]];
            a0 += p[0];
            a1 += p[1];
            a2 += p[2];
        }
        f[3 * r] = a0;
        f[3 * r + 1] = a1;
        f[3 * r + 2] = a2;
    }
    // surface pressure forces, in surface order
    const int ns = 3 * s->nps;
    for (int64_t si = 0; si < s->S; si++) {
        double xs[18], fs[18];
        uint8_t bc[18];
        for (int a = 0; a < s->nps; a++) {
            const int32_t nd2 = s->sconn[si * s->nps + a];
            for (int c = 0; c < 3; c++) {
                xs[3 * a + c] = x[3 * (int64_t)nd2 + c];
                bc[3 * a + c] = (s->bcmask[nd2] >> c) & 1;
            }
        }
        const double p = s->pressures[s->sgroup[si]];
        if (s->nps == 3)
            orc_t3_pressure_forces(xs, bc, p, fs);
        else
            orc_t6_pressure_forces(xs, bc, p, fs);
        for (int a = 0; a < s->nps; a++)
            for (int c = 0; c < 3; c++) f[3 * (int64_t)s->sconn[si * s->nps + a] + c] += fs[3 * a + c];
        (void)ns;
    }
    for (int64_t r = 0; r < s->N; r++)
        for (int c = 0; c < 3; c++)
            if ((s->bcmask[r] >> c) & 1) f[3 * r + c] = 0.0;
    return ORC_SUCCESS;
}

// block values = FD element Jacobians (reference layout) + cm M, identity on constrained DOFs
int tangent(orc_cpu_solver* s, const double* x, double time, double h, double cm) {
    const int npe = s->npe, nd = 3 * npe;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t e = 0; e < s->E; e++) {
        double xe[30];
        uint8_t bc[30];
        gather_elem(s, e, x, xe, bc);
        const orc_material* m = &s->mats[s->matidx[e]];
        const int rc = npe == 10 ? orc_t10_fd_tangent(xe, &s->dNdX[150 * e], &s->bases[45 * e], s->w[e], bc, m, time,
                                                      s->tact[e], h, &s->Ke[(int64_t)nd * nd * e])
                                 : orc_t4_fd_tangent(xe, &s->dNdX[12 * e], &s->bases[9 * e], s->w[e], bc, m, time,
                                                     s->tact[e], h, &s->Ke[(int64_t)nd * nd * e]);
        bad += rc != ORC_SUCCESS;
    }
    if (bad) return ORC_CORRUPT_ELEMENT;
    std::fill(s->vals.begin(), s->vals.end(), 0.0);
    // scatter-add in element order (ADD_VALUES of the adapter, CBElementAdapter.cpp:146-163)
    for (int64_t e = 0; e < s->E; e++) {
        const double* K = &s->Ke[(int64_t)nd * nd * e];
        for (int a = 0; a < npe; a++)
            for (int b = 0; b < npe; b++) {
                double* v = &s->vals[9 * s->slot[((int64_t)e * npe + a) * npe + b]];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) v[3 * i + j] += K[(3 * a + i) * nd + 3 * b + j];
            }
    }
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < s->N; r++)
        for (int64_t k = s->rowptr[r]; k < s->rowptr[r + 1]; k++) {
            double* v = &s->vals[9 * k];
            const int32_t col = s->colind[k];
            for (int i = 0; i < 3; i++) v[4 * i] += cm * s->mval[k];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    if (((s->bcmask[r] >> i) & 1) || ((s->bcmask[col] >> j) & 1))
                        v[3 * i + j] = (r == col && i == j) ? 1.0 : 0.0;
        }
    return ORC_SUCCESS;
}

void spmv(const orc_cpu_solver* s, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < s->N; r++) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int64_t k = s->rowptr[r]; k < s->rowptr[r + 1]; k++) {
            const double* v = &s->vals[9 * k];
            const double* xc = x + 3 * (int64_t)s->colind[k];
            a0 += v[0] * xc[0] + v[1] * xc[1] + v[2] * xc[2];
            a1 += v[3] * xc[0] + v[4] * xc[1] + v[5] * xc[2];
            a2 += v[6] * xc[0] + v[7] * xc[1] + v[8] * xc[2];
        }
        y[3 * r] = a0;
        y[3 * r + 1] = a1;
        y[3 * r + 2] = a2;
    }
}

void bj_setup(orc_cpu_solver* s) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < s->N; r++) {
        const int32_t* b = s->colind.data() + s->rowptr[r];
        const int32_t* e = s->colind.data() + s->rowptr[r + 1];
        const int64_t k = s->rowptr[r] + (std::lower_bound(b, e, (int32_t)r) - b);
        const double* a = &s->vals[9 * k];
        double* o = &s->dinv[9 * r];
        const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                           a[2] * (a[3] * a[7] - a[4] * a[6]);
        const double id = 1.0 / det;
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
}

void bj_apply(const orc_cpu_solver* s, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < s->N; r++) {
        const double* a = &s->dinv[9 * r];
        const double* q = x + 3 * r;
        y[3 * r] = a[0] * q[0] + a[1] * q[1] + a[2] * q[2];
        y[3 * r + 1] = a[3] * q[0] + a[4] * q[1] + a[5] * q[2];
        y[3 * r + 2] = a[6] * q[0] + a[7] * q[1] + a[8] * q[2];
    }
}

double dot(int64_t n, const double* a, const double* b) {
    double t = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : t)
    for (int64_t i = 0; i < n; i++) t += a[i] * b[i];
    return t;
}

// BiCGStab with right block-Jacobi preconditioning (van der Vorst 1992): A x = b from x = 0
int bicgstab(orc_cpu_solver* s, const double* b, double* x, double rtol, int maxit, int* its) {
    const int64_t n = s->n;
    std::vector<double> r(b, b + n), rh(b, b + n), p(n, 0.0), v(n, 0.0), ph(n), sv(n), sh(n), t(n);
    std::fill(x, x + n, 0.0);
    bj_setup(s);
    const double bn = std::sqrt(dot(n, b, b));
    const double target = rtol * bn;
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    int it = 0;
    double rn = bn;
    while (rn > target && it < maxit) {
        const double rho1 = dot(n, rh.data(), r.data());
        if (!(std::fabs(rho1) > 0.0)) break;
        const double beta = (rho1 / rho) * (alpha / omega);
        rho = rho1;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) p[i] = it == 0 ? r[i] : r[i] + beta * (p[i] - omega * v[i]);
        bj_apply(s, p.data(), ph.data());
        spmv(s, ph.data(), v.data());
        alpha = rho / dot(n, rh.data(), v.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) sv[i] = r[i] - alpha * v[i];
        bj_apply(s, sv.data(), sh.data());
        spmv(s, sh.data(), t.data());
        const double tt = dot(n, t.data(), t.data());
        omega = tt > 0.0 ? dot(n, t.data(), sv.data()) / tt : 0.0;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            x[i] += alpha * ph[i] + omega * sh[i];
            r[i] = sv[i] - omega * t[i];
        }
        rn = std::sqrt(dot(n, r.data(), r.data()));
        it++;
        if (omega == 0.0) break;
    }
    *its = it;
    return rn <= target ? ORC_SUCCESS : ORC_FAILED;
}

// Newton on r(d) = f(nodes + d) [+ cm M (d - d~)] = 0 from the current d, the same
// rules as the GPU path (abstol / rtol / stol = Solver.Precision, Eisenstat-Walker
// forcing with eta_0 = 0.1, eta <= 0.5 and Kelley's safeguard, backtracking on ||r||)
int newton(orc_cpu_solver* s, double time, double cm, double eps, bool static_h, double atol, double rtol,
           double stol, int maxit, double* stats) {
    const int64_t n = s->n;
    std::vector<double> x(n), r(n), delta(n), rhs(n), md(n);
    double r0 = 0.0, rprev = 0.0, rn = 0.0, snorm = 0.0, lambda = 1.0, eta = 0.1;
    double t_tan = 0.0, t_lin = 0.0;
    int it = 0, kry = 0, reason = 0;
    auto eval = [&](double* out) -> int {
        for (int64_t i = 0; i < n; i++) x[i] = s->nodes[i] + s->d[i];
        int rc = residual(s, x.data(), time, out);
        if (rc != ORC_SUCCESS) return rc;
        if (cm != 0.0) {
            for (int64_t i = 0; i < n; i++) md[i] = s->d[i] - s->dpred[i];
#pragma omp parallel for schedule(static)
            for (int64_t rr = 0; rr < s->N; rr++) {
                double a0 = 0.0, a1 = 0.0, a2 = 0.0;
                for (int64_t k = s->rowptr[rr]; k < s->rowptr[rr + 1]; k++) {
                    const double mv = s->mval[k];
                    const double* q = &md[3 * (int64_t)s->colind[k]];
                    a0 += mv * q[0];
                    a1 += mv * q[1];
                    a2 += mv * q[2];
                }
                const uint8_t m = s->bcmask[rr];
                if (!(m & 1)) out[3 * rr] += cm * a0;
                if (!(m & 2)) out[3 * rr + 1] += cm * a1;
                if (!(m & 4)) out[3 * rr + 2] += cm * a2;
            }
        }
        return ORC_SUCCESS;
    };
    for (;; it++) {
        const int rc = eval(r.data());
        if (rc != ORC_SUCCESS) {
            if (it > 0 && lambda > 1.0 / (1 << 20)) {
                lambda *= 0.5;
                for (int64_t i = 0; i < n; i++) s->d[i] -= lambda * delta[i];
                it--;
                continue;
            }
            reason = -6;
            break;
        }
        rn = std::sqrt(dot(n, r.data(), r.data()));
        if (it == 0) {
            r0 = rn;
        } else if (rn > rprev && lambda > 1.0 / 64 && rn > atol) {
            lambda *= 0.5;
            for (int64_t i = 0; i < n; i++) s->d[i] -= lambda * delta[i];
            it--;
            continue;
        }
        if (rn < atol) { reason = 2; break; }
        if (it > 0 && rn < rtol * r0) { reason = 3; break; }
        const double dn = std::sqrt(dot(n, s->d.data(), s->d.data()));
        if (it > 0 && rn <= rprev && snorm * lambda < stol * dn) { reason = 4; break; }
        if (it >= maxit) { reason = -5; break; }
        if (it > 0) eta = std::min(0.5, std::max(1e-6, 0.9 * (rn / rprev) * (rn / rprev)));
        const double eta_k = std::min(0.5, std::max(eta, 0.5 * (atol + rtol * r0) / rn));
        rprev = rn;
        const double ta = now_s();
        const double h = static_h ? (1.0 + dn) * eps : eps;
        if (tangent(s, x.data(), time, h, cm) != ORC_SUCCESS) { reason = -6; break; }
        t_tan += now_s() - ta;
        const double tl = now_s();
        for (int64_t i = 0; i < n; i++) rhs[i] = -r[i];
        int k = 0;
        bicgstab(s, rhs.data(), delta.data(), eta_k, 20000, &k);
        kry += k;
        t_lin += now_s() - tl;
        lambda = 1.0;
        for (int64_t i = 0; i < n; i++) s->d[i] += delta[i];
        snorm = std::sqrt(dot(n, delta.data(), delta.data()));
        if (!(snorm > 0.0) || !std::isfinite(snorm)) { reason = -3; break; }
    }
    stats[0] = it;
    stats[1] = kry;
    stats[2] = reason;
    stats[3] = rn;
    stats[4] = r0;
    stats[5] = t_tan;
    stats[6] = t_lin;
    return reason > 0 ? ORC_SUCCESS : ORC_FAILED;
}

}  // namespace

extern "C" {

orc_cpu_solver* orc_cpu_solver_create(int32_t npe, int64_t n_elem, const int32_t* conn, int64_t n_nodes,
                                      const double* coords, const double* dNdX, const double* bases, const double* w,
                                      const uint8_t* bcmask, const orc_material* mats, int32_t n_mats,
                                      const int32_t* matidx, const double* tact, int32_t nps, int64_t n_surf,
                                      const int32_t* sconn, const int32_t* sgroup, int32_t n_groups) {
    auto* s = new orc_cpu_solver();
    s->npe = npe;
    s->nps = nps;
    s->E = n_elem;
    s->N = n_nodes;
    s->S = n_surf;
    s->n = 3 * n_nodes;
    const int nq = npe == 10 ? 150 : 12, nb = npe == 10 ? 45 : 9;
    s->conn.assign(conn, conn + n_elem * npe);
    s->dNdX.assign(dNdX, dNdX + n_elem * nq);
    s->bases.assign(bases, bases + n_elem * nb);
    s->w.assign(w, w + n_elem);
    s->tact.assign(tact, tact + n_elem);
    s->matidx.assign(matidx, matidx + n_elem);
    s->mats.assign(mats, mats + n_mats);
    s->bcmask.assign(bcmask, bcmask + n_nodes);
    s->sconn.assign(sconn, sconn + n_surf * nps);
    s->sgroup.assign(sgroup, sgroup + n_surf);
    s->pressures.assign(std::max(1, n_groups), 0.0);
    // node -> (element, local node) incidences, element order
    s->inc_ptr.assign(n_nodes + 1, 0);
    for (int64_t i = 0; i < n_elem * npe; i++) s->inc_ptr[conn[i] + 1]++;
    for (int64_t r = 0; r < n_nodes; r++) s->inc_ptr[r + 1] += s->inc_ptr[r];
    s->inc.resize(n_elem * npe);
    std::vector<int64_t> fill(s->inc_ptr.begin(), s->inc_ptr.end() - 1);
    for (int64_t i = 0; i < n_elem * npe; i++) s->inc[fill[conn[i]]++] = (int32_t)i;
    // block pattern: nodes sharing an element
    s->rowptr.assign(n_nodes + 1, 0);
    std::vector<int32_t> tmp;
    for (int64_t r = 0; r < n_nodes; r++) {
        tmp.clear();
        for (int64_t k = s->inc_ptr[r]; k < s->inc_ptr[r + 1]; k++) {
            const int64_t e = s->inc[k] / npe;
            for (int b = 0; b < npe; b++) tmp.push_back(conn[e * npe + b]);
        }
        if (tmp.empty()) tmp.push_back((int32_t)r);
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        s->colind.insert(s->colind.end(), tmp.begin(), tmp.end());
        s->rowptr[r + 1] = (int64_t)s->colind.size();
    }
    s->nnzb = (int64_t)s->colind.size();
    s->slot.resize(n_elem * npe * npe);
    for (int64_t e = 0; e < n_elem; e++)
        for (int a = 0; a < npe; a++) {
            const int32_t ra = conn[e * npe + a];
            const int32_t* b0 = &s->colind[s->rowptr[ra]];
            const int32_t* b1 = &s->colind[0] + s->rowptr[ra + 1];
            for (int b = 0; b < npe; b++)
                s->slot[(e * npe + a) * npe + b] =
                    (int32_t)(s->rowptr[ra] + (std::lower_bound(b0, b1, conn[e * npe + b]) - b0));
        }
    // consistent mass, one scalar per block (orc_t10_mass_consistent / orc_t4_mass_consistent)
    s->mval.assign(s->nnzb, 0.0);
    const int nd = 3 * npe;
    std::vector<double> Me(nd * nd);
    uint8_t bc0[30] = {};
    for (int64_t e = 0; e < n_elem; e++) {
        const double rho = mats[matidx[e]].density;
        if (npe == 10)
            orc_t10_mass_consistent(w[e], rho, bc0, Me.data());
        else
            orc_t4_mass_consistent(w[e], rho, bc0, Me.data());
        for (int a = 0; a < npe; a++)
            for (int b = 0; b < npe; b++) s->mval[s->slot[(e * npe + a) * npe + b]] += Me[(3 * a) * nd + 3 * b];
    }
    s->nodes.assign(coords, coords + 3 * n_nodes);
    for (auto* v : {&s->vel, &s->acc, &s->d, &s->dpred, &s->guess}) v->assign(s->n, 0.0);
    s->fe.assign(n_elem * nd, 0.0);
    s->Ke.assign((size_t)n_elem * nd * nd, 0.0);
    s->vals.assign(9 * s->nnzb, 0.0);
    s->dinv.assign(9 * n_nodes, 0.0);
    return s;
}

void orc_cpu_solver_destroy(orc_cpu_solver* s) { delete s; }

void orc_cpu_set_pressures(orc_cpu_solver* s, const double* p) {
    std::copy(p, p + s->pressures.size(), s->pressures.begin());
}

void orc_cpu_get_state(const orc_cpu_solver* s, double* nodes, double* vel, double* acc) {
    if (nodes) std::copy(s->nodes.begin(), s->nodes.end(), nodes);
    if (vel) std::copy(s->vel.begin(), s->vel.end(), vel);
    if (acc) std::copy(s->acc.begin(), s->acc.end(), acc);
}

void orc_cpu_set_state(orc_cpu_solver* s, const double* nodes, const double* vel, const double* acc) {
    if (nodes) std::copy(nodes, nodes + s->n, s->nodes.begin());
    if (vel) std::copy(vel, vel + s->n, s->vel.begin());
    if (acc) std::copy(acc, acc + s->n, s->acc.begin());
}

// stats[8]: Newton iterations, Krylov iterations, reason, ||r||, ||r0||, tangent s, linear s, wall s
int32_t orc_cpu_newmark_step(orc_cpu_solver* s, double time, double dt, double beta, double gamma, double eps,
                             double atol, double rtol, double stol, int32_t maxit, double* stats) {
    const double t0 = now_s();
    const int64_t n = s->n;
    std::vector<double> vpred(n);
    for (int64_t i = 0; i < n; i++) {
        s->dpred[i] = dt * s->vel[i] + dt * dt * 0.5 * (1 - 2 * beta) * s->acc[i];
        vpred[i] = s->vel[i] + (1 - gamma) * dt * s->acc[i];
        s->d[i] = ((s->bcmask[i / 3] >> (i % 3)) & 1) ? 0.0 : s->dpred[i];
    }
    const int rc = newton(s, time, 1.0 / (beta * dt * dt), eps, false, atol, rtol, stol, maxit, stats);
    if (rc == ORC_SUCCESS) {
        for (int64_t i = 0; i < n; i++) {
            s->acc[i] = (s->d[i] - s->dpred[i]) / (beta * dt * dt);
            s->vel[i] = vpred[i] + gamma * dt * s->acc[i];
            s->nodes[i] += s->d[i];
        }
    }
    stats[7] = now_s() - t0;
    return rc;
}

int32_t orc_cpu_static_step(orc_cpu_solver* s, double time, double dt, double eps, double atol, double rtol,
                            double stol, int32_t maxit, double* stats) {
    const double t0 = now_s();
    const int64_t n = s->n;
    // UpdateInitialGuess (CBSolverEquilibrium.cpp:234-263)
    if (time > s->time_last) {
        if (s->step_last > 0.0) {
            for (int64_t i = 0; i < n; i++) {
                s->guess[i] = s->d[i] / s->step_last;
                s->d[i] = s->d[i] * dt / s->step_last;
            }
        } else {
            std::fill(s->d.begin(), s->d.end(), 0.0);
        }
    } else if (time == s->time_last) {
        std::fill(s->d.begin(), s->d.end(), 0.0);
    } else {
        for (int64_t i = 0; i < n; i++) s->d[i] = s->guess[i] * dt;
    }
    s->time_last = time;
    s->step_last = dt;
    for (int64_t i = 0; i < n; i++)
        if ((s->bcmask[i / 3] >> (i % 3)) & 1) s->d[i] = 0.0;
    const int rc = newton(s, time, 0.0, eps, true, atol, rtol, stol, maxit, stats);
    if (rc == ORC_SUCCESS)
        for (int64_t i = 0; i < n; i++) s->nodes[i] += s->d[i];
    else
        std::fill(s->d.begin(), s->d.end(), 0.0);
    stats[7] = now_s() - t0;
    return rc;
}

// CBSolver::CalcFiniteDifferencesEpsilon (CBSolver.cpp:1610-1625): (1 + ||a||_2) eps
double orc_fd_epsilon(const double* a, int64_t n, double eps) {
    double t = 0.0;
    for (int64_t i = 0; i < n; i++) t += a[i] * a[i];
    double dx = 1.0 + std::sqrt(t);
    if (std::fabs(dx) < 1e-16) dx = (dx < 0.0 ? -1.0 : 1.0) * 1e-1;
    return dx * eps;
}

}  // extern "C"
