// Launchers of the parameter-set batch kernels (batch_kernels.hip).  Internal.
#pragma once

#include <algorithm>

#include "cmg_internal.h"

namespace cmg {

constexpr int kMaxSets = 256;

// one scalar per set, passed by value (kernel argument space, no copy to the device)
struct SetCoefs {
    double c[kMaxSets];
};

// a preconditioner application z = P^-1 u launched on the caller's stream
typedef void (*PrecondApply)(void* apply_ctx, const double* u, double* z);

// Device state of S preconditioned conjugate-gradient solves (all arrays device memory).
struct BatchCG {
    int S = 0;
    double *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;   // [S][n]
    double *rz = nullptr, *rr = nullptr, *rr0 = nullptr, *alpha = nullptr, *beta = nullptr, *tol2 = nullptr;  // [S]
    int32_t *done = nullptr, *it = nullptr;   // [S]: 0 running, 1 converged / idle, 2 breakdown
    double* part = nullptr;                   // [S][part_ld] partial sums
    int64_t part_ld = 0;
};

// Device state of S right block-Jacobi preconditioned BiCGStab solves (van der
// Vorst 1992), scalars on the device: one iteration is eight launches and no host
// synchronisation.  done: 0 running, 1 converged by the recurrence residual,
// 2 breakdown (the host confirms with the true residual and restarts when needed).
struct BatchBiCG {
    int S = 0;
    double *x = nullptr, *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr;   // [S][n]
    double *s = nullptr, *t = nullptr, *ph = nullptr, *sh = nullptr;              // [S][n]
    double *rho = nullptr, *alpha = nullptr, *omega = nullptr, *beta = nullptr;    // [S]
    double *rr = nullptr, *tol2 = nullptr;                                        // [S]
    int32_t *done = nullptr, *it = nullptr, *fin = nullptr;                       // [S]
    double* part = nullptr;                                                       // [S][part_ld]
    int64_t part_ld = 0;
};
int64_t bicg_part_ld(int64_t N);   // partial-sum slots per set the kernels need

// r must hold b - A x: r^ = r, p = r, rho = r.r, tol2 = (rtol ||b||)^2 for the active sets
void launch_bicg_init(const BatchBiCG& k, int64_t N, const double* b, const double* rtol, const int32_t* active,
                      hipStream_t s);
// one BiCGStab iteration of every running set; chunked: 144-byte value chunks (values
// 16-byte aligned per set and padded, as for launch_spmv).  The preconditioner is
// block Jacobi with dinv, or -- one set only -- apply(apply_ctx, u, z) launching
// z = P^-1 u on the stream (dinv then null)
void launch_bicg_iteration(const BatchBiCG& k, int lanes, bool chunked, int64_t N, const int64_t* rowptr,
                           const int32_t* colind, const void* vals, bool fp32, int64_t set_vals, const double* dinv,
                           hipStream_t s, PrecondApply apply = nullptr, void* apply_ctx = nullptr);

int batch_dot_blocks(int64_t n);
int batch_spmv_blocks(int64_t N, int lanes);
int batch_node_blocks(int64_t N);

// x = 0, r = b, z = p = D^-1 b and the tolerances ||r|| <= rtol[s] ||b_s|| for the active sets
// apply: z = P^-1 u for all sets (u, z [S][n]) in place of the block Jacobi with dinv
void launch_bcg_init(const BatchCG& cg, int64_t N, const double* b, const double* dinv, const double* rtol,
                     const int32_t* active, hipStream_t s, PrecondApply apply = nullptr, void* apply_ctx = nullptr);
// one CG iteration of every running set (5 launches, no host synchronisation)
void launch_bcg_iteration(const BatchCG& cg, int lanes, int64_t N, const int64_t* rowptr, const int32_t* colind,
                          const void* vals, bool fp32, int64_t set_vals, const double* dinv, hipStream_t s,
                          PrecondApply apply = nullptr, void* apply_ctx = nullptr);
// only the q = A p (+ p.q partials) step, for timing the lane count
void launch_bcg_spmv_dot(const BatchCG& cg, int lanes, int64_t N, const int64_t* rowptr, const int32_t* colind,
                         const void* vals, bool fp32, int64_t set_vals, hipStream_t s);
// out[s] = x_s . y_s (vectors of length n, set-major)
void launch_bdot(int S, int64_t n, const double* x, const double* y, double* partial, int64_t part_ld, double* out,
                 hipStream_t s);
// y_s += a[s] x_s ; y_s *= a[s]
void launch_baxpy(int S, int64_t n, const SetCoefs& a, const double* x, double* y, hipStream_t s);
void launch_bscale(int S, int64_t n, const SetCoefs& a, double* y, hipStream_t s);

}  // namespace cmg
