/*
 * CPU FP64 oracle for the cardiac-mechanics hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C++ on the CPU, the arithmetic of the
 * reference's element, material and surface-load routines so that the HIP
 * product path can be checked against it.  Only tests/, the smoke check of
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  The
 * product library never links or calls it.
 *
 * Every routine cites the reference file:line it follows.  Each routine is
 * templated on the scalar so that the same code runs in double and in
 * std::complex<double>: the complex instantiation gives complex-step
 * derivatives (Im f(x + i h e_k) / h), i.e. Jacobians exact to rounding, which
 * pin the product's hand-derived analytic tangent independently of how that
 * tangent was derived.
 *
 * Material record layout (shared with the product's cmg_material by meaning,
 * not by code): see struct orc_material below.
 */
#ifndef CMG_ORACLE_H
#define CMG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_LAW_GUCCIONE = 0,
    ORC_LAW_HOLZAPFEL = 1
};

enum {
    ORC_FN_NONE = 0,
    ORC_FN_LINEAR = 1,
    ORC_FN_SINUS = 2,
    ORC_FN_SINUSDRIVER = 3,
    ORC_FN_DOUBLEHILL = 4,
    ORC_FN_LAND17 = 6,       /* Land17 cross-bridge model; parameters and per-element state through
                                orc_land17_configure */
    ORC_FN_TANH = 5          /* TanH tension model; fn = cycleLength, l0, t_emd, S_peak, t_dur,
                                tau_c0, ld, ld_up, tau_r, kff, startTime */
};

/* CBStatus values (mechanics/src/CBStatus/CBStatus.h:20-37). */
enum {
    ORC_SUCCESS = 2,
    ORC_FAILED = -3,
    ORC_CORRUPT_ELEMENT = -6
};

typedef struct orc_material {
    int32_t law;             /* ORC_LAW_* */
    int32_t ignore_corrupt;  /* Materials.Mat_<id>.IgnoreCorruptElements */
    int32_t tension_fn;      /* ORC_FN_* of the FromFunction tension model */
    int32_t pad_;
    double density;          /* Materials.Mat_<id>.Density */
    /* Guccione: C, bf, bt, bfs, K (aScale/bScale already applied)
     * Holzapfel: a, b, af, bf, as, bs, afs, bfs, kappa, k              */
    double p[12];
    double tension_max;      /* Materials.Mat_<id>.TensionMax */
    /* analytic function: startTime, stopTime, keepMaxValue, omega, phi,
     * DoubleHill period, contrTimeOffset, relaxTimeOffset, contrRateConst,
     * relaxRateConst, onsetTime, driverMax                              */
    double fn[12];
} orc_material;

/* ---- analytic time functions (CBAnalyticFunction.h:53-176) ------------ */
double orc_analytic(int32_t type, const double* fn, double time, double prev_time);

/* ---- constitutive laws, local fibre frame, row-major 3x3 -------------- */
int32_t orc_pk2(const orc_material* m, const double* F, double* S);
int32_t orc_energy(const orc_material* m, const double* F, double* W);
/* complex-step derivative of the energy w.r.t. F, dW/dF (row-major) */
int32_t orc_energy_dF_cs(const orc_material* m, const double* F, double* dWdF);
/* complex-step dS/dF: out[9*a + b] = dS_a / dF_b */
int32_t orc_pk2_dF_cs(const orc_material* m, const double* F, double* dSdF);
void orc_active_stress(const double* F, double tension, double* Sa);
/* CBTensionModelTanH::CalcActiveTension (CBTensionModelTanH.cpp:71-97): tension at
 * fibre stretch lambda, time t, element activation time tact */
double orc_tanh_tension(const orc_material* m, double lambda, double time, double tact);
/* active PK2 of the element's tension model at the local F (FromFunction or TanH) */
void orc_active_stress_model(const orc_material* m, const double* F, double time, double tact, double* Sa);

/* ---- T10 (CBElementSolidT10.cpp) ------------------------------------- */
void orc_t10_shape(const double* X30, double* dNdX150, double* detJ);
int32_t orc_t10_residual(const double* x30, const double* dNdX150, const double* bases45,
                         double detJ, const uint8_t* bc30, const orc_material* m,
                         double time, double tact, double* f30);
int32_t orc_t10_fd_tangent(const double* x30, const double* dNdX150, const double* bases45,
                           double detJ, const uint8_t* bc30, const orc_material* m,
                           double time, double tact, double eps, double* K900);
int32_t orc_t10_cs_tangent(const double* x30, const double* dNdX150, const double* bases45,
                           double detJ, const uint8_t* bc30, const orc_material* m,
                           double time, double tact, double* K900);
void orc_t10_mass_consistent(double detJ, double density, const uint8_t* bc30, double* M900);
void orc_t10_mass_lumped(double volume, double density, const uint8_t* bc30, double* M900);
/* energy of one element: V/4 * sum_q W(F_q), for the energy-consistency pin */
int32_t orc_t10_energy(const double* x30, const double* dNdX150, const double* bases45,
                       double detJ, const orc_material* m, double* W);
int32_t orc_t10_energy_grad_cs(const double* x30, const double* dNdX150, const double* bases45,
                               double detJ, const orc_material* m, double* g30);

/* ---- T4 (CBElementSolidT4.cpp) ---------------------------------------- */
void orc_t4_shape(const double* X12, double* dNdX12, double* detJ);
int32_t orc_t4_residual(const double* x12, const double* dNdX12, const double* basis9,
                        double volume, const uint8_t* bc12, const orc_material* m,
                        double time, double tact, double* f12);
int32_t orc_t4_fd_tangent(const double* x12, const double* dNdX12, const double* basis9,
                          double volume, const uint8_t* bc12, const orc_material* m,
                          double time, double tact, double eps, double* K144);
int32_t orc_t4_cs_tangent(const double* x12, const double* dNdX12, const double* basis9,
                          double volume, const uint8_t* bc12, const orc_material* m,
                          double time, double tact, double* K144);
void orc_t4_mass_consistent(double volume, double density, const uint8_t* bc12, double* M144);
void orc_t4_mass_lumped(double volume, double density, const uint8_t* bc12, double* M144);

/* ---- surface loads (CBElementSurfaceT3.cpp / CBElementSurfaceT6.cpp) -- */
void orc_t3_pressure_forces(const double* x9, const uint8_t* bc9, double p, double* f9);
void orc_t3_pressure_cs_tangent(const double* x9, const uint8_t* bc9, double p, double* K81);
double orc_t3_volume(const double* x9, const double* ref3);
void orc_t6_pressure_forces(const double* x18, const uint8_t* bc18, double p, double* f18);
void orc_t6_pressure_fd_tangent(const double* x18, const uint8_t* bc18, double p, double eps, double* K324);
void orc_t6_pressure_cs_tangent(const double* x18, const uint8_t* bc18, double p, double* K324);
double orc_t6_volume(const double* x18, const double* ref3);

/* ---- mesh-wide batches (OpenMP over elements) ------------------------- */
/* npe = 10 (T10) or 4 (T4); bases per element: 45 (T10) / 9 (T4);
 * dNdX per element: 150 (T10) / 12 (T4); w = detJ (T10) or volume (T4).  */
int32_t orc_residual_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* coords,
                           const double* dNdX, const double* bases, const double* w,
                           const uint8_t* bcmask, const orc_material* mats, const int32_t* matidx,
                           const double* tact, double time, double* fe, int32_t* status);
int32_t orc_fd_tangent_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* coords,
                             const double* dNdX, const double* bases, const double* w,
                             const uint8_t* bcmask, const orc_material* mats, const int32_t* matidx,
                             const double* tact, double time, double eps, double* Ke, int32_t* status);
int32_t orc_cs_tangent_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* coords,
                             const double* dNdX, const double* bases, const double* w,
                             const uint8_t* bcmask, const orc_material* mats, const int32_t* matidx,
                             const double* tact, double time, double* Ke, int32_t* status);
int32_t orc_num_threads(void);
void orc_shape_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* coords, double* dNdX,
                     double* detJ);
void orc_mass_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* w, const double* density,
                    const uint8_t* bcmask, int32_t consistent, double* Me);

/* ---- CPU solver steps with the reference's finite-difference tangent (cpu_solver.cpp) */
typedef struct orc_cpu_solver orc_cpu_solver;
orc_cpu_solver* orc_cpu_solver_create(int32_t npe, int64_t n_elem, const int32_t* conn, int64_t n_nodes,
                                      const double* coords, const double* dNdX, const double* bases, const double* w,
                                      const uint8_t* bcmask, const orc_material* mats, int32_t n_mats,
                                      const int32_t* matidx, const double* tact, int32_t nps, int64_t n_surf,
                                      const int32_t* sconn, const int32_t* sgroup, int32_t n_groups);
void orc_cpu_solver_destroy(orc_cpu_solver* s);
void orc_cpu_set_pressures(orc_cpu_solver* s, const double* p);
void orc_cpu_get_state(const orc_cpu_solver* s, double* nodes, double* vel, double* acc);
void orc_cpu_set_state(orc_cpu_solver* s, const double* nodes, const double* vel, const double* acc);
int32_t orc_cpu_newmark_step(orc_cpu_solver* s, double time, double dt, double beta, double gamma, double eps,
                             double atol, double rtol, double stol, int32_t maxit, double* stats);
int32_t orc_cpu_static_step(orc_cpu_solver* s, double time, double dt, double eps, double atol, double rtol,
                            double stol, int32_t maxit, double* stats);
double orc_fd_epsilon(const double* a, int64_t n, double eps);

/* ---- Land17 tension model (CBTensionModelLand17.cpp) -------------------- */
void orc_land17_configure(const double* params, const double* state, double dt_ms);
double orc_land17_cai(const double* p, double time, double tact);
double orc_land17_step(const double* p, double tmax, const double* state, double lambda, double dt_ms, double cai,
                       double* next, double* dTdl);
void orc_land17_commit_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* coords,
                             const double* dNdX, const double* bases, const orc_material* mats,
                             const int32_t* matidx, const double* tact, const double* params, double time,
                             double dt_ms, double* state);

#ifdef __cplusplus
}
#endif
#endif
