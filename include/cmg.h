/*
 * cmg.h -- C ABI of the MI355X (gfx950) mechanics hot path.
 *
 * One context owns every device buffer of one mesh on one GPU: coordinates,
 * connectivity, reference shape-function derivatives, fibre bases, the fixed
 * 3x3-block sparsity pattern, the deterministic assembly maps and the solver
 * workspace.  Plain pointers and sizes only; no exceptions cross this ABI.
 * Return values are CBStatus codes (mechanics/src/CBStatus/CBStatus.h:20-37)
 * or the negative CMG_ERR_* argument/runtime codes below.
 *
 * Conventions (mechanics/src/CBElement/CBElementAdapter.h:55-60): 0-based node
 * ids, DOF = 3*node + component, coordinates in metres, SI units throughout.
 * Host-pointer entry points copy; *_device entry points take device pointers
 * and run stream-ordered on the context's stream (they do not synchronise).
 * Calls on one context are not re-entrant.
 */
#ifndef CMG_H
#define CMG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CMG_API_VERSION 1

/* CBStatus values */
#define CMG_SUCCESS 2
#define CMG_NOTHING_DONE 0
#define CMG_REPEAT (-2)
#define CMG_FAILED (-3)
#define CMG_INFINITIVE (-4)
#define CMG_NOT_A_NUMBER (-5)
#define CMG_CORRUPT_ELEMENT (-6)
#define CMG_SYSTEM_WRONG_STATE (-8)
/* ABI-level errors */
#define CMG_ERR_INVALID (-100)
#define CMG_ERR_HIP (-101)
#define CMG_ERR_NOMEM (-102)

#define CMG_LAW_GUCCIONE 0
#define CMG_LAW_HOLZAPFEL 1

#define CMG_FN_NONE 0
#define CMG_FN_LINEAR 1
#define CMG_FN_SINUS 2
#define CMG_FN_SINUSDRIVER 3
#define CMG_FN_DOUBLEHILL 4
#define CMG_FN_TANH 5   /* the TanH tension model (CBTensionModelTanH.cpp), not a time function */
#define CMG_FN_LAND17 6 /* the Land17 cross-bridge model (CBTensionModelLand17.cpp): per-element ODE
                           state, parameters through cmg_set_land17_parameters */

#define CMG_LAND17_NPARAMS 40

#define CMG_PREC_FP64 0
#define CMG_PREC_MIXED 1

typedef struct cmg_ctx cmg_ctx;

/* One material: CBMaterial (CBMaterial.cpp:19-74) + its constitutive model
 * (CBConstitutiveModelGuccione.cpp:20-71 / CBConstitutiveModelHolzapfel.cpp:18-85)
 * + the FromFunction tension model (CBTensionFromFunction.cpp:36-56).
 *   Guccione  params: C, bf, bt, bfs, K          (aScale/bScale already applied)
 *   Holzapfel params: a, b, af, bf, as, bs, afs, bfs, kappa, k
 *   fn: startTime, stopTime, keepMaxValue, omega, phi, DoubleHill period,
 *       contrTimeOffset, relaxTimeOffset, contrRateConst, relaxRateConst,
 *       onsetTime, driverMax
 * tension_fn = CMG_FN_TANH selects the TanH tension model instead
 * (CBTensionModelTanH.cpp:71-97): T = Tmax max(S_peak phi tanh^2(t_s/tau_c)
 * tanh^2((t_dur - t_s)/tau_r), 0) with phi = tanh(ld (lambda_f - l0)),
 * tau_c = tau_c0 + ld_up (1 - phi), t_s = clamp(t - t_act - startTime - t_emd, 0, t_dur)
 * (t wrapped into one cycle), active PK2 kff T / lambda_f on ff;
 *   fn: cycleLength, MinFiberStretch l0, EMD t_emd, PeakTension S_peak, Duration t_dur,
 *       ContractionTime tau_c0, DegreeOfLengthDependence ld,
 *       LengthDependenceUpstrokeTime ld_up, RelaxationTime tau_r, kff, startTime     */
typedef struct cmg_material {
    int32_t law;
    int32_t ignore_corrupt;   /* Materials.Mat_<id>.IgnoreCorruptElements */
    int32_t tension_fn;       /* CMG_FN_* */
    int32_t reserved;
    double density;           /* Materials.Mat_<id>.Density (kg/m^3) */
    double params[12];
    double tension_max;       /* Materials.Mat_<id>.TensionMax (Pa) */
    double fn[12];
} cmg_material;

/* Mesh as the Tetgen loader hands it to the solver (CBModelLoaderTetgen.cpp:59-547). */
typedef struct cmg_mesh {
    int64_t n_nodes;
    int64_t n_elements;
    int64_t n_surfaces;
    int32_t nodes_per_element;  /* 10 (T10) or 4 (T4) */
    int32_t nodes_per_surface;  /* 3 (T3), 6 (T6), 0 without surfaces */
    int32_t n_groups;           /* pressure groups (cavities) */
    int32_t reserved;
    const double* ref_coords;       /* [n_nodes*3] reference coordinates */
    const int32_t* elements;        /* [n_elements*npe] */
    const double* bases;            /* [n_elements*nb*9], nb = 5 (T10: centroid, QP1..4) or 1 (T4);
                                       row-major Matrix3 with columns f, s, n */
    const int32_t* material_index;  /* [n_elements] or NULL (all 0) */
    const double* activation_time;  /* [n_elements] or NULL (all 0) */
    const uint8_t* bc;              /* [n_nodes] bitmask 1=x 2=y 4=z, or NULL */
    const int32_t* surfaces;        /* [n_surfaces*nps] */
    const int32_t* surface_group;   /* [n_surfaces] pressure group of each surface element */
} cmg_mesh;

typedef struct cmg_options {
    int32_t precision;          /* CMG_PREC_FP64 (default) or CMG_PREC_MIXED: FP32 tangent
                                   values, FP32 preconditioned Krylov operator; residuals,
                                   det J / penalty and convergence checks stay FP64 */
    int32_t transpose_tangent;  /* 1: assemble each T10 element tangent transposed, the
                                   reference convention (CBElementSolidT10.cpp:559-565); T4
                                   tangents are never transposed (CBElementSolidT4.cpp:340-349) */
    int32_t pressure_tangent;   /* 0: none (the reference's T3 tangent is zero,
                                   CBElementSurfaceT3.cpp:268-269); 1: exact (default) */
    int32_t consistent_mass;    /* Solver.NewmarkBeta.ConsistentMassMatrix (default 1) */
    int32_t newton_maxit;       /* Solver.MaxIterations (default 50) */
    int32_t krylov_restart;     /* GMRES restart length (default 80) */
    int32_t krylov_maxit;       /* Krylov iterations per linear solve (default 4000) */
    int32_t device;             /* HIP device ordinal (default 0) */
    int32_t element_variant;    /* residual+tangent kernel: 2 shared quadrature stage, 3 the same in
                                   two node passes, 4 the same with two lanes per element and
                                   spatial pair, 1 per-lane quadrature points, 5 / 6 variants 2 / 4
                                   with 14 elements per workgroup, 7 / 8 variants 6 / 5 with the
                                   element tangents written through LDS as contiguous 16-byte
                                   runs; 0 (default): 6 in FP64, 5 in mixed precision.  A context
                                   with a Land17 material always runs the shape of variant 6 */
    int32_t krylov_method;      /* 0 GMRES(restart), 1 BiCGStab with device-side scalars and graph-replayed
                                   iterations, 2 BiCGStab with host-side scalars (default); all right
                                   preconditioned (see preconditioner) */
    double newton_atol;         /* Solver.Precision used as SNES abstol/rtol/stol (default 1e-8) */
    double newton_rtol;
    double newton_stol;
    double krylov_rtol;         /* linear-solve relative tolerance (default 1e-6) */
    double newmark_beta;        /* Solver.NewmarkBeta.Beta  (default 0.25) */
    double newmark_gamma;       /* Solver.NewmarkBeta.Gamma (default 0.5) */
    double fd_epsilon;          /* Solver.Epsilon (default 1e-12), reported only */
    double rayleigh_alpha;      /* Materials.Global.Damping.Rayleigh.Alpha (1/s, default 0) */
    double rayleigh_beta;       /* Materials.Global.Damping.Rayleigh.Beta (s, default 0): Newmark steps
                                   add C (v~ + gamma/(beta dt) (d - d~)) to the residual and
                                   gamma/(beta dt) C to the tangent, C = alpha M + beta_R K(x_n) with
                                   K the element tangent at the step's start configuration in the
                                   reference's finite-difference layout (K_e^T per T10 element,
                                   CBElementSolidT10.cpp:559-565; K_e per T4) whatever
                                   transpose_tangent says, rebuilt whenever that configuration
                                   changes (CBSolverNewmarkBeta.cpp:384-390, 622-626, 657-702,
                                   857-866); off when both are 0 */
    int32_t preconditioner;     /* of the Krylov solves: 1 3x3 nodal block Jacobi, 2 vertex-patch
                                   additive Schwarz (T10: each element vertex with the midside
                                   nodes of its edges, patch matrices inverted on the device;
                                   block Jacobi for T4); 0 (default):
                                   2 where it applies */
    int32_t precond_reuse;      /* 1 (default): the Schwarz preconditioner is built from the first
                                   tangent of each Newton solve and reused by its later linear solves
                                   (the reference's SNESSetLagPreconditioner(-2),
                                   CBSolverNewmarkBeta.cpp:173-174); 0: built from every tangent.
                                   Block Jacobi is always built from every tangent */
    double krylov_eta_max;      /* upper bound of the Eisenstat-Walker forcing term of the Newton
                                   linear solves (default 0.5; 0 reads as the default) */
    int32_t newton_stall;       /* k > 0: a Newton solve whose residual norm falls by less than 10 %
                                   in each of k consecutive iterations ends as diverged (reason -7)
                                   instead of running on to Solver.MaxIterations; 0: off */
    int32_t spmv_choice;        /* 0 (default): the fastest of the nine SpMV layouts, timed on the
                                   context's first solve; 1-9: that layout without timing (1-3 one
                                   3x3 block per step at 4/8/16 lanes, 4-6 144-byte chunks at 4/8/16
                                   lanes, 7-9 the flat word stream at 16/32/64 lanes), so that two
                                   contexts do the same arithmetic */
} cmg_options;

typedef struct cmg_solve_stats {
    int32_t iterations;
    int32_t converged;
    double initial_residual;
    double final_residual;
} cmg_solve_stats;

typedef struct cmg_step_stats {
    int32_t newton_iterations;   /* SNESGetIterationNumber */
    int32_t krylov_iterations;   /* KSPGetTotalIterations */
    int32_t converged_reason;    /* >0 converged (2 atol, 3 rtol, 4 stol), <0 diverged (-3 the Krylov
                                    solve gave no usable update, -4 non-finite residual, -5
                                    Solver.MaxIterations, -6 an element stays inverted, -7 the
                                    residual stalled (cmg_options.newton_stall), -9
                                    ||r|| > 1e4 ||r_0||) */
    int32_t corrupt_element;     /* first element with det F <= 0, or -1 */
    double residual_norm;        /* final ||r||_2 */
    double initial_residual_norm;
    double step_norm;
    double wall_seconds;
    double tangent_seconds;
    double krylov_seconds;
    int32_t guess_source;        /* Newton initial guess: 0 predictor / rescaled previous increment,
                                    1 cmg_set_initial_guess, 2 proposed guess kept (its residual was
                                    below the predictor's), 3 proposed guess rejected, 4 the
                                    initial guess inverted an element and Newton restarted from a
                                    zero increment, 5 the pending trial increment
                                    (cmg_newmark_solve with guess 1) */
    int32_t reserved_s;
    double guess_residual_ratio; /* ||r(proposed)|| / ||r(predictor)|| when a guess was proposed */
} cmg_step_stats;

void cmg_default_options(cmg_options* opt);
const char* cmg_last_error(void);
int cmg_device_count(void);

/* Replaces CBSolver::Init's mesh distribution + CBSolver::PrepareElements
 * (CBSolver.cpp:1222-1275, 2157-2168) and CBModel::CalcNonZeroes (CBModel.cpp:43-83):
 * uploads the mesh, computes dN/dX, the block pattern and the assembly maps. */
int cmg_create(const cmg_mesh* mesh, const cmg_material* materials, int32_t n_materials,
               const cmg_options* options, cmg_ctx** out);
void cmg_destroy(cmg_ctx* ctx);

/* sizes: n_nodes, n_elements, n_dofs, number of nonzero 3x3 blocks */
int cmg_sizes(cmg_ctx* ctx, int64_t* n_nodes, int64_t* n_elements, int64_t* n_dofs, int64_t* nnz_blocks);
/* The fixed block-CSR pattern: rowptr [n_nodes+1], colind [nnz_blocks]; values of
 * every matrix of this ABI are [nnz_blocks][3][3] row-major blocks. */
int cmg_pattern(cmg_ctx* ctx, int64_t* rowptr, int32_t* colind);
int cmg_set_pressures(cmg_ctx* ctx, const double* pressures /* [n_groups] */);
int cmg_synchronize(cmg_ctx* ctx);
/* opaque stream handle (hipStream_t) the *_device entry points run on */
void* cmg_stream(cmg_ctx* ctx);

/* ---- Seam 1: element loop, CBFormulation::CalcNodalForces /
 *      CalcNodalForcesJacobian (CBFormulation.h:30-41, CBFormulationTotalLagrangian.cpp:18-40) */

/* Internal nodal forces of all solid elements at current coordinates x [3N],
 * summed per DOF, BC components zero (CBElementSolidT10.cpp:620-693 + adapter
 * AddNodalForcesComponents, CBElementAdapter.cpp:138-141).  Surface pressure
 * forces of the current pressures are added when include_pressure != 0
 * (CBCirculationCavity::ApplyPressure, CBCirculationCavity.cpp:108-172). */
int cmg_residual(cmg_ctx* ctx, const double* x, double time, int32_t include_pressure, double* f);
/* Assembled tangent d f / d x in the block pattern, identity on constrained DOFs
 * (CBSolver.cpp:413-435), values [nnz_blocks*9] (double; float in mixed precision). */
int cmg_tangent(cmg_ctx* ctx, const double* x, double time, int32_t include_pressure, void* values);
/* per-element outputs for parity checks: fe [E][3npe]; Ke [E][3npe][3npe] with
 * Ke[r][c] = d fe_r / d x_c (no BC masking; double or float per precision) */
int cmg_element_residuals(cmg_ctx* ctx, const double* x, double time, double* fe);
int cmg_element_tangents(cmg_ctx* ctx, const double* x, double time, void* Ke);
/* surface pressure forces [3N] at x for the given group pressures, and the
 * cavity volume per group about the origin (CBCirculationCavity::CalcVolume, :77-94) */
int cmg_surface_forces(cmg_ctx* ctx, const double* x, const double* pressures, double* f);
int cmg_cavity_volumes(cmg_ctx* ctx, const double* x, double* volumes);
/* consistent (or lumped) mass matrix in the block pattern
 * (CBElementSolidT10.cpp:352-422, CBSolverNewmarkBeta.cpp:704-787) */
int cmg_mass_matrix(cmg_ctx* ctx, double* values);
/* Per-element state for result export (CBSolver::Export, CBSolver.cpp:538-585), at x
 * (host [3N]) or, with x = NULL, at the current nodes.  out[18][E] (row r, element e
 * at out[r*E + e]): rows 0-8 the deformation gradient at the element centroid in the
 * global frame (CBElementSolidT10::GetDeformationTensor, CBElementSolidT10.cpp:764-777,
 * before its rotation into the fibre frame; T4: CBElementSolidT4.cpp:232-245), rows
 * 9-17 the passive PK2 stress in the fibre frame averaged over the quadrature points
 * (CBElementSolidT10::GetPK2Stress, :817-838), NaN for an element with det F <= 0.
 * J, fibre/sheet/normal stretch (Lambda), C, E and the rotated fibre follow from F
 * and the centroid basis on the host. */
int cmg_element_export(cmg_ctx* ctx, const double* x, double* out);

/* device-pointer variants (stream ordered) */
int cmg_residual_device(cmg_ctx* ctx, const double* x_dev, double time, int32_t include_pressure, double* f_dev);
int cmg_residual_tangent_device(cmg_ctx* ctx, const double* x_dev, double time, int32_t include_pressure,
                                double* f_dev, void* values_dev);
/* element kernel only: per-element residuals and tangents into the context's
 * element buffers (the unit measured as element residual+tangent throughput) */
int cmg_element_kernel_device(cmg_ctx* ctx, const double* x_dev, double time, int32_t with_tangent);

/* No counterpart in the reference: cmg_residual_device for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and result as cmg_residual_device, but its scratch (the
 * corrupt-element flag, the element forces [E][3 npe] and, with pressure, the surface
 * forces [S][3 nps]) is one fixed 64 MiB device array per GPU, shared by every call on
 * that GPU: meshes whose forces need more (E*3*npe + S*3*nps > 8 Mi - 32 doubles, about
 * 279k T10 elements) are rejected with CMG_ERR_INVALID, and calls that overlap on one GPU
 * race.  Synchronises the context's stream for the corrupt-element check. */
int cmg_residual_device_noalloc(cmg_ctx* ctx, const double* x_dev, double time, int32_t include_pressure,
                                double* f_dev);

/* Device copy of n doubles on the context's stream (a plain streaming kernel; used as
 * the known-size reference when calibrating HBM traffic counters). */
int cmg_copy_device(cmg_ctx* ctx, const double* src_dev, double* dst_dev, int64_t n);

/* ---- linear algebra on the context's pattern ---------------------------- */
int cmg_spmv(cmg_ctx* ctx, const double* values, const double* x, double* y);
/* the same product with an explicit kernel layout (the Krylov solvers pick the fastest
 * on their first solve): layout 0 one 3x3 block per lane step, 1 the values read in
 * 144-byte chunks (both with lanes = 4, 8 or 16 per block row), 2 the block row's
 * values read as one coalesced word stream (lanes = 16, 32 or 64) */
int cmg_spmv_layout(cmg_ctx* ctx, const double* values, const double* x, double* y, int32_t lanes,
                    int32_t layout);
/* the SpMV tuning of the context's first Krylov solve: ms[9] = time per product of
 * layouts 0, 1 at 4/8/16 lanes and layout 2 at 16/32/64 lanes; the chosen layout and
 * lanes (all zero before the first solve) */
int cmg_spmv_tuning(cmg_ctx* ctx, double* ms, int32_t* layout, int32_t* lanes);
/* BiCGStab or GMRES(m) (cmg_options.krylov_method), right preconditioned
 * (cmg_options.preconditioner); replaces the KSP/PC of CBSolverNewmarkBeta.cpp:166-196
 * (a Krylov solve instead of LU). */
int cmg_linear_solve(cmg_ctx* ctx, const double* values, const double* b, double* x, double rtol,
                     cmg_solve_stats* stats);

/* ---- Seam 2: solver step, CBSolver::SolverStep (CBSolver.h:214) ----------- */
/* Current node coordinates (CBSolver nodes_, the configuration at the start of the
 * next step), velocity and acceleration [3N] each; NULL leaves a field unchanged. */
int cmg_set_state(cmg_ctx* ctx, const double* nodes, const double* velocity, const double* acceleration);
int cmg_get_state(cmg_ctx* ctx, double* nodes, double* velocity, double* acceleration, double* displacement);
/* Initial Newton guess for the next step's increment (overrides the predictor /
 * rescaled-previous-increment guess once). */
int cmg_set_initial_guess(cmg_ctx* ctx, const double* increment);
/* CBSolverEquilibrium::SolverStep (CBSolverEquilibrium.cpp:265-363): quasi-static
 * Newton for the increment at time t with step dt; on success nodes += increment. */
int cmg_static_step(cmg_ctx* ctx, double time, double dt, cmg_step_stats* stats);
/* CBSolverNewmarkBeta::SolverStep (CBSolverNewmarkBeta.cpp:829-1019): predictor,
 * Newton on M(d - d~)/(beta dt^2) + f_int + f_p = 0, corrector. */
int cmg_newmark_step(cmg_ctx* ctx, double time, double dt, cmg_step_stats* stats);
/* Energies the Newmark corrector records (CBSolverNewmarkBeta.cpp:990-1003): the
 * kinetic energy v^T M v / 2 after the last committed step, and the Rayleigh damping
 * dissipation, the sum over committed steps of (C v_{n+1}) . d_{n+1} (0 without
 * damping); the reference skips that sum while preparing the simulation, this
 * library counts every committed step. */
int cmg_newmark_energies(cmg_ctx* ctx, double* kinetic, double* damping_dissipated);
/* The same step in two halves, for plugins that analyse the converged state before it
 * is accepted (CBSolverNewmarkBeta.cpp:928-976; CBCirculation::AnalyzeResults,
 * CBCirculation.cpp): cmg_newmark_solve runs the predictor and Newton and keeps the
 * converged increment pending without changing nodes, velocity or acceleration;
 * cmg_trial_volumes gives the cavity volume per pressure group at nodes + increment;
 * cmg_commit_step applies the corrector and nodes += increment.  A plugin's REPEAT is
 * another cmg_newmark_solve of the same step (nothing to undo).  guess: 0 the
 * predictor (as the reference), 1 the pending trial increment when its dt equals
 * this dt (else the predictor; stats->guess_source = 5 when it was used).
 * cmg_newmark_step = cmg_newmark_solve(guess 0) + cmg_commit_step. */
int cmg_newmark_solve(cmg_ctx* ctx, double time, double dt, int32_t guess, cmg_step_stats* stats);
int cmg_trial_volumes(cmg_ctx* ctx, double* volumes /* [n_groups] */);
int cmg_commit_step(cmg_ctx* ctx);

/* ---- Land17 tension model (CBTensionModelLand17.cpp) -------------------------
 * A material with tension_fn = CMG_FN_LAND17 carries one set of cross-bridge states per
 * element (XS XW TRPN TmBlocked ZetaS ZetaW Cd lambda): every evaluation at time t
 * advances them by one explicit Euler step of (t - t_commit) from the committed state,
 * at each quadrature point's fibre stretch (CalcActiveTension, :298-372), and the
 * accepted static or Newmark step commits the state of the element's last quadrature
 * point (SaveStateVariablesAsNeeded, :274-293).  Active PK2 kff T / lambda_f on ff.
 * params[40]: TRPNn TRPNk Ca50 ku nTM TRPN50 kuw kws rs rw gammas gammaw phi Aeff beta0
 * beta1 Tref a k eta_l eta_s xi (0-21, InitParamsFromXml :428-500), startTime
 * cycleLength, calcium transient (0 Coppini, 1 Brixius), rateDependancy (0 OFF, 1 ON),
 * kff (22-26), initial XS XW TRPN TmBlocked ZetaS ZetaW Cd (27-33); the reference's
 * defaults until set.  Setting a material's parameters resets its elements' state. */
int cmg_set_land17_parameters(cmg_ctx* ctx, int32_t material, const double* params);
/* state [8][E] (row k of element e at k*E + e) and the time of the last commit */
int cmg_get_land17_state(cmg_ctx* ctx, double* state, double* commit_time);
int cmg_set_land17_state(cmg_ctx* ctx, const double* state, double commit_time);

/* ---- Parameter-set batch ----------------------------------------------------
 * No counterpart in the reference, which solves one parameter set per run: S
 * copies of a context's quasi-static problem with their own material tables, one
 * mesh.  Each set follows cmg_static_step (CBSolverEquilibrium.cpp:234-363);
 * every kernel launch covers all sets, and the linear solve is a conjugate gradient,
 * preconditioned like the context's solves (vertex-patch Schwarz with one set of patch
 * inverses per parameter set, or 3x3 block Jacobi), whose scalars stay on the device
 * (runs of iterations are replayed as one HIP graph).  CG needs symmetric matrices: the
 * assembled follower-pressure tangent of a closed T3 cavity whose rim is clamped
 * (the LV endocardium) is symmetric and is included when the context's
 * pressure_tangent option is set; for open or T6 pressure surfaces create the
 * context with pressure_tangent = 0 (the reference's T3 tangent is zero,
 * CBElementSurfaceT3.cpp:331-372). */
typedef struct cmg_batch cmg_batch;

typedef struct cmg_batch_info {
    int32_t n_sets;
    int32_t element_variant;   /* chosen by the autotuner (or fixed by cmg_batch_configure) */
    int32_t spmv_lanes;        /* lanes per block row of the batched SpMV */
    int32_t cg_chunk;          /* CG iterations per graph replay */
    int32_t use_graph;
    int32_t tuned;             /* 1 once the autotuner has run */
    int32_t preconditioner;    /* of the batched CG: 1 block Jacobi, 2 vertex-patch Schwarz */
    int32_t reserved;
    double element_ms[8];      /* measured launch time of element variants 1..8 (0: not timed) */
    double spmv_ms[3];         /* measured batched SpMV time at 4, 8, 16 lanes */
    double tune_seconds;
} cmg_batch_info;

/* materials: [n_sets][n_materials of ctx], same order as at cmg_create; Guccione only
 * (CMG_ERR_INVALID otherwise: the Holzapfel tangent as coded is unsymmetric).  The batch
 * shares ctx's mesh, pattern and pressures (cmg_set_pressures on ctx); ctx must
 * outlive it.  Every set starts from ctx's current node coordinates. */
int cmg_batch_create(cmg_ctx* ctx, int32_t n_sets, const cmg_material* materials, cmg_batch** out);
int cmg_batch_destroy(cmg_batch* batch);
/* use_graph 0/1; spmv_lanes 0 (autotune) or 4/8/16; element_variant 0 (autotune) or 1..8;
 * cg_chunk >= 1.  Call before the first step. */
int cmg_batch_configure(cmg_batch* batch, int32_t use_graph, int32_t spmv_lanes, int32_t element_variant,
                        int32_t cg_chunk);
int cmg_batch_info_get(cmg_batch* batch, cmg_batch_info* info);
/* static load step of every set; stats: [n_sets] (may be NULL).  Returns CMG_SUCCESS
 * if every set converged, CMG_FAILED otherwise (failed sets keep their nodes, as
 * cmg_static_step). */
int cmg_batch_static_step(cmg_batch* batch, double time, double dt, cmg_step_stats* stats);
/* node coordinates of every set [n_sets][3N] (host) */
int cmg_batch_get_nodes(cmg_batch* batch, double* nodes);
int cmg_batch_set_nodes(cmg_batch* batch, const double* nodes);
/* assembled residual (internal + pressure forces) of every set at its own
 * coordinates x [n_sets][3N] (host) -> f [n_sets][3N] (host) */
int cmg_batch_residual(cmg_batch* batch, const double* x, double time, double* f);
/* cavity volumes of every set at its current nodes [n_sets][n_groups] */
int cmg_batch_cavity_volumes(cmg_batch* batch, double* volumes);

/* ---- Learned Newton warm start -----------------------------------------------
 * No counterpart in the reference (its Newmark Newton starts from the predictor
 * d~, CBSolverNewmarkBeta.cpp:845-855).  A per-node MLP 8 -> 128 -> 128 -> 3 (tanh)
 * maps the features of the next Newmark step of a context
 *   x = (s_inc * increment of the last step [3], s_inc * increment of the step
 *        before [3], s_p * pressure rate, s_t * (time - mean activation time of
 *        the node's elements))
 * to s_inc * that step's increment.  Trained on the context's own trajectories by
 * the library's kernels (FP32, mean squared error, Adam).  All float arrays of
 * these entry points are DEVICE pointers, row-major [rows][8] / [rows][3]. */
typedef struct cmg_mlp cmg_mlp;
int32_t cmg_mlp_num_weights(int32_t hidden);   /* -1 unless hidden == 128 */
/* Glorot-uniform weights from a seeded generator; runs on ctx's device and stream */
int cmg_mlp_create(cmg_ctx* ctx, int32_t hidden, uint64_t seed, cmg_mlp** out);
int cmg_mlp_destroy(cmg_mlp* mlp);
/* flat FP32 weights W1[8][128] b1[128] W2[128][128] b2[128] W3[128][3] b3[3] (host arrays) */
int cmg_mlp_get_weights(cmg_mlp* mlp, float* weights);
int cmg_mlp_set_weights(cmg_mlp* mlp, const float* weights);   /* also resets the Adam state */
int cmg_mlp_set_scales(cmg_mlp* mlp, double inc_scale, double pressure_rate_scale, double phase_scale);
/* features of the next step at `time` (before cmg_newmark_step) -> X [N][8] */
int cmg_mlp_features(cmg_mlp* mlp, double time, double pressure_rate, float* X);
/* scaled increment of the last accepted Newmark step -> T [N][3] */
int cmg_mlp_targets(cmg_mlp* mlp, float* T);
/* Y = net(X); bf16 = 1: inputs, weights and hidden activations in bfloat16, the 128 x 128
 * layer on the matrix cores (FP32 accumulation) */
int cmg_mlp_forward(cmg_mlp* mlp, const float* X, int64_t rows, int32_t bf16, float* Y);
/* mean squared error of net(X) against T and its gradient (kept for cmg_mlp_adam_step;
 * copied to grad [num_weights] on the host when grad is not NULL) */
int cmg_mlp_loss_grad(cmg_mlp* mlp, const float* X, const float* T, int64_t rows, double* loss, float* grad);
/* one Adam update (beta1 0.9, beta2 0.999, eps 1e-8) with the last gradient */
int cmg_mlp_adam_step(cmg_mlp* mlp, float lr);
/* propose Y [N][3] (unscaled by 1/s_inc, constrained DOFs zeroed) as the Newton
 * initial guess of the next cmg_newmark_step, which keeps it only when its residual
 * norm is below the predictor's (cmg_step_stats.guess_source 2 or 3) */
int cmg_mlp_propose_guess(cmg_mlp* mlp, const float* Y);

#ifdef __cplusplus
}
#endif
#endif /* CMG_H */
