/*
 * exahype_amd -- MI355X (gfx950) ADER-DG and finite-volume kernels behind a
 * C ABI.
 *
 * Two layers:
 *
 *  1. Per-cell / per-face drop-ins for the kernel entry points of the
 *     reference's solver plugin surface.  Arrays are host pointers in the
 *     reference's own layouts (row-major, variable index fastest, nData =
 *     nVar + nPar entries per node for state arrays, nVar for fluxes):
 *       luh    [z][y][x][nData]            lduh [z][y][x][nVar]
 *       lQhbnd [face][a][b][nData]         lFhbnd [face][a][b][nVar]
 *       face order x-,x+,y-,y+,z-,z+; (a,b) = (z,y), (z,x), (y,x)
 *       lQi [z][y][x][t][nData]            lFi [t][z][y][x][4][nVar]
 *       lQhi [z][y][x][nData]              lFhi = x|y|z|source blocks of B^3*nVar,
 *                                          x:[z][y][x], y:[z][x][y], z:[y][x][z]
 *     Each call is synchronous and re-entrant (one stream per solver handle,
 *     serialised by a mutex).
 *
 *  2. A device-resident grid: cells of a regular periodic mesh kept in HBM in
 *     structure-of-arrays blocks, advanced by whole time steps (predictor,
 *     Riemann solves, surface integral + update + CFL reduction) without host
 *     round trips.  The step is the reference's non-fused stepping
 *     (E/runners/Runner.cpp:1125-1190) or, on request, its fused stepping's
 *     time step control (exa_grid_set_fused_time_stepping).
 *
 * Every entry point returns EXA_OK (0) or a negative status; exa_last_error()
 * gives the message of the last failure on the calling thread.  No entry
 * point falls back to a CPU implementation: without a usable gfx950 device
 * they return EXA_ERR_HIP.
 */
#ifndef EXAHYPE_AMD_H
#define EXAHYPE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define EXA_OK 0
#define EXA_ERR_ARG (-1)
#define EXA_ERR_HIP (-2)
#define EXA_ERR_UNSUPPORTED (-3)
#define EXA_ERR_NOMEM (-4)
#define EXA_ERR_NUMERICS (-5) /* the state is not finite or has a negative pressure: its stable dt
                                  is not a positive number (reported by the dt and sync calls)  */

/* predictor precision */
#define EXA_PREC_FP64 0       /* plain method: every operation in FP64                        */
#define EXA_PREC_FP32 1       /* Picard sweeps in FP32 (vector ALU), then FP64 correction sweeps */
#define EXA_PREC_FP32_MFMA 2  /* sweeps' spatial contractions on FP32 MFMA, FP32 elsewhere, then
                                 FP64 correction sweeps                                        */
#define EXA_PREC_BF16_MFMA 3  /* as EXA_PREC_FP32_MFMA with BF16 contraction operands (FP32
                                 accumulation)                                                 */
/* On the adaptive mesh (exa_amr_grid_*) with EXA_PREC_FP32, a cell whose
   nodes differ from one state by less than FP32 resolves (4 eps max|u|)
   takes the FP64 sweeps alone, with the FP64 predictor's stopping rule and
   outputs. */

/* PDEs compiled into the kernels */
#define EXA_PDE_EULER 0   /* 3D compressible Euler, 5 variables (gamma from config) */
#define EXA_PDE_EULER_GRAVITY 2  /* 3D Euler with constant gravity g along -z: algebraic
                                    source S = (0, 0, 0, -rho g, -m_z g), the useSource arm of
                                    spaceTimePredictorNonlinear.cpph:227-252 and of
                                    volumeIntegralNonlinear.cpph (FP64 predictor only) */
/* Linear advection of 5 variables with the velocity of the config, in the
 * flux form F_d = a_d Q and in the non-conservative form (no flux,
 * nonConservativeProduct = a . grad Q): the useFlux = false / useNCP = true
 * arms of spaceTimePredictorNonlinear.cpph:42-252 and
 * riemannSolverNonlinear.cpph:95-113.  FP64, orders 1..7. */
#define EXA_PDE_ADVECTION 3
#define EXA_PDE_ADVECTION_NCP 4
/* 3D Euler with the test non-conservative product of the reference's Euler
 * kernel test, BgradQ = (Q0, Q3, 3, dQ0/dx, 0.7)
 * (E/tests/kernels/c/GenericEulerKernelTest3D.cpp:84-106): the useNCP arm on
 * a nonlinear flux.  FP64, orders 1..7. */
#define EXA_PDE_EULER_NCP_TEST 5

const char* exa_last_error(void);

/* Version string of the library build. */
const char* exa_version(void);

/* Selects the device for the calling thread and uploads the DG tables
 * (idempotent per device). */
int exa_init(int device);

/* Waits for all work on the current device (every stream). */
int exa_device_synchronize(void);

/* Device memory of the current device for callers without their own HIP
 * allocator (the batched entry points take device pointers).  No counterpart
 * in the reference, whose cell data lives in host heaps. */
int exa_device_malloc(void** ptr, long long bytes);
int exa_device_free(void* ptr);
int exa_memcpy_htod(void* dst, const void* src, long long bytes);
int exa_memcpy_dtoh(void* dst, const void* src, long long bytes);
/* Device-to-device copy by a plain streaming kernel (32-byte aligned
 * pointers); *ms receives its event-timed duration.  The HBM reference of
 * the benchmarks. */
int exa_device_copy(void* dst, const void* src, long long bytes, float* ms);

/* Self-test of the matrix-core operand layouts used by the mixed-precision
 * predictor: D = A * B for 16x16 row-major FP32 matrices through one
 * wavefront of 16x16x4 FP32 MFMAs (kind 0) or one 16x16x16 BF16 MFMA (kind 1,
 * operands rounded to BF16).  No counterpart in the reference. */
int exa_selftest_mfma(int kind, const float* A, const float* B, float* D);

/* ---------------------------------------------------------------------- */
/* Solver handle: the compile-time constants of the reference's generated
 * Abstract<Solver> (TT/AbstractGenericADERDGSolverHeader.template:31-34). */
typedef struct exa_solver exa_solver;

typedef struct {
  int order;        /* N, 1..9 (8, 9: EXA_PREC_FP64 or EXA_PREC_FP32) */
  int nvar;         /* must be 5 for EXA_PDE_EULER                 */
  int npar;         /* material parameters (passed through)        */
  double cfl;       /* CFL factor (0.9 in the reference)           */
  double gamma;     /* ratio of specific heats                     */
  int pde;          /* EXA_PDE_*                                   */
  int precision;    /* EXA_PREC_*                                  */
  int fp64_sweeps;  /* mixed precisions: FP64 Picard sweeps after the low-precision phase (>=1) */
  int waves_per_simd; /* predictor occupancy target: 0 = tuned default per order, 1 or 2 */
  int contraction;  /* EXA_CONTRACT_*: form of the sweeps' spatial contraction */
  double gravity;   /* g of EXA_PDE_EULER_GRAVITY (ignored otherwise)          */
  double velocity[3]; /* advection velocity of EXA_PDE_ADVECTION(_NCP) (ignored otherwise) */
  int linear;       /* 1: the linear kernel family, the generated glue's isLinear branch
                       (TT/AbstractGenericADERDGSolverImplementation.template:55-192):
                       Cauchy-Kowalevski predictor (spaceTimePredictorLinear.cpph),
                       volumeIntegralLinear, surfaceIntegralLinear (fluctuations),
                       riemannSolverLinear (coefficient matrix).  For
                       EXA_PDE_ADVECTION_NCP and EXA_PDE_EULER_NCP_TEST, FP64,
                       orders 1..9.  For a linear solver, spaceTimePredictor's lQi
                       receives the B+1 time derivatives [t][z][y][x][nData] and lFi
                       the products [t][z][y][x][3][nVar]; lFhi is [4][z][y][x][nVar]
                       (block 3 zero) and volumeIntegral reads its first 3 blocks. */
  int point_source; /* 1 (linear solvers only): usePointSource -- the point-source arm of
                       spaceTimePredictorLinear.cpph:78-102, 287-302, with the user
                       pointSource (E/solvers/UserSolverInterface.h:91-95) compiled as a
                       Ricker wavelet force source_amp[v] (1 - 2a) exp(-a),
                       a = (pi source_f0 (t - source_t0))^2, located at source_x0.  Grids
                       evaluate it per node at the time of each step;
                       exa_aderdg_point_source gives the forces of one cell. */
  double source_x0[3];
  double source_amp[5];
  double source_f0;
  double source_t0;
} exa_aderdg_config;

#define EXA_CONTRACT_DEFAULT 0  /* tuned per order (the staged form for FP64 at N = 7)    */
#define EXA_CONTRACT_PENCIL 1   /* pencil owners contract, two barriers per time slice    */
#define EXA_CONTRACT_NODE 2     /* node owners contract their pencils, one barrier per slice */
#define EXA_CONTRACT_PENCIL_PAIR 3 /* pencil owners, two slices per phase, contraction in place */
#define EXA_CONTRACT_STAGED 4   /* space-time iterate staged in global memory (FP64, or
                                   EXA_PREC_FP32: an FP32 phase on FP32 staging arrays, then
                                   FP64 correction sweeps; orders >= 5; the only form of
                                   orders 8 and 9)                                          */

int exa_aderdg_create(const exa_aderdg_config* cfg, exa_solver** out);

/* Autotuner: times the predictor of one order and precision in every launch
 * variant (1 or 2 waves per SIMD x EXA_CONTRACT_PENCIL, _NODE, _PENCIL_PAIR)
 * on an ncells^3 vortex grid and records the fastest for this device; solvers
 * created later with waves_per_simd = 0 and contraction = 0 use it.  Reports
 * the choice and the six average launch times (ms, waves-major: w1-pencil,
 * w1-node, w1-pair, w2-pencil, w2-node, w2-pair).  No counterpart in the
 * reference. */
int exa_autotune(int order, int precision, int ncells, int* waves_per_simd, int* contraction, double ms[6]);
int exa_aderdg_destroy(exa_solver* s);

/* ADERDGSolver::spaceTimePredictor (E/solvers/ADERDGSolver.h:1049-1059;
 * kernel kernels/aderdg/generic/c/3d/spaceTimePredictorNonlinear.cpph:524-577).
 * lQhi/lFhi receive what the reference writes into tempUnknowns and
 * tempFluxUnknowns.  lQi, lFi, lQhi, lFhi and picard_iterations may be NULL. */
int exa_aderdg_space_time_predictor(exa_solver* s, double* lQhbnd, double* lFhbnd,
                                    double* lQi, double* lFi, double* lQhi, double* lFhi,
                                    const double* luh, const double dx[3], double dt,
                                    int* picard_iterations);

/* ADERDGSolver::spaceTimePredictor with its pointForceSources argument
 * (E/solvers/ADERDGSolver.h:1092-1103): point_forces, the forces of
 * exa_aderdg_point_source [B+1][z][y][x][nVar], enter the linear predictor's
 * point-source arm (spaceTimePredictorLinear.cpph:78-102, 287-302); NULL
 * behaves as exa_aderdg_space_time_predictor.  Non-NULL point_forces need a
 * linear solver. */
int exa_aderdg_space_time_predictor_ps(exa_solver* s, double* lQhbnd, double* lFhbnd,
                                       double* lQi, double* lFi, double* lQhi, double* lFhi,
                                       const double* luh, const double dx[3], double dt,
                                       const double* point_forces, int* picard_iterations);

/* ADERDGSolver::pointSource (E/solvers/ADERDGSolver.h:1195-1200; kernel
 * kernels/aderdg/generic/c/3d/pointSource.cpph:28-145): the point forces of
 * the cell with the given centre and size at time t, step dt, projected onto
 * the nodal basis: point_forces [B+1][z][y][x][nVar], level 0 at t and level
 * n >= 1 at the GL time node n-1 of [t, t+dt].  Needs config point_source. */
int exa_aderdg_point_source(exa_solver* s, double t, double dt, const double centre[3], const double dx[3],
                            double* point_forces);

/* ADERDGSolver::volumeIntegral (E/solvers/ADERDGSolver.h:1015-1017;
 * kernels/aderdg/generic/c/3d/volumeIntegralNonlinear.cpph:26-125) */
int exa_aderdg_volume_integral(exa_solver* s, double* lduh, const double* lFhi, const double dx[3]);

/* ADERDGSolver::surfaceIntegral (E/solvers/ADERDGSolver.h:1027-1029;
 * kernels/aderdg/generic/c/3d/surfaceIntegralNonlinear.cpph:21-86) */
int exa_aderdg_surface_integral(exa_solver* s, double* lduh, const double* lFhbnd, const double dx[3]);

/* ADERDGSolver::riemannSolver (E/solvers/ADERDGSolver.h:1073-1080;
 * kernels/aderdg/generic/c/3d/riemannSolverNonlinear.cpph:38-120).  FL and FR
 * enter holding the extrapolated normal fluxes and leave holding the
 * Rusanov flux. */
int exa_aderdg_riemann_solver(exa_solver* s, double* FL, double* FR, const double* QL,
                              const double* QR, double dt, int normalNonZero);

/* ADERDGSolver::solutionUpdate (E/solvers/ADERDGSolver.h:1005;
 * kernels/aderdg/generic/c/3d/solutionUpdate.cpph:23-47) */
int exa_aderdg_solution_update(exa_solver* s, double* luh, const double* lduh, double dt);

/* ADERDGSolver::stableTimeStepSize (E/solvers/ADERDGSolver.h:1112-1115;
 * kernels/aderdg/generic/c/3d/stableTimeStepSize.cpph:21-61) */
int exa_aderdg_stable_time_step_size(exa_solver* s, const double* luh, const double dx[3], double* dt);

/* ---------------------------------------------------------------------- */
/* Initial conditions evaluated at the Gauss-Legendre nodes, the user
 * adjustPointSolution of the reference (E/solvers/ADERDGSolver.h:1183;
 * kernels/aderdg/generic/c/3d/solutionAdjustment.cpph:17-52). */
typedef struct {
  int kind;              /* 0 isentropic vortex + density modes, 1 Sod sphere,
                            2 constant state, 3 sine modes Q_v = 1 + (v+1) sum_m a_m
                            sin(2 pi k_m.(x - uinf t) + phi_m) (linear-advection PDEs),
                            4 entropy wave,
                            5 Gaussian density pulse rho_inf + amp exp(-|x-xc|^2/L) */
  double xc[3];          /* vortex / sphere centre                       */
  double L, eps;         /* vortex length scale and strength             */
  double uinf[3];        /* background velocity                          */
  double rho_inf, p_inf; /* background density and pressure              */
  double amp;            /* density-mode amplitude                       */
  int nmodes;            /* <= 8                                         */
  double k[8][3], a[8], phi[8];
  double r0, rho_in, p_in, rho_out, p_out;
  double q0[8];
} exa_ic;

/* ADERDGSolver::adjustSolution with a pointwise initial condition. */
int exa_aderdg_adjust_solution(exa_solver* s, double* luh, const double centre[3],
                               const double dx[3], double t, const exa_ic* ic);

/* Boundary values of the Euler solvers, the user boundaryValues that the
 * generated solver calls per face node (GenericADERDGSolverInCUserCode.template:44-52):
 *   EXA_BC_OUTFLOW    outside state and normal flux = the inside ones (the
 *                     EulerFV demonstrator's choice, Demonstrators/EulerFV/MyEulerSolver.cpp:87-98);
 *   EXA_BC_REFLECTING the inside state with its normal momentum negated, and
 *                     that state's normal flux (a slip wall);
 *   EXA_BC_EXACT      the state of `exact` at the face node, state and normal
 *                     flux averaged over [t, t+dt] with the order's
 *                     Gauss-Legendre rule.
 * REFLECTING and EXACT need the Euler PDEs (EXA_PDE_EULER, EXA_PDE_EULER_GRAVITY). */
#define EXA_BC_OUTFLOW 1
#define EXA_BC_REFLECTING 2
#define EXA_BC_EXACT 3
typedef struct {
  int kind;
  exa_ic exact; /* EXA_BC_EXACT only */
} exa_bc;

/* EXA_BC_USER: the boundaryValues of a PDE registered through
 * exahype_amd_pde.h (exa_pde_register). */
#define EXA_BC_USER 4

/* ADERDGSolver::boundaryConditions (E/solvers/ADERDGSolver.h:1061-1070;
 * kernels/aderdg/generic/c/3d/boundaryConditions.cpph:24-85): the outside
 * state QOut [B^2][nData] and normal flux FOut [B^2][nVar] of one domain
 * face (faceIndex 0..5 = 2 normalNonZero + side) of the cell at centre with
 * extent size, from the inside QIn, FIn; node (i, j) sits at the face's
 * GL point with i along the slower, j along the faster tangential axis. */
int exa_aderdg_boundary_conditions(exa_solver* s, const exa_bc* bc, double* FOut, double* QOut, const double* FIn,
                                   const double* QIn, const double centre[3], const double size[3], double t,
                                   double dt, int faceIndex, int normalNonZero);

/* Deferred per-cell entry points.  The reference calls its kernels cell by
 * cell inside one mesh traversal and reads the results only in the next one;
 * these calls copy their inputs, queue the work and return, and
 * exa_aderdg_batch_flush -- called where the traversal ends (the mappings'
 * endIteration, e.g. E/mappings/Prediction.cpp:105-108) -- runs every queued
 * item of a kind in one launch and writes the outputs to the arrays given at
 * queue time (which must stay valid until then).  Kinds run in the order
 * prediction, Riemann, update; queued items must not depend on each other.
 *   batch_prediction: spaceTimePredictor + volumeIntegral of one cell
 *     (performPredictionAndVolumeIntegral, E/solvers/ADERDGSolver.cpp:1818-1946):
 *     lQhbnd, lFhbnd and lduh of luh over dt (the predictor time step).
 *   batch_riemann: riemannSolver of one face (solveRiemannProblemAtInterface,
 *     E/solvers/ADERDGSolver.cpp:2744-2825; arguments as exa_aderdg_riemann_solver).
 *   batch_update: surfaceIntegral + solutionUpdate of one cell (updateSolution,
 *     E/solvers/ADERDGSolver.cpp:2100-2179): luh += dt/w * (lduh - surface
 *     fluxes); lduh receives the surface-integrated update.
 * Same layouts as the per-cell calls above. */
int exa_aderdg_batch_prediction(exa_solver* s, double* lQhbnd, double* lFhbnd, double* lduh, const double* luh,
                                const double dx[3], double dt);
int exa_aderdg_batch_riemann(exa_solver* s, double* FL, double* FR, const double* QL, const double* QR, double dt,
                             int normalNonZero);
int exa_aderdg_batch_update(exa_solver* s, double* luh, double* lduh, const double* lFhbnd, const double dx[3],
                            double dt);
/* queued items per kind: counts[0] predictions, [1] Riemann, [2] updates */
int exa_aderdg_batch_pending(exa_solver* s, long long counts[3]);
int exa_aderdg_batch_flush(exa_solver* s);

/* ---------------------------------------------------------------------- */
/* PDEs compiled by a project into the kernels (exahype_amd_pde.h).  No
 * counterpart in the C ABI of the reference, whose generated solver passes
 * the user class to the kernel templates (TT/GenericADERDGSolverHeader.template:29-128):
 * here the project's translation unit instantiates the kernel templates for
 * its PDE and hands the kernels to the library, which dispatches to them for
 * the returned PDE id.  Fill the table with exa::register_pde<PDE>() rather
 * than by hand; the entries are __global__ function pointers of that unit,
 * indexed by order (1..max_order), and init uploads the DG tables given to it
 * into the unit's constant memory on the current device. */
#define EXA_PDE_USER_FIRST 16  /* ids of registered PDEs: EXA_PDE_USER_FIRST + slot */
#define EXA_PDE_USER_SLOTS 8
#define EXA_PDE_KERNELS_VERSION 2
typedef struct {
  int version;    /* EXA_PDE_KERNELS_VERSION */
  int max_order;  /* highest order instantiated (<= 7) */
  int (*init)(const void* tables);
  const void* predictor[10][2];  /* [order][debug]: FP64 Picard predictor + volume integral */
  const void* riemann[10];       /* every face of a grid */
  const void* riemann_list[10];  /* a list of independent faces */
  const void* update_dt[10];     /* CFL time step */
  const void* update_full[10];   /* surface integral + update + CFL time step */
  const void* initial_conditions[10]; /* adjustPointSolution on a grid */
  const void* boundary_faces[10];     /* boundaryValues + Riemann on a grid's domain faces */
  const void* boundary_ref[10];       /* boundaryValues of one face, reference layout */
  int predictor_lds_bytes[10];
} exa_pde_kernels;
/* Registers (or, under a name already registered, replaces) a PDE; *pde
 * receives its id.  Call after exa_init on every device that will run it. */
int exa_pde_register(const char* name, const exa_pde_kernels* kernels, int* pde);
/* The id of a registered PDE (EXA_ERR_ARG if the name is unknown). */
int exa_pde_lookup(const char* name, int* pde);

/* ---------------------------------------------------------------------- */
/* Device-resident periodic grid. */
typedef struct exa_grid exa_grid;

int exa_grid_create(exa_solver* s, int nx, int ny, int nz, const double offset[3],
                    const double width[3], exa_grid** out);
int exa_grid_destroy(exa_grid* g);
int exa_grid_num_cells(const exa_grid* g, long long* n);
/* luh of every cell, cells ordered (z,y,x) with x fastest, each in the
 * reference layout [z][y][x][nData]. */
int exa_grid_upload(exa_grid* g, const double* luh);
int exa_grid_download(exa_grid* g, double* luh);
/* ic may be null for a registered PDE: its adjustPointSolution then sets the state. */
int exa_grid_set_initial_conditions(exa_grid* g, const exa_ic* ic);
/* Domain boundaries: axis d stays periodic where periodic[d] != 0 (the
 * default); on the two domain faces of every other axis each step calls
 * boundaryConditions with bc and solves the Riemann problem with the
 * outside state on the outer side (applyBoundaryConditions,
 * E/solvers/ADERDGSolver.cpp:2854-2931).  The Euler PDEs take the
 * EXA_BC_OUTFLOW..EXA_BC_EXACT kinds, a registered PDE EXA_BC_USER (its
 * boundaryValues); other PDEs stay periodic (EXA_ERR_ARG).  Rebuilds the
 * step's graph. */
int exa_grid_set_boundary(exa_grid* g, const int periodic[3], const exa_bc* bc);
/* The values a Cartesian-vertices plotter writes
 * (exahype::plotters::ADERDG2CartesianVTK::plotVertexData,
 * E/plotters/VTK/ADERDG2CartesianVTK.cpp:288-336): the solution of every
 * cell at its (N+1)^3 equidistant vertices offset + (i, j, k) dx / N,
 * computed on the device; out (host) [cell][vertex, x fastest][nVar]. */
int exa_grid_sample_cartesian(exa_grid* g, double* out);
/* Stable dt of the current state (global minimum over cells); also
 * becomes the dt of the next step.  Synchronous. */
int exa_grid_compute_dt(exa_grid* g, double* dt);
/* Overrides the dt of the next step (with fused stepping also the
 * optimistic one). */
int exa_grid_set_dt(exa_grid* g, double dt);
/* Fused time stepping, the reference's `fuse-algorithmic-steps = on` with
 * `fuse-algorithmic-steps-factor = weight` in (0, 1] (global time stepping;
 * exahype::runners::Runner::runOneTimeStepWithFusedAlgorithmicSteps,
 * E/runners/Runner.cpp:1069-1122; the time step data of
 * TimeStepSizeComputation.cpp:144-187, ADERDGSolver.cpp:556-570 and
 * FusedTimeSteppingInitialisation.cpp:80-100).  Starting (here with a dt in
 * hand, or at the next exa_grid_compute_dt) the first step takes
 * 0.5 (s + weight s), s the stable dt; after each step, with `used` the
 * step's optimistic predictor dt and `stable` that of the new state: used >
 * stable -> the next step takes weight * stable (where the reference reruns
 * the predictor; counted by exa_grid_fused_reruns), otherwise it takes
 * `used` and the next optimistic dt is 0.5 (used + weight stable).  The
 * predictor runs once per step with the final dt: the minimum over the
 * cells is known on the device before it is launched.  enable = 0 returns
 * to the step-by-step stable dt. */
int exa_grid_set_fused_time_stepping(exa_grid* g, int enable, double weight);
/* Steps since fused stepping started whose optimistic dt exceeded the stable
 * one (the reference's predictor reruns). */
int exa_grid_fused_reruns(exa_grid* g, long long* reruns);
/* Enqueues nsteps time steps; each uses the stable dt of its start state
 * (fused stepping: the dt chosen as above). */
int exa_grid_step(exa_grid* g, int nsteps);
/* Waits for queued work; returns simulated time and the dt of the next step. */
int exa_grid_sync(exa_grid* g, double* time, double* next_dt);
/* Picard iterations executed by the last predictor of each cell (int per cell). */
int exa_grid_picard_iterations(exa_grid* g, int* iters);
/* 1: replay each time step as one captured hipGraph; 0: plain launches. */
int exa_grid_use_graph(exa_grid* g, int enable);
/* Kernel timing: when enabled, HIP events bracket every launch of each
 * phase on the grid's stream; times[k] = summed milliseconds, counts[k] =
 * launches, for k = 0 predictor, 1 riemann, 2 update(+surface+dt).  The
 * Euler PDEs solve every interior face's Riemann problem inside the update
 * launch (one kernel per step for Riemann, surface integral, update and
 * CFL): phase 1 then holds the domain faces of closed axes only.
 * Only meaningful with graphs disabled. */
int exa_grid_enable_timing(exa_grid* g, int enable);
int exa_grid_kernel_times(exa_grid* g, double times[3], long long counts[3]);
/* Picard sweeps summed over all cells and all steps since the previous call
 * (a device counter; the call resets it). */
int exa_grid_total_picard_iterations(exa_grid* g, long long* total);

/* ---------------------------------------------------------------------- */
/* Ensemble of independent periodic grids of uneven sizes, orders and gammas.
 * The instances of one order share every launch of a step (one predictor,
 * one Riemann and one update launch over all their cells; each cell reads
 * its instance's dt, mesh width and gamma), and a step of all orders is one
 * hipGraph.  Each instance keeps its own time and CFL time step.  No single
 * counterpart in the reference, which runs one grid per process. */
typedef struct exa_ensemble exa_ensemble;

typedef struct {
  int order;          /* 1..7                       */
  int nx, ny, nz;     /* cells per direction        */
  double offset[3];   /* domain origin              */
  double width[3];    /* domain extent              */
  double gamma;       /* ratio of specific heats    */
} exa_ensemble_member;

/* cfg supplies precision, fp64_sweeps, cfl, waves_per_simd and contraction
 * (order, gamma and the PDE fields are taken per member).  Fails with
 * EXA_ERR_NOMEM when the members do not fit the free device memory. */
int exa_ensemble_create(const exa_ensemble_member* members, int n, const exa_aderdg_config* cfg,
                        exa_ensemble** out);
int exa_ensemble_destroy(exa_ensemble* e);
/* ics: one initial condition per member */
int exa_ensemble_set_initial_conditions(exa_ensemble* e, const exa_ic* ics);
/* stable dt of every member's current state becomes its next dt */
int exa_ensemble_compute_dt(exa_ensemble* e);
int exa_ensemble_use_graph(exa_ensemble* e, int enable);
int exa_ensemble_step(exa_ensemble* e, int nsteps);
/* times/dts: per member (either may be NULL); waits for the ensemble's work */
int exa_ensemble_sync(exa_ensemble* e, double* times, double* dts);
/* one member's luh, cells ordered (z,y,x), reference layout per cell */
int exa_ensemble_download(exa_ensemble* e, int member, double* luh);
/* Picard sweeps summed over all cells and steps since the last call */
int exa_ensemble_total_picard_iterations(exa_ensemble* e, long long* total);

/* ---------------------------------------------------------------------- */
/* Finite volumes: the kernel surface of exahype::solvers::FiniteVolumesSolver
 * (E/solvers/FiniteVolumesSolver.h:235-400).  A patch is stored with its
 * ghost layers, reference layout [(z)][y][x][nData], side patch+2*ghost. */
#define EXA_FV_GODUNOV 0        /* first order, ghost width 1 (2D and 3D)        */
#define EXA_FV_MUSCLHANCOCK 1   /* second order, ghost width 2, minmod (3D only) */
#define EXA_PDE_EULERFV2D 1     /* the EulerFV demonstrator's 2D Euler (rho, j[3], E) */

typedef struct exa_fv_solver exa_fv_solver;

typedef struct {
  int dim;        /* 2 or 3                                        */
  int patch;      /* subcells per direction (PatchSize)            */
  int scheme;     /* EXA_FV_GODUNOV | EXA_FV_MUSCLHANCOCK          */
  int nvar;       /* 5                                             */
  int npar;       /* 0                                             */
  double cfl;     /* CFL of stableTimeStepSize (0.9 in the reference) */
  double gamma;
  int pde;        /* EXA_PDE_EULER (3D) | EXA_PDE_EULERFV2D (2D)   */
} exa_fv_config;

int exa_fv_create(const exa_fv_config* cfg, exa_fv_solver** out);
int exa_fv_destroy(exa_fv_solver* s);

/* FiniteVolumesSolver::solutionUpdate (E/solvers/FiniteVolumesSolver.h:300-306;
 * kernels/finitevolumes/godunov/c/{2d,3d}/godunov.cpph,
 * kernels/finitevolumes/musclhancock/c/3d/musclhancock.cpph).  luh_new holds a
 * copy of luh on entry, as FiniteVolumesSolver::updateSolution passes it
 * (E/solvers/FiniteVolumesSolver.cpp:519-583); max_admissible_dt receives the
 * kernel's face-speed time step (MUSCL-Hancock uses its own CFL 0.99). */
int exa_fv_solution_update(exa_fv_solver* s, double* luh_new, const double* luh, const double dx[3], double dt,
                           double* max_admissible_dt);

/* FiniteVolumesSolver::stableTimeStepSize (E/solvers/FiniteVolumesSolver.h:252-255;
 * kernels/finitevolumes/commons/c/commons.cpph:26-67, :410-463) */
int exa_fv_stable_time_step_size(exa_fv_solver* s, const double* luh, const double dx[3], double* dt);

/* The boundary hooks of one patch (luh with ghost layers, reference layout
 * [(z)][y][x][nData]); positions are unit vectors towards the face, faces
 * blocks hold G layers: 3D x faces [z][y][G], y faces [z][G][x], z faces
 * [G][y][x]; 2D x faces [y][G], y faces [G][x] (each of nData).
 * FiniteVolumesSolver::ghostLayerFilling (E/solvers/FiniteVolumesSolver.h:264-267;
 * kernels/finitevolumes/commons/c/commons.cpph:338-403, :877-982): my ghost
 * layers from the neighbour's interior layers next to the shared face. */
int exa_fv_ghost_layer_filling(exa_fv_solver* s, double* luh, const double* luhNeighbour,
                               const int neighbourPosition[3]);
/* ::ghostLayerFillingAtBoundary (FiniteVolumesSolver.h:277-280; commons.cpph:267-336, :615-725) */
int exa_fv_ghost_layer_filling_at_boundary(exa_fv_solver* s, double* luh, const double* luhbnd,
                                           const int boundaryPosition[3]);
/* ::boundaryLayerExtraction (FiniteVolumesSolver.h:289-292; commons.cpph:100-173, :503-613),
 * including the reference's 3D y- and z- faces, which copy layer
 * patchBegin into every layer of the block (commons.cpph:561, :591). */
int exa_fv_boundary_layer_extraction(exa_fv_solver* s, double* luhbnd, const double* luh,
                                     const int boundaryPosition[3]);
/* ::boundaryConditions (FiniteVolumesSolver.h:306-312; commons.cpph:175-265,
 * :727-875): stateOut layer k = boundaryValues of stateIn layer G-1-k at the
 * outside subcell centre the reference computes (the 3D x+ face scales its
 * y coordinate by size[2], commons.cpph:780).  FV boundary values: outflow
 * copies, reflecting negates the normal momentum, exact gives the state of
 * bc->exact at (x, t). */
int exa_fv_boundary_conditions(exa_fv_solver* s, const exa_bc* bc, double* stateOut, const double* stateIn,
                               const double centre[3], const double size[3], double t, double dt, int faceIndex,
                               int normalNonZero);

/* FiniteVolumesSolver::riemannSolver (E/solvers/FiniteVolumesSolver.h:333-336;
 * kernels/finitevolumes/riemannsolvers/c/rusanov.cpph:24-156); returns s_max. */
int exa_fv_riemann_solver(exa_fv_solver* s, double* fL, double* fR, const double* qL, const double* qR,
                          int normalNonZero, double* s_max);

/* Device-resident periodic grid of patches (nz = 1 in 2D).  One step is the
 * reference's FV step: ghost-layer filling from the neighbours
 * (E/solvers/FiniteVolumesSolver.cpp:659-687), solution update with the dt of
 * the step (:519-583), then the stable dt of the new state (:443-466). */
typedef struct exa_fv_grid exa_fv_grid;

int exa_fv_grid_create(exa_fv_solver* s, int nx, int ny, int nz, const double offset[3], const double width[3],
                       exa_fv_grid** out);
int exa_fv_grid_destroy(exa_fv_grid* g);
/* all patches with ghost layers, patches ordered (z,y,x) x fastest */
int exa_fv_grid_upload(exa_fv_grid* g, const double* luh);
int exa_fv_grid_download(exa_fv_grid* g, double* luh);
int exa_fv_grid_set_initial_conditions(exa_fv_grid* g, const exa_ic* ic);
/* Closes the axes with periodic[d] == 0: their domain faces are filled as
 * mergeWithBoundaryData does (E/solvers/FiniteVolumesSolver.cpp:689-728),
 * boundaryLayerExtraction, boundaryConditions with bc at the patch's time
 * stamp, ghostLayerFillingAtBoundary. */
int exa_fv_grid_set_boundary(exa_fv_grid* g, const int periodic[3], const exa_bc* bc);
int exa_fv_grid_compute_dt(exa_fv_grid* g, double* dt);
int exa_fv_grid_step(exa_fv_grid* g, int nsteps);
/* simulated time, dt of the next step, min over the last step's patches of
 * the kernels' admissible dt */
int exa_fv_grid_sync(exa_fv_grid* g, double* time, double* next_dt, double* admissible_dt);
/* Per-phase kernel times of the steps taken with timing enabled (HIP events
 * on the grid's stream), summed since the previous call, which resets them:
 * k = 0 ghost layers, 1 (reserved, 0), 2 update (Godunov, or MUSCL-Hancock
 * slopes + half step + Riemann in one kernel), 3 CFL time step.  No
 * counterpart in the reference (measurement). */
int exa_fv_grid_enable_timing(exa_fv_grid* g, int enable);
int exa_fv_grid_kernel_times(exa_fv_grid* g, double times[4], long long counts[4]);

/* ---------------------------------------------------------------------- */
/* Subcell limiter (orders 1..9, 2N+1 subcells per direction, 5 variables).
 * luh [z][y][x][nVar]; lim [z][y][x][nVar] of side 2N+1+2*ghost, of which the
 * interior is read or written. */

/* kernels::limiter::generic::c::projectOnFVLimiterSpace
 * (kernels/limiter/generic/c/Limiter.cpp:23-66) */
int exa_limiter_project_on_fv(int order, const double* luh, int ghost, double* lim);

/* kernels::limiter::generic::c::projectOnDGSpace (kernels/limiter/generic/c/Limiter.cpp:69-112) */
int exa_limiter_project_on_dg(int order, const double* lim, int ghost, double* luh);

/* kernels::limiter::generic::c::discreteMaximumPrincipleAndMinAndMaxSearch
 * (kernels/limiter/generic/c/Limiter.cpp:114-165) with the observables the
 * first nobs variables.  bmin/bmax: [6][nobs] face-neighbour minima/maxima,
 * updated as the reference does (block 0 <- cell min/max, block 1 <- boundary
 * min/max); *admissible = 1 if the DMP holds. */
int exa_limiter_dmp(int order, const double* luh, int nobs, double relax, double scaling, double* bmin,
                    double* bmax, int* admissible);

/* ---- Limiter on a device grid (orders 1..9, Euler observables) ----------
 * The per-step troubled-cell work of LimitingADERDGSolver on a single-level
 * periodic grid (E/solvers/LimitingADERDGSolver.cpp:838-1048), one launch
 * per stage over all cells:
 *   snapshot: every cell's min/max of the first nobs variables (GL nodes,
 *     Gauss-Lobatto nodes, subcell averages: Limiter.cpp:210-368) of the
 *     current state, merged over the cell and its six face neighbours (the
 *     boundary extrema the reference merges over faces);
 *   check (after a step): the new min/max against the snapshot's merged
 *     extrema with the relaxed DMP (Limiter.cpp:114-165), plus physical
 *     admissibility (density and, for nobs = 5, energy minima > 0);
 *     *ntroubled receives the count, flags (host, one int per cell, may be
 *     NULL) 1 for troubled cells;
 *   project: the troubled cells' DG solutions projected onto their
 *     (2N+1)^3 subcells (projectOnFVLimiterSpace, Limiter.cpp:23-66), one
 *     workgroup per troubled cell, into the device array d_lim of capacity
 *     doubles: patch i = [v][z][y][x] of side 2N+1+2*ghost (interior
 *     written) for the cell cells[i] (host, may be NULL; the order of the
 *     list is not deterministic).
 * The limiter status machine, rollback and FV recomputation are not part of
 * this layer. */
int exa_grid_limiter_snapshot(exa_grid* g, int nobs);
int exa_grid_limiter_check(exa_grid* g, double relax, double scaling, long long* ntroubled, int* flags);
int exa_grid_limiter_project(exa_grid* g, int ghost, double* d_lim, long long capacity, int* cells);

/* ---- AMR level transfers (orders 0..9, 1..10 levels) ---------------------
 * Between a coarse cell/face and a descendant fineGridLevel - coarseGridLevel
 * levels finer; the sub-index of the descendant counts 0..3^levels-1 per axis
 * (x first).  Prolongation overwrites the fine data, restriction ADDS the L2
 * projection of the fine data into the coarse data, as the reference does.
 * Face arrays [a][b][n] (b is sub-index axis 0), volume arrays [z][y][x][n]. */

/* faceUnknownsProlongation (E/solvers/ADERDGSolver.h:1229; kernel
 * kernels/aderdg/generic/c/3d/amrRoutines.cpph:45-144; called from
 * E/solvers/ADERDGSolver.cpp:2315) */
int exa_amr_face_unknowns_prolongation(int order, int nvar, int npar, double* lQhbndFine, double* lFhbndFine,
                                       const double* lQhbndCoarse, const double* lFhbndCoarse, int coarseGridLevel,
                                       int fineGridLevel, const int subfaceIndex[2]);
/* faceUnknownsRestriction (E/solvers/ADERDGSolver.h:1244; kernel
 * kernels/aderdg/generic/c/3d/amrRoutines.cpph:182-258; called from
 * E/solvers/ADERDGSolver.cpp:2438) */
int exa_amr_face_unknowns_restriction(int order, int nvar, int npar, double* lQhbndCoarse, double* lFhbndCoarse,
                                      const double* lQhbndFine, const double* lFhbndFine, int coarseGridLevel,
                                      int fineGridLevel, const int subfaceIndex[2]);
/* volumeUnknownsProlongation (E/solvers/ADERDGSolver.h:1259; kernel
 * kernels/aderdg/generic/c/3d/amrRoutines.cpph:289-374; called from
 * E/solvers/ADERDGSolver.cpp:1415) */
int exa_amr_volume_unknowns_prolongation(int order, int nvar, int npar, double* luhFine, const double* luhCoarse,
                                         int coarseGridLevel, int fineGridLevel, const int subcellIndex[3]);
/* volumeUnknownsRestriction (E/solvers/ADERDGSolver.h:1272; kernel
 * kernels/aderdg/generic/c/3d/amrRoutines.cpph:410-493; called from
 * E/solvers/ADERDGSolver.cpp:1688) */
int exa_amr_volume_unknowns_restriction(int order, int nvar, int npar, double* luhCoarse, const double* luhFine,
                                        int coarseGridLevel, int fineGridLevel, const int subcellIndex[3]);

/* One transfer of a batch: block offsets in doubles into the fine and coarse
 * device arrays, the level difference and the sub-index (index[2] unused for
 * faces). */
typedef struct {
  long long fine;
  long long coarse;
  int levels;
  int index[3];
} exa_amr_transfer;

/* Batched forms of the four transfers above over DEVICE arrays (d_fine of
 * n_fine doubles, d_coarse of n_coarse doubles), dim 2 = faces, 3 = volumes,
 * ndata values per node; the transfer list is a host array.  One workgroup per
 * output block: prolongations must write disjoint fine blocks; restrictions
 * into one coarse block must be consecutive in the list and are added in list
 * order.  Input and output arrays must not overlap.  Synchronous.  (No single
 * counterpart in the reference, which calls the per-cell kernels from its
 * mesh traversal, E/solvers/ADERDGSolver.cpp.) */
int exa_amr_prolongation_batched(int dim, int order, int ndata, double* d_fine, long long n_fine,
                                 const double* d_coarse, long long n_coarse, const exa_amr_transfer* transfers,
                                 int count);
int exa_amr_restriction_batched(int dim, int order, int ndata, double* d_coarse, long long n_coarse,
                                const double* d_fine, long long n_fine, const exa_amr_transfer* transfers,
                                int count);
/* exa_amr_restriction_batched without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, checks and results as
 * exa_amr_restriction_batched; the tables and the transfer list are staged in
 * a fixed device array set up by exa_init (call exa_init on the device
 * first), so at most EXA_AMR_NOALLOC_MAX_TRANSFERS transfers per call (more:
 * EXA_ERR_ARG).  Calls are serialised within the process; the staging area
 * keeps the last call's contents, and calls from two processes on one GPU at
 * once would race on it.  Synchronous; records no kernel time. */
#define EXA_AMR_NOALLOC_MAX_TRANSFERS 16384
int exa_amr_restriction_batched_noalloc(int dim, int order, int ndata, double* d_coarse, long long n_coarse,
                                        const double* d_fine, long long n_fine, const exa_amr_transfer* transfers,
                                        int count);
/* Duration (HIP events) of the transfer kernel of the last batched call on
 * the current device.  No counterpart in the reference (measurement). */
int exa_amr_last_kernel_ms(double* ms);

/* ---------------------------------------------------------------------------
 * Limiting ADER-DG on a static adaptive mesh: the step of
 * exahype::solvers::LimitingADERDGSolver (E/solvers/LimitingADERDGSolver.cpp:
 * 721-1490) over all leaves of a spacetree of up to three levels, the
 * limiter (MUSCL-Hancock on 2N+1 subcells per direction) on the finest
 * level, one hipGraph per step.  It replaces the engine's traversals of a
 * non-fused step (E/runners/Runner.cpp:1123-1140) with the limiter's local
 * recomputation (Runner.cpp:886-955):
 *   prediction on every leaf; face data of virtual cells by face
 *   prolongation / restriction (E/solvers/ADERDGSolver.cpp:2240-2470);
 *   merging of limiter status, DMP extrema and Riemann problems per the
 *   limiter status (LimitingADERDGSolver.cpp:1641-1850); the update switch,
 *   DMP + admissibility and the new status (:721-1048); on a newly troubled
 *   cell, status spreading (five rounds), rollback and FV recomputation
 *   (:369-379, :1283-1441); the global ADER-DG time step.
 * The mesh is built once: base_cells^3 cells over the unit cube (periodic),
 * a cell of level l < levels-1 refined by 3 where its box meets the shell
 * radius - width[l] <= |x - centre| <= radius + width[l], 2:1 balanced.  It
 * never changes: a troubled cell below the finest level and a limiter
 * status above Ok next to a coarser cell are counted (the reference would
 * refine there) and those cells keep their ADER-DG update.  Euler, orders
 * 1..7, the solver's predictor precision; FP64 update and time step. */
typedef struct exa_amr_grid exa_amr_grid;
typedef struct {
  int base_cells;            /* cells per direction of level 0                       */
  int levels;                /* 1..3                                                 */
  double centre[3];          /* refinement shell                                     */
  double radius;
  double width[2];           /* half width of the shell refining level 0 and level 1 */
  double dmp_relaxation;     /* the spec's dmp-relaxation-parameter (Parser.cpp:780) */
  double dmp_scaling;        /* dmp-difference-scaling (Parser.cpp:797)              */
  int steps_till_cured;      /* steps-till-cured (Parser.cpp:831)                    */
  long long patch_capacity;  /* limiter patches kept on the device; 0: a quarter of
                                the finest-level cells (at least 64)                 */
} exa_amr_config;
/* counters of exa_amr_grid_counters, summed over the steps */
#define EXA_AMR_CNT_IRREGULAR_STEPS 0  /* steps with a newly troubled cell (recomputed)        */
#define EXA_AMR_CNT_REFINE_REQUESTS 1  /* troubled cells below the finest level                */
#define EXA_AMR_CNT_COARSE_FINE 2      /* finest cells in the limiter domain next to a coarser cell */
#define EXA_AMR_CNT_MIRRORED 3         /* FV ghost faces filled from the patch itself          */
#define EXA_AMR_CNT_CANNOT_COMMUNICATE 4 /* face pairs on which the reference aborts          */
#define EXA_AMR_CNT_PATCH_OVERFLOW 5   /* cells that found no free limiter patch (an error)    */
#define EXA_AMR_CNT_TROUBLED 6         /* troubled finest cells                                */
#define EXA_AMR_CNT_FV_UPDATES 7       /* FV updates of finest cells                           */
/* The leaves of the mesh a configuration builds, [n][4] = level, ix, iy, iz
 * in storage order (host only: no device needed); cells may be NULL. */
int exa_amr_mesh_cells(const exa_amr_config* cfg, int* cells, long long capacity, long long* n);
int exa_amr_grid_create(exa_solver* s, const exa_amr_config* cfg, exa_amr_grid** out);
int exa_amr_grid_destroy(exa_amr_grid* g);
int exa_amr_grid_num_cells(const exa_amr_grid* g, long long* n);
/* leaves in storage order: [n][4] = level, ix, iy, iz (index at that level) */
int exa_amr_grid_cells(const exa_amr_grid* g, int* cells);
/* solution of every leaf, [leaf][z][y][x][5]; setting it (or the initial
 * conditions) starts the limiter as setInitialConditions does
 * (LimitingADERDGSolver.cpp:692-715, :906-930): admissibility only, then
 * spreading; the limiter patches then hold the FV initial conditions at the
 * subcell centres (exa_amr_grid_set_initial_conditions) or the projection
 * of the uploaded solution (exa_amr_grid_upload) */
int exa_amr_grid_upload(exa_amr_grid* g, const double* luh);
int exa_amr_grid_download(exa_amr_grid* g, double* luh);
int exa_amr_grid_set_initial_conditions(exa_amr_grid* g, const exa_ic* ic);
int exa_amr_grid_compute_dt(exa_amr_grid* g, double* dt);
int exa_amr_grid_use_graph(exa_amr_grid* g, int enable);
int exa_amr_grid_enable_timing(exa_amr_grid* g, int enable);
int exa_amr_grid_step(exa_amr_grid* g, int nsteps);
int exa_amr_grid_sync(exa_amr_grid* g, double* time, double* next_dt);
int exa_amr_grid_limiter_status(exa_amr_grid* g, int* status);
/* the interior subcells of the limiter patches, [k][z][y][x][5] (2N+1 per
 * direction) of the k-th cell listed in cells (up to capacity patches) */
int exa_amr_grid_patches(exa_amr_grid* g, double* patches, int* cells, long long capacity, long long* n);
int exa_amr_grid_counters(exa_amr_grid* g, long long counts[8]);
/* per step (timing on, graph off): predictor, face transfers + merges,
 * Riemann, ADER-DG update, limiter (FV, status, recomputation), time step (ms) */
int exa_amr_grid_phase_times(exa_amr_grid* g, double ms[6], long long* steps);
/* Picard sweeps of all predictor launches so far, summed over the leaves
 * (measurement; no counterpart in the reference) */
int exa_amr_grid_total_picard_iterations(exa_amr_grid* g, long long* total);

#ifdef __cplusplus
}
#endif
#endif
