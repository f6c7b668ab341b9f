/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the reference's generic CPU kernels for the 3D
 * nonlinear ADER-DG scheme and the finite-volume (Godunov / MUSCL-Hancock)
 * schemes, used to check the HIP product path.  Only the tests, the smoke
 * check and the CPU-baseline leg of bench.py may load it; the product never
 * calls it.  Every function keeps the reference's array layout (row-major,
 * variable index fastest) and its loop and summation order, so that on the
 * same inputs it reproduces the reference's arithmetic.  The file:line each
 * function follows is cited beside it (paths relative to the reference repo).
 *
 * Pins (what checks each function): see tests/test_oracle_*.py.
 *   flux/eigenvalues      golden testPDEFluxes (generic_euler_testdata.cpp)
 *   volume integral       golden testVolumeIntegralNonlinear (N=3)
 *   surface integral      golden testSurfaceIntegralNonlinear (N=3)
 *   riemann solver        golden testRiemannSolverNonlinear (N=3, random states)
 *   solution update       golden testSolutionUpdate (N=3)
 *   space-time predictor  golden testSpaceTimePredictorNonlinear (N=3, constant
 *                         state) + exactness for linear advection (closed form)
 *   stable dt             closed form
 *   whole scheme          convergence order on the isentropic vortex
 *   FV kernels            closed forms, conservation, symmetry
 *   NCP arm               flux form == non-conservative form (advection)
 *   linear family         goldens testVolumeIntegralLinear, testSurfaceIntegralLinear,
 *                         testRiemannSolverLinear; predictor exact for polynomial data
 *   point sources         Lagrange basis by interpolation uniqueness; projection by
 *                         node sources and quadrature of the force; predictor arm by
 *                         closed forms at zero velocity
 */
#ifndef EXA_ORACLE_H
#define EXA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int order;      /* N; B = N+1 nodes per axis                       */
  int nvar;       /* conserved variables (5 for 3D Euler)            */
  int npar;       /* material parameters carried in luh (0 for Euler) */
  double gamma;   /* ratio of specific heats                         */
  double cfl;     /* CFL factor (0.9 for ADER-DG)                    */
  int pde;        /* 0: 3D Euler; 1: linear advection test PDE; 2: EulerFV demonstrator (FV only);
                     3: 3D Euler with constant gravity (algebraic source, the useSource arm);
                     4: linear advection in non-conservative form (no flux, B_d = a_d I: the
                        useFlux = false, useNCP = true arms);
                     5: 3D Euler with the test non-conservative product of the reference's
                        Euler kernel test (the useNCP arm on a nonlinear flux) */
  double adv[3];  /* advection velocity for pdes 1 and 4             */
  double grav;    /* gravity along -z for pde==3                      */
  int linear;     /* 1: the linear kernel family (pdes 4, 5): Cauchy-Kowalevski
                     predictor, fluctuation-form integrals, coefficient-matrix Riemann */
  int point_source;      /* 1 (linear only): the usePointSource arm, with the user
                            pointSource of orc_user_point_source */
  double src_x0[3];      /* position of the point source                     */
  double src_amp[5];     /* force amplitude per variable                     */
  double src_f0, src_t0; /* Ricker wavelet: peak frequency and delay         */
} orc_cfg;

/* user PDE (E/tests/kernels/c/GenericEulerKernelTest3D.cpp:26-82) */
void orc_flux(const orc_cfg* c, const double* Q, double* F /* [3][nvar] */);
void orc_eigenvalues(const orc_cfg* c, const double* Q, int d, double* lambda);
/* algebraic source S(Q) (the reference's fusedSource without NCP); zero
 * except for pde==3: S = (0, 0, 0, -rho g, -m_z g) */
void orc_source(const orc_cfg* c, const double* Q, double* S);
int orc_has_source(const orc_cfg* c);
/* the PDE's feature arms: flux (all but pde 4), non-conservative product (pdes 4, 5) */
int orc_has_flux(const orc_cfg* c);
int orc_has_ncp(const orc_cfg* c);
/* nonConservativeProduct (UserSolverInterface.h), gradQ [3][nvar] */
void orc_ncp(const orc_cfg* c, const double* Q, const double* gradQ, double* BgradQ);
/* kernels::fusedSource (kernels/fusedMethods.cpph:17-32): source - ncp */
void orc_fused_source(const orc_cfg* c, const double* Q, const double* gradQ, double* S);

/* ADER-DG kernels, kernels/aderdg/generic/c/3d/ */
int orc_space_time_predictor(const orc_cfg* c, double* lQhbnd, double* lFhbnd,
                             double* lQi, double* lFi, double* lQhi, double* lFhi,
                             const double* luh, const double dx[3], double dt);
void orc_volume_integral(const orc_cfg* c, double* lduh, const double* lFhi, const double dx[3]);
void orc_surface_integral(const orc_cfg* c, double* lduh, const double* lFhbnd, const double dx[3]);
void orc_riemann_solver(const orc_cfg* c, double* FL, double* FR, const double* QL,
                        const double* QR, double dt, int direction);
void orc_solution_update(const orc_cfg* c, double* luh, const double* lduh, double dt);
double orc_stable_time_step_size(const orc_cfg* c, const double* luh, const double dx[3]);

/* Cartesian-vertex sampling of one cell (ADERDG2CartesianVTK::plotVertexData):
 * out [(N+1)^3 vertices, x fastest][nvar] */
void orc_sample_cartesian_vertices(const orc_cfg* c, const double* luh, double* out);

/* the linear kernels of kernels/aderdg/generic/c/3d (oracle/linear.c) */
void orc_ncp_linear(const orc_cfg* c, const double* Q, const double* gradQ, double* T /* [3][nvar] */);
void orc_coefficient_matrix(const orc_cfg* c, const double* Q, int d, double* Bn /* [nvar][nvar] */);
void orc_space_time_predictor_linear(const orc_cfg* c, double* lQbnd, double* lFbnd, double* lQi, double* lFi,
                                     double* lQh, double* lFh, const double* luh, const double dx[3], double dt,
                                     const double* PSi /* NULL: no point source */);
/* point sources (kernels/aderdg/generic/c/3d/pointSource.cpph) */
double orc_lagrange(int order, int i, double xi);
void orc_user_point_source(const orc_cfg* c, const double x[3], double t, double dt, double* force, double x0[3]);
void orc_point_source(const orc_cfg* c, double t, double dt, const double centre[3], const double dx[3],
                      double* PSi /* [B+1][z][y][x][nVar] */);
void orc_volume_integral_linear(const orc_cfg* c, double* lduh, const double* lFhi, const double dx[3]);
void orc_surface_integral_linear(const orc_cfg* c, double* lduh, const double* lFbnd, const double dx[3]);
void orc_riemann_solver_linear(const orc_cfg* c, double* FL, double* FR, const double* QL, const double* QR,
                               double dt, int direction);

/* initial conditions at GL nodes (kernels/aderdg/generic/c/3d/solutionAdjustment.cpph:17-52) */
typedef struct {
  int kind;            /* 0 vortex+modes, 1 Sod sphere, 2 constant, 3 advected sine, 4 entropy wave,
                          5 Gaussian density pulse (EulerFV demonstrator setup) */
  double xc[3];        /* vortex centre / sphere centre                  */
  double L, eps;       /* vortex length scale and strength               */
  double uinf[3];      /* background velocity                            */
  double rho_inf, p_inf;
  double amp;          /* density-mode amplitude                         */
  int nmodes;
  double k[8][3], a[8], phi[8];
  double r0, rho_in, p_in, rho_out, p_out;  /* Sod sphere */
  double q0[8];        /* constant state (kind 2)                        */
} orc_ic;
void orc_ic_point(const orc_cfg* c, const orc_ic* ic, const double x[3], double t, double* Q);
void orc_solution_adjustment(const orc_cfg* c, const orc_ic* ic, double* luh,
                             const double centre[3], const double dx[3], double t);

/* The compiled user boundaryValues of the Euler solvers (the body of the
 * generated GenericADERDGSolverInCUserCode.template:44-52): kind 1 outflow
 * (copy, as the EulerFV demonstrator, Demonstrators/EulerFV/MyEulerSolver.cpp:87-98),
 * 2 reflecting wall (normal momentum negated, flux of that state), 3 the
 * exact solution of ic averaged over the step at the GL time nodes. */
typedef struct {
  int kind;
  orc_ic ic;
} orc_bc;
void orc_boundary_values(const orc_cfg* c, const orc_bc* bc, const double x[3], double t, double dt, int faceIndex,
                         int d, const double* Fin, const double* Qin, double* Fout, double* Qout);
/* boundaryConditions, kernels/aderdg/generic/c/3d/boundaryConditions.cpph:24-85 */
void orc_boundary_conditions(const orc_cfg* c, const orc_bc* bc, double* FOut, double* QOut, const double* FIn,
                             const double* QIn, const double centre[3], const double size[3], double t, double dt,
                             int faceIndex, int normalNonZero);

/* regular grid driver (non-fused stepping, E/runners/Runner.cpp:1125-1190);
 * periodic along an axis unless periodic[d] == 0, where the domain faces take
 * boundaryConditions + Riemann (applyBoundaryConditions,
 * E/solvers/ADERDGSolver.cpp:2854-2931) */
typedef struct {
  int nx, ny, nz;
  double offset[3], width[3];
  double* luh;      /* [cell][B^3][nData]   */
  double* lduh;     /* [cell][B^3][nvar]    */
  double* lQhbnd;   /* [cell][6][B^2][nData]*/
  double* lFhbnd;   /* [cell][6][B^2][nvar] */
  int* iters;       /* Picard iterations of the last predictor per cell */
  double t;         /* time at the start of the next step (advanced by orc_grid_step) */
  int periodic[3];
  orc_bc bc;
} orc_grid;
int orc_grid_alloc(const orc_cfg* c, orc_grid* g, int nx, int ny, int nz);
void orc_grid_free(orc_grid* g);
void orc_grid_init(const orc_cfg* c, orc_grid* g, const orc_ic* ic);
double orc_grid_stable_dt(const orc_cfg* c, const orc_grid* g, int nthreads);
/* one step with dt; returns the stable dt of the new state */
double orc_grid_step(const orc_cfg* c, orc_grid* g, double dt, int nthreads);
/* predictor+volume integral on a subset of cells (CPU baseline sample) */
void orc_grid_predict_cells(const orc_cfg* c, orc_grid* g, double dt, int first, int count, int nthreads);

/* FV kernels, kernels/finitevolumes/ */
double orc_fv_rusanov(const orc_cfg* c, int dim, double* fL, double* fR,
                      const double* qL, const double* qR, int normal);
double orc_fv_stable_dt(const orc_cfg* c, int dim, int patch, int ghost,
                        const double* luh, const double* dx, double cfl);
double orc_godunov_update(const orc_cfg* c, int dim, int patch, double* luh_new,
                          const double* luh, const double* dx, double dt, double cfl);
double orc_musclhancock_update(const orc_cfg* c, int patch, double* luh_new,
                               const double* luh, const double* dx, double dt, double cfl);

/* periodic grid of FV patches with face ghost layers (width 1 Godunov, 2 MUSCL-Hancock) */
typedef struct {
  int dim, scheme, patch, ghost;  /* scheme 0 Godunov, 1 MUSCL-Hancock */
  int nx, ny, nz;
  double offset[3], width[3];
  double* luh;      /* [patch][(z)][y][x][nData] including ghosts */
  double* luh_new;
  int periodic[3];  /* 0: the domain faces of that axis take boundary values */
  orc_bc bc;
  double t;         /* time stamp of the patches (advanced by each step)    */
} orc_fv_grid;
void orc_fv_patch_ic(const orc_cfg* c, const orc_ic* ic, int P, int G, const double centre[3], const double dx[3],
                     double* luh);
/* The FV boundary hooks of one patch (P^dim subcells, ghost width G; layouts
 * of kernels/finitevolumes/commons/c/commons.cpph, 2D :99-335, 3D :503-982):
 * pos / faceIndex as FiniteVolumesSolver::mergeWithBoundaryData passes them
 * (E/solvers/FiniteVolumesSolver.cpp:689-728). luhbnd blocks: x faces
 * [(z)][y][G], y faces [(z)][G][x], z faces [G][y][x], each of nData. */
void orc_fv_boundary_layer_extraction(const orc_cfg* c, int dim, int P, int G, double* luhbnd, const double* luh,
                                      const int pos[3]);
void orc_fv_ghost_layer_filling_at_boundary(const orc_cfg* c, int dim, int P, int G, double* luh, const double* luhbnd,
                                            const int pos[3]);
void orc_fv_ghost_layer_filling(const orc_cfg* c, int dim, int P, int G, double* luh, const double* luhNb,
                                const int pos[3]);
/* boundaryValues of the Euler FV solvers at an outside subcell centre: kind 1
 * copy, 2 normal momentum negated, 3 the state of bc->ic at (x, t). */
void orc_fv_boundary_values(const orc_cfg* c, const orc_bc* bc, const double x[3], double t, double dt,
                            int faceIndex, int d, const double* Qin, double* Qout);
void orc_fv_boundary_conditions(const orc_cfg* c, const orc_bc* bc, int dim, int P, int G, double* stateOut,
                                const double* stateIn, const double centre[3], const double size[3], double t,
                                double dt, int faceIndex, int normalNonZero);
int orc_fv_grid_alloc(const orc_cfg* c, orc_fv_grid* g, int dim, int scheme, int patch, int nx, int ny, int nz);
void orc_fv_grid_free(orc_fv_grid* g);
void orc_fv_grid_init(const orc_cfg* c, orc_fv_grid* g, const orc_ic* ic);
double orc_fv_grid_stable_dt(const orc_cfg* c, const orc_fv_grid* g, int nthreads);
double orc_fv_grid_step(const orc_cfg* c, orc_fv_grid* g, double dt, double* dt_admissible, int nthreads);

/* subcell limiter, kernels/limiter/generic/c/Limiter.cpp and kernels/LimiterProjectionMatrices.cpp */
int orc_limiter_matrices(int N, double* uh2lim, double* uh2lob, double* lim2uh);
void orc_project_on_fv(int N, int nvar, int ghost, const double* luh, double* lim);
void orc_project_on_dg(int N, int nvar, int ghost, const double* lim, double* luh);
void orc_cell_min_max(int N, int nvar, int nobs, const double* luh, double* mn, double* mx);
int orc_dmp(int N, int nvar, int nobs, const double* luh, double relax, double scaling, double* bmin, double* bmax);

/* AMR level transfers, kernels/aderdg/generic/c/3d/amrRoutines.cpph */
void orc_face_unknowns_prolongation(int N, int nvar, int npar, double* lQhbndFine, double* lFhbndFine,
                                    const double* lQhbndCoarse, const double* lFhbndCoarse, int coarseLevel,
                                    int fineLevel, const int subfaceIndex[2]);
void orc_face_unknowns_restriction(int N, int nvar, int npar, double* lQhbndCoarse, double* lFhbndCoarse,
                                   const double* lQhbndFine, const double* lFhbndFine, int coarseLevel,
                                   int fineLevel, const int subfaceIndex[2]);
void orc_volume_unknowns_prolongation(int N, int ndata, double* luhFine, const double* luhCoarse, int coarseLevel,
                                      int fineLevel, const int subcellIndex[3]);
void orc_volume_unknowns_restriction(int N, int ndata, double* luhCoarse, const double* luhFine, int coarseLevel,
                                     int fineLevel, const int subcellIndex[3]);

#ifdef __cplusplus
}
#endif
#endif
