/*
 * oracle.h -- FP64 CPU restatement of PhySim's quasi-static Newton inner loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path (libpgn.so) never links it.
 *
 * Every function cites the reference file:line it restates.  Parity pins:
 *   - material Psi / P / dP/dF  : golden vectors from the reference's own Maple
 *                                 kernels (oracle/_ref) + closed-form invariants
 *   - Tet4 element E/g/H        : finite differences (reference eps/tol), brute
 *                                 force identities (sum of g = 0, H*translation=0)
 *   - mesh generator            : node/tet counts and volumes of the grid beam
 *   - assembly / fixed DoFs     : dense brute-force sum on tiny meshes
 *   - PCG                       : scipy/numpy dense solves on small SPD systems
 *   - Newton driver             : converged equilibrium (unique) and the
 *                                 Euler-Bernoulli small-load limit.
 *     Iteration counts / trajectories of the reference's Eigen build are
 *     "parity unpinned" (Eigen is unpinned and absent).
 */
#ifndef PHYSIM_ORACLE_H
#define PHYSIM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_CONH = 0, ORC_STVK = 1 };

/* ---- geometry ---------------------------------------------------------- */
/* MeshUtils.cpp:927-991 CreateTetMesh_Grid.  V: 3*numNodes (node-major xyz),
 * T: 4*numTets.  Returns number of tets. */
int orc_tet_grid(const double scale[3], const int dims[3], double* V, int* T);
int orc_tet_grid_counts(const int dims[3], int* num_nodes, int* num_tets);

/* ParameterSet.cpp:54-70 */
void orc_lame_from_young_poisson(double young, double poisson, double* lame1,
                                 double* lame2);

/* ---- material models (column-major vec(F), MathUtils.h:97-111) --------- */
/* MaterialModels.h:166-230 (StVK), :235-299 (CoNH) */
double orc_psi(int kind, const double* Fv, double lame1, double lame2);
void orc_pk1(int kind, const double* Fv, double lame1, double lame2,
             double* P);
void orc_dpdf(int kind, const double* Fv, double lame1, double lame2,
              double* C);

/* ---- Tet4 element, 1-point quadrature ---------------------------------- */
/* ShapeFunctions.cpp:89-133 + EnergyElement_FEM.cpp:59-122.  X: 12 rest
 * coords (node-major).  DFDx: 9x12 row-major.  *V0 = (1/6) det(J). */
void orc_tet_rest(const double X[12], double DFDx[108], double* V0);
/* EnergyElement_FEM.cpp:125-222.  Returns 1 if the element is inverted (then
 * E = HUGE_VAL and g/H are still evaluated like the reference does). */
int orc_tet_egh(const double x[12], const double DFDx[108], double V0,
                int kind, double lame1, double lame2, double* E, double g[12],
                double H[144]);

/* ---- model (Simulable_FEM + MassElement_Lumped + BCs) ------------------ */
typedef struct orc_model orc_model;

orc_model* orc_model_create(int n_nodes, const double* X0, int n_tets,
                            const int* T, const unsigned char* kind,
                            const double* lame1, const double* lame2,
                            const double* rho);
/* Mixed element kinds over one node set: kind 0 CoNH tet, 1 StVK tet,
 * 2 StVK membrane triangle (T row = v0 v1 v2 v2, extra = thickness, rho per
 * unit area), 3 discrete-shell hinge (T row = edge v0 v1, opposite v2 v3,
 * extra = bending stiffness kB; lame / rho unused).  extra may be NULL when
 * every element is a tet. */
orc_model* orc_model_create_mixed(int n_nodes, const double* X0, int n_elems,
                                  const int* T, const unsigned char* kind,
                                  const double* lame1, const double* lame2,
                                  const double* rho, const double* extra);
void orc_model_destroy(orc_model* m);
int orc_model_num_dofs(const orc_model* m);
void orc_model_set_fixed(orc_model* m, const unsigned char* mask);
/* BC_Gravity + VectorAnimator_LerpValues; stages >= 1.  The current value is
 * lerp(g_ini, g_end, cur/stages); cur starts at 0. */
void orc_model_set_gravity(orc_model* m, const double g_ini[3],
                           const double g_end[3], int stages);
/* BC_Force (EnergyElement_Force: E = +x.f, g = +f), lerp over stages. */
void orc_model_set_forces(orc_model* m, const double* f_ini,
                          const double* f_end, int stages);
int orc_model_step_load(orc_model* m);
int orc_model_fully_loaded(const orc_model* m);
void orc_model_set_x(orc_model* m, const double* x);
void orc_model_get_x(const orc_model* m, double* x);
/* lumped node masses (3n, replicated per component) */
void orc_model_get_mass(const orc_model* m, double* mass3n);

/* Simulable::AddEnergy (elements in order, then mass, then BC energies) */
double orc_model_energy(orc_model* m);
/* |elastic energy| + |load energy| of the last orc_model_energy call */
double orc_model_energy_scale(const orc_model* m);
/* Simulable::GetGradient + FixVectorDoFs */
void orc_model_gradient(orc_model* m, double* g);
/* lower-triangular CSR (row >= col) pattern size */
long orc_model_nnz_lower(orc_model* m);
/* Simulable::GetHessian + FixMatrixDoFs, lower CSR. rowptr: 3n+1. */
void orc_model_hessian_lower(orc_model* m, long* rowptr, int* col,
                             double* val);
/* per-element local outputs (IEnergyElement::GetLocal*) */
void orc_model_element_egh(orc_model* m, double* E, double* g, double* H);

/* ---- linear solve: Eigen ConjugateGradient<Lower, Diagonal> ------------ */
/* A given as lower CSR of a symmetric matrix.  Returns iterations, writes the
 * relative residual to *err. */
int orc_cg(int n, const long* rowptr, const int* col, const double* val,
           const double* b, double* x, double tol, int maxit, double* err);
/* LinearSolver_EigenCG::SolveInternal (+ LinearSolver::GetMatrixData):
 * b.x < 0 => val += regT*10^(i-1) on the diagonal (cumulative), <= regIters.
 * Returns 0 on SUCCESS, 1 on FAILURE.  *tot_iters accumulates CG iterations.
 * val is modified in place like the reference's mA. */
int orc_linear_solve(int n, const long* rowptr, const int* col, double* val,
                     const double* b, double* x, double tol, int maxit,
                     int reg_iters, int* tot_iters, int* reg_used);

/* ---- Newton driver ----------------------------------------------------- */
/* OptimSolverOptions (OptimSolver.h:23-71) */
typedef struct {
  double tolMaxError;
  double tolMinStep;
  double tolMinImpr;
  int minStepMaxCount;
  int minImprMaxCount;
  int numMaxIters;
  double maxStepSize;
  int trySteepest;
  int lsearch_numTry;
  double lSearch_alpha;
  int lsNumMaxIters;
  double lsTolMaxError;
  int lsNumRegTrials;
  /* Not in the reference (0 = reference semantics): also accept a line-search
   * step whose energy rises by at most ls_noise_guard * Ebar (Ebar = |elastic
   * energy| + |load energy|, the scale of the FP64 rounding of the total) when
   * the gradient norm decreases. */
  double ls_noise_guard;
  /* OptimSolverOptions::lsSolverType (OptimSolver.h:40,58) with the values of
   * LSSolverType (BasicTypes.h:36-46): 1 = EigenCG (Jacobi CG, the default
   * here), 2 = EigenLDLT (the reference's own default) */
  int ls_solver;
  /* OptimSolverOptions::qpSolverType (OptimSolver.h:30) with the values of
   * QPSolverType (BasicTypes.h:48-55): 0 = Steepest, 1 = Newton (default),
   * 2 = LBFGS (ComputeStep_LBFGS, OptimSolver.cpp:522-591, corrected: see
   * orc_lbfgs_direction) */
  int qpSolverType;
} orc_newton_opts;

/* The LBFGS direction dx = -H g by the two-loop recursion (Nocedal & Wright,
 * Numerical Optimization, 2nd ed., Alg. 7.4) over the `count` newest of the
 * pairs (s_i, y_i) held in a ring of M columns (column-major N x M arrays S,
 * Y), the newest in column `newest`; H0 = (s.y / y.y) I of the newest pair,
 * as ComputeStep_LBFGS scales it (OptimSolver.cpp:568-580).  The reference's
 * own recursion differs in three places, corrected here: its previous x and
 * g are taken once (OptimSolver.cpp:524-525) instead of at every step, its
 * second loop advances the ring index modulo the problem size (:587)
 * instead of M, and its pair count min(M, K + 1) uses the ring slot K, so
 * that it falls back to one pair whenever the ring wraps (:567). */
void orc_lbfgs_direction(int N, int M, int count, int newest, const double* S,
                         const double* Y, const double* g, double* dx);

/* OSResult (BasicTypes.h:69-77) */
enum {
  ORC_OS_ONGOING = 0,
  ORC_OS_SUCCESS,
  ORC_OS_MINSTEP,
  ORC_OS_MINIMPR,
  ORC_OS_FAILURE,
  ORC_OS_MAXITER,
  ORC_OS_NONDESC
};

typedef struct {
  int result;
  int iterations;
  int cg_iterations;
  int line_search_bisections;
  double objective;
  double optimality;
} orc_newton_result;

void orc_newton_default_opts(orc_newton_opts* o);

/* A problem the generic Newton driver runs on (IOptimProblem +
 * ISimulable::UpdateKinematics + the linear solver). */
typedef struct {
  void* p;
  int N;
  double (*energy)(void*);
  double (*energy_scale)(void*);
  void (*gradient)(void*, double*);
  /* Hessian at the current state; solve H dx = b with warm start dx;
   * returns linear iterations */
  int (*solve)(void*, const double* b, double* dx, const orc_newton_opts*);
  int (*step_load)(void*);
  int (*fully_loaded)(void*);
  void (*get_x)(void*, double*);
  void (*set_x)(void*, const double*);
  void (*iter_finish)(void*);
  /* optional (may be NULL): the solver's previous x kept across SolveFull
   * calls of one solver object (time steps); used as the initial v0 and
   * written back at the end */
  double* v0_keep;
} orc_problem;
void orc_newton_generic(orc_problem* P, const orc_newton_opts* o,
                        orc_newton_result* r);
/* OptimSolver::Init + SolveFull (OptimSolver.cpp:36-55,163-264) with
 * OptimSolver_USQP_LS::SolveStepInternal and OptimProblem_BasicStatic. */
void orc_newton_solve(orc_model* m, const orc_newton_opts* o,
                      orc_newton_result* r);

/* Implicit dynamics: OptimProblem_BasicDynamic (OptimProblem_BasicDynamic.cpp:
 * 20-170) over an orc_model, one backward-Euler step per orc_dynamic_step. */
typedef struct orc_dynamic orc_dynamic;
orc_dynamic* orc_dynamic_create(orc_model* m);
void orc_dynamic_destroy(orc_dynamic* d);
void orc_dynamic_set_velocity(orc_dynamic* d, const double* v);
void orc_dynamic_get_velocity(const orc_dynamic* d, double* v);
/* OnSolveStart (:133-146): x0 = x, v0 = v, loads stepped, x = x0 + dt v0 */
void orc_dynamic_begin(orc_dynamic* d, double dt);
/* the step's objective at the current x: GetEnergy (:68-86), GetGradient
 * (:88-107, fixed DoFs zero), GetHessian (:109-121, lower CSR of the model's
 * pattern, fixed rows / columns identity) */
double orc_dynamic_energy(orc_dynamic* d);
void orc_dynamic_gradient(orc_dynamic* d, double* g);
void orc_dynamic_hessian_lower(orc_dynamic* d, long* rowptr, int* col, double* val);
/* SolveFull on the step's objective (IsFullyConstrained :127-129) */
void orc_dynamic_solve(orc_dynamic* d, const orc_newton_opts* o, orc_newton_result* r);
/* PosPerformStep (:159-169): v = (x - x0) / dt */
void orc_dynamic_end(orc_dynamic* d);
/* begin + solve + end */
void orc_dynamic_step(orc_dynamic* d, double dt, const orc_newton_opts* o,
                      orc_newton_result* r);

/* ---- thin shells (shells.c) -------------------------------------------- */
/* hinge over x = [x0 x1 x2 x3] (edge x0-x1, opposite x2 and x3) */
double orc_hinge_theta(const double x[12]);
double orc_hinge_shape(const double x[12]);
void orc_hinge_dtheta(const double x[12], double d[12]);
double orc_hinge_egh(const double x[12], double kB, double theta0, double shape0,
                     double* g12, double* H144);
/* membrane triangle: rest Dm^-1 (2x2 row-major) and area; E/g/H over 9 DoFs */
void orc_membrane_rest(const double X[9], double Dminv[4], double* area);
double orc_membrane_egh(const double x[9], const double Dminv[4], double A0t,
                        double lame1, double lame2, double* g9, double* H81);
/* 2D StVK over a row-major 3x2 F (FEMMem_StVK_*.mcg layout) */
double orc_mem_psi(const double* Fv, double lame1, double lame2);
void orc_mem_pk1(const double* Fv, double lame1, double lame2, double* P);
void orc_mem_dpdf(const double* Fv, double lame1, double lame2, double* C);

/* ---- discrete elastic rods (rods.c) ------------------------------------ */
/* bend-twist element: prm = shear[2], young[2], radw[2], radh[2], L0, L1,
 * K0[2], twist0, nRef1[3], tRef1[3], nRef2[3], tRef2[3], rt (27 values, the
 * getBT*OPT argument order); g8 / H64 over (e1, e2, theta1, theta2) */
double orc_rod_bt(const double* prm, const double* e1, const double* e2,
                  double th1, double th2, double* g8, double* H64);
double orc_rod_stretch(const double* x12, double L0, double k, double* g6,
                       double* H36);
double orc_reference_twist_change(const double* tan1, const double* nor1,
                                  const double* tan2, const double* nor2,
                                  double rt);
typedef struct orc_rod orc_rod;
/* one open rod of n vertices (edges i = (i, i+1)); frame0_nor: normal of
 * the first edge's reference frame (projected onto the plane normal to it);
 * uniform material and cross-section; lumped node mass line_density*L/2.
 * DoFs: 3n positions, then n-1 edge angles. */
orc_rod* orc_rod_create(int n, const double* X0, const double* frame0_nor,
                        double young, double shear, double rh, double rw,
                        double line_density);
void orc_rod_destroy(orc_rod* R);
int orc_rod_num_dofs(const orc_rod* R);
void orc_rod_set_fixed(orc_rod* R, const unsigned char* mask);
void orc_rod_set_gravity(orc_rod* R, const double* g_ini, const double* g_end,
                         int stages);
int orc_rod_step_load(orc_rod* R);
int orc_rod_fully_loaded(const orc_rod* R);
void orc_rod_get_x(const orc_rod* R, double* x);
void orc_rod_set_x(orc_rod* R, const double* x);
void orc_rod_get_frames(const orc_rod* R, double* tan, double* nor, double* rt);
double orc_rod_energy(orc_rod* R);
double orc_rod_energy_scale(const orc_rod* R);
void orc_rod_gradient(orc_rod* R, double* g);
void orc_rod_hessian(orc_rod* R, double* H); /* dense N x N */
void orc_rod_update_kinematics(orc_rod* R);
int orc_dense_ldlt_solve(int N, double* A, const double* b, double* x,
                         double maxError, int regIters, int* reg_used);
int orc_ldlt_solve_norm(int N, double* A, double normA, const double* b, double* x,
                        double maxError, int regIters, int* reg_used);
/* LinearSolver_EigenLDLT on the lower CSR Hessian of orc_model_hessian_lower
 * (normA over the stored entries, AMatrixSd's set; solved densely, small
 * models only).  Returns 0 on SUCCESS. */
int orc_sparse_ldlt_solve(int n, const long* rowptr, const int* col, const double* val,
                          const double* b, double* x, double maxError, int regIters,
                          int* reg_used);
void orc_rod_newton(orc_rod* R, const orc_newton_opts* o, orc_newton_result* r);

#ifdef __cplusplus
}
#endif
#endif
