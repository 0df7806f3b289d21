/*
 * pgn.h -- C ABI of the MI355X-native PhySim Newton inner loop (libpgn.so).
 *
 * The library replaces, for the quasi-static FEM path, the C++ virtual
 * interfaces PhySim's optimizer calls:
 *   ISimulable       include/PhySim/Physics/Simulables/Simulable.h:29-102
 *   IEnergyElement   include/PhySim/Physics/Elements/IEnergyElement.h:20-67
 *   IOptimProblem    include/PhySim/Solvers/OptimProblem.h:61-97
 *   ILinearSolver    include/PhySim/Solvers/LinearSolver.h:38-52
 *   OptimSolver      include/PhySim/Solvers/OptimSolver.h:23-71 (options),
 *                    source/PhySim/Solvers/OptimSolver.cpp:163 (SolveFull)
 * Each entry point below names the reference method it stands in for.
 *
 * Conventions (kept from the reference):
 *   - DoF vectors are FP64, node-major xyz (KinematicsTree.cpp:33-40);
 *   - gradients have fixed entries zeroed (Simulable.cpp:735-744);
 *   - Hessians have fixed rows/cols replaced by identity (Simulable.cpp:746-760);
 *   - an inverted element has energy +HUGE_VAL (EnergyElement_FEM.cpp:126-131).
 *
 * Errors: every int-returning call returns PGN_OK (0) or a negative code; the
 * message of the last failure on the calling thread is pgn_last_error().
 * Ownership: the library owns all device memory; host buffers passed in are
 * copied before the call returns.  Threading: one context = one HIP stream;
 * calls on one context (and its models) must be externally serialized.
 */
#ifndef PGN_H
#define PGN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGN_OK 0
#define PGN_ERR_INVALID (-1)  /* bad argument / size / state */
#define PGN_ERR_HIP (-2)      /* HIP runtime failure (message has details) */
#define PGN_ERR_NOMEM (-3)    /* device or host allocation failed */
#define PGN_ERR_NODEVICE (-4) /* no HIP device / extension unusable */

/* linear solvers, values of LSSolverType (BasicTypes.h:36-46) */
enum pgn_ls_solver {
  PGN_LS_EIGENCG = 1,   /* LinearSolver_EigenCG: (block-)Jacobi PCG (default here) */
  PGN_LS_EIGENLDLT = 2  /* LinearSolver_EigenLDLT: FP64 LDL^T, the reference's
                           default (OptimSolver.h:58); libpgn factors in the
                           narrowest of three node orders (as given, reverse
                           Cuthill-McKee, sorted along the longest axis) and
                           holds a half-bandwidth of at most 139 DoFs (one
                           workgroup's LDS): small models */
};

/* element material (MaterialModels.h:166 StVK, :235 CoNH) */
enum pgn_material { PGN_CONH = 0, PGN_STVK = 1 };

/* arithmetic of a model's Newton path */
enum pgn_precision {
  PGN_FP64 = 0,  /* plain path: every kernel and the PCG in FP64 */
  PGN_MIXED = 1  /* FP32 element Hessians + FP32 PCG; FP64 energy, gradient,
                    residual norm and iterative-refinement residual */
};

/* Hessian export layouts */
enum pgn_layout {
  PGN_LAYOUT_LOWER = 0, /* the entry set AMatrixSd stores (BasicTypes.cpp:18-33):
                           node blocks below the diagonal plus whole 3x3
                           diagonal node blocks; row-major CSR, cols sorted */
  PGN_LAYOUT_FULL = 1,  /* full symmetric scalar CSR */
  PGN_LAYOUT_BSR3 = 2   /* full symmetric 3x3-block CSR, blocks row-major */
};

typedef struct pgn_ctx pgn_ctx;
typedef struct pgn_model pgn_model;

/* ---- context ------------------------------------------------------------ */
const char* pgn_last_error(void);
int pgn_ctx_create(int device, pgn_ctx** out);
void pgn_ctx_destroy(pgn_ctx* ctx);
/* the hipStream_t every call of this context is issued on */
void* pgn_ctx_stream(pgn_ctx* ctx);
int pgn_ctx_synchronize(pgn_ctx* ctx);

/* ---- model setup -------------------------------------------------------- */
/* Simulable_Mesh::InitInternal (Simulable_Mesh.cpp:68) +
 * Simulable_FEM::CreateEnergyElements (Simulable_FEM.cpp:152) +
 * EnergyElement_FEM::Init (EnergyElement_FEM.cpp:59) +
 * MassElement_Lumped::ComputeAndStore_Mass (MassElement_Lumped.cpp:47).
 * X0: 3*n_nodes rest positions; T: 4*n_tets node ids; material/lame1/lame2/
 * density: per tet (ParameterSet::Param_Lame1/Lame2/Density). */
int pgn_model_create_tet(pgn_ctx* ctx, int n_nodes, const double* X0,
                         int n_tets, const int* T,
                         const unsigned char* material, const double* lame1,
                         const double* lame2, const double* density,
                         int precision, pgn_model** out);
/* Thin shell (Simulable_ThinShell, Simulable_ThinShell.cpp:53-120) on one
 * node set: StVK membrane triangles (EnergyElement_FEM on Tri3 with
 * MaterialModel_2Din3D_StVK, MaterialModels.h:376-440; energy = rest area x
 * thickness x Psi) and discrete-shell hinges (EnergyElement_DiscreteShells,
 * EnergyElement_DiscreteShells.cpp:22-198; E = kB shape0 (theta - theta0)^2,
 * SPD Hessian).  F: 3 node ids per triangle; lame1/lame2/thickness/density per
 * triangle (node masses density * area / 3, MassElement_Lumped with the
 * triangle's area as volume basis); hinges: 4 node ids each (the hinge edge
 * v0 v1, then the vertex opposite it in the face where the edge runs v1 -> v0,
 * then the vertex opposite it in the other face: the half-edge order of
 * EnergyElement_DiscreteShells.cpp:30-34); bending_k: kB per hinge
 * (Mat_BendingK).  Rest dihedral and shape factor come from X0. */
int pgn_model_create_shell(pgn_ctx* ctx, int n_nodes, const double* X0, int n_tris,
                           const int* F, const double* lame1, const double* lame2,
                           const double* thickness, const double* density,
                           int n_hinges, const int* hinges, const double* bending_k,
                           int precision, pgn_model** out);
void pgn_model_destroy(pgn_model* model);
int pgn_model_num_dofs(const pgn_model* model);
int pgn_model_num_elements(const pgn_model* model);
/* bytes of device memory the model holds */
size_t pgn_model_device_bytes(const pgn_model* model);

/* BC_FixDoF hard fixation (BC_FixDoF.cpp:51-77): mask has 3*n_nodes bytes */
int pgn_model_set_fixed(pgn_model* model, const unsigned char* mask);
/* BC_Gravity (BC_Gravity.cpp:34) with VectorAnimator_LerpValues
 * (VectorAnimators.h:36-69): value at stage k is lerp(g_ini, g_end, k/stages) */
int pgn_model_set_gravity(pgn_model* model, const double g_ini[3],
                          const double g_end[3], int num_stages);
/* BC_Force (BC_Force.cpp) / EnergyElement_Force (E = +x.f, g = +f), lerped */
int pgn_model_set_forces(pgn_model* model, const double* f_ini,
                         const double* f_end, int num_stages);
/* Simulable::StepBoundaryConditions (Simulable.cpp:1496) */
int pgn_model_step_load(pgn_model* model, int* stepped);
/* Simulable::BoundaryConditionsLoaded (Simulable.cpp:1476) */
int pgn_model_is_loaded(const pgn_model* model);
/* x := X0, velocity := 0 and every load stage := 0 */
int pgn_model_reset(pgn_model* model);

/* ISimulable::SetDOFVector / GetDOFVector (Position_X) */
int pgn_model_set_x(pgn_model* model, const double* x);
int pgn_model_get_x(pgn_model* model, double* x);
/* ISimulable::SetDOFVector / GetDOFVector (Tag::Velocity, Simulable.cpp:
 * 150-170); a model whose velocity was never set is at rest */
int pgn_model_set_velocity(pgn_model* model, const double* v);
int pgn_model_get_velocity(pgn_model* model, double* v);

/* ---- evaluation (IOptimProblem / ISimulable) ---------------------------- */
/* OptimProblem_BasicStatic::GetEnergy -> Simulable::AddEnergy (:321) */
int pgn_model_get_energy(pgn_model* model, double* energy);
/* OptimProblem_BasicStatic::GetGradient -> GetGradient + FixVectorDoFs */
int pgn_model_get_gradient(pgn_model* model, double* gradient);
/* Hessian pattern of a layout: nnz (scalar entries; blocks for BSR3) */
int pgn_model_hessian_nnz(pgn_model* model, int layout, long* nnz);
/* rowptr: 3n+1 (n+1 for BSR3) entries; col: nnz entries */
int pgn_model_hessian_pattern(pgn_model* model, int layout, long* rowptr,
                              int* col);
/* OptimProblem_BasicStatic::GetHessian -> GetHessian + FixMatrixDoFs.
 * Values in the model's Hessian precision, widened to FP64 (nnz*9 for BSR3) */
int pgn_model_get_hessian(pgn_model* model, int layout, double* values);
/* y = H p with the Hessian of the current x (assembled first, fixed rows and
 * columns projected; in the model's Hessian precision, p and y FP64 at the
 * boundary): the product of each CG iteration (LinearSolver_EigenCG.cpp:89,
 * Eigen's ConjugateGradient).  matrix_free = 0: the assembled BSR3 matrix;
 * 1: the sum of the element products over the stored element Hessians
 * (tet models), as the PCG forms it under PGN_MATFREE=1. */
int pgn_model_hessian_product(pgn_model* model, int matrix_free, const double* p,
                              double* y);
/* The preconditioner of LinearSolver_EigenCG's ConjugateGradient
 * (LinearSolver_EigenCG.cpp:68-89: Eigen's DiagonalPreconditioner), and the
 * two MI355X alternatives, applied to r with the Hessian of the current x:
 * z = M^-1 r (FP64 at the boundary, the model's Hessian precision inside).
 * precond 0: Jacobi; 1: 3x3 block Jacobi B; 2: two-level,
 * z = B^-1 r + P Ac^-1 P^T r (see pgn_model_coarse_space). */
int pgn_model_precond_apply(pgn_model* model, int precond, const double* r,
                            double* z);
/* The coarse space of the two-level preconditioner (no reference
 * counterpart; it replaces the diagonal preconditioner of the CG above when
 * precond = 2).  Nodes are aggregated by their rest positions into a grid
 * of boxes over the bounding box (about 128 of them, shaped to its extents;
 * PGN_COARSE_AGGS in the environment sets the count, at most 341); each
 * aggregate spans the three translations and the three infinitesimal
 * rotations about its centroid over its free DoFs (columns dependent on the
 * kept ones are dropped).  n_agg: aggregates; agg_of (n, may be NULL): the
 * aggregate of each node; cmask (n_agg, may be NULL): kept columns (bits 0-2
 * translations x y z, 3-5 rotations about x y z); coarse and coarse_inv
 * (36 n_agg^2 each, may be NULL): Ac = P^T H P for the Hessian of the
 * current x, dense row-major over coarse DoFs 6a + c with identity rows for
 * dropped columns, and the inverse the PCG applies. */
int pgn_model_coarse_space(pgn_model* model, int* n_agg, int* agg_of,
                           unsigned char* cmask, double* coarse,
                           double* coarse_inv);
/* IEnergyElement::GetLocalEnergy/Gradient/Hessian for every element at the
 * current x, from the fused E+g+H element kernels of the given precision
 * (PGN_FP64 or PGN_MIXED = all-FP32 kernels).  Any pointer may be NULL.
 * E: m values; g: 12 per element; H: full 12x12 row-major per element, over
 * the element's four connectivity slots (shell models: triangles first, each
 * with zeros in its fourth slot, then hinges). */
int pgn_model_element_egh(pgn_model* model, int precision, double* E,
                          double* g, double* H);

/* ---- linear solve (ILinearSolver / LinearSolver_EigenCG) ---------------- */
typedef struct {
  double tol;      /* relative residual |r| <= tol |b| (lsTolMaxError)   */
  int max_iters;   /* lsNumMaxIters                                       */
  int reg_iters;   /* LinearSolverOptions::regIters                       */
  int precond;     /* 0 = Jacobi (reference), 1 = 3x3 block Jacobi, 2 =
                      two-level (block Jacobi + aggregated rigid-body coarse
                      space, pgn_model_coarse_space; Chronopoulos-Gear PCG
                      over the assembled matrix only)                    */
  double inner_tol;/* MIXED only: FP32 inner relative tolerance per IR sweep */
  int solver;      /* LinearSolverOptions::type: PGN_LS_EIGENCG (default) or
                      PGN_LS_EIGENLDLT; with LDL^T, tol is maxError of the
                      exactness check and max_iters, precond, inner_tol are
                      not used                                             */
} pgn_linear_opts;

typedef struct {
  int status;      /* 0 = SUCCESS, 1 = FAILURE (LSResult)                  */
  int iterations;  /* PCG iterations over all regularization tries (0: LDL^T) */
  int reg_used;    /* regularization try that succeeded                   */
  int refinements; /* MIXED: outer FP64 refinement sweeps                  */
  double rel_residual;
} pgn_linear_result;

void pgn_linear_default_opts(pgn_linear_opts* o);
/* LinearSolver_EigenCG::SolveInternal (LinearSolver_EigenCG.cpp:51-125) or,
 * with solver = PGN_LS_EIGENLDLT, LinearSolver_EigenLDLT::SolveInternal
 * (LinearSolver_EigenLDLT.cpp:49-153) on the Hessian assembled at the
 * current x: solves H x = b, x is the warm start (CG) or kept until a
 * factorization succeeds (LDL^T).  A model whose band exceeds the direct
 * solver's limit gets PGN_ERR_INVALID. */
int pgn_linear_solve(pgn_model* model, const double* b, double* x,
                     const pgn_linear_opts* opts, pgn_linear_result* res);

/* ---- linear solve on a caller's matrix (ILinearSolver::Solve) ----------- */
/* ILinearSolver::Solve(MatrixSd& A, const VectorXd& b, VectorXd& x)
 * (LinearSolver.h:38-52) -> LinearSolver_EigenCG::SolveInternal
 * (LinearSolver_EigenCG.cpp:51-125) on a matrix the caller assembled, with
 * no device model: FP64 Jacobi PCG (Eigen's ConjugateGradient with
 * DiagonalPreconditioner: 1 / a_ii, 1 where a_ii is 0; relative residual
 * |r| <= tol |b|; warm start x; Eigen's iteration count) and the
 * regularization loop (b.x < 0 -> A += regT 10^(i-1) I, cumulative,
 * regT = 1e-6 |A|_F over the stored entries as given, as GetMatrixData
 * computes it, LinearSolver.cpp:57-66).  A is symmetric, given as row-major
 * CSR over n rows (rowptr: n+1 entries, nnz = rowptr[n]) either as its lower
 * part (PGN_LAYOUT_LOWER: AMatrixSd's entry set, whose whole 3x3 diagonal
 * node blocks put some entries above the diagonal; as with Eigen's
 * selfadjointView<Lower> only entries with col <= row enter the product) or
 * in full (PGN_LAYOUT_FULL); every row stores its diagonal entry.  Like the
 * reference (which adds the shifts to mA), the diagonal entries of val keep
 * the shifts of the tries that ran.  opts: tol, max_iters, reg_iters;
 * precond must be 0 (Jacobi) and solver PGN_LS_EIGENCG.  res->iterations
 * counts the PCG iterations of every try.  Host arrays. */
int pgn_csr_linear_solve(pgn_ctx* ctx, int n, int layout, const long* rowptr,
                         const int* col, double* val, const double* b,
                         double* x, const pgn_linear_opts* opts,
                         pgn_linear_result* res);

/* Most rows pgn_csr_linear_solve_noalloc accepts. */
#define PGN_CSR_NOALLOC_MAX_ROWS 262144
/* No counterpart in the reference: pgn_csr_linear_solve for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same computation and kernels,
 * same results bit for bit, but every array is a DEVICE pointer on the
 * context's device, A is given in full (PGN_LAYOUT_FULL only), n is at most
 * PGN_CSR_NOALLOC_MAX_ROWS, and the work space (FP64 vectors, partial sums,
 * state) is a static device array of 14.7 MB per device, held until the
 * process ends, shared by every call on that device: calls that overlap on
 * one GPU (from several threads or contexts) race on it and must be
 * serialized by the caller.  Calls no hipMalloc / hipFree; reads a few state
 * words back to the host between kernel chunks.  val's diagonal keeps the
 * regularization shifts, as in pgn_csr_linear_solve. */
int pgn_csr_linear_solve_noalloc(pgn_ctx* ctx, int n, const long* rowptr,
                                 const int* col, double* val, const double* b,
                                 double* x, const pgn_linear_opts* opts,
                                 pgn_linear_result* res);

/* ---- Newton solve (OptimSolver::SolveFull) ------------------------------ */
typedef struct {
  /* OptimSolverOptions (OptimSolver.h:23-71) */
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
  /* MI355X path controls (no reference counterpart) */
  int precond;        /* 0 = Jacobi, 1 = 3x3 block Jacobi, 2 = two-level
                         (fused batches: one coarse level per instance)    */
  double inner_tol;   /* MIXED: FP32 inner tolerance per refinement sweep */
  int adaptive_tol;   /* 0: every linear solve to lsTolMaxError (reference).
                         1: Eisenstat-Walker forcing (choice 2, gamma
                         ew_gamma, alpha 2) with Kelley's safeguard
                         (at least 0.5 tolMaxError / |g|), capped at
                         ew_max, floored at lsTolMaxError; a step whose
                         loosened solve gives a direction the line search
                         cannot descend along is taken again at
                         lsTolMaxError, as is every later step of the
                         solve (per instance in a batch)                    */
  double ls_noise_guard; /* 0 = reference line search.  > 0: also accept a
                         step whose energy rises by at most ls_noise_guard *
                         (|E_elastic| + |E_load|), i.e. within FP64 rounding
                         of the total, when |g| decreases                   */
  double ew_gamma;    /* Eisenstat-Walker gamma (default 0.9)              */
  double ew_max;      /* largest forcing term (default 0.2)                */
  /* OptimSolverOptions::lsSolverType (OptimSolver.h:40): PGN_LS_EIGENCG
     (default here) or PGN_LS_EIGENLDLT (the reference's default; linear
     solves are then exact and adaptive_tol is not used)                  */
  int lsSolverType;
  /* OptimSolverOptions::qpSolverType (OptimSolver.h:30), QPSolverType values
     (BasicTypes.h:48-55): PGN_QP_NEWTON (default), PGN_QP_STEEPEST or
     PGN_QP_LBFGS (single models only; ComputeStep_LBFGS,
     OptimSolver.cpp:522-591, with a history of 3 pairs, corrected: the
     previous x and g are updated at every step, the second loop wraps
     modulo the history length and the pair count saturates at it, where
     the reference takes them once, wraps modulo the problem size and falls
     back to one pair when its ring wraps; the history starts afresh with
     every solve)                                                         */
  int qpSolverType;
  /* MI355X path control (no reference counterpart): fused batches with
     precond = 2 only.  An instance whose linear solve takes more than
     coarse_giveup PCG iterations continues its solve with block Jacobi (the
     aggregated rigid-body coarse level then does not hold what limits its
     convergence, as on buckling instances, and the cheaper iteration wins);
     a solve where no active instance keeps the level runs block Jacobi
     outright.  0 (default): every instance keeps it                      */
  int coarse_giveup;
} pgn_newton_opts;

/* QPSolverType (BasicTypes.h:48-55) */
enum pgn_qp_solver { PGN_QP_STEEPEST = 0, PGN_QP_NEWTON = 1, PGN_QP_LBFGS = 2 };

/* OSResult (BasicTypes.h:69-77) */
enum pgn_os_result {
  PGN_OS_ONGOING = 0,
  PGN_OS_SUCCESS,
  PGN_OS_MINSTEP,
  PGN_OS_MINIMPR,
  PGN_OS_FAILURE,
  PGN_OS_MAXITER,
  PGN_OS_NONDESC
};

typedef struct {
  int result;         /* pgn_os_result */
  int iterations;     /* Newton iterations (SolveFull itCount)          */
  int pcg_iterations; /* total PCG iterations                           */
  int line_search_bisections;
  int energy_evals;   /* E+g evaluations                                 */
  int hessian_evals;
  double objective;
  double optimality;  /* |g_free| */
  double seconds;     /* wall time of the solve                          */
} pgn_newton_result;

void pgn_newton_default_opts(pgn_newton_opts* o);
/* OptimSolver::Init + SolveFull (OptimSolver.cpp:36-55,163-264) with
 * OptimSolver_USQP_LS::SolveStepInternal (OptimSolver_USQP_LS.cpp:18-68),
 * Newton steps (ComputeStep_Newton :468) and LineSearch_Simple (:688). */
int pgn_newton_solve(pgn_model* model, const pgn_newton_opts* opts,
                     pgn_newton_result* res);
/* One implicit (backward-Euler) time step: OptimProblem_BasicDynamic
 * (OptimProblem_BasicDynamic.cpp:20-170) minimised by SolveFull.  The step
 * keeps x0 = x and v0 = velocity, steps the boundary conditions once
 * (OnSolveStart :133-146), starts from x0 + dt v0 and minimises
 *   E(x) + dt^2/2 a.M a,  a = (x - x0)/dt^2 - v0/dt   (GetEnergy :68-86)
 * with gradient g + M a (:88-107) and Hessian H + M/dt^2 (:109-121), M the
 * lumped masses; every iteration is fully constrained (:127-129).  After the
 * solve the velocity is (x - x0)/dt (PosPerformStep :159-169).  The
 * solver's previous x (warm start of the linear solve) carries over from one
 * step to the next until pgn_model_reset. */
int pgn_dynamic_step(pgn_model* model, double dt, const pgn_newton_opts* opts,
                     pgn_newton_result* res);

/* ---- fused batches of independent solves -------------------------------- */
/* Many independent quasi-static problems (tet meshes and thin shells) solved
 * together: pgn_batch_build uploads them as ONE device model over the
 * disjoint union of their nodes and elements, so every element, gather,
 * assembly and PCG launch covers the whole batch.  Each instance keeps the
 * reference's own Newton state (load stages, LineSearch_Simple, the
 * SolveFull status machine: OptimSolver.cpp:163-264, 688-736); the linear
 * solve is one PCG over the block-diagonal batch Hessian.  Instance DoF
 * vectors use the single-model layout (3 per node, node-major). */
typedef struct pgn_batch pgn_batch;
int pgn_batch_create(pgn_ctx* ctx, int precision, pgn_batch** out);
void pgn_batch_destroy(pgn_batch* batch);
/* stage one instance (arguments as pgn_model_create_tet / _shell) */
int pgn_batch_add_tet(pgn_batch* batch, int n_nodes, const double* X0, int n_tets,
                      const int* T, const unsigned char* material, const double* lame1,
                      const double* lame2, const double* density, int* instance);
int pgn_batch_add_shell(pgn_batch* batch, int n_nodes, const double* X0, int n_tris,
                        const int* F, const double* lame1, const double* lame2,
                        const double* thickness, const double* density, int n_hinges,
                        const int* hinges, const double* bending_k, int* instance);
/* boundary conditions of one staged instance (as pgn_model_set_*) */
int pgn_batch_set_fixed(pgn_batch* batch, int instance, const unsigned char* mask);
int pgn_batch_set_gravity(pgn_batch* batch, int instance, const double g_ini[3],
                          const double g_end[3], int num_stages);
int pgn_batch_set_forces(pgn_batch* batch, int instance, const double* f_ini,
                         const double* f_end, int num_stages);
/* builds the fused device model; no instance can be added afterwards */
int pgn_batch_build(pgn_batch* batch);
int pgn_batch_num_instances(const pgn_batch* batch);
size_t pgn_batch_device_bytes(const pgn_batch* batch);
/* x := X0 and load stages := 0 for every instance */
int pgn_batch_reset(pgn_batch* batch);
/* OptimSolver::SolveFull per instance; res: one result per instance */
int pgn_batch_newton(pgn_batch* batch, const pgn_newton_opts* opts,
                     pgn_newton_result* res);
int pgn_batch_get_x(pgn_batch* batch, int instance, double* x);
/* per-instance total energy at the current x and load stages */
int pgn_batch_get_energy(pgn_batch* batch, double* energy);

/* ---- request loop (serving) ---------------------------------------------- */
/* A persistent loop that solves independent requests on one GPU with
 * continuous batching.  Requests are copied at submit and queued.  Each lane
 * (a worker thread with its own HIP stream) keeps up to cohorts_per_lane
 * fused batches resident (pgn_batch semantics: one launch per kernel covers
 * a batch, every request keeps its own Newton state) and advances them one
 * Newton iteration at a time; between iterations it admits queued requests
 * as a new batch while their estimated device footprint fits
 * hbm_budget_bytes (by default, group_by_size, the cheapest queued request
 * and requests of at most twice its footprint; a request of solo_nodes or
 * more alone, solved by the single-model PCG), hands each request back the
 * moment its own solve ends, and frees a batch's memory when its last
 * request has ended.  A request arriving while others run starts within one
 * Newton iteration of a lane; solves overlap on the device across lanes.
 * No reference counterpart: PhySim solves one Simulable per OptimSolver
 * call (Simulator.cpp:38-72). */
typedef struct {
  size_t hbm_budget_bytes; /* device memory the resident requests may hold */
  int lanes;               /* concurrent batches (streams), 1..16          */
  int max_batch;           /* requests per batch                           */
  int precision;           /* PGN_FP64 or PGN_MIXED                        */
  pgn_newton_opts newton;  /* solver options of every request              */
  int group_by_size;       /* 1 (default): batches of similar footprint,
                              cheapest first; 0: arrival order              */
  double max_wait_s;       /* group_by_size: a request queued this long is
                              taken before cheaper ones (default 2 s)       */
  int cohorts_per_lane;    /* fused batches a lane keeps resident at once
                              (default 4; 1: a lane's batch ends before the
                              lane admits more)                             */
  long solo_nodes;         /* requests of at least this many nodes run as a
                              batch of their own, solved by the single-model
                              PCG (default 32768)                           */
} pgn_server_opts;
typedef struct pgn_server pgn_server;
void pgn_server_default_opts(pgn_server_opts* o);
int pgn_server_create(int device, const pgn_server_opts* opts, pgn_server** out);
/* stops the lanes after their current Newton iteration; requests not yet
 * solved end with PGN_ERR_INVALID ("server destroyed"), their waiters wake */
void pgn_server_destroy(pgn_server* server);
/* submit a tet or thin-shell request (mesh arguments as pgn_model_create_*);
 * fixed: 3 per node or NULL; gravity: 3 or NULL; forces: 3 per node or NULL;
 * both loads ramped from zero over `stages` load stages */
int pgn_server_submit_tet(pgn_server* server, int n_nodes, const double* X0, int n_tets,
                          const int* T, const unsigned char* material,
                          const double* lame1, const double* lame2,
                          const double* density, const unsigned char* fixed,
                          const double* gravity, const double* forces, int stages,
                          long* ticket);
int pgn_server_submit_shell(pgn_server* server, int n_nodes, const double* X0,
                            int n_tris, const int* F, const double* lame1,
                            const double* lame2, const double* thickness,
                            const double* density, int n_hinges, const int* hinges,
                            const double* bending_k, const unsigned char* fixed,
                            const double* gravity, const double* forces, int stages,
                            long* ticket);
int pgn_server_poll(pgn_server* server, long ticket, int* done);
/* blocks until the request is solved; x: 3 per node (or NULL); latency_s:
 * submit -> result on host; queue_s: submit -> its batch started;
 * res->seconds: its batch start -> its own solve's end */
int pgn_server_wait(pgn_server* server, long ticket, double* x, pgn_newton_result* res,
                    double* latency_s, double* queue_s);
int pgn_server_stats(pgn_server* server, long* batches, long* solved,
                     size_t* bytes_in_use);

/* ---- measurement hooks -------------------------------------------------- */
/* Runs `reps` launches of one kernel on the model's current state, on the
 * context stream, and returns the mean launch time (HIP events) in ms.
 * which: 0 = element E+g+H (model precision), 1 = element E+g (FP64),
 *        2 = BSR Hessian gather+fix, 3 = BSR SpMV (one PCG matrix product),
 *        4 = element E+g+H FP64 (plain path kernel), 5 = element E+g+H FP32,
 *        7 = calibration copy of 256 MiB (FP32)                            */
int pgn_time_kernel(pgn_model* model, int which, int reps, double* ms_mean);
/* algorithmic bytes one launch of `which` moves (for roofline accounting);
 * also 7 = the calibration copy of pgn_time_kernel, 8 = PCG SpMV launch (K1),
 * 9 = PCG vector-update launch (K2) */
int pgn_kernel_bytes(pgn_model* model, int which, double* bytes);
/* Live timing of the PCG launches inside linear / Newton solves: when on,
 * the middle SpMV launch (K1) and vector-update launch (K2) of every queued
 * chunk of PCG iterations (8 to 64 iterations) carry HIP events stamped by
 * their own dispatch (hipExtLaunchKernel start / stop events) on the context
 * stream, a sample that leaves the solve's own timing unperturbed.  Switching on resets the counters; switching off keeps them
 * for pgn_model_get_profile. */
int pgn_model_set_profiling(pgn_model* model, int on);
/* which: 0 = K1, 1 = K2: sampled launches that did work and their summed ms;
 * 2, 3, 4: the SpMV time (us, in total_ms) with 4, 8, 16 lanes per block
 * row, timed once per model, and in count the lanes per row in use (16 from
 * 12 mean blocks per row, 8 from 6, else 4; PGN_PCG_LPR overrides);
 * 5: in count the bytes per block column of the SELL-64 copy (2: 16-bit
 * offsets from the row, 4: 32-bit columns, 0: no SELL-64 copy);
 * 6, 7, 8: in count the tile of the tet element kernel the in-library
 * autotuner chose (timed on the model's own elements at the first launch;
 * -1 before it; 0..5 = workgroup size x elements per lane: 256x1, 128x1,
 * 64x1, 256x2, 128x2, 64x2) for the Hessian (model precision), FP64 E+g
 * and E+g+H (model precision), in total_ms its time per launch in us;
 * 9 + t: the Hessian launch's time (us) with tile t (0 if not timed) */
int pgn_model_get_profile(pgn_model* model, int which, long* count,
                          double* total_ms);

/* ---- scene ingestion ----------------------------------------------------- */
/* TetGen volume meshes, the format of Serializer_Mesh::ReadVolume_TET
 * (Serializer_Mesh.h:46-52; declared, never implemented in the reference):
 * base.node and base.ele (base may carry either extension), Tet4 only,
 * indices from 0 or 1; a tet listed with negative volume gets its last two
 * corners swapped.  X: 3 per node; T: 4 per tet, 0-based. */
int pgn_tetgen_sizes(const char* base, int* n_nodes, int* n_tets);
int pgn_tetgen_read(const char* base, double* X, int* T);

/* A scene document in the reference's XML vocabulary: a <scene> root whose
 * children are the elements Serializer_XML reads (<simulable> as
 * LoadSimulable_VolumeFEM, Serializer_XML.cpp:1199-1325; <boundary> with
 * <bc> children as LoadBoundary, :1507-1647; <problem> as
 * LoadOptimProblem, :1327-1354; <solver> as LoadOptimSolver, :1356-1505).
 * Volume Tet4 models with CONH / STVK materials, Gravity and hard
 * FixPosition conditions, Static / Dynamic problems, Newton / Steepest /
 * LBFGS steps with EigenCG / EigenLDLT; anything else fails with
 * PGN_ERR_INVALID naming it.  meshInit FILEVOLUME reads its `path`
 * attribute (relative to the document; the reference reads `mesh`), GRID
 * builds CreateTetMesh_Grid's 5-tet split (a TODO in the reference's tet
 * reader).  Host only: no device is used until pgn_scene_create_model. */
typedef struct pgn_scene pgn_scene;
int pgn_scene_load(const char* xml_path, pgn_scene** out);
void pgn_scene_destroy(pgn_scene* scene);
int pgn_scene_sizes(const pgn_scene* scene, int* n_nodes, int* n_tets);
/* X0: rest positions (meshRest, else meshInit), X: initial positions
 * (meshInit), 3 per node; T: 4 per tet; per tet material kind (pgn_material),
 * Lame parameters (from YOUNG / POISSON by ParameterSet::
 * InitLinearFromYoungPoisson, ParameterSet.cpp:54-70) and density.  Any
 * pointer may be NULL. */
int pgn_scene_mesh(const pgn_scene* scene, double* X0, double* X, int* T,
                   unsigned char* material, double* lame1, double* lame2,
                   double* density);
/* fixed: 3 per node; gravity as BC_Gravity without an animator applies it:
 * the same value at every stage (g_ini = g_end), stages = numStep */
int pgn_scene_boundary(const pgn_scene* scene, unsigned char* fixed,
                       double g_ini[3], double g_end[3], int* stages,
                       int* has_gravity);
/* the solver options of <solver> over pgn_newton_default_opts; dynamic and
 * its time step from <problem> */
int pgn_scene_solver(const pgn_scene* scene, pgn_newton_opts* opts, int* dynamic,
                     double* dt);
/* pgn_model_create_tet + pgn_model_set_fixed / _set_gravity (+ set_x when
 * meshInit differs from meshRest) */
int pgn_scene_create_model(pgn_ctx* ctx, const pgn_scene* scene, int precision,
                           pgn_model** out);

/* ---- discrete elastic rods (Simulable_DER) ------------------------------ */
/* A batch of open rods solved together (every rod keeps its own Newton
 * state).  Reference DoF layout per rod: 3n positions then n-1 edge angles
 * (KinematicsTree.cpp:33-40); batch vectors concatenate the rods.
 * Setup: Simulable_DER::CreateKinematicsEle / CreateEnergyElements
 * (Simulable_DER.cpp:66-150), Mesh_DER::Init (Mesh_DER.cpp:97-164).
 * X0: 3 per vertex, all rods; frame0_normal: 3 per rod, the normal of the
 * first edge's reference frame (projected to be orthogonal to the edge);
 * per rod: Young's modulus (Param_Young), shear modulus (Param_Lame2),
 * cross-section radii (Size_0 = h, Size_1 = w) and density (kg per metre of
 * rod: node masses are density * L / 2 per edge). */
typedef struct pgn_rods pgn_rods;
int pgn_rods_create(pgn_ctx* ctx, int n_rods, const int* n_verts,
                    const double* X0, const double* frame0_normal,
                    const double* young, const double* shear,
                    const double* radius_h, const double* radius_w,
                    const double* density, int precision, pgn_rods** out);
void pgn_rods_destroy(pgn_rods* rods);
long pgn_rods_num_dofs(const pgn_rods* rods);
int pgn_rods_set_fixed(pgn_rods* rods, const unsigned char* mask);
int pgn_rods_set_gravity(pgn_rods* rods, const double g_ini[3],
                         const double g_end[3], int num_stages);
int pgn_rods_reset(pgn_rods* rods);
int pgn_rods_set_x(pgn_rods* rods, const double* x);
int pgn_rods_get_x(pgn_rods* rods, double* x);
/* per-rod energies (n_rods) / batch gradient, fixed entries zeroed */
int pgn_rods_get_energy(pgn_rods* rods, double* energy);
int pgn_rods_get_gradient(pgn_rods* rods, double* gradient);
/* dense Hessian of one rod in its reference layout ((4n-1)^2 values, fixed
 * rows / columns identity) */
int pgn_rods_get_hessian_dense(pgn_rods* rods, int rod, double* H);
/* reference frames (tangent, normal: 3 per edge) and reference twist (per
 * vertex) of every rod, edges / vertices concatenated in rod order */
int pgn_rods_get_frames(pgn_rods* rods, double* tan, double* nor, double* rt);
int pgn_rods_set_frames(pgn_rods* rods, const double* tan, const double* nor,
                        const double* rt);
/* Simulable::UpdateKinematics: KEleDERMatTwist / KEleDERRefTwist
 * (KEleDER.cpp:94-98,216-220) */
int pgn_rods_update_kinematics(pgn_rods* rods);
/* bend-twist element outputs at every vertex (zero where no element): E,
 * g over (e1, e2, theta1, theta2) (8), H (64, row-major; a PGN_MIXED batch
 * evaluates its Hessians in FP32, energies and gradients in FP64) */
int pgn_rods_element_bt(pgn_rods* rods, double* E, double* g8, double* H64);
/* EnergyElement_DER_BendTwist::ComputeAndStore_{Energy,Gradient,Hessian}
 * (EnergyElement_DER_BendTwist.cpp:67-249) on caller-given element states,
 * evaluated by the device kernel: per element the 27 parameters of the
 * reference's getBT{Energy,Gradient,Hessian}OPT kernels (shear and Young's
 * moduli, radii and lengths of both edges, rest curvatures K0 K1, rest
 * twist, frames n1 u1 n2 u2, reference twist, one unused), the edge
 * vectors e1 e2 (3 each) and the material angles (2).  precision PGN_MIXED
 * evaluates in FP32 (the mixed batches' Hessian path). */
int pgn_rod_bt_eval(pgn_ctx* ctx, int count, const double* prm, const double* e1,
                    const double* e2, const double* theta, int precision, double* E,
                    double* g8, double* H64);
/* OptimSolver::SolveFull for every rod (LinearSolver_EigenLDLT semantics on
 * each rod's banded Hessian); res: n_rods results */
int pgn_rods_newton(pgn_rods* rods, const pgn_newton_opts* opts,
                    pgn_newton_result* res);

/* ---- host-only helpers (no GPU needed; used by CPU tests) --------------- */
/* Builds the BSR3 pattern and the element->block slot map of a tet mesh the
 * way pgn_model_create_tet does.  nb_out: number of 3x3 blocks. */
int pgn_host_bsr_pattern(int n_nodes, int n_tets, const int* T, long* nb_out,
                         long* rowptr /* n+1 or NULL */,
                         int* col /* nb or NULL */);

#ifdef __cplusplus
}
#endif
#endif
