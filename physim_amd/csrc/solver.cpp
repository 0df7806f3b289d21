// solver.cpp -- linear solve (ILinearSolver) and Newton driver (OptimSolver)
// over a device-resident model.  Host C++ orchestrates; every vector stays in
// HBM and only scalars (energy, norms, PCG status) come back to the host.
#include "solver.hpp"

#include "ldlt.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>

namespace pgn {

// Runs PCG in the model's Hessian precision on rhs `pb` with solution /
// warm start in the PCG x vector.  Returns iterations (Eigen's count).
static int pcg_run(pgn_model& M, double tol, int maxit, int precond, bool x_zero,
                   double& rr, double& bb) {
  hipStream_t s = M.stream();
  const PcgArgs a = M.pcg_args(tol, precond);
  launch_pcg_init(s, M.hfp32, a, M.pb.p, x_zero ? 1 : 0);
  if (a.coarse) {
    M.coarse_join();  // the coarse operator built on the side stream
    launch_coarse_apply(s, M.hfp32, a);
  }
  if (M.profiling && M.prof_ev.size() < 4) {
    while (M.prof_ev.size() < 4) {
      hipEvent_t e;
      PGN_HIP(hipEventCreate(&e));
      M.prof_ev.push_back(e);
    }
  }
  // chunks after the first (whose iteration 0 starts the recurrences) replay
  // a captured graph of their launches when they have the full length
  PcgArgs ga = a;
  ga.tol = 0.0;  // read at iteration 0 only
  ga.kctl = M.kctl.p;
  const bool graphs = !M.profiling && (M.use_graphs >= 1 ||
                                        (M.use_graphs < 0 && M.n <= pgn::GRAPH_AUTO_MAX_NODES));
  int k = 0, chunk = 8;
  for (;;) {
    const int c = std::min(chunk, maxit - k);
    const int js = c / 2;  // sampled iteration of the chunk (mid-stream)
    hipGraphExec_t ex = nullptr;
    if (graphs && k > 0 && c == chunk)
      ex = M.pcg_graphs.get(s, c, k & 1, ga, [&](hipStream_t cs) {
        for (int j = 0; j < c; ++j) {
          launch_pcg_k1a(cs, M.hfp32, ga, j);
          launch_pcg_k1(cs, M.hfp32, ga, j, nullptr, nullptr);
          launch_pcg_k2(cs, M.hfp32, ga, j, nullptr, nullptr);
          if (ga.coarse) launch_coarse_apply(cs, M.hfp32, ga);
        }
        launch_pcg_status(cs, ga, c, 0);
      });
    if (ex) {
      M.hkctl[0] = k;
      M.hkctl[1] = maxit;
      PGN_HIP(hipMemcpyAsync(M.kctl.p, M.hkctl, 2 * sizeof(int), hipMemcpyHostToDevice, s));
      PGN_HIP(hipGraphLaunch(ex, s));
      k += c;
    } else {
      for (int j = 0; j < c; ++j) {
        // profiling brackets one K1 and K2 of every chunk with events: a
        // sample of the launches, so the timed solve itself stays unperturbed
        const bool ev = M.profiling && j == js;
        launch_pcg_k1a(s, M.hfp32, a, k + j);
        launch_pcg_k1(s, M.hfp32, a, k + j, ev ? M.prof_ev[0] : nullptr,
                      ev ? M.prof_ev[1] : nullptr);
        launch_pcg_k2(s, M.hfp32, a, k + j, ev ? M.prof_ev[2] : nullptr,
                      ev ? M.prof_ev[3] : nullptr);
        if (a.coarse) launch_coarse_apply(s, M.hfp32, a);
      }
      k += c;
      launch_pcg_status(s, a, k, maxit);
    }
    PGN_HIP(hipMemcpyAsync(M.hsc, M.sc.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    M.sync();
    // launches after convergence return at once: the sample counts only if
    // its iteration still did work
    if (M.profiling && (!M.hsc->done || M.hsc->iters + 1 > k - c + js)) {
      float t1 = 0.f, t2 = 0.f;
      PGN_HIP(hipEventElapsedTime(&t1, M.prof_ev[0], M.prof_ev[1]));
      PGN_HIP(hipEventElapsedTime(&t2, M.prof_ev[2], M.prof_ev[3]));
      M.prof_ms[0] += t1;
      M.prof_ms[1] += t2;
      M.prof_n[0] += 1;
      M.prof_n[1] += 1;
    }
    if (M.hsc->done || k >= maxit) break;
    chunk = std::min(2 * chunk, M.pcg_chunk_max);
  }
  rr = M.hsc->rr;
  bb = M.hsc->bb;
  return M.hsc->iters;
}

// One unregularized solve of A x = b (x: FP64 warm start, in/out).
static void solve_once(pgn_model& M, const double* b, double* x, const pgn_linear_opts& o,
                       int& iters, int& refinements, double& relres) {
  hipStream_t s = M.stream();
  const long N = 3L * M.n;
  const size_t ts = M.tsize();
  char* px = M.pv.p;  // PCG solution vector
  if (!M.hfp32) {
    // plain FP64: Eigen ConjugateGradient semantics on (b, x) directly
    PGN_HIP(hipMemcpyAsync(M.pb.p, b, N * ts, hipMemcpyDeviceToDevice, s));
    PGN_HIP(hipMemcpyAsync(px, x, N * ts, hipMemcpyDeviceToDevice, s));
    double rr = 0, bb = 0;
    iters = pcg_run(M, o.tol, o.max_iters, o.precond, false, rr, bb);
    if (bb == 0.0) {
      PGN_HIP(hipMemsetAsync(x, 0, N * sizeof(double), s));
      relres = 0.0;
    } else {
      PGN_HIP(hipMemcpyAsync(x, px, N * ts, hipMemcpyDeviceToDevice, s));
      relres = std::sqrt(rr / bb);
    }
    refinements = 0;
    return;
  }
  // mixed: FP32 PCG corrections inside an FP64 iterative refinement
  const double bb = M.dot(b, b, N);
  iters = 0;
  refinements = 0;
  if (bb == 0.0) {
    PGN_HIP(hipMemsetAsync(x, 0, N * sizeof(double), s));
    relres = 0.0;
    return;
  }
  const double thr = o.tol * o.tol * bb;
  double rr = 0.0;
  // the FP64 residual reads the SELL-64 copy when the PCG does (packed for
  // this try), else the BSR3 matrix
  const PcgArgs ra = M.pcg_args(0.0, 0);
  for (;;) {
    if (ra.sell)
      launch_residual64_sell(s, true, ra, x, b, M.r64.p, M.red64.p, M.scal.p + 4);
    else
      launch_residual64(s, true, M.n, M.rowptr.p, M.bcol.p, M.bval.p, x, b, M.r64.p,
                        M.red64.p, M.scal.p + 4);
    PGN_HIP(hipMemcpyAsync(M.hpin + 4, M.scal.p + 4, sizeof(double), hipMemcpyDeviceToHost, s));
    M.sync();
    rr = M.hpin[4];
    // also stops on NaN
    if (!(rr >= thr) || iters >= o.max_iters || refinements >= MAX_REFINEMENTS) break;
    const double inner = std::max(o.inner_tol, o.tol * std::sqrt(bb / rr));
    launch_cast_f64_to_f32(s, N, M.r64.p, reinterpret_cast<float*>(M.pb.p));
    double irr = 0, ibb = 0;
    const int it = pcg_run(M, inner, o.max_iters - iters, o.precond, true, irr, ibb);
    launch_axpy_f32_to_f64(s, N, reinterpret_cast<const float*>(px), x);
    iters += it;
    ++refinements;
    // The residual rounded to zero in FP32: no FP32 correction can follow.
    // From x = 0 the PCG otherwise always takes a step (inner <= 1 here), even
    // when Eigen's count says 0: that count is 0 also when one iteration ran
    // and met the tolerance.
    if (ibb == 0.0) break;
  }
  relres = std::sqrt(rr / bb);
}

// LinearSolver_EigenCG::SolveInternal (LinearSolver_EigenCG.cpp:51-125):
// regularize A += regT 10^(i-1) I (cumulative) while b.x < 0.
void linear_solve(pgn_model& M, const double* b, double* x, const pgn_linear_opts& o,
                  pgn_linear_result& res) {
  if (o.solver == PGN_LS_EIGENLDLT) {
    ldlt_linear_solve(M, b, x, o, res);
    return;
  }
  PGN_CHECK(M.hess_valid, "Hessian not assembled");
  PGN_CHECK(o.precond >= 0 && o.precond <= 2,
            "precond: 0 (Jacobi), 1 (block Jacobi) or 2 (two-level)");
  const bool coarse = o.precond == 2;
  if (coarse) {
    PGN_CHECK(M.pcg_cg1 && !M.use_matfree,
              "the two-level preconditioner needs the Chronopoulos-Gear PCG over the "
              "assembled matrix (PGN_PCG and PGN_MATFREE unset)");
    M.coarse_prepare();
  }
  M.tune();
  if (M.use_matfree) M.mf_args();  // its element product buffer
  hipStream_t s = M.stream();
  const long N = 3L * M.n;
  double* xprev = M.tmp64b.p;
  PGN_HIP(hipMemcpyAsync(xprev, x, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  res.status = 1;
  res.iterations = 0;
  res.reg_used = 0;
  res.refinements = 0;
  res.rel_residual = 0.0;
  double regT = -1.0;
  const int regs = std::max(o.reg_iters, 1);
  for (int i = 0; i < regs; ++i) {
    M.coarse_join();  // the last try's coarse build reads the matrix changed below
    if (i != 0) {
      if (regT < 0.0) regT = 1e-6 * M.frob_lower();  // LinearSolver.cpp:57-66
      launch_add_diag(s, M.hfp32, M.n, M.diag.p, M.bval.p, regT * std::pow(10.0, i - 1));
      M.mf_shift += regT * std::pow(10.0, i - 1);
      M.update_precond();
      res.reg_used = i;
    }
    PGN_HIP(hipMemcpyAsync(x, xprev, N * sizeof(double), hipMemcpyDeviceToDevice, s));
    // this try's coarse operator (side stream), then the SpMV's copy of the
    // matrix, which overlaps it
    if (coarse) M.coarse_update(false, true);
    M.sell_pack();
    int it = 0, refs = 0;
    double rel = 0.0;
    solve_once(M, b, x, o, it, refs, rel);
    res.iterations += it;
    res.refinements += refs;
    res.rel_residual = rel;
    if (M.dot(b, x, N) < 0.0) continue;  // indefinite (regSign > 0)
    res.status = 0;
    break;
  }
  M.coarse_join();
}

// ---- Newton -----------------------------------------------------------------

void newton_solve(pgn_model& M, const pgn_newton_opts& o, pgn_newton_result& R, bool dynamic) {
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = M.stream();
  const long N = 3L * M.n;
  M.nw_xsave.ensure(N);
  M.nw_v0.ensure(N);
  M.nw_dx.ensure(N);
  M.nw_b.ensure(N);
  double* x = M.x.p;
  double* xsave = M.nw_xsave.p;
  double* v0 = M.nw_v0.p;
  double* dx = M.nw_dx.p;
  double* bvec = M.nw_b.p;
  pgn_linear_opts lo;
  pgn_linear_default_opts(&lo);
  lo.tol = o.lsTolMaxError;
  lo.max_iters = o.lsNumMaxIters;
  lo.precond = o.precond;
  lo.inner_tol = o.inner_tol;
  PGN_CHECK(o.lsSolverType == PGN_LS_EIGENCG || o.lsSolverType == PGN_LS_EIGENLDLT,
            "lsSolverType: PGN_LS_EIGENCG or PGN_LS_EIGENLDLT");
  lo.solver = o.lsSolverType;
  PGN_CHECK(o.qpSolverType == PGN_QP_NEWTON || o.qpSolverType == PGN_QP_STEEPEST ||
                o.qpSolverType == PGN_QP_LBFGS,
            "qpSolverType: PGN_QP_STEEPEST, PGN_QP_NEWTON or PGN_QP_LBFGS");
  const bool direct = o.lsSolverType == PGN_LS_EIGENLDLT && o.qpSolverType == PGN_QP_NEWTON;
  if (direct) M.ldlt_prepare();  // a band beyond the limit fails before any step
  // LBFGS history (ComputeStep_LBFGS, OptimSolver.cpp:522-591), fresh per solve
  const bool lbfgs = o.qpSolverType == PGN_QP_LBFGS;
  bool lb_init = false;
  int lb_count = 0, lb_newest = 0;
  if (lbfgs) {
    M.lb_xp.ensure(N);
    M.lb_gp.ensure(N);
    M.lb_S.ensure((size_t)pgn_model::LBFGS_M * N);
    M.lb_Y.ensure((size_t)pgn_model::LBFGS_M * N);
  }
  R = pgn_newton_result{};

  auto d2d = [&](double* dst, const double* src) {
    PGN_HIP(hipMemcpyAsync(dst, src, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  };

  // PGN_TRACE=1 prints one line per Newton step (diagnostics only)
  const bool trace = std::getenv("PGN_TRACE") != nullptr;
  // OptimSolver::Init (OptimSolver.cpp:36-55); a time step keeps the previous
  // x of the solver's last step instead (the solver is initialised once and
  // SolveFull runs per time step)
  if (!dynamic) d2d(v0, x);
  double obj = 0.0, gg = 0.0;
  M.eval_eg(obj, gg);
  R.energy_evals++;
  double opt = std::sqrt(gg);
  bool g_valid = true;  // M.g holds the gradient at the current x and loads
  int result = PGN_OS_SUCCESS;
  int minImpr = 0, minStep = 0;
  bool steepest = false;
  int itCount = 0;
  // Eisenstat-Walker forcing (choice 2, gamma 0.9, alpha 2) when adaptive_tol:
  // the linear tolerance follows the Newton residual's own convergence rate.
  double eta = o.ew_max, g_prev = -1.0;
  // A step whose loosened linear solve gave a direction the line search
  // could not descend along is taken again with lsTolMaxError, and so is
  // every later step of the solve: on buckling-prone models the inexact
  // directions otherwise end the solve with NONDESC where the exact ones
  // converge.
  bool ew_off = false;
  bool inexact = false;  // the current step's linear solve was looser than lsTolMaxError

  // OptimSolver_USQP_LS::SolveStepInternal (OptimSolver_USQP_LS.cpp:18-68)
  auto solve_step = [&](bool use_steepest) {
    if (!dynamic) M.step_load();   // OnIterStart: StepBoundaryConditions (static)
    d2d(xsave, x);                 //              GetState
    M.eval_eg(obj, gg);            // GetEnergy + GetGradient
    R.energy_evals++;
    g_valid = true;
    opt = std::sqrt(gg);
    if (opt < o.tolMaxError) {
      result = PGN_OS_SUCCESS;
      return;
    }
    // ComputeStep (OptimSolver.cpp:313-352); trial 0 always succeeds
    if (!use_steepest && lbfgs) {
      // ComputeStep_LBFGS, corrected: the newest pair is (x - x_prev,
      // g - g_prev) of the last LBFGS step, then the two-loop recursion
      // (Nocedal & Wright, Numerical Optimization, Alg. 7.4)
      inexact = false;
      if (!lb_init) {
        d2d(M.lb_xp.p, xsave);
        d2d(M.lb_gp.p, M.g.p);
        launch_lincomb64(s, N, -1.0, M.g.p, 0.0, M.g.p, dx);  // steepest first
        lb_init = true;
      } else {
        const int Mh = pgn_model::LBFGS_M;
        const int slot = lb_count == 0 ? 0 : (lb_newest + 1) % Mh;
        double* S = M.lb_S.p;
        double* Y = M.lb_Y.p;
        launch_lincomb64(s, N, 1.0, xsave, -1.0, M.lb_xp.p, S + slot * N);
        launch_lincomb64(s, N, 1.0, M.g.p, -1.0, M.lb_gp.p, Y + slot * N);
        d2d(M.lb_xp.p, xsave);
        d2d(M.lb_gp.p, M.g.p);
        lb_newest = slot;
        if (lb_count < Mh) lb_count++;
        double a[pgn_model::LBFGS_M];
        launch_lincomb64(s, N, -1.0, M.g.p, 0.0, M.g.p, dx);
        int j = lb_newest;
        for (int i = 0; i < lb_count; ++i) {  // newest to oldest
          const double* sj = S + j * N;
          const double* yj = Y + j * N;
          a[j] = M.dot(sj, dx, N) / M.dot(yj, sj, N);
          launch_lincomb64(s, N, 1.0, dx, -a[j], yj, dx);
          if (i + 1 < lb_count) j = (j + Mh - 1) % Mh;
        }
        const double* sn = S + lb_newest * N;
        const double* yn = Y + lb_newest * N;
        const double gam = M.dot(sn, yn, N) / M.dot(yn, yn, N);
        launch_lincomb64(s, N, gam, dx, 0.0, dx, dx);
        for (int i = 0; i < lb_count; ++i) {  // oldest to newest
          const double* sj = S + j * N;
          const double* yj = Y + j * N;
          const double beta = M.dot(yj, dx, N) / M.dot(yj, sj, N);
          launch_lincomb64(s, N, 1.0, dx, a[j] - beta, sj, dx);
          j = (j + 1) % Mh;
        }
      }
    } else if (!use_steepest && o.qpSolverType == PGN_QP_NEWTON) {
      M.eval_hessian();
      R.hessian_evals++;
      launch_lincomb64(s, N, 1.0, xsave, -1.0, v0, dx);  // dx0 = x - x_prev (:488)
      launch_lincomb64(s, N, -1.0, M.g.p, 0.0, M.g.p, bvec);
      lo.tol = o.lsTolMaxError;
      if (o.adaptive_tol && !ew_off && !direct) {
        if (g_prev > 0.0) {
          const double r = opt / g_prev;
          double e = o.ew_gamma * r * r;
          const double guard = o.ew_gamma * eta * eta;
          if (guard > 0.1) e = std::max(e, guard);
          eta = e;
        }
        // Kelley's safeguard (Iterative Methods for Linear and Nonlinear
        // Equations, SIAM 1995, sec. 6.3): a step whose linear residual is
        // 0.5 tolMaxError already meets the Newton tolerance, so no linear
        // solve is taken tighter than that
        eta = std::min(o.ew_max, std::max(eta, 0.5 * o.tolMaxError / opt));
        lo.tol = std::max(o.lsTolMaxError, eta);
        g_prev = opt;
      }
      inexact = lo.tol > o.lsTolMaxError;
      pgn_linear_result lr;
      linear_solve(M, bvec, dx, lo, lr);
      R.pcg_iterations += lr.iterations;
      if (trace) {
        int cok = -1;  // two-level: the coarse operator of the last try was SPD
        if (lo.precond == 2) {
          PGN_HIP(hipMemcpyAsync(&cok, M.cg_ok.p, sizeof(int), hipMemcpyDeviceToHost, s));
          M.sync();
        }
        std::fprintf(stderr,
                     "[pgn] newton %d |g| %.3e lin tol %.1e pcg %d refine %d reg %d relres "
                     "%.2e coarse %d\n",
                     itCount, opt, lo.tol, lr.iterations, lr.refinements, lr.reg_used,
                     lr.rel_residual, cok);
      }
    } else {
      inexact = false;
      launch_lincomb64(s, N, -1.0, M.g.p, 0.0, M.g.p, dx);
    }
    double len = std::sqrt(M.dot(dx, dx, N));
    if (len > o.maxStepSize) {
      launch_lincomb64(s, N, o.maxStepSize / len, dx, 0.0, dx, dx);
      len = o.maxStepSize;
    }
    // LineSearch_Simple (OptimSolver.cpp:688-736)
    double factor = 1.0, objNew = 0.0, ggNew = 0.0, stepLen = 0.0, escale = 0.0;
    int numBis = -1;
    bool guarded = false;
    do {
      ++numBis;
      stepLen = factor * len;
      launch_lincomb64(s, N, 1.0, xsave, factor, dx, x);  // restore + IncVariables
      M.eval_eg(objNew, ggNew, &escale);
      R.energy_evals++;
      g_valid = true;
      if (objNew < obj) break;
      if (o.ls_noise_guard > 0.0 && objNew <= obj + o.ls_noise_guard * escale &&
          ggNew < gg) {
        // energy change below FP64 rounding: accept on the gradient norm and
        // book it as an improvement of zero (the decrease branch below)
        guarded = true;
        break;
      }
      factor *= o.lSearch_alpha;
      d2d(x, xsave);  // PrePerformStep
      g_valid = false;
    } while (stepLen > o.tolMinStep && numBis <= o.lsearch_numTry);
    R.line_search_bisections += numBis;
    // UpdateStepResult (OptimSolver.cpp:402-466)
    double impr = obj - objNew;
    if (guarded && impr < 0) impr = 0;
    if (impr > 0 || guarded) {
      const double optNew = std::sqrt(ggNew);
      if (impr < o.tolMinImpr) {
        minImpr++;
        result = PGN_OS_MINIMPR;
      } else if (stepLen < o.tolMinStep) {
        minStep++;
        result = PGN_OS_MINSTEP;
      } else if (optNew <= o.tolMaxError) {
        result = PGN_OS_SUCCESS;
        minImpr = 0;
        minStep = 0;
      } else {
        result = PGN_OS_ONGOING;
        minImpr = 0;
        minStep = 0;
      }
      d2d(v0, xsave);
      obj = objNew;
      opt = optNew;
    } else if (inexact) {  // retry with the fixed tolerance
      ew_off = true;
      result = PGN_OS_ONGOING;
    } else {
      result = PGN_OS_NONDESC;
    }
    // OnIterFinish: UpdateKinematics (no kinematic state for tets)
  };

  // OptimSolver::SolveFull (OptimSolver.cpp:163-264)
  for (;;) {
    ++itCount;
    solve_step(steepest);
    if (!g_valid) {
      double e2 = 0.0;
      M.eval_eg(e2, gg);
      R.energy_evals++;
      g_valid = true;
    }
    opt = std::sqrt(gg);
    // IsFullyConstrained: the dynamic problem always is
    // (OptimProblem_BasicDynamic.cpp:127-129)
    if (dynamic || M.is_loaded()) {
      if (result == PGN_OS_NONDESC) {
        if (!o.trySteepest) break;
        if (steepest) break;
        steepest = true;
      }
      if (result == PGN_OS_MINIMPR) {
        if (minImpr == o.minImprMaxCount) break;
        if (o.trySteepest) steepest = !steepest;
      }
      if (result == PGN_OS_MINSTEP) {
        if (minStep == o.minStepMaxCount) break;
        if (o.trySteepest) steepest = !steepest;
      }
      if (result == PGN_OS_SUCCESS) {
        if (opt <= o.tolMaxError) break;
        result = PGN_OS_ONGOING;
      }
      if (result == PGN_OS_ONGOING) steepest = false;
    }
    if (itCount + 1 > o.numMaxIters) {
      result = PGN_OS_MAXITER;
      break;
    }
  }
  M.sync();
  R.result = result;
  R.iterations = itCount;
  R.objective = obj;
  R.optimality = opt;
  R.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---- implicit dynamics ---------------------------------------------------------
// One backward-Euler time step: OptimProblem_BasicDynamic
// (OptimProblem_BasicDynamic.cpp:20-170) minimised by SolveFull.
void dynamic_step(pgn_model& M, double dt, const pgn_newton_opts& o, pgn_newton_result& R) {
  hipStream_t s = M.stream();
  const long N = 3L * M.n;
  if (M.vel.n != (size_t)N) {
    M.vel.alloc(N);
    PGN_HIP(hipMemsetAsync(M.vel.p, 0, N * sizeof(double), s));
  }
  M.xhat.ensure(N);
  M.xstep0.ensure(N);
  M.nw_v0.ensure(N);
  if (!M.dyn_v0_valid) {  // OptimSolver::Init: m_v0 = x
    PGN_HIP(hipMemcpyAsync(M.nw_v0.p, M.x.p, N * sizeof(double), hipMemcpyDeviceToDevice, s));
    M.dyn_v0_valid = true;
  }
  // OnSolveStart (:133-146): x0, v0 kept, boundary conditions stepped, x set
  // to the predictor x0 + dt v0
  PGN_HIP(hipMemcpyAsync(M.xstep0.p, M.x.p, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  M.step_load();
  launch_lincomb64(s, N, 1.0, M.xstep0.p, dt, M.vel.p, M.xhat.p);
  PGN_HIP(hipMemcpyAsync(M.x.p, M.xhat.p, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  M.dyn_dt = dt;
  M.hess_valid = false;
  try {
    newton_solve(M, o, R, true);
  } catch (...) {
    M.dyn_dt = 0.0;
    M.hess_valid = false;
    throw;
  }
  M.dyn_dt = 0.0;
  M.hess_valid = false;
  // PosPerformStep (:159-169): v = (x - x0) / dt
  launch_lincomb64(s, N, 1.0 / dt, M.x.p, -1.0 / dt, M.xstep0.p, M.vel.p);
  M.sync();
}

}  // namespace pgn
