// batch.hpp -- a fused batch of independent quasi-static solves.
//
// Instances (tet meshes and thin shells) are staged on the host, then built
// into ONE device model over the disjoint union of their nodes and elements:
// every element kernel, gather, assembly and PCG launch covers the whole
// batch.  Each instance keeps its own Newton state (load stage, line-search
// factor, status machine of OptimSolver::SolveFull); the linear solve is one
// PCG over the block-diagonal batch Hessian, whose blocks do not couple.
#pragma once

#include <memory>

#include "model.hpp"

struct pgn_batch_inst {
  int n = 0;
  std::vector<double> X0;            // 3n
  std::vector<int> T[3];             // per class (tets, membranes, hinges), 4 per element
  std::vector<uint8_t> kind[3];
  std::vector<double> l1[3], l2[3], rho[3], extra[3];
  std::vector<uint8_t> fixed;        // 3n
  bool has_g = false;
  double g_ini[3] = {0, 0, 0}, g_end[3] = {0, 0, 0};
  int g_stages = 1;
  bool has_f = false;
  std::vector<double> f_ini, f_end;  // 3n
  int f_stages = 1;
  size_t footprint() const;          // device bytes once built (estimate)
};

struct pgn_batch {
  pgn_ctx* ctx = nullptr;
  int prec = PGN_FP64;
  std::vector<pgn_batch_inst> inst;
  std::unique_ptr<pgn_model> M;
  std::vector<long> hnoff, heoff;
  std::vector<int> g_cur, f_cur;
  pgn::DBuf<long> noff, eoff;
  pgn::DBuf<int> node_inst;
  pgn::DBuf<double> grav, falpha, fac, elemE, nodeE, nodeG2, out, n2;
  pgn::DBuf<uint8_t> mode, mask;
  pgn::DBuf<double> xsave, v0, dx, bvec;
  // segmented PCG: workgroup table per instance (planned for a lanes-per-row)
  int seg_lpr = 0, seg_W = 0;
  std::vector<uint8_t> seg_act;  // instances the current table serves
  pgn::DBuf<int> wg_seg, wg_loc, seg_G, seg_wg0;
  pgn::DBuf<long> seg_row0, seg_row1;
  pgn::DBuf<double> seg_tol, seg_part1, seg_part2, seg_partb, kvec, kvec2, xprev;
  pgn::DBuf<pgn::PcgScalars> seg_sc;
  pgn::ChunkGraphs<pgn::SegPcgArgs> seg_graphs;  // replayable segmented PCG chunks
  pgn::DBuf<short> bcol16;  // 16-bit column offsets for the segmented SpMV (empty: 32-bit)
  bool bcol16_done = false;
  std::vector<pgn::PcgScalars> hseg_sc;

  int K() const { return (int)inst.size(); }
  bool built() const { return (bool)M; }
  void build();
  pgn::BatchArgs args() const;
  void upload_loads();
  bool step_load(int k);
  bool is_loaded(int k) const;
  // per-instance {E, |g|^2, energy scale, elastic energy} at the current x;
  // with act, only the instances it marks are evaluated (their element and
  // node ranges alone, the same bits as a whole-batch evaluation) and the
  // others' entries keep their last values
  void eval_eg(std::vector<double>& out4, const std::vector<uint8_t>* act = nullptr);
  pgn::DBuf<uint8_t> eg_mask;
  // runs [k0, k1) of consecutive instances marked in act
  std::vector<std::pair<int, int>> active_runs(const std::vector<uint8_t>& act) const;
  // element Hessians, BSR3 gather and block-Jacobi inverses of the instances
  // marked in act only (their element, block and node ranges); many runs or
  // all instances: the whole-batch evaluation
  void eval_hessian(const std::vector<uint8_t>& act);
  // FP64 residual r64 = b - A x over the rows of the instances in act
  void residual64(const std::vector<uint8_t>& act, const double* x, const double* b);
  // per-instance reductions over the instances in act (others keep out[k]):
  // out[k] = x_k . y_k, or |A_k|_F^2 over the stored lower entry set
  void inst_dot(const std::vector<uint8_t>& act, const double* x, const double* y, double* out);
  void inst_frob(const std::vector<uint8_t>& act, double* out);
  bool sr_ready = false;
  pgn::SegRedTable tab_dof, tab_blk, tab_eg;
  pgn::DBuf<int> sr_inst_e, sr_wg0_e;
  pgn::DBuf<long> sr_lo_e, sr_hi_e;
  pgn::DBuf<uint8_t> sr_kind_e;
  pgn::DBuf<double> sr_part_e;
  pgn::DBuf<int> sr_inst_d, sr_wg0_d, sr_inst_b, sr_wg0_b;
  pgn::DBuf<long> sr_lo_d, sr_hi_d, sr_lo_b, sr_hi_b;
  pgn::DBuf<double> sr_part_d, sr_part_b;
  pgn::DBuf<uint8_t> rmask;
  void seg_red_prepare();
  void reset();
  void seg_plan(const std::vector<uint8_t>& act);
  pgn::SegPcgArgs seg_args(int precond) const;
  // Two-level preconditioner of the segmented PCG (precond 2): aggregates of
  // each instance's rest positions (about one per sc_nodes nodes, 2 to
  // SEG_COARSE_MAXAGG per instance; PGN_SEG_COARSE_NODES in the environment
  // sets sc_nodes), built on the first solve that asks for it; each
  // instance's coarse operator is assembled and inverted for the instances of
  // the current workgroup table before every solve try (seg_coarse_update)
  bool sc_ready = false;
  int sc_nodes = 1000;
  int sc_na = 0, sc_ncb = 0, sc_npiece = 0;
  std::vector<int> h_sc_aoff;  // K + 1
  pgn::DBuf<int> sc_agg_ptr, sc_agg_nodes, sc_agg_of, sc_agg_inst, sc_aoff;
  pgn::DBuf<long> sc_moff;
  pgn::DBuf<int> sc_cb_a, sc_cb_b, sc_pair_piece0, sc_piece_lo, sc_cb_fine, sc_ok;
  pgn::DBuf<char> sc_rel;
  pgn::DBuf<uint8_t> sc_cmask;
  pgn::DBuf<double> sc_A, sc_piece_sum, sc_rc, sc_uc, sc_cdot;
  void seg_coarse_prepare();
  void seg_coarse_update();
  // PCG of every instance at once on the PCG rhs (model precision, M->pb) and
  // solution (M->pv x); tol[k] = +inf leaves instance k untouched.  Returns
  // per-instance iteration counts.
  void seg_pcg(const std::vector<double>& tol, int maxit, int precond, bool x_zero,
               std::vector<int>& iters);
  // LinearSolver_EigenCG::SolveInternal per instance on the batch Hessian:
  // active[k] selects the instances solved; tol per instance.  With precond
  // 2, coarse (or null: all) marks the instances that keep the coarse level;
  // the others run block Jacobi (a solve where none keeps it runs precond 1).
  void linear_solve(const double* b, double* x, const std::vector<uint8_t>& active,
                    const std::vector<double>& tol, const pgn_linear_opts& o,
                    std::vector<int>& iters, const std::vector<uint8_t>* coarse = nullptr);
  pgn::DBuf<uint8_t> sc_use;  // device copy of the coarse mask of the current solve
  bool sc_use_on = false;     // the mask applies (some instances run without the level)
  void newton(const pgn_newton_opts& o, pgn_newton_result* R);
  // The same run one Newton iteration at a time: newton_begin starts it
  // (OptimSolver::Init), newton_step advances every active instance by one
  // iteration and returns the instances that finished in it, whose results
  // (newton_result) and positions are then final while the others go on.
  struct NewtonRun;
  std::unique_ptr<NewtonRun> run;
  void newton_begin(const pgn_newton_opts& o);
  std::vector<int> newton_step();
  bool newton_active() const;
  void newton_result(int k, pgn_newton_result& r) const;
  pgn_batch();
  ~pgn_batch();

  // EigenLDLT: one banded LDL^T per instance, one workgroup each (band
  // orders planned on first use)
  bool ldlt_ready = false;
  long ldlt_maxband = 0;
  pgn::DBuf<pgn::LdltProb> ldlt_prob;
  pgn::DBuf<int> ldlt_perm, ldlt_iperm, ldlt_list, ldlt_status;
  pgn::DBuf<double> ldlt_band, ldlt_D, ldlt_sinT, ldlt_y;
  void ldlt_prepare();
  pgn::LdltArgs ldlt_args() const;
  // LinearSolver_EigenLDLT::SolveInternal per active instance
  void ldlt_solve(const double* b, double* x, const std::vector<uint8_t>& active,
                  const std::vector<double>& tol, const pgn_linear_opts& o);
};
