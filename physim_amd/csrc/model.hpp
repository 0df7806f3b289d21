// model.hpp -- device-resident tet FEM model (the ISimulable of libpgn).
#pragma once

#include <cstdlib>
#include <cstring>
#include <memory>

#include "graph.hpp"
#include "kernels.hpp"
#include "pgn_internal.hpp"

namespace pgn {

// 0 = cg1, 1 = split, 2 = fused (from PGN_PCG; unset or unknown: cg1)
inline int pcg_variant() {
  const char* v = std::getenv("PGN_PCG");
  if (v && std::strcmp(v, "split") == 0) return 1;
  if (v && std::strcmp(v, "fused") == 0) return 2;
  return 0;
}

// Host-side pattern of the BSR3 Hessian and the element->block slot map.
struct BsrPattern {
  int n = 0;
  long nb = 0;
  std::vector<long> rowptr;   // n+1
  std::vector<int> bcol;      // nb
  std::vector<int> brow;      // nb
  std::vector<long> diag;     // n: index of block (i, i)
  std::vector<long> cptr;     // nb+1
  std::vector<uint32_t> code; // element*16 + 4a + b, element order per block
  std::vector<int> inc_ptr;   // n+1: node -> incident (element, local) list
  std::vector<int> inc;       // element*4 + a
};
// Builds the pattern of a tet mesh (throws on bad node ids).
void build_bsr_pattern(int n, int m, const int* T, BsrPattern& P);

// Aggregation of the two-level preconditioner (coarse.cpp): the n rest
// positions X binned into about target (at most max_aggs) boxes shaped to
// their bounding box; fills agg_of (n), rel (3n, offset from the aggregate's
// centroid) and appends one column mask per aggregate to cmask (bit c: column
// c of [translations | rotations] kept; fix3, 3n or null, marks fixed DoFs).
// Returns the number of aggregates; bins_out (or null) receives the grid.
int bin_aggregates(const double* X, const uint8_t* fix3, long n, int target, int max_aggs,
                   int* agg_of, double* rel, std::vector<uint8_t>& cmask, int bins_out[3]);
// Coarse block pairs (a <= b) reached by the lower BSR3 blocks of P, the
// blocks of each pair (cb_fine, in block order) and the pair's pieces of at
// most COARSE_PIECE blocks (pair_piece0: first piece of each pair; piece_lo:
// first entry of cb_fine of each piece, closed by the total).
struct CoarsePairs {
  std::vector<int> cb_a, cb_b, pair_piece0, piece_lo, cb_fine;
};
void coarse_pairs(const BsrPattern& P, const std::vector<int>& agg_of, int na, CoarsePairs& out);

}  // namespace pgn

struct pgn_model {
  pgn_ctx* ctx = nullptr;
  int n = 0, m = 0;
  int m_tet = 0, m_mem = 0, m_hinge = 0;  // elements [tets | membranes | hinges]
  int prec = PGN_FP64;
  bool hfp32 = false;  // Hessian / PCG precision is FP32

  pgn::BsrPattern pat;
  std::vector<double> hX0;
  std::vector<uint8_t> hfix;  // 3n

  // geometry and element parameters
  pgn::DBuf<double> X0, x;
  pgn::DBuf<int4> tets;
  pgn::DBuf<uint8_t> kind;
  pgn::DBuf<double> prm64;  // 12 x m SoA
  pgn::DBuf<float> prm32;   // 12 x m SoA
  pgn::DBuf<double> mass;   // n
  pgn::DBuf<uint8_t> fixmask;  // n
  pgn::DBuf<int> inc_ptr, inc;
  // BSR Hessian
  pgn::DBuf<long> rowptr, cptr, diag;
  pgn::DBuf<int> bcol, brow;
  pgn::DBuf<uint32_t> code;
  pgn::DBuf<char> bval, dinv, binv;  // precision of the Hessian
  // element outputs
  pgn::DBuf<double> eg64;
  pgn::DBuf<float> eg32;
  pgn::DBuf<char> eH;     // Hessian precision
  pgn::DBuf<char> eH_alt; // the other precision (timing / parity only)
  pgn::DBuf<double> blockE, nodeP, scal;  // scal: {E, gg, frob, dot, rr, ...}
  pgn::DBuf<double> g;    // 3n gradient at x
  // boundary conditions
  bool has_g = false;
  double g_ini[3] = {0, 0, 0}, g_end[3] = {0, 0, 0};
  int g_stages = 1, g_cur = 0;
  bool has_f = false;
  pgn::DBuf<double> f_ini, f_end;
  int f_stages = 1, f_cur = 0;
  // PCG work space (Hessian precision) + FP64 refinement vectors
  pgn::DBuf<char> pv;     // 6 vectors of 3n: x r z p0 p1 q
  pgn::DBuf<char> pb;     // rhs in Hessian precision
  pgn::DBuf<double> part1, part2, partb, red64;
  int pcg_lpr = 8;  // lanes per block row of the PCG SpMV (tuned per model)
  // PCG launch structure, PGN_PCG in the environment: "cg1" (default:
  // Chronopoulos-Gear, one SpMV on z + one streaming launch per iteration),
  // "split" (streaming p update, plain SpMV, update launch) or "fused" (p
  // update folded into the SpMV's column gathers, update launch)
  bool pcg_cg1 = pgn::pcg_variant() == 0;
  bool pcg_split = pgn::pcg_variant() != 2;
  // PGN_MATFREE=1 in the environment: the PCG product of a tet model's
  // static solves is the matrix-free one over the stored element Hessians
  // (textbook split form, direct launches; the assembled matrix still
  // supplies the preconditioner); mf_shift tracks the regularization shift
  // the assembled diagonal received
  bool use_matfree = false;
  double mf_shift = 0.0;
  pgn::DBuf<char> mf_ye;
  pgn::MatFreeArgs mf_args();
  pgn::DBuf<double> cgs;  // its 4 scalars
  // SELL-64 copy of the matrix for the C-G SpMV (PGN_SELL=0 in the
  // environment keeps the BSR3 SpMV); built on the first PCG solve, packed
  // from the BSR3 values before every solve try
  bool use_sell = pcg_cg1 && !(std::getenv("PGN_SELL") && std::strcmp(std::getenv("PGN_SELL"), "0") == 0);
  bool sell_ready = false;
  long sell_ns = 0;
  int sell_spw = 0, sell_g = 0;
  pgn::DBuf<long> sell_so;
  pgn::DBuf<int> sell_len, sell_col;
  pgn::DBuf<short> sell_c16;  // 16-bit column deltas, used instead of sell_col when all fit
  pgn::DBuf<char> sell_val;
  void sell_prepare();
  void sell_pack();
  bool tuned = false;
  double tune_us[3] = {0, 0, 0};  // SpMV time for 4, 8, 16 lanes per row
  pgn::DBuf<pgn::PcgScalars> sc;
  // Two-level preconditioner (precond 2): aggregates of nodes binned by rest
  // position (about coarse_target of them), built on the first solve that
  // asks for it and again after the fixed DoFs change; the coarse operator is
  // assembled and inverted before every solve try (coarse.cpp)
  bool coarse_ready = false;
  int coarse_target = 128;
  int coarse_bins[3] = {1, 1, 1};
  pgn::CoarseArgs cga;  // device pointers of the application (R set per build)
  int coarse_ncb = 0;
  int coarse_npiece = 0;
  pgn::DBuf<int> cg_agg_ptr, cg_agg_nodes, cg_agg_of;
  pgn::DBuf<int> cg_cb_a, cg_cb_b, cg_pair_piece0, cg_piece_lo, cg_cb_fine;
  pgn::DBuf<char> cg_rel;
  pgn::DBuf<uint8_t> cg_cmask;
  pgn::DBuf<double> cg_A, cg_D, cg_piece_sum, cg_rc, cg_uc, cg_cpart, cg_cdot;
  pgn::DBuf<int> cg_ok;
  std::vector<int> h_agg_of;  // host copy of the aggregation
  void coarse_prepare();
  // assembles Ac from the current BSR3 values and inverts it (no sync);
  // assemble_only leaves Ac itself in the first buffer (tests).  async: on
  // the side stream, after the work queued so far on the model's stream, so
  // that the sweeps (a few workgroups at a time) overlap the SELL pack, the
  // FP64 residual and the PCG init; coarse_join() makes the model's stream
  // wait for it before the first coarse application
  double* coarse_update(bool assemble_only = false, bool async = false);
  void coarse_join();
  hipStream_t side = nullptr;
  hipEvent_t ev_hess = nullptr, ev_coarse = nullptr;
  bool coarse_pending = false;
  pgn::DBuf<double> r64, tmp64a, tmp64b;
  pgn::DBuf<double> nw_xsave, nw_v0, nw_dx, nw_b;  // Newton work vectors
  // LBFGS (QPSolverType::LBFGS): x and g of the last LBFGS step, a ring of
  // LBFGS_M pairs (s, y), 3n each
  static constexpr int LBFGS_M = 3;
  pgn::DBuf<double> lb_xp, lb_gp, lb_S, lb_Y;
  // implicit dynamics (OptimProblem_BasicDynamic): while dyn_dt > 0 the
  // objective carries the inertia term of one backward-Euler step of length
  // dyn_dt toward the predictor xhat
  double dyn_dt = 0.0;
  pgn::DBuf<double> vel, xhat, xstep0;  // 3n: velocity, predictor, x at step start
  bool dyn_v0_valid = false;  // nw_v0 carries the solver's previous x across steps
  // EigenLDLT (FP64 banded LDL^T): band order of the nodes, built on the
  // first direct solve
  bool ldlt_ready = false;
  int ldlt_bw = 0;
  pgn::DBuf<pgn::LdltProb> ldlt_prob;
  pgn::DBuf<int> ldlt_perm, ldlt_iperm, ldlt_nprob, ldlt_list, ldlt_status;
  pgn::DBuf<double> ldlt_band, ldlt_D, ldlt_sinT, ldlt_y;
  void ldlt_prepare();
  pgn::LdltArgs ldlt_args() const;
  double* hpin = nullptr;  // pinned host scalars
  pgn::PcgScalars* hsc = nullptr;
  bool hess_valid = false;
  int pcg_chunk_max = 64;
  // PCG chunks after the first replay captured graphs (PGN_GRAPH=0: direct
  // launches; unset: models of at most GRAPH_AUTO_MAX_NODES nodes); kctl = {first iteration of the chunk, iteration cap}
  int use_graphs = pgn::graphs_level();
  pgn::DBuf<int> kctl;
  int* hkctl = nullptr;  // pinned host copy
  pgn::ChunkGraphs<pgn::PcgArgs> pcg_graphs;
  // live per-kernel timing of the PCG launches (HIP events on the stream)
  bool profiling = false;
  std::vector<hipEvent_t> prof_ev;
  long prof_n[2] = {0, 0};     // K1 (SpMV), K2 (update + precond)
  double prof_ms[2] = {0, 0};

  ~pgn_model();
  // kindh: 0 CoNH tet, 1 StVK tet, 2 StVK membrane triangle (row v0 v1 v2 v2),
  // 3 discrete-shell hinge (edge v0 v1, opposite v2 v3), ordered tets,
  // membranes, hinges.  extra: thickness (membranes) / kB (hinges), may be
  // null for tet-only models.
  void setup(int n_nodes, const double* X0h, int n_elems, const int* Th,
             const uint8_t* kindh, const double* l1, const double* l2,
             const double* rho, const double* extra, int precision);
  hipStream_t stream() const { return ctx->stream; }
  size_t tsize() const { return hfp32 ? sizeof(float) : sizeof(double); }

  // BCs
  void cur_gravity(double gv[3]) const;
  double cur_force_alpha() const;
  bool step_load();
  bool is_loaded() const;

  pgn::TetArgs tet_args(bool fp32) const;
  pgn::ShellArgs shell_args(bool fp32, int kind) const;
  // every element kernel of the model (tets, membranes, hinges) with the
  // given flags; blockE receives elem_partials() FP64 energy partials
  void launch_elements(bool fp32, int flags, double* blockE, double* elemE, void* eg,
                       void* eH);
  // the tet elements alone, in the tile tuned for (precision, flags): the
  // first launch of a flag set of the Newton path on a model of at least
  // TUNE_MIN_TETS tets times every tile (the real outputs, written again by
  // each) and keeps the fastest; PGN_TET_TILE in the environment fixes it
  void launch_tet(bool fp32, int flags, double* blockE, double* elemE, void* eg, void* eH);
  int tet_tile_choice[2][8] = {{-1, -1, -1, -1, -1, -1, -1, -1},
                               {-1, -1, -1, -1, -1, -1, -1, -1}};
  float tet_tile_us[2][8][pgn::TET_TILES] = {};  // the timings of the choice
  static constexpr int TUNE_MIN_TETS = 16384;
  int elem_partials() const;
  pgn::NodeArgs node_args() const;
  pgn::BsrArgs bsr_args() const;
  pgn::PcgArgs pcg_args(double tol, int precond) const;

  // E+g (FP64) at the current x; returns E, |g|^2 and the energy scale
  // |E_elastic| + |E_load| (synchronizes)
  void eval_eg(double& E, double& gg, double* escale = nullptr);
  // launches E+g (FP64) without reading back
  void launch_eg();
  // element Hessians (model precision) + BSR gather + preconditioners
  void eval_hessian();
  void update_precond();
  // Frobenius norm of the stored lower entry set (synchronizes)
  double frob_lower();
  double dot(const double* a, const double* b, long n);  // synchronizes
  // picks the SpMV lanes-per-row of this matrix by timing (once per model)
  void tune();
  void sync();
};
