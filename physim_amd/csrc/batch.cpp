// batch.cpp -- fused batch of independent quasi-static solves (pgn_batch_*).
//
// Build: the instances' nodes are concatenated in instance order; elements
// are concatenated per class (all tets, then all membranes, then all
// hinges, each class in instance order) so that one launch per element
// class covers the batch, and the BSR pattern of the union is the union of
// the instances' patterns.  Newton: OptimSolver::SolveFull
// (OptimSolver.cpp:163-264) with OptimSolver_USQP_LS::SolveStepInternal
// (OptimSolver_USQP_LS.cpp:18-68) per instance, all active instances
// advancing one iteration per batch iteration; per-instance energies and
// gradient norms come from segmented reductions.
#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <new>

#include "ldlt.hpp"
#include "solver.hpp"

using namespace pgn;

namespace pgn {
int api_guard_set(const char* msg, int code);
}

size_t pgn_batch_inst::footprint() const {
  // per node: x, X0, g, PCG and Newton vectors (~30 x 8 B); per element:
  // connectivity + parameters + outputs (~16 + 96 + 96 + 78 x 8 B); per
  // block (~15 per node): values + indices (9 x 8 + 24 B), and the SpMV's
  // sliced copy of a batch of one (9 x 8 + 4 B)
  long m = 0;
  for (int c = 0; c < 3; ++c) m += (long)T[c].size() / 4;
  const double ts = 8.0;
  return (size_t)(n * (30.0 * ts + 15.0 * (18 * ts + 28)) + m * (16 + 24 * ts + 78 * ts + 40));
}

void pgn_batch::build() {
  PGN_CHECK(!inst.empty(), "empty batch");
  const int Kn = K();
  hnoff.assign(Kn + 1, 0);
  for (int k = 0; k < Kn; ++k) hnoff[k + 1] = hnoff[k] + inst[k].n;
  const long n = hnoff[Kn];
  PGN_CHECK(n < (1L << 31), "batch too large for 32-bit node ids");
  heoff.assign(3 * (Kn + 1), 0);
  for (int c = 0; c < 3; ++c) {
    long base = (c == 0) ? 0 : heoff[c * (Kn + 1) - 1];
    heoff[c * (Kn + 1)] = base;
    for (int k = 0; k < Kn; ++k)
      heoff[c * (Kn + 1) + k + 1] = heoff[c * (Kn + 1) + k] + (long)inst[k].T[c].size() / 4;
  }
  const long m = heoff[3 * (Kn + 1) - 1];
  PGN_CHECK(m < (1L << 28), "batch too large for 32-bit slot codes");
  std::vector<double> X0(3 * n), l1(m), l2(m), rho(m), extra(m);
  std::vector<int> T(4 * m), ninst(n);
  std::vector<uint8_t> kind(m), fixed(3 * n);
  bool any_f = false;
  for (int k = 0; k < Kn; ++k) {
    const pgn_batch_inst& I = inst[k];
    std::copy(I.X0.begin(), I.X0.end(), X0.begin() + 3 * hnoff[k]);
    std::copy(I.fixed.begin(), I.fixed.end(), fixed.begin() + 3 * hnoff[k]);
    for (long i = 0; i < I.n; ++i) ninst[hnoff[k] + i] = k;
    any_f |= I.has_f;
    for (int c = 0; c < 3; ++c) {
      const long e0 = heoff[c * (Kn + 1) + k];
      const long mc = (long)I.T[c].size() / 4;
      for (long e = 0; e < mc; ++e) {
        for (int a = 0; a < 4; ++a) T[4 * (e0 + e) + a] = I.T[c][4 * e + a] + (int)hnoff[k];
        kind[e0 + e] = I.kind[c][e];
        l1[e0 + e] = I.l1[c][e];
        l2[e0 + e] = I.l2[c][e];
        rho[e0 + e] = I.rho[c][e];
        extra[e0 + e] = I.extra[c][e];
      }
    }
  }
  hipStream_t s = ctx->stream;
  auto model = std::make_unique<pgn_model>();
  model->ctx = ctx;
  model->setup((int)n, X0.data(), (int)m, T.data(), kind.data(), l1.data(), l2.data(),
               rho.data(), extra.data(), prec);
  model->hfix = fixed;
  std::vector<uint8_t> fm(n);
  for (long i = 0; i < n; ++i)
    fm[i] = (fixed[3 * i] ? 1 : 0) | (fixed[3 * i + 1] ? 2 : 0) | (fixed[3 * i + 2] ? 4 : 0);
  model->fixmask.upload(fm, s);
  if (any_f) {
    std::vector<double> fi(3 * n, 0.0), fe(3 * n, 0.0);
    for (int k = 0; k < Kn; ++k)
      if (inst[k].has_f) {
        std::copy(inst[k].f_ini.begin(), inst[k].f_ini.end(), fi.begin() + 3 * hnoff[k]);
        std::copy(inst[k].f_end.begin(), inst[k].f_end.end(), fe.begin() + 3 * hnoff[k]);
      }
    model->f_ini.upload(fi, s);
    model->f_end.upload(fe, s);
    model->has_f = true;
  }
  M = std::move(model);
  noff.upload(hnoff, s);
  eoff.upload(heoff, s);
  node_inst.upload(ninst, s);
  grav.alloc(3 * Kn);
  falpha.alloc(Kn);
  fac.alloc(Kn);
  out.alloc(4 * Kn);
  n2.alloc(Kn);
  mode.alloc(Kn);
  mask.alloc(Kn);
  elemE.alloc(std::max<long>(m, 1));
  nodeE.alloc(n);
  nodeG2.alloc(n);
  xsave.alloc(3 * n);
  v0.alloc(3 * n);
  dx.alloc(3 * n);
  bvec.alloc(3 * n);
  g_cur.assign(Kn, 0);
  f_cur.assign(Kn, 0);
  PGN_HIP(hipStreamSynchronize(s));
}

BatchArgs pgn_batch::args() const {
  BatchArgs a;
  a.K = K();
  a.noff = noff.p;
  a.eoff = eoff.p;
  a.node_inst = node_inst.p;
  return a;
}

bool pgn_batch::step_load(int k) {
  bool st = false;
  const pgn_batch_inst& I = inst[k];
  if (I.has_g && g_cur[k] < I.g_stages) {
    ++g_cur[k];
    st = true;
  }
  if (I.has_f && f_cur[k] < I.f_stages) {
    ++f_cur[k];
    st = true;
  }
  return st;
}

bool pgn_batch::is_loaded(int k) const {
  const pgn_batch_inst& I = inst[k];
  if (I.has_g && g_cur[k] != I.g_stages) return false;
  if (I.has_f && f_cur[k] != I.f_stages) return false;
  return true;
}

// BCondition::GetCurValues with VectorAnimator_LerpValues per instance
void pgn_batch::upload_loads() {
  const int Kn = K();
  std::vector<double> g(3 * Kn, 0.0), fa(Kn, 0.0);
  for (int k = 0; k < Kn; ++k) {
    const pgn_batch_inst& I = inst[k];
    if (I.has_g) {
      const double a = (double)g_cur[k] / I.g_stages;
      for (int i = 0; i < 3; ++i) g[3 * k + i] = I.g_ini[i] + (I.g_end[i] - I.g_ini[i]) * a;
    }
    if (I.has_f) fa[k] = (double)f_cur[k] / I.f_stages;
  }
  grav.upload(g, ctx->stream);
  falpha.upload(fa, ctx->stream);
}

std::vector<std::pair<int, int>> pgn_batch::active_runs(const std::vector<uint8_t>& act) const {
  std::vector<std::pair<int, int>> runs;
  for (int k = 0; k < K(); ++k) {
    if (!act[k]) continue;
    if (!runs.empty() && runs.back().second == k) runs.back().second = k + 1;
    else runs.emplace_back(k, k + 1);
  }
  return runs;
}

void pgn_batch::eval_hessian(const std::vector<uint8_t>& act) {
  const int Kn = K();
  const auto runs = active_runs(act);
  if (runs.size() > 16 || (runs.size() == 1 && runs[0].first == 0 && runs[0].second == Kn)) {
    M->eval_hessian();
    return;
  }
  hipStream_t s = ctx->stream;
  const bool f32 = M->hfp32;
  const size_t ts = M->tsize();
  const int tile = std::max(0, M->tet_tile_choice[f32 ? 1 : 0][EVAL_H]);
  const BsrArgs ba0 = M->bsr_args();
  for (const auto& rk : runs) {
    const int k0 = rk.first, k1 = rk.second;
    for (int c = 0; c < 3; ++c) {
      const long e0 = heoff[c * (Kn + 1) + k0], e1 = heoff[c * (Kn + 1) + k1];
      if (e1 <= e0) continue;
      if (c == 0) {
        TetArgs ta = M->tet_args(f32);
        ta.m = (int)(e1 - e0);
        ta.tets += e0;
        ta.kind += e0;
        ta.prm = reinterpret_cast<const char*>(ta.prm) + e0 * ts;
        launch_tet_eval(s, f32, EVAL_H, ta, nullptr, nullptr, nullptr,
                        M->eH.p + e0 * TET_H_PACK * ts, tile);
      } else {
        const int kind = c == 1 ? SHELL_MEMBRANE : SHELL_HINGE;
        ShellArgs sa = M->shell_args(f32, kind);
        sa.first = (int)e0;
        sa.count = (int)(e1 - e0);
        launch_shell_eval(s, f32, kind, EVAL_H, sa, nullptr, nullptr, nullptr, M->eH.p);
      }
    }
    const long n0 = hnoff[k0], n1 = hnoff[k1];
    BsrArgs ba = ba0;
    ba.s0 = M->pat.rowptr[n0];
    ba.nb = M->pat.rowptr[n1] - ba.s0;
    launch_bsr_gather(s, f32, ba, M->eH.p, M->bval.p);
    launch_precond(s, f32, (int)(n1 - n0), M->diag.p, M->bval.p, M->dinv.p, M->binv.p, n0);
  }
  M->mf_shift = 0.0;
  M->hess_valid = true;
}

void pgn_batch::seg_red_prepare() {
  if (sr_ready) return;
  const int Kn = K();
  hipStream_t s = ctx->stream;
  auto table = [&](const std::vector<long>& r, long chunk, DBuf<int>& inst, DBuf<int>& wg0,
                   DBuf<long>& lo, DBuf<long>& hi, DBuf<double>& part, SegRedTable& t) {
    std::vector<int> vi, vw(Kn + 1, 0);
    std::vector<long> vl, vh;
    for (int k = 0; k < Kn; ++k) {
      vw[k] = (int)vi.size();
      for (long a = r[k]; a < r[k + 1]; a += chunk) {
        vi.push_back(k);
        vl.push_back(a);
        vh.push_back(std::min(a + chunk, r[k + 1]));
      }
    }
    vw[Kn] = (int)vi.size();
    inst.upload(vi, s);
    wg0.upload(vw, s);
    lo.upload(vl, s);
    hi.upload(vh, s);
    part.ensure(std::max<size_t>(vi.size(), 1));
    PGN_HIP(hipStreamSynchronize(s));  // the host vectors are released on return
    t.W = (int)vi.size();
    t.inst = inst.p;
    t.lo = lo.p;
    t.hi = hi.p;
    t.wg0 = wg0.p;
    t.part = part.p;
  };
  std::vector<long> dofs(Kn + 1), blks(Kn + 1);
  for (int k = 0; k <= Kn; ++k) {
    dofs[k] = 3 * hnoff[k];
    blks[k] = M->pat.rowptr[hnoff[k]];
  }
  table(dofs, 4096, sr_inst_d, sr_wg0_d, sr_lo_d, sr_hi_d, sr_part_d, tab_dof);
  table(blks, 2048, sr_inst_b, sr_wg0_b, sr_lo_b, sr_hi_b, sr_part_b, tab_blk);
  // energy table: each instance's element chunks (tets, membranes, hinges),
  // then its node chunks
  {
    constexpr long CH = 4096;
    std::vector<int> vi, vw(Kn + 1, 0);
    std::vector<long> vl, vh;
    std::vector<uint8_t> vk;
    auto add = [&](int k, long a0, long a1, uint8_t kind) {
      for (long a = a0; a < a1; a += CH) {
        vi.push_back(k);
        vl.push_back(a);
        vh.push_back(std::min(a + CH, a1));
        vk.push_back(kind);
      }
    };
    for (int k = 0; k < Kn; ++k) {
      vw[k] = (int)vi.size();
      for (int c = 0; c < 3; ++c) add(k, heoff[c * (Kn + 1) + k], heoff[c * (Kn + 1) + k + 1], 0);
      add(k, hnoff[k], hnoff[k + 1], 1);
    }
    vw[Kn] = (int)vi.size();
    sr_inst_e.upload(vi, s);
    sr_wg0_e.upload(vw, s);
    sr_lo_e.upload(vl, s);
    sr_hi_e.upload(vh, s);
    sr_kind_e.upload(vk, s);
    sr_part_e.ensure(3 * std::max<size_t>(vi.size(), 1));
    PGN_HIP(hipStreamSynchronize(s));
    tab_eg.W = (int)vi.size();
    tab_eg.inst = sr_inst_e.p;
    tab_eg.lo = sr_lo_e.p;
    tab_eg.hi = sr_hi_e.p;
    tab_eg.wg0 = sr_wg0_e.p;
    tab_eg.part = sr_part_e.p;
    tab_eg.kind = sr_kind_e.p;
  }
  sr_ready = true;
}

void pgn_batch::residual64(const std::vector<uint8_t>& act, const double* x, const double* b) {
  hipStream_t s = ctx->stream;
  const auto runs = active_runs(act);
  if (runs.size() > 16) {
    launch_residual64(s, M->hfp32, M->n, M->rowptr.p, M->bcol.p, M->bval.p, x, b, M->r64.p,
                      M->red64.p, M->scal.p + 4);
    return;
  }
  for (const auto& rk : runs)
    launch_residual64(s, M->hfp32, (int)(hnoff[rk.second] - hnoff[rk.first]), M->rowptr.p,
                      M->bcol.p, M->bval.p, x, b, M->r64.p, M->red64.p, M->scal.p + 4,
                      hnoff[rk.first]);
}

void pgn_batch::inst_dot(const std::vector<uint8_t>& act, const double* x, const double* y,
                         double* out) {
  seg_red_prepare();
  rmask.upload(act, ctx->stream);
  launch_seg_dot(ctx->stream, K(), tab_dof, rmask.p, x, y, out);
}

void pgn_batch::inst_frob(const std::vector<uint8_t>& act, double* out) {
  seg_red_prepare();
  rmask.upload(act, ctx->stream);
  launch_seg_frob(ctx->stream, M->hfp32, K(), tab_blk, rmask.p, M->brow.p, M->bcol.p,
                  M->bval.p, out);
}

void pgn_batch::eval_eg(std::vector<double>& o4, const std::vector<uint8_t>* act) {
  hipStream_t s = ctx->stream;
  const int Kn = K();
  NodeArgs na = M->node_args();
  na.node_inst = node_inst.p;
  na.grav_inst = grav.p;
  na.falpha_inst = falpha.p;
  na.nodeE = nodeE.p;
  na.nodeG2 = nodeG2.p;
  // runs of consecutive marked instances (their elements and nodes are
  // contiguous); many runs, or all instances: one whole-batch launch each
  std::vector<std::pair<int, int>> runs;
  if (act) {
    for (int k = 0; k < Kn; ++k) {
      if (!(*act)[k]) continue;
      if (!runs.empty() && runs.back().second == k) runs.back().second = k + 1;
      else runs.emplace_back(k, k + 1);
    }
  }
  const bool whole = !act || runs.size() > 16 || (runs.size() == 1 && runs[0].first == 0 &&
                                                  runs[0].second == Kn);
  if (whole) {
    M->launch_elements(false, EVAL_E | EVAL_G, nullptr, elemE.p, M->eg64.p, nullptr);
    launch_node_gather(s, false, na, M->eg64.p, M->g.p, M->nodeP.p);
    seg_red_prepare();
    launch_seg_energy(s, Kn, tab_eg, nullptr, elemE.p, nodeE.p, nodeG2.p, out.p);
  } else {
    const int tile = std::max(0, M->tet_tile_choice[0][EVAL_E | EVAL_G]);
    for (const auto& rk : runs) {
      const int k0 = rk.first, k1 = rk.second;
      for (int c = 0; c < 3; ++c) {
        const long e0 = heoff[c * (Kn + 1) + k0], e1 = heoff[c * (Kn + 1) + k1];
        if (e1 <= e0) continue;
        if (c == 0) {
          TetArgs ta = M->tet_args(false);
          ta.m = (int)(e1 - e0);
          ta.tets += e0;
          ta.kind += e0;
          ta.prm = reinterpret_cast<const double*>(ta.prm) + e0;
          launch_tet_eval(s, false, EVAL_E | EVAL_G, ta, nullptr, elemE.p + e0,
                          M->eg64.p + 12 * e0, nullptr, tile);
        } else {
          const int kind = c == 1 ? SHELL_MEMBRANE : SHELL_HINGE;
          ShellArgs sa = M->shell_args(false, kind);
          sa.first = (int)e0;
          sa.count = (int)(e1 - e0);
          launch_shell_eval(s, false, kind, EVAL_E | EVAL_G, sa, nullptr, elemE.p, M->eg64.p,
                            nullptr);
        }
      }
      NodeArgs nr = na;
      nr.i0 = hnoff[k0];
      nr.n = (int)(hnoff[k1] - hnoff[k0]);
      launch_node_gather(s, false, nr, M->eg64.p, M->g.p, M->nodeP.p);
    }
    eg_mask.upload(*act, s);
    seg_red_prepare();
    launch_seg_energy(s, Kn, tab_eg, eg_mask.p, elemE.p, nodeE.p, nodeG2.p, out.p);
  }
  o4.resize(4 * Kn);
  out.download(o4.data(), o4.size(), s);
  PGN_HIP(hipStreamSynchronize(s));
}

void pgn_batch::reset() {
  PGN_HIP(hipMemcpyAsync(M->x.p, M->X0.p, M->X0.bytes(), hipMemcpyDeviceToDevice, ctx->stream));
  PGN_HIP(hipStreamSynchronize(ctx->stream));
  std::fill(g_cur.begin(), g_cur.end(), 0);
  std::fill(f_cur.begin(), f_cur.end(), 0);
  M->hess_valid = false;
}

// ---- segmented PCG and linear solve -------------------------------------------

// Workgroup table of the segmented PCG over the instances in `act` (the
// others get no workgroups: launches then cost only what active instances
// need, however many have already converged).
void pgn_batch::seg_plan(const std::vector<uint8_t>& act) {
  const int Kn = K();
  const int lpr = M->pcg_lpr;
  if (lpr == seg_lpr && act == seg_act) return;
  const int rpw = seg_rows_per_wg(lpr);
  std::vector<int> G(Kn), wg0(Kn), ws, wl;
  std::vector<long> r0(Kn), r1(Kn);
  int W = 0;
  for (int k = 0; k < Kn; ++k) {
    const long rows = hnoff[k + 1] - hnoff[k];
    long g = (rows + rpw - 1) / rpw;
    g = std::max(1L, std::min<long>(g, SEG_MAXG));
    if (!act[k]) g = 0;
    G[k] = (int)g;
    wg0[k] = W;
    r0[k] = hnoff[k];
    r1[k] = hnoff[k + 1];
    for (int j = 0; j < G[k]; ++j) {
      ws.push_back(k);
      wl.push_back(j);
    }
    W += G[k];
  }
  hipStream_t s = ctx->stream;
  wg_seg.upload(ws, s);
  wg_loc.upload(wl, s);
  seg_G.upload(G, s);
  seg_wg0.upload(wg0, s);
  seg_row0.upload(r0, s);
  seg_row1.upload(r1, s);
  if (seg_part1.n < (size_t)std::max(W, 1)) {
    seg_part1.alloc(std::max(W, 1));
    seg_part2.alloc(4L * std::max(W, 1));
    seg_partb.alloc(std::max(W, 1));
  }
  if (seg_sc.n < (size_t)Kn) {
    seg_sc.alloc(Kn);
    seg_tol.alloc(Kn);
  }
  hseg_sc.resize(Kn);
  if (!bcol16_done) {
    // block columns as 16-bit offsets from their rows when every block lies
    // within 32767 nodes of the diagonal (instances never couple, so the
    // widest instance band decides); PGN_SELL_C16=0 keeps 32-bit columns
    bcol16_done = true;
    const char* e = std::getenv("PGN_SELL_C16");
    bool fit = !(e && std::strcmp(e, "0") == 0);
    const auto& P = M->pat;
    std::vector<short> c16;
    if (fit) c16.resize(P.nb);
    for (int r = 0; r < P.n && fit; ++r)
      for (long q = P.rowptr[r]; q < P.rowptr[r + 1]; ++q) {
        const long d = (long)P.bcol[q] - r;
        if (d > 32767 || d < -32767) {
          fit = false;
          break;
        }
        c16[q] = (short)d;
      }
    if (fit && P.nb > 0) {
      bcol16.upload(c16, s);
      PGN_HIP(hipStreamSynchronize(s));  // c16 is released at the end of this block
    }
  }
  seg_W = W;
  seg_lpr = lpr;
  seg_act = act;
  PGN_HIP(hipStreamSynchronize(s));
}

SegPcgArgs pgn_batch::seg_args(int precond) const {
  SegPcgArgs a;
  std::memset(&a, 0, sizeof(a));  // padding too: chunk graphs compare argument bytes
  a.K = K();
  a.W = seg_W;
  a.wg_seg = wg_seg.p;
  a.wg_loc = wg_loc.p;
  a.seg_G = seg_G.p;
  a.seg_wg0 = seg_wg0.p;
  a.seg_row0 = seg_row0.p;
  a.seg_row1 = seg_row1.p;
  a.tol = seg_tol.p;
  a.rowptr = M->rowptr.p;
  a.bcol = M->bcol.p;
  a.bcol16 = bcol16.p;
  a.bval = M->bval.p;
  a.dinv = M->dinv.p;
  a.binv = M->binv.p;
  a.precond = precond >= 1 ? 1 : 0;
  a.lpr = seg_lpr;
  const size_t v = 3L * M->n * M->tsize();
  char* base = M->pv.p;
  a.x = base;
  a.r = base + v;
  a.z = base + 2 * v;
  a.p0 = base + 3 * v;
  a.p1 = base + 4 * v;
  a.q = base + 5 * v;
  a.part1 = seg_part1.p;
  a.part2 = seg_part2.p;
  a.partb = seg_partb.p;
  a.sc = seg_sc.p;
  if (precond == 2 && sc_ready) {
    a.coarse = 1;
    SegCoarseArgs& c = a.cg;
    c.NA = sc_na;
    c.agg_ptr = sc_agg_ptr.p;
    c.agg_nodes = sc_agg_nodes.p;
    c.agg_of = sc_agg_of.p;
    c.agg_inst = sc_agg_inst.p;
    c.inst_aoff = sc_aoff.p;
    c.inst_moff = sc_moff.p;
    c.rel = sc_rel.p;
    c.cmask = sc_cmask.p;
    c.fixmask = M->fixmask.p;
    c.R = sc_A.p;
    c.ok = sc_ok.p;
    c.rc = sc_rc.p;
    c.uc = sc_uc.p;
    c.cdot = sc_cdot.p;
    c.use = sc_use_on ? sc_use.p : nullptr;
  }
  return a;
}

void pgn_batch::seg_coarse_prepare() {
  if (sc_ready) return;
  if (const char* e = std::getenv("PGN_SEG_COARSE_NODES")) sc_nodes = std::max(1, std::atoi(e));
  const int Kn = K();
  const long n = M->n;
  const uint8_t* fix = M->hfix.size() == 3u * n ? M->hfix.data() : nullptr;
  std::vector<int> agg_of(n), agg_inst;
  std::vector<double> rel(3 * n);
  std::vector<uint8_t> cmask;
  std::vector<long> moff(Kn + 1, 0);
  h_sc_aoff.assign(Kn + 1, 0);
  for (int k = 0; k < Kn; ++k) {
    const long o = hnoff[k], nk = hnoff[k + 1] - o;
    int na = 0;
    if (nk > 0) {
      const int target =
          (int)std::min<long>(std::max<long>(std::lround((double)nk / sc_nodes), 2), SEG_COARSE_MAXAGG);
      na = bin_aggregates(M->hX0.data() + 3 * o, fix ? fix + 3 * o : nullptr, nk, target,
                          SEG_COARSE_MAXAGG, agg_of.data() + o, rel.data() + 3 * o, cmask,
                          nullptr);
      for (long i = o; i < o + nk; ++i) agg_of[i] += h_sc_aoff[k];
    }
    h_sc_aoff[k + 1] = h_sc_aoff[k] + na;
    agg_inst.insert(agg_inst.end(), na, k);
    moff[k + 1] = moff[k] + 36L * na * na;
  }
  const int NA = h_sc_aoff[Kn];
  std::vector<int> agg_ptr(NA + 1, 0);
  for (long i = 0; i < n; ++i) agg_ptr[agg_of[i] + 1]++;
  for (int a = 0; a < NA; ++a) agg_ptr[a + 1] += agg_ptr[a];
  std::vector<int> nodes(n), cur(agg_ptr.begin(), agg_ptr.end() - 1);
  for (long i = 0; i < n; ++i) nodes[cur[agg_of[i]]++] = (int)i;
  CoarsePairs cp;
  coarse_pairs(M->pat, agg_of, NA, cp);
  hipStream_t s = ctx->stream;
  sc_agg_ptr.upload(agg_ptr, s);
  sc_agg_nodes.upload(nodes, s);
  sc_agg_of.upload(agg_of, s);
  sc_agg_inst.upload(agg_inst, s);
  sc_aoff.upload(h_sc_aoff, s);
  sc_moff.upload(moff, s);
  sc_cmask.upload(cmask, s);
  sc_cb_a.upload(cp.cb_a, s);
  sc_cb_b.upload(cp.cb_b, s);
  sc_pair_piece0.upload(cp.pair_piece0, s);
  sc_piece_lo.upload(cp.piece_lo, s);
  sc_cb_fine.upload(cp.cb_fine, s);
  std::vector<float> r32;
  if (M->hfp32) {
    r32.assign(rel.begin(), rel.end());
    sc_rel.upload(reinterpret_cast<const char*>(r32.data()), r32.size() * sizeof(float), s);
  } else {
    sc_rel.upload(reinterpret_cast<const char*>(rel.data()), rel.size() * sizeof(double), s);
  }
  sc_A.ensure(std::max<long>(moff[Kn], 1));
  sc_ok.ensure(Kn);
  sc_npiece = (int)cp.piece_lo.size() - 1;
  sc_ncb = (int)cp.cb_a.size();
  sc_piece_sum.ensure(36L * std::max(sc_npiece, 1));
  sc_rc.ensure(6L * std::max(NA, 1));
  sc_uc.ensure(6L * std::max(NA, 1));
  sc_cdot.ensure(2L * Kn);
  PGN_HIP(hipMemsetAsync(sc_cdot.p, 0, sc_cdot.bytes(), s));
  sc_na = NA;
  PGN_HIP(hipStreamSynchronize(s));  // the host vectors above are released on return
  sc_ready = true;
}

void pgn_batch::seg_coarse_update() {
  seg_coarse_prepare();
  CoarseBuild b{};
  b.ncb = sc_ncb;
  b.npiece = sc_npiece;
  b.cb_a = sc_cb_a.p;
  b.cb_b = sc_cb_b.p;
  b.pair_piece0 = sc_pair_piece0.p;
  b.piece_lo = sc_piece_lo.p;
  b.cb_fine = sc_cb_fine.p;
  b.piece_sum = sc_piece_sum.p;
  b.brow = M->brow.p;
  b.bcol = M->bcol.p;
  b.bval = M->bval.p;
  b.ok = sc_ok.p;
  launch_seg_coarse_operator(ctx->stream, M->hfp32, seg_args(2), b, sc_A.p, sc_ok.p);
  PGN_HIP(hipGetLastError());
}

void pgn_batch::seg_pcg(const std::vector<double>& tol, int maxit, int precond, bool x_zero,
                        std::vector<int>& iters) {
  hipStream_t s = ctx->stream;
  const int Kn = K();
  const bool f32 = M->hfp32;
  seg_tol.upload(tol, s);
  const SegPcgArgs a = seg_args(precond);
  if (a.W == 0) {
    iters.assign(Kn, 0);
    return;
  }
  launch_seg_init(s, f32, a, M->pb.p, x_zero ? 1 : 0);
  if (a.coarse) launch_seg_coarse_apply(s, f32, a, -1);
  // chunks after the first replay a captured graph when they have the full
  // length (launch iterations relative to the chunk, read from kctl)
  SegPcgArgs ga = a;
  ga.kctl = M->kctl.p;
  int k = 0, chunk = 8;
  for (;;) {
    const int c = std::min(chunk, maxit - k);
    hipGraphExec_t ex = nullptr;
    if (M->use_graphs >= 2 && k > 0 && c == chunk)
      ex = seg_graphs.get(s, c, k & 1, ga, [&](hipStream_t cs) {
        for (int j = 0; j < c; ++j) {
          launch_seg_k1(cs, f32, ga, j);
          launch_seg_k2(cs, f32, ga, j);
          if (ga.coarse) launch_seg_coarse_apply(cs, f32, ga, j);
        }
        launch_seg_status(cs, ga, c, 0);
      });
    if (ex) {
      M->hkctl[0] = k;
      M->hkctl[1] = maxit;
      PGN_HIP(hipMemcpyAsync(M->kctl.p, M->hkctl, 2 * sizeof(int), hipMemcpyHostToDevice, s));
      PGN_HIP(hipGraphLaunch(ex, s));
      k += c;
    } else {
      for (int j = 0; j < c; ++j) {
        launch_seg_k1(s, f32, a, k + j);
        launch_seg_k2(s, f32, a, k + j);
        if (a.coarse) launch_seg_coarse_apply(s, f32, a, k + j);
      }
      k += c;
      launch_seg_status(s, a, k, maxit);
    }
    seg_sc.download(hseg_sc.data(), Kn, s);
    PGN_HIP(hipStreamSynchronize(s));
    bool all = true;
    for (int q = 0; q < Kn; ++q) all &= (!seg_act[q] || hseg_sc[q].done != 0);
    if (all || k >= maxit) break;
    chunk = std::min(2 * chunk, 64);
  }
  iters.resize(Kn);
  for (int q = 0; q < Kn; ++q) iters[q] = seg_act[q] ? hseg_sc[q].iters : 0;
}

void pgn_batch::linear_solve(const double* b, double* x, const std::vector<uint8_t>& active,
                             const std::vector<double>& tol, const pgn_linear_opts& oin,
                             std::vector<int>& iters, const std::vector<uint8_t>* coarse) {
  hipStream_t s = ctx->stream;
  const int Kn = K();
  // instances that gave the coarse level up run block Jacobi; a solve where
  // none of the active ones keeps it runs precond 1 outright
  pgn_linear_opts o = oin;
  sc_use_on = false;
  if (o.precond == 2 && coarse) {
    bool any = false, all = true;
    for (int k = 0; k < Kn; ++k)
      if (active[k]) {
        any |= (*coarse)[k] != 0;
        all &= (*coarse)[k] != 0;
      }
    if (!any) {
      o.precond = 1;
    } else if (!all) {
      sc_use.upload(*coarse, s);
      sc_use_on = true;
    }
  }
  const long n = M->n, N = 3 * n;
  const BatchArgs ba = args();
  const double inf = std::numeric_limits<double>::infinity();
  M->tune();
  if (xprev.n < (size_t)N) xprev.alloc(N);
  if (kvec.n < (size_t)Kn) kvec.alloc(Kn);
  if (kvec2.n < (size_t)Kn) kvec2.alloc(Kn);
  PGN_CHECK(o.precond >= 0 && o.precond <= 2, "precond: 0 (none), 1 (block Jacobi) or 2 (two-level)");
  if (o.solver == PGN_LS_EIGENLDLT) {
    iters.assign(Kn, 0);
    ldlt_solve(b, x, active, tol, o);
    return;
  }
  if (Kn == 1) {
    // one instance: the single-model PCG (SELL-64 SpMV, captured chunks)
    // solves the same system with the same regularization tries
    iters.assign(1, 0);
    if (!active[0]) return;
    pgn_linear_opts o1 = o;
    o1.tol = tol[0];
    pgn_linear_result r1{};
    pgn::linear_solve(*M, b, x, o1, r1);
    iters[0] = r1.iterations;
    return;
  }
  PGN_HIP(hipMemcpyAsync(xprev.p, x, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  iters.assign(Kn, 0);
  std::vector<uint8_t> F = active;
  std::vector<double> regT(Kn, -1.0), h(Kn), shift(Kn);
  bool reg_known = false;
  auto download_k = [&](const double* d, std::vector<double>& out) {
    PGN_HIP(hipMemcpyAsync(out.data(), d, Kn * sizeof(double), hipMemcpyDeviceToHost, s));
    PGN_HIP(hipStreamSynchronize(s));
  };
  const int regs = std::max(o.reg_iters, 1);
  for (int it = 0; it < regs; ++it) {
    bool any = false;
    for (int k = 0; k < Kn; ++k) any |= F[k] != 0;
    if (!any) break;
    if (it != 0) {
      // LinearSolver_EigenCG.cpp:73 + LinearSolver.cpp:57-66 per instance:
      // A += 1e-6 |A_lower|_F 10^(it-1) I, cumulative over the tries
      if (!reg_known) {  // of the unshifted matrix, for the instances still trying
        inst_frob(F, kvec.p);
        download_k(kvec.p, h);
        for (int k = 0; k < Kn; ++k)
          if (F[k]) regT[k] = 1e-6 * std::sqrt(h[k]);
        reg_known = true;
      }
      for (int k = 0; k < Kn; ++k) shift[k] = F[k] ? regT[k] * std::pow(10.0, it - 1) : 0.0;
      kvec2.upload(shift.data(), Kn, s);
      launch_batch_add_diag(s, M->hfp32, ba, n, M->diag.p, M->bval.p, kvec2.p);
      M->update_precond();
    }
    mask.upload(F, s);
    launch_batch_copy(s, ba, n, mask.p, xprev.p, x);
    seg_plan(F);
    if (o.precond == 2) seg_coarse_update();  // the instances of F, this try's matrix
    // ---- one unregularized solve of the instances in F ----
    if (!M->hfp32) {
      std::vector<double> tv(Kn);
      for (int k = 0; k < Kn; ++k) tv[k] = F[k] ? tol[k] : inf;
      PGN_HIP(hipMemcpyAsync(M->pb.p, b, N * sizeof(double), hipMemcpyDeviceToDevice, s));
      PGN_HIP(hipMemcpyAsync(M->pv.p, x, N * sizeof(double), hipMemcpyDeviceToDevice, s));
      std::vector<int> itk;
      seg_pcg(tv, o.max_iters, o.precond, false, itk);
      PGN_HIP(hipMemcpyAsync(x, M->pv.p, N * sizeof(double), hipMemcpyDeviceToDevice, s));
      for (int k = 0; k < Kn; ++k)
        if (F[k]) iters[k] += itk[k];
    } else {
      // FP32 PCG corrections inside FP64 iterative refinement, per instance
      inst_dot(F, b, b, kvec.p);
      std::vector<double> bb(Kn), rr(Kn);
      download_k(kvec.p, bb);
      std::vector<uint8_t> act(Kn, 0), zero(Kn, 0);
      bool anyz = false;
      for (int k = 0; k < Kn; ++k) {
        if (!F[k]) continue;
        if (bb[k] == 0.0) {
          zero[k] = 1;  // b = 0: x = 0
          anyz = true;
        } else {
          act[k] = 1;
        }
      }
      if (anyz) {
        mask.upload(zero, s);
        launch_batch_copy(s, ba, n, mask.p, b, x);
      }
      std::vector<int> nref(Kn, 0);
      for (;;) {
        residual64(act, x, b);
        inst_dot(act, M->r64.p, M->r64.p, kvec.p);
        download_k(kvec.p, rr);
        bool any_act = false;
        int min_it = o.max_iters;
        std::vector<double> inner(Kn, inf);
        for (int k = 0; k < Kn; ++k) {
          if (!act[k]) continue;
          const double thr = tol[k] * tol[k] * bb[k];
          // also stops on NaN
          if (!(rr[k] >= thr) || iters[k] >= o.max_iters || nref[k] >= MAX_REFINEMENTS) {
            act[k] = 0;
            continue;
          }
          ++nref[k];
          inner[k] = std::max(o.inner_tol, tol[k] * std::sqrt(bb[k] / rr[k]));
          min_it = std::min(min_it, iters[k]);
          any_act = true;
        }
        if (!any_act) break;
        launch_cast_f64_to_f32(s, N, M->r64.p, reinterpret_cast<float*>(M->pb.p));
        std::vector<int> itk;
        seg_plan(act);
        seg_pcg(inner, std::max(1, o.max_iters - min_it), o.precond, true, itk);
        mask.upload(act, s);
        launch_batch_axpy_f32(s, ba, n, mask.p, reinterpret_cast<const float*>(M->pv.p), x);
        for (int k = 0; k < Kn; ++k) {
          if (!act[k]) continue;
          iters[k] += itk[k];
          // the residual rounded to zero in FP32: no FP32 correction can
          // follow (a count of 0 alone can still be one iteration taken)
          if (hseg_sc[k].bb == 0.0) act[k] = 0;
        }
      }
    }
    // indefinite (b.x < 0): try again with a larger shift
    inst_dot(F, b, x, kvec.p);
    download_k(kvec.p, h);
    for (int k = 0; k < Kn; ++k)
      if (F[k] && !(h[k] < 0.0)) F[k] = 0;
  }
}

// One SolveFull run over the batch (OptimSolver.cpp:163-264 per instance),
// resumable one Newton iteration at a time so that a caller can hand back
// the instances that finish while the others go on.
struct pgn_batch::NewtonRun {
  pgn_newton_opts o{};
  bool direct = false;
  pgn_linear_opts lo{};
  std::chrono::steady_clock::time_point t0;
  std::vector<int> st, result, minImpr, minStep, steep, its, bis, evals, hevals, pcg, nbis,
      searching, guarded;
  std::vector<uint8_t> hmode, hmask;
  std::vector<double> obj, opt, optStart, objNew, optNew, f, len, clamp, stepLen, hfac, eta,
      gprev, hn2, secs;
  // per instance: fixed lsTolMaxError after an inexact step failed its line
  // search (the retry rule of the single-model solver); inexact: this step's
  // linear solve was looser than lsTolMaxError
  std::vector<uint8_t> ew_off, inexact;
  // two-level: instances still keeping the coarse level (pgn_newton_opts::
  // coarse_giveup; PGN_SEG_COARSE_GIVEUP in the environment overrides it)
  std::vector<uint8_t> coarse_on;
  int coarse_giveup = 0;
  std::vector<double> E;
};

pgn_batch::pgn_batch() = default;
pgn_batch::~pgn_batch() = default;

void pgn_batch::newton_begin(const pgn_newton_opts& o) {
  PGN_CHECK(o.qpSolverType == PGN_QP_NEWTON, "fused batches take Newton steps (PGN_QP_NEWTON)");
  PGN_CHECK(o.lsSolverType == PGN_LS_EIGENCG || o.lsSolverType == PGN_LS_EIGENLDLT,
            "lsSolverType: PGN_LS_EIGENCG or PGN_LS_EIGENLDLT");
  const bool direct = o.lsSolverType == PGN_LS_EIGENLDLT;
  if (direct) ldlt_prepare();  // an instance beyond the band limit fails before any step
  run = std::make_unique<NewtonRun>();
  NewtonRun& r = *run;
  r.o = o;
  r.direct = direct;
  r.t0 = std::chrono::steady_clock::now();
  hipStream_t s = ctx->stream;
  const int Kn = K();
  const long N = 3L * M->n;
  pgn_linear_default_opts(&r.lo);
  r.lo.tol = o.lsTolMaxError;
  r.lo.max_iters = o.lsNumMaxIters;
  r.lo.precond = o.precond;
  r.lo.inner_tol = o.inner_tol;
  r.lo.solver = o.lsSolverType;
  for (auto* v : {&r.st, &r.minImpr, &r.minStep, &r.steep, &r.its, &r.bis, &r.evals, &r.hevals,
                  &r.pcg, &r.nbis, &r.searching, &r.guarded})
    v->assign(Kn, 0);
  r.result.assign(Kn, PGN_OS_SUCCESS);
  r.hmode.assign(Kn, 0);
  r.hmask.assign(Kn, 0);
  for (auto* v : {&r.obj, &r.opt, &r.optStart, &r.objNew, &r.optNew, &r.f, &r.len, &r.clamp,
                  &r.stepLen, &r.hfac, &r.hn2, &r.secs})
    v->assign(Kn, 0.0);
  r.eta.assign(Kn, o.ew_max);
  r.gprev.assign(Kn, -1.0);
  r.ew_off.assign(Kn, 0);
  r.coarse_on.assign(Kn, 1);
  r.coarse_giveup = o.coarse_giveup;
  if (const char* e = std::getenv("PGN_SEG_COARSE_GIVEUP")) r.coarse_giveup = std::atoi(e);
  r.inexact.assign(Kn, 0);
  // OptimSolver::Init (OptimSolver.cpp:36-55)
  PGN_HIP(hipMemcpyAsync(v0.p, M->x.p, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  upload_loads();
  eval_eg(r.E);
  for (int k = 0; k < Kn; ++k) {
    r.obj[k] = r.E[4 * k];
    r.opt[k] = std::sqrt(r.E[4 * k + 1]);
    r.evals[k]++;
  }
}

bool pgn_batch::newton_active() const {
  if (!run) return false;
  for (int v : run->st)
    if (v == 0) return true;
  return false;
}

void pgn_batch::newton_result(int k, pgn_newton_result& R) const {
  PGN_CHECK(run && k >= 0 && k < K(), "no Newton run / instance out of range");
  const NewtonRun& r = *run;
  R = pgn_newton_result{};
  R.result = r.result[k];
  R.iterations = r.its[k];
  R.pcg_iterations = r.pcg[k];
  R.line_search_bisections = r.bis[k];
  R.energy_evals = r.evals[k];
  R.hessian_evals = r.hevals[k];
  R.objective = r.obj[k];
  R.optimality = r.opt[k];
  R.seconds = r.secs[k];
}

std::vector<int> pgn_batch::newton_step() {
  std::vector<int> finished;
  PGN_CHECK(run, "newton_step before newton_begin");
  NewtonRun& r = *run;
  const pgn_newton_opts& o = r.o;
  const bool direct = r.direct;
  const pgn_linear_opts& lo = r.lo;
  hipStream_t s = ctx->stream;
  const int Kn = K();
  const long n = M->n, N = 3 * n;
  const BatchArgs ba = args();
  enum { ACTIVE = 0, DONE = 1 };
  auto& st = r.st;
  auto& result = r.result;
  auto& minImpr = r.minImpr;
  auto& minStep = r.minStep;
  auto& steep = r.steep;
  auto& its = r.its;
  auto& bis = r.bis;
  auto& evals = r.evals;
  auto& hevals = r.hevals;
  auto& pcg = r.pcg;
  auto& nbis = r.nbis;
  auto& searching = r.searching;
  auto& guarded = r.guarded;
  auto& hmode = r.hmode;
  auto& hmask = r.hmask;
  auto& obj = r.obj;
  auto& opt = r.opt;
  auto& optStart = r.optStart;
  auto& objNew = r.objNew;
  auto& optNew = r.optNew;
  auto& f = r.f;
  auto& len = r.len;
  auto& clamp = r.clamp;
  auto& stepLen = r.stepLen;
  auto& hfac = r.hfac;
  auto& eta = r.eta;
  auto& gprev = r.gprev;
  auto& hn2 = r.hn2;
  auto& ew_off = r.ew_off;
  auto& inexact = r.inexact;
  auto& coarse_on = r.coarse_on;
  auto& E = r.E;
  auto d2d = [&](double* dst, const double* src) {
    PGN_HIP(hipMemcpyAsync(dst, src, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  };
  bool any = false;
  for (int k = 0; k < Kn; ++k) any |= (st[k] == ACTIVE);
  if (!any) return finished;
  // OnIterStart: StepBoundaryConditions of every active instance
  for (int k = 0; k < Kn; ++k)
    if (st[k] == ACTIVE) step_load(k);
  upload_loads();
  d2d(xsave.p, M->x.p);
  std::vector<uint8_t> live(Kn);
  for (int k = 0; k < Kn; ++k) live[k] = st[k] == ACTIVE;
  eval_eg(E, &live);
  bool need_h = false;
  std::vector<double> tol_lin(Kn, o.lsTolMaxError);
  for (int k = 0; k < Kn; ++k) {
    hmode[k] = 0;
    if (st[k] != ACTIVE) continue;
    its[k]++;
    evals[k]++;
    obj[k] = E[4 * k];
    opt[k] = optStart[k] = std::sqrt(E[4 * k + 1]);
    if (opt[k] < o.tolMaxError) {
      result[k] = PGN_OS_SUCCESS;
      continue;
    }
    hmode[k] = steep[k] ? 2 : 1;
    inexact[k] = 0;
    if (hmode[k] == 1) {
      need_h = true;
      hevals[k]++;
      double lt = o.lsTolMaxError;
      if (o.adaptive_tol && !ew_off[k] && !direct) {  // Eisenstat-Walker choice 2
        if (gprev[k] > 0.0) {
          const double r = opt[k] / gprev[k];
          double e = o.ew_gamma * r * r;
          const double guard = o.ew_gamma * eta[k] * eta[k];
          if (guard > 0.1) e = std::max(e, guard);
          eta[k] = e;
        }
        // Kelley's safeguard: no tighter than the absolute tolerance needs
        eta[k] = std::min(o.ew_max, std::max(eta[k], 0.5 * o.tolMaxError / opt[k]));
        gprev[k] = opt[k];
        lt = std::max(o.lsTolMaxError, eta[k]);
      }
      tol_lin[k] = lt;
      inexact[k] = lt > o.lsTolMaxError;
    }
  }
  mode.upload(hmode, s);
  std::vector<uint8_t> hact(Kn), stepping(Kn);
  for (int k = 0; k < Kn; ++k) {
    hact[k] = hmode[k] == 1;
    stepping[k] = hmode[k] != 0;
  }
  if (need_h) eval_hessian(hact);
  launch_batch_rhs(s, ba, n, mode.p, M->g.p, xsave.p, v0.p, bvec.p, dx.p);
  if (need_h) {
    std::vector<int> its_lin;
    linear_solve(bvec.p, dx.p, hact, tol_lin, lo, its_lin, &coarse_on);
    for (int k = 0; k < Kn; ++k)
      if (hmode[k] == 1) {
        pcg[k] += its_lin[k];
        if (r.coarse_giveup > 0 && its_lin[k] > r.coarse_giveup) coarse_on[k] = 0;
      }
  }
  inst_dot(stepping, dx.p, dx.p, n2.p);
  n2.download(hn2.data(), Kn, s);
  PGN_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < Kn; ++k) {
    searching[k] = 0;
    if (!hmode[k]) continue;
    len[k] = std::sqrt(hn2[k]);
    clamp[k] = 1.0;
    if (len[k] > o.maxStepSize) {  // OptimSolver.cpp:346
      clamp[k] = o.maxStepSize / len[k];
      len[k] = o.maxStepSize;
    }
    f[k] = 1.0;
    nbis[k] = -1;
    searching[k] = 1;
    guarded[k] = 0;
  }
  // LineSearch_Simple (OptimSolver.cpp:688-736), instances in lockstep
  for (;;) {
    bool anys = false;
    for (int k = 0; k < Kn; ++k) {
      hfac[k] = -1.0;
      if (searching[k]) {
        nbis[k]++;
        stepLen[k] = f[k] * len[k];
        hfac[k] = f[k] * clamp[k];
        anys = true;
      }
    }
    if (!anys) break;
    fac.upload(hfac, s);
    launch_batch_step(s, ba, n, fac.p, xsave.p, dx.p, M->x.p);
    std::vector<uint8_t> srch(searching.begin(), searching.end());
    eval_eg(E, &srch);
    bool restore = false;
    for (int k = 0; k < Kn; ++k) {
      hfac[k] = -1.0;
      if (!searching[k]) continue;
      evals[k]++;
      const double en = E[4 * k], gn = std::sqrt(E[4 * k + 1]);
      objNew[k] = en;
      optNew[k] = gn;
      if (en < obj[k]) {
        searching[k] = 0;
        continue;
      }
      if (o.ls_noise_guard > 0.0 && en <= obj[k] + o.ls_noise_guard * E[4 * k + 2] &&
          gn < opt[k]) {
        guarded[k] = 1;  // booked as an improvement of zero
        searching[k] = 0;
        continue;
      }
      f[k] *= o.lSearch_alpha;
      if (!(stepLen[k] > o.tolMinStep && nbis[k] <= o.lsearch_numTry)) {
        searching[k] = 0;  // exhausted: back to the step start
        hfac[k] = 0.0;
        optNew[k] = optStart[k];
        restore = true;
      }
    }
    if (restore) {
      fac.upload(hfac, s);
      launch_batch_step(s, ba, n, fac.p, xsave.p, dx.p, M->x.p);
    }
  }
  // UpdateStepResult (OptimSolver.cpp:402-466)
  bool any_impr = false;
  for (int k = 0; k < Kn; ++k) {
    hmask[k] = 0;
    if (!hmode[k]) continue;
    bis[k] += nbis[k];
    double impr = obj[k] - objNew[k];
    if (guarded[k] && impr < 0) impr = 0;
    if (impr > 0 || guarded[k]) {
      if (impr < o.tolMinImpr) {
        minImpr[k]++;
        result[k] = PGN_OS_MINIMPR;
      } else if (stepLen[k] < o.tolMinStep) {
        minStep[k]++;
        result[k] = PGN_OS_MINSTEP;
      } else if (optNew[k] <= o.tolMaxError) {
        result[k] = PGN_OS_SUCCESS;
        minImpr[k] = minStep[k] = 0;
      } else {
        result[k] = PGN_OS_ONGOING;
        minImpr[k] = minStep[k] = 0;
      }
      obj[k] = objNew[k];
      opt[k] = optNew[k];
      hmask[k] = 1;  // v0 := the step start (OptimSolver.cpp:488 warm start)
      any_impr = true;
    } else if (inexact[k]) {  // retry with the fixed tolerance
      ew_off[k] = 1;
      result[k] = PGN_OS_ONGOING;
      opt[k] = optStart[k];
    } else {
      result[k] = PGN_OS_NONDESC;
      opt[k] = optStart[k];
    }
  }
  if (any_impr) {
    mask.upload(hmask, s);
    launch_batch_copy(s, ba, n, mask.p, xsave.p, v0.p);
  }
  // SolveFull status machine per instance
  for (int k = 0; k < Kn; ++k) {
    if (st[k] != ACTIVE) continue;
    bool stop = false;
    if (is_loaded(k)) {
      int& res = result[k];
      if (res == PGN_OS_NONDESC) {
        if (!o.trySteepest || steep[k]) stop = true;
        else steep[k] = 1;
      }
      if (!stop && res == PGN_OS_MINIMPR) {
        if (minImpr[k] == o.minImprMaxCount) stop = true;
        else if (o.trySteepest) steep[k] = !steep[k];
      }
      if (!stop && res == PGN_OS_MINSTEP) {
        if (minStep[k] == o.minStepMaxCount) stop = true;
        else if (o.trySteepest) steep[k] = !steep[k];
      }
      if (!stop && res == PGN_OS_SUCCESS) {
        if (opt[k] <= o.tolMaxError) stop = true;
        else res = PGN_OS_ONGOING;
      }
      if (!stop && res == PGN_OS_ONGOING) steep[k] = 0;
    }
    if (!stop && its[k] + 1 > o.numMaxIters) {
      result[k] = PGN_OS_MAXITER;
      stop = true;
    }
    if (stop) {
      st[k] = DONE;
      r.secs[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - r.t0).count();
      finished.push_back(k);
    }
  }
  return finished;
}

void pgn_batch::newton(const pgn_newton_opts& o, pgn_newton_result* R) {
  newton_begin(o);
  while (newton_active()) newton_step();
  PGN_HIP(hipStreamSynchronize(ctx->stream));
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - run->t0).count();
  for (int k = 0; k < K(); ++k) {
    newton_result(k, R[k]);
    R[k].seconds = secs;
  }
}

// ---- EigenLDLT per instance ------------------------------------------------------

void pgn_batch::ldlt_prepare() {
  if (ldlt_ready) return;
  const int Kn = K();
  const int n = M->n;
  std::vector<int> perm(n), iperm(n);
  std::vector<LdltProb> P(Kn);
  long band0 = 0;
  ldlt_maxband = 0;
  for (int k = 0; k < Kn; ++k) {
    const int n0 = (int)hnoff[k], n1 = (int)hnoff[k + 1];
    const int bw = ldlt_plan(n0, n1, M->pat.rowptr, M->pat.bcol, M->hX0.data(), perm, iperm);
    PGN_CHECK(bw + 1 <= LDLT_MAXW,
              "EigenLDLT: instance " + std::to_string(k) + " has a half-bandwidth of " +
                  std::to_string(bw) + " DoFs after reordering; the direct solver holds at most " +
                  std::to_string(LDLT_MAXW - 1) + " (use EigenCG)");
    P[k].band0 = band0;
    P[k].dof0 = 3 * n0;
    P[k].N = 3 * (n1 - n0);
    P[k].bw = bw;
    P[k].node0 = n0;
    const long len = (long)P[k].N * (bw + 1);
    band0 += len;
    ldlt_maxband = std::max(ldlt_maxband, len);
  }
  hipStream_t s = ctx->stream;
  ldlt_prob.upload(P, s);
  ldlt_perm.upload(perm, s);
  ldlt_iperm.upload(iperm, s);
  ldlt_list.alloc(Kn);
  ldlt_status.alloc(2L * Kn);
  ldlt_band.alloc(std::max<long>(band0, 1));
  ldlt_D.alloc(3L * n);
  ldlt_y.alloc(3L * n);
  ldlt_sinT.alloc(Kn);
  PGN_HIP(hipStreamSynchronize(s));
  ldlt_ready = true;
}

LdltArgs pgn_batch::ldlt_args() const {
  LdltArgs a;
  a.prob = ldlt_prob.p;
  a.perm = ldlt_perm.p;
  a.iperm = ldlt_iperm.p;
  a.nprob = node_inst.p;
  a.band = ldlt_band.p;
  a.D = ldlt_D.p;
  a.sinT = ldlt_sinT.p;
  a.status = ldlt_status.p;
  return a;
}

// The tries of ldlt_linear_solve (LinearSolver_EigenLDLT.cpp:49-153) for
// every active instance at once: per-instance normA (its stored lower
// entries), shifts, pivot checks, indefiniteness and exactness checks; the
// factorizations of one try run side by side, one workgroup per instance.
void pgn_batch::ldlt_solve(const double* b, double* x, const std::vector<uint8_t>& active,
                           const std::vector<double>& tol, const pgn_linear_opts& o) {
  ldlt_prepare();
  hipStream_t s = ctx->stream;
  const int Kn = K();
  const long n = M->n;
  const BatchArgs ba = args();
  const LdltArgs la = ldlt_args();
  if (kvec.n < (size_t)Kn) kvec.alloc(Kn);
  if (kvec2.n < (size_t)Kn) kvec2.alloc(Kn);
  auto download_k = [&](const double* d, std::vector<double>& out) {
    out.resize(Kn);
    PGN_HIP(hipMemcpyAsync(out.data(), d, Kn * sizeof(double), hipMemcpyDeviceToHost, s));
    PGN_HIP(hipStreamSynchronize(s));
  };
  std::vector<double> h, bb, xx, bx, rr, sinT(Kn), regT(Kn), shift(Kn);
  launch_batch_frob(s, M->hfp32, ba, M->rowptr.p, M->brow.p, M->bcol.p, M->bval.p, kvec.p);
  download_k(kvec.p, h);
  for (int k = 0; k < Kn; ++k) {
    const double nA = std::sqrt(h[k]);
    sinT[k] = std::max(std::min(1e-6 * nA, 1e-6), 1e-9);
    regT[k] = 1e-6 * nA;
  }
  ldlt_sinT.upload(sinT.data(), Kn, s);
  launch_batch_dot(s, ba, b, b, kvec.p);
  download_k(kvec.p, bb);
  std::vector<uint8_t> F = active;
  std::vector<int> list, st(2 * Kn);
  const int regs = std::max(o.reg_iters, 1);
  for (int it = 0; it < regs; ++it) {
    list.clear();
    for (int k = 0; k < Kn; ++k)
      if (F[k]) list.push_back(k);
    if (list.empty()) break;
    if (it != 0) {
      for (int k = 0; k < Kn; ++k) shift[k] = F[k] ? regT[k] * std::pow(10.0, it - 1) : 0.0;
      kvec2.upload(shift.data(), Kn, s);
      launch_batch_add_diag(s, M->hfp32, ba, n, M->diag.p, M->bval.p, kvec2.p);
    }
    PGN_HIP(hipMemcpyAsync(ldlt_list.p, list.data(), list.size() * sizeof(int),
                           hipMemcpyHostToDevice, s));
    mask.upload(F, s);
    launch_ldlt_gather(s, M->hfp32, la, ldlt_list.p, (int)list.size(), ldlt_maxband, M->pat.nb,
                       M->brow.p, M->bcol.p, M->bval.p, mask.p);
    launch_ldlt_factor(s, la, ldlt_list.p, (int)list.size());
    ldlt_status.download(st.data(), 2L * Kn, s);
    PGN_HIP(hipStreamSynchronize(s));
    std::vector<int> ok;
    for (int k : list)
      if (st[2 * k] < 0 && st[2 * k + 1] == 0) ok.push_back(k);  // else: zero / singular pivot
    if (ok.empty()) continue;
    PGN_HIP(hipMemcpyAsync(ldlt_list.p, ok.data(), ok.size() * sizeof(int),
                           hipMemcpyHostToDevice, s));
    launch_ldlt_solve(s, la, ldlt_list.p, (int)ok.size(), b, x, ldlt_y.p);
    launch_batch_dot(s, ba, x, x, kvec.p);
    download_k(kvec.p, xx);
    launch_batch_dot(s, ba, b, x, kvec.p);
    download_k(kvec.p, bx);
    launch_residual64(s, M->hfp32, (int)n, M->rowptr.p, M->bcol.p, M->bval.p, x, b, M->r64.p,
                      M->red64.p, M->scal.p + 4);
    launch_batch_norm2(s, ba, M->r64.p, kvec.p);
    download_k(kvec.p, rr);
    for (int k : ok) {
      const double nb = std::sqrt(bb[k]), nx = std::sqrt(xx[k]);
      if (nb > 0.0 && nx > 0.0 && bx[k] / (nb * nx) < -regT[k]) continue;  // indefinite
      const double ae = std::sqrt(rr[k]);
      const double re = nb > 0.0 ? ae / nb : std::numeric_limits<double>::quiet_NaN();
      if (ae > tol[k] && re > tol[k]) continue;  // inexact
      F[k] = 0;
    }
  }
}

// ---- C ABI ---------------------------------------------------------------------

namespace {
template <class F>
int bguard(F&& f) {
  try {
    f();
    return PGN_OK;
  } catch (const Error& e) {
    return api_guard_set(e.what(), e.code);
  } catch (const std::bad_alloc&) {
    return api_guard_set("host allocation failed", PGN_ERR_NOMEM);
  } catch (const std::exception& e) {
    return api_guard_set(e.what(), PGN_ERR_INVALID);
  }
}

pgn_batch_inst& staged(pgn_batch* B, int k) {
  PGN_CHECK(B, "null batch");
  PGN_CHECK(!B->built(), "batch already built");
  PGN_CHECK(k >= 0 && k < B->K(), "instance index out of range");
  return B->inst[k];
}

void check_ids(const int* ids, long count, int n) {
  for (long q = 0; q < count; ++q) PGN_CHECK(ids[q] >= 0 && ids[q] < n, "node id out of range");
}
}  // namespace

extern "C" {

int pgn_batch_create(pgn_ctx* ctx, int precision, pgn_batch** out) {
  return bguard([&] {
    PGN_CHECK(ctx && out, "null argument");
    PGN_CHECK(precision == PGN_FP64 || precision == PGN_MIXED, "unknown precision");
    auto* B = new pgn_batch();
    B->ctx = ctx;
    B->prec = precision;
    *out = B;
  });
}

void pgn_batch_destroy(pgn_batch* B) {
  if (!B) return;
  (void)hipSetDevice(B->ctx->device);
  (void)hipStreamSynchronize(B->ctx->stream);
  delete B;
}

int pgn_batch_add_tet(pgn_batch* B, int n_nodes, const double* X0, int n_tets, const int* T,
                      const unsigned char* material, const double* lame1, const double* lame2,
                      const double* density, int* instance) {
  return bguard([&] {
    PGN_CHECK(B && X0 && T && material && lame1 && lame2 && density, "null argument");
    PGN_CHECK(!B->built(), "batch already built");
    PGN_CHECK(n_nodes > 0 && n_tets > 0, "empty instance");
    check_ids(T, 4L * n_tets, n_nodes);
    pgn_batch_inst I;
    I.n = n_nodes;
    I.X0.assign(X0, X0 + 3L * n_nodes);
    I.T[0].assign(T, T + 4L * n_tets);
    for (int e = 0; e < n_tets; ++e)
      PGN_CHECK(material[e] <= 1, "unknown material kind at element " + std::to_string(e));
    I.kind[0].assign(material, material + n_tets);
    I.l1[0].assign(lame1, lame1 + n_tets);
    I.l2[0].assign(lame2, lame2 + n_tets);
    I.rho[0].assign(density, density + n_tets);
    I.extra[0].assign(n_tets, 0.0);
    I.fixed.assign(3L * n_nodes, 0);
    B->inst.push_back(std::move(I));
    if (instance) *instance = B->K() - 1;
  });
}

int pgn_batch_add_shell(pgn_batch* B, int n_nodes, const double* X0, int n_tris, const int* F,
                        const double* lame1, const double* lame2, const double* thickness,
                        const double* density, int n_hinges, const int* hinges,
                        const double* bending_k, int* instance) {
  return bguard([&] {
    PGN_CHECK(B && X0, "null argument");
    PGN_CHECK(!B->built(), "batch already built");
    PGN_CHECK(n_nodes > 0 && n_tris >= 0 && n_hinges >= 0 && n_tris + n_hinges > 0,
              "empty instance");
    PGN_CHECK((n_tris == 0 || (F && lame1 && lame2 && thickness && density)) &&
                  (n_hinges == 0 || (hinges && bending_k)),
              "null input array");
    check_ids(F, 3L * n_tris, n_nodes);
    check_ids(hinges, 4L * n_hinges, n_nodes);
    pgn_batch_inst I;
    I.n = n_nodes;
    I.X0.assign(X0, X0 + 3L * n_nodes);
    for (long t = 0; t < n_tris; ++t) {
      for (int a = 0; a < 3; ++a) I.T[1].push_back(F[3 * t + a]);
      I.T[1].push_back(F[3 * t + 2]);
      I.kind[1].push_back(2);
      I.l1[1].push_back(lame1[t]);
      I.l2[1].push_back(lame2[t]);
      I.rho[1].push_back(density[t]);
      I.extra[1].push_back(thickness[t]);
    }
    for (long h = 0; h < n_hinges; ++h) {
      for (int a = 0; a < 4; ++a) I.T[2].push_back(hinges[4 * h + a]);
      I.kind[2].push_back(3);
      I.l1[2].push_back(0.0);
      I.l2[2].push_back(0.0);
      I.rho[2].push_back(0.0);
      I.extra[2].push_back(bending_k[h]);
    }
    I.fixed.assign(3L * n_nodes, 0);
    B->inst.push_back(std::move(I));
    if (instance) *instance = B->K() - 1;
  });
}

int pgn_batch_set_fixed(pgn_batch* B, int k, const unsigned char* mask) {
  return bguard([&] {
    pgn_batch_inst& I = staged(B, k);
    PGN_CHECK(mask, "null argument");
    I.fixed.assign(mask, mask + 3L * I.n);
  });
}

int pgn_batch_set_gravity(pgn_batch* B, int k, const double g_ini[3], const double g_end[3],
                          int num_stages) {
  return bguard([&] {
    pgn_batch_inst& I = staged(B, k);
    PGN_CHECK(g_ini && g_end && num_stages >= 1, "bad argument");
    I.has_g = true;
    std::copy(g_ini, g_ini + 3, I.g_ini);
    std::copy(g_end, g_end + 3, I.g_end);
    I.g_stages = num_stages;
  });
}

int pgn_batch_set_forces(pgn_batch* B, int k, const double* f_ini, const double* f_end,
                         int num_stages) {
  return bguard([&] {
    pgn_batch_inst& I = staged(B, k);
    PGN_CHECK(f_ini && f_end && num_stages >= 1, "bad argument");
    I.has_f = true;
    I.f_ini.assign(f_ini, f_ini + 3L * I.n);
    I.f_end.assign(f_end, f_end + 3L * I.n);
    I.f_stages = num_stages;
  });
}

int pgn_batch_build(pgn_batch* B) {
  return bguard([&] {
    PGN_CHECK(B, "null batch");
    PGN_CHECK(!B->built(), "batch already built");
    PGN_HIP(hipSetDevice(B->ctx->device));
    B->build();
  });
}

int pgn_batch_num_instances(const pgn_batch* B) { return B ? B->K() : 0; }

size_t pgn_batch_device_bytes(const pgn_batch* B) {
  if (!B || !B->built()) return 0;
  return pgn_model_device_bytes(B->M.get()) +
         sum_bytes(B->noff, B->eoff, B->node_inst, B->grav, B->falpha, B->fac, B->elemE,
                   B->nodeE, B->nodeG2, B->out, B->n2, B->mode, B->mask, B->xsave, B->v0,
                   B->dx, B->bvec, B->wg_seg, B->wg_loc, B->seg_G, B->seg_wg0, B->seg_row0,
                   B->seg_row1, B->seg_tol, B->seg_part1, B->seg_part2, B->seg_partb,
                   B->kvec, B->kvec2, B->xprev, B->seg_sc, B->bcol16, B->ldlt_prob,
                   B->ldlt_perm, B->ldlt_iperm, B->ldlt_list, B->ldlt_status, B->ldlt_band,
                   B->ldlt_D, B->ldlt_sinT, B->ldlt_y, B->eg_mask) +
         sum_bytes(B->sc_agg_ptr, B->sc_agg_nodes, B->sc_agg_of, B->sc_agg_inst, B->sc_aoff,
                   B->sc_moff, B->sc_cb_a, B->sc_cb_b, B->sc_pair_piece0, B->sc_piece_lo,
                   B->sc_cb_fine, B->sc_ok, B->sc_rel, B->sc_cmask, B->sc_A, B->sc_piece_sum,
                   B->sc_rc, B->sc_uc, B->sc_cdot, B->sc_use) +
         sum_bytes(B->sr_inst_d, B->sr_wg0_d, B->sr_inst_b, B->sr_wg0_b, B->sr_lo_d, B->sr_hi_d,
                   B->sr_lo_b, B->sr_hi_b, B->sr_part_d, B->sr_part_b, B->rmask, B->sr_inst_e,
                   B->sr_wg0_e, B->sr_lo_e, B->sr_hi_e, B->sr_kind_e, B->sr_part_e);
}

int pgn_batch_reset(pgn_batch* B) {
  return bguard([&] {
    PGN_CHECK(B && B->built(), "batch not built");
    PGN_HIP(hipSetDevice(B->ctx->device));
    B->reset();
  });
}

int pgn_batch_newton(pgn_batch* B, const pgn_newton_opts* opts, pgn_newton_result* res) {
  return bguard([&] {
    PGN_CHECK(B && res, "null argument");
    PGN_CHECK(B->built(), "batch not built");
    PGN_HIP(hipSetDevice(B->ctx->device));
    pgn_newton_opts o;
    pgn_newton_default_opts(&o);
    if (opts) o = *opts;
    B->newton(o, res);
  });
}

int pgn_batch_get_x(pgn_batch* B, int k, double* x) {
  return bguard([&] {
    PGN_CHECK(B && x, "null argument");
    PGN_CHECK(B->built(), "batch not built");
    PGN_CHECK(k >= 0 && k < B->K(), "instance index out of range");
    PGN_HIP(hipSetDevice(B->ctx->device));
    const long n0 = B->hnoff[k], nk = B->hnoff[k + 1] - n0;
    PGN_HIP(hipMemcpyAsync(x, B->M->x.p + 3 * n0, 3 * nk * sizeof(double),
                           hipMemcpyDeviceToHost, B->ctx->stream));
    PGN_HIP(hipStreamSynchronize(B->ctx->stream));
  });
}

int pgn_batch_get_energy(pgn_batch* B, double* energy) {
  return bguard([&] {
    PGN_CHECK(B && energy, "null argument");
    PGN_CHECK(B->built(), "batch not built");
    PGN_HIP(hipSetDevice(B->ctx->device));
    B->upload_loads();
    std::vector<double> E;
    B->eval_eg(E);
    for (int k = 0; k < B->K(); ++k) energy[k] = E[4 * k];
  });
}

}  // extern "C"
