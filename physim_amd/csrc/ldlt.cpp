// ldlt.cpp -- the EigenLDLT linear solver on the device: band order, the
// regularization tries and their checks around the banded LDL^T kernels.
#include "ldlt.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <string>

namespace pgn {

namespace {

// Breadth-first level structure from r over nodes [n0, n1); returns the
// eccentricity and leaves the last level in `last`.  stamp/mark avoid
// clearing the visit array between searches.
int bfs_levels(int r, int n0, const std::vector<long>& rp, const std::vector<int>& bc,
               std::vector<int>& mark, int stamp, std::vector<int>& last) {
  std::vector<int> cur{r}, next;
  mark[r - n0] = stamp;
  int ecc = 0;
  for (;;) {
    next.clear();
    for (int u : cur)
      for (long e = rp[u]; e < rp[u + 1]; ++e) {
        const int v = bc[e];
        if (mark[v - n0] != stamp) {
          mark[v - n0] = stamp;
          next.push_back(v);
        }
      }
    if (next.empty()) break;
    ++ecc;
    cur.swap(next);
  }
  last = cur;
  return ecc;
}

}  // namespace

// Reverse Cuthill-McKee order of nodes [n0, n1) of the block pattern (the
// pattern must not couple them to nodes outside the range): every connected
// component from a pseudo-peripheral node (George and Liu's level-structure
// search), neighbours taken by increasing degree, the whole order reversed.
void rcm_order(int n0, int n1, const std::vector<long>& rp, const std::vector<int>& bc,
               std::vector<int>& order) {
  const int n = n1 - n0;
  order.clear();
  order.reserve(n);
  std::vector<int> deg(n), mark(n, 0), last, nbr;
  std::vector<char> seen(n, 0);
  for (int i = n0; i < n1; ++i) deg[i - n0] = (int)(rp[i + 1] - rp[i]);
  int stamp = 0;
  for (int s0 = n0; s0 < n1; ++s0) {
    if (seen[s0 - n0]) continue;
    // lowest-degree node of this component as the first candidate
    int r = s0;
    {
      bfs_levels(s0, n0, rp, bc, mark, ++stamp, last);
      for (int i = n0; i < n1; ++i)
        if (mark[i - n0] == stamp && deg[i - n0] < deg[r - n0]) r = i;
    }
    int ecc = bfs_levels(r, n0, rp, bc, mark, ++stamp, last);
    for (int it = 0; it < 8; ++it) {
      int c = last[0];
      for (int v : last)
        if (deg[v - n0] < deg[c - n0]) c = v;
      const int e2 = bfs_levels(c, n0, rp, bc, mark, ++stamp, last);
      if (e2 <= ecc) break;
      r = c;
      ecc = e2;
    }
    // Cuthill-McKee sweep from r
    const size_t head0 = order.size();
    order.push_back(r);
    seen[r - n0] = 1;
    for (size_t h = head0; h < order.size(); ++h) {
      const int u = order[h];
      nbr.clear();
      for (long e = rp[u]; e < rp[u + 1]; ++e) {
        const int v = bc[e];
        if (!seen[v - n0]) {
          seen[v - n0] = 1;
          nbr.push_back(v);
        }
      }
      std::stable_sort(nbr.begin(), nbr.end(),
                       [&](int a, int b) { return deg[a - n0] < deg[b - n0]; });
      order.insert(order.end(), nbr.begin(), nbr.end());
    }
  }
  std::reverse(order.begin(), order.end());
}

int ldlt_plan(int n0, int n1, const std::vector<long>& rp, const std::vector<int>& bc,
              const double* X, std::vector<int>& perm, std::vector<int>& iperm) {
  const int n = n1 - n0;
  // node half-bandwidth of an order (order[q] = node at band position q)
  std::vector<int> pos(n);
  auto band_of = [&](const std::vector<int>& order) {
    for (int q = 0; q < n; ++q) pos[order[q] - n0] = q;
    int bn = 0;
    for (int i = n0; i < n1; ++i)
      for (long e = rp[i]; e < rp[i + 1]; ++e)
        bn = std::max(bn, std::abs(pos[i - n0] - pos[bc[e] - n0]));
    return bn;
  };
  std::vector<std::vector<int>> cand;
  // the given node order (mesh generators number nodes section by section)
  cand.emplace_back(n);
  for (int q = 0; q < n; ++q) cand[0][q] = n0 + q;
  // reverse Cuthill-McKee
  cand.emplace_back();
  rcm_order(n0, n1, rp, bc, cand[1]);
  // rest positions sorted along the longest box axis, then the others
  if (X && n > 0) {
    double lo[3], hi[3];
    for (int c = 0; c < 3; ++c) lo[c] = hi[c] = X[3L * n0 + c];
    for (int i = n0; i < n1; ++i)
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], X[3L * i + c]);
        hi[c] = std::max(hi[c], X[3L * i + c]);
      }
    int ax[3] = {0, 1, 2};
    std::sort(ax, ax + 3, [&](int a, int b) { return hi[a] - lo[a] > hi[b] - lo[b]; });
    std::vector<int> o(n);
    for (int q = 0; q < n; ++q) o[q] = n0 + q;
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) {
      for (int c : ax)
        if (X[3L * a + c] != X[3L * b + c]) return X[3L * a + c] < X[3L * b + c];
      return false;
    });
    cand.push_back(std::move(o));
  }
  int best = 0, bbn = band_of(cand[0]);
  for (size_t c = 1; c < cand.size(); ++c) {
    const int bn = band_of(cand[c]);
    if (bn < bbn) {
      bbn = bn;
      best = (int)c;
    }
  }
  for (int q = 0; q < n; ++q) {
    perm[cand[best][q]] = n0 + q;
    iperm[n0 + q] = cand[best][q];
  }
  return 3 * bbn + 2;
}

}  // namespace pgn

using namespace pgn;

void pgn_model::ldlt_prepare() {
  if (ldlt_ready) return;
  std::vector<int> perm(n), iperm(n);
  const int bw = ldlt_plan(0, n, pat.rowptr, pat.bcol, hX0.data(), perm, iperm);
  PGN_CHECK(bw + 1 <= LDLT_MAXW,
            "EigenLDLT: the half-bandwidth after reordering is " + std::to_string(bw) +
                " DoFs; the direct solver holds at most " + std::to_string(LDLT_MAXW - 1) +
                " (use EigenCG)");
  hipStream_t s = stream();
  LdltProb P;
  P.band0 = 0;
  P.dof0 = 0;
  P.N = 3 * n;
  P.bw = bw;
  P.node0 = 0;
  ldlt_prob.upload(&P, 1, s);
  ldlt_perm.upload(perm, s);
  ldlt_iperm.upload(iperm, s);
  std::vector<int> zeros(n, 0);
  ldlt_nprob.upload(zeros, s);
  const int list0 = 0;
  ldlt_list.upload(&list0, 1, s);
  ldlt_band.alloc((size_t)P.N * (bw + 1));
  ldlt_D.alloc(P.N);
  ldlt_y.alloc(P.N);
  ldlt_sinT.alloc(1);
  ldlt_status.alloc(2);
  sync();
  ldlt_bw = bw;
  ldlt_ready = true;
}

LdltArgs pgn_model::ldlt_args() const {
  LdltArgs a;
  a.prob = ldlt_prob.p;
  a.perm = ldlt_perm.p;
  a.iperm = ldlt_iperm.p;
  a.nprob = ldlt_nprob.p;
  a.band = ldlt_band.p;
  a.D = ldlt_D.p;
  a.sinT = ldlt_sinT.p;
  a.status = ldlt_status.p;
  return a;
}

namespace pgn {

// LinearSolver_EigenLDLT::SolveInternal (LinearSolver_EigenLDLT.cpp:49-153)
// with LinearSolver::GetMatrixData (LinearSolver.cpp:57-66): normA is the
// Frobenius norm of the stored lower entries, sinT = max(min(1e-6 normA,
// 1e-6), 1e-9), regT = 1e-6 normA.  Try i adds regT 10^(i-1) to the
// diagonal (cumulative); a try fails on a zero pivot, on a pivot below sinT
// (singular), on normalized(b).normalized(x) < -regT (indefinite) or on
// |b - A x| above o.tol both absolutely and relative to |b| (inexact).  x
// keeps its input until a factorization succeeds.
void ldlt_linear_solve(pgn_model& M, const double* b, double* x, const pgn_linear_opts& o,
                       pgn_linear_result& res) {
  PGN_CHECK(M.hess_valid, "Hessian not assembled");
  M.ldlt_prepare();
  hipStream_t s = M.stream();
  const long N = 3L * M.n;
  const double normA = M.frob_lower();
  const double sinT = std::max(std::min(1e-6 * normA, 1e-6), 1e-9);
  const double regT = 1e-6 * normA;
  M.ldlt_sinT.upload(&sinT, 1, s);
  res.status = 1;
  res.iterations = 0;
  res.reg_used = 0;
  res.refinements = 0;
  res.rel_residual = 0.0;
  const double nb = std::sqrt(M.dot(b, b, N));
  const LdltArgs a = M.ldlt_args();
  const int regs = std::max(o.reg_iters, 1);
  for (int i = 0; i < regs; ++i) {
    if (i != 0) {
      launch_add_diag(s, M.hfp32, M.n, M.diag.p, M.bval.p, regT * std::pow(10.0, i - 1));
      res.reg_used = i;
    }
    launch_ldlt_gather(s, M.hfp32, a, M.ldlt_list.p, 1, N * (M.ldlt_bw + 1), M.pat.nb,
                       M.brow.p, M.bcol.p, M.bval.p, nullptr);
    launch_ldlt_factor(s, a, M.ldlt_list.p, 1);
    int st[2];
    M.ldlt_status.download(st, 2, s);
    M.sync();
    if (st[0] >= 0 || st[1] > 0) continue;  // zero or singular pivot
    launch_ldlt_solve(s, a, M.ldlt_list.p, 1, b, x, M.ldlt_y.p);
    const double nx = std::sqrt(M.dot(x, x, N));
    if (nb > 0.0 && nx > 0.0 && M.dot(b, x, N) / (nb * nx) < -regT) continue;
    launch_residual64(s, M.hfp32, M.n, M.rowptr.p, M.bcol.p, M.bval.p, x, b, M.r64.p,
                      M.red64.p, M.scal.p + 4);
    PGN_HIP(hipMemcpyAsync(M.hpin + 4, M.scal.p + 4, sizeof(double), hipMemcpyDeviceToHost, s));
    M.sync();
    const double ae = std::sqrt(M.hpin[4]);
    const double re = nb > 0.0 ? ae / nb : std::numeric_limits<double>::quiet_NaN();
    res.rel_residual = nb > 0.0 ? re : 0.0;
    if (ae > o.tol && re > o.tol) continue;  // NaN compares false, as in the reference
    res.status = 0;
    break;
  }
}

}  // namespace pgn
