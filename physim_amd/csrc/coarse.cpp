// coarse.cpp -- host setup of the two-level preconditioner's coarse space.
//
// Aggregates: the rest positions binned into a grid of boxes over the model's
// bounding box, about coarse_target boxes shaped to its extents (empty boxes
// dropped, ids in box order).  Each aggregate spans the translations and the
// infinitesimal rotations about its centroid over its free DoFs; a column is
// kept only where it is independent of the kept ones (an aggregate whose
// free nodes are all fixed, collinear or too few keeps fewer than six).
// PGN_COARSE_AGGS in the environment sets the target (at most 341, the
// coarse dense inverse's capacity).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "model.hpp"

namespace pgn {

namespace {

// keeps column c of the Gram matrix G (6x6) when its pivot, after the kept
// columns, is a fair part of its own diagonal (incremental Cholesky)
uint8_t independent_columns(const double G[36]) {
  double L[36] = {0};
  int kept[6], nk = 0;
  uint8_t mask = 0;
  for (int c = 0; c < 6; ++c) {
    const double gcc = G[6 * c + c];
    if (!(gcc > 0.0)) continue;
    double y[6];
    double d = gcc;
    for (int k = 0; k < nk; ++k) {
      double v = G[6 * kept[k] + c];
      for (int m = 0; m < k; ++m) v -= L[6 * k + m] * y[m];
      y[k] = v / L[6 * k + k];
      d -= y[k] * y[k];
    }
    if (d > 1e-8 * gcc) {
      for (int m = 0; m < nk; ++m) L[6 * nk + m] = y[m];
      L[6 * nk + nk] = std::sqrt(d);
      kept[nk++] = c;
      mask |= (uint8_t)(1u << c);
    }
  }
  return mask;
}

}  // namespace

int bin_aggregates(const double* X, const uint8_t* fix3, long n, int target, int max_aggs,
                   int* agg_of, double* rel, std::vector<uint8_t>& cmask, int bins_out[3]) {
  target = std::max(1, std::min(target, max_aggs));
  // bins shaped to the bounding box
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) lo[k] = hi[k] = X[k];
  for (long i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], X[3 * i + k]);
      hi[k] = std::max(hi[k], X[3 * i + k]);
    }
  double ext[3], emax = 0.0;
  for (int k = 0; k < 3; ++k) {
    ext[k] = hi[k] - lo[k];
    emax = std::max(emax, ext[k]);
  }
  int bins[3] = {1, 1, 1};
  if (emax > 0.0) {
    double e[3], vol = 1.0;
    for (int k = 0; k < 3; ++k) {
      e[k] = std::max(ext[k], 1e-3 * emax);
      vol *= e[k];
    }
    const double s = std::cbrt((double)target / vol);
    for (int k = 0; k < 3; ++k) bins[k] = std::max(1, (int)std::lround(s * e[k]));
    while ((long)bins[0] * bins[1] * bins[2] > max_aggs) {
      const int k = (bins[0] >= bins[1] && bins[0] >= bins[2]) ? 0 : (bins[1] >= bins[2] ? 1 : 2);
      bins[k]--;
    }
  }
  if (bins_out) std::copy(bins, bins + 3, bins_out);
  std::vector<long> key(n);
  for (long i = 0; i < n; ++i) {
    long b[3];
    for (int k = 0; k < 3; ++k) {
      long v = 0;
      if (ext[k] > 0.0) v = (long)std::floor((X[3 * i + k] - lo[k]) / ext[k] * bins[k]);
      b[k] = std::min<long>(std::max<long>(v, 0), bins[k] - 1);
    }
    key[i] = (b[2] * bins[1] + b[1]) * bins[0] + b[0];
  }
  std::vector<long> keys(key);
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  const int na = (int)keys.size();
  for (long i = 0; i < n; ++i)
    agg_of[i] = (int)(std::lower_bound(keys.begin(), keys.end(), key[i]) - keys.begin());
  // centroids, offsets, column masks
  std::vector<double> cen(3L * na, 0.0);
  std::vector<int> cnt(na, 0);
  for (long i = 0; i < n; ++i) {
    const int a = agg_of[i];
    cnt[a]++;
    for (int k = 0; k < 3; ++k) cen[3L * a + k] += X[3 * i + k];
  }
  for (int a = 0; a < na; ++a)
    for (int k = 0; k < 3; ++k) cen[3L * a + k] /= cnt[a];
  for (long i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) rel[3 * i + k] = X[3 * i + k] - cen[3L * agg_of[i] + k];
  // Gram matrix sum_i P_i^T F_i P_i per aggregate, rotation columns scaled by
  // the aggregate's rms offset (so that the independence test is unit-free)
  std::vector<double> rms(na, 0.0);
  for (long i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) rms[agg_of[i]] += rel[3 * i + k] * rel[3 * i + k];
  for (int a = 0; a < na; ++a) rms[a] = std::sqrt(rms[a] / cnt[a]);
  std::vector<double> G(36L * na, 0.0);
  for (long i = 0; i < n; ++i) {
    const int a = agg_of[i];
    const double sc = rms[a] > 0.0 ? 1.0 / rms[a] : 0.0;
    const double x = rel[3 * i] * sc, y = rel[3 * i + 1] * sc, z = rel[3 * i + 2] * sc;
    // rows of P_i: component p of (t + w x rel)
    const double P[3][6] = {{1, 0, 0, 0, z, -y}, {0, 1, 0, -z, 0, x}, {0, 0, 1, y, -x, 0}};
    for (int p = 0; p < 3; ++p) {
      if (fix3 && fix3[3 * i + p]) continue;
      for (int u = 0; u < 6; ++u)
        for (int v = 0; v < 6; ++v) G[36L * a + 6 * u + v] += P[p][u] * P[p][v];
    }
  }
  for (int a = 0; a < na; ++a) cmask.push_back(independent_columns(&G[36L * a]));
  return na;
}

void coarse_pairs(const BsrPattern& P, const std::vector<int>& agg_of, int na, CoarsePairs& out) {
  // per aggregate a, the aggregates b >= a its blocks reach (few: a box and
  // its neighbours), with their block counts
  std::vector<std::vector<std::pair<int, int>>> reach(na);
  auto slot = [&](int a, int b) -> std::pair<int, int>& {
    for (auto& e : reach[a])
      if (e.first == b) return e;
    reach[a].emplace_back(b, 0);
    return reach[a].back();
  };
  for (long f = 0; f < P.nb; ++f) {
    const int a = agg_of[P.brow[f]], b = agg_of[P.bcol[f]];
    if (a <= b) slot(a, b).second++;
  }
  out.cb_a.clear();
  out.cb_b.clear();
  std::vector<int> ptr(1, 0);
  std::vector<std::vector<int>> index(na);
  for (int a = 0; a < na; ++a) {
    std::sort(reach[a].begin(), reach[a].end());
    for (const auto& e : reach[a]) {
      index[a].push_back((int)out.cb_a.size());
      out.cb_a.push_back(a);
      out.cb_b.push_back(e.first);
      ptr.push_back(ptr.back() + e.second);
    }
  }
  out.cb_fine.assign(ptr.back(), 0);
  std::vector<int> pos(ptr.begin(), ptr.end() - 1);
  for (long f = 0; f < P.nb; ++f) {
    const int a = agg_of[P.brow[f]], b = agg_of[P.bcol[f]];
    if (a > b) continue;
    const auto& r = reach[a];
    const int j = (int)(std::lower_bound(r.begin(), r.end(), std::make_pair(b, 0)) - r.begin());
    out.cb_fine[pos[index[a][j]]++] = (int)f;
  }
  // pieces of at most COARSE_PIECE blocks, pair by pair
  out.pair_piece0.assign(out.cb_a.size() + 1, 0);
  out.piece_lo.clear();
  for (size_t c = 0; c < out.cb_a.size(); ++c) {
    out.pair_piece0[c] = (int)out.piece_lo.size();
    for (int t = ptr[c]; t < ptr[c + 1]; t += COARSE_PIECE) out.piece_lo.push_back(t);
  }
  out.pair_piece0[out.cb_a.size()] = (int)out.piece_lo.size();
  out.piece_lo.push_back(ptr.back());
}

}  // namespace pgn

using namespace pgn;

void pgn_model::coarse_prepare() {
  if (coarse_ready) return;
  if (const char* e = std::getenv("PGN_COARSE_AGGS")) coarse_target = std::max(1, std::atoi(e));
  h_agg_of.assign(n, 0);
  std::vector<double> rel(3L * n);
  std::vector<uint8_t> cmask;
  const int na = bin_aggregates(hX0.data(), hfix.size() == 3u * n ? hfix.data() : nullptr, n,
                                coarse_target, COARSE_MAX_PAD / 6, h_agg_of.data(), rel.data(),
                                cmask, coarse_bins);
  // nodes by aggregate
  std::vector<int> agg_ptr(na + 1, 0);
  for (int i = 0; i < n; ++i) agg_ptr[h_agg_of[i] + 1]++;
  for (int a = 0; a < na; ++a) agg_ptr[a + 1] += agg_ptr[a];
  std::vector<int> nodes(n), cur(agg_ptr.begin(), agg_ptr.end() - 1);
  for (int i = 0; i < n; ++i) nodes[cur[h_agg_of[i]]++] = i;
  CoarsePairs cp;
  coarse_pairs(pat, h_agg_of, na, cp);
  const std::vector<int>& cb_a = cp.cb_a;
  const std::vector<int>& cb_b = cp.cb_b;
  const std::vector<int>& pair_piece0 = cp.pair_piece0;
  const std::vector<int>& piece_lo = cp.piece_lo;
  const std::vector<int>& cb_fine = cp.cb_fine;
  // device copies
  hipStream_t s = stream();
  cg_agg_ptr.upload(agg_ptr, s);
  cg_agg_nodes.upload(nodes, s);
  cg_agg_of.upload(h_agg_of, s);
  cg_cmask.upload(cmask, s);
  cg_cb_a.upload(cb_a, s);
  cg_cb_b.upload(cb_b, s);
  cg_pair_piece0.upload(pair_piece0, s);
  cg_piece_lo.upload(piece_lo, s);
  cg_cb_fine.upload(cb_fine, s);
  std::vector<float> r32;  // alive until the sync below
  if (hfp32) {
    r32.assign(rel.begin(), rel.end());
    cg_rel.upload(reinterpret_cast<const char*>(r32.data()), r32.size() * sizeof(float), s);
  } else {
    cg_rel.upload(reinterpret_cast<const char*>(rel.data()), rel.size() * sizeof(double), s);
  }
  const int npad = (6 * na + COARSE_TILE - 1) / COARSE_TILE * COARSE_TILE;
  cg_A.ensure(2L * npad * npad);
  cg_D.ensure(2L * COARSE_TILE * COARSE_TILE);
  cg_ok.ensure(1);
  coarse_npiece = (int)piece_lo.size() - 1;
  cg_piece_sum.ensure(36L * std::max(coarse_npiece, 1));
  cg_rc.ensure((size_t)COARSE_SPLIT * npad);
  PGN_HIP(hipMemsetAsync(cg_rc.p, 0, cg_rc.bytes(), s));  // padding stays zero
  cg_uc.ensure(npad);
  const int gg = coarse_gemv_grid(npad);
  cg_cpart.ensure(gg);
  cg_cdot.ensure(1);
  PGN_HIP(hipMemsetAsync(cg_cdot.p, 0, sizeof(double), s));
  coarse_ncb = (int)cb_a.size();
  cga = CoarseArgs{};
  cga.na = na;
  cga.npad = npad;
  cga.gemv_g = gg;
  cga.agg_ptr = cg_agg_ptr.p;
  cga.agg_nodes = cg_agg_nodes.p;
  cga.agg_of = cg_agg_of.p;
  cga.rel = cg_rel.p;
  cga.cmask = cg_cmask.p;
  cga.fixmask = fixmask.p;
  cga.R = nullptr;
  cga.ok = cg_ok.p;
  cga.rc = cg_rc.p;
  cga.uc = cg_uc.p;
  cga.cpart = cg_cpart.p;
  cga.cdot = cg_cdot.p;
  sync();
  coarse_ready = true;
}

double* pgn_model::coarse_update(bool assemble_only, bool async) {
  coarse_prepare();
  hipStream_t st = stream();
  if (async) {
    if (!side) {
      PGN_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
      PGN_HIP(hipEventCreateWithFlags(&ev_hess, hipEventDisableTiming));
      PGN_HIP(hipEventCreateWithFlags(&ev_coarse, hipEventDisableTiming));
    }
    PGN_HIP(hipEventRecord(ev_hess, stream()));
    PGN_HIP(hipStreamWaitEvent(side, ev_hess, 0));
    st = side;
  } else {
    coarse_join();
  }
  CoarseBuild b;
  b.ncb = coarse_ncb;
  b.npiece = coarse_npiece;
  b.cb_a = cg_cb_a.p;
  b.cb_b = cg_cb_b.p;
  b.pair_piece0 = cg_pair_piece0.p;
  b.piece_lo = cg_piece_lo.p;
  b.cb_fine = cg_cb_fine.p;
  b.piece_sum = cg_piece_sum.p;
  b.brow = brow.p;
  b.bcol = bcol.p;
  b.bval = bval.p;
  const long nn = (long)cga.npad * cga.npad;
  b.A0 = cg_A.p;
  b.A1 = cg_A.p + nn;
  b.D = cg_D.p;
  b.ok = cg_ok.p;
  double* R = launch_coarse_operator(st, hfp32, cga, b, assemble_only);
  PGN_HIP(hipGetLastError());
  if (!assemble_only) cga.R = R;
  if (async) {
    PGN_HIP(hipEventRecord(ev_coarse, side));
    coarse_pending = true;
  }
  return R;
}

void pgn_model::coarse_join() {
  if (!coarse_pending) return;
  PGN_HIP(hipStreamWaitEvent(stream(), ev_coarse, 0));
  coarse_pending = false;
}
