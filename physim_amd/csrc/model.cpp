// model.cpp -- setup and evaluation of a device-resident tet FEM model.
//
// Setup restates Simulable_Mesh::InitInternal (Simulable_Mesh.cpp:68),
// EnergyElement_FEM::Init (EnergyElement_FEM.cpp:59-122: Dm^-1 from the rest
// shape, weight (1/6) det Dm) and MassElement_Lumped::ComputeAndStore_Mass
// (MassElement_Lumped.cpp:47-54).  The BSR3 pattern and the element->block
// slot map replace AMatrixSd's triplet assembly + pointer map
// (BasicTypes.cpp:18-164, DerivativeAssembler.cpp:205-301).
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace pgn {

void build_bsr_pattern(int n, int m, const int* T, BsrPattern& P) {
  PGN_CHECK(n > 0, "model needs at least one node");
  PGN_CHECK(m >= 0, "negative element count");
  PGN_CHECK((long)m * 16 < (1L << 32), "too many elements for 32-bit slot codes");
  for (long q = 0; q < 4L * m; ++q)
    PGN_CHECK(T[q] >= 0 && T[q] < n,
              "element " + std::to_string(q / 4) + " has node id out of range");
  P.n = n;
  // node -> (element, local) incidence, element order
  P.inc_ptr.assign(n + 1, 0);
  for (long q = 0; q < 4L * m; ++q) P.inc_ptr[T[q] + 1]++;
  for (int i = 0; i < n; ++i) P.inc_ptr[i + 1] += P.inc_ptr[i];
  P.inc.resize(4L * m);
  {
    std::vector<int> cur(P.inc_ptr.begin(), P.inc_ptr.end() - 1);
    for (int e = 0; e < m; ++e)
      for (int a = 0; a < 4; ++a) P.inc[cur[T[4L * e + a]]++] = 4 * e + a;
  }
  // rows: sorted unique neighbour nodes (including the node itself)
  P.rowptr.assign(n + 1, 0);
  std::vector<int> mark(n, -1), nb;
  std::vector<int> cols;
  cols.reserve(27L * n);
  for (int i = 0; i < n; ++i) {
    nb.clear();
    mark[i] = i;
    nb.push_back(i);
    for (int p = P.inc_ptr[i]; p < P.inc_ptr[i + 1]; ++p) {
      const int e = P.inc[p] >> 2;
      for (int b = 0; b < 4; ++b) {
        const int j = T[4L * e + b];
        if (mark[j] != i) {
          mark[j] = i;
          nb.push_back(j);
        }
      }
    }
    std::sort(nb.begin(), nb.end());
    cols.insert(cols.end(), nb.begin(), nb.end());
    P.rowptr[i + 1] = (long)cols.size();
  }
  P.nb = (long)cols.size();
  P.bcol.swap(cols);
  P.brow.resize(P.nb);
  P.diag.resize(n);
  for (int i = 0; i < n; ++i)
    for (long s = P.rowptr[i]; s < P.rowptr[i + 1]; ++s) {
      P.brow[s] = i;
      if (P.bcol[s] == i) P.diag[i] = s;
    }
  auto find = [&](int i, int j) -> long {
    const int* b = P.bcol.data() + P.rowptr[i];
    const int* e = P.bcol.data() + P.rowptr[i + 1];
    const int* it = std::lower_bound(b, e, j);
    return P.rowptr[i] + (it - b);
  };
  // contributions per block, in element order
  P.cptr.assign(P.nb + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int p = P.inc_ptr[i]; p < P.inc_ptr[i + 1]; ++p) {
      const int e = P.inc[p] >> 2;
      for (int b = 0; b < 4; ++b) P.cptr[find(i, T[4L * e + b]) + 1]++;
    }
  for (long s = 0; s < P.nb; ++s) P.cptr[s + 1] += P.cptr[s];
  P.code.resize(P.cptr[P.nb]);
  std::vector<long> cur(P.cptr.begin(), P.cptr.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int p = P.inc_ptr[i]; p < P.inc_ptr[i + 1]; ++p) {
      const int e = P.inc[p] >> 2, a = P.inc[p] & 3;
      for (int b = 0; b < 4; ++b) {
        const long s = find(i, T[4L * e + b]);
        P.code[cur[s]++] = (uint32_t)e * 16u + (uint32_t)(4 * a + b);
      }
    }
}

}  // namespace pgn

using namespace pgn;

pgn_model::~pgn_model() {
  if (side) {
    (void)hipStreamSynchronize(side);
    (void)hipStreamDestroy(side);
  }
  if (ev_hess) (void)hipEventDestroy(ev_hess);
  if (ev_coarse) (void)hipEventDestroy(ev_coarse);
  pcg_graphs.clear();
  if (hkctl) (void)hipHostFree(hkctl);
  for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
  if (hpin) (void)hipHostFree(hpin);
  if (hsc) (void)hipHostFree(hsc);
}

namespace {

double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
void cross3(const double* a, const double* b, double* r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// EnergyElement_DiscreteShells::Init (EnergyElement_DiscreteShells.cpp:50-54):
// rest dihedral (DiscreteShellsBending_Dihedral) and shape factor
// |e|^2 / (4 (|e x (x2-x0)| + |(x3-x0) x e|)) (DiscreteShellsBending_Shape)
void hinge_rest(const double* p0, const double* p1, const double* p2, const double* p3,
                double& theta, double& shape) {
  double e[3], a2[3], a3[3], N0[3], N1[3];
  for (int i = 0; i < 3; ++i) {
    e[i] = p1[i] - p0[i];
    a2[i] = p2[i] - p0[i];
    a3[i] = p3[i] - p0[i];
  }
  cross3(e, a2, N0);
  cross3(a3, e, N1);
  const double l0 = std::sqrt(dot3(N0, N0)), l1 = std::sqrt(dot3(N1, N1));
  const double le = std::sqrt(dot3(e, e));
  PGN_CHECK(l0 > 0.0 && l1 > 0.0 && le > 0.0, "degenerate rest hinge");
  double n0[3], mv[3], mn[3];
  for (int i = 0; i < 3; ++i) {
    n0[i] = N0[i] / l0;
    mv[i] = n0[i] + N1[i] / l1;
  }
  const double lm = std::sqrt(dot3(mv, mv));
  PGN_CHECK(lm > 0.0, "rest hinge folded flat onto itself");
  for (int i = 0; i < 3; ++i) mv[i] /= lm;
  cross3(mv, n0, mn);
  theta = 2.0 * std::atan(dot3(e, mn) / le / dot3(mv, n0));
  shape = dot3(e, e) / (4.0 * l0 + 4.0 * l1);
}

// rest triangle in its own orthonormal 2D frame (t1 along edge 01, t2 =
// n x t1): Dm^-1 (row-major 2x2) and area
void membrane_rest(const double* p0, const double* p1, const double* p2, double* Dinv,
                   double& area) {
  double e1[3], e2[3], nv[3], t1[3], t2[3];
  for (int i = 0; i < 3; ++i) {
    e1[i] = p1[i] - p0[i];
    e2[i] = p2[i] - p0[i];
  }
  cross3(e1, e2, nv);
  const double ln = std::sqrt(dot3(nv, nv)), l1 = std::sqrt(dot3(e1, e1));
  PGN_CHECK(ln > 0.0 && l1 > 0.0, "degenerate rest triangle");
  for (int i = 0; i < 3; ++i) {
    nv[i] /= ln;
    t1[i] = e1[i] / l1;
  }
  cross3(nv, t1, t2);
  const double a = dot3(e1, t1), b = dot3(e2, t1), c = dot3(e1, t2), d = dot3(e2, t2);
  const double det = a * d - b * c;
  Dinv[0] = d / det;
  Dinv[1] = -b / det;
  Dinv[2] = -c / det;
  Dinv[3] = a / det;
  area = 0.5 * std::fabs(det);
}

// nodes of an element that carry lumped mass (a membrane's fourth slot
// repeats its third node; hinges carry none)
int mass_nodes(int kind) { return kind <= 1 ? 4 : (kind == 2 ? 3 : 0); }

}  // namespace

void pgn_model::setup(int n_nodes, const double* X0h, int n_elems, const int* Th,
                      const uint8_t* kindh, const double* l1, const double* l2,
                      const double* rho, const double* extra, int precision) {
  PGN_CHECK(precision == PGN_FP64 || precision == PGN_MIXED, "unknown precision");
  PGN_CHECK(X0h && (n_elems == 0 || (Th && kindh && l1 && l2 && rho)),
            "null input array");
  n = n_nodes;
  m = n_elems;
  prec = precision;
  hfp32 = (precision == PGN_MIXED);
  hipStream_t s = stream();
  // element order: tets, membranes, hinges
  m_tet = m_mem = m_hinge = 0;
  for (int e = 0; e < m; ++e) {
    const int k = kindh[e];
    PGN_CHECK(k <= 3, "unknown element kind at element " + std::to_string(e));
    const int cls = k <= 1 ? 0 : k - 1;
    const int prev = (m_hinge > 0) ? 2 : (m_mem > 0 ? 1 : 0);
    PGN_CHECK(cls >= prev, "elements must be ordered tets, membranes, hinges");
    if (cls == 0) ++m_tet;
    else if (cls == 1) ++m_mem;
    else ++m_hinge;
  }
  PGN_CHECK(!(m_mem + m_hinge) || extra, "shell elements need thickness / kB values");
  {
    const char* mf = std::getenv("PGN_MATFREE");
    use_matfree = mf && std::strcmp(mf, "1") == 0 && m_mem == 0 && m_hinge == 0 && m > 0;
    if (use_matfree) {  // the textbook split form with direct launches
      pcg_cg1 = false;
      pcg_split = true;
      use_graphs = 0;
    }
  }
  build_bsr_pattern(n, m, Th, pat);
  hX0.assign(X0h, X0h + 3L * n);
  hfix.assign(3L * n, 0);

  // element parameters (12 SoA rows).  Tets: rows of Dm^-1, V0 = det(Dm)/6,
  // lame1, lame2.  Membranes / hinges: see ShellArgs.
  std::vector<double> p64(12L * m, 0.0);
  std::vector<float> p32(12L * m);
  std::vector<double> emass(m, 0.0), hmass(n, 0.0);
  std::vector<int4> ht(m);
  for (int e = 0; e < m; ++e) {
    const int* t = Th + 4L * e;
    ht[e] = make_int4(t[0], t[1], t[2], t[3]);
    if (kindh[e] == 2) {
      PGN_CHECK(t[3] == t[2], "membrane element " + std::to_string(e) +
                                  " must repeat its third node in the fourth slot");
      double Dinv[4], area;
      membrane_rest(X0h + 3L * t[0], X0h + 3L * t[1], X0h + 3L * t[2], Dinv, area);
      for (int c = 0; c < 4; ++c) p64[(long)c * m + e] = Dinv[c];
      p64[9L * m + e] = area * extra[e];
      p64[10L * m + e] = l1[e];
      p64[11L * m + e] = l2[e];
      // MassElement_Lumped: rho * VolumeBasis / nodes, VolumeBasis of a
      // triangle = its area (Face_Tri.cpp:33-39)
      emass[e] = rho[e] * area / 3.0;
      continue;
    }
    if (kindh[e] == 3) {
      double th, sh;
      hinge_rest(X0h + 3L * t[0], X0h + 3L * t[1], X0h + 3L * t[2], X0h + 3L * t[3], th, sh);
      p64[0L * m + e] = extra[e];
      p64[1L * m + e] = th;
      p64[2L * m + e] = sh;
      continue;
    }
    double Dm[3][3];  // Dm(i, a) = X_{a+1}(i) - X_0(i)
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < 3; ++i) Dm[i][a] = X0h[3L * t[a + 1] + i] - X0h[3L * t[0] + i];
    const double c00 = Dm[1][1] * Dm[2][2] - Dm[1][2] * Dm[2][1];
    const double c01 = Dm[1][2] * Dm[2][0] - Dm[1][0] * Dm[2][2];
    const double c02 = Dm[1][0] * Dm[2][1] - Dm[1][1] * Dm[2][0];
    const double det = Dm[0][0] * c00 + Dm[0][1] * c01 + Dm[0][2] * c02;
    PGN_CHECK(det != 0.0 && std::isfinite(det),
              "degenerate rest element " + std::to_string(e));
    const double id = 1.0 / det;
    double Bi[3][3];
    Bi[0][0] = c00 * id;
    Bi[0][1] = (Dm[0][2] * Dm[2][1] - Dm[0][1] * Dm[2][2]) * id;
    Bi[0][2] = (Dm[0][1] * Dm[1][2] - Dm[0][2] * Dm[1][1]) * id;
    Bi[1][0] = c01 * id;
    Bi[1][1] = (Dm[0][0] * Dm[2][2] - Dm[0][2] * Dm[2][0]) * id;
    Bi[1][2] = (Dm[0][2] * Dm[1][0] - Dm[0][0] * Dm[1][2]) * id;
    Bi[2][0] = c02 * id;
    Bi[2][1] = (Dm[0][1] * Dm[2][0] - Dm[0][0] * Dm[2][1]) * id;
    Bi[2][2] = (Dm[0][0] * Dm[1][1] - Dm[0][1] * Dm[1][0]) * id;
    const double V0 = det / 6.0;
    for (int c = 0; c < 9; ++c) p64[(long)c * m + e] = Bi[c / 3][c % 3];
    p64[9L * m + e] = V0;
    p64[10L * m + e] = l1[e];
    p64[11L * m + e] = l2[e];
    emass[e] = rho[e] * V0 / 4.0;
  }
  for (long q = 0; q < 12L * m; ++q) p32[q] = (float)p64[q];
  // lumped masses, summed per node in element order
  for (int i = 0; i < n; ++i)
    for (int p = pat.inc_ptr[i]; p < pat.inc_ptr[i + 1]; ++p) {
      const int e = pat.inc[p] >> 2, a = pat.inc[p] & 3;
      if (a < mass_nodes(kindh[e])) hmass[i] += emass[e];
    }

  X0.upload(X0h, 3L * n, s);
  x.upload(X0h, 3L * n, s);
  tets.upload(ht, s);
  kind.upload(kindh, m, s);
  prm64.upload(p64, s);
  prm32.upload(p32, s);
  mass.upload(hmass, s);
  std::vector<uint8_t> fm(n, 0);
  fixmask.upload(fm, s);
  inc_ptr.upload(pat.inc_ptr, s);
  inc.upload(pat.inc, s);
  rowptr.upload(pat.rowptr, s);
  cptr.upload(pat.cptr, s);
  diag.upload(pat.diag, s);
  bcol.upload(pat.bcol, s);
  brow.upload(pat.brow, s);
  code.upload(pat.code, s);
  const size_t ts = tsize();
  bval.alloc(9 * pat.nb * ts);
  dinv.alloc(3L * n * ts);
  binv.alloc(6L * n * ts);
  eg64.alloc(12L * std::max(m, 1));
  eH.alloc((size_t)TET_H_PACK * std::max(m, 1) * ts);
  blockE.alloc(elem_partials() + 1);
  nodeP.alloc(2L * node_gather_blocks(n) + 2);
  scal.alloc(16);
  g.alloc(3L * n);
  pv.alloc(6 * 3L * n * ts);
  pb.alloc(3L * n * ts);
  part1.alloc(PCG_MAXP);
  part2.alloc(4L * PCG_MAXP);
  partb.alloc(PCG_MAXP);
  red64.alloc(PCG_MAXP + 8);
  sc.alloc(1);
  cgs.alloc(8);
  r64.alloc(3L * n);
  tmp64a.alloc(3L * n);
  tmp64b.alloc(3L * n);
  PGN_HIP(hipMemsetAsync(part1.p, 0, part1.bytes(), s));
  PGN_HIP(hipMemsetAsync(part2.p, 0, part2.bytes(), s));
  PGN_HIP(hipMemsetAsync(red64.p, 0, red64.bytes(), s));
  PGN_HIP(hipMemsetAsync(sc.p, 0, sc.bytes(), s));
  PGN_HIP(hipMemsetAsync(pv.p, 0, pv.bytes(), s));
  PGN_HIP(hipHostMalloc((void**)&hpin, 16 * sizeof(double), hipHostMallocDefault));
  PGN_HIP(hipHostMalloc((void**)&hsc, sizeof(PcgScalars), hipHostMallocDefault));
  PGN_HIP(hipHostMalloc((void**)&hkctl, 2 * sizeof(int), hipHostMallocDefault));
  kctl.alloc(2);
  sync();
}

void pgn_model::sync() { PGN_HIP(hipStreamSynchronize(stream())); }

// BCondition.h:140-155 GetCurValues with VectorAnimator_LerpValues
void pgn_model::cur_gravity(double gv[3]) const {
  if (!has_g) {
    gv[0] = gv[1] = gv[2] = 0.0;
    return;
  }
  const double a = (double)g_cur / (double)g_stages;
  for (int i = 0; i < 3; ++i) gv[i] = g_ini[i] + (g_end[i] - g_ini[i]) * a;
}

double pgn_model::cur_force_alpha() const {
  return has_f ? (double)f_cur / (double)f_stages : 0.0;
}

// BCondition.cpp:42 StepLoading for every BC
bool pgn_model::step_load() {
  bool stepped = false;
  if (has_g && g_cur < g_stages) {
    ++g_cur;
    stepped = true;
  }
  if (has_f && f_cur < f_stages) {
    ++f_cur;
    stepped = true;
  }
  return stepped;
}

bool pgn_model::is_loaded() const {
  if (has_g && g_cur != g_stages) return false;
  if (has_f && f_cur != f_stages) return false;
  return true;
}

TetArgs pgn_model::tet_args(bool fp32) const {
  TetArgs a;
  a.m = m_tet;
  a.stride = m;
  a.tets = tets.p;
  a.kind = kind.p;
  a.prm = fp32 ? (const void*)prm32.p : (const void*)prm64.p;
  a.x = x.p;
  a.X0 = X0.p;
  return a;
}

ShellArgs pgn_model::shell_args(bool fp32, int k) const {
  ShellArgs a;
  a.first = (k == SHELL_MEMBRANE) ? m_tet : m_tet + m_mem;
  a.count = (k == SHELL_MEMBRANE) ? m_mem : m_hinge;
  a.stride = m;
  a.elems = tets.p;
  a.prm = fp32 ? (const void*)prm32.p : (const void*)prm64.p;
  a.prm64 = prm64.p;
  a.x = x.p;
  a.X0 = X0.p;
  return a;
}

static int parts(int count) { return (count + TET_BLOCK - 1) / TET_BLOCK; }
static int tet_parts(int count) { return (count + 63) / 64; }  // one per wavefront

int pgn_model::elem_partials() const {
  return tet_parts(m_tet) + parts(m_mem) + parts(m_hinge);
}

void pgn_model::launch_tet(bool fp32, int flags, double* bE, double* elemE, void* eg,
                           void* eHp) {
  hipStream_t s = stream();
  const TetArgs ta = tet_args(fp32);
  const bool tunable = flags == (EVAL_E | EVAL_G) || flags == EVAL_H ||
                       flags == (EVAL_E | EVAL_G | EVAL_H);
  int& tile = tet_tile_choice[fp32 ? 1 : 0][flags & 7];
  if (tile < 0) {
    tile = 0;
    if (const char* e = std::getenv("PGN_TET_TILE")) {
      tile = std::max(0, std::min(TET_TILES - 1, std::atoi(e)));
    } else if (tunable && m_tet >= TUNE_MIN_TETS) {
      hipEvent_t e0, e1;
      PGN_HIP(hipEventCreate(&e0));
      PGN_HIP(hipEventCreate(&e1));
      float best = 0.0f;
      for (int t = 0; t < TET_TILES; ++t) {
        launch_tet_eval(s, fp32, flags, ta, bE, elemE, eg, eHp, t);  // warm-up
        PGN_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r) launch_tet_eval(s, fp32, flags, ta, bE, elemE, eg, eHp, t);
        PGN_HIP(hipEventRecord(e1, s));
        PGN_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        PGN_HIP(hipEventElapsedTime(&ms, e0, e1));
        tet_tile_us[fp32 ? 1 : 0][flags & 7][t] = 1e3f * ms / 3.0f;
        if (t == 0 || ms < best) {
          best = ms;
          tile = t;
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
  }
  launch_tet_eval(s, fp32, flags, ta, bE, elemE, eg, eHp, tile);
}

void pgn_model::launch_elements(bool fp32, int flags, double* bE, double* elemE, void* eg,
                                void* eHp) {
  hipStream_t s = stream();
  launch_tet(fp32, flags, bE, elemE, eg, eHp);
  double* b2 = bE ? bE + tet_parts(m_tet) : nullptr;
  launch_shell_eval(s, fp32, SHELL_MEMBRANE, flags, shell_args(fp32, SHELL_MEMBRANE), b2, elemE,
                    eg, eHp);
  double* b3 = bE ? b2 + parts(m_mem) : nullptr;
  launch_shell_eval(s, fp32, SHELL_HINGE, flags, shell_args(fp32, SHELL_HINGE), b3, elemE, eg,
                    eHp);
}

NodeArgs pgn_model::node_args() const {
  NodeArgs a;
  a.n = n;
  a.inc_ptr = inc_ptr.p;
  a.inc = inc.p;
  a.mass = mass.p;
  a.fixmask = fixmask.p;
  a.x = x.p;
  cur_gravity(a.grav);
  a.f_ini = has_f ? f_ini.p : nullptr;
  a.f_end = has_f ? f_end.p : nullptr;
  a.f_alpha = cur_force_alpha();
  if (dyn_dt > 0.0) {
    a.xhat = xhat.p;
    a.inv_dt2 = 1.0 / (dyn_dt * dyn_dt);
  }
  return a;
}

BsrArgs pgn_model::bsr_args() const {
  BsrArgs a;
  a.nb = pat.nb;
  a.brow = brow.p;
  a.bcol = bcol.p;
  a.cptr = cptr.p;
  a.code = code.p;
  a.fixmask = fixmask.p;
  if (dyn_dt > 0.0) {
    a.mass = mass.p;
    a.inv_dt2 = 1.0 / (dyn_dt * dyn_dt);
  }
  return a;
}

void pgn_model::sell_prepare() {
  if (sell_ready || !use_sell) return;
  const long ns = (n + 63) / 64;
  std::vector<long> so(ns + 1, 0);
  std::vector<int> len(n);
  for (int i = 0; i < n; ++i) len[i] = (int)(pat.rowptr[i + 1] - pat.rowptr[i]);
  for (long sl = 0; sl < ns; ++sl) {
    int w = 0;
    for (long r = sl * 64; r < std::min<long>(n, sl * 64 + 64); ++r) w = std::max(w, len[r]);
    so[sl + 1] = so[sl] + 64L * w;
  }
  const long slots = so[ns];
  // Columns as 16-bit offsets from the slot's row when every block of the
  // pattern lies within 32767 nodes of the diagonal (PGN_SELL_C16=0 keeps
  // 32-bit columns).  Padding slots hold the row itself, or, for the lanes
  // past the last row, an offset to the last row, so that the kernel's
  // unconditional first-block loads stay in bounds.
  bool c16 = !(std::getenv("PGN_SELL_C16") && std::strcmp(std::getenv("PGN_SELL_C16"), "0") == 0);
  for (int r = 0; r < n && c16; ++r)
    for (long e = pat.rowptr[r]; e < pat.rowptr[r + 1]; ++e)
      if (std::abs((long)pat.bcol[e] - r) > 32767) {
        c16 = false;
        break;
      }
  hipStream_t s = stream();
  if (c16) {
    std::vector<short> col(slots, 0);
    for (long sl = 0; sl < ns; ++sl)
      for (long q = 0; q < 64; ++q) {
        const long r = sl * 64 + q;
        if (r >= n) {
          const long w = (so[sl + 1] - so[sl]) / 64;
          for (long j = 0; j < w; ++j) col[so[sl] + 64L * j + q] = (short)(n - 1 - r);
          continue;
        }
        for (int j = 0; j < len[r]; ++j)
          col[so[sl] + 64L * j + q] = (short)(pat.bcol[pat.rowptr[r] + j] - r);
      }
    sell_c16.upload(col, s);
  } else {
    std::vector<int> col(slots, 0);  // padding slots: column 0, never used
    for (long sl = 0; sl < ns; ++sl)
      for (long r = sl * 64; r < std::min<long>(n, sl * 64 + 64); ++r)
        for (int j = 0; j < len[r]; ++j)
          col[so[sl] + 64L * j + (r - sl * 64)] = pat.bcol[pat.rowptr[r] + j];
    sell_col.upload(col, s);
  }
  sell_so.upload(so, s);
  sell_len.upload(len, s);
  sell_val.alloc((size_t)9 * std::max<long>(slots, 1) * tsize());
  sell_ns = ns;
  sell_grid(ns, sell_spw, sell_g);
  sync();
  sell_ready = true;
}

void pgn_model::sell_pack() {
  if (!use_sell) return;
  sell_prepare();
  launch_sell_pack(stream(), hfp32, pat.nb, brow.p, rowptr.p, sell_so.p, bval.p, sell_val.p);
}

PcgArgs pgn_model::pcg_args(double tol, int precond) const {
  PcgArgs a;
  std::memset(&a, 0, sizeof(a));  // padding too: chunk graphs compare argument bytes
  a.lpr = pcg_lpr;
  a.split = pcg_split ? 1 : 0;
  a.cg1 = pcg_cg1 ? 1 : 0;
  a.cgs = cgs.p;
  a.g1 = pcg_grid1(n, pcg_lpr);
  if (use_sell && sell_ready) {  // init and the SpMV share the SELL grid
    a.sell = 1;
    a.g1 = sell_g;
    a.sell_spw = sell_spw;
    a.sell_ns = sell_ns;
    a.sell_so = sell_so.p;
    a.sell_len = sell_len.p;
    a.sell_col = sell_col.p;
    a.sell_c16 = sell_c16.p;
    a.sell_val = sell_val.p;
  }
  a.g2 = pcg_grid2(n);
  // PGN_K2_GRID overrides the update launch's grid (diagnostics: it changes
  // only the summation order of the r.r / r.z partials)
  if (const char* e = std::getenv("PGN_K2_GRID"))
    a.g2 = std::max(1, std::min(PCG_GRID_CAP, std::atoi(e)));
  a.tol = tol;
  a.partb = partb.p;
  const size_t v = 3L * n * tsize();
  char* base = pv.p;
  a.n = n;
  a.rowptr = rowptr.p;
  a.bcol = bcol.p;
  a.bval = bval.p;
  a.dinv = dinv.p;
  a.binv = binv.p;
  // precond 2: block Jacobi in the update launch plus the coarse correction
  a.precond = precond >= 1 ? 1 : 0;
  if (precond == 2) {
    PGN_CHECK(coarse_ready && cga.R, "two-level preconditioner: coarse operator not built");
    a.coarse = 1;
    a.cg = cga;
  }
  a.x = base;
  a.r = base + v;
  a.z = base + 2 * v;
  a.p0 = base + 3 * v;
  a.p1 = base + 4 * v;
  a.q = base + 5 * v;
  a.part1 = part1.p;
  a.part2 = part2.p;
  a.sc = sc.p;
  // the matrix-free product for static solves (the inertia of a time step
  // lives in the assembled diagonal only)
  if (use_matfree && dyn_dt == 0.0 && mf_ye.p) {
    a.mf = 1;
    a.mfa.m = m_tet;
    a.mfa.n = n;
    a.mfa.tets = tets.p;
    a.mfa.eH = eH.p;
    a.mfa.ye = mf_ye.p;
    a.mfa.inc_ptr = inc_ptr.p;
    a.mfa.inc = inc.p;
    a.mfa.fixmask = fixmask.p;
    a.mfa.shift = mf_shift;
  }
  return a;
}

void pgn_model::launch_eg() {
  hipStream_t s = stream();
  launch_elements(false, EVAL_E | EVAL_G, blockE.p, nullptr, eg64.p, nullptr);
  const NodeArgs na = node_args();
  launch_node_gather(s, false, na, eg64.p, g.p, nodeP.p);
  launch_energy_final(s, blockE.p, elem_partials(), nodeP.p, node_gather_blocks(n), scal.p);
}

void pgn_model::eval_eg(double& E, double& gg, double* escale) {
  launch_eg();
  PGN_HIP(hipMemcpyAsync(hpin, scal.p, 3 * sizeof(double), hipMemcpyDeviceToHost, stream()));
  sync();
  E = hpin[0];
  gg = hpin[1];
  if (escale) *escale = hpin[2];
}

void pgn_model::update_precond() {
  launch_precond(stream(), hfp32, n, diag.p, bval.p, dinv.p, binv.p);
}

void pgn_model::eval_hessian() {
  hipStream_t s = stream();
  launch_elements(hfp32, EVAL_H, nullptr, nullptr, nullptr, eH.p);
  launch_bsr_gather(s, hfp32, bsr_args(), eH.p, bval.p);
  update_precond();
  mf_shift = 0.0;
  hess_valid = true;
}

MatFreeArgs pgn_model::mf_args() {
  MatFreeArgs a;
  std::memset(&a, 0, sizeof(a));
  a.m = m_tet;
  a.n = n;
  a.tets = tets.p;
  a.eH = eH.p;
  mf_ye.ensure(12L * std::max<long>(m_tet, 1) * tsize());
  a.ye = mf_ye.p;
  a.inc_ptr = inc_ptr.p;
  a.inc = inc.p;
  a.fixmask = fixmask.p;
  a.shift = mf_shift;
  return a;
}

double pgn_model::frob_lower() {
  launch_frob_lower(stream(), hfp32, pat.nb, brow.p, bcol.p, bval.p, red64.p, scal.p + 5);
  PGN_HIP(hipMemcpyAsync(hpin + 5, scal.p + 5, sizeof(double), hipMemcpyDeviceToHost, stream()));
  sync();
  return std::sqrt(hpin[5]);
}

double pgn_model::dot(const double* a, const double* b, long len) {
  launch_dot64(stream(), len, a, b, red64.p, scal.p + 3);
  PGN_HIP(hipMemcpyAsync(hpin + 3, scal.p + 3, sizeof(double), hipMemcpyDeviceToHost, stream()));
  sync();
  return hpin[3];
}

// Autotuner: time the BSR SpMV with 4, 8 and 16 lanes per block row on this
// model's own matrix (assembled or not, the access pattern is the same) and
// keep the fastest for the PCG.
// Lanes per block row of the PCG SpMV, from the mean blocks per row: 16 from
// 12 blocks (tet meshes, ~18), 8 from 6 (shells), else 4.  The choice fixes
// the FP32 summation order of every row, so it is made from the matrix
// alone and a solve gives the same bits on every run; the three widths are
// still timed once (10 launches each) for the report.  PGN_PCG_LPR=4|8|16
// in the environment overrides the rule.
void pgn_model::tune() {
  if (tuned) return;
  if (const char* env = std::getenv("PGN_PCG_LPR")) {
    const int l = std::atoi(env);
    if (l == 4 || l == 8 || l == 16) {
      pcg_lpr = l;
      tuned = true;
      return;
    }
  }
  hipStream_t s = stream();
  const size_t v = 3L * n * tsize();
  char* base = pv.p;
  const int lprs[3] = {4, 8, 16};
  hipEvent_t e0, e1;
  PGN_HIP(hipEventCreate(&e0));
  PGN_HIP(hipEventCreate(&e1));
  for (int c = 0; c < 3; ++c) {
    launch_spmv(s, hfp32, n, rowptr.p, bcol.p, bval.p, base + 4 * v, base + 5 * v, lprs[c]);
    PGN_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < 10; ++r)
      launch_spmv(s, hfp32, n, rowptr.p, bcol.p, bval.p, base + 4 * v, base + 5 * v, lprs[c]);
    PGN_HIP(hipEventRecord(e1, s));
    PGN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    PGN_HIP(hipEventElapsedTime(&ms, e0, e1));
    tune_us[c] = 1e3 * ms / 10;
  }
  const double mean_blocks = n > 0 ? (double)pat.nb / n : 0.0;
  pcg_lpr = mean_blocks >= 12.0 ? 16 : (mean_blocks >= 6.0 ? 8 : 4);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  tuned = true;
}
