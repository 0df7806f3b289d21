// rod_model.cpp -- a batch of discrete elastic rods resident on the GPU, and
// the batched Newton solve over it.
//
// Setup restates Simulable_DER / Mesh_DER (Simulable_DER.cpp:66-170,
// Mesh_DER.cpp:97-164, Mesh_Frame.cpp:117-195): twist-free reference frames,
// rest lengths, rest bend strain (EnergyElement_DER_BendTwist.cpp:59-65,
// :285-302), zero rest twist, lumped node masses rho L / 2
// (MassElement_Lumped.cpp:47-54 on edges).
//
// The Newton solve runs OptimSolver::SolveFull (OptimSolver.cpp:163-264) for
// every rod at once: each rod keeps its own state machine (load stage, line
// search factor, MINIMPR / MINSTEP / NONDESC counters), while every GPU launch
// (element evaluation, band assembly, LDL^T, frame update) covers all rods.
#include "rod_model.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace pgn {

__global__ void rod_select_kernel(long nd, const int* rod_of, const unsigned char* sel,
                                  const double* g, double* dx) {
  const long d = blockIdx.x * 256L + threadIdx.x;
  if (d >= nd) return;
  if (sel[rod_of[d >> 2]]) dx[d] = -g[d];
}

__global__ void rod_neg_kernel(long nd, const double* g, double* b) {
  const long d = blockIdx.x * 256L + threadIdx.x;
  if (d < nd) b[d] = -g[d];
}

// per-rod |v|^2 (one workgroup per rod, fixed order)
__global__ __launch_bounds__(256) void rod_norm2_kernel(const long* voff, const double* v,
                                                        double* out) {
  __shared__ double red[8];
  const int r = blockIdx.x;
  const long d0 = 4 * voff[r], d1 = 4 * voff[r + 1];
  double s = 0.0;
  for (long d = d0 + threadIdx.x; d < d1; d += 256) s += v[d] * v[d];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) out[r] = s;
}

}  // namespace pgn

using namespace pgn;

static void pt_h(const double* u, const double* t1, const double* t2, double* r) {
  const double d = t1[0] * t2[0] + t1[1] * t2[1] + t1[2] * t2[2];
  const double b[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2],
                       t1[0] * t2[1] - t1[1] * t2[0]};
  const double bu[3] = {b[1] * u[2] - b[2] * u[1], b[2] * u[0] - b[0] * u[2],
                        b[0] * u[1] - b[1] * u[0]};
  const double k = (b[0] * u[0] + b[1] * u[1] + b[2] * u[2]) / (1.0 + d);
  for (int i = 0; i < 3; ++i) r[i] = u[i] * d + b[i] * k + bu[i];
}

void pgn_rods::setup(int n_rods, const int* nverts, const double* X0h, const double* nor0,
                     const double* young, const double* shear, const double* rh,
                     const double* rw, const double* rho, int precision) {
  PGN_CHECK(n_rods > 0, "need at least one rod");
  PGN_CHECK(precision == PGN_FP64 || precision == PGN_MIXED, "unknown precision");
  nrods = n_rods;
  prec = precision;
  hvoff.assign(nrods + 1, 0);
  for (int r = 0; r < nrods; ++r) {
    PGN_CHECK(nverts[r] >= 3, "a rod needs at least 3 vertices");
    hvoff[r + 1] = hvoff[r] + nverts[r];
  }
  nv = hvoff[nrods];
  const double pi = 3.14159265358979323846;
  std::vector<double> hx(4 * nv, 0.0), hL0(nv, 0.0), hk(nv, 0.0), hbtc(6 * nv, 0.0),
      hfr(6 * nv, 0.0), hmass(nv, 0.0);
  hrod_of.assign(nv, 0);
  hflags.assign(nv, 0);
  hfixmask.assign(nv, 0);
  for (int r = 0; r < nrods; ++r) {
    const long v0 = hvoff[r], v1 = hvoff[r + 1];
    for (long v = v0; v < v1; ++v) {
      hrod_of[v] = r;
      for (int i = 0; i < 3; ++i) hx[4 * v + i] = X0h[3 * v + i];
      uint8_t f = 0;
      if (v + 1 < v1) f |= ROD_HAS_EDGE;
      if (v > v0) f |= ROD_HAS_PREV_EDGE;
      if (v > v0 && v + 1 < v1) f |= ROD_HAS_BT;
      hflags[v] = f;
    }
    hfixmask[v1 - 1] = 8;  // dummy angle of the last vertex
    // edges: rest length, stretch stiffness, twist-free reference frames
    for (long v = v0; v + 1 < v1; ++v) {
      double t[3];
      for (int i = 0; i < 3; ++i) t[i] = X0h[3 * (v + 1) + i] - X0h[3 * v + i];
      const double l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
      PGN_CHECK(l > 0.0, "zero-length rod edge");
      for (int i = 0; i < 3; ++i) t[i] /= l;
      hL0[v] = l;
      hk[v] = young[r] * pi * rh[r] * rw[r];
      hmass[v] += rho[r] * l / 2.0;
      hmass[v + 1] += rho[r] * l / 2.0;
      double n[3];
      if (v == v0) {
        const double* a = nor0 + 3 * r;
        const double k = a[0] * t[0] + a[1] * t[1] + a[2] * t[2];
        for (int i = 0; i < 3; ++i) n[i] = a[i] - k * t[i];
        const double ln = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        PGN_CHECK(ln > 1e-12, "first frame normal is parallel to the first edge");
        for (int i = 0; i < 3; ++i) n[i] /= ln;
      } else {
        pt_h(&hfr[6 * (v - 1) + 3], &hfr[6 * (v - 1)], t, n);
      }
      for (int i = 0; i < 3; ++i) {
        hfr[6 * v + i] = t[i];
        hfr[6 * v + 3 + i] = n[i];
      }
    }
    // bend-twist: stiffnesses and rest strains
    for (long v = v0 + 1; v + 1 < v1; ++v) {
      const double lbar = 0.5 * (hL0[v - 1] + hL0[v]);
      const double* te = &hfr[6 * (v - 1)];
      const double* ne = te + 3;
      const double* tf = &hfr[6 * v];
      const double* nf = tf + 3;
      double kb[3] = {te[1] * tf[2] - te[2] * tf[1], te[2] * tf[0] - te[0] * tf[2],
                      te[0] * tf[1] - te[1] * tf[0]};
      const double den = 2.0 / (1.0 + te[0] * tf[0] + te[1] * tf[1] + te[2] * tf[2]);
      for (int i = 0; i < 3; ++i) kb[i] *= den;
      const double be[3] = {te[1] * ne[2] - te[2] * ne[1], te[2] * ne[0] - te[0] * ne[2],
                            te[0] * ne[1] - te[1] * ne[0]};
      const double bf[3] = {tf[1] * nf[2] - tf[2] * nf[1], tf[2] * nf[0] - tf[0] * nf[2],
                            tf[0] * nf[1] - tf[1] * nf[0]};
      double K0 = 0.0, K1 = 0.0;
      for (int i = 0; i < 3; ++i) {
        K0 += 0.5 * (be[i] + bf[i]) * kb[i];
        K1 += -0.5 * (ne[i] + nf[i]) * kb[i];
      }
      double* p = &hbtc[6 * v];
      p[0] = pi / 8.0 * young[r] * rw[r] * rh[r] * rh[r] * rh[r] / lbar;
      p[1] = pi / 8.0 * young[r] * rw[r] * rw[r] * rw[r] * rh[r] / lbar;
      p[2] = pi / 8.0 * shear[r] * rw[r] * rh[r] * (rw[r] * rw[r] + rh[r] * rh[r]) / lbar;
      p[3] = K0;
      p[4] = K1;
      p[5] = 0.0;
    }
  }
  hframe0 = hfr;
  hx0 = hx;
  hipStream_t s = stream();
  voff.upload(hvoff, s);
  rod_of.upload(hrod_of, s);
  // the LDL^T is a chain of steps per rod: the longest chains start first so
  // that no long rod waits behind short ones for a slot (stable: ties in
  // rod order)
  std::vector<int> order(nrods);
  for (int r = 0; r < nrods; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return hvoff[a + 1] - hvoff[a] > hvoff[b + 1] - hvoff[b];
  });
  ldlt_order.upload(order, s);
  PGN_HIP(hipStreamSynchronize(s));  // order is released at the end of this function
  flags.upload(hflags, s);
  fixmask.upload(hfixmask, s);
  x.upload(hx, s);
  L0.upload(hL0, s);
  kst.upload(hk, s);
  btc.upload(hbtc, s);
  frame.upload(hfr, s);
  std::vector<double> zeros(nv, 0.0);
  rt.upload(zeros, s);
  mass.upload(hmass, s);
  es.alloc(nv);
  gs.alloc(3 * nv);
  Js.alloc(6 * nv);
  eb.alloc(nv);
  gb.alloc(8 * nv);
  Hb.alloc(36 * nv * sizeof(double));
  g.alloc(4 * nv);
  vE.alloc(3 * nv);
  rodE.alloc(4 * nrods);
  band.alloc((ROD_BW + 1) * 4 * nv);
  work.alloc((ROD_BW + 1) * 4 * nv);
  status.alloc(nrods);
  xsave.alloc(4 * nv);
  dx.alloc(4 * nv);
  rhs.alloc(4 * nv);
  fac.alloc(nrods);
  sel.alloc(nrods);
  rn2.alloc(nrods);
  sync();
}

void pgn_rods::sync() { PGN_HIP(hipStreamSynchronize(stream())); }

RodArgs pgn_rods::args() const {
  RodArgs a;
  a.nv = nv;
  a.nrods = nrods;
  a.voff = voff.p;
  a.rod_of = rod_of.p;
  a.flags = flags.p;
  a.fixmask = fixmask.p;
  a.x = x.p;
  a.L0 = L0.p;
  a.kst = kst.p;
  a.btc = btc.p;
  a.frame = frame.p;
  a.rt = rt.p;
  a.mass = mass.p;
  const double al = has_g ? (double)g_cur / g_stages : 0.0;
  for (int i = 0; i < 3; ++i) a.grav[i] = has_g ? g_ini[i] + (g_end[i] - g_ini[i]) * al : 0.0;
  return a;
}

RodBufs pgn_rods::bufs() {
  RodBufs b;
  b.es = es.p;
  b.gs = gs.p;
  b.Js = Js.p;
  b.eb = eb.p;
  b.gb = gb.p;
  b.Hb = Hb.p;
  b.g = g.p;
  b.vE = vE.p;
  b.rodE = rodE.p;
  b.band = band.p;
  b.work = work.p;
  b.status = status.p;
  return b;
}

void pgn_rods::eval(bool hess, const std::vector<unsigned char>* live) {
  RodBufs b = bufs();
  // PGN_MIXED: the bend-twist Hessians in FP32 (energy and gradient FP64)
  const bool fp32 = prec == PGN_MIXED;
  RodArgs a = args();
  if (live) {
    livem.upload(reinterpret_cast<const uint8_t*>(live->data()), live->size(), stream());
    a.live = livem.p;
  }
  rod_launch_eval(stream(), a, hess, fp32, b);
  if (hess) rod_launch_band(stream(), a, fp32, b);
}

void pgn_rods::download_rodE(std::vector<double>& out) {
  out.resize(4 * nrods);
  PGN_HIP(hipMemcpyAsync(out.data(), rodE.p, 4 * nrods * sizeof(double), hipMemcpyDeviceToHost,
                         stream()));
  sync();
}

// reference layout <-> internal vertex-major layout
long pgn_rods::ref_dofs() const { return 4 * nv - nrods; }

void pgn_rods::to_ref(const double* in4, double* out) const {
  long o = 0;
  for (int r = 0; r < nrods; ++r) {
    const long v0 = hvoff[r], n = hvoff[r + 1] - v0;
    for (long i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) out[o + 3 * i + c] = in4[4 * (v0 + i) + c];
    for (long i = 0; i + 1 < n; ++i) out[o + 3 * n + i] = in4[4 * (v0 + i) + 3];
    o += 4 * n - 1;
  }
}

void pgn_rods::from_ref(const double* in, double* out4) const {
  long o = 0;
  for (int r = 0; r < nrods; ++r) {
    const long v0 = hvoff[r], n = hvoff[r + 1] - v0;
    for (long i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) out4[4 * (v0 + i) + c] = in[o + 3 * i + c];
    for (long i = 0; i + 1 < n; ++i) out4[4 * (v0 + i) + 3] = in[o + 3 * n + i];
    out4[4 * (v0 + n - 1) + 3] = 0.0;
    o += 4 * n - 1;
  }
}

long pgn_rods::ref_index(int r, long i4) const {
  // internal DoF (global) -> reference index within the rod's block
  const long v0 = hvoff[r], n = hvoff[r + 1] - v0;
  const long v = i4 / 4 - v0;
  const int c = (int)(i4 % 4);
  if (c < 3) return 3 * v + c;
  return (v + 1 < n) ? 3 * n + v : -1;
}

// ---- batched Newton -----------------------------------------------------------------

void pgn_rods::newton(const pgn_newton_opts& o, pgn_newton_result* R) {
  PGN_CHECK(o.qpSolverType == PGN_QP_NEWTON, "rod batches take Newton steps (PGN_QP_NEWTON)");
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = stream();
  const long nd = 4 * nv;
  const unsigned gd = (unsigned)((nd + 255) / 256);
  std::vector<double> E;
  enum { ACTIVE = 0, DONE = 1 };
  std::vector<int> st(nrods, ACTIVE), result(nrods, PGN_OS_SUCCESS), minImpr(nrods, 0),
      minStep(nrods, 0), steep(nrods, 0), its(nrods, 0), bis(nrods, 0), evals(nrods, 0),
      hevals(nrods, 0), stepping(nrods, 0), nbis(nrods, 0), searching(nrods, 0),
      guarded(nrods, 0);
  std::vector<double> obj(nrods), opt(nrods), objNew(nrods), optNew(nrods), f(nrods),
      len(nrods), stepLen(nrods), escale(nrods);
  std::vector<unsigned char> hsel(nrods), hact(nrods);
  std::vector<double> hfac(nrods);

  // OptimSolver::Init
  eval(false);
  download_rodE(E);
  for (int r = 0; r < nrods; ++r) {
    obj[r] = E[4 * r];
    opt[r] = std::sqrt(E[4 * r + 1]);
    evals[r]++;
  }
  auto upload_fac = [&]() {
    PGN_HIP(hipMemcpyAsync(fac.p, hfac.data(), nrods * sizeof(double), hipMemcpyHostToDevice, s));
  };
  for (;;) {
    bool any = false;
    for (int r = 0; r < nrods; ++r) any |= (st[r] == ACTIVE);
    if (!any) break;
    // OnIterStart: every active rod has taken the same number of steps, so
    // one load stage counter serves all of them
    if (has_g && g_cur < g_stages) ++g_cur;
    PGN_HIP(hipMemcpyAsync(xsave.p, x.p, nd * sizeof(double), hipMemcpyDeviceToDevice, s));
    std::vector<unsigned char> live(nrods);
    for (int r = 0; r < nrods; ++r) live[r] = st[r] == ACTIVE;
    eval(false, &live);
    download_rodE(E);
    bool need_h = false, need_steep = false;
    for (int r = 0; r < nrods; ++r) {
      stepping[r] = 0;
      if (st[r] != ACTIVE) continue;
      its[r]++;
      evals[r]++;
      obj[r] = E[4 * r];
      opt[r] = std::sqrt(E[4 * r + 1]);
      if (opt[r] < o.tolMaxError) {
        result[r] = PGN_OS_SUCCESS;
        continue;
      }
      stepping[r] = 1;
      if (steep[r]) need_steep = true;
      else need_h = true;
    }
    if (need_h) {
      for (int r = 0; r < nrods; ++r) {
        hact[r] = stepping[r] && !steep[r];
        if (hact[r]) hevals[r]++;
      }
      eval(true, &hact);
      hipLaunchKernelGGL(rod_neg_kernel, dim3(gd), dim3(256), 0, s, nd, g.p, rhs.p);
      PGN_HIP(hipMemcpyAsync(sel.p, hact.data(), nrods, hipMemcpyHostToDevice, s));
      RodBufs b = bufs();
      rod_launch_ldlt(s, args(), b, rhs.p, dx.p, std::max(o.lsNumRegTrials, 1), o.lsTolMaxError,
                      ldlt_order.p, sel.p);
    }
    if (need_steep) {
      for (int r = 0; r < nrods; ++r) hsel[r] = stepping[r] && steep[r];
      PGN_HIP(hipMemcpyAsync(sel.p, hsel.data(), nrods, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(rod_select_kernel, dim3(gd), dim3(256), 0, s, nd, rod_of.p, sel.p, g.p,
                         dx.p);
    }
    hipLaunchKernelGGL(rod_norm2_kernel, dim3(nrods), dim3(256), 0, s, voff.p, dx.p, rn2.p);
    std::vector<double> hn2(nrods);
    PGN_HIP(hipMemcpyAsync(hn2.data(), rn2.p, nrods * sizeof(double), hipMemcpyDeviceToHost, s));
    sync();
    // maxStepSize clamp folded into the line-search factor
    std::vector<double> clamp(nrods, 1.0);
    for (int r = 0; r < nrods; ++r) {
      if (!stepping[r]) continue;
      len[r] = std::sqrt(hn2[r]);
      if (len[r] > o.maxStepSize) {
        clamp[r] = o.maxStepSize / len[r];
        len[r] = o.maxStepSize;
      }
      f[r] = 1.0;
      nbis[r] = -1;
      searching[r] = 1;
      guarded[r] = 0;
    }
    // LineSearch_Simple, all rods in lockstep
    for (;;) {
      bool anys = false;
      for (int r = 0; r < nrods; ++r) {
        hfac[r] = -1.0;
        if (stepping[r] && searching[r]) {
          nbis[r]++;
          stepLen[r] = f[r] * len[r];
          hfac[r] = f[r] * clamp[r];
          anys = true;
        }
      }
      if (!anys) break;
      upload_fac();
      rod_launch_step(s, args(), fac.p, xsave.p, dx.p, x.p);
      for (int r = 0; r < nrods; ++r) live[r] = stepping[r] && searching[r];
      eval(false, &live);
      download_rodE(E);
      bool restore = false;
      for (int r = 0; r < nrods; ++r) {
        hfac[r] = -1.0;
        if (!(stepping[r] && searching[r])) continue;
        evals[r]++;
        const double en = E[4 * r], gn = std::sqrt(E[4 * r + 1]);
        objNew[r] = en;
        optNew[r] = gn;
        if (en < obj[r]) {
          searching[r] = 0;
          continue;
        }
        if (o.ls_noise_guard > 0.0 && en <= obj[r] + o.ls_noise_guard * E[4 * r + 2] &&
            gn < opt[r]) {
          guarded[r] = 1;  // booked as an improvement of zero
          searching[r] = 0;
          continue;
        }
        f[r] *= o.lSearch_alpha;
        if (!(stepLen[r] > o.tolMinStep && nbis[r] <= o.lsearch_numTry)) {
          searching[r] = 0;  // exhausted: back to the saved state
          hfac[r] = 0.0;
          restore = true;
        }
      }
      if (restore) {
        upload_fac();
        rod_launch_step(s, args(), fac.p, xsave.p, dx.p, x.p);
      }
    }
    // UpdateStepResult
    for (int r = 0; r < nrods; ++r) {
      if (!stepping[r]) continue;
      bis[r] += nbis[r];
      double impr = obj[r] - objNew[r];
      if (guarded[r] && impr < 0) impr = 0;
      if (impr > 0 || guarded[r]) {
        if (impr < o.tolMinImpr) {
          minImpr[r]++;
          result[r] = PGN_OS_MINIMPR;
        } else if (stepLen[r] < o.tolMinStep) {
          minStep[r]++;
          result[r] = PGN_OS_MINSTEP;
        } else if (optNew[r] <= o.tolMaxError) {
          result[r] = PGN_OS_SUCCESS;
          minImpr[r] = minStep[r] = 0;
        } else {
          result[r] = PGN_OS_ONGOING;
          minImpr[r] = minStep[r] = 0;
        }
        obj[r] = objNew[r];
        opt[r] = optNew[r];
      } else {
        result[r] = PGN_OS_NONDESC;
      }
    }
    // OnIterFinish: UpdateKinematics (frames and reference twist)
    rod_launch_kinematics(s, args(), frame.p, rt.p);
    // SolveFull state machine (the gradient at the current state is the
    // line search's, or the step start's after a restore)
    const bool loaded = !has_g || g_cur == g_stages;
    for (int r = 0; r < nrods; ++r) {
      if (st[r] != ACTIVE) continue;
      bool stop = false;
      if (loaded) {
        int& res = result[r];
        if (res == PGN_OS_NONDESC) {
          if (!o.trySteepest || steep[r]) stop = true;
          else steep[r] = 1;
        }
        if (!stop && res == PGN_OS_MINIMPR) {
          if (minImpr[r] == o.minImprMaxCount) stop = true;
          else if (o.trySteepest) steep[r] = !steep[r];
        }
        if (!stop && res == PGN_OS_MINSTEP) {
          if (minStep[r] == o.minStepMaxCount) stop = true;
          else if (o.trySteepest) steep[r] = !steep[r];
        }
        if (!stop && res == PGN_OS_SUCCESS) {
          if (opt[r] <= o.tolMaxError) stop = true;
          else res = PGN_OS_ONGOING;
        }
        if (!stop && res == PGN_OS_ONGOING) steep[r] = 0;
      }
      if (!stop && its[r] + 1 > o.numMaxIters) {
        result[r] = PGN_OS_MAXITER;
        stop = true;
      }
      if (stop) st[r] = DONE;
    }
  }
  sync();
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int r = 0; r < nrods; ++r) {
    R[r] = pgn_newton_result{};
    R[r].result = result[r];
    R[r].iterations = its[r];
    R[r].line_search_bisections = bis[r];
    R[r].energy_evals = evals[r];
    R[r].hessian_evals = hevals[r];
    R[r].objective = obj[r];
    R[r].optimality = opt[r];
    R[r].seconds = secs;
  }
}
