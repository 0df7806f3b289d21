// rod_api.cpp -- extern "C" entry points of the rod batch (include/pgn.h).
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "rod_model.hpp"

using namespace pgn;

namespace pgn {
int api_guard_set(const char* msg, int code);
}

namespace {
template <class F>
int rguard(F&& f) {
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
}  // namespace

extern "C" {

int pgn_rods_create(pgn_ctx* ctx, int n_rods, const int* n_verts, const double* X0,
                    const double* frame0_normal, const double* young, const double* shear,
                    const double* radius_h, const double* radius_w, const double* density,
                    int precision, pgn_rods** out) {
  return rguard([&] {
    PGN_CHECK(ctx && out && n_verts && X0 && frame0_normal && young && shear && radius_h &&
                  radius_w && density,
              "null argument");
    *out = nullptr;
    PGN_HIP(hipSetDevice(ctx->device));
    auto* R = new pgn_rods();
    R->ctx = ctx;
    try {
      R->setup(n_rods, n_verts, X0, frame0_normal, young, shear, radius_h, radius_w, density,
               precision);
    } catch (...) {
      delete R;
      throw;
    }
    *out = R;
  });
}

void pgn_rods_destroy(pgn_rods* R) {
  if (!R) return;
  (void)hipStreamSynchronize(R->stream());
  delete R;
}

long pgn_rods_num_dofs(const pgn_rods* R) { return R ? R->ref_dofs() : 0; }

int pgn_rods_set_fixed(pgn_rods* R, const unsigned char* mask) {
  return rguard([&] {
    PGN_CHECK(R && mask, "null argument");
    std::vector<uint8_t> fm(R->nv, 0);
    long o = 0;
    for (int r = 0; r < R->nrods; ++r) {
      const long v0 = R->hvoff[r], n = R->hvoff[r + 1] - v0;
      for (long i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c)
          if (mask[o + 3 * i + c]) fm[v0 + i] |= (uint8_t)(1u << c);
      for (long i = 0; i + 1 < n; ++i)
        if (mask[o + 3 * n + i]) fm[v0 + i] |= 8;
      fm[v0 + n - 1] |= 8;
      o += 4 * n - 1;
    }
    R->hfixmask = fm;
    R->fixmask.upload(fm, R->stream());
    R->sync();
  });
}

int pgn_rods_set_gravity(pgn_rods* R, const double g_ini[3], const double g_end[3],
                         int num_stages) {
  return rguard([&] {
    PGN_CHECK(R && g_ini && g_end && num_stages >= 1, "bad argument");
    R->has_g = true;
    for (int i = 0; i < 3; ++i) {
      R->g_ini[i] = g_ini[i];
      R->g_end[i] = g_end[i];
    }
    R->g_stages = num_stages;
    R->g_cur = 0;
  });
}

int pgn_rods_reset(pgn_rods* R) {
  return rguard([&] {
    PGN_CHECK(R, "null argument");
    R->x.upload(R->hx0, R->stream());
    R->frame.upload(R->hframe0, R->stream());
    std::vector<double> z(R->nv, 0.0);
    R->rt.upload(z, R->stream());
    R->sync();
    R->g_cur = 0;
  });
}

int pgn_rods_set_x(pgn_rods* R, const double* x) {
  return rguard([&] {
    PGN_CHECK(R && x, "null argument");
    std::vector<double> h(4 * R->nv);
    R->from_ref(x, h.data());
    R->x.upload(h, R->stream());
    R->sync();
  });
}

int pgn_rods_get_x(pgn_rods* R, double* x) {
  return rguard([&] {
    PGN_CHECK(R && x, "null argument");
    std::vector<double> h(4 * R->nv);
    R->x.download(h.data(), h.size(), R->stream());
    R->sync();
    R->to_ref(h.data(), x);
  });
}

int pgn_rods_get_energy(pgn_rods* R, double* energy) {
  return rguard([&] {
    PGN_CHECK(R && energy, "null argument");
    R->eval(false);
    std::vector<double> E;
    R->download_rodE(E);
    for (int r = 0; r < R->nrods; ++r) energy[r] = E[4 * r];
  });
}

int pgn_rods_get_gradient(pgn_rods* R, double* gradient) {
  return rguard([&] {
    PGN_CHECK(R && gradient, "null argument");
    R->eval(false);
    std::vector<double> h(4 * R->nv);
    R->g.download(h.data(), h.size(), R->stream());
    R->sync();
    R->to_ref(h.data(), gradient);
  });
}

int pgn_rods_get_hessian_dense(pgn_rods* R, int rod, double* H) {
  return rguard([&] {
    PGN_CHECK(R && H && rod >= 0 && rod < R->nrods, "bad argument");
    R->eval(true);
    constexpr int W = ROD_BW + 1;
    std::vector<double> band(W * 4 * R->nv);
    R->band.download(band.data(), band.size(), R->stream());
    R->sync();
    const long v0 = R->hvoff[rod], n = R->hvoff[rod + 1] - v0;
    const long N = 4 * n - 1;
    std::memset(H, 0, sizeof(double) * N * N);
    for (long d = 4 * v0; d < 4 * (v0 + n); ++d) {
      const long rd = R->ref_index(rod, d);
      if (rd < 0) continue;
      for (int k = 0; k < W; ++k) {
        const long c = d - k;
        if (c < 4 * v0) break;
        const long rc = R->ref_index(rod, c);
        if (rc < 0) continue;
        H[rd * N + rc] = band[W * d + k];
        H[rc * N + rd] = band[W * d + k];
      }
    }
  });
}

int pgn_rods_get_frames(pgn_rods* R, double* tan, double* nor, double* rt) {
  return rguard([&] {
    PGN_CHECK(R, "null argument");
    std::vector<double> fr(6 * R->nv), t(R->nv);
    R->frame.download(fr.data(), fr.size(), R->stream());
    R->rt.download(t.data(), t.size(), R->stream());
    R->sync();
    long e = 0;
    for (int r = 0; r < R->nrods; ++r)
      for (long v = R->hvoff[r]; v + 1 < R->hvoff[r + 1]; ++v, ++e)
        for (int i = 0; i < 3; ++i) {
          if (tan) tan[3 * e + i] = fr[6 * v + i];
          if (nor) nor[3 * e + i] = fr[6 * v + 3 + i];
        }
    if (rt)
      for (long v = 0; v < R->nv; ++v) rt[v] = t[v];
  });
}

int pgn_rods_set_frames(pgn_rods* R, const double* tan, const double* nor, const double* rt) {
  return rguard([&] {
    PGN_CHECK(R && tan && nor && rt, "null argument");
    std::vector<double> fr(6 * R->nv, 0.0), t(rt, rt + R->nv);
    long e = 0;
    for (int r = 0; r < R->nrods; ++r)
      for (long v = R->hvoff[r]; v + 1 < R->hvoff[r + 1]; ++v, ++e)
        for (int i = 0; i < 3; ++i) {
          fr[6 * v + i] = tan[3 * e + i];
          fr[6 * v + 3 + i] = nor[3 * e + i];
        }
    R->frame.upload(fr, R->stream());
    R->rt.upload(t, R->stream());
    R->sync();
  });
}

int pgn_rods_update_kinematics(pgn_rods* R) {
  return rguard([&] {
    PGN_CHECK(R, "null argument");
    rod_launch_kinematics(R->stream(), R->args(), R->frame.p, R->rt.p);
    R->sync();
  });
}

int pgn_rod_bt_eval(pgn_ctx* ctx, int count, const double* prm, const double* e1,
                    const double* e2, const double* theta, int precision, double* E, double* g8,
                    double* H64) {
  return rguard([&] {
    PGN_CHECK(ctx && prm && e1 && e2 && theta && E && g8 && H64, "null argument");
    PGN_CHECK(count >= 0, "negative element count");
    PGN_CHECK(precision == PGN_FP64 || precision == PGN_MIXED, "unknown precision");
    PGN_HIP(hipSetDevice(ctx->device));
    if (count == 0) return;
    hipStream_t s = ctx->stream;
    DBuf<double> dp, d1, d2, dt, dE, dg;
    DBuf<char> dH;
    dp.upload(prm, 27L * count, s);
    d1.upload(e1, 3L * count, s);
    d2.upload(e2, 3L * count, s);
    dt.upload(theta, 2L * count, s);
    dE.alloc(count);
    dg.alloc(8L * count);
    const bool fp32 = precision == PGN_MIXED;
    dH.alloc(36L * count * (fp32 ? sizeof(float) : sizeof(double)));
    rod_launch_bt_raw(s, count, dp.p, d1.p, d2.p, dt.p, fp32, dE.p, dg.p, dH.p);
    PGN_HIP(hipGetLastError());
    std::vector<double> h(36L * count);
    dE.download(E, count, s);
    dg.download(g8, 8L * count, s);
    if (fp32) {
      std::vector<float> hf(h.size());
      PGN_HIP(hipMemcpyAsync(hf.data(), dH.p, hf.size() * sizeof(float), hipMemcpyDeviceToHost,
                             s));
      PGN_HIP(hipStreamSynchronize(s));
      std::copy(hf.begin(), hf.end(), h.begin());
    } else {
      PGN_HIP(hipMemcpyAsync(h.data(), dH.p, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                             s));
      PGN_HIP(hipStreamSynchronize(s));
    }
    for (long k = 0; k < count; ++k) {
      int q = 0;
      for (int i = 0; i < 8; ++i)
        for (int j = 0; j <= i; ++j, ++q) {
          H64[64 * k + 8 * i + j] = h[36 * k + q];
          H64[64 * k + 8 * j + i] = h[36 * k + q];
        }
    }
  });
}

int pgn_rods_element_bt(pgn_rods* R, double* E, double* g8, double* H64) {
  return rguard([&] {
    PGN_CHECK(R, "null argument");
    R->eval(true);
    std::vector<double> e(R->nv), g(8 * R->nv), h(36 * R->nv);
    R->eb.download(e.data(), e.size(), R->stream());
    R->gb.download(g.data(), g.size(), R->stream());
    if (R->prec == PGN_MIXED) {  // FP32 Hessians
      std::vector<float> hf(h.size());
      PGN_HIP(hipMemcpyAsync(hf.data(), R->Hb.p, hf.size() * sizeof(float),
                             hipMemcpyDeviceToHost, R->stream()));
      R->sync();
      std::copy(hf.begin(), hf.end(), h.begin());
    } else {
      PGN_HIP(hipMemcpyAsync(h.data(), R->Hb.p, h.size() * sizeof(double),
                             hipMemcpyDeviceToHost, R->stream()));
    }
    R->sync();
    for (long v = 0; v < R->nv; ++v) {
      const bool has = R->hflags[v] & ROD_HAS_BT;
      if (E) E[v] = has ? e[v] : 0.0;
      if (g8)
        for (int i = 0; i < 8; ++i) g8[8 * v + i] = has ? g[8 * v + i] : 0.0;
      if (H64) {
        int k = 0;
        for (int i = 0; i < 8; ++i)
          for (int j = 0; j <= i; ++j, ++k) {
            const double val = has ? h[36 * v + k] : 0.0;
            H64[64 * v + 8 * i + j] = val;
            H64[64 * v + 8 * j + i] = val;
          }
      }
    }
  });
}

int pgn_rods_newton(pgn_rods* R, const pgn_newton_opts* opts, pgn_newton_result* res) {
  return rguard([&] {
    PGN_CHECK(R && res, "null argument");
    pgn_newton_opts o;
    pgn_newton_default_opts(&o);
    if (opts) o = *opts;
    R->newton(o, res);
  });
}

}  // extern "C"
