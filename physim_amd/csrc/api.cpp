// api.cpp -- the extern "C" boundary of libpgn (declared in include/pgn.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <new>

#include "solver.hpp"

using namespace pgn;

namespace {

thread_local std::string g_err;

template <class F>
int guard(F&& f) {
  try {
    f();
    return PGN_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return PGN_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return PGN_ERR_INVALID;
  }
}

void use_device(const pgn_ctx* c) { PGN_HIP(hipSetDevice(c->device)); }

}  // namespace

namespace pgn {
void csr_scratch_setup(int device);  // csr_solve.hip
// error reporting shared by every extern "C" translation unit
int api_guard_set(const char* msg, int code) {
  g_err = msg;
  return code;
}
}  // namespace pgn

namespace {

// unpack the 78-value element Hessian into a full 12x12 row-major matrix
template <class T>
void unpack_tet_h(const T* h, double* H) {
  for (int a = 1; a < 4; ++a)
    for (int b = 0; b < a; ++b) {
      const T* o = h + 9 * (a * (a - 1) / 2 + b);
      for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
          H[(3 * a + i) * 12 + 3 * b + k] = o[3 * i + k];
          H[(3 * b + k) * 12 + 3 * a + i] = o[3 * i + k];
        }
    }
  for (int a = 0; a < 4; ++a) {
    const T* d = h + 54 + 6 * a;
    const int idx[3][3] = {{0, 1, 3}, {1, 2, 4}, {3, 4, 5}};
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) H[(3 * a + i) * 12 + 3 * a + k] = d[idx[i][k]];
  }
}

template <class T>
void bsr_to_layout(const pgn_model* M, const T* bv, int layout, double* out) {
  const BsrPattern& P = M->pat;
  if (layout == PGN_LAYOUT_BSR3) {
    for (long q = 0; q < 9 * P.nb; ++q) out[q] = (double)bv[q];
    return;
  }
  long w = 0;
  for (int i = 0; i < P.n; ++i)
    for (int p = 0; p < 3; ++p)
      for (long s = P.rowptr[i]; s < P.rowptr[i + 1]; ++s) {
        if (layout == PGN_LAYOUT_LOWER && P.bcol[s] > i) continue;
        for (int q = 0; q < 3; ++q) out[w++] = (double)bv[9 * s + 3 * p + q];
      }
}

long layout_nnz(const pgn_model* M, int layout) {
  const BsrPattern& P = M->pat;
  if (layout == PGN_LAYOUT_BSR3) return P.nb;
  if (layout == PGN_LAYOUT_FULL) return 9 * P.nb;
  long c = 0;
  for (int i = 0; i < P.n; ++i)
    for (long s = P.rowptr[i]; s < P.rowptr[i + 1]; ++s)
      if (P.bcol[s] <= i) c += 9;
  return c;
}

}  // namespace

extern "C" {

const char* pgn_last_error(void) { return g_err.c_str(); }

int pgn_ctx_create(int device, pgn_ctx** out) {
  return guard([&] {
    PGN_CHECK(out, "null output pointer");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      throw Error(PGN_ERR_NODEVICE, "no HIP device available: libpgn needs an MI355X (gfx950)");
    PGN_CHECK(device >= 0 && device < count, "device index out of range");
    PGN_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    PGN_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      throw Error(PGN_ERR_NODEVICE, std::string("libpgn is built for gfx950, device is ") +
                                        prop.gcnArchName);
    csr_scratch_setup(device);  // the allocation-free CSR solve's work space
    auto* c = new pgn_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      throw Error(PGN_ERR_HIP, "hipStreamCreate failed");
    }
    *out = c;
  });
}

void pgn_ctx_destroy(pgn_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
}

void* pgn_ctx_stream(pgn_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int pgn_ctx_synchronize(pgn_ctx* ctx) {
  return guard([&] {
    PGN_CHECK(ctx, "null context");
    PGN_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int pgn_model_create_tet(pgn_ctx* ctx, int n_nodes, const double* X0, int n_tets,
                         const int* T, const unsigned char* material,
                         const double* lame1, const double* lame2,
                         const double* density, int precision, pgn_model** out) {
  return guard([&] {
    PGN_CHECK(ctx && out, "null context or output pointer");
    *out = nullptr;
    use_device(ctx);
    auto* M = new pgn_model();
    M->ctx = ctx;
    try {
      for (int e = 0; e < n_tets && material; ++e)
        PGN_CHECK(material[e] <= 1, "unknown material kind at element " + std::to_string(e));
      M->setup(n_nodes, X0, n_tets, T, material, lame1, lame2, density, nullptr, precision);
    } catch (...) {
      delete M;
      throw;
    }
    *out = M;
  });
}

int pgn_model_create_shell(pgn_ctx* ctx, int n_nodes, const double* X0, int n_tris,
                           const int* F, const double* lame1, const double* lame2,
                           const double* thickness, const double* density, int n_hinges,
                           const int* hinges, const double* bending_k, int precision,
                           pgn_model** out) {
  return guard([&] {
    PGN_CHECK(ctx && out, "null context or output pointer");
    PGN_CHECK(n_tris >= 0 && n_hinges >= 0, "negative element count");
    PGN_CHECK((n_tris == 0 || (F && lame1 && lame2 && thickness && density)) &&
                  (n_hinges == 0 || (hinges && bending_k)),
              "null input array");
    *out = nullptr;
    use_device(ctx);
    const long m = (long)n_tris + n_hinges;
    std::vector<int> T(4 * m);
    std::vector<unsigned char> kind(m);
    std::vector<double> l1(m, 0.0), l2(m, 0.0), rho(m, 0.0), extra(m, 0.0);
    for (long t = 0; t < n_tris; ++t) {
      for (int a = 0; a < 3; ++a) T[4 * t + a] = F[3 * t + a];
      T[4 * t + 3] = F[3 * t + 2];
      kind[t] = 2;
      l1[t] = lame1[t];
      l2[t] = lame2[t];
      rho[t] = density[t];
      extra[t] = thickness[t];
    }
    for (long h = 0; h < n_hinges; ++h) {
      const long e = n_tris + h;
      for (int a = 0; a < 4; ++a) T[4 * e + a] = hinges[4 * h + a];
      kind[e] = 3;
      extra[e] = bending_k[h];
    }
    auto* M = new pgn_model();
    M->ctx = ctx;
    try {
      M->setup(n_nodes, X0, (int)m, T.data(), kind.data(), l1.data(), l2.data(), rho.data(),
               extra.data(), precision);
    } catch (...) {
      delete M;
      throw;
    }
    *out = M;
  });
}

void pgn_model_destroy(pgn_model* model) {
  if (!model) return;
  (void)hipSetDevice(model->ctx->device);
  (void)hipStreamSynchronize(model->ctx->stream);
  delete model;
}

int pgn_model_num_dofs(const pgn_model* model) { return model ? 3 * model->n : 0; }
int pgn_model_num_elements(const pgn_model* model) { return model ? model->m : 0; }

size_t pgn_model_device_bytes(const pgn_model* M) {
  if (!M) return 0;
  // every device allocation the model holds (its matrix copies, PCG, Newton,
  // time-step and direct-solver buffers included)
  return sum_bytes(M->X0, M->x, M->tets, M->kind, M->prm64, M->prm32, M->mass, M->fixmask,
                   M->inc_ptr, M->inc, M->rowptr, M->cptr, M->diag, M->bcol, M->brow,
                   M->code, M->bval, M->dinv, M->binv, M->eg64, M->eg32, M->eH, M->eH_alt,
                   M->blockE, M->nodeP, M->scal, M->g, M->f_ini, M->f_end, M->pv, M->pb,
                   M->part1, M->part2, M->partb, M->red64, M->cgs, M->sell_so, M->sell_len,
                   M->sell_col, M->sell_c16, M->sell_val, M->sc, M->r64, M->tmp64a,
                   M->tmp64b, M->nw_xsave, M->nw_v0, M->nw_dx, M->nw_b, M->lb_xp, M->lb_gp,
                   M->lb_S, M->lb_Y, M->vel, M->xhat,
                   M->xstep0, M->ldlt_prob, M->ldlt_perm, M->ldlt_iperm, M->ldlt_nprob,
                   M->ldlt_list, M->ldlt_status, M->ldlt_band, M->ldlt_D, M->ldlt_sinT,
                   M->ldlt_y, M->cg_agg_ptr, M->cg_agg_nodes, M->cg_agg_of, M->cg_cb_a,
                   M->cg_cb_b, M->cg_pair_piece0, M->cg_piece_lo, M->cg_cb_fine, M->cg_rel,
                   M->cg_cmask, M->cg_A, M->cg_D, M->cg_piece_sum, M->cg_rc, M->cg_uc,
                   M->cg_cpart, M->cg_cdot, M->cg_ok);
}

int pgn_model_set_fixed(pgn_model* M, const unsigned char* mask) {
  return guard([&] {
    PGN_CHECK(M && mask, "null argument");
    use_device(M->ctx);
    M->hfix.assign(mask, mask + 3L * M->n);
    std::vector<uint8_t> fm(M->n);
    for (int i = 0; i < M->n; ++i)
      fm[i] = (mask[3L * i] ? 1 : 0) | (mask[3L * i + 1] ? 2 : 0) | (mask[3L * i + 2] ? 4 : 0);
    M->fixmask.upload(fm, M->stream());
    M->sync();
    M->hess_valid = false;
    M->coarse_ready = false;  // the coarse columns depend on the free DoFs
  });
}

int pgn_model_set_gravity(pgn_model* M, const double g_ini[3], const double g_end[3],
                          int num_stages) {
  return guard([&] {
    PGN_CHECK(M && g_ini && g_end, "null argument");
    PGN_CHECK(num_stages >= 1, "num_stages must be >= 1");
    M->has_g = true;
    std::copy(g_ini, g_ini + 3, M->g_ini);
    std::copy(g_end, g_end + 3, M->g_end);
    M->g_stages = num_stages;
    M->g_cur = 0;
  });
}

int pgn_model_set_forces(pgn_model* M, const double* f_ini, const double* f_end,
                         int num_stages) {
  return guard([&] {
    PGN_CHECK(M && f_ini && f_end, "null argument");
    PGN_CHECK(num_stages >= 1, "num_stages must be >= 1");
    use_device(M->ctx);
    M->f_ini.upload(f_ini, 3L * M->n, M->stream());
    M->f_end.upload(f_end, 3L * M->n, M->stream());
    M->sync();
    M->has_f = true;
    M->f_stages = num_stages;
    M->f_cur = 0;
  });
}

int pgn_model_step_load(pgn_model* M, int* stepped) {
  return guard([&] {
    PGN_CHECK(M, "null model");
    const bool s = M->step_load();
    if (stepped) *stepped = s ? 1 : 0;
  });
}

int pgn_model_is_loaded(const pgn_model* M) { return M && M->is_loaded() ? 1 : 0; }

int pgn_model_reset(pgn_model* M) {
  return guard([&] {
    PGN_CHECK(M, "null model");
    use_device(M->ctx);
    PGN_HIP(hipMemcpyAsync(M->x.p, M->X0.p, M->X0.bytes(), hipMemcpyDeviceToDevice,
                           M->stream()));
    if (M->vel.p)
      PGN_HIP(hipMemsetAsync(M->vel.p, 0, M->vel.bytes(), M->stream()));
    M->sync();
    M->g_cur = 0;
    M->f_cur = 0;
    M->hess_valid = false;
    M->dyn_v0_valid = false;
  });
}

int pgn_model_set_x(pgn_model* M, const double* x) {
  return guard([&] {
    PGN_CHECK(M && x, "null argument");
    use_device(M->ctx);
    M->x.upload(x, 3L * M->n, M->stream());
    M->sync();
    M->hess_valid = false;
  });
}

int pgn_model_get_x(pgn_model* M, double* x) {
  return guard([&] {
    PGN_CHECK(M && x, "null argument");
    use_device(M->ctx);
    M->x.download(x, 3L * M->n, M->stream());
    M->sync();
  });
}

int pgn_model_set_velocity(pgn_model* M, const double* v) {
  return guard([&] {
    PGN_CHECK(M && v, "null argument");
    use_device(M->ctx);
    M->vel.upload(v, 3L * M->n, M->stream());
    M->sync();
  });
}

int pgn_model_get_velocity(pgn_model* M, double* v) {
  return guard([&] {
    PGN_CHECK(M && v, "null argument");
    use_device(M->ctx);
    const long N = 3L * M->n;
    if (M->vel.n == (size_t)N) {
      M->vel.download(v, N, M->stream());
      M->sync();
    } else {
      std::fill(v, v + N, 0.0);  // never set: at rest
    }
  });
}

int pgn_model_get_energy(pgn_model* M, double* energy) {
  return guard([&] {
    PGN_CHECK(M && energy, "null argument");
    use_device(M->ctx);
    double gg;
    M->eval_eg(*energy, gg);
  });
}

int pgn_model_get_gradient(pgn_model* M, double* gradient) {
  return guard([&] {
    PGN_CHECK(M && gradient, "null argument");
    use_device(M->ctx);
    double E, gg;
    M->eval_eg(E, gg);
    M->g.download(gradient, 3L * M->n, M->stream());
    M->sync();
  });
}

int pgn_model_hessian_nnz(pgn_model* M, int layout, long* nnz) {
  return guard([&] {
    PGN_CHECK(M && nnz, "null argument");
    PGN_CHECK(layout >= 0 && layout <= 2, "unknown layout");
    *nnz = layout_nnz(M, layout);
  });
}

int pgn_model_hessian_pattern(pgn_model* M, int layout, long* rowptr, int* col) {
  return guard([&] {
    PGN_CHECK(M && rowptr && col, "null argument");
    PGN_CHECK(layout >= 0 && layout <= 2, "unknown layout");
    const BsrPattern& P = M->pat;
    if (layout == PGN_LAYOUT_BSR3) {
      std::copy(P.rowptr.begin(), P.rowptr.end(), rowptr);
      std::copy(P.bcol.begin(), P.bcol.end(), col);
      return;
    }
    long w = 0;
    rowptr[0] = 0;
    for (int i = 0; i < P.n; ++i)
      for (int p = 0; p < 3; ++p) {
        for (long s = P.rowptr[i]; s < P.rowptr[i + 1]; ++s) {
          if (layout == PGN_LAYOUT_LOWER && P.bcol[s] > i) continue;
          for (int q = 0; q < 3; ++q) col[w++] = 3 * P.bcol[s] + q;
        }
        rowptr[3 * i + p + 1] = w;
      }
  });
}

int pgn_model_get_hessian(pgn_model* M, int layout, double* values) {
  return guard([&] {
    PGN_CHECK(M && values, "null argument");
    PGN_CHECK(layout >= 0 && layout <= 2, "unknown layout");
    use_device(M->ctx);
    M->eval_hessian();
    const long nv = 9 * M->pat.nb;
    if (M->hfp32) {
      std::vector<float> h(nv);
      PGN_HIP(hipMemcpyAsync(h.data(), M->bval.p, nv * sizeof(float), hipMemcpyDeviceToHost,
                             M->stream()));
      M->sync();
      bsr_to_layout(M, h.data(), layout, values);
    } else {
      std::vector<double> h(nv);
      PGN_HIP(hipMemcpyAsync(h.data(), M->bval.p, nv * sizeof(double), hipMemcpyDeviceToHost,
                             M->stream()));
      M->sync();
      bsr_to_layout(M, h.data(), layout, values);
    }
  });
}

int pgn_model_hessian_product(pgn_model* M, int matrix_free, const double* p, double* y) {
  return guard([&] {
    PGN_CHECK(M && p && y, "null argument");
    PGN_CHECK(!matrix_free || (M->m_mem == 0 && M->m_hinge == 0),
              "the matrix-free product covers tet models");
    use_device(M->ctx);
    M->eval_hessian();
    hipStream_t s = M->stream();
    const long N = 3L * M->n;
    const size_t ts = M->tsize();
    char* pin = M->pv.p + 3 * N * ts;   // the PCG's p0 and q slots
    char* qout = M->pv.p + 5 * N * ts;
    if (M->hfp32) {
      PGN_HIP(hipMemcpyAsync(M->tmp64a.p, p, N * sizeof(double), hipMemcpyHostToDevice, s));
      launch_cast_f64_to_f32(s, N, M->tmp64a.p, reinterpret_cast<float*>(pin));
    } else {
      PGN_HIP(hipMemcpyAsync(pin, p, N * sizeof(double), hipMemcpyHostToDevice, s));
    }
    if (matrix_free)
      launch_matfree_product(s, M->hfp32, M->mf_args(), pin, qout, nullptr,
                             pcg_grid1(M->n, M->pcg_lpr));
    else
      launch_spmv(s, M->hfp32, M->n, M->rowptr.p, M->bcol.p, M->bval.p, pin, qout, M->pcg_lpr);
    PGN_HIP(hipGetLastError());
    if (M->hfp32) {
      std::vector<float> h(N);
      PGN_HIP(hipMemcpyAsync(h.data(), qout, N * sizeof(float), hipMemcpyDeviceToHost, s));
      M->sync();
      for (long k = 0; k < N; ++k) y[k] = h[k];
    } else {
      PGN_HIP(hipMemcpyAsync(y, qout, N * sizeof(double), hipMemcpyDeviceToHost, s));
      M->sync();
    }
  });
}

int pgn_model_coarse_space(pgn_model* M, int* n_agg, int* agg_of, unsigned char* cmask,
                           double* coarse, double* coarse_inv) {
  return guard([&] {
    PGN_CHECK(M && n_agg, "null argument");
    use_device(M->ctx);
    M->coarse_prepare();
    const int na = M->cga.na;
    *n_agg = na;
    hipStream_t s = M->stream();
    if (agg_of) std::copy(M->h_agg_of.begin(), M->h_agg_of.end(), agg_of);
    if (cmask) {
      M->cg_cmask.download(cmask, na, s);
      M->sync();
    }
    if (!coarse && !coarse_inv) return;
    M->eval_hessian();
    const int nc = 6 * na, np = M->cga.npad;
    std::vector<double> h((size_t)np * np);
    auto take = [&](const double* dev, double* out, double sign) {
      PGN_HIP(hipMemcpyAsync(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      M->sync();
      for (int i = 0; i < nc; ++i)
        for (int j = 0; j < nc; ++j) out[(size_t)i * nc + j] = sign * h[(size_t)i * np + j];
    };
    if (coarse) take(M->coarse_update(true), coarse, 1.0);
    if (coarse_inv) {
      const double* R = M->coarse_update(false);
      int ok = 0;
      PGN_HIP(hipMemcpyAsync(&ok, M->cg_ok.p, sizeof(int), hipMemcpyDeviceToHost, s));
      M->sync();
      PGN_CHECK(ok, "coarse operator not positive definite");
      take(R, coarse_inv, -1.0);  // R holds -Ac^-1
    }
  });
}

int pgn_model_precond_apply(pgn_model* M, int precond, const double* r, double* z) {
  return guard([&] {
    PGN_CHECK(M && r && z, "null argument");
    PGN_CHECK(precond >= 0 && precond <= 2, "precond: 0, 1 or 2");
    use_device(M->ctx);
    M->eval_hessian();
    if (precond == 2) {
      M->coarse_prepare();
      M->coarse_update();
    }
    hipStream_t s = M->stream();
    const long N = 3L * M->n;
    const size_t ts = M->tsize();
    PGN_HIP(hipMemcpyAsync(M->tmp64a.p, r, N * sizeof(double), hipMemcpyHostToDevice, s));
    launch_narrow_or_copy(s, M->hfp32, N, M->tmp64a.p, M->pb.p);
    // the PCG's init from x = 0: r_pcg = r, z = B^-1 r, then the coarse part
    const PcgArgs a = M->pcg_args(0.0, precond);
    launch_pcg_init(s, M->hfp32, a, M->pb.p, 1);
    if (a.coarse) launch_coarse_apply(s, M->hfp32, a);
    PGN_HIP(hipGetLastError());
    launch_widen(s, M->hfp32, N, M->pv.p + 2 * N * ts, M->tmp64b.p);
    PGN_HIP(hipMemcpyAsync(z, M->tmp64b.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
    M->sync();
  });
}

int pgn_model_element_egh(pgn_model* M, int precision, double* E, double* g, double* H) {
  return guard([&] {
    PGN_CHECK(M, "null model");
    PGN_CHECK(precision == PGN_FP64 || precision == PGN_MIXED, "unknown precision");
    use_device(M->ctx);
    const bool f32 = precision == PGN_MIXED;
    const long m = M->m;
    if (m == 0) return;
    hipStream_t s = M->stream();
    DBuf<double> dE;
    dE.alloc(m);
    void* eg;
    void* eH;
    if (f32) {
      if (M->eg32.n < (size_t)(12 * m)) M->eg32.alloc(12 * m);
      eg = M->eg32.p;
    } else {
      eg = M->eg64.p;
    }
    const size_t ts = f32 ? sizeof(float) : sizeof(double);
    if (f32 == M->hfp32) {
      eH = M->eH.p;
    } else {
      if (M->eH_alt.n < (size_t)TET_H_PACK * m * ts) M->eH_alt.alloc((size_t)TET_H_PACK * m * ts);
      eH = M->eH_alt.p;
    }
    M->launch_elements(f32, EVAL_E | EVAL_G | EVAL_H, nullptr, dE.p, eg, eH);
    PGN_HIP(hipGetLastError());
    if (E) dE.download(E, m, s);
    std::vector<char> hg(12 * m * ts), hh((size_t)TET_H_PACK * m * ts);
    if (g) PGN_HIP(hipMemcpyAsync(hg.data(), eg, hg.size(), hipMemcpyDeviceToHost, s));
    if (H) PGN_HIP(hipMemcpyAsync(hh.data(), eH, hh.size(), hipMemcpyDeviceToHost, s));
    M->sync();
    for (long e = 0; e < m; ++e) {
      if (g)
        for (int q = 0; q < 12; ++q)
          g[12 * e + q] = f32 ? (double)reinterpret_cast<float*>(hg.data())[12 * e + q]
                              : reinterpret_cast<double*>(hg.data())[12 * e + q];
      if (H) {
        if (f32)
          unpack_tet_h(reinterpret_cast<float*>(hh.data()) + TET_H_PACK * e, H + 144 * e);
        else
          unpack_tet_h(reinterpret_cast<double*>(hh.data()) + TET_H_PACK * e, H + 144 * e);
      }
    }
    M->hess_valid = M->hess_valid && (f32 != M->hfp32);
  });
}

void pgn_linear_default_opts(pgn_linear_opts* o) {
  if (!o) return;
  o->tol = 1e-9;        // OptimSolverOptions::lsTolMaxError
  o->max_iters = 1000;  // lsNumMaxIters
  o->reg_iters = 10;    // LinearSolverOptions::regIters
  o->precond = 0;
  o->inner_tol = 1e-5;
  o->solver = PGN_LS_EIGENCG;
}

int pgn_linear_solve(pgn_model* M, const double* b, double* x, const pgn_linear_opts* opts,
                     pgn_linear_result* res) {
  return guard([&] {
    PGN_CHECK(M && b && x && res, "null argument");
    use_device(M->ctx);
    pgn_linear_opts o;
    pgn_linear_default_opts(&o);
    if (opts) o = *opts;
    PGN_CHECK(o.max_iters >= 0 && o.tol >= 0.0, "bad linear options");
    PGN_CHECK(o.solver == PGN_LS_EIGENCG || o.solver == PGN_LS_EIGENLDLT,
              "linear solver: PGN_LS_EIGENCG or PGN_LS_EIGENLDLT");
    M->eval_hessian();
    const long N = 3L * M->n;
    M->nw_b.alloc(N);
    M->nw_dx.alloc(N);
    M->nw_b.upload(b, N, M->stream());
    M->nw_dx.upload(x, N, M->stream());
    linear_solve(*M, M->nw_b.p, M->nw_dx.p, o, *res);
    M->nw_dx.download(x, N, M->stream());
    M->sync();
    M->hess_valid = false;  // regularization may have modified it
  });
}

void pgn_newton_default_opts(pgn_newton_opts* o) {
  if (!o) return;
  // OptimSolverOptions() (OptimSolver.h:47-70)
  o->tolMaxError = 1e-6;
  o->tolMinStep = 1e-5;
  o->tolMinImpr = 1e-9;
  o->minStepMaxCount = 3;
  o->minImprMaxCount = 3;
  o->numMaxIters = 100;
  o->maxStepSize = 1000;
  o->trySteepest = 1;
  o->lsearch_numTry = 32;
  o->lSearch_alpha = 0.5;
  o->lsNumMaxIters = 1000;
  o->lsTolMaxError = 1e-9;
  o->lsNumRegTrials = 10;
  o->precond = 0;
  o->inner_tol = 1e-5;
  o->adaptive_tol = 0;
  o->ls_noise_guard = 0.0;
  o->ew_gamma = 0.9;
  o->ew_max = 0.2;
  o->lsSolverType = PGN_LS_EIGENCG;
  o->qpSolverType = PGN_QP_NEWTON;
  o->coarse_giveup = 0;
}

int pgn_newton_solve(pgn_model* M, const pgn_newton_opts* opts, pgn_newton_result* res) {
  return guard([&] {
    PGN_CHECK(M && res, "null argument");
    use_device(M->ctx);
    pgn_newton_opts o;
    pgn_newton_default_opts(&o);
    if (opts) o = *opts;
    newton_solve(*M, o, *res);
  });
}

int pgn_dynamic_step(pgn_model* M, double dt, const pgn_newton_opts* opts,
                     pgn_newton_result* res) {
  return guard([&] {
    PGN_CHECK(M && res, "null argument");
    PGN_CHECK(dt > 0.0 && std::isfinite(dt), "time step must be positive and finite");
    use_device(M->ctx);
    pgn_newton_opts o;
    pgn_newton_default_opts(&o);
    if (opts) o = *opts;
    dynamic_step(*M, dt, o, *res);
  });
}

int pgn_model_set_profiling(pgn_model* M, int on) {
  return guard([&] {
    PGN_CHECK(M, "null model");
    M->profiling = on != 0;
    if (on) {
      M->prof_n[0] = M->prof_n[1] = 0;
      M->prof_ms[0] = M->prof_ms[1] = 0.0;
    }
  });
}

int pgn_model_get_profile(pgn_model* M, int which, long* count, double* total_ms) {
  return guard([&] {
    PGN_CHECK(M && count && total_ms, "null argument");
    PGN_CHECK(which >= 0 && which <= 8 + TET_TILES, "unknown profile slot");
    if (which >= 6) {  // element kernel tiles: 6 = Hessian (model precision),
                       // 7 = FP64 E+g, 8 = E+g+H (model precision); 9.. = the
                       // timings (us) of the Hessian's tiles
      const int fp = M->hfp32 ? 1 : 0;
      if (which >= 9) {
        *count = which - 9;
        *total_ms = M->tet_tile_us[fp][EVAL_H][which - 9];
        return;
      }
      const int slot[3][2] = {{fp, EVAL_H}, {0, EVAL_E | EVAL_G}, {fp, EVAL_E | EVAL_G | EVAL_H}};
      const int t = M->tet_tile_choice[slot[which - 6][0]][slot[which - 6][1]];
      *count = t;  // -1: not launched yet
      *total_ms = t < 0 ? 0.0 : M->tet_tile_us[slot[which - 6][0]][slot[which - 6][1]][t];
    } else if (which <= 1) {
      *count = M->prof_n[which];
      *total_ms = M->prof_ms[which];
    } else if (which == 5) {  // bytes per SELL-64 block column, 0 without SELL
      if (M->pcg_cg1 && M->use_sell) M->sell_prepare();
      *count = !(M->pcg_cg1 && M->use_sell) ? 0 : (M->sell_c16.p ? 2 : 4);
      *total_ms = 0.0;
    } else {  // SpMV us with 4 / 8 / 16 lanes per row; count = width in use
      *count = M->pcg_lpr;
      *total_ms = M->tune_us[which - 2];
    }
  });
}

int pgn_kernel_bytes(pgn_model* M, int which, double* bytes) {
  return guard([&] {
    PGN_CHECK(M && bytes, "null argument");
    const double m = M->m, n = M->n, nb = (double)M->pat.nb;
    const double nc = (double)M->pat.code.size();
    const double hs = M->hfp32 ? 4.0 : 8.0;
    // compulsory traffic: every input read once, every output written once
    switch (which) {
      case 0:  // E+g+H, model precision: tets 16 + kind 1 + prm 12*hs, g 12*hs, H 78*hs
        *bytes = m * (17 + 12 * hs + 12 * hs + 78 * hs) + n * 48;
        break;
      case 1:  // E+g FP64
        *bytes = m * (17 + 96 + 96) + n * 48;
        break;
      case 2:  // BSR gather: codes, ranges, rows/cols, element H, blocks out
        *bytes = nc * 4 + nb * (8 + 8 + 9 * hs) + m * 78 * hs + n;
        break;
      case 3:  // SpMV: values, block cols, row pointers, x, y
        *bytes = nb * (9 * hs + 4) + (n + 1) * 8 + 2 * 3 * n * hs;
        break;
      case 4:  // E+g+H FP64
        *bytes = m * (17 + 96 + 96 + 624) + n * 48;
        break;
      case 5:  // E+g+H FP32
        *bytes = m * (17 + 48 + 48 + 312) + n * 48;
        break;
      case 7:  // calibration copy: 256 MiB read + 256 MiB written
        *bytes = 2.0 * 256.0 * 1024 * 1024;
        break;
      case 8:  // the PCG SpMV launch.  Fused K1: values + block cols + row
               // pointers, z and p_{k-1} read, p_k and q written.
               // Chronopoulos-Gear (default) and split: the plain SpMV,
               // values + cols + row pointers, one vector read, one written.
        // SELL-64 (the C-G default): values + cols of the rows' own blocks
        // (padding slots are not read), a 4 B length per row and an 8 B
        // offset per slice instead of the row pointers; z and q read, q
        // (= A z + beta q) written.  Columns take 2 B where the layout
        // stores them as 16-bit offsets from the row.
        if (M->pcg_cg1 && M->use_sell) {
          M->sell_prepare();
          const double cb = M->sell_c16.p ? 2.0 : 4.0;
          *bytes = nb * (9 * hs + cb) + n * 4.0 + (double)(((long)n + 63) / 64) * 8.0 +
                   3 * 3 * n * hs;
        }
        else if (M->pcg_split || M->pcg_cg1)
          *bytes = nb * (9 * hs + 4) + (n + 1) * 8 + 2 * 3 * n * hs;
        else
          *bytes = nb * (9 * hs + 4) + (n + 1) * 8 + 4 * 3 * n * hs;
        break;
      case 9:  // the PCG update launch with the block-Jacobi preconditioner
               // (6 values per node: the symmetric inverse's upper triangle).
               // Chronopoulos-Gear: z, p, w, q, x, r read, p, q, x, r, z
               // written; split and fused K2: p, q, x, r read, x, r, z written.
        // With the SELL-64 SpMV forming q: z, p, q, x, r read, p, x, r, z
        // written.
        *bytes = (M->pcg_cg1 ? (M->use_sell ? 9.0 : 11.0) : 7.0) * 3 * n * hs + 6.0 * n * hs;
        break;
      default:
        throw Error(PGN_ERR_INVALID, "unknown kernel id");
    }
  });
}

int pgn_time_kernel(pgn_model* M, int which, int reps, double* ms_mean) {
  return guard([&] {
    PGN_CHECK(M && ms_mean && reps > 0, "bad argument");
    use_device(M->ctx);
    hipStream_t s = M->stream();
    const long m = M->m;
    const size_t h64 = (size_t)TET_H_PACK * std::max<long>(m, 1) * sizeof(double);
    const size_t h32 = (size_t)TET_H_PACK * std::max<long>(m, 1) * sizeof(float);
    auto ensure_eH = [&](bool f32) -> void* {
      if (f32 == M->hfp32) return M->eH.p;
      const size_t need = f32 ? h32 : h64;
      if (M->eH_alt.n < need) M->eH_alt.alloc(need);
      return M->eH_alt.p;
    };
    auto ensure_eg32 = [&]() -> void* {
      if (M->eg32.n < (size_t)(12 * std::max<long>(m, 1))) M->eg32.alloc(12 * std::max<long>(m, 1));
      return M->eg32.p;
    };
    std::function<void()> launch;
    switch (which) {
      case 0: {
        const bool f = M->hfp32;
        void* eg = f ? ensure_eg32() : (void*)M->eg64.p;
        void* eH = M->eH.p;
        launch = [=] {
          M->launch_tet(f, EVAL_E | EVAL_G | EVAL_H, M->blockE.p, nullptr, eg, eH);
        };
        break;
      }
      case 1:
        launch = [=] {
          M->launch_tet(false, EVAL_E | EVAL_G, M->blockE.p, nullptr, M->eg64.p, nullptr);
        };
        break;
      case 2:
        M->eval_hessian();
        launch = [=] { launch_bsr_gather(s, M->hfp32, M->bsr_args(), M->eH.p, M->bval.p); };
        break;
      case 3: {
        M->eval_hessian();
        const size_t v = 3L * M->n * M->tsize();
        char* base = M->pv.p;
        launch = [=] {
          launch_spmv(s, M->hfp32, M->n, M->rowptr.p, M->bcol.p, M->bval.p, base + 4 * v,
                      base + 5 * v, M->pcg_lpr);
        };
        break;
      }
      case 4: {
        void* eH = ensure_eH(false);
        launch = [=] {
          M->launch_tet(false, EVAL_E | EVAL_G | EVAL_H, M->blockE.p, nullptr, M->eg64.p, eH);
        };
        break;
      }
      case 5: {
        void* eH = ensure_eH(true);
        void* eg = ensure_eg32();
        launch = [=] {
          M->launch_tet(true, EVAL_E | EVAL_G | EVAL_H, M->blockE.p, nullptr, eg, eH);
        };
        break;
      }
      case 7: {
        const long nf = 64L * 1024 * 1024;  // 256 MiB of floats
        auto buf = std::make_shared<DBuf<float>>();
        buf->alloc(2 * nf);
        launch = [=] { launch_copy_f32(s, nf, buf->p, buf->p + nf); };
        break;
      }
      default:
        throw Error(PGN_ERR_INVALID, "unknown kernel id");
    }
    launch();  // warm-up
    PGN_HIP(hipGetLastError());
    hipEvent_t e0, e1;
    PGN_HIP(hipEventCreate(&e0));
    PGN_HIP(hipEventCreate(&e1));
    PGN_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) launch();
    PGN_HIP(hipEventRecord(e1, s));
    PGN_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    PGN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_mean = (double)ms / reps;
    if (which == 2 || which == 3) M->hess_valid = false;
  });
}

int pgn_host_bsr_pattern(int n_nodes, int n_tets, const int* T, long* nb_out, long* rowptr,
                         int* col) {
  return guard([&] {
    PGN_CHECK(T && nb_out, "null argument");
    BsrPattern P;
    build_bsr_pattern(n_nodes, n_tets, T, P);
    *nb_out = P.nb;
    if (rowptr) std::copy(P.rowptr.begin(), P.rowptr.end(), rowptr);
    if (col) std::copy(P.bcol.begin(), P.bcol.end(), col);
  });
}

}  // extern "C"
