// rod_model.hpp -- device-resident batch of discrete elastic rods.
#pragma once

#include <vector>

#include "pgn_internal.hpp"
#include "reduce.hpp"
#include "rods.hpp"

struct pgn_rods {
  pgn_ctx* ctx = nullptr;
  int nrods = 0;
  long nv = 0;
  int prec = PGN_FP64;
  std::vector<long> hvoff;
  std::vector<int> hrod_of;
  std::vector<uint8_t> hflags, hfixmask;
  std::vector<double> hframe0, hx0;
  pgn::DBuf<long> voff;
  pgn::DBuf<int> rod_of;
  pgn::DBuf<int> ldlt_order;  // rods by decreasing length: the LDL^T workgroup order
  pgn::DBuf<uint8_t> flags, fixmask, sel;
  pgn::DBuf<double> x, L0, kst, btc, frame, rt, mass;
  pgn::DBuf<double> es, gs, Js, eb, gb, g, vE, rodE, band, work, xsave, dx, rhs, fac, rn2;
  pgn::DBuf<char> Hb;
  pgn::DBuf<int> status;
  bool has_g = false;
  double g_ini[3] = {0, 0, 0}, g_end[3] = {0, 0, 0};
  int g_stages = 1, g_cur = 0;

  void setup(int n_rods, const int* nverts, const double* X0, const double* nor0,
             const double* young, const double* shear, const double* rh, const double* rw,
             const double* rho, int precision);
  hipStream_t stream() const { return ctx->stream; }
  void sync();
  pgn::RodArgs args() const;
  pgn::RodBufs bufs();
  // live (nrods) or null: only the marked rods are evaluated, the others'
  // per-rod outputs keep their values
  void eval(bool hess, const std::vector<unsigned char>* live = nullptr);
  pgn::DBuf<uint8_t> livem;
  void download_rodE(std::vector<double>& out);
  long ref_dofs() const;
  void to_ref(const double* in4, double* out) const;
  void from_ref(const double* in, double* out4) const;
  long ref_index(int r, long i4) const;
  void newton(const pgn_newton_opts& o, pgn_newton_result* R);
};
