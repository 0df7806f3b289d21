// rods.hpp -- device data and launchers of the discrete-elastic-rod path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pgn {

constexpr int ROD_BW = 10;  // half-bandwidth of a rod Hessian (vertex-major DoFs)

enum : uint8_t {
  ROD_HAS_EDGE = 1,       // edge (v, v+1) exists
  ROD_HAS_BT = 2,         // bend-twist element centred at v exists
  ROD_HAS_PREV_EDGE = 4,  // edge (v-1, v) exists
};

struct RodBtConst {
  double u1[3], n1[3], u2[3], n2[3];
  double kb0, kb1, kt, K0, K1, tau0, rt, crt, srt;
};

struct RodArgs {
  long nv;              // vertices of all rods
  int nrods;
  const long* voff;     // nrods + 1 vertex offsets
  const int* rod_of;    // nv
  const uint8_t* flags; // nv
  const uint8_t* fixmask;  // nv: bit c => DoF 4v+c fixed
  const double* x;      // 4 nv
  const double* L0;     // nv (edge slot)
  const double* kst;    // nv (edge slot): E pi rh rw
  const double* btc;    // 6 nv (bend-twist slot): kb0 kb1 kt K0 K1 tau0
  const double* frame;  // 6 nv (edge slot): tangent, normal of the reference frame
  const double* rt;     // nv (bend-twist slot): reference twist
  const double* mass;   // nv
  double grav[3];
  const uint8_t* live = nullptr;  // nrods or null: rods evaluated (others keep their outputs)
};

struct RodBufs {
  double *es, *gs, *Js;   // stretch: energy, 3-gradient, 6-Hessian per edge slot
  double *eb, *gb;        // bend-twist: energy, 8-gradient per slot
  void* Hb;               // bend-twist: 36 (lower 8x8) per slot, FP64 or FP32
  double* g;              // 4 nv gradient
  double* vE;             // 3 nv: elastic, load energy, |g|^2 per vertex
  double* rodE;           // 4 nrods: E, |g|^2, energy scale
  double *band, *work;    // (ROD_BW+1) * 4 nv
  int* status;            // nrods
};

void rod_launch_eval(hipStream_t s, const RodArgs& a, bool hess, bool fp32_hess, RodBufs& b);
void rod_launch_band(hipStream_t s, const RodArgs& a, bool fp32_hess, RodBufs& b);
// order (nrods or null): the rod of each workgroup (longest first)
void rod_launch_ldlt(hipStream_t s, const RodArgs& a, RodBufs& b, const double* rhs, double* x,
                     int reg_iters, double max_error, const int* order,
                     const unsigned char* active);
// bend-twist elements on given inputs (27 parameters, e1, e2, two angles per
// element); H: 36 packed values per element in FP32 (fp32) or FP64
void rod_launch_bt_raw(hipStream_t s, long count, const double* prm, const double* e1,
                       const double* e2, const double* th, bool fp32, double* E, double* g,
                       void* H);
void rod_launch_kinematics(hipStream_t s, const RodArgs& a, double* frame, double* rt);
void rod_launch_step(hipStream_t s, const RodArgs& a, const double* fac, const double* xsave,
                     const double* dx, double* x);

}  // namespace pgn
