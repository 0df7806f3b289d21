// kernels.hpp -- host-callable launchers of libpgn's gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pgn {

enum { EVAL_E = 1, EVAL_G = 2, EVAL_H = 4 };

// Number of unique values of one packed tet Hessian: 6 off-diagonal 3x3
// node blocks (a > b, row-major) followed by 4 diagonal blocks stored as
// their lower triangles (6 values each).
constexpr int TET_H_PACK = 78;
constexpr int TET_BLOCK = 256;  // threads per element-kernel workgroup

// Element inputs of a tet mesh.  prm holds 12 SoA arrays of m values:
// rows b1, b2, b3 of Dm^-1 (9), V0, lame1, lame2.
struct TetArgs {
  int m;                 // tets (elements [0, m) of the model)
  long stride;           // SoA stride of prm: all elements of the model
  const int4* tets;
  const uint8_t* kind;
  const void* prm;       // float* or double* (matches the kernel precision)
  const double* x;       // 3n current positions
  const double* X0;      // 3n rest positions
};

// Element evaluation.  blockE: one FP64 partial energy per 64 consecutive
// elements (ceil(m / 64) entries) or null; elemE: per-element FP64 energy or
// null; eg: 12 values per element (precision T); eH: TET_H_PACK per element.
// tile: workgroup size x elements per lane, TET_TILES of them (0 = 256 x 1);
// every tile gives the same results bit for bit.
constexpr int TET_TILES = 6;
constexpr int TET_TILE_BS[TET_TILES] = {256, 128, 64, 256, 128, 64};
constexpr int TET_TILE_EPL[TET_TILES] = {1, 1, 1, 2, 2, 2};
void launch_tet_eval(hipStream_t s, bool fp32, int flags, const TetArgs& a,
                     double* blockE, double* elemE, void* eg, void* eH, int tile = 0);

// Thin-shell elements (membrane triangles, hinges) evaluated into the same
// per-element outputs as tets.  Elements [first, first + count) of the model;
// prm: 12 SoA arrays of `stride` values.  Membrane: prm rows 0-3 = Dm^-1
// (2x2 row-major), 9 = rest area x thickness, 10 = lambda, 11 = mu.  Hinge:
// 0 = kB, 1 = theta0, 2 = shape0.  blockE receives ceil(count / TET_BLOCK)
// partials.
enum { SHELL_MEMBRANE = 2, SHELL_HINGE = 3 };
struct ShellArgs {
  int first, count;
  long stride;
  const int4* elems;
  const void* prm;
  const double* prm64;   // the FP64 table (the rest dihedral is read from it)
  const double* x;
  const double* X0;
};
void launch_shell_eval(hipStream_t s, bool fp32, int kind, int flags, const ShellArgs& a,
                       double* blockE, double* elemE, void* eg, void* eH);

// Node gather of element gradients + lumped gravity + nodal forces, fixed
// DoFs zeroed; also node energy terms and |g|^2, reduced with the element
// energy partials into out2 = {E, |g|^2}.
struct NodeArgs {
  int n;
  const int* inc_ptr;    // n+1
  const int* inc;        // element*4 + local node, element order per node
  const double* mass;    // n lumped masses
  const uint8_t* fixmask;// n: bit k set => DoF 3i+k fixed
  const double* x;       // 3n
  double grav[3];
  const double* f_ini;   // 3n or null
  const double* f_end;   // 3n or null
  double f_alpha;
  // batches of independent instances: when node_inst is set, node i takes
  // gravity grav_inst[3 inst] and force ramp falpha_inst[inst] of its
  // instance, and its load energy and |g_i|^2 go to nodeE / nodeG2
  const int* node_inst = nullptr;
  const double* grav_inst = nullptr;
  const double* falpha_inst = nullptr;
  double* nodeE = nullptr;
  double* nodeG2 = nullptr;
  // implicit dynamics: inertia term m/(2 dt^2) |x_i - xhat_i|^2 of one
  // backward-Euler step toward the predictor xhat (inv_dt2 = 1/dt^2)
  const double* xhat = nullptr;
  double inv_dt2 = 0.0;
  // nodes [i0, i0 + n) of the model (a batch gathers its active instances'
  // node ranges alone)
  long i0 = 0;
};
void launch_node_gather(hipStream_t s, bool fp32_eg, const NodeArgs& a,
                        const void* eg, double* g, double* node_partials);
// Reduces n_elem_partials element partials and n_node_partials node partials
// (pairs {E, gg}) into out[0] = E, out[1] = |g|^2, out[2] = |E_elastic| +
// |E_load|.
void launch_energy_final(hipStream_t s, const double* elem_partials,
                         int n_elem_partials, const double* node_partials,
                         int n_node_partials, double* out2);
int node_gather_blocks(int n);

// BSR3 Hessian gather from packed element Hessians, with the fixed-DoF
// projection and the Jacobi / block-Jacobi preconditioner of the diagonal.
struct BsrArgs {
  long nb;
  const int* brow;       // nb: block row (node) of each block
  const int* bcol;       // nb
  const long* cptr;      // nb+1 contribution ranges
  const uint32_t* code;  // element*16 + 4*a + b
  const uint8_t* fixmask;
  // implicit dynamics: m_i / dt^2 added to the diagonal of block (i, i)
  const double* mass = nullptr;
  double inv_dt2 = 0.0;
  long s0 = 0;  // blocks [s0, s0 + nb) are gathered
};
void launch_bsr_gather(hipStream_t s, bool fp32, const BsrArgs& a,
                       const void* eH, void* bval);
// diag: n block indexes of the diagonal blocks.  Writes dinv (3n, Jacobi) and
// binv (6n, block-Jacobi inverse as upper triangles) in precision T.
// (nodes [i0, i0 + n))
void launch_precond(hipStream_t s, bool fp32, int n, const long* diag,
                    const void* bval, void* dinv, void* binv, long i0 = 0);
// bval[diag blocks] += shift on the scalar diagonal
void launch_add_diag(hipStream_t s, bool fp32, int n, const long* diag,
                     void* bval, double shift);
// Frobenius norm^2 of the reference's stored lower entry set, reduced into
// out[0] (partials: PCG_MAXP + 4 doubles, shared with residual / dot).
int frob_blocks();
void launch_frob_lower(hipStream_t s, bool fp32, long nb, const int* brow,
                       const int* bcol, const void* bval, double* partials,
                       double* out);

// ---- PCG ------------------------------------------------------------------
// Reductions are fence-free: each producer workgroup writes a partial sum,
// and every consumer workgroup of the next launch re-sums the partials in a
// fixed order (identical bits everywhere; kernel boundaries give coherence).
constexpr int PCG_MAXP = 2048;      // capacity: producer workgroups of one partial set
// cap on FP32 corrections per mixed-precision linear solve
constexpr int MAX_REFINEMENTS = 64;
constexpr int PCG_GRID_CAP = 1024;  // grid cap of the BSR3 SpMV / update / residual launches

struct PcgScalars {
  double thr;   // max(tol^2 |b|^2, DBL_MIN)
  double bb;    // |b|^2
  double rr;    // |r|^2 at the last check
  int done;
  int iters;    // Eigen's iteration count once done
  int pad[2];
};

// Matrix-free Hessian product over the stored packed element Hessians (tet
// models): y_e = H_e p_e per tet into ye (12 per tet), then per node the
// sum over its incidences with the fixed rows' identity and the diagonal
// shift (matfree_kernels.hip).
struct MatFreeArgs {
  long m;                // tets
  int n;                 // nodes
  const int4* tets;
  const void* eH;        // TET_H_PACK per tet, Hessian precision
  void* ye;              // 12 per tet, Hessian precision
  const int* inc_ptr;
  const int* inc;
  const uint8_t* fixmask;
  double shift;          // regularization shift on every diagonal entry
};
// q = H p (p, q: 3n in the Hessian precision); part: one p.q partial per
// workgroup of the node launch (grid of them, rows blocked per XCD as the
// PCG SpMV) or null
void launch_matfree_product(hipStream_t s, bool fp32, const MatFreeArgs& a, const void* p,
                            void* q, double* part, int grid);

// Two-level preconditioner (precond 2): block Jacobi plus a coarse
// correction, M^-1 r = B^-1 r + P Ac^-1 P^T r.  The coarse space is spanned,
// per aggregate of nodes (rest positions binned into boxes), by the three
// translations and the three infinitesimal rotations about the aggregate's
// centroid, restricted to the free DoFs: node i of aggregate a contributes
// P_i = F_i [I | W_i] with W_i w = w x rel_i (rel_i = X0_i - centroid_a) and
// F_i zeroing its fixed components.  Ac = P^T A P (6 x 6 blocks, FP64) is
// inverted densely on the device by block Gauss-Jordan sweeps; its inverse is
// applied to the restricted residual each PCG iteration (coarse_kernels.hip).
constexpr int COARSE_TILE = 32;     // Gauss-Jordan tile
constexpr int COARSE_MAX_PAD = 2048;  // coarse DoFs, padded to the tile, at most
constexpr int COARSE_PIECE = 2048;  // BSR3 blocks per coarse assembly workgroup at most
constexpr int COARSE_SPLIT = 4;     // restriction workgroups per aggregate
struct CoarseArgs {
  int na = 0;                  // aggregates
  int npad = 0;                // 6 na padded to a multiple of COARSE_TILE
  int gemv_g = 0;              // workgroups of the coarse product
  const int* agg_ptr = nullptr;   // na + 1: ranges into agg_nodes
  const int* agg_nodes = nullptr; // n: nodes ordered by aggregate, ascending
  const int* agg_of = nullptr;    // n
  const void* rel = nullptr;      // 3n, Hessian precision
  const uint8_t* cmask = nullptr; // na: active coarse columns (bits 0-5)
  const uint8_t* fixmask = nullptr;
  const double* R = nullptr;      // npad x npad: -Ac^-1
  const int* ok = nullptr;        // 0: the sweeps met a non-positive pivot
  double* rc = nullptr;           // COARSE_SPLIT x npad parts of P^T r (padding stays zero)
  double* uc = nullptr;           // npad: Ac^-1 P^T r
  double* cpart = nullptr;        // gemv_g partials of rc.uc
  double* cdot = nullptr;         // their sum: r.(P uc) of the last application
};

struct PcgArgs {
  int n;                 // block rows (nodes)
  const long* rowptr;    // n+1
  const int* bcol;
  const void* bval;
  const void* dinv;      // Jacobi (3n)
  const void* binv;      // block Jacobi (6n: upper triangles)
  int precond;           // 0 Jacobi, 1 block Jacobi
  int lpr;               // lanes per block row in the SpMV (4, 8 or 16)
  int split;             // K1 as two launches (p update, then plain SpMV)
  int cg1;               // Chronopoulos-Gear form: K1 = SpMV on z, K2 = every
                         // vector update (p, q, x, r, z); overrides split
  double* cgs;           // cg1 scalars: 2 parities x {alpha, r.z}, then {beta, r.z}
                         // of the current iteration (SELL form)
  int g1, g2;            // workgroups of K1/init and of K2
  double tol;
  void* x;
  void* r;
  void* z;
  void* p0;
  void* p1;
  void* q;
  double* part1;         // PCG_MAXP: p.q of K1
  double* part2;         // 2 parities x PCG_MAXP x {r.r, r.z}
  double* partb;         // PCG_MAXP: |b|^2 of init
  PcgScalars* sc;
  // SELL-64 copy of the matrix for the Chronopoulos-Gear SpMV: slices of 64
  // rows, one row per lane, slot (slice, j, lane) at sell_so[slice] + 64 j +
  // lane, the 9 values of a slot component-major per (slice, j) so that each
  // load of a wavefront reads 64 consecutive values
  int sell = 0;          // the SpMV launch uses it (g1 = its grid) and also
                         // forms beta and q = w + beta q for its rows (cgs[4],
                         // cgs[5] = beta, r.z for the update launch)
  int sell_spw = 0;      // slices per workgroup (a multiple of 8)
  long sell_ns = 0;      // slices
  const long* sell_so = nullptr;   // ns + 1 slot offsets
  const int* sell_len = nullptr;   // n: blocks of each row
  const int* sell_col = nullptr;   // per slot: block column (null when sell_c16 is set)
  const short* sell_c16 = nullptr; // per slot: block column minus the slot's row
  const void* sell_val = nullptr;  // 9 per slot, Hessian precision
  // Launches inside a captured chunk graph take their iteration relative to
  // the chunk: k = kctl[0] + argument, and the status launch's iteration
  // cap is kctl[1] (both written by the host before each replay).  Null:
  // the arguments are absolute.
  const int* kctl = nullptr;
  // split form only: the SpMV launch replaced by the matrix-free product
  int mf = 0;
  MatFreeArgs mfa;
  // Chronopoulos-Gear form only: the coarse correction follows every z =
  // B^-1 r (init and update launches); r.z adds cg.cdot
  int coarse = 0;
  CoarseArgs cg;
};
int pcg_grid1(int n, int lpr);
// SELL-64 layout: slices per workgroup and grid (<= PCG_MAXP) for ns slices
void sell_grid(long ns, int& spw, int& grid);
// copies the BSR3 values into the SELL-64 layout (brow: block row of each
// block; padding slots are left as they are)
void launch_sell_pack(hipStream_t s, bool fp32, long nb, const int* brow, const long* rowptr,
                      const long* so, const void* bval, void* sval);
int pcg_grid2(int n);
// r = b - A x (b in T); z = M^-1 r; p = 0.  x_zero: start from x = 0.
void launch_pcg_init(hipStream_t s, bool fp32, const PcgArgs& a, const void* b, int x_zero);
// iteration k: K1 (beta, p_k, q = A p_k, p.q; one launch, or two when split)
// and K2 (alpha, x, r, z, r.r, r.z)
// split mode: K1a (beta, p update) must precede K1 (then the plain SpMV)
void launch_pcg_k1a(hipStream_t s, bool fp32, const PcgArgs& a, int k);
// e0 / e1 (optional): events stamped with the launch's own start and end
void launch_pcg_k1(hipStream_t s, bool fp32, const PcgArgs& a, int k, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr);
void launch_pcg_k2(hipStream_t s, bool fp32, const PcgArgs& a, int k, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr);
// after iterations [0, k_next): convergence test, sets done / iters / rr
void launch_pcg_status(hipStream_t s, const PcgArgs& a, int k_next, int maxit);
// plain device copy of n floats (bandwidth calibration)
void launch_copy_f32(hipStream_t s, long n, const float* a, float* b);

// ---- segmented PCG: one independent PCG per instance of a batch --------------
// Segment k (an instance) owns block rows [row0[k], row1[k]) and workgroups
// [wg0[k], wg0[k] + G[k]) of every launch (the WG table is wg_seg / wg_loc);
// each segment has its own tolerance, scalars (alpha, beta from its own
// partial sums) and convergence state, and stops on its own.
// Coarse level of the segmented PCG (precond 2 in a batch): each instance
// has its own aggregates (a contiguous range of the global ids, never
// spanning two instances) and its own dense coarse operator, so the coarse
// correction stays per instance like every other part of the iteration.
constexpr int SEG_COARSE_MAXAGG = 24;  // aggregates per instance at most
struct SegCoarseArgs {
  int NA;                   // aggregates over all instances
  const int* agg_ptr;       // NA + 1: ranges into agg_nodes
  const int* agg_nodes;     // n: nodes ordered by aggregate, ascending
  const int* agg_of;        // n: global aggregate of each node
  const int* agg_inst;      // NA: instance of each aggregate
  const int* inst_aoff;     // K + 1: aggregate range of each instance
  const long* inst_moff;    // K: offset of each instance's dense nc x nc operator
  const void* rel;          // 3n, Hessian precision: node offset from its centroid
  const uint8_t* cmask;     // NA: active coarse columns (bits 0-5)
  const uint8_t* fixmask;   // n
  const double* R;          // per instance -Ac^-1 (nc = 6 x its aggregates)
  const int* ok;            // K: 0 when the instance's sweep met a non-positive pivot
  double* rc;               // 6 NA: P^T r
  double* uc;               // 6 NA: Ac^-1 P^T r
  double* cdot;             // 2 x K: r.(P uc) by iteration parity
  const uint8_t* use;       // K or null: instances with the coarse level (others: uc = 0)
};
struct SegPcgArgs {
  int K;                 // segments
  int W;                 // workgroups of every launch (sum of G)
  const int* wg_seg;     // W: segment of each workgroup
  const int* wg_loc;     // W: index of the workgroup within its segment
  const int* seg_G;      // K
  const int* seg_wg0;    // K
  const long* seg_row0;  // K
  const long* seg_row1;  // K
  const double* tol;     // K relative tolerances
  const long* rowptr;
  const int* bcol;
  const short* bcol16 = nullptr;  // block columns minus their rows, when all fit in 16 bits
  const void* bval;
  const void* dinv;
  const void* binv;
  int precond;
  int lpr;
  void* x;
  void* r;
  void* z;
  void* p0;
  void* p1;
  void* q;
  double* part1;         // W
  double* part2;         // 2 parities x 2 x W
  double* partb;         // W
  PcgScalars* sc;        // K
  const int* kctl = nullptr;  // as PcgArgs::kctl
  int coarse = 0;        // 1: two-level (precond 1 + cg)
  SegCoarseArgs cg;
};
constexpr int SEG_MAXG = 256;  // workgroups per segment at most
void launch_seg_init(hipStream_t s, bool fp32, const SegPcgArgs& a, const void* b, int x_zero);
void launch_seg_k1(hipStream_t s, bool fp32, const SegPcgArgs& a, int k);
void launch_seg_k2(hipStream_t s, bool fp32, const SegPcgArgs& a, int k);
void launch_seg_status(hipStream_t s, const SegPcgArgs& a, int k_next, int maxit);
// block rows of one K1 workgroup pass for `lpr` lanes per row
int seg_rows_per_wg(int lpr);

// r64 = b64 - A x64 with A in precision T and FP64 accumulation; writes
// |r|^2 into out[0].  partials: PCG_MAXP doubles + one unsigned ticket.
// the same over the SELL-64 copy described by a (a.sell set; grid a.g1 <= PCG_MAXP)
void launch_residual64_sell(hipStream_t s, bool fp32_mat, const PcgArgs& a, const double* x,
                            const double* b, double* r, double* partials, double* out);
// (rows [row0, row0 + n))
void launch_residual64(hipStream_t s, bool fp32_mat, int n, const long* rowptr,
                       const int* bcol, const void* bval, const double* x,
                       const double* b, double* r, double* partials,
                       double* out, long row0 = 0);
// y = A x in precision T with `lpr` lanes per block row (tuning, timing)
void launch_spmv(hipStream_t s, bool fp32, int n, const long* rowptr,
                 const int* bcol, const void* bval, const void* x, void* y, int lpr);

// ---- batches of independent instances ---------------------------------------
// Instance k owns nodes [noff[k], noff[k+1]) and, per element class c
// (0 tets, 1 membranes, 2 hinges), elements [eoff[c*(K+1)+k], eoff[c*(K+1)+k+1]).
struct BatchArgs {
  int K;
  const long* noff;
  const long* eoff;
  const int* node_inst;
};
// out[k] = |v_k|^2 over the instance's 3 DoFs per node
void launch_batch_norm2(hipStream_t s, const BatchArgs& a, const double* v, double* out);
// x = xsave + fac[k] dx on instances with fac[k] >= 0 (others untouched)
void launch_batch_step(hipStream_t s, const BatchArgs& a, long n_nodes, const double* fac,
                       const double* xsave, const double* dx, double* x);
// per instance mode[k]: 0 -> b = 0, dx = 0; 1 (Newton) -> b = -g,
// dx = xsave - v0 (warm start); 2 (steepest) -> b = 0, dx = -g
void launch_batch_rhs(hipStream_t s, const BatchArgs& a, long n_nodes, const uint8_t* mode,
                      const double* g, const double* xsave, const double* v0, double* b,
                      double* dx);
// out[k] = a_k . b_k over each instance's DoFs (FP64)
void launch_batch_dot(hipStream_t s, const BatchArgs& a, const double* x, const double* y,
                      double* out);
// bval[diagonal block of node i] += shift[instance of i] I
void launch_batch_add_diag(hipStream_t s, bool fp32, const BatchArgs& a, long n_nodes,
                           const long* diag, void* bval, const double* shift);
// out[k] = |A_k|_F^2 over the stored lower entry set of instance k's rows
void launch_batch_frob(hipStream_t s, bool fp32, const BatchArgs& a, const long* rowptr,
                       const int* brow, const int* bcol, const void* bval, double* out);
// x += d (FP32 correction) on instances with mask[k] != 0
void launch_batch_axpy_f32(hipStream_t s, const BatchArgs& a, long n_nodes, const uint8_t* mask,
                           const float* d, double* x);
// dst = src on instances with mask[k] != 0
void launch_batch_copy(hipStream_t s, const BatchArgs& a, long n_nodes, const uint8_t* mask,
                       const double* src, double* dst);

// Per-instance reductions over a static workgroup table (each instance's
// DoF or block range cut into chunks): one partial per workgroup, then each
// instance's partials summed in order by a second launch.  Instances with
// mask[k] == 0 (mask may be null) are skipped and keep out[k].
struct SegRedTable {
  int W = 0;                  // workgroups
  const int* inst = nullptr;  // W: instance of each workgroup
  const long* lo = nullptr;   // W: chunk start (DoF, or BSR3 block)
  const long* hi = nullptr;   // W: chunk end
  const int* wg0 = nullptr;   // K + 1: first workgroup of each instance
  double* part = nullptr;     // W partials (3 W for the energy table)
  const uint8_t* kind = nullptr;  // energy table: 0 element chunk, 1 node chunk
};
// out[4k..4k+3] = {E_k, |g_k|^2, |E_el| + |E_load|, E_el} over the energy
// table (element chunks of elemE, node chunks of nodeE / nodeG2)
void launch_seg_energy(hipStream_t s, int K, const SegRedTable& t, const uint8_t* mask,
                       const double* elemE, const double* nodeE, const double* nodeG2,
                       double* out);
// out[k] = x_k . y_k (FP64) over the DoF table
void launch_seg_dot(hipStream_t s, int K, const SegRedTable& t, const uint8_t* mask,
                    const double* x, const double* y, double* out);
// out[k] = |A_k|_F^2 over the stored lower entry set (blocks with row >= column)
// over the block table
void launch_seg_frob(hipStream_t s, bool fp32, int K, const SegRedTable& t, const uint8_t* mask,
                     const int* brow, const int* bcol, const void* bval, double* out);

// small vector helpers
void launch_cast_f64_to_f32(hipStream_t s, long n, const double* a, float* b);
void launch_axpy_f32_to_f64(hipStream_t s, long n, const float* d, double* x);
// out = a*x + b*y in FP64 (out may alias x or y)
void launch_lincomb64(hipStream_t s, long n, double a, const double* x, double b,
                      const double* y, double* out);
// out[0] = dot(a, b) (FP64); partials: PCG_MAXP doubles + unsigned ticket
void launch_dot64(hipStream_t s, long n, const double* a, const double* b,
                  double* partials, double* out);
// cast a T vector to FP64
void launch_widen(hipStream_t s, bool fp32, long n, const void* a, double* b);
void launch_narrow_or_copy(hipStream_t s, bool fp32, long n, const double* a,
                           void* b);

// ---- FP64 banded LDL^T (the EigenLDLT linear solver) ------------------------
// One factorization per problem (a model, or one instance of a batch), in a
// band order of its nodes: DoF I = 3 q + a for the node at position q.  Its
// band holds N rows of bw + 1 doubles, row I = entries (I, I - d), d = 0..bw:
// the lower band of the matrix before factorization, L below the diagonal
// after (the diagonal keeps A; D is stored apart).
constexpr int LDLT_MAXW = 140;  // bw + 1 at most: the LDS window is W x W doubles
struct LdltProb {
  long band0;  // offset of the band (doubles)
  int dof0;    // offset of the problem's DoFs in D and the solve scratch
  int N;       // DoFs
  int bw;      // half-bandwidth in DoFs (3 x node half-bandwidth + 2)
  int node0;   // first node of the problem in the model's node numbering
};
struct LdltArgs {
  const LdltProb* prob;
  const int* perm;     // node -> node0 + its band position
  const int* iperm;    // node0 + band position -> node
  const int* nprob;    // node -> problem
  double* band;
  double* D;           // pivots, band order
  const double* sinT;  // per problem: pivots below it are singular
  int* status;         // per problem {first zero pivot or -1, pivots below sinT}
};
// zero the bands of the listed problems, then copy the lower BSR3 entries of
// problems with pmask set (all when pmask is null) into their bands
void launch_ldlt_gather(hipStream_t s, bool fp32, const LdltArgs& a, const int* list,
                        int nlist, long max_band, long nb, const int* brow, const int* bcol,
                        const void* bval, const uint8_t* pmask);
// factor the listed problems, one workgroup each
void launch_ldlt_factor(hipStream_t s, const LdltArgs& a, const int* list, int nlist);
// x = A^-1 b for the listed problems (b, x: the model's node-major DoF
// vectors); y: scratch of the problems' DoF count
void launch_ldlt_solve(hipStream_t s, const LdltArgs& a, const int* list, int nlist,
                       const double* b, double* x, double* y);

// ---- two-level preconditioner: coarse operator and its application ----------
// Pattern of Ac = P^T A P: coarse block pairs (a, b), a <= b, each with the
// BSR3 blocks (i in a, j in b) that sum into it, in block order.
struct CoarseBuild {
  int ncb;                // coarse block pairs
  int npiece;             // pieces of at most COARSE_PIECE blocks, pair by pair
  const int* cb_a;        // ncb
  const int* cb_b;        // ncb
  const int* pair_piece0; // ncb + 1: pieces of each pair
  const int* piece_lo;    // npiece + 1: ranges into cb_fine
  const int* cb_fine;     // BSR3 block indexes
  double* piece_sum;      // npiece x 36 partial 6x6 blocks
  const int* brow;        // BSR3 block row / column (nodes)
  const int* bcol;
  const void* bval;       // BSR3 values, Hessian precision
  double* A0;             // npad x npad: Ac, then the sweeps' ping-pong pair
  double* A1;
  double* D;              // 2 x 32 x 32: inverses of the next pivot tiles
  int* ok;
  // segmented builds: pieces of instances without workgroups (live[instance]
  // == 0) are skipped
  const int* agg_inst = nullptr;
  const int* live = nullptr;
};
// Ac into A0 (dense FP64: inactive and padding DoFs get identity rows), then
// -Ac^-1 by block Gauss-Jordan sweeps; returns the buffer that holds it.
// ok is cleared when a pivot is not positive (Ac not SPD to FP64 precision).
double* launch_coarse_operator(hipStream_t s, bool fp32, const CoarseArgs& c,
                               const CoarseBuild& b, bool assemble_only);
// z += P Ac^-1 P^T r after z = B^-1 r, and cdot = (P^T r).(Ac^-1 P^T r);
// returns at once when the PCG has converged
void launch_coarse_apply(hipStream_t s, bool fp32, const PcgArgs& a);
// Segmented versions over the instances with workgroups in a (seg_G > 0):
// Ac per instance from the pieces of b (b.A0..D unused) into c.R's storage
// (A, the sum of nc^2), inverted in place by one workgroup per instance.
// The application after iteration k (k = -1 after the init) is two
// launches: P^T r per aggregate, then uc = Ac^-1 rc and cdot[(k & 1) K +
// instance] per instance; the segmented SpMV of iteration k + 1 adds P uc to
// z where it reads z.
void launch_seg_coarse_operator(hipStream_t s, bool fp32, const SegPcgArgs& a,
                                const CoarseBuild& b, double* A, int* ok);
void launch_seg_coarse_apply(hipStream_t s, bool fp32, const SegPcgArgs& a, int k);
// workgroups of the coarse product for npad coarse DoFs
int coarse_gemv_grid(int npad);

}  // namespace pgn
