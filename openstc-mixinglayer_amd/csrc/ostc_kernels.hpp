// ostc_kernels.hpp -- host-side launcher declarations for the HIP kernels.
#pragma once
#include <algorithm>
#include "ostc_ops.hpp"
#include "ostc_parallel.hpp"

namespace ostc {

// per-block (reference-shaped) path, ostc_kernels_plain.hip
void plain_rhs(const DevBlock& B, const Params& P, hipStream_t st, long long* n);
void plain_update(const DevBlock& B, double dt, int rk, hipStream_t st, long long* n);
void plain_local_dt(const DevBlock& B, const Params& P, hipStream_t st, long long* n);
void min_reduce(const double* in, long long len, double* scratch, int n_first, hipStream_t st, long long* n);
void padded_pass(const DevBlock& B, int slot, int what, hipStream_t st, long long* n);
void count_pass(unsigned long long* pass, int lo, int hi, hipStream_t st, long long* n);
void sponge_mean(const DevBlock& B, const unsigned long long* iter, hipStream_t st, long long* n);
void ops_launch(const DevBlock* d_blocks, int n_blocks, const DevOp* d_ops, int op_lo, int op_hi, long long total, int slot,
                int src, hipStream_t st, long long* n, bool props = false);
// parallel faces (ring slot 0): pack into / unpack from buf + f.offset
void par_pack(const DevBlock& B, const ParFace& f, double* buf, hipStream_t st);
void par_unpack(const DevBlock& B, const ParFace& f, const double* recv, const double* sent, hipStream_t st);
void store_last(const DevBlock& B, hipStream_t st);
void residual_partial(const DevBlock& B, double* out4, int nblk, hipStream_t st);

// dual time stepping + DPLUR on the per-block path, ostc_kernels_dual.hip
struct DualBlock {
  double* qn1;      // q^(n-1), 5 padded planes
  double* qn_star;  // J/(2 dt) (4 q^n - q^(n-1)), 5 interior planes
  double* sr;       // inviscid spectral radii along xi, eta, zeta, 3 interior planes
  double* dq0;      // dt dq / diag, 5 interior planes
  double* dqk;      // the DPLUR iterate, 5 interior planes
  double* last;     // rho, |V|, p, T of the last inner iteration, then their squared change
};
void dual_qn_star(const DevBlock& B, const DualBlock& D, double dt, hipStream_t st, long long* n);
void dual_pre(const DevBlock& B, const Params& P, const DualBlock& D, double dt, double diag_factor, hipStream_t st,
              long long* n);
void dual_dplur(const DevBlock& B, const DualBlock& D, double diag_factor, hipStream_t st, long long* n);
void dual_update(const DevBlock& B, hipStream_t st, long long* n);
void dual_last_store(const DevBlock& B, const DualBlock& D, hipStream_t st);
void dual_last_square(const DevBlock& B, const DualBlock& D, double* out4, int nblk, hipStream_t st);

// fused multi-block path, ostc_kernels_fused.hip
struct FusedArgs {
  const DevBlock* blocks;
  const int* tile_block;     // block of each tile
  int n_tiles;
  int in, out;               // ring slots read / written
  double ra, rb, rc;         // SSP-RK3 coefficients of this stage
  const double* dt;          // device dt (bit pattern of a positive double)
  Params P;
  int xcd_remap;             // 1: contiguous tile chunks per XCD
  int target_block;          // >= 0: only this block's tiles (dq mode)
};

struct TileShape { int tx, ty, tz, waves_per_eu; };  // waves_per_eu: occupancy hint given to the compiler
int fused_shape_count(int scheme);
TileShape fused_shape(int scheme, int index);
// Launches the fused kernel of table entry `index`; mode 0 = RK update in
// one kernel, mode 1 = write dq only, mode 2 = RK update split into an
// inviscid kernel and a viscous+update kernel.  Returns the dynamic LDS bytes
// used, or -1.
int fused_launch(int index, int mode, const FusedArgs& a, int n_wg, hipStream_t st);
int fused_lds_bytes(int scheme, int index);
void dt_rows(const FusedArgs& a, int n_blocks, int n_wg, double* dt_bits, hipStream_t st, long long* n);

// reverse mode, ostc_kernels_adjoint.hip
struct VjpArgs {
  int mx, my, mz, g, sx, sy;
  long long sz, n0;
  const double* cv;      // state the RHS was evaluated at (5 planes, padded)
  const double* bv;      // its primitives (6 planes, padded)
  const double* mul;     // its viscosity (padded)
  const double* jac;
  const double* met;
  const double* w;       // dq cotangent (5 planes, interior)
  double* lam;           // cv cotangent (5 planes, padded), accumulated
  double* fs;            // viscous face cotangents: 3 directions x 16 planes
  long long fs_stride;   // plane stride of fs, >= vjp_face_stride()
  double* gs;            // partial gathers of the viscous cotangents: 8 planes over [-1, m]^3
  long long gs_stride;   // plane stride of gs, >= vjp_gather_stride()
  Params P;
};
long long vjp_face_stride(int mx, int my, int mz);
long long vjp_gather_stride(int mx, int my, int mz);
void rhs_vjp(const VjpArgs& a, hipStream_t st, long long* n);
void adj_stage_pre(const DevBlock& B, double* lam, double* lam_qn, double* w, double ra, double rb, double cdt,
                   hipStream_t st);
void adj_stage_qn(const DevBlock& B, double* lam, double* lam_qn, hipStream_t st);
void adj_bc_op(const DevBlock& B, const DevOp& op, double* lam, hipStream_t st);

constexpr int kMaxModes = 64;
struct FourierModes {        // see ostc_fourier in abi.h
  int n;
  int comp[kMaxModes], kx[kMaxModes], kz[kMaxModes], trig[kMaxModes];
  double amp[kMaxModes];
  double y0, dx, dy, dz, lx, lz, delta, yc, du;
};
struct ThetaParams {         // see ostc_theta in abi.h
  double u1, u2, rho0, du, dy;
  int ni, nk;
};
void fourier_add(const DevBlock& B, const FourierModes& F, double* bv, hipStream_t st);
int fourier_grad_parts();
void fourier_grad(const DevBlock& B, const FourierModes& F, const double* lam, const double* cv, double* part,
                  double* grad, hipStream_t st);
void theta_eval(const DevBlock& B, const ThetaParams& T, const double* cv, double* rows, double* coef, double* theta,
                hipStream_t st);
void theta_seed(const DevBlock& B, const ThetaParams& T, const double* coef, double* lam, hipStream_t st);
// The momentum thickness with its row sums, coefficients and result in a
// file-scope device array (4 kThetaMaxRows + 1 doubles per GPU) instead of
// library-allocated memory; `scratch_ptr` comes from theta_scratch_ptr.
constexpr int kThetaMaxRows = 16384;
int theta_scratch_ptr(int device, double** scratch_ptr);

}  // namespace ostc
