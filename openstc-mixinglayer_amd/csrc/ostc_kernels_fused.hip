// ostc_kernels_fused.hip -- one persistent kernel per RK stage over the tiles
// of every block: inviscid WENO fluxes along x, y, z and second-order viscous
// fluxes staged through LDS, the residual dq kept in registers, and the
// SSP-RK3 update, primitives and Sutherland viscosity written out.  Reads
// ring slot `in` (state incl. ghosts) and qn = slot 0, writes the interior of
// slot `out`; ghost layers are filled afterwards by the op launches.
//
// Work decomposition: a tile is TX x TY x TZ cells of one block, one thread
// per cell.  Each direction phase computes split fluxes F+/F- for the tile's
// (T_d + 2R) x (columns) points, then one WENO flux per face, then each cell
// subtracts its face difference.  Accumulation order follows the reference:
// dq = ((0 - dFx) - dFy) - dFz, then + dGx + dGy + dGz (src/RK.cuh:113-139).
#include "ostc_kernels.hpp"

namespace ostc {

template <int TX, int TY, int TZ, int D>
struct Dir {
  static constexpr int NT = TX * TY * TZ;
  static constexpr int TD = D == 0 ? TX : (D == 1 ? TY : TZ);
  static constexpr int PC = NT / TD;  // columns
  static constexpr int NF = (TD + 1) * PC;
  static constexpr int FPT = (NF + NT - 1) / NT;
  // column -> the two tile coordinates other than D
  __device__ static void col_coords(int col, int& a, int& b) {
    if (D == 0) { a = col % TY; b = col / TY; }   // (tj, tk)
    else { a = col % TX; b = col / TX; }          // (ti, tk) or (ti, tj)
  }
  __device__ static int my_col(int ti, int tj, int tk) {
    return D == 0 ? tj + TY * tk : (D == 1 ? ti + TX * tk : ti + TX * tj);
  }
  __device__ static int my_pos(int ti, int tj, int tk) { return D == 0 ? ti : (D == 1 ? tj : tk); }
  // index of (column, position s) in a per-column array of length LEN
  template <int LEN>
  __device__ static int at(int col, int s) { return D == 0 ? col * LEN + s : s * PC + col; }
};

template <int TX, int TY, int TZ, int SCHEME>
struct FusedLds {
  static constexpr int NT = TX * TY * TZ;
  static constexpr int R = Stencil<SCHEME>::R;
  static constexpr int pts(int td) { return (td + 2 * R) * (NT / td); }
  static constexpr int mx3(int a, int b, int c) { return a > b ? (a > c ? a : c) : (b > c ? b : c); }
  static constexpr int NPTS = mx3(pts(TX), pts(TY), pts(TZ));
  static constexpr int NB = (TX + 2) * (TY + 2) * (TZ + 2);
  static constexpr int NFMAX = mx3((TX + 1) * (NT / TX), (TY + 1) * (NT / TY), (TZ + 1) * (NT / TZ));
  static constexpr int VWORDS = 5 * NB + 10 * NT + 4 * NFMAX;  // viscous phase
  static constexpr int WORDS = (10 * NPTS > VWORDS) ? 10 * NPTS : VWORDS;
  static constexpr int BYTES = WORDS * 8;
  static constexpr int VBYTES = VWORDS * 8;
};

struct TileCtx {
  int i0, j0, k0;      // tile origin
  int ti, tj, tk;      // this thread's cell in the tile
  bool valid;          // cell inside the block
};

template <int TX, int TY, int TZ, int SCHEME, int D>
__device__ __forceinline__ void inviscid_phase(double* smem, const DevBlock& B, const Params& P, int slot,
                                               const TileCtx& t, double dq[5]) {
  using Dr = Dir<TX, TY, TZ, D>;
  constexpr int NT = Dr::NT, R = Stencil<SCHEME>::R, W = Stencil<SCHEME>::W;
  constexpr int TD = Dr::TD, PC = Dr::PC, LP = TD + 2 * R, NPTS = LP * PC, NF = Dr::NF, FPT = Dr::FPT;
  const int tid = threadIdx.x;
  const int m = D == 0 ? B.mx : (D == 1 ? B.my : B.mz);
  const int origin = D == 0 ? t.i0 : (D == 1 ? t.j0 : t.k0);
  const double* cv = B.cv[slot];
  const double* bv = B.bv[slot];
  const long long s = B.sz;
  double* sFp = smem;
  double* sFm = smem + 5 * NPTS;

  // split fluxes at the points of every column
  for (int p = tid; p < NPTS; p += NT) {
    int col, sp;
    if (D == 0) { col = p / LP; sp = p % LP; } else { sp = p / PC; col = p % PC; }
    int a, b;
    Dr::col_coords(col, a, b);
    int ijk[3];
    if (D == 0) { ijk[1] = t.j0 + a; ijk[2] = t.k0 + b; }
    if (D == 1) { ijk[0] = t.i0 + a; ijk[2] = t.k0 + b; }
    if (D == 2) { ijk[0] = t.i0 + a; ijk[1] = t.j0 + b; }
    const int pos = origin + sp - R;
    ijk[D] = pos;
    const bool ok = ijk[0] < B.mx + (D == 0 ? R : 0) && ijk[1] < B.my + (D == 1 ? R : 0) &&
                    ijk[2] < B.mz + (D == 2 ? R : 0);
    if (ok) {
      const long long q = pidx(B, ijk[0], ijk[1], ijk[2]);
      double c5[5];
#pragma unroll
      for (int l = 0; l < 5; ++l) c5[l] = cv[q + l * s];
      const double c = sqrt(kGamma * kRAir * bv[q + 5 * s]);
      double fp[5], fm[5];
      lf_split5(c5, bv[q + 4 * s], c, B.met[q + (3 * D) * s], B.met[q + (3 * D + 1) * s],
                B.met[q + (3 * D + 2) * s], B.jac[q], fp, fm);
#pragma unroll
      for (int l = 0; l < 5; ++l) { sFp[l * NPTS + p] = fp[l]; sFm[l * NPTS + p] = fm[l]; }
    }
  }
  __syncthreads();

  double fres[FPT][5];
#pragma unroll
  for (int u = 0; u < FPT; ++u) {
    const int f = tid + u * NT;
    if (f < NF) {
      int col, sf;
      if (D == 0) { col = f / (TD + 1); sf = f % (TD + 1); } else { sf = f / PC; col = f % PC; }
      int a, b;
      Dr::col_coords(col, a, b);
      int ijk[3];
      if (D == 0) { ijk[1] = t.j0 + a; ijk[2] = t.k0 + b; }
      if (D == 1) { ijk[0] = t.i0 + a; ijk[2] = t.k0 + b; }
      if (D == 2) { ijk[0] = t.i0 + a; ijk[1] = t.j0 + b; }
      const int L = origin + sf - 1;  // face L+1/2
      ijk[D] = L;
      const bool ok = L <= m - 1 && ijk[0] < B.mx + (D == 0 ? 1 : 0) && ijk[1] < B.my + (D == 1 ? 1 : 0) &&
                      ijk[2] < B.mz + (D == 2 ? 1 : 0);
      if (ok) {
        const long long p1 = pidx(B, ijk[0], ijk[1], ijk[2]);
        const long long p2 = p1 + (D == 0 ? 1 : (D == 1 ? (long long)B.sx : (long long)B.sx * B.sy));
        const double j1 = B.jac[p1], j2 = B.jac[p2];
        double eps3[3];
        weno_eps(P, B.met[p1 + 3 * D * s] * j1, B.met[p1 + (3 * D + 1) * s] * j1, B.met[p1 + (3 * D + 2) * s] * j1,
                 B.met[p2 + 3 * D * s] * j2, B.met[p2 + (3 * D + 1) * s] * j2, B.met[p2 + (3 * D + 2) * s] * j2,
                 eps3);
        const int sL = sf - 1 + R;
#pragma unroll
        for (int l = 0; l < 5; ++l) {
          double vp[W], vm[W];
#pragma unroll
          for (int r = 0; r < W; ++r) {
            vp[r] = sFp[l * NPTS + Dr::template at<LP>(col, sL - W / 2 + r)];
            vm[r] = sFm[l * NPTS + Dr::template at<LP>(col, sL - W / 2 + 1 + r)];
          }
          const double e = l == 0 ? eps3[0] : (l == 4 ? eps3[2] : eps3[1]);
          fres[u][l] = weno_face<SCHEME>(vp, vm, e);
        }
      }
    }
  }
  __syncthreads();
  double* sFc = smem;  // aliases the split fluxes, which are no longer read
#pragma unroll
  for (int u = 0; u < FPT; ++u) {
    const int f = tid + u * NT;
    if (f < NF) {
#pragma unroll
      for (int l = 0; l < 5; ++l) sFc[l * NF + f] = fres[u][l];
    }
  }
  __syncthreads();
  if (t.valid) {
    const int col = Dr::my_col(t.ti, t.tj, t.tk), pos = Dr::my_pos(t.ti, t.tj, t.tk);
    const int fL = Dr::template at<TD + 1>(col, pos), fR = Dr::template at<TD + 1>(col, pos + 1);
#pragma unroll
    for (int l = 0; l < 5; ++l) dq[l] -= sFc[l * NF + fR] - sFc[l * NF + fL];
  }
  __syncthreads();
}

template <int TX, int TY, int TZ, int D>
__device__ __forceinline__ void viscous_dir(double* smem, const DevBlock& B, const Params& P, const TileCtx& t,
                                            double dq[5]) {
  using Dr = Dir<TX, TY, TZ, D>;
  constexpr int NT = Dr::NT, TD = Dr::TD, PC = Dr::PC, NF = Dr::NF, FPT = Dr::FPT;
  constexpr int BX = TX + 2, BY = TY + 2, NB = (TX + 2) * (TY + 2) * (TZ + 2);
  const int tid = threadIdx.x;
  const int m = D == 0 ? B.mx : (D == 1 ? B.my : B.mz);
  const int origin = D == 0 ? t.i0 : (D == 1 ? t.j0 : t.k0);
  const long long s = B.sz;
  const double* sV = smem;          // [5][NB]: u, v, w, T, mu on the tile box
  const double* sM = smem + 5 * NB; // [10][NT]: metric (9) and Jacobian of the tile cells
  double* sFv = smem + 5 * NB + 10 * NT;  // [4][NF]
  auto pv = [&](int i, int j, int k, int l) {
    const int q = l == 5 ? 3 : l - 1;
    return sV[q * NB + ((k - t.k0 + 1) * BY + (j - t.j0 + 1)) * BX + (i - t.i0 + 1)];
  };
  // one face at a time, each flux written straight to LDS (sFv does not
  // alias the staged box, so no barrier is needed before the writes)
#pragma unroll 1
  for (int u = 0; u < FPT; ++u) {
    const int f = tid + u * NT;
    if (f < NF) {
      int col, sf;
      if (D == 0) { col = f / (TD + 1); sf = f % (TD + 1); } else { sf = f / PC; col = f % PC; }
      int a, b;
      Dr::col_coords(col, a, b);
      int ijk[3];
      if (D == 0) { ijk[1] = t.j0 + a; ijk[2] = t.k0 + b; }
      if (D == 1) { ijk[0] = t.i0 + a; ijk[2] = t.k0 + b; }
      if (D == 2) { ijk[0] = t.i0 + a; ijk[1] = t.j0 + b; }
      const int L = origin + sf - 1;
      ijk[D] = L;
      const bool ok = L <= m - 1 && ijk[0] < B.mx + (D == 0 ? 1 : 0) && ijk[1] < B.my + (D == 1 ? 1 : 0) &&
                      ijk[2] < B.mz + (D == 2 ? 1 : 0);
      if (ok) {
        const int i = ijk[0], j = ijk[1], k = ijk[2];
        double d[4][3];
        const int var[4] = {1, 2, 3, 5};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int l = var[q];
          if (D == 0) {
            d[q][0] = pv(i + 1, j, k, l) - pv(i, j, k, l);
            d[q][1] = 0.25 * (pv(i, j + 1, k, l) - pv(i, j - 1, k, l) + pv(i + 1, j + 1, k, l) - pv(i + 1, j - 1, k, l));
            d[q][2] = 0.25 * (pv(i, j, k + 1, l) - pv(i, j, k - 1, l) + pv(i + 1, j, k + 1, l) - pv(i + 1, j, k - 1, l));
          } else if (D == 1) {
            d[q][0] = 0.25 * (pv(i + 1, j, k, l) - pv(i - 1, j, k, l) + pv(i + 1, j + 1, k, l) - pv(i - 1, j + 1, k, l));
            d[q][1] = pv(i, j + 1, k, l) - pv(i, j, k, l);
            d[q][2] = 0.25 * (pv(i, j, k + 1, l) - pv(i, j, k - 1, l) + pv(i, j + 1, k + 1, l) - pv(i, j + 1, k - 1, l));
          } else {
            d[q][0] = 0.25 * (pv(i + 1, j, k, l) - pv(i - 1, j, k, l) + pv(i + 1, j, k + 1, l) - pv(i - 1, j, k + 1, l));
            d[q][1] = 0.25 * (pv(i, j + 1, k, l) - pv(i, j - 1, k, l) + pv(i, j + 1, k + 1, l) - pv(i, j - 1, k + 1, l));
            d[q][2] = pv(i, j, k + 1, l) - pv(i, j, k, l);
          }
        }
        const long long pA = pidx(B, i, j, k);
        const long long pB = pA + (D == 0 ? 1 : (D == 1 ? (long long)B.sx : (long long)B.sx * B.sy));
        const int iB = i + (D == 0), jB = j + (D == 1), kB = k + (D == 2);
        // metric and Jacobian of the two face cells: from the tile's LDS copy,
        // or from global memory for the cell beyond either end of a column
        double mA[9], mB[9], jacA, jacB;
        const int cA = ((k - t.k0) * TY + (j - t.j0)) * TX + (i - t.i0);
        const int cB = ((kB - t.k0) * TY + (jB - t.j0)) * TX + (iB - t.i0);
        const bool inA = (D == 0 ? i >= t.i0 : (D == 1 ? j >= t.j0 : k >= t.k0));
        const bool inB = (D == 0 ? iB < t.i0 + TX : (D == 1 ? jB < t.j0 + TY : kB < t.k0 + TZ));
        if (inA) {
#pragma unroll
          for (int c = 0; c < 9; ++c) mA[c] = sM[c * NT + cA];
          jacA = sM[9 * NT + cA];
        } else {
#pragma unroll
          for (int c = 0; c < 9; ++c) mA[c] = B.met[pA + c * s];
          jacA = B.jac[pA];
        }
        if (inB) {
#pragma unroll
          for (int c = 0; c < 9; ++c) mB[c] = sM[c * NT + cB];
          jacB = sM[9 * NT + cB];
        } else {
#pragma unroll
          for (int c = 0; c < 9; ++c) mB[c] = B.met[pB + c * s];
          jacB = B.jac[pB];
        }
        const double um = 0.5 * (pv(i, j, k, 1) + pv(iB, jB, kB, 1));
        const double vm = 0.5 * (pv(i, j, k, 2) + pv(iB, jB, kB, 2));
        const double wm = 0.5 * (pv(i, j, k, 3) + pv(iB, jB, kB, 3));
        const double muA = sV[4 * NB + ((k - t.k0 + 1) * BY + (j - t.j0 + 1)) * BX + (i - t.i0 + 1)];
        const double muB = sV[4 * NB + ((kB - t.k0 + 1) * BY + (jB - t.j0 + 1)) * BX + (iB - t.i0 + 1)];
        double fv[4];
        viscous_face_flux<D>(d, mA, mB, jacA, jacB, muA, muB, um, vm, wm, P.prandtl, fv);
#pragma unroll
        for (int l = 0; l < 4; ++l) sFv[l * NF + f] = fv[l];
      }
    }
  }
  __syncthreads();
  if (t.valid) {
    const int col = Dr::my_col(t.ti, t.tj, t.tk), pos = Dr::my_pos(t.ti, t.tj, t.tk);
    const int fL = Dr::template at<TD + 1>(col, pos), fR = Dr::template at<TD + 1>(col, pos + 1);
#pragma unroll
    for (int l = 0; l < 4; ++l) dq[l + 1] += sFv[l * NF + fR] - sFv[l * NF + fL];
  }
  __syncthreads();
}

template <int TX, int TY, int TZ>
__device__ __forceinline__ void viscous_phase(double* smem, const DevBlock& B, const Params& P, int slot,
                                              const TileCtx& t, double dq[5]) {
  constexpr int NT = TX * TY * TZ, BX = TX + 2, BY = TY + 2, NB = (TX + 2) * (TY + 2) * (TZ + 2);
  const double* bv = B.bv[slot];
  const double* mu = B.mul[slot];
  const long long s = B.sz;
  for (int p = threadIdx.x; p < NB; p += NT) {
    const int a = p % BX, b = (p / BX) % BY, c = p / (BX * BY);
    const int i = t.i0 - 1 + a, j = t.j0 - 1 + b, k = t.k0 - 1 + c;
    if (i <= B.mx && j <= B.my && k <= B.mz) {
      const long long q = pidx(B, i, j, k);
      smem[p] = bv[q + s];
      smem[NB + p] = bv[q + 2 * s];
      smem[2 * NB + p] = bv[q + 3 * s];
      smem[3 * NB + p] = bv[q + 5 * s];
      smem[4 * NB + p] = mu[q];
    }
  }
  {
    // this thread's cell: metric and Jacobian, read once per stage
    const int i = t.i0 + t.ti, j = t.j0 + t.tj, k = t.k0 + t.tk;
    double* sM = smem + 5 * NB;
    const int c = threadIdx.x;
    if (i <= B.mx && j <= B.my && k <= B.mz) {
      const long long q = pidx(B, i, j, k);
#pragma unroll
      for (int m = 0; m < 9; ++m) sM[m * NT + c] = B.met[q + m * s];
      sM[9 * NT + c] = B.jac[q];
    }
  }
  __syncthreads();
  viscous_dir<TX, TY, TZ, 0>(smem, B, P, t, dq);
  viscous_dir<TX, TY, TZ, 1>(smem, B, P, t, dq);
  viscous_dir<TX, TY, TZ, 2>(smem, B, P, t, dq);
}

// MODE 0: inviscid + viscous + RK update; MODE 1: write dq (inviscid, and
// viscous when VISC) without updating; MODE 2: viscous + RK update starting
// from the inviscid dq that a MODE 1 / VISC 0 launch stored (a stage split in
// two kernels of lower register and LDS pressure).
template <int TX, int TY, int TZ, int WPE, int SCHEME, int VISC, int MODE>
__global__ __launch_bounds__(TX * TY * TZ) __attribute__((amdgpu_waves_per_eu(WPE))) void k_fused(FusedArgs A) {
  constexpr int NT = TX * TY * TZ;
  extern __shared__ double smem[];
  const int tid = threadIdx.x;
  TileCtx t;
  t.ti = tid % TX;
  t.tj = (tid / TX) % TY;
  t.tk = tid / (TX * TY);
  // tile assignment.  Round-robin dispatch puts workgroup w on XCD w % 8;
  // with xcd_remap the workgroups of one XCD take consecutive work items
  // (XCD x owns items [x q + min(x, r), ...) of G = 8 q + r), so tiles that
  // run together on one XCD are neighbours in the block and share their
  // halo points through that XCD's L2.  One workgroup per tile: item = tile;
  // a persistent grid: item = a chunk of consecutive tiles.
  const int G = gridDim.x;
  int lid = blockIdx.x;
  if (A.xcd_remap) {
    const int x = blockIdx.x % 8, q = G / 8, r = G % 8;
    lid = x * q + min(x, r) + blockIdx.x / 8;
  }
  int first, step, last;
  if (G < A.n_tiles && A.xcd_remap) {
    const int chunk = (A.n_tiles + G - 1) / G;
    first = lid * chunk;
    last = min(A.n_tiles, first + chunk);
    step = 1;
  } else {
    first = lid;
    last = A.n_tiles;
    step = G;
  }
  for (int tile = first; tile < last; tile += step) {
    const int b = A.tile_block[tile];
    if (A.target_block >= 0 && b != A.target_block) continue;
    // a reference into global memory: the ring slots are picked by runtime
    // indices, which would push a private copy of the descriptor to scratch
    const DevBlock& B = A.blocks[b];
    const int local = tile - B.tile_base;
    t.i0 = (local % B.ntx) * TX;
    t.j0 = ((local / B.ntx) % B.nty) * TY;
    t.k0 = (local / (B.ntx * B.nty)) * TZ;
    const int i = t.i0 + t.ti, j = t.j0 + t.tj, k = t.k0 + t.tk;
    t.valid = i < B.mx && j < B.my && k < B.mz;
    double dq[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (MODE != 2) {
      inviscid_phase<TX, TY, TZ, SCHEME, 0>(smem, B, A.P, A.in, t, dq);
      inviscid_phase<TX, TY, TZ, SCHEME, 1>(smem, B, A.P, A.in, t, dq);
      inviscid_phase<TX, TY, TZ, SCHEME, 2>(smem, B, A.P, A.in, t, dq);
    } else if (t.valid) {
      // split stage: the inviscid part of dq was written by the MODE 1 launch
      const long long p0 = pidx0(B, i, j, k);
#pragma unroll
      for (int l = 0; l < 5; ++l) dq[l] = B.dq[p0 + l * B.n0];
    }
    if (VISC) viscous_phase<TX, TY, TZ>(smem, B, A.P, A.in, t, dq);
    if (!t.valid) continue;
    const long long p = pidx(B, i, j, k), s = B.sz;
    if (MODE == 1) {
      const long long p0 = pidx0(B, i, j, k);
#pragma unroll
      for (int l = 0; l < 5; ++l) B.dq[p0 + l * B.n0] = dq[l];
      continue;
    }
    // update_cv_and_bv_rk<Air> (src/RK.cuh:230-284) and the interior part of
    // update_physical_properties<Air> (src/FieldOperation.cuh:107-140)
    const double dt = *A.dt;
    const double dt_div_jac = dt / B.jac[p];
    const double* qn = B.cv[0];
    const double* cvi = B.cv[A.in];
    double* cvo = B.cv[A.out];
    double c[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      c[l] = A.ra * qn[p + l * s] + A.rb * cvi[p + l * s] + A.rc * dt_div_jac * dq[l];
      cvo[p + l * s] = c[l];
    }
    double* bvo = B.bv[A.out];
    const double density_inv = 1.0 / c[0];
    const double u = c[1] * density_inv, v = c[2] * density_inv, w = c[3] * density_inv;
    const double V2 = u * u + v * v + w * w;
    const double pr = (kGamma - 1) * (c[4] - 0.5 * c[0] * V2);
    const double T = pr * kMwAir * density_inv / kRu;
    bvo[p] = c[0];
    bvo[p + s] = u;
    bvo[p + 2 * s] = v;
    bvo[p + 3 * s] = w;
    bvo[p + 4 * s] = pr;
    bvo[p + 5 * s] = T;
    constexpr double c_temp{kGamma * kRu / kMwAir};
    B.mul[A.out][p] = sutherland(T);
    B.mach[p] = sqrt(V2) / sqrt(c_temp * T);
  }
}

// local_time_step + global MIN (src/TimeAdvanceFunc.cuh:28-83,
// src/TimeAdvanceFunc.cu:39-89) over every block's interior, reduced with
// 64-lane shuffles and one 64-bit atomicMin per workgroup on the bit pattern
// of the (positive) time step.
__device__ double local_dt_point_f(const DevBlock& B, const Params& P, long long p) {
  const long long s = B.sz;
  const double* m = B.met;
  const double gxi = sqrt(m[p] * m[p] + m[p + s] * m[p + s] + m[p + 2 * s] * m[p + 2 * s]);
  const double geta = sqrt(m[p + 3 * s] * m[p + 3 * s] + m[p + 4 * s] * m[p + 4 * s] + m[p + 5 * s] * m[p + 5 * s]);
  const double gzeta = sqrt(m[p + 6 * s] * m[p + 6 * s] + m[p + 7 * s] * m[p + 7 * s] + m[p + 8 * s] * m[p + 8 * s]);
  const double* bv = B.bv[0];
  const double u = bv[p + s], v = bv[p + 2 * s], w = bv[p + 3 * s];
  const double U = u * m[p] + v * m[p + s] + w * m[p + 2 * s];
  const double V = u * m[p + 3 * s] + v * m[p + 4 * s] + w * m[p + 5 * s];
  const double Wc = u * m[p + 6 * s] + v * m[p + 7 * s] + w * m[p + 8 * s];
  const double a = sqrt(kGamma * kRu * bv[p + 5 * s] / kMwAir);
  const double sr_inv = (fabs(U) + a * gxi) + (fabs(V) + a * geta) + (fabs(Wc) + a * gzeta);
  const double coeff_1 = (kGamma > 4.0 / 3.0 ? kGamma : 4.0 / 3.0) / bv[p];
  const double coeff_2 = B.mul[0][p] / P.prandtl;
  double sr_visc = gxi * gxi + geta * geta;
  sr_visc += gzeta * gzeta;
  sr_visc *= coeff_1 * coeff_2;
  return P.cfl / (sr_inv + sr_visc);
}

__global__ void k_dt_init(unsigned long long* bits) { *bits = (unsigned long long)__double_as_longlong(1e6); }

template <int TX, int TY, int TZ>
__global__ __launch_bounds__(TX * TY * TZ) void k_dt_tiles(FusedArgs A, unsigned long long* bits) {
  constexpr int NT = TX * TY * TZ;
  __shared__ double part[NT / 64];
  const int tid = threadIdx.x;
  const int ti = tid % TX, tj = (tid / TX) % TY, tk = tid / (TX * TY);
  double v = 1e6;
  for (int tile = blockIdx.x; tile < A.n_tiles; tile += gridDim.x) {
    const DevBlock& B = A.blocks[A.tile_block[tile]];
    const int local = tile - B.tile_base;
    const int i = (local % B.ntx) * TX + ti, j = ((local / B.ntx) % B.nty) * TY + tj,
              k = (local / (B.ntx * B.nty)) * TZ + tk;
    if (i < B.mx && j < B.my && k < B.mz) v = fmin(v, local_dt_point_f(B, A.P, pidx(B, i, j, k)));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  if ((tid & 63) == 0) part[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < NT / 64; ++w) v = fmin(v, part[w]);
    v = fmin(v, part[0]);
    atomicMin(bits, (unsigned long long)__double_as_longlong(v));
  }
}

// ---------------------------------------------------------- shape table
#define OSTC_SHAPES(X) X(16, 4, 4, 1) X(8, 8, 4, 1) X(8, 4, 4, 1) X(32, 4, 2, 1) X(16, 8, 2, 1) X(8, 8, 8, 1)

static const TileShape kShapes[] = {
#define X(a, b, c, w) {a, b, c, w},
    OSTC_SHAPES(X)
#undef X
};
static constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

int fused_shape_count(int) { return kNumShapes; }
TileShape fused_shape(int, int index) { return kShapes[index]; }

template <int TX, int TY, int TZ, int WPE, int SCHEME>
static int launch_shape(int mode, int visc, const FusedArgs& a, int n_wg, hipStream_t st) {
  constexpr int bytes = FusedLds<TX, TY, TZ, SCHEME>::BYTES;
  constexpr int vbytes = FusedLds<TX, TY, TZ, SCHEME>::VBYTES;
  if (bytes > 160 * 1024) return -1;
  static bool attr_set = false;  // dynamic LDS above 64 KiB has to be enabled per kernel
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fused<TX, TY, TZ, WPE, SCHEME, 1, 2>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, vbytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fused<TX, TY, TZ, WPE, SCHEME, 1, 0>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fused<TX, TY, TZ, WPE, SCHEME, 0, 0>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fused<TX, TY, TZ, WPE, SCHEME, 1, 1>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fused<TX, TY, TZ, WPE, SCHEME, 0, 1>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    attr_set = true;
  }
  const dim3 grid(n_wg), block(TX * TY * TZ);
  if (mode == 2 && visc) {  // split stage
    hipLaunchKernelGGL((k_fused<TX, TY, TZ, WPE, SCHEME, 0, 1>), grid, block, bytes, st, a);
    hipLaunchKernelGGL((k_fused<TX, TY, TZ, WPE, SCHEME, 1, 2>), grid, block, vbytes, st, a);
    return bytes;
  }
  if (mode == 2) mode = 0;
  if (mode == 0 && visc) hipLaunchKernelGGL((k_fused<TX, TY, TZ, WPE, SCHEME, 1, 0>), grid, block, bytes, st, a);
  else if (mode == 0) hipLaunchKernelGGL((k_fused<TX, TY, TZ, WPE, SCHEME, 0, 0>), grid, block, bytes, st, a);
  else if (visc) hipLaunchKernelGGL((k_fused<TX, TY, TZ, WPE, SCHEME, 1, 1>), grid, block, bytes, st, a);
  else hipLaunchKernelGGL((k_fused<TX, TY, TZ, WPE, SCHEME, 0, 1>), grid, block, bytes, st, a);
  return bytes;
}

template <int TX, int TY, int TZ, int WPE>
static int launch_scheme(int mode, const FusedArgs& a, int n_wg, hipStream_t st) {
  switch (a.P.scheme) {
    case 51: return launch_shape<TX, TY, TZ, WPE, 51>(mode, a.P.visc, a, n_wg, st);
    case 50: return launch_shape<TX, TY, TZ, WPE, 50>(mode, a.P.visc, a, n_wg, st);
    case 71: return launch_shape<TX, TY, TZ, WPE, 71>(mode, a.P.visc, a, n_wg, st);
    default: return -1;
  }
}

int fused_launch(int index, int mode, const FusedArgs& a, int n_wg, hipStream_t st) {
  int i = 0;
#define X(a_, b_, c_, w_) \
  if (index == i) return launch_scheme<a_, b_, c_, w_>(mode, a, n_wg, st); \
  ++i;
  OSTC_SHAPES(X)
#undef X
  return -1;
}

int fused_lds_bytes(int scheme, int index) {
  int i = 0;
#define X(a_, b_, c_, w_)                                                        \
  if (index == i) {                                                          \
    if (scheme == 71) return FusedLds<a_, b_, c_, 71>::BYTES;                \
    return FusedLds<a_, b_, c_, 51>::BYTES;                                  \
  }                                                                          \
  ++i;
  OSTC_SHAPES(X)
#undef X
  return -1;
}

void dt_batched(const FusedArgs& a, int tx, int ty, int tz, int n_wg, double* dt_bits, hipStream_t st,
                long long* n) {
  auto* bits = reinterpret_cast<unsigned long long*>(dt_bits);
  hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, st, bits);
  int i = 0;
  bool done = false;
#define X(a_, b_, c_, w_)                                                                              \
  if (!done && tx == a_ && ty == b_ && tz == c_) {                                                 \
    hipLaunchKernelGGL((k_dt_tiles<a_, b_, c_>), dim3(n_wg), dim3(a_ * b_ * c_), 0, st, a, bits); \
    done = true;                                                                                   \
  }                                                                                                \
  ++i;
  OSTC_SHAPES(X)
#undef X
  (void)i;
  *n += 2;
}

}  // namespace ostc
