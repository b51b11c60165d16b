// ostc_kernels_fused.hip -- one persistent kernel per RK stage over the tiles
// of every block: inviscid WENO fluxes along x, y, z and second-order viscous
// fluxes staged through LDS, the residual dq kept in registers, and the
// SSP-RK3 update, primitives and Sutherland viscosity written out (with the
// sponge layer's relaxation term first where a block has one).  Reads
// ring slot `in` (state incl. ghosts) and qn = slot 0, writes the interior of
// slot `out`; ghost layers are filled afterwards by the op launches.
//
// Work decomposition: a tile is TX x TY x TZ cells of one block, one thread
// per cell in the cell phases.  Each inviscid direction phase computes split
// fluxes F+/F- for the tile's (T_d + 2R) x (columns) points, then the WENO
// flux of every (variable, face) item, then each cell subtracts its face
// difference.  The viscous phase computes the faces of all three directions
// as one work list.  Items rather than faces or cells are dealt to the
// threads, so no wave runs a tail round alone while the others wait at the
// next barrier.  Accumulation order follows the reference:
// dq = ((0 - dFx) - dFy) - dFz, then + dGx + dGy + dGz (src/RK.cuh:113-139).
#include "ostc_kernels.hpp"

namespace ostc {

template <int TX, int TY, int TZ, int D>
struct Dir {
  static constexpr int NT = TX * TY * TZ;
  static constexpr int TD = D == 0 ? TX : (D == 1 ? TY : TZ);
  static constexpr int PC = NT / TD;  // columns
  static constexpr int NF = (TD + 1) * PC;
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
  // face f of the tile -> (column, position of the face in the column)
  __device__ static void face(int f, int& col, int& sf) {
    if (D == 0) { col = f / (TD + 1); sf = f % (TD + 1); } else { sf = f / PC; col = f % PC; }
  }
};

// LDS words of one workgroup.  Inviscid phase of direction d: F+ and F- of
// the 5 variables at the (T_d + 2R) x columns points, plus the WENO epsilon
// of every face.  Viscous phase: u, v, w, T, mu on the tile box (radius 1)
// and the 4 viscous fluxes of the faces of all three directions.
//
// Two layouts of the inviscid phase.  Compact: the face fluxes overwrite the
// split fluxes after a barrier, so each thread keeps its items' results in
// registers until then.  Separate (SEP): the face fluxes and the three
// per-class epsilons of every face have their own words, so each item writes
// its result straight to LDS.  And two of the viscous phase: one work list
// for the faces of all three directions, or (VSPLIT) the x and y faces, then
// the z faces, with flux words for the larger pass only.  Each choice takes
// the smaller layout only when the larger one would hold fewer workgroups on
// a CU than the register budget WPE (waves per SIMD) allows.
template <int TX, int TY, int TZ, int SCHEME, int WPE>
struct FusedLds {
  static constexpr int NT = TX * TY * TZ;
  static constexpr int R = Stencil<SCHEME>::R;
  static constexpr int pts(int td) { return (td + 2 * R) * (NT / td); }
  static constexpr int nf(int td) { return (td + 1) * (NT / td); }
  static constexpr int mx3(int a, int b, int c) { return a > b ? (a > c ? a : c) : (b > c ? b : c); }
  static constexpr int mx2(int a, int b) { return a > b ? a : b; }
  static constexpr int mn2(int a, int b) { return a < b ? a : b; }
  static constexpr int CU_BYTES = 160 * 1024;
  static constexpr int WGS = mx2(1, WPE * 4 * 64 / NT);  // workgroups per CU the registers allow
  static constexpr int wgs(int words) { return mn2(WGS, CU_BYTES / (words * 8)); }
  static constexpr int IWORDS = mx3(10 * pts(TX) + nf(TX), 10 * pts(TY) + nf(TY), 10 * pts(TZ) + nf(TZ));
  static constexpr int IWORDS_SEP =
      mx3(10 * pts(TX) + 8 * nf(TX), 10 * pts(TY) + 8 * nf(TY), 10 * pts(TZ) + 8 * nf(TZ));
  static constexpr int NB = (TX + 2) * (TY + 2) * (TZ + 2);
  static constexpr int NFV = nf(TX) + nf(TY) + nf(TZ);
  static constexpr int VWORDS_ONE = 5 * NB + 4 * NFV;
  static constexpr int VWORDS_TWO = 5 * NB + 4 * mx2(nf(TX) + nf(TY), nf(TZ));
  static constexpr bool VSPLIT = wgs(mx2(IWORDS, VWORDS_TWO)) > wgs(mx2(IWORDS, VWORDS_ONE));
  static constexpr int VWORDS = VSPLIT ? VWORDS_TWO : VWORDS_ONE;
  static constexpr bool SEP = wgs(mx2(IWORDS_SEP, VWORDS)) >= wgs(mx2(IWORDS, VWORDS));
  // characteristic schemes (52, 72): split fluxes, the face states (AirFace:
  // 9 words, and the WENO epsilon) and the 5 characteristic face fluxes
  static constexpr bool CH = SCHEME == 52 || SCHEME == 72;
  static constexpr int CH_STATE = 10;
  static constexpr int IWORDS_CH = mx3(10 * pts(TX) + (CH_STATE + 5) * nf(TX), 10 * pts(TY) + (CH_STATE + 5) * nf(TY),
                                       10 * pts(TZ) + (CH_STATE + 5) * nf(TZ));
  static constexpr int WORDS = mx2(CH ? IWORDS_CH : (SEP ? IWORDS_SEP : IWORDS), VWORDS);
  static constexpr int BYTES = WORDS * 8;
  static constexpr int VBYTES = VWORDS * 8;
};

struct TileCtx {
  int i0, j0, k0;      // tile origin
  int ti, tj, tk;      // this thread's cell in the tile
  bool valid;          // cell inside the block
};

template <class T>
using gptr = __attribute__((address_space(1))) T*;

// What one stage reads and writes of a block, with the arrays as pointers
// into global memory.  The block table is itself read from memory, so the
// pointers it holds are generic to the compiler, which would then access
// every array with flat instructions: those wait on the LDS counter as well
// and need 64-bit addresses per lane.
struct StageBlock {
  int mx, my, mz, g, sx, sy, sz;  // sz fits 32 bits: the host refuses 9 sz >= 2^31
  int ntx, nty, tile_base, sp_jac;
  long long n0;
  gptr<const double> cvi, bvi, muli, qn, jac, met, sp_sigma, sp_mean;
  gptr<double> cvo, bvo, mulo, mach, dq;
};

__device__ __forceinline__ StageBlock stage_block(const DevBlock& B, int in, int out) {
  StageBlock S;
  S.mx = B.mx; S.my = B.my; S.mz = B.mz; S.g = B.g; S.sx = B.sx; S.sy = B.sy; S.sz = (int)B.sz;
  S.ntx = B.ntx; S.nty = B.nty; S.tile_base = B.tile_base; S.sp_jac = B.sp_jac;
  S.n0 = B.n0;
  S.cvi = (gptr<const double>)B.cv[in];
  S.bvi = (gptr<const double>)B.bv[in];
  S.muli = (gptr<const double>)B.mul[in];
  S.qn = (gptr<const double>)B.cv[0];
  S.jac = (gptr<const double>)B.jac;
  S.met = (gptr<const double>)B.met;
  S.sp_sigma = (gptr<const double>)B.sp_sigma;
  S.sp_mean = (gptr<const double>)B.sp_mean;
  S.cvo = (gptr<double>)B.cv[out];
  S.bvo = (gptr<double>)B.bv[out];
  S.mulo = (gptr<double>)B.mul[out];
  S.mach = (gptr<double>)B.mach;
  S.dq = (gptr<double>)B.dq;
  return S;
}

// 32-bit offset of node (i, j, k) in a padded plane; the host refuses blocks
// whose 9 metric planes would not fit 2^31 doubles.
__device__ __forceinline__ int poff(const StageBlock& b, int i, int j, int k) {
  return ((k + b.g) * b.sy + (j + b.g)) * b.sx + (i + b.g);
}

// The block coordinates of face f of direction D (its lower node L along D,
// face L+1/2) and whether the face lies in the block (faces past the end of a
// column, or in columns outside the block, are never read).
template <int TX, int TY, int TZ, int D>
__device__ __forceinline__ bool face_node(const StageBlock& B, const TileCtx& t, int f, int ijk[3]) {
  using Dr = Dir<TX, TY, TZ, D>;
  int col, sf, a, b;
  Dr::face(f, col, sf);
  Dr::col_coords(col, a, b);
  if (D == 0) { ijk[1] = t.j0 + a; ijk[2] = t.k0 + b; }
  if (D == 1) { ijk[0] = t.i0 + a; ijk[2] = t.k0 + b; }
  if (D == 2) { ijk[0] = t.i0 + a; ijk[1] = t.j0 + b; }
  const int origin = D == 0 ? t.i0 : (D == 1 ? t.j0 : t.k0);
  const int m = D == 0 ? B.mx : (D == 1 ? B.my : B.mz);
  ijk[D] = origin + sf - 1;
  return ijk[D] <= m - 1 && ijk[0] < B.mx + (D == 0 ? 1 : 0) && ijk[1] < B.my + (D == 1 ? 1 : 0) &&
         ijk[2] < B.mz + (D == 2 ? 1 : 0);
}

// Characteristic WENO along direction D (schemes 52 / 72,
// compute_weno_flux_ch<Air> "method 1", src/WENO.cu:1147-1453), after the
// split fluxes are staged: (1) the Roe-averaged state of every face
// (AirFace) and its epsilon into LDS; (2) (field, face) items: the left
// eigenvector row of the field projects the W split fluxes of each side, and
// WENO-Z gives the characteristic flux; (3) each cell forms the flux of its
// two faces from the right eigenvectors and subtracts their difference.  The
// arithmetic is that of the per-block kernel (k_weno_pencil), element by
// element.
template <int TX, int TY, int TZ, int SCHEME, int D>
__device__ __forceinline__ void characteristic_faces(double* smem, const StageBlock& B, const Params& P,
                                                     const TileCtx& t, double dq[5]) {
  using Dr = Dir<TX, TY, TZ, D>;
  constexpr int NT = Dr::NT, R = Stencil<SCHEME>::R, W = Stencil<SCHEME>::W;
  constexpr int TD = Dr::TD, LP = TD + 2 * R, NPTS = LP * Dr::PC, NF = Dr::NF, NI = 5 * NF;
  constexpr int ST = 9;  // words of a face state: vel[3], enthalpy, c_sq, c, n[3]
  static_assert(FusedLds<TX, TY, TZ, SCHEME, 1>::CH_STATE == ST + 1, "LDS words per face: state and epsilon");
  const int tid = threadIdx.x;
  const int s = B.sz;
  const double* sFp = smem;
  const double* sFm = smem + 5 * NPTS;
  double* sSt = smem + 10 * NPTS;        // [ST][NF] face states
  double* sEps = sSt + ST * NF;          // [NF] WENO epsilon of each face
  double* sCh = sEps + NF;               // [5][NF] characteristic face fluxes
  for (int f = tid; f < NF; f += NT) {
    int ijk[3];
    if (!face_node<TX, TY, TZ, D>(B, t, f, ijk)) continue;
    const int p1 = poff(B, ijk[0], ijk[1], ijk[2]);
    const int p2 = p1 + (D == 0 ? 1 : (D == 1 ? B.sx : B.sx * B.sy));
    double cvl[6], cvr[6];
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      cvl[l] = B.cvi[p1 + l * s];
      cvr[l] = B.cvi[p2 + l * s];
    }
    cvl[5] = B.bvi[p1 + 4 * s];
    cvr[5] = B.bvi[p2 + 4 * s];
    const double m1[3] = {B.met[p1 + 3 * D * s], B.met[p1 + (3 * D + 1) * s], B.met[p1 + (3 * D + 2) * s]};
    const double m2[3] = {B.met[p2 + 3 * D * s], B.met[p2 + (3 * D + 1) * s], B.met[p2 + (3 * D + 2) * s]};
    const double j1 = B.jac[p1], j2 = B.jac[p2];
    AirFace F;
    air_face_state(cvl, cvr, m1, m2, j1, j2, F);
    sSt[0 * NF + f] = F.vel[0];
    sSt[1 * NF + f] = F.vel[1];
    sSt[2 * NF + f] = F.vel[2];
    sSt[3 * NF + f] = F.enthalpy;
    sSt[4 * NF + f] = F.c_sq;
    sSt[5 * NF + f] = F.c;
    sSt[6 * NF + f] = F.n[0];
    sSt[7 * NF + f] = F.n[1];
    sSt[8 * NF + f] = F.n[2];
    constexpr double eps{1e-40};
    const double ax = m1[0] * j1 + m2[0] * j2, ay = m1[1] * j1 + m2[1] * j2, az = m1[2] * j1 + m2[2] * j2;
    sEps[f] = eps * P.weno_eps_scale * 0.25 * (ax * ax + ay * ay + az * az);
  }
  __syncthreads();
  auto load_face = [&](int f, AirFace& F) {
    F.vel[0] = sSt[0 * NF + f];
    F.vel[1] = sSt[1 * NF + f];
    F.vel[2] = sSt[2 * NF + f];
    F.enthalpy = sSt[3 * NF + f];
    F.c_sq = sSt[4 * NF + f];
    F.c = sSt[5 * NF + f];
    F.n[0] = sSt[6 * NF + f];
    F.n[1] = sSt[7 * NF + f];
    F.n[2] = sSt[8 * NF + f];
  };
#pragma unroll 1
  for (int it = tid; it < NI; it += NT) {
    const int l = it / NF, f = it - l * NF;
    double val = 0.0;
    int ijk[3];
    if (face_node<TX, TY, TZ, D>(B, t, f, ijk)) {
      int col, sf;
      Dr::face(f, col, sf);
      const int sL = sf - 1 + R;
      AirFace F;
      load_face(f, F);
      double Lr[5];
      air_left_row(F, l, Lr);
      double vp[W], vm[W];
#pragma unroll
      for (int r = 0; r < W; ++r) {
        const int ip = Dr::template at<LP>(col, sL - W / 2 + r), im = Dr::template at<LP>(col, sL - W / 2 + 1 + r);
        vp[r] = Lr[0] * sFp[ip];
        vm[r] = Lr[0] * sFm[im];
#pragma unroll
        for (int n = 1; n < 5; ++n) {
          vp[r] += Lr[n] * sFp[n * NPTS + ip];
          vm[r] += Lr[n] * sFm[n * NPTS + im];
        }
      }
      if constexpr (SCHEME == 72) val = weno7z(vp, vm, sEps[f]);
      else val = weno5z(vp, vm, sEps[f]);
    }
    sCh[it] = val;
  }
  __syncthreads();
  if (t.valid) {
    const int col = Dr::my_col(t.ti, t.tj, t.tk), pos = Dr::my_pos(t.ti, t.tj, t.tk);
    const int fs[2] = {Dr::template at<TD + 1>(col, pos), Dr::template at<TD + 1>(col, pos + 1)};
    double fc[2][5];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int f = fs[side];
      AirFace F;
      load_face(f, F);
      double fch[5];
#pragma unroll
      for (int w = 0; w < 5; ++w) fch[w] = sCh[w * NF + f];
      double Rc[5];
      air_right_column(F, 0, Rc);
#pragma unroll
      for (int r = 0; r < 5; ++r) fc[side][r] = Rc[r] * fch[0];
#pragma unroll
      for (int w = 1; w < 5; ++w) {
        air_right_column(F, w, Rc);
#pragma unroll
        for (int r = 0; r < 5; ++r) fc[side][r] += Rc[r] * fch[w];
      }
    }
#pragma unroll
    for (int l = 0; l < 5; ++l) dq[l] -= fc[1][l] - fc[0][l];
  }
  __syncthreads();
}

// Inviscid fluxes along direction D, subtracted from dq.  The face work is
// spread over (variable, face) items, 5 NF of them for NT threads, so every
// wave carries the same share whatever NF mod NT is.
template <int TX, int TY, int TZ, int SCHEME, int WPE, int D>
__device__ __forceinline__ void inviscid_phase(double* smem, const StageBlock& B, const Params& P,
                                               const TileCtx& t, double dq[5]) {
  using Dr = Dir<TX, TY, TZ, D>;
  constexpr int NT = Dr::NT, R = Stencil<SCHEME>::R, W = Stencil<SCHEME>::W;
  constexpr int TD = Dr::TD, PC = Dr::PC, LP = TD + 2 * R, NPTS = LP * PC, NF = Dr::NF;
  constexpr int NI = 5 * NF, IPT = (NI + NT - 1) / NT;
  constexpr bool SEP = FusedLds<TX, TY, TZ, SCHEME, WPE>::SEP;
  const int tid = threadIdx.x;
  const int origin = D == 0 ? t.i0 : (D == 1 ? t.j0 : t.k0);
  const gptr<const double> cv = B.cvi;
  const gptr<const double> bv = B.bvi;
  const int s = B.sz;
  double* sFp = smem;
  double* sFm = smem + 5 * NPTS;
  double* sEps = smem + 10 * NPTS;  // SEP: [3][NF], the epsilon of rho, of the momenta and of the energy

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
    ijk[D] = origin + sp - R;
    const bool ok = ijk[0] < B.mx + (D == 0 ? R : 0) && ijk[1] < B.my + (D == 1 ? R : 0) &&
                    ijk[2] < B.mz + (D == 2 ? R : 0);
    if (ok) {
      const int q = poff(B, ijk[0], ijk[1], ijk[2]);
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
  if constexpr (Stencil<SCHEME>::CH) {
    characteristic_faces<TX, TY, TZ, SCHEME, D>(smem, B, P, t, dq);
    return;
  }
  // the epsilon of every face (src/WENO.cu:1469-1482), shared by its 5 items
  for (int f = tid; f < NF; f += NT) {
    int ijk[3];
    if (face_node<TX, TY, TZ, D>(B, t, f, ijk)) {
      const int p1 = poff(B, ijk[0], ijk[1], ijk[2]);
      const int p2 = p1 + (D == 0 ? 1 : (D == 1 ? B.sx : B.sx * B.sy));
      const double j1 = B.jac[p1], j2 = B.jac[p2];
      double eps3[3];
      weno_eps(P, B.met[p1 + 3 * D * s] * j1, B.met[p1 + (3 * D + 1) * s] * j1, B.met[p1 + (3 * D + 2) * s] * j1,
               B.met[p2 + 3 * D * s] * j2, B.met[p2 + (3 * D + 1) * s] * j2, B.met[p2 + (3 * D + 2) * s] * j2,
               eps3);
      sEps[f] = eps3[0];
      if (SEP) {
        sEps[NF + f] = eps3[1];
        sEps[2 * NF + f] = eps3[2];
      }
    }
  }
  __syncthreads();

  if constexpr (SEP) {
    double* sFc = smem + 10 * NPTS + 3 * NF;  // [5][NF]
#pragma unroll 1
    for (int it = tid; it < NI; it += NT) {
      const int l = it / NF, f = it - l * NF;
      double val = 0.0;
      int ijk[3];
      if (face_node<TX, TY, TZ, D>(B, t, f, ijk)) {
        int col, sf;
        Dr::face(f, col, sf);
        const int sL = sf - 1 + R;
        const double e = sEps[(l == 0 ? 0 : (l == 4 ? 2 : 1)) * NF + f];
        const double* fpl = sFp + l * NPTS;
        const double* fml = sFm + l * NPTS;
        double vp[W], vm[W];
#pragma unroll
        for (int r = 0; r < W; ++r) {
          vp[r] = fpl[Dr::template at<LP>(col, sL - W / 2 + r)];
          vm[r] = fml[Dr::template at<LP>(col, sL - W / 2 + 1 + r)];
        }
        val = weno_face<SCHEME>(vp, vm, e);
      }
      sFc[it] = val;
    }
    __syncthreads();
    if (t.valid) {
      const int col = Dr::my_col(t.ti, t.tj, t.tk), pos = Dr::my_pos(t.ti, t.tj, t.tk);
      const int fL = Dr::template at<TD + 1>(col, pos), fR = Dr::template at<TD + 1>(col, pos + 1);
#pragma unroll
      for (int l = 0; l < 5; ++l) dq[l] -= sFc[l * NF + fR] - sFc[l * NF + fL];
    }
    __syncthreads();
    return;
  }

  // one item per iteration (the loop is not unrolled, so the WENO evaluations
  // of different items do not overlap in registers); each result is kept in
  // its own register through a chain of selects
  double fres[IPT];
#pragma unroll
  for (int u = 0; u < IPT; ++u) fres[u] = 0.0;
#pragma unroll 1
  for (int u = 0; u < IPT; ++u) {
    const int it = tid + u * NT;
    double val = 0.0;
    if (it < NI) {
      const int l = it / NF, f = it - l * NF;
      int ijk[3];
      if (face_node<TX, TY, TZ, D>(B, t, f, ijk)) {
        int col, sf;
        Dr::face(f, col, sf);
        const int sL = sf - 1 + R;
        // eps_scaled of src/WENO.cu:1478-1481 for variable l
        const double e0 = sEps[f];
        const double e = l == 0 ? e0 : (l == 4 ? e0 * P.v_ref * P.v_ref * P.v_ref * P.v_ref : e0 * P.v_ref * P.v_ref);
        const double* fpl = sFp + l * NPTS;
        const double* fml = sFm + l * NPTS;
        double vp[W], vm[W];
#pragma unroll
        for (int r = 0; r < W; ++r) {
          vp[r] = fpl[Dr::template at<LP>(col, sL - W / 2 + r)];
          vm[r] = fml[Dr::template at<LP>(col, sL - W / 2 + 1 + r)];
        }
        val = weno_face<SCHEME>(vp, vm, e);
      }
    }
#pragma unroll
    for (int q = 0; q < IPT; ++q) fres[q] = q == u ? val : fres[q];
  }
  __syncthreads();
  double* sFc = smem;  // [5][NF], aliases the split fluxes, which are no longer read
#pragma unroll
  for (int u = 0; u < IPT; ++u) {
    const int it = tid + u * NT;
    if (it < NI) sFc[it] = fres[u];
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

// Viscous flux of face f of direction D (src/ViscousScheme.cuh:23-909),
// written to sFv[l * NF + f]: u, v, w, T and mu from the staged tile box,
// the metric and Jacobian of the two face cells from global memory.
template <int TX, int TY, int TZ, int D>
__device__ __forceinline__ void viscous_face(const double* sV, double* sFv, const StageBlock& B, const Params& P,
                                             const TileCtx& t, int f) {
  using Dr = Dir<TX, TY, TZ, D>;
  constexpr int NF = Dr::NF, BX = TX + 2, BY = TY + 2, NB = (TX + 2) * (TY + 2) * (TZ + 2);
  int ijk[3];
  if (!face_node<TX, TY, TZ, D>(B, t, f, ijk)) return;
  const int s = B.sz;
  auto box = [&](int i, int j, int k) { return ((k - t.k0 + 1) * BY + (j - t.j0 + 1)) * BX + (i - t.i0 + 1); };
  auto pv = [&](int i, int j, int k, int l) {
    const int q = l == 5 ? 3 : l - 1;
    return sV[q * NB + box(i, j, k)];
  };
  const int i = ijk[0], j = ijk[1], k = ijk[2];
  const int iB = i + (D == 0), jB = j + (D == 1), kB = k + (D == 2);
  const int pA = poff(B, i, j, k);
  const int pB = pA + (D == 0 ? 1 : (D == 1 ? B.sx : B.sx * B.sy));
  double fm[9], kr[3];
  viscous_face_metric<D>(B.met, s, pA, pB, B.jac[pA], B.jac[pB], fm, kr);
  // one variable at a time: its differences along xi, eta, zeta, then its
  // physical gradient, so only the gradients stay live
  double g[4][3];
  const int var[4] = {1, 2, 3, 5};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int l = var[q];
    double d[3];
    if (D == 0) {
      d[0] = pv(i + 1, j, k, l) - pv(i, j, k, l);
      d[1] = 0.25 * (pv(i, j + 1, k, l) - pv(i, j - 1, k, l) + pv(i + 1, j + 1, k, l) - pv(i + 1, j - 1, k, l));
      d[2] = 0.25 * (pv(i, j, k + 1, l) - pv(i, j, k - 1, l) + pv(i + 1, j, k + 1, l) - pv(i + 1, j, k - 1, l));
    } else if (D == 1) {
      d[0] = 0.25 * (pv(i + 1, j, k, l) - pv(i - 1, j, k, l) + pv(i + 1, j + 1, k, l) - pv(i - 1, j + 1, k, l));
      d[1] = pv(i, j + 1, k, l) - pv(i, j, k, l);
      d[2] = 0.25 * (pv(i, j, k + 1, l) - pv(i, j, k - 1, l) + pv(i, j + 1, k + 1, l) - pv(i, j + 1, k - 1, l));
    } else {
      d[0] = 0.25 * (pv(i + 1, j, k, l) - pv(i - 1, j, k, l) + pv(i + 1, j, k + 1, l) - pv(i - 1, j, k + 1, l));
      d[1] = 0.25 * (pv(i, j + 1, k, l) - pv(i, j - 1, k, l) + pv(i, j + 1, k + 1, l) - pv(i, j - 1, k + 1, l));
      d[2] = pv(i, j, k + 1, l) - pv(i, j, k, l);
    }
    viscous_grad(d, fm, g[q]);
    // keep the scheduler from hoisting the next variable's LDS reads above
    // this one's arithmetic (they would all be live at once)
    __builtin_amdgcn_sched_barrier(0);
  }
  const double um = 0.5 * (pv(i, j, k, 1) + pv(iB, jB, kB, 1));
  const double vm = 0.5 * (pv(i, j, k, 2) + pv(iB, jB, kB, 2));
  const double wm = 0.5 * (pv(i, j, k, 3) + pv(iB, jB, kB, 3));
  const double muA = sV[4 * NB + box(i, j, k)];
  const double muB = sV[4 * NB + box(iB, jB, kB)];
  double fv[4];
  viscous_flux_from_grad(g, kr, muA, muB, um, vm, wm, P.prandtl, fv);
#pragma unroll
  for (int l = 0; l < 4; ++l) sFv[l * NF + f] = fv[l];
}

// Viscous fluxes of the three directions, added to dq in the reference's
// order (x, then y, then z).  The faces of all three directions form one
// work list, so the threads share them evenly (VSPLIT: two lists, the x and y
// faces, then the z faces, reusing the flux words).
template <int TX, int TY, int TZ, bool VSPLIT>
__device__ __forceinline__ void viscous_phase(double* smem, const StageBlock& B, const Params& P,
                                              const TileCtx& t, double dq[5]) {
  constexpr int NT = TX * TY * TZ, BX = TX + 2, BY = TY + 2, NB = (TX + 2) * (TY + 2) * (TZ + 2);
  using D0 = Dir<TX, TY, TZ, 0>;
  using D1 = Dir<TX, TY, TZ, 1>;
  using D2 = Dir<TX, TY, TZ, 2>;
  constexpr int NF0 = D0::NF, NF1 = D1::NF, NF2 = D2::NF, NFV = NF0 + NF1 + NF2;
  const gptr<const double> bv = B.bvi;
  const gptr<const double> mu = B.muli;
  const int s = B.sz;
  for (int p = threadIdx.x; p < NB; p += NT) {
    const int a = p % BX, b = (p / BX) % BY, c = p / (BX * BY);
    const int i = t.i0 - 1 + a, j = t.j0 - 1 + b, k = t.k0 - 1 + c;
    if (i <= B.mx && j <= B.my && k <= B.mz) {
      const int q = poff(B, i, j, k);
      smem[p] = bv[q + s];
      smem[NB + p] = bv[q + 2 * s];
      smem[2 * NB + p] = bv[q + 3 * s];
      smem[3 * NB + p] = bv[q + 5 * s];
      smem[4 * NB + p] = mu[q];
    }
  }
  __syncthreads();
  double* sF0 = smem + 5 * NB;
  double* sF1 = sF0 + 4 * NF0;
  double* sF2 = VSPLIT ? sF0 : sF1 + 4 * NF1;
  if constexpr (VSPLIT) {
#pragma unroll 1
    for (int it = threadIdx.x; it < NF0 + NF1; it += NT) {
      if (it < NF0) viscous_face<TX, TY, TZ, 0>(smem, sF0, B, P, t, it);
      else viscous_face<TX, TY, TZ, 1>(smem, sF1, B, P, t, it - NF0);
    }
    __syncthreads();
    if (t.valid) {
      {
        const int col = D0::my_col(t.ti, t.tj, t.tk), pos = D0::my_pos(t.ti, t.tj, t.tk);
        const int fL = D0::template at<D0::TD + 1>(col, pos), fR = D0::template at<D0::TD + 1>(col, pos + 1);
#pragma unroll
        for (int l = 0; l < 4; ++l) dq[l + 1] += sF0[l * NF0 + fR] - sF0[l * NF0 + fL];
      }
      {
        const int col = D1::my_col(t.ti, t.tj, t.tk), pos = D1::my_pos(t.ti, t.tj, t.tk);
        const int fL = D1::template at<D1::TD + 1>(col, pos), fR = D1::template at<D1::TD + 1>(col, pos + 1);
#pragma unroll
        for (int l = 0; l < 4; ++l) dq[l + 1] += sF1[l * NF1 + fR] - sF1[l * NF1 + fL];
      }
    }
    __syncthreads();
#pragma unroll 1
    for (int it = threadIdx.x; it < NF2; it += NT) viscous_face<TX, TY, TZ, 2>(smem, sF2, B, P, t, it);
    __syncthreads();
    if (t.valid) {
      const int col = D2::my_col(t.ti, t.tj, t.tk), pos = D2::my_pos(t.ti, t.tj, t.tk);
      const int fL = D2::template at<D2::TD + 1>(col, pos), fR = D2::template at<D2::TD + 1>(col, pos + 1);
#pragma unroll
      for (int l = 0; l < 4; ++l) dq[l + 1] += sF2[l * NF2 + fR] - sF2[l * NF2 + fL];
    }
    __syncthreads();
    return;
  }
#pragma unroll 1
  for (int it = threadIdx.x; it < NFV; it += NT) {
    if (it < NF0) viscous_face<TX, TY, TZ, 0>(smem, sF0, B, P, t, it);
    else if (it < NF0 + NF1) viscous_face<TX, TY, TZ, 1>(smem, sF1, B, P, t, it - NF0);
    else viscous_face<TX, TY, TZ, 2>(smem, sF2, B, P, t, it - NF0 - NF1);
  }
  __syncthreads();
  if (t.valid) {
    {
      const int col = D0::my_col(t.ti, t.tj, t.tk), pos = D0::my_pos(t.ti, t.tj, t.tk);
      const int fL = D0::template at<D0::TD + 1>(col, pos), fR = D0::template at<D0::TD + 1>(col, pos + 1);
#pragma unroll
      for (int l = 0; l < 4; ++l) dq[l + 1] += sF0[l * NF0 + fR] - sF0[l * NF0 + fL];
    }
    {
      const int col = D1::my_col(t.ti, t.tj, t.tk), pos = D1::my_pos(t.ti, t.tj, t.tk);
      const int fL = D1::template at<D1::TD + 1>(col, pos), fR = D1::template at<D1::TD + 1>(col, pos + 1);
#pragma unroll
      for (int l = 0; l < 4; ++l) dq[l + 1] += sF1[l * NF1 + fR] - sF1[l * NF1 + fL];
    }
    {
      const int col = D2::my_col(t.ti, t.tj, t.tk), pos = D2::my_pos(t.ti, t.tj, t.tk);
      const int fL = D2::template at<D2::TD + 1>(col, pos), fR = D2::template at<D2::TD + 1>(col, pos + 1);
#pragma unroll
      for (int l = 0; l < 4; ++l) dq[l + 1] += sF2[l * NF2 + fR] - sF2[l * NF2 + fL];
    }
  }
  __syncthreads();
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
    // the descriptor is read from memory once per tile into the stage's view
    const StageBlock B = stage_block(A.blocks[b], A.in, A.out);
    const int local = tile - B.tile_base;
    t.i0 = (local % B.ntx) * TX;
    t.j0 = ((local / B.ntx) % B.nty) * TY;
    t.k0 = (local / (B.ntx * B.nty)) * TZ;
    const int i = t.i0 + t.ti, j = t.j0 + t.tj, k = t.k0 + t.tk;
    t.valid = i < B.mx && j < B.my && k < B.mz;
    const long long q0 = ((long long)k * B.my + j) * B.mx + i;  // interior index (dq, sponge planes)
    double dq[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (MODE != 2 && B.sp_sigma && t.valid) {
      // sponge_layer_source (src/SpongeLayer.cu:68-160), the first term of the
      // stage's dq as in compute_source: dq -= sigma (cv - mean)
      const int p = poff(B, i, j, k), s = B.sz;
      const double sigma = B.sp_jac ? B.sp_sigma[q0] * B.jac[p] : B.sp_sigma[q0];
#pragma unroll
      for (int l = 0; l < 5; ++l) dq[l] -= sigma * (B.cvi[p + l * s] - B.sp_mean[q0 + l * B.n0]);
    }
    if (MODE != 2) {
      inviscid_phase<TX, TY, TZ, SCHEME, WPE, 0>(smem, B, A.P, t, dq);
      inviscid_phase<TX, TY, TZ, SCHEME, WPE, 1>(smem, B, A.P, t, dq);
      inviscid_phase<TX, TY, TZ, SCHEME, WPE, 2>(smem, B, A.P, t, dq);
    } else if (t.valid) {
      // split stage: the inviscid part of dq was written by the MODE 1 launch
#pragma unroll
      for (int l = 0; l < 5; ++l) dq[l] = B.dq[q0 + l * B.n0];
    }
    if (VISC) viscous_phase<TX, TY, TZ, FusedLds<TX, TY, TZ, SCHEME, WPE>::VSPLIT>(smem, B, A.P, t, dq);
    if (!t.valid) continue;
    const int p = poff(B, i, j, k), s = B.sz;
    if (MODE == 1) {
#pragma unroll
      for (int l = 0; l < 5; ++l) B.dq[q0 + l * B.n0] = dq[l];
      continue;
    }
    // update_cv_and_bv_rk<Air> (src/RK.cuh:230-284) and the interior part of
    // update_physical_properties<Air> (src/FieldOperation.cuh:107-140)
    const double dt = *A.dt;
    const double dt_div_jac = dt / B.jac[p];
    double c[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      c[l] = A.ra * B.qn[p + l * s] + A.rb * B.cvi[p + l * s] + A.rc * dt_div_jac * dq[l];
      B.cvo[p + l * s] = c[l];
    }
    const double density_inv = 1.0 / c[0];
    const double u = c[1] * density_inv, v = c[2] * density_inv, w = c[3] * density_inv;
    const double V2 = u * u + v * v + w * w;
    const double pr = (kGamma - 1) * (c[4] - 0.5 * c[0] * V2);
    const double T = pr * kMwAir * density_inv / kRu;
    B.bvo[p] = c[0];
    B.bvo[p + s] = u;
    B.bvo[p + 2 * s] = v;
    B.bvo[p + 3 * s] = w;
    B.bvo[p + 4 * s] = pr;
    B.bvo[p + 5 * s] = T;
    constexpr double c_temp{kGamma * kRu / kMwAir};
    B.mulo[p] = sutherland(T);
    B.mach[p] = sqrt(V2) / sqrt(c_temp * T);
  }
}

// local_time_step + global MIN (src/TimeAdvanceFunc.cuh:28-83,
// src/TimeAdvanceFunc.cu:39-89) over every block's interior, reduced with
// 64-lane shuffles and one 64-bit atomicMin per workgroup on the bit pattern
// of the (positive) time step.
__device__ double local_dt_point_f(const StageBlock& B, const Params& P, int p) {
  const int s = B.sz;
  const gptr<const double> m = B.met;
  const double gxi = sqrt(m[p] * m[p] + m[p + s] * m[p + s] + m[p + 2 * s] * m[p + 2 * s]);
  const double geta = sqrt(m[p + 3 * s] * m[p + 3 * s] + m[p + 4 * s] * m[p + 4 * s] + m[p + 5 * s] * m[p + 5 * s]);
  const double gzeta = sqrt(m[p + 6 * s] * m[p + 6 * s] + m[p + 7 * s] * m[p + 7 * s] + m[p + 8 * s] * m[p + 8 * s]);
  const gptr<const double> bv = B.bvi;
  const double u = bv[p + s], v = bv[p + 2 * s], w = bv[p + 3 * s];
  const double U = u * m[p] + v * m[p + s] + w * m[p + 2 * s];
  const double V = u * m[p + 3 * s] + v * m[p + 4 * s] + w * m[p + 5 * s];
  const double Wc = u * m[p + 6 * s] + v * m[p + 7 * s] + w * m[p + 8 * s];
  const double a = sqrt(kGamma * kRu * bv[p + 5 * s] / kMwAir);
  const double sr_inv = (fabs(U) + a * gxi) + (fabs(V) + a * geta) + (fabs(Wc) + a * gzeta);
  const double coeff_1 = (kGamma > 4.0 / 3.0 ? kGamma : 4.0 / 3.0) / bv[p];
  const double coeff_2 = B.muli[p] / P.prandtl;
  double sr_visc = gxi * gxi + geta * geta;
  sr_visc += gzeta * gzeta;
  sr_visc *= coeff_1 * coeff_2;
  return P.cfl / (sr_inv + sr_visc);
}

__global__ void k_dt_init(unsigned long long* bits) { *bits = (unsigned long long)__double_as_longlong(1e6); }

// Every block's interior walked row by row (consecutive lanes, consecutive i:
// whole cache lines of the metric and primitive planes; tiles would read
// 16-double pieces of rows that straddle lines), grid-stride over the cells of
// one block after another.
__global__ __launch_bounds__(256) void k_dt_rows(FusedArgs A, int n_blocks, unsigned long long* bits) {
  __shared__ double part[4];
  const int tid = threadIdx.x;
  double v = 1e6;
  const unsigned stride = gridDim.x * 256u;
  for (int b = 0; b < n_blocks; ++b) {
    const StageBlock B = stage_block(A.blocks[b], 0, 0);  // the step's state, slot 0
    const unsigned row = (unsigned)B.mx, plane = row * (unsigned)B.my, n = plane * (unsigned)B.mz;
    for (unsigned q = blockIdx.x * 256u + tid; q < n; q += stride) {
      const unsigned k = q / plane, r = q - k * plane, j = r / row, i = r - j * row;
      v = fmin(v, local_dt_point_f(B, A.P, poff(B, (int)i, (int)j, (int)k)));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  if ((tid & 63) == 0) part[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < 4; ++w) v = fmin(v, part[w]);
    v = fmin(v, part[0]);
    atomicMin(bits, (unsigned long long)__double_as_longlong(v));
  }
}

// ---------------------------------------------------------- shape table
// (TX, TY, TZ, waves per SIMD the register allocation must allow).  Shapes
// whose LDS admits three workgroups per CU are compiled for 3 waves per SIMD.
#define OSTC_SHAPES(X) X(8, 8, 4, 3) X(16, 4, 4, 2) X(8, 4, 4, 3) X(4, 8, 8, 3) X(16, 8, 2, 2) X(8, 8, 8, 2)

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
  constexpr int bytes = FusedLds<TX, TY, TZ, SCHEME, WPE>::BYTES;
  constexpr int vbytes = FusedLds<TX, TY, TZ, SCHEME, WPE>::VBYTES;
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
    case 52: return launch_shape<TX, TY, TZ, WPE, 52>(mode, a.P.visc, a, n_wg, st);
    case 72: return launch_shape<TX, TY, TZ, WPE, 72>(mode, a.P.visc, a, n_wg, st);
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
    if (scheme == 71) return FusedLds<a_, b_, c_, 71, w_>::BYTES;            \
    if (scheme == 52) return FusedLds<a_, b_, c_, 52, w_>::BYTES;            \
    if (scheme == 72) return FusedLds<a_, b_, c_, 72, w_>::BYTES;            \
    return FusedLds<a_, b_, c_, 51, w_>::BYTES;                              \
  }                                                                          \
  ++i;
  OSTC_SHAPES(X)
#undef X
  return -1;
}

void dt_rows(const FusedArgs& a, int n_blocks, int n_wg, double* dt_bits, hipStream_t st, long long* n) {
  auto* bits = reinterpret_cast<unsigned long long*>(dt_bits);
  hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, st, bits);
  hipLaunchKernelGGL(k_dt_rows, dim3(n_wg), dim3(256), 0, st, a, n_blocks, bits);
  *n += 2;
}

}  // namespace ostc
