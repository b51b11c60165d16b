// ostc_kernels_adjoint.hip -- reverse mode of the air RHS + SSP-RK3 stage:
// vector-Jacobian products (VJPs) of the WENO-Z5 / WENO5-JS reconstruction
// with point-local Lax-Friedrichs splitting (src/WENO.cu:697-720,
// :1582-1617) and of the second-order viscous fluxes
// (src/ViscousScheme.cuh:23-909), the transposed RK3 update
// (src/RK.cuh:230-284) and boundary copies (src/BoundCond.cuh:205-233,
// :1443-1486), plus the pieces of the differentiation workload (Fourier-mode
// perturbation, momentum-thickness objective and seed, gradient projection).
//
// Every cotangent is accumulated in a fixed order (gathers, no atomics), so a
// reverse sweep is bitwise reproducible run to run.
#include <mutex>
#include <type_traits>

#include "ostc_kernels.hpp"

namespace ostc {
namespace {

constexpr int kT = 256;

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ------------------------------------------------------------ point VJPs
// Upwind side of WENO-Z5 (JS = false) or WENO5-JS (JS = true) on the stencil
// u[0..4] (u[2] is the upwind cell; linear weights 0.3, 0.6, 0.1 on the
// candidates q0, q1, q2): ub[r] += fbar * df/du[r].  The downwind side of a
// face is the same function of the reversed stencil.
template <bool JS>
__device__ __forceinline__ void weno5_side_vjp(const double* u, double eps, double fbar, double* ub) {
  constexpr double c6 = 1.0 / 6, c13 = 13.0 / 12;
  const double q0 = c6 * (2 * u[2] + 5 * u[3] - u[4]);
  const double q1 = c6 * (-u[1] + 5 * u[2] + 2 * u[3]);
  const double q2 = c6 * (2 * u[0] - 7 * u[1] + 11 * u[2]);
  const double s0 = u[2] + u[4] - 2 * u[3], t0 = 3 * u[2] - 4 * u[3] + u[4];
  const double s1 = u[1] + u[3] - 2 * u[2], t1 = u[1] - u[3];
  const double s2 = u[0] + u[2] - 2 * u[1], t2 = u[0] - 4 * u[1] + 3 * u[2];
  const double b0 = c13 * s0 * s0 + 0.25 * t0 * t0;
  const double b1 = c13 * s1 * s1 + 0.25 * t1 * t1;
  const double b2 = c13 * s2 * s2 + 0.25 * t2 * t2;
  const double e0 = 1 / (eps + b0), e1 = 1 / (eps + b1), e2 = 1 / (eps + b2);
  const double w0 = 0.3 * e0 * e0, w1 = 0.6 * e1 * e1, w2 = 0.1 * e2 * e2;
  double a0, a1, a2, tau = 0;
  if (JS) {
    a0 = w0; a1 = w1; a2 = w2;
  } else {
    tau = (b0 - b2) * (b0 - b2);
    a0 = 0.3 + tau * w0; a1 = 0.6 + tau * w1; a2 = 0.1 + tau * w2;
  }
  const double iA = 1 / (a0 + a1 + a2);
  const double f = (a0 * q0 + a1 * q1 + a2 * q2) * iA;
  const double g = fbar * iA;
  const double qb0 = g * a0, qb1 = g * a1, qb2 = g * a2;
  const double ab0 = g * (q0 - f), ab1 = g * (q1 - f), ab2 = g * (q2 - f);
  // a_k = d_k e_k^2 (JS) or d_k (1 + tau e_k^2) (Z): da/dbeta = -2 d tau e^3, -2 d e^3
  const double tz = JS ? 1.0 : tau;
  double bb0 = -2 * ab0 * tz * w0 * e0, bb1 = -2 * ab1 * tz * w1 * e1, bb2 = -2 * ab2 * tz * w2 * e2;
  if (!JS) {
    const double taub = ab0 * w0 + ab1 * w1 + ab2 * w2;
    const double d = 2 * (b0 - b2) * taub;
    bb0 += d;
    bb2 -= d;
  }
  ub[0] += c6 * (2 * qb2) + bb2 * (2 * c13 * s2 + 0.5 * t2);
  ub[1] += c6 * (-qb1 - 7 * qb2) + bb1 * (2 * c13 * s1 + 0.5 * t1) + bb2 * (-4 * c13 * s2 - 2 * t2);
  ub[2] += c6 * (2 * qb0 + 5 * qb1 + 11 * qb2) + bb0 * (2 * c13 * s0 + 1.5 * t0) + bb1 * (-4 * c13 * s1) +
           bb2 * (2 * c13 * s2 + 1.5 * t2);
  ub[3] += c6 * (5 * qb0 + 2 * qb1) + bb0 * (-4 * c13 * s0 - 2 * t0) + bb1 * (2 * c13 * s1 - 0.5 * t1);
  ub[4] += c6 * (-qb0) + bb0 * (2 * c13 * s0 + 0.5 * t0);
}

// One face of WENO5 (Z or JS): gp[r] = fbar df/dvp[r], gm[r] = fbar df/dvm[r].
template <bool JS>
__device__ __forceinline__ void weno5_face_vjp(const double* vp, const double* vm, double eps, double fbar, double* gp,
                                               double* gm) {
#pragma unroll
  for (int r = 0; r < 5; ++r) gp[r] = gm[r] = 0;
  weno5_side_vjp<JS>(vp, eps, fbar, gp);
  double u[5] = {vm[4], vm[3], vm[2], vm[1], vm[0]}, ub[5] = {0, 0, 0, 0, 0};
  weno5_side_vjp<JS>(u, eps, fbar, ub);
#pragma unroll
  for (int r = 0; r < 5; ++r) gm[r] = ub[4 - r];
}

// Lax-Friedrichs split fluxes F+- = J/2 (F(q) +- lambda q), lambda = |U| + c|m|:
// cotangents fpb, fmb -> cotangents of q (cv, direct), p and c.
__device__ __forceinline__ void lf_split_vjp(const double* q, double p, double c, const double* m, double J,
                                             const double* fpb, const double* fmb, double* qb, double& pb,
                                             double& cb) {
  const double iq0 = 1 / q[0];
  const double Uk = (q[1] * m[0] + q[2] * m[1] + q[3] * m[2]) * iq0;
  const double gm = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const double lam = fabs(Uk) + c * gm;
  double Fb[5], Gb[5];
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    Fb[l] = 0.5 * J * (fpb[l] + fmb[l]);
    Gb[l] = 0.5 * J * (fpb[l] - fmb[l]);
  }
  double Ub = Fb[0] * q[0] + Fb[1] * q[1] + Fb[2] * q[2] + Fb[3] * q[3] + Fb[4] * (q[4] + p);
  const double lb = Gb[0] * q[0] + Gb[1] * q[1] + Gb[2] * q[2] + Gb[3] * q[3] + Gb[4] * q[4];
#pragma unroll
  for (int l = 0; l < 5; ++l) qb[l] = Fb[l] * Uk + Gb[l] * lam;
  pb = Fb[1] * m[0] + Fb[2] * m[1] + Fb[3] * m[2] + Fb[4] * Uk;
  Ub += Uk < 0 ? -lb : lb;  // d|U|/dU = sign(U), +1 at 0
  cb = lb * gm;
  qb[1] += Ub * m[0] * iq0;
  qb[2] += Ub * m[1] * iq0;
  qb[3] += Ub * m[2] * iq0;
  qb[0] -= Ub * Uk * iq0;
}

// Cotangents of the primitives (u, v, w, p, T) at a point -> added to the
// cotangents of cv: u_i = m_i/rho, p = (gamma-1)(E - |m|^2/(2 rho)), T = p/(rho R).
__device__ __forceinline__ void prim_vjp(const double* q, double ub, double vb, double wb, double pb, double Tb,
                                         double* cb) {
  const double ir = 1 / q[0];
  const double u = q[1] * ir, v = q[2] * ir, w = q[3] * ir;
  const double p = (kGamma - 1) * (q[4] - 0.5 * q[0] * (u * u + v * v + w * w));
  const double T = p * kMwAir * ir / kRu;
  const double pt = pb + Tb * kMwAir * ir / kRu;
  const double gp = (kGamma - 1) * pt;
  cb[0] += -Tb * T * ir + gp * 0.5 * (u * u + v * v + w * w) - (ub * u + vb * v + wb * w) * ir;
  cb[1] += -gp * u + ub * ir;
  cb[2] += -gp * v + vb * ir;
  cb[3] += -gp * w + wb * ir;
  cb[4] += gp;
}

struct Geo {
  int mx, my, mz, g, sx, sy;
  long long sz, n0;
};
__device__ __forceinline__ long long gidx(const Geo& G, int i, int j, int k) {
  return ((long long)(k + G.g) * G.sy + (j + G.g)) * G.sx + (i + G.g);
}
__device__ __forceinline__ long long gidx0(const Geo& G, int i, int j, int k) {
  return ((long long)k * G.my + j) * G.mx + i;
}
__host__ __device__ inline Geo geo_of(const VjpArgs& a) { return Geo{a.mx, a.my, a.mz, a.g, a.sx, a.sy, a.sz, a.n0}; }
__host__ __device__ inline Geo geo_of(const DevBlock& B) { return Geo{B.mx, B.my, B.mz, B.g, B.sx, B.sy, B.sz, B.n0}; }

// ----------------------------------------------------------- inviscid VJP
// Pencils along DIR: a workgroup owns NP consecutive padded points of NQ
// neighbouring pencils.  It stages the split fluxes of NP+10 points and the
// cotangent products of the NP+5 faces that touch its points in LDS; every
// owned point then gathers its flux cotangents from those faces and applies
// the VJP of the splitting.  DIR = 0 maps lanes along the pencil, DIR = 1, 2
// across pencils (along i), so global accesses are unit-stride in i.
template <int DIR, int NP, int NQ, bool JS>
__global__ __launch_bounds__(kT) void k_inviscid_vjp(VjpArgs A) {
  static_assert(NP * NQ == kT, "one owned point per thread");
  constexpr int W = NP + 10, NF = NP + 5;
  __shared__ double sF[10][NQ][W];
  __shared__ double sG[10][NQ][NF];
  __shared__ double sE[NQ][NF];
  const Geo G = geo_of(A);
  const int m = DIR == 0 ? G.mx : (DIR == 1 ? G.my : G.mz);
  const int nq = DIR == 0 ? G.my : G.mx;  // pencils across
  const int S0 = -G.g + (int)blockIdx.x * NP;
  const int q0 = (int)blockIdx.y * NQ;
  const int zc = blockIdx.z;
  const long long dstride = DIR == 0 ? 1 : (DIR == 1 ? (long long)G.sx : (long long)G.sx * G.sy);
  const long long dstride0 = DIR == 0 ? 1 : (DIR == 1 ? (long long)G.mx : (long long)G.mx * G.my);
  // padded index of (pencil q, position s along DIR)
  auto pt_index = [&](int q, int s) -> long long {
    const int a = q0 + q;
    if (DIR == 0) return gidx(G, s, a, zc);
    if (DIR == 1) return gidx(G, a, s, zc);
    return gidx(G, a, zc, s);
  };
  auto in_index = [&](int q, int s) -> long long {
    const int a = q0 + q;
    if (DIR == 0) return gidx0(G, s, a, zc);
    if (DIR == 1) return gidx0(G, a, s, zc);
    return gidx0(G, a, zc, s);
  };
  const int tid = threadIdx.x;
  const double* met = A.met + 3 * DIR * G.sz;

  // phase A: split fluxes of the loaded points
  for (int e = tid; e < NQ * W; e += kT) {
    const int q = DIR == 0 ? e / W : e % NQ;
    const int x = DIR == 0 ? e % W : e / NQ;
    const int s = S0 - 5 + x;
    double fp[5] = {0, 0, 0, 0, 0}, fm[5] = {0, 0, 0, 0, 0};
    if (q0 + q < nq && s >= -G.g && s < m + G.g) {
      const long long p = pt_index(q, s);
      double cv[5];
#pragma unroll
      for (int l = 0; l < 5; ++l) cv[l] = A.cv[p + l * G.sz];
      const double pr = A.bv[p + 4 * G.sz];
      const double c = sqrt(kGamma * kRAir * A.bv[p + 5 * G.sz]);
      const double mr[3] = {met[p], met[p + G.sz], met[p + 2 * G.sz]};
      const double J = A.jac[p];
      const double Uk = (cv[1] * mr[0] + cv[2] * mr[1] + cv[3] * mr[2]) / cv[0];
      const double lam = fabs(Uk) + c * sqrt(mr[0] * mr[0] + mr[1] * mr[1] + mr[2] * mr[2]);
      const double F[5] = {Uk * cv[0], Uk * cv[1] + pr * mr[0], Uk * cv[2] + pr * mr[1], Uk * cv[3] + pr * mr[2],
                           Uk * (cv[4] + pr)};
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        fp[l] = 0.5 * J * (F[l] + lam * cv[l]);
        fm[l] = 0.5 * J * (F[l] - lam * cv[l]);
      }
    }
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      sF[l][q][x] = fp[l];
      sF[5 + l][q][x] = fm[l];
    }
  }
  // WENO epsilon of each face t + 1/2 (geometry only), src/WENO.cu:1500-1520
  for (int e = tid; e < NQ * NF; e += kT) {
    const int q = DIR == 0 ? e / NF : e % NQ;
    const int f = DIR == 0 ? e % NF : e / NQ;
    const int t = S0 - 3 + f;
    double eps = 0;
    if (q0 + q < nq && t >= -1 && t <= m - 1) {
      const long long p1 = pt_index(q, t), p2 = p1 + dstride;
      const double j1 = A.jac[p1], j2 = A.jac[p2];
      double acc = 0;
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const double v = met[p1 + cc * G.sz] * j1 + met[p2 + cc * G.sz] * j2;
        acc += v * v;
      }
      eps = 1e-20 * A.P.weno_eps_scale * 0.25 * acc;
    }
    sE[q][f] = eps;
  }
  __syncthreads();

  const int oq = DIR == 0 ? tid / NP : tid % NQ;
  const int os = DIR == 0 ? tid % NP : tid / NQ;
  double fpb[5], fmb[5];
  const double v2 = A.P.v_ref * A.P.v_ref;
  for (int l = 0; l < 5; ++l) {
    for (int e = tid; e < NQ * NF; e += kT) {
      const int q = DIR == 0 ? e / NF : e % NQ;
      const int f = DIR == 0 ? e % NF : e / NQ;
      const int t = S0 - 3 + f;
      double gp[5] = {0, 0, 0, 0, 0}, gm[5] = {0, 0, 0, 0, 0};
      if (q0 + q < nq && t >= -1 && t <= m - 1) {
        // dq(t) -= fc(t+1/2) - fc(t-1/2): fc_bar(t+1/2) = w(t+1) - w(t)
        const long long pi = in_index(q, t < 0 ? 0 : t);
        double fbar = 0;
        if (t >= 0) fbar -= A.w[pi + l * G.n0];
        if (t + 1 <= m - 1) fbar += A.w[(t >= 0 ? pi + dstride0 : pi) + l * G.n0];
        double eps = sE[q][f];
        if (l >= 1 && l <= 3) eps = eps * v2;
        else if (l == 4) eps = eps * v2 * v2;
        double vp[5], vm[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
          vp[r] = sF[l][q][f + r];
          vm[r] = sF[5 + l][q][f + 1 + r];
        }
        weno5_face_vjp<JS>(vp, vm, eps, fbar, gp, gm);
      }
#pragma unroll
      for (int r = 0; r < 5; ++r) {
        sG[r][q][f] = gp[r];
        sG[5 + r][q][f] = gm[r];
      }
    }
    __syncthreads();
    // owned point s = S0 + os: faces t = s+2-r (F+ slot r), t = s+1-r (F- slot r)
    double a = 0, b = 0;
#pragma unroll
    for (int r = 0; r < 5; ++r) {
      a += sG[r][oq][os + 5 - r];
      b += sG[5 + r][oq][os + 4 - r];
    }
    fpb[l] = a;
    fmb[l] = b;
    __syncthreads();
  }

  const int s = S0 + os;
  if (q0 + oq < nq && s >= -G.g && s < m + G.g) {
    const long long p = pt_index(oq, s);
    double cv[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) cv[l] = A.cv[p + l * G.sz];
    const double pr = A.bv[p + 4 * G.sz], T = A.bv[p + 5 * G.sz];
    const double c = sqrt(kGamma * kRAir * T);
    const double mr[3] = {met[p], met[p + G.sz], met[p + 2 * G.sz]};
    double qb[5], pb, cb;
    lf_split_vjp(cv, pr, c, mr, A.jac[p], fpb, fmb, qb, pb, cb);
    // c = sqrt(gamma R T): dc/dT = c / (2T)
    prim_vjp(cv, 0, 0, 0, pb, cb * c / (2 * T), qb);
#pragma unroll
    for (int l = 0; l < 5; ++l) A.lam[p + l * G.sz] += qb[l];
  }
}

// ------------------------------------------------------------ viscous VJP
// Face (DIR, base B) lies between B and B + e_DIR; B_DIR in [-1, m-1], the
// other two coordinates interior.  Faces are numbered with the index along
// DIR shifted by one.
template <int DIR>
__device__ __forceinline__ long long face_id(const Geo& G, int i, int j, int k) {
  if (DIR == 0) return ((long long)k * G.my + j) * (G.mx + 1) + (i + 1);
  if (DIR == 1) return ((long long)k * (G.my + 1) + (j + 1)) * G.mx + i;
  return ((long long)(k + 1) * G.my + j) * G.mx + i;
}
template <int DIR>
__device__ __forceinline__ bool face_exists(const Geo& G, int i, int j, int k) {
  const int c[3] = {i, j, k}, m[3] = {G.mx, G.my, G.mz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int lo = a == DIR ? -1 : 0;
    if (c[a] < lo || c[a] > m[a] - 1) return false;
  }
  return true;
}

// Cotangents of one viscous face: 12 computational-space differences
// D[q][r] (q = u, v, w, T; r = xi, eta, zeta), the three face-mean
// velocities and the face viscosity, written as 16 planes of fs.
template <int DIR>
__global__ __launch_bounds__(kT) void k_visc_face_vjp(VjpArgs A) {
  const Geo G = geo_of(A);
  const int ex = DIR == 0 ? G.mx + 1 : G.mx, ey = DIR == 1 ? G.my + 1 : G.my, ez = DIR == 2 ? G.mz + 1 : G.mz;
  const long long nf = (long long)ex * ey * ez;
  const long long id = (long long)blockIdx.x * kT + threadIdx.x;
  if (id >= nf) return;
  int i = (int)(id % ex), j = (int)((id / ex) % ey), k = (int)(id / ((long long)ex * ey));
  if (DIR == 0) i -= 1;
  if (DIR == 1) j -= 1;
  if (DIR == 2) k -= 1;
  const int o[3] = {DIR == 0, DIR == 1, DIR == 2};
  const long long p = gidx(G, i, j, k), p1 = gidx(G, i + o[0], j + o[1], k + o[2]);
  const int m_d = DIR == 0 ? G.mx : (DIR == 1 ? G.my : G.mz);
  const int t = DIR == 0 ? i : (DIR == 1 ? j : k);
  // dq(t) += fv(t+1/2) - fv(t-1/2): fv_bar = w(t) - w(t+1)
  double fb[4] = {0, 0, 0, 0};
  const long long s0 = DIR == 0 ? 1 : (DIR == 1 ? (long long)G.mx : (long long)G.mx * G.my);
  const long long q0 = gidx0(G, i < 0 ? 0 : i, j < 0 ? 0 : j, k < 0 ? 0 : k);
#pragma unroll
  for (int l = 1; l < 5; ++l) {
    double v = 0;
    if (t >= 0) v += A.w[q0 + l * G.n0];
    if (t + 1 <= m_d - 1) v -= A.w[(t >= 0 ? q0 + s0 : q0) + l * G.n0];
    fb[l - 1] = v;
  }
  const double* bv = A.bv;
  const long long sz = G.sz;
  auto pv = [&](int di, int dj, int dk, int l) { return bv[gidx(G, i + di, j + dj, k + dk) + l * sz]; };
  double ma[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) ma[r][c] = 0.5 * (A.met[p + (3 * r + c) * sz] + A.met[p1 + (3 * r + c) * sz]);
  double D[4][3];
  const int var[4] = {1, 2, 3, 5};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int l = var[q];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      if (r == DIR) {
        D[q][r] = pv(o[0], o[1], o[2], l) - pv(0, 0, 0, l);
      } else {
        const int e[3] = {r == 0, r == 1, r == 2};
        D[q][r] = 0.25 * (pv(e[0], e[1], e[2], l) - pv(-e[0], -e[1], -e[2], l) +
                          pv(o[0] + e[0], o[1] + e[1], o[2] + e[2], l) - pv(o[0] - e[0], o[1] - e[1], o[2] - e[2], l));
      }
    }
  }
  double gq[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) gq[q][c] = D[q][0] * ma[0][c] + D[q][1] * ma[1][c] + D[q][2] * ma[2][c];
  const double ux = gq[0][0], uy = gq[0][1], uz = gq[0][2];
  const double vx = gq[1][0], vy = gq[1][1], vz = gq[1][2];
  const double wx = gq[2][0], wy = gq[2][1], wz = gq[2][2];
  const double mu = 0.5 * (A.mul[p] + A.mul[p1]);
  const double sxx = (4 * ux - 2 * vy - 2 * wz) / 3.0, syy = (4 * vy - 2 * ux - 2 * wz) / 3.0,
               szz = (4 * wz - 2 * ux - 2 * vy) / 3.0;
  const double sxy = uy + vx, sxz = uz + wx, syz = vz + wy;
  const double Jp = A.jac[p], Jp1 = A.jac[p1];
  const double kx = 0.5 * (A.met[p + (3 * DIR) * sz] * Jp + A.met[p1 + (3 * DIR) * sz] * Jp1);
  const double ky = 0.5 * (A.met[p + (3 * DIR + 1) * sz] * Jp + A.met[p1 + (3 * DIR + 1) * sz] * Jp1);
  const double kz = 0.5 * (A.met[p + (3 * DIR + 2) * sz] * Jp + A.met[p1 + (3 * DIR + 2) * sz] * Jp1);
  const double fv1 = mu * (kx * sxx + ky * sxy + kz * sxz);
  const double fv2 = mu * (kx * sxy + ky * syy + kz * syz);
  const double fv3 = mu * (kx * sxz + ky * syz + kz * szz);
  const double um = 0.5 * (pv(0, 0, 0, 1) + pv(o[0], o[1], o[2], 1));
  const double vm = 0.5 * (pv(0, 0, 0, 2) + pv(o[0], o[1], o[2], 2));
  const double wm = 0.5 * (pv(0, 0, 0, 3) + pv(o[0], o[1], o[2], 3));
  constexpr double cp = kGamma * kRu / kMwAir / (kGamma - 1);
  const double kap = mu / A.P.prandtl * cp;
  const double kt = kx * gq[3][0] + ky * gq[3][1] + kz * gq[3][2];
  // reverse
  const double f1b = fb[0] + fb[3] * um, f2b = fb[1] + fb[3] * vm, f3b = fb[2] + fb[3] * wm;
  const double umb = fb[3] * fv1, vmb = fb[3] * fv2, wmb = fb[3] * fv3;
  const double kapb = fb[3] * kt;
  const double txb = fb[3] * kap * kx, tyb = fb[3] * kap * ky, tzb = fb[3] * kap * kz;
  const double axx = f1b * kx, axy = f1b * ky + f2b * kx, axz = f1b * kz + f3b * kx;
  const double ayy = f2b * ky, ayz = f2b * kz + f3b * ky, azz = f3b * kz;
  const double mub = axx * sxx + ayy * syy + azz * szz + axy * sxy + axz * sxz + ayz * syz + kapb * cp / A.P.prandtl;
  double gb[4][3];
  gb[0][0] = mu * (4 * axx - 2 * ayy - 2 * azz) / 3.0;
  gb[1][1] = mu * (4 * ayy - 2 * axx - 2 * azz) / 3.0;
  gb[2][2] = mu * (4 * azz - 2 * axx - 2 * ayy) / 3.0;
  gb[0][1] = gb[1][0] = mu * axy;
  gb[0][2] = gb[2][0] = mu * axz;
  gb[1][2] = gb[2][1] = mu * ayz;
  gb[3][0] = txb;
  gb[3][1] = tyb;
  gb[3][2] = tzb;
  double* fs = A.fs + (long long)DIR * 16 * A.fs_stride;
  const long long fid = face_id<DIR>(G, i, j, k);
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      fs[(q * 3 + r) * A.fs_stride + fid] = gb[q][0] * ma[r][0] + gb[q][1] * ma[r][1] + gb[q][2] * ma[r][2];
  fs[12 * A.fs_stride + fid] = umb;
  fs[13 * A.fs_stride + fid] = vmb;
  fs[14 * A.fs_stride + fid] = wmb;
  fs[15 * A.fs_stride + fid] = mub;
}

// Gathers the viscous face cotangents onto the points [-1, m]^3 and chains
// them through Sutherland's law and the primitive variables to cv.  One pass
// per face direction D: pass 0 starts the point's sums (vb: velocity and
// temperature, mb: the face-mean terms), passes 1 and 2 continue them from
// A.gs in the same order, and pass 2 chains them to lam.  The sums are those of
// one pass over all three directions, term for term; split, a pass holds a
// third of the 144 face loads in flight (all of them take 302 VGPRs, one
// wave per SIMD).
template <int D>
__global__ __launch_bounds__(kT) void k_visc_gather(VjpArgs A) {
  const Geo G = geo_of(A);
  const int ex = G.mx + 2, ey = G.my + 2, ez = G.mz + 2;
  const long long id = (long long)blockIdx.x * kT + threadIdx.x;
  if (id >= (long long)ex * ey * ez) return;
  const int i = (int)(id % ex) - 1, j = (int)((id / ex) % ey) - 1, k = (int)(id / ((long long)ex * ey)) - 1;
  double vb[4] = {0, 0, 0, 0}, mb[4] = {0, 0, 0, 0};
  if (D > 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      vb[q] = A.gs[q * A.gs_stride + id];
      mb[q] = A.gs[(4 + q) * A.gs_stride + id];
    }
  }
  auto face = [&](int plane, int a, int b, int c) -> double {
    if (!face_exists<D>(G, a, b, c)) return 0.0;
    return A.fs[((long long)D * 16 + plane) * A.fs_stride + face_id<D>(G, a, b, c)];
  };
  const int o[3] = {D == 0, D == 1, D == 2};
#pragma unroll
  for (int q = 0; q < 4; ++q) vb[q] += face(3 * q + D, i - o[0], j - o[1], k - o[2]) - face(3 * q + D, i, j, k);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    if (r == D) continue;
    const int e[3] = {r == 0, r == 1, r == 2};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      vb[q] += 0.25 * (face(3 * q + r, i - e[0], j - e[1], k - e[2]) - face(3 * q + r, i + e[0], j + e[1], k + e[2]) +
                       face(3 * q + r, i - e[0] - o[0], j - e[1] - o[1], k - e[2] - o[2]) -
                       face(3 * q + r, i + e[0] - o[0], j + e[1] - o[1], k + e[2] - o[2]));
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) mb[r] += 0.5 * (face(12 + r, i, j, k) + face(12 + r, i - o[0], j - o[1], k - o[2]));
  if (D < 2) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      A.gs[q * A.gs_stride + id] = vb[q];
      A.gs[(4 + q) * A.gs_stride + id] = mb[q];
    }
    return;
  }
  // points no face reaches (the corners of the shell) carry no cotangent
  if (vb[0] == 0 && vb[1] == 0 && vb[2] == 0 && vb[3] == 0 && mb[0] == 0 && mb[1] == 0 && mb[2] == 0 && mb[3] == 0)
    return;
  const long long p = gidx(G, i, j, k);
  double cv[5];
#pragma unroll
  for (int l = 0; l < 5; ++l) cv[l] = A.cv[p + l * G.sz];
  const double T = A.bv[p + 5 * G.sz], mu = A.mul[p];
  // Sutherland: dmu/dT = mu (1.5/T - 1/(T + 110))
  const double Tb = vb[3] + mb[3] * mu * (1.5 / T - 1.0 / (T + 110));
  double cb[5] = {0, 0, 0, 0, 0};
  prim_vjp(cv, vb[0] + mb[0], vb[1] + mb[1], vb[2] + mb[2], 0, Tb, cb);
#pragma unroll
  for (int l = 0; l < 5; ++l) A.lam[p + l * G.sz] += cb[l];
}

// -------------------------------------------------------- RK3 transposes
// Interior: w = c dt/J lam (dq cotangent), lam_qn += a lam, lam = b lam.
__global__ __launch_bounds__(kT) void k_stage_pre(Geo G, const double* jac, double* lam, double* lam_qn, double* w,
                                                  double ra, double rb, double cdt) {
  const long long id = (long long)blockIdx.x * kT + threadIdx.x;
  if (id >= G.n0) return;
  const int i = (int)(id % G.mx), j = (int)((id / G.mx) % G.my), k = (int)(id / ((long long)G.mx * G.my));
  const long long p = gidx(G, i, j, k);
  const double f = cdt / jac[p];
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    const double L = lam[p + l * G.sz];
    w[id + l * G.n0] = f * L;
    lam_qn[p + l * G.sz] += ra * L;
    lam[p + l * G.sz] = rb * L;
  }
}

// Interior: lam += lam_qn, lam_qn = 0 (qn is the step's initial state).
__global__ __launch_bounds__(kT) void k_stage_qn(Geo G, double* lam, double* lam_qn) {
  const long long id = (long long)blockIdx.x * kT + threadIdx.x;
  if (id >= G.n0) return;
  const int i = (int)(id % G.mx), j = (int)((id / G.mx) % G.my), k = (int)(id / ((long long)G.mx * G.my));
  const long long p = gidx(G, i, j, k);
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    lam[p + l * G.sz] += lam_qn[p + l * G.sz];
    lam_qn[p + l * G.sz] = 0;
  }
}

// Transpose of one boundary op (ghost = copy of a source cell):
// lam[src] += lam[ghost], lam[ghost] = 0.  Sources and ghosts of one op are
// disjoint, and the ghosts sharing a source belong to one thread.
__global__ __launch_bounds__(kT) void k_bc_adj(Geo G, DevOp op, double* lam) {
  const long long t = (long long)blockIdx.x * kT + threadIdx.x;
  if (t >= op.count) return;
  const int i = op.lo[0] + (int)(t % op.n[0]);
  const int j = op.lo[1] + (int)((t / op.n[0]) % op.n[1]);
  const int k = op.lo[2] + (int)(t / ((long long)op.n[0] * op.n[1]));
  int d[3] = {0, 0, 0};
  d[op.axis] = op.side;
  int so[3] = {i, j, k};
  if (op.type == OP_BC_PERIODIC) {
    const int m = op.axis == 0 ? G.mx : (op.axis == 1 ? G.my : G.mz);
    so[op.axis] = op.side < 0 ? m - 1 : 0;
  }
  for (int g = 1; g <= G.g; ++g) {
    const long long pg = gidx(G, i + g * d[0], j + g * d[1], k + g * d[2]);
    const long long ps = op.type == OP_BC_PERIODIC ? gidx(G, so[0] + g * d[0], so[1] + g * d[1], so[2] + g * d[2])
                                                   : gidx(G, i, j, k);
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      lam[ps + l * G.sz] += lam[pg + l * G.sz];
      lam[pg + l * G.sz] = 0;
    }
  }
}

// ------------------------------------------------- differentiation workload
__device__ __forceinline__ double mode_value(const FourierModes& F, int m, int i, int j, int k) {
  const double y = F.y0 + j * F.dy;
  const double yr = (y - F.yc) / F.delta;
  const double env = F.du * exp(-yr * yr);
  const double ph = 6.283185307179586 * (F.kx[m] * (i * F.dx) / F.lx + F.kz[m] * (k * F.dz) / F.lz);
  return env * (F.trig[m] ? cos(ph) : sin(ph));
}

__global__ __launch_bounds__(kT) void k_fourier_add(Geo G, FourierModes F, double* bv) {
  const long long id = (long long)blockIdx.x * kT + threadIdx.x;
  if (id >= G.sz) return;
  const int i = (int)(id % G.sx) - G.g, j = (int)((id / G.sx) % G.sy) - G.g, k = (int)(id / ((long long)G.sx * G.sy)) - G.g;
  double add[3] = {0, 0, 0};
  for (int m = 0; m < F.n; ++m) add[F.comp[m]] += F.amp[m] * mode_value(F, m, i, j, k);
#pragma unroll
  for (int c = 0; c < 3; ++c) bv[id + (1 + c) * G.sz] += add[c];
}

// grad_m partials: sum over the padded points of phi_m * d(cv)/d(u_c)^T lam,
// d(rho u_c)/du_c = rho, d(E)/du_c = rho u_c.  One row of n partials per workgroup.
__global__ __launch_bounds__(kT) void k_fourier_grad(Geo G, FourierModes F, const double* lam, const double* cv,
                                                     double* part) {
  __shared__ double sw[kT / 64][kMaxModes];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  double acc[kMaxModes];
  for (int m = 0; m < F.n; ++m) acc[m] = 0;
  for (long long id = (long long)blockIdx.x * kT + threadIdx.x; id < G.sz; id += (long long)gridDim.x * kT) {
    const int i = (int)(id % G.sx) - G.g, j = (int)((id / G.sx) % G.sy) - G.g,
              k = (int)(id / ((long long)G.sx * G.sy)) - G.g;
    const double rho = cv[id];
    double gc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) gc[c] = rho * lam[id + (1 + c) * G.sz] + cv[id + (1 + c) * G.sz] * lam[id + 4 * G.sz];
    if (gc[0] == 0 && gc[1] == 0 && gc[2] == 0) continue;
    for (int m = 0; m < F.n; ++m) acc[m] += gc[F.comp[m]] * mode_value(F, m, i, j, k);
  }
  for (int m = 0; m < F.n; ++m) {
    const double v = wave_sum64(acc[m]);
    if (lane == 0) sw[wv][m] = v;
  }
  __syncthreads();
  if (threadIdx.x < F.n) {
    double s = 0;
    for (int q = 0; q < kT / 64; ++q) s += sw[q][threadIdx.x];
    part[(long long)blockIdx.x * kMaxModes + threadIdx.x] = s;
  }
}

__global__ void k_fourier_grad_final(int n, int n_part, const double* part, double* grad) {
  const int m = threadIdx.x;
  if (m >= n) return;
  double s = 0;
  for (int q = 0; q < n_part; ++q) s += part[(long long)q * kMaxModes + m];
  grad[m] = s;
}

// Row sums S0_j = sum rho, S1_j = sum rho u over i < ni, k < nk.
__global__ __launch_bounds__(kT) void k_theta_rows(Geo G, const double* cv, int ni, int nk, double* rows) {
  __shared__ double s0w[kT / 64], s1w[kT / 64];
  const int j = blockIdx.x;
  double a = 0, b = 0;
  for (long long e = threadIdx.x; e < (long long)ni * nk; e += kT) {
    const int i = (int)(e % ni), k = (int)(e / ni);
    const long long p = gidx(G, i, j, k);
    a += cv[p];
    b += cv[p + G.sz];
  }
  a = wave_sum64(a);
  b = wave_sum64(b);
  if (threadIdx.x % 64 == 0) {
    s0w[threadIdx.x / 64] = a;
    s1w[threadIdx.x / 64] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0, y = 0;
    for (int q = 0; q < kT / 64; ++q) {
      x += s0w[q];
      y += s1w[q];
    }
    rows[2 * j] = x;
    rows[2 * j + 1] = y;
  }
}

// theta and its derivatives with respect to the row sums (one thread).
__global__ void k_theta_final(int my, ThetaParams T, const double* rows, double* coef, double* theta) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double N = (double)T.ni * T.nk;
  double th = 0;
  for (int j = 0; j < my; ++j) {
    const double S0 = rows[2 * j], S1 = rows[2 * j + 1];
    const double rho = S0 / N, U = S1 / S0;
    const double C = ((j == 0 || j == my - 1) ? 0.5 * T.dy : T.dy) / (T.rho0 * T.du * T.du);
    th += C * rho * (T.u1 - U) * (U - T.u2);
    const double d_rho = C * (T.u1 - U) * (U - T.u2);
    const double d_U = C * rho * (T.u1 + T.u2 - 2 * U);
    coef[2 * j] = d_rho / N - d_U * S1 / (S0 * S0);
    coef[2 * j + 1] = d_U / S0;
  }
  *theta = th;
}

// lam = d theta / d cv on the padded domain.
__global__ __launch_bounds__(kT) void k_theta_seed(Geo G, int ni, int nk, const double* coef, double* lam) {
  const long long id = (long long)blockIdx.x * kT + threadIdx.x;
  if (id >= G.sz) return;
  const int i = (int)(id % G.sx) - G.g, j = (int)((id / G.sx) % G.sy) - G.g, k = (int)(id / ((long long)G.sx * G.sy)) - G.g;
  const bool in = i >= 0 && i < ni && k >= 0 && k < nk && j >= 0 && j < G.my;
  lam[id] = in ? coef[2 * j] : 0.0;
  lam[id + G.sz] = in ? coef[2 * j + 1] : 0.0;
#pragma unroll
  for (int l = 2; l < 5; ++l) lam[id + l * G.sz] = 0.0;
}

unsigned blocks_for(long long n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace

// =============================================================== launchers
void rhs_vjp(const VjpArgs& a, hipStream_t st, long long* n) {
  const int g = a.g;
  const bool js = a.P.scheme == 50;
  {
    dim3 grid((a.mx + 2 * g + 63) / 64, (a.my + 3) / 4, a.mz);
    if (js) k_inviscid_vjp<0, 64, 4, true><<<grid, kT, 0, st>>>(a);
    else k_inviscid_vjp<0, 64, 4, false><<<grid, kT, 0, st>>>(a);
  }
  {
    dim3 grid((a.my + 2 * g + 15) / 16, (a.mx + 15) / 16, a.mz);
    if (js) k_inviscid_vjp<1, 16, 16, true><<<grid, kT, 0, st>>>(a);
    else k_inviscid_vjp<1, 16, 16, false><<<grid, kT, 0, st>>>(a);
  }
  {
    dim3 grid((a.mz + 2 * g + 15) / 16, (a.mx + 15) / 16, a.my);
    if (js) k_inviscid_vjp<2, 16, 16, true><<<grid, kT, 0, st>>>(a);
    else k_inviscid_vjp<2, 16, 16, false><<<grid, kT, 0, st>>>(a);
  }
  *n += 3;
  if (!a.P.visc) return;
  const long long f0 = (long long)(a.mx + 1) * a.my * a.mz, f1 = (long long)a.mx * (a.my + 1) * a.mz,
                  f2 = (long long)a.mx * a.my * (a.mz + 1);
  k_visc_face_vjp<0><<<blocks_for(f0), kT, 0, st>>>(a);
  k_visc_face_vjp<1><<<blocks_for(f1), kT, 0, st>>>(a);
  k_visc_face_vjp<2><<<blocks_for(f2), kT, 0, st>>>(a);
  const long long gp = vjp_gather_stride(a.mx, a.my, a.mz);
  k_visc_gather<0><<<blocks_for(gp), kT, 0, st>>>(a);
  k_visc_gather<1><<<blocks_for(gp), kT, 0, st>>>(a);
  k_visc_gather<2><<<blocks_for(gp), kT, 0, st>>>(a);
  *n += 6;
}

long long vjp_face_stride(int mx, int my, int mz) {
  return std::max({(long long)(mx + 1) * my * mz, (long long)mx * (my + 1) * mz, (long long)mx * my * (mz + 1)});
}

long long vjp_gather_stride(int mx, int my, int mz) { return (long long)(mx + 2) * (my + 2) * (mz + 2); }

void adj_stage_pre(const DevBlock& B, double* lam, double* lam_qn, double* w, double ra, double rb, double cdt,
                   hipStream_t st) {
  k_stage_pre<<<blocks_for(B.n0), kT, 0, st>>>(geo_of(B), B.jac, lam, lam_qn, w, ra, rb, cdt);
}

void adj_stage_qn(const DevBlock& B, double* lam, double* lam_qn, hipStream_t st) {
  k_stage_qn<<<blocks_for(B.n0), kT, 0, st>>>(geo_of(B), lam, lam_qn);
}

void adj_bc_op(const DevBlock& B, const DevOp& op, double* lam, hipStream_t st) {
  if (op.type != OP_BC_OUTFLOW && op.type != OP_BC_PERIODIC) return;
  k_bc_adj<<<blocks_for(op.count), kT, 0, st>>>(geo_of(B), op, lam);
}

void fourier_add(const DevBlock& B, const FourierModes& F, double* bv, hipStream_t st) {
  k_fourier_add<<<blocks_for(B.sz), kT, 0, st>>>(geo_of(B), F, bv);
}

int fourier_grad_parts() { return 1024; }

void fourier_grad(const DevBlock& B, const FourierModes& F, const double* lam, const double* cv, double* part,
                  double* grad, hipStream_t st) {
  const int np = (int)std::min<long long>(fourier_grad_parts(), blocks_for(B.sz));
  k_fourier_grad<<<np, kT, 0, st>>>(geo_of(B), F, lam, cv, part);
  k_fourier_grad_final<<<1, kMaxModes, 0, st>>>(F.n, np, part, grad);
}

void theta_eval(const DevBlock& B, const ThetaParams& T, const double* cv, double* rows, double* coef, double* theta,
                hipStream_t st) {
  k_theta_rows<<<B.my, kT, 0, st>>>(geo_of(B), cv, T.ni, T.nk, rows);
  k_theta_final<<<1, 64, 0, st>>>(B.my, T, rows, coef, theta);
}

void theta_seed(const DevBlock& B, const ThetaParams& T, const double* coef, double* lam, hipStream_t st) {
  k_theta_seed<<<blocks_for(B.sz), kT, 0, st>>>(geo_of(B), T.ni, T.nk, coef, lam);
}

}  // namespace ostc

// ------------------------------------------------ allocation-free objective
// Row sums (2 my), theta coefficients (2 my) and theta of the momentum
// thickness for blocks of up to kThetaMaxRows rows, as one file-scope device
// array: every GPU the library runs on holds its own copy until the process
// ends.  k_theta_rows writes every row it covers and k_theta_final every
// coefficient and theta before anything reads them, so nothing needs
// clearing; calls that overlap on one GPU share the array and race.

namespace ostc {

static __device__ __align__(256) double scratch[4 * kThetaMaxRows + 1];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current: the
// array's host-side name is not a device address, so kernels and copies
// take this pointer.  Looked up once per device and kept.
int theta_scratch_ptr(int device, double** scratch_ptr) {
  if (device < 0 || device >= kMaxDevices || !scratch_ptr) return (int)hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *scratch_ptr = g_scratch_ptr[device];
  return (int)hipSuccess;
}

}  // namespace ostc
