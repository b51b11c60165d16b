// ostc_fr_kernels.hip -- HIP kernels of the finite-rate mixture path
// (the reference's MixtureModel::FR, laminar), for gfx950.
//
// Per RK stage and block: k_source (dq = J omega, or 0), three inviscid
// sweeps, three viscous sweeps, k_update (SSP-RK3 + primitives + T-Newton),
// then ghost ops and k_props over the padded domain.  dt lives on the device
// (k_local_dt reduces it with 64-lane shuffles and one atomicMin per workgroup), so
// a whole step has no host synchronisation and can be captured in a graph.
//
// Inviscid sweeps stage the split fluxes of a tile in LDS: a tile is TW
// columns x (TS + 2 ngg) points along the sweep, one thread per face
// (TS + 1 faces per column, 256 faces per tile).  Points are loaded with
// consecutive threads on consecutive i (coalesced), F+ and F- are stored as
// structure-of-arrays planes (bank-conflict-free face reads), face fluxes
// overwrite the same LDS, and each cell subtracts its face difference from dq.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ostc_fr.hpp"
#include "ostc_inflow.hpp"

namespace ostc {
namespace fr {
namespace {

constexpr int kT = 256;

__device__ __forceinline__ long long pid(const FrBlockDev& B, int i, int j, int k) {
  return ((long long)(k + B.g) * B.sy + (j + B.g)) * B.sx + (i + B.g);
}
__device__ __forceinline__ long long pid0(const FrBlockDev& B, int i, int j, int k) {
  return ((long long)k * B.my + j) * B.mx + i;
}

// WENO-Z5, src/WENO.cu:1582-1617
__device__ __forceinline__ double weno5z(const double* vp, const double* vm, double eps) {
  constexpr double one6th{1.0 / 6};
  double v0{one6th * (2 * vp[2] + 5 * vp[3] - vp[4])};
  double v1{one6th * (-vp[1] + 5 * vp[2] + 2 * vp[3])};
  double v2{one6th * (2 * vp[0] - 7 * vp[1] + 11 * vp[2])};
  constexpr double c1312{13.0 / 12};
  double b0 = c1312 * (vp[2] + vp[4] - 2 * vp[3]) * (vp[2] + vp[4] - 2 * vp[3]) +
              0.25 * (3 * vp[2] - 4 * vp[3] + vp[4]) * (3 * vp[2] - 4 * vp[3] + vp[4]);
  double b1 = c1312 * (vp[1] + vp[3] - 2 * vp[2]) * (vp[1] + vp[3] - 2 * vp[2]) +
              0.25 * (vp[1] - vp[3]) * (vp[1] - vp[3]);
  double b2 = c1312 * (vp[0] + vp[2] - 2 * vp[1]) * (vp[0] + vp[2] - 2 * vp[1]) +
              0.25 * (vp[0] - 4 * vp[1] + 3 * vp[2]) * (vp[0] - 4 * vp[1] + 3 * vp[2]);
  double tau{(b0 - b2) * (b0 - b2)};
  double a0{0.3 + 0.3 * tau / ((eps + b0) * (eps + b0))};
  double a1{0.6 + 0.6 * tau / ((eps + b1) * (eps + b1))};
  double a2{0.1 + 0.1 * tau / ((eps + b2) * (eps + b2))};
  const double fp{(a0 * v0 + a1 * v1 + a2 * v2) / (a0 + a1 + a2)};
  v0 = one6th * (11 * vm[2] - 7 * vm[3] + 2 * vm[4]);
  v1 = one6th * (2 * vm[1] + 5 * vm[2] - vm[3]);
  v2 = one6th * (-vm[0] + 5 * vm[1] + 2 * vm[2]);
  b0 = c1312 * (vm[2] + vm[4] - 2 * vm[3]) * (vm[2] + vm[4] - 2 * vm[3]) +
       0.25 * (3 * vm[2] - 4 * vm[3] + vm[4]) * (3 * vm[2] - 4 * vm[3] + vm[4]);
  b1 = c1312 * (vm[1] + vm[3] - 2 * vm[2]) * (vm[1] + vm[3] - 2 * vm[2]) + 0.25 * (vm[1] - vm[3]) * (vm[1] - vm[3]);
  b2 = c1312 * (vm[0] + vm[2] - 2 * vm[1]) * (vm[0] + vm[2] - 2 * vm[1]) +
       0.25 * (vm[0] - 4 * vm[1] + 3 * vm[2]) * (vm[0] - 4 * vm[1] + 3 * vm[2]);
  tau = (b0 - b2) * (b0 - b2);
  a0 = 0.1 + 0.1 * tau / ((eps + b0) * (eps + b0));
  a1 = 0.6 + 0.6 * tau / ((eps + b1) * (eps + b1));
  a2 = 0.3 + 0.3 * tau / ((eps + b2) * (eps + b2));
  const double fm{(a0 * v0 + a1 * v1 + a2 * v2) / (a0 + a1 + a2)};
  return fp + fm;
}

// WENO-Z7, src/WENO.cu:1619-1693.  vp = F+ at i-3..i+3, vm = F- at i-2..i+4.
__device__ __forceinline__ double weno7z(const double* vp, const double* vm, double eps) {
  constexpr double one6th{1.0 / 6}, d12{13.0 / 12.0}, d13{1043.0 / 960}, d14{1.0 / 12};
  constexpr double c0{1.0 / 35}, c1{12.0 / 35}, c2{18.0 / 35}, c3{4.0 / 35}, one12th{1.0 / 12};
  double s1 = one6th * (-2 * vp[0] + 9 * vp[1] - 18 * vp[2] + 11 * vp[3]);
  double s2 = -vp[0] + 4 * vp[1] - 5 * vp[2] + 2 * vp[3];
  double s3 = -vp[0] + 3 * vp[1] - 3 * vp[2] + vp[3];
  double b0 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  s1 = one6th * (vp[1] - 6 * vp[2] + 3 * vp[3] + 2 * vp[4]);
  s2 = vp[2] - 2 * vp[3] + vp[4];
  s3 = -vp[1] + 3 * vp[2] - 3 * vp[3] + vp[4];
  double b1 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  s1 = one6th * (-2 * vp[2] - 3 * vp[3] + 6 * vp[4] - vp[5]);
  s3 = -vp[2] + 3 * vp[3] - 3 * vp[4] + vp[5];
  double b2 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  s1 = one6th * (-11 * vp[3] + 18 * vp[4] - 9 * vp[5] + 2 * vp[6]);
  s2 = 2 * vp[3] - 5 * vp[4] + 4 * vp[5] - vp[6];
  s3 = -vp[3] + 3 * vp[4] - 3 * vp[5] + vp[6];
  double b3 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  double tau = (b0 - b3) * (b0 - b3);
  double a0 = c0 + c0 * tau / ((eps + b0) * (eps + b0));
  double a1 = c1 + c1 * tau / ((eps + b1) * (eps + b1));
  double a2 = c2 + c2 * tau / ((eps + b2) * (eps + b2));
  double a3 = c3 + c3 * tau / ((eps + b3) * (eps + b3));
  double v0 = one12th * (-3 * vp[0] + 13 * vp[1] - 23 * vp[2] + 25 * vp[3]);
  double v1 = one12th * (vp[1] - 5 * vp[2] + 13 * vp[3] + 3 * vp[4]);
  double v2 = one12th * (-vp[2] + 7 * vp[3] + 7 * vp[4] - vp[5]);
  double v3 = one12th * (3 * vp[3] + 13 * vp[4] - 5 * vp[5] + vp[6]);
  const double fp = (a0 * v0 + a1 * v1 + a2 * v2 + a3 * v3) / (a0 + a1 + a2 + a3);

  s1 = one6th * (-2 * vm[6] + 9 * vm[5] - 18 * vm[4] + 11 * vm[3]);
  s2 = -vm[6] + 4 * vm[5] - 5 * vm[4] + 2 * vm[3];
  s3 = -vm[6] + 3 * vm[5] - 3 * vm[4] + vm[3];
  b0 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  s1 = one6th * (vm[5] - 6 * vm[4] + 3 * vm[3] + 2 * vm[2]);
  s2 = vm[4] - 2 * vm[3] + vm[2];
  s3 = -vm[5] + 3 * vm[4] - 3 * vm[3] + vm[2];
  b1 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  s1 = one6th * (-2 * vm[4] - 3 * vm[3] + 6 * vm[2] - vm[1]);
  s3 = -vm[4] + 3 * vm[3] - 3 * vm[2] + vm[1];
  b2 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  s1 = one6th * (-11 * vm[3] + 18 * vm[2] - 9 * vm[1] + 2 * vm[0]);
  s2 = 2 * vm[3] - 5 * vm[2] + 4 * vm[1] - vm[0];
  s3 = -vm[3] + 3 * vm[2] - 3 * vm[1] + vm[0];
  b3 = s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
  tau = (b0 - b3) * (b0 - b3);
  a0 = c0 + c0 * tau / ((eps + b0) * (eps + b0));
  a1 = c1 + c1 * tau / ((eps + b1) * (eps + b1));
  a2 = c2 + c2 * tau / ((eps + b2) * (eps + b2));
  a3 = c3 + c3 * tau / ((eps + b3) * (eps + b3));
  v0 = one12th * (-3 * vm[6] + 13 * vm[5] - 23 * vm[4] + 25 * vm[3]);
  v1 = one12th * (vm[5] - 5 * vm[4] + 13 * vm[3] + 3 * vm[2]);
  v2 = one12th * (-vm[4] + 7 * vm[3] + 7 * vm[2] - vm[1]);
  v3 = one12th * (3 * vm[3] + 13 * vm[2] - 5 * vm[1] + vm[0]);
  const double fm = (a0 * v0 + a1 * v1 + a2 * v2 + a3 * v3) / (a0 + a1 + a2 + a3);
  return fp + fm;
}

// cv of one point from bv and sv (compute_cv_from_bv_1_point<FR>,
// src/FieldOperation.cuh:78-105)
template <int NS>
__device__ __forceinline__ void cv_point(const FrArgs& A, long long p) {
  const FrBlockDev& B = A.B;
  const long long s = B.sz;
  const double rho = B.bv[p];
  B.cv[p] = rho;
  B.cv[p + s] = rho * B.bv[p + s];
  B.cv[p + 2 * s] = rho * B.bv[p + 2 * s];
  B.cv[p + 3 * s] = rho * B.bv[p + 3 * s];
  double Y[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    Y[l] = B.sv[p + l * s];
    B.cv[p + (5 + l) * s] = rho * Y[l];
  }
  B.cv[p + 4 * s] =
      total_energy<NS>(A.T, rho, B.bv[p + s], B.bv[p + 2 * s], B.bv[p + 3 * s], B.bv[p + 4 * s], B.bv[p + 5 * s], Y);
}

// The block and block-local workgroup of a batched launch (FrBatch): a
// linear search of the uniform prefix table, at most kFrBatchMax entries.
struct WgPos { int b, x, y, z; };
__device__ __forceinline__ WgPos wg_pos(const FrBatch& W) {
  const int w = blockIdx.x;
  int b = 0;
  while (b + 1 < W.nb && W.first[b + 1] <= w) ++b;
  const int local = w - W.first[b], gx = W.gx[b], gy = W.gy[b];
  return WgPos{b, local % gx, (local / gx) % gy, local / (gx * gy)};
}

// ------------------------------------------------------------ point kernels
template <int NS>
__global__ __launch_bounds__(kT) void k_cv_from_bv(FrArgs A) {
  const long long p = (long long)blockIdx.x * kT + threadIdx.x;
  if (p < A.B.sz) cv_point<NS>(A, p);
}

// Parallel faces (faces shared with another rank), mixture: per face point
// the 6 primitives and ns mass fractions of the face node and of the ngg
// nodes inside it (setup_data_to_be_sent, src/DataCommunication.cu:3-40) ...
template <int NS>
__global__ __launch_bounds__(kParThreads) void k_par_pack(FrArgs A, ParFace f, double* buf) {
  const FrBlockDev& B = A.B;
  const long long s = B.sz;
  par_pack_chunk<6 + NS>(f, B.sx, B.sy, B.g, buf,
                         [&](int v, long long p) { return v < 6 ? B.bv[p + v * s] : B.sv[p + (v - 6) * s]; });
}

// ... and from the partner's buffer: the face node takes the mean of its own
// and the received values, the ngg ghost nodes beyond it the received inner
// ones, each with cv from its primitives (assign_data_received,
// src/DataCommunication.cuh:201-240; cv as compute_cv_from_bv_1_point).
template <int NS>
__global__ __launch_bounds__(kParThreads) void k_par_unpack(FrArgs A, ParFace f, const double* recv,
                                                            const double* sent) {
  const FrBlockDev& B = A.B;
  const long long s = B.sz;
  par_unpack_chunk<6 + NS>(
      f, B.sx, B.sy, B.g, recv, sent,
      [&](int v, long long p) { return v < 6 ? B.bv[p + v * s] : B.sv[p + (v - 6) * s]; },
      [&](int v, long long p, double x) {
        if (v < 6)
          B.bv[p + v * s] = x;
        else
          B.sv[p + (v - 6) * s] = x;
      },
      [&](long long p, const double* q) {
        const double rho = q[0];
        B.cv[p] = rho;
        B.cv[p + s] = rho * q[1];
        B.cv[p + 2 * s] = rho * q[2];
        B.cv[p + 3 * s] = rho * q[3];
        double Y[NS];
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          Y[l] = q[6 + l];
          B.cv[p + (5 + l) * s] = rho * Y[l];
        }
        B.cv[p + 4 * s] = total_energy<NS>(A.T, rho, q[1], q[2], q[3], q[4], q[5], Y);
      });
}

// update_physical_properties<FR>, src/FieldOperation.cuh:107-140, over the
// padded domain
template <int NS>
__device__ __forceinline__ void props_body(const FrArgs& A, const int wx, const int wy, const int wz) {
  const FrBlockDev& B = A.B;
  const long long p = (long long)wx * kT + threadIdx.x;
  if (p >= B.sz) return;
  const long long s = B.sz;
  const double t = B.bv[p + 5 * s];
  const double u = B.bv[p + s], v = B.bv[p + 2 * s], w = B.bv[p + 3 * s];
  const double V = sqrt(u * u + v * v + w * w);
  double mw{0}, cp_tot{0}, cvv{0}, cp[NS], Y[NS];
  heat_capacity<NS>(A.T, t, cp);
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    Y[l] = B.sv[p + l * s];
    mw += Y[l] / A.T->mw[l];
    cp_tot += Y[l] * cp[l];
    cvv += Y[l] * (cp[l] - kRu / A.T->mw[l]);
  }
  mw = 1 / mw;
  const double gam = cp_tot / cvv;
  const double c = sqrt(gam * kRu * t / mw);
  B.cp[p] = cp_tot;
  B.gam[p] = gam;
  B.cc[p] = c;
  double mu, lam, rd[NS];
  transport<NS>(A.T, A.tlog, t, B.bv[p + 4 * s], B.bv[p], Y, mw, cp, &mu, &lam, rd);
  B.mul[p] = mu;
  B.lam[p] = lam;
#pragma unroll
  for (int l = 0; l < NS; ++l) B.rhod[p + l * s] = rd[l];
  B.mach[p] = V / c;
}
template <int NS>
__global__ __launch_bounds__(kT) void k_props(FrArgs A) {
  props_body<NS>(A, blockIdx.x, blockIdx.y, blockIdx.z);
}
template <int NS>
__global__ __launch_bounds__(kT) void k_props_batched(FrBatch W) {
  const WgPos w = wg_pos(W);
  FrArgs A = W.common;
  A.B = W.blocks[w.b];
  props_body<NS>(A, w.x, w.y, w.z);
}

// dq = 0, then the chemical source J omega on the species rows
// (cudaMemset + compute_source<FR>, src/RK.cuh:114-118, src/SourceTerm.cuh:11-35)
template <int NS>
__device__ __forceinline__ void source_body(const FrArgs& A, const int wx, const int wy, const int wz) {
  const FrBlockDev& B = A.B;
  const long long q = (long long)wx * kT + threadIdx.x;
  if (q >= B.n0) return;
  const int i = (int)(q % B.mx), j = (int)((q / B.mx) % B.my), k = (int)(q / ((long long)B.mx * B.my));
  const long long p = pid(B, i, j, k);
  const long long n0 = B.n0;
#pragma unroll
  for (int l = 0; l < 5; ++l) B.dq[q + l * n0] = 0;
  if (A.reaction) {
    double Y[NS], om[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) Y[l] = B.sv[p + l * B.sz];
    chemistry<NS>(A.T, B.bv[p], B.bv[p + 5 * B.sz], Y, om);
    const double J = B.jac[p];
#pragma unroll
    for (int l = 0; l < NS; ++l) B.dq[q + (5 + l) * n0] = 0.0 + J * om[l];
  } else {
#pragma unroll
    for (int l = 0; l < NS; ++l) B.dq[q + (5 + l) * n0] = 0;
  }
}
template <int NS>
__global__ __launch_bounds__(kT) void k_source(FrArgs A) {
  source_body<NS>(A, blockIdx.x, blockIdx.y, blockIdx.z);
}
template <int NS>
__global__ __launch_bounds__(kT) void k_source_batched(FrBatch W) {
  const WgPos w = wg_pos(W);
  FrArgs A = W.common;
  A.B = W.blocks[w.b];
  source_body<NS>(A, w.x, w.y, w.z);
}

__global__ void k_dt_set(double* dt, double v) { *dt = v; }

// local_time_step<FR, Laminar>, src/TimeAdvanceFunc.cuh:28-83, and the MIN of
// global_time_step: dt_local per cell, a 64-lane shuffle MIN per wave, the
// waves' MIN through LDS and one 64-bit atomicMin per workgroup on the bit
// pattern of a positive double (one per wave put ~260k atomics on one
// address for 16.5 M cells, most of the kernel's time).
__device__ __forceinline__ void local_dt_body(const FrArgs& A, unsigned long long* dt_bits, const int wx, const int wy, const int wz) {
  const FrBlockDev& B = A.B;
  const long long q = (long long)wx * kT + threadIdx.x;
  double dtl = 1e6;
  if (q < B.n0) {
    const int i = (int)(q % B.mx), j = (int)((q / B.mx) % B.my), k = (int)(q / ((long long)B.mx * B.my));
    const long long p = pid(B, i, j, k), s = B.sz;
    auto m = [&](int r, int c) { return B.met[p + (3 * (r - 1) + (c - 1)) * s]; };
    const double gxi = sqrt(m(1, 1) * m(1, 1) + m(1, 2) * m(1, 2) + m(1, 3) * m(1, 3));
    const double geta = sqrt(m(2, 1) * m(2, 1) + m(2, 2) * m(2, 2) + m(2, 3) * m(2, 3));
    const double gzeta = sqrt(m(3, 1) * m(3, 1) + m(3, 2) * m(3, 2) + m(3, 3) * m(3, 3));
    const double u = B.bv[p + s], v = B.bv[p + 2 * s], w = B.bv[p + 3 * s];
    const double U = u * m(1, 1) + v * m(1, 2) + w * m(1, 3);
    const double V = u * m(2, 1) + v * m(2, 2) + w * m(2, 3);
    const double W = u * m(3, 1) + v * m(3, 2) + w * m(3, 3);
    const double a = B.cc[p];
    const double r0 = fabs(U) + a * gxi, r1 = fabs(V) + a * geta, r2 = fabs(W) + a * gzeta;
    const double sr_inv = r0 + r1 + r2;
    const double coeff_1 = fmax(B.gam[p], 4.0 / 3.0) / B.bv[p];
    const double coeff_2 = B.mul[p] / A.prandtl;
    double sr_visc = gxi * gxi + geta * geta;
    sr_visc += gzeta * gzeta;
    sr_visc *= coeff_1 * coeff_2;
    dtl = A.cfl / (sr_inv + sr_visc);
    B.dtl[q] = dtl;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) dtl = fmin(dtl, __shfl_xor(dtl, o, 64));
  __shared__ double s_min[kT / 64];
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = dtl;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = s_min[0];
#pragma unroll
    for (int w = 1; w < kT / 64; ++w) m = fmin(m, s_min[w]);
    atomicMin(dt_bits, (unsigned long long)__double_as_longlong(m));
  }
}
__global__ __launch_bounds__(kT) void k_local_dt(FrArgs A, unsigned long long* dt_bits) {
  local_dt_body(A, dt_bits, blockIdx.x, blockIdx.y, blockIdx.z);
}
__global__ __launch_bounds__(kT) void k_local_dt_batched(FrBatch W, unsigned long long* dt_bits) {
  const WgPos w = wg_pos(W);
  FrArgs A = W.common;
  A.B = W.blocks[w.b];
  local_dt_body(A, dt_bits, w.x, w.y, w.z);
}

// update_cv_and_bv_rk<FR>, src/RK.cuh:230-284 + compute_temperature_and_pressure
template <int NS>
__device__ __forceinline__ void update_body(const FrArgs& A, double ra, double rb, double rc, const int wx, const int wy, const int wz) {
  const FrBlockDev& B = A.B;
  const long long q = (long long)wx * kT + threadIdx.x;
  if (q >= B.n0) return;
  const int i = (int)(q % B.mx), j = (int)((q / B.mx) % B.my), k = (int)(q / ((long long)B.mx * B.my));
  const long long p = pid(B, i, j, k), s = B.sz, n0 = B.n0;
  const double dt = *A.dt;
  const double dt_div_jac = dt / B.jac[p];
  double cv[5 + NS];
#pragma unroll
  for (int l = 0; l < 5 + NS; ++l) {
    cv[l] = ra * B.qn[p + l * s] + rb * B.cv[p + l * s] + rc * dt_div_jac * B.dq[q + l * n0];
    B.cv[p + l * s] = cv[l];
  }
  const double density_inv = 1.0 / cv[0];
  const double u = cv[1] * density_inv, v = cv[2] * density_inv, w = cv[3] * density_inv;
  B.bv[p] = cv[0];
  B.bv[p + s] = u;
  B.bv[p + 2 * s] = v;
  B.bv[p + 3 * s] = w;
  double Y[NS], mw{0};
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    Y[l] = cv[5 + l] * density_inv;
    B.sv[p + l * s] = Y[l];
  }
#pragma unroll
  for (int l = 0; l < NS; ++l) mw += Y[l] / A.T->mw[l];
  mw = 1 / mw;
  const double gas_const = kRu / mw;
  const double e = cv[4] / cv[0] - 0.5 * (u * u + v * v + w * w);
  const double t = temperature<NS>(A.T, e, Y, gas_const, B.bv[p + 5 * s]);
  B.bv[p + 5 * s] = t;
  B.bv[p + 4 * s] = cv[0] * t * gas_const;
}
template <int NS>
__global__ __launch_bounds__(kT) void k_update(FrArgs A, double ra, double rb, double rc) {
  update_body<NS>(A, ra, rb, rc, blockIdx.x, blockIdx.y, blockIdx.z);
}
template <int NS>
__global__ __launch_bounds__(kT) void k_update_batched(FrBatch W, double ra, double rb, double rc) {
  const WgPos w = wg_pos(W);
  FrArgs A = W.common;
  A.B = W.blocks[w.b];
  update_body<NS>(A, ra, rb, rc, w.x, w.y, w.z);
}

// apply_outflow / apply_periodic with scalars, src/BoundCond.cuh:205-233,
// :1443-1486: one thread per face point over the face's padded tangential
// extent, looping over the ngg ghost layers.
// apply_farfield<FR>, src/BoundCond.cuh:755-1022, at face point ix: the
// face point's normal Mach number (gamma from the property field) picks
// supersonic inflow (the far field on the face point and its ghosts),
// supersonic outflow (copy to the ghosts) or the subsonic state of the two
// Riemann invariants, whose gamma is iterated from 1.4 with cp of the
// boundary temperature and composition until two values agree to 1e-4.
// F: density, u, v, w, pressure, temperature, acoustic_speed, entropy,
// specific_heat_ratio, mw, sv[16].
template <int NS>
__device__ void farfield_point(const FrArgs& A, int axis, int side, const double* __restrict__ F, const int* ix) {
  const FrBlockDev& B = A.B;
  const long long s = B.sz;
  const long long p = pid(B, ix[0], ix[1], ix[2]);
  double nx = B.met[p + (3 * axis) * s], ny = B.met[p + (3 * axis + 1) * s], nz = B.met[p + (3 * axis + 2) * s];
  const double grad_n_inv = side / sqrt(nx * nx + ny * ny + nz * nz);
  nx *= grad_n_inv;
  ny *= grad_n_inv;
  nz *= grad_n_inv;
  const double u_b = B.bv[p + s], v_b = B.bv[p + 2 * s], w_b = B.bv[p + 3 * s];
  const double u_face = nx * u_b + ny * v_b + nz * w_b;
  double sv_b[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) sv_b[l] = B.sv[p + l * s];
  const double gamma_b = B.gam[p];
  double mw_inv{0};
#pragma unroll
  for (int l = 0; l < NS; ++l) mw_inv += sv_b[l] / A.T->mw[l];
  double mw = 1.0 / mw_inv;
  const double p_b = B.bv[p + 4 * s], rho_b = B.bv[p];
  const double a_b = sqrt(gamma_b * p_b / rho_b);
  const double mach_b = u_face / a_b;
  double st[6];
  bool face_too = true;
  if (mach_b <= -1) {
#pragma unroll
    for (int l = 0; l < 6; ++l) st[l] = F[l];
#pragma unroll
    for (int l = 0; l < NS; ++l) sv_b[l] = F[10 + l];
  } else if (mach_b >= 1) {
#pragma unroll
    for (int l = 0; l < 6; ++l) st[l] = B.bv[p + l * s];
    face_too = false;
  } else {
    const double riemann_pos = u_face + 2 * a_b / (gamma_b - 1);
    const double u_inf = F[1] * nx + F[2] * ny + F[3] * nz;
    const double riemann_neg = u_inf - 2 * F[6] / (F[8] - 1);
    const double Un = 0.5 * (riemann_pos + riemann_neg);
    if (mach_b < 0) {
      st[1] = F[1] + (Un - u_inf) * nx;
      st[2] = F[2] + (Un - u_inf) * ny;
      st[3] = F[3] + (Un - u_inf) * nz;
#pragma unroll
      for (int l = 0; l < NS; ++l) sv_b[l] = F[10 + l];
      mw = F[9];
    } else {
      st[1] = u_b + (Un - u_face) * nx;
      st[2] = v_b + (Un - u_face) * ny;
      st[3] = w_b + (Un - u_face) * nz;
    }
    double gamma{1.4}, err{1};
    double density = 0, pressure = 0, temperature = 0;
    // the iteration is a fixed point of a smooth contraction; the bound only
    // guarantees every lane leaves the loop
    for (int it = 0; it < 1000 && err > 1e-4; ++it) {
      const double gamma_last = gamma;
      const double c_b = 0.25 * (gamma - 1) * (riemann_pos - riemann_neg);
      const double s_b = mach_b <= 0 ? F[7] : p_b / pow(rho_b, gamma);
      density = pow(c_b * c_b / (gamma * s_b), 1 / (gamma - 1));
      pressure = density * c_b * c_b / gamma;
      temperature = pressure * mw / (density * kRu);
      double cp[NS];
      heat_capacity<NS>(A.T, temperature, cp);
      double cp_tot{0};
#pragma unroll
      for (int l = 0; l < NS; ++l) cp_tot += cp[l] * sv_b[l];
      gamma = cp_tot / (cp_tot - kRu / mw);
      err = fabs(1 - gamma / gamma_last);
    }
    st[0] = density;
    st[4] = pressure;
    st[5] = temperature;
  }
  if (face_too) {
#pragma unroll
    for (int l = 0; l < 6; ++l) B.bv[p + l * s] = st[l];
#pragma unroll
    for (int l = 0; l < NS; ++l) B.sv[p + l * s] = sv_b[l];
    cv_point<NS>(A, p);
  }
  for (int g = 1; g <= B.g; ++g) {
    int gi[3] = {ix[0], ix[1], ix[2]};
    gi[axis] += g * side;
    const long long pg = pid(B, gi[0], gi[1], gi[2]);
#pragma unroll
    for (int l = 0; l < 6; ++l) B.bv[pg + l * s] = st[l];
#pragma unroll
    for (int l = 0; l < NS; ++l) B.sv[pg + l * s] = sv_b[l];
    cv_point<NS>(A, pg);
  }
}

template <int NS>
__global__ __launch_bounds__(kT) void k_bc(FrArgs A, FaceBox fb, int type, const double* ff, const double* y,
                                           const unsigned long long* pass) {
  const FrBlockDev& B = A.B;
  const int ext[3] = {B.mx, B.my, B.mz};
  const int axis = fb.axis, side = fb.side;
  const int q = blockIdx.x * kT + threadIdx.x;
  if (q >= fb.n[0] * fb.n[1] * fb.n[2]) return;
  int ix[3];
  ix[0] = fb.lo[0] + q % fb.n[0];
  ix[1] = fb.lo[1] + (q / fb.n[0]) % fb.n[1];
  ix[2] = fb.lo[2] + q / (fb.n[0] * fb.n[1]);
  if (type == 4) {
    farfield_point<NS>(A, axis, side, ff, ix);
    return;
  }
  if (type == 5) {  // apply_inflow, src/BoundCond.cuh:235-612
    double st[6];
    const bool upper = ostc::inflow_state(ff, y[q], *pass, ostc::inflow_point(axis, ix[0], ix[1], ix[2], B.mx, B.my, B.g), st);
    const double* sv_in = ff + (upper ? ostc::kInflowSvUpper : ostc::kInflowSvLower);
    const long long s = B.sz;
    for (int g = 0; g <= B.g; ++g) {
      int gi[3] = {ix[0], ix[1], ix[2]};
      gi[axis] += g * side;
      const long long pg = pid(B, gi[0], gi[1], gi[2]);
#pragma unroll
      for (int l = 0; l < 6; ++l) B.bv[pg + l * s] = st[l];
#pragma unroll
      for (int l = 0; l < NS; ++l) B.sv[pg + l * s] = sv_in[l];
      cv_point<NS>(A, pg);
    }
    return;
  }
  int src[3] = {ix[0], ix[1], ix[2]};
  if (type == 8) src[axis] = side < 0 ? ext[axis] - 1 : 0;
  const long long s = B.sz;
  for (int g = 1; g <= B.g; ++g) {
    int gi[3] = {ix[0], ix[1], ix[2]}, si[3] = {src[0], src[1], src[2]};
    gi[axis] += g * side;
    if (type == 8) si[axis] += g * side;
    const long long pg = pid(B, gi[0], gi[1], gi[2]), ps = pid(B, si[0], si[1], si[2]);
#pragma unroll
    for (int l = 0; l < 6; ++l) B.bv[pg + l * s] = B.bv[ps + l * s];
#pragma unroll
    for (int l = 0; l < NS; ++l) B.sv[pg + l * s] = B.sv[ps + l * s];
    cv_point<NS>(A, pg);
  }
}

// inner_communication with scalars, src/DataCommunication.cuh:60-105: the
// shared node layer takes the mean of both blocks (in both), the ghost layers
// the target's values.
template <int NS>
__global__ __launch_bounds__(kT) void k_comm(FrArgs A, FrBlockDev Tg, CommMap m) {
  const FrBlockDev& B = A.B;
  const long long q = (long long)blockIdx.x * kT + threadIdx.x;
  if (q >= (long long)m.n[0] * m.n[1] * m.n[2]) return;
  const int i = m.lo[0] + (int)(q % m.n[0]);
  const int j = m.lo[1] + (int)((q / m.n[0]) % m.n[1]);
  const int k = m.lo[2] + (int)(q / ((long long)m.n[0] * m.n[1]));
  auto src_index = [&](int code) {
    const int a = (code < 0 ? -code : code) - 1;
    const int v = a == 0 ? i : (a == 1 ? j : k);
    return code < 0 ? -v : v;
  };
  const int tx0 = m.off[0] + src_index(m.code[0]), tx1 = m.off[1] + src_index(m.code[1]),
            tx2 = m.off[2] + src_index(m.code[2]);
  const bool shared = (m.axis == 0 ? i : (m.axis == 1 ? j : k)) == m.shared;
  const long long p = pid(B, i, j, k), pt = pid(Tg, tx0, tx1, tx2);
  const long long s = B.sz, ts = Tg.sz;
  if (shared) {
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      const double ave{0.5 * (Tg.bv[pt + l * ts] + B.bv[p + l * s])};
      B.bv[p + l * s] = ave;
      Tg.bv[pt + l * ts] = ave;
    }
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const double ave{0.5 * (Tg.sv[pt + l * ts] + B.sv[p + l * s])};
      B.sv[p + l * s] = ave;
      Tg.sv[pt + l * ts] = ave;
    }
  } else {
#pragma unroll
    for (int l = 0; l < 6; ++l) B.bv[p + l * s] = Tg.bv[pt + l * ts];
#pragma unroll
    for (int l = 0; l < NS; ++l) B.sv[p + l * s] = Tg.sv[pt + l * ts];
  }
  cv_point<NS>(A, p);
}

// ------------------------------------------------------------ tile geometry
// DIR 0: TW = 4 columns along j, 64 faces per column; DIR 1, 2: TW = 16
// columns along i, 16 faces per column.
template <int DIR> struct Tile;
template <> struct Tile<0> { static constexpr int TW = 4, F = 64; };
template <> struct Tile<1> { static constexpr int TW = 16, F = 16; };
template <> struct Tile<2> { static constexpr int TW = 16, F = 16; };

// Coordinates of tile position (col, sweep index t) in block indices.
template <int DIR>
__device__ __forceinline__ void tile_idx(int col0, int t, int col, int third, int* idx) {
  if (DIR == 0) {
    idx[0] = t;
    idx[1] = col0 + col;
    idx[2] = third;
  } else if (DIR == 1) {
    idx[0] = col0 + col;
    idx[1] = t;
    idx[2] = third;
  } else {
    idx[0] = col0 + col;
    idx[1] = third;
    idx[2] = t;
  }
}

// LDS slot of face f of column col among a tile's face fluxes: along x the
// lanes of a wave take consecutive faces of a column, along y and z
// consecutive columns, so consecutive lanes always hit consecutive slots
// (a column-major slot along y/z put the lanes F doubles apart: bank
// conflicts on every face-flux store and read).
template <int DIR, int TW, int F>
__device__ __forceinline__ int face_slot(int col, int f) {
  return DIR == 0 ? col * F + f : f * TW + col;
}

// Positivity-preserving limiter on the species part of one face,
// src/WENO.cu:1538-1580.
template <int NS>
__device__ __forceinline__ void pp_limit(double* fc, const double* f1, const double* ql, const double* qr, double jl,
                                         double jr, double dt, int idx, int max_extent) {
  const double alpha = 1.0 / 3.0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    double theta_p = 1.0, theta_m = 1.0;
    if (idx > -1) {
      const double up = 0.5 * alpha * ql[l] * jl - dt * fc[l];
      if (up < 0) {
        const double up_lf = 0.5 * alpha * ql[l] * jl - dt * f1[l];
        if (fabs(up - up_lf) > 1e-20) {
          theta_p = (0 - up_lf) / (up - up_lf);
          if (theta_p > 1) theta_p = 1.0;
          else if (theta_p < 0) theta_p = 0;
        }
      }
    }
    if (idx < max_extent - 1) {
      const double um = 0.5 * alpha * qr[l] * jr + dt * fc[l];
      if (um < 0) {
        const double um_lf = 0.5 * alpha * qr[l] * jr + dt * f1[l];
        if (fabs(um - um_lf) > 1e-20) {
          theta_m = (0 - um_lf) / (um - um_lf);
          if (theta_m > 1) theta_m = 1.0;
          else if (theta_m < 0) theta_m = 0;
        }
      }
    }
    fc[l] = fmin(theta_p, theta_m) * (fc[l] - f1[l]) + f1[l];
  }
}

// Point-local Lax-Friedrichs split fluxes F+ and F- (compute_flux,
// src/WENO.cu:697-720) of the TW columns x P points of a sweep tile, cells
// t0-G-LEFT .. t0+P-G-LEFT-1 along DIR, into LDS planes [NV][P * TW]; points
// past the end of the padded block hold zeros, and a point before its start
// (t < -G, only with LEFT > 0) holds the fluxes of the outermost ghost point.
template <int DIR, int NS, int G, int TW, int P, int LEFT = 0>
__device__ __forceinline__ void split_fluxes(const FrArgs& A, double* Fp, double* Fm, int col0, int third, int t0) {
  constexpr int NV = 5 + NS, NP = P * TW;
  const FrBlockDev& B = A.B;
  const int ext[3] = {B.mx, B.my, B.mz};
  const int m = ext[DIR];
  const int across = DIR == 0 ? 1 : 0;
  const long long s = B.sz;
  const int tid = threadIdx.x;
  for (int e = tid; e < NP; e += kT) {
    int col, sp;
    if (DIR == 0) {
      col = e / P;
      sp = e % P;
    } else {
      col = e % TW;
      sp = e / TW;
    }
    const int t = t0 - G - LEFT + sp;
    int idx[3];
    tile_idx<DIR>(col0, t < -G ? -G : t, col, third, idx);
    const int slot = col * P + sp;
    if (idx[across] < ext[across] && t < m + G) {
      const long long p = pid(B, idx[0], idx[1], idx[2]);
      double q[NV + 2], kk[3];
#pragma unroll
      for (int l = 0; l < NV; ++l) q[l] = B.cv[p + l * s];
      q[NV] = B.bv[p + 4 * s];
      q[NV + 1] = B.cc[p];
#pragma unroll
      for (int c = 0; c < 3; ++c) kk[c] = B.met[p + (3 * DIR + c) * s];
      const double jac = B.jac[p];
      // compute_flux, src/WENO.cu:697-720
      const double Uk{(q[1] * kk[0] + q[2] * kk[1] + q[3] * kk[2]) / q[0]};
      const double pk{q[NV]};
      const double cGradK = q[NV + 1] * sqrt(kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2]);
      const double lambda0 = fabs(Uk) + cGradK;
      Fp[0 * NP + slot] = 0.5 * jac * (Uk * q[0] + lambda0 * q[0]);
      Fp[1 * NP + slot] = 0.5 * jac * (Uk * q[1] + pk * kk[0] + lambda0 * q[1]);
      Fp[2 * NP + slot] = 0.5 * jac * (Uk * q[2] + pk * kk[1] + lambda0 * q[2]);
      Fp[3 * NP + slot] = 0.5 * jac * (Uk * q[3] + pk * kk[2] + lambda0 * q[3]);
      Fp[4 * NP + slot] = 0.5 * jac * (Uk * (q[4] + pk) + lambda0 * q[4]);
      Fm[0 * NP + slot] = 0.5 * jac * (Uk * q[0] - lambda0 * q[0]);
      Fm[1 * NP + slot] = 0.5 * jac * (Uk * q[1] + pk * kk[0] - lambda0 * q[1]);
      Fm[2 * NP + slot] = 0.5 * jac * (Uk * q[2] + pk * kk[1] - lambda0 * q[2]);
      Fm[3 * NP + slot] = 0.5 * jac * (Uk * q[3] + pk * kk[2] - lambda0 * q[3]);
      Fm[4 * NP + slot] = 0.5 * jac * (Uk * (q[4] + pk) - lambda0 * q[4]);
#pragma unroll
      for (int l = 5; l < NV; ++l) {
        Fp[l * NP + slot] = 0.5 * jac * (Uk * q[l] + lambda0 * q[l]);
        Fm[l * NP + slot] = 0.5 * jac * (Uk * q[l] - lambda0 * q[l]);
      }
    } else {
#pragma unroll
      for (int l = 0; l < NV; ++l) Fp[l * NP + slot] = Fm[l * NP + slot] = 0;
    }
  }
}

// One inviscid sweep along DIR (compute_convective_term_weno_{x,y,z}<FR>,
// src/WENO.cu:8-461) with the component-wise WENO-Z5 (51) or WENO-Z7 (71)
// flux, src/WENO.cu:1455-1536; the characteristic schemes run in
// k_inviscid_ch.
template <int DIR, int NS, int SCHEME>
__device__ __forceinline__ void inviscid_body(const FrArgs& A, const int wx, const int wy, const int wz) {
  // WENO-Z5 (51, 52): 3 ghost layers, 5-point candidates; WENO-Z7 (71, 72): 4 and 7
  constexpr int G = (SCHEME == 71 || SCHEME == 72) ? 4 : 3, W = 2 * G - 1;
  constexpr int NV = 5 + NS;
  constexpr int TW = Tile<DIR>::TW, F = Tile<DIR>::F, TS = F - 1, P = TS + 2 * G, NP = P * TW;
  extern __shared__ double lds[];
  double* Fp = lds;            // [NV][NP]
  double* Fm = lds + NV * NP;  // [NV][NP]
  const FrBlockDev& B = A.B;
  const int ext[3] = {B.mx, B.my, B.mz};
  const int m = ext[DIR];
  const int across = DIR == 0 ? 1 : 0;
  const int col0 = wy * TW;
  const int third = wz;
  const int t0 = wx * TS;  // first cell of the tile along DIR
  const long long s = B.sz;
  const int tid = threadIdx.x;

  // 1. split fluxes of the tile's points (cells t0-G .. t0+TS+G-1)
  split_fluxes<DIR, NS, G, TW, P>(A, Fp, Fm, col0, third, t0);
  __syncthreads();

  // 2. one face per thread: face between cells t0-1+f and t0+f
  // face f of column col; along x consecutive lanes take consecutive faces,
  // along y and z consecutive columns (consecutive i), so the global loads of
  // a wave are contiguous in both cases
  const int f = DIR == 0 ? tid % F : tid / TW, col = DIR == 0 ? tid / F : tid % TW;
  const int tl = t0 - 1 + f;  // left cell of the face
  int idx[3];
  tile_idx<DIR>(col0, tl, col, third, idx);
  const bool active = idx[across] < ext[across] && tl < m;
  double fc[NV];
  if (active) {
    const int sl = col * P + (f + G - 1);  // LDS slot of the left cell
    const long long pl = pid(B, idx[0], idx[1], idx[2]);
    int idr[3] = {idx[0], idx[1], idx[2]};
    idr[DIR] += 1;
    const long long pr = pid(B, idr[0], idr[1], idr[2]);
    double ml[3], mr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ml[c] = B.met[pl + (3 * DIR + c) * s];
      mr[c] = B.met[pr + (3 * DIR + c) * s];
    }
    const double jac1 = B.jac[pl], jac2 = B.jac[pr];
    double f1[NS];
    {
      constexpr double eps{1e-20};
      const double eps_ref = eps * A.weno_eps_scale * 0.25 *
                             ((ml[0] * jac1 + mr[0] * jac2) * (ml[0] * jac1 + mr[0] * jac2) +
                              (ml[1] * jac1 + mr[1] * jac2) * (ml[1] * jac1 + mr[1] * jac2) +
                              (ml[2] * jac1 + mr[2] * jac2) * (ml[2] * jac1 + mr[2] * jac2));
      double eps_scaled[3];
      eps_scaled[0] = eps_ref;
      eps_scaled[1] = eps_ref * A.v_ref * A.v_ref;
      eps_scaled[2] = eps_scaled[1] * A.v_ref * A.v_ref;
      if (A.pp) {
#pragma unroll
        for (int l = 0; l < NS; ++l) f1[l] = 0.5 * Fp[(l + 5) * NP + sl] + 0.5 * Fm[(l + 5) * NP + sl + 1];
      }
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        const double eh = (l == 1 || l == 2 || l == 3) ? eps_scaled[1] : (l == 4 ? eps_scaled[2] : eps_scaled[0]);
        double vp[W], vm[W];
#pragma unroll
        for (int r = 0; r < W; ++r) {
          vp[r] = Fp[l * NP + sl - W / 2 + r];
          vm[r] = Fm[l * NP + sl - W / 2 + 1 + r];
        }
        if constexpr (W == 7) fc[l] = weno7z(vp, vm, eh);  // compute_weno_flux_cp, 71 (src/WENO.cu:1516-1534)
        else fc[l] = weno5z(vp, vm, eh);
      }
      if (A.pp) {
        double ql[NS], qr[NS];
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          ql[l] = B.cv[pl + (5 + l) * s];
          qr[l] = B.cv[pr + (5 + l) * s];
        }
        pp_limit<NS>(fc + 5, f1, ql, qr, jac1, jac2, *A.dt, tl, m);
      }
    }
  }
  __syncthreads();  // every face has read its stencil; the LDS now holds face fluxes
  double* fcs = lds;  // [NV][F * TW], face slot face_slot(col, f)
  if (active) {
#pragma unroll
    for (int l = 0; l < NV; ++l) fcs[l * (F * TW) + face_slot<DIR, TW, F>(col, f)] = fc[l];
  }
  __syncthreads();

  // 3. cell t0-1+f (f >= 1): dq -= fc(right) - fc(left)
  if (f >= 1 && active) {
    const long long q0 = pid0(B, idx[0], idx[1], idx[2]);
    const long long n0 = B.n0;
#pragma unroll
    for (int l = 0; l < NV; ++l)
      B.dq[q0 + l * n0] -= fcs[l * (F * TW) + face_slot<DIR, TW, F>(col, f)] -
                           fcs[l * (F * TW) + face_slot<DIR, TW, F>(col, f - 1)];
  }
}
template <int DIR, int NS, int SCHEME>
__global__ __launch_bounds__(kT) void k_inviscid(FrArgs A) {
  inviscid_body<DIR, NS, SCHEME>(A, blockIdx.x, blockIdx.y, blockIdx.z);
}
template <int DIR, int NS, int SCHEME>
__global__ __launch_bounds__(kT) void k_inviscid_batched(FrBatch W) {
  const WgPos w = wg_pos(W);
  FrArgs A = W.common;
  A.B = W.blocks[w.b];
  inviscid_body<DIR, NS, SCHEME>(A, w.x, w.y, w.z);
}

// Tiles of the characteristic sweep: 128 faces per 256-thread workgroup,
// two threads per face.
template <int DIR> struct TileCh;
template <> struct TileCh<0> { static constexpr int TW = 2, F = 64; };
template <> struct TileCh<1> { static constexpr int TW = 8, F = 16; };
template <> struct TileCh<2> { static constexpr int TW = 8, F = 16; };

// One characteristic WENO sweep along DIR (compute_weno_flux_ch<FR>, "method
// 1", src/WENO.cu:722-1143; 72: 7 points and WENO-Z7).  Two threads share a
// face: waves 0-1 (half 0) project and reconstruct the acoustic/entropy fields
// 0-2 and the first NS/2 species fields, waves 2-3 (half 1) fields 3-4 and the
// other species; the characteristic fluxes meet in LDS, half 0 back-projects
// the five flow components and half 1 the species (with the positivity
// limiter), each in the reference's term order, so the face flux is the
// same as one thread computing every field.  Both halves form the face's Roe
// state themselves.
template <int DIR, int NS, int SCHEME>
__device__ __forceinline__ void inviscid_ch_body(const FrArgs& A, const int wx, const int wy, const int wz) {
  constexpr int G = SCHEME == 72 ? 4 : 3, W = 2 * G - 1;
  // scheme 52 keeps one more point on the left for the reference's stencil
  // of the acoustic/entropy fields (A.ch_ref: F+ at i-3..i+1, F- at i-2..i+2
  // for the five conservative components, src/WENO.cu:880-881)
  constexpr int LEFT = SCHEME == 52 ? 1 : 0;
  constexpr int NV = 5 + NS, SH = NS / 2;  // species fields < SH belong to half 0
  constexpr int TW = TileCh<DIR>::TW, F = TileCh<DIR>::F, TS = F - 1, P = TS + 2 * G + LEFT, NP = P * TW;
  constexpr int NF = F * TW;  // faces of the tile (128)
  extern __shared__ double lds[];
  double* Fp = lds;                // [NV][NP]
  double* Fm = lds + NV * NP;      // [NV][NP]
  double* fch = lds + 2 * NV * NP; // [NV][NF] characteristic fluxes of every face
  const FrBlockDev& B = A.B;
  const int ext[3] = {B.mx, B.my, B.mz};
  const int m = ext[DIR];
  const int across = DIR == 0 ? 1 : 0;
  const int col0 = wy * TW;
  const int third = wz;
  const int t0 = wx * TS;
  const long long s = B.sz;
  const int tid = threadIdx.x;

  split_fluxes<DIR, NS, G, TW, P, LEFT>(A, Fp, Fm, col0, third, t0);
  __syncthreads();
  const int sh = SCHEME == 52 ? A.ch_ref : 0;  // shift of the conservative components' stencil

  const int half = tid / NF, q = tid % NF;  // wave-uniform half
  const int f = DIR == 0 ? q % F : q / TW, col = DIR == 0 ? q / F : q % TW;
  const int tl = t0 - 1 + f;
  int idx[3];
  tile_idx<DIR>(col0, tl, col, third, idx);
  const bool active = idx[across] < ext[across] && tl < m;
  double fc[NV];
  CharFace<NS> C;
  long long pl = 0, pr = 0;
  double jac1 = 0, jac2 = 0, f1[NS];
  const int sl = col * P + (f + G - 1 + LEFT);  // LDS slot of the left cell
  if (active) {
    pl = pid(B, idx[0], idx[1], idx[2]);
    int idr[3] = {idx[0], idx[1], idx[2]};
    idr[DIR] += 1;
    pr = pid(B, idr[0], idr[1], idr[2]);
    double ml[3], mr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ml[c] = B.met[pl + (3 * DIR + c) * s];
      mr[c] = B.met[pr + (3 * DIR + c) * s];
    }
    jac1 = B.jac[pl];
    jac2 = B.jac[pr];
    {
      double cvl[NV + 2], cvr[NV + 2];
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        cvl[l] = B.cv[pl + l * s];
        cvr[l] = B.cv[pr + l * s];
      }
      cvl[NV] = B.bv[pl + 4 * s];
      cvr[NV] = B.bv[pr + 4 * s];
      cvl[NV + 1] = B.cc[pl];
      cvr[NV + 1] = B.cc[pr];
      char_face<NS>(A.T, cvl, cvr, ml, mr, jac1, jac2, C);
      if (A.pp && half == 1) {
        const double srThis = fabs((ml[0] * cvl[1] + ml[1] * cvl[2] + ml[2] * cvl[3]) / cvl[0] +
                                   cvl[NV + 1] * sqrt(ml[0] * ml[0] + ml[1] * ml[1] + ml[2] * ml[2]));
        const double srNext = fabs((mr[0] * cvr[1] + mr[1] * cvr[2] + mr[2] * cvr[3]) / cvr[0] +
                                   cvr[NV + 1] * sqrt(mr[0] * mr[0] + mr[1] * mr[1] + mr[2] * mr[2]));
#pragma unroll
        for (int l = 0; l < NS; ++l)
          f1[l] = 0.5 * (Fp[(l + 5) * NP + sl] + srThis * cvl[l + 5] * jac1) +
                  0.5 * (Fp[(l + 5) * NP + sl + 1] - srNext * cvr[l + 5] * jac2);
      }
    }
    constexpr double eps{1e-40};
    const double eps_scaled = eps * A.weno_eps_scale * 0.25 * (C.nj[0] * C.nj[0] + C.nj[1] * C.nj[1] + C.nj[2] * C.nj[2]);
    // acoustic and entropy fields of this half
    const int l_lo = half == 0 ? 0 : 3, l_hi = half == 0 ? 3 : 5;
#pragma unroll 1
    for (int l = l_lo; l < l_hi; ++l) {
      double L[5];
      const double cas = char_row<NS>(C, l, L);
      double vPlus[W], vMinus[W];
#pragma unroll 1
      for (int r = 0; r < W; ++r) {
        double ap = 0, am = 0;
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          ap += L[n] * Fp[n * NP + sl - W / 2 + r - sh];
          am += L[n] * Fm[n * NP + sl - W / 2 + 1 + r - sh];
        }
#pragma unroll
        for (int n = 0; n < NS; ++n) {
          ap += cas * C.coupling[n] * Fp[(5 + n) * NP + sl - W / 2 + r];
          am += cas * C.coupling[n] * Fm[(5 + n) * NP + sl - W / 2 + 1 + r];
        }
#pragma unroll
        for (int qq = 0; qq < W; ++qq) {
          vPlus[qq] = (qq == r) ? ap : vPlus[qq];
          vMinus[qq] = (qq == r) ? am : vMinus[qq];
        }
      }
      double v;
      if constexpr (W == 7) v = weno7z(vPlus, vMinus, eps_scaled);
      else v = weno5z(vPlus, vMinus, eps_scaled);
      fch[l * NF + q] = v;
    }
    // species fields of this half (unrolled: CharFace stays in registers)
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      if ((l < SH) != (half == 0)) continue;
      double vPlus[W], vMinus[W];
      const double y = C.y[l];
#pragma unroll
      for (int r = 0; r < W; ++r) {
        vPlus[r] = -y * Fp[sl - W / 2 + r] + Fp[(5 + l) * NP + sl - W / 2 + r];
        vMinus[r] = -y * Fm[sl - W / 2 + 1 + r] + Fm[(5 + l) * NP + sl - W / 2 + 1 + r];
      }
      double v;
      if constexpr (W == 7) v = weno7z(vPlus, vMinus, eps_scaled);
      else v = weno5z(vPlus, vMinus, eps_scaled);
      fch[(5 + l) * NF + q] = v;
    }
  }
  __syncthreads();
  if (active) {
    // back projection (src/WENO.cu:1121-1142): flux of variable r =
    // sum over the waves w = 0..4, in order, of R[r][w] times the wave's
    // characteristic flux (the first term initialises the sum)
    const double cw[5] = {fch[0 * NF + q], fch[1 * NF + q], fch[2 * NF + q], fch[3 * NF + q], fch[4 * NF + q]};
    double f0 = cw[0];  // R[0][0] = 1: the product is exact
#pragma unroll
    for (int w = 1; w < 5; ++w) f0 = f0 + (w == 4 ? cw[4] : pick3(C.n, w - 1) * cw[w]);
    if (half == 0) {
      double col[5], f[5];
      char_column<NS>(C, 0, col);
#pragma unroll
      for (int r = 1; r < 5; ++r) f[r] = col[r] * cw[0];
#pragma unroll
      for (int w = 1; w < 5; ++w) {
        char_column<NS>(C, w, col);
#pragma unroll
        for (int r = 1; r < 5; ++r) f[r] = f[r] + col[r] * cw[w];
      }
      double add{0};
#pragma unroll
      for (int l = 0; l < NS; ++l) add += C.coupling[l] * fch[(5 + l) * NF + q];
      fc[0] = f0;
      fc[1] = f[1];
      fc[2] = f[2];
      fc[3] = f[3];
      fc[4] = f[4] - add * C.c * C.c / C.gm1;
    } else {
#pragma unroll
      for (int l = 0; l < NS; ++l) fc[5 + l] = C.y[l] * f0 + fch[(5 + l) * NF + q];
      if (A.pp) {
        double ql[NS], qr[NS];
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          ql[l] = B.cv[pl + (5 + l) * s];
          qr[l] = B.cv[pr + (5 + l) * s];
        }
        pp_limit<NS>(fc + 5, f1, ql, qr, jac1, jac2, *A.dt, tl, m);
      }
    }
  }
  __syncthreads();  // every face has read F+/F- and fch; the LDS now holds face fluxes
  double* fcs = lds;  // [NV][NF], face slot face_slot(col, f)
  const int v_lo = half == 0 ? 0 : 5, v_hi = half == 0 ? 5 : NV;
  if (active) {
#pragma unroll
    for (int l = 0; l < NV; ++l)
      if (l >= v_lo && l < v_hi) fcs[l * NF + face_slot<DIR, TW, F>(col, f)] = fc[l];
  }
  __syncthreads();
  if (f >= 1 && active) {
    const long long q0 = pid0(B, idx[0], idx[1], idx[2]);
    const long long n0 = B.n0;
#pragma unroll
    for (int l = 0; l < NV; ++l)
      if (l >= v_lo && l < v_hi)
        B.dq[q0 + l * n0] -= fcs[l * NF + face_slot<DIR, TW, F>(col, f)] - fcs[l * NF + face_slot<DIR, TW, F>(col, f - 1)];
  }
}
template <int DIR, int NS, int SCHEME>
__global__ __launch_bounds__(kT) void k_inviscid_ch(FrArgs A) {
  inviscid_ch_body<DIR, NS, SCHEME>(A, blockIdx.x, blockIdx.y, blockIdx.z);
}
template <int DIR, int NS, int SCHEME>
__global__ __launch_bounds__(kT) void k_inviscid_ch_batched(FrBatch W) {
  const WgPos w = wg_pos(W);
  FrArgs A = W.common;
  A.B = W.blocks[w.b];
  inviscid_ch_body<DIR, NS, SCHEME>(A, w.x, w.y, w.z);
}

// One inviscid sweep along DIR with reconstructed primitives and an
// approximate Riemann solver: compute_convective_term_pv_1D<FR>,
// src/InviscidScheme.cu:50-165 (AUSM+ for scheme 3, HLLC for 4).  The tile
// holds rho, u, v, w, p, Y_k of TS + 4 points per column in LDS.
template <int DIR, int NS>
__device__ __forceinline__ void inviscid_pv_body(const FrArgs& A, const int wx, const int wy, const int wz) {
  constexpr int NV = 5 + NS;
  constexpr int TW = Tile<DIR>::TW, F = Tile<DIR>::F, TS = F - 1, P = TS + 4, NP = P * TW;
  extern __shared__ double lds[];
  double* s_delta = lds + NV * (P > F ? P : F) * TW;  // [NP] entropy-fix deltas (Roe)
  const FrBlockDev& B = A.B;
  const int ext[3] = {B.mx, B.my, B.mz};
  const int m = ext[DIR];
  const int across = DIR == 0 ? 1 : 0;
  const int col0 = wy * TW;
  const int third = wz;
  const int t0 = wx * TS;
  const long long s = B.sz;
  const int tid = threadIdx.x;
  for (int e = tid; e < NP; e += kT) {  // points t0-2 .. t0+TS+1
    int col, sp;
    if (DIR == 0) {
      col = e / P;
      sp = e % P;
    } else {
      col = e % TW;
      sp = e / TW;
    }
    const int t = t0 - 2 + sp;
    int idx[3];
    tile_idx<DIR>(col0, t, col, third, idx);
    const int slot = col * P + sp;
    if (idx[across] < ext[across] && t < m + 2) {
      const long long p = pid(B, idx[0], idx[1], idx[2]);
#pragma unroll
      for (int l = 0; l < 5; ++l) lds[l * NP + slot] = B.bv[p + l * s];
#pragma unroll
      for (int l = 0; l < NS; ++l) lds[(5 + l) * NP + slot] = B.sv[p + l * s];
      if (A.scheme == 2) {  // the point's entropy-fix delta from its full metric, once for both of its faces
        double mm[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) mm[c] = B.met[p + c * s];
        s_delta[slot] = entropy_delta_fr(B.bv[p + s], B.bv[p + 2 * s], B.bv[p + 3 * s], B.cc[p], mm, A.entropy_fix);
      }
    } else {
#pragma unroll
      for (int l = 0; l < NV; ++l) lds[l * NP + slot] = 1.0;
    }
  }
  __syncthreads();
  // face f of column col; along x consecutive lanes take consecutive faces,
  // along y and z consecutive columns (consecutive i), so the global loads of
  // a wave are contiguous in both cases
  const int f = DIR == 0 ? tid % F : tid / TW, col = DIR == 0 ? tid / F : tid % TW;
  const int tl = t0 - 1 + f;
  int idx[3];
  tile_idx<DIR>(col0, tl, col, third, idx);
  const bool active = idx[across] < ext[across] && tl < m;
  double fc[NV];
  if (active) {
    const int sl = col * P + f + 1;  // slot of the left cell
    double pl[NV + 2], pr[NV + 2];
#pragma unroll
    for (int l = 0; l < NV; ++l)
      reconstruct1(A.recon, lds[l * NP + sl - 1], lds[l * NP + sl], lds[l * NP + sl + 1], lds[l * NP + sl + 2], &pl[l],
                   &pr[l]);
    pv_energy<NS>(A.T, pl);
    pv_energy<NS>(A.T, pr);
    const long long p0 = pid(B, idx[0], idx[1], idx[2]);
    int idr[3] = {idx[0], idx[1], idx[2]};
    idr[DIR] += 1;
    const long long p1 = pid(B, idr[0], idr[1], idr[2]);
    double ml[3], mr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ml[c] = B.met[p0 + (3 * DIR + c) * s];
      mr[c] = B.met[p1 + (3 * DIR + c) * s];
    }
    if (A.scheme == 2) {
      roe_fr<NS>(A.T, pl, pr, ml, mr, B.jac[p0], B.jac[p1], s_delta[sl], s_delta[sl + 1], fc);
    } else if (A.scheme == 3)
      ausm_plus<NS>(pl, pr, ml, mr, B.jac[p0], B.jac[p1], fc);
    else
      hllc<NS>(A.T, pl, pr, ml, mr, B.jac[p0], B.jac[p1], fc);
  }
  __syncthreads();
  double* fcs = lds;
  if (active) {
#pragma unroll
    for (int l = 0; l < NV; ++l) fcs[l * (F * TW) + face_slot<DIR, TW, F>(col, f)] = fc[l];
  }
  __syncthreads();
  if (f >= 1 && active) {
    const long long q0 = pid0(B, idx[0], idx[1], idx[2]);
    const long long n0 = B.n0;
#pragma unroll
    for (int l = 0; l < NV; ++l)
      B.dq[q0 + l * n0] -= fcs[l * (F * TW) + face_slot<DIR, TW, F>(col, f)] -
                           fcs[l * (F * TW) + face_slot<DIR, TW, F>(col, f - 1)];
  }
}
template <int DIR, int NS>
__global__ __launch_bounds__(kT) void k_inviscid_pv(FrArgs A) {
  inviscid_pv_body<DIR, NS>(A, blockIdx.x, blockIdx.y, blockIdx.z);
}
template <int DIR, int NS>
__global__ __launch_bounds__(kT) void k_inviscid_pv_batched(FrBatch W) {
  const WgPos w = wg_pos(W);
  FrArgs A = W.common;
  A.B = W.blocks[w.b];
  inviscid_pv_body<DIR, NS>(A, w.x, w.y, w.z);
}

// One viscous sweep along DIR: compute_{fv,gv,hv}_2nd_order<FR, Laminar>,
// src/ViscousScheme.cuh:23-318, :320-616, :618-909; one thread per face,
// face fluxes through LDS, dq += fv(right) - fv(left).
template <int DIR, int NS>
__device__ __forceinline__ void viscous_body(const FrArgs& A, const int wx, const int wy, const int wz) {
  constexpr int NV = 5 + NS;
  constexpr int TW = Tile<DIR>::TW, F = Tile<DIR>::F, TS = F - 1;
  extern __shared__ double lds[];
  const FrBlockDev& B = A.B;
  const int ext[3] = {B.mx, B.my, B.mz};
  const int m = ext[DIR];
  const int across = DIR == 0 ? 1 : 0;
  const int col0 = wy * TW;
  const int third = wz;
  const int t0 = wx * TS;
  const long long s = B.sz;
  const int tid = threadIdx.x;
  // face f of column col; along x consecutive lanes take consecutive faces,
  // along y and z consecutive columns (consecutive i), so the global loads of
  // a wave are contiguous in both cases
  const int f = DIR == 0 ? tid % F : tid / TW, col = DIR == 0 ? tid / F : tid % TW;
  const int tl = t0 - 1 + f;
  int idx[3];
  tile_idx<DIR>(col0, tl, col, third, idx);
  const bool active = idx[across] < ext[across] && tl < m;
  double fv[NV];
  if (active) {
    const int i = idx[0], j = idx[1], k = idx[2];
    int o[3] = {0, 0, 0};
    o[DIR] = 1;
    const long long p = pid(B, i, j, k), p1 = pid(B, i + o[0], j + o[1], k + o[2]);
    auto M = [&](long long pt, int r, int c) { return B.met[pt + (3 * (r - 1) + (c - 1)) * s]; };
    const double xi_x = 0.5 * (M(p, 1, 1) + M(p1, 1, 1));
    const double xi_y = 0.5 * (M(p, 1, 2) + M(p1, 1, 2));
    const double xi_z = 0.5 * (M(p, 1, 3) + M(p1, 1, 3));
    const double eta_x = 0.5 * (M(p, 2, 1) + M(p1, 2, 1));
    const double eta_y = 0.5 * (M(p, 2, 2) + M(p1, 2, 2));
    const double eta_z = 0.5 * (M(p, 2, 3) + M(p1, 2, 3));
    const double zeta_x = 0.5 * (M(p, 3, 1) + M(p1, 3, 1));
    const double zeta_y = 0.5 * (M(p, 3, 2) + M(p1, 3, 2));
    const double zeta_z = 0.5 * (M(p, 3, 3) + M(p1, 3, 3));
    // normal derivative: 2 points; tangential: average of the two 2-point
    // central differences at the face's cells
    auto deriv = [&](const double* fp_, double* d) {
      auto Fv = [&](int ii, int jj, int kk) { return fp_[pid(B, ii, jj, kk)]; };
      if (DIR == 0) {
        d[0] = Fv(i + 1, j, k) - Fv(i, j, k);
        d[1] = 0.25 * (Fv(i, j + 1, k) - Fv(i, j - 1, k) + Fv(i + 1, j + 1, k) - Fv(i + 1, j - 1, k));
        d[2] = 0.25 * (Fv(i, j, k + 1) - Fv(i, j, k - 1) + Fv(i + 1, j, k + 1) - Fv(i + 1, j, k - 1));
      } else if (DIR == 1) {
        d[0] = 0.25 * (Fv(i + 1, j, k) - Fv(i - 1, j, k) + Fv(i + 1, j + 1, k) - Fv(i - 1, j + 1, k));
        d[1] = Fv(i, j + 1, k) - Fv(i, j, k);
        d[2] = 0.25 * (Fv(i, j, k + 1) - Fv(i, j, k - 1) + Fv(i, j + 1, k + 1) - Fv(i, j + 1, k - 1));
      } else {
        d[0] = 0.25 * (Fv(i + 1, j, k) - Fv(i - 1, j, k) + Fv(i + 1, j, k + 1) - Fv(i - 1, j, k + 1));
        d[1] = 0.25 * (Fv(i, j + 1, k) - Fv(i, j - 1, k) + Fv(i, j + 1, k + 1) - Fv(i, j - 1, k + 1));
        d[2] = Fv(i, j, k + 1) - Fv(i, j, k);
      }
    };
    double du[3], dv[3], dw[3], dT[3];
    deriv(B.bv + 1 * s, du);
    deriv(B.bv + 2 * s, dv);
    deriv(B.bv + 3 * s, dw);
    deriv(B.bv + 5 * s, dT);
    const double u_x = du[0] * xi_x + du[1] * eta_x + du[2] * zeta_x;
    const double u_y = du[0] * xi_y + du[1] * eta_y + du[2] * zeta_y;
    const double u_z = du[0] * xi_z + du[1] * eta_z + du[2] * zeta_z;
    const double v_x = dv[0] * xi_x + dv[1] * eta_x + dv[2] * zeta_x;
    const double v_y = dv[0] * xi_y + dv[1] * eta_y + dv[2] * zeta_y;
    const double v_z = dv[0] * xi_z + dv[1] * eta_z + dv[2] * zeta_z;
    const double w_x = dw[0] * xi_x + dw[1] * eta_x + dw[2] * zeta_x;
    const double w_y = dw[0] * xi_y + dw[1] * eta_y + dw[2] * zeta_y;
    const double w_z = dw[0] * xi_z + dw[1] * eta_z + dw[2] * zeta_z;
    const double mul = 0.5 * (B.mul[p] + B.mul[p1]);
    const double viscosity = mul + 0.0;
    const double tau_xx = viscosity * (4 * u_x - 2 * v_y - 2 * w_z) / 3.0;
    const double tau_yy = viscosity * (4 * v_y - 2 * u_x - 2 * w_z) / 3.0;
    const double tau_zz = viscosity * (4 * w_z - 2 * u_x - 2 * v_y) / 3.0;
    const double tau_xy = viscosity * (u_y + v_x);
    const double tau_xz = viscosity * (u_z + w_x);
    const double tau_yz = viscosity * (v_z + w_y);
    constexpr int r = DIR + 1;
    const double kx = 0.5 * (M(p, r, 1) * B.jac[p] + M(p1, r, 1) * B.jac[p1]);
    const double ky = 0.5 * (M(p, r, 2) * B.jac[p] + M(p1, r, 2) * B.jac[p1]);
    const double kz = 0.5 * (M(p, r, 3) * B.jac[p] + M(p1, r, 3) * B.jac[p1]);
    fv[0] = 0;
    fv[1] = kx * tau_xx + ky * tau_xy + kz * tau_xz;
    fv[2] = kx * tau_xy + ky * tau_yy + kz * tau_yz;
    fv[3] = kx * tau_xz + ky * tau_yz + kz * tau_zz;
    const double um = 0.5 * (B.bv[p + s] + B.bv[p1 + s]);
    const double vm = 0.5 * (B.bv[p + 2 * s] + B.bv[p1 + 2 * s]);
    const double wm = 0.5 * (B.bv[p + 3 * s] + B.bv[p1 + 3 * s]);
    const double t_x = dT[0] * xi_x + dT[1] * eta_x + dT[2] * zeta_x;
    const double t_y = dT[0] * xi_y + dT[1] * eta_y + dT[2] * zeta_y;
    const double t_z = dT[0] * xi_z + dT[1] * eta_z + dT[2] * zeta_z;
    const double conductivity = 0.5 * (B.lam[p] + B.lam[p1]);
    fv[4] = um * fv[1] + vm * fv[2] + wm * fv[3] + conductivity * (kx * t_x + ky * t_y + kz * t_z);
    double diffusivity[NS], yk[NS], ddf[NS];
    double sum_grad_over_w{0}, sum_rhoDkYk{0}, corr{0}, mw_tot{0};
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      yk[l] = 0.5 * (B.sv[p + l * s] + B.sv[p1 + l * s]);
      diffusivity[l] = 0.5 * (B.rhod[p + l * s] + B.rhod[p1 + l * s]) + 0.0;
      double dy[3];
      deriv(B.sv + l * s, dy);
      const double y_x = dy[0] * xi_x + dy[1] * eta_x + dy[2] * zeta_x;
      const double y_y = dy[0] * xi_y + dy[1] * eta_y + dy[2] * zeta_y;
      const double y_z = dy[0] * xi_z + dy[1] * eta_z + dy[2] * zeta_z;
      const double gy = kx * y_x + ky * y_y + kz * y_z;
      ddf[l] = gy;
      corr += diffusivity[l] * gy;
      sum_grad_over_w += gy / A.T->mw[l];
      mw_tot += yk[l] / A.T->mw[l];
      sum_rhoDkYk += diffusivity[l] * yk[l];
    }
    mw_tot = 1.0 / mw_tot;
    corr -= mw_tot * sum_rhoDkYk * sum_grad_over_w;
    if (A.grad_p) {
      double dp[3];
      deriv(B.bv + 4 * s, dp);
      const double p_x{dp[0] * xi_x + dp[1] * eta_x + dp[2] * zeta_x};
      const double p_y{dp[0] * xi_y + dp[1] * eta_y + dp[2] * zeta_y};
      const double p_z{dp[0] * xi_z + dp[1] * eta_z + dp[2] * zeta_z};
      const double gp = (kx * p_x + ky * p_y + kz * p_z) / (0.5 * (B.bv[p1 + 4 * s] + B.bv[p + 4 * s]));
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        ddf[l] += (mw_tot / A.T->mw[l] - 1) * yk[l] * gp;
        corr += (mw_tot / A.T->mw[l] - 1) * yk[l] * gp * diffusivity[l];
      }
    }
    double h[NS];
    const double tm = 0.5 * (B.bv[p + 5 * s] + B.bv[p1 + 5 * s]);
    enthalpy<NS>(A.T, tm, h);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const double flux{diffusivity[l] * (ddf[l] - mw_tot * yk[l] * sum_grad_over_w) - yk[l] * corr};
      fv[5 + l] = flux;
      fv[4] += h[l] * flux;
    }
#pragma unroll
    for (int l = 0; l < NV; ++l) lds[l * (F * TW) + face_slot<DIR, TW, F>(col, f)] = fv[l];
  }
  __syncthreads();
  if (f >= 1 && active) {
    const long long q0 = pid0(B, idx[0], idx[1], idx[2]);
    const long long n0 = B.n0;
#pragma unroll
    for (int l = 0; l < NV; ++l)
      B.dq[q0 + l * n0] += lds[l * (F * TW) + face_slot<DIR, TW, F>(col, f)] -
                           lds[l * (F * TW) + face_slot<DIR, TW, F>(col, f - 1)];
  }
}
// At most 256 VGPRs, two waves per SIMD: the y and z sweeps need 258-260
// unconstrained, which leaves one wave per SIMD; capped they spill 2-4
// registers and take 12.3 instead of 17.0 ms per step on 512x256x128.
template <int DIR, int NS>
__global__ __launch_bounds__(kT, 2) void k_viscous(FrArgs A) {
  viscous_body<DIR, NS>(A, blockIdx.x, blockIdx.y, blockIdx.z);
}
template <int DIR, int NS>
__global__ __launch_bounds__(kT, 2) void k_viscous_batched(FrBatch W) {
  const WgPos w = wg_pos(W);
  FrArgs A = W.common;
  A.B = W.blocks[w.b];
  viscous_body<DIR, NS>(A, w.x, w.y, w.z);
}

unsigned nblk(long long n) { return (unsigned)((n + kT - 1) / kT); }

// sponge_layer_source (src/SpongeLayer.cu:68-160), after the chemistry source
// of the same stage: dq -= sigma (cv - mean) on every conservative variable,
// sigma times the Jacobian when sp_jac is set (as the chemistry's J omega).
template <int NS>
__global__ __launch_bounds__(kT) void k_sponge_source(FrBlockDev B) {
  const long long t = (long long)blockIdx.x * kT + threadIdx.x;
  if (t >= B.n0) return;
  const int i = (int)(t % B.mx), j = (int)((t / B.mx) % B.my), k = (int)(t / ((long long)B.mx * B.my));
  const long long p = pid(B, i, j, k);
  const double sigma = B.sp_jac ? B.sp_sigma[t] * B.jac[p] : B.sp_sigma[t];
#pragma unroll
  for (int l = 0; l < 5 + NS; ++l) B.dq[t + l * B.n0] -= sigma * (B.cv[p + l * B.sz] - B.sp_mean[t + l * B.n0]);
}

// Grids and LDS bytes of the sweeps: a component-wise or viscous tile is
// TW columns x (F - 1) cells of one pencil (Tile<D>), a characteristic tile
// TileCh<D>; the third index of the grid runs over the remaining axis.
template <int D, class T>
dim3 sweep_grid(const FrBlockDev& B) {
  constexpr int TW = T::TW, TS = T::F - 1;
  const int ext[3] = {B.mx, B.my, B.mz};
  const int across = D == 0 ? 1 : 0, third = D == 2 ? 1 : 2;
  return dim3((ext[D] + TS - 1) / TS, (ext[across] + TW - 1) / TW, ext[third]);
}
dim3 point_grid(long long n) { return dim3(nblk(n)); }

template <int D, int NS>
size_t inviscid_lds(int scheme) {
  if (scheme == 52 || scheme == 72) {
    constexpr int TWc = TileCh<D>::TW, Fc = TileCh<D>::F, TSc = Fc - 1;
    const int Pc = TSc + 2 * (scheme == 72 ? 4 : 3) + (scheme == 52 ? 1 : 0);
    return ((size_t)2 * (5 + NS) * Pc * TWc + (size_t)(5 + NS) * Fc * TWc) * sizeof(double);
  }
  constexpr int TW = Tile<D>::TW, F = Tile<D>::F, TS = F - 1;
  if (scheme == 51 || scheme == 71) return (size_t)2 * (5 + NS) * (TS + 2 * (scheme == 71 ? 4 : 3)) * TW * sizeof(double);
  // primitive reconstruction (2, 3, 4): the fields, then one entropy-fix delta per point
  return ((size_t)(5 + NS) * std::max(TS + 4, F) * TW + (size_t)(TS + 4) * TW) * sizeof(double);
}

// The FrBatch of one kernel kind over the nb (<= kFrBatchMax) blocks of h.
template <class GridOf>
FrBatch make_batch(const FrArgs* h, int nb, GridOf grid_of, unsigned* total) {
  FrBatch W{};
  W.common = h[0];
  W.nb = nb;
  long long acc = 0;
  for (int b = 0; b < nb; ++b) {
    const dim3 g = grid_of(h[b].B);
    W.blocks[b] = h[b].B;
    W.first[b] = (int)acc;
    W.gx[b] = (int)g.x;
    W.gy[b] = (int)g.y;
    acc += (long long)g.x * g.y * g.z;
  }
  W.first[nb] = (int)acc;
  *total = (unsigned)acc;
  return W;
}

// The stage residual of the nb blocks of h: one launch per kernel kind and
// block (batched false, the reference's loop) or per kernel kind and
// kFrBatchMax blocks (batched true).
template <int NS>
void launch_rhs_t(const FrArgs* h, int nb, bool batched, hipStream_t st, long long* n, KernelTimer* tm) {
  const FrArgs& A = h[0];  // scheme switches are the same for every block
  const int step = batched ? kFrBatchMax : 1;
  auto rec = [&](int which, bool end) {
    if (tm) tm->mark(which, end, st);
  };
  unsigned total = 0;
  for (int c0 = 0; c0 < nb; c0 += step) {
    const int nc = std::min(step, nb - c0);
    const FrArgs* hc = h + c0;
    rec(KT_SOURCE, false);
    if (batched) {
      const FrBatch W = make_batch(hc, nc, [](const FrBlockDev& B) { return point_grid(B.n0); }, &total);
      hipLaunchKernelGGL(k_source_batched<NS>, dim3(total), dim3(kT), 0, st, W);
    } else {
      hipLaunchKernelGGL(k_source<NS>, point_grid(hc[0].B.n0), dim3(kT), 0, st, hc[0]);
    }
    ++*n;
    for (int b = 0; b < nc; ++b)
      if (hc[b].B.sp_sigma) {
        hipLaunchKernelGGL(k_sponge_source<NS>, point_grid(hc[b].B.n0), dim3(kT), 0, st, hc[b].B);
        ++*n;
      }
    rec(KT_SOURCE, true);
    auto inv = [&](auto dirc) {
      constexpr int D = decltype(dirc)::value;
      const size_t lds = inviscid_lds<D, NS>(A.scheme);
      rec(KT_INV_X + D, false);
      const bool ch = A.scheme == 52 || A.scheme == 72;
      FrBatch W{};
      if (batched) {
        W = ch ? make_batch(hc, nc, sweep_grid<D, TileCh<D>>, &total)
               : make_batch(hc, nc, sweep_grid<D, Tile<D>>, &total);
      }
      const FrArgs& Ab = hc[0];
      const dim3 grid = batched ? dim3(total) : (ch ? sweep_grid<D, TileCh<D>>(Ab.B) : sweep_grid<D, Tile<D>>(Ab.B));
      switch (A.scheme) {
        case 52:
          if (batched) hipLaunchKernelGGL((k_inviscid_ch_batched<D, NS, 52>), grid, dim3(kT), lds, st, W);
          else hipLaunchKernelGGL((k_inviscid_ch<D, NS, 52>), grid, dim3(kT), lds, st, Ab);
          break;
        case 72:
          if (batched) hipLaunchKernelGGL((k_inviscid_ch_batched<D, NS, 72>), grid, dim3(kT), lds, st, W);
          else hipLaunchKernelGGL((k_inviscid_ch<D, NS, 72>), grid, dim3(kT), lds, st, Ab);
          break;
        case 51:
          if (batched) hipLaunchKernelGGL((k_inviscid_batched<D, NS, 51>), grid, dim3(kT), lds, st, W);
          else hipLaunchKernelGGL((k_inviscid<D, NS, 51>), grid, dim3(kT), lds, st, Ab);
          break;
        case 71:
          if (batched) hipLaunchKernelGGL((k_inviscid_batched<D, NS, 71>), grid, dim3(kT), lds, st, W);
          else hipLaunchKernelGGL((k_inviscid<D, NS, 71>), grid, dim3(kT), lds, st, Ab);
          break;
        default:  // primitive reconstruction + Riemann solver (2, 3, 4)
          if (batched) hipLaunchKernelGGL((k_inviscid_pv_batched<D, NS>), grid, dim3(kT), lds, st, W);
          else hipLaunchKernelGGL((k_inviscid_pv<D, NS>), grid, dim3(kT), lds, st, Ab);
      }
      rec(KT_INV_X + D, true);
      ++*n;
    };
    inv(std::integral_constant<int, 0>{});
    inv(std::integral_constant<int, 1>{});
    inv(std::integral_constant<int, 2>{});
    if (A.viscous) {
      auto vis = [&](auto dirc) {
        constexpr int D = decltype(dirc)::value;
        const size_t lds = (size_t)(5 + NS) * Tile<D>::F * Tile<D>::TW * sizeof(double);
        rec(KT_VIS_X + D, false);
        if (batched) {
          const FrBatch W = make_batch(hc, nc, sweep_grid<D, Tile<D>>, &total);
          hipLaunchKernelGGL((k_viscous_batched<D, NS>), dim3(total), dim3(kT), lds, st, W);
        } else {
          const dim3 grid = sweep_grid<D, Tile<D>>(hc[0].B);
          hipLaunchKernelGGL((k_viscous<D, NS>), grid, dim3(kT), lds, st, hc[0]);
        }
        rec(KT_VIS_X + D, true);
        ++*n;
      };
      vis(std::integral_constant<int, 0>{});
      vis(std::integral_constant<int, 1>{});
      vis(std::integral_constant<int, 2>{});
    }
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers
bool supported_ns(int ns) { return ns == 9; }

#define OSTC_FR_DISPATCH(NSV, CALL) \
  switch (NSV) {                    \
    case 9: {                       \
      constexpr int NS = 9;         \
      CALL;                         \
    } break;                        \
    default: break;                 \
  }

void launch_rhs(const FrArgs& A, int ns, hipStream_t st, long long* n, KernelTimer* tm) {
  OSTC_FR_DISPATCH(ns, launch_rhs_t<NS>(&A, 1, false, st, n, tm))
}

void launch_rhs_batched(const FrArgs* h, int nb, int ns, hipStream_t st, long long* n, KernelTimer* tm) {
  OSTC_FR_DISPATCH(ns, launch_rhs_t<NS>(h, nb, true, st, n, tm))
}

void launch_update_batched(const FrArgs* h, int nb, int ns, double ra, double rb, double rc, hipStream_t st,
                           long long* n, KernelTimer* tm) {
  if (tm) tm->mark(KT_UPDATE, false, st);
  for (int c0 = 0; c0 < nb; c0 += kFrBatchMax) {
    unsigned total = 0;
    const FrBatch W = make_batch(h + c0, std::min(kFrBatchMax, nb - c0),
                                 [](const FrBlockDev& B) { return point_grid(B.n0); }, &total);
    OSTC_FR_DISPATCH(ns, hipLaunchKernelGGL(k_update_batched<NS>, dim3(total), dim3(kT), 0, st, W, ra, rb, rc))
    ++*n;
  }
  if (tm) tm->mark(KT_UPDATE, true, st);
}

void launch_props_batched(const FrArgs* h, int nb, int ns, hipStream_t st, long long* n, KernelTimer* tm) {
  if (tm) tm->mark(KT_PROPS, false, st);
  for (int c0 = 0; c0 < nb; c0 += kFrBatchMax) {
    unsigned total = 0;
    const FrBatch W = make_batch(h + c0, std::min(kFrBatchMax, nb - c0),
                                 [](const FrBlockDev& B) { return point_grid(B.sz); }, &total);
    OSTC_FR_DISPATCH(ns, hipLaunchKernelGGL(k_props_batched<NS>, dim3(total), dim3(kT), 0, st, W))
    ++*n;
  }
  if (tm) tm->mark(KT_PROPS, true, st);
}

void launch_local_dt_batched(const FrArgs* h, int nb, unsigned long long* dt_bits, hipStream_t st, long long* n) {
  for (int c0 = 0; c0 < nb; c0 += kFrBatchMax) {
    unsigned total = 0;
    const FrBatch W = make_batch(h + c0, std::min(kFrBatchMax, nb - c0),
                                 [](const FrBlockDev& B) { return point_grid(B.n0); }, &total);
    hipLaunchKernelGGL(k_local_dt_batched, dim3(total), dim3(kT), 0, st, W, dt_bits);
    ++*n;
  }
}

void launch_props(const FrArgs& A, int ns, hipStream_t st, long long* n, KernelTimer* tm) {
  if (tm) tm->mark(KT_PROPS, false, st);
  OSTC_FR_DISPATCH(ns, hipLaunchKernelGGL(k_props<NS>, dim3(nblk(A.B.sz)), dim3(kT), 0, st, A))
  if (tm) tm->mark(KT_PROPS, true, st);
  ++*n;
}

void launch_par_pack(const FrArgs& A, int ns, const ParFace& f, double* buf, hipStream_t st) {
  if (f.points <= 0) return;
  OSTC_FR_DISPATCH(ns, hipLaunchKernelGGL(k_par_pack<NS>, dim3(par_chunks(f)), dim3(kParThreads), 0, st, A, f, buf))
}

void launch_par_unpack(const FrArgs& A, int ns, const ParFace& f, const double* recv, const double* sent,
                       hipStream_t st) {
  if (f.points <= 0) return;
  OSTC_FR_DISPATCH(ns, hipLaunchKernelGGL(k_par_unpack<NS>, dim3(par_chunks(f)), dim3(kParThreads), 0, st, A, f, recv,
                                          sent))
}

void launch_cv_from_bv(const FrArgs& A, int ns, hipStream_t st, long long* n) {
  OSTC_FR_DISPATCH(ns, hipLaunchKernelGGL(k_cv_from_bv<NS>, dim3(nblk(A.B.sz)), dim3(kT), 0, st, A))
  ++*n;
}

void launch_update(const FrArgs& A, int ns, double ra, double rb, double rc, hipStream_t st, long long* n,
                   KernelTimer* tm) {
  if (tm) tm->mark(KT_UPDATE, false, st);
  OSTC_FR_DISPATCH(ns, hipLaunchKernelGGL(k_update<NS>, dim3(nblk(A.B.n0)), dim3(kT), 0, st, A, ra, rb, rc))
  if (tm) tm->mark(KT_UPDATE, true, st);
  ++*n;
}

void launch_local_dt(const FrArgs& A, unsigned long long* dt_bits, hipStream_t st, long long* n) {
  hipLaunchKernelGGL(k_local_dt, dim3(nblk(A.B.n0)), dim3(kT), 0, st, A, dt_bits);
  ++*n;
}

void launch_dt_set(double* dt, double v, hipStream_t st, long long* n) {
  hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(1), 0, st, dt, v);
  ++*n;
}

void launch_bc(const FrArgs& A, int ns, const FaceBox& fb, int type, const double* ff, const double* y,
               const unsigned long long* pass, hipStream_t st, long long* n) {
  const long long cnt = (long long)fb.n[0] * fb.n[1] * fb.n[2];
  OSTC_FR_DISPATCH(ns, hipLaunchKernelGGL(k_bc<NS>, dim3(nblk(cnt)), dim3(kT), 0, st, A, fb, type, ff, y, pass))
  ++*n;
}

// One more boundary-condition pass of a block: the counter that keys the
// inflow's white noise (a kernel, so that a captured graph repeats it).
__global__ void k_count_pass(unsigned long long* pass) {
  if (threadIdx.x == 0) *pass += 1;
}

void launch_count_pass(unsigned long long* pass, hipStream_t st, long long* n) {
  hipLaunchKernelGGL(k_count_pass, dim3(1), dim3(64), 0, st, pass);
  ++*n;
}

template <int NS>
__global__ __launch_bounds__(kT) void k_sponge_mean(FrBlockDev B, const unsigned long long* iter) {
  const long long t = (long long)blockIdx.x * kT + threadIdx.x;
  if (t >= B.n0) return;
  const int i = (int)(t % B.mx), j = (int)((t / B.mx) % B.my), k = (int)(t / ((long long)B.mx * B.my));
  const long long p = pid(B, i, j, k);
  const double n = (double)*iter;
#pragma unroll
  for (int l = 0; l < 5 + NS; ++l) {
    double& m = B.sp_mean[t + l * B.n0];
    m = (m * n + B.cv[p + l * B.sz]) / (n + 1);
  }
}

void launch_sponge_mean(const FrArgs& A, int ns, const unsigned long long* iter, hipStream_t st, long long* n) {
  OSTC_FR_DISPATCH(ns, hipLaunchKernelGGL(k_sponge_mean<NS>, dim3(nblk(A.B.n0)), dim3(kT), 0, st, A.B, iter))
  ++*n;
}

void launch_comm(const FrArgs& A, const FrBlockDev& Tg, int ns, const CommMap& m, hipStream_t st, long long* n) {
  const long long cnt = (long long)m.n[0] * m.n[1] * m.n[2];
  OSTC_FR_DISPATCH(ns, hipLaunchKernelGGL(k_comm<NS>, dim3(nblk(cnt)), dim3(kT), 0, st, A, Tg, m))
  ++*n;
}

// store_last_step and compute_square_of_dbv + reduction_of_dv_squared<4>
// (src/TimeAdvanceFunc.cu:6-37, src/Residual.cuh:12-69) of the interior.
__global__ __launch_bounds__(256) void k_store_last(FrBlockDev B) {
  const long long t = blockIdx.x * 256ll + threadIdx.x;
  if (t >= B.n0) return;
  const int i = (int)(t % B.mx), j = (int)((t / B.mx) % B.my), k = (int)(t / ((long long)B.mx * B.my));
  const long long p = pid(B, i, j, k), s = B.sz;
  const double u = B.bv[p + s], v = B.bv[p + 2 * s], w = B.bv[p + 3 * s];
  B.bvl[t] = B.bv[p];
  B.bvl[t + B.n0] = sqrt(u * u + v * v + w * w);
  B.bvl[t + 2 * B.n0] = B.bv[p + 4 * s];
  B.bvl[t + 3 * B.n0] = B.bv[p + 5 * s];
}

__global__ __launch_bounds__(256) void k_residual(FrBlockDev B, double* out4) {
  __shared__ double part[4][4];
  double acc[4] = {0, 0, 0, 0};
  const long long s = B.sz;
  for (long long t = blockIdx.x * 256ll + threadIdx.x; t < B.n0; t += 256ll * gridDim.x) {
    const int i = (int)(t % B.mx), j = (int)((t / B.mx) % B.my), k = (int)(t / ((long long)B.mx * B.my));
    const long long p = pid(B, i, j, k);
    const double u = B.bv[p + s], v = B.bv[p + 2 * s], w = B.bv[p + 3 * s];
    const double vel = sqrt(u * u + v * v + w * w);
    const double d0 = B.bv[p] - B.bvl[t], d1 = vel - B.bvl[t + B.n0];
    const double d2 = B.bv[p + 4 * s] - B.bvl[t + 2 * B.n0], d3 = B.bv[p + 5 * s] - B.bvl[t + 3 * B.n0];
    acc[0] += d0 * d0;
    acc[1] += d1 * d1;
    acc[2] += d2 * d2;
    acc[3] += d3 * d3;
  }
  for (int l = 0; l < 4; ++l) {
    double v = acc[l];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) part[l][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int l = threadIdx.x;
    out4[blockIdx.x * 4 + l] = part[l][0] + part[l][1] + part[l][2] + part[l][3];
  }
}

void launch_store_last(const FrBlockDev& B, hipStream_t st) {
  hipLaunchKernelGGL(k_store_last, dim3((unsigned)((B.n0 + 255) / 256)), dim3(256), 0, st, B);
}

void launch_residual(const FrBlockDev& B, double* out4, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(k_residual, dim3(nblk), dim3(256), 0, st, B, out4);
}

}  // namespace fr
}  // namespace ostc
