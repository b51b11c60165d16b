// ostc_kernels_dual.hip -- dual time stepping with the DPLUR inner solver on
// the per-block air path (dual_time_stepping, src/DualTimeStepping.cuh:33-236,
// src/DualTimeStepping.cu; DPLUR, src/DPLUR.cuh:14-327, src/DPLUR.cu).
//
// One inner iteration of one block: the stage residual dq of the RK path's
// kernels, then k_dual_pre (local pseudo time step and inviscid spectral
// radii, the BDF2 terms of the physical time derivative, dq0 = dt dq / diag),
// DPLUR sweeps (k_dplur, dqk from the neighbours' dq), and k_dual_update
// (dq / dt back, q += dt / J dq, primitives).  Every kernel is one thread per
// interior node with the reference's arithmetic, operation by operation.
#include "ostc_kernels.hpp"

namespace ostc {
namespace {

constexpr int kT = 256;

__device__ __forceinline__ bool node_of(const DevBlock& B, long long t, int& i, int& j, int& k) {
  if (t >= B.n0) return false;
  i = (int)(t % B.mx);
  j = (int)((t / B.mx) % B.my);
  k = (int)(t / ((long long)B.mx * B.my));
  return true;
}

// compute_qn_star (src/DualTimeStepping.cu:3-15): qn* = J/(2 dt) (4 q^n - q^(n-1))
__global__ __launch_bounds__(kT) void k_qn_star(DevBlock B, DualBlock D, double dt) {
  int i, j, k;
  const long long t = (long long)blockIdx.x * kT + threadIdx.x;
  if (!node_of(B, t, i, j, k)) return;
  const long long p = pidx(B, i, j, k), s = B.sz;
  const double factor = 0.5 * B.jac[p] / dt;
#pragma unroll
  for (int l = 0; l < 5; ++l) D.qn_star[t + l * B.n0] = factor * (4 * B.cv[0][p + l * s] - D.qn1[p + l * s]);
}

// local_time_step (src/TimeAdvanceFunc.cuh:28-83) keeping the inviscid
// spectral radii, compute_modified_rhs (src/DualTimeStepping.cu:17-31) and
// compute_DQ_0<Air> (src/DPLUR.cuh:14-45): dt_local, sr, dq += qn* - 3J/(2dt) q,
// dq = dt dq / diag, dq0 = dq.  The reference's viscous spectral radii are an
// array it allocates zeroed and never writes (src/Field.cu:553-566), so the
// diagonal's viscous term adds 2 dt (0 + 0 + 0).
__global__ __launch_bounds__(kT) void k_dual_pre(DevBlock B, Params P, DualBlock D, double dt, double diag_factor) {
  int i, j, k;
  const long long t = (long long)blockIdx.x * kT + threadIdx.x;
  if (!node_of(B, t, i, j, k)) return;
  const long long p = pidx(B, i, j, k), s = B.sz, n0 = B.n0;
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
  const double r0 = fabs(U) + a * gxi, r1 = fabs(V) + a * geta, r2 = fabs(Wc) + a * gzeta;
  const double sr_inv = r0 + r1 + r2;
  const double coeff_1 = (kGamma > 4.0 / 3.0 ? kGamma : 4.0 / 3.0) / bv[p];
  const double coeff_2 = B.mul[0][p] / P.prandtl;
  double sr_visc = gxi * gxi + geta * geta;
  sr_visc += gzeta * gzeta;
  sr_visc *= coeff_1 * coeff_2;
  const double dtl = P.cfl / (sr_inv + sr_visc);
  B.dtl[t] = dtl;
  D.sr[t] = r0;
  D.sr[t + n0] = r1;
  D.sr[t + 2 * n0] = r2;
  const double factor = 1.5 * B.jac[p] / dt;
  const double diag = 1 + diag_factor * dtl + dtl * (r0 + r1 + r2) + 2 * dtl * (0.0 + 0.0 + 0.0);
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    double d = B.dq[t + l * n0] + (D.qn_star[t + l * n0] - factor * B.cv[0][p + l * s]);
    d *= dtl;
    d /= diag;
    B.dq[t + l * n0] = d;
    D.dq0[t + l * n0] = d;
  }
}

// compute_jacobian_times_dq<Air> (src/DPLUR.cuh:98-160) at node (p, q) along
// direction d: (A_d + lambda) dq, A_d the inviscid flux Jacobian along the
// (not Jacobian-weighted) metric row d.
__device__ __forceinline__ void jac_times_dq(const DevBlock& B, const double* dq, long long p, long long q, int d,
                                             double lambda, double out[5]) {
  const long long s = B.sz, n0 = B.n0;
  const double kx = B.met[p + (3 * d) * s], ky = B.met[p + (3 * d + 1) * s], kz = B.met[p + (3 * d + 2) * s];
  const double* bv = B.bv[0];
  const double u = bv[p + s], v = bv[p + 2 * s], w = bv[p + 3 * s];
  const double U = kx * u + ky * v + kz * w;
  const double lmd1 = U + lambda;
  const double e = 0.5 * (u * u + v * v + w * w);
  constexpr double gamma = kGamma;
  const double b3 = 0, b4 = 0;
  const double h = gamma / (gamma - 1) * bv[p + 4 * s] / bv[p] + e;
  double x[5];
#pragma unroll
  for (int l = 0; l < 5; ++l) x[l] = dq[q + l * n0];
  const double b1 = kx * x[1] + ky * x[2] + kz * x[3] - U * x[0];
  const double b2 = (gamma - 1) * (e * x[0] - u * x[1] - v * x[2] - w * x[3] + x[4]);
  out[0] = b1 + lmd1 * x[0];
  out[1] = u * b1 + kx * b2 + lmd1 * x[1] + kx * (b3 - b4);
  out[2] = v * b1 + ky * b2 + lmd1 * x[2] + ky * (b3 - b4);
  out[3] = w * b1 + kz * b2 + lmd1 * x[3] + kz * (b3 - b4);
  out[4] = h * b1 + U * b2 + lmd1 * x[4] + U * (b3 - b4);
}

// DPLUR_inner_iteration<Air> (src/DPLUR.cuh:162-266): dqk = dq0 + dt/diag *
// sum of +-1/2 (A +- lambda) dq over the neighbours inside the block, in the
// order i-1, j-1, k-1, i+1, j+1, k+1.
__global__ __launch_bounds__(kT) void k_dplur(DevBlock B, DualBlock D, const double* dq, double* dqk,
                                              double diag_factor) {
  int i, j, k;
  const long long t = (long long)blockIdx.x * kT + threadIdx.x;
  if (!node_of(B, t, i, j, k)) return;
  const long long n0 = B.n0;
  const int ijk[3] = {i, j, k}, ext[3] = {B.mx, B.my, B.mz};
  double total[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int side = -1; side <= 1; side += 2) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (side < 0 ? ijk[d] == 0 : ijk[d] == ext[d] - 1) continue;
      int n[3] = {i, j, k};
      n[d] += side;
      const long long pn = pidx(B, n[0], n[1], n[2]), qn = pidx0(B, n[0], n[1], n[2]);
      const double visc = 0.0;  // the zero viscous spectral radius, as in k_dual_pre
      const double sr = D.sr[qn + d * n0];
      const double lambda = side < 0 ? sr + 2 * visc : -sr - 2 * visc;
      double out[5];
      jac_times_dq(B, dq, pn, qn, d, lambda, out);
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        if (side < 0) total[l] += 0.5 * out[l];
        else total[l] -= 0.5 * out[l];
      }
    }
  }
  const double dtl = B.dtl[t];
  const double diag = 1 + diag_factor * dtl + dtl * (D.sr[t] + D.sr[t + n0] + D.sr[t + 2 * n0]) +
                      2 * dtl * (0.0 + 0.0 + 0.0);
#pragma unroll
  for (int l = 0; l < 5; ++l) dqk[t + l * n0] = D.dq0[t + l * n0] + dtl * total[l] / diag;
}

// convert_dq_back_to_dqDt (src/DPLUR.cu:5-17) and update_cv_and_bv<Air>
// (src/FieldOperation.cuh:188-239): dq /= dt_local; q += dt_local / J dq;
// primitives from q.
__global__ __launch_bounds__(kT) void k_dual_update(DevBlock B) {
  int i, j, k;
  const long long t = (long long)blockIdx.x * kT + threadIdx.x;
  if (!node_of(B, t, i, j, k)) return;
  const long long p = pidx(B, i, j, k), s = B.sz, n0 = B.n0;
  double* cv = B.cv[0];
  double* bv = B.bv[0];
  const double dtl = B.dtl[t];
  const double dt_div_jac = dtl / B.jac[p];
  double c[5];
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    const double d = B.dq[t + l * n0] / dtl;
    B.dq[t + l * n0] = d;
    c[l] = cv[p + l * s] + dt_div_jac * d;
    cv[p + l * s] = c[l];
  }
  bv[p] = c[0];
  const double density_inv = 1.0 / c[0];
  const double u = c[1] * density_inv, v = c[2] * density_inv, w = c[3] * density_inv;
  bv[p + s] = u;
  bv[p + 2 * s] = v;
  bv[p + 3 * s] = w;
  const double V2 = u * u + v * v + w * w;
  const double pr = (kGamma - 1) * (c[4] - 0.5 * c[0] * V2);
  bv[p + 4 * s] = pr;
  bv[p + 5 * s] = pr * kMwAir * density_inv / kRu;
}

// store_last_iter (src/DualTimeStepping.cu:155-171): rho, |V|, p, T
__global__ __launch_bounds__(kT) void k_last_store(DevBlock B, DualBlock D) {
  int i, j, k;
  const long long t = (long long)blockIdx.x * kT + threadIdx.x;
  if (!node_of(B, t, i, j, k)) return;
  const long long p = pidx(B, i, j, k), s = B.sz, n0 = B.n0;
  const double* bv = B.bv[0];
  D.last[t] = bv[p];
  D.last[t + n0] = sqrt(bv[p + s] * bv[p + s] + bv[p + 2 * s] * bv[p + 2 * s] + bv[p + 3 * s] * bv[p + 3 * s]);
  D.last[t + 2 * n0] = bv[p + 4 * s];
  D.last[t + 3 * n0] = bv[p + 5 * s];
}

// compute_square_of_dbv_wrt_last_inner_iter (src/DualTimeStepping.cu:137-153)
// and the per-workgroup partial sums of its reduction (out4[4 w + l]).
__global__ __launch_bounds__(kT) void k_last_square(DevBlock B, DualBlock D, double* out4) {
  __shared__ double part[4][kT / 64];
  const long long s = B.sz, n0 = B.n0;
  const double* bv = B.bv[0];
  double acc[4] = {0, 0, 0, 0};
  for (long long t = (long long)blockIdx.x * kT + threadIdx.x; t < n0; t += (long long)kT * gridDim.x) {
    int i, j, k;
    node_of(B, t, i, j, k);
    const long long p = pidx(B, i, j, k);
    const double vel = sqrt(bv[p + s] * bv[p + s] + bv[p + 2 * s] * bv[p + 2 * s] + bv[p + 3 * s] * bv[p + 3 * s]);
    const double cur[4] = {bv[p], vel, bv[p + 4 * s], bv[p + 5 * s]};
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const double d = cur[l] - D.last[t + l * n0];
      D.last[t + l * n0] = d * d;
      acc[l] += d * d;
    }
  }
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    double v = acc[l];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) part[l][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double v = 0;
    for (int w = 0; w < kT / 64; ++w) v += part[threadIdx.x][w];
    out4[4 * blockIdx.x + threadIdx.x] = v;
  }
}

unsigned grid_n0(const DevBlock& B) { return (unsigned)((B.n0 + kT - 1) / kT); }

}  // namespace

void dual_qn_star(const DevBlock& B, const DualBlock& D, double dt, hipStream_t st, long long* n) {
  hipLaunchKernelGGL(k_qn_star, dim3(grid_n0(B)), dim3(kT), 0, st, B, D, dt);
  *n += 1;
}

void dual_pre(const DevBlock& B, const Params& P, const DualBlock& D, double dt, double diag_factor, hipStream_t st,
              long long* n) {
  hipLaunchKernelGGL(k_dual_pre, dim3(grid_n0(B)), dim3(kT), 0, st, B, P, D, dt, diag_factor);
  *n += 1;
}

void dual_dplur(const DevBlock& B, const DualBlock& D, double diag_factor, hipStream_t st, long long* n) {
  hipLaunchKernelGGL(k_dplur, dim3(grid_n0(B)), dim3(kT), 0, st, B, D, B.dq, D.dqk, diag_factor);
  *n += 1;
}

void dual_update(const DevBlock& B, hipStream_t st, long long* n) {
  hipLaunchKernelGGL(k_dual_update, dim3(grid_n0(B)), dim3(kT), 0, st, B);
  *n += 1;
}

void dual_last_store(const DevBlock& B, const DualBlock& D, hipStream_t st) {
  hipLaunchKernelGGL(k_last_store, dim3(grid_n0(B)), dim3(kT), 0, st, B, D);
}

void dual_last_square(const DevBlock& B, const DualBlock& D, double* out4, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(k_last_square, dim3(nblk), dim3(kT), 0, st, B, D, out4);
}

}  // namespace ostc
