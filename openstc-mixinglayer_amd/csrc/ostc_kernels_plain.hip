// ostc_kernels_plain.hip -- per-block kernels launched the way the reference
// launches them (src/RK.cuh:85-226): 64-thread pencils per direction for the
// inviscid and viscous fluxes with dq read-modify-write in HBM, an (8,8,4)
// interior grid for the update and local time step, a two-pass MIN with a
// host read-back for dt, and one launch per boundary or inner face.  This is
// the drop-in path; the fused batched kernel lives in ostc_kernels_fused.hip.
#include "ostc_kernels.hpp"

namespace ostc {

// compute_convective_term_weno_{x,y,z} (src/WENO.cu:8-461): one 64-thread
// block makes 64 face fluxes and updates 63 cells along DIR.  Launch grid
// (segments along DIR, other axis a, other axis b).
template <int DIR, int SCHEME>
__global__ __launch_bounds__(64) void k_weno_pencil(DevBlock B, Params P, int slot) {
  constexpr int R = Stencil<SCHEME>::R, W = Stencil<SCHEME>::W;
  constexpr int NP = 63 + 2 * R;
  __shared__ double sFp[5][NP], sFm[5][NP], sFc[5][64];
  const int m = DIR == 0 ? B.mx : (DIR == 1 ? B.my : B.mz);
  const int c0 = 63 * blockIdx.x;
  const int tid = threadIdx.x;
  int base[3];
  if (DIR == 0) { base[1] = blockIdx.y; base[2] = blockIdx.z; }
  if (DIR == 1) { base[0] = blockIdx.y; base[2] = blockIdx.z; }
  if (DIR == 2) { base[0] = blockIdx.y; base[1] = blockIdx.z; }
  const double* cv = B.cv[slot];
  const double* bv = B.bv[slot];
  const long long s = B.sz;
  for (int pt = tid; pt < NP; pt += 64) {
    const int pos = c0 - R + pt;
    if (pos > m - 1 + R) continue;
    int ijk[3] = {base[0], base[1], base[2]};
    ijk[DIR] = pos;
    const long long p = pidx(B, ijk[0], ijk[1], ijk[2]);
    double q[5];
    for (int l = 0; l < 5; ++l) q[l] = cv[p + l * s];
    const double c = sqrt(kGamma * kRAir * bv[p + 5 * s]);
    double fp[5], fm[5];
    lf_split5(q, bv[p + 4 * s], c, B.met[p + (3 * DIR) * s], B.met[p + (3 * DIR + 1) * s],
              B.met[p + (3 * DIR + 2) * s], B.jac[p], fp, fm);
    for (int l = 0; l < 5; ++l) { sFp[l][pt] = fp[l]; sFm[l][pt] = fm[l]; }
  }
  __syncthreads();
  const int L = c0 - 1 + tid;  // face L+1/2
  if (L <= m - 1) {
    int ijk[3] = {base[0], base[1], base[2]};
    ijk[DIR] = L;
    const long long p1 = pidx(B, ijk[0], ijk[1], ijk[2]);
    ijk[DIR] = L + 1;
    const long long p2 = pidx(B, ijk[0], ijk[1], ijk[2]);
    const double j1 = B.jac[p1], j2 = B.jac[p2];
    const int sL = tid + R - 1;  // LDS index of cell L
    if constexpr (SCHEME == 52 || SCHEME == 72) {
      // compute_weno_flux_ch<Air>, "method 1" (src/WENO.cu:1147-1453): the
      // W split fluxes of each side projected on the left eigenvectors of
      // the Roe-averaged face state, WENO-Z5 (52) or WENO-Z7 (72) per
      // characteristic field, and the face flux projected back with the right
      // eigenvectors.
      double cvl[6], cvr[6];
      for (int l = 0; l < 5; ++l) {
        cvl[l] = cv[p1 + l * s];
        cvr[l] = cv[p2 + l * s];
      }
      cvl[5] = bv[p1 + 4 * s];
      cvr[5] = bv[p2 + 4 * s];
      const double m1[3] = {B.met[p1 + 3 * DIR * s], B.met[p1 + (3 * DIR + 1) * s], B.met[p1 + (3 * DIR + 2) * s]};
      const double m2[3] = {B.met[p2 + 3 * DIR * s], B.met[p2 + (3 * DIR + 1) * s], B.met[p2 + (3 * DIR + 2) * s]};
      CharAir M;
      char_air(cvl, cvr, m1, m2, j1, j2, M);
      constexpr double eps{1e-40};
      const double ax = m1[0] * j1 + m2[0] * j2, ay = m1[1] * j1 + m2[1] * j2, az = m1[2] * j1 + m2[2] * j2;
      const double e = eps * P.weno_eps_scale * 0.25 * (ax * ax + ay * ay + az * az);
      double fChar[5];
      for (int l = 0; l < 5; ++l) {
        double vp[W], vm[W];
        for (int r = 0; r < W; ++r) {
          vp[r] = M.L[l][0] * sFp[0][sL - W / 2 + r];
          vm[r] = M.L[l][0] * sFm[0][sL - W / 2 + 1 + r];
          for (int n = 1; n < 5; ++n) {
            vp[r] += M.L[l][n] * sFp[n][sL - W / 2 + r];
            vm[r] += M.L[l][n] * sFm[n][sL - W / 2 + 1 + r];
          }
        }
        if constexpr (SCHEME == 72) fChar[l] = weno7z(vp, vm, e);
        else fChar[l] = weno5z(vp, vm, e);
      }
      for (int l = 0; l < 5; ++l) {
        double f = M.R[l][0] * fChar[0];
        for (int q = 1; q < 5; ++q) f += M.R[l][q] * fChar[q];
        sFc[l][tid] = f;
      }
    } else {
    double eps3[3];
    weno_eps(P, B.met[p1 + 3 * DIR * s] * j1, B.met[p1 + (3 * DIR + 1) * s] * j1, B.met[p1 + (3 * DIR + 2) * s] * j1,
             B.met[p2 + 3 * DIR * s] * j2, B.met[p2 + (3 * DIR + 1) * s] * j2, B.met[p2 + (3 * DIR + 2) * s] * j2,
             eps3);
    for (int l = 0; l < 5; ++l) {
      double vp[W], vm[W];
      for (int r = 0; r < W; ++r) {
        vp[r] = sFp[l][sL - (W / 2) + r];
        vm[r] = sFm[l][sL - (W / 2) + 1 + r];
      }
      const double e = l == 0 ? eps3[0] : (l == 4 ? eps3[2] : eps3[1]);
      sFc[l][tid] = weno_face<SCHEME>(vp, vm, e);
    }
    }
  }
  __syncthreads();
  if (tid > 0 && L <= m - 1) {
    int ijk[3] = {base[0], base[1], base[2]};
    ijk[DIR] = L;
    const long long p0 = pidx0(B, ijk[0], ijk[1], ijk[2]);
    for (int l = 0; l < 5; ++l) B.dq[p0 + l * B.n0] -= sFc[l][tid] - sFc[l][tid - 1];
  }
}

// compute_convective_term_ep_1D<Air> (src/InviscidScheme.cu:550-768), the
// sixth-order energy-preserving scheme (inviscid_scheme 6; Pirozzoli,
// J. Comput. Phys. 230 (2011) 2997-3014): rho, u, v, w, p, H = (E + p)/rho,
// the metric row, J and UJ = (k . u) J of 69 pencil points in LDS; the
// two-point fluxes T(i, l), l = 1..3, of 66 points; 64 face fluxes
// 2 (3/4 T(i,1) - 3/20 (T(i,2) + T(i-1,2)) + 1/60 (T(i,3) + T(i-1,3) +
// T(i-2,3))); 63 cells updated.
template <int DIR>
__global__ __launch_bounds__(64) void k_ep_pencil(DevBlock B, int slot) {
  constexpr int R = 3, NP = 63 + 2 * R, NT = 66;
  __shared__ double sP[11][NP], sT[3][5][NT], sFc[5][64];
  const int m = DIR == 0 ? B.mx : (DIR == 1 ? B.my : B.mz);
  const int c0 = 63 * blockIdx.x;
  const int tid = threadIdx.x;
  int base[3];
  if (DIR == 0) { base[1] = blockIdx.y; base[2] = blockIdx.z; }
  if (DIR == 1) { base[0] = blockIdx.y; base[2] = blockIdx.z; }
  if (DIR == 2) { base[0] = blockIdx.y; base[1] = blockIdx.z; }
  const double* cv = B.cv[slot];
  const double* bv = B.bv[slot];
  const long long s = B.sz;
  for (int pt = tid; pt < NP; pt += 64) {
    const int pos = c0 - R + pt;
    if (pos > m - 1 + R) continue;
    int ijk[3] = {base[0], base[1], base[2]};
    ijk[DIR] = pos;
    const long long p = pidx(B, ijk[0], ijk[1], ijk[2]);
#pragma unroll
    for (int l = 0; l < 5; ++l) sP[l][pt] = bv[p + l * s];
    sP[5][pt] = (cv[p + 4 * s] + bv[p + 4 * s]) / bv[p];
    const double kx = B.met[p + (3 * DIR) * s], ky = B.met[p + (3 * DIR + 1) * s], kz = B.met[p + (3 * DIR + 2) * s];
    sP[6][pt] = kx;
    sP[7][pt] = ky;
    sP[8][pt] = kz;
    const double jac = B.jac[p];
    sP[9][pt] = jac;
    double uk = kx * bv[p + s] + ky * bv[p + 2 * s] + kz * bv[p + 3 * s];
    uk *= jac;
    sP[10][pt] = uk;
  }
  __syncthreads();
  for (int t = tid; t < NT; t += 64) {
    const int pos = c0 - R + t;
#pragma unroll
    for (int l = 1; l <= 3; ++l) {
      if (pos + l > m - 1 + R) continue;
      const int a = t, b = t + l;
      const double weight = 0.125 * (sP[10][a] + sP[10][b]) * (sP[0][a] + sP[0][b]);
      const double pWeight = 0.25 * (sP[4][a] + sP[4][b]);
      sT[l - 1][0][t] = 2 * weight;
      sT[l - 1][1][t] = weight * (sP[1][a] + sP[1][b]) + pWeight * (sP[6][a] * sP[9][a] + sP[6][b] * sP[9][b]);
      sT[l - 1][2][t] = weight * (sP[2][a] + sP[2][b]) + pWeight * (sP[7][a] * sP[9][a] + sP[7][b] * sP[9][b]);
      sT[l - 1][3][t] = weight * (sP[3][a] + sP[3][b]) + pWeight * (sP[8][a] * sP[9][a] + sP[8][b] * sP[9][b]);
      sT[l - 1][4][t] = weight * (sP[5][a] + sP[5][b]);
    }
  }
  __syncthreads();
  const int L = c0 - 1 + tid;  // face L+1/2
  if (L <= m - 1) {
    constexpr double central_1[3] = {0.75, -3.0 / 20, 1.0 / 60};
    const int t = tid + R - 1;  // point L
#pragma unroll
    for (int l = 0; l < 5; ++l)
      sFc[l][tid] = 2 * (central_1[0] * sT[0][l][t] + central_1[1] * (sT[1][l][t] + sT[1][l][t - 1]) +
                         central_1[2] * (sT[2][l][t] + sT[2][l][t - 1] + sT[2][l][t - 2]));
  }
  __syncthreads();
  if (tid > 0 && L <= m - 1) {
    int ijk[3] = {base[0], base[1], base[2]};
    ijk[DIR] = L;
    const long long p0 = pidx0(B, ijk[0], ijk[1], ijk[2]);
    for (int l = 0; l < 5; ++l) B.dq[p0 + l * B.n0] -= sFc[l][tid] - sFc[l][tid - 1];
  }
}

// compute_convective_term_pv_1D<Air> (src/InviscidScheme.cu:50-165): the
// primitive variables of 67 pencil points staged in LDS, MUSCL / NND2 /
// first-order reconstruction (src/Reconstruction.cu:5-46) and an LF, AUSM+ or
// HLLC flux (src/RiemannSolver.cuh) at 64 faces, 63 cells updated.
template <int DIR>
__global__ __launch_bounds__(64) void k_pv_pencil(DevBlock B, Params P, int slot) {
  constexpr int NP = 67;  // points c0-2 .. c0+64
  __shared__ double sPv[5][NP], sK[3][NP], sJ[NP], sFc[5][64], sDelta[NP];
  const int m = DIR == 0 ? B.mx : (DIR == 1 ? B.my : B.mz);
  const int c0 = 63 * blockIdx.x;
  const int tid = threadIdx.x;
  int base[3];
  if (DIR == 0) { base[1] = blockIdx.y; base[2] = blockIdx.z; }
  if (DIR == 1) { base[0] = blockIdx.y; base[2] = blockIdx.z; }
  if (DIR == 2) { base[0] = blockIdx.y; base[1] = blockIdx.z; }
  const double* bv = B.bv[slot];
  const long long s = B.sz;
  for (int pt = tid; pt < NP; pt += 64) {
    const int pos = c0 - 2 + pt;
    if (pos > m + 1) continue;  // the faces need points -2 .. m+1 (ngg >= 2)
    int ijk[3] = {base[0], base[1], base[2]};
    ijk[DIR] = pos;
    const long long p = pidx(B, ijk[0], ijk[1], ijk[2]);
    for (int l = 0; l < 5; ++l) sPv[l][pt] = bv[p + l * s];
    for (int c = 0; c < 3; ++c) sK[c][pt] = B.met[p + (3 * DIR + c) * s];
    sJ[pt] = B.jac[p];
    if (P.scheme == 2) {  // the point's entropy-fix delta from its full metric, once for both of its faces
      double mm[9], qq[5];
      for (int cc = 0; cc < 9; ++cc) mm[cc] = B.met[p + cc * s];
      for (int l = 0; l < 5; ++l) qq[l] = bv[p + l * s];
      sDelta[pt] = entropy_delta_air(qq, mm, P.entropy_fix);
    }
  }
  __syncthreads();
  const int L = c0 - 1 + tid;  // face L+1/2 between LDS points tid+1 and tid+2
  if (L <= m - 1) {
    double pl[6], pr[6];
    for (int l = 0; l < 5; ++l)
      reconstruct_var(P.recon, sPv[l][tid], sPv[l][tid + 1], sPv[l][tid + 2], sPv[l][tid + 3], &pl[l], &pr[l]);
    // total energy of the reconstructed states, reconstruction<Air> (src/InviscidScheme.cu:210-215)
    const double el = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
    const double er = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
    pl[5] = el * pl[0] + pl[4] / (kGamma - 1);
    pr[5] = er * pr[0] + pr[4] / (kGamma - 1);
    const double ml[3] = {sK[0][tid + 1], sK[1][tid + 1], sK[2][tid + 1]};
    const double mr[3] = {sK[0][tid + 2], sK[1][tid + 2], sK[2][tid + 2]};
    double f[5];
    if (P.scheme == 2) {
      roe_air(pl, pr, ml, mr, sJ[tid + 1], sJ[tid + 2], sDelta[tid + 1], sDelta[tid + 2], f);
    } else {
      riemann_air(P.scheme, pl, pr, ml, mr, sJ[tid + 1], sJ[tid + 2], f);
    }
    for (int l = 0; l < 5; ++l) sFc[l][tid] = f[l];
  }
  __syncthreads();
  if (tid > 0 && L <= m - 1) {
    int ijk[3] = {base[0], base[1], base[2]};
    ijk[DIR] = L;
    const long long p0 = pidx0(B, ijk[0], ijk[1], ijk[2]);
    for (int l = 0; l < 5; ++l) B.dq[p0 + l * B.n0] -= sFc[l][tid] - sFc[l][tid - 1];
  }
}

// Gradients in computational space at face (i,j,k)+1/2 along DIR from a bv
// accessor, src/ViscousScheme.cuh:41-56 (fv), :338-353 (gv), :636-651 (hv).
template <int DIR, class PV>
__device__ __forceinline__ void face_gradients(PV pv, int i, int j, int k, double d[4][3]) {
  const int var[4] = {1, 2, 3, 5};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int l = var[q];
    if (DIR == 0) {
      d[q][0] = pv(i + 1, j, k, l) - pv(i, j, k, l);
      d[q][1] = 0.25 * (pv(i, j + 1, k, l) - pv(i, j - 1, k, l) + pv(i + 1, j + 1, k, l) - pv(i + 1, j - 1, k, l));
      d[q][2] = 0.25 * (pv(i, j, k + 1, l) - pv(i, j, k - 1, l) + pv(i + 1, j, k + 1, l) - pv(i + 1, j, k - 1, l));
    } else if (DIR == 1) {
      d[q][0] = 0.25 * (pv(i + 1, j, k, l) - pv(i - 1, j, k, l) + pv(i + 1, j + 1, k, l) - pv(i - 1, j + 1, k, l));
      d[q][1] = pv(i, j + 1, k, l) - pv(i, j, k, l);
      d[q][2] = 0.25 * (pv(i, j, k + 1, l) - pv(i, j, k - 1, l) + pv(i, j + 1, k + 1, l) - pv(i, j + 1, k - 1, l));
    } else {
      d[q][0] = 0.25 * (pv(i + 1, j, k, l) - pv(i - 1, j, k, l) + pv(i + 1, j, k + 1, l) - pv(i - 1, j, k + 1, l));
      d[q][1] = 0.25 * (pv(i, j + 1, k, l) - pv(i, j - 1, k, l) + pv(i, j + 1, k + 1, l) - pv(i, j - 1, k + 1, l));
      d[q][2] = pv(i, j, k + 1, l) - pv(i, j, k, l);
    }
  }
}

// viscous_flux_{fv,gv,hv} (src/SchemeSelector.cuh:69-154).
template <int DIR>
__global__ __launch_bounds__(64) void k_visc_pencil(DevBlock B, Params P, int slot) {
  __shared__ double sFv[4][64];
  const int m = DIR == 0 ? B.mx : (DIR == 1 ? B.my : B.mz);
  const int tid = threadIdx.x;
  const int L = 63 * blockIdx.x - 1 + tid;
  int ijk[3];
  if (DIR == 0) { ijk[1] = blockIdx.y; ijk[2] = blockIdx.z; }
  if (DIR == 1) { ijk[0] = blockIdx.y; ijk[2] = blockIdx.z; }
  if (DIR == 2) { ijk[0] = blockIdx.y; ijk[1] = blockIdx.z; }
  ijk[DIR] = L;
  const double* bv = B.bv[slot];
  const double* mu = B.mul[slot];
  const long long s = B.sz;
  if (L <= m - 1) {
    auto pv = [&](int a, int b, int c, int l) { return bv[pidx(B, a, b, c) + l * s]; };
    double d[4][3];
    face_gradients<DIR>(pv, ijk[0], ijk[1], ijk[2], d);
    int o[3] = {0, 0, 0};
    o[DIR] = 1;
    const long long pA = pidx(B, ijk[0], ijk[1], ijk[2]);
    const long long pB = pidx(B, ijk[0] + o[0], ijk[1] + o[1], ijk[2] + o[2]);
    double mA[9], mB[9];
    for (int c = 0; c < 9; ++c) { mA[c] = B.met[pA + c * s]; mB[c] = B.met[pB + c * s]; }
    const double um = 0.5 * (bv[pA + s] + bv[pB + s]);
    const double vm = 0.5 * (bv[pA + 2 * s] + bv[pB + 2 * s]);
    const double wm = 0.5 * (bv[pA + 3 * s] + bv[pB + 3 * s]);
    double fv[4];
    viscous_face_flux<DIR>(d, mA, mB, B.jac[pA], B.jac[pB], mu[pA], mu[pB], um, vm, wm, P.prandtl, fv);
    for (int l = 0; l < 4; ++l) sFv[l][tid] = fv[l];
  }
  __syncthreads();
  if (tid > 0 && L <= m - 1) {
    const long long p0 = pidx0(B, ijk[0], ijk[1], ijk[2]);
    for (int l = 0; l < 4; ++l) B.dq[p0 + (l + 1) * B.n0] += sFv[l][tid] - sFv[l][tid - 1];
  }
}

// update_cv_and_bv_rk<Air> (src/RK.cuh:230-284), interior cells, tpb (8,8,4).
__global__ __launch_bounds__(256) void k_update(DevBlock B, double dt, double ra, double rb, double rc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z * blockDim.z + threadIdx.z;
  if (i >= B.mx || j >= B.my || k >= B.mz) return;
  double* cv = B.cv[0];
  double* bv = B.bv[0];
  const double* qn = B.cv[1];
  const long long p = pidx(B, i, j, k), p0 = pidx0(B, i, j, k), s = B.sz;
  const double dt_div_jac = dt / B.jac[p];
  double c[5];
  for (int l = 0; l < 5; ++l) {
    c[l] = ra * qn[p + l * s] + rb * cv[p + l * s] + rc * dt_div_jac * B.dq[p0 + l * B.n0];
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

// local_time_step<Air, Laminar>, src/TimeAdvanceFunc.cuh:28-83.
__device__ __forceinline__ double local_dt_point(const DevBlock& B, const Params& P, int slot, long long p) {
  const long long s = B.sz;
  const double* m = B.met;
  const double gxi = sqrt(m[p] * m[p] + m[p + s] * m[p + s] + m[p + 2 * s] * m[p + 2 * s]);
  const double geta = sqrt(m[p + 3 * s] * m[p + 3 * s] + m[p + 4 * s] * m[p + 4 * s] + m[p + 5 * s] * m[p + 5 * s]);
  const double gzeta = sqrt(m[p + 6 * s] * m[p + 6 * s] + m[p + 7 * s] * m[p + 7 * s] + m[p + 8 * s] * m[p + 8 * s]);
  const double* bv = B.bv[slot];
  const double u = bv[p + s], v = bv[p + 2 * s], w = bv[p + 3 * s];
  const double U = u * m[p] + v * m[p + s] + w * m[p + 2 * s];
  const double V = u * m[p + 3 * s] + v * m[p + 4 * s] + w * m[p + 5 * s];
  const double Wc = u * m[p + 6 * s] + v * m[p + 7 * s] + w * m[p + 8 * s];
  const double a = sqrt(kGamma * kRu * bv[p + 5 * s] / kMwAir);
  const double sr_inv = (fabs(U) + a * gxi) + (fabs(V) + a * geta) + (fabs(Wc) + a * gzeta);
  const double coeff_1 = (kGamma > 4.0 / 3.0 ? kGamma : 4.0 / 3.0) / bv[p];
  const double coeff_2 = B.mul[slot][p] / P.prandtl;
  double sr_visc = gxi * gxi + geta * geta;
  sr_visc += gzeta * gzeta;
  sr_visc *= coeff_1 * coeff_2;
  return P.cfl / (sr_inv + sr_visc);
}

__global__ __launch_bounds__(256) void k_local_dt(DevBlock B, Params P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z * blockDim.z + threadIdx.z;
  if (i >= B.mx || j >= B.my || k >= B.mz) return;
  B.dtl[pidx0(B, i, j, k)] = local_dt_point(B, P, 0, pidx(B, i, j, k));
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// min_of_arr (src/TimeAdvanceFunc.cu:70-89): grid-stride MIN starting from
// 1e6, reduced across the 64-lane wavefront with shuffles, then across waves.
__global__ __launch_bounds__(256) void k_min_pass(const double* in, long long n, double* out) {
  __shared__ double part[4];
  double v = 1e6;
  for (long long t = blockIdx.x * 256ll + threadIdx.x; t < n; t += 256ll * gridDim.x) v = fmin(v, in[t]);
  v = wave_min(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = fmin(fmin(part[0], part[1]), fmin(part[2], part[3]));
}

// Padded-domain kernels, tpb (8,8,4): compute_cv_from_bv and
// update_physical_properties (src/FieldOperation.cuh:40-76, :107-140).
__global__ __launch_bounds__(256) void k_padded(DevBlock B, int slot, int what) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x - B.g;
  const int j = blockIdx.y * blockDim.y + threadIdx.y - B.g;
  const int k = blockIdx.z * blockDim.z + threadIdx.z - B.g;
  if (i >= B.mx + B.g || j >= B.my + B.g || k >= B.mz + B.g) return;
  DevOp op{};
  op.type = what;
  op.block = 0;
  op.lo[0] = i; op.lo[1] = j; op.lo[2] = k;
  op.n[0] = op.n[1] = op.n[2] = 1;
  op_point(&B, op, 0, slot, slot);
}

// One launch of a range of independent ops.  Thread -> op by binary search
// over the ops' running point offsets.
// The launch's op offsets and the block descriptors are first staged in LDS
// (when they fit), so a thread finds its op and its arrays without a chain of
// dependent global loads.
constexpr int kOpsInLds = 256;
constexpr int kBlocksInLds = 64;

__global__ __launch_bounds__(256) void k_ops(const DevBlock* blocks, int n_blocks, const DevOp* ops, int op_lo,
                                             int op_hi, long long total, int slot, int src, int props) {
  __shared__ long long s_begin[kOpsInLds];
  __shared__ DevBlock s_blocks[kBlocksInLds];
  const int n_ops = op_hi - op_lo;
  const bool ops_lds = n_ops <= kOpsInLds, blk_lds = n_blocks <= kBlocksInLds;
  const long long base = ops[op_lo].begin;
  if (ops_lds)
    for (int t = threadIdx.x; t < n_ops; t += 256) s_begin[t] = ops[op_lo + t].begin - base;
  if (blk_lds) {
    static_assert(sizeof(DevBlock) % 8 == 0, "DevBlock is copied as 8-byte words");
    const int words = n_blocks * (int)(sizeof(DevBlock) / 8);
    const long long* src_w = reinterpret_cast<const long long*>(blocks);
    long long* dst_w = reinterpret_cast<long long*>(s_blocks);
    for (int w = threadIdx.x; w < words; w += 256) dst_w[w] = src_w[w];
  }
  __syncthreads();
  const long long gid = blockIdx.x * 256ll + threadIdx.x;
  if (gid >= total) return;
  int lo = 0, hi = n_ops - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const long long b = ops_lds ? s_begin[mid] : ops[op_lo + mid].begin - base;
    if (b <= gid) lo = mid; else hi = mid - 1;
  }
  const DevOp op = ops[op_lo + lo];
  op_point(blk_lds ? s_blocks : blocks, op, gid - (op.begin - base), slot, src, props != 0);
}

// store_last_step and compute_square_of_dbv + reduction_of_dv_squared<4>
// (src/TimeAdvanceFunc.cu:6-37, src/Residual.cuh:12-69).
__global__ __launch_bounds__(256) void k_store_last(DevBlock B) {
  const long long t = blockIdx.x * 256ll + threadIdx.x;
  if (t >= B.n0) return;
  const int i = (int)(t % B.mx), j = (int)((t / B.mx) % B.my), k = (int)(t / ((long long)B.mx * B.my));
  const long long p = pidx(B, i, j, k), s = B.sz;
  const double* bv = B.bv[0];
  const double u = bv[p + s], v = bv[p + 2 * s], w = bv[p + 3 * s];
  B.bv_last[t] = bv[p];
  B.bv_last[t + B.n0] = sqrt(u * u + v * v + w * w);
  B.bv_last[t + 2 * B.n0] = bv[p + 4 * s];
  B.bv_last[t + 3 * B.n0] = bv[p + 5 * s];
}

__global__ __launch_bounds__(256) void k_residual(DevBlock B, double* out4) {
  __shared__ double part[4][4];
  double acc[4] = {0, 0, 0, 0};
  const long long s = B.sz;
  const double* bv = B.bv[0];
  for (long long t = blockIdx.x * 256ll + threadIdx.x; t < B.n0; t += 256ll * gridDim.x) {
    const int i = (int)(t % B.mx), j = (int)((t / B.mx) % B.my), k = (int)(t / ((long long)B.mx * B.my));
    const long long p = pidx(B, i, j, k);
    const double u = bv[p + s], v = bv[p + 2 * s], w = bv[p + 3 * s];
    const double vel = sqrt(u * u + v * v + w * w);
    const double d0 = bv[p] - B.bv_last[t], d1 = vel - B.bv_last[t + B.n0];
    const double d2 = bv[p + 4 * s] - B.bv_last[t + 2 * B.n0], d3 = bv[p + 5 * s] - B.bv_last[t + 3 * B.n0];
    acc[0] += d0 * d0; acc[1] += d1 * d1; acc[2] += d2 * d2; acc[3] += d3 * d3;
  }
  for (int l = 0; l < 4; ++l) {
    const double v = wave_sum(acc[l]);
    if ((threadIdx.x & 63) == 0) part[l][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int l = threadIdx.x;
    out4[blockIdx.x * 4 + l] = part[l][0] + part[l][1] + part[l][2] + part[l][3];
  }
}

// ------------------------------------------------------------- launchers
static dim3 pencil_grid(const DevBlock& B, int dir) {
  const int m = dir == 0 ? B.mx : (dir == 1 ? B.my : B.mz);
  const unsigned seg = (m - 1) / 63 + 1;
  if (dir == 0) return dim3(seg, B.my, B.mz);
  if (dir == 1) return dim3(seg, B.mx, B.mz);
  return dim3(seg, B.mx, B.my);
}

template <int SCHEME>
static void launch_weno(const DevBlock& B, const Params& P, hipStream_t st, long long* n) {
  hipLaunchKernelGGL((k_weno_pencil<0, SCHEME>), pencil_grid(B, 0), dim3(64), 0, st, B, P, 0);
  hipLaunchKernelGGL((k_weno_pencil<1, SCHEME>), pencil_grid(B, 1), dim3(64), 0, st, B, P, 0);
  hipLaunchKernelGGL((k_weno_pencil<2, SCHEME>), pencil_grid(B, 2), dim3(64), 0, st, B, P, 0);
  *n += 3;
}

// sponge_layer_source (src/SpongeLayer.cu:68-160), first term of the stage's
// dq as in compute_source: dq -= sigma (cv - mean), sigma times the Jacobian
// when sp_jac is set.
__global__ __launch_bounds__(256) void k_sponge_source(DevBlock B) {
  const long long t = blockIdx.x * 256ll + threadIdx.x;
  if (t >= B.n0) return;
  const int i = (int)(t % B.mx), j = (int)((t / B.mx) % B.my), k = (int)(t / ((long long)B.mx * B.my));
  const long long p = pidx(B, i, j, k);
  const double sigma = B.sp_jac ? B.sp_sigma[t] * B.jac[p] : B.sp_sigma[t];
  for (int l = 0; l < 5; ++l) B.dq[t + l * B.n0] -= sigma * (B.cv[0][p + l * B.sz] - B.sp_mean[t + l * B.n0]);
}

// update_sponge_layer_value (src/SpongeLayer.cu:28-49): mean = (mean n + cv) / (n + 1).
__global__ __launch_bounds__(256) void k_sponge_mean(DevBlock B, const unsigned long long* iter) {
  const long long t = blockIdx.x * 256ll + threadIdx.x;
  if (t >= B.n0) return;
  const int i = (int)(t % B.mx), j = (int)((t / B.mx) % B.my), k = (int)(t / ((long long)B.mx * B.my));
  const long long p = pidx(B, i, j, k);
  const double n = (double)*iter;
  for (int l = 0; l < 5; ++l) {
    double& m = B.sp_mean[t + l * B.n0];
    m = (m * n + B.cv[0][p + l * B.sz]) / (n + 1);
  }
}

void sponge_mean(const DevBlock& B, const unsigned long long* iter, hipStream_t st, long long* n) {
  hipLaunchKernelGGL(k_sponge_mean, dim3((unsigned)((B.n0 + 255) / 256)), dim3(256), 0, st, B, iter);
  *n += 1;
}

void plain_rhs(const DevBlock& B, const Params& P, hipStream_t st, long long* n) {
  if (B.sp_sigma) {
    hipLaunchKernelGGL(k_sponge_source, dim3((unsigned)((B.n0 + 255) / 256)), dim3(256), 0, st, B);
    *n += 1;
  }
  if (P.scheme == 71) launch_weno<71>(B, P, st, n);
  else if (P.scheme == 50) launch_weno<50>(B, P, st, n);
  else if (P.scheme == 52) launch_weno<52>(B, P, st, n);
  else if (P.scheme == 72) launch_weno<72>(B, P, st, n);
  else if (P.scheme == 6) {
    hipLaunchKernelGGL(k_ep_pencil<0>, pencil_grid(B, 0), dim3(64), 0, st, B, 0);
    hipLaunchKernelGGL(k_ep_pencil<1>, pencil_grid(B, 1), dim3(64), 0, st, B, 0);
    hipLaunchKernelGGL(k_ep_pencil<2>, pencil_grid(B, 2), dim3(64), 0, st, B, 0);
    *n += 3;
  }
  else if (P.scheme == 1 || P.scheme == 2 || P.scheme == 3 || P.scheme == 4) {
    hipLaunchKernelGGL(k_pv_pencil<0>, pencil_grid(B, 0), dim3(64), 0, st, B, P, 0);
    hipLaunchKernelGGL(k_pv_pencil<1>, pencil_grid(B, 1), dim3(64), 0, st, B, P, 0);
    hipLaunchKernelGGL(k_pv_pencil<2>, pencil_grid(B, 2), dim3(64), 0, st, B, P, 0);
    *n += 3;
  } else launch_weno<51>(B, P, st, n);
  if (P.visc) {
    hipLaunchKernelGGL(k_visc_pencil<0>, pencil_grid(B, 0), dim3(64), 0, st, B, P, 0);
    hipLaunchKernelGGL(k_visc_pencil<1>, pencil_grid(B, 1), dim3(64), 0, st, B, P, 0);
    hipLaunchKernelGGL(k_visc_pencil<2>, pencil_grid(B, 2), dim3(64), 0, st, B, P, 0);
    *n += 3;
  }
}

static dim3 interior_grid(const DevBlock& B) {
  return dim3((B.mx - 1) / 8 + 1, (B.my - 1) / 8 + 1, (B.mz - 1) / 4 + 1);
}

void plain_update(const DevBlock& B, double dt, int rk, hipStream_t st, long long* n) {
  const RkCoef c = rk_coef(rk);
  hipLaunchKernelGGL(k_update, interior_grid(B), dim3(8, 8, 4), 0, st, B, dt, c.a, c.b, c.c);
  *n += 1;
}

void plain_local_dt(const DevBlock& B, const Params& P, hipStream_t st, long long* n) {
  hipLaunchKernelGGL(k_local_dt, interior_grid(B), dim3(8, 8, 4), 0, st, B, P);
  *n += 1;
}

void min_reduce(const double* in, long long len, double* scratch, int n_first, hipStream_t st, long long* n) {
  const int nb = (int)std::min<long long>(n_first, (len + 255) / 256);
  hipLaunchKernelGGL(k_min_pass, dim3(nb), dim3(256), 0, st, in, len, scratch + 1);
  hipLaunchKernelGGL(k_min_pass, dim3(1), dim3(256), 0, st, scratch + 1, (long long)nb, scratch);
  *n += 2;
}

void padded_pass(const DevBlock& B, int slot, int what, hipStream_t st, long long* n) {
  const dim3 g((B.mx + 2 * B.g - 1) / 8 + 1, (B.my + 2 * B.g - 1) / 8 + 1, (B.mz + 2 * B.g - 1) / 4 + 1);
  hipLaunchKernelGGL(k_padded, g, dim3(8, 8, 4), 0, st, B, slot, what);
  *n += 1;
}

void ops_launch(const DevBlock* d_blocks, int n_blocks, const DevOp* d_ops, int op_lo, int op_hi, long long total, int slot,
                int src, hipStream_t st, long long* n, bool props) {
  if (total <= 0) return;
  const unsigned nb = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(k_ops, dim3(nb), dim3(256), 0, st, d_blocks, n_blocks, d_ops, op_lo, op_hi, total, slot, src,
                     props ? 1 : 0);
  *n += 1;
}

// Parallel faces (faces shared with another rank), air: the 6 primitives of
// the face node and the ngg nodes inside it, per face point, into `buf`
// (setup_data_to_be_sent, src/DataCommunication.cu:3-40) ...
__global__ __launch_bounds__(kParThreads) void k_par_pack(DevBlock B, ParFace f, double* buf) {
  const double* bv = B.bv[0];
  const long long s = B.sz;
  par_pack_chunk<6>(f, B.sx, B.sy, B.g, buf, [&](int v, long long p) { return bv[p + v * s]; });
}

// ... and from the partner's buffer: the face node takes the mean of its own
// and the received primitives, the ngg ghost nodes beyond it the received
// interior ones, each with cv from its primitives (assign_data_received,
// src/DataCommunication.cuh:201-240; cv as compute_cv_from_bv_1_point<Air>).
__global__ __launch_bounds__(kParThreads) void k_par_unpack(DevBlock B, ParFace f, const double* recv,
                                                            const double* sent) {
  double* bv = B.bv[0];
  double* cv = B.cv[0];
  const long long s = B.sz;
  par_unpack_chunk<6>(
      f, B.sx, B.sy, B.g, recv, sent, [&](int v, long long p) { return bv[p + v * s]; },
      [&](int v, long long p, double x) { bv[p + v * s] = x; },
      [&](long long p, const double* q) {
        const double rho = q[0];
        cv[p] = rho;
        cv[p + s] = rho * q[1];
        cv[p + 2 * s] = rho * q[2];
        cv[p + 3 * s] = rho * q[3];
        cv[p + 4 * s] = total_energy(rho, q[1], q[2], q[3], q[4]);
      });
}

void par_pack(const DevBlock& B, const ParFace& f, double* buf, hipStream_t st) {
  if (f.points <= 0) return;
  hipLaunchKernelGGL(k_par_pack, dim3(par_chunks(f)), dim3(kParThreads), 0, st, B, f, buf);
}

void par_unpack(const DevBlock& B, const ParFace& f, const double* recv, const double* sent, hipStream_t st) {
  if (f.points <= 0) return;
  hipLaunchKernelGGL(k_par_unpack, dim3(par_chunks(f)), dim3(kParThreads), 0, st, B, f, recv, sent);
}

// One more boundary-condition pass for blocks lo..hi-1 (the counter that
// keys the inflow's white noise).  A kernel rather than a memory operation, so
// that it records into a captured graph like the ops it follows.
__global__ void k_count_pass(unsigned long long* pass, int lo, int hi) {
  const int b = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (b < hi) pass[b] += 1;
}

void count_pass(unsigned long long* pass, int lo, int hi, hipStream_t st, long long* n) {
  if (hi <= lo) return;
  hipLaunchKernelGGL(k_count_pass, dim3((unsigned)((hi - lo + 63) / 64)), dim3(64), 0, st, pass, lo, hi);
  *n += 1;
}

void store_last(const DevBlock& B, hipStream_t st) {
  hipLaunchKernelGGL(k_store_last, dim3((unsigned)((B.n0 + 255) / 256)), dim3(256), 0, st, B);
}

void residual_partial(const DevBlock& B, double* out4, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(k_residual, dim3(nblk), dim3(256), 0, st, B, out4);
}

}  // namespace ostc
