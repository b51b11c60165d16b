// ostc_oracle.cpp -- CPU oracle for OpenSTC's per-stage RHS + SSP-RK3 path (air).
//
// TEST INFRASTRUCTURE ONLY: a host-C++ restatement, OpenMP over cells, of the
// reference routines listed beside each function.  Expression order follows
// the reference operation by operation; it is built with -ffp-contract=off so
// every product and sum is rounded separately, as written.  Only tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py use it; the
// HIP library never links or calls it.
//
// Parity pins (see tests/test_oracle_*.py): WENO-Z5/Z7/JS closed forms
// (polynomial exactness, constant-data weights, mirror symmetry, 50-digit
// mpmath evaluation), LF split identities, free-stream preservation,
// telescoping conservation on a periodic box, RK3 amplification polynomial,
// Sutherland and dt closed forms, viscous linear-profile exactness, and an
// independent numpy restatement of the whole RHS on tiny grids.
#include "ostc_oracle.h"
#include "ostc_oracle_faces.h"
#include "ostc_oracle_inflow.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#ifdef ORC_DUAL
#include "dual.h"
#endif
#ifdef ORC_COUNT
#include "count.h"
#endif

namespace {

// State values are `real`: double, or a {value, tangent} dual number in the
// tangent-linear build (-DORC_DUAL).  Geometry values (grid, metric,
// Jacobian) are `geo`: double in both, and in the operation-counting build
// (-DORC_COUNT) state and geometry are both counted scalars.  The time step
// stays double.
#ifdef ORC_DUAL
using real = orc::Dual;
using geo = double;
using orc::tangent_of;
using orc::value_of;
constexpr int kWidth = 2;
#elif defined(ORC_COUNT)
using real = orc::Counted;
using geo = orc::Counted;
using orc::value_of;
inline double tangent_of(const real&) { return 0.0; }
constexpr int kWidth = 1;
#else
using real = double;
using geo = double;
inline double value_of(double a) { return a; }
inline double tangent_of(double) { return 0.0; }
constexpr int kWidth = 1;
#endif
using std::abs;
using std::pow;
using std::sqrt;
using std::tanh;

// src/Constants.h:17-34
constexpr double kRu = 8314.462618;
constexpr double kGammaAir = 1.4;
constexpr double kMwAir = 28.9647;
constexpr double kRAir = kRu / kMwAir;

// src/RK.cuh:10-14
constexpr double kRkA[3] = {1.0, 0.75, 1.0 / 3.0};
constexpr double kRkB[3] = {0.0, 0.25, 2.0 / 3.0};
constexpr double kRkC[3] = {1.0, 0.25, 2.0 / 3.0};

struct Face {       // physical boundary face, src/Mesh.cpp:14-32
  int axis, side, type;
  orc_farfield ff;   // type ORC_BC_FARFIELD: the state of its label
  int lo[3], hi[3];  // the face layer; tangential node range widened by ngg
  orc_inflow in;     // type ORC_BC_INFLOW: the inflow of its label
};

using InnerFace = orc_faces::Inner;  // 1-to-1 inner face, src/Mesh.cpp:38-176

struct Block {
  int mx, my, mz, g;
  long long sx, sy, sz;  // padded strides for g ghosts
  std::vector<real> cv, bv, qn, dq, mul, mach, dtl, bv_last;
  // dual time stepping + DPLUR (allocated by orc_dual_time_init): the state
  // of the previous physical step (padded, n_var planes), qn_star, the
  // inviscid spectral radii (3 planes), dq0 and dqk (interior), and the
  // primitives of the last inner iteration (4 planes)
  std::vector<real> qn1, qn_star, sr, dq0, dqk, last_iter;
  std::vector<geo> jac, metric;
  std::vector<double> x, y, z;
  std::vector<Face> faces;
  std::vector<InnerFace> inner;
  unsigned long long bc_passes = 0;  // boundary-condition passes so far (keys the inflow noise)
  bool sponge = false;               // sponge layer on this block
  orc_sponge sp{};
  std::vector<real> sponge_mean_cv;  // [5][n0]

  long long id(int i, int j, int k) const { return (k + g) * sx * sy + (j + g) * sx + (i + g); }
  long long id0(int i, int j, int k) const { return (long long)k * mx * my + (long long)j * mx + i; }
  // grid arrays carry g+1 ghost layers (src/Mesh.cpp:687)
  long long idg(int i, int j, int k) const {
    const long long gx = mx + 2 * (g + 1), gy = my + 2 * (g + 1);
    return (k + g + 1) * gx * gy + (j + g + 1) * gx + (i + g + 1);
  }
  long long n0() const { return (long long)mx * my * mz; }
};

struct Ctx {
  orc_setup s;
  int ngg;
  std::vector<Block> blocks;
  double weno_eps_scale;
  double dt_cur = 1e6;
  unsigned long long sponge_iter = 0;  // steps averaged into the sponge means
  orc_dual_time dual{};                // dual time stepping settings and counters
  bool dual_ready = false;
  double dual_res_scale[4] = {1, 1, 1, 1};
};

inline double sqr(double a) { return a * a; }

// src/Transport.cu:6-9
inline real sutherland(const real& t) {
  return 1.7894e-5 * pow(t / 288.16, 1.5) * (288.16 + 110) / (t + 110);
}

// ---------------------------------------------------------------- WENO kernels
// Smoothness indicators of the three 3-point candidates of a 5-point stencil
// v[0..4] (src/WENO.cu:1587-1592; the same formulas serve the F- side).  The
// reference writes each squared factor twice; it is computed once here, which
// rounds identically.
inline void beta5(const real* v, real b[3]) {
  constexpr double c1312{13.0 / 12};
  real s = v[2] + v[4] - 2 * v[3], t = 3 * v[2] - 4 * v[3] + v[4];
  b[0] = c1312 * s * s + 0.25 * t * t;
  s = v[1] + v[3] - 2 * v[2];
  t = v[1] - v[3];
  b[1] = c1312 * s * s + 0.25 * t * t;
  s = v[0] + v[2] - 2 * v[1];
  t = v[0] - 4 * v[1] + 3 * v[2];
  b[2] = c1312 * s * s + 0.25 * t * t;
}

// WENO-Z5 weights alpha_k = d_k (1 + tau5 / (eps + beta_k)^2), tau5 = (beta_0 - beta_2)^2.
inline void alpha5z(const real b[3], geo eps, double d0, double d1, double d2, real a[3]) {
  const real db = b[0] - b[2];
  const real tau = db * db;
  const real e0 = eps + b[0], e1 = eps + b[1], e2 = eps + b[2];
  a[0] = d0 + d0 * tau / (e0 * e0);
  a[1] = d1 + d1 * tau / (e1 * e1);
  a[2] = d2 + d2 * tau / (e2 * e2);
}

// WENO-Z5, src/WENO.cu:1582-1617
real weno5z(const real* vp, const real* vm, geo eps) {
  constexpr double one6th{1.0 / 6};
  real v0{one6th * (2 * vp[2] + 5 * vp[3] - vp[4])};
  real v1{one6th * (-vp[1] + 5 * vp[2] + 2 * vp[3])};
  real v2{one6th * (2 * vp[0] - 7 * vp[1] + 11 * vp[2])};
  real b[3], a[3];
  beta5(vp, b);
  alpha5z(b, eps, 0.3, 0.6, 0.1, a);
  const real fp{(a[0] * v0 + a[1] * v1 + a[2] * v2) / (a[0] + a[1] + a[2])};

  v0 = one6th * (11 * vm[2] - 7 * vm[3] + 2 * vm[4]);
  v1 = one6th * (2 * vm[1] + 5 * vm[2] - vm[3]);
  v2 = one6th * (-vm[0] + 5 * vm[1] + 2 * vm[2]);
  beta5(vm, b);
  alpha5z(b, eps, 0.1, 0.6, 0.3, a);
  const real fm{(a[0] * v0 + a[1] * v1 + a[2] * v2) / (a[0] + a[1] + a[2])};
  return fp + fm;
}

// WENO5-JS (Jiang & Shu 1996, J. Comput. Phys. 126:202): the same candidates
// and smoothness indicators as WENO-Z5 above, weights alpha_k = d_k/(eps+beta_k)^2.
// Not present in the reference; selected by inviscid_scheme = 50.
real weno5js(const real* vp, const real* vm, geo eps) {
  constexpr double one6th{1.0 / 6};
  real v0{one6th * (2 * vp[2] + 5 * vp[3] - vp[4])};
  real v1{one6th * (-vp[1] + 5 * vp[2] + 2 * vp[3])};
  real v2{one6th * (2 * vp[0] - 7 * vp[1] + 11 * vp[2])};
  real b[3];
  beta5(vp, b);
  real e0 = eps + b[0], e1 = eps + b[1], e2 = eps + b[2];
  real a0{0.3 / (e0 * e0)};
  real a1{0.6 / (e1 * e1)};
  real a2{0.1 / (e2 * e2)};
  const real fp{(a0 * v0 + a1 * v1 + a2 * v2) / (a0 + a1 + a2)};

  v0 = one6th * (11 * vm[2] - 7 * vm[3] + 2 * vm[4]);
  v1 = one6th * (2 * vm[1] + 5 * vm[2] - vm[3]);
  v2 = one6th * (-vm[0] + 5 * vm[1] + 2 * vm[2]);
  beta5(vm, b);
  e0 = eps + b[0];
  e1 = eps + b[1];
  e2 = eps + b[2];
  a0 = 0.1 / (e0 * e0);
  a1 = 0.6 / (e1 * e1);
  a2 = 0.3 / (e2 * e2);
  const real fm{(a0 * v0 + a1 * v1 + a2 * v2) / (a0 + a1 + a2)};
  return fp + fm;
}

// One WENO-Z weight d (1 + tau / (eps + beta)^2), (eps + beta) formed once.
inline real zweight(double d, const real& tau, geo eps, const real& beta) {
  const real e = eps + beta;
  return d + d * tau / (e * e);
}

// WENO-Z7, src/WENO.cu:1619-1693
real weno7z(const real* vp, const real* vm, geo eps) {
  constexpr double one6th{1.0 / 6};
  constexpr double d12{13.0 / 12.0}, d13{1043.0 / 960}, d14{1.0 / 12};
  real s1{one6th * (-2 * vp[0] + 9 * vp[1] - 18 * vp[2] + 11 * vp[3])};
  real s2{-vp[0] + 4 * vp[1] - 5 * vp[2] + 2 * vp[3]};
  real s3{-vp[0] + 3 * vp[1] - 3 * vp[2] + vp[3]};
  real b0{s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3};
  s1 = one6th * (vp[1] - 6 * vp[2] + 3 * vp[3] + 2 * vp[4]);
  s2 = vp[2] - 2 * vp[3] + vp[4];
  s3 = -vp[1] + 3 * vp[2] - 3 * vp[3] + vp[4];
  real b1{s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3};
  s1 = one6th * (-2 * vp[2] - 3 * vp[3] + 6 * vp[4] - vp[5]);
  s3 = -vp[2] + 3 * vp[3] - 3 * vp[4] + vp[5];
  real b2{s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3};
  s1 = one6th * (-11 * vp[3] + 18 * vp[4] - 9 * vp[5] + 2 * vp[6]);
  s2 = 2 * vp[3] - 5 * vp[4] + 4 * vp[5] - vp[6];
  s3 = -vp[3] + 3 * vp[4] - 3 * vp[5] + vp[6];
  real b3{s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3};
  const real db{b0 - b3};
  real tau{db * db};
  constexpr double c0{1.0 / 35}, c1{12.0 / 35}, c2{18.0 / 35}, c3{4.0 / 35};
  real a0{zweight(c0, tau, eps, b0)};
  real a1{zweight(c1, tau, eps, b1)};
  real a2{zweight(c2, tau, eps, b2)};
  real a3{zweight(c3, tau, eps, b3)};
  constexpr double one12th{1.0 / 12};
  real v0{one12th * (-3 * vp[0] + 13 * vp[1] - 23 * vp[2] + 25 * vp[3])};
  real v1{one12th * (vp[1] - 5 * vp[2] + 13 * vp[3] + 3 * vp[4])};
  real v2{one12th * (-vp[2] + 7 * vp[3] + 7 * vp[4] - vp[5])};
  real v3{one12th * (3 * vp[3] + 13 * vp[4] - 5 * vp[5] + vp[6])};
  const real fp{(a0 * v0 + a1 * v1 + a2 * v2 + a3 * v3) / (a0 + a1 + a2 + a3)};

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
  const real dbm{b0 - b3};
  tau = dbm * dbm;
  a0 = zweight(c0, tau, eps, b0);
  a1 = zweight(c1, tau, eps, b1);
  a2 = zweight(c2, tau, eps, b2);
  a3 = zweight(c3, tau, eps, b3);
  v0 = one12th * (-3 * vm[6] + 13 * vm[5] - 23 * vm[4] + 25 * vm[3]);
  v1 = one12th * (vm[5] - 5 * vm[4] + 13 * vm[3] + 3 * vm[2]);
  v2 = one12th * (-vm[4] + 7 * vm[3] + 7 * vm[2] - vm[1]);
  v3 = one12th * (3 * vm[3] + 13 * vm[2] - 5 * vm[1] + vm[0]);
  const real fm{(a0 * v0 + a1 * v1 + a2 * v2 + a3 * v3) / (a0 + a1 + a2 + a3)};
  return fp + fm;
}

// Point-local Lax-Friedrichs splitting, src/WENO.cu:697-720.
// q = [cv(0..nv-1), p, c]; k3 = metric row of the sweep direction.
void lf_split(const real* q, int nv, const geo* m, geo jac, real* fp, real* fm) {
  const real Uk{(q[1] * m[0] + q[2] * m[1] + q[3] * m[2]) / q[0]};
  const real pk{q[nv]};
  const real cGradK = q[nv + 1] * sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const real lambda0 = abs(Uk) + cGradK;
  const geo hj = 0.5 * jac;  // the reference's 0.5 * jac of every line, formed once
  fp[0] = hj * (Uk * q[0] + lambda0 * q[0]);
  fp[1] = hj * (Uk * q[1] + pk * m[0] + lambda0 * q[1]);
  fp[2] = hj * (Uk * q[2] + pk * m[1] + lambda0 * q[2]);
  fp[3] = hj * (Uk * q[3] + pk * m[2] + lambda0 * q[3]);
  fp[4] = hj * (Uk * (q[4] + pk) + lambda0 * q[4]);
  fm[0] = hj * (Uk * q[0] - lambda0 * q[0]);
  fm[1] = hj * (Uk * q[1] + pk * m[0] - lambda0 * q[1]);
  fm[2] = hj * (Uk * q[2] + pk * m[1] - lambda0 * q[2]);
  fm[3] = hj * (Uk * q[3] + pk * m[2] - lambda0 * q[3]);
  fm[4] = hj * (Uk * (q[4] + pk) - lambda0 * q[4]);
  for (int l = 5; l < nv; ++l) {
    fp[l] = hj * (Uk * q[l] + lambda0 * q[l]);
    fm[l] = hj * (Uk * q[l] - lambda0 * q[l]);
  }
}

// ------------------------------------------------------------------ operators
// compute_cv_from_bv / compute_cv_from_bv_1_point / compute_total_energy<Air>,
// src/FieldOperation.cuh:17-105
inline void cv_from_bv_point(Block& b, long long p) {
  const long long s = b.sz;
  const real rho = b.bv[p];
  b.cv[p] = rho;
  b.cv[p + s] = rho * b.bv[p + s];
  b.cv[p + 2 * s] = rho * b.bv[p + 2 * s];
  b.cv[p + 3 * s] = rho * b.bv[p + 3 * s];
  const real u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
  const real V2 = u * u + v * v + w * w;
  real te = 0.5 * V2;
  te *= rho;
  te += b.bv[p + 4 * s] / (kGammaAir - 1);
  b.cv[p + 4 * s] = te;
}

void cv_from_bv(Block& b) {
  const int g = b.g;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = -g; k < b.mz + g; ++k)
    for (int j = -g; j < b.my + g; ++j)
      for (int i = -g; i < b.mx + g; ++i) cv_from_bv_point(b, b.id(i, j, k));
}

// update_physical_properties<Air>, src/FieldOperation.cuh:107-140
void properties(Block& b) {
  const int g = b.g;
  const long long s = b.sz;
  constexpr double c_temp{kGammaAir * kRu / kMwAir};
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = -g; k < b.mz + g; ++k)
    for (int j = -g; j < b.my + g; ++j)
      for (int i = -g; i < b.mx + g; ++i) {
        const long long p = b.id(i, j, k);
        const real t = b.bv[p + 5 * s];
        const real u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
        const real V = sqrt(u * u + v * v + w * w);
        b.mul[p] = sutherland(t);
        b.mach[p] = V / sqrt(c_temp * t);
      }
}

// local_time_step<Air, Laminar>, src/TimeAdvanceFunc.cuh:28-83
void local_dt(const Ctx& c, Block& b) {
  const long long s = b.sz, ms = b.sz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k);
        const geo* M = &b.metric[0];
        auto m = [&](int r, int cc) { return M[p + (3 * (r - 1) + (cc - 1)) * ms]; };
        const geo gxi = sqrt(m(1, 1) * m(1, 1) + m(1, 2) * m(1, 2) + m(1, 3) * m(1, 3));
        const geo geta = sqrt(m(2, 1) * m(2, 1) + m(2, 2) * m(2, 2) + m(2, 3) * m(2, 3));
        const geo gzeta = sqrt(m(3, 1) * m(3, 1) + m(3, 2) * m(3, 2) + m(3, 3) * m(3, 3));
        const real u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
        const real U = u * m(1, 1) + v * m(1, 2) + w * m(1, 3);
        const real V = u * m(2, 1) + v * m(2, 2) + w * m(2, 3);
        const real W = u * m(3, 1) + v * m(3, 2) + w * m(3, 3);
        const real a = sqrt(kGammaAir * kRu * b.bv[p + 5 * s] / kMwAir);
        const real r0 = abs(U) + a * gxi;
        const real r1 = abs(V) + a * geta;
        const real r2 = abs(W) + a * gzeta;
        const real sr_inv = r0 + r1 + r2;
        const real coeff_1 = std::max(kGammaAir, 4.0 / 3.0) / b.bv[p];
        const real coeff_2 = b.mul[p] / c.s.prandtl;
        real sr_visc = gxi * gxi + geta * geta;
        sr_visc += gzeta * gzeta;
        sr_visc *= coeff_1 * coeff_2;
        b.dtl[b.id0(i, j, k)] = c.s.cfl / (sr_inv + sr_visc);
      }
}


// ------------------------------------------------ characteristic WENO (air)
// compute_weno_flux_ch<Air>, "method 1", src/WENO.cu:1147-1453: Roe-averaged
// left eigenvectors at the face (LR filled as L, then as R), the split fluxes
// of the five points projected, WENO-Z5 per field, projected back.
struct CharAir {
  real L[5][5], R[5][5];
};

void char_air(const real* cvl, const real* cvr, const geo* m1, const geo* m2, geo jac1, geo jac2,
              CharAir& M) {
  const real rlc{sqrt(cvl[0]) / (sqrt(cvl[0]) + sqrt(cvr[0]))};
  const real rrc{sqrt(cvr[0]) / (sqrt(cvl[0]) + sqrt(cvr[0]))};
  const real um{rlc * cvl[1] / cvl[0] + rrc * cvr[1] / cvr[0]};
  const real vm{rlc * cvl[2] / cvl[0] + rrc * cvr[2] / cvr[0]};
  const real wm{rlc * cvl[3] / cvl[0] + rrc * cvr[3] / cvr[0]};
  const real ekm{0.5 * (um * um + vm * vm + wm * wm)};
  constexpr double gm1{kGammaAir - 1};
  const real hl{(cvl[4] + cvl[5]) / cvl[0]};
  const real hr{(cvr[4] + cvr[5]) / cvr[0]};
  const real hm{rlc * hl + rrc * hr};
  const real cm2{gm1 * (hm - ekm)};
  const real cm{sqrt(cm2)};
  geo kx{(jac1 * m1[0] + jac2 * m2[0]) / (jac1 + jac2)};
  geo ky{(jac1 * m1[1] + jac2 * m2[1]) / (jac1 + jac2)};
  geo kz{(jac1 * m1[2] + jac2 * m2[2]) / (jac1 + jac2)};
  const geo gradK{sqrt(kx * kx + ky * ky + kz * kz)};
  kx /= gradK;
  ky /= gradK;
  kz /= gradK;
  const real Uk_bar{kx * um + ky * vm + kz * wm};
  const real alpha{gm1 * ekm};
  auto& L = M.L;
  L[0][0] = (alpha + Uk_bar * cm) / cm2 * 0.5;
  L[0][1] = -(gm1 * um + kx * cm) / cm2 * 0.5;
  L[0][2] = -(gm1 * vm + ky * cm) / cm2 * 0.5;
  L[0][3] = -(gm1 * wm + kz * cm) / cm2 * 0.5;
  L[0][4] = gm1 / cm2 * 0.5;
  L[1][0] = kx * (1 - alpha / cm2) - (kz * vm - ky * wm) / cm;
  L[1][1] = kx * gm1 * um / cm2;
  L[1][2] = (kx * gm1 * vm + kz * cm) / cm2;
  L[1][3] = (kx * gm1 * wm - ky * cm) / cm2;
  L[1][4] = -kx * gm1 / cm2;
  L[2][0] = ky * (1 - alpha / cm2) - (kx * wm - kz * um) / cm;
  L[2][1] = (ky * gm1 * um - kz * cm) / cm2;
  L[2][2] = ky * gm1 * vm / cm2;
  L[2][3] = (ky * gm1 * wm + kx * cm) / cm2;
  L[2][4] = -ky * gm1 / cm2;
  L[3][0] = kz * (1 - alpha / cm2) - (ky * um - kx * vm) / cm;
  L[3][1] = (kz * gm1 * um + ky * cm) / cm2;
  L[3][2] = (kz * gm1 * vm - kx * cm) / cm2;
  L[3][3] = kz * gm1 * wm / cm2;
  L[3][4] = -kz * gm1 / cm2;
  L[4][0] = (alpha - Uk_bar * cm) / cm2 * 0.5;
  L[4][1] = -(gm1 * um - kx * cm) / cm2 * 0.5;
  L[4][2] = -(gm1 * vm - ky * cm) / cm2 * 0.5;
  L[4][3] = -(gm1 * wm - kz * cm) / cm2 * 0.5;
  L[4][4] = gm1 / cm2 * 0.5;
  auto& R = M.R;
  R[0][0] = 1.0;
  R[0][1] = kx;
  R[0][2] = ky;
  R[0][3] = kz;
  R[0][4] = 1.0;
  R[1][0] = um - kx * cm;
  R[1][1] = kx * um;
  R[1][2] = ky * um - kz * cm;
  R[1][3] = kz * um + ky * cm;
  R[1][4] = um + kx * cm;
  R[2][0] = vm - ky * cm;
  R[2][1] = kx * vm + kz * cm;
  R[2][2] = ky * vm;
  R[2][3] = kz * vm - kx * cm;
  R[2][4] = vm + ky * cm;
  R[3][0] = wm - kz * cm;
  R[3][1] = kx * wm - ky * cm;
  R[3][2] = ky * wm + kx * cm;
  R[3][3] = kz * wm;
  R[3][4] = wm + kz * cm;
  R[4][0] = hm - Uk_bar * cm;
  R[4][1] = kx * alpha / gm1 + (kz * vm - ky * wm) * cm;
  R[4][2] = ky * alpha / gm1 + (kx * wm - kz * um) * cm;
  R[4][3] = kz * alpha / gm1 + (ky * um - kx * vm) * cm;
  R[4][4] = hm + Uk_bar * cm;
}

// ---------------------------------------- primitive reconstruction (air)
// gxl::sgn (src/gxl_lib/Math.hpp:13-15): -1 for negative, +1 otherwise
inline int sgn(const real& a) { return a < 0.0 ? -1 : 1; }
// minmod, src/Limiter.cuh:7-15
inline real minmod(const real& a, const real& b) {
  return 0.5 * (sgn(a) + sgn(b)) * std::min(abs(a), abs(b));
}

// src/Reconstruction.cu:5-46, face between pencil points is and is+1;
// pv holds 5 values per point; pl/pr get [rho,u,v,w,p,E] (E of
// reconstruction<Air>, src/InviscidScheme.cu:210-215)
void reconstruct_air(const real* pv, int is, int recon, real* pl, real* pr) {
  for (int l = 0; l < 5; ++l) {
    const real a0 = pv[(is - 1) * 5 + l], a1 = pv[is * 5 + l], a2 = pv[(is + 1) * 5 + l];
    if (recon == 2 || recon == 3) {
      const real a3 = pv[(is + 2) * 5 + l];
      const real delta_i{a1 - a0}, delta_i1{a2 - a1}, delta_i2{a3 - a2};
      if (recon == 2) {
        constexpr double kappa{1.0 / 3.0};
        const real neg_l = minmod(delta_i, delta_i1), pos_l = minmod(delta_i1, delta_i);
        const real neg_r = minmod(delta_i1, delta_i2), pos_r = minmod(delta_i2, delta_i1);
        pl[l] = a1 + 0.25 * ((1 - kappa) * neg_l + (1 + kappa) * pos_l);
        pr[l] = a2 - 0.25 * ((1 - kappa) * pos_r + (1 + kappa) * neg_r);
      } else {
        pl[l] = a1 + 0.5 * minmod(delta_i, delta_i1);
        pr[l] = a2 - 0.5 * minmod(delta_i1, delta_i2);
      }
    } else {
      pl[l] = a1;
      pr[l] = a2;
    }
  }
  const real el = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
  const real er = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
  pl[5] = el * pl[0] + pl[4] / (kGammaAir - 1);
  pr[5] = er * pr[0] + pr[4] / (kGammaAir - 1);
}

// riemannSolver_ausmPlus<Air>, src/RiemannSolver.cuh:9-96
void ausm_air(const real* pl, const real* pr, const geo* ml, const geo* mr, geo jl, geo jr, real* fci) {
  const geo k1 = 0.5 * (jl * ml[0] + jr * mr[0]);
  const geo k2 = 0.5 * (jl * ml[1] + jr * mr[1]);
  const geo k3 = 0.5 * (jl * ml[2] + jr * mr[2]);
  const geo grad_k_div_jac = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const real ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) / grad_k_div_jac;
  const real ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) / grad_k_div_jac;
  const real p_l = pl[4], p_r = pr[4], rho_l = pl[0], rho_r = pr[0];
  const real c = 0.5 * (sqrt(kGammaAir * p_l / rho_l) + sqrt(kGammaAir * p_r / rho_r));
  const real mach_l = ul / c, mach_r = ur / c;
  real mlp, mrn, plp, prn;
  constexpr double alpha{3 / 16.0};
  if (abs(mach_l) > 1.0) {
    mlp = 0.5 * (mach_l + abs(mach_l));
    plp = mlp / mach_l;
  } else {
    const real a = (mach_l + 1) * (mach_l + 1) * 0.25;
    const real b = (mach_l * mach_l - 1) * (mach_l * mach_l - 1);
    mlp = a + 0.125 * b;
    plp = a * (2 - mach_l) + alpha * mach_l * b;
  }
  if (abs(mach_r) > 1.0) {
    mrn = 0.5 * (mach_r - abs(mach_r));
    prn = mrn / mach_r;
  } else {
    const real a = (mach_r - 1) * (mach_r - 1) * 0.25;
    const real b = (mach_r * mach_r - 1) * (mach_r * mach_r - 1);
    mrn = -a - 0.125 * b;
    prn = a * (2 + mach_r) - alpha * mach_r * b;
  }
  const real p_coeff = plp * p_l + prn * p_r;
  const real m_half = mlp + mrn;
  const real mach_pos = 0.5 * (m_half + abs(m_half));
  const real mach_neg = 0.5 * (m_half - abs(m_half));
  const real mass_flux_half = c * (rho_l * mach_pos + rho_r * mach_neg);
  const real coeff = mass_flux_half * grad_k_div_jac;
  const real* q = mass_flux_half >= 0.0 ? pl : pr;
  fci[0] = coeff;
  fci[1] = coeff * q[1] + p_coeff * k1;
  fci[2] = coeff * q[2] + p_coeff * k2;
  fci[3] = coeff * q[3] + p_coeff * k3;
  fci[4] = coeff * (q[5] + q[4]) / q[0];
}

// riemannSolver_hllc<Air>, src/RiemannSolver.cuh:98-230
void hllc_air(const real* pl, const real* pr, const geo* ml, const geo* mr, geo jl, geo jr, real* fci) {
  const real rl_c{sqrt(pl[0]) / (sqrt(pl[0]) + sqrt(pr[0]))}, rr_c{sqrt(pr[0]) / (sqrt(pl[0]) + sqrt(pr[0]))};
  const real u_tilde{pl[1] * rl_c + pr[1] * rr_c};
  const real v_tilde{pl[2] * rl_c + pr[2] * rr_c};
  const real w_tilde{pl[3] * rl_c + pr[3] * rr_c};
  const double gamma{kGammaAir};
  const real hl{(pl[5] + pl[4]) / pl[0]};
  const real hr{(pr[5] + pr[4]) / pr[0]};
  const real h_tilde{hl * rl_c + hr * rr_c};
  const real ek_tilde{0.5 * (u_tilde * u_tilde + v_tilde * v_tilde + w_tilde * w_tilde)};
  const real c_tilde = sqrt((gamma - 1) * (h_tilde - ek_tilde));
  const geo kx = 0.5 * (ml[0] + mr[0]);
  const geo ky = 0.5 * (ml[1] + mr[1]);
  const geo kz = 0.5 * (ml[2] + mr[2]);
  const geo gradK = sqrt(kx * kx + ky * ky + kz * kz);
  const real U_tilde_bar{(kx * u_tilde + ky * v_tilde + kz * w_tilde) / gradK};
  const geo jac_ave{0.5 * (jl + jr)};
  const real Ul{kx * pl[1] + ky * pl[2] + kz * pl[3]};
  const real cl{sqrt(kGammaAir * pl[4] / pl[0])};
  const real sl{std::min(Ul / gradK - cl, U_tilde_bar - c_tilde)};
  const real Ur{kx * pr[1] + ky * pr[2] + kz * pr[3]};
  const real cr{sqrt(kGammaAir * pr[4] / pr[0])};
  const real sr{std::max(Ur / gradK + cr, U_tilde_bar + c_tilde)};
  if (sl >= 0.0 || sr < 0.0) {
    const real* q = sl >= 0.0 ? pl : pr;
    const real U = sl >= 0.0 ? Ul : Ur;
    const real rhoUk{q[0] * U};
    fci[0] = jac_ave * rhoUk;
    fci[1] = jac_ave * (rhoUk * q[1] + q[4] * kx);
    fci[2] = jac_ave * (rhoUk * q[2] + q[4] * ky);
    fci[3] = jac_ave * (rhoUk * q[3] + q[4] * kz);
    fci[4] = jac_ave * ((q[5] + q[4]) * U);
    return;
  }
  const real sm{((pr[0] * Ur * (sr - Ur / gradK) - pl[0] * Ul * (sl - Ul / gradK)) / gradK + pl[4] - pr[4]) /
                (pr[0] * (sr - Ur / gradK) - pl[0] * (sl - Ul / gradK))};
  const real pm{pl[0] * (sl - Ul / gradK) * (sm - Ul / gradK) + pl[4]};
  const bool left = sm >= 0.0;
  const real* q = left ? pl : pr;
  const real sw = left ? sl : sr;
  const real U = left ? Ul : Ur;
  const real pCoeff{1.0 / (sw - sm)};
  const real QCoeff{jac_ave * pCoeff * sm * (sw * gradK - U) * q[0]};
  fci[0] = QCoeff;
  const real dP{(sw * pm - sm * q[4]) * pCoeff * jac_ave};
  fci[1] = QCoeff * q[1] + dP * kx;
  fci[2] = QCoeff * q[2] + dP * ky;
  fci[3] = QCoeff * q[3] + dP * kz;
  fci[4] = QCoeff * q[5] / q[0] + pCoeff * jac_ave * (sw * pm * sm * gradK - sm * q[4] * U);
}

// riemannSolver_laxFriedrich<Air>, src/RiemannSolver.cuh:426-461
void lf_air(const real* pl, const real* pr, const geo* ml, const geo* mr, geo jl, geo jr, real* fci) {
  const geo kx = 0.5 * (ml[0] + mr[0]);
  const geo ky = 0.5 * (ml[1] + mr[1]);
  const geo kz = 0.5 * (ml[2] + mr[2]);
  const geo gradK = sqrt(kx * kx + ky * ky + kz * kz);
  const real Ukl{pl[1] * kx + pl[2] * ky + pl[3] * kz};
  const real Ukr{pr[1] * kx + pr[2] * ky + pr[3] * kz};
  const real cl{sqrt(kGammaAir * pl[4] / pl[0])};
  const real cr{sqrt(kGammaAir * pr[4] / pr[0])};
  const real spectral_radius{std::max(abs(Ukl) + cl * gradK, abs(Ukr) + cr * gradK)};
  const geo half_jac_ave{0.5 * 0.5 * (jl + jr)};
  const real rhoUl{pl[0] * Ukl};
  const real rhoUr{pr[0] * Ukr};
  fci[0] = (rhoUl + rhoUr - spectral_radius * (pr[0] - pl[0])) * half_jac_ave;
  fci[1] = (rhoUl * pl[1] + rhoUr * pr[1] + kx * (pl[4] + pr[4]) - spectral_radius * (pr[1] * pr[0] - pl[1] * pl[0])) *
           half_jac_ave;
  fci[2] = (rhoUl * pl[2] + rhoUr * pr[2] + ky * (pl[4] + pr[4]) - spectral_radius * (pr[2] * pr[0] - pl[2] * pl[0])) *
           half_jac_ave;
  fci[3] = (rhoUl * pl[3] + rhoUr * pr[3] + kz * (pl[4] + pr[4]) - spectral_radius * (pr[3] * pr[0] - pl[3] * pl[0])) *
           half_jac_ave;
  fci[4] = ((pl[5] + pl[4]) * Ukl + (pr[5] + pr[4]) * Ukr - spectral_radius * (pr[5] - pl[5])) * half_jac_ave;
}

// compute_entropy_fix_delta<Air>, src/InviscidScheme.cu:421-453 (3D), with
// the entropy_fix_factor of the setup (src/Parameter.cpp:474, default 0.125).
real entropy_delta(const real* q, const geo* m, double factor) {
  const real U = abs(q[1] * m[0] + q[2] * m[1] + q[3] * m[2]);
  const real V = abs(q[1] * m[3] + q[2] * m[4] + q[3] * m[5]);
  const real W = abs(q[1] * m[6] + q[2] * m[7] + q[3] * m[8]);
  const geo kx = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const geo ky = sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
  const geo kz = sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8]);
  const real c = sqrt(kGammaAir * q[4] / q[0]);
  return factor * (U + V + W + c * (kx + ky + kz) / 3.0);
}

// riemannSolver_Roe<Air>, src/RiemannSolver.cuh:279-417, with
// compute_half_sum_left_right_flux :234-277.
void roe_air(const real* pl, const real* pr, const geo* ml, const geo* mr, geo jl, geo jr, real del_l,
             real del_r, real* fci) {
  geo JacKx = jl * ml[0], JacKy = jl * ml[1], JacKz = jl * ml[2];
  real Uk = pl[1] * JacKx + pl[2] * JacKy + pl[3] * JacKz;
  real coeff = Uk * pl[0];
  fci[0] = 0.5 * coeff;
  fci[1] = 0.5 * (coeff * pl[1] + pl[4] * JacKx);
  fci[2] = 0.5 * (coeff * pl[2] + pl[4] * JacKy);
  fci[3] = 0.5 * (coeff * pl[3] + pl[4] * JacKz);
  fci[4] = 0.5 * Uk * (pl[4] + pl[5]);
  JacKx = jr * mr[0];
  JacKy = jr * mr[1];
  JacKz = jr * mr[2];
  Uk = pr[1] * JacKx + pr[2] * JacKy + pr[3] * JacKz;
  coeff = Uk * pr[0];
  fci[0] += 0.5 * coeff;
  fci[1] += 0.5 * (coeff * pr[1] + pr[4] * JacKx);
  fci[2] += 0.5 * (coeff * pr[2] + pr[4] * JacKy);
  fci[3] += 0.5 * (coeff * pr[3] + pr[4] * JacKz);
  fci[4] += 0.5 * Uk * (pr[4] + pr[5]);

  const real dl = sqrt(pl[0]), dr = sqrt(pr[0]);
  const real inv_denominator = 1.0 / (dl + dr);
  const real u = (dl * pl[1] + dr * pr[1]) * inv_denominator;
  const real v = (dl * pl[2] + dr * pr[2]) * inv_denominator;
  const real w = (dl * pl[3] + dr * pr[3]) * inv_denominator;
  const real ek = 0.5 * (u * u + v * v + w * w);
  const real hl = (pl[5] + pl[4]) / pl[0];
  const real hr = (pr[5] + pr[4]) / pr[0];
  const real h = (dl * hl + dr * hr) * inv_denominator;
  const double gamma = kGammaAir;
  const real c = sqrt((gamma - 1) * (h - ek));
  geo kx = 0.5 * (ml[0] + mr[0]);
  geo ky = 0.5 * (ml[1] + mr[1]);
  geo kz = 0.5 * (ml[2] + mr[2]);
  const geo gradK = sqrt(kx * kx + ky * ky + kz * kz);
  real Ukk = kx * u + ky * v + kz * w;
  real characteristic[3] = {Ukk - gradK * c, Ukk, Ukk + gradK * c};
  const real delta_ave = 0.5 * (del_l + del_r);
  for (auto& cc : characteristic) {
    cc = abs(cc);
    if (cc < delta_ave) cc = 0.5 * (cc * cc / delta_ave + delta_ave);
  }
  kx /= gradK;
  ky /= gradK;
  kz /= gradK;
  Ukk /= gradK;
  const geo jac_ave = 0.5 * (jl + jr);
  real dq[5];
  dq[0] = jac_ave * (pr[0] - pl[0]);
  for (int l = 1; l < 5; ++l) dq[l] = jac_ave * (pr[0] * pr[l] - pl[0] * pl[l]);
  dq[4] = jac_ave * (pr[5] - pl[5]);
  real c1 = (gamma - 1) * (ek * dq[0] - u * dq[1] - v * dq[2] - w * dq[3] + dq[4]) / (c * c);
  real c2 = (kx * dq[1] + ky * dq[2] + kz * dq[3] - Ukk * dq[0]) / c;
  real c3 = dq[0] - c1;
  real LDq[5];
  LDq[0] = 0.5 * (c1 - c2);
  LDq[1] = kx * c3 - ((kz * v - ky * w) * dq[0] - kz * dq[2] + ky * dq[3]) / c;
  LDq[2] = ky * c3 - ((kx * w - kz * u) * dq[0] - kx * dq[3] + kz * dq[1]) / c;
  LDq[3] = kz * c3 - ((ky * u - kx * v) * dq[0] - ky * dq[1] + kx * dq[2]) / c;
  LDq[4] = 0.5 * (c1 + c2);
  real bb[5];
  bb[0] = -characteristic[0] * LDq[0];
  for (int l = 1; l < 5; ++l) bb[l] = -characteristic[1] * LDq[l];
  bb[4] = -characteristic[2] * LDq[4];
  const real c0 = kx * bb[1] + ky * bb[2] + kz * bb[3];
  c1 = c0 + bb[0] + bb[4];
  c2 = c * (bb[4] - bb[0]);
  c3 = 0;
  fci[0] += 0.5 * c1;
  fci[1] += 0.5 * (u * c1 + kx * c2 - c * (kz * bb[2] - ky * bb[3]));
  fci[2] += 0.5 * (v * c1 + ky * c2 - c * (kx * bb[3] - kz * bb[1]));
  fci[3] += 0.5 * (w * c1 + kz * c2 - c * (ky * bb[1] - kx * bb[2]));
  fci[4] += 0.5 * (h * c1 + Ukk * c2 - c * c * c0 / (gamma - 1) +
                   c * ((kz * v - ky * w) * bb[1] + (kx * w - kz * u) * bb[2] + (ky * u - kx * v) * bb[3]) + c3);
}

void riemann_air(int scheme, const real* pl, const real* pr, const geo* ml, const geo* mr, geo jl, geo jr,
                 real* fci) {
  if (scheme == 1) lf_air(pl, pr, ml, mr, jl, jr, fci);
  else if (scheme == 3) ausm_air(pl, pr, ml, mr, jl, jr, fci);
  else hllc_air(pl, pr, ml, mr, jl, jr, fci);
}

// compute_convective_term_pv_1D<Air>, src/InviscidScheme.cu:50-165
void inviscid_pv(const Ctx& c, Block& b) {
  const int g = b.g;
  const long long s = b.sz, n0 = b.n0();
  const int ext[3] = {b.mx, b.my, b.mz};
  for (int dir = 0; dir < 3; ++dir) {
    const int m = ext[dir];
    const int oa = (dir == 0) ? 1 : 0, ob = (dir == 2) ? 1 : 2;
    const int na = ext[oa], nb = ext[ob];
#pragma omp parallel
    {
      const int np = m + 2 * g;
      std::vector<real> PV(np * 5), fc((m + 1) * 5), D(np);
      std::vector<geo> K(np * 3), J(np);
#pragma omp for collapse(2) schedule(static)
      for (int bb = 0; bb < nb; ++bb)
        for (int aa = 0; aa < na; ++aa) {
          int idx[3];
          idx[oa] = aa;
          idx[ob] = bb;
          for (int t = -g; t < m + g; ++t) {
            idx[dir] = t;
            const long long p = b.id(idx[0], idx[1], idx[2]);
            const int ps = t + g;
            for (int l = 0; l < 5; ++l) PV[ps * 5 + l] = b.bv[p + l * s];
            for (int cc = 0; cc < 3; ++cc) K[ps * 3 + cc] = b.metric[p + (3 * dir + cc) * s];
            J[ps] = b.jac[p];
            if (c.s.inviscid_scheme == 2) {
              geo mm[9];
              for (int cc = 0; cc < 9; ++cc) mm[cc] = b.metric[p + cc * s];
              D[ps] = entropy_delta(&PV[ps * 5], mm, c.s.entropy_fix_factor);
            }
          }
          for (int t = -1; t < m; ++t) {
            const int is = t + g;
            real pl[6], pr[6];
            reconstruct_air(PV.data(), is, c.s.reconstruction, pl, pr);
            if (c.s.inviscid_scheme == 2)
              roe_air(pl, pr, &K[is * 3], &K[(is + 1) * 3], J[is], J[is + 1], D[is], D[is + 1], &fc[(t + 1) * 5]);
            else
              riemann_air(c.s.inviscid_scheme, pl, pr, &K[is * 3], &K[(is + 1) * 3], J[is], J[is + 1],
                          &fc[(t + 1) * 5]);
          }
          for (int t = 0; t < m; ++t) {
            idx[dir] = t;
            const long long p0 = b.id0(idx[0], idx[1], idx[2]);
            for (int l = 0; l < 5; ++l) b.dq[p0 + l * n0] -= fc[(t + 1) * 5 + l] - fc[t * 5 + l];
          }
        }
    }
  }
}

// |m1 J1 + m2 J2|^2 of the two points' Jacobian-weighted metric rows, the
// factor of the WENO epsilon (src/WENO.cu:1469-1476); each sum formed once.
inline geo face_metric_sq(const geo* m1, const geo* m2, geo jac1, geo jac2) {
  const geo ax = m1[0] * jac1 + m2[0] * jac2, ay = m1[1] * jac1 + m2[1] * jac2, az = m1[2] * jac1 + m2[2] * jac2;
  return ax * ax + ay * ay + az * az;
}

// compute_convective_term_ep<Air> / _ep_1D, src/InviscidScheme.cu:550-768:
// the sixth-order energy-preserving central scheme (Pirozzoli, J. Comput.
// Phys. 230 (2011) 2997-3014).  Two-point fluxes between points i and i+l,
// l = 1..3, with the Jacobian-weighted contravariant velocity UJ and the
// total enthalpy H = (E + p)/rho:
//   w = (UJ_i + UJ_i+l)(rho_i + rho_i+l)/8,  pw = (p_i + p_i+l)/4,
//   T = (2w, w(u_i + u_i+l) + pw(k_x J|_i + k_x J|_i+l), ..., w(H_i + H_i+l));
// the face flux at i+1/2 is 2 (3/4 T(i,1) - 3/20 (T(i,2) + T(i-1,2))
// + 1/60 (T(i,3) + T(i-1,3) + T(i-2,3))), and dq(i) -= fc(i+1/2) - fc(i-1/2).
void inviscid_ep(const Ctx& c, Block& b) {
  const int nv = c.s.n_var, g = b.g;
  const long long s = b.sz;
  const int ext[3] = {b.mx, b.my, b.mz};
  constexpr double central_1[3] = {0.75, -3.0 / 20, 1.0 / 60};
  for (int dir = 0; dir < 3; ++dir) {
    const int m = ext[dir];
    const int oa = (dir == 0) ? 1 : 0, ob = (dir == 2) ? 1 : 2;
    const int na = ext[oa], nb = ext[ob];
#pragma omp parallel
    {
      const int np = m + 2 * g;
      // per pencil point: rho, u, v, w, p, H, k_x, k_y, k_z, J, UJ
      std::vector<real> pt(np * 11);
      std::vector<real> tilde(np * 3 * 5), fc((m + 1) * 5);
#pragma omp for collapse(2) schedule(static)
      for (int bb = 0; bb < nb; ++bb)
        for (int aa = 0; aa < na; ++aa) {
          int idx[3];
          idx[oa] = aa;
          idx[ob] = bb;
          for (int t = -g; t < m + g; ++t) {
            idx[dir] = t;
            const long long p = b.id(idx[0], idx[1], idx[2]);
            real* P = &pt[(t + g) * 11];
            for (int l = 0; l < 5; ++l) P[l] = b.bv[p + l * s];
            P[5] = (b.cv[p + 4 * s] + b.bv[p + 4 * s]) / b.bv[p];
            for (int cc = 0; cc < 3; ++cc) P[6 + cc] = b.metric[p + (3 * dir + cc) * s];
            P[9] = b.jac[p];
            real uk = P[6] * P[1] + P[7] * P[2] + P[8] * P[3];
            uk *= P[9];
            P[10] = uk;
          }
          // two-point fluxes T(i, l), i = -g .. m+g-1-l
          for (int t = -g; t < m + g; ++t)
            for (int l = 1; l <= 3; ++l) {
              if (t + l > m + g - 1) continue;
              const real* A = &pt[(t + g) * 11];
              const real* Bq = &pt[(t + l + g) * 11];
              const real weight = 0.125 * (A[10] + Bq[10]) * (A[0] + Bq[0]);
              const real pWeight = 0.25 * (A[4] + Bq[4]);
              real* T = &tilde[((t + g) * 3 + (l - 1)) * 5];
              T[0] = 2 * weight;
              T[1] = weight * (A[1] + Bq[1]) + pWeight * (A[6] * A[9] + Bq[6] * Bq[9]);
              T[2] = weight * (A[2] + Bq[2]) + pWeight * (A[7] * A[9] + Bq[7] * Bq[9]);
              T[3] = weight * (A[3] + Bq[3]) + pWeight * (A[8] * A[9] + Bq[8] * Bq[9]);
              T[4] = weight * (A[5] + Bq[5]);
            }
          auto Tq = [&](int t, int l, int v) { return tilde[((t + g) * 3 + (l - 1)) * 5 + v]; };
          for (int t = -1; t < m; ++t)
            for (int v = 0; v < 5; ++v)
              fc[(t + 1) * 5 + v] = 2 * (central_1[0] * Tq(t, 1, v) + central_1[1] * (Tq(t, 2, v) + Tq(t - 1, 2, v)) +
                                         central_1[2] * (Tq(t, 3, v) + Tq(t - 1, 3, v) + Tq(t - 2, 3, v)));
          for (int t = 0; t < m; ++t) {
            idx[dir] = t;
            const long long q0 = b.id0(idx[0], idx[1], idx[2]);
            for (int v = 0; v < nv; ++v) b.dq[q0 + v * b.n0()] -= fc[(t + 1) * 5 + v] - fc[t * 5 + v];
          }
        }
    }
  }
}

// compute_convective_term_weno_{x,y,z} + compute_weno_flux_cp,
// src/WENO.cu:8-461 (pencil kernels), :1455-1536 (cp flux), :463-491 (launcher).
// dq(i) -= fc(i+1/2) - fc(i-1/2), directions x, y, z in that order.
void inviscid(const Ctx& c, Block& b) {
  if (c.s.inviscid_scheme == 1 || c.s.inviscid_scheme == 2 || c.s.inviscid_scheme == 3 ||
      c.s.inviscid_scheme == 4) {
    inviscid_pv(c, b);
    return;
  }
  if (c.s.inviscid_scheme == 6) {
    inviscid_ep(c, b);
    return;
  }
  const int nv = c.s.n_var, g = b.g;
  const int scheme = c.s.inviscid_scheme;
  const long long s = b.sz;
  const int ext[3] = {b.mx, b.my, b.mz};
  for (int dir = 0; dir < 3; ++dir) {
    const int m = ext[dir];
    const int oa = (dir == 0) ? 1 : 0, ob = (dir == 2) ? 1 : 2;  // the two other axes
    const int na = ext[oa], nb = ext[ob];
#pragma omp parallel
    {
      const int np = m + 2 * g;
      std::vector<real> Fp(np * nv), Fm(np * nv), fc((m + 1) * nv);
      std::vector<geo> kj(np * 3);
      std::vector<real> q(nv + 2);
#pragma omp for collapse(2) schedule(static)
      for (int bb = 0; bb < nb; ++bb)
        for (int aa = 0; aa < na; ++aa) {
          int idx[3];
          idx[oa] = aa;
          idx[ob] = bb;
          // split fluxes at every point of the pencil, ghosts included
          for (int t = -g; t < m + g; ++t) {
            idx[dir] = t;
            const long long p = b.id(idx[0], idx[1], idx[2]);
            for (int l = 0; l < nv; ++l) q[l] = b.cv[p + l * s];
            q[nv] = b.bv[p + 4 * s];
            q[nv + 1] = sqrt(kGammaAir * kRAir * b.bv[p + 5 * s]);
            geo mrow[3];
            for (int cc = 0; cc < 3; ++cc) mrow[cc] = b.metric[p + (3 * dir + cc) * s];
            const geo jac = b.jac[p];
            const int ps = t + g;
            lf_split(q.data(), nv, mrow, jac, &Fp[ps * nv], &Fm[ps * nv]);
            for (int cc = 0; cc < 3; ++cc) kj[ps * 3 + cc] = mrow[cc];
          }
          // faces t+1/2, t = -1 .. m-1
          for (int t = -1; t < m; ++t) {
            const int is = t + g;
            idx[dir] = t;
            const long long p1 = b.id(idx[0], idx[1], idx[2]);
            idx[dir] = t + 1;
            const long long p2 = b.id(idx[0], idx[1], idx[2]);
            const geo jac1{b.jac[p1]}, jac2{b.jac[p2]};
            const geo* m1 = &kj[is * 3];
            const geo* m2 = &kj[(is + 1) * 3];
            if (scheme == 52 || scheme == 72) {
              // compute_weno_flux_ch<Air>, "method 1", src/WENO.cu:1147-1453;
              // 72 projects 7 points and applies WENO-Z7 to the five fields
              // (:1309-1320; the reference's field loop there runs to 7 over
              // its 5x5 matrix and 5-entry fChar, past their ends -- only
              // the five defined fields enter the back projection)
              real cvl[6], cvr[6];
              for (int l = 0; l < 5; ++l) {
                cvl[l] = b.cv[p1 + l * s];
                cvr[l] = b.cv[p2 + l * s];
              }
              cvl[5] = b.bv[p1 + 4 * s];
              cvr[5] = b.bv[p2 + 4 * s];
              CharAir M;
              char_air(cvl, cvr, m1, m2, jac1, jac2, M);
              constexpr double eps{1e-40};
              const geo eps_scaled = eps * c.weno_eps_scale * 0.25 * face_metric_sq(m1, m2, jac1, jac2);
              real fChar[5];
              for (int l = 0; l < 5; ++l) {
                if (scheme == 72) {
                  real vPlus[7] = {}, vMinus[7] = {};
                  for (int r = 0; r < 7; ++r)
                    for (int n = 0; n < 5; ++n) {
                      vPlus[r] += M.L[l][n] * Fp[(is - 3 + r) * nv + n];
                      vMinus[r] += M.L[l][n] * Fm[(is - 2 + r) * nv + n];
                    }
                  fChar[l] = weno7z(vPlus, vMinus, eps_scaled);
                  continue;
                }
                real vPlus[5] = {}, vMinus[5] = {};
                for (int r = 0; r < 5; ++r)
                  for (int n = 0; n < 5; ++n) {
                    vPlus[r] += M.L[l][n] * Fp[(is - 2 + r) * nv + n];
                    vMinus[r] += M.L[l][n] * Fm[(is - 1 + r) * nv + n];
                  }
                fChar[l] = weno5z(vPlus, vMinus, eps_scaled);
              }
              for (int l = 0; l < 5; ++l) {
                real f{0};
                for (int q = 0; q < 5; ++q) f += M.R[l][q] * fChar[q];
                fc[(t + 1) * nv + l] = f;
              }
              continue;
            }
            constexpr double eps{1e-20};
            const geo eps_ref = eps * c.weno_eps_scale * 0.25 * face_metric_sq(m1, m2, jac1, jac2);
            geo eps_scaled[3];
            eps_scaled[0] = eps_ref;
            eps_scaled[1] = eps_ref * c.s.v_ref * c.s.v_ref;
            eps_scaled[2] = eps_scaled[1] * c.s.v_ref * c.s.v_ref;
            for (int l = 0; l < nv; ++l) {
              geo eh{eps_scaled[0]};
              if (l == 1 || l == 2 || l == 3) eh = eps_scaled[1];
              else if (l == 4) eh = eps_scaled[2];
              real f;
              if (scheme == 71) {
                real vp[7], vm[7];
                for (int r = 0; r < 7; ++r) {
                  vp[r] = Fp[(is - 3 + r) * nv + l];
                  vm[r] = Fm[(is - 2 + r) * nv + l];
                }
                f = weno7z(vp, vm, eh);
              } else {
                real vp[5], vm[5];
                for (int r = 0; r < 5; ++r) {
                  vp[r] = Fp[(is - 2 + r) * nv + l];
                  vm[r] = Fm[(is - 1 + r) * nv + l];
                }
                f = (scheme == 50) ? weno5js(vp, vm, eh) : weno5z(vp, vm, eh);
              }
              fc[(t + 1) * nv + l] = f;
            }
          }
          for (int t = 0; t < m; ++t) {
            idx[dir] = t;
            const long long p0 = b.id0(idx[0], idx[1], idx[2]);
            const long long n0 = b.n0();
            for (int l = 0; l < nv; ++l)
              b.dq[p0 + l * n0] -= fc[(t + 1) * nv + l] - fc[t * nv + l];
          }
        }
    }
  }
}

// compute_{fv,gv,hv}_2nd_order<Air, Laminar>, src/ViscousScheme.cuh:23-318,
// :320-616, :618-909; launched per direction by viscous_flux_{fv,gv,hv},
// src/SchemeSelector.cuh:47-154.  dq(i) += fv(i+1/2) - fv(i-1/2).
void viscous_face(const Ctx& c, const Block& b, int dir, int i, int j, int k, real* fv) {
  const long long s = b.sz;
  int o[3] = {0, 0, 0};
  o[dir] = 1;
  const long long p = b.id(i, j, k), p1 = b.id(i + o[0], j + o[1], k + o[2]);
  auto M = [&](long long pt, int r, int cc) { return b.metric[pt + (3 * (r - 1) + (cc - 1)) * s]; };
  const geo xi_x = 0.5 * (M(p, 1, 1) + M(p1, 1, 1));
  const geo xi_y = 0.5 * (M(p, 1, 2) + M(p1, 1, 2));
  const geo xi_z = 0.5 * (M(p, 1, 3) + M(p1, 1, 3));
  const geo eta_x = 0.5 * (M(p, 2, 1) + M(p1, 2, 1));
  const geo eta_y = 0.5 * (M(p, 2, 2) + M(p1, 2, 2));
  const geo eta_z = 0.5 * (M(p, 2, 3) + M(p1, 2, 3));
  const geo zeta_x = 0.5 * (M(p, 3, 1) + M(p1, 3, 1));
  const geo zeta_y = 0.5 * (M(p, 3, 2) + M(p1, 3, 2));
  const geo zeta_z = 0.5 * (M(p, 3, 3) + M(p1, 3, 3));
  auto pv = [&](int ii, int jj, int kk, int l) { return b.bv[b.id(ii, jj, kk) + l * s]; };
  // derivative along the face normal: 2-point; tangential: 4-point average
  real d[4][3];  // [u,v,w,T][xi,eta,zeta]
  const int var[4] = {1, 2, 3, 5};
  for (int q = 0; q < 4; ++q) {
    const int l = var[q];
    if (dir == 0) {
      d[q][0] = pv(i + 1, j, k, l) - pv(i, j, k, l);
      d[q][1] = 0.25 * (pv(i, j + 1, k, l) - pv(i, j - 1, k, l) + pv(i + 1, j + 1, k, l) - pv(i + 1, j - 1, k, l));
      d[q][2] = 0.25 * (pv(i, j, k + 1, l) - pv(i, j, k - 1, l) + pv(i + 1, j, k + 1, l) - pv(i + 1, j, k - 1, l));
    } else if (dir == 1) {
      d[q][0] = 0.25 * (pv(i + 1, j, k, l) - pv(i - 1, j, k, l) + pv(i + 1, j + 1, k, l) - pv(i - 1, j + 1, k, l));
      d[q][1] = pv(i, j + 1, k, l) - pv(i, j, k, l);
      d[q][2] = 0.25 * (pv(i, j, k + 1, l) - pv(i, j, k - 1, l) + pv(i, j + 1, k + 1, l) - pv(i, j + 1, k - 1, l));
    } else {
      d[q][0] = 0.25 * (pv(i + 1, j, k, l) - pv(i - 1, j, k, l) + pv(i + 1, j, k + 1, l) - pv(i - 1, j, k + 1, l));
      d[q][1] = 0.25 * (pv(i, j + 1, k, l) - pv(i, j - 1, k, l) + pv(i, j + 1, k + 1, l) - pv(i, j - 1, k + 1, l));
      d[q][2] = pv(i, j, k + 1, l) - pv(i, j, k, l);
    }
  }
  const real u_x = d[0][0] * xi_x + d[0][1] * eta_x + d[0][2] * zeta_x;
  const real u_y = d[0][0] * xi_y + d[0][1] * eta_y + d[0][2] * zeta_y;
  const real u_z = d[0][0] * xi_z + d[0][1] * eta_z + d[0][2] * zeta_z;
  const real v_x = d[1][0] * xi_x + d[1][1] * eta_x + d[1][2] * zeta_x;
  const real v_y = d[1][0] * xi_y + d[1][1] * eta_y + d[1][2] * zeta_y;
  const real v_z = d[1][0] * xi_z + d[1][1] * eta_z + d[1][2] * zeta_z;
  const real w_x = d[2][0] * xi_x + d[2][1] * eta_x + d[2][2] * zeta_x;
  const real w_y = d[2][0] * xi_y + d[2][1] * eta_y + d[2][2] * zeta_y;
  const real w_z = d[2][0] * xi_z + d[2][1] * eta_z + d[2][2] * zeta_z;
  const real mul = 0.5 * (b.mul[p] + b.mul[p1]);
  const double mut{0};
  const real viscosity = mul + mut;
  const real tau_xx = viscosity * (4 * u_x - 2 * v_y - 2 * w_z) / 3.0;
  const real tau_yy = viscosity * (4 * v_y - 2 * u_x - 2 * w_z) / 3.0;
  const real tau_zz = viscosity * (4 * w_z - 2 * u_x - 2 * v_y) / 3.0;
  const real tau_xy = viscosity * (u_y + v_x);
  const real tau_xz = viscosity * (u_z + w_x);
  const real tau_yz = viscosity * (v_z + w_y);
  const int r = dir + 1;
  const geo kx = 0.5 * (M(p, r, 1) * b.jac[p] + M(p1, r, 1) * b.jac[p1]);
  const geo ky = 0.5 * (M(p, r, 2) * b.jac[p] + M(p1, r, 2) * b.jac[p1]);
  const geo kz = 0.5 * (M(p, r, 3) * b.jac[p] + M(p1, r, 3) * b.jac[p1]);
  fv[0] = 0;
  fv[1] = kx * tau_xx + ky * tau_xy + kz * tau_xz;
  fv[2] = kx * tau_xy + ky * tau_yy + kz * tau_yz;
  fv[3] = kx * tau_xz + ky * tau_yz + kz * tau_zz;
  const real um = 0.5 * (pv(i, j, k, 1) + pv(i + o[0], j + o[1], k + o[2], 1));
  const real vm = 0.5 * (pv(i, j, k, 2) + pv(i + o[0], j + o[1], k + o[2], 2));
  const real wm = 0.5 * (pv(i, j, k, 3) + pv(i + o[0], j + o[1], k + o[2], 3));
  const real t_x = d[3][0] * xi_x + d[3][1] * eta_x + d[3][2] * zeta_x;
  const real t_y = d[3][0] * xi_y + d[3][1] * eta_y + d[3][2] * zeta_y;
  const real t_z = d[3][0] * xi_z + d[3][1] * eta_z + d[3][2] * zeta_z;
  constexpr double cp{kGammaAir * kRu / kMwAir / (kGammaAir - 1)};
  const double prt = 0.9;  // turbulent_prandtl_number default; multiplies mut = 0
  const real conductivity = (mul / c.s.prandtl + mut / prt) * cp;
  fv[4] = um * fv[1] + vm * fv[2] + wm * fv[3] + conductivity * (kx * t_x + ky * t_y + kz * t_z);
}

void viscous(const Ctx& c, Block& b) {
  if (c.s.viscous_order == 0) return;
  const int nv = c.s.n_var;
  const long long n0 = b.n0();
  const int ext[3] = {b.mx, b.my, b.mz};
  for (int dir = 0; dir < 3; ++dir) {
    const int m = ext[dir];
    const int oa = (dir == 0) ? 1 : 0, ob = (dir == 2) ? 1 : 2;
    const int na = ext[oa], nb = ext[ob];
#pragma omp parallel
    {
      std::vector<real> fv((m + 1) * 5);
#pragma omp for collapse(2) schedule(static)
      for (int bb = 0; bb < nb; ++bb)
        for (int aa = 0; aa < na; ++aa) {
          int idx[3];
          idx[oa] = aa;
          idx[ob] = bb;
          for (int t = -1; t < m; ++t) {
            idx[dir] = t;
            viscous_face(c, b, dir, idx[0], idx[1], idx[2], &fv[(t + 1) * 5]);
          }
          for (int t = 0; t < m; ++t) {
            idx[dir] = t;
            const long long p0 = b.id0(idx[0], idx[1], idx[2]);
            for (int l = 0; l < nv; ++l) b.dq[p0 + l * n0] += fv[(t + 1) * 5 + l] - fv[t * 5 + l];
          }
        }
    }
  }
}

// update_cv_and_bv_rk<Air>, src/RK.cuh:230-284
void update(const Ctx& c, Block& b, double dt, int rk) {
  const int nv = c.s.n_var;
  const long long s = b.sz, n0 = b.n0();
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), p0 = b.id0(i, j, k);
        const geo dt_div_jac = dt / b.jac[p];
        for (int l = 0; l < nv; ++l)
          b.cv[p + l * s] = kRkA[rk] * b.qn[p + l * s] + kRkB[rk] * b.cv[p + l * s] +
                            kRkC[rk] * dt_div_jac * b.dq[p0 + l * n0];
        b.bv[p] = b.cv[p];
        const real density_inv = 1.0 / b.cv[p];
        b.bv[p + s] = b.cv[p + s] * density_inv;
        b.bv[p + 2 * s] = b.cv[p + 2 * s] * density_inv;
        b.bv[p + 3 * s] = b.cv[p + 3 * s] * density_inv;
        const real V2 = b.bv[p + s] * b.bv[p + s] + b.bv[p + 2 * s] * b.bv[p + 2 * s] +
                          b.bv[p + 3 * s] * b.bv[p + 3 * s];
        b.bv[p + 4 * s] = (kGammaAir - 1) * (b.cv[p + 4 * s] - 0.5 * b.bv[p] * V2);
        b.bv[p + 5 * s] = b.bv[p + 4 * s] * kMwAir * density_inv / kRu;
      }
}

// apply_outflow / apply_periodic, src/BoundCond.cuh:205-233, :1443-1486.
// Ranges: the face plus ngg on the tangential directions (src/Mesh.cpp:14-32).
void apply_face(Block& b, const Face& f) {
  const int g = b.g;
  const long long s = b.sz;
  const int ext[3] = {b.mx, b.my, b.mz};
  const int *lo = f.lo, *hi = f.hi;
  int dir[3] = {0, 0, 0};
  dir[f.axis] = f.side;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        int other[3] = {i, j, k};
        if (f.type == ORC_BC_PERIODIC) other[f.axis] = (f.side < 0) ? ext[f.axis] - 1 : 0;
        for (int gg = 1; gg <= g; ++gg) {
          const long long pg = b.id(i + gg * dir[0], j + gg * dir[1], k + gg * dir[2]);
          long long ps;
          if (f.type == ORC_BC_PERIODIC)
            ps = b.id(other[0] + gg * dir[0], other[1] + gg * dir[1], other[2] + gg * dir[2]);
          else
            ps = b.id(i, j, k);
          for (int l = 0; l < 6; ++l) b.bv[pg + l * s] = b.bv[ps + l * s];
          cv_from_bv_point(b, pg);
        }
      }
}

// apply_farfield<Air>, src/BoundCond.cuh:755-1022.  Each point of the face
// (ngg points beyond it on the tangential directions) decides from its own
// normal Mach number: supersonic inflow takes the far-field state on the face
// point and its ghosts, supersonic outflow copies the face point to the
// ghosts, subsonic faces take the state from the two Riemann invariants
// (outgoing from the face point, incoming from the far field) with the
// far-field entropy and tangential velocity on inflow, the face point's on
// outflow.
void apply_farfield(Block& b, const Face& f) {
  const int g = b.g;
  const long long s = b.sz;
  const int *lo = f.lo, *hi = f.hi;
  int dir[3] = {0, 0, 0};
  dir[f.axis] = f.side;
  const orc_farfield& ff = f.ff;
  constexpr double gamma_air{kGammaAir};
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        const long long p = b.id(i, j, k);
        geo nx = b.metric[p + (3 * f.axis) * s], ny = b.metric[p + (3 * f.axis + 1) * s],
               nz = b.metric[p + (3 * f.axis + 2) * s];
        const geo grad_n_inv = f.side / sqrt(nx * nx + ny * ny + nz * nz);
        nx *= grad_n_inv;
        ny *= grad_n_inv;
        nz *= grad_n_inv;
        const real u_b{b.bv[p + s]}, v_b{b.bv[p + 2 * s]}, w_b{b.bv[p + 3 * s]};
        const real u_face{nx * u_b + ny * v_b + nz * w_b};
        const real p_b{b.bv[p + 4 * s]}, rho_b{b.bv[p]};
        const real a_b{sqrt(gamma_air * p_b / rho_b)};
        const real mach_b{u_face / a_b};
        real st[6];
        bool face_too = true;
        if (mach_b <= -1) {
          st[0] = ff.density; st[1] = ff.u; st[2] = ff.v; st[3] = ff.w;
          st[4] = ff.pressure; st[5] = ff.temperature;
        } else if (mach_b >= 1) {
          for (int l = 0; l < 6; ++l) st[l] = b.bv[p + l * s];
          face_too = false;
        } else {
          const real riemann_pos{u_face + 2 * a_b / (gamma_air - 1)};
          const geo u_inf{ff.u * nx + ff.v * ny + ff.w * nz};
          const geo riemann_neg{u_inf - 2 * ff.acoustic_speed / (ff.specific_heat_ratio - 1)};
          const real Un{0.5 * (riemann_pos + riemann_neg)};
          const real c_b{0.25 * (gamma_air - 1) * (riemann_pos - riemann_neg)};
          real s_b;
          if (mach_b <= 0) {
            s_b = ff.entropy;
            st[1] = ff.u + (Un - u_inf) * nx;
            st[2] = ff.v + (Un - u_inf) * ny;
            st[3] = ff.w + (Un - u_inf) * nz;
          } else {
            s_b = p_b / pow(rho_b, gamma_air);
            st[1] = u_b + (Un - u_face) * nx;
            st[2] = v_b + (Un - u_face) * ny;
            st[3] = w_b + (Un - u_face) * nz;
          }
          st[0] = pow(c_b * c_b / (gamma_air * s_b), 1 / (gamma_air - 1));
          st[4] = st[0] * c_b * c_b / gamma_air;
          st[5] = st[4] * kMwAir / (st[0] * kRu);
        }
        if (face_too) {
          for (int l = 0; l < 6; ++l) b.bv[p + l * s] = st[l];
          cv_from_bv_point(b, p);
        }
        for (int gg = 1; gg <= g; ++gg) {
          const long long pg = b.id(i + gg * dir[0], j + gg * dir[1], k + gg * dir[2]);
          for (int l = 0; l < 6; ++l) b.bv[pg + l * s] = st[l];
          cv_from_bv_point(b, pg);
        }
      }
}

// apply_inflow<Air>, src/BoundCond.cuh:235-612 (inflow_type 2: the
// mixing-layer profile, :420-482; otherwise the constant state, :483-522),
// fluctuation_type 1 (white noise, the random states indexed by the point's
// position in the padded face plane).  The face point and its ghost layers
// take the same state.
void apply_inflow(Block& b, const Face& f) {
  const orc_inflow& in = f.in;
  const int g = b.g;
  const long long s = b.sz;
  int dir[3] = {0, 0, 0};
  dir[f.axis] = f.side;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = f.lo[2]; k <= f.hi[2]; ++k)
    for (int j = f.lo[1]; j <= f.hi[1]; ++j)
      for (int i = f.lo[0]; i <= f.hi[0]; ++i) {
        const double y = b.y[b.idg(i, j, k)];
        double u;
        bool upper = true;
        if (in.inflow_type == 2) {
          u = 0.5 * (in.upper[1] + in.lower[1]) + 0.5 * (in.upper[1] - in.lower[1]) * std::tanh(2 * y / in.delta_omega);
          upper = y >= 0;
        } else {
          u = in.upper[1];
        }
        const double* S = upper ? in.upper : in.lower;
        const double density = S[0], w = S[3], p = S[4], T = S[5];
        double v = S[2];
        if (in.fluctuation_type == 1) {
          const double vel = sqrt(u * u + v * v + w * w);
          const uint64_t point = f.axis == 0 ? (uint64_t)(k + g) * (b.my + 2 * g) + (j + g)
                                 : f.axis == 1 ? (uint64_t)(k + g) * (b.mx + 2 * g) + (i + g)
                                               : (uint64_t)(j + g) * (b.mx + 2 * g) + (i + g);
          const double rms = in.fluctuation_intensity;
          if (in.inflow_type == 2) {
            if (y < 4 * in.delta_omega && y > -4 * in.delta_omega) {
              double n1, n2;
              orc_inflow_noise::normal_pair(in.seed, b.bc_passes, point, &n1, &n2);
              u += n1 * rms * vel;
              v += n2 * rms * vel;
            }
          } else {
            double n1, n2;
            orc_inflow_noise::normal_pair(in.seed, b.bc_passes, point, &n1, &n2);
            u += n1 * rms * vel;
          }
        }
        for (int gg = 0; gg <= g; ++gg) {
          const long long q = b.id(i + gg * dir[0], j + gg * dir[1], k + gg * dir[2]);
          b.bv[q] = density;
          b.bv[q + s] = u;
          b.bv[q + 2 * s] = v;
          b.bv[q + 3 * s] = w;
          b.bv[q + 4 * s] = p;
          b.bv[q + 5 * s] = T;
          cv_from_bv_point(b, q);
        }
      }
}

// sponge_layer_source<Air>, src/SpongeLayer.cu:68-160, called from
// compute_source before the inviscid terms: sigma from the enabled sides
// (x-, x+, y-, y+, z-, z+ in this order; sponge_function 0), then
// dq -= sigma (cv - mean), sigma times J when jacobian_scaled.
void sponge_source(Block& b) {
  const orc_sponge& sp = b.sp;
  const long long s = b.sz, n0 = (long long)b.mx * b.my * b.mz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), t = ((long long)k * b.my + j) * b.mx + i, q = b.idg(i, j, k);
        const double pos[3] = {b.x[q], b.y[q], b.z[q]};
        geo sigma = 0;
        for (int side = 0; side < 6; ++side) {
          if (!sp.enabled[side]) continue;
          const geo x = pos[side / 2], x1 = sp.start[side];
          if (!(side % 2 == 0 ? x < x1 : x > x1)) continue;
          geo xi = (x - x1) / (sp.end[side] - x1);
          if (xi > 1) xi = 1;
          const geo xi3 = xi * xi * xi;
          sigma += sp.sigma[side] * (6 * xi3 * xi * xi - 15 * xi3 * xi + 10 * xi3);
        }
        if (sp.jacobian_scaled) sigma = sigma * b.jac[p];
        for (int l = 0; l < 5; ++l) b.dq[t + l * n0] -= sigma * (b.cv[p + l * s] - b.sponge_mean_cv[t + l * n0]);
      }
}

// update_sponge_layer_value, src/SpongeLayer.cu:28-49
void sponge_mean(Block& b, unsigned long long iter) {
  const long long s = b.sz, n0 = (long long)b.mx * b.my * b.mz;
  const double n = static_cast<double>(iter);
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), t = ((long long)k * b.my + j) * b.mx + i;
        for (int l = 0; l < 5; ++l) {
          real& m = b.sponge_mean_cv[t + l * n0];
          m = (m * n + b.cv[p + l * s]) / (n + 1);
        }
      }
}

void apply_bc(Block& b) {
  // priority: farfield(4), outflow(6), inflow(5), ..., periodic(8) (src/BoundCond.cuh:1488-1678)
  const int order[] = {ORC_BC_FARFIELD, ORC_BC_OUTFLOW, ORC_BC_INFLOW, ORC_BC_PERIODIC};
  for (int t : order)
    for (const Face& f : b.faces)
      if (f.type == t) {
        if (t == ORC_BC_FARFIELD) apply_farfield(b, f);
        else if (t == ORC_BC_INFLOW) apply_inflow(b, f);
        else apply_face(b, f);
      }
  ++b.bc_passes;
}

// inner_communication, src/DataCommunication.cuh:60-105: the shared node
// layer takes the mean of both blocks (in both), the ghost layers take the
// target's values.
void communicate_face(Ctx& c, Block& b, const InnerFace& f) {
  Block& t = c.blocks[f.target];
  const long long s = b.sz, ts = t.sz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int n2 = 0; n2 < f.n_point[2]; ++n2)
    for (int n1 = 0; n1 < f.n_point[1]; ++n1)
      for (int n0 = 0; n0 < f.n_point[0]; ++n0) {
        const int n[3] = {n0, n1, n2};
        int idx[3], tid[3];
        bool shared;
        orc_faces::point_of(f, n, idx, tid, &shared);
        const long long p = b.id(idx[0], idx[1], idx[2]), pt = t.id(tid[0], tid[1], tid[2]);
        if (shared) {
          for (int l = 0; l < 6; ++l) {
            const real ave{0.5 * (t.bv[pt + l * ts] + b.bv[p + l * s])};
            b.bv[p + l * s] = ave;
            t.bv[pt + l * ts] = ave;
          }
        } else {
          for (int l = 0; l < 6; ++l) b.bv[p + l * s] = t.bv[pt + l * ts];
        }
        cv_from_bv_point(b, p);
      }
}

void communicate(Ctx& c) {
  for (auto& b : c.blocks)
    for (const auto& f : b.inner) communicate_face(c, b, f);
}

// global_time_step, src/TimeAdvanceFunc.cu:39-64 (min_of_arr starts from 1e6)
double global_dt(Ctx& c) {
  double dt{1e6};
  for (auto& b : c.blocks) {
    double m{1e6};  // the tangent build treats dt as a constant
    for (const real& v : b.dtl) m = std::min(m, value_of(v));
    dt = std::min(dt, m);
  }
  return dt;
}

void alloc_block(Block& b, int nv) {
  const int g = b.g;
  b.sx = b.mx + 2 * g;
  b.sy = b.my + 2 * g;
  b.sz = b.sx * b.sy * (b.mz + 2 * g);
  b.cv.assign(b.sz * nv, 0.0);
  b.qn.assign(b.sz * nv, 0.0);
  b.bv.assign(b.sz * 6, 0.0);
  b.mul.assign(b.sz, 0.0);
  b.mach.assign(b.sz, 0.0);
  b.jac.assign(b.sz, 0.0);
  b.metric.assign(b.sz * 9, 0.0);
  b.dq.assign(b.n0() * nv, 0.0);
  b.dtl.assign(b.n0(), 0.0);
  b.bv_last.assign(b.n0() * 4, 0.0);
  const long long gs = (long long)(b.mx + 2 * (g + 1)) * (b.my + 2 * (g + 1)) * (b.mz + 2 * (g + 1));
  b.x.assign(gs, 0.0);
  b.y.assign(gs, 0.0);
  b.z.assign(gs, 0.0);
}

// ------------------------------------------------ dual time stepping + DPLUR
// local_time_step (src/TimeAdvanceFunc.cuh:28-83) keeping the inviscid
// spectral radii for the implicit operator: sr[d] = |U_d| + c |grad xi_d|.
void local_dt_sr(const Ctx& c, Block& b) {
  local_dt(c, b);
  const long long s = b.sz, n0 = b.n0();
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), q = b.id0(i, j, k);
        const geo* M = &b.metric[0];
        auto m = [&](int r, int cc) { return M[p + (3 * (r - 1) + (cc - 1)) * s]; };
        const geo gxi = sqrt(m(1, 1) * m(1, 1) + m(1, 2) * m(1, 2) + m(1, 3) * m(1, 3));
        const geo geta = sqrt(m(2, 1) * m(2, 1) + m(2, 2) * m(2, 2) + m(2, 3) * m(2, 3));
        const geo gzeta = sqrt(m(3, 1) * m(3, 1) + m(3, 2) * m(3, 2) + m(3, 3) * m(3, 3));
        const real u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
        const real U = u * m(1, 1) + v * m(1, 2) + w * m(1, 3);
        const real V = u * m(2, 1) + v * m(2, 2) + w * m(2, 3);
        const real W = u * m(3, 1) + v * m(3, 2) + w * m(3, 3);
        const real a = sqrt(kGammaAir * kRu * b.bv[p + 5 * s] / kMwAir);
        b.sr[q] = abs(U) + a * gxi;
        b.sr[q + n0] = abs(V) + a * geta;
        b.sr[q + 2 * n0] = abs(W) + a * gzeta;
      }
}

// compute_qn_star, src/DualTimeStepping.cu:3-15: qn* = J/(2 dt) (4 q^n - q^(n-1))
void dual_qn_star(const Ctx& c, Block& b, double dt) {
  const int nv = c.s.n_var;
  const long long s = b.sz, n0 = b.n0();
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), q = b.id0(i, j, k);
        const real factor = 0.5 * b.jac[p] / dt;
        for (int l = 0; l < nv; ++l) b.qn_star[q + l * n0] = factor * (4 * b.cv[p + l * s] - b.qn1[p + l * s]);
      }
}

// compute_modified_rhs, src/DualTimeStepping.cu:17-31: dq += qn* - 3J/(2 dt) q
void dual_modified_rhs(const Ctx& c, Block& b, double dt) {
  const int nv = c.s.n_var;
  const long long s = b.sz, n0 = b.n0();
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), q = b.id0(i, j, k);
        const real factor = 1.5 * b.jac[p] / dt;
        for (int l = 0; l < nv; ++l) b.dq[q + l * n0] += b.qn_star[q + l * n0] - factor * b.cv[p + l * s];
      }
}

// The DPLUR diagonal, 1 + diag_factor dt + dt (sr_xi + sr_eta + sr_zeta).  The
// reference adds 2 dt times the viscous spectral radii, an array it
// allocates zeroed and never writes (src/Field.cu:553-566), so that term is 0.
inline real dplur_diag(const Block& b, long long q, double diag_factor) {
  const long long n0 = b.n0();
  const real dtl = b.dtl[q];
  const real visc = 2 * dtl * (0.0 + 0.0 + 0.0);
  return 1 + diag_factor * dtl + dtl * (b.sr[q] + b.sr[q + n0] + b.sr[q + 2 * n0]) + visc;
}

// compute_DQ_0<Air>, src/DPLUR.cuh:14-45: dq = dt dq / diag, then dq0 = dq
void dplur_dq0(const Ctx& c, Block& b, double diag_factor) {
  const int nv = c.s.n_var;
  const long long n0 = b.n0();
  for (long long q = 0; q < n0; ++q) {
    const real dtl = b.dtl[q];
    for (int l = 0; l < nv; ++l) b.dq[q + l * n0] *= dtl;
    const real diag = dplur_diag(b, q, diag_factor);
    for (int l = 0; l < nv; ++l) b.dq[q + l * n0] /= diag;
  }
  b.dq0 = b.dq;
}

// compute_jacobian_times_dq<Air>, src/DPLUR.cuh:98-160: (A_d + lambda) dq at
// one point, A_d the inviscid flux Jacobian along direction d (metric row
// d+1, not Jacobian-weighted), lambda = +-(spectral radius).
void jacobian_times_dq(const geo* k3, const real* pv, real lambda, const real* dq, real* out) {
  const real u = pv[1], v = pv[2], w = pv[3];
  const real U = k3[0] * u + k3[1] * v + k3[2] * w;
  const real lmd1 = U + lambda;
  const real e = 0.5 * (u * u + v * v + w * w);
  const real gamma = kGammaAir;
  const real b3 = 0, b4 = 0;
  const real h = gamma / (gamma - 1) * pv[4] / pv[0] + e;
  const real b1 = k3[0] * dq[1] + k3[1] * dq[2] + k3[2] * dq[3] - U * dq[0];
  const real b2 = (gamma - 1) * (e * dq[0] - u * dq[1] - v * dq[2] - w * dq[3] + dq[4]);
  out[0] = b1 + lmd1 * dq[0];
  out[1] = u * b1 + k3[0] * b2 + lmd1 * dq[1] + k3[0] * (b3 - b4);
  out[2] = v * b1 + k3[1] * b2 + lmd1 * dq[2] + k3[1] * (b3 - b4);
  out[3] = w * b1 + k3[2] * b2 + lmd1 * dq[3] + k3[2] * (b3 - b4);
  out[4] = h * b1 + U * b2 + lmd1 * dq[4] + U * (b3 - b4);
}

// DPLUR_inner_iteration<Air>, src/DPLUR.cuh:162-266: dqk = dq0 + dt/diag *
// sum over the six neighbours inside the block of +-1/2 (A +- lambda) dq.
void dplur_sweep(const Ctx& c, Block& b, double diag_factor) {
  const int nv = c.s.n_var;
  const long long s = b.sz, n0 = b.n0();
  const int ext[3] = {b.mx, b.my, b.mz};
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const int ijk[3] = {i, j, k};
        real total[5] = {0, 0, 0, 0, 0};
        auto neighbour = [&](int d, int side) {
          int n[3] = {i, j, k};
          n[d] += side;
          const long long pn = b.id(n[0], n[1], n[2]), qn = b.id0(n[0], n[1], n[2]);
          geo k3[3];
          for (int cc = 0; cc < 3; ++cc) k3[cc] = b.metric[pn + (3 * d + cc) * s];
          real pv[6], dqn[5], out[5];
          for (int l = 0; l < 6; ++l) pv[l] = b.bv[pn + l * s];
          for (int l = 0; l < nv; ++l) dqn[l] = b.dq[qn + l * n0];
          const real visc = 0.0;  // the zero viscous spectral radius, as above
          const real lambda = side < 0 ? b.sr[qn + d * n0] + 2 * visc : -b.sr[qn + d * n0] - 2 * visc;
          jacobian_times_dq(k3, pv, lambda, dqn, out);
          for (int l = 0; l < nv; ++l) {
            if (side < 0) total[l] += 0.5 * out[l];
            else total[l] -= 0.5 * out[l];
          }
        };
        for (int d = 0; d < 3; ++d)
          if (ijk[d] > 0) neighbour(d, -1);
        for (int d = 0; d < 3; ++d)
          if (ijk[d] != ext[d] - 1) neighbour(d, +1);
        const long long q = b.id0(i, j, k);
        const real dtl = b.dtl[q];
        const real diag = dplur_diag(b, q, diag_factor);
        for (int l = 0; l < nv; ++l) b.dqk[q + l * n0] = b.dq0[q + l * n0] + dtl * total[l] / diag;
      }
}

// update_cv_and_bv<Air>, src/FieldOperation.cuh:188-239: q += dt_local / J dq
void dual_update(const Ctx& c, Block& b) {
  const int nv = c.s.n_var;
  const long long s = b.sz, n0 = b.n0();
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), q = b.id0(i, j, k);
        const real dt_div_jac = b.dtl[q] / b.jac[p];
        for (int l = 0; l < nv; ++l) b.cv[p + l * s] += dt_div_jac * b.dq[q + l * n0];
        b.bv[p] = b.cv[p];
        const real density_inv = 1.0 / b.cv[p];
        b.bv[p + s] = b.cv[p + s] * density_inv;
        b.bv[p + 2 * s] = b.cv[p + 2 * s] * density_inv;
        b.bv[p + 3 * s] = b.cv[p + 3 * s] * density_inv;
        const real V2 = b.bv[p + s] * b.bv[p + s] + b.bv[p + 2 * s] * b.bv[p + 2 * s] +
                        b.bv[p + 3 * s] * b.bv[p + 3 * s];
        b.bv[p + 4 * s] = (kGammaAir - 1) * (b.cv[p + 4 * s] - 0.5 * b.bv[p] * V2);
        b.bv[p + 5 * s] = b.bv[p + 4 * s] * kMwAir * density_inv / kRu;
      }
}

// store_last_iter and compute_square_of_dbv_wrt_last_inner_iter
// (src/DualTimeStepping.cu:137-171), one block: last = (rho, |V|, p, T), then
// last = (new - last)^2, summed into acc.
void last_iter_store(Block& b) {
  const long long s = b.sz, n0 = b.n0();
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), q = b.id0(i, j, k);
        b.last_iter[q] = b.bv[p];
        b.last_iter[q + n0] = sqrt(b.bv[p + s] * b.bv[p + s] + b.bv[p + 2 * s] * b.bv[p + 2 * s] +
                                   b.bv[p + 3 * s] * b.bv[p + 3 * s]);
        b.last_iter[q + 2 * n0] = b.bv[p + 4 * s];
        b.last_iter[q + 3 * n0] = b.bv[p + 5 * s];
      }
}

void last_iter_square(Block& b, real acc[4]) {
  const long long s = b.sz, n0 = b.n0();
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), q = b.id0(i, j, k);
        const real vel = sqrt(b.bv[p + s] * b.bv[p + s] + b.bv[p + 2 * s] * b.bv[p + 2 * s] +
                              b.bv[p + 3 * s] * b.bv[p + 3 * s]);
        const real cur[4] = {b.bv[p], vel, b.bv[p + 4 * s], b.bv[p + 5 * s]};
        for (int l = 0; l < 4; ++l) {
          const real d = cur[l] - b.last_iter[q + l * n0];
          b.last_iter[q + l * n0] = d * d;
          acc[l] += b.last_iter[q + l * n0];
        }
      }
}

// inner_converged, src/DualTimeStepping.cu:33-134 (one rank): the RMS change
// of (rho, |V|, p, T) over the inner iteration, scaled by its value at
// iteration 1; at the last iteration a residual above 1e-3 adds 5 iterations
// (up to max_inner_iteration); below 1e-3 the step's iteration count drops to
// this iteration and the inner loop ends.  Returns 1 to end the loop, -1 on NaN.
int inner_converged(Ctx& c, int iter, int* inner_iter) {
  real acc[4] = {0, 0, 0, 0};
  long long n_total = 0;
  for (auto& b : c.blocks) {
    last_iter_square(b, acc);
    n_total += b.n0();
  }
  double res[4];
  for (int l = 0; l < 4; ++l) res[l] = std::sqrt(value_of(acc[l]) / (double)n_total);
  if (iter == 1)
    for (int l = 0; l < 4; ++l) c.dual_res_scale[l] = res[l] < 1e-20 ? 1e-20 : res[l];
  for (int l = 0; l < 4; ++l) res[l] /= c.dual_res_scale[l];
  double err_max = res[0];
  for (int l = 1; l < 4; ++l)
    if (std::abs(res[l]) > err_max) err_max = res[l];
  if (std::isnan(err_max)) return -1;
  if (iter == *inner_iter) {
    const int max_iter = c.dual.max_inner_iteration;
    if (err_max > 1e-3 && *inner_iter < max_iter) *inner_iter = std::min(*inner_iter + 5, max_iter);
  }
  if (err_max < 1e-3) {
    if (*inner_iter - iter > 0) *inner_iter = iter;
    return 1;
  }
  return 0;
}

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// A whole face as the boundary N.txt record that names it.
void whole_face_record(const Block& b, int axis, int side, int r[6]) {
  const int ext[3] = {b.mx, b.my, b.mz};
  for (int i = 0; i < 3; ++i) {
    r[2 * i] = 0;
    r[2 * i + 1] = ext[i] - 1;
  }
  r[2 * axis] = r[2 * axis + 1] = side < 0 ? 0 : ext[axis] - 1;
}

int add_face(Ctx* c, int block, const int range[6], int type, const orc_farfield* ff) {
  Block& b = c->blocks[block];
  const int ext[3] = {b.mx, b.my, b.mz};
  orc_faces::Boundary r;
  if (!orc_faces::register_boundary(range, c->ngg, ext, &r)) return -1;
  Face f{r.face, r.direction, type, ff ? *ff : orc_farfield{}, {}, {}, {}};
  for (int i = 0; i < 3; ++i) {
    f.lo[i] = r.range_start[i];
    f.hi[i] = r.range_end[i];
  }
  b.faces.push_back(f);
  return 0;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

void* orc_create(const orc_setup* s) {
  auto* c = new Ctx;
  c->s = *s;
  c->ngg = (s->inviscid_scheme == 71 || s->inviscid_scheme == 72) ? 4 : 3;  // src/Parameter.cpp:213-238
  // src/DParameter.cu:26-29
  c->weno_eps_scale = s->rho_ref * s->v_ref * s->rho_ref * s->v_ref;
#ifdef _OPENMP
  if (s->n_threads > 0) omp_set_num_threads(s->n_threads);
#endif
  return c;
}

void orc_destroy(void* ctx) { delete static_cast<Ctx*>(ctx); }

int orc_ngg(void* ctx) { return static_cast<Ctx*>(ctx)->ngg; }

int orc_add_block(void* ctx, int mx, int my, int mz) {
  auto* c = static_cast<Ctx*>(ctx);
  if (mx < 2 || my < 2 || mz < 2) return -1;
  Block b;
  b.mx = mx; b.my = my; b.mz = mz; b.g = c->ngg;
  alloc_block(b, c->s.n_var);
  c->blocks.push_back(std::move(b));
  return (int)c->blocks.size() - 1;
}

int orc_add_boundary(void* ctx, int block, int axis, int side, int type) {
  auto* c = static_cast<Ctx*>(ctx);
  if (block < 0 || block >= (int)c->blocks.size()) return -1;
  int r[6];
  whole_face_record(c->blocks[block], axis, side, r);
  return add_face(c, block, r, type, nullptr);
}

int orc_add_farfield(void* ctx, int block, int axis, int side, const orc_farfield* ff) {
  auto* c = static_cast<Ctx*>(ctx);
  if (block < 0 || block >= (int)c->blocks.size() || !ff) return -1;
  int r[6];
  whole_face_record(c->blocks[block], axis, side, r);
  return add_face(c, block, r, ORC_BC_FARFIELD, ff);
}

int orc_set_sponge(void* ctx, int block, const orc_sponge* sp) {
  auto* c = static_cast<Ctx*>(ctx);
  if (block < 0 || block >= (int)c->blocks.size() || !sp) return -1;
  Block& b = c->blocks[block];
  b.sponge = true;
  b.sp = *sp;
  b.sponge_mean_cv.assign((size_t)5 * b.mx * b.my * b.mz, real{0.0});
  return 0;
}

int orc_source(void* ctx, int block) {
  Block& b = static_cast<Ctx*>(ctx)->blocks[block];
  if (b.sponge) sponge_source(b);
  return 0;
}

int orc_add_inflow(void* ctx, int block, const int range[6], const orc_inflow* in) {
  auto* c = static_cast<Ctx*>(ctx);
  if (block < 0 || block >= (int)c->blocks.size() || !in) return -1;
  if (add_face(c, block, range, ORC_BC_INFLOW, nullptr) != 0) return -1;
  c->blocks[block].faces.back().in = *in;
  return 0;
}

int orc_add_boundary_range(void* ctx, int block, const int range[6], int type, const orc_farfield* ff) {
  auto* c = static_cast<Ctx*>(ctx);
  if (block < 0 || block >= (int)c->blocks.size() || (type == ORC_BC_FARFIELD && !ff)) return -1;
  return add_face(c, block, range, type, ff);
}

// FarField::FarField for air (n_spec = 0, laminar), src/BoundCond.cpp:586-700
int orc_farfield_state_air(const orc_farfield_spec* in, orc_farfield* out) {
  if (!in || !out) return -1;
  orc_farfield f{};
  double mach = in->mach, pressure = in->pressure, temperature = in->temperature, velocity = in->velocity,
         density = in->density;
  double u = in->u, v = in->v, w = in->w;
  const double mw = kMwAir;
  if (temperature < 0) temperature = pressure * mw / (density * kRu);
  const double viscosity = value_of(sutherland(temperature));
  const double gamma = kGammaAir;
  const double acoustic_speed = sqrt(gamma * kRu / mw * temperature);
  if (mach < 0) mach = velocity / acoustic_speed;
  else velocity = mach * acoustic_speed;
  u *= velocity;
  v *= velocity;
  w *= velocity;
  if (pressure < 0) {
    density = viscosity * in->reynolds_number / velocity;
    pressure = density * temperature * kRu / mw;
  }
  if (density < 0) density = pressure * mw / (kRu * temperature);
  f.density = density;
  f.u = u;
  f.v = v;
  f.w = w;
  f.pressure = pressure;
  f.temperature = temperature;
  f.acoustic_speed = acoustic_speed;
  f.entropy = pressure / std::pow(density, gamma);
  f.specific_heat_ratio = gamma;
  f.mw = mw;
  *out = f;
  return 0;
}

int orc_add_inner_face_range(void* ctx, int block, const int range[6], const int target_range[6], int target) {
  auto* c = static_cast<Ctx*>(ctx);
  if (block < 0 || block >= (int)c->blocks.size() || target < 0 || target >= (int)c->blocks.size()) return -1;
  const Block &b = c->blocks[block], &t = c->blocks[target];
  const int ext[3] = {b.mx, b.my, b.mz}, text[3] = {t.mx, t.my, t.mz};
  InnerFace f;
  if (!orc_faces::register_inner(range, target_range, c->ngg, target, &f)) return -1;
  if (!orc_faces::inner_in_bounds(f, ext, text, c->ngg)) return -1;
  c->blocks[block].inner.push_back(f);
  return 0;
}

int orc_add_inner_face(void* ctx, int block, int axis, int side, int target) {
  auto* c = static_cast<Ctx*>(ctx);
  if (block < 0 || block >= (int)c->blocks.size() || target < 0 || target >= (int)c->blocks.size()) return -1;
  const Block &b = c->blocks[block], &t = c->blocks[target];
  const int ext[3] = {b.mx, b.my, b.mz}, text[3] = {t.mx, t.my, t.mz};
  int src[6], tgt[6];
  orc_faces::aligned_records(axis, side, ext, text, src, tgt);
  return orc_add_inner_face_range(ctx, block, src, tgt, target);
}

double* orc_array(void* ctx, int block, int which, long long* n_out) {
  auto* c = static_cast<Ctx*>(ctx);
  Block& b = c->blocks[block];
  std::vector<real>* v = nullptr;
  std::vector<geo>* g = nullptr;
  std::vector<double>* xyz = nullptr;
  switch (which) {
    case ORC_CV: v = &b.cv; break;
    case ORC_BV: v = &b.bv; break;
    case ORC_QN: v = &b.qn; break;
    case ORC_DQ: v = &b.dq; break;
    case ORC_MUL: v = &b.mul; break;
    case ORC_MACH: v = &b.mach; break;
    case ORC_JAC: g = &b.jac; break;
    case ORC_METRIC: g = &b.metric; break;
    case ORC_DT_LOCAL: v = &b.dtl; break;
    case ORC_X: xyz = &b.x; break;
    case ORC_Y: xyz = &b.y; break;
    case ORC_Z: xyz = &b.z; break;
    case ORC_BV_LAST: v = &b.bv_last; break;
    default: return nullptr;
  }
  if (xyz) {
    if (n_out) *n_out = (long long)xyz->size();
    return xyz->data();
  }
  if (g) {  // geometry: double, or a counted scalar of the same layout
    static_assert(sizeof(geo) == sizeof(double), "geometry layout");
    if (n_out) *n_out = (long long)g->size();
    return reinterpret_cast<double*>(g->data());
  }
  // state arrays: kWidth doubles per value ({value, tangent} in the tangent build)
  if (n_out) *n_out = (long long)v->size() * kWidth;
  return reinterpret_cast<double*>(v->data());
}

// Analytic uniform Cartesian grid over all ngg+1 ghost layers.  The reference
// mirrors interior nodes into the ghost layers (src/Mesh.cpp:893-1031); on a
// uniform grid that gives the same coordinates up to rounding.
int orc_uniform_grid(void* ctx, int block, double x0, double y0, double z0, double dx, double dy, double dz) {
  auto* c = static_cast<Ctx*>(ctx);
  Block& b = c->blocks[block];
  const int G = b.g + 1;
  for (int k = -G; k < b.mz + G; ++k)
    for (int j = -G; j < b.my + G; ++j)
      for (int i = -G; i < b.mx + G; ++i) {
        const long long p = b.idg(i, j, k);
        b.x[p] = x0 + i * dx;
        b.y[p] = y0 + j * dy;
        b.z[p] = z0 + k * dz;
      }
  return 0;
}

// Block::compute_jac_metric, src/Mesh.cpp:241-300 (the interior formula; the
// ghost-layer loops at :300-502 use the same expressions).
int orc_compute_metric(void* ctx, int block) {
  auto* c = static_cast<Ctx*>(ctx);
  Block& b = c->blocks[block];
  const int g = b.g;
  const long long s = b.sz;
  for (int k = -g; k < b.mz + g; ++k)
    for (int j = -g; j < b.my + g; ++j)
      for (int i = -g; i < b.mx + g; ++i) {
        auto X = [&](int ii, int jj, int kk) { return b.x[b.idg(ii, jj, kk)]; };
        auto Y = [&](int ii, int jj, int kk) { return b.y[b.idg(ii, jj, kk)]; };
        auto Z = [&](int ii, int jj, int kk) { return b.z[b.idg(ii, jj, kk)]; };
        const geo dxd1 = (X(i + 1, j, k) - X(i - 1, j, k)) * 0.5, dyd1 = (Y(i + 1, j, k) - Y(i - 1, j, k)) * 0.5,
                     dzd1 = (Z(i + 1, j, k) - Z(i - 1, j, k)) * 0.5;
        const geo dxd2 = (X(i, j + 1, k) - X(i, j - 1, k)) * 0.5, dyd2 = (Y(i, j + 1, k) - Y(i, j - 1, k)) * 0.5,
                     dzd2 = (Z(i, j + 1, k) - Z(i, j - 1, k)) * 0.5;
        const geo dxd3 = (X(i, j, k + 1) - X(i, j, k - 1)) * 0.5, dyd3 = (Y(i, j, k + 1) - Y(i, j, k - 1)) * 0.5,
                     dzd3 = (Z(i, j, k + 1) - Z(i, j, k - 1)) * 0.5;
        const geo jac = dxd1 * dyd2 * dzd3 + dxd2 * dyd3 * dzd1 + dxd3 * dyd1 * dzd2 - dxd1 * dyd3 * dzd2 -
                           dxd2 * dyd1 * dzd3 - dxd3 * dyd2 * dzd1;
        if (jac <= 0) return -2;
        const long long p = b.id(i, j, k);
        b.jac[p] = jac;
        geo* M = b.metric.data();
        M[p + 0 * s] = (dyd2 * dzd3 - dyd3 * dzd2) / jac;
        M[p + 1 * s] = (dxd3 * dzd2 - dxd2 * dzd3) / jac;
        M[p + 2 * s] = (dxd2 * dyd3 - dxd3 * dyd2) / jac;
        M[p + 3 * s] = (dyd3 * dzd1 - dyd1 * dzd3) / jac;
        M[p + 4 * s] = (dxd1 * dzd3 - dxd3 * dzd1) / jac;
        M[p + 5 * s] = (dxd3 * dyd1 - dxd1 * dyd3) / jac;
        M[p + 6 * s] = (dyd1 * dzd2 - dyd2 * dzd1) / jac;
        M[p + 7 * s] = (dxd2 * dzd1 - dxd1 * dzd2) / jac;
        M[p + 8 * s] = (dxd1 * dyd2 - dxd2 * dyd1) / jac;
      }
  return 0;
}

double orc_sutherland(double t) { return value_of(sutherland(t)); }

// get_mixing_layer_info for air (n_spec = 0), src/MixingLayer.cu:9-190.
// Only the branches the configurations use: velocity_ratio <= 0 with the upper
// Mach number given (:104-107 for upper_faster, :108-110 otherwise).
void orc_mixing_layer_streams(double ma_c, double ma_upper, double t_upper, double t_lower, double p_upper,
                              double p_lower, int upper_faster, orc_ml_streams* out) {
  const double c_upper = sqrt(kGammaAir * kRu / kMwAir * t_upper);
  const double c_lower = sqrt(kGammaAir * kRu / kMwAir * t_lower);
  const double rho_upper = p_upper * kMwAir / (kRu * t_upper);
  const double rho_lower = p_lower * kMwAir / (kRu * t_lower);
  double ma_lower;
  if (upper_faster) ma_lower = (ma_upper - ma_c) * c_upper / c_lower - ma_c;
  else ma_lower = (ma_c + ma_upper) * c_upper / c_lower + ma_c;
  const double u_upper{ma_upper * c_upper}, u_lower{ma_lower * c_lower};
  out->upper[0] = rho_upper; out->upper[1] = u_upper; out->upper[2] = 0; out->upper[3] = 0;
  out->upper[4] = p_upper; out->upper[5] = t_upper;
  out->lower[0] = rho_lower; out->lower[1] = u_lower; out->lower[2] = 0; out->lower[3] = 0;
  out->lower[4] = p_lower; out->lower[5] = t_lower;
}

// initialize_mixing_layer_with_info, src/Initialize.cu:259-325 (air), on the
// padded domain, plus a seeded velocity perturbation that the reference does
// not have: amplitude * U(-1,1) from SplitMix64 of (seed, global node, comp).
// Global node indices wrap with the given periods (<=0: no wrap).
int orc_mixing_layer_ic(void* ctx, int block, const orc_ml_streams* st, double delta_omega, int gi0, int gj0, int gk0,
                        int period_i, int period_j, int period_k, uint64_t seed, double amplitude) {
  auto* c = static_cast<Ctx*>(ctx);
  Block& b = c->blocks[block];
  const int g = b.g;
  const long long s = b.sz;
  auto wrap = [](long long v, int per) -> long long {
    if (per <= 0) return v;
    long long r = v % per;
    return r < 0 ? r + per : r;
  };
  for (int k = -g; k < b.mz + g; ++k)
    for (int j = -g; j < b.my + g; ++j)
      for (int i = -g; i < b.mx + g; ++i) {
        const long long p = b.id(i, j, k);
        const double y = b.y[b.idg(i, j, k)];
        const double u_upper = st->upper[1], u_lower = st->lower[1];
        b.bv[p + s] = 0.5 * (u_upper + u_lower) + 0.5 * (u_upper - u_lower) * std::tanh(2 * y / delta_omega);
        const double* var = (y >= 0) ? st->upper : st->lower;
        b.bv[p] = var[0];
        b.bv[p + 2 * s] = var[2];
        b.bv[p + 3 * s] = var[3];
        b.bv[p + 4 * s] = var[4];
        b.bv[p + 5 * s] = var[5];
        if (amplitude != 0) {
          const long long gi = wrap((long long)gi0 + i, period_i), gj = wrap((long long)gj0 + j, period_j),
                          gk = wrap((long long)gk0 + k, period_k);
          const uint64_t node = (uint64_t)((gk * 1000003LL + gj) * 1000033LL + gi);
          for (int comp = 0; comp < 3; ++comp) {
            const uint64_t h = splitmix64(seed ^ splitmix64(node * 4 + (uint64_t)comp));
            const double r = (double)(h >> 11) * (1.0 / 9007199254740992.0);
            b.bv[p + (1 + comp) * s] += amplitude * (2.0 * r - 1.0);
          }
        }
      }
  return 0;
}

// Fourier-mode velocity perturbation of the differentiation workload (no
// counterpart in the reference): bv[1 + comp_m] += a_m * phi_m over the padded
// domain, phi_m = du * exp(-((y - yc)/delta)^2) * trig_m(2 pi (kx_m (x - x0)/lx +
// kz_m (z - z0)/lz)), trig = sin (0) or cos (1).  In the tangent build the
// amplitudes carry the tangents amp_dot.
int orc_fourier_perturbation(void* ctx, int block, const orc_fourier* f) {
  auto* c = static_cast<Ctx*>(ctx);
  Block& b = c->blocks[block];
  const int g = b.g;
  const long long s = b.sz;
  const double two_pi = 6.283185307179586;
  for (int m = 0; m < f->n_modes; ++m)
    if (f->comp[m] < 0 || f->comp[m] > 2) return -1;
  for (int k = -g; k < b.mz + g; ++k)
    for (int j = -g; j < b.my + g; ++j)
      for (int i = -g; i < b.mx + g; ++i) {
        const long long p = b.id(i, j, k), q = b.idg(i, j, k);
        const double x = b.x[q], y = b.y[q], z = b.z[q];
        const double yr = (y - f->yc) / f->delta;
        const double env = f->du * std::exp(-yr * yr);
        for (int m = 0; m < f->n_modes; ++m) {
          const double ph = two_pi * (f->kx[m] * (x - f->x0) / f->lx + f->kz[m] * (z - f->z0) / f->lz);
          const double phi = env * (f->trig[m] ? std::cos(ph) : std::sin(ph));
#ifdef ORC_DUAL
          const real a(f->amp[m], f->amp_dot ? f->amp_dot[m] : 0.0);
#else
          const real a = f->amp[m];
#endif
          b.bv[p + (1 + f->comp[m]) * s] += a * phi;
        }
      }
  return 0;
}

// Momentum thickness of the differentiation workload (no counterpart in the
// reference): theta = sum_j w_j rho_j (u1 - U_j)(U_j - u2) / (rho0 du^2), where
// rho_j = <rho>, U_j = <rho u>/<rho> average over the nodes i < ni, k < nk of
// row j, and w_j = dy (dy/2 at j = 0 and my-1, the trapezoid rule).
double orc_momentum_thickness(void* ctx, int block, double u1, double u2, double rho0, double du, double dy,
                              int ni, int nk, double* dtheta) {
  auto* c = static_cast<Ctx*>(ctx);
  const Block& b = c->blocks[block];
  const long long s = b.sz;
  real theta = 0.0;
  for (int j = 0; j < b.my; ++j) {
    real s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < nk; ++k)
      for (int i = 0; i < ni; ++i) {
        const long long p = b.id(i, j, k);
        s0 += b.cv[p];
        s1 += b.cv[p + s];
      }
    const real rho = s0 / (double)((long long)ni * nk);
    const real U = s1 / s0;
    const double w = (j == 0 || j == b.my - 1) ? 0.5 * dy : dy;
    theta += w * rho * (u1 - U) * (U - u2) / (rho0 * du * du);
  }
  if (dtheta) *dtheta = tangent_of(theta);
  return value_of(theta);
}

int orc_cv_from_bv(void* ctx, int block) { cv_from_bv(static_cast<Ctx*>(ctx)->blocks[block]); return 0; }

int orc_copy_qn(void* ctx, int block) {
  Block& b = static_cast<Ctx*>(ctx)->blocks[block];
  b.qn = b.cv;
  return 0;
}

int orc_zero_dq(void* ctx, int block) {
  Block& b = static_cast<Ctx*>(ctx)->blocks[block];
  std::fill(b.dq.begin(), b.dq.end(), 0.0);
  return 0;
}

int orc_inviscid(void* ctx, int block) {
  auto* c = static_cast<Ctx*>(ctx);
  inviscid(*c, c->blocks[block]);
  return 0;
}

int orc_viscous(void* ctx, int block) {
  auto* c = static_cast<Ctx*>(ctx);
  viscous(*c, c->blocks[block]);
  return 0;
}

int orc_local_dt(void* ctx, int block) {
  auto* c = static_cast<Ctx*>(ctx);
  local_dt(*c, c->blocks[block]);
  return 0;
}

double orc_global_dt(void* ctx) { return global_dt(*static_cast<Ctx*>(ctx)); }

int orc_update(void* ctx, int block, double dt, int rk) {
  auto* c = static_cast<Ctx*>(ctx);
  update(*c, c->blocks[block], dt, rk);
  return 0;
}

int orc_apply_bc(void* ctx, int block) {
  apply_bc(static_cast<Ctx*>(ctx)->blocks[block]);
  return 0;
}

int orc_communicate(void* ctx) {
  communicate(*static_cast<Ctx*>(ctx));
  return 0;
}

int orc_properties(void* ctx, int block) {
  properties(static_cast<Ctx*>(ctx)->blocks[block]);
  return 0;
}

// Driver::initialize_computation (src/Driver.cu:115-199) then the RK3
// prologue compute_cv_from_bv (src/RK.cuh:38-49).
int orc_initialize(void* ctx) {
  auto* c = static_cast<Ctx*>(ctx);
  for (auto& b : c->blocks) apply_bc(b);
  communicate(*c);
  for (auto& b : c->blocks) properties(b);
  for (auto& b : c->blocks) cv_from_bv(b);
  return 0;
}

// One SSP-RK3 step, src/RK.cuh:85-226 (no source term for air/laminar).
int orc_step(void* ctx, int n_steps, double* last_dt) {
  auto* c = static_cast<Ctx*>(ctx);
  double dt = c->s.fixed_time_step ? c->s.dt : 1e6;
  for (int st = 0; st < n_steps; ++st) {
    for (auto& b : c->blocks) b.qn = b.cv;
    bool any_sponge = false;
    for (auto& b : c->blocks)
      if (b.sponge) {
        sponge_mean(b, c->sponge_iter);
        any_sponge = true;
      }
    if (any_sponge) ++c->sponge_iter;
    for (int rk = 0; rk < 3; ++rk) {
      for (auto& b : c->blocks) std::fill(b.dq.begin(), b.dq.end(), 0.0);
      if (!c->s.fixed_time_step && rk == 0) {
        for (auto& b : c->blocks) local_dt(*c, b);
        dt = global_dt(*c);
      }
      for (auto& b : c->blocks) {
        if (b.sponge) sponge_source(b);
        inviscid(*c, b);
        viscous(*c, b);
        update(*c, b, dt, rk);
        apply_bc(b);
      }
      communicate(*c);
      for (auto& b : c->blocks) properties(b);
    }
  }
  c->dt_cur = dt;
  if (last_dt) *last_dt = dt;
  return 0;
}

// dual_time_stepping's set-up (src/DualTimeStepping.cuh:33-76): the
// settings, and q^(n-1) = the current cv.
int orc_dual_time_init(void* ctx, const orc_dual_time* d) {
  auto* c = static_cast<Ctx*>(ctx);
  if (!d || !(d->dt > 0) || d->inner_iteration < 1 || d->dplur_inner_step < 0) return -1;
  c->dual = *d;
  const int nv = c->s.n_var;
  for (auto& b : c->blocks) {
    const long long n0 = b.n0();
    b.qn1 = b.cv;
    b.qn_star.assign(n0 * nv, 0.0);
    b.sr.assign(n0 * 3, 0.0);
    b.dq0.assign(n0 * nv, 0.0);
    b.dqk.assign(n0 * nv, 0.0);
    b.last_iter.assign(n0 * 4, 0.0);
  }
  c->dual_ready = true;
  return 0;
}

// Physical steps of dual_time_stepping with the DPLUR inner solver
// (src/DualTimeStepping.cuh:33-236, dual_time_stepping_implicit_treat,
// src/DPLUR.cuh:271-327), air, laminar.  Returns the number of inner
// iterations the next step will take; -1 on a NaN residual.
int orc_dual_time_step(void* ctx, int n_steps, int* inner_iteration) {
  auto* c = static_cast<Ctx*>(ctx);
  if (!c->dual_ready) return -2;
  const double dt = c->dual.dt;
  const double diag_factor = 1.5 / dt;
  for (int st = 0; st < n_steps; ++st) {
    const int step = ++c->dual.step;
    for (auto& b : c->blocks) {
      dual_qn_star(*c, b, dt);
      b.qn1 = b.cv;
    }
    // (the reference's dual time loop does not update the sponge layer's
    // running mean; compute_source reads it as it stands)
    const bool monitor = c->dual.iteration_adjust_step > 0 && step % c->dual.iteration_adjust_step == 0;
    for (int iter = 1; iter <= c->dual.inner_iteration; ++iter) {
      for (auto& b : c->blocks) {
        if (monitor) last_iter_store(b);
        std::fill(b.dq.begin(), b.dq.end(), 0.0);
        if (b.sponge) sponge_source(b);
        inviscid(*c, b);
        viscous(*c, b);
        local_dt_sr(*c, b);
        dual_modified_rhs(*c, b, dt);
        dplur_dq0(*c, b, diag_factor);
        for (int it = 0; it < c->dual.dplur_inner_step; ++it) {
          dplur_sweep(*c, b, diag_factor);
          b.dq = b.dqk;
        }
        const long long n0 = b.n0();
        for (int l = 0; l < c->s.n_var; ++l)
          for (long long q = 0; q < n0; ++q) b.dq[q + l * n0] /= b.dtl[q];
        dual_update(*c, b);
        apply_bc(b);
      }
      communicate(*c);
      for (auto& b : c->blocks) properties(b);
      if (monitor) {
        const int r = inner_converged(*c, iter, &c->dual.inner_iteration);
        if (r < 0) return -1;
        if (r > 0) break;
      }
    }
  }
  if (inner_iteration) *inner_iteration = c->dual.inner_iteration;
  return c->dual.inner_iteration;
}

// compute_jacobian_times_dq<Air> at one point (test entry point): pv =
// (rho, u, v, w, p, T), k3 the metric row, lambda the signed spectral term.
void orc_jacobian_times_dq(const double* k3, const double* pv, double lambda, const double* dq, double* out) {
  geo k[3] = {k3[0], k3[1], k3[2]};
  real p[6], d[5], o[5];
  for (int l = 0; l < 6; ++l) p[l] = pv[l];
  for (int l = 0; l < 5; ++l) d[l] = dq[l];
  jacobian_times_dq(k, p, lambda, d, o);
  for (int l = 0; l < 5; ++l) out[l] = value_of(o[l]);
}

// store_last_step, src/TimeAdvanceFunc.cu:6-19
int orc_store_last_step(void* ctx) {
  auto* c = static_cast<Ctx*>(ctx);
  for (auto& b : c->blocks) {
    const long long s = b.sz, n0 = b.n0();
    for (int k = 0; k < b.mz; ++k)
      for (int j = 0; j < b.my; ++j)
        for (int i = 0; i < b.mx; ++i) {
          const long long p = b.id(i, j, k), q = b.id0(i, j, k);
          const real u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
          b.bv_last[q] = b.bv[p];
          b.bv_last[q + n0] = sqrt(u * u + v * v + w * w);
          b.bv_last[q + 2 * n0] = b.bv[p + 4 * s];
          b.bv_last[q + 3 * n0] = b.bv[p + 5 * s];
        }
  }
  return 0;
}

// compute_square_of_dbv + reduction + sqrt(sum/N), src/TimeAdvanceFunc.cu:21-37,
// src/Residual.cuh:75-122 (unscaled: res_scale division is left to the caller).
int orc_residual(void* ctx, double res[4]) {
  auto* c = static_cast<Ctx*>(ctx);
  real acc[4] = {0, 0, 0, 0};
  long long n_total = 0;
  for (auto& b : c->blocks) {
    const long long s = b.sz, n0 = b.n0();
    n_total += n0;
    for (int k = 0; k < b.mz; ++k)
      for (int j = 0; j < b.my; ++j)
        for (int i = 0; i < b.mx; ++i) {
          const long long p = b.id(i, j, k), q = b.id0(i, j, k);
          const real u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
          const real vel = sqrt(u * u + v * v + w * w);
          acc[0] += (b.bv[p] - b.bv_last[q]) * (b.bv[p] - b.bv_last[q]);
          acc[1] += (vel - b.bv_last[q + n0]) * (vel - b.bv_last[q + n0]);
          acc[2] += (b.bv[p + 4 * s] - b.bv_last[q + 2 * n0]) * (b.bv[p + 4 * s] - b.bv_last[q + 2 * n0]);
          acc[3] += (b.bv[p + 5 * s] - b.bv_last[q + 3 * n0]) * (b.bv[p + 5 * s] - b.bv_last[q + 3 * n0]);
        }
  }
  for (int l = 0; l < 4; ++l) res[l] = std::sqrt(value_of(acc[l]) / (double)n_total);
  return 0;
}

// point-function entry points of the tests (double in, double out)
double orc_weno5z(const double* vp, const double* vm, double eps) {
  real a[5], b[5];
  std::copy(vp, vp + 5, a);
  std::copy(vm, vm + 5, b);
  return value_of(weno5z(a, b, eps));
}
double orc_weno7z(const double* vp, const double* vm, double eps) {
  real a[7], b[7];
  std::copy(vp, vp + 7, a);
  std::copy(vm, vm + 7, b);
  return value_of(weno7z(a, b, eps));
}
double orc_weno5js(const double* vp, const double* vm, double eps) {
  real a[5], b[5];
  std::copy(vp, vp + 5, a);
  std::copy(vm, vm + 5, b);
  return value_of(weno5js(a, b, eps));
}
int orc_riemann(int scheme, const double* ql, const double* qr, const double* kl, const double* kr, double jl,
                double jr, double* fc) {
  real pl[6], pr[6], f[5];
  real pv[4 * 5];
  for (int l = 0; l < 5; ++l) {
    pv[5 + l] = ql[l];
    pv[10 + l] = qr[l];
    pv[l] = ql[l];
    pv[15 + l] = qr[l];
  }
  reconstruct_air(pv, 1, 1, pl, pr);
  const geo ml[3] = {kl[0], kl[1], kl[2]}, mr[3] = {kr[0], kr[1], kr[2]};
  riemann_air(scheme, pl, pr, ml, mr, jl, jr, f);
  for (int l = 0; l < 5; ++l) fc[l] = value_of(f[l]);
  return 0;
}

int orc_char_matrices(const double* ql, const double* qr, const double* kl, const double* kr, double jl, double jr,
                      double* L, double* R) {
  real a[6], b[6];
  for (int l = 0; l < 6; ++l) {
    a[l] = ql[l];
    b[l] = qr[l];
  }
  CharAir M;
  const geo ml[3] = {kl[0], kl[1], kl[2]}, mr[3] = {kr[0], kr[1], kr[2]};
  char_air(a, b, ml, mr, jl, jr, M);
  for (int r = 0; r < 5; ++r)
    for (int q = 0; q < 5; ++q) {
      L[r * 5 + q] = value_of(M.L[r][q]);
      R[r * 5 + q] = value_of(M.R[r][q]);
    }
  return 0;
}

int orc_roe(const double* ql, const double* qr, const double* kl, const double* kr, double jl, double jr, double dl,
            double dr, double* f) {
  real a[6], b[6], out[5];
  for (int l = 0; l < 6; ++l) {
    a[l] = ql[l];
    b[l] = qr[l];
  }
  const geo ml[3] = {kl[0], kl[1], kl[2]}, mr[3] = {kr[0], kr[1], kr[2]};
  roe_air(a, b, ml, mr, jl, jr, dl, dr, out);
  for (int l = 0; l < 5; ++l) f[l] = value_of(out[l]);
  return 0;
}

void orc_lf_split(const double* q, int n_var, const double* k3, double jac, double* fp, double* fm) {
  std::vector<real> qq(q, q + n_var + 2), p(n_var), m(n_var);
  const geo k[3] = {k3[0], k3[1], k3[2]};
  lf_split(qq.data(), n_var, k, jac, p.data(), m.data());
  for (int l = 0; l < n_var; ++l) {
    fp[l] = value_of(p[l]);
    fm[l] = value_of(m[l]);
  }
}

#ifdef ORC_COUNT
// Operation counts since the last reset: additions and subtractions,
// multiplications, divisions, square roots, transcendental calls.
void orc_op_counts(unsigned long long out[5]) {
  out[0] = orc::g_ops.add;
  out[1] = orc::g_ops.mul;
  out[2] = orc::g_ops.div;
  out[3] = orc::g_ops.sqrt;
  out[4] = orc::g_ops.trans;
}
void orc_op_counts_reset(void) { orc::g_ops = orc::OpCounts{}; }
#endif

int orc_scalar_width(void) { return kWidth; }

}  // extern "C"
