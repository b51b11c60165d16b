// ostc_oracle_fr.cpp -- CPU oracle for the finite-rate mixture path.
//
// TEST INFRASTRUCTURE ONLY: a host-C++ restatement, OpenMP over cells, of
// the reference's MixtureModel::FR routines listed beside each function,
// built with -ffp-contract=off so every product and sum is rounded
// separately, in the written order.  The HIP library never links or calls
// it; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
//
// Parity pins (tests/test_oracle_fr.py): NASA-7 identities (dh/dT = cp,
// h continuity at T_mid, g/RT = h/RT - s/R evaluated independently),
// chemistry mass and element conservation, equilibrium constants against an
// independent 50-digit evaluation, detailed balance at a computed
// equilibrium state, Troe limits, pure-species transport limits, T-Newton
// inversion of e(T), characteristic L R = I, free-stream preservation and
// telescoping conservation of the whole mixture RHS, and an independent
// numpy restatement of the chemistry source and transport.
//
// The characteristic WENO-Z5 (scheme 52) of the five acoustic/entropy fields
// has two stencils (setup key ch_stencil).  0: i-2..i+2 / i-1..i+3, the
// stencil of the species fields, of the air specialisation and of WENO-Z7.
// 1: the mixture code of the reference, which reads i-3..i+1 / i-2..i+2 for
// the five conservative components of those fields (src/WENO.cu:880-881); at
// the first face of every pencil the reference's point i-3 lies outside its
// shared-memory arrays, and at the first face of a block it lies beyond the
// ghost layers: there the outermost ghost point stands in for it.
#include "ostc_oracle_fr.h"
#include "ostc_oracle_faces.h"
#include "ostc_oracle_inflow.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using std::abs;
using std::exp;
using std::log;
using std::pow;
using std::sqrt;

// src/Constants.h:17-34
constexpr double kRu = 8314.462618;
constexpr double kPatm = 101325;
// src/RK.cuh:10-14
constexpr double kRkA[3] = {1.0, 0.75, 1.0 / 3.0};
constexpr double kRkB[3] = {0.0, 0.25, 2.0 / 3.0};
constexpr double kRkC[3] = {1.0, 0.25, 2.0 / 3.0};
constexpr int MS = OSTC_MAX_SPEC;
constexpr int MR = OSTC_MAX_REAC;

using Tab = ostc_mech_tables;

// ------------------------------------------------------------- thermodynamics
// compute_enthalpy (Combustion2Part), src/Thermo.cu:130-148
void enthalpy(const Tab& T, double t, double* h) {
  const double t2{t * t}, t3{t2 * t}, t4{t3 * t}, t5{t4 * t};
  for (int i = 0; i < T.n_spec; ++i) {
    if (t < T.t_low[i]) {
      const double tt = T.t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double* c = T.low_coeff[i];
      h[i] = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      const double cp = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      h[i] += cp * (t - tt);
    } else {
      const double* c = t < T.t_mid[i] ? T.low_coeff[i] : T.high_coeff[i];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
    }
    h[i] *= kRu / T.mw[i];
  }
}

// compute_enthalpy_and_cp (Combustion2Part), src/Thermo.cu:150-172
void enthalpy_cp(const Tab& T, double t, double* h, double* cp) {
  const double t2{t * t}, t3{t2 * t}, t4{t3 * t}, t5{t4 * t};
  for (int i = 0; i < T.n_spec; ++i) {
    if (t < T.t_low[i]) {
      const double tt = T.t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double* c = T.low_coeff[i];
      h[i] = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      h[i] += cp[i] * (t - tt);
    } else {
      const double* c = t < T.t_mid[i] ? T.low_coeff[i] : T.high_coeff[i];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    h[i] *= kRu / T.mw[i];
    cp[i] *= kRu / T.mw[i];
  }
}

// compute_cp (Combustion2Part), src/Thermo.cu:174-188
void heat_capacity(const Tab& T, double t, double* cp) {
  const double t2{t * t}, t3{t2 * t}, t4{t3 * t};
  for (int i = 0; i < T.n_spec; ++i) {
    if (t < T.t_low[i]) {
      const double tt = T.t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
      const double* c = T.low_coeff[i];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
    } else {
      const double* c = t < T.t_mid[i] ? T.low_coeff[i] : T.high_coeff[i];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    cp[i] *= kRu / T.mw[i];
  }
}

// compute_gibbs_div_rt (Combustion2Part), src/Thermo.cu:190-208
void gibbs_rt(const Tab& T, double t, double* g) {
  const double t2{t * t}, t3{t2 * t}, t4{t3 * t}, t_inv{1 / t}, log_t{log(t)};
  for (int i = 0; i < T.n_spec; ++i) {
    if (t < T.t_low[i]) {
      const double tt = T.t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt_inv = 1 / tt, log_tt = log(tt);
      const double* c = T.low_coeff[i];
      g[i] = c[0] * (1.0 - log_tt) - 0.5 * c[1] * tt - c[2] * tt2 / 6.0 - c[3] * tt3 / 12.0 - c[4] * tt4 * 0.05 +
             c[5] * tt_inv - c[6];
    } else {
      const double* c = t < T.t_mid[i] ? T.low_coeff[i] : T.high_coeff[i];
      g[i] = c[0] * (1.0 - log_t) - 0.5 * c[1] * t - c[2] * t2 / 6.0 - c[3] * t3 / 12.0 - c[4] * t4 * 0.05 +
             c[5] * t_inv - c[6];
    }
  }
}

// ------------------------------------------------------------------ transport
// compute_Omega_D, src/Transport.cu:168-170
inline double omega_d(double t_red) { return 1.0 / pow(t_red, 0.145) + 1.0 / ((t_red + 0.5) * (t_red + 0.5)); }

// compute_transport_property, src/Transport.cu:43-136 (the mixture-averaged
// diffusivity branch; the commented-out conductivity corrections are off)
void transport(const Tab& T, double t, double p, double rho, const double* Y, double mw_total, const double* cp,
               double* mu_out, double* lambda_out, double* rho_d) {
  const int ns = T.n_spec;
  double X[MS], vis[MS], dij[MS * MS];
  for (int l = 0; l < ns; ++l) {
    X[l] = Y[l] * mw_total / T.mw[l];
    const double t_dl{t * T.lj_potent_inv[l]};
    const double collision_integral{1.147 * pow(t_dl, -0.145) + pow(t_dl + 0.5, -2)};
    vis[l] = T.vis_coeff[l] * sqrt(t) / collision_integral;
  }
  const double t_3over2_over_p{t * sqrt(t) / p};
  for (int m = 0; m < ns; ++m)
    for (int n = m + 1; n < ns; ++n) {  // the diagonal is never used
      const double t_red{t * T.kb_over_eps_jk[m][n]};
      dij[m * ns + n] = T.binary_diffusivity_coeff[m][n] * t_3over2_over_p / omega_d(t_red);
      dij[n * ns + m] = dij[m * ns + n];
    }
  double viscosity = 0, conductivity = 0;
  for (int m = 0; m < ns; ++m) {
    double vis_temp{0};
    for (int n = 0; n < ns; ++n) {
      double partition_func{1.0};
      if (m != n) {
        const double numerator{1 + sqrt(vis[m] / vis[n]) * T.wj_div_wi_to_one4th[m][n]};
        partition_func = numerator * numerator * T.sqrt_wi_div_wj_pl1_mul8[m][n];
      }
      vis_temp += partition_func * X[n];
    }
    const double cond_temp = 1.065 * vis_temp - 0.065 * X[m];
    viscosity += vis[m] * X[m] / vis_temp;
    const double lambda = vis[m] * (cp[m] + 1.25 * kRu / T.mw[m]);
    conductivity += lambda * X[m] / cond_temp;
  }
  *mu_out = viscosity;
  *lambda_out = conductivity;
  constexpr double eps{1e-12};
  for (int l = 0; l < ns; ++l) {
    double num{0}, den{0};
    for (int n = 0; n < ns; ++n)
      if (l != n) {
        num += (X[n] + eps) * T.mw[n];
        den += (X[n] + eps) / dij[l * ns + n];
      }
    rho_d[l] = rho * num / (den * mw_total);
  }
}

// ------------------------------------------------------------------ chemistry
// arrhenius, src/FiniteRateChem.cu:104-106
inline double arrhenius(double t, double A, double b, double Ea) { return A * pow(t, b) * exp(-Ea / t); }

// forward_reaction_rate, src/FiniteRateChem.cu:60-102
void forward_rate(const Tab& T, double t, const double* c, double* kf) {
  for (int i = 0; i < T.n_reac; ++i) {
    const int type = T.label[i];
    kf[i] = arrhenius(t, T.A[i], T.b[i], T.Ea[i]);
    if (type == 3) {
      kf[i] += arrhenius(t, T.A2[i], T.b2[i], T.Ea2[i]);
    } else if (type > 3) {
      double cc{0};
      for (int l = 0; l < T.n_spec; ++l) cc += c[l] * T.third_body_coeff[i][l];
      if (type == 4) {
        kf[i] *= cc;
      } else {
        const double kf_low = arrhenius(t, T.A2[i], T.b2[i], T.Ea2[i]);
        const double kf_high = kf[i];
        const double reduced_pressure = kf_low * cc / kf_high;
        double F = 1.0;
        if (type > 5) {
          double f_cent = (1 - T.troe_alpha[i]) * exp(-t / T.troe_t3[i]) + T.troe_alpha[i] * exp(-t / T.troe_t1[i]);
          if (type == 7) f_cent += exp(-T.troe_t2[i] / t);
          const double logFc = std::log10(f_cent);
          const double cc2 = -0.4 - 0.67 * logFc;
          const double n = 0.75 - 1.27 * logFc;
          const double logPr = std::log10(reduced_pressure);
          const double tempo = (logPr + cc2) / (n - 0.14 * (logPr + cc2));
          const double pw = logFc / (1.0 + tempo * tempo);
          F = pow(10, pw);
        }
        kf[i] = kf_high * reduced_pressure / (1.0 + reduced_pressure) * F;
      }
    }
  }
}

// backward_reaction_rate, src/FiniteRateChem.cu:108-150
void backward_rate(const Tab& T, double t, const double* kf, const double* c, double* kb) {
  int n_gibbs{T.n_reac};
  for (int i = 0; i < T.n_reac; ++i) {
    kb[i] = 0;
    if (T.label[i] == 0) {
      --n_gibbs;
    } else if (T.rev_type[i] == 1) {
      kb[i] = arrhenius(t, T.A2[i], T.b2[i], T.Ea2[i]);
      if (T.label[i] == 4) {
        double cc{0};
        for (int l = 0; l < T.n_spec; ++l) cc += c[l] * T.third_body_coeff[i][l];
        kb[i] *= cc;
      }
      --n_gibbs;
    }
  }
  if (n_gibbs < 1) return;
  double g[MS];
  gibbs_rt(T, t, g);
  constexpr double temp_p = kPatm / kRu * 1e-3;
  const double temp_t = temp_p / t;
  for (int i = 0; i < T.n_reac; ++i) {
    if (T.label[i] != 2 && T.label[i] != 0) {
      double d_gibbs{0};
      for (int l = 0; l < T.n_spec; ++l) d_gibbs += g[l] * (T.stoi_b[i][l] - T.stoi_f[i][l]);
      const double kc{pow(temp_t, T.order[i]) * exp(-d_gibbs)};
      kb[i] = kf[i] / kc;
    }
  }
}

// finite_rate_chemistry: rate_of_progress + chemical_source,
// src/FiniteRateChem.cu:7-58, :152-195 (explicit treatment, chemSrcMethod 0)
void chemistry(const Tab& T, double rho, double t, const double* Y, double* omega, double* kf_out, double* kb_out) {
  const int ns = T.n_spec, nr = T.n_reac;
  double c[MS], kf[MR], kb[MR] = {}, q1[MR], q2[MR];
  for (int l = 0; l < ns; ++l) c[l] = rho * Y[l] / T.mw[l] * 1e-3;
  forward_rate(T, t, c, kf);
  backward_rate(T, t, kf, c, kb);
  for (int i = 0; i < nr; ++i) {
    if (T.label[i] != 0) {
      q1[i] = 1.0;
      q2[i] = 1.0;
      for (int j = 0; j < ns; ++j) {
        q1[i] *= pow(c[j], T.stoi_f[i][j]);
        q2[i] *= pow(c[j], T.stoi_b[i][j]);
      }
      q1[i] *= kf[i];
      q2[i] *= kb[i];
    } else {
      q1[i] = 1.0;
      q2[i] = 0.0;
      for (int j = 0; j < ns; ++j) q1[i] *= pow(c[j], T.stoi_f[i][j]);
      q1[i] *= kf[i];
    }
  }
  for (int i = 0; i < ns; ++i) {
    double creation = 0, destruction = 0;
    for (int j = 0; j < nr; ++j) {
      creation += q2[j] * T.stoi_f[j][i] + q1[j] * T.stoi_b[j][i];
      destruction += q1[j] * T.stoi_f[j][i] + q2[j] * T.stoi_b[j][i];
    }
    creation *= T.mw[i] * 1e+3;
    destruction *= T.mw[i] * 1e+3;
    omega[i] = creation - destruction;
  }
  if (kf_out) std::copy(kf, kf + nr, kf_out);
  if (kb_out) std::copy(kb, kb + nr, kb_out);
}

// compute_temperature_and_pressure, src/FieldOperation.cu:3-40 (Newton on
// e(T) = h(T) - R T, stop at |1 - T_new/T| <= 1e-3 or 1000 iterations)
double temperature(const Tab& T, double e, const double* Y, double gas_const, double t, int* iters) {
  double err{1};
  constexpr int max_iter{1000};
  constexpr double eps{1e-3};
  int iter = 0;
  double h_i[MS], cp_i[MS];
  while (err > eps && iter++ < max_iter) {
    enthalpy_cp(T, t, h_i, cp_i);
    double cp_tot{0}, h{0};
    for (int l = 0; l < T.n_spec; ++l) {
      cp_tot += cp_i[l] * Y[l];
      h += h_i[l] * Y[l];
    }
    const double e_t = h - gas_const * t;
    const double cv = cp_tot - gas_const;
    const double t1 = t - (e_t - e) / cv;
    err = abs(1 - t1 / t);
    t = t1;
  }
  if (iters) *iters = iter;
  return t;
}

// ------------------------------------------------------------ WENO kernels
// WENO-Z5, src/WENO.cu:1582-1617 (f+ on vp[0..4], f- on vm[0..4])
double weno5z(const double* vp, const double* vm, double eps) {
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

// WENO-Z7, src/WENO.cu:1619-1693
double weno7z(const double* vp, const double* vm, double eps) {
  constexpr double one6th{1.0 / 6};
  constexpr double d12{13.0 / 12.0}, d13{1043.0 / 960}, d14{1.0 / 12};
  double s1{one6th * (-2 * vp[0] + 9 * vp[1] - 18 * vp[2] + 11 * vp[3])};
  double s2{-vp[0] + 4 * vp[1] - 5 * vp[2] + 2 * vp[3]};
  double s3{-vp[0] + 3 * vp[1] - 3 * vp[2] + vp[3]};
  double b0{s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3};
  s1 = one6th * (vp[1] - 6 * vp[2] + 3 * vp[3] + 2 * vp[4]);
  s2 = vp[2] - 2 * vp[3] + vp[4];
  s3 = -vp[1] + 3 * vp[2] - 3 * vp[3] + vp[4];
  double b1{s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3};
  s1 = one6th * (-2 * vp[2] - 3 * vp[3] + 6 * vp[4] - vp[5]);
  s3 = -vp[2] + 3 * vp[3] - 3 * vp[4] + vp[5];
  double b2{s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3};
  s1 = one6th * (-11 * vp[3] + 18 * vp[4] - 9 * vp[5] + 2 * vp[6]);
  s2 = 2 * vp[3] - 5 * vp[4] + 4 * vp[5] - vp[6];
  s3 = -vp[3] + 3 * vp[4] - 3 * vp[5] + vp[6];
  double b3{s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3};
  double tau{(b0 - b3) * (b0 - b3)};
  constexpr double c0{1.0 / 35}, c1{12.0 / 35}, c2{18.0 / 35}, c3{4.0 / 35};
  double a0{c0 + c0 * tau / ((eps + b0) * (eps + b0))};
  double a1{c1 + c1 * tau / ((eps + b1) * (eps + b1))};
  double a2{c2 + c2 * tau / ((eps + b2) * (eps + b2))};
  double a3{c3 + c3 * tau / ((eps + b3) * (eps + b3))};
  constexpr double one12th{1.0 / 12};
  double v0{one12th * (-3 * vp[0] + 13 * vp[1] - 23 * vp[2] + 25 * vp[3])};
  double v1{one12th * (vp[1] - 5 * vp[2] + 13 * vp[3] + 3 * vp[4])};
  double v2{one12th * (-vp[2] + 7 * vp[3] + 7 * vp[4] - vp[5])};
  double v3{one12th * (3 * vp[3] + 13 * vp[4] - 5 * vp[5] + vp[6])};
  const double fp{(a0 * v0 + a1 * v1 + a2 * v2 + a3 * v3) / (a0 + a1 + a2 + a3)};

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
  const double fm{(a0 * v0 + a1 * v1 + a2 * v2 + a3 * v3) / (a0 + a1 + a2 + a3)};
  return fp + fm;
}

// compute_flux (LF split), src/WENO.cu:697-720; q = [cv(0..nv-1), p, c]
void lf_split(const double* q, int nv, const double* m, double jac, double* fp, double* fm) {
  const double Uk{(q[1] * m[0] + q[2] * m[1] + q[3] * m[2]) / q[0]};
  const double pk{q[nv]};
  const double cGradK = q[nv + 1] * sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const double lambda0 = abs(Uk) + cGradK;
  fp[0] = 0.5 * jac * (Uk * q[0] + lambda0 * q[0]);
  fp[1] = 0.5 * jac * (Uk * q[1] + pk * m[0] + lambda0 * q[1]);
  fp[2] = 0.5 * jac * (Uk * q[2] + pk * m[1] + lambda0 * q[2]);
  fp[3] = 0.5 * jac * (Uk * q[3] + pk * m[2] + lambda0 * q[3]);
  fp[4] = 0.5 * jac * (Uk * (q[4] + pk) + lambda0 * q[4]);
  fm[0] = 0.5 * jac * (Uk * q[0] - lambda0 * q[0]);
  fm[1] = 0.5 * jac * (Uk * q[1] + pk * m[0] - lambda0 * q[1]);
  fm[2] = 0.5 * jac * (Uk * q[2] + pk * m[1] - lambda0 * q[2]);
  fm[3] = 0.5 * jac * (Uk * q[3] + pk * m[2] - lambda0 * q[3]);
  fm[4] = 0.5 * jac * (Uk * (q[4] + pk) - lambda0 * q[4]);
  for (int l = 5; l < nv; ++l) {
    fp[l] = 0.5 * jac * (Uk * q[l] + lambda0 * q[l]);
    fm[l] = 0.5 * jac * (Uk * q[l] - lambda0 * q[l]);
  }
}

// Roe-averaged state and characteristic quantities of compute_weno_flux_ch
// (mixture), src/WENO.cu:744-815.
struct CharFace {
  double um, vm, wm, hm, cm, gm1, kx, ky, kz, Uk_bar, alpha, cm2_inv;
  double svm[MS], alpha_l[MS];
  double kxJ, kyJ, kzJ;
};

void char_face(const Tab& T, int nv, const double* cvl, const double* cvr, const double* m_l, const double* m_r,
               double jac_l, double jac_r, CharFace& F) {
  const int ns = T.n_spec;
  const double rhoL_inv{1.0 / cvl[0]}, rhoR_inv{1.0 / cvr[0]};
  const double rlc{sqrt(cvl[0]) / (sqrt(cvl[0]) + sqrt(cvr[0]))};
  const double rrc{sqrt(cvr[0]) / (sqrt(cvl[0]) + sqrt(cvr[0]))};
  F.um = rlc * cvl[1] * rhoL_inv + rrc * cvr[1] * rhoR_inv;
  F.vm = rlc * cvl[2] * rhoL_inv + rrc * cvr[2] * rhoR_inv;
  F.wm = rlc * cvl[3] * rhoL_inv + rrc * cvr[3] * rhoR_inv;
  const double ekm{0.5 * (F.um * F.um + F.vm * F.vm + F.wm * F.wm)};
  const double hl{(cvl[4] + cvl[nv]) * rhoL_inv};
  const double hr{(cvr[4] + cvr[nv]) * rhoR_inv};
  F.hm = rlc * hl + rrc * hr;
  for (int l = 0; l < ns; ++l) F.svm[l] = rlc * cvl[l + 5] * rhoL_inv + rrc * cvr[l + 5] * rhoR_inv;
  double mw_inv = 0;
  for (int l = 0; l < ns; ++l) mw_inv += F.svm[l] / T.mw[l];
  const double tl{cvl[nv] * rhoL_inv};
  const double tr{cvr[nv] * rhoR_inv};
  const double tm = (rlc * tl + rrc * tr) / (kRu * mw_inv);
  double cp_i[MS], h_i[MS];
  enthalpy_cp(T, tm, h_i, cp_i);
  double cp{0}, cv_tot{0};
  for (int l = 0; l < ns; ++l) {
    cp += F.svm[l] * cp_i[l];
    cv_tot += F.svm[l] * (cp_i[l] - kRu / T.mw[l]);
  }
  const double gamma = cp / cv_tot;
  F.cm = sqrt(gamma * kRu * mw_inv * tm);
  F.gm1 = gamma - 1;
  F.kxJ = m_l[0] * jac_l + m_r[0] * jac_r;
  F.kyJ = m_l[1] * jac_l + m_r[1] * jac_r;
  F.kzJ = m_l[2] * jac_l + m_r[2] * jac_r;
  F.kx = F.kxJ / (jac_l + jac_r);
  F.ky = F.kyJ / (jac_l + jac_r);
  F.kz = F.kzJ / (jac_l + jac_r);
  const double gradK{sqrt(F.kx * F.kx + F.ky * F.ky + F.kz * F.kz)};
  F.kx /= gradK;
  F.ky /= gradK;
  F.kz /= gradK;
  F.Uk_bar = F.kx * F.um + F.ky * F.vm + F.kz * F.wm;
  F.alpha = F.gm1 * ekm;
  F.cm2_inv = 1.0 / (F.cm * F.cm);
  for (int l = 0; l < ns; ++l) {
    F.alpha_l[l] = gamma * kRu / T.mw[l] * tm - (gamma - 1) * h_i[l];
    F.alpha_l[l] *= F.cm2_inv;
  }
}

// Row l (0..4) of the left eigenvector matrix, src/WENO.cu:828-870; returns
// the coefficient of the species coupling term.
double char_row(const CharFace& F, int l, double* L) {
  const double um = F.um, vm = F.vm, wm = F.wm, cm = F.cm, gm1 = F.gm1, kx = F.kx, ky = F.ky, kz = F.kz;
  const double cm2_inv = F.cm2_inv, alpha = F.alpha, Uk_bar = F.Uk_bar;
  double coeff_alpha_s{0.5};
  switch (l) {
    case 0:
      L[0] = (alpha + Uk_bar * cm) * cm2_inv * 0.5;
      L[1] = -(gm1 * um + kx * cm) * cm2_inv * 0.5;
      L[2] = -(gm1 * vm + ky * cm) * cm2_inv * 0.5;
      L[3] = -(gm1 * wm + kz * cm) * cm2_inv * 0.5;
      L[4] = gm1 * cm2_inv * 0.5;
      break;
    case 1:
      coeff_alpha_s = -kx;
      L[0] = kx * (1 - alpha * cm2_inv) - (kz * vm - ky * wm) / cm;
      L[1] = kx * gm1 * um * cm2_inv;
      L[2] = (kx * gm1 * vm + kz * cm) * cm2_inv;
      L[3] = (kx * gm1 * wm - ky * cm) * cm2_inv;
      L[4] = -kx * gm1 * cm2_inv;
      break;
    case 2:
      coeff_alpha_s = -ky;
      L[0] = ky * (1 - alpha * cm2_inv) - (kx * wm - kz * um) / cm;
      L[1] = (ky * gm1 * um - kz * cm) * cm2_inv;
      L[2] = ky * gm1 * vm * cm2_inv;
      L[3] = (ky * gm1 * wm + kx * cm) * cm2_inv;
      L[4] = -ky * gm1 * cm2_inv;
      break;
    case 3:
      coeff_alpha_s = -kz;
      L[0] = kz * (1 - alpha * cm2_inv) - (ky * um - kx * vm) / cm;
      L[1] = (kz * gm1 * um + ky * cm) * cm2_inv;
      L[2] = (kz * gm1 * vm - kx * cm) * cm2_inv;
      L[3] = kz * gm1 * wm * cm2_inv;
      L[4] = -kz * gm1 * cm2_inv;
      break;
    default:
      L[0] = (alpha - Uk_bar * cm) * cm2_inv * 0.5;
      L[1] = -(gm1 * um - kx * cm) * cm2_inv * 0.5;
      L[2] = -(gm1 * vm - ky * cm) * cm2_inv * 0.5;
      L[3] = -(gm1 * wm - kz * cm) * cm2_inv * 0.5;
      L[4] = gm1 * cm2_inv * 0.5;
      break;
  }
  return coeff_alpha_s;
}

// Back projection to physical space, src/WENO.cu:1121-1142
void char_back(const CharFace& F, int ns, const double* fChar, double* fci) {
  const double um = F.um, vm = F.vm, wm = F.wm, cm = F.cm, gm1 = F.gm1, kx = F.kx, ky = F.ky, kz = F.kz;
  const double hm = F.hm, Uk_bar = F.Uk_bar;
  fci[0] = fChar[0] + kx * fChar[1] + ky * fChar[2] + kz * fChar[3] + fChar[4];
  fci[1] = (um - kx * cm) * fChar[0] + kx * um * fChar[1] + (ky * um - kz * cm) * fChar[2] +
           (kz * um + ky * cm) * fChar[3] + (um + kx * cm) * fChar[4];
  fci[2] = (vm - ky * cm) * fChar[0] + (kx * vm + kz * cm) * fChar[1] + ky * vm * fChar[2] +
           (kz * vm - kx * cm) * fChar[3] + (vm + ky * cm) * fChar[4];
  fci[3] = (wm - kz * cm) * fChar[0] + (kx * wm - ky * cm) * fChar[1] + (ky * wm + kx * cm) * fChar[2] +
           kz * wm * fChar[3] + (wm + kz * cm) * fChar[4];
  fci[4] = (hm - Uk_bar * cm) * fChar[0] + (kx * (hm - cm * cm / gm1) + (kz * vm - ky * wm) * cm) * fChar[1] +
           (ky * (hm - cm * cm / gm1) + (kx * wm - kz * um) * cm) * fChar[2] +
           (kz * (hm - cm * cm / gm1) + (ky * um - kx * vm) * cm) * fChar[3] + (hm + Uk_bar * cm) * fChar[4];
  double add{0};
  const double coeff_add = fChar[0] + kx * fChar[1] + ky * fChar[2] + kz * fChar[3] + fChar[4];
  for (int l = 0; l < ns; ++l) {
    fci[5 + l] = F.svm[l] * coeff_add + fChar[l + 5];
    add += F.alpha_l[l] * fChar[l + 5];
  }
  fci[4] -= add * cm * cm / gm1;
}

// --------------------------------------------------------------- structures
struct Face {
  int axis, side, type;
  orf_farfield ff;   // type 4: the state of its label
  int lo[3], hi[3];  // the face layer; tangential node range widened by ngg (src/Mesh.cpp:14-32)
  orf_inflow in;     // type 5: the inflow of its label
  std::vector<double> y;  // type 5: y of the face points, i fastest over lo..hi
};
using InnerFace = orc_faces::Inner;  // 1-to-1 inner face, src/Mesh.cpp:38-176

struct Block {
  int mx, my, mz, g;
  long long sx, sy, sz;
  std::vector<double> cv, bv, sv, qn, dq, mul, lam, cp, gam, cc, mach, rhod, jac, metric, dtl, bv_last;
  std::vector<Face> faces;
  std::vector<InnerFace> inner;
  unsigned long long bc_passes = 0;  // boundary-condition passes so far (keys the inflow noise)
  bool sponge = false;                      // sponge layer on this block
  orf_sponge sp{};
  std::vector<double> xs, ys, zs;           // interior node coordinates (sponge layer)
  std::vector<double> sponge_mean_cv;       // [nv][n0]
  long long id(int i, int j, int k) const { return (k + g) * sx * sy + (j + g) * sx + (i + g); }
  long long id0(int i, int j, int k) const { return (long long)k * mx * my + (long long)j * mx + i; }
  long long n0() const { return (long long)mx * my * mz; }
};

struct Ctx {
  orf_setup s;
  Tab T;
  int ngg = 3, nv = 5, ns = 0;
  double weno_eps_scale = 1;
  double dt_cur = 0;
  std::vector<Block> blocks;
  unsigned long long sponge_iter = 0;  // steps averaged into the sponge means
};

// compute_total_energy<FR> + compute_cv_from_bv_1_point, src/FieldOperation.cuh:17-105
void cv_point(const Ctx& c, Block& b, long long p) {
  const long long s = b.sz;
  const double rho = b.bv[p];
  b.cv[p] = rho;
  b.cv[p + s] = rho * b.bv[p + s];
  b.cv[p + 2 * s] = rho * b.bv[p + 2 * s];
  b.cv[p + 3 * s] = rho * b.bv[p + 3 * s];
  for (int l = 0; l < c.ns; ++l) b.cv[p + (5 + l) * s] = rho * b.sv[p + l * s];
  const double u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
  const double V2 = u * u + v * v + w * w;
  double te = 0.5 * V2;
  double h[MS];
  enthalpy(c.T, b.bv[p + 5 * s], h);
  for (int l = 0; l < c.ns; ++l) te += h[l] * b.sv[p + l * s];
  te *= rho;
  te -= b.bv[p + 4 * s];
  b.cv[p + 4 * s] = te;
}

void cv_from_bv(const Ctx& c, Block& b) {
  const int g = b.g;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = -g; k < b.mz + g; ++k)
    for (int j = -g; j < b.my + g; ++j)
      for (int i = -g; i < b.mx + g; ++i) cv_point(c, b, b.id(i, j, k));
}

// update_physical_properties<FR>, src/FieldOperation.cuh:107-140
void properties(const Ctx& c, Block& b) {
  const int g = b.g, ns = c.ns;
  const long long s = b.sz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = -g; k < b.mz + g; ++k)
    for (int j = -g; j < b.my + g; ++j)
      for (int i = -g; i < b.mx + g; ++i) {
        const long long p = b.id(i, j, k);
        const double t = b.bv[p + 5 * s];
        const double u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
        const double V = sqrt(u * u + v * v + w * w);
        double mw{0}, cp_tot{0}, cvv{0}, cp[MS], Y[MS];
        heat_capacity(c.T, t, cp);
        for (int l = 0; l < ns; ++l) {
          Y[l] = b.sv[p + l * s];
          mw += Y[l] / c.T.mw[l];
          cp_tot += Y[l] * cp[l];
          cvv += Y[l] * (cp[l] - kRu / c.T.mw[l]);
        }
        mw = 1 / mw;
        b.cp[p] = cp_tot;
        b.gam[p] = cp_tot / cvv;
        b.cc[p] = sqrt(b.gam[p] * kRu * t / mw);
        double mu, lam, rd[MS];
        transport(c.T, t, b.bv[p + 4 * s], b.bv[p], Y, mw, cp, &mu, &lam, rd);
        b.mul[p] = mu;
        b.lam[p] = lam;
        for (int l = 0; l < ns; ++l) b.rhod[p + l * s] = rd[l];
        b.mach[p] = V / b.cc[p];
      }
}

// compute_source<FR> -> finite_rate_chemistry, src/SourceTerm.cuh:11-35
// sponge_function 0 (Nektar++), src/SpongeLayer.cu:162-170
double sponge_function(double xi) {
  const double xi3 = xi * xi * xi;
  return 6 * xi3 * xi * xi - 15 * xi3 * xi + 10 * xi3;
}

// sponge_layer_source, src/SpongeLayer.cu:68-160: sigma from the sides that
// are on (x-, x+, y-, y+, z-, z+ in this order), then dq -= sigma (cv - mean).
void sponge_source(const Ctx& c, Block& b) {
  const orf_sponge& sp = b.sp;
  const long long s = b.sz, n0 = b.n0();
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), p0 = b.id0(i, j, k);
        const double pos[3] = {b.xs[p0], b.ys[p0], b.zs[p0]};
        double sigma = 0;
        for (int side = 0; side < 6; ++side) {
          if (!sp.enabled[side]) continue;
          const double x = pos[side / 2], x1 = sp.start[side];
          const bool inside = side % 2 == 0 ? x < x1 : x > x1;
          if (!inside) continue;
          const double x2 = sp.end[side];
          double xi = (x - x1) / (x2 - x1);
          if (xi > 1) xi = 1;
          sigma += sp.sigma[side] * sponge_function(xi);
        }
        // the reference's term carries no Jacobian (its rate is sigma / J);
        // jacobian_scaled multiplies it in, as the chemistry source does
        if (sp.jacobian_scaled) sigma = sigma * b.jac[p];
        for (int l = 0; l < c.nv; ++l)
          b.dq[p0 + l * n0] -= sigma * (b.cv[p + l * s] - b.sponge_mean_cv[p0 + l * n0]);
      }
}

// update_sponge_layer_value, src/SpongeLayer.cu:28-49 (the scalar counters
// equal the step counter in a new run)
void sponge_mean(const Ctx& c, Block& b) {
  const long long s = b.sz, n0 = b.n0();
  const double n = static_cast<double>(c.sponge_iter);
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), p0 = b.id0(i, j, k);
        for (int l = 0; l < c.nv; ++l) {
          double& m = b.sponge_mean_cv[p0 + l * n0];
          m = (m * n + b.cv[p + l * s]) / (n + 1);
        }
      }
}

void source(const Ctx& c, Block& b) {
  if (b.sponge && (!c.s.reaction || c.T.n_reac == 0)) sponge_source(c, b);
  if (!c.s.reaction || c.T.n_reac == 0) return;
  const long long s = b.sz, n0 = b.n0();
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), p0 = b.id0(i, j, k);
        double Y[MS], om[MS];
        for (int l = 0; l < c.ns; ++l) Y[l] = b.sv[p + l * s];
        chemistry(c.T, b.bv[p], b.bv[p + 5 * s], Y, om, nullptr, nullptr);
        for (int l = 0; l < c.ns; ++l) b.dq[p0 + (l + 5) * n0] += b.jac[p] * om[l];
      }
  if (b.sponge) sponge_source(c, b);
}

// local_time_step<FR, Laminar>, src/TimeAdvanceFunc.cuh:28-83
void local_dt(const Ctx& c, Block& b) {
  const long long s = b.sz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k);
        auto m = [&](int r, int cc) { return b.metric[p + (3 * (r - 1) + (cc - 1)) * s]; };
        const double gxi = sqrt(m(1, 1) * m(1, 1) + m(1, 2) * m(1, 2) + m(1, 3) * m(1, 3));
        const double geta = sqrt(m(2, 1) * m(2, 1) + m(2, 2) * m(2, 2) + m(2, 3) * m(2, 3));
        const double gzeta = sqrt(m(3, 1) * m(3, 1) + m(3, 2) * m(3, 2) + m(3, 3) * m(3, 3));
        const double u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
        const double U = u * m(1, 1) + v * m(1, 2) + w * m(1, 3);
        const double V = u * m(2, 1) + v * m(2, 2) + w * m(2, 3);
        const double W = u * m(3, 1) + v * m(3, 2) + w * m(3, 3);
        const double a = b.cc[p];
        const double r0 = abs(U) + a * gxi, r1 = abs(V) + a * geta, r2 = abs(W) + a * gzeta;
        const double sr_inv = r0 + r1 + r2;
        const double coeff_1 = std::max(b.gam[p], 4.0 / 3.0) / b.bv[p];
        const double coeff_2 = b.mul[p] / c.s.prandtl;
        double sr_visc = gxi * gxi + geta * geta;
        sr_visc += gzeta * gzeta;
        sr_visc *= coeff_1 * coeff_2;
        b.dtl[b.id0(i, j, k)] = c.s.cfl / (sr_inv + sr_visc);
      }
}

// positive_preserving_limiter, src/WENO.cu:1538-1580; fc, f1 are the
// species parts of one face, ql/qr the species densities of its two cells.
void pp_limiter(int ns, double* fc, const double* f1, const double* ql, const double* qr, double jl, double jr,
                double dt, int idx, int max_extent) {
  const double alpha = 1.0 / 3.0;  // three-dimensional blocks
  for (int l = 0; l < ns; ++l) {
    double theta_p = 1.0, theta_m = 1.0;
    if (idx > -1) {
      const double up = 0.5 * alpha * ql[l] * jl - dt * fc[l];
      if (up < 0) {
        const double up_lf = 0.5 * alpha * ql[l] * jl - dt * f1[l];
        if (abs(up - up_lf) > 1e-20) {
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
        if (abs(um - um_lf) > 1e-20) {
          theta_m = (0 - um_lf) / (um - um_lf);
          if (theta_m > 1) theta_m = 1.0;
          else if (theta_m < 0) theta_m = 0;
        }
      }
    }
    fc[l] = std::min(theta_p, theta_m) * (fc[l] - f1[l]) + f1[l];
  }
}

// ------------------------------------------- primitive reconstruction path
// (inviscid_type 0: compute_convective_term_pv, src/InviscidScheme.cu:12-215)

// gxl::sgn (src/gxl_lib/Math.hpp:13-15): -1 for negative, +1 otherwise (also 0)
inline int sgn(double a) { return a < 0 ? -1 : 1; }
// minmod limiter, src/Limiter.cuh:7-15
inline double minmod(double a, double b) { return 0.5 * (sgn(a) + sgn(b)) * std::min(abs(a), abs(b)); }

// first_order_reconstruct / MUSCL_reconstruct / NND2_reconstruct,
// src/Reconstruction.cu:5-46; pv holds nv values per point, face between
// points `is` and `is`+1.
void reconstruct(const double* pv, int nv, int is, int recon, double* pl, double* pr) {
  for (int l = 0; l < nv; ++l) {
    const double a0 = pv[(is - 1) * nv + l], a1 = pv[is * nv + l], a2 = pv[(is + 1) * nv + l];
    if (recon == 2 || recon == 3) {
      const double a3 = pv[(is + 2) * nv + l];
      const double delta_i{a1 - a0};
      const double delta_i1{a2 - a1};
      const double delta_i2{a3 - a2};
      if (recon == 2) {
        constexpr double kappa{1.0 / 3.0};
        const double neg_l = minmod(delta_i, delta_i1), pos_l = minmod(delta_i1, delta_i);
        const double neg_r = minmod(delta_i1, delta_i2), pos_r = minmod(delta_i2, delta_i1);
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
}

// total energy and gamma of a reconstructed mixture state (the mixture part
// of reconstruction<>, src/InviscidScheme.cu:180-209): q[nv] = E, q[nv+1] = gamma
void pv_energy(const Tab& T, int nv, double* q) {
  const int ns = T.n_spec;
  double e = 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  double mw_inv{0.0};
  for (int l = 0; l < ns; ++l) mw_inv += q[5 + l] / T.mw[l];
  const double t = q[4] / (q[0] * kRu * mw_inv);
  double h[MS], cpi[MS];
  enthalpy_cp(T, t, h, cpi);
  double cp{0}, cv{0};
  for (int l = 0; l < ns; ++l) {
    cp += cpi[l] * q[l + 5];
    cv += q[l + 5] * (cpi[l] - kRu / T.mw[l]);
    e += h[l] * q[l + 5];
  }
  q[nv] = q[0] * e - q[4];
  q[nv + 1] = cp / cv;
}

// riemannSolver_ausmPlus (mixture), src/RiemannSolver.cuh:9-96
void ausm_plus(int nv, const double* pl, const double* pr, const double* ml, const double* mr, double jl, double jr,
               double* fci) {
  const double k1 = 0.5 * (jl * ml[0] + jr * mr[0]);
  const double k2 = 0.5 * (jl * ml[1] + jr * mr[1]);
  const double k3 = 0.5 * (jl * ml[2] + jr * mr[2]);
  const double grad_k_div_jac = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const double ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) / grad_k_div_jac;
  const double ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) / grad_k_div_jac;
  const double p_l = pl[4], p_r = pr[4], rho_l = pl[0], rho_r = pr[0];
  const double gam_l = pl[nv + 1], gam_r = pr[nv + 1];
  const double c = 0.5 * (sqrt(gam_l * p_l / rho_l) + sqrt(gam_r * p_r / rho_r));
  const double mach_l = ul / c, mach_r = ur / c;
  double mlp, mrn, plp, prn;
  constexpr double alpha{3 / 16.0};
  if (abs(mach_l) > 1) {
    mlp = 0.5 * (mach_l + abs(mach_l));
    plp = mlp / mach_l;
  } else {
    const double a = (mach_l + 1) * (mach_l + 1) * 0.25;
    const double b = (mach_l * mach_l - 1) * (mach_l * mach_l - 1);
    mlp = a + 0.125 * b;
    plp = a * (2 - mach_l) + alpha * mach_l * b;
  }
  if (abs(mach_r) > 1) {
    mrn = 0.5 * (mach_r - abs(mach_r));
    prn = mrn / mach_r;
  } else {
    const double a = (mach_r - 1) * (mach_r - 1) * 0.25;
    const double b = (mach_r * mach_r - 1) * (mach_r * mach_r - 1);
    mrn = -a - 0.125 * b;
    prn = a * (2 + mach_r) - alpha * mach_r * b;
  }
  const double p_coeff = plp * p_l + prn * p_r;
  const double m_half = mlp + mrn;
  const double mach_pos = 0.5 * (m_half + abs(m_half));
  const double mach_neg = 0.5 * (m_half - abs(m_half));
  const double mass_flux_half = c * (rho_l * mach_pos + rho_r * mach_neg);
  const double coeff = mass_flux_half * grad_k_div_jac;
  const double* q = mass_flux_half >= 0 ? pl : pr;
  fci[0] = coeff;
  fci[1] = coeff * q[1] + p_coeff * k1;
  fci[2] = coeff * q[2] + p_coeff * k2;
  fci[3] = coeff * q[3] + p_coeff * k3;
  fci[4] = coeff * (q[nv] + q[4]) / q[0];
  for (int l = 5; l < nv; ++l) fci[l] = coeff * q[l];
}

// riemannSolver_hllc (mixture), src/RiemannSolver.cuh:98-230
void hllc(const Tab& T, int nv, const double* pl, const double* pr, const double* ml, const double* mr, double jl,
          double jr, double* fci) {
  const int ns = T.n_spec;
  const double rl_c{sqrt(pl[0]) / (sqrt(pl[0]) + sqrt(pr[0]))}, rr_c{sqrt(pr[0]) / (sqrt(pl[0]) + sqrt(pr[0]))};
  const double u_tilde{pl[1] * rl_c + pr[1] * rr_c};
  const double v_tilde{pl[2] * rl_c + pr[2] * rr_c};
  const double w_tilde{pl[3] * rl_c + pr[3] * rr_c};
  double svm[MS];
  for (int l = 0; l < ns; ++l) svm[l] = rl_c * pl[l + 5] + rr_c * pr[l + 5];
  double mw_inv{0};
  for (int l = 0; l < ns; ++l) mw_inv += svm[l] / T.mw[l];
  const double tl{pl[4] / pl[0]};
  const double tr{pr[4] / pr[0]};
  const double t{(rl_c * tl + rr_c * tr) / (kRu * mw_inv)};
  double cp_i[MS], h_i[MS];
  enthalpy_cp(T, t, h_i, cp_i);
  double cv{0}, cp{0};
  for (int l = 0; l < ns; ++l) {
    cp += cp_i[l] * svm[l];
    cv += svm[l] * (cp_i[l] - kRu / T.mw[l]);
  }
  const double gamma = cp / cv;
  const double c_tilde = sqrt(gamma * kRu * mw_inv * t);
  const double kx = 0.5 * (ml[0] + mr[0]);
  const double ky = 0.5 * (ml[1] + mr[1]);
  const double kz = 0.5 * (ml[2] + mr[2]);
  const double gradK = sqrt(kx * kx + ky * ky + kz * kz);
  const double U_tilde_bar{(kx * u_tilde + ky * v_tilde + kz * w_tilde) / gradK};
  const double jac_ave{0.5 * (jl + jr)};
  const double gamma_l = pl[nv + 1], gamma_r = pr[nv + 1];
  const double Ul{kx * pl[1] + ky * pl[2] + kz * pl[3]};
  const double cl{sqrt(gamma_l * pl[4] / pl[0])};
  const double sl{std::min(Ul / gradK - cl, U_tilde_bar - c_tilde)};
  if (sl >= 0) {
    const double rhoUk{pl[0] * Ul};
    fci[0] = jac_ave * rhoUk;
    fci[1] = jac_ave * (rhoUk * pl[1] + pl[4] * kx);
    fci[2] = jac_ave * (rhoUk * pl[2] + pl[4] * ky);
    fci[3] = jac_ave * (rhoUk * pl[3] + pl[4] * kz);
    fci[4] = jac_ave * ((pl[nv] + pl[4]) * Ul);
    for (int l = 5; l < nv; ++l) fci[l] = jac_ave * rhoUk * pl[l];
    return;
  }
  const double Ur{kx * pr[1] + ky * pr[2] + kz * pr[3]};
  const double cr{sqrt(gamma_r * pr[4] / pr[0])};
  const double sr{std::max(Ur / gradK + cr, U_tilde_bar + c_tilde)};
  if (sr < 0) {
    const double rhoUk{pr[0] * Ur};
    fci[0] = jac_ave * rhoUk;
    fci[1] = jac_ave * (rhoUk * pr[1] + pr[4] * kx);
    fci[2] = jac_ave * (rhoUk * pr[2] + pr[4] * ky);
    fci[3] = jac_ave * (rhoUk * pr[3] + pr[4] * kz);
    fci[4] = jac_ave * ((pr[nv] + pr[4]) * Ur);
    for (int l = 5; l < nv; ++l) fci[l] = jac_ave * rhoUk * pr[l];
    return;
  }
  const double sm{((pr[0] * Ur * (sr - Ur / gradK) - pl[0] * Ul * (sl - Ul / gradK)) / gradK + pl[4] - pr[4]) /
                  (pr[0] * (sr - Ur / gradK) - pl[0] * (sl - Ul / gradK))};
  const double pm{pl[0] * (sl - Ul / gradK) * (sm - Ul / gradK) + pl[4]};
  if (sm >= 0) {
    const double pCoeff{1.0 / (sl - sm)};
    const double QlCoeff{jac_ave * pCoeff * sm * (sl * gradK - Ul) * pl[0]};
    fci[0] = QlCoeff;
    const double dP{(sl * pm - sm * pl[4]) * pCoeff * jac_ave};
    fci[1] = QlCoeff * pl[1] + dP * kx;
    fci[2] = QlCoeff * pl[2] + dP * ky;
    fci[3] = QlCoeff * pl[3] + dP * kz;
    fci[4] = QlCoeff * pl[nv] / pl[0] + pCoeff * jac_ave * (sl * pm * sm * gradK - sm * pl[4] * Ul);
    for (int l = 5; l < nv; ++l) fci[l] = QlCoeff * pl[l];
  } else {
    const double pCoeff{1.0 / (sr - sm)};
    const double QrCoeff{jac_ave * pCoeff * sm * (sr * gradK - Ur) * pr[0]};
    fci[0] = QrCoeff;
    const double dP{(sr * pm - sm * pr[4]) * pCoeff * jac_ave};
    fci[1] = QrCoeff * pr[1] + dP * kx;
    fci[2] = QrCoeff * pr[2] + dP * ky;
    fci[3] = QrCoeff * pr[3] + dP * kz;
    fci[4] = QrCoeff * pr[nv] / pr[0] + pCoeff * jac_ave * (sr * pm * sm * gradK - sm * pr[4] * Ur);
    for (int l = 5; l < nv; ++l) fci[l] = QrCoeff * pr[l];
  }
}

// compute_entropy_fix_delta (mixture, 3D), src/InviscidScheme.cu:420-453:
// entropy_fix_factor (src/Parameter.cpp:474, default 0.125) times (|U| + |V| +
// |W| + c (|grad xi| + |grad eta| + |grad zeta|) / 3) at one point, m the 9
// metric entries.
double entropy_delta(double u, double v, double w, double c, const double* m, double factor) {
  const double U = fabs(u * m[0] + v * m[1] + w * m[2]);
  const double V = fabs(u * m[3] + v * m[4] + w * m[5]);
  const double W = fabs(u * m[6] + v * m[7] + w * m[8]);
  const double kx = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const double ky = sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
  const double kz = sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8]);
  return factor * (U + V + W + c * (kx + ky + kz) / 3.0);
}

// riemannSolver_Roe (mixture), src/RiemannSolver.cuh:234-417:
// compute_half_sum_left_right_flux (:234-276) plus -1/2 R |Lambda| L dQ.
// The Roe state is the sqrt(rho)-weighted mean of u, v, w, H, Y_k and p/rho;
// T = mean(p/rho) / (R_u sum Y_k / W_k); gamma, c, h_k from NASA-7 at T.
// |Lambda| of U - c|k|, U, U + c|k| with Harten's fix of width
// (delta_l + delta_r) / 2.  pl, pr: rho, u, v, w, p, Y_k, E.
void roe(const Tab& T, int nv, const double* pl, const double* pr, const double* ml, const double* mr, double jl,
         double jr, double delta_l, double delta_r, double* fci) {
  const int ns = T.n_spec;
  const double* ps[2] = {pl, pr};
  const double* ms[2] = {ml, mr};
  const double js[2] = {jl, jr};
  for (int l = 0; l < nv; ++l) fci[l] = 0;
  for (int side = 0; side < 2; ++side) {
    const double* q = ps[side];
    const double ax = js[side] * ms[side][0], ay = js[side] * ms[side][1], az = js[side] * ms[side][2];
    const double Uk = q[1] * ax + q[2] * ay + q[3] * az;
    const double mflux = Uk * q[0];
    fci[0] += 0.5 * mflux;
    fci[1] += 0.5 * (mflux * q[1] + q[4] * ax);
    fci[2] += 0.5 * (mflux * q[2] + q[4] * ay);
    fci[3] += 0.5 * (mflux * q[3] + q[4] * az);
    fci[4] += 0.5 * Uk * (q[4] + q[nv]);
    for (int l = 5; l < nv; ++l) fci[l] += 0.5 * mflux * q[l];
  }
  const double wl = sqrt(pl[0]), wr = sqrt(pr[0]);
  const double wsum_inv = 1.0 / (wl + wr);
  auto avg = [&](double a, double b) { return (wl * a + wr * b) * wsum_inv; };
  const double u = avg(pl[1], pr[1]), v = avg(pl[2], pr[2]), w = avg(pl[3], pr[3]);
  const double ek = 0.5 * (u * u + v * v + w * w);
  const double H = avg((pl[nv] + pl[4]) / pl[0], (pr[nv] + pr[4]) / pr[0]);
  double Y[MS], hk[MS], cpk[MS];
  double r_mix{0};
  for (int l = 0; l < ns; ++l) {
    Y[l] = avg(pl[5 + l], pr[5 + l]);
    r_mix += Y[l] / T.mw[l];
  }
  const double t = avg(pl[4] / pl[0], pr[4] / pr[0]) / (kRu * r_mix);
  enthalpy_cp(T, t, hk, cpk);
  double cp_m{0}, cv_m{0};
  for (int l = 0; l < ns; ++l) {
    cp_m += cpk[l] * Y[l];
    cv_m += Y[l] * (cpk[l] - kRu / T.mw[l]);
  }
  const double gm = cp_m / cv_m, gm1 = gm - 1;
  const double c = sqrt(gm * kRu * r_mix * t);
  const double W_mix = 1.0 / r_mix;
  double nx = 0.5 * (ml[0] + mr[0]), ny = 0.5 * (ml[1] + mr[1]), nz = 0.5 * (ml[2] + mr[2]);
  const double kn = sqrt(nx * nx + ny * ny + nz * nz);
  double Un = nx * u + ny * v + nz * w;
  double lam[3] = {Un - kn * c, Un, Un + kn * c};
  const double dlt = 0.5 * (delta_l + delta_r);
  for (double& x : lam) {
    x = fabs(x);
    if (x < dlt) x = 0.5 * (x * x / dlt + dlt);
  }
  nx /= kn;
  ny /= kn;
  nz /= kn;
  Un /= kn;
  // jump of the conservative variables (times the mean Jacobian)
  const double jm = 0.5 * (jl + jr);
  double dQ[5 + MS];
  dQ[0] = jm * (pr[0] - pl[0]);
  for (int l = 1; l < nv; ++l) dQ[l] = jm * (pr[0] * pr[l] - pl[0] * pl[l]);
  dQ[4] = jm * (pr[nv] - pl[nv]);
  // wave strengths L dQ
  double a1 = gm1 * (ek * dQ[0] - u * dQ[1] - v * dQ[2] - w * dQ[3] + dQ[4]) / (c * c);
  const double a2 = (nx * dQ[1] + ny * dQ[2] + nz * dQ[3] - Un * dQ[0]) / c;
  for (int l = 0; l < ns; ++l) a1 += (W_mix / T.mw[l] - hk[l] * gm1 / (c * c)) * dQ[5 + l];
  const double a3 = dQ[0] - a1;
  double bw[5 + MS];
  bw[0] = -lam[0] * (0.5 * (a1 - a2));
  bw[1] = -lam[1] * (nx * a3 - ((nz * v - ny * w) * dQ[0] - nz * dQ[2] + ny * dQ[3]) / c);
  bw[2] = -lam[1] * (ny * a3 - ((nx * w - nz * u) * dQ[0] - nx * dQ[3] + nz * dQ[1]) / c);
  bw[3] = -lam[1] * (nz * a3 - ((ny * u - nx * v) * dQ[0] - ny * dQ[1] + nx * dQ[2]) / c);
  bw[4] = -lam[2] * (0.5 * (a1 + a2));
  for (int l = 0; l < ns; ++l) bw[5 + l] = -lam[1] * (dQ[5 + l] - Y[l] * dQ[0]);
  // R |Lambda| L dQ
  const double e0 = nx * bw[1] + ny * bw[2] + nz * bw[3];
  const double e1 = e0 + bw[0] + bw[4];
  const double e2 = c * (bw[4] - bw[0]);
  double e3 = 0;
  for (int l = 0; l < ns; ++l) e3 += (hk[l] - W_mix / T.mw[l] * c * c / gm1) * bw[5 + l];
  fci[0] += 0.5 * e1;
  fci[1] += 0.5 * (u * e1 + nx * e2 - c * (nz * bw[2] - ny * bw[3]));
  fci[2] += 0.5 * (v * e1 + ny * e2 - c * (nx * bw[3] - nz * bw[1]));
  fci[3] += 0.5 * (w * e1 + nz * e2 - c * (ny * bw[1] - nx * bw[2]));
  fci[4] += 0.5 * (H * e1 + Un * e2 - c * c * e0 / gm1 +
                   c * ((nz * v - ny * w) * bw[1] + (nx * w - nz * u) * bw[2] + (ny * u - nx * v) * bw[3]) + e3);
  for (int l = 0; l < ns; ++l) fci[5 + l] += 0.5 * (bw[5 + l] + Y[l] * e1);
}

// compute_convective_term_pv_1D<FR>, src/InviscidScheme.cu:50-165: per
// direction x, y, z, reconstructed primitives -> Riemann flux,
// dq(i) -= fc(i+1/2) - fc(i-1/2).
void inviscid_pv(const Ctx& c, Block& b) {
  const int nv = c.nv, ns = c.ns, g = b.g;
  const long long s = b.sz, n0 = b.n0();
  const int ext[3] = {b.mx, b.my, b.mz};
  for (int dir = 0; dir < 3; ++dir) {
    const int m = ext[dir];
    const int oa = (dir == 0) ? 1 : 0, ob = (dir == 2) ? 1 : 2;
    const int na = ext[oa], nb = ext[ob];
#pragma omp parallel
    {
      const int np = m + 2 * g;
      std::vector<double> PV(np * nv), K(np * 3), J(np), D(np), fc((m + 1) * nv);
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
            for (int l = 0; l < 5; ++l) PV[ps * nv + l] = b.bv[p + l * s];
            for (int l = 0; l < ns; ++l) PV[ps * nv + 5 + l] = b.sv[p + l * s];
            for (int cc = 0; cc < 3; ++cc) K[ps * 3 + cc] = b.metric[p + (3 * dir + cc) * s];
            J[ps] = b.jac[p];
            if (c.s.inviscid_scheme == 2) {
              double mm[9];
              for (int cc = 0; cc < 9; ++cc) mm[cc] = b.metric[p + cc * s];
              D[ps] = entropy_delta(b.bv[p + s], b.bv[p + 2 * s], b.bv[p + 3 * s], b.cc[p], mm, c.s.entropy_fix_factor);
            }
          }
          for (int t = -1; t < m; ++t) {
            const int is = t + g;
            double pl[5 + MS + 2], pr[5 + MS + 2];
            reconstruct(PV.data(), nv, is, c.s.reconstruction, pl, pr);
            pv_energy(c.T, nv, pl);
            pv_energy(c.T, nv, pr);
            double* fci = &fc[(t + 1) * nv];
            if (c.s.inviscid_scheme == 2)
              roe(c.T, nv, pl, pr, &K[is * 3], &K[(is + 1) * 3], J[is], J[is + 1], D[is], D[is + 1], fci);
            else if (c.s.inviscid_scheme == 3)
              ausm_plus(nv, pl, pr, &K[is * 3], &K[(is + 1) * 3], J[is], J[is + 1], fci);
            else
              hllc(c.T, nv, pl, pr, &K[is * 3], &K[(is + 1) * 3], J[is], J[is + 1], fci);
          }
          for (int t = 0; t < m; ++t) {
            idx[dir] = t;
            const long long p0 = b.id0(idx[0], idx[1], idx[2]);
            for (int l = 0; l < nv; ++l) b.dq[p0 + l * n0] -= fc[(t + 1) * nv + l] - fc[t * nv + l];
          }
        }
    }
  }
}

// compute_convective_term_weno_{x,y,z}<FR> with compute_weno_flux_cp (51)
// or compute_weno_flux_ch (52, "method 1"), src/WENO.cu:8-461, :722-1143,
// :1455-1536.  dq(i) -= fc(i+1/2) - fc(i-1/2) for x, y, z in that order.
void inviscid(const Ctx& c, Block& b, double dt) {
  if (c.s.inviscid_scheme == 2 || c.s.inviscid_scheme == 3 || c.s.inviscid_scheme == 4) {
    inviscid_pv(c, b);
    return;
  }
  const int nv = c.nv, ns = c.ns, g = b.g, scheme = c.s.inviscid_scheme;
  const long long s = b.sz, n0 = b.n0();
  const int ext[3] = {b.mx, b.my, b.mz};
  const bool pp = c.s.positive_preserving != 0;
  for (int dir = 0; dir < 3; ++dir) {
    const int m = ext[dir];
    const int oa = (dir == 0) ? 1 : 0, ob = (dir == 2) ? 1 : 2;
    const int na = ext[oa], nb = ext[ob];
#pragma omp parallel
    {
      const int np = m + 2 * g;
      std::vector<double> Q(np * (nv + 2)), Fp(np * nv), Fm(np * nv), fc((m + 1) * nv), f1((m + 1) * MS);
      std::vector<double> K(np * 3), J(np);
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
            double* q = &Q[ps * (nv + 2)];
            for (int l = 0; l < nv; ++l) q[l] = b.cv[p + l * s];
            q[nv] = b.bv[p + 4 * s];
            q[nv + 1] = b.cc[p];
            for (int cc = 0; cc < 3; ++cc) K[ps * 3 + cc] = b.metric[p + (3 * dir + cc) * s];
            J[ps] = b.jac[p];
            lf_split(q, nv, &K[ps * 3], J[ps], &Fp[ps * nv], &Fm[ps * nv]);
          }
          for (int t = -1; t < m; ++t) {
            const int is = t + g;
            const double *m1 = &K[is * 3], *m2 = &K[(is + 1) * 3];
            const double jac1 = J[is], jac2 = J[is + 1];
            double* fci = &fc[(t + 1) * nv];
            double* f1i = &f1[(t + 1) * MS];
            if (scheme == 52 || scheme == 72) {
              const double* cvl = &Q[is * (nv + 2)];
              const double* cvr = &Q[(is + 1) * (nv + 2)];
              CharFace F;
              char_face(c.T, nv, cvl, cvr, m1, m2, jac1, jac2, F);
              constexpr double eps{1e-40};
              const double eps_scaled = eps * c.weno_eps_scale * 0.25 * (F.kxJ * F.kxJ + F.kyJ * F.kyJ + F.kzJ * F.kzJ);
              if (pp) {
                const double srThis =
                    abs((m1[0] * cvl[1] + m1[1] * cvl[2] + m1[2] * cvl[3]) / cvl[0] + cvl[nv + 1] * sqrt(m1[0] * m1[0] + m1[1] * m1[1] + m1[2] * m1[2]));
                const double srNext =
                    abs((m2[0] * cvr[1] + m2[1] * cvr[2] + m2[2] * cvr[3]) / cvr[0] + cvr[nv + 1] * sqrt(m2[0] * m2[0] + m2[1] * m2[1] + m2[2] * m2[2]));
                for (int l = 0; l < ns; ++l)
                  f1i[l] = 0.5 * (Fp[is * nv + l + 5] + srThis * cvl[l + 5] * jac1) +
                           0.5 * (Fp[(is + 1) * nv + l + 5] - srNext * cvr[l + 5] * jac2);
              }
              double fChar[5 + MS];
              if (scheme == 72) {
                // WENO-Z7 characteristic, src/WENO.cu:898-965: 7 points
                // i-3..i+3 (F+) and i-2..i+4 (F-) for every field
                for (int l = 0; l < 5; ++l) {
                  double L[5];
                  const double cas = char_row(F, l, L);
                  double vPlus[7] = {}, vMinus[7] = {};
                  for (int r = 0; r < 7; ++r) {
                    for (int n = 0; n < 5; ++n) {
                      vPlus[r] += L[n] * Fp[(is - 3 + r) * nv + n];
                      vMinus[r] += L[n] * Fm[(is - 2 + r) * nv + n];
                    }
                    for (int n = 0; n < ns; ++n) {
                      vPlus[r] += cas * F.alpha_l[n] * Fp[(is - 3 + r) * nv + 5 + n];
                      vMinus[r] += cas * F.alpha_l[n] * Fm[(is - 2 + r) * nv + 5 + n];
                    }
                  }
                  fChar[l] = weno7z(vPlus, vMinus, eps_scaled);
                }
                for (int l = 0; l < ns; ++l) {
                  double vPlus[7], vMinus[7];
                  for (int r = 0; r < 7; ++r) {
                    vPlus[r] = -F.svm[l] * Fp[(is - 3 + r) * nv] + Fp[(is - 3 + r) * nv + 5 + l];
                    vMinus[r] = -F.svm[l] * Fm[(is - 2 + r) * nv] + Fm[(is - 2 + r) * nv + 5 + l];
                  }
                  fChar[5 + l] = weno7z(vPlus, vMinus, eps_scaled);
                }
              } else {
              // the five conservative components of the acoustic/entropy fields
              // read F+ at i-2..i+2 and F- at i-1..i+3, or with ch_stencil 1 the
              // reference's i-3..i+1 and i-2..i+2 (src/WENO.cu:880-881); the
              // point i-3 of the block's first face lies beyond the ghost
              // layers, and the outermost ghost point stands in for it
              const int sh = c.s.ch_stencil ? 1 : 0;
              for (int l = 0; l < 5; ++l) {
                double L[5];
                const double cas = char_row(F, l, L);
                double vPlus[5] = {}, vMinus[5] = {};
                for (int r = 0; r < 5; ++r) {
                  const int ip = std::max(is - 2 - sh + r, 0), im = is - 1 - sh + r;
                  for (int n = 0; n < 5; ++n) {
                    vPlus[r] += L[n] * Fp[ip * nv + n];
                    vMinus[r] += L[n] * Fm[im * nv + n];
                  }
                  for (int n = 0; n < ns; ++n) {
                    vPlus[r] += cas * F.alpha_l[n] * Fp[(is - 2 + r) * nv + 5 + n];
                    vMinus[r] += cas * F.alpha_l[n] * Fm[(is - 1 + r) * nv + 5 + n];
                  }
                }
                fChar[l] = weno5z(vPlus, vMinus, eps_scaled);
              }
              for (int l = 0; l < ns; ++l) {
                double vPlus[5], vMinus[5];
                for (int r = 0; r < 5; ++r) {
                  vPlus[r] = -F.svm[l] * Fp[(is - 2 + r) * nv] + Fp[(is - 2 + r) * nv + 5 + l];
                  vMinus[r] = -F.svm[l] * Fm[(is - 1 + r) * nv] + Fm[(is - 1 + r) * nv + 5 + l];
                }
                fChar[5 + l] = weno5z(vPlus, vMinus, eps_scaled);
              }
              }
              char_back(F, ns, fChar, fci);
            } else {
              constexpr double eps{1e-20};
              const double eps_ref = eps * c.weno_eps_scale * 0.25 *
                                     ((m1[0] * jac1 + m2[0] * jac2) * (m1[0] * jac1 + m2[0] * jac2) +
                                      (m1[1] * jac1 + m2[1] * jac2) * (m1[1] * jac1 + m2[1] * jac2) +
                                      (m1[2] * jac1 + m2[2] * jac2) * (m1[2] * jac1 + m2[2] * jac2));
              double eps_scaled[3];
              eps_scaled[0] = eps_ref;
              eps_scaled[1] = eps_ref * c.s.v_ref * c.s.v_ref;
              eps_scaled[2] = eps_scaled[1] * c.s.v_ref * c.s.v_ref;
              if (pp)
                for (int l = 0; l < ns; ++l) f1i[l] = 0.5 * Fp[is * nv + l + 5] + 0.5 * Fm[(is + 1) * nv + l + 5];
              for (int l = 0; l < nv; ++l) {
                double eh{eps_scaled[0]};
                if (l == 1 || l == 2 || l == 3) eh = eps_scaled[1];
                else if (l == 4) eh = eps_scaled[2];
                if (scheme == 71) {
                  // WENO-Z7 component-wise, src/WENO.cu:1516-1534
                  double vp[7], vm[7];
                  for (int r = 0; r < 7; ++r) {
                    vp[r] = Fp[(is - 3 + r) * nv + l];
                    vm[r] = Fm[(is - 2 + r) * nv + l];
                  }
                  fci[l] = weno7z(vp, vm, eh);
                  continue;
                }
                double vp[5], vm[5];
                for (int r = 0; r < 5; ++r) {
                  vp[r] = Fp[(is - 2 + r) * nv + l];
                  vm[r] = Fm[(is - 1 + r) * nv + l];
                }
                fci[l] = weno5z(vp, vm, eh);
              }
            }
            if (pp) {
              const double* cvl = &Q[is * (nv + 2)];
              const double* cvr = &Q[(is + 1) * (nv + 2)];
              pp_limiter(ns, fci + 5, f1i, cvl + 5, cvr + 5, jac1, jac2, dt, t, m);
            }
          }
          for (int t = 0; t < m; ++t) {
            idx[dir] = t;
            const long long p0 = b.id0(idx[0], idx[1], idx[2]);
            for (int l = 0; l < nv; ++l) b.dq[p0 + l * n0] -= fc[(t + 1) * nv + l] - fc[t * nv + l];
          }
        }
    }
  }
}

// compute_{fv,gv,hv}_2nd_order<FR, Laminar>, src/ViscousScheme.cuh:23-318,
// :320-616, :618-909 (gradPInDiffusionFlux optional)
void viscous_face(const Ctx& c, const Block& b, int dir, int i, int j, int k, double* fv) {
  const long long s = b.sz;
  const int ns = c.ns;
  int o[3] = {0, 0, 0};
  o[dir] = 1;
  const long long p = b.id(i, j, k), p1 = b.id(i + o[0], j + o[1], k + o[2]);
  auto M = [&](long long pt, int r, int cc) { return b.metric[pt + (3 * (r - 1) + (cc - 1)) * s]; };
  const double xi_x = 0.5 * (M(p, 1, 1) + M(p1, 1, 1));
  const double xi_y = 0.5 * (M(p, 1, 2) + M(p1, 1, 2));
  const double xi_z = 0.5 * (M(p, 1, 3) + M(p1, 1, 3));
  const double eta_x = 0.5 * (M(p, 2, 1) + M(p1, 2, 1));
  const double eta_y = 0.5 * (M(p, 2, 2) + M(p1, 2, 2));
  const double eta_z = 0.5 * (M(p, 2, 3) + M(p1, 2, 3));
  const double zeta_x = 0.5 * (M(p, 3, 1) + M(p1, 3, 1));
  const double zeta_y = 0.5 * (M(p, 3, 2) + M(p1, 3, 2));
  const double zeta_z = 0.5 * (M(p, 3, 3) + M(p1, 3, 3));
  // derivative of a padded plane f along (xi, eta, zeta) at the face
  auto deriv = [&](const double* f, double* d) {
    auto F = [&](int ii, int jj, int kk) { return f[b.id(ii, jj, kk)]; };
    if (dir == 0) {
      d[0] = F(i + 1, j, k) - F(i, j, k);
      d[1] = 0.25 * (F(i, j + 1, k) - F(i, j - 1, k) + F(i + 1, j + 1, k) - F(i + 1, j - 1, k));
      d[2] = 0.25 * (F(i, j, k + 1) - F(i, j, k - 1) + F(i + 1, j, k + 1) - F(i + 1, j, k - 1));
    } else if (dir == 1) {
      d[0] = 0.25 * (F(i + 1, j, k) - F(i - 1, j, k) + F(i + 1, j + 1, k) - F(i - 1, j + 1, k));
      d[1] = F(i, j + 1, k) - F(i, j, k);
      d[2] = 0.25 * (F(i, j, k + 1) - F(i, j, k - 1) + F(i, j + 1, k + 1) - F(i, j + 1, k - 1));
    } else {
      d[0] = 0.25 * (F(i + 1, j, k) - F(i - 1, j, k) + F(i + 1, j, k + 1) - F(i - 1, j, k + 1));
      d[1] = 0.25 * (F(i, j + 1, k) - F(i, j - 1, k) + F(i, j + 1, k + 1) - F(i, j - 1, k + 1));
      d[2] = F(i, j, k + 1) - F(i, j, k);
    }
  };
  double du[3], dv[3], dw[3], dT[3];
  deriv(&b.bv[1 * s], du);
  deriv(&b.bv[2 * s], dv);
  deriv(&b.bv[3 * s], dw);
  deriv(&b.bv[5 * s], dT);
  const double u_x = du[0] * xi_x + du[1] * eta_x + du[2] * zeta_x;
  const double u_y = du[0] * xi_y + du[1] * eta_y + du[2] * zeta_y;
  const double u_z = du[0] * xi_z + du[1] * eta_z + du[2] * zeta_z;
  const double v_x = dv[0] * xi_x + dv[1] * eta_x + dv[2] * zeta_x;
  const double v_y = dv[0] * xi_y + dv[1] * eta_y + dv[2] * zeta_y;
  const double v_z = dv[0] * xi_z + dv[1] * eta_z + dv[2] * zeta_z;
  const double w_x = dw[0] * xi_x + dw[1] * eta_x + dw[2] * zeta_x;
  const double w_y = dw[0] * xi_y + dw[1] * eta_y + dw[2] * zeta_y;
  const double w_z = dw[0] * xi_z + dw[1] * eta_z + dw[2] * zeta_z;
  const double mul = 0.5 * (b.mul[p] + b.mul[p1]);
  const double mut{0};
  const double viscosity = mul + mut;
  const double tau_xx = viscosity * (4 * u_x - 2 * v_y - 2 * w_z) / 3.0;
  const double tau_yy = viscosity * (4 * v_y - 2 * u_x - 2 * w_z) / 3.0;
  const double tau_zz = viscosity * (4 * w_z - 2 * u_x - 2 * v_y) / 3.0;
  const double tau_xy = viscosity * (u_y + v_x);
  const double tau_xz = viscosity * (u_z + w_x);
  const double tau_yz = viscosity * (v_z + w_y);
  const int r = dir + 1;
  const double kx = 0.5 * (M(p, r, 1) * b.jac[p] + M(p1, r, 1) * b.jac[p1]);
  const double ky = 0.5 * (M(p, r, 2) * b.jac[p] + M(p1, r, 2) * b.jac[p1]);
  const double kz = 0.5 * (M(p, r, 3) * b.jac[p] + M(p1, r, 3) * b.jac[p1]);
  fv[0] = 0;
  fv[1] = kx * tau_xx + ky * tau_xy + kz * tau_xz;
  fv[2] = kx * tau_xy + ky * tau_yy + kz * tau_yz;
  fv[3] = kx * tau_xz + ky * tau_yz + kz * tau_zz;
  const double um = 0.5 * (b.bv[p + s] + b.bv[p1 + s]);
  const double vm = 0.5 * (b.bv[p + 2 * s] + b.bv[p1 + 2 * s]);
  const double wm = 0.5 * (b.bv[p + 3 * s] + b.bv[p1 + 3 * s]);
  const double t_x = dT[0] * xi_x + dT[1] * eta_x + dT[2] * zeta_x;
  const double t_y = dT[0] * xi_y + dT[1] * eta_y + dT[2] * zeta_y;
  const double t_z = dT[0] * xi_z + dT[1] * eta_z + dT[2] * zeta_z;
  const double conductivity = 0.5 * (b.lam[p] + b.lam[p1]);
  fv[4] = um * fv[1] + vm * fv[2] + wm * fv[3] + conductivity * (kx * t_x + ky * t_y + kz * t_z);

  double diffusivity[MS], yk[MS], ddf[MS];
  double sum_grad_over_w{0}, sum_rhoDkYk{0}, corr{0}, mw_tot{0};
  for (int l = 0; l < ns; ++l) {
    yk[l] = 0.5 * (b.sv[p + l * s] + b.sv[p1 + l * s]);
    diffusivity[l] = 0.5 * (b.rhod[p + l * s] + b.rhod[p1 + l * s]) + 0.0;
    double dy[3];
    deriv(&b.sv[l * s], dy);
    const double y_x = dy[0] * xi_x + dy[1] * eta_x + dy[2] * zeta_x;
    const double y_y = dy[0] * xi_y + dy[1] * eta_y + dy[2] * zeta_y;
    const double y_z = dy[0] * xi_z + dy[1] * eta_z + dy[2] * zeta_z;
    const double gy = kx * y_x + ky * y_y + kz * y_z;
    ddf[l] = gy;
    corr += diffusivity[l] * gy;
    sum_grad_over_w += gy / c.T.mw[l];
    mw_tot += yk[l] / c.T.mw[l];
    sum_rhoDkYk += diffusivity[l] * yk[l];
  }
  mw_tot = 1.0 / mw_tot;
  corr -= mw_tot * sum_rhoDkYk * sum_grad_over_w;
  if (c.s.grad_p_diffusion) {
    double dp[3];
    deriv(&b.bv[4 * s], dp);
    const double p_x{dp[0] * xi_x + dp[1] * eta_x + dp[2] * zeta_x};
    const double p_y{dp[0] * xi_y + dp[1] * eta_y + dp[2] * zeta_y};
    const double p_z{dp[0] * xi_z + dp[1] * eta_z + dp[2] * zeta_z};
    const double gp = (kx * p_x + ky * p_y + kz * p_z) / (0.5 * (b.bv[p1 + 4 * s] + b.bv[p + 4 * s]));
    for (int l = 0; l < ns; ++l) {
      ddf[l] += (mw_tot / c.T.mw[l] - 1) * yk[l] * gp;
      corr += (mw_tot / c.T.mw[l] - 1) * yk[l] * gp * diffusivity[l];
    }
  }
  double h[MS];
  const double tm = 0.5 * (b.bv[p + 5 * s] + b.bv[p1 + 5 * s]);
  enthalpy(c.T, tm, h);
  for (int l = 0; l < ns; ++l) {
    const double flux{diffusivity[l] * (ddf[l] - mw_tot * yk[l] * sum_grad_over_w) - yk[l] * corr};
    fv[5 + l] = flux;
    fv[4] += h[l] * flux;
  }
}

void viscous(const Ctx& c, Block& b) {
  if (c.s.viscous_order == 0) return;
  const int nv = c.nv;
  const long long n0 = b.n0();
  const int ext[3] = {b.mx, b.my, b.mz};
  for (int dir = 0; dir < 3; ++dir) {
    const int m = ext[dir];
    const int oa = (dir == 0) ? 1 : 0, ob = (dir == 2) ? 1 : 2;
    const int na = ext[oa], nb = ext[ob];
#pragma omp parallel
    {
      std::vector<double> fv((m + 1) * nv);
#pragma omp for collapse(2) schedule(static)
      for (int bb = 0; bb < nb; ++bb)
        for (int aa = 0; aa < na; ++aa) {
          int idx[3];
          idx[oa] = aa;
          idx[ob] = bb;
          for (int t = -1; t < m; ++t) {
            idx[dir] = t;
            viscous_face(c, b, dir, idx[0], idx[1], idx[2], &fv[(t + 1) * nv]);
          }
          for (int t = 0; t < m; ++t) {
            idx[dir] = t;
            const long long p0 = b.id0(idx[0], idx[1], idx[2]);
            for (int l = 0; l < nv; ++l) b.dq[p0 + l * n0] += fv[(t + 1) * nv + l] - fv[t * nv + l];
          }
        }
    }
  }
}

// update_cv_and_bv_rk<FR>, src/RK.cuh:230-284, with the T-Newton of
// compute_temperature_and_pressure (starting from the cell's current T)
void update(const Ctx& c, Block& b, double dt, int rk) {
  const int nv = c.nv, ns = c.ns;
  const long long s = b.sz, n0 = b.n0();
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = b.id(i, j, k), p0 = b.id0(i, j, k);
        const double dt_div_jac = dt / b.jac[p];
        for (int l = 0; l < nv; ++l)
          b.cv[p + l * s] =
              kRkA[rk] * b.qn[p + l * s] + kRkB[rk] * b.cv[p + l * s] + kRkC[rk] * dt_div_jac * b.dq[p0 + l * n0];
        b.bv[p] = b.cv[p];
        const double density_inv = 1.0 / b.cv[p];
        b.bv[p + s] = b.cv[p + s] * density_inv;
        b.bv[p + 2 * s] = b.cv[p + 2 * s] * density_inv;
        b.bv[p + 3 * s] = b.cv[p + 3 * s] * density_inv;
        double Y[MS];
        for (int l = 0; l < ns; ++l) {
          b.sv[p + l * s] = b.cv[p + (5 + l) * s] * density_inv;
          Y[l] = b.sv[p + l * s];
        }
        double mw{0};
        for (int l = 0; l < ns; ++l) mw += Y[l] / c.T.mw[l];
        mw = 1 / mw;
        const double gas_const = kRu / mw;
        const double u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
        const double e = b.cv[p + 4 * s] / b.bv[p] - 0.5 * (u * u + v * v + w * w);
        const double t = temperature(c.T, e, Y, gas_const, b.bv[p + 5 * s], nullptr);
        b.bv[p + 5 * s] = t;
        b.bv[p + 4 * s] = b.bv[p] * t * gas_const;
      }
}

int bc_priority(int type) {  // src/BoundCond.cuh:1488-1678: farfield, outflow, inflow, ..., periodic
  return type == 4 ? 0 : (type == 6 ? 1 : (type == 5 ? 2 : 3));
}

// apply_inflow with mass fractions, src/BoundCond.cuh:235-612: inflow_type 2
// (the mixing-layer profile, :420-482) or the constant state (:483-522),
// white noise (fluctuation_type 1) keyed by the point's position in the
// padded face plane; the face point and its ghost layers take the same state.
void inflow_point(const Ctx& c, Block& b, const Face& f, int i, int j, int k, double y) {
  const orf_inflow& in = f.in;
  const int g = b.g;
  const long long s = b.sz;
  double u;
  bool upper = true;
  if (in.inflow_type == 2) {
    u = 0.5 * (in.upper[1] + in.lower[1]) + 0.5 * (in.upper[1] - in.lower[1]) * std::tanh(2 * y / in.delta_omega);
    upper = y >= 0;
  } else {
    u = in.upper[1];
  }
  const double* S = upper ? in.upper : in.lower;
  const double* sv = upper ? in.sv_upper : in.sv_lower;
  double v = S[2];
  const double w = S[3];
  if (in.fluctuation_type == 1) {
    const double vel = std::sqrt(u * u + v * v + w * w);
    const uint64_t point = f.axis == 0 ? (uint64_t)(k + g) * (b.my + 2 * g) + (j + g)
                           : f.axis == 1 ? (uint64_t)(k + g) * (b.mx + 2 * g) + (i + g)
                                         : (uint64_t)(j + g) * (b.mx + 2 * g) + (i + g);
    const double rms = in.fluctuation_intensity;
    double n1, n2;
    if (in.inflow_type == 2) {
      if (y < 4 * in.delta_omega && y > -4 * in.delta_omega) {
        orc_inflow_noise::normal_pair(in.seed, b.bc_passes, point, &n1, &n2);
        u += n1 * rms * vel;
        v += n2 * rms * vel;
      }
    } else {
      orc_inflow_noise::normal_pair(in.seed, b.bc_passes, point, &n1, &n2);
      u += n1 * rms * vel;
    }
  }
  int dir[3] = {0, 0, 0};
  dir[f.axis] = f.side;
  for (int gg = 0; gg <= g; ++gg) {
    const long long q = b.id(i + gg * dir[0], j + gg * dir[1], k + gg * dir[2]);
    b.bv[q] = S[0];
    b.bv[q + s] = u;
    b.bv[q + 2 * s] = v;
    b.bv[q + 3 * s] = w;
    b.bv[q + 4 * s] = S[4];
    b.bv[q + 5 * s] = S[5];
    for (int l = 0; l < c.ns; ++l) b.sv[q + l * s] = sv[l];
    cv_point(c, b, q);
  }
}

// apply_farfield<FR>, src/BoundCond.cuh:755-1022, one face point (i, j, k)
// and its ngg ghosts.  The face point's gamma comes from the property field;
// the subsonic state iterates gamma (from 1.4) until two successive values
// agree to 1e-4, with cp of the boundary temperature and composition.
void farfield_point(const Ctx& c, Block& b, const Face& f, int i, int j, int k) {
  const long long s = b.sz;
  const int ns = c.ns;
  const orf_farfield& ff = f.ff;
  int dir[3] = {0, 0, 0};
  dir[f.axis] = f.side;
  const long long p = b.id(i, j, k);
  double nx = b.metric[p + (3 * f.axis) * s], ny = b.metric[p + (3 * f.axis + 1) * s],
         nz = b.metric[p + (3 * f.axis + 2) * s];
  const double grad_n_inv = f.side / std::sqrt(nx * nx + ny * ny + nz * nz);
  nx *= grad_n_inv;
  ny *= grad_n_inv;
  nz *= grad_n_inv;
  const double u_b{b.bv[p + s]}, v_b{b.bv[p + 2 * s]}, w_b{b.bv[p + 3 * s]};
  const double u_face{nx * u_b + ny * v_b + nz * w_b};
  double sv_b[MS];
  for (int l = 0; l < ns; ++l) sv_b[l] = b.sv[p + l * s];
  const double gamma_b{b.gam[p]};
  double mw_inv{0};
  for (int l = 0; l < ns; ++l) mw_inv += sv_b[l] / c.T.mw[l];
  double mw = 1.0 / mw_inv;
  const double p_b{b.bv[p + 4 * s]}, rho_b{b.bv[p]};
  const double a_b{std::sqrt(gamma_b * p_b / rho_b)};
  const double mach_b{u_face / a_b};
  double st[6];
  bool face_too = true;
  if (mach_b <= -1) {
    st[0] = ff.density; st[1] = ff.u; st[2] = ff.v; st[3] = ff.w; st[4] = ff.pressure; st[5] = ff.temperature;
    for (int l = 0; l < ns; ++l) sv_b[l] = ff.sv[l];
  } else if (mach_b >= 1) {
    for (int l = 0; l < 6; ++l) st[l] = b.bv[p + l * s];
    face_too = false;
  } else {
    const double riemann_pos{u_face + 2 * a_b / (gamma_b - 1)};
    const double u_inf{ff.u * nx + ff.v * ny + ff.w * nz};
    const double riemann_neg{u_inf - 2 * ff.acoustic_speed / (ff.specific_heat_ratio - 1)};
    const double Un{0.5 * (riemann_pos + riemann_neg)};
    if (mach_b < 0) {
      st[1] = ff.u + (Un - u_inf) * nx;
      st[2] = ff.v + (Un - u_inf) * ny;
      st[3] = ff.w + (Un - u_inf) * nz;
      for (int l = 0; l < ns; ++l) sv_b[l] = ff.sv[l];
      mw = ff.mw;
    } else {
      st[1] = u_b + (Un - u_face) * nx;
      st[2] = v_b + (Un - u_face) * ny;
      st[3] = w_b + (Un - u_face) * nz;
    }
    double gamma{1.4}, err{1}, gamma_last;
    double density = 0, pressure = 0, temperature = 0;
    while (err > 1e-4) {
      gamma_last = gamma;
      const double c_b{0.25 * (gamma - 1) * (riemann_pos - riemann_neg)};
      const double s_b = mach_b <= 0 ? ff.entropy : p_b / std::pow(rho_b, gamma);
      density = std::pow(c_b * c_b / (gamma * s_b), 1 / (gamma - 1));
      pressure = density * c_b * c_b / gamma;
      temperature = pressure * mw / (density * kRu);
      double cp[MS];
      heat_capacity(c.T, temperature, cp);
      double cp_tot{0};
      for (int l = 0; l < ns; ++l) cp_tot += cp[l] * sv_b[l];
      gamma = cp_tot / (cp_tot - kRu / mw);
      err = std::abs(1 - gamma / gamma_last);
    }
    st[0] = density;
    st[4] = pressure;
    st[5] = temperature;
  }
  if (face_too) {
    for (int l = 0; l < 6; ++l) b.bv[p + l * s] = st[l];
    for (int l = 0; l < ns; ++l) b.sv[p + l * s] = sv_b[l];
    cv_point(c, b, p);
  }
  for (int g = 1; g <= b.g; ++g) {
    const long long pg = b.id(i + g * dir[0], j + g * dir[1], k + g * dir[2]);
    for (int l = 0; l < 6; ++l) b.bv[pg + l * s] = st[l];
    for (int l = 0; l < ns; ++l) b.sv[pg + l * s] = sv_b[l];
    cv_point(c, b, pg);
  }
}

// apply_outflow / apply_periodic with scalars, src/BoundCond.cuh:205-233, :1443-1486
void apply_bc(const Ctx& c, Block& b) {
  const long long s = b.sz;
  std::vector<Face> fs = b.faces;
  std::stable_sort(fs.begin(), fs.end(), [](const Face& x, const Face& y) {
    return bc_priority(x.type) < bc_priority(y.type);
  });
  const int ext[3] = {b.mx, b.my, b.mz};
  for (const Face& f : fs) {
    const int a = f.axis, o1 = (a + 1) % 3, o2 = (a + 2) % 3;
    const int face_idx = f.lo[a];
#pragma omp parallel for collapse(2) schedule(static)
    for (int t2 = f.lo[o2]; t2 <= f.hi[o2]; ++t2)
      for (int t1 = f.lo[o1]; t1 <= f.hi[o1]; ++t1) {
        int ix[3];
        ix[a] = face_idx;
        ix[o1] = t1;
        ix[o2] = t2;
        if (f.type == 4) {
          farfield_point(c, b, f, ix[0], ix[1], ix[2]);
          continue;
        }
        if (f.type == 5) {
          const long long q = ((long long)(ix[2] - f.lo[2]) * (f.hi[1] - f.lo[1] + 1) + (ix[1] - f.lo[1])) *
                                  (f.hi[0] - f.lo[0] + 1) + (ix[0] - f.lo[0]);
          inflow_point(c, b, f, ix[0], ix[1], ix[2], f.y[q]);
          continue;
        }
        int src[3] = {ix[0], ix[1], ix[2]};
        if (f.type == 8) src[a] = f.side < 0 ? ext[a] - 1 : 0;
        for (int g = 1; g <= b.g; ++g) {
          int gi[3] = {ix[0], ix[1], ix[2]}, si[3] = {src[0], src[1], src[2]};
          gi[a] += g * f.side;
          if (f.type == 8) si[a] += g * f.side;
          const long long pg = b.id(gi[0], gi[1], gi[2]), ps = b.id(si[0], si[1], si[2]);
          for (int l = 0; l < 6; ++l) b.bv[pg + l * s] = b.bv[ps + l * s];
          for (int l = 0; l < c.ns; ++l) b.sv[pg + l * s] = b.sv[ps + l * s];
          cv_point(c, b, pg);
        }
      }
  }
  ++b.bc_passes;
}

// inner_communication with scalars, src/DataCommunication.cuh:60-105: per
// face (blocks, then faces) the shared node layer takes the mean of both
// blocks, the ghost layers the target's values.
void communicate(Ctx& c) {
  for (int bi = 0; bi < (int)c.blocks.size(); ++bi) {
    Block& b = c.blocks[bi];
    for (const InnerFace& f : b.inner) {
      Block& t = c.blocks[f.target];
      const long long s = b.sz, ts = t.sz;
#pragma omp parallel for collapse(2) schedule(static)
      for (int n2 = 0; n2 < f.n_point[2]; ++n2)
        for (int n1 = 0; n1 < f.n_point[1]; ++n1)
          for (int n0 = 0; n0 < f.n_point[0]; ++n0) {
            const int n[3] = {n0, n1, n2};
            int ix[3], tx[3];
            bool shared;
            orc_faces::point_of(f, n, ix, tx, &shared);
            const long long p = b.id(ix[0], ix[1], ix[2]), pt = t.id(tx[0], tx[1], tx[2]);
            if (shared) {
              for (int l = 0; l < 6; ++l) {
                const double ave{0.5 * (t.bv[pt + l * ts] + b.bv[p + l * s])};
                b.bv[p + l * s] = ave;
                t.bv[pt + l * ts] = ave;
              }
              for (int l = 0; l < c.ns; ++l) {
                const double ave{0.5 * (t.sv[pt + l * ts] + b.sv[p + l * s])};
                b.sv[p + l * s] = ave;
                t.sv[pt + l * ts] = ave;
              }
            } else {
              for (int l = 0; l < 6; ++l) b.bv[p + l * s] = t.bv[pt + l * ts];
              for (int l = 0; l < c.ns; ++l) b.sv[p + l * s] = t.sv[pt + l * ts];
            }
            cv_point(c, b, p);
          }
    }
  }
}

// global_time_step, src/TimeAdvanceFunc.cu:39-64 (min_of_arr starts from 1e6)
double global_dt(Ctx& c) {
  double dt = 1e6;
  for (auto& b : c.blocks)
    for (double v : b.dtl) dt = std::min(dt, v);
  return dt;
}

void step(Ctx& c, double* dt_io) {
  double dt = c.s.fixed_time_step ? c.s.dt : c.dt_cur;
  for (auto& b : c.blocks) b.qn = b.cv;
  bool any_sponge = false;
  for (auto& b : c.blocks)
    if (b.sponge) {
      sponge_mean(c, b);
      any_sponge = true;
    }
  if (any_sponge) ++c.sponge_iter;
  for (int rk = 0; rk < 3; ++rk) {
    for (auto& b : c.blocks) {
      std::fill(b.dq.begin(), b.dq.end(), 0.0);
      source(c, b);
    }
    if (!c.s.fixed_time_step && rk == 0) {
      for (auto& b : c.blocks) local_dt(c, b);
      dt = global_dt(c);
      c.dt_cur = dt;
    }
    for (auto& b : c.blocks) {
      inviscid(c, b, dt);
      viscous(c, b);
      update(c, b, dt, rk);
      apply_bc(c, b);
    }
    communicate(c);
    for (auto& b : c.blocks) properties(c, b);
  }
  if (dt_io) *dt_io = dt;
}

Ctx* C(void* p) { return static_cast<Ctx*>(p); }

// Registers a boundary N.txt record (Boundary::register_boundary, src/Mesh.cpp:14-32).
int add_face(Ctx* c, int block, const int range[6], int type, const orf_farfield* ff) {
  if (block < 0 || block >= (int)c->blocks.size() || (type == 4 && !ff)) return -1;
  Block& b = c->blocks[block];
  const int ext[3] = {b.mx, b.my, b.mz};
  orc_faces::Boundary r;
  if (!orc_faces::register_boundary(range, c->ngg, ext, &r)) return -1;
  Face f{r.face, r.direction, type, ff ? *ff : orf_farfield{}, {}, {}, {}, {}};
  for (int i = 0; i < 3; ++i) {
    f.lo[i] = r.range_start[i];
    f.hi[i] = r.range_end[i];
  }
  b.faces.push_back(f);
  return 0;
}

void whole_face_record(const Block& b, int axis, int side, int r[6]) {
  const int ext[3] = {b.mx, b.my, b.mz};
  for (int i = 0; i < 3; ++i) {
    r[2 * i] = 0;
    r[2 * i + 1] = ext[i] - 1;
  }
  r[2 * axis] = r[2 * axis + 1] = side < 0 ? 0 : ext[axis] - 1;
}

}  // namespace

extern "C" {

void* orf_create(const orf_setup* s, const ostc_mech_tables* t) {
  if (!s || !t) return nullptr;
  const int sc = s->inviscid_scheme;
  if (sc != 51 && sc != 52 && sc != 71 && sc != 72 && sc != 2 && sc != 3 && sc != 4) return nullptr;
  Ctx* c = new Ctx();
  c->s = *s;
  c->T = *t;
  c->ns = t->n_spec;
  c->nv = 5 + c->ns;
  c->ngg = (s->inviscid_scheme == 71 || s->inviscid_scheme == 72) ? 4 : 3;  // src/Parameter.cpp:227-238
  c->weno_eps_scale = (s->rho_ref * s->v_ref) * (s->rho_ref * s->v_ref);
#ifdef _OPENMP
  if (s->n_threads > 0) omp_set_num_threads(s->n_threads);
#endif
  return c;
}

void orf_destroy(void* ctx) { delete C(ctx); }

int orf_add_block(void* ctx, int mx, int my, int mz) {
  Ctx* c = C(ctx);
  Block b;
  b.mx = mx;
  b.my = my;
  b.mz = mz;
  b.g = c->ngg;
  b.sx = mx + 2 * b.g;
  b.sy = my + 2 * b.g;
  b.sz = b.sx * b.sy * (mz + 2 * b.g);
  const long long sz = b.sz, n0 = b.n0();
  b.cv.assign(sz * c->nv, 0);
  b.qn.assign(sz * c->nv, 0);
  b.bv.assign(sz * 6, 0);
  b.sv.assign(sz * c->ns, 0);
  b.dq.assign(n0 * c->nv, 0);
  b.bv_last.assign(n0 * 4, 0);
  b.mul.assign(sz, 0);
  b.lam.assign(sz, 0);
  b.cp.assign(sz, 0);
  b.gam.assign(sz, 0);
  b.cc.assign(sz, 0);
  b.mach.assign(sz, 0);
  b.rhod.assign(sz * c->ns, 0);
  b.jac.assign(sz, 0);
  b.metric.assign(sz * 9, 0);
  b.dtl.assign(n0, 0);
  c->blocks.push_back(std::move(b));
  return (int)c->blocks.size() - 1;
}

int orf_ngg(void* ctx) { return C(ctx)->ngg; }

int orf_add_boundary(void* ctx, int block, int axis, int side, int type) {
  int r[6];
  whole_face_record(C(ctx)->blocks[block], axis, side, r);
  return add_face(C(ctx), block, r, type, nullptr);
}

int orf_add_farfield(void* ctx, int block, int axis, int side, const orf_farfield* ff) {
  int r[6];
  whole_face_record(C(ctx)->blocks[block], axis, side, r);
  return add_face(C(ctx), block, r, 4, ff);
}

int orf_add_inflow(void* ctx, int block, const int range[6], const orf_inflow* in, const double* y) {
  Ctx* c = C(ctx);
  if (!in || !y || add_face(c, block, range, 5, nullptr) != 0) return -1;
  Block& b = c->blocks[block];
  Face& f = b.faces.back();
  f.in = *in;
  const long long sx = b.mx + 2 * b.g, sy = b.my + 2 * b.g;
  for (int k = f.lo[2]; k <= f.hi[2]; ++k)
    for (int j = f.lo[1]; j <= f.hi[1]; ++j)
      for (int i = f.lo[0]; i <= f.hi[0]; ++i)
        f.y.push_back(y[((long long)(k + b.g) * sy + (j + b.g)) * sx + (i + b.g)]);
  return 0;
}

int orf_set_sponge(void* ctx, int block, const orf_sponge* sp, const double* x, const double* y, const double* z) {
  Ctx* c = C(ctx);
  if (block < 0 || block >= (int)c->blocks.size() || !sp || !x || !y || !z) return -1;
  Block& b = c->blocks[block];
  b.sponge = true;
  b.sp = *sp;
  const long long n0 = b.n0(), sx = b.mx + 2 * b.g, sy = b.my + 2 * b.g;
  b.xs.assign(n0, 0);
  b.ys.assign(n0, 0);
  b.zs.assign(n0, 0);
  b.sponge_mean_cv.assign(n0 * c->nv, 0.0);
  for (int k = 0; k < b.mz; ++k)
    for (int j = 0; j < b.my; ++j)
      for (int i = 0; i < b.mx; ++i) {
        const long long p = ((long long)(k + b.g) * sy + (j + b.g)) * sx + (i + b.g);
        b.xs[b.id0(i, j, k)] = x[p];
        b.ys[b.id0(i, j, k)] = y[p];
        b.zs[b.id0(i, j, k)] = z[p];
      }
  return 0;
}

int orf_add_boundary_range(void* ctx, int block, const int range[6], int type, const orf_farfield* ff) {
  return add_face(C(ctx), block, range, type, ff);
}

// compute_viscosity (host), src/Transport.cu:11-41
double mixture_viscosity(const Tab& T, double temperature, double mw_total, const double* Y) {
  const int ns = T.n_spec;
  double x[MS], vis[MS];
  for (int i = 0; i < ns; ++i) {
    x[i] = Y[i] * mw_total / T.mw[i];
    const double t_dl{temperature * T.lj_potent_inv[i]};
    const double collision_integral{1.147 * std::pow(t_dl, -0.145) + std::pow(t_dl + 0.5, -2)};
    vis[i] = T.vis_coeff[i] * std::sqrt(temperature) / collision_integral;
  }
  double viscosity{0};
  for (int i = 0; i < ns; ++i) {
    double vis_temp{0};
    for (int j = 0; j < ns; ++j) {
      double phi = 1.0;
      if (i != j) {
        const double numerator{1 + std::sqrt(vis[i] / vis[j]) * T.wj_div_wi_to_one4th[i][j]};
        phi = numerator * numerator * T.sqrt_wi_div_wj_pl1_mul8[i][j];
      }
      vis_temp += phi * x[j];
    }
    viscosity += vis[i] * x[i] / vis_temp;
  }
  return viscosity;
}

// FarField::FarField with species (laminar), src/BoundCond.cpp:586-700
int orf_farfield_state(const ostc_mech_tables* t, const orf_farfield_spec* in, const double* y, orf_farfield* out) {
  const Tab& T = *t;
  const int ns = T.n_spec;
  orf_farfield f{};
  double mach = in->mach, pressure = in->pressure, temperature = in->temperature, velocity = in->velocity,
         density = in->density;
  double u = in->u, v = in->v, w = in->w;
  for (int i = 0; i < ns; ++i) f.sv[i] = y[i];
  double mw = 0;
  for (int i = 0; i < ns; ++i) mw += f.sv[i] / T.mw[i];
  mw = 1 / mw;
  if (temperature < 0) temperature = pressure * mw / (density * kRu);
  const double viscosity = mixture_viscosity(T, temperature, mw, f.sv);
  double cpi[MS];
  heat_capacity(T, temperature, cpi);
  double cp{0}, cv{0};
  for (int i = 0; i < ns; ++i) {
    cp += f.sv[i] * cpi[i];
    cv += f.sv[i] * (cpi[i] - kRu / T.mw[i]);
  }
  const double gamma = cp / cv;
  const double acoustic_speed = std::sqrt(gamma * kRu / mw * temperature);
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

int orf_add_inner_face_range(void* ctx, int block, const int range[6], const int target_range[6], int target) {
  Ctx* c = C(ctx);
  if (block < 0 || block >= (int)c->blocks.size() || target < 0 || target >= (int)c->blocks.size()) return -1;
  const Block &b = c->blocks[block], &t = c->blocks[target];
  const int ext[3] = {b.mx, b.my, b.mz}, text[3] = {t.mx, t.my, t.mz};
  InnerFace f;
  if (!orc_faces::register_inner(range, target_range, c->ngg, target, &f)) return -1;
  if (!orc_faces::inner_in_bounds(f, ext, text, c->ngg)) return -1;
  c->blocks[block].inner.push_back(f);
  return 0;
}

int orf_add_inner_face(void* ctx, int block, int axis, int side, int target) {
  Ctx* c = C(ctx);
  if (block < 0 || block >= (int)c->blocks.size() || target < 0 || target >= (int)c->blocks.size()) return -1;
  const Block &b = c->blocks[block], &t = c->blocks[target];
  const int ext[3] = {b.mx, b.my, b.mz}, text[3] = {t.mx, t.my, t.mz};
  int src[6], tgt[6];
  orc_faces::aligned_records(axis, side, ext, text, src, tgt);
  return orf_add_inner_face_range(ctx, block, src, tgt, target);
}

double* orf_array(void* ctx, int block, int which, long long* n) {
  Block& b = C(ctx)->blocks[block];
  std::vector<double>* v = nullptr;
  switch (which) {
    case ORF_CV: v = &b.cv; break;
    case ORF_BV: v = &b.bv; break;
    case ORF_SV: v = &b.sv; break;
    case ORF_QN: v = &b.qn; break;
    case ORF_DQ: v = &b.dq; break;
    case ORF_MUL: v = &b.mul; break;
    case ORF_LAMBDA: v = &b.lam; break;
    case ORF_CP: v = &b.cp; break;
    case ORF_GAMMA: v = &b.gam; break;
    case ORF_C: v = &b.cc; break;
    case ORF_MACH: v = &b.mach; break;
    case ORF_RHOD: v = &b.rhod; break;
    case ORF_JAC: v = &b.jac; break;
    case ORF_METRIC: v = &b.metric; break;
    case ORF_DT_LOCAL: v = &b.dtl; break;
    default: return nullptr;
  }
  if (n) *n = (long long)v->size();
  return v->data();
}

// store_last_step and compute_residual (src/TimeAdvanceFunc.cu:6-37,
// src/Residual.cuh:12-187): RMS over every block's interior of the change of
// rho, |V|, p and T since the last store.
int orf_store_last_step(void* ctx) {
  for (auto& b : C(ctx)->blocks) {
    const long long s = b.sz, n0 = b.n0();
    for (int k = 0; k < b.mz; ++k)
      for (int j = 0; j < b.my; ++j)
        for (int i = 0; i < b.mx; ++i) {
          const long long p = b.id(i, j, k), q = b.id0(i, j, k);
          const double u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
          b.bv_last[q] = b.bv[p];
          b.bv_last[q + n0] = std::sqrt(u * u + v * v + w * w);
          b.bv_last[q + 2 * n0] = b.bv[p + 4 * s];
          b.bv_last[q + 3 * n0] = b.bv[p + 5 * s];
        }
  }
  return 0;
}

int orf_residual(void* ctx, double res[4]) {
  double acc[4] = {0, 0, 0, 0};
  long long n_total = 0;
  for (auto& b : C(ctx)->blocks) {
    const long long s = b.sz, n0 = b.n0();
    n_total += n0;
    for (int k = 0; k < b.mz; ++k)
      for (int j = 0; j < b.my; ++j)
        for (int i = 0; i < b.mx; ++i) {
          const long long p = b.id(i, j, k), q = b.id0(i, j, k);
          const double u = b.bv[p + s], v = b.bv[p + 2 * s], w = b.bv[p + 3 * s];
          const double d[4] = {b.bv[p] - b.bv_last[q], std::sqrt(u * u + v * v + w * w) - b.bv_last[q + n0],
                               b.bv[p + 4 * s] - b.bv_last[q + 2 * n0], b.bv[p + 5 * s] - b.bv_last[q + 3 * n0]};
          for (int l = 0; l < 4; ++l) acc[l] += d[l] * d[l];
        }
  }
  for (int l = 0; l < 4; ++l) res[l] = std::sqrt(acc[l] / (double)n_total);
  return 0;
}

int orf_initialize(void* ctx) {
  Ctx* c = C(ctx);
  for (auto& b : c->blocks) cv_from_bv(*c, b);
  // the farfield faces read gamma of their face points; the reference reads
  // a property field nothing has written yet, here it holds the initial state's
  for (auto& b : c->blocks) properties(*c, b);
  for (auto& b : c->blocks) apply_bc(*c, b);
  communicate(*c);
  for (auto& b : c->blocks) properties(*c, b);
  return 0;
}

int orf_step(void* ctx, int n, double* last_dt) {
  Ctx* c = C(ctx);
  for (int i = 0; i < n; ++i) step(*c, last_dt);
  return 0;
}

int orf_cv_from_bv(void* ctx, int b) { cv_from_bv(*C(ctx), C(ctx)->blocks[b]); return 0; }
int orf_properties(void* ctx, int b) { properties(*C(ctx), C(ctx)->blocks[b]); return 0; }
int orf_zero_dq(void* ctx, int b) {
  auto& d = C(ctx)->blocks[b].dq;
  std::fill(d.begin(), d.end(), 0.0);
  return 0;
}
int orf_source(void* ctx, int b) { source(*C(ctx), C(ctx)->blocks[b]); return 0; }
int orf_inviscid(void* ctx, int b, double dt) { inviscid(*C(ctx), C(ctx)->blocks[b], dt); return 0; }
int orf_viscous(void* ctx, int b) { viscous(*C(ctx), C(ctx)->blocks[b]); return 0; }
int orf_local_dt(void* ctx, int b) { local_dt(*C(ctx), C(ctx)->blocks[b]); return 0; }
double orf_global_dt(void* ctx) { return global_dt(*C(ctx)); }
int orf_update(void* ctx, int b, double dt, int rk) { update(*C(ctx), C(ctx)->blocks[b], dt, rk); return 0; }
int orf_apply_bc(void* ctx, int b) { apply_bc(*C(ctx), C(ctx)->blocks[b]); return 0; }
int orf_communicate(void* ctx) { communicate(*C(ctx)); return 0; }

void orf_enthalpy_cp(const ostc_mech_tables* t, double temp, double* h, double* cp) { enthalpy_cp(*t, temp, h, cp); }
void orf_gibbs_rt(const ostc_mech_tables* t, double temp, double* g) { gibbs_rt(*t, temp, g); }
void orf_chemical_source(const ostc_mech_tables* t, double rho, double temp, const double* y, double* omega,
                         double* kf, double* kb) {
  chemistry(*t, rho, temp, y, omega, kf, kb);
}
void orf_transport(const ostc_mech_tables* t, double temp, double p, double rho, const double* y, double* mu,
                   double* lambda, double* rho_d) {
  double cp[MS], mw = 0;
  heat_capacity(*t, temp, cp);
  for (int l = 0; l < t->n_spec; ++l) mw += y[l] / t->mw[l];
  transport(*t, temp, p, rho, y, 1 / mw, cp, mu, lambda, rho_d);
}
double orf_temperature(const ostc_mech_tables* t, double e, const double* y, double t_start, int* iters) {
  double mw = 0;
  for (int l = 0; l < t->n_spec; ++l) mw += y[l] / t->mw[l];
  return temperature(*t, e, y, kRu * mw, t_start, iters);
}
int orf_riemann(const ostc_mech_tables* t, int scheme, const double* ql, const double* qr, const double* kl,
                const double* kr, double jl, double jr, double* fc) {
  const int nv = 5 + t->n_spec;
  double pl[5 + MS + 2], pr[5 + MS + 2];
  for (int l = 0; l < nv; ++l) {
    pl[l] = ql[l];
    pr[l] = qr[l];
  }
  pv_energy(*t, nv, pl);
  pv_energy(*t, nv, pr);
  if (scheme == 3) ausm_plus(nv, pl, pr, kl, kr, jl, jr, fc);
  else hllc(*t, nv, pl, pr, kl, kr, jl, jr, fc);
  return 0;
}

// one Roe face flux from the primitives rho, u, v, w, p, Y_k of both sides,
// the face-normal metric rows, the Jacobians and the entropy-fix deltas
int orf_roe(const ostc_mech_tables* t, const double* ql, const double* qr, const double* kl, const double* kr,
            double jl, double jr, double delta_l, double delta_r, double* fc) {
  const int nv = 5 + t->n_spec;
  double pl[5 + MS + 2], pr[5 + MS + 2];
  for (int l = 0; l < nv; ++l) {
    pl[l] = ql[l];
    pr[l] = qr[l];
  }
  pv_energy(*t, nv, pl);
  pv_energy(*t, nv, pr);
  roe(*t, nv, pl, pr, kl, kr, jl, jr, delta_l, delta_r, fc);
  return 0;
}

void orf_reconstruct(const double* pv, int nv, int recon, double* pl, double* pr) {
  reconstruct(pv, nv, 1, recon, pl, pr);
}

int orf_char_matrices(const ostc_mech_tables* t, const double* ql, const double* qr, const double* kl,
                      const double* kr, double jl, double jr, double* L, double* R) {
  const int ns = t->n_spec, nv = 5 + ns;
  CharFace F;
  char_face(*t, nv, ql, qr, kl, kr, jl, jr, F);
  for (int l = 0; l < nv * nv; ++l) L[l] = 0;
  for (int l = 0; l < 5; ++l) {
    double row[5];
    const double cas = char_row(F, l, row);
    for (int n = 0; n < 5; ++n) L[l * nv + n] = row[n];
    for (int n = 0; n < ns; ++n) L[l * nv + 5 + n] = cas * F.alpha_l[n];
  }
  for (int l = 0; l < ns; ++l) {
    L[(5 + l) * nv + 0] = -F.svm[l];
    L[(5 + l) * nv + 5 + l] = 1;
  }
  // columns of R: the back projection of unit characteristic fluxes
  std::vector<double> e(nv), col(nv);
  for (int c = 0; c < nv; ++c) {
    std::fill(e.begin(), e.end(), 0.0);
    e[c] = 1;
    char_back(F, ns, e.data(), col.data());
    for (int r = 0; r < nv; ++r) R[r * nv + c] = col[r];
  }
  return 0;
}

}  // extern "C"
