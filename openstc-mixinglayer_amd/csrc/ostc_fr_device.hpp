// ostc_fr_device.hpp -- point physics of the finite-rate mixture path on the
// device: NASA-7 thermodynamics, mixture-averaged transport, finite-rate
// chemistry, the temperature Newton iteration and the Roe-averaged
// characteristic decomposition of a face.  Templated on the species count NS
// so per-species temporaries live in registers; the coefficient tables
// (ostc_mech_tables) are read through a kernel-argument pointer, which the
// compiler turns into scalar loads wherever the index is wave-uniform.
//
// Arithmetic follows the reference routine named beside each function, in
// its expression order (the library is built with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>

#include "openstc/mech.h"
#include "ostc_select.hpp"

namespace ostc {
namespace fr {

constexpr double kRu = 8314.462618;   // src/Constants.h:19
constexpr double kPatm = 101325;      // src/Constants.h:26

using Tab = ostc_mech_tables;

// compute_enthalpy (Combustion2Part), src/Thermo.cu:130-148
template <int NS>
__device__ __forceinline__ void enthalpy(const Tab* __restrict__ T, double t, double* h) {
  const double t2{t * t}, t3{t2 * t}, t4{t3 * t}, t5{t4 * t};
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    if (t < T->t_low[i]) {
      const double tt = T->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double* c = T->low_coeff[i];
      h[i] = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      const double cp = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      h[i] += cp * (t - tt);
    } else {
      const double* c = t < T->t_mid[i] ? T->low_coeff[i] : T->high_coeff[i];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
    }
    h[i] *= kRu / T->mw[i];
  }
}

// compute_enthalpy_and_cp (Combustion2Part), src/Thermo.cu:150-172
template <int NS>
__device__ __forceinline__ void enthalpy_cp(const Tab* __restrict__ T, double t, double* h, double* cp) {
  const double t2{t * t}, t3{t2 * t}, t4{t3 * t}, t5{t4 * t};
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    if (t < T->t_low[i]) {
      const double tt = T->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double* c = T->low_coeff[i];
      h[i] = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      h[i] += cp[i] * (t - tt);
    } else {
      const double* c = t < T->t_mid[i] ? T->low_coeff[i] : T->high_coeff[i];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    h[i] *= kRu / T->mw[i];
    cp[i] *= kRu / T->mw[i];
  }
}

// compute_cp (Combustion2Part), src/Thermo.cu:174-188
template <int NS>
__device__ __forceinline__ void heat_capacity(const Tab* __restrict__ T, double t, double* cp) {
  const double t2{t * t}, t3{t2 * t}, t4{t3 * t};
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    if (t < T->t_low[i]) {
      const double tt = T->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
      const double* c = T->low_coeff[i];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
    } else {
      const double* c = t < T->t_mid[i] ? T->low_coeff[i] : T->high_coeff[i];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    cp[i] *= kRu / T->mw[i];
  }
}

// compute_gibbs_div_rt (Combustion2Part), src/Thermo.cu:190-208
template <int NS>
__device__ __forceinline__ void gibbs_rt(const Tab* __restrict__ T, double t, double* g) {
  const double t2{t * t}, t3{t2 * t}, t4{t3 * t}, t_inv{1 / t}, log_t{log(t)};
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    if (t < T->t_low[i]) {
      const double tt = T->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt_inv = 1 / tt, log_tt = log(tt);
      const double* c = T->low_coeff[i];
      g[i] = c[0] * (1.0 - log_tt) - 0.5 * c[1] * tt - c[2] * tt2 / 6.0 - c[3] * tt3 / 12.0 - c[4] * tt4 * 0.05 +
             c[5] * tt_inv - c[6];
    } else {
      const double* c = t < T->t_mid[i] ? T->low_coeff[i] : T->high_coeff[i];
      g[i] = c[0] * (1.0 - log_t) - 0.5 * c[1] * t - c[2] * t2 / 6.0 - c[3] * t3 / 12.0 - c[4] * t4 * 0.05 +
             c[5] * t_inv - c[6];
    }
  }
}

// compute_transport_property, src/Transport.cu:43-136, with compute_Omega_D
// (:168-170).  Only the off-diagonal binary diffusivities are formed: the
// mixture-averaged rho D_k never reads the diagonal.  The reference's powers
// are evaluated as exp(b (log T + log c)) with log c tabulated per species and
// species pair (tlog: log(lj_potent_inv[l]), then log(kb_over_eps_jk[m][n])
// at 16 + 16 m + n), and (x + 0.5)^-2 as 1 / ((x + 0.5)^2): one logarithm per
// point instead of 54 pow calls, within a few ulp of pow.
template <int NS>
__device__ __forceinline__ void transport(const Tab* __restrict__ T, const double* __restrict__ tlog, double t,
                                          double p, double rho, const double* Y, double mw_total, const double* cp,
                                          double* mu_out, double* lambda_out, double* rho_d) {
  double X[NS], vis[NS];
  const double log_t = log(t);
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    X[l] = Y[l] * mw_total / T->mw[l];
    const double t_dl{t * T->lj_potent_inv[l]};
    const double collision_integral{1.147 * exp(-0.145 * (log_t + tlog[l])) + 1.0 / ((t_dl + 0.5) * (t_dl + 0.5))};
    vis[l] = T->vis_coeff[l] * sqrt(t) / collision_integral;
  }
  double viscosity = 0, conductivity = 0;
#pragma unroll
  for (int m = 0; m < NS; ++m) {
    double vis_temp{0};
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      double partition_func{1.0};
      if (m != n) {
        const double numerator{1 + sqrt(vis[m] / vis[n]) * T->wj_div_wi_to_one4th[m][n]};
        partition_func = numerator * numerator * T->sqrt_wi_div_wj_pl1_mul8[m][n];
      }
      vis_temp += partition_func * X[n];
    }
    const double cond_temp = 1.065 * vis_temp - 0.065 * X[m];
    viscosity += vis[m] * X[m] / vis_temp;
    const double lambda = vis[m] * (cp[m] + 1.25 * kRu / T->mw[m]);
    conductivity += lambda * X[m] / cond_temp;
  }
  *mu_out = viscosity;
  *lambda_out = conductivity;
  const double t_3over2_over_p{t * sqrt(t) / p};
  double dij[NS][NS];  // binary diffusivities D_mn, symmetric
#pragma unroll
  for (int m = 0; m < NS; ++m)
#pragma unroll
    for (int n = m + 1; n < NS; ++n) {
      const double t_red{t * T->kb_over_eps_jk[m][n]};
      const double om = exp(-0.145 * (log_t + tlog[16 + 16 * m + n])) + 1.0 / ((t_red + 0.5) * (t_red + 0.5));
      const double d = T->binary_diffusivity_coeff[m][n] * t_3over2_over_p / om;
      dij[m][n] = d;
      dij[n][m] = d;
    }
  constexpr double eps{1e-12};
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    double num{0}, den{0};
#pragma unroll
    for (int n = 0; n < NS; ++n)
      if (l != n) {
        num += (X[n] + eps) * T->mw[n];
        den += (X[n] + eps) / dij[l][n];
      }
    rho_d[l] = rho * num / (den * mw_total);
  }
}

// arrhenius, src/FiniteRateChem.cu:104-106: A * T^b * exp(-Ea/T), evaluated
// as (A * pow(T, b)) * exp(-Ea / T).  A zero exponent or activation energy
// makes its factor exactly 1, and multiplying by 1 is exact, so those factors
// are skipped (the tables make the test wave-uniform).
__device__ __forceinline__ double arrhenius(double t, double A, double b, double Ea) {
  double r = A;
  if (b != 0) r = A * pow(t, b);
  if (Ea != 0) r = r * exp(-Ea / t);
  return r;
}

// c^nu for a stoichiometric coefficient: pow(c, 0) = 1 and pow(c, 1) = c
// exactly; c^2 is the correctly rounded square.
__device__ __forceinline__ double stoich_pow(double c, int nu) {
  if (nu == 1) return c;
  if (nu == 2) return c * c;
  return pow(c, (double)nu);
}

// finite_rate_chemistry (explicit), src/FiniteRateChem.cu:7-58 with
// forward_reaction_rate :60-102, backward_reaction_rate :108-150,
// rate_of_progress :152-177, chemical_source :179-195.  The reaction loop
// runs over the runtime reaction count; its tables are wave-uniform.
template <int NS>
__device__ void chemistry(const Tab* __restrict__ T, double rho, double t, const double* Y, double* omega) {
  const int nr = T->n_reac;
  double c[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) c[l] = rho * Y[l] / T->mw[l] * 1e-3;
  double g[NS];
  gibbs_rt<NS>(T, t, g);
  constexpr double temp_p = kPatm / kRu * 1e-3;
  const double temp_t = temp_p / t;
  double creation[NS], destruction[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) creation[l] = destruction[l] = 0;
  // equilibrium kb is used only if some reaction is neither irreversible nor
  // REV (src/FiniteRateChem.cu:110-128)
  bool use_kc = false;
  for (int i = 0; i < nr; ++i) use_kc |= (T->label[i] != 0 && T->rev_type[i] != 1);
  for (int i = 0; i < nr; ++i) {
    const int type = T->label[i];
    double kf = arrhenius(t, T->A[i], T->b[i], T->Ea[i]);
    if (type == 3) {
      kf += arrhenius(t, T->A2[i], T->b2[i], T->Ea2[i]);
    } else if (type > 3) {
      double cc{0};
#pragma unroll
      for (int l = 0; l < NS; ++l) cc += c[l] * T->third_body_coeff[i][l];
      if (type == 4) {
        kf *= cc;
      } else {
        const double kf_low = arrhenius(t, T->A2[i], T->b2[i], T->Ea2[i]);
        const double kf_high = kf;
        const double reduced_pressure = kf_low * cc / kf_high;
        double F = 1.0;
        if (type > 5) {
          double f_cent = (1 - T->troe_alpha[i]) * exp(-t / T->troe_t3[i]) + T->troe_alpha[i] * exp(-t / T->troe_t1[i]);
          if (type == 7) f_cent += exp(-T->troe_t2[i] / t);
          const double logFc = log10(f_cent);
          const double cc2 = -0.4 - 0.67 * logFc;
          const double n = 0.75 - 1.27 * logFc;
          const double logPr = log10(reduced_pressure);
          const double tempo = (logPr + cc2) / (n - 0.14 * (logPr + cc2));
          const double pw = logFc / (1.0 + tempo * tempo);
          F = pow(10.0, pw);
        }
        kf = kf_high * reduced_pressure / (1.0 + reduced_pressure) * F;
      }
    }
    double kb = 0;
    if (type != 0) {
      if (T->rev_type[i] == 1) {
        kb = arrhenius(t, T->A2[i], T->b2[i], T->Ea2[i]);
        if (type == 4) {
          double cc{0};
#pragma unroll
          for (int l = 0; l < NS; ++l) cc += c[l] * T->third_body_coeff[i][l];
          kb *= cc;
        }
      }
      // then every reversible label takes kb from equilibrium (the
      // reference's test at src/FiniteRateChem.cu:140 only excludes labels 0
      // and 2), REV included
      if (use_kc) {
        double d_gibbs{0};
#pragma unroll
        for (int l = 0; l < NS; ++l) d_gibbs += g[l] * (T->stoi_b[i][l] - T->stoi_f[i][l]);
        const int ord = T->order[i];
        const double tp = ord == 0 ? 1.0 : (ord == 1 ? temp_t : pow(temp_t, (double)ord));
        const double kc{tp * exp(-d_gibbs)};
        kb = kf / kc;
      }
    }
    // rate of progress: factors with a zero coefficient are exactly 1
    double q1 = 1.0, q2 = type != 0 ? 1.0 : 0.0;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int nf = T->stoi_f[i][j], nb = T->stoi_b[i][j];
      if (nf != 0) q1 *= stoich_pow(c[j], nf);
      if (type != 0 && nb != 0) q2 *= stoich_pow(c[j], nb);
    }
    q1 *= kf;
    if (type != 0) q2 *= kb;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      creation[l] += q2 * T->stoi_f[i][l] + q1 * T->stoi_b[i][l];
      destruction[l] += q1 * T->stoi_f[i][l] + q2 * T->stoi_b[i][l];
    }
  }
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    const double cr = creation[l] * (T->mw[l] * 1e+3);
    const double de = destruction[l] * (T->mw[l] * 1e+3);
    omega[l] = cr - de;
  }
}

// compute_temperature_and_pressure, src/FieldOperation.cu:3-40
template <int NS>
__device__ __forceinline__ double temperature(const Tab* __restrict__ T, double e, const double* Y, double gas_const,
                                              double t) {
  double err{1};
  int iter = 0;
  while (err > 1e-3 && iter++ < 1000) {
    double h_i[NS], cp_i[NS];
    enthalpy_cp<NS>(T, t, h_i, cp_i);
    double cp_tot{0}, h{0};
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cp_tot += cp_i[l] * Y[l];
      h += h_i[l] * Y[l];
    }
    const double e_t = h - gas_const * t;
    const double cv = cp_tot - gas_const;
    const double t1 = t - (e_t - e) / cv;
    err = fabs(1 - t1 / t);
    t = t1;
  }
  return t;
}

// total energy of a mixture point, compute_total_energy<FR>,
// src/FieldOperation.cuh:17-38
template <int NS>
__device__ __forceinline__ double total_energy(const Tab* __restrict__ T, double rho, double u, double v, double w,
                                               double p, double t, const double* Y) {
  double h[NS];
  enthalpy<NS>(T, t, h);
  const double V2 = u * u + v * v + w * w;
  double te = 0.5 * V2;
#pragma unroll
  for (int l = 0; l < NS; ++l) te += h[l] * Y[l];
  te *= rho;
  te -= p;
  return te;
}

}  // namespace fr
}  // namespace ostc
