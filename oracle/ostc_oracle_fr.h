/*
 * ostc_oracle_fr.h -- C interface of the CPU oracle for the finite-rate
 * mixture path (the reference's MixtureModel::FR with laminar transport).
 *
 * TEST INFRASTRUCTURE ONLY.  A host-C++ (OpenMP over cells) restatement of
 * the reference's mixture routines: NASA-7 thermodynamics, mixture-averaged
 * transport, finite-rate chemistry, the temperature Newton iteration,
 * component-wise and characteristic WENO-Z fluxes with the positivity
 * limiter, second-order viscous fluxes with species diffusion, the SSP-RK3
 * update, boundary copies and inner-face exchange.  Nothing in the product
 * links, loads or calls it; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Layout: as the air oracle (ostc_oracle.h) -- per-variable planes, i
 * fastest, g = ngg ghost layers on cv [n_var], bv [6], sv [n_spec], qn, the
 * property planes and geometry; g = 0 for dq [n_var] and dt_local.
 * n_var = 5 + n_spec.
 */
#ifndef OSTC_ORACLE_FR_H
#define OSTC_ORACLE_FR_H
#include <stdint.h>
#include "openstc/mech.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int    inviscid_scheme;      /* 51 / 71 WENO-Z5 / WENO-Z7 component-wise, 52 / 72 characteristic,
                                  3 AUSM+, 4 HLLC on reconstructed primitives */
  int    viscous_order;        /* 0 or 2 */
  int    reaction;             /* 1: finite-rate chemistry source */
  int    positive_preserving;  /* 1: limiter on the species fluxes */
  int    grad_p_diffusion;     /* gradPInDiffusionFlux */
  int    fixed_time_step;
  int    reconstruction;       /* schemes 3/4: 1 first order, 2 MUSCL, 3 NND2 (minmod limiter) */
  double dt, cfl;
  double prandtl;              /* prandtl_number (viscous spectral radius of the time step) */
  double rho_ref, v_ref;       /* weno_eps_scale = (rho_ref v_ref)^2 */
  int    n_threads;
  double entropy_fix_factor;   /* Roe flux (scheme 2); the reference's default 0.125 */
  int    ch_stencil;           /* scheme 52: 0 the i-2..i+2 / i-1..i+3 stencil for every field, 1 the
                                  reference's i-3..i+1 / i-2..i+2 for the five conservative components of
                                  the acoustic/entropy fields (src/WENO.cu:880-881) */
} orf_setup;

/* arrays */
enum { ORF_CV = 0, ORF_BV, ORF_SV, ORF_QN, ORF_DQ, ORF_MUL, ORF_LAMBDA, ORF_CP, ORF_GAMMA, ORF_C,
       ORF_MACH, ORF_RHOD, ORF_JAC, ORF_METRIC, ORF_DT_LOCAL };

void*   orf_create(const orf_setup* s, const ostc_mech_tables* t);
void    orf_destroy(void* ctx);
int     orf_add_block(void* ctx, int mx, int my, int mz);
int     orf_ngg(void* ctx);
int     orf_add_boundary(void* ctx, int block, int axis, int side, int type);  /* 6 outflow, 8 periodic */
int     orf_add_inner_face(void* ctx, int block, int axis, int side, int target_block);
/* connectivity records: a boundary N.txt line, an inner N.txt line pair */
int     orf_add_inner_face_range(void* ctx, int block, const int range[6], const int target_range[6],
                                 int target_block);

/* Far-field state of one label (FarField after its constructor,
 * src/BoundCond.h:129-149) and the constructor's inputs (-1 = not given,
 * velocity 0, direction (1, 0, 0)); the same layouts as the air oracle's. */
typedef struct {
  double density, u, v, w, pressure, temperature;
  double acoustic_speed, entropy, specific_heat_ratio, mw;
  double sv[OSTC_MAX_SPEC];
} orf_farfield;
typedef struct {
  double mach, pressure, temperature, velocity, density, u, v, w, reynolds_number;
} orf_farfield_spec;
/* FarField::FarField with species (src/BoundCond.cpp:586-700): mass fractions y */
int     orf_farfield_state(const ostc_mech_tables* t, const orf_farfield_spec* in, const double* y,
                           orf_farfield* out);
/* farfield face (label 4), apply_farfield<FR> (src/BoundCond.cuh:755-1022) */
int     orf_add_farfield(void* ctx, int block, int axis, int side, const orf_farfield* ff);
int     orf_add_boundary_range(void* ctx, int block, const int range[6], int type, const orf_farfield* ff);
/* Inflow face (label 5): the state in the layout of orc_inflow (oracle/ostc_oracle.h) and
 * the block's node y coordinates on the padded domain (the bv plane layout). */
typedef struct {
  int inflow_type, fluctuation_type;
  double delta_omega, fluctuation_intensity;
  unsigned long long seed;
  double upper[6], lower[6];
  double sv_upper[16], sv_lower[16];
} orf_inflow;
int     orf_add_inflow(void* ctx, int block, const int range[6], const orf_inflow* in, const double* y);
/* Sponge layer of one block (src/SpongeLayer.cu:28-172): x, y, z on the padded domain. */
typedef struct {
  int enabled[6];
  double start[6], end[6], sigma[6];
  int function;
  int jacobian_scaled;  /* 1: the term times the Jacobian; 0: the reference's term as written */
} orf_sponge;
int     orf_set_sponge(void* ctx, int block, const orf_sponge* sp, const double* x, const double* y,
                       const double* z);
double* orf_array(void* ctx, int block, int which, long long* n_out);

int     orf_initialize(void* ctx);            /* cv from bv, properties, BCs, exchange, properties */
int     orf_step(void* ctx, int n_steps, double* last_dt);
/* stage pieces */
int     orf_cv_from_bv(void* ctx, int block);
int     orf_properties(void* ctx, int block);
int     orf_zero_dq(void* ctx, int block);
int     orf_source(void* ctx, int block);
int     orf_inviscid(void* ctx, int block, double dt);
int     orf_viscous(void* ctx, int block);
int     orf_local_dt(void* ctx, int block);
double  orf_global_dt(void* ctx);
int     orf_update(void* ctx, int block, double dt, int rk);
int     orf_apply_bc(void* ctx, int block);
int     orf_communicate(void* ctx);
/* store_last_step / compute_residual: RMS change of rho, |V|, p, T since the store */
int     orf_store_last_step(void* ctx);
int     orf_residual(void* ctx, double res[4]);

/* point routines, for closed-form and identity tests */
void    orf_enthalpy_cp(const ostc_mech_tables* t, double temp, double* h, double* cp);
void    orf_gibbs_rt(const ostc_mech_tables* t, double temp, double* g);
/* omega [ns] (kg/m3/s) for density rho, temperature t, mass fractions y */
void    orf_chemical_source(const ostc_mech_tables* t, double rho, double temp, const double* y, double* omega,
                            double* kf, double* kb);
/* mu, lambda, rho_D [ns] of a mixture at (T, p, rho, Y) */
void    orf_transport(const ostc_mech_tables* t, double temp, double p, double rho, const double* y, double* mu,
                      double* lambda, double* rho_d);
/* T-Newton: returns T, *iters receives the iteration count */
double  orf_temperature(const ostc_mech_tables* t, double e, const double* y, double t_start, int* iters);
/* Riemann flux (scheme 3 AUSM+, 4 HLLC) between primitive states q = [rho, u,
 * v, w, p, Y_1..Y_ns], metric rows kl/kr and Jacobians jl/jr */
int     orf_riemann(const ostc_mech_tables* t, int scheme, const double* ql, const double* qr, const double* kl,
                    const double* kr, double jl, double jr, double* fc);
/* Roe flux with Harten's entropy fix (scheme 2) between primitive states q;
 * delta_l/delta_r: the entropy-fix widths of the two points */
int     orf_roe(const ostc_mech_tables* t, const double* ql, const double* qr, const double* kl, const double* kr,
                double jl, double jr, double delta_l, double delta_r, double* fc);
/* reconstruction of the face between points 1 and 2 of pv[4][nv] */
void    orf_reconstruct(const double* pv, int nv, int recon, double* pl, double* pr);
/* left (5+ns)x(5+ns) characteristic matrix rows L and the back projection R
 * of compute_weno_flux_ch at a face, for the L R = I identity */
int     orf_char_matrices(const ostc_mech_tables* t, const double* ql, const double* qr, const double* kl,
                          const double* kr, double jl, double jr, double* L, double* R);

#ifdef __cplusplus
}
#endif
#endif
