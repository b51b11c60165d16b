/*
 * ostc_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a host-C++ (OpenMP over cells)
 * restatement of OpenSTC's per-stage RHS + SSP-RK3 path for calorically
 * perfect air.  It exists to check the HIP library; nothing in the product
 * links, loads or calls it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Arrays follow the reference's ColMajor SoA ghost-padded layout
 * (src/gxl_lib/Array.cuh:171-177, allocation :209-230):
 *   idx(i,j,k,l) = (k+g)*(mx+2g)*(my+2g) + (j+g)*(mx+2g) + (i+g) + l*sz,
 *   sz = (mx+2g)(my+2g)(mz+2g),  g = ngg for cv/bv/qn/mul/mach/jac/metric,
 *   g = 0 for dq and dt_local (src/Field.cu:505-550),
 *   g = ngg+1 for the grid coordinates x/y/z (src/Mesh.cpp:687).
 * The metric is stored as 9 planes, plane 3*(r-1)+(c-1) holding the
 * reference's 1-based m(r,c) (src/gxl_lib/Matrix.hpp:53-72; rows xi, eta,
 * zeta, columns x, y, z).
 */
#ifndef OSTC_ORACLE_H
#define OSTC_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int    n_var;            /* 5 for air */
  int    inviscid_scheme;  /* 51 WENO-Z5 cp, 52 WENO-Z5 characteristic, 71 WENO-Z7 cp, 72 WENO-Z7
                              characteristic, 50 WENO5-JS cp
                              (not in the reference); 1 LF, 3 AUSM+, 4 HLLC on reconstructed primitives */
  int    viscous_order;    /* 0 inviscid, 2 second-order central */
  int    fixed_time_step;  /* 1: use dt below; 0: CFL-based global dt */
  double dt;
  double cfl;
  double prandtl;          /* Pr */
  double rho_ref;          /* rho_inf of the reference state */
  double v_ref;            /* v_inf of the reference state */
  int    n_threads;        /* OpenMP threads, <=0: library default */
  int    reconstruction;   /* schemes 1/3/4: 1 first order, 2 MUSCL, 3 NND2 (minmod) */
  double entropy_fix_factor;  /* Roe flux (scheme 2); the reference's default 0.125 */
} orc_setup;

/* Mixing-layer stream states, as produced by get_mixing_layer_info for air
 * (src/MixingLayer.cu:9-190): [rho,u,v,w,p,T] of the upper and lower streams. */
typedef struct {
  double upper[6];
  double lower[6];
} orc_ml_streams;

enum { ORC_CV = 0, ORC_BV, ORC_QN, ORC_DQ, ORC_MUL, ORC_MACH, ORC_JAC, ORC_METRIC,
       ORC_DT_LOCAL, ORC_X, ORC_Y, ORC_Z, ORC_BV_LAST };

enum { ORC_BC_FARFIELD = 4, ORC_BC_INFLOW = 5, ORC_BC_OUTFLOW = 6, ORC_BC_PERIODIC = 8 };

/* Inflow of one boundary label (Inflow members, src/BoundCond.cpp:10-321):
 * inflow_type 2 mixing layer / 0 constant, fluctuation_type 0 none / 1 white
 * noise from the counter-based generator seeded by `seed`. */
typedef struct {
  int inflow_type, fluctuation_type;
  double delta_omega, fluctuation_intensity;
  unsigned long long seed;
  double upper[6], lower[6];
  double sv_upper[16], sv_lower[16];
} orc_inflow;

/* Far-field state of one boundary label: the members of FarField after its
 * constructor (src/BoundCond.h:129-149).  sv holds mass fractions on the
 * mixture path. */
typedef struct {
  double density, u, v, w, pressure, temperature;
  double acoustic_speed, entropy, specific_heat_ratio, mw;
  double sv[16];
} orc_farfield;

/* Constructor inputs of FarField (src/BoundCond.cpp:586-700) with the
 * reference's defaults for keys not given: mach, pressure, temperature,
 * density, reynolds_number = -1, velocity = 0, (u, v, w) = (1, 0, 0), the
 * velocity direction.  Combinations: (mach, pressure, temperature),
 * (density, velocity, pressure), (mach, temperature, reynolds_number). */
typedef struct {
  double mach, pressure, temperature, velocity, density, u, v, w, reynolds_number;
} orc_farfield_spec;

void*   orc_create(const orc_setup* s);
void    orc_destroy(void* ctx);
int     orc_add_block(void* ctx, int mx, int my, int mz);
int     orc_ngg(void* ctx);
/* axis 0/1/2, side -1 (index 0) or +1 (index m-1); type = ORC_BC_*.  Faces
 * are applied in the reference's priority order by type, then in the order
 * they were added (src/BoundCond.cuh:1488-1678). */
int     orc_add_boundary(void* ctx, int block, int axis, int side, int type);
/* 1-to-1 conforming face between block (axis, side) and target block on the
 * opposite side of the same axis, same index orientation, full face.  Each
 * side of an interface is registered separately, as in the reference's
 * inner N.txt files (src/Mesh.cpp:38-176). */
int     orc_add_inner_face(void* ctx, int block, int axis, int side, int target_block);
/* Connectivity records: one boundary N.txt line (0-based node range, ff for
 * type ORC_BC_FARFIELD) and one inner N.txt line pair (1-based, the paired
 * tangential axis marked negative; target_block 0-based). */
int     orc_add_boundary_range(void* ctx, int block, const int range[6], int type, const orc_farfield* ff);
/* Inflow face (label 5) over a boundary N.txt range; y from the block's own coordinates. */
int     orc_add_inflow(void* ctx, int block, const int range[6], const orc_inflow* in);
/* Sponge layer of one block (src/SpongeLayer.cu:28-172), sigma from the block's own coordinates;
 * the layout of orf_sponge (oracle/ostc_oracle_fr.h). */
typedef struct {
  int enabled[6];
  double start[6], end[6], sigma[6];
  int function;
  int jacobian_scaled;
} orc_sponge;
int     orc_set_sponge(void* ctx, int block, const orc_sponge* sp);
/* compute_source of air: the sponge term of the stage (nothing without a sponge layer) */
int     orc_source(void* ctx, int block);
int     orc_add_inner_face_range(void* ctx, int block, const int range[6], const int target_range[6],
                                 int target_block);
/* farfield face (label 4) with its state, apply_farfield<Air> (src/BoundCond.cuh:755-1022) */
int     orc_add_farfield(void* ctx, int block, int axis, int side, const orc_farfield* ff);
/* FarField constructor for air (n_spec = 0), src/BoundCond.cpp:586-700 */
int     orc_farfield_state_air(const orc_farfield_spec* in, orc_farfield* out);
double* orc_array(void* ctx, int block, int which, long long* n_out);

/* setup */
int     orc_uniform_grid(void* ctx, int block, double x0, double y0, double z0,
                         double dx, double dy, double dz);
int     orc_compute_metric(void* ctx, int block);
void    orc_mixing_layer_streams(double ma_c, double ma_upper, double t_upper, double t_lower,
                                 double p_upper, double p_lower, int upper_faster,
                                 orc_ml_streams* out);
double  orc_sutherland(double t);
int     orc_mixing_layer_ic(void* ctx, int block, const orc_ml_streams* st, double delta_omega,
                            int gi0, int gj0, int gk0, int period_i, int period_j, int period_k,
                            uint64_t seed, double amplitude);

/* Fourier-mode velocity perturbation of the differentiation workload:
 * bv[1 + comp[m]] += amp[m] * du * exp(-((y - yc)/delta)^2)
 *                    * (trig[m] ? cos : sin)(2 pi (kx[m] (x - x0)/lx + kz[m] (z - z0)/lz))
 * over the padded domain (coordinates from the grid arrays).  amp_dot (may be
 * NULL) gives the amplitude tangents in the tangent build. */
typedef struct {
  int           n_modes;
  const int*    comp;
  const int*    kx;
  const int*    kz;
  const int*    trig;
  const double* amp;
  const double* amp_dot;
  double        x0, z0, lx, lz, delta, yc, du;
} orc_fourier;
int     orc_fourier_perturbation(void* ctx, int block, const orc_fourier* f);
/* theta = sum_j w_j <rho>_j (u1 - U_j)(U_j - u2) / (rho0 du^2), U_j = <rho u>_j/<rho>_j,
 * averages over nodes i < ni, k < nk; w_j = dy, dy/2 at the two end rows.
 * dtheta (may be NULL) receives the tangent in the tangent build, else 0. */
double  orc_momentum_thickness(void* ctx, int block, double u1, double u2, double rho0, double du,
                               double dy, int ni, int nk, double* dtheta);
/* 1: state arrays hold doubles; 2: {value, tangent} pairs (tangent build) */
int     orc_scalar_width(void);

/* RK3 prologue/epilogue and whole steps (src/Driver.cu:115-199, src/RK.cuh:20-228) */
int     orc_initialize(void* ctx);
int     orc_step(void* ctx, int n_steps, double* last_dt);

/* Dual time stepping with the DPLUR inner solver (temporal_scheme 2,
 * implicit_method 1; src/DualTimeStepping.cuh:33-236, src/DPLUR.cuh:14-327):
 * the setup.txt keys dt, inner_iteration, max_inner_iteration,
 * iteration_adjust_step, DPLUR_inner_step and the step counter of the run. */
typedef struct {
  double dt;
  int inner_iteration, max_inner_iteration, iteration_adjust_step, dplur_inner_step;
  int step;
} orc_dual_time;
int     orc_dual_time_init(void* ctx, const orc_dual_time* d);
int     orc_dual_time_step(void* ctx, int n_steps, int* inner_iteration);
void    orc_jacobian_times_dq(const double* k3, const double* pv, double lambda, const double* dq, double* out);

/* stage-level operators */
int     orc_cv_from_bv(void* ctx, int block);
int     orc_copy_qn(void* ctx, int block);
int     orc_zero_dq(void* ctx, int block);
int     orc_inviscid(void* ctx, int block);
int     orc_viscous(void* ctx, int block);
int     orc_local_dt(void* ctx, int block);
double  orc_global_dt(void* ctx);
int     orc_update(void* ctx, int block, double dt, int rk);
int     orc_apply_bc(void* ctx, int block);
int     orc_communicate(void* ctx);
int     orc_properties(void* ctx, int block);
int     orc_store_last_step(void* ctx);
int     orc_residual(void* ctx, double res[4]);

/* point kernels, exposed for closed-form tests */
double  orc_weno5z(const double* vp, const double* vm, double eps);
double  orc_weno7z(const double* vp, const double* vm, double eps);
double  orc_weno5js(const double* vp, const double* vm, double eps);
/* Riemann flux between primitive states [rho,u,v,w,p] (scheme 1 LF, 3 AUSM+,
 * 4 HLLC), metric rows kl/kr, Jacobians jl/jr; and the characteristic WENO-Z5
 * left/right matrices of a face between conservative states q = [cv(5), p]. */
/* riemannSolver_Roe<Air> on reconstructed states [rho,u,v,w,p,E] with the two points' entropy-fix deltas */
int     orc_roe(const double* ql, const double* qr, const double* kl, const double* kr, double jl, double jr,
                double dl, double dr, double* f);
int     orc_riemann(int scheme, const double* ql, const double* qr, const double* kl, const double* kr, double jl,
                    double jr, double* fc);
int     orc_char_matrices(const double* ql, const double* qr, const double* kl, const double* kr, double jl,
                          double jr, double* L, double* R);
void    orc_lf_split(const double* q /* n_var cv, p, c */, int n_var, const double* k3, double jac,
                     double* fp, double* fm);

#ifdef __cplusplus
}
#endif
#endif
