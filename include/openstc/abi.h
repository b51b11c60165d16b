/*
 * openstc/abi.h -- C ABI of the MI355X-native RHS + SSP-RK3 library
 * (libostc_hip.so).  Plain pointers and sizes only; every function returns
 * an int status (OSTC_OK = 0, negative on error) and never aborts; the text
 * of the last error is available from ostc_last_error().
 *
 * The reference (OpenSTC-MixingLayer) has no plugin or FFI layer: its
 * operator surface is a set of C++ templates called from RK3
 * (src/RK.cuh:20-228).  Each entry point below names the reference routine
 * it stands in for.
 *
 * Layout contract (the reference's DZone/Field, src/Field.h:12-90,
 * src/gxl_lib/Array.cuh:162-230): per-variable planes, i fastest, ngg ghost
 * layers on every side,
 *     idx(i,j,k,l) = (k+g)*(mx+2g)*(my+2g) + (j+g)*(mx+2g) + (i+g) + l*sz,
 *     sz = (mx+2g)(my+2g)(mz+2g),
 * with g = ngg for cv [n_var], bv [6: rho,u,v,w,p,T], mul, mach, jac and the
 * 9 metric planes, and g = 0 for dq [n_var] and dt_local.
 *
 * Threading: one host thread per context; all device work is queued on the
 * stream given at creation.  Host synchronisation happens only where a
 * function returns device data to the host (dt, residuals, downloads).
 */
#ifndef OPENSTC_ABI_H
#define OPENSTC_ABI_H
#include <stdint.h>

#include "openstc/mech.h"
#ifdef __cplusplus
extern "C" {
#endif

#define OSTC_OK 0
#define OSTC_ERR_ARG (-1)      /* bad argument, size or index */
#define OSTC_ERR_HIP (-2)      /* HIP runtime error */
#define OSTC_ERR_STATE (-3)    /* call not valid in the current state */
#define OSTC_ERR_NUMERIC (-4)  /* NaN / non-positive density or pressure detected */
#define OSTC_ERR_UNSUPPORTED (-5)

/* boundary labels of the reference (src/BoundCond.cuh:1488-1678) */
#define OSTC_BC_FARFIELD 4
#define OSTC_BC_OUTFLOW 6
#define OSTC_BC_PERIODIC 8
#define OSTC_BC_INFLOW 5

/* time-step modes of ostc_step */
#define OSTC_MODE_PLAIN 0    /* per-block launches shaped like the reference */
#define OSTC_MODE_BATCHED 1  /* fused RHS+update over a block work-list, graph-captured */
#define OSTC_MODE_BATCHED_TIMED 2  /* the batched kernels launched directly, with HIP events
                                      around every fused launch (ostc_fused_kernel_ms) */

/* properties of the ghost layers after a batched stage (ostc_set_ghost_props) */
#define OSTC_GHOST_PROPS_IN_OPS 0  /* each ghost op recomputes the nodes it writes (default) */
#define OSTC_GHOST_PROPS_PASS 1    /* one more launch over every ghost layer and shared node layer */

/* fields for ostc_get_field / ostc_set_field */
#define OSTC_FIELD_CV 0
#define OSTC_FIELD_BV 1
#define OSTC_FIELD_MUL 2
#define OSTC_FIELD_MACH 3
#define OSTC_FIELD_DQ 4
#define OSTC_FIELD_DT_LOCAL 5
#define OSTC_FIELD_JAC 6
#define OSTC_FIELD_METRIC 7
#define OSTC_FIELD_QN 8

typedef struct ostc_ctx_s* ostc_ctx;

/* Scheme and physics keys of setup.txt (src/Parameter.cpp:187-437,
 * src/DParameter.cu:7-283).  Air (calorically perfect, gamma 1.4,
 * Sutherland viscosity) is the mixture model on this path. */
typedef struct {
  int n_var;            /* 5 */
  int inviscid_scheme;  /* 51 WENO-Z5 cp, 71 WENO-Z7 cp, 50 WENO5-JS cp (no reference counterpart);
                           OSTC_MODE_PLAIN only: 52 / 72 WENO-Z5 / WENO-Z7 characteristic (src/WENO.cu:1147-1453),
                           1 LF, 2 Roe (entropy fix below), 3 AUSM+, 4 HLLC on reconstructed primitives
                           (src/InviscidScheme.cu:12-215, src/RiemannSolver.cuh:279-417) */
  int viscous_order;    /* 0 inviscid, 2 second-order central */
  int fixed_time_step;  /* 1: use dt below, 0: CFL-based global dt */
  double dt;
  double cfl;
  double prandtl;       /* prandtl_number */
  double rho_ref;       /* rho_inf of the reference state (src/BoundCond.cpp:300-321) */
  double v_ref;         /* v_inf of the reference state */
  int reconstruction;   /* schemes 1/3/4: 1 first order, 2 MUSCL, 3 NND2, minmod limiter
                           (src/Reconstruction.cu:5-46, src/Limiter.cuh:7-41) */
  double entropy_fix_factor;  /* entropy_fix_factor of the Roe flux (scheme 2), >= 0; the reference's
                                 default is 0.125 (src/Parameter.cpp:474, src/DParameter.cu:16) */
} ostc_setup;

typedef struct {
  int mx, my, mz;       /* interior node counts (mz > 1: three-dimensional) */
  /* Optional caller-owned device arrays of the block, the reference's DZone
   * arrays (Field::setup_device_memory, src/Field.cu:469-647) in the layout
   * above for the context's ghost depth.  NULL (the zero-initialised rest of
   * the struct): the library allocates, zeroes and owns the array.  What the
   * caller passes stays the caller's: the library never frees it, and it must
   * stay valid until the context is destroyed.
   *   cv   n_var planes: the block's conservative state (the state the
   *        batched step reads and leaves its result in)
   *   bv   6 planes: rho, u, v, w, p, T
   *   sv   n_spec planes of mass fractions (mixture contexts; ignored for air)
   *   qn   n_var planes: the state at the start of the step (per-block steps);
   *        batched steps use the array as intermediate stage storage
   *   dq   n_var planes, no ghosts
   *   mul  viscosity (air contexts: the state the batched step leaves)
   *   jac  Jacobian */
  double* cv;
  double* bv;
  double* sv;
  double* qn;
  double* dq;
  double* mul;
  double* jac;
  /* Optional device metric, read by finalize into the library's own SoA
   * planes: component c (0..8, c = 3(r-1)+(c'-1)) of padded point p at
   * metric[p*metric_elem_stride + c*metric_comp_stride + metric_offset]
   * (doubles), so SoA planes (1, sz, 0) and the reference's 80-byte
   * gxl::Matrix AoS (10, 1, 1; src/gxl_lib/Matrix.hpp:53-72) are read where
   * they lie.  NULL: give the geometry with ostc_set_geometry. */
  const double* metric;
  long long metric_elem_stride, metric_comp_stride, metric_offset;
} ostc_block_desc;

/* Device view of one block, the DZone equivalent (src/Field.h:12-90). */
typedef struct {
  int mx, my, mz, ngg;
  long long sz;         /* padded plane size */
  double* cv;           /* n_var planes */
  double* bv;           /* 6 planes */
  double* mul;
  double* mach;
  double* jac;
  double* metric;       /* 9 planes, plane 3(r-1)+(c-1) = m(r,c) */
  double* dq;           /* n_var planes, no ghosts */
  double* dt_local;     /* no ghosts */
} ostc_zone;

/* Driver/Field construction: src/Driver.cu:12-113, src/Field.cu:469-647 */
int ostc_create(const ostc_setup* s, void* hip_stream, ostc_ctx* out);
int ostc_destroy(ostc_ctx ctx);
int ostc_add_block(ostc_ctx ctx, const ostc_block_desc* d, int* block_id);
/* Boundary face (Boundary::register_boundary, src/Mesh.cpp:14-32):
 * axis 0/1/2, side -1 (index 0) or +1 (index m-1), type OSTC_BC_*. */
int ostc_add_boundary(ostc_ctx ctx, int block, int axis, int side, int type);

/* Far-field state of one boundary label: the members of FarField after its
 * constructor (src/BoundCond.h:129-149).  sv: mass fractions (mixture
 * contexts; ignored for air). */
typedef struct {
  double density, u, v, w, pressure, temperature;
  double acoustic_speed, entropy, specific_heat_ratio, mw;
  double sv[OSTC_MAX_SPEC];
} ostc_farfield;

/* Inputs of the FarField constructor (src/BoundCond.cpp:586-700) with the
 * reference's defaults for keys not given: mach, pressure, temperature,
 * density, reynolds_number = -1, velocity = 0, (u, v, w) = (1, 0, 0) as the
 * velocity direction.  Combinations: (mach, pressure, temperature),
 * (density, velocity, pressure), (mach, temperature, reynolds_number). */
typedef struct {
  double mach, pressure, temperature, velocity, density, u, v, w, reynolds_number;
} ostc_farfield_spec;

/* FarField::FarField for air (n_spec = 0, laminar), src/BoundCond.cpp:586-700.
 * Host-only; no context needed. */
int ostc_farfield_state_air(const ostc_farfield_spec* in, ostc_farfield* out);

/* Farfield face (label 4) with the state of its label: apply_farfield<Air>,
 * src/BoundCond.cuh:755-1022 (characteristic state from the face point's
 * normal Mach number).  Replaces ostc_add_boundary(..., OSTC_BC_FARFIELD). */
int ostc_add_farfield(ostc_ctx ctx, int block, int axis, int side, const ostc_farfield* ff);
/* 1-to-1 conforming inner face with aligned index directions
 * (InnerFace::register_boundary, src/Mesh.cpp:38-176).  Register each side
 * of an interface, as the reference's inner N.txt files do. */
int ostc_add_inner_face(ostc_ctx ctx, int block, int axis, int side, int target_block);
/* Connectivity records of the reference's input files (Mesh::read_boundary,
 * Mesh::read_inner_interface, src/Mesh.cpp:790-843).
 * ostc_add_boundary_range: one boundary N.txt line, range = {x1, x2, y1, y2,
 * z1, z2} in 0-based node indices, the normal axis having equal start and end
 * (Boundary::register_boundary, src/Mesh.cpp:14-32); the face covers that node
 * range widened by ngg on the tangential axes.  type OSTC_BC_*; ff is the
 * far-field state of type OSTC_BC_FARFIELD (NULL otherwise).
 * ostc_add_inner_face_range: one inner N.txt line pair, 1-based node indices of
 * this block's face (range) and of the target block's face (target_range), the
 * tangential axis that pairs with the other side's marked axis given as
 * negative indices; the two faces may have any relative orientation (index
 * axes permuted and reversed; InnerFace::register_boundary,
 * src/Mesh.cpp:38-176).  target_block is 0-based (the file's id minus 1). */
int ostc_add_boundary_range(ostc_ctx ctx, int block, const int range[6], int type, const ostc_farfield* ff);
int ostc_add_inner_face_range(ostc_ctx ctx, int block, const int range[6], const int target_range[6],
                              int target_block);

/* Inflow face (label 5), apply_inflow, src/BoundCond.cuh:235-612, with the
 * members of Inflow after its constructor (src/BoundCond.cpp:10-321).
 * inflow_type 2 (mixing layer): u = (u1+u2)/2 + (u1-u2)/2 tanh(2y/delta_omega)
 * at every face point, the other primitives (and mass fractions) of the
 * upper stream where y >= 0 and of the lower stream below; inflow_type 0
 * (constant): the upper state everywhere.  The ghost layers repeat the face
 * point.  fluctuation_type 1 (white noise) adds intensity*|V| times standard
 * normal deviates to u and v (type 2, only where |y| < 4 delta_omega) or to u
 * (type 0).  The reference draws them from cuRAND states seeded by the wall
 * clock; here they come from a counter-based generator (SplitMix64 finalizer
 * and the Box-Muller transform) keyed by seed, the face point and the count
 * of the block's boundary-condition passes, so a run is reproducible. */
typedef struct {
  int inflow_type;            /* 2 mixing layer, 0 constant */
  int fluctuation_type;       /* 0 none, 1 white noise */
  double delta_omega;
  double fluctuation_intensity;
  unsigned long long seed;
  double upper[6], lower[6];  /* rho, u, v, w, p, T */
  double sv_upper[OSTC_MAX_SPEC], sv_lower[OSTC_MAX_SPEC];
} ostc_inflow;
/* range: a boundary N.txt record of the face (0-based, as ostc_add_boundary_range);
 * y: the block's node y coordinates on the padded domain (host, the bv plane layout). */
int ostc_add_inflow(ostc_ctx ctx, int block, const int range[6], const ostc_inflow* in, const double* y);
/* Allocates device memory and builds the work-lists; blocks and faces are
 * fixed afterwards. */
int ostc_finalize(ostc_ctx ctx);
int ostc_ngg(ostc_ctx ctx, int* ngg);
int ostc_zone_view(ostc_ctx ctx, int block, ostc_zone* out);

/* Geometry: host arrays on the padded domain.  metric component c (0..8,
 * c = 3(r-1)+(c'-1)) of point p is read at
 * metric[p*elem_stride + c*comp_stride + offset] (in doubles), so both SoA
 * planes (1, sz, 0) and the reference's 80-byte gxl::Matrix AoS (10, 1, 1;
 * src/gxl_lib/Matrix.hpp:53-72) can be passed. */
int ostc_set_geometry(ostc_ctx ctx, int block, const double* jac, const double* metric,
                      long long elem_stride, long long comp_stride, long long offset);
/* Host <-> device copies of a field in the layout above (synchronous). */
int ostc_set_field(ostc_ctx ctx, int block, int field, const double* host);
int ostc_get_field(ostc_ctx ctx, int block, int field, double* host);

/* Driver::initialize_computation (src/Driver.cu:115-199) followed by the
 * RK3 prologue compute_cv_from_bv (src/RK.cuh:38-49): boundary conditions,
 * inner-face exchange, properties, cv from bv over the padded domain. */
int ostc_initialize(ostc_ctx ctx);

/* n_steps SSP-RK3 steps (src/RK.cuh:85-226).  last_dt may be NULL. */
int ostc_step(ostc_ctx ctx, int n_steps, int mode, double* last_dt);

/* Faces shared with a block of another rank (one GPU per rank), for a
 * driver that runs the stage operators (ostc_stage_rhs ... ostc_exchange,
 * then the parallel exchange, then ostc_update_properties) and moves the
 * bytes itself (RCCL send/recv, a peer copy, MPI).  parallel_communication,
 * src/DataCommunication.cuh:108-199, with the record format of
 * Mesh::read_parallel_interface / ParallelFace::register_boundary
 * (src/Mesh.cpp:846-876, :177-231).  ostc_step refuses a context with
 * parallel faces. */
typedef struct {
  int block, rank, flag_send, flag_recv;
  long long offset;  /* first double of the face in the send and receive buffers */
  long long count;   /* doubles: 6 primitives x (ngg + 1) layers x points */
  long long points;  /* face nodes */
} ostc_parallel_info;
/* One parallel N.txt record, before ostc_finalize: range = iMin iMax jMin
 * jMax kMin kMax as in the file (1-based; the normal axis has equal ends; the
 * tangential axis listed second in the buffer is given negated). */
int ostc_add_parallel_face(ostc_ctx ctx, int block, const int range[6], int rank, int flag_send, int flag_recv);
int ostc_parallel_face_count(ostc_ctx ctx, int* n);
int ostc_parallel_face_info(ostc_ctx ctx, int face, ostc_parallel_info* out);
/* The device send and receive buffers (total doubles each), owned by the context. */
int ostc_parallel_buffers(ostc_ctx ctx, double** send, double** recv, long long* total);
/* setup_data_to_be_sent (src/DataCommunication.cu:3-40) of every parallel face
 * into the send buffer: per face point the face node, then the ngg nodes
 * inside it, 6 primitives each. */
int ostc_pack_parallel(ostc_ctx ctx);
/* assign_data_received (src/DataCommunication.cuh:201-240) of every parallel
 * face, in face order, from the receive buffer (the partner's send buffer
 * slice of that face): the face node takes the mean of its own and the
 * partner's primitives, the ngg ghost nodes beyond it the partner's interior
 * ones; cv from them.  As in the reference nothing may change the fields
 * between ostc_pack_parallel and this call: the own values of a face come from
 * the send buffer, except for a face that shares nodes with an earlier face
 * of its block, which reads the field (and so sees the earlier face's
 * means).  Returns OSTC_ERR_STATE unless ostc_pack_parallel ran since the last
 * unpack or ostc_set_field. */
int ostc_unpack_parallel(ostc_ctx ctx);
/* Host copies for drivers without device transport: which = 0 copies the send
 * buffer to host, which = 1 the host array into the receive buffer (total
 * doubles; synchronous). */
int ostc_parallel_copy(ostc_ctx ctx, int which, double* host);

/* Where a batched stage recomputes the viscosity and Mach number of the
 * ghost layers (update_physical_properties over the padded domain,
 * src/RK.cuh:216-223): OSTC_GHOST_PROPS_IN_OPS or OSTC_GHOST_PROPS_PASS.
 * Both give bit-identical states; the first saves a pass over the ghost
 * layers.  No counterpart in the reference (a launch-shape choice). */
int ostc_set_ghost_props(ostc_ctx ctx, int mode);

/* Dual time stepping with the DPLUR inner solver (temporal_scheme 2,
 * implicit_method 1: dual_time_stepping, src/DualTimeStepping.cuh:33-236,
 * DPLUR, src/DPLUR.cuh:14-327), air, on the per-block kernels.  Keys of
 * setup.txt: the physical time step dt, inner_iteration,
 * max_inner_iteration, iteration_adjust_step (every this many steps the inner
 * iterations are monitored: the count grows by 5 while the last iteration
 * still changes the state by more than 1e-3 of the first iteration's change,
 * and drops to the iteration where the change falls below; 0: never),
 * DPLUR_inner_step, and the step counter of the run ("step").  The pseudo time
 * step is the local CFL step (cfl of ostc_setup). */
typedef struct {
  double dt;
  int inner_iteration;
  int max_inner_iteration;
  int iteration_adjust_step;
  int dplur_inner_step;
  int step;
} ostc_dual_time;
/* Arms dual time stepping; q^(n-1) becomes the current state. */
int ostc_set_dual_time(ostc_ctx ctx, const ostc_dual_time* d);
/* n_steps physical steps; *inner_iteration (may be NULL) receives the inner
 * iteration count the next step will take.  OSTC_ERR_NUMERIC on a NaN in a
 * monitored inner-iteration residual. */
int ostc_dual_time_step(ostc_ctx ctx, int n_steps, int* inner_iteration);

/* Stage-level operators, shaped like the reference's per-block launches. */
int ostc_stage_rhs(ostc_ctx ctx, int block);                 /* dq = 0; inviscid; viscous (src/RK.cuh:113-139) */
int ostc_compute_dt(ostc_ctx ctx, double* dt_out);            /* local_time_step + global_time_step */
int ostc_stage_update(ostc_ctx ctx, int block, int rk, double dt); /* update_cv_and_bv_rk (src/RK.cuh:230-284) */
int ostc_apply_bc(ostc_ctx ctx, int block);                   /* DBoundCond::apply_boundary_conditions */
int ostc_exchange(ostc_ctx ctx);                              /* data_communication, inner faces */
int ostc_update_properties(ostc_ctx ctx, int block);          /* update_physical_properties<Air> */
int ostc_copy_qn(ostc_ctx ctx, int block);                    /* qn <- cv (src/RK.cuh:99-102) */
/* Batched fused RHS of one block written to dq (no update), for checking
 * the fused kernel's residual against the per-direction kernels. */
int ostc_batched_rhs(ostc_ctx ctx, int block);

/* Residual norms sqrt(sum(dv^2)/N) of (rho, |V|, p, T) against the last
 * ostc_store_last_step (src/TimeAdvanceFunc.cu:6-37, src/Residual.cuh:75-122). */
int ostc_store_last_step(ostc_ctx ctx);
int ostc_residual(ostc_ctx ctx, double res[4]);

/* Launch configuration of the fused kernel: tile index into the compiled
 * table, workgroups per CU of a persistent grid (0: one workgroup per tile),
 * and split = 1 to run each stage as an inviscid kernel followed by a
 * viscous + update kernel.  ostc_autotune times every combination on the
 * current blocks and keeps the fastest. */
int ostc_tile_count(int* n);
int ostc_tile_shape(int index, int* tx, int* ty, int* tz);
int ostc_set_tile(ostc_ctx ctx, int index, int waves_per_cu, int split);
int ostc_tile_config(ostc_ctx ctx, int* index, int* waves_per_cu, int* split);
int ostc_autotune(ostc_ctx ctx, int reps, int* best_index, double* best_ms);

/* Device timing of the last ostc_step call: total milliseconds measured
 * with HIP events on the context's stream, and the launch count. */
int ostc_last_step_timing(ostc_ctx ctx, double* ms, long long* launches);
/* Average and maximum duration (ms) of the fused RHS+update kernel over the
 * launches of the last OSTC_MODE_BATCHED_TIMED ostc_step call, from HIP
 * events recorded around each launch on the context's stream. */
int ostc_fused_kernel_ms(ostc_ctx ctx, double* avg_ms, double* max_ms, long long* count);
int ostc_sync(ostc_ctx ctx);
const char* ostc_last_error(ostc_ctx ctx);

/* Setup helpers (host code, src/Mesh.cpp:241-300 and src/Initialize.cu:259-325):
 * a uniform Cartesian grid's jac and metric on the padded domain (SoA), and
 * the air mixing-layer initial state bv on the padded domain plus a seeded
 * velocity perturbation amplitude*U(-1,1) from SplitMix64(seed, node, comp). */
int ostc_util_uniform_geometry(int mx, int my, int mz, int ngg, double x0, double y0, double z0,
                               double dx, double dy, double dz, double* jac, double* metric);
/* Block::compute_jac_metric (src/Mesh.cpp:241-502) from node coordinates: x, y, z
 * on the domain padded by ngg+1 ghost layers, jac and metric (SoA) on the
 * domain padded by ngg.  OSTC_ERR_NUMERIC on a non-positive Jacobian. */
int ostc_util_metric(int mx, int my, int mz, int ngg, const double* x, const double* y, const double* z,
                     double* jac, double* metric);
int ostc_util_mixing_layer_bv(int mx, int my, int mz, int ngg, double y0, double dy,
                              const double upper[6], const double lower[6], double delta_omega,
                              int gi0, int gj0, int gk0, int period_i, int period_j, int period_k,
                              uint64_t seed, double amplitude, double* bv);

/* ------------------------------------------------------------ reverse mode
 * The reference has no adjoint; these entry points differentiate its RHS
 * (src/WENO.cu:697-720, :1582-1617; src/ViscousScheme.cuh:23-909) and its
 * SSP-RK3 stage (src/RK.cuh:85-284) for callers that need gradients.
 * Supported: WENO-Z5 (51) and WENO5-JS (50); the RK3 reverse sweep needs a
 * single block with outflow/periodic faces and fixed_time_step = 1. */

/* Library-owned cotangent fields of each block (allocated on first use):
 * dq_bar [n_var planes, no ghosts] and cv_bar [n_var planes, padded]. */
#define OSTC_FIELD_DQ_BAR 9
#define OSTC_FIELD_CV_BAR 10

/* cv_bar = (d dq / d cv)^T dq_bar for block `block` at its current state
 * (cv, bv, mul as left by ostc_initialize or a step), over the padded domain
 * (ghost cells included; the caller folds ghost cotangents back through its
 * own ghost-cell rules).  dq_bar and cv_bar are device pointers in the
 * layout above; NULL selects OSTC_FIELD_DQ_BAR / OSTC_FIELD_CV_BAR.
 * cv_bar is overwritten. */
int ostc_rhs_vjp(ostc_ctx ctx, int block, const double* dq_bar, double* cv_bar);

/* Fourier-mode velocity perturbation of the differentiation workload:
 * bv[1 + comp[m]] += amp[m] * du * exp(-((y - yc)/delta)^2)
 *                    * (trig[m] ? cos : sin)(2 pi (kx[m] i dx / lx + kz[m] k dz / lz)),
 * y = y0 + j dy, over the padded domain of the block's slot-0 bv (before
 * ostc_initialize). */
#define OSTC_MAX_MODES 64
typedef struct {
  int n_modes;
  int comp[OSTC_MAX_MODES];  /* 0 u, 1 v, 2 w */
  int kx[OSTC_MAX_MODES];
  int kz[OSTC_MAX_MODES];
  int trig[OSTC_MAX_MODES];  /* 0 sin, 1 cos */
  double amp[OSTC_MAX_MODES];
  double y0;                 /* y of node j = 0 */
  double dx, dy, dz;         /* uniform node spacing */
  double lx, lz;             /* periods of the phase */
  double delta;              /* envelope width */
  double yc;                 /* envelope centre */
  double du;                 /* velocity scale */
} ostc_fourier;
int ostc_perturb_fourier(ostc_ctx ctx, int block, const ostc_fourier* f);

/* Momentum thickness of a temporal mixing layer,
 * theta = sum_j w_j <rho>_j (u1 - U_j)(U_j - u2) / (rho0 du^2),
 * U_j = <rho u>_j / <rho>_j, averages over nodes i < ni, k < nk of row j,
 * w_j = dy (dy/2 on the first and last row). */
typedef struct {
  double u1, u2;             /* upper and lower stream velocities */
  double rho0, du;           /* normalising density and velocity difference */
  double dy;
  int ni, nk;
} ostc_theta;
int ostc_momentum_thickness(ostc_ctx ctx, int block, const ostc_theta* t, double* theta);
/* No counterpart in the reference.  The same momentum thickness, bit for bit,
 * for callers that cannot allocate device memory (code under stream capture,
 * tight optimisation loops): it calls no hipMalloc/hipFree and keeps its row
 * sums in a file-scope device array of 4 * 16384 + 1 doubles that each GPU
 * holds until the process ends.  Blocks of more than 16384 rows (my) return
 * OSTC_ERR_ARG; the context's device must be current (else OSTC_ERR_STATE).
 * Calls that overlap on one GPU share the array and race. */
int ostc_momentum_thickness_noalloc(ostc_ctx ctx, int block, const ostc_theta* t, double* theta);

/* Forward half of the differentiated run: n_steps batched SSP-RK3 steps from
 * the current (initialised) state with the fixed dt, keeping the state every
 * checkpoint_every steps (and the stage states of the last segment) on the
 * device, then theta of the final state.  Replaces ostc_step for the run. */
int ostc_adjoint_forward(ostc_ctx ctx, int n_steps, int checkpoint_every, const ostc_theta* t,
                         double* theta);
/* Reverse sweep of the last ostc_adjoint_forward: segments are recomputed
 * from their checkpoints, stages are reversed with ostc_rhs_vjp's kernels,
 * and grad[m] = d theta / d amp[m] for the perturbation f that produced the
 * initial state (f may be NULL: then only OSTC_FIELD_CV_BAR, d theta / d cv
 * of the initial state after its boundary copies are folded back, is kept). */
int ostc_adjoint_backward(ostc_ctx ctx, const ostc_fourier* f, double* grad);
/* Milliseconds of the last forward and backward calls (HIP events). */
int ostc_adjoint_timing(ostc_ctx ctx, double* forward_ms, double* backward_ms);

/* ---------------------------------------------------- finite-rate mixtures
 * The reference's MixtureModel::FR instantiation of the same operators
 * (species = 1, reaction = 0/1; laminar): NASA-7 thermodynamics,
 * mixture-averaged transport (src/Transport.cu:43-136), finite-rate
 * chemistry (src/FiniteRateChem.cu:7-195), the temperature Newton iteration
 * (src/FieldOperation.cu:3-40), component-wise (51) and characteristic (52)
 * WENO-Z5 with Lax-Friedrichs-type splitting (src/WENO.cu:697-1143,
 * :1455-1536), the positivity-preserving limiter (:1538-1580) and viscous
 * fluxes with species diffusion (src/ViscousScheme.cuh:23-909).
 * n_var = 5 + n_spec; the species counts compiled into the library are
 * reported by ostc_fr_supported (others return OSTC_ERR_UNSUPPORTED).
 *
 * Extra planes of a block (reference layout): sv [n_spec] (mass fractions),
 * thermal conductivity, cp, gamma, acoustic speed, rho_D [n_spec]; cv/qn/dq
 * carry n_var planes.  All state stays on the device; a step makes no host
 * synchronisation, so OSTC_FR_MODE_GRAPH captures it once and replays it. */
typedef struct ostc_fr_ctx_s* ostc_fr_ctx;

typedef struct {
  int inviscid_scheme;       /* 51 / 71 WENO-Z5 / WENO-Z7 component-wise, 52 / 72 characteristic,
                                2 Roe (entropy fix below), 3 AUSM+, 4 HLLC on reconstructed primitives
                                (inviscid_type 0) */
  int viscous_order;         /* 0 or 2 */
  int reaction;              /* 1: finite-rate chemistry source */
  int positive_preserving;   /* positive_preserving: limiter on the species fluxes */
  int grad_p_diffusion;      /* gradPInDiffusionFlux */
  int fixed_time_step;
  int reconstruction;        /* schemes 3/4: 1 first order, 2 MUSCL, 3 NND2, minmod limiter
                                (src/Reconstruction.cu:5-46, src/Limiter.cuh:7-41) */
  double dt, cfl;
  double prandtl;            /* prandtl_number (viscous spectral radius of the time step) */
  double rho_ref, v_ref;     /* reference state: weno_eps_scale = (rho_ref v_ref)^2 */
  double entropy_fix_factor; /* Roe flux (scheme 2), >= 0; reference default 0.125 (src/Parameter.cpp:474) */
  int ch_stencil;            /* scheme 52, points of F+/F- that the five acoustic/entropy fields project:
                                0: i-2..i+2 / i-1..i+3 for every field (the stencil of the species fields,
                                of the air specialisation and of WENO-Z7); 1: the mixture code of the
                                reference, i-3..i+1 / i-2..i+2 for the five conservative components
                                (src/WENO.cu:880-881) -- at the first face of a block, whose point i-3 lies
                                beyond the ghost layers, the outermost ghost point stands in */
} ostc_fr_setup;

#define OSTC_FR_FIELD_CV 0
#define OSTC_FR_FIELD_BV 1
#define OSTC_FR_FIELD_SV 2
#define OSTC_FR_FIELD_QN 3
#define OSTC_FR_FIELD_DQ 4
#define OSTC_FR_FIELD_MUL 5
#define OSTC_FR_FIELD_LAMBDA 6
#define OSTC_FR_FIELD_CP 7
#define OSTC_FR_FIELD_GAMMA 8
#define OSTC_FR_FIELD_C 9
#define OSTC_FR_FIELD_MACH 10
#define OSTC_FR_FIELD_RHOD 11
#define OSTC_FR_FIELD_JAC 12
#define OSTC_FR_FIELD_METRIC 13
#define OSTC_FR_FIELD_DT_LOCAL 14

#define OSTC_FR_MODE_DIRECT 0   /* kernels launched one by one */
#define OSTC_FR_MODE_GRAPH 1    /* one step captured in a hipGraph, replayed */
#define OSTC_FR_MODE_TIMED 2    /* direct, with HIP events around every kernel (ostc_fr_kernel_ms) */
/* One launch per kernel kind over a work-list of every block's tiles (source, sweeps, update,
 * properties, local dt), the step captured in a hipGraph and replayed; boundary conditions and
 * inner faces stay per face.  Same results as the per-block modes. */
#define OSTC_FR_MODE_BATCHED 3
#define OSTC_FR_MODE_BATCHED_TIMED 4  /* the batched step launched directly, with HIP events per kernel kind */

/* kernel kinds of ostc_fr_kernel_ms */
#define OSTC_FR_K_SOURCE 0
#define OSTC_FR_K_INVISCID_X 1
#define OSTC_FR_K_INVISCID_Y 2
#define OSTC_FR_K_INVISCID_Z 3
#define OSTC_FR_K_VISCOUS_X 4
#define OSTC_FR_K_VISCOUS_Y 5
#define OSTC_FR_K_VISCOUS_Z 6
#define OSTC_FR_K_UPDATE 7
#define OSTC_FR_K_PROPS 8
#define OSTC_FR_K_COUNT 9

int ostc_fr_supported(int n_spec);  /* 1 if kernels for n_spec species are compiled in */
/* Driver/Field construction for FR (src/Driver.cu:12-113, src/Field.cu:469-647) */
int ostc_fr_create(const ostc_fr_setup* s, const ostc_mech_tables* mech, void* hip_stream, ostc_fr_ctx* out);
int ostc_fr_destroy(ostc_fr_ctx ctx);
int ostc_fr_add_block(ostc_fr_ctx ctx, const ostc_block_desc* d, int* block_id);
int ostc_fr_add_boundary(ostc_fr_ctx ctx, int block, int axis, int side, int type);  /* outflow, periodic */
int ostc_fr_add_inner_face(ostc_fr_ctx ctx, int block, int axis, int side, int target_block);
/* Connectivity records of the reference's input files, as ostc_add_boundary_range
 * and ostc_add_inner_face_range (src/Mesh.cpp:14-176, :790-843): a boundary
 * N.txt line (0-based node range, ff for a farfield label) and an inner N.txt
 * line pair (1-based, the paired tangential axis marked negative; any relative
 * orientation of the two blocks' index axes). */
int ostc_fr_add_boundary_range(ostc_fr_ctx ctx, int block, const int range[6], int type, const ostc_farfield* ff);
int ostc_fr_add_inner_face_range(ostc_fr_ctx ctx, int block, const int range[6], const int target_range[6],
                                 int target_block);
/* Inflow face (label 5) of a mixture, apply_inflow with mass fractions; as ostc_add_inflow. */
int ostc_fr_add_inflow(ostc_fr_ctx ctx, int block, const int range[6], const ostc_inflow* in, const double* y);

/* Sponge layer (src/SpongeLayer.cu:28-172): at the start of every step the
 * running mean of cv over the steps so far is updated, mean = (mean n + cv) / (n + 1)
 * (update_sponge_layer_value), and in every stage after the chemistry source
 * dq -= sigma (cv - mean) on every conservative variable (sponge_layer_source),
 * sigma = sum over the enabled sides of sigma_d S(xi), xi = (s - start_d)/(end_d - start_d)
 * clipped at 1 where the coordinate s lies beyond start_d (below it on the
 * minus sides x-, y-, z-, above it on the plus sides), S(xi) = 6 xi^5 - 15 xi^4 + 10 xi^3
 * (sponge_function 0).  sigma_d = spongeCoefficient * v_char / |start_d - end_d|
 * (DParameter, src/DParameter.cu:258-282).  The mean starts at zero (a new run). */
typedef struct {
  int enabled[6];             /* x-, x+, y-, y+, z-, z+ (sponge_layer_direction 0..5) */
  double start[6], end[6];    /* spongeXMinusStart, spongeXMinusEnd, spongeXPlusStart, ... */
  double sigma[6];            /* sponge_sigma0..5 */
  int function;               /* sponge_function: 0 */
  int jacobian_scaled;        /* 1: dq -= (sigma J)(cv - mean), the scaling of every other source term
                                 (the chemistry's J omega, src/FiniteRateChem.cu:36), so that the
                                 relaxation rate is sigma; 0: the reference's term as written, without J
                                 (src/SpongeLayer.cu:155), whose rate is sigma / J */
} ostc_sponge;
/* Air contexts: the same sponge layer on the five conservative variables,
 * with OSTC_MODE_PLAIN and the stage operators (the fused batched kernel
 * refuses it); before ostc_finalize, per block, as ostc_fr_set_sponge. */
int ostc_set_sponge(ostc_ctx ctx, int block, const ostc_sponge* sp, const double* x, const double* y,
                    const double* z);
/* Before ostc_fr_finalize, per block: x, y, z are the block's node coordinates
 * on the padded domain (host, the bv plane layout); the library keeps sigma of
 * every interior node. */
int ostc_fr_set_sponge(ostc_fr_ctx ctx, int block, const ostc_sponge* sp, const double* x, const double* y,
                       const double* z);
/* ghost layers of every array of the context: 3, or 4 for the WENO-Z7 schemes 71/72 (src/Parameter.cpp:227-238) */
int ostc_fr_ngg(ostc_fr_ctx ctx, int* ngg);
/* store_last_step / compute_residual (src/TimeAdvanceFunc.cu:6-37, src/Residual.cuh:12-187): unscaled
 * RMS of the change of rho, |V|, p and T over every block's interior since the last store. */
int ostc_fr_store_last_step(ostc_fr_ctx ctx);
int ostc_fr_residual(ostc_fr_ctx ctx, double res[4]);
/* Farfield face (label 4) of a mixture, apply_farfield<FR> (src/BoundCond.cuh:755-1022): the subsonic
 * state iterates gamma with cp of the boundary temperature and composition. */
int ostc_fr_add_farfield(ostc_fr_ctx ctx, int block, int axis, int side, const ostc_farfield* ff);
/* FarField::FarField with species (src/BoundCond.cpp:586-700): mass fractions y[n_spec]; the
 * (mach, temperature, reynolds_number) combination uses the host mixture viscosity
 * (compute_viscosity, src/Transport.cu:11-41).  Host-only. */
int ostc_fr_farfield_state(const ostc_mech_tables* tables, const ostc_farfield_spec* in, const double* y,
                           ostc_farfield* out);
int ostc_fr_finalize(ostc_fr_ctx ctx);
int ostc_fr_set_geometry(ostc_fr_ctx ctx, int block, const double* jac, const double* metric, long long elem_stride,
                         long long comp_stride, long long offset);
int ostc_fr_set_field(ostc_fr_ctx ctx, int block, int field, const double* host);
int ostc_fr_get_field(ostc_fr_ctx ctx, int block, int field, double* host);
/* initialize_computation + compute_cv_from_bv for FR: cv from bv/sv, boundary
 * conditions, inner-face exchange, properties */
int ostc_fr_initialize(ostc_fr_ctx ctx);
/* n_steps SSP-RK3 steps (src/RK.cuh:85-226, MixtureModel::FR) */
int ostc_fr_step(ostc_fr_ctx ctx, int n_steps, int mode, double* last_dt);
/* stage pieces: dq = J omega (source); inviscid x, y, z; viscous x, y, z,
 * with the dt the positivity limiter reads (src/RK.cuh:113-139) */
int ostc_fr_stage_rhs(ostc_fr_ctx ctx, int block, double dt);
int ostc_fr_compute_dt(ostc_fr_ctx ctx, double* dt_out);   /* local_time_step + global MIN */
int ostc_fr_stage_update(ostc_fr_ctx ctx, int block, int rk, double dt);
int ostc_fr_update_properties(ostc_fr_ctx ctx, int block);
/* qn = cv (src/RK.cuh:99-102), apply_boundary_conditions of one block
 * (src/BoundCond.cuh:1488-1678) and data_communication over the inner faces
 * (src/DataCommunication.cuh:29-58): the rest of a stage for a driver that
 * runs the pieces itself, e.g. around the parallel exchange below. */
int ostc_fr_copy_qn(ostc_fr_ctx ctx, int block);
int ostc_fr_apply_bc(ostc_fr_ctx ctx, int block);
int ostc_fr_exchange(ostc_fr_ctx ctx);
/* Faces shared with another rank, as ostc_add_parallel_face ... ostc_parallel_copy
 * for the air context; each face point carries 6 primitives and ns mass
 * fractions per layer.  ostc_fr_step refuses a context with parallel faces. */
int ostc_fr_add_parallel_face(ostc_fr_ctx ctx, int block, const int range[6], int rank, int flag_send,
                              int flag_recv);
int ostc_fr_parallel_face_count(ostc_fr_ctx ctx, int* n);
int ostc_fr_parallel_face_info(ostc_fr_ctx ctx, int face, ostc_parallel_info* out);
int ostc_fr_parallel_buffers(ostc_fr_ctx ctx, double** send, double** recv, long long* total);
int ostc_fr_pack_parallel(ostc_fr_ctx ctx);
int ostc_fr_unpack_parallel(ostc_fr_ctx ctx);
int ostc_fr_parallel_copy(ostc_fr_ctx ctx, int which, double* host);
/* timing of the last ostc_fr_step: device ms and launches; per kernel kind
 * (OSTC_FR_K_*) the summed ms and launch counts of the last TIMED call */
int ostc_fr_last_step_timing(ostc_fr_ctx ctx, double* ms, long long* launches);
int ostc_fr_kernel_ms(ostc_fr_ctx ctx, double ms[OSTC_FR_K_COUNT], long long counts[OSTC_FR_K_COUNT]);
int ostc_fr_sync(ostc_fr_ctx ctx);
const char* ostc_fr_last_error(ostc_fr_ctx ctx);
/* species of the mixing-layer initial state: sv[l] = y >= 0 ? y_upper[l] : y_lower[l]
 * over the padded domain, y = y0 + j dy (src/Initialize.cu:259-325) */
int ostc_util_mixing_layer_sv(int mx, int my, int mz, int ngg, int n_spec, double y0, double dy,
                              const double* y_upper, const double* y_lower, double* sv);

#ifdef __cplusplus
}
#endif
#endif
