// ostc_fr_api.hip -- host runtime behind the finite-rate mixture entry points
// of the C ABI (ostc_fr_*, include/openstc/abi.h): block and face registry,
// device memory in the reference layout, the RK3 step of
// RK3<MixtureModel::FR> (src/RK.cuh:85-226) with the dt kept on the device,
// hipGraph capture of a step, and per-kernel event timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "openstc/abi.h"
#include "ostc_faces.hpp"
#include "ostc_fr.hpp"
#include "ostc_inflow.hpp"
#include "ostc_nanscan.hpp"
#include "ostc_repack.hpp"
#include "ostc_sponge.hpp"

using namespace ostc::fr;
using ostc::FaceMap;
using ostc::FaceRange;
using ostc::ParFace;
using ostc::parallel_face_decode;

namespace {

struct FaceDesc {
  int type, ff;  // ff: index of the far-field state (type 4) or of the inflow state (type 5)
  FaceRange r;   // the face layer and its tangential range widened by ngg
  std::vector<double> y;  // type 5: y of the face points, i fastest over r
  size_t y_off = 0;       // type 5: offset of y in d_inflow
};

struct HBlock {
  int mx, my, mz;
  FrBlockDev d{};
  std::vector<FaceDesc> faces;
  std::vector<FaceMap> inner;
  std::vector<double> sp_sigma;  // sponge sigma of every interior node (empty: none)
  bool sp_jac = true;            // the sponge term carries the Jacobian
  ostc_block_desc desc{};        // the caller's descriptor: optional device arrays it owns
  int e[3];  // mx, my, mz
  const int* ext() const { return e; }
};

constexpr int kResidualBlocks = 256;                  // workgroups of the residual reduction
constexpr double kRkA[3] = {1.0, 0.75, 1.0 / 3.0};  // src/RK.cuh:10-14
constexpr double kRkB[3] = {0.0, 0.25, 2.0 / 3.0};
constexpr double kRkC[3] = {1.0, 0.25, 2.0 / 3.0};

}  // namespace

struct ostc_fr_ctx_s {
  ostc_fr_setup s{};
  ostc_mech_tables tab{};
  Tab* d_tab = nullptr;
  double* d_tlog = nullptr;        // log tables of the transport kernel (16 + 256 doubles)
  int ns = 0, nv = 5, ngg = 3;
  hipStream_t st = nullptr;
  bool own_stream = false;
  std::vector<HBlock> blocks;
  std::vector<ostc_farfield> farfields;  // states of the farfield faces, in registration order
  double* d_ff = nullptr;
  std::vector<double> inflows;           // packed inflow states (kInflowDoubles each)
  bool inflow_noise = false;             // some inflow face has white noise
  double* d_inflow = nullptr;            // inflow states, then the y of every inflow face point
  unsigned long long* d_pass = nullptr;  // boundary-condition passes per block
  bool sponge = false;                   // a sponge layer is set on some block
  unsigned long long* d_sp_iter = nullptr;  // steps averaged into the sponge mean
  bool finalized = false;
  double* d_dt = nullptr;          // dt of the current step (one double)
  double* d_res = nullptr;         // per-workgroup partial sums of the residual
  double weno_eps_scale = 1;
  std::vector<void*> allocs;
  std::string err;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long long graph_launches = 0;
  // batched step: the per-block kernel arguments and its own graph
  std::vector<FrArgs> h_args;
  hipGraph_t bgraph = nullptr;
  hipGraphExec_t bgexec = nullptr;
  long long bgraph_launches = 0;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  double last_ms = 0;
  long long last_launches = 0;
  KernelTimer* timer = nullptr;
  double kms[KT_COUNT] = {};
  long long kcount[KT_COUNT] = {};
  // faces shared with other ranks (parallel N.txt) and the send / receive buffers
  std::vector<ParFace> par_faces;
  double *d_par_send = nullptr, *d_par_recv = nullptr;
  long long par_total = 0;
  bool par_packed = false;  // a pack has run since the last unpack or host write of a field
};

namespace {

int fail(ostc_fr_ctx c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}

int hipfail(ostc_fr_ctx c, hipError_t e, const char* what) {
  return fail(c, OSTC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define FR_HIP(c, call)                                  \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return hipfail((c), e_, #call); \
  } while (0)

// Device allocation zeroed on the context's stream.
int dalloc(ostc_fr_ctx c, double** p, size_t n) {
  FR_HIP(c, hipMalloc((void**)p, n * sizeof(double)));
  c->allocs.push_back(*p);
  FR_HIP(c, hipMemsetAsync(*p, 0, n * sizeof(double), c->st));
  return OSTC_OK;
}

FrArgs args(ostc_fr_ctx c, int b) {
  FrArgs a{};
  a.B = c->blocks[b].d;
  a.T = c->d_tab;
  a.tlog = c->d_tlog;
  a.scheme = c->s.inviscid_scheme;
  a.viscous = c->s.viscous_order != 0;
  a.reaction = c->s.reaction != 0;
  a.pp = c->s.positive_preserving != 0;
  a.grad_p = c->s.grad_p_diffusion != 0;
  a.recon = c->s.reconstruction;
  a.cfl = c->s.cfl;
  a.prandtl = c->s.prandtl;
  a.weno_eps_scale = c->weno_eps_scale;
  a.v_ref = c->s.v_ref;
  a.entropy_fix = c->s.entropy_fix_factor;
  a.ch_ref = c->s.ch_stencil;
  a.dt = c->d_dt;
  return a;
}

int check_ready(ostc_fr_ctx c) {
  if (!c) return OSTC_ERR_ARG;
  if (!c->finalized) return fail(c, OSTC_ERR_STATE, "context not finalized");
  return OSTC_OK;
}

// src/BoundCond.cuh:1488-1678: farfield, outflow, inflow, ..., periodic
int bc_priority(int type) {
  return type == OSTC_BC_FARFIELD ? 0 : (type == OSTC_BC_OUTFLOW ? 1 : (type == OSTC_BC_INFLOW ? 2 : 3));
}

// Boundary conditions of one block (priority order, then registration
// order), src/BoundCond.cuh:1488-1678.
void apply_bc(ostc_fr_ctx c, int b, long long* n) {
  std::vector<FaceDesc> fs = c->blocks[b].faces;
  std::stable_sort(fs.begin(), fs.end(),
                   [](const FaceDesc& x, const FaceDesc& y) { return bc_priority(x.type) < bc_priority(y.type); });
  const FrArgs a = args(c, b);
  for (const FaceDesc& f : fs) {
    FaceBox fb{};
    for (int d = 0; d < 3; ++d) {
      fb.lo[d] = f.r.lo[d];
      fb.n[d] = f.r.hi[d] - f.r.lo[d] + 1;
    }
    fb.axis = f.r.axis;
    fb.side = f.r.side;
    const double* ff = f.type == OSTC_BC_FARFIELD ? c->d_ff + (size_t)f.ff * 26
                       : f.type == OSTC_BC_INFLOW ? c->d_inflow + (size_t)f.ff * ostc::kInflowDoubles
                                                  : nullptr;
    launch_bc(a, c->ns, fb, f.type, ff, f.type == OSTC_BC_INFLOW ? c->d_inflow + f.y_off : nullptr, c->d_pass + b,
              c->st, n);
  }
  if (c->inflow_noise) launch_count_pass(c->d_pass + b, c->st, n);
}

// data_communication, inner faces in block order then face order
// (src/DataCommunication.cuh:29-58).
void exchange(ostc_fr_ctx c, long long* n) {
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    const FrArgs a = args(c, (int)b);
    for (const FaceMap& f : c->blocks[b].inner) {
      CommMap m{};
      for (int d = 0; d < 3; ++d) {
        m.lo[d] = f.lo[d];
        m.n[d] = f.hi[d] - f.lo[d] + 1;
        m.code[d] = f.code[d];
        m.off[d] = f.off[d];
      }
      m.axis = f.axis;
      m.shared = f.shared;
      launch_comm(a, c->blocks[f.target].d, c->ns, m, c->st, n);
    }
  }
}

// One SSP-RK3 step, src/RK.cuh:85-226 (MixtureModel::FR, laminar).
void step(ostc_fr_ctx c, long long* n, KernelTimer* tm) {
  const int nb = (int)c->blocks.size();
  if (c->s.fixed_time_step) launch_dt_set(c->d_dt, c->s.dt, c->st, n);
  for (int b = 0; b < nb; ++b) {  // qn <- cv
    const FrBlockDev& d = c->blocks[b].d;
    (void)hipMemcpyAsync(d.qn, d.cv, (size_t)d.sz * c->nv * sizeof(double), hipMemcpyDeviceToDevice, c->st);
    ++*n;
  }
  if (c->sponge) {  // update_sponge_layer_value, then update_sponge_iter (src/RK.cuh:104-109)
    for (int b = 0; b < nb; ++b)
      if (c->blocks[b].d.sp_sigma) launch_sponge_mean(args(c, b), c->ns, c->d_sp_iter, c->st, n);
    launch_count_pass(c->d_sp_iter, c->st, n);
  }
  for (int rk = 0; rk < 3; ++rk) {
    if (!c->s.fixed_time_step && rk == 0) {
      // dt = global MIN of local_time_step (properties of the step's state)
      launch_dt_set(c->d_dt, 1e6, c->st, n);
      for (int b = 0; b < nb; ++b)
        launch_local_dt(args(c, b), reinterpret_cast<unsigned long long*>(c->d_dt), c->st, n);
    }
    for (int b = 0; b < nb; ++b) {
      const FrArgs a = args(c, b);
      launch_rhs(a, c->ns, c->st, n, tm);
      launch_update(a, c->ns, kRkA[rk], kRkB[rk], kRkC[rk], c->st, n, tm);
      apply_bc(c, b, n);
    }
    exchange(c, n);
    for (int b = 0; b < nb; ++b) launch_props(args(c, b), c->ns, c->st, n, tm);
  }
}

// The same step with one launch per kernel kind over every block
// (source, sweeps, update, properties, local dt); boundary conditions and
// inner faces stay per face.  Each block's stage only reads and writes its
// own arrays until the boundary conditions, so running all residuals, then
// all updates, then all boundary conditions gives the per-block order's
// result.
void step_batched(ostc_fr_ctx c, long long* n, KernelTimer* tm) {
  const int nb = (int)c->blocks.size();
  const FrArgs* h = c->h_args.data();
  if (c->s.fixed_time_step) launch_dt_set(c->d_dt, c->s.dt, c->st, n);
  for (int b = 0; b < nb; ++b) {  // qn <- cv
    const FrBlockDev& d = c->blocks[b].d;
    (void)hipMemcpyAsync(d.qn, d.cv, (size_t)d.sz * c->nv * sizeof(double), hipMemcpyDeviceToDevice, c->st);
    ++*n;
  }
  if (c->sponge) {
    for (int b = 0; b < nb; ++b)
      if (c->blocks[b].d.sp_sigma) launch_sponge_mean(h[b], c->ns, c->d_sp_iter, c->st, n);
    launch_count_pass(c->d_sp_iter, c->st, n);
  }
  for (int rk = 0; rk < 3; ++rk) {
    if (!c->s.fixed_time_step && rk == 0) {
      launch_dt_set(c->d_dt, 1e6, c->st, n);
      launch_local_dt_batched(h, nb, reinterpret_cast<unsigned long long*>(c->d_dt), c->st, n);
    }
    launch_rhs_batched(h, nb, c->ns, c->st, n, tm);
    launch_update_batched(h, nb, c->ns, kRkA[rk], kRkB[rk], kRkC[rk], c->st, n, tm);
    for (int b = 0; b < nb; ++b) apply_bc(c, b, n);
    exchange(c, n);
    launch_props_batched(h, nb, c->ns, c->st, n, tm);
  }
}

void invalidate_graph(ostc_fr_ctx c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->bgexec) (void)hipGraphExecDestroy(c->bgexec);
  if (c->bgraph) (void)hipGraphDestroy(c->bgraph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->bgexec = nullptr;
  c->bgraph = nullptr;
}

int build_graph(ostc_fr_ctx c, bool batched) {
  if (batched ? c->bgexec : c->gexec) return OSTC_OK;
  long long n = 0;
  FR_HIP(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  if (batched) step_batched(c, &n, nullptr);
  else step(c, &n, nullptr);
  hipGraph_t g = nullptr;
  FR_HIP(c, hipStreamEndCapture(c->st, &g));
  (batched ? c->bgraph : c->graph) = g;
  FR_HIP(c, hipGraphInstantiate(batched ? &c->bgexec : &c->gexec, g, nullptr, nullptr, 0));
  (batched ? c->bgraph_launches : c->graph_launches) = n;
  return OSTC_OK;
}

int field_ptr(ostc_fr_ctx c, int b, int f, double** p, size_t* n) {
  const FrBlockDev& d = c->blocks[b].d;
  const size_t sz = (size_t)d.sz, n0 = (size_t)d.n0;
  switch (f) {
    case OSTC_FR_FIELD_CV: *p = d.cv; *n = sz * c->nv; break;
    case OSTC_FR_FIELD_BV: *p = d.bv; *n = sz * 6; break;
    case OSTC_FR_FIELD_SV: *p = d.sv; *n = sz * c->ns; break;
    case OSTC_FR_FIELD_QN: *p = d.qn; *n = sz * c->nv; break;
    case OSTC_FR_FIELD_DQ: *p = d.dq; *n = n0 * c->nv; break;
    case OSTC_FR_FIELD_MUL: *p = d.mul; *n = sz; break;
    case OSTC_FR_FIELD_LAMBDA: *p = d.lam; *n = sz; break;
    case OSTC_FR_FIELD_CP: *p = d.cp; *n = sz; break;
    case OSTC_FR_FIELD_GAMMA: *p = d.gam; *n = sz; break;
    case OSTC_FR_FIELD_C: *p = d.cc; *n = sz; break;
    case OSTC_FR_FIELD_MACH: *p = d.mach; *n = sz; break;
    case OSTC_FR_FIELD_RHOD: *p = d.rhod; *n = sz * c->ns; break;
    case OSTC_FR_FIELD_JAC: *p = d.jac; *n = sz; break;
    case OSTC_FR_FIELD_METRIC: *p = d.met; *n = sz * 9; break;
    case OSTC_FR_FIELD_DT_LOCAL: *p = d.dtl; *n = n0; break;
    default: return fail(c, OSTC_ERR_ARG, "unknown field");
  }
  return OSTC_OK;
}

int check_block(ostc_fr_ctx c, int b) {
  if (b < 0 || b >= (int)c->blocks.size()) return fail(c, OSTC_ERR_ARG, "block index out of range");
  return OSTC_OK;
}

}  // namespace

extern "C" {

int ostc_fr_supported(int n_spec) { return supported_ns(n_spec) ? 1 : 0; }

int ostc_fr_create(const ostc_fr_setup* s, const ostc_mech_tables* mech, void* stream, ostc_fr_ctx* out) {
  if (!s || !mech || !out) return OSTC_ERR_ARG;
  *out = nullptr;
  const int sc = s->inviscid_scheme;
  if (sc != 51 && sc != 52 && sc != 71 && sc != 72 && sc != 2 && sc != 3 && sc != 4) return OSTC_ERR_UNSUPPORTED;
  if ((sc == 2 || sc == 3 || sc == 4) && (s->reconstruction < 1 || s->reconstruction > 3))
    return OSTC_ERR_UNSUPPORTED;
  if (s->viscous_order != 0 && s->viscous_order != 2) return OSTC_ERR_UNSUPPORTED;
  if (!supported_ns(mech->n_spec)) return OSTC_ERR_UNSUPPORTED;
  if (s->cfl <= 0 && !s->fixed_time_step) return OSTC_ERR_ARG;
  if (s->fixed_time_step && !(s->dt > 0)) return OSTC_ERR_ARG;
  if (!(s->entropy_fix_factor >= 0) || (s->ch_stencil != 0 && s->ch_stencil != 1)) return OSTC_ERR_ARG;
  auto* c = new ostc_fr_ctx_s();
  c->s = *s;
  c->tab = *mech;
  c->ns = mech->n_spec;
  c->nv = 5 + c->ns;
  c->ngg = (sc == 71 || sc == 72) ? 4 : 3;  // src/Parameter.cpp:227-238
  c->weno_eps_scale = (s->rho_ref * s->v_ref) * (s->rho_ref * s->v_ref);  // src/DParameter.cu:26-29
  if (stream) {
    c->st = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      return OSTC_ERR_HIP;
    }
    c->own_stream = true;
  }
  *out = c;
  return OSTC_OK;
}

int ostc_fr_destroy(ostc_fr_ctx c) {
  if (!c) return OSTC_ERR_ARG;
  if (c->st) (void)hipStreamSynchronize(c->st);
  invalidate_graph(c);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->d_tab) (void)hipFree(c->d_tab);
  if (c->ev_a) (void)hipEventDestroy(c->ev_a);
  if (c->ev_b) (void)hipEventDestroy(c->ev_b);
  if (c->timer) {
    for (int i = 0; i < KernelTimer::kMax; ++i) (void)hipEventDestroy(c->timer->ev[i]);
    delete c->timer;
  }
  if (c->own_stream) (void)hipStreamDestroy(c->st);
  delete c;
  return OSTC_OK;
}

int ostc_fr_add_block(ostc_fr_ctx c, const ostc_block_desc* d, int* id) {
  if (!c || !d) return OSTC_ERR_ARG;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "blocks are fixed after finalize");
  if (d->mz < 2) return fail(c, OSTC_ERR_UNSUPPORTED, "two-dimensional blocks are not supported");
  if (d->mx < 2 || d->my < 2) return fail(c, OSTC_ERR_ARG, "block too small");
  if (d->metric && (d->metric_elem_stride < 1 || d->metric_comp_stride < 1 || d->metric_offset < 0))
    return fail(c, OSTC_ERR_ARG, "device metric: strides must be positive and the offset non-negative");
  HBlock h;
  h.mx = h.e[0] = d->mx;
  h.my = h.e[1] = d->my;
  h.mz = h.e[2] = d->mz;
  h.desc = *d;
  c->blocks.push_back(h);
  if (id) *id = (int)c->blocks.size() - 1;
  return OSTC_OK;
}

// Registers a boundary face of `type` over the points of `r`; `ff` is the
// far-field state of type 4.
static int add_face(ostc_fr_ctx c, int b, const FaceRange& r, int type, const ostc_farfield* ff) {
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "faces are fixed after ostc_fr_finalize");
  if (type == OSTC_BC_FARFIELD) {
    if (!ff) return fail(c, OSTC_ERR_ARG, "farfield faces need their state: use ostc_fr_add_farfield");
    if (!(ff->density > 0) || !(ff->pressure > 0) || !(ff->acoustic_speed > 0) ||
        !(ff->specific_heat_ratio > 1) || !(ff->entropy > 0) || !(ff->mw > 0))
      return fail(c, OSTC_ERR_ARG,
                  "far-field state: density, pressure, acoustic speed, entropy, mw > 0, gamma > 1");
    c->blocks[b].faces.push_back({OSTC_BC_FARFIELD, (int)c->farfields.size(), r});
    c->farfields.push_back(*ff);
    return OSTC_OK;
  }
  if (type == OSTC_BC_INFLOW)
    return fail(c, OSTC_ERR_ARG, "inflow faces (label 5) need their state: use ostc_fr_add_inflow");
  if (type != OSTC_BC_OUTFLOW && type != OSTC_BC_PERIODIC)
    return fail(c, OSTC_ERR_UNSUPPORTED, "mixture path: boundary label " + std::to_string(type) +
                                             " not implemented (farfield 4, inflow 5, outflow 6, periodic 8 are)");
  c->blocks[b].faces.push_back({type, -1, r});
  return OSTC_OK;
}

int ostc_fr_add_boundary(ostc_fr_ctx c, int b, int axis, int side, int type) {
  if (!c) return OSTC_ERR_ARG;
  if (int r = check_block(c, b)) return r;
  if (axis < 0 || axis > 2 || (side != -1 && side != 1)) return fail(c, OSTC_ERR_ARG, "bad face");
  if (type == OSTC_BC_FARFIELD)
    return fail(c, OSTC_ERR_ARG, "farfield faces need their state: use ostc_fr_add_farfield");
  return add_face(c, b, ostc::whole_face(axis, side, c->blocks[b].ext(), c->ngg), type, nullptr);
}

int ostc_fr_add_farfield(ostc_fr_ctx c, int b, int axis, int side, const ostc_farfield* ff) {
  if (!c) return OSTC_ERR_ARG;
  if (int r = check_block(c, b)) return r;
  if (axis < 0 || axis > 2 || (side != -1 && side != 1) || !ff) return fail(c, OSTC_ERR_ARG, "bad face");
  return add_face(c, b, ostc::whole_face(axis, side, c->blocks[b].ext(), c->ngg), OSTC_BC_FARFIELD, ff);
}

int ostc_fr_add_inflow(ostc_fr_ctx c, int b, const int range[6], const ostc_inflow* in, const double* y) {
  if (!c) return OSTC_ERR_ARG;
  if (int r = check_block(c, b)) return r;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "faces are fixed after ostc_fr_finalize");
  if (!range || !in || !y) return fail(c, OSTC_ERR_ARG, "bad inflow face");
  if ((in->inflow_type != 0 && in->inflow_type != 2) || (in->fluctuation_type != 0 && in->fluctuation_type != 1))
    return fail(c, OSTC_ERR_UNSUPPORTED, "inflow_type 0/2 and fluctuation_type 0/1 are implemented");
  if (in->inflow_type == 2 && !(in->delta_omega > 0)) return fail(c, OSTC_ERR_ARG, "inflow: delta_omega > 0");
  for (const double* S : {in->upper, in->lower})
    if (!(S[0] > 0) || !(S[4] > 0) || !(S[5] > 0)) return fail(c, OSTC_ERR_ARG, "inflow: rho, p, T > 0");
  const HBlock& h = c->blocks[b];
  FaceRange fr{};
  std::string err;
  if (!ostc::boundary_range(range, h.ext(), c->ngg, &fr, &err)) return fail(c, OSTC_ERR_ARG, err);
  FaceDesc fd{OSTC_BC_INFLOW, (int)(c->inflows.size() / ostc::kInflowDoubles), fr, {}, 0};
  const int G = c->ngg;
  const long long sx = h.mx + 2 * G, sy = h.my + 2 * G;
  for (int k = fr.lo[2]; k <= fr.hi[2]; ++k)
    for (int j = fr.lo[1]; j <= fr.hi[1]; ++j)
      for (int i = fr.lo[0]; i <= fr.hi[0]; ++i) fd.y.push_back(y[((long long)(k + G) * sy + (j + G)) * sx + (i + G)]);
  c->inflows.resize(c->inflows.size() + ostc::kInflowDoubles);
  ostc::pack_inflow(*in, c->inflows.data() + c->inflows.size() - ostc::kInflowDoubles);
  if (in->fluctuation_type == 1) c->inflow_noise = true;
  c->blocks[b].faces.push_back(std::move(fd));
  return OSTC_OK;
}

int ostc_fr_set_sponge(ostc_fr_ctx c, int b, const ostc_sponge* sp, const double* x, const double* y,
                       const double* z) {
  if (!c) return OSTC_ERR_ARG;
  if (int r = check_block(c, b)) return r;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "the sponge layer is fixed after ostc_fr_finalize");
  if (!sp || !x || !y || !z) return fail(c, OSTC_ERR_ARG, "bad sponge layer");
  if (sp->function != 0) return fail(c, OSTC_ERR_UNSUPPORTED, "sponge_function 0 is implemented");
  for (int d = 0; d < 6; ++d)
    if (sp->enabled[d] && !(std::fabs(sp->start[d] - sp->end[d]) >= 1e-10))
      return fail(c, OSTC_ERR_ARG, "sponge layer: start and end of an enabled side coincide");
  HBlock& h = c->blocks[b];
  const int G = c->ngg;
  const long long sx = h.mx + 2 * G, sy = h.my + 2 * G;
  h.sp_sigma.assign((size_t)h.mx * h.my * h.mz, 0.0);
  h.sp_jac = sp->jacobian_scaled != 0;
  for (int k = 0; k < h.mz; ++k)
    for (int j = 0; j < h.my; ++j)
      for (int i = 0; i < h.mx; ++i) {
        const long long p = ((long long)(k + G) * sy + (j + G)) * sx + (i + G);
        h.sp_sigma[((size_t)k * h.my + j) * h.mx + i] = ostc::sponge_sigma_at(*sp, x[p], y[p], z[p]);
      }
  c->sponge = true;
  return OSTC_OK;
}

int ostc_fr_add_boundary_range(ostc_fr_ctx c, int b, const int range[6], int type, const ostc_farfield* ff) {
  if (!c) return OSTC_ERR_ARG;
  if (int r = check_block(c, b)) return r;
  if (!range) return fail(c, OSTC_ERR_ARG, "bad face");
  FaceRange fr{};
  std::string err;
  if (!ostc::boundary_range(range, c->blocks[b].ext(), c->ngg, &fr, &err)) return fail(c, OSTC_ERR_ARG, err);
  return add_face(c, b, fr, type, ff);
}

// compute_viscosity (host), src/Transport.cu:11-41
static double host_mixture_viscosity(const ostc_mech_tables& T, double temperature, double mw_total,
                                     const double* Y) {
  const int ns = T.n_spec;
  double x[OSTC_MAX_SPEC], vis[OSTC_MAX_SPEC];
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

// Species::compute_cp (Combustion2Part), src/ChemData.cpp:108-150
static void host_cp(const ostc_mech_tables& T, double temp, double* cp) {
  const double t2{temp * temp}, t3{t2 * temp}, t4{t3 * temp};
  for (int i = 0; i < T.n_spec; ++i) {
    if (temp < T.t_low[i]) {
      const double tt = T.t_low[i];
      const double tt2{tt * tt}, tt3{tt2 * tt}, tt4{tt3 * tt};
      const double* c = T.low_coeff[i];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
    } else {
      const double* c = temp < T.t_mid[i] ? T.low_coeff[i] : T.high_coeff[i];
      cp[i] = c[0] + c[1] * temp + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    cp[i] *= kRu / T.mw[i];
  }
}

// FarField::FarField with species (laminar), src/BoundCond.cpp:586-700
int ostc_fr_farfield_state(const ostc_mech_tables* tables, const ostc_farfield_spec* in, const double* y,
                           ostc_farfield* out) {
  if (!tables || !in || !y || !out) return OSTC_ERR_ARG;
  const ostc_mech_tables& T = *tables;
  const int ns = T.n_spec;
  if (ns < 1 || ns > OSTC_MAX_SPEC) return OSTC_ERR_ARG;
  ostc_farfield f{};
  double mach = in->mach, pressure = in->pressure, temperature = in->temperature, velocity = in->velocity,
         density = in->density;
  double u = in->u, v = in->v, w = in->w;
  for (int i = 0; i < ns; ++i) f.sv[i] = y[i];
  double mw = 0;
  for (int i = 0; i < ns; ++i) mw += f.sv[i] / T.mw[i];
  mw = 1 / mw;
  if (temperature < 0) {
    if (!(pressure > 0) || !(density > 0)) return OSTC_ERR_ARG;
    temperature = pressure * mw / (density * kRu);
  }
  const double viscosity = host_mixture_viscosity(T, temperature, mw, f.sv);
  double cpi[OSTC_MAX_SPEC];
  host_cp(T, temperature, cpi);
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
    if (!(in->reynolds_number > 0) || !(velocity > 0)) return OSTC_ERR_ARG;
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
  return OSTC_OK;
}

int ostc_fr_ngg(ostc_fr_ctx c, int* ngg) {
  if (!c || !ngg) return OSTC_ERR_ARG;
  *ngg = c->ngg;
  return OSTC_OK;
}

int ostc_fr_add_inner_face(ostc_fr_ctx c, int b, int axis, int side, int target) {
  if (!c) return OSTC_ERR_ARG;
  if (int r = check_block(c, b)) return r;
  if (int r = check_block(c, target)) return r;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "faces are fixed after ostc_fr_finalize");
  if (axis < 0 || axis > 2 || (side != -1 && side != 1)) return fail(c, OSTC_ERR_ARG, "bad face");
  const HBlock &h = c->blocks[b], &t = c->blocks[target];
  for (int a = 0; a < 3; ++a)
    if (a != axis && h.ext()[a] != t.ext()[a]) return fail(c, OSTC_ERR_ARG, "inner face is not 1-to-1");
  if (t.ext()[axis] < c->ngg + 1) return fail(c, OSTC_ERR_ARG, "target block thinner than ngg+1 layers");
  c->blocks[b].inner.push_back(ostc::aligned_face(axis, side, h.ext(), t.ext(), c->ngg, target));
  return OSTC_OK;
}

int ostc_fr_add_inner_face_range(ostc_fr_ctx c, int b, const int range[6], const int target_range[6], int target) {
  if (!c) return OSTC_ERR_ARG;
  if (int r = check_block(c, b)) return r;
  if (int r = check_block(c, target)) return r;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "faces are fixed after ostc_fr_finalize");
  if (!range || !target_range) return fail(c, OSTC_ERR_ARG, "bad face");
  FaceMap m{};
  std::string err;
  if (!ostc::inner_face_map(range, target_range, c->blocks[b].ext(), c->blocks[target].ext(), c->ngg, target, &m,
                            &err))
    return fail(c, OSTC_ERR_ARG, err);
  c->blocks[b].inner.push_back(m);
  return OSTC_OK;
}

int ostc_fr_add_parallel_face(ostc_fr_ctx c, int b, const int range[6], int rank, int flag_send, int flag_recv) {
  if (!c) return OSTC_ERR_ARG;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "faces are fixed after ostc_fr_finalize");
  if (b < 0 || b >= (int)c->blocks.size() || !range) return fail(c, OSTC_ERR_ARG, "bad parallel face");
  if (c->ngg + 1 > ostc::kParMaxLayers) return fail(c, OSTC_ERR_UNSUPPORTED, "parallel faces: ngg above 4");
  ParFace f{};
  std::string err;
  if (!parallel_face_decode(range, c->blocks[b].ext(), &f, &err)) return fail(c, OSTC_ERR_ARG, err);
  f.block = b;
  f.rank = rank;
  f.flag_send = flag_send;
  f.flag_recv = flag_recv;
  c->par_faces.push_back(f);
  return OSTC_OK;
}

int ostc_fr_finalize(ostc_fr_ctx c) {
  if (!c) return OSTC_ERR_ARG;
  if (c->finalized) return OSTC_OK;
  if (c->blocks.empty()) return fail(c, OSTC_ERR_STATE, "no blocks");
  FR_HIP(c, hipMalloc((void**)&c->d_tab, sizeof(Tab)));
  FR_HIP(c, hipMemcpyAsync(c->d_tab, &c->tab, sizeof(Tab), hipMemcpyHostToDevice, c->st));
  {
    // logarithms of the temperature scales of the collision integrals
    static_assert(OSTC_MAX_SPEC == 16, "log table layout");
    std::vector<double> h(16 + 16 * 16, 0.0);  // the copy below completes before h goes out of scope
    for (int l = 0; l < c->ns; ++l) h[l] = std::log(c->tab.lj_potent_inv[l]);
    for (int m = 0; m < c->ns; ++m)
      for (int n = 0; n < c->ns; ++n)
        if (m != n) h[16 + 16 * m + n] = std::log(c->tab.kb_over_eps_jk[m][n]);
    if (int r = dalloc(c, &c->d_tlog, h.size())) return r;
    FR_HIP(c, hipMemcpyAsync(c->d_tlog, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    FR_HIP(c, hipStreamSynchronize(c->st));
  }
  const int g = c->ngg, nv = c->nv, ns = c->ns;
  for (auto& h : c->blocks) {
    FrBlockDev& d = h.d;
    d.mx = h.mx;
    d.my = h.my;
    d.mz = h.mz;
    d.g = g;
    d.sx = h.mx + 2 * g;
    d.sy = h.my + 2 * g;
    d.sz = (long long)d.sx * d.sy * (h.mz + 2 * g);
    d.n0 = (long long)h.mx * h.my * h.mz;
    const size_t sz = (size_t)d.sz, n0 = (size_t)d.n0;
    // the caller's arrays where it gave them, the library's own otherwise
    const ostc_block_desc& u = h.desc;
    auto take = [&](double* mine, double** p, size_t n) { return mine ? (*p = mine, 0) : dalloc(c, p, n); };
    int r = OSTC_OK;
    if ((r = take(u.cv, &d.cv, sz * nv)) || (r = take(u.qn, &d.qn, sz * nv)) || (r = take(u.bv, &d.bv, sz * 6)) ||
        (r = take(u.sv, &d.sv, sz * ns)) || (r = take(u.dq, &d.dq, n0 * nv)) || (r = take(u.mul, &d.mul, sz)) ||
        (r = dalloc(c, &d.lam, sz)) || (r = dalloc(c, &d.cp, sz)) || (r = dalloc(c, &d.gam, sz)) ||
        (r = dalloc(c, &d.cc, sz)) || (r = dalloc(c, &d.mach, sz)) || (r = dalloc(c, &d.rhod, sz * ns)) ||
        (r = take(u.jac, &d.jac, sz)) || (r = dalloc(c, &d.met, sz * 9)) || (r = dalloc(c, &d.dtl, n0)) ||
        (r = dalloc(c, &d.bvl, n0 * 4)))
      return r;
    if (u.metric)
      ostc::repack_metric(u.metric, u.metric_elem_stride, u.metric_comp_stride, u.metric_offset, d.met, d.sz, c->st);
  }
  FR_HIP(c, hipGetLastError());  // the metric copies of caller-owned geometry
  if (int r = dalloc(c, &c->d_dt, 1)) return r;
  if (int r = dalloc(c, &c->d_res, 4 * kResidualBlocks)) return r;
  if (!c->farfields.empty()) {
    static_assert(sizeof(ostc_farfield) == 26 * sizeof(double), "far-field layout");
    if (int r = dalloc(c, &c->d_ff, c->farfields.size() * 26)) return r;
    FR_HIP(c, hipMemcpyAsync(c->d_ff, c->farfields.data(), c->farfields.size() * sizeof(ostc_farfield),
                             hipMemcpyHostToDevice, c->st));
    FR_HIP(c, hipStreamSynchronize(c->st));
  }
  if (!c->inflows.empty()) {
    std::vector<double> h = c->inflows;  // states, then the y arrays of the inflow faces
    for (auto& B : c->blocks)
      for (auto& f : B.faces)
        if (f.type == OSTC_BC_INFLOW) {
          f.y_off = h.size();
          h.insert(h.end(), f.y.begin(), f.y.end());
        }
    if (int r = dalloc(c, &c->d_inflow, h.size())) return r;
    FR_HIP(c, hipMemcpyAsync(c->d_inflow, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    FR_HIP(c, hipStreamSynchronize(c->st));
  }
  {
    double* p = nullptr;  // one 8-byte counter per block, then the sponge step counter
    if (int r = dalloc(c, &p, c->blocks.size() + 1)) return r;
    c->d_pass = reinterpret_cast<unsigned long long*>(p);
    c->d_sp_iter = c->d_pass + c->blocks.size();
  }
  for (auto& h : c->blocks) {
    if (h.sp_sigma.empty()) continue;
    if (int r = dalloc(c, &h.d.sp_sigma, h.sp_sigma.size())) return r;
    if (int r = dalloc(c, &h.d.sp_mean, h.sp_sigma.size() * c->nv)) return r;
    h.d.sp_jac = h.sp_jac ? 1 : 0;
    FR_HIP(c, hipMemcpyAsync(h.d.sp_sigma, h.sp_sigma.data(), h.sp_sigma.size() * sizeof(double),
                             hipMemcpyHostToDevice, c->st));
    FR_HIP(c, hipStreamSynchronize(c->st));
  }
  if (!c->par_faces.empty()) {  // (6 + ns) transferred values x (ngg + 1) layers per face point
    c->par_total = 0;
    for (ParFace& f : c->par_faces) {
      f.offset = c->par_total;
      c->par_total += (long long)(6 + ns) * (g + 1) * f.points;
    }
    ostc::par_assign_own_source(c->par_faces.data(), (int)c->par_faces.size());
    if (int r = dalloc(c, &c->d_par_send, (size_t)c->par_total)) return r;
    if (int r = dalloc(c, &c->d_par_recv, (size_t)c->par_total)) return r;
  }
  c->h_args.clear();
  for (int b = 0; b < (int)c->blocks.size(); ++b) c->h_args.push_back(args(c, b));
  long long n = 0;
  launch_dt_set(c->d_dt, c->s.fixed_time_step ? c->s.dt : 1e6, c->st, &n);
  FR_HIP(c, hipEventCreate(&c->ev_a));
  FR_HIP(c, hipEventCreate(&c->ev_b));
  FR_HIP(c, hipStreamSynchronize(c->st));
  c->finalized = true;
  return OSTC_OK;
}

int ostc_fr_set_geometry(ostc_fr_ctx c, int b, const double* jac, const double* metric, long long es, long long cs,
                         long long off) {
  if (int r = check_ready(c)) return r;
  if (int r = check_block(c, b)) return r;
  if (!jac || !metric) return fail(c, OSTC_ERR_ARG, "null geometry");
  const FrBlockDev& d = c->blocks[b].d;
  std::vector<double> planes((size_t)d.sz * 9);
  for (long long p = 0; p < d.sz; ++p)
    for (int k = 0; k < 9; ++k) planes[(size_t)k * d.sz + p] = metric[p * es + k * cs + off];
  FR_HIP(c, hipMemcpyAsync(d.jac, jac, (size_t)d.sz * sizeof(double), hipMemcpyHostToDevice, c->st));
  FR_HIP(c, hipMemcpyAsync(d.met, planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
  FR_HIP(c, hipStreamSynchronize(c->st));
  invalidate_graph(c);
  return OSTC_OK;
}

int ostc_fr_set_field(ostc_fr_ctx c, int b, int f, const double* host) {
  if (int r = check_ready(c)) return r;
  if (int r = check_block(c, b)) return r;
  if (!host) return fail(c, OSTC_ERR_ARG, "null host pointer");
  c->par_packed = false;
  double* p = nullptr;
  size_t n = 0;
  if (int r = field_ptr(c, b, f, &p, &n)) return r;
  FR_HIP(c, hipMemcpyAsync(p, host, n * sizeof(double), hipMemcpyHostToDevice, c->st));
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_fr_get_field(ostc_fr_ctx c, int b, int f, double* host) {
  if (int r = check_ready(c)) return r;
  if (int r = check_block(c, b)) return r;
  if (!host) return fail(c, OSTC_ERR_ARG, "null host pointer");
  double* p = nullptr;
  size_t n = 0;
  if (int r = field_ptr(c, b, f, &p, &n)) return r;
  FR_HIP(c, hipMemcpyAsync(host, p, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_fr_initialize(ostc_fr_ctx c) {
  if (int r = check_ready(c)) return r;
  long long n = 0;
  for (int b = 0; b < (int)c->blocks.size(); ++b) launch_cv_from_bv(args(c, b), c->ns, c->st, &n);
  // farfield faces read gamma of their face points: give the property fields
  // the initial state's values before the first boundary pass
  if (!c->farfields.empty())
    for (int b = 0; b < (int)c->blocks.size(); ++b) launch_props(args(c, b), c->ns, c->st, &n, nullptr);
  for (int b = 0; b < (int)c->blocks.size(); ++b) apply_bc(c, b, &n);
  exchange(c, &n);
  for (int b = 0; b < (int)c->blocks.size(); ++b) launch_props(args(c, b), c->ns, c->st, &n, nullptr);
  FR_HIP(c, hipGetLastError());
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_fr_step(ostc_fr_ctx c, int n_steps, int mode, double* last_dt) {
  if (int r = check_ready(c)) return r;
  if (!c->par_faces.empty())
    return fail(c, OSTC_ERR_STATE,
                "parallel faces: run the stage operators, with ostc_fr_pack_parallel / the transfer / "
                "ostc_fr_unpack_parallel after ostc_fr_exchange");
  if (n_steps < 0) return fail(c, OSTC_ERR_ARG, "negative step count");
  if (mode != OSTC_FR_MODE_DIRECT && mode != OSTC_FR_MODE_GRAPH && mode != OSTC_FR_MODE_TIMED &&
      mode != OSTC_FR_MODE_BATCHED && mode != OSTC_FR_MODE_BATCHED_TIMED)
    return fail(c, OSTC_ERR_ARG, "unknown mode");
  long long n = 0;
  KernelTimer* tm = nullptr;
  const bool batched = mode == OSTC_FR_MODE_BATCHED || mode == OSTC_FR_MODE_BATCHED_TIMED;
  const bool graph = mode == OSTC_FR_MODE_GRAPH || mode == OSTC_FR_MODE_BATCHED;
  if (mode == OSTC_FR_MODE_TIMED || mode == OSTC_FR_MODE_BATCHED_TIMED) {
    if (!c->timer) {
      c->timer = new KernelTimer();
      for (int i = 0; i < KernelTimer::kMax; ++i) FR_HIP(c, hipEventCreate(&c->timer->ev[i]));
    }
    tm = c->timer;
    tm->n = 0;
    tm->ok = true;
  }
  if (graph) {
    if (int r = build_graph(c, batched)) return r;
  }
  FR_HIP(c, hipEventRecord(c->ev_a, c->st));
  for (int i = 0; i < n_steps; ++i) {
    if (graph) {
      FR_HIP(c, hipGraphLaunch(batched ? c->bgexec : c->gexec, c->st));
      n += batched ? c->bgraph_launches : c->graph_launches;
    } else if (batched) {
      step_batched(c, &n, tm);
    } else {
      step(c, &n, tm);
    }
  }
  FR_HIP(c, hipEventRecord(c->ev_b, c->st));
  FR_HIP(c, hipGetLastError());
  FR_HIP(c, hipEventSynchronize(c->ev_b));
  float ms = 0;
  FR_HIP(c, hipEventElapsedTime(&ms, c->ev_a, c->ev_b));
  c->last_ms = ms;
  c->last_launches = n;
  if (tm) {
    if (!tm->ok) return fail(c, OSTC_ERR_STATE, "too many timed launches in one call (use fewer steps)");
    std::fill(c->kms, c->kms + KT_COUNT, 0.0);
    std::fill(c->kcount, c->kcount + KT_COUNT, 0LL);
    for (int e = 0; e + 1 < tm->n; e += 2) {
      float k = 0;
      FR_HIP(c, hipEventElapsedTime(&k, tm->ev[e], tm->ev[e + 1]));
      c->kms[tm->kind[e / 2]] += k;
      c->kcount[tm->kind[e / 2]] += 1;
    }
  }
  if (last_dt) FR_HIP(c, hipMemcpy(last_dt, c->d_dt, sizeof(double), hipMemcpyDeviceToHost));
  return OSTC_OK;
}

int ostc_fr_stage_rhs(ostc_fr_ctx c, int b, double dt) {
  if (int r = check_ready(c)) return r;
  if (int r = check_block(c, b)) return r;
  long long n = 0;
  launch_dt_set(c->d_dt, dt, c->st, &n);
  launch_rhs(args(c, b), c->ns, c->st, &n, nullptr);
  FR_HIP(c, hipGetLastError());
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_fr_compute_dt(ostc_fr_ctx c, double* dt_out) {
  if (int r = check_ready(c)) return r;
  long long n = 0;
  launch_dt_set(c->d_dt, 1e6, c->st, &n);
  for (int b = 0; b < (int)c->blocks.size(); ++b)
    launch_local_dt(args(c, b), reinterpret_cast<unsigned long long*>(c->d_dt), c->st, &n);
  FR_HIP(c, hipGetLastError());
  double dt = 0;
  FR_HIP(c, hipMemcpyAsync(&dt, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->st));
  FR_HIP(c, hipStreamSynchronize(c->st));
  if (dt_out) *dt_out = dt;
  return OSTC_OK;
}

int ostc_fr_stage_update(ostc_fr_ctx c, int b, int rk, double dt) {
  if (int r = check_ready(c)) return r;
  if (int r = check_block(c, b)) return r;
  if (rk < 0 || rk > 2) return fail(c, OSTC_ERR_ARG, "rk must be 0, 1 or 2");
  long long n = 0;
  launch_dt_set(c->d_dt, dt, c->st, &n);
  launch_update(args(c, b), c->ns, kRkA[rk], kRkB[rk], kRkC[rk], c->st, &n, nullptr);
  FR_HIP(c, hipGetLastError());
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_fr_copy_qn(ostc_fr_ctx c, int b) {
  if (int r = check_ready(c)) return r;
  if (int r = check_block(c, b)) return r;
  const FrBlockDev& d = c->blocks[b].d;
  FR_HIP(c, hipMemcpyAsync(d.qn, d.cv, (size_t)d.sz * c->nv * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_fr_apply_bc(ostc_fr_ctx c, int b) {
  if (int r = check_ready(c)) return r;
  if (int r = check_block(c, b)) return r;
  long long n = 0;
  apply_bc(c, b, &n);
  FR_HIP(c, hipGetLastError());
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_fr_exchange(ostc_fr_ctx c) {
  if (int r = check_ready(c)) return r;
  long long n = 0;
  exchange(c, &n);
  FR_HIP(c, hipGetLastError());
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_fr_update_properties(ostc_fr_ctx c, int b) {
  if (int r = check_ready(c)) return r;
  if (int r = check_block(c, b)) return r;
  long long n = 0;
  launch_props(args(c, b), c->ns, c->st, &n, nullptr);
  FR_HIP(c, hipGetLastError());
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_fr_store_last_step(ostc_fr_ctx c) {
  if (int r = check_ready(c)) return r;
  for (auto& B : c->blocks) launch_store_last(B.d, c->st);
  FR_HIP(c, hipGetLastError());
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_fr_residual(ostc_fr_ctx c, double res[4]) {
  if (int r = check_ready(c)) return r;
  if (!res) return OSTC_ERR_ARG;
  double acc[4] = {0, 0, 0, 0};
  long long n_total = 0;
  std::vector<double> h(4 * kResidualBlocks);
  for (auto& B : c->blocks) {
    launch_residual(B.d, c->d_res, kResidualBlocks, c->st);
    FR_HIP(c, hipMemcpyAsync(h.data(), c->d_res, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    FR_HIP(c, hipStreamSynchronize(c->st));
    for (int q = 0; q < kResidualBlocks; ++q)
      for (int l = 0; l < 4; ++l) acc[l] += h[4 * q + l];
    n_total += B.d.n0;
  }
  for (int l = 0; l < 4; ++l) res[l] = std::sqrt(acc[l] / (double)n_total);
  if (std::isnan(res[0] + res[1] + res[2] + res[3])) {
    std::string at;  // the first node of the first block whose primitives hold a NaN
    auto* word = reinterpret_cast<unsigned long long*>(c->d_res);
    for (size_t b = 0; b < c->blocks.size() && at.empty(); ++b) {
      const FrBlockDev& d = c->blocks[b].d;
      at = ostc::first_nan_text(d.bv, d.sz, d.mx, d.my, d.mz, d.g, word, c->st);
      if (!at.empty()) at = ": block " + std::to_string(b) + ", " + at;
    }
    return fail(c, OSTC_ERR_NUMERIC, "NaN in the residual" + at);
  }
  return OSTC_OK;
}

int ostc_fr_last_step_timing(ostc_fr_ctx c, double* ms, long long* launches) {
  if (!c) return OSTC_ERR_ARG;
  if (ms) *ms = c->last_ms;
  if (launches) *launches = c->last_launches;
  return OSTC_OK;
}

int ostc_fr_kernel_ms(ostc_fr_ctx c, double ms[OSTC_FR_K_COUNT], long long counts[OSTC_FR_K_COUNT]) {
  if (!c) return OSTC_ERR_ARG;
  static_assert(KT_COUNT == OSTC_FR_K_COUNT, "kernel kinds");
  for (int k = 0; k < KT_COUNT; ++k) {
    if (ms) ms[k] = c->kms[k];
    if (counts) counts[k] = c->kcount[k];
  }
  return OSTC_OK;
}

int ostc_fr_parallel_face_count(ostc_fr_ctx c, int* n) {
  if (!c || !n) return OSTC_ERR_ARG;
  *n = (int)c->par_faces.size();
  return OSTC_OK;
}

int ostc_fr_parallel_face_info(ostc_fr_ctx c, int face, ostc_parallel_info* out) {
  if (int r = check_ready(c)) return r;
  if (!out || face < 0 || face >= (int)c->par_faces.size()) return fail(c, OSTC_ERR_ARG, "bad parallel face index");
  const ParFace& f = c->par_faces[face];
  out->block = f.block;
  out->rank = f.rank;
  out->flag_send = f.flag_send;
  out->flag_recv = f.flag_recv;
  out->offset = f.offset;
  out->count = (long long)(6 + c->ns) * (c->ngg + 1) * f.points;
  out->points = f.points;
  return OSTC_OK;
}

int ostc_fr_parallel_buffers(ostc_fr_ctx c, double** send, double** recv, long long* total) {
  if (int r = check_ready(c)) return r;
  if (send) *send = c->d_par_send;
  if (recv) *recv = c->d_par_recv;
  if (total) *total = c->par_total;
  return OSTC_OK;
}

int ostc_fr_pack_parallel(ostc_fr_ctx c) {
  if (int r = check_ready(c)) return r;
  for (const ParFace& f : c->par_faces) launch_par_pack(c->h_args[f.block], c->ns, f, c->d_par_send, c->st);
  FR_HIP(c, hipGetLastError());
  c->par_packed = true;
  return OSTC_OK;
}

int ostc_fr_unpack_parallel(ostc_fr_ctx c) {
  if (int r = check_ready(c)) return r;
  if (!c->par_packed && !c->par_faces.empty())
    return fail(c, OSTC_ERR_STATE, "ostc_fr_unpack_parallel: call ostc_fr_pack_parallel first (the face nodes take "
                                   "the mean of what was sent and what was received)");
  c->par_packed = false;
  for (const ParFace& f : c->par_faces)
    launch_par_unpack(c->h_args[f.block], c->ns, f, c->d_par_recv, c->d_par_send, c->st);
  FR_HIP(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_fr_parallel_copy(ostc_fr_ctx c, int which, double* host) {
  if (int r = check_ready(c)) return r;
  if (!host || (which != 0 && which != 1)) return fail(c, OSTC_ERR_ARG, "bad parallel buffer copy");
  const size_t bytes = (size_t)c->par_total * sizeof(double);
  if (which == 0) FR_HIP(c, hipMemcpyAsync(host, c->d_par_send, bytes, hipMemcpyDeviceToHost, c->st));
  else FR_HIP(c, hipMemcpyAsync(c->d_par_recv, host, bytes, hipMemcpyHostToDevice, c->st));
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_fr_sync(ostc_fr_ctx c) {
  if (!c) return OSTC_ERR_ARG;
  FR_HIP(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

const char* ostc_fr_last_error(ostc_fr_ctx c) { return c ? c->err.c_str() : "null context"; }

int ostc_util_mixing_layer_sv(int mx, int my, int mz, int g, int ns, double y0, double dy, const double* y_upper,
                              const double* y_lower, double* sv) {
  if (mx < 2 || my < 2 || mz < 2 || g < 1 || ns < 1 || !y_upper || !y_lower || !sv) return OSTC_ERR_ARG;
  const long long sx = mx + 2 * g, sy = my + 2 * g, sz = sx * sy * (mz + 2 * g);
  for (int k = -g; k < mz + g; ++k)
    for (int j = -g; j < my + g; ++j) {
      const double y = y0 + j * dy;
      const double* Y = y >= 0 ? y_upper : y_lower;
      for (int i = -g; i < mx + g; ++i) {
        const long long p = ((long long)(k + g) * sy + (j + g)) * sx + (i + g);
        for (int l = 0; l < ns; ++l) sv[p + l * sz] = Y[l];
      }
    }
  return OSTC_OK;
}

}  // extern "C"
