// ostc_api.hip -- host runtime behind the C ABI (include/openstc/abi.h):
// block and face registry, device memory, the ghost-layer work-lists, the
// per-block (reference-shaped) step and the batched, graph-captured step.
#include "openstc/abi.h"
#include "ostc_faces.hpp"
#include "ostc_sponge.hpp"
#include "ostc_kernels.hpp"
#include "ostc_nanscan.hpp"
#include "ostc_repack.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace ostc;

namespace {

struct FaceDesc {
  int type, ff;  // ff: index of the far-field state (type 4) or of the inflow state (type 5)
  FaceRange r;   // the face layer and its tangential range widened by ngg
  std::vector<double> y;  // type 5: y of the face points, i fastest over r
  size_t y_off = 0;       // type 5: offset of y in d_inflow
};
struct Box { int lo[3], hi[3]; };  // inclusive bounds

bool box_ok(const Box& b) { return b.lo[0] <= b.hi[0] && b.lo[1] <= b.hi[1] && b.lo[2] <= b.hi[2]; }
bool box_hit(const Box& a, const Box& b) {
  if (!box_ok(a) || !box_ok(b)) return false;
  for (int d = 0; d < 3; ++d)
    if (a.hi[d] < b.lo[d] || b.hi[d] < a.lo[d]) return false;
  return true;
}
// The parts of `b` outside the interior [0, m-1]^3.
std::vector<Box> minus_interior(Box r, const int m[3]) {
  std::vector<Box> out;
  for (int a = 0; a < 3 && box_ok(r); ++a) {
    if (r.lo[a] < 0) {
      Box p = r;
      p.hi[a] = std::min(r.hi[a], -1);
      if (box_ok(p)) out.push_back(p);
      r.lo[a] = std::max(r.lo[a], 0);
    }
    if (r.hi[a] > m[a] - 1) {
      Box p = r;
      p.lo[a] = std::max(r.lo[a], m[a]);
      if (box_ok(p)) out.push_back(p);
      r.hi[a] = std::min(r.hi[a], m[a] - 1);
    }
  }
  return out;
}

struct HBlock {
  int m[3];
  std::vector<FaceDesc> faces;
  std::vector<FaceMap> inner;
  std::vector<double> sp_sigma;  // sponge sigma of every interior node (empty: none)
  int sp_jac = 1;
  ostc_block_desc desc{};  // the caller's descriptor: optional device arrays it owns
  DevBlock d{};
};

struct Group { int lo = 0, hi = 0; long long total = 0; };

struct Access { int block; Box box; };

int bc_priority(int type) {  // src/BoundCond.cuh:1488-1678 (farfield, outflow, ..., periodic)
  switch (type) {
    case OSTC_BC_FARFIELD: return 0;
    case OSTC_BC_OUTFLOW: return 1;
    case OSTC_BC_INFLOW: return 2;
    case OSTC_BC_PERIODIC: return 7;
    default: return 99;
  }
}

}  // namespace

struct ostc_ctx_s {
  ostc_setup s{};
  Params P{};
  int ngg = 3;
  hipStream_t st = nullptr;
  bool own_stream = false;
  int cus = 256;
  int dev = 0;  // the device current at ostc_create
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
  DevBlock* d_blocks = nullptr;
  std::vector<DevOp> ops;
  DevOp* d_ops = nullptr;
  std::vector<Group> bc_rounds, comm_levels;
  Group props, rim;
  std::vector<std::vector<int>> plain_bc;
  std::vector<int> plain_comm;
  int tile_index = 0, n_wg = 0, xcd_remap = 1, n_tiles = 0, waves = 0, split = 0;
  int* d_tile_block = nullptr;
  double* d_scratch = nullptr;
  int n_first = 1024;
  double* d_dt = nullptr;
  // properties of the ghost layers after a batched stage: computed by the
  // ghost ops themselves (OSTC_GHOST_PROPS_IN_OPS) or by one more pass
  int ghost_props = OSTC_GHOST_PROPS_IN_OPS;
  // faces shared with other ranks (parallel N.txt) and the send / receive buffers
  std::vector<ParFace> par_faces;
  double *d_par_send = nullptr, *d_par_recv = nullptr;
  long long par_total = 0;
  bool par_packed = false;  // a pack has run since the last unpack or host write of a field
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_valid = false;
  std::vector<hipEvent_t> evpool;  // pairs around fused launches (timed mode)
  long long timed_launches = 0;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  long long graph_launches = 0;
  double last_ms = 0;
  long long last_launches = 0;
  std::vector<void*> allocs;
  std::string err;
  // reverse mode
  std::vector<double*> dq_bar, cv_bar;   // per block, allocated on first use
  double* d_fs = nullptr;                // viscous face cotangents
  long long fs_stride = 0;
  double* d_gs = nullptr;                // partial gathers of the viscous cotangents
  long long gs_stride = 0;
  double* d_lam = nullptr;               // cv cotangent of the reverse sweep (padded)
  double* d_lam_qn = nullptr;            // cotangent of the step's initial state, through qn
  double* d_w = nullptr;                 // dq cotangent of one stage
  double* d_obj = nullptr;               // theta row sums, their cotangents, theta
  double* d_part = nullptr;              // gradient partials
  std::vector<double*> ckpt, stage;      // saved states, 12 planes each (cv, bv, mul)
  int adj_steps = 0, adj_every = 0;
  bool adj_ready = false;
  ostc_theta adj_theta{};
  hipEvent_t ev_f0 = nullptr, ev_f1 = nullptr, ev_b0 = nullptr, ev_b1 = nullptr;
  // dual time stepping + DPLUR (per-block path), armed by ostc_set_dual_time
  ostc_dual_time dual{};
  std::vector<DualBlock> dual_blocks;
  bool dual_ready = false;
  double dual_res_scale[4] = {1, 1, 1, 1};
};


namespace {

int fail(ostc_ctx c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail((c), OSTC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device allocation, zeroed on the context's stream: the stream is
// non-blocking, so a memset on the null stream would not be ordered before
// the work queued on it next.
int dalloc(ostc_ctx c, double** p, size_t n) {
  void* v = nullptr;
  HIPCHK(c, hipMalloc(&v, n * sizeof(double)));
  HIPCHK(c, hipMemsetAsync(v, 0, n * sizeof(double), c->st));
  c->allocs.push_back(v);
  *p = static_cast<double*>(v);
  return OSTC_OK;
}

DevOp make_op(int type, int block, const Box& b) {
  DevOp op{};
  op.type = type;
  op.block = block;
  op.target = block;
  for (int d = 0; d < 3; ++d) {
    op.lo[d] = b.lo[d];
    op.n[d] = b.hi[d] - b.lo[d] + 1;
  }
  op.count = (long long)op.n[0] * op.n[1] * op.n[2];
  return op;
}

// The points of a boundary face's range with the normal index running lo..hi.
Box face_box(const FaceRange& r, int lo, int hi) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = r.lo[d];
    b.hi[d] = r.hi[d];
  }
  b.lo[r.axis] = lo;
  b.hi[r.axis] = hi;
  return b;
}

// Image of a source box of an inner face in the target block.
Box map_box(const FaceMap& m, const Box& b) {
  Box t;
  for (int d = 0; d < 3; ++d) {
    const int c = m.code[d], p = (c < 0 ? -c : c) - 1;
    const int x = m.off[d] + (c < 0 ? -b.lo[p] : b.lo[p]), y = m.off[d] + (c < 0 ? -b.hi[p] : b.hi[p]);
    t.lo[d] = std::min(x, y);
    t.hi[d] = std::max(x, y);
  }
  return t;
}

Box box_and(const Box& a, const Box& b) {
  Box r;
  for (int d = 0; d < 3; ++d) {
    r.lo[d] = std::max(a.lo[d], b.lo[d]);
    r.hi[d] = std::min(a.hi[d], b.hi[d]);
  }
  return r;
}

// Work-lists for the ghost layers.  Boundary ops of one block keep the
// reference's sequential order (one launch "round" per position in the
// block's priority-sorted list; blocks are independent).  Inner-face ops keep
// the reference's order (blocks, then faces) only between ops whose cell
// sets conflict: each op gets level 1 + the highest level of an earlier
// conflicting op, and one launch runs one level.
int build_ops(ostc_ctx c) {
  const int G = c->ngg;
  const int nb = (int)c->blocks.size();
  std::vector<std::vector<DevOp>> bc_of_block(nb);
  std::vector<DevOp> rim_ops, props_ops;
  for (int b = 0; b < nb; ++b) {
    HBlock& B = c->blocks[b];
    std::vector<FaceDesc> f = B.faces;
    std::stable_sort(f.begin(), f.end(),
                     [](const FaceDesc& x, const FaceDesc& y) { return bc_priority(x.type) < bc_priority(y.type); });
    for (const FaceDesc& fd : f) {
      const int a = fd.r.axis, m = B.m[a];
      const int idx = fd.r.side < 0 ? 0 : m - 1;
      const int ot = fd.type == OSTC_BC_PERIODIC   ? OP_BC_PERIODIC
                     : fd.type == OSTC_BC_FARFIELD ? OP_BC_FARFIELD
                     : fd.type == OSTC_BC_INFLOW   ? OP_BC_INFLOW
                                                   : OP_BC_OUTFLOW;
      DevOp op = make_op(ot, b, face_box(fd.r, idx, idx));
      op.axis = a;
      op.side = fd.r.side;
      if (fd.type == OSTC_BC_FARFIELD) op.ff = c->d_ff + (size_t)fd.ff * kFarfieldDoubles;
      if (fd.type == OSTC_BC_INFLOW) {
        op.ff = c->d_inflow + (size_t)fd.ff * kInflowDoubles;
        op.aux = c->d_inflow + fd.y_off;
        op.pass = c->d_pass;
      }
      bc_of_block[b].push_back(op);
      // ghost cells this op reads before the same stage can have written them
      Box rd = fd.type == OSTC_BC_PERIODIC
                   ? (fd.r.side < 0 ? face_box(fd.r, m - 1 - G, m - 2) : face_box(fd.r, 1, G))
                   : face_box(fd.r, idx, idx);
      for (const Box& rb : minus_interior(rd, B.m)) rim_ops.push_back(make_op(OP_COPY_BV, b, rb));
      if (fd.type == OSTC_BC_FARFIELD || fd.type == OSTC_BC_INFLOW) {
        // the boundary state replaces the face points after the fused update
        // computed their properties: recompute them with the ghost layers
        Box fp = face_box(fd.r, idx, idx);
        for (int d = 0; d < 3; ++d) {
          fp.lo[d] = std::max(fp.lo[d], 0);
          fp.hi[d] = std::min(fp.hi[d], B.m[d] - 1);
        }
        props_ops.push_back(make_op(OP_PROPS, b, fp));
      }
    }
    Box pad;
    for (int d = 0; d < 3; ++d) { pad.lo[d] = -G; pad.hi[d] = B.m[d] - 1 + G; }
    for (const Box& sb : minus_interior(pad, B.m)) props_ops.push_back(make_op(OP_PROPS, b, sb));
  }
  // Inner faces.  The per-block path launches one op per face in the
  // reference's order (blocks, then faces).  The batched path splits every
  // face of a two-sided interface into
  //  * a bulk part, the face without a frame of ngg+1 cells along its
  //    tangential rims: no other face's op reads or writes these cells, so
  //    all bulk parts run in one launch; one side of the pair (the high face
  //    of a low/high pair, else the face registered first) averages the
  //    shared node layer and recomputes cv in both blocks, the other side
  //    copies its ghost layers only;
  //  * the frame, whose cells neighbouring faces also touch: frame ops keep
  //    the reference's order wherever their cell sets conflict (level =
  //    1 + the highest level of an earlier conflicting frame op).
  // Ghost layers read only the target's interior layers beyond its shared
  // layer, which no inner-face op writes.
  struct FaceGeom { int A, T; FaceMap M; Box box, shA; int partner; };
  std::vector<FaceGeom> geo;
  for (int b = 0; b < nb; ++b) {
    for (const FaceMap& f : c->blocks[b].inner) {
      FaceGeom F;
      F.A = b;
      F.T = f.target;
      F.M = f;
      for (int d = 0; d < 3; ++d) {
        F.box.lo[d] = F.shA.lo[d] = f.lo[d];
        F.box.hi[d] = F.shA.hi[d] = f.hi[d];
      }
      F.shA.lo[f.axis] = F.shA.hi[f.axis] = f.shared;
      F.partner = -1;
      geo.push_back(F);
      props_ops.push_back(make_op(OP_PROPS, b, F.shA));  // shared layers after averaging
    }
  }
  // the partner of a face: the target's face back to this block over the
  // same shared nodes
  for (size_t q = 0; q < geo.size(); ++q) {
    const Box shT = map_box(geo[q].M, geo[q].shA);
    for (size_t r = 0; r < geo.size() && geo[q].partner < 0; ++r) {
      if (r == q || geo[r].A != geo[q].T || geo[r].T != geo[q].A) continue;
      const Box& o = geo[r].shA;
      bool same = true;
      for (int d = 0; d < 3; ++d) same = same && o.lo[d] == shT.lo[d] && o.hi[d] == shT.hi[d];
      if (same) geo[q].partner = (int)r;
    }
  }
  auto comm_op = [&](const FaceGeom& F, const Box& bx) {
    DevOp op = make_op(OP_COMM, F.A, bx);
    op.target = F.T;
    op.axis = F.M.axis;
    op.side = F.M.side;
    op.shared = F.M.shared;
    for (int d = 0; d < 3; ++d) {
      op.tcode[d] = F.M.code[d];
      op.toff[d] = F.M.off[d];
    }
    return op;
  };
  std::vector<DevOp> plain_ops, bulk_ops;
  struct CommOp { DevOp op; std::vector<Access> w, r; int level; };
  std::vector<CommOp> edge;
  for (size_t q = 0; q < geo.size(); ++q) {
    const FaceGeom& F = geo[q];
    plain_ops.push_back(comm_op(F, F.box));
    const int a = F.M.axis, W = G + 1;
    Box core = F.box;
    for (int d = 0; d < 3; ++d)
      if (d != a) { core.lo[d] = F.box.lo[d] + W; core.hi[d] = F.box.hi[d] - W; }
    std::vector<Box> frame;
    if (F.partner >= 0 && box_ok(core)) {
      const FaceGeom& P = geo[F.partner];
      const bool averages = F.M.side != P.M.side ? F.M.side > 0 : (int)q < F.partner;
      Box cb = core;
      if (!averages) {  // ghost layers only; the partner averages
        if (F.M.side < 0) cb.hi[a] = F.M.shared - 1;
        else cb.lo[a] = F.M.shared + 1;
      }
      DevOp op = comm_op(F, cb);
      if (averages) op.flags = 1;  // recompute the target block's cv on the shared layer
      bulk_ops.push_back(op);
      // frame = box minus core, as slabs along the first tangential axis
      // and then the second
      const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
      Box r = F.box;
      Box p = r; p.hi[t1] = F.box.lo[t1] + W - 1; frame.push_back(p);
      p = r; p.lo[t1] = F.box.hi[t1] - W + 1; frame.push_back(p);
      r.lo[t1] = core.lo[t1]; r.hi[t1] = core.hi[t1];
      p = r; p.hi[t2] = F.box.lo[t2] + W - 1; frame.push_back(p);
      p = r; p.lo[t2] = F.box.hi[t2] - W + 1; frame.push_back(p);
    } else {
      frame.push_back(F.box);
    }
    for (const Box& eb : frame) {
      CommOp co;
      co.op = comm_op(F, eb);
      const Box sh = box_and(F.shA, eb);
      co.w = {{F.A, eb}, {F.T, map_box(F.M, sh)}};
      co.r = {{F.T, map_box(F.M, eb)}, {F.A, sh}};
      co.level = 0;
      for (auto& e : edge) {
        auto meet = [](const std::vector<Access>& x, const std::vector<Access>& y) {
          for (auto& u : x)
            for (auto& v : y)
              if (u.block == v.block && box_hit(u.box, v.box)) return true;
          return false;
        };
        if (meet(e.w, co.w) || meet(e.w, co.r) || meet(e.r, co.w)) co.level = std::max(co.level, e.level + 1);
      }
      edge.push_back(co);
    }
  }
  // assemble: [bc rounds][comm levels][props][rim]
  c->ops.clear();
  c->bc_rounds.clear();
  c->comm_levels.clear();
  c->plain_bc.assign(nb, {});
  size_t rounds = 0;
  for (auto& v : bc_of_block) rounds = std::max(rounds, v.size());
  auto close_group = [&](Group& g, int lo) {
    g.lo = lo;
    g.hi = (int)c->ops.size();
    long long acc = 0;
    for (int i = g.lo; i < g.hi; ++i) { c->ops[i].begin = acc; acc += c->ops[i].count; }
    g.total = acc;
  };
  for (size_t r = 0; r < rounds; ++r) {
    const int lo = (int)c->ops.size();
    for (int b = 0; b < nb; ++b)
      if (r < bc_of_block[b].size()) {
        c->plain_bc[b].push_back((int)c->ops.size());
        c->ops.push_back(bc_of_block[b][r]);
      }
    Group g;
    close_group(g, lo);
    c->bc_rounds.push_back(g);
  }
  if (!bulk_ops.empty()) {
    const int lo = (int)c->ops.size();
    for (auto& o : bulk_ops) c->ops.push_back(o);
    Group g;
    close_group(g, lo);
    c->comm_levels.push_back(g);
  }
  int max_level = -1;
  for (auto& e : edge) max_level = std::max(max_level, e.level);
  for (int L = 0; L <= max_level; ++L) {
    const int lo = (int)c->ops.size();
    for (size_t q = 0; q < edge.size(); ++q)
      if (edge[q].level == L) c->ops.push_back(edge[q].op);
    Group g;
    close_group(g, lo);
    c->comm_levels.push_back(g);
  }
  c->plain_comm.clear();
  for (auto& o : plain_ops) {
    // one launch per face: each op is a group of its own (begin = 0)
    c->plain_comm.push_back((int)c->ops.size());
    o.begin = 0;
    c->ops.push_back(o);
  }
  {
    const int lo = (int)c->ops.size();
    for (auto& o : props_ops) c->ops.push_back(o);
    close_group(c->props, lo);
  }
  {
    const int lo = (int)c->ops.size();
    for (auto& o : rim_ops) c->ops.push_back(o);
    close_group(c->rim, lo);
  }
  if (!c->ops.empty()) {
    void* v = nullptr;
    HIPCHK(c, hipMalloc(&v, c->ops.size() * sizeof(DevOp)));
    c->allocs.push_back(v);
    c->d_ops = static_cast<DevOp*>(v);
    HIPCHK(c, hipMemcpy(c->d_ops, c->ops.data(), c->ops.size() * sizeof(DevOp), hipMemcpyHostToDevice));
  }
  return OSTC_OK;
}

void invalidate_graph(ostc_ctx c) {
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_valid = false;
}

int upload_blocks(ostc_ctx c) {
  std::vector<DevBlock> h;
  for (auto& B : c->blocks) h.push_back(B.d);
  HIPCHK(c, hipMemcpy(c->d_blocks, h.data(), h.size() * sizeof(DevBlock), hipMemcpyHostToDevice));
  return OSTC_OK;
}

// The fused batched kernel and the reverse mode implement the component-wise
// WENO schemes 50/51/71 and the characteristic WENO-Z 52/72; the primitive
// reconstruction + Riemann schemes (1/2/3/4) and EP6 run on the per-block kernels.
bool fused_scheme(const ostc_ctx c) {
  const int sc = c->P.scheme;
  return sc == 51 || sc == 50 || sc == 71 || sc == 52 || sc == 72;
}

int need_fused(ostc_ctx c) {
  for (auto& B : c->blocks)  // the fused kernel indexes a block's planes with 32-bit offsets
    if (9 * B.d.sz >= (1LL << 31))
      return fail(c, OSTC_ERR_UNSUPPORTED, "block too large for the batched kernel (9 padded planes >= 2^31 doubles)");
  if (fused_scheme(c)) return OSTC_OK;
  return fail(c, OSTC_ERR_UNSUPPORTED,
              "inviscid_scheme 1/2/3/4/6 run with OSTC_MODE_PLAIN and the stage operators only");
}

int set_tiles(ostc_ctx c, int index, int n_wg) {
  const TileShape t = fused_shape(c->P.scheme, index);
  const int lds = fused_lds_bytes(c->P.scheme, index);
  if (lds < 0 || lds > 160 * 1024) return fail(c, OSTC_ERR_ARG, "tile shape does not fit in LDS");
  // the tile list and block table are replaced below; queued kernels read them
  HIPCHK(c, hipStreamSynchronize(c->st));
  std::vector<int> tb;
  int base = 0;
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    DevBlock& d = c->blocks[b].d;
    d.ntx = (d.mx + t.tx - 1) / t.tx;
    d.nty = (d.my + t.ty - 1) / t.ty;
    d.ntz = (d.mz + t.tz - 1) / t.tz;
    d.tile_base = base;
    const int n = d.ntx * d.nty * d.ntz;
    for (int i = 0; i < n; ++i) tb.push_back((int)b);
    base += n;
  }
  c->n_tiles = base;
  if (c->d_tile_block) hipFree(c->d_tile_block);
  void* v = nullptr;
  HIPCHK(c, hipMalloc(&v, tb.size() * sizeof(int)));
  c->d_tile_block = static_cast<int*>(v);
  HIPCHK(c, hipMemcpy(c->d_tile_block, tb.data(), tb.size() * sizeof(int), hipMemcpyHostToDevice));
  c->tile_index = index;
  c->n_wg = n_wg > 0 ? std::min(n_wg, c->n_tiles) : c->n_tiles;
  invalidate_graph(c);
  return upload_blocks(c);
}

FusedArgs fused_args(ostc_ctx c, int in, int out, int rk) {
  FusedArgs a{};
  a.blocks = c->d_blocks;
  a.tile_block = c->d_tile_block;
  a.n_tiles = c->n_tiles;
  a.in = in;
  a.out = out;
  const RkCoef k = rk_coef(rk);
  a.ra = k.a;
  a.rb = k.b;
  a.rc = k.c;
  a.dt = c->d_dt;
  a.P = c->P;
  a.xcd_remap = c->xcd_remap;
  a.target_block = -1;
  return a;
}

void launch_group(ostc_ctx c, const Group& g, int slot, int src, long long* n, bool props = false) {
  if (g.hi > g.lo)
    ops_launch(c->d_blocks, (int)c->blocks.size(), c->d_ops, g.lo, g.hi, g.total, slot, src, c->st, n, props);
}

void launch_op(ostc_ctx c, int idx, int slot, long long* n) {
  ops_launch(c->d_blocks, (int)c->blocks.size(), c->d_ops, idx, idx + 1, c->ops[idx].count, slot, slot, c->st, n);
}

// ---------------------------------------------------- per-block path pieces
int plain_bc(ostc_ctx c, int b, long long* n) {
  for (int idx : c->plain_bc[b]) launch_op(c, idx, 0, n);
  if (c->inflow_noise) count_pass(c->d_pass, b, b + 1, c->st, n);
  return OSTC_OK;
}

int plain_exchange(ostc_ctx c, long long* n) {
  for (int idx : c->plain_comm) launch_op(c, idx, 0, n);
  return OSTC_OK;
}

int plain_dt(ostc_ctx c, double* dt, long long* n) {
  double v = 1e6;
  for (auto& B : c->blocks) {
    plain_local_dt(B.d, c->P, c->st, n);
    min_reduce(B.d.dtl, B.d.n0, c->d_scratch, c->n_first, c->st, n);
    double h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, c->d_scratch, sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    v = std::min(v, h);
  }
  *dt = v;
  return OSTC_OK;
}

// One SSP-RK3 step shaped like src/RK.cuh:85-226.
int plain_step(ostc_ctx c, double* dt_io, long long* n) {
  for (auto& B : c->blocks) {
    HIPCHK(c, hipMemcpyAsync(B.d.cv[1], B.d.cv[0], 5 * B.d.sz * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    ++*n;
  }
  if (c->sponge) {  // update_sponge_layer_value, then update_sponge_iter (src/RK.cuh:104-109)
    for (auto& B : c->blocks)
      if (B.d.sp_sigma) sponge_mean(B.d, c->d_sp_iter, c->st, n);
    count_pass(c->d_sp_iter, 0, 1, c->st, n);
  }
  double dt = *dt_io;
  for (int rk = 0; rk < 3; ++rk) {
    for (auto& B : c->blocks) {
      HIPCHK(c, hipMemsetAsync(B.d.dq, 0, 5 * B.d.n0 * sizeof(double), c->st));
      ++*n;
    }
    if (!c->s.fixed_time_step && rk == 0) {
      int r = plain_dt(c, &dt, n);
      if (r) return r;
    }
    for (size_t b = 0; b < c->blocks.size(); ++b) {
      const DevBlock& B = c->blocks[b].d;
      plain_rhs(B, c->P, c->st, n);
      plain_update(B, dt, rk, c->st, n);
      plain_bc(c, (int)b, n);
    }
    plain_exchange(c, n);
    for (auto& B : c->blocks) padded_pass(B.d, 0, OP_PROPS, c->st, n);
  }
  *dt_io = dt;
  return OSTC_OK;
}

// ------------------------------------------------ dual time stepping
// inner_converged (src/DualTimeStepping.cu:33-134, one rank): RMS change of
// (rho, |V|, p, T) over the inner iteration, scaled by its value at
// iteration 1; at the last iteration a change above 1e-3 adds 5 iterations (up
// to max_inner_iteration); below 1e-3 the count drops to this iteration and
// the loop ends (*stop = true).
int dual_converged(ostc_ctx c, int iter, bool* stop) {
  *stop = false;
  constexpr int nblk = 256;
  double acc[4] = {0, 0, 0, 0};
  long long n_total = 0;
  std::vector<double> h(4 * nblk);
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    const DevBlock& B = c->blocks[b].d;
    dual_last_square(B, c->dual_blocks[b], c->d_scratch + 1, nblk, c->st);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_scratch + 1, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int q = 0; q < nblk; ++q)
      for (int l = 0; l < 4; ++l) acc[l] += h[4 * q + l];
    n_total += B.n0;
  }
  double res[4];
  for (int l = 0; l < 4; ++l) res[l] = std::sqrt(acc[l] / (double)n_total);
  if (iter == 1)
    for (int l = 0; l < 4; ++l) c->dual_res_scale[l] = res[l] < 1e-20 ? 1e-20 : res[l];
  for (int l = 0; l < 4; ++l) res[l] /= c->dual_res_scale[l];
  double err_max = res[0];
  for (int l = 1; l < 4; ++l)
    if (std::abs(res[l]) > err_max) err_max = res[l];
  if (std::isnan(err_max))
    return fail(c, OSTC_ERR_NUMERIC, "NaN in the inner-iteration residual of dual time step " +
                                         std::to_string(c->dual.step) + ", iteration " + std::to_string(iter));
  int& inner = c->dual.inner_iteration;
  if (iter == inner && err_max > 1e-3 && inner < c->dual.max_inner_iteration)
    inner = std::min(inner + 5, c->dual.max_inner_iteration);
  if (err_max < 1e-3) {
    if (inner - iter > 0) inner = iter;
    *stop = true;
  }
  return OSTC_OK;
}

// One physical step of dual_time_stepping (src/DualTimeStepping.cuh:33-236)
// with dual_time_stepping_implicit_treat and DPLUR (:239-249,
// src/DPLUR.cuh:271-327), blocks in order as the reference loops over them.
int dual_step(ostc_ctx c, long long* n) {
  const double dt = c->dual.dt, diag_factor = 1.5 / dt;
  const int step = ++c->dual.step;
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    const DevBlock& B = c->blocks[b].d;
    const DualBlock& D = c->dual_blocks[b];
    dual_qn_star(B, D, dt, c->st, n);
    HIPCHK(c, hipMemcpyAsync(D.qn1, B.cv[0], 5 * (size_t)B.sz * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  }
  const bool monitor = c->dual.iteration_adjust_step > 0 && step % c->dual.iteration_adjust_step == 0;
  for (int iter = 1; iter <= c->dual.inner_iteration; ++iter) {
    for (size_t b = 0; b < c->blocks.size(); ++b) {
      const DevBlock& B = c->blocks[b].d;
      const DualBlock& D = c->dual_blocks[b];
      if (monitor) dual_last_store(B, D, c->st);
      HIPCHK(c, hipMemsetAsync(B.dq, 0, 5 * B.n0 * sizeof(double), c->st));
      plain_rhs(B, c->P, c->st, n);
      dual_pre(B, c->P, D, dt, diag_factor, c->st, n);
      for (int it = 0; it < c->dual.dplur_inner_step; ++it) {
        dual_dplur(B, D, diag_factor, c->st, n);
        HIPCHK(c, hipMemcpyAsync(B.dq, D.dqk, 5 * B.n0 * sizeof(double), hipMemcpyDeviceToDevice, c->st));
      }
      dual_update(B, c->st, n);
      plain_bc(c, (int)b, n);
    }
    plain_exchange(c, n);
    for (auto& B : c->blocks) padded_pass(B.d, 0, OP_PROPS, c->st, n);
    if (monitor) {
      bool stop = false;
      const int r = dual_converged(c, iter, &stop);
      if (r) return r;
      if (stop) break;
    }
  }
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

// ------------------------------------------------------- batched step
// One batched step.  ev_base >= 0: record event pairs evpool[ev_base + 2 rk],
// evpool[ev_base + 2 rk + 1] around the fused kernel of stage rk.
// Stage rk of a batched step: ring slot rk -> slot (rk+1) % 3, then the
// ghost layers of the new slot.  ev_base >= 0: events evpool[ev_base + 2 rk],
// evpool[ev_base + 2 rk + 1] around the fused kernel.
int batched_stage(ostc_ctx c, int rk, long long* n, int ev_base) {
  const bool events = ev_base >= 0;
  const int in = rk, out = (rk + 1) % kRing;
  FusedArgs a = fused_args(c, in, out, rk);
  if (events) HIPCHK(c, hipEventRecord(c->evpool[ev_base + 2 * rk], c->st));
  if (fused_launch(c->tile_index, c->split ? 2 : 0, a, c->n_wg, c->st) < 0)
    return fail(c, OSTC_ERR_UNSUPPORTED, "fused kernel: scheme/tile not available");
  *n += c->split ? 2 : 1;
  if (events) HIPCHK(c, hipEventRecord(c->evpool[ev_base + 2 * rk + 1], c->st));
  const bool fold = c->ghost_props == OSTC_GHOST_PROPS_IN_OPS;
  launch_group(c, c->rim, out, in, n, fold);
  for (auto& g : c->bc_rounds) launch_group(c, g, out, out, n, fold);
  if (c->inflow_noise) count_pass(c->d_pass, 0, (int)c->blocks.size(), c->st, n);
  for (auto& g : c->comm_levels) launch_group(c, g, out, out, n, fold);
  if (!fold) launch_group(c, c->props, out, out, n);
  return OSTC_OK;
}

int batched_record(ostc_ctx c, long long* n, int ev_base) {
  if (c->sponge) {  // update_sponge_layer_value, then update_sponge_iter (src/RK.cuh:104-109), on slot 0
    for (auto& B : c->blocks)
      if (B.d.sp_sigma) sponge_mean(B.d, c->d_sp_iter, c->st, n);
    count_pass(c->d_sp_iter, 0, 1, c->st, n);
  }
  if (!c->s.fixed_time_step) {
    FusedArgs a = fused_args(c, 0, 1, 0);
    dt_rows(a, (int)c->blocks.size(), c->cus * 4, c->d_dt, c->st, n);
  }
  for (int rk = 0; rk < 3; ++rk) {
    const int r = batched_stage(c, rk, n, ev_base);
    if (r) return r;
  }
  return OSTC_OK;
}

int build_graph(ostc_ctx c) {
  invalidate_graph(c);
  long long n = 0;
  HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  int r = batched_record(c, &n, -1);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->st, &g);
  if (r == OSTC_OK && e == hipSuccess) {
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e == hipSuccess) {
      c->graph = g;
      c->gexec = ex;
      c->graph_valid = true;
      c->graph_launches = n;
      return OSTC_OK;
    }
  }
  if (g) hipGraphDestroy(g);
  (void)hipGetLastError();
  if (r) return r;
  return fail(c, OSTC_ERR_HIP, std::string("graph capture failed: ") + hipGetErrorString(e));
}

int check_ready(ostc_ctx c) {
  if (!c) return OSTC_ERR_ARG;
  if (!c->finalized) return fail(c, OSTC_ERR_STATE, "ostc_finalize has not been called");
  return OSTC_OK;
}

int check_block(ostc_ctx c, int b) {
  int r = check_ready(c);
  if (r) return r;
  if (b < 0 || b >= (int)c->blocks.size()) return fail(c, OSTC_ERR_ARG, "block index out of range");
  return OSTC_OK;
}

}  // namespace

// =================================================================== ABI
extern "C" {

int ostc_create(const ostc_setup* s, void* hip_stream, ostc_ctx* out) {
  if (!s || !out) return OSTC_ERR_ARG;
  *out = nullptr;
  if (s->n_var != 5) return OSTC_ERR_UNSUPPORTED;
  const int sc = s->inviscid_scheme;
  if (sc != 51 && sc != 50 && sc != 71 && sc != 52 && sc != 72 && sc != 1 && sc != 2 && sc != 3 && sc != 4 &&
      sc != 6)
    return OSTC_ERR_UNSUPPORTED;
  if ((sc == 1 || sc == 2 || sc == 3 || sc == 4) && (s->reconstruction < 1 || s->reconstruction > 3))
    return OSTC_ERR_UNSUPPORTED;
  if (s->viscous_order != 0 && s->viscous_order != 2) return OSTC_ERR_UNSUPPORTED;
  if (!(s->cfl > 0) || !(s->prandtl > 0)) return OSTC_ERR_ARG;
  if (!(s->entropy_fix_factor >= 0)) return OSTC_ERR_ARG;
  auto* c = new ostc_ctx_s;
  c->s = *s;
  c->ngg = (sc == 71 || sc == 72) ? 4 : 3;  // src/Parameter.cpp:227-238
  c->P.n_var = 5;
  c->P.scheme = s->inviscid_scheme;
  c->P.visc = s->viscous_order == 2;
  c->P.recon = s->reconstruction;
  c->P.cfl = s->cfl;
  c->P.prandtl = s->prandtl;
  c->P.weno_eps_scale = s->rho_ref * s->v_ref * s->rho_ref * s->v_ref;  // src/DParameter.cu:26-29
  c->P.v_ref = s->v_ref;
  c->P.entropy_fix = s->entropy_fix_factor;
  if (hip_stream) {
    c->st = static_cast<hipStream_t>(hip_stream);
  } else {
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      return OSTC_ERR_HIP;
    }
    c->own_stream = true;
  }
  hipGetDevice(&c->dev);
  hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, c->dev);
  *out = c;
  return OSTC_OK;
}

int ostc_destroy(ostc_ctx c) {
  if (!c) return OSTC_ERR_ARG;
  if (c->st) hipStreamSynchronize(c->st);
  invalidate_graph(c);
  for (auto& e : c->evpool) hipEventDestroy(e);
  if (c->ev_a) hipEventDestroy(c->ev_a);
  if (c->ev_b) hipEventDestroy(c->ev_b);
  for (hipEvent_t e : {c->ev_f0, c->ev_f1, c->ev_b0, c->ev_b1})
    if (e) hipEventDestroy(e);
  for (void* p : c->allocs) hipFree(p);
  if (c->d_tile_block) hipFree(c->d_tile_block);
  if (c->own_stream) hipStreamDestroy(c->st);
  delete c;
  return OSTC_OK;
}

int ostc_add_block(ostc_ctx c, const ostc_block_desc* d, int* id) {
  if (!c || !d) return OSTC_ERR_ARG;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "blocks are fixed after ostc_finalize");
  if (d->mx < c->ngg + 2 || d->my < c->ngg + 2 || d->mz < c->ngg + 2)
    return fail(c, OSTC_ERR_UNSUPPORTED, "each extent must be at least ngg+2 (three-dimensional blocks only)");
  if (d->metric && (d->metric_elem_stride < 1 || d->metric_comp_stride < 1 || d->metric_offset < 0))
    return fail(c, OSTC_ERR_ARG, "device metric: strides must be positive and the offset non-negative");
  HBlock B;
  B.m[0] = d->mx;
  B.m[1] = d->my;
  B.m[2] = d->mz;
  B.desc = *d;
  c->blocks.push_back(B);
  if (id) *id = (int)c->blocks.size() - 1;
  return OSTC_OK;
}

// Registers a boundary face of `type` over the points of `r`; `ff` is the
// far-field state of type 4.
static int add_face(ostc_ctx c, int block, const FaceRange& r, int type, const ostc_farfield* ff) {
  if (type == OSTC_BC_FARFIELD) {
    if (!ff) return fail(c, OSTC_ERR_ARG, "farfield faces need their state");
    if (!(ff->density > 0) || !(ff->pressure > 0) || !(ff->acoustic_speed > 0) || !(ff->specific_heat_ratio > 1) ||
        !(ff->entropy > 0))
      return fail(c, OSTC_ERR_ARG, "far-field state: density, pressure, acoustic speed, entropy > 0 and gamma > 1");
    c->blocks[block].faces.push_back({OSTC_BC_FARFIELD, (int)c->farfields.size(), r});
    c->farfields.push_back(*ff);
    return OSTC_OK;
  }
  if (type == OSTC_BC_INFLOW) return fail(c, OSTC_ERR_ARG, "inflow faces (label 5) need their state: use ostc_add_inflow");
  if (type != OSTC_BC_OUTFLOW && type != OSTC_BC_PERIODIC)
    return fail(c, OSTC_ERR_UNSUPPORTED, "boundary label " + std::to_string(type) +
                                             " not implemented (farfield 4, inflow 5, outflow 6, periodic 8 are)");
  c->blocks[block].faces.push_back({type, -1, r});
  return OSTC_OK;
}

static int check_face_args(ostc_ctx c, int block, int axis, int side) {
  if (!c) return OSTC_ERR_ARG;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "faces are fixed after ostc_finalize");
  if (block < 0 || block >= (int)c->blocks.size() || axis < 0 || axis > 2 || (side != -1 && side != 1))
    return fail(c, OSTC_ERR_ARG, "bad boundary face");
  return OSTC_OK;
}

int ostc_add_boundary(ostc_ctx c, int block, int axis, int side, int type) {
  if (int r = check_face_args(c, block, axis, side)) return r;
  if (type == OSTC_BC_FARFIELD)
    return fail(c, OSTC_ERR_ARG, "farfield faces need their state: use ostc_add_farfield");
  return add_face(c, block, whole_face(axis, side, c->blocks[block].m, c->ngg), type, nullptr);
}

int ostc_add_farfield(ostc_ctx c, int block, int axis, int side, const ostc_farfield* ff) {
  if (int r = check_face_args(c, block, axis, side)) return r;
  if (!ff) return fail(c, OSTC_ERR_ARG, "bad farfield face");
  return add_face(c, block, whole_face(axis, side, c->blocks[block].m, c->ngg), OSTC_BC_FARFIELD, ff);
}

int ostc_add_inflow(ostc_ctx c, int block, const int range[6], const ostc_inflow* in, const double* y) {
  if (!c) return OSTC_ERR_ARG;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "faces are fixed after ostc_finalize");
  if (block < 0 || block >= (int)c->blocks.size() || !range || !in || !y) return fail(c, OSTC_ERR_ARG, "bad inflow face");
  if ((in->inflow_type != 0 && in->inflow_type != 2) || (in->fluctuation_type != 0 && in->fluctuation_type != 1))
    return fail(c, OSTC_ERR_UNSUPPORTED, "inflow_type 0/2 and fluctuation_type 0/1 are implemented");
  if (in->inflow_type == 2 && !(in->delta_omega > 0)) return fail(c, OSTC_ERR_ARG, "inflow: delta_omega > 0");
  for (const double* S : {in->upper, in->lower})
    if (!(S[0] > 0) || !(S[4] > 0) || !(S[5] > 0)) return fail(c, OSTC_ERR_ARG, "inflow: rho, p, T > 0");
  const HBlock& B = c->blocks[block];
  FaceRange r{};
  std::string err;
  if (!boundary_range(range, B.m, c->ngg, &r, &err)) return fail(c, OSTC_ERR_ARG, err);
  FaceDesc fd{OSTC_BC_INFLOW, (int)(c->inflows.size() / kInflowDoubles), r, {}};
  const int G = c->ngg;
  const long long sx = B.m[0] + 2 * G, sy = B.m[1] + 2 * G;
  for (int k = r.lo[2]; k <= r.hi[2]; ++k)
    for (int j = r.lo[1]; j <= r.hi[1]; ++j)
      for (int i = r.lo[0]; i <= r.hi[0]; ++i) fd.y.push_back(y[((long long)(k + G) * sy + (j + G)) * sx + (i + G)]);
  c->inflows.resize(c->inflows.size() + kInflowDoubles);
  pack_inflow(*in, c->inflows.data() + c->inflows.size() - kInflowDoubles);
  if (in->fluctuation_type == 1) c->inflow_noise = true;
  c->blocks[block].faces.push_back(std::move(fd));
  return OSTC_OK;
}

int ostc_set_sponge(ostc_ctx c, int block, const ostc_sponge* sp, const double* x, const double* y,
                    const double* z) {
  if (!c) return OSTC_ERR_ARG;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "the sponge layer is fixed after ostc_finalize");
  if (block < 0 || block >= (int)c->blocks.size() || !sp || !x || !y || !z)
    return fail(c, OSTC_ERR_ARG, "bad sponge layer");
  if (sp->function != 0) return fail(c, OSTC_ERR_UNSUPPORTED, "sponge_function 0 is implemented");
  for (int d = 0; d < 6; ++d)
    if (sp->enabled[d] && !(std::fabs(sp->start[d] - sp->end[d]) >= 1e-10))
      return fail(c, OSTC_ERR_ARG, "sponge layer: start and end of an enabled side coincide");
  HBlock& B = c->blocks[block];
  const int G = c->ngg;
  const long long sx = B.m[0] + 2 * G, sy = B.m[1] + 2 * G;
  B.sp_sigma.assign((size_t)B.m[0] * B.m[1] * B.m[2], 0.0);
  B.sp_jac = sp->jacobian_scaled != 0;
  for (int k = 0; k < B.m[2]; ++k)
    for (int j = 0; j < B.m[1]; ++j)
      for (int i = 0; i < B.m[0]; ++i) {
        const long long p = ((long long)(k + G) * sy + (j + G)) * sx + (i + G);
        B.sp_sigma[((size_t)k * B.m[1] + j) * B.m[0] + i] = sponge_sigma_at(*sp, x[p], y[p], z[p]);
      }
  c->sponge = true;
  return OSTC_OK;
}

int ostc_add_boundary_range(ostc_ctx c, int block, const int range[6], int type, const ostc_farfield* ff) {
  if (!c) return OSTC_ERR_ARG;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "faces are fixed after ostc_finalize");
  if (block < 0 || block >= (int)c->blocks.size() || !range) return fail(c, OSTC_ERR_ARG, "bad boundary face");
  FaceRange r{};
  std::string err;
  if (!boundary_range(range, c->blocks[block].m, c->ngg, &r, &err)) return fail(c, OSTC_ERR_ARG, err);
  return add_face(c, block, r, type, ff);
}

// FarField::FarField for air (n_spec = 0, laminar), src/BoundCond.cpp:586-700.
int ostc_farfield_state_air(const ostc_farfield_spec* in, ostc_farfield* out) {
  if (!in || !out) return OSTC_ERR_ARG;
  ostc_farfield f{};
  double mach = in->mach, pressure = in->pressure, temperature = in->temperature, velocity = in->velocity,
         density = in->density;
  double u = in->u, v = in->v, w = in->w;
  const double mw = kMwAir;
  if (temperature < 0) {
    if (!(pressure > 0) || !(density > 0)) return OSTC_ERR_ARG;
    temperature = pressure * mw / (density * kRu);
  }
  const double viscosity = 1.7894e-5 * std::pow(temperature / 288.16, 1.5) * (288.16 + 110) / (temperature + 110);
  const double gamma = kGamma;
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

int ostc_add_inner_face(ostc_ctx c, int block, int axis, int side, int target) {
  if (!c) return OSTC_ERR_ARG;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "faces are fixed after ostc_finalize");
  const int nb = (int)c->blocks.size();
  if (block < 0 || block >= nb || target < 0 || target >= nb || axis < 0 || axis > 2 || (side != -1 && side != 1))
    return fail(c, OSTC_ERR_ARG, "bad inner face");
  const HBlock &A = c->blocks[block], &T = c->blocks[target];
  for (int d = 0; d < 3; ++d)
    if (d != axis && A.m[d] != T.m[d])
      return fail(c, OSTC_ERR_ARG, "inner face: tangential extents of the two blocks differ");
  if (T.m[axis] < c->ngg + 1) return fail(c, OSTC_ERR_ARG, "inner face: target block thinner than ngg+1 layers");
  c->blocks[block].inner.push_back(aligned_face(axis, side, A.m, T.m, c->ngg, target));
  return OSTC_OK;
}

int ostc_add_inner_face_range(ostc_ctx c, int block, const int range[6], const int target_range[6], int target) {
  if (!c) return OSTC_ERR_ARG;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "faces are fixed after ostc_finalize");
  const int nb = (int)c->blocks.size();
  if (block < 0 || block >= nb || target < 0 || target >= nb || !range || !target_range)
    return fail(c, OSTC_ERR_ARG, "bad inner face");
  FaceMap m{};
  std::string err;
  if (!inner_face_map(range, target_range, c->blocks[block].m, c->blocks[target].m, c->ngg, target, &m, &err))
    return fail(c, OSTC_ERR_ARG, err);
  c->blocks[block].inner.push_back(m);
  return OSTC_OK;
}

int ostc_add_parallel_face(ostc_ctx c, int block, const int range[6], int rank, int flag_send, int flag_recv) {
  if (!c) return OSTC_ERR_ARG;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "faces are fixed after ostc_finalize");
  if (block < 0 || block >= (int)c->blocks.size() || !range) return fail(c, OSTC_ERR_ARG, "bad parallel face");
  if (c->ngg + 1 > ostc::kParMaxLayers) return fail(c, OSTC_ERR_UNSUPPORTED, "parallel faces: ngg above 4");
  ParFace f{};
  std::string err;
  if (!parallel_face_decode(range, c->blocks[block].m, &f, &err)) return fail(c, OSTC_ERR_ARG, err);
  f.block = block;
  f.rank = rank;
  f.flag_send = flag_send;
  f.flag_recv = flag_recv;
  c->par_faces.push_back(f);
  return OSTC_OK;
}

int ostc_finalize(ostc_ctx c) {
  if (!c) return OSTC_ERR_ARG;
  if (c->finalized) return fail(c, OSTC_ERR_STATE, "already finalized");
  if (c->blocks.empty()) return fail(c, OSTC_ERR_STATE, "no blocks");
  const int G = c->ngg;
  for (auto& B : c->blocks) {
    DevBlock& d = B.d;
    d.mx = B.m[0];
    d.my = B.m[1];
    d.mz = B.m[2];
    d.g = G;
    d.sx = d.mx + 2 * G;
    d.sy = d.my + 2 * G;
    d.sz = (long long)d.sx * d.sy * (d.mz + 2 * G);
    d.n0 = (long long)d.mx * d.my * d.mz;
    // the caller's arrays where it gave them (slot 0 is the state, qn is
    // slot 1), the library's own otherwise
    const ostc_block_desc& u = B.desc;
    auto take = [&](double* mine, double** p, size_t n) { return mine ? (*p = mine, 0) : dalloc(c, p, n); };
    int r = 0;
    for (int q = 0; q < kRing && !r; ++q) {
      r = take(q == 0 ? u.cv : (q == 1 ? u.qn : nullptr), &d.cv[q], 5 * d.sz);
      if (!r) r = take(q == 0 ? u.bv : nullptr, &d.bv[q], 6 * d.sz);
      if (!r) r = take(q == 0 ? u.mul : nullptr, &d.mul[q], d.sz);
    }
    double *jac = nullptr, *met = nullptr;
    if (!r) r = dalloc(c, &d.mach, d.sz);
    if (!r) r = take(u.jac, &jac, d.sz);
    if (!r) r = dalloc(c, &met, 9 * d.sz);
    if (!r && u.metric) repack_metric(u.metric, u.metric_elem_stride, u.metric_comp_stride, u.metric_offset, met, d.sz,
                                      c->st);
    if (!r) r = take(u.dq, &d.dq, 5 * d.n0);
    if (!r) r = dalloc(c, &d.dtl, d.n0);
    if (!r) r = dalloc(c, &d.bv_last, 4 * d.n0);
    if (r) return r;
    d.jac = jac;
    d.met = met;
  }
  {
    void* v = nullptr;
    HIPCHK(c, hipMalloc(&v, c->blocks.size() * sizeof(DevBlock)));
    c->allocs.push_back(v);
    c->d_blocks = static_cast<DevBlock*>(v);
  }
  if (!c->farfields.empty()) {
    static_assert(sizeof(ostc_farfield) == kFarfieldDoubles * sizeof(double), "far-field layout");
    void* v = nullptr;
    HIPCHK(c, hipMalloc(&v, c->farfields.size() * sizeof(ostc_farfield)));
    c->allocs.push_back(v);
    c->d_ff = static_cast<double*>(v);
    HIPCHK(c, hipMemcpy(c->d_ff, c->farfields.data(), c->farfields.size() * sizeof(ostc_farfield),
                        hipMemcpyHostToDevice));
  }
  if (!c->inflows.empty()) {
    std::vector<double> h = c->inflows;
    for (auto& B : c->blocks)
      for (auto& f : B.faces)
        if (f.type == OSTC_BC_INFLOW) {
          f.y_off = h.size();
          h.insert(h.end(), f.y.begin(), f.y.end());
        }
    int r = dalloc(c, &c->d_inflow, h.size());
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(c->d_inflow, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  {
    double* p = nullptr;  // one 8-byte counter per block, then the sponge step counter
    if (int r = dalloc(c, &p, c->blocks.size() + 1)) return r;
    c->d_pass = reinterpret_cast<unsigned long long*>(p);
    c->d_sp_iter = c->d_pass + c->blocks.size();
  }
  for (auto& B : c->blocks) {
    if (B.sp_sigma.empty()) continue;
    double *sig = nullptr, *mean = nullptr;
    if (int r = dalloc(c, &sig, B.sp_sigma.size())) return r;
    if (int r = dalloc(c, &mean, 5 * B.sp_sigma.size())) return r;
    HIPCHK(c, hipMemcpyAsync(sig, B.sp_sigma.data(), B.sp_sigma.size() * sizeof(double), hipMemcpyHostToDevice,
                             c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    B.d.sp_sigma = sig;
    B.d.sp_mean = mean;
    B.d.sp_jac = B.sp_jac;
  }
  HIPCHK(c, hipGetLastError());  // the metric copies of caller-owned geometry
  int r = dalloc(c, &c->d_scratch, 1 + 4 * (size_t)c->n_first);
  if (!r) r = dalloc(c, &c->d_dt, 1);
  if (!r && !c->par_faces.empty()) {  // n_trans (6) x (ngg + 1) layers per face point
    c->par_total = 0;
    for (ParFace& f : c->par_faces) {
      f.offset = c->par_total;
      c->par_total += 6ll * (G + 1) * f.points;
    }
    par_assign_own_source(c->par_faces.data(), (int)c->par_faces.size());
    r = dalloc(c, &c->d_par_send, (size_t)c->par_total);
    if (!r) r = dalloc(c, &c->d_par_recv, (size_t)c->par_total);
  }
  if (!r) r = build_ops(c);
  if (r) return r;
  HIPCHK(c, hipEventCreate(&c->ev_a));
  HIPCHK(c, hipEventCreate(&c->ev_b));
  for (hipEvent_t* e : {&c->ev_f0, &c->ev_f1, &c->ev_b0, &c->ev_b1}) HIPCHK(c, hipEventCreate(e));
  c->dq_bar.assign(c->blocks.size(), nullptr);
  c->cv_bar.assign(c->blocks.size(), nullptr);
  if (c->s.fixed_time_step) HIPCHK(c, hipMemcpy(c->d_dt, &c->s.dt, sizeof(double), hipMemcpyHostToDevice));
  c->finalized = true;
  return set_tiles(c, 0, 0);
}

int ostc_ngg(ostc_ctx c, int* ngg) {
  if (!c || !ngg) return OSTC_ERR_ARG;
  *ngg = c->ngg;
  return OSTC_OK;
}

int ostc_zone_view(ostc_ctx c, int b, ostc_zone* z) {
  int r = check_block(c, b);
  if (r) return r;
  if (!z) return OSTC_ERR_ARG;
  const DevBlock& d = c->blocks[b].d;
  z->mx = d.mx; z->my = d.my; z->mz = d.mz; z->ngg = d.g; z->sz = d.sz;
  z->cv = d.cv[0]; z->bv = d.bv[0]; z->mul = d.mul[0]; z->mach = d.mach;
  z->jac = const_cast<double*>(d.jac); z->metric = const_cast<double*>(d.met);
  z->dq = d.dq; z->dt_local = d.dtl;
  return OSTC_OK;
}

int ostc_set_geometry(ostc_ctx c, int b, const double* jac, const double* metric, long long es, long long cs,
                      long long off) {
  int r = check_block(c, b);
  if (r) return r;
  if (!jac || !metric) return fail(c, OSTC_ERR_ARG, "null geometry");
  const DevBlock& d = c->blocks[b].d;
  HIPCHK(c, hipStreamSynchronize(c->st));  // kernels in flight may read the old geometry
  std::vector<double> m(9 * (size_t)d.sz);
  for (long long p = 0; p < d.sz; ++p) {
    if (!(jac[p] > 0)) return fail(c, OSTC_ERR_NUMERIC, "non-positive Jacobian");
    for (int q = 0; q < 9; ++q) m[q * d.sz + p] = metric[p * es + q * cs + off];
  }
  HIPCHK(c, hipMemcpy(const_cast<double*>(d.jac), jac, d.sz * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(const_cast<double*>(d.met), m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice));
  return OSTC_OK;
}

// The cotangent fields of block b (allocated on first use).
static int adj_fields(ostc_ctx c, int b) {
  const DevBlock& d = c->blocks[b].d;
  int r = OSTC_OK;
  if (!c->dq_bar[b]) r = dalloc(c, &c->dq_bar[b], 5 * d.n0);
  if (!r && !c->cv_bar[b]) r = dalloc(c, &c->cv_bar[b], 5 * d.sz);
  return r;
}

static int field_ptr(ostc_ctx c, int b, int f, int slot, double** p, size_t* n) {
  const DevBlock& d = c->blocks[b].d;
  if (f == OSTC_FIELD_DQ_BAR || f == OSTC_FIELD_CV_BAR) {
    const int r = adj_fields(c, b);
    if (r) return r;
  }
  switch (f) {
    case OSTC_FIELD_CV: *p = d.cv[slot]; *n = 5 * d.sz; break;
    case OSTC_FIELD_BV: *p = d.bv[slot]; *n = 6 * d.sz; break;
    case OSTC_FIELD_MUL: *p = d.mul[slot]; *n = d.sz; break;
    case OSTC_FIELD_MACH: *p = d.mach; *n = d.sz; break;
    case OSTC_FIELD_DQ: *p = d.dq; *n = 5 * d.n0; break;
    case OSTC_FIELD_DT_LOCAL: *p = d.dtl; *n = d.n0; break;
    case OSTC_FIELD_JAC: *p = const_cast<double*>(d.jac); *n = d.sz; break;
    case OSTC_FIELD_METRIC: *p = const_cast<double*>(d.met); *n = 9 * d.sz; break;
    case OSTC_FIELD_QN: *p = d.cv[1]; *n = 5 * d.sz; break;
    case OSTC_FIELD_DQ_BAR: *p = c->dq_bar[b]; *n = 5 * d.n0; break;
    case OSTC_FIELD_CV_BAR: *p = c->cv_bar[b]; *n = 5 * d.sz; break;
    default: return fail(c, OSTC_ERR_ARG, "unknown field");
  }
  return OSTC_OK;
}

int ostc_set_field(ostc_ctx c, int b, int f, const double* host) {
  int r = check_block(c, b);
  if (r) return r;
  if (!host) return OSTC_ERR_ARG;
  c->par_packed = false;
  const bool ring = f == OSTC_FIELD_CV || f == OSTC_FIELD_BV || f == OSTC_FIELD_MUL;
  for (int slot = 0; slot < (ring ? kRing : 1); ++slot) {
    double* p;
    size_t n;
    r = field_ptr(c, b, f, slot, &p, &n);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(p, host, n * sizeof(double), hipMemcpyHostToDevice, c->st));
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_get_field(ostc_ctx c, int b, int f, double* host) {
  int r = check_block(c, b);
  if (r) return r;
  if (!host) return OSTC_ERR_ARG;
  double* p;
  size_t n;
  r = field_ptr(c, b, f, 0, &p, &n);
  if (r) return r;
  HIPCHK(c, hipMemcpyAsync(host, p, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_initialize(ostc_ctx c) {
  int r = check_ready(c);
  if (r) return r;
  long long n = 0;
  for (size_t b = 0; b < c->blocks.size(); ++b) plain_bc(c, (int)b, &n);
  plain_exchange(c, &n);
  for (auto& B : c->blocks) padded_pass(B.d, 0, OP_PROPS, c->st, &n);
  for (auto& B : c->blocks) padded_pass(B.d, 0, OP_CV_FROM_BV, c->st, &n);
  // every ring slot starts identical, so ghost cells that no operator
  // rewrites hold the same value whichever slot a stage reads
  for (auto& B : c->blocks)
    for (int q = 1; q < kRing; ++q) {
      HIPCHK(c, hipMemcpyAsync(B.d.cv[q], B.d.cv[0], 5 * B.d.sz * 8, hipMemcpyDeviceToDevice, c->st));
      HIPCHK(c, hipMemcpyAsync(B.d.bv[q], B.d.bv[0], 6 * B.d.sz * 8, hipMemcpyDeviceToDevice, c->st));
      HIPCHK(c, hipMemcpyAsync(B.d.mul[q], B.d.mul[0], B.d.sz * 8, hipMemcpyDeviceToDevice, c->st));
    }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_step(ostc_ctx c, int n_steps, int mode, double* last_dt) {
  int r = check_ready(c);
  if (r) return r;
  if (!c->par_faces.empty())
    return fail(c, OSTC_ERR_STATE,
                "parallel faces: run the stage operators, with ostc_pack_parallel / the transfer / "
                "ostc_unpack_parallel after ostc_exchange");
  if (n_steps < 0) return fail(c, OSTC_ERR_ARG, "negative step count");
  long long n = 0;
  HIPCHK(c, hipEventRecord(c->ev_a, c->st));
  double dt = c->s.fixed_time_step ? c->s.dt : 1e6;
  if (mode == OSTC_MODE_PLAIN) {
    for (int s = 0; s < n_steps; ++s) {
      r = plain_step(c, &dt, &n);
      if (r) return r;
    }
  } else if (mode == OSTC_MODE_BATCHED) {
    if ((r = need_fused(c))) return r;
    if (!c->graph_valid) {
      r = build_graph(c);
      if (r) return r;
    }
    for (int s = 0; s < n_steps; ++s) HIPCHK(c, hipGraphLaunch(c->gexec, c->st));
    n = c->graph_launches * n_steps;
    if (last_dt && !c->s.fixed_time_step)
      HIPCHK(c, hipMemcpyAsync(&dt, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->st));
  } else if (mode == OSTC_MODE_BATCHED_TIMED) {
    if ((r = need_fused(c))) return r;
    while ((long long)c->evpool.size() < 6ll * n_steps) {
      hipEvent_t e;
      HIPCHK(c, hipEventCreate(&e));
      c->evpool.push_back(e);
    }
    for (int s = 0; s < n_steps; ++s) {
      r = batched_record(c, &n, 6 * s);
      if (r) return r;
    }
    c->timed_launches = 3ll * n_steps;
    if (last_dt && !c->s.fixed_time_step)
      HIPCHK(c, hipMemcpyAsync(&dt, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->st));
  } else {
    return fail(c, OSTC_ERR_ARG, "unknown mode");
  }
  HIPCHK(c, hipEventRecord(c->ev_b, c->st));
  HIPCHK(c, hipGetLastError());
  c->last_launches = n;
  if (last_dt) {
    HIPCHK(c, hipStreamSynchronize(c->st));
    *last_dt = dt;
  }
  return OSTC_OK;
}

int ostc_parallel_face_count(ostc_ctx c, int* n) {
  if (!c || !n) return OSTC_ERR_ARG;
  *n = (int)c->par_faces.size();
  return OSTC_OK;
}

int ostc_parallel_face_info(ostc_ctx c, int face, ostc_parallel_info* out) {
  int r = check_ready(c);
  if (r) return r;
  if (!out || face < 0 || face >= (int)c->par_faces.size()) return fail(c, OSTC_ERR_ARG, "bad parallel face index");
  const ParFace& f = c->par_faces[face];
  out->block = f.block;
  out->rank = f.rank;
  out->flag_send = f.flag_send;
  out->flag_recv = f.flag_recv;
  out->offset = f.offset;
  out->count = 6ll * (c->ngg + 1) * f.points;
  out->points = f.points;
  return OSTC_OK;
}

int ostc_parallel_buffers(ostc_ctx c, double** send, double** recv, long long* total) {
  int r = check_ready(c);
  if (r) return r;
  if (send) *send = c->d_par_send;
  if (recv) *recv = c->d_par_recv;
  if (total) *total = c->par_total;
  return OSTC_OK;
}

int ostc_pack_parallel(ostc_ctx c) {
  int r = check_ready(c);
  if (r) return r;
  for (const ParFace& f : c->par_faces) par_pack(c->blocks[f.block].d, f, c->d_par_send, c->st);
  HIPCHK(c, hipGetLastError());
  c->par_packed = true;
  return OSTC_OK;
}

int ostc_unpack_parallel(ostc_ctx c) {
  int r = check_ready(c);
  if (r) return r;
  if (!c->par_packed && !c->par_faces.empty())
    return fail(c, OSTC_ERR_STATE, "ostc_unpack_parallel: call ostc_pack_parallel first (the face nodes take the "
                                   "mean of what was sent and what was received)");
  c->par_packed = false;
  for (const ParFace& f : c->par_faces) par_unpack(c->blocks[f.block].d, f, c->d_par_recv, c->d_par_send, c->st);
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_parallel_copy(ostc_ctx c, int which, double* host) {
  int r = check_ready(c);
  if (r) return r;
  if (!host || (which != 0 && which != 1)) return fail(c, OSTC_ERR_ARG, "bad parallel buffer copy");
  const size_t bytes = (size_t)c->par_total * sizeof(double);
  if (which == 0) HIPCHK(c, hipMemcpyAsync(host, c->d_par_send, bytes, hipMemcpyDeviceToHost, c->st));
  else HIPCHK(c, hipMemcpyAsync(c->d_par_recv, host, bytes, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_set_ghost_props(ostc_ctx c, int mode) {
  if (!c) return OSTC_ERR_ARG;
  if (mode != OSTC_GHOST_PROPS_IN_OPS && mode != OSTC_GHOST_PROPS_PASS)
    return fail(c, OSTC_ERR_ARG, "unknown ghost properties mode");
  if (mode != c->ghost_props) invalidate_graph(c);
  c->ghost_props = mode;
  return OSTC_OK;
}

int ostc_last_step_timing(ostc_ctx c, double* ms, long long* launches) {
  int r = check_ready(c);
  if (r) return r;
  HIPCHK(c, hipEventSynchronize(c->ev_b));
  float f = 0;
  HIPCHK(c, hipEventElapsedTime(&f, c->ev_a, c->ev_b));
  if (ms) *ms = f;
  if (launches) *launches = c->last_launches;
  return OSTC_OK;
}

int ostc_fused_kernel_ms(ostc_ctx c, double* avg_ms, double* max_ms, long long* count) {
  int r = check_ready(c);
  if (r) return r;
  if (c->timed_launches <= 0) return fail(c, OSTC_ERR_STATE, "no OSTC_MODE_BATCHED_TIMED step has run");
  HIPCHK(c, hipEventSynchronize(c->evpool[2 * c->timed_launches - 1]));
  double sum = 0, mx = 0;
  for (long long q = 0; q < c->timed_launches; ++q) {
    float f = 0;
    HIPCHK(c, hipEventElapsedTime(&f, c->evpool[2 * q], c->evpool[2 * q + 1]));
    sum += f;
    mx = std::max(mx, (double)f);
  }
  if (avg_ms) *avg_ms = sum / (double)c->timed_launches;
  if (max_ms) *max_ms = mx;
  if (count) *count = c->timed_launches;
  return OSTC_OK;
}

int ostc_stage_rhs(ostc_ctx c, int b) {
  int r = check_block(c, b);
  if (r) return r;
  long long n = 0;
  const DevBlock& d = c->blocks[b].d;
  HIPCHK(c, hipMemsetAsync(d.dq, 0, 5 * d.n0 * sizeof(double), c->st));
  plain_rhs(d, c->P, c->st, &n);
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_batched_rhs(ostc_ctx c, int b) {
  int r = check_block(c, b);
  if (r) return r;
  if ((r = need_fused(c))) return r;
  FusedArgs a = fused_args(c, 0, 1, 0);
  a.target_block = b;
  if (fused_launch(c->tile_index, 1, a, c->n_wg, c->st) < 0)
    return fail(c, OSTC_ERR_UNSUPPORTED, "fused kernel: scheme/tile not available");
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_compute_dt(ostc_ctx c, double* dt_out) {
  int r = check_ready(c);
  if (r) return r;
  long long n = 0;
  double dt = 0;
  r = plain_dt(c, &dt, &n);
  if (r) return r;
  if (!(dt > 0) || std::isnan(dt)) return fail(c, OSTC_ERR_NUMERIC, "non-positive or NaN time step");
  HIPCHK(c, hipMemcpyAsync(c->d_dt, &dt, sizeof(double), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (dt_out) *dt_out = dt;
  return OSTC_OK;
}

int ostc_stage_update(ostc_ctx c, int b, int rk, double dt) {
  int r = check_block(c, b);
  if (r) return r;
  if (rk < 0 || rk > 2) return fail(c, OSTC_ERR_ARG, "rk must be 0, 1 or 2");
  long long n = 0;
  plain_update(c->blocks[b].d, dt, rk, c->st, &n);
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_apply_bc(ostc_ctx c, int b) {
  int r = check_block(c, b);
  if (r) return r;
  long long n = 0;
  plain_bc(c, b, &n);
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_exchange(ostc_ctx c) {
  int r = check_ready(c);
  if (r) return r;
  long long n = 0;
  plain_exchange(c, &n);
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_update_properties(ostc_ctx c, int b) {
  int r = check_block(c, b);
  if (r) return r;
  long long n = 0;
  padded_pass(c->blocks[b].d, 0, OP_PROPS, c->st, &n);
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_copy_qn(ostc_ctx c, int b) {
  int r = check_block(c, b);
  if (r) return r;
  const DevBlock& d = c->blocks[b].d;
  HIPCHK(c, hipMemcpyAsync(d.cv[1], d.cv[0], 5 * d.sz * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  return OSTC_OK;
}

int ostc_store_last_step(ostc_ctx c) {
  int r = check_ready(c);
  if (r) return r;
  for (auto& B : c->blocks) store_last(B.d, c->st);
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_residual(ostc_ctx c, double res[4]) {
  int r = check_ready(c);
  if (r) return r;
  if (!res) return OSTC_ERR_ARG;
  double acc[4] = {0, 0, 0, 0};
  long long n_total = 0;
  const int nblk = 256;
  std::vector<double> h(4 * nblk);
  for (auto& B : c->blocks) {
    residual_partial(B.d, c->d_scratch + 1, nblk, c->st);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_scratch + 1, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int q = 0; q < nblk; ++q)
      for (int l = 0; l < 4; ++l) acc[l] += h[4 * q + l];
    n_total += B.d.n0;
  }
  for (int l = 0; l < 4; ++l) res[l] = std::sqrt(acc[l] / (double)n_total);
  if (std::isnan(res[0] + res[1] + res[2] + res[3])) {
    // where: the first node of the first block whose primitives hold a NaN
    std::string at;
    auto* word = reinterpret_cast<unsigned long long*>(c->d_scratch);
    for (size_t b = 0; b < c->blocks.size() && at.empty(); ++b) {
      const DevBlock& d = c->blocks[b].d;
      at = first_nan_text(d.bv[0], d.sz, d.mx, d.my, d.mz, d.g, word, c->st);
      if (!at.empty()) at = ": block " + std::to_string(b) + ", " + at;
    }
    return fail(c, OSTC_ERR_NUMERIC, "NaN in the residual" + at);
  }
  return OSTC_OK;
}

int ostc_tile_count(int* n) {
  if (!n) return OSTC_ERR_ARG;
  *n = fused_shape_count(51);
  return OSTC_OK;
}

int ostc_tile_shape(int index, int* tx, int* ty, int* tz) {
  if (index < 0 || index >= fused_shape_count(51)) return OSTC_ERR_ARG;
  const TileShape t = fused_shape(51, index);
  if (tx) *tx = t.tx;
  if (ty) *ty = t.ty;
  if (tz) *tz = t.tz;
  return OSTC_OK;
}

int ostc_set_tile(ostc_ctx c, int index, int waves_per_cu, int split) {
  int r = check_ready(c);
  if (r) return r;
  if ((r = need_fused(c))) return r;
  if (index < 0 || index >= fused_shape_count(c->P.scheme)) return fail(c, OSTC_ERR_ARG, "tile index out of range");
  // waves_per_cu <= 0: one workgroup per tile; otherwise a persistent grid of
  // waves_per_cu workgroups per compute unit.  Either way the workgroups of
  // one XCD take consecutive tiles.
  c->xcd_remap = 1;
  c->waves = waves_per_cu > 0 ? waves_per_cu : 0;
  c->split = split ? 1 : 0;
  return set_tiles(c, index, waves_per_cu > 0 ? waves_per_cu * c->cus : 0);
}

int ostc_tile_config(ostc_ctx c, int* index, int* waves_per_cu, int* split) {
  int r = check_ready(c);
  if (r) return r;
  if (index) *index = c->tile_index;
  if (waves_per_cu) *waves_per_cu = c->waves;
  if (split) *split = c->split;
  return OSTC_OK;
}

int ostc_autotune(ostc_ctx c, int reps, int* best_index, double* best_ms) {
  int r = check_ready(c);
  if (r) return r;
  if ((r = need_fused(c))) return r;
  if (reps < 1) reps = 3;
  double best = 1e30;
  int bi = c->tile_index, bw = 0, bs = 0;
  const int waves[] = {0, 1, 2, 4};
  for (int idx = 0; idx < fused_shape_count(c->P.scheme); ++idx) {
    const int lds = fused_lds_bytes(c->P.scheme, idx);
    if (lds < 0 || lds > 160 * 1024) continue;
    for (int sp = 0; sp < (c->P.visc ? 2 : 1); ++sp)
      for (int w : waves) {
        r = ostc_set_tile(c, idx, w, sp);
        if (r) return r;
        FusedArgs a = fused_args(c, 0, 1, 0);
        const int mode = sp ? 2 : 0;
        if (fused_launch(idx, mode, a, c->n_wg, c->st) < 0) continue;  // warm-up
        HIPCHK(c, hipEventRecord(c->ev_a, c->st));
        for (int q = 0; q < reps; ++q) fused_launch(idx, mode, a, c->n_wg, c->st);
        HIPCHK(c, hipEventRecord(c->ev_b, c->st));
        HIPCHK(c, hipEventSynchronize(c->ev_b));
        HIPCHK(c, hipGetLastError());
        float f = 0;
        HIPCHK(c, hipEventElapsedTime(&f, c->ev_a, c->ev_b));
        const double ms = f / reps;
        if (ms < best) { best = ms; bi = idx; bw = w; bs = sp; }
      }
  }
  r = ostc_set_tile(c, bi, bw, bs);
  if (r) return r;
  if (best_index) *best_index = bi;
  if (best_ms) *best_ms = best;
  return OSTC_OK;
}

int ostc_sync(ostc_ctx c) {
  if (!c) return OSTC_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

const char* ostc_last_error(ostc_ctx c) { return c ? c->err.c_str() : "null context"; }

// ------------------------------------------------------------ reverse mode
}  // extern "C"

namespace {

FourierModes to_modes(const ostc_fourier& f) {
  FourierModes F{};
  F.n = f.n_modes;
  for (int m = 0; m < f.n_modes; ++m) {
    F.comp[m] = f.comp[m];
    F.kx[m] = f.kx[m];
    F.kz[m] = f.kz[m];
    F.trig[m] = f.trig[m];
    F.amp[m] = f.amp[m];
  }
  F.y0 = f.y0; F.dx = f.dx; F.dy = f.dy; F.dz = f.dz;
  F.lx = f.lx; F.lz = f.lz; F.delta = f.delta; F.yc = f.yc; F.du = f.du;
  return F;
}

int check_modes(ostc_ctx c, const ostc_fourier* f) {
  if (!f || f->n_modes < 1 || f->n_modes > OSTC_MAX_MODES) return fail(c, OSTC_ERR_ARG, "n_modes must be 1..64");
  for (int m = 0; m < f->n_modes; ++m)
    if (f->comp[m] < 0 || f->comp[m] > 2) return fail(c, OSTC_ERR_ARG, "mode component must be 0, 1 or 2");
  if (!(f->lx > 0) || !(f->lz > 0) || !(f->delta > 0)) return fail(c, OSTC_ERR_ARG, "lx, lz, delta must be positive");
  return OSTC_OK;
}

ThetaParams to_theta(const ostc_theta& t) { return ThetaParams{t.u1, t.u2, t.rho0, t.du, t.dy, t.ni, t.nk}; }

int check_theta(ostc_ctx c, const DevBlock& B, const ostc_theta* t) {
  if (!t || t->ni < 1 || t->ni > B.mx || t->nk < 1 || t->nk > B.mz || !(t->rho0 > 0) || t->du == 0)
    return fail(c, OSTC_ERR_ARG, "bad momentum-thickness parameters");
  return OSTC_OK;
}

// Buffers shared by every VJP of the context (face cotangents sized for the
// largest block) and the objective's row sums.
int adj_common(ostc_ctx c) {
  long long need = 0, gneed = 0;
  int my = 0;
  for (auto& B : c->blocks) {
    need = std::max(need, vjp_face_stride(B.d.mx, B.d.my, B.d.mz));
    gneed = std::max(gneed, vjp_gather_stride(B.d.mx, B.d.my, B.d.mz));
    my = std::max(my, B.d.my);
  }
  if (need > c->fs_stride) {
    int r = dalloc(c, &c->d_fs, 48 * (size_t)need);
    if (r) return r;
    c->fs_stride = need;
  }
  if (gneed > c->gs_stride) {
    int r = dalloc(c, &c->d_gs, 8 * (size_t)gneed);
    if (r) return r;
    c->gs_stride = gneed;
  }
  if (!c->d_obj) {
    int r = dalloc(c, &c->d_obj, 4 * (size_t)my + 1);
    if (r) return r;
  }
  return OSTC_OK;
}

VjpArgs vjp_args(ostc_ctx c, const DevBlock& B, const double* cv, const double* bv, const double* mul, const double* w,
                 double* lam) {
  VjpArgs a{};
  a.mx = B.mx; a.my = B.my; a.mz = B.mz; a.g = B.g; a.sx = B.sx; a.sy = B.sy; a.sz = B.sz; a.n0 = B.n0;
  a.cv = cv; a.bv = bv; a.mul = mul; a.jac = B.jac; a.met = B.met;
  a.w = w; a.lam = lam; a.fs = c->d_fs; a.fs_stride = c->fs_stride; a.P = c->P;
  a.gs = c->d_gs; a.gs_stride = c->gs_stride;
  return a;
}

// A saved state: planes cv (5), bv (6), mul (1) of one ring slot.
int save_state(ostc_ctx c, const DevBlock& B, int slot, double* to) {
  const size_t s = (size_t)B.sz * sizeof(double);
  HIPCHK(c, hipMemcpyAsync(to, B.cv[slot], 5 * s, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(to + 5 * B.sz, B.bv[slot], 6 * s, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(to + 11 * B.sz, B.mul[slot], s, hipMemcpyDeviceToDevice, c->st));
  return OSTC_OK;
}

int restore_state(ostc_ctx c, const DevBlock& B, const double* from) {
  const size_t s = (size_t)B.sz * sizeof(double);
  HIPCHK(c, hipMemcpyAsync(B.cv[0], from, 5 * s, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(B.bv[0], from + 5 * B.sz, 6 * s, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(B.mul[0], from + 11 * B.sz, s, hipMemcpyDeviceToDevice, c->st));
  return OSTC_OK;
}

// Transposed boundary copies of block b, in the reverse of their order.
void bc_transpose(ostc_ctx c, int b, double* lam) {
  const DevBlock& B = c->blocks[b].d;
  for (auto it = c->plain_bc[b].rbegin(); it != c->plain_bc[b].rend(); ++it) adj_bc_op(B, c->ops[*it], lam, c->st);
}

}  // namespace

extern "C" {

int ostc_rhs_vjp(ostc_ctx c, int b, const double* dq_bar, double* cv_bar) {
  int r = check_block(c, b);
  if (r) return r;
  if (c->P.scheme != 51 && c->P.scheme != 50)
    return fail(c, OSTC_ERR_UNSUPPORTED, "reverse mode covers inviscid_scheme 51 and 50");
  r = adj_common(c);
  if (!r && (!dq_bar || !cv_bar)) r = adj_fields(c, b);
  if (r) return r;
  const DevBlock& B = c->blocks[b].d;
  const double* w = dq_bar ? dq_bar : c->dq_bar[b];
  double* lam = cv_bar ? cv_bar : c->cv_bar[b];
  HIPCHK(c, hipMemsetAsync(lam, 0, 5 * (size_t)B.sz * sizeof(double), c->st));
  long long n = 0;
  rhs_vjp(vjp_args(c, B, B.cv[0], B.bv[0], B.mul[0], w, lam), c->st, &n);
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_perturb_fourier(ostc_ctx c, int b, const ostc_fourier* f) {
  int r = check_block(c, b);
  if (!r) r = check_modes(c, f);
  if (r) return r;
  const DevBlock& B = c->blocks[b].d;
  fourier_add(B, to_modes(*f), B.bv[0], c->st);
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_set_dual_time(ostc_ctx c, const ostc_dual_time* d) {
  int r = check_ready(c);
  if (r) return r;
  if (!d || !(d->dt > 0) || d->inner_iteration < 1 || d->max_inner_iteration < 1 || d->dplur_inner_step < 0 ||
      d->iteration_adjust_step < 0)
    return fail(c, OSTC_ERR_ARG, "bad dual time stepping settings");
  if (c->P.n_var != 5) return fail(c, OSTC_ERR_UNSUPPORTED, "dual time stepping covers air");
  if (c->dual_blocks.empty()) {
    for (auto& B : c->blocks) {
      const size_t sz = (size_t)B.d.sz, n0 = (size_t)B.d.n0;
      DualBlock D{};
      if ((r = dalloc(c, &D.qn1, 5 * sz)) || (r = dalloc(c, &D.qn_star, 5 * n0)) || (r = dalloc(c, &D.sr, 3 * n0)) ||
          (r = dalloc(c, &D.dq0, 5 * n0)) || (r = dalloc(c, &D.dqk, 5 * n0)) || (r = dalloc(c, &D.last, 4 * n0)))
        return r;
      c->dual_blocks.push_back(D);
    }
  }
  c->dual = *d;
  // q^(n-1) starts as the current state (src/DualTimeStepping.cuh:66-71)
  for (size_t b = 0; b < c->blocks.size(); ++b)
    HIPCHK(c, hipMemcpyAsync(c->dual_blocks[b].qn1, c->blocks[b].d.cv[0], 5 * (size_t)c->blocks[b].d.sz * sizeof(double),
                             hipMemcpyDeviceToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->dual_ready = true;
  return OSTC_OK;
}

int ostc_dual_time_step(ostc_ctx c, int n_steps, int* inner_iteration) {
  int r = check_ready(c);
  if (r) return r;
  if (!c->dual_ready) return fail(c, OSTC_ERR_STATE, "ostc_set_dual_time has not been called");
  if (n_steps < 0) return fail(c, OSTC_ERR_ARG, "negative step count");
  invalidate_graph(c);  // the batched graph's state assumptions do not survive other steps
  long long n = 0;
  HIPCHK(c, hipEventRecord(c->ev_a, c->st));
  for (int i = 0; i < n_steps; ++i)
    if ((r = dual_step(c, &n))) return r;
  HIPCHK(c, hipEventRecord(c->ev_b, c->st));
  HIPCHK(c, hipEventSynchronize(c->ev_b));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev_a, c->ev_b));
  c->last_ms = ms;
  c->last_launches = n;
  if (inner_iteration) *inner_iteration = c->dual.inner_iteration;
  return OSTC_OK;
}

int ostc_momentum_thickness(ostc_ctx c, int b, const ostc_theta* t, double* theta) {
  int r = check_block(c, b);
  if (r) return r;
  const DevBlock& B = c->blocks[b].d;
  r = check_theta(c, B, t);
  if (!r) r = adj_common(c);
  if (r) return r;
  if (!theta) return fail(c, OSTC_ERR_ARG, "null theta");
  double* obj = c->d_obj;
  theta_eval(B, to_theta(*t), B.cv[0], obj, obj + 2 * B.my, obj + 4 * B.my, c->st);
  HIPCHK(c, hipMemcpyAsync(theta, obj + 4 * B.my, sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_momentum_thickness_noalloc(ostc_ctx c, int b, const ostc_theta* t, double* theta) {
  int r = check_block(c, b);
  if (r) return r;
  const DevBlock& B = c->blocks[b].d;
  r = check_theta(c, B, t);
  if (r) return r;
  if (!theta) return fail(c, OSTC_ERR_ARG, "null theta");
  if (B.my > kThetaMaxRows)
    return fail(c, OSTC_ERR_ARG, "block " + std::to_string(b) + " has " + std::to_string(B.my) +
                                     " rows; the allocation-free momentum thickness covers at most " +
                                     std::to_string(kThetaMaxRows));
  int dev = -1;
  HIPCHK(c, hipGetDevice(&dev));
  if (dev != c->dev)
    return fail(c, OSTC_ERR_STATE, "device " + std::to_string(dev) + " is current, the context runs on device " +
                                       std::to_string(c->dev));
  double* obj = nullptr;
  const int e = theta_scratch_ptr(dev, &obj);
  if (e != (int)hipSuccess)
    return fail(c, OSTC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString((hipError_t)e));
  theta_eval(B, to_theta(*t), B.cv[0], obj, obj + 2 * B.my, obj + 4 * B.my, c->st);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(theta, obj + 4 * B.my, sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return OSTC_OK;
}

int ostc_adjoint_forward(ostc_ctx c, int n_steps, int every, const ostc_theta* t, double* theta) {
  int r = check_ready(c);
  if (r) return r;
  if (c->blocks.size() != 1 || !c->blocks[0].inner.empty())
    return fail(c, OSTC_ERR_UNSUPPORTED, "the RK3 reverse sweep covers a single block without inner faces");
  for (const FaceDesc& f : c->blocks[0].faces)
    if (f.type != OSTC_BC_OUTFLOW && f.type != OSTC_BC_PERIODIC)
      return fail(c, OSTC_ERR_UNSUPPORTED, "the RK3 reverse sweep covers outflow and periodic faces");
  if (c->sponge) return fail(c, OSTC_ERR_UNSUPPORTED, "the RK3 reverse sweep does not cover the sponge layer");
  if (!c->s.fixed_time_step) return fail(c, OSTC_ERR_UNSUPPORTED, "the RK3 reverse sweep needs fixed_time_step = 1");
  if (c->P.scheme != 51 && c->P.scheme != 50)
    return fail(c, OSTC_ERR_UNSUPPORTED, "reverse mode covers inviscid_scheme 51 and 50");
  if (n_steps < 1 || every < 1) return fail(c, OSTC_ERR_ARG, "n_steps and checkpoint_every must be positive");
  if (!theta) return fail(c, OSTC_ERR_ARG, "null theta");
  const DevBlock& B = c->blocks[0].d;
  r = check_theta(c, B, t);
  if (!r) r = adj_common(c);
  if (r) return r;
  const size_t plane = 12 * (size_t)B.sz;
  const int n_seg = (n_steps + every - 1) / every;
  const int seg_len = std::min(every, n_steps);
  if (!c->d_lam) {
    r = dalloc(c, &c->d_lam, 5 * (size_t)B.sz);
    if (!r) r = dalloc(c, &c->d_lam_qn, 5 * (size_t)B.sz);
    if (!r) r = dalloc(c, &c->d_w, 5 * (size_t)B.n0);
    if (!r) r = dalloc(c, &c->d_part, (size_t)fourier_grad_parts() * kMaxModes + kMaxModes);
    if (r) return r;
  }
  while ((int)c->ckpt.size() < n_seg) {
    double* p = nullptr;
    r = dalloc(c, &p, plane);
    if (r) return r;
    c->ckpt.push_back(p);
  }
  while ((int)c->stage.size() < 3 * seg_len) {
    double* p = nullptr;
    r = dalloc(c, &p, plane);
    if (r) return r;
    c->stage.push_back(p);
  }
  c->adj_ready = false;
  HIPCHK(c, hipEventRecord(c->ev_f0, c->st));
  const int last0 = (n_seg - 1) * every;
  long long n = 0;
  for (int s = 0; s < n_steps; ++s) {
    if (s % every == 0) save_state(c, B, 0, c->ckpt[s / every]);
    for (int rk = 0; rk < 3; ++rk) {
      if (s >= last0) save_state(c, B, rk, c->stage[3 * (s - last0) + rk]);
      r = batched_stage(c, rk, &n, -1);
      if (r) return r;
    }
  }
  double* obj = c->d_obj;
  theta_eval(B, to_theta(*t), B.cv[0], obj, obj + 2 * B.my, obj + 4 * B.my, c->st);
  HIPCHK(c, hipEventRecord(c->ev_f1, c->st));
  HIPCHK(c, hipMemcpyAsync(theta, obj + 4 * B.my, sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipGetLastError());
  c->adj_steps = n_steps;
  c->adj_every = every;
  c->adj_theta = *t;
  c->adj_ready = true;
  return OSTC_OK;
}

int ostc_adjoint_backward(ostc_ctx c, const ostc_fourier* f, double* grad) {
  int r = check_ready(c);
  if (r) return r;
  if (!c->adj_ready) return fail(c, OSTC_ERR_STATE, "ostc_adjoint_forward has not run");
  if (f) {
    r = check_modes(c, f);
    if (r) return r;
    if (!grad) return fail(c, OSTC_ERR_ARG, "null grad");
  }
  r = adj_fields(c, 0);
  if (r) return r;
  const DevBlock& B = c->blocks[0].d;
  const int n_steps = c->adj_steps, every = c->adj_every;
  const int n_seg = (n_steps + every - 1) / every;
  const double dt = c->s.dt;
  HIPCHK(c, hipEventRecord(c->ev_b0, c->st));
  theta_seed(B, to_theta(c->adj_theta), c->d_obj + 2 * B.my, c->d_lam, c->st);
  HIPCHK(c, hipMemsetAsync(c->d_lam_qn, 0, 5 * (size_t)B.sz * sizeof(double), c->st));
  long long n = 0;
  for (int seg = n_seg - 1; seg >= 0; --seg) {
    const int s0 = seg * every, s1 = std::min(n_steps, s0 + every);
    if (seg != n_seg - 1) {
      // recompute the segment from its checkpoint, keeping every stage's input
      r = restore_state(c, B, c->ckpt[seg]);
      if (r) return r;
      for (int s = s0; s < s1; ++s)
        for (int rk = 0; rk < 3; ++rk) {
          save_state(c, B, rk, c->stage[3 * (s - s0) + rk]);
          r = batched_stage(c, rk, &n, -1);
          if (r) return r;
        }
    }
    for (int s = s1 - 1; s >= s0; --s)
      for (int rk = 2; rk >= 0; --rk) {
        const double* S = c->stage[3 * (s - s0) + rk];
        bc_transpose(c, 0, c->d_lam);
        const RkCoef k = rk_coef(rk);
        adj_stage_pre(B, c->d_lam, c->d_lam_qn, c->d_w, k.a, k.b, k.c * dt, c->st);
        rhs_vjp(vjp_args(c, B, S, S + 5 * B.sz, S + 11 * B.sz, c->d_w, c->d_lam), c->st, &n);
        if (rk == 0) adj_stage_qn(B, c->d_lam, c->d_lam_qn, c->st);
      }
  }
  // the initial state: cv_from_bv after the boundary copies of ostc_initialize
  bc_transpose(c, 0, c->d_lam);
  HIPCHK(c, hipMemcpyAsync(c->cv_bar[0], c->d_lam, 5 * (size_t)B.sz * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  double* d_grad = c->d_part + (size_t)fourier_grad_parts() * kMaxModes;
  if (f) fourier_grad(B, to_modes(*f), c->d_lam, c->ckpt[0], c->d_part, d_grad, c->st);
  HIPCHK(c, hipEventRecord(c->ev_b1, c->st));
  if (f) HIPCHK(c, hipMemcpyAsync(grad, d_grad, f->n_modes * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipGetLastError());
  return OSTC_OK;
}

int ostc_adjoint_timing(ostc_ctx c, double* forward_ms, double* backward_ms) {
  int r = check_ready(c);
  if (r) return r;
  if (!c->adj_ready) return fail(c, OSTC_ERR_STATE, "ostc_adjoint_forward has not run");
  float f = 0, b = 0;
  HIPCHK(c, hipEventElapsedTime(&f, c->ev_f0, c->ev_f1));
  if (backward_ms) {
    if (hipEventElapsedTime(&b, c->ev_b0, c->ev_b1) != hipSuccess) {
      (void)hipGetLastError();
      b = -1;  // no backward since the last forward
    }
    *backward_ms = b;
  }
  if (forward_ms) *forward_ms = f;
  return OSTC_OK;
}

// ----------------------------------------------------------- setup helpers
// Block::compute_jac_metric (src/Mesh.cpp:241-300) on an analytic uniform
// grid x = x0 + i dx (ghost nodes included), SoA output on the padded domain.
int ostc_util_uniform_geometry(int mx, int my, int mz, int g, double x0, double y0, double z0, double dx, double dy,
                               double dz, double* jac, double* metric) {
  if (mx < 2 || my < 2 || mz < 2 || g < 1 || !jac || !metric) return OSTC_ERR_ARG;
  const long long sx = mx + 2 * g, sy = my + 2 * g, sz = sx * sy * (mz + 2 * g);
  for (int k = -g; k < mz + g; ++k)
    for (int j = -g; j < my + g; ++j)
      for (int i = -g; i < mx + g; ++i) {
        auto X = [&](int a) { return x0 + a * dx; };
        auto Y = [&](int a) { return y0 + a * dy; };
        auto Z = [&](int a) { return z0 + a * dz; };
        const double dxd1 = (X(i + 1) - X(i - 1)) * 0.5, dyd1 = (Y(j) - Y(j)) * 0.5, dzd1 = (Z(k) - Z(k)) * 0.5;
        const double dxd2 = (X(i) - X(i)) * 0.5, dyd2 = (Y(j + 1) - Y(j - 1)) * 0.5, dzd2 = (Z(k) - Z(k)) * 0.5;
        const double dxd3 = (X(i) - X(i)) * 0.5, dyd3 = (Y(j) - Y(j)) * 0.5, dzd3 = (Z(k + 1) - Z(k - 1)) * 0.5;
        const double J = dxd1 * dyd2 * dzd3 + dxd2 * dyd3 * dzd1 + dxd3 * dyd1 * dzd2 - dxd1 * dyd3 * dzd2 -
                         dxd2 * dyd1 * dzd3 - dxd3 * dyd2 * dzd1;
        if (!(J > 0)) return OSTC_ERR_NUMERIC;
        const long long p = ((long long)(k + g) * sy + (j + g)) * sx + (i + g);
        jac[p] = J;
        metric[p + 0 * sz] = (dyd2 * dzd3 - dyd3 * dzd2) / J;
        metric[p + 1 * sz] = (dxd3 * dzd2 - dxd2 * dzd3) / J;
        metric[p + 2 * sz] = (dxd2 * dyd3 - dxd3 * dyd2) / J;
        metric[p + 3 * sz] = (dyd3 * dzd1 - dyd1 * dzd3) / J;
        metric[p + 4 * sz] = (dxd1 * dzd3 - dxd3 * dzd1) / J;
        metric[p + 5 * sz] = (dxd3 * dyd1 - dxd1 * dyd3) / J;
        metric[p + 6 * sz] = (dyd1 * dzd2 - dyd2 * dzd1) / J;
        metric[p + 7 * sz] = (dxd2 * dzd1 - dxd1 * dzd2) / J;
        metric[p + 8 * sz] = (dxd1 * dyd2 - dxd2 * dyd1) / J;
      }
  return OSTC_OK;
}

// Block::compute_jac_metric (src/Mesh.cpp:241-502) from node coordinates:
// x, y, z on the domain padded by g+1 ghost layers (the half-central
// differences of the outermost ghost layer read one layer beyond it), jac and
// metric SoA on the domain padded by g.
int ostc_util_metric(int mx, int my, int mz, int g, const double* x, const double* y, const double* z, double* jac,
                     double* metric) {
  if (mx < 2 || my < 2 || mz < 2 || g < 1 || !x || !y || !z || !jac || !metric) return OSTC_ERR_ARG;
  const long long sx = mx + 2 * g, sy = my + 2 * g, sz = sx * sy * (mz + 2 * g);
  const int G = g + 1;
  const long long cx = mx + 2 * G, cy = my + 2 * G;
  auto at = [&](const double* a, int i, int j, int k) { return a[((long long)(k + G) * cy + (j + G)) * cx + (i + G)]; };
  for (int k = -g; k < mz + g; ++k)
    for (int j = -g; j < my + g; ++j)
      for (int i = -g; i < mx + g; ++i) {
        const double dxd1 = (at(x, i + 1, j, k) - at(x, i - 1, j, k)) * 0.5,
                     dyd1 = (at(y, i + 1, j, k) - at(y, i - 1, j, k)) * 0.5,
                     dzd1 = (at(z, i + 1, j, k) - at(z, i - 1, j, k)) * 0.5;
        const double dxd2 = (at(x, i, j + 1, k) - at(x, i, j - 1, k)) * 0.5,
                     dyd2 = (at(y, i, j + 1, k) - at(y, i, j - 1, k)) * 0.5,
                     dzd2 = (at(z, i, j + 1, k) - at(z, i, j - 1, k)) * 0.5;
        const double dxd3 = (at(x, i, j, k + 1) - at(x, i, j, k - 1)) * 0.5,
                     dyd3 = (at(y, i, j, k + 1) - at(y, i, j, k - 1)) * 0.5,
                     dzd3 = (at(z, i, j, k + 1) - at(z, i, j, k - 1)) * 0.5;
        const double J = dxd1 * dyd2 * dzd3 + dxd2 * dyd3 * dzd1 + dxd3 * dyd1 * dzd2 - dxd1 * dyd3 * dzd2 -
                         dxd2 * dyd1 * dzd3 - dxd3 * dyd2 * dzd1;
        if (!(J > 0)) return OSTC_ERR_NUMERIC;
        const long long p = ((long long)(k + g) * sy + (j + g)) * sx + (i + g);
        jac[p] = J;
        metric[p + 0 * sz] = (dyd2 * dzd3 - dyd3 * dzd2) / J;
        metric[p + 1 * sz] = (dxd3 * dzd2 - dxd2 * dzd3) / J;
        metric[p + 2 * sz] = (dxd2 * dyd3 - dxd3 * dyd2) / J;
        metric[p + 3 * sz] = (dyd3 * dzd1 - dyd1 * dzd3) / J;
        metric[p + 4 * sz] = (dxd1 * dzd3 - dxd3 * dzd1) / J;
        metric[p + 5 * sz] = (dxd3 * dyd1 - dxd1 * dyd3) / J;
        metric[p + 6 * sz] = (dyd1 * dzd2 - dyd2 * dzd1) / J;
        metric[p + 7 * sz] = (dxd2 * dzd1 - dxd1 * dzd2) / J;
        metric[p + 8 * sz] = (dxd1 * dyd2 - dxd2 * dyd1) / J;
      }
  return OSTC_OK;
}

static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// initialize_mixing_layer_with_info for air (src/Initialize.cu:259-325) on the
// padded domain, y = y0 + j dy, plus the seeded velocity perturbation.
int ostc_util_mixing_layer_bv(int mx, int my, int mz, int g, double y0, double dy, const double upper[6],
                              const double lower[6], double delta_omega, int gi0, int gj0, int gk0, int period_i,
                              int period_j, int period_k, uint64_t seed, double amplitude, double* bv) {
  if (mx < 2 || my < 2 || mz < 2 || g < 1 || !upper || !lower || !bv) return OSTC_ERR_ARG;
  const long long sx = mx + 2 * g, sy = my + 2 * g, sz = sx * sy * (mz + 2 * g);
  auto wrap = [](long long v, int per) -> long long {
    if (per <= 0) return v;
    long long r = v % per;
    return r < 0 ? r + per : r;
  };
  for (int k = -g; k < mz + g; ++k)
    for (int j = -g; j < my + g; ++j)
      for (int i = -g; i < mx + g; ++i) {
        const long long p = ((long long)(k + g) * sy + (j + g)) * sx + (i + g);
        const double y = y0 + j * dy;
        bv[p + sz] = 0.5 * (upper[1] + lower[1]) + 0.5 * (upper[1] - lower[1]) * std::tanh(2 * y / delta_omega);
        const double* var = y >= 0 ? upper : lower;
        bv[p] = var[0];
        bv[p + 2 * sz] = var[2];
        bv[p + 3 * sz] = var[3];
        bv[p + 4 * sz] = var[4];
        bv[p + 5 * sz] = var[5];
        if (amplitude != 0) {
          const long long gi = wrap((long long)gi0 + i, period_i), gj = wrap((long long)gj0 + j, period_j),
                          gk = wrap((long long)gk0 + k, period_k);
          const uint64_t node = (uint64_t)((gk * 1000003LL + gj) * 1000033LL + gi);
          for (int comp = 0; comp < 3; ++comp) {
            const uint64_t h = splitmix64(seed ^ splitmix64(node * 4 + (uint64_t)comp));
            const double r = (double)(h >> 11) * (1.0 / 9007199254740992.0);
            bv[p + (1 + comp) * sz] += amplitude * (2.0 * r - 1.0);
          }
        }
      }
  return OSTC_OK;
}

}  // extern "C"
