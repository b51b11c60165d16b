// ostc_ops.hpp -- ghost-layer work items (boundary conditions, inner-face
// exchange, properties, cv-from-bv) as one list of boxes, executed either one
// op per launch (the reference's per-face launches) or many independent ops
// per launch (the batched path).
#pragma once
#include "ostc_device.hpp"
#include "ostc_inflow.hpp"

namespace ostc {

enum OpType : int {
  OP_BC_OUTFLOW = 0,   // apply_outflow,  src/BoundCond.cuh:205-233
  OP_BC_PERIODIC = 1,  // apply_periodic, src/BoundCond.cuh:1443-1486
  OP_COMM = 2,         // inner_communication, src/DataCommunication.cuh:60-105
  OP_PROPS = 3,        // update_physical_properties<Air>, src/FieldOperation.cuh:107-140
  OP_CV_FROM_BV = 4,   // compute_cv_from_bv, src/FieldOperation.cuh:40-76
  OP_COPY_BV = 5,      // bv of slot `src` -> slot `slot` (ghost rims read by boundary ops)
  OP_BC_FARFIELD = 6,  // apply_farfield<Air>, src/BoundCond.cuh:755-1022
  OP_BC_INFLOW = 7,    // apply_inflow<Air>, src/BoundCond.cuh:235-612
};

// Doubles of one far-field state on the device, the layout of ostc_farfield:
// density, u, v, w, pressure, temperature, acoustic_speed, entropy,
// specific_heat_ratio, mw, sv[16].
constexpr int kFarfieldDoubles = 26;

struct DevOp {
  int type;
  int block, target;
  int axis, side;
  int lo[3];           // box origin (block indices, ghosts negative)
  int n[3];            // box extents
  int shared;          // OP_COMM: index along axis of the shared node layer
  int tcode[3];        // OP_COMM: target index along d = toff[d] + sgn * (index along p), tcode[d] = sgn * (p + 1)
  int toff[3];
  int flags;           // OP_COMM bit 0: also recompute the target's cv on the shared layer
  long long begin;     // first point of this op in its launch range
  long long count;     // n[0]*n[1]*n[2]
  const double* ff;    // OP_BC_FARFIELD: the state of the face's label (kFarfieldDoubles);
                       // OP_BC_INFLOW: the inflow state (kInflowDoubles)
  const double* aux;   // OP_BC_INFLOW: y of every op point (the face layer), in op point order
  const unsigned long long* pass;  // OP_BC_INFLOW: boundary-condition passes per block so far
};

// compute_cv_from_bv_1_point<Air> (src/FieldOperation.cuh:40-76) in ring slot `slot`.
__device__ __forceinline__ void air_cv_point(const DevBlock& Z, long long p, int slot) {
  const double* zbv = Z.bv[slot];
  double* zcv = Z.cv[slot];
  const long long zs = Z.sz;
  const double rho = zbv[p];
  zcv[p] = rho;
  zcv[p + zs] = rho * zbv[p + zs];
  zcv[p + 2 * zs] = rho * zbv[p + 2 * zs];
  zcv[p + 3 * zs] = rho * zbv[p + 3 * zs];
  zcv[p + 4 * zs] = total_energy(rho, zbv[p + zs], zbv[p + 2 * zs], zbv[p + 3 * zs], zbv[p + 4 * zs]);
}

// update_physical_properties<Air> at one point (src/FieldOperation.cuh:107-140):
// Sutherland viscosity of ring slot `slot` and the Mach number.
__device__ __forceinline__ void props_point(const DevBlock& Z, long long p, int slot) {
  const double* zbv = Z.bv[slot];
  const long long zs = Z.sz;
  const double t = zbv[p + 5 * zs];
  const double u = zbv[p + zs], v = zbv[p + 2 * zs], w = zbv[p + 3 * zs];
  const double V = sqrt(u * u + v * v + w * w);
  constexpr double c_temp{kGamma * kRu / kMwAir};
  Z.mul[slot][p] = sutherland(t);
  Z.mach[p] = V / sqrt(c_temp * t);
}

// Writes the primitives q[6] of node p in ring slot `slot` and, from the
// same registers, its cv (compute_cv_from_bv_1_point<Air>) and, with
// `props`, its viscosity and Mach number: the values air_cv_point and
// props_point would compute after reading q back.  An op loads everything it
// reads before its first store, so a point costs one memory round trip
// rather than one per store the next load might alias.
__device__ __forceinline__ void put_node(const DevBlock& Z, long long p, int slot, const double q[6], bool with_cv,
                                         bool props) {
  double* zbv = Z.bv[slot];
  const long long zs = Z.sz;
#pragma unroll
  for (int l = 0; l < 6; ++l) zbv[p + l * zs] = q[l];
  if (with_cv) {
    double* zcv = Z.cv[slot];
    const double rho = q[0];
    zcv[p] = rho;
    zcv[p + zs] = rho * q[1];
    zcv[p + 2 * zs] = rho * q[2];
    zcv[p + 3 * zs] = rho * q[3];
    zcv[p + 4 * zs] = total_energy(rho, q[1], q[2], q[3], q[4]);
  }
  if (props) {
    const double t = q[5];
    const double u = q[1], v = q[2], w = q[3];
    const double V = sqrt(u * u + v * v + w * w);
    constexpr double c_temp{kGamma * kRu / kMwAir};
    Z.mul[slot][p] = sutherland(t);
    Z.mach[p] = V / sqrt(c_temp * t);
  }
}

// Largest ngg (WENO-Z7); the ghost loops below unroll to it.
constexpr int kMaxGhost = 4;

// One op point.  Every op reads and writes the ring slot `slot` of cv/bv/mul.
// With `props`, each op also recomputes the properties (props_point) of every
// node whose primitives it writes, and the rim copy carries the viscosity of
// slot `src` along: then no separate properties pass over the ghost layers is
// needed (a node no op writes keeps the primitives and properties every slot
// was initialised with).
__device__ __forceinline__ void op_point(const DevBlock* blocks, const DevOp& op, long long local, int slot,
                                         int src, bool props = false) {
  const int i = op.lo[0] + (int)(local % op.n[0]);
  const int j = op.lo[1] + (int)((local / op.n[0]) % op.n[1]);
  const int k = op.lo[2] + (int)(local / ((long long)op.n[0] * op.n[1]));
  const DevBlock& B = blocks[op.block];
  // unit step along the op's axis (selects, not a private array indexed at
  // run time, which would live in scratch memory)
  const int ax = op.axis;
  const int di = ax == 0 ? op.side : 0, dj = ax == 1 ? op.side : 0, dk = ax == 2 ? op.side : 0;
  double* bv = B.bv[slot];
  double* cv = B.cv[slot];
  const long long s = B.sz;
  auto cv_point = [&](const DevBlock& Z, long long p) { air_cv_point(Z, p, slot); };
  switch (op.type) {
    case OP_BC_OUTFLOW:
    case OP_BC_PERIODIC: {
      int si = i, sj = j, sk = k;
      if (op.type == OP_BC_PERIODIC) {
        const int m = ax == 0 ? B.mx : (ax == 1 ? B.my : B.mz);
        const int src = op.side < 0 ? m - 1 : 0;
        si = ax == 0 ? src : i;
        sj = ax == 1 ? src : j;
        sk = ax == 2 ? src : k;
      }
      double q[kMaxGhost][6];
#pragma unroll
      for (int g = 1; g <= kMaxGhost; ++g) {
        if (g > B.g) break;
        const long long ps = (op.type == OP_BC_PERIODIC) ? pidx(B, si + g * di, sj + g * dj, sk + g * dk)
                                                         : pidx(B, i, j, k);
#pragma unroll
        for (int l = 0; l < 6; ++l) q[g - 1][l] = bv[ps + l * s];
      }
#pragma unroll
      for (int g = 1; g <= kMaxGhost; ++g) {
        if (g > B.g) break;
        put_node(B, pidx(B, i + g * di, j + g * dj, k + g * dk), slot, q[g - 1], true, props);
      }
      break;
    }
    case OP_BC_INFLOW: {
      double st[6];
      inflow_state(op.ff, op.aux[local], op.pass[op.block], inflow_point(ax, i, j, k, B.mx, B.my, B.g), st);
      for (int g = 0; g <= B.g; ++g) put_node(B, pidx(B, i + g * di, j + g * dj, k + g * dk), slot, st, true, props);
      break;
    }
    case OP_BC_FARFIELD: {
      // Each face point (and ngg beyond on the tangential directions) decides
      // from its own normal Mach number: supersonic inflow takes the far field
      // on the face point and its ghosts, supersonic outflow copies the face
      // point to the ghosts, subsonic faces take the state from the outgoing
      // Riemann invariant of the face point and the incoming one of the far
      // field, with the far field's entropy and tangential velocity on
      // inflow and the face point's on outflow.
      const double* F = op.ff;
      const long long p = pidx(B, i, j, k);
      double nx = B.met[p + (3 * op.axis) * s], ny = B.met[p + (3 * op.axis + 1) * s],
             nz = B.met[p + (3 * op.axis + 2) * s];
      const double grad_n_inv = op.side / sqrt(nx * nx + ny * ny + nz * nz);
      nx *= grad_n_inv;
      ny *= grad_n_inv;
      nz *= grad_n_inv;
      const double u_b = bv[p + s], v_b = bv[p + 2 * s], w_b = bv[p + 3 * s];
      const double u_face = nx * u_b + ny * v_b + nz * w_b;
      const double p_b = bv[p + 4 * s], rho_b = bv[p];
      const double a_b = sqrt(kGamma * p_b / rho_b);
      const double mach_b = u_face / a_b;
      double st[6];
      bool face_too = true;
      if (mach_b <= -1) {
        for (int l = 0; l < 6; ++l) st[l] = F[l];
      } else if (mach_b >= 1) {
        for (int l = 0; l < 6; ++l) st[l] = bv[p + l * s];
        face_too = false;
      } else {
        const double riemann_pos = u_face + 2 * a_b / (kGamma - 1);
        const double u_inf = F[1] * nx + F[2] * ny + F[3] * nz;
        const double riemann_neg = u_inf - 2 * F[6] / (F[8] - 1);
        const double Un = 0.5 * (riemann_pos + riemann_neg);
        const double c_b = 0.25 * (kGamma - 1) * (riemann_pos - riemann_neg);
        double s_b;
        if (mach_b <= 0) {
          s_b = F[7];
          st[1] = F[1] + (Un - u_inf) * nx;
          st[2] = F[2] + (Un - u_inf) * ny;
          st[3] = F[3] + (Un - u_inf) * nz;
        } else {
          s_b = p_b / pow(rho_b, kGamma);
          st[1] = u_b + (Un - u_face) * nx;
          st[2] = v_b + (Un - u_face) * ny;
          st[3] = w_b + (Un - u_face) * nz;
        }
        st[0] = pow(c_b * c_b / (kGamma * s_b), 1 / (kGamma - 1));
        st[4] = st[0] * c_b * c_b / kGamma;
        st[5] = st[4] * kMwAir / (st[0] * kRu);
      }
      if (face_too) put_node(B, p, slot, st, true, props);
      for (int g = 1; g <= B.g; ++g) put_node(B, pidx(B, i + g * di, j + g * dj, k + g * dk), slot, st, true, props);
      break;
    }
    case OP_COMM: {
      const DevBlock& T = blocks[op.target];
      const int idx_a = ax == 0 ? i : (ax == 1 ? j : k);
      auto src_index = [&](int code) {
        const int q = (code < 0 ? -code : code) - 1;
        const int v = q == 0 ? i : (q == 1 ? j : k);
        return code < 0 ? -v : v;
      };
      const long long p = pidx(B, i, j, k);
      const long long pt = pidx(T, op.toff[0] + src_index(op.tcode[0]), op.toff[1] + src_index(op.tcode[1]),
                                op.toff[2] + src_index(op.tcode[2]));
      const double* tbv = T.bv[slot];
      const long long ts = T.sz;
      double q[6];
#pragma unroll
      for (int l = 0; l < 6; ++l) q[l] = tbv[pt + l * ts];
      if (idx_a == op.shared) {
        double own[6];
#pragma unroll
        for (int l = 0; l < 6; ++l) own[l] = bv[p + l * s];
#pragma unroll
        for (int l = 0; l < 6; ++l) q[l] = 0.5 * (q[l] + own[l]);
        put_node(T, pt, slot, q, (op.flags & 1) != 0, props);
      }
      put_node(B, p, slot, q, true, props);
      break;
    }
    case OP_PROPS: {
      props_point(B, pidx(B, i, j, k), slot);
      break;
    }
    case OP_CV_FROM_BV: {
      cv_point(B, pidx(B, i, j, k));
      break;
    }
    case OP_COPY_BV: {
      const long long p = pidx(B, i, j, k);
      const double* sbv = B.bv[src];
      double q[6];
#pragma unroll
      for (int l = 0; l < 6; ++l) q[l] = sbv[p + l * s];
      const double mu = props ? B.mul[src][p] : 0.0;
#pragma unroll
      for (int l = 0; l < 6; ++l) bv[p + l * s] = q[l];
      if (props) B.mul[slot][p] = mu;
      break;
    }
    default:
      break;
  }
  (void)cv;
}

}  // namespace ostc
