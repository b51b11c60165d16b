// ostc_ops.hpp -- ghost-layer work items (boundary conditions, inner-face
// exchange, properties, cv-from-bv) as one list of boxes, executed either one
// op per launch (the reference's per-face launches) or many independent ops
// per launch (the batched path).
#pragma once
#include "ostc_device.hpp"

namespace ostc {

enum OpType : int {
  OP_BC_OUTFLOW = 0,   // apply_outflow,  src/BoundCond.cuh:205-233
  OP_BC_PERIODIC = 1,  // apply_periodic, src/BoundCond.cuh:1443-1486
  OP_COMM = 2,         // inner_communication, src/DataCommunication.cuh:60-105
  OP_PROPS = 3,        // update_physical_properties<Air>, src/FieldOperation.cuh:107-140
  OP_CV_FROM_BV = 4,   // compute_cv_from_bv, src/FieldOperation.cuh:40-76
  OP_COPY_BV = 5,      // bv of slot `src` -> slot `slot` (ghost rims read by boundary ops)
};

struct DevOp {
  int type;
  int block, target;
  int axis, side;
  int lo[3];           // box origin (block indices, ghosts negative)
  int n[3];            // box extents
  int shared;          // OP_COMM: index along axis of the shared node layer
  int tstart;          // OP_COMM: target index matching lo[axis]
  int flags;           // OP_COMM bit 0: also recompute the target's cv on the shared layer
  long long begin;     // first point of this op in its launch range
  long long count;     // n[0]*n[1]*n[2]
};

// One op point.  Every op reads and writes the ring slot `slot` of cv/bv/mul.
__device__ __forceinline__ void op_point(const DevBlock* blocks, const DevOp& op, long long local, int slot,
                                         int src) {
  const int i = op.lo[0] + (int)(local % op.n[0]);
  const int j = op.lo[1] + (int)((local / op.n[0]) % op.n[1]);
  const int k = op.lo[2] + (int)(local / ((long long)op.n[0] * op.n[1]));
  const DevBlock& B = blocks[op.block];
  double* bv = B.bv[slot];
  double* cv = B.cv[slot];
  const long long s = B.sz;
  auto cv_point = [&](const DevBlock& Z, long long p) {
    double* zbv = Z.bv[slot];
    double* zcv = Z.cv[slot];
    const long long zs = Z.sz;
    const double rho = zbv[p];
    zcv[p] = rho;
    zcv[p + zs] = rho * zbv[p + zs];
    zcv[p + 2 * zs] = rho * zbv[p + 2 * zs];
    zcv[p + 3 * zs] = rho * zbv[p + 3 * zs];
    zcv[p + 4 * zs] = total_energy(rho, zbv[p + zs], zbv[p + 2 * zs], zbv[p + 3 * zs], zbv[p + 4 * zs]);
  };
  switch (op.type) {
    case OP_BC_OUTFLOW:
    case OP_BC_PERIODIC: {
      int d[3] = {0, 0, 0};
      d[op.axis] = op.side;
      int so[3] = {i, j, k};
      if (op.type == OP_BC_PERIODIC) {
        const int m = op.axis == 0 ? B.mx : (op.axis == 1 ? B.my : B.mz);
        so[op.axis] = op.side < 0 ? m - 1 : 0;
      }
      for (int g = 1; g <= B.g; ++g) {
        const long long pg = pidx(B, i + g * d[0], j + g * d[1], k + g * d[2]);
        const long long ps = (op.type == OP_BC_PERIODIC)
                                 ? pidx(B, so[0] + g * d[0], so[1] + g * d[1], so[2] + g * d[2])
                                 : pidx(B, i, j, k);
        for (int l = 0; l < 6; ++l) bv[pg + l * s] = bv[ps + l * s];
        cv_point(B, pg);
      }
      break;
    }
    case OP_COMM: {
      const DevBlock& T = blocks[op.target];
      int t[3] = {i, j, k};
      const int idx_a = op.axis == 0 ? i : (op.axis == 1 ? j : k);
      t[op.axis] = op.tstart + (idx_a - op.lo[op.axis]);
      const long long p = pidx(B, i, j, k), pt = pidx(T, t[0], t[1], t[2]);
      double* tbv = T.bv[slot];
      const long long ts = T.sz;
      if (idx_a == op.shared) {
        for (int l = 0; l < 6; ++l) {
          const double ave = 0.5 * (tbv[pt + l * ts] + bv[p + l * s]);
          bv[p + l * s] = ave;
          tbv[pt + l * ts] = ave;
        }
        if (op.flags & 1) cv_point(T, pt);
      } else {
        for (int l = 0; l < 6; ++l) bv[p + l * s] = tbv[pt + l * ts];
      }
      cv_point(B, p);
      break;
    }
    case OP_PROPS: {
      const long long p = pidx(B, i, j, k);
      const double t = bv[p + 5 * s];
      const double u = bv[p + s], v = bv[p + 2 * s], w = bv[p + 3 * s];
      const double V = sqrt(u * u + v * v + w * w);
      constexpr double c_temp{kGamma * kRu / kMwAir};
      B.mul[slot][p] = sutherland(t);
      B.mach[p] = V / sqrt(c_temp * t);
      break;
    }
    case OP_CV_FROM_BV: {
      cv_point(B, pidx(B, i, j, k));
      break;
    }
    case OP_COPY_BV: {
      const long long p = pidx(B, i, j, k);
      const double* sbv = B.bv[src];
      for (int l = 0; l < 6; ++l) bv[p + l * s] = sbv[p + l * s];
      break;
    }
    default:
      break;
  }
  (void)cv;
}

}  // namespace ostc
