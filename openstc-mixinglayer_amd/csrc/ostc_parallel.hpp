// ostc_parallel.hpp -- faces a block shares with a block of another rank
// (parallel N.txt): decoding of one record on the host, and the node of a
// face point for the pack / unpack kernels.  The transfer itself (MPI in the
// reference, RCCL or a peer copy here) is the caller's; the library fills
// and drains one send and one receive buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>

namespace ostc {

// One parallel face after ParallelFace::register_boundary (src/Mesh.cpp:177-231).
// Buffer layout (setup_data_to_be_sent, src/DataCommunication.cu:3-40): face
// point t = n[order1] * npt[order2] + n[order2] holds (ngg + 1) layers of nv
// doubles, layer 0 the face node, layer g the node g steps inside the block.
struct ParFace {
  int block;
  int start[3];           // first node, 0-based
  int step[3];            // +1 / -1 along the tangential axes, 0 along the normal
  int npt[3];             // nodes along each axis (1 along the normal)
  int order1, order2;     // the tangential axes (order2 varies fastest)
  int normal, direction;  // normal axis; -1 for the low face, +1 for the high one
  int rank, flag_send, flag_recv;
  long long offset;       // first double of this face in the send / receive buffers
  long long points;       // npt[order1] * npt[order2]
  int own_from_send;      // 1: the unpack takes the face nodes' own values from the send buffer
};

// Decodes a record as read from parallel N.txt (1-based node ranges; the
// normal axis has equal ends, 1 meaning the low face; the tangential axis
// that varies fastest in the buffer is given negated) for a block of m[3]
// nodes.  Checks what register_boundary assumes and the reference leaves
// unchecked: one normal axis on a block face, one negated tangential axis,
// ranges inside the block.
inline bool parallel_face_decode(const int rec[6], const int m[3], ParFace* f, std::string* err) {
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = rec[2 * d];
    hi[d] = rec[2 * d + 1];
  }
  int normal = -1;
  for (int d = 0; d < 3 && normal < 0; ++d)
    if (lo[d] == hi[d]) normal = d;
  if (normal < 0) {
    *err = "parallel face: no axis with equal ends (the face normal)";
    return false;
  }
  const int at = lo[normal];
  if (at != 1 && at != m[normal]) {
    *err = "parallel face: the normal index is not on a block face";
    return false;
  }
  f->normal = normal;
  f->direction = at == 1 ? -1 : 1;
  f->start[normal] = at - 1;
  f->step[normal] = 0;
  f->npt[normal] = 1;
  int neg = -1;
  for (int d = 0; d < 3 && neg < 0; ++d)
    if (d != normal && lo[d] < 0) neg = d;
  const int other = neg < 0 ? -1 : 3 - normal - neg;
  if (neg < 0 || lo[other] < 0 || hi[other] < 0 || hi[neg] > 0) {
    *err = "parallel face: exactly one tangential axis must be given negated";
    return false;
  }
  const int a[2] = {other, neg};
  for (int q = 0; q < 2; ++q) {
    const int d = a[q];
    const int s0 = (q == 1 ? -lo[d] : lo[d]) - 1, s1 = (q == 1 ? -hi[d] : hi[d]) - 1;
    if (s0 < 0 || s1 < 0 || s0 >= m[d] || s1 >= m[d]) {
      *err = "parallel face: tangential range outside the block";
      return false;
    }
    f->start[d] = s0;
    f->step[d] = s0 < s1 ? 1 : -1;
    f->npt[d] = std::abs(s1 - s0) + 1;
  }
  f->order1 = other;
  f->order2 = neg;
  f->points = (long long)f->npt[other] * f->npt[neg];
  return true;
}

// Whether the face nodes of a and b (same block) have a node in common.
inline bool par_faces_touch(const ParFace& a, const ParFace& b) {
  for (int d = 0; d < 3; ++d) {
    const int a0 = a.start[d], a1 = a.start[d] + (a.npt[d] - 1) * a.step[d];
    const int b0 = b.start[d], b1 = b.start[d] + (b.npt[d] - 1) * b.step[d];
    const int alo = a0 < a1 ? a0 : a1, ahi = a0 < a1 ? a1 : a0;
    const int blo = b0 < b1 ? b0 : b1, bhi = b0 < b1 ? b1 : b0;
    if (ahi < blo || bhi < alo) return false;
  }
  return true;
}

// Where each face's unpack reads the face nodes' own values.  Unpacks run in
// face order after every face has been packed (parallel_communication,
// src/DataCommunication.cuh:108-199), so until an earlier face of the same
// block rewrites a node the field still holds what the pack sent: those faces
// read the send buffer (in order, a fraction of the cost of one double per
// field line); a face that shares nodes with an earlier one reads the field.
inline void par_assign_own_source(ParFace* faces, int n) {
  for (int i = 0; i < n; ++i) {
    faces[i].own_from_send = 1;
    for (int j = 0; j < i; ++j)
      if (faces[j].block == faces[i].block && par_faces_touch(faces[j], faces[i])) faces[i].own_from_send = 0;
  }
}

// Face node (layer 0) of point t.
__host__ __device__ inline void par_face_node(const ParFace& f, long long t, int idx[3]) {
  const int n1 = (int)(t / f.npt[f.order2]), n2 = (int)(t % f.npt[f.order2]);
  idx[f.normal] = f.start[f.normal];
  idx[f.order1] = f.start[f.order1] + n1 * f.step[f.order1];
  idx[f.order2] = f.start[f.order2] + n2 * f.step[f.order2];
}

// Pack / unpack kernels: one workgroup of kParThreads per kParChunk
// consecutive face points.  The chunk's slice of the buffer is contiguous
// (chunk x (ngg + 1) x nv doubles) and is written or read in order; the field
// side goes through LDS and is walked along the axis that is contiguous in
// memory where the face has one: the layers when the face is normal to x
// (ngg + 1 neighbouring doubles per point and variable), the face points when
// x is the tangential axis that varies fastest.  Nodes use the padded
// index ((k + g) * sy + (j + g)) * sx + (i + g) of both contexts.
constexpr int kParThreads = 256;
constexpr int kParChunk = 32;
constexpr int kParMaxLayers = 5;  // ngg + 1 for ngg <= 4

// Workgroups of a pack / unpack launch.
inline unsigned par_chunks(const ParFace& f) { return (unsigned)((f.points + kParChunk - 1) / kParChunk); }

__device__ __forceinline__ bool par_points_fastest(const ParFace& f) { return f.order2 == 0; }

// Node (pt, layer) and variable v of element e of a chunk of np points.
__device__ __forceinline__ void par_walk(const ParFace& f, int g1, int np, int e, int* pt, int* layer, int* v) {
  if (par_points_fastest(f)) {
    *pt = e % np;
    const int r = e / np;
    *layer = r % g1;
    *v = r / g1;
  } else {
    *layer = e % g1;
    const int r = e / g1;
    *pt = r % np;
    *v = r / np;
  }
}

// First point, point count and face-node indices (into base[]) of this
// workgroup's chunk; the caller synchronises before reading base[].
__device__ __forceinline__ int par_chunk(const ParFace& f, int sx, int sy, int g, long long* t0, long long* base) {
  *t0 = (long long)blockIdx.x * kParChunk;
  const long long left = f.points - *t0;
  const int np = left < kParChunk ? (int)left : kParChunk;
  for (int q = threadIdx.x; q < np; q += kParThreads) {
    int idx[3];
    par_face_node(f, *t0 + q, idx);
    base[q] = ((long long)(idx[2] + g) * sy + (idx[1] + g)) * sx + (idx[0] + g);
  }
  return np;
}

// Index step of one node along the outward normal.
__device__ __forceinline__ long long par_outward(const ParFace& f, int sx, int sy) {
  const long long s = f.normal == 0 ? 1 : (f.normal == 1 ? (long long)sx : (long long)sx * sy);
  return f.direction * s;
}

// Elements of a chunk per thread, at most.
template <int NV>
constexpr int par_per_thread() {
  return (kParChunk * kParMaxLayers * NV + kParThreads - 1) / kParThreads;
}

// setup_data_to_be_sent (src/DataCommunication.cu:3-40) of one chunk:
// load(v, p) reads variable v of node p.  Every gather of a thread is issued
// before the first is stored, so that they are in flight together.
template <int NV, class Load>
__device__ __forceinline__ void par_pack_chunk(const ParFace& f, int sx, int sy, int g, double* buf, Load load) {
  constexpr int IT = par_per_thread<NV>();
  __shared__ double stage[kParChunk * kParMaxLayers * NV];
  __shared__ long long base[kParChunk];
  const int g1 = g + 1;
  long long t0;
  const int np = par_chunk(f, sx, sy, g, &t0, base);
  const long long outward = par_outward(f, sx, sy);
  __syncthreads();
  const int n = np * g1 * NV;
  double x[IT];
  int slot[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = threadIdx.x + it * kParThreads;
    slot[it] = -1;
    if (e < n) {
      int pt, layer, v;
      par_walk(f, g1, np, e, &pt, &layer, &v);
      slot[it] = (pt * g1 + layer) * NV + v;
      x[it] = load(v, base[pt] - layer * outward);
    }
  }
#pragma unroll
  for (int it = 0; it < IT; ++it)
    if (slot[it] >= 0) stage[slot[it]] = x[it];
  __syncthreads();
  double* out = buf + f.offset + t0 * g1 * NV;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = threadIdx.x + it * kParThreads;
    if (e < n) out[e] = stage[e];
  }
}

// assign_data_received (src/DataCommunication.cuh:201-240) of one chunk:
// the face node takes 0.5 (own + received), the ngg nodes beyond it the
// received ones.  The own values come from `sent` (this face's slice of the
// send buffer, layer 0) when f.own_from_send, else from get(v, p).
// get(v, p) / set(v, p, x) read and write variable v of node p; cv(p, q)
// writes the conservative variables of node p from its nv primitives q[] (in
// LDS).
template <int NV, class Get, class Set, class Cv>
__device__ __forceinline__ void par_unpack_chunk(const ParFace& f, int sx, int sy, int g, const double* recv,
                                                 const double* sent, Get get, Set set, Cv cv) {
  __shared__ double stage[kParChunk * kParMaxLayers * NV];
  __shared__ double own_sent[kParChunk * NV];
  __shared__ long long base[kParChunk];
  const int g1 = g + 1;
  long long t0;
  const int np = par_chunk(f, sx, sy, g, &t0, base);
  const long long outward = par_outward(f, sx, sy);
  const int n = np * g1 * NV;
  const long long first = f.offset + t0 * g1 * NV;
  const double* in = recv + first;
  const double* mine = sent + first;
  const bool from_send = f.own_from_send != 0;
  constexpr int IT = par_per_thread<NV>();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = threadIdx.x + it * kParThreads;
    if (e < n) {
      stage[e] = in[e];
      const int pl = e / NV;  // point * g1 + layer
      if (from_send && pl % g1 == 0) own_sent[(pl / g1) * NV + (e - pl * NV)] = mine[e];
    }
  }
  __syncthreads();
  // own values of the face nodes first (all in flight together), then the
  // writes of every element
  double own[IT];
  long long node[IT];
  int slot[IT], var[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = threadIdx.x + it * kParThreads;
    slot[it] = -1;
    if (e < n) {
      int pt, layer, v;
      par_walk(f, g1, np, e, &pt, &layer, &v);
      node[it] = base[pt] + layer * outward;
      slot[it] = (pt * g1 + layer) * NV + v;
      var[it] = layer == 0 ? v : -1 - v;
      if (layer == 0 && !from_send) own[it] = get(v, node[it]);
    }
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    if (slot[it] < 0) continue;
    double x = stage[slot[it]];
    if (var[it] >= 0) {
      const double mine_v = from_send ? own_sent[(slot[it] / (g1 * NV)) * NV + var[it]] : own[it];
      x = 0.5 * (mine_v + x);
      stage[slot[it]] = x;
    }
    set(var[it] >= 0 ? var[it] : -1 - var[it], node[it], x);
  }
  __syncthreads();
  const int nodes = np * g1;
  for (int e = threadIdx.x; e < nodes; e += kParThreads) {
    const int pt = par_points_fastest(f) ? e % np : e / g1;
    const int layer = par_points_fastest(f) ? e / np : e % g1;
    cv(base[pt] + layer * outward, &stage[(pt * g1 + layer) * NV]);
  }
}

}  // namespace ostc
