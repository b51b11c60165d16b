// ostc_oracle_faces.h -- test infrastructure only (the CPU oracle).
// Restates how the reference turns its connectivity records into loop ranges:
// Boundary::register_boundary (src/Mesh.cpp:14-32) for one boundary N.txt
// line and InnerFace::register_boundary (src/Mesh.cpp:38-176, the
// three-dimensional branch) for one inner N.txt line pair, step by step, and
// the index loop of inner_communication (src/DataCommunication.cuh:60-105).
#pragma once

#include <cstdlib>
#include <utility>

namespace orc_faces {

inline int sgn(int v) { return (v > 0) - (v < 0); }

// Boundary after register_boundary: normal `face`, `direction` -1 for the
// index-0 face, range over the face layer and the tangential range widened
// by ngg (0-based node indices in the record).
struct Boundary {
  int face = 0, direction = 1;
  int range_start[3], range_end[3];
};

inline bool register_boundary(const int r[6], int ngg, const int ext[3], Boundary* out) {
  Boundary b;
  for (int i = 0; i < 3; ++i) {
    b.range_start[i] = r[2 * i];
    b.range_end[i] = r[2 * i + 1];
  }
  int face = -1;
  for (int i = 0; i < 3; ++i)
    if (b.range_start[i] == b.range_end[i]) {
      face = i;
      if (b.range_start[i] == 0) b.direction = -1;
      break;
    }
  if (face < 0) return false;
  b.face = face;
  // the record must name a face of the block (the reference assumes it)
  if (b.direction > 0 && b.range_start[face] != ext[face] - 1) return false;
  for (int i = 0; i < 3; ++i) {
    if (i == face) continue;
    if (b.range_end[i] < b.range_start[i]) std::swap(b.range_end[i], b.range_start[i]);
    if (b.range_start[i] < 0 || b.range_end[i] > ext[i] - 1) return false;
    b.range_start[i] -= ngg;
    b.range_end[i] += ngg;
  }
  *out = b;
  return true;
}

// InnerFace after register_boundary.
struct Inner {
  int target = 0;
  int face = 0, direction = 1, target_face = 0, target_direction = 1;
  int range_start[3], range_end[3], target_start[3], target_end[3];
  int src_tar[3] = {0, 0, 0}, loop_dir[3] = {1, 1, 1}, target_loop_dir[3] = {1, 1, 1}, n_point[3];
};

inline bool register_inner(const int src[6], const int tgt[6], int ngg, int target, Inner* out) {
  Inner f;
  f.target = target;
  for (int i = 0; i < 3; ++i) {
    f.range_start[i] = src[2 * i];
    f.range_end[i] = src[2 * i + 1];
    f.target_start[i] = tgt[2 * i];
    f.target_end[i] = tgt[2 * i + 1];
  }
  int face = -1, target_face = -1;
  for (int i = 0; i < 3; ++i)
    if (f.range_start[i] == f.range_end[i]) {
      face = i;
      if (f.range_start[i] == 1) f.direction = -1;
      f.range_start[i] -= 1;
      f.range_end[i] -= 1;
      break;
    }
  for (int i = 0; i < 3; ++i)
    if (f.target_start[i] == f.target_end[i]) {
      target_face = i;
      if (f.target_start[i] == 1) f.target_direction = -1;
      f.target_start[i] -= 1;
      f.target_end[i] -= 1;
      break;
    }
  if (face < 0 || target_face < 0) return false;
  f.face = face;
  f.target_face = target_face;
  int src_tar[3] = {0, 0, 0};
  src_tar[target_face] = -(f.direction * f.target_direction) * (face + 1);
  int next_face = -1, target_next_face = -1;
  for (int i = 0; i < 3; ++i)
    if (f.range_start[i] < 0) {
      next_face = i;
      f.range_start[i] = -f.range_start[i] - 1;
      f.range_end[i] = -f.range_end[i] - 1;
      break;
    }
  for (int i = 0; i < 3; ++i)
    if (f.target_start[i] < 0) {
      target_next_face = i;
      f.target_start[i] = -f.target_start[i] - 1;
      f.target_end[i] = -f.target_end[i] - 1;
      break;
    }
  if (next_face < 0 || target_next_face < 0 || next_face == face || target_next_face == target_face) return false;
  src_tar[target_next_face] = sgn((f.target_end[target_next_face] - f.target_start[target_next_face]) *
                                  (f.range_end[next_face] - f.range_start[next_face])) *
                              (next_face + 1);
  next_face = 3 - face - next_face;
  target_next_face = 3 - target_face - target_next_face;
  f.range_start[next_face] -= 1;
  f.range_end[next_face] -= 1;
  f.target_start[target_next_face] -= 1;
  f.target_end[target_next_face] -= 1;
  src_tar[target_next_face] = sgn((f.target_end[target_next_face] - f.target_start[target_next_face]) *
                                  (f.range_end[next_face] - f.range_start[next_face])) *
                              (next_face + 1);
  // the loop covers the shared node layer and ngg ghost layers
  f.range_start[face] = f.direction < 0 ? -ngg : f.range_end[face];
  f.range_end[face] = f.direction < 0 ? 0 : f.range_end[face] + ngg;
  const int sgn_tar = sgn(src_tar[target_face]);
  if (f.target_direction > 0)
    f.target_start[target_face] = sgn_tar > 0 ? f.target_end[target_face] - ngg : f.target_end[target_face];
  else
    f.target_start[target_face] = sgn_tar > 0 ? 0 : ngg;
  for (int i = 0; i < 3; ++i) {
    if (src_tar[i] == 0) return false;  // a tangential range of one node
    f.loop_dir[i] = f.range_start[i] <= f.range_end[i] ? 1 : -1;
    f.target_loop_dir[i] = sgn(src_tar[i]);
    f.src_tar[i] = std::abs(src_tar[i]) - 1;
    f.n_point[i] = std::abs(f.range_start[i] - f.range_end[i]) + 1;
  }
  *out = f;
  return true;
}

// The point pairs of inner_communication: fn(idx, idx_tar, on_shared_layer)
// for every n in [0, n_point).
inline void point_of(const Inner& f, const int n[3], int idx[3], int idx_tar[3], bool* shared) {
  int d_idx[3];
  for (int i = 0; i < 3; ++i) {
    d_idx[i] = f.loop_dir[i] * n[i];
    idx[i] = f.range_start[i] + d_idx[i];
  }
  for (int i = 0; i < 3; ++i) idx_tar[i] = f.target_start[i] + f.target_loop_dir[i] * d_idx[f.src_tar[i]];
  const int face_dir = f.direction > 0 ? f.range_start[f.face] : f.range_end[f.face];
  *shared = idx[f.face] == face_dir;
}

// All mapped points inside the blocks' padded index ranges (ngg ghosts)?
inline bool inner_in_bounds(const Inner& f, const int ext[3], const int text[3], int ngg) {
  for (int c = 0; c < 8; ++c) {
    int n[3], idx[3], tid[3];
    for (int i = 0; i < 3; ++i) n[i] = (c >> i & 1) ? f.n_point[i] - 1 : 0;
    bool sh;
    point_of(f, n, idx, tid, &sh);
    for (int i = 0; i < 3; ++i)
      if (idx[i] < -ngg || idx[i] > ext[i] - 1 + ngg || tid[i] < 0 || tid[i] > text[i] - 1) return false;
  }
  return true;
}

// The inner N.txt records of an aligned 1-to-1 face: this block's face on
// `axis` at `side`, the target's opposite face, the first tangential axis
// marked on both sides.
inline void aligned_records(int axis, int side, const int ext[3], const int text[3], int src[6], int tgt[6]) {
  const int t1 = axis == 0 ? 1 : 0;
  for (int i = 0; i < 3; ++i) {
    const int mk = i == t1 ? -1 : 1;
    src[2 * i] = tgt[2 * i] = mk;
    src[2 * i + 1] = mk * ext[i];
    tgt[2 * i + 1] = mk * text[i];
  }
  src[2 * axis] = src[2 * axis + 1] = side < 0 ? 1 : ext[axis];
  tgt[2 * axis] = tgt[2 * axis + 1] = side < 0 ? text[axis] : 1;
}

}  // namespace orc_faces
