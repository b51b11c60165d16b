// ostc_faces.hpp -- host-side geometry of boundary and inner faces, built
// either from whole-face arguments (axis, side) or from the records of the
// reference's connectivity files: one line of boundary N.txt
// (x1 x2 y1 y2 z1 z2 label, 0-based node indices; Boundary::register_boundary,
// src/Mesh.cpp:14-32) and one line pair of inner N.txt (1-based node indices,
// the tangential axis that pairs with the other block's marked one given as
// negative numbers; InnerFace::register_boundary, src/Mesh.cpp:38-176).
#pragma once

#include <algorithm>
#include <cstdlib>
#include <string>

namespace ostc {

// Points of a boundary face: the face layer and its tangential range widened
// by ngg on both sides.
struct FaceRange {
  int axis, side;    // normal axis; side -1: the index-0 face, +1: the index m-1 face
  int lo[3], hi[3];  // inclusive; lo[axis] = hi[axis] = the face layer
};

// One side of a 1-to-1 inner face.  The source points lo..hi are the shared
// node layer and the ngg ghost layers beyond it along `axis`, and the face's
// node range on the two tangential axes.  A source point (i, j, k) meets the
// target point whose index along axis d is off[d] + sgn * (source index along
// axis p), with code[d] = sgn * (p + 1).
struct FaceMap {
  int target;
  int axis, side;    // source face: normal axis and side
  int taxis, tside;  // target face
  int lo[3], hi[3];  // source points, inclusive
  int shared;        // index along `axis` of the node layer both blocks hold
  int code[3], off[3];
  int to_target(int d, const int idx[3]) const {
    const int c = code[d], p = (c < 0 ? -c : c) - 1;
    return off[d] + (c < 0 ? -idx[p] : idx[p]);
  }
};

inline FaceRange whole_face(int axis, int side, const int ext[3], int ngg) {
  FaceRange f{};
  f.axis = axis;
  f.side = side;
  for (int d = 0; d < 3; ++d) {
    f.lo[d] = -ngg;
    f.hi[d] = ext[d] - 1 + ngg;
  }
  f.lo[axis] = f.hi[axis] = side < 0 ? 0 : ext[axis] - 1;
  return f;
}

// A boundary N.txt record r = {x1, x2, y1, y2, z1, z2}: the normal is the first
// axis whose two indices agree; 0 there is the low face, otherwise it must be
// the last node.  Tangential ranges may be given in either order.
inline bool boundary_range(const int r[6], const int ext[3], int ngg, FaceRange* f, std::string* err) {
  int axis = -1;
  for (int d = 0; d < 3 && axis < 0; ++d)
    if (r[2 * d] == r[2 * d + 1]) axis = d;
  if (axis < 0) { *err = "boundary range: no axis with equal start and end (not a face)"; return false; }
  const int v = r[2 * axis];
  if (v != 0 && v != ext[axis] - 1) { *err = "boundary range: the face index is neither 0 nor m-1"; return false; }
  f->axis = axis;
  f->side = v == 0 ? -1 : 1;
  for (int d = 0; d < 3; ++d) {
    int a = r[2 * d], b = r[2 * d + 1];
    if (a > b) std::swap(a, b);
    if (a < 0 || b > ext[d] - 1) { *err = "boundary range outside the block"; return false; }
    if (d == axis) {
      f->lo[d] = f->hi[d] = v;
    } else {
      f->lo[d] = a - ngg;
      f->hi[d] = b + ngg;
    }
  }
  return true;
}

inline FaceMap aligned_face(int axis, int side, const int ext[3], const int text[3], int ngg, int target) {
  FaceMap m{};
  m.target = target;
  m.axis = m.taxis = axis;
  m.side = side;
  m.tside = -side;
  for (int d = 0; d < 3; ++d) {
    m.lo[d] = 0;
    m.hi[d] = ext[d] - 1;
    m.code[d] = d + 1;
    m.off[d] = 0;
  }
  m.shared = side < 0 ? 0 : ext[axis] - 1;
  m.lo[axis] = side < 0 ? -ngg : m.shared;
  m.hi[axis] = side < 0 ? 0 : m.shared + ngg;
  m.off[axis] = (side < 0 ? text[axis] - 1 : 0) - m.shared;
  return m;
}

namespace faces_detail {
// Normal axis and side of one inner N.txt record (1-based; first axis with
// equal start and end; index 1 is the low face, the last node the high one).
inline bool normal_of(const int r[6], const int ext[3], int* axis, int* side) {
  for (int d = 0; d < 3; ++d)
    if (r[2 * d] == r[2 * d + 1]) {
      *axis = d;
      if (r[2 * d] == 1) *side = -1;
      else if (r[2 * d] == ext[d]) *side = 1;
      else return false;
      return true;
    }
  return false;
}
// The tangential axis marked by negative indices, and the other one.
inline bool tangential_of(const int r[6], int axis, int* marked, int* plain) {
  *marked = *plain = -1;
  for (int d = 0; d < 3; ++d) {
    if (d == axis) continue;
    const bool neg_a = r[2 * d] < 0, neg_b = r[2 * d + 1] < 0;
    if (neg_a != neg_b) return false;
    if (neg_a) {
      if (*marked >= 0) return false;
      *marked = d;
    } else {
      if (*plain >= 0) return false;
      *plain = d;
    }
  }
  return *marked >= 0 && *plain >= 0;
}
inline int zero_based(int v) { return v < 0 ? -v - 1 : v - 1; }
}  // namespace faces_detail

// An inner N.txt line pair: src = {x1, x2, y1, y2, z1, z2} of this block's
// face, tgt = the same of the target block's face.  The normals pair with each
// other, the marked tangential axes with each other and the remaining two with
// each other; a pair runs the same way (sgn +1) when both ranges ascend or both
// descend.  Along the normals the source's shared layer meets the target's,
// and the ghost layers continue into the target's interior (sgn -1 when both
// faces are low faces or both high faces).
inline bool inner_face_map(const int src[6], const int tgt[6], const int ext[3], const int text[3], int ngg,
                           int target, FaceMap* m, std::string* err) {
  using namespace faces_detail;
  FaceMap f{};
  f.target = target;
  if (!normal_of(src, ext, &f.axis, &f.side) || !normal_of(tgt, text, &f.taxis, &f.tside)) {
    *err = "inner face: the normal index must be 1 or the last node on both sides";
    return false;
  }
  int sm, sp, tm, tp;
  if (!tangential_of(src, f.axis, &sm, &sp) || !tangential_of(tgt, f.taxis, &tm, &tp)) {
    *err = "inner face: exactly one tangential axis of each side must be marked by negative indices";
    return false;
  }
  if (text[f.taxis] < ngg + 1) { *err = "inner face: target block thinner than ngg+1 layers"; return false; }
  const int pairs[2][2] = {{sm, tm}, {sp, tp}};
  for (const auto& pr : pairs) {
    const int s = pr[0], t = pr[1];
    const int s0 = zero_based(src[2 * s]), s1 = zero_based(src[2 * s + 1]);
    const int t0 = zero_based(tgt[2 * t]), t1 = zero_based(tgt[2 * t + 1]);
    if (std::abs(s1 - s0) != std::abs(t1 - t0) || s0 == s1) {
      *err = "inner face: the paired tangential ranges differ in length (not 1-to-1)";
      return false;
    }
    if (std::min(s0, s1) < 0 || std::max(s0, s1) > ext[s] - 1 || std::min(t0, t1) < 0 ||
        std::max(t0, t1) > text[t] - 1) {
      *err = "inner face: range outside its block";
      return false;
    }
    const int sgn = ((t1 - t0) > 0) == ((s1 - s0) > 0) ? 1 : -1;
    f.code[t] = sgn * (s + 1);
    f.off[t] = t0 - sgn * s0;
    f.lo[s] = std::min(s0, s1);
    f.hi[s] = std::max(s0, s1);
  }
  const int sgn = -(f.side * f.tside);
  f.shared = f.side < 0 ? 0 : ext[f.axis] - 1;
  const int tshared = f.tside < 0 ? 0 : text[f.taxis] - 1;
  f.code[f.taxis] = sgn * (f.axis + 1);
  f.off[f.taxis] = tshared - sgn * f.shared;
  f.lo[f.axis] = f.side < 0 ? -ngg : f.shared;
  f.hi[f.axis] = f.side < 0 ? 0 : f.shared + ngg;
  *m = f;
  return true;
}

}  // namespace ostc
