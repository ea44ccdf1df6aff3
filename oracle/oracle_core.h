/*
 * CPU ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A plain single-threaded (optionally OpenMP) restatement of MID-Fusion's
 * per-frame dense path, used as the checker for the HIP library and as the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * The scalar type R is float by default (reference semantics: the same IEEE
 * operations in the same order as the reference C++).  Building with
 * -DORC_REAL=double gives the FP64 oracle used for pose/TSDF error reports.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 */
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL R;

namespace orc {

static const int INVALID = -2;                 // include/commons.h:54

struct V3 { R x, y, z; };
struct V2 { R x, y; };
static inline V3 v3(R x, R y, R z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator*(R s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline V3 operator*(V3 a, R s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 cmul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 operator/(V3 a, R s) { return v3(a.x / s, a.y / s, a.z / s); }
// cutil-style dot: ((a.x*b.x) + a.y*b.y) + a.z*b.z
static inline R dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline R rsqrt_(R x) { return R(1) / std::sqrt(x); }   // utils/cutil_math.h:59-61
static inline V3 normalize(V3 v) { R inv = rsqrt_(dot(v, v)); return v * inv; }
static inline R length(V3 v) { return std::sqrt(dot(v, v)); }
static inline R sqf(float r) { return (R)(r * r); }               // utils/math_utils.h:92 (float only)

// Row-major 4x4 (utils/math_utils.h:47-49: four float4 rows)
struct M4 { R m[4][4]; };
static inline M4 ident() { M4 r; for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) r.m[i][j] = (i == j); return r; }
// operator*(Matrix4, float3): affine (math_utils.h:105-110)
static inline V3 affine(const M4& M, V3 v) {
  return v3((M.m[0][0] * v.x + M.m[0][1] * v.y + M.m[0][2] * v.z) + M.m[0][3],
            (M.m[1][0] * v.x + M.m[1][1] * v.y + M.m[1][2] * v.z) + M.m[1][3],
            (M.m[2][0] * v.x + M.m[2][1] * v.y + M.m[2][2] * v.z) + M.m[2][3]);
}
// rotate(Matrix4, float3) (math_utils.h:112-116)
static inline V3 rotate(const M4& M, V3 v) {
  return v3(M.m[0][0] * v.x + M.m[0][1] * v.y + M.m[0][2] * v.z,
            M.m[1][0] * v.x + M.m[1][1] * v.y + M.m[1][2] * v.z,
            M.m[2][0] * v.x + M.m[2][1] * v.y + M.m[2][2] * v.z);
}
// rotate(float3, Matrix4): dot with columns (math_utils.h:117-121)
static inline V3 rotate_t(V3 v, const M4& M) {
  return v3(M.m[0][0] * v.x + M.m[1][0] * v.y + M.m[2][0] * v.z,
            M.m[0][1] * v.x + M.m[1][1] * v.y + M.m[2][1] * v.z,
            M.m[0][2] * v.x + M.m[1][2] * v.y + M.m[2][2] * v.z);
}
static inline V3 translation(const M4& M) { return v3(M.m[0][3], M.m[1][3], M.m[2][3]); }
// Matrix4*Matrix4 (commons.h:657-662, TooN product, dot accumulated k=0..3)
static inline M4 mul(const M4& A, const M4& B) {
  M4 C;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      R s = A.m[i][0] * B.m[0][j];
      s = s + A.m[i][1] * B.m[1][j];
      s = s + A.m[i][2] * B.m[2][j];
      s = s + A.m[i][3] * B.m[3][j];
      C.m[i][j] = s;
    }
  return C;
}
// inverse(Matrix4) (commons.h:649-655): Gauss-Jordan elimination with partial
// pivoting in the working precision (TooN gaussian_elimination; rounding of
// TooN itself is parity-unpinned).
static inline M4 inverse(const M4& In) {
  R A[4][4], b[4][4];
  for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) { A[i][j] = In.m[i][j]; b[i][j] = (i == j); }
  for (int i = 0; i < 4; i++) {
    int am = i; R mx = std::fabs(A[i][i]);
    for (int ii = i + 1; ii < 4; ii++) { R v = std::fabs(A[ii][i]); if (v > mx) { mx = v; am = ii; } }
    R inv = R(1) / A[am][i];
    if (am != i) {
      for (int j = i; j < 4; j++) std::swap(A[i][j], A[am][j]);
      for (int j = 0; j < 4; j++) std::swap(b[i][j], b[am][j]);
    }
    for (int j = i + 1; j < 4; j++) A[i][j] = A[i][j] * inv;
    for (int j = 0; j < 4; j++) b[i][j] = b[i][j] * inv;
    for (int u = i + 1; u < 4; u++) {
      R f = A[u][i];
      for (int j = i + 1; j < 4; j++) A[u][j] = A[u][j] - f * A[i][j];
      for (int j = 0; j < 4; j++) b[u][j] = b[u][j] - f * b[i][j];
    }
  }
  for (int i = 3; i > 0; i--)
    for (int u = 0; u < i; u++)
      for (int j = 0; j < 4; j++) b[u][j] = b[u][j] - A[u][i] * b[i][j];
  M4 r; for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) r.m[i][j] = b[i][j];
  return r;
}
// getCameraMatrix / getInverseCameraMatrix (commons.h:631-647)
static inline M4 camK(R fx, R fy, R cx, R cy) {
  M4 K = ident(); K.m[0][0] = fx; K.m[0][2] = cx; K.m[1][1] = fy; K.m[1][2] = cy; return K;
}
static inline M4 camKinv(R fx, R fy, R cx, R cy) {
  M4 K = ident(); K.m[0][0] = R(1) / fx; K.m[0][2] = -cx / fx; K.m[1][1] = R(1) / fy; K.m[1][2] = -cy / fy; return K;
}

// TrackData (commons.h:624-628)
struct TrackRow { int result; R error; R J[6]; };

// bilinear_interp<T> (math_utils.h:149-166)
static inline R bilin(const R* f, int w, R px, R py) {
  int bx = (int)std::floor(px), by = (int)std::floor(py);
  R fx = px - std::floor(px), fy = py - std::floor(py);
  R a = f[bx + by * w], b = f[bx + 1 + by * w], c = f[bx + (by + 1) * w], d = f[bx + 1 + (by + 1) * w];
  return a * (R(1) - fx) * (R(1) - fy) + b * (R(1) - fy) * fx + c * (R(1) - fx) * fy + d * fx * fy;
}
static inline V3 bilin3(const V3* f, int w, R px, R py) {
  int bx = (int)std::floor(px), by = (int)std::floor(py);
  R fx = px - std::floor(px), fy = py - std::floor(py);
  V3 a = f[bx + by * w], b = f[bx + 1 + by * w], c = f[bx + (by + 1) * w], d = f[bx + 1 + (by + 1) * w];
  R wa0 = R(1) - fx, wa1 = R(1) - fy;
  V3 r;
  r.x = a.x * wa0 * wa1 + b.x * wa1 * fx + c.x * wa0 * fy + d.x * fx * fy;
  r.y = a.y * wa0 * wa1 + b.y * wa1 * fx + c.y * wa0 * fy + d.y * fx * fy;
  r.z = a.z * wa0 * wa1 + b.z * wa1 * fx + c.z * wa0 * fy + d.z * fx * fy;
  return r;
}

}  // namespace orc
