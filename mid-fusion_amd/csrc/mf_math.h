// Host/device POD math for the MID-Fusion dense path on gfx950.
//
// Every expression keeps the operand order of the reference C++ so that,
// compiled with -ffp-contract=off (no FMA contraction) and IEEE division and
// square root, results are bit-identical to the reference semantics where the
// path needs it (block keys).  Reference files: utils/math_utils.h,
// utils/cutil_math.h (semantics only), include/commons.h:624-708.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MF_HD __host__ __device__ __forceinline__

namespace mf {

static constexpr int INVALID = -2;   // include/commons.h:54

struct f3 { float x, y, z; };
MF_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
MF_HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
MF_HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
MF_HD f3 operator*(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
MF_HD f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
MF_HD f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
MF_HD f3 cmul(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
MF_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MF_HD f3 cross(f3 a, f3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
MF_HD float length(f3 v) { return sqrtf(dot(v, v)); }
// normalize(v) = v * rsqrtf(dot(v,v)) with host rsqrtf = 1/sqrtf (cutil_math.h:59-61)
MF_HD f3 normalize(f3 v) { float inv = 1.0f / sqrtf(dot(v, v)); return v * inv; }
// Host fminf/fmaxf of the reference (cutil_math.h:43-49): a<b?a:b, a>b?a:b.
MF_HD float fmin_r(float a, float b) { return a < b ? a : b; }
MF_HD float fmax_r(float a, float b) { return a > b ? a : b; }

// Row-major 4x4 (Matrix4, math_utils.h:47-49)
struct Mat4 { float m[4][4]; };
MF_HD Mat4 ident() { Mat4 r; for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) r.m[i][j] = (i == j) ? 1.f : 0.f; return r; }
// operator*(Matrix4, float3) (math_utils.h:105-110)
MF_HD f3 affine(const Mat4& M, f3 v) {
  return mk3((M.m[0][0] * v.x + M.m[0][1] * v.y + M.m[0][2] * v.z) + M.m[0][3],
             (M.m[1][0] * v.x + M.m[1][1] * v.y + M.m[1][2] * v.z) + M.m[1][3],
             (M.m[2][0] * v.x + M.m[2][1] * v.y + M.m[2][2] * v.z) + M.m[2][3]);
}
// rotate(Matrix4, float3) (math_utils.h:112-116)
MF_HD f3 rotate(const Mat4& M, f3 v) {
  return mk3(M.m[0][0] * v.x + M.m[0][1] * v.y + M.m[0][2] * v.z,
             M.m[1][0] * v.x + M.m[1][1] * v.y + M.m[1][2] * v.z,
             M.m[2][0] * v.x + M.m[2][1] * v.y + M.m[2][2] * v.z);
}
MF_HD f3 translation(const Mat4& M) { return mk3(M.m[0][3], M.m[1][3], M.m[2][3]); }
// Matrix4*Matrix4 (commons.h:657-662): dot accumulated k = 0..3
MF_HD Mat4 mul(const Mat4& A, const Mat4& B) {
  Mat4 C;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      float s = A.m[i][0] * B.m[0][j];
      s = s + A.m[i][1] * B.m[1][j];
      s = s + A.m[i][2] * B.m[2][j];
      s = s + A.m[i][3] * B.m[3][j];
      C.m[i][j] = s;
    }
  return C;
}
// inverse(Matrix4) (commons.h:649-655): Gauss-Jordan with partial pivoting
// in float.
MF_HD Mat4 inverse(const Mat4& In) {
  float A[4][4], b[4][4];
  for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) { A[i][j] = In.m[i][j]; b[i][j] = (i == j) ? 1.f : 0.f; }
  for (int i = 0; i < 4; i++) {
    int am = i; float mx = fabsf(A[i][i]);
    for (int ii = i + 1; ii < 4; ii++) { float v = fabsf(A[ii][i]); if (v > mx) { mx = v; am = ii; } }
    float inv = 1.0f / A[am][i];
    if (am != i) {
      for (int j = 0; j < 4; j++) { float t = A[i][j]; A[i][j] = A[am][j]; A[am][j] = t; }
      for (int j = 0; j < 4; j++) { float t = b[i][j]; b[i][j] = b[am][j]; b[am][j] = t; }
    }
    for (int j = i + 1; j < 4; j++) A[i][j] = A[i][j] * inv;
    for (int j = 0; j < 4; j++) b[i][j] = b[i][j] * inv;
    for (int u = i + 1; u < 4; u++) {
      float f = A[u][i];
      for (int j = i + 1; j < 4; j++) A[u][j] = A[u][j] - f * A[i][j];
      for (int j = 0; j < 4; j++) b[u][j] = b[u][j] - f * b[i][j];
    }
  }
  for (int i = 3; i > 0; i--)
    for (int u = 0; u < i; u++)
      for (int j = 0; j < 4; j++) b[u][j] = b[u][j] - A[u][i] * b[i][j];
  Mat4 r;
  for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) r.m[i][j] = b[i][j];
  return r;
}
// getCameraMatrix / getInverseCameraMatrix (commons.h:631-647)
MF_HD Mat4 camK(float fx, float fy, float cx, float cy) {
  Mat4 K = ident(); K.m[0][0] = fx; K.m[0][2] = cx; K.m[1][1] = fy; K.m[1][2] = cy; return K;
}
MF_HD Mat4 camKinv(float fx, float fy, float cx, float cy) {
  Mat4 K = ident(); K.m[0][0] = 1.0f / fx; K.m[0][2] = -cx / fx; K.m[1][1] = 1.0f / fy; K.m[1][2] = -cy / fy; return K;
}

// ---- Morton keys (lib/utils/morton_utils.hpp:38-73, lib/octant_ops.hpp:51-55)
MF_HD uint64_t expand21(uint64_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}
MF_HD uint64_t compact21(uint64_t v) {
  uint64_t x = v & 0x1249249249249249ULL;
  x = (x | x >> 2) & 0x10c30c30c30c30c3ULL;
  x = (x | x >> 4) & 0x100f00f00f00f00fULL;
  x = (x | x >> 8) & 0x1f0000ff0000ffULL;
  x = (x | x >> 16) & 0x1f00000000ffffULL;
  x = (x | x >> 32) & 0x1fffffULL;
  return x;
}
MF_HD uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) { return expand21(x) | (expand21(y) << 1) | (expand21(z) << 2); }
// expand21 for v < 1024 in 32-bit arithmetic (block coordinates of volumes
// up to 8192^3: the same bits, interleaved into a 30-bit Morton index)
MF_HD uint32_t expand10(uint32_t v) {
  uint32_t x = v & 0x3ffu;
  x = (x | x << 16) & 0x030000ffu;
  x = (x | x << 8) & 0x0300f00fu;
  x = (x | x << 4) & 0x030c30c3u;
  x = (x | x << 2) & 0x09249249u;
  return x;
}
MF_HD uint32_t morton3_10(uint32_t x, uint32_t y, uint32_t z) { return expand10(x) | (expand10(y) << 1) | (expand10(z) << 2); }

// Per-pixel TrackData (commons.h:624-628)
struct TrackRow { int result; float error; float J[6]; };

}  // namespace mf
