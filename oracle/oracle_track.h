/*
 * CPU ORACLE (test infrastructure) — tracking residuals, reduction, solve.
 * Restates apps/kfusion/src/cpp/tracking/tracking.cpp and include/commons.h
 * (makeJTJ/solve/toMatrix4).  TooN's GR_SVD and SE3::exp are third-party
 * (TooN @ 92241416d2a4874fd2334e08a5d417dfea6a1a3f, not vendored); their
 * published algorithms are restated here (SVD pseudo-inverse with condition
 * 1e6; SE3 exponential with translation first).
 */
#pragma once
#include "oracle_core.h"

namespace orc {

enum Robust { ROBUST_NONE = 0, ROBUST_HUBER = 1, ROBUST_TUKEY = 2, ROBUST_CAUCHY = 3 };

// Tracking::calc_robust_weight (tracking.cpp:1502-1532).  The reference's
// sq() takes float, so (residual/k) is computed in double, squared in float.
static inline R robust_weight(R r, int fn, double k) {
  double w = 1.0;
  switch (fn) {
    case ROBUST_NONE: w = 1.0; break;
    case ROBUST_HUBER: w = (std::fabs(r) <= k) ? 1.0 : k / std::fabs((double)r); break;
    case ROBUST_TUKEY:
      if (std::fabs(r) <= k) { float q = (float)(sqf((float)r) / sqf((float)k)); w = (double)sqf((float)(1.0 - (double)q)); }
      else w = 0;
      break;
    case ROBUST_CAUCHY: {
      R q = sqf((float)((double)r / k));
      w = (double)(R(1) / (R(1) + q));
      break;
    }
  }
  return (R)w;
}
// Tracking::calc_robust_residual (tracking.cpp:1474-1500)
static inline R robust_loss(R r, int fn, double k) {
  double loss = 0;
  switch (fn) {
    case ROBUST_NONE: loss = 0.5 * r * r; break;
    case ROBUST_HUBER:
      if (std::fabs(r) <= k) loss = 0.5 * r * r;
      else loss = k * (std::fabs((double)r) - 0.5 * k);
      break;
    case ROBUST_TUKEY:
      if (std::fabs(r) <= k) loss = k * k / 6.0 * (1.0 - std::pow(1.0 - (double)sqf((float)((double)r / k)), 3));
      else loss = k * k / 6.0;
      break;
    case ROBUST_CAUCHY: loss = 0.5 * k * k * std::log(1.0 + (double)sqf((float)((double)r / k))); break;
  }
  return (R)loss;
}

// Per-pixel ICP row, Tracking::track_ICP_Kernel (mask overload,
// tracking.cpp:923-1020).
static inline void icp_row(TrackRow& row, int px, int py, const V3* inV, const V3* inN, int iw,
                           const V3* refV, const V3* refN, int rw, int rh, const M4& Ttrack,
                           const M4& view, R dist_th, R normal_th, const V3* cov, const uint8_t* mask) {
  const int i = px + py * iw;
  if (mask && mask[i] != 0) { row.result = -6; return; }
  if (inN[i].x == INVALID) { row.result = -1; return; }
  const V3 pv = affine(Ttrack, inV[i]);
  const V3 pp = affine(view, pv);
  const R ux = pp.x / pp.z, uy = pp.y / pp.z;
  if (ux < 1 || ux > R(rw - 2) || uy < 1 || uy > R(rh - 2) || std::isnan(ux) || std::isnan(uy)) {
    row.result = -2; return;
  }
  const unsigned rx = (unsigned)ux, ry = (unsigned)uy;
  const V3 rn = refN[rx + ry * rw];
  if (rn.x == INVALID) { row.result = -3; return; }
  const V3 diff = refV[rx + ry * rw] - pv;
  const V3 pn = rotate(Ttrack, inN[i]);
  if (length(diff) > dist_th) { row.result = -4; return; }
  if (dot(pn, rn) < normal_th) { row.result = -5; return; }
  row.result = 1;
  const V3 P = cov[i];
  const R sig = rn.x * rn.x * P.x + rn.y * rn.y * P.y + rn.z * rn.z * P.z;
  const R w = std::sqrt((R)(1.0 / (double)sig));
  row.error = w * dot(rn, diff);
  const V3 j0 = w * rn, j1 = w * cross(pv, rn);
  row.J[0] = j0.x; row.J[1] = j0.y; row.J[2] = j0.z; row.J[3] = j1.x; row.J[4] = j1.y; row.J[5] = j1.z;
}

// Per-pixel photometric row, Tracking::trackRGB (mask overload,
// tracking.cpp:1023-1111) with warp (:1114-1121), rgb_residual (:1180-1186)
// and rgb_jacobian (:1123-1150).
static inline void rgb_row(TrackRow& row, int px, int py, const V3* rVr, const V3* rVl, const R* rI, int rw,
                           const R* lI, int lw, int lh, const R* gx, const R* gy, const M4& T_w_r,
                           const M4& T_w_l, const M4& inv_T_w_l, R fx, R fy, R cx, R cy, R inv_cov,
                           R occl, const uint8_t* mask) {
  const int i = px + py * rw;
  if (mask && mask[i] != 0) { row.result = -6; return; }
  const V3 vr = rVr[i], vl = rVl[i];
  if (vr.z <= 0 || vr.z == INVALID) { row.result = -1; return; }
  if (length(vr - vl) > occl) { if (vl.z > 0) { row.result = -3; return; } }
  const V3 wv = affine(T_w_r, vr);
  const V3 lv = affine(inv_T_w_l, wv);
  const V3 pj = rotate(camK(fx, fy, cx, cy), lv);
  const R ux = pj.x / pj.z, uy = pj.y / pj.z;
  if (ux < 1 || ux > R(lw - 2) || uy < 1 || uy > R(lh - 2) || std::isnan(ux) || std::isnan(uy)) {
    row.result = -2; return;
  }
  const R li = bilin(lI, lw, ux, uy);
  const R res = rI[i] - li;
  const R g_x = bilin(gx, lw, ux, uy), g_y = bilin(gy, lw, ux, uy);
  const R s = R(1) / lv.z;
  V3 jt = s * v3(g_x * fx, g_y * fy, -(g_x * lv.x * fx + g_y * lv.y * fy) / lv.z);
  jt = R(-1) * rotate(T_w_l, jt);      // rotate(v, transpose(T)) == rotate(T, v)
  const V3 jr = cross(wv, jt);
  const V3 j0 = inv_cov * jt, j1 = inv_cov * jr;
  row.J[0] = j0.x; row.J[1] = j0.y; row.J[2] = j0.z; row.J[3] = j1.x; row.J[4] = j1.y; row.J[5] = j1.z;
  row.error = inv_cov * res;
  row.result = 1;
}

// Object-frame ICP row, Tracking::track_Obj_ICP_Kernel (tracking.cpp:392-486)
// with obj_warp (:359-367) and obj_icp_residual (:370-390).
static inline void obj_icp_row(TrackRow& row, int px, int py, const V3* cV, const V3* cN, int iw,
                               const V3* refV, const V3* refN, int rw, int rh, const M4& T_c2_o2,
                               const M4& inv_T_c2_o2, const M4& T_c1_o1, const M4& inv_T_w_o1,
                               const M4& K, R dist_th, R normal_th, const V3* cov, const uint8_t* mask) {
  const int i = px + py * iw;
  if (mask && mask[i] != 0) { row.result = -6; return; }
  if (cN[i].x == INVALID) { row.result = -1; return; }
  const V3 c2v = cV[i];
  if (c2v.x == 0 && c2v.y == 0 && c2v.z == 0) { row.result = -1; return; }
  const V3 o2v = affine(inv_T_c2_o2, c2v);
  const V3 c1v = affine(T_c1_o1, o2v);
  const V3 pj = rotate(K, c1v);
  const R ux = pj.x / pj.z, uy = pj.y / pj.z;
  if (ux < 1 || ux > R(rw - 2) || uy < 1 || uy > R(rh - 2) || std::isnan(ux) || std::isnan(uy)) {
    row.result = -2; return;
  }
  const V3 o2n = rotate(inv_T_c2_o2, cN[i]);
  const V3 wrv = bilin3(refV, rw, ux, uy);
  const V3 wrn = bilin3(refN, rw, ux, uy);
  if (wrn.x == INVALID) { row.result = -3; return; }
  const V3 o1rv = affine(inv_T_w_o1, wrv);
  const V3 o1rn = rotate(inv_T_w_o1, wrn);
  const V3 diff = o2v - o1rv;
  const R residual = dot(o1rn, diff);
  if (length(diff) > dist_th) { row.result = -4; return; }
  if (dot(o2n, o1rn) < normal_th) { row.result = -5; return; }
  const V3 P = cov[i];
  const R sig = o1rn.x * o1rn.x * P.x + o1rn.y * o1rn.y * P.y + o1rn.z * o1rn.z * P.z;
  const R w = std::sqrt((R)(1.0 / (double)sig));
  row.error = w * residual;
  const V3 jt = rotate(T_c2_o2, o1rn);
  const V3 j0 = (R(-1) * w) * (R(-1) * jt);   // -1.f * inv_cov * -1.f * Jtrans
  const V3 j1 = w * cross(c2v, jt);
  row.J[0] = j0.x; row.J[1] = j0.y; row.J[2] = j0.z; row.J[3] = j1.x; row.J[4] = j1.y; row.J[5] = j1.z;
  row.result = 1;
}

// Object photometric row, Tracking::track_obj_RGB_kernel (tracking.cpp:491-578)
// with obj_rgb_jacobian (:1152-1178).  The warp is obj_warp(T_c2_o2, T_c1_o1)
// with the argument roles as called at :536-537.
static inline void obj_rgb_row(TrackRow& row, int px, int py, const V3* rVr, const V3* rVl, const R* rI, int rw,
                               const R* lI, int lw, int lh, const R* gx, const R* gy, const M4& T_c2_o2,
                               const M4& inv_T_c1_o1, R fx, R fy, R cx, R cy, R inv_cov, R occl,
                               const uint8_t* mask) {
  const int i = px + py * rw;
  if (mask && mask[i] != 0) { row.result = -6; return; }
  const V3 vr = rVr[i], vl = rVl[i];
  if (vr.z <= 0 || vr.z == INVALID) { row.result = -1; return; }
  if (length(vr - vl) > occl) { if (vl.z > 0) { row.result = -3; return; } }
  const V3 o1v = affine(inv_T_c1_o1, vr);
  const V3 c2v = affine(T_c2_o2, o1v);
  const V3 pj = rotate(camK(fx, fy, cx, cy), c2v);
  const R ux = pj.x / pj.z, uy = pj.y / pj.z;
  if (ux < 1 || ux > R(lw - 2) || uy < 1 || uy > R(lh - 2)) { row.result = -2; return; }
  const R li = bilin(lI, lw, ux, uy);
  const R res = rI[i] - li;
  const R g_x = bilin(gx, lw, ux, uy), g_y = bilin(gy, lw, ux, uy);
  const R s = R(1) / c2v.z;
  const V3 jt = s * v3(g_x * fx, g_y * fy, -(g_x * c2v.x * fx + g_y * c2v.y * fy) / c2v.z);
  const V3 jr = cross(c2v, jt);
  const V3 j0 = inv_cov * jt, j1 = inv_cov * jr;
  row.J[0] = j0.x; row.J[1] = j0.y; row.J[2] = j0.z; row.J[3] = j1.x; row.J[4] = j1.y; row.J[5] = j1.z;
  row.error = inv_cov * res;
  row.result = 1;
}

// Tracking::reduce / reduceKernel (tracking.cpp:1191-1352): 8 row-interleaved
// partials, then summed into row 0.  stack rows at offset jw*jh.
static inline void reduce(R out[8 * 32], const TrackRow* J, int jw, int jh, int w, int h, int stack, int fn,
                          double k) {
#pragma omp parallel for
  for (int b = 0; b < 8; b++) {
    R s[32];
    for (int i = 0; i < 32; i++) s[i] = 0;
    for (int y = b; y < h; y += 8)
      for (int x = 0; x < w; x++)
        for (int st = 0; st < stack; st++) {
          const TrackRow& row = J[(size_t)st * jw * jh + x + y * jw];
          if (row.result < 1) {
            s[29] += row.result == -4 ? 1 : 0;
            s[30] += row.result == -5 ? 1 : 0;
            s[31] += row.result > -4 ? 1 : 0;
            continue;
          }
          const R wgt = robust_weight(row.error, fn, k);
          s[0] += robust_loss(row.error, fn, k);
          for (int i = 0; i < 6; i++) s[1 + i] += wgt * row.error * row.J[i];
          int c = 7;
          for (int a = 0; a < 6; a++)
            for (int bb = a; bb < 6; bb++) s[c++] += wgt * row.J[a] * row.J[bb];
          s[28] += 1;
        }
    for (int i = 0; i < 32; i++) out[b * 32 + i] = s[i];
  }
  for (int j = 1; j < 8; j++)
    for (int i = 0; i < 32; i++) out[i] += out[j * 32 + i];
}

// ---- FP64 6x6 symmetric pseudo-inverse (restates TooN GR_SVD backsub with
// condition 1e6, commons.h:689-696): for a symmetric matrix the SVD is the
// eigendecomposition with |lambda|, computed here with cyclic Jacobi.
static inline void sym_eig6(const double Ain[6][6], double lam[6], double V[6][6]) {
  double A[6][6];
  for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) { A[i][j] = Ain[i][j]; V[i][j] = (i == j); }
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0;
    for (int p = 0; p < 6; p++) for (int q = p + 1; q < 6; q++) off += A[p][q] * A[p][q];
    if (off == 0) break;
    for (int p = 0; p < 6; p++)
      for (int q = p + 1; q < 6; q++) {
        if (A[p][q] == 0) continue;
        double theta = (A[q][q] - A[p][p]) / (2 * A[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < 6; k++) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq; A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 6; k++) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk; A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 6; k++) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq; V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < 6; i++) lam[i] = A[i][i];
}
// makeJTJ (commons.h:672-687) + solve (:689-696)
static inline void solve6(const double s27[27], double x[6]) {
  double b[6], C[6][6];
  for (int i = 0; i < 6; i++) b[i] = s27[i];
  int c = 6;
  for (int i = 0; i < 6; i++) for (int j = i; j < 6; j++) { C[i][j] = s27[c]; C[j][i] = s27[c]; c++; }
  double lam[6], V[6][6];
  sym_eig6(C, lam, V);
  double mx = 0;
  for (int i = 0; i < 6; i++) mx = std::max(mx, std::fabs(lam[i]));
  for (int i = 0; i < 6; i++) x[i] = 0;
  for (int i = 0; i < 6; i++) {
    double sv = std::fabs(lam[i]);
    if (sv * 1e6 <= mx) continue;
    double p = 0;
    for (int k = 0; k < 6; k++) p += V[k][i] * b[k];
    p /= lam[i];
    for (int k = 0; k < 6; k++) x[k] += p * V[k][i];
  }
}
// SE3 exponential, x[0..2] translation, x[3..5] rotation (TooN SE3::exp as
// used at tracking.cpp:1361): Rodrigues + left Jacobian, in double.
static inline void se3_exp(const double x[6], double Rm[3][3], double t[3]) {
  const double w0 = x[3], w1 = x[4], w2 = x[5];
  const double th2 = w0 * w0 + w1 * w1 + w2 * w2;
  double A, B, C;
  if (th2 < 1e-8) { A = 1.0 - th2 / 6.0; B = 0.5 - th2 / 24.0; C = 1.0 / 6.0 - th2 / 120.0; }
  else if (th2 < 1e-6) { C = (1.0 / 6.0) * (1.0 - th2 / 20.0); A = 1.0 - th2 * C; B = 0.5 - th2 / 24.0; }
  else { const double th = std::sqrt(th2); A = std::sin(th) / th; B = (1.0 - std::cos(th)) / th2; C = (1.0 - A) / th2; }
  const double W[3][3] = {{0, -w2, w1}, {w2, 0, -w0}, {-w1, w0, 0}};
  double W2[3][3];
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) { double s = 0; for (int k = 0; k < 3; k++) s += W[i][k] * W[k][j]; W2[i][j] = s; }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Rm[i][j] = (i == j) + A * W[i][j] + B * W2[i][j];
  for (int i = 0; i < 3; i++) {
    double s = x[i];
    for (int k = 0; k < 3; k++) s += (B * W[i][k] + C * W2[i][k]) * x[k];
    t[i] = s;
  }
}
// Tracking::solvePoseKernel (tracking.cpp:1354-1374).  Returns converged.
static inline bool solve_pose(M4& upd, const double sums[32], double xout[6]) {
  double x[6];
  solve6(sums + 1, x);
  double Rm[3][3], t[3];
  se3_exp(x, Rm, t);
  upd = ident();
  for (int i = 0; i < 3; i++) { for (int j = 0; j < 3; j++) upd.m[i][j] = (R)Rm[i][j]; upd.m[i][3] = (R)t[i]; }
  if (xout) for (int i = 0; i < 6; i++) xout[i] = x[i];
  double nx = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
  const R s0 = (R)sums[0], s28 = (R)sums[28];
  // icp_threshold is a float constant (tracking.h:387) compared with the
  // double norm(x) (tracking.cpp:1365): its value is (double)1e-5f
  return (nx < (double)1e-5f) && ((double)std::sqrt(s0 / s28) < 2e-3);
}

}  // namespace orc
