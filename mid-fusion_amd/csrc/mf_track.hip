// ICP + photometric Gauss-Newton tracking on gfx950.
//
// One kernel per GN iteration computes every pixel's ICP and RGB rows
// (Tracking::track_ICP_Kernel / trackRGB / track_Obj_ICP_Kernel /
// track_obj_RGB_kernel, tracking.cpp:392-578, 923-1111) in registers and
// reduces the 32 robust-weighted sums (Tracking::reduce, tracking.cpp:
// 1208-1352) per workgroup with a register butterfly; rows never reach HBM.
// A one-workgroup-per-instance kernel then accumulates the workgroup partials
// in FP64, solves the 6x6 normal equations in FP64 and applies the reference
// GN control flow (step-back, break-without-update, zero-inlier break;
// tracking.cpp:300-334, 664-688), so a whole pyramid of iterations is
// enqueued without host synchronisation.  A launch covers any number of instances (camera or objects)
// through its grid's y dimension.
#include "mf_kernels.h"

namespace mf {

static constexpr int TB = 256;   // threads per tracking workgroup (4 waves)
static constexpr int NW = TB / 64;

// Level planes (live / reference pyramid level, stride lstride) and the
// instance's model vertex (planes 0-2) and normal (planes 3-5) maps.
__device__ __forceinline__ f3 lv3(const TrackLaunch& p, int pl, size_t i) {
  const float* b = p.live + (size_t)pl * p.lstride;
  return mk3(ldx(b, (uint32_t)i), ldx(b + p.lstride, (uint32_t)i), ldx(b + 2 * (size_t)p.lstride, (uint32_t)i));
}
__device__ __forceinline__ f3 rf3(const TrackLaunch& p, int pl, size_t i) {
  const float* b = p.ref + (size_t)pl * p.lstride;
  return mk3(ldx(b, (uint32_t)i), ldx(b + p.lstride, (uint32_t)i), ldx(b + 2 * (size_t)p.lstride, (uint32_t)i));
}
__device__ __forceinline__ f3 mdl3(const TrackLaunch& p, const TrackInst& in, int pl, size_t i) {
  const size_t n = (size_t)p.W * p.H;
  const float* b = in.mvn + pl * n;
  return mk3(ldx(b, (uint32_t)i), ldx(b + n, (uint32_t)i), ldx(b + 2 * n, (uint32_t)i));
}

// bilinear_interp<T> (utils/math_utils.h:149-166)
__device__ __forceinline__ float bilin(const float* __restrict__ f, int w, float px, float py) {
  const float fpx = floorf(px), fpy = floorf(py);
  const int bx = (int)fpx, by = (int)fpy;
  const float fx = px - fpx, fy = py - fpy;
  const uint32_t i0 = (uint32_t)bx + (uint32_t)by * (uint32_t)w;
  const float a = ldx(f, i0), b = ldx(f, i0 + 1), c = ldx(f, i0 + w), d = ldx(f, i0 + w + 1);
  return a * (1.f - fx) * (1.f - fy) + b * (1.f - fy) * fx + c * (1.f - fx) * fy + d * fx * fy;
}
__device__ __forceinline__ float bmix(float a, float b, float c, float d, float w0, float w1, float fx, float fy) {
  return a * w0 * w1 + b * w1 * fx + c * w0 * fy + d * fx * fy;
}
__device__ __forceinline__ f3 bilin3(const float* f0, size_t stride, int w, float px, float py) {
  const float fpx = floorf(px), fpy = floorf(py);
  const int bx = (int)fpx, by = (int)fpy;
  const float fx = px - fpx, fy = py - fpy;
  const uint32_t i0 = (uint32_t)bx + (uint32_t)by * (uint32_t)w;
  const float w0 = 1.f - fx, w1 = 1.f - fy;
  f3 r;
  float* rr = &r.x;
  for (int k = 0; k < 3; k++) {
    const float* p = f0 + k * stride;
    rr[k] = bmix(ldx(p, i0), ldx(p, i0 + 1), ldx(p, i0 + w), ldx(p, i0 + w + 1), w0, w1, fx, fy);
  }
  return r;
}
// bilin3 over taps that all hold the same value c per component (the same
// arithmetic, so the same result bits as loading them)
__device__ __forceinline__ f3 bilin3_const(f3 c, float px, float py) {
  const float fpx = floorf(px), fpy = floorf(py);
  const float fx = px - fpx, fy = py - fpy;
  const float w0 = 1.f - fx, w1 = 1.f - fy;
  return mk3(bmix(c.x, c.x, c.x, c.x, w0, w1, fx, fy), bmix(c.y, c.y, c.y, c.y, w0, w1, fx, fy),
             bmix(c.z, c.z, c.z, c.z, w0, w1, fx, fy));
}
// True when the 2x2 bilinear taps at (px, py) (inside [1, W-2] x [1, H-2])
// all lie in model-map tiles holding only raycast misses.
__device__ __forceinline__ bool taps_all_miss(const uint8_t* tiles, int W, float px, float py) {
  const int bx = (int)floorf(px), by = (int)floorf(py);
  const int ntx = (W + MAP_TILE - 1) / MAP_TILE;
  const int tx0 = bx / MAP_TILE, tx1 = (bx + 1) / MAP_TILE, ty0 = by / MAP_TILE, ty1 = (by + 1) / MAP_TILE;
  return (ldx(tiles, (uint32_t)(ty0 * ntx + tx0)) | ldx(tiles, (uint32_t)(ty0 * ntx + tx1)) |
          ldx(tiles, (uint32_t)(ty1 * ntx + tx0)) | ldx(tiles, (uint32_t)(ty1 * ntx + tx1))) == 0;
}

// True when the 2x2 bilinear taps at (px, py) all lie in tiles outside the
// box tb = {x0, y0, x1, y1} (in tiles) that holds every tile with a surface.
__device__ __forceinline__ bool taps_outside_box(const int (&tb)[4], float px, float py) {
  const int bx = (int)floorf(px), by = (int)floorf(py);
  const int tx0 = bx / MAP_TILE, tx1 = (bx + 1) / MAP_TILE, ty0 = by / MAP_TILE, ty1 = (by + 1) / MAP_TILE;
  return tx1 < tb[0] || tx0 > tb[2] || ty1 < tb[1] || ty0 > tb[3];
}

// calc_robust_weight / calc_robust_residual (tracking.cpp:1474-1532).  The
// reference's sq() takes float: (e/k) in double, squared in float.
// Cauchy, the reference's default (tracking.h:342-345):
__device__ __forceinline__ void robust_cauchy(float e, double k, float& w, float& loss) {
  const float r = (float)((double)e / k);
  const float q = r * r;
  w = 1.0f / (1.0f + q);
  loss = (float)(0.5 * k * k * log(1.0 + (double)q));
}
__device__ __forceinline__ void robust(float e, int fn, double k, float& w, float& loss) {
  switch (fn) {
    case 0: w = 1.f; loss = (float)(0.5 * e * e); break;
    case 1:
      if (fabsf(e) <= k) { w = 1.f; loss = (float)(0.5 * e * e); }
      else { w = (float)(k / fabs((double)e)); loss = (float)(k * (fabs((double)e) - 0.5 * k)); }
      break;
    case 2:
      if (fabsf(e) <= k) {
        const float q = (e * e) / ((float)k * (float)k);
        const float a = (float)(1.0 - (double)q);
        w = a * a;
        const float r = (float)((double)e / k);
        loss = (float)(k * k / 6.0 * (1.0 - pow(1.0 - (double)(r * r), 3)));
      } else { w = 0.f; loss = (float)(k * k / 6.0); }
      break;
    default: robust_cauchy(e, k, w, loss);
  }
}

// Pose matrices a row reads, staged in LDS by the workgroup and re-read from
// LDS (broadcast) inside the row loop: holding 60 uniform floats in scalar
// registers spills them, and in vector registers costs occupancy.
// `mov` has bit l set when label l is a moving object (camera re-track).
struct Mats { Mat4 T, invT, A, B, C; uint64_t mov; int tb[4]; };

// Marks loaded values as used at this point, so that loads issued above it
// complete under one wait instead of being sunk, one by one, behind the
// validity tests that read them: a row then costs one round trip for its
// pixel's own inputs and one for its gathers.
__device__ __forceinline__ void hold(float v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void hold(int v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void hold(const f3& v) { hold(v.x); hold(v.y); hold(v.z); }

// Camera re-track outlier mask from the raycast labels of moving objects,
// resized as cv::resize does for a 0/255 mask (tracking.cpp:1782-1788): the
// label taps of pixel (x, y) at the launch's level (1 at level 0, else 2x2).
struct LabelTaps { int l[4]; };
__device__ __forceinline__ int label_taps(const TrackLaunch& p, int x, int y, LabelTaps& t) {
  const int s = 1 << p.level;
  const int o = (p.level <= 1) ? 0 : (s / 2 - 1);
  const int n = (p.level == 0) ? 1 : 2;
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int i = 0; i < 2; i++)
      t.l[2 * j + i] = (i < n && j < n) ? ldx(p.label, (uint32_t)(x * s + o + i) + (uint32_t)(y * s + o + j) * (uint32_t)p.W) : 0;
  return n;
}
__device__ __forceinline__ bool taps_moving(const TrackLaunch& p, const LabelTaps& t, uint64_t mov) {
  bool m = false;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int l = t.l[k];
    if (l > 0 && l < p.n_labels && ((mov >> l) & 1)) m = true;
  }
  return m;
}

// ICP row (camera), track_ICP_Kernel mask overload (tracking.cpp:923-1020)
template <class MT>
__device__ __forceinline__ int icp_row(const TrackLaunch& p, const TrackInst& in, const MT& M, int x, int y,
                                       size_t i, float& e, float J[6]) {
  const f3 nl = lv3(p, PL_N, i), vl = lv3(p, PL_V, i), P = lv3(p, PL_COV, i);
  LabelTaps lt;
  if (p.label) label_taps(p, x, y, lt);
  hold(nl); hold(vl); hold(P);
  if (p.label) {
#pragma unroll
    for (int k = 0; k < 4; k++) hold(lt.l[k]);
    if (taps_moving(p, lt, M.mov)) return -6;
  }
  if (nl.x == INVALID) return -1;
  const f3 pv = affine(M.T, vl);
  const f3 pp = affine(M.B, pv);
  const float ux = pp.x / pp.z, uy = pp.y / pp.z;
  if (ux < 1 || ux > (float)(p.W - 2) || uy < 1 || uy > (float)(p.H - 2) || isnan(ux) || isnan(uy)) return -2;
  const unsigned rx = (unsigned)ux, ry = (unsigned)uy;
  const size_t ri = rx + (size_t)ry * p.W;
  const f3 rn = mdl3(p, in, 3, ri), rv = mdl3(p, in, 0, ri);
  hold(rn); hold(rv);
  if (rn.x == INVALID) return -3;
  const f3 diff = rv - pv;
  const f3 pn = rotate(M.T, nl);
  if (length(diff) > 0.1f) return -4;
  if (dot(pn, rn) < 0.8f) return -5;
  const float sig = rn.x * rn.x * P.x + rn.y * rn.y * P.y + rn.z * rn.z * P.z;
  const float w = sqrtf((float)(1.0 / (double)sig));
  e = w * dot(rn, diff);
  const f3 j0 = w * rn, j1 = w * cross(pv, rn);
  J[0] = j0.x; J[1] = j0.y; J[2] = j0.z; J[3] = j1.x; J[4] = j1.y; J[5] = j1.z;
  return 1;
}

// Object ICP row, track_Obj_ICP_Kernel (tracking.cpp:392-486), from the live
// pixel's normal, vertex and covariance.
template <class MT>
__device__ __forceinline__ int obj_icp_core(const TrackLaunch& p, const TrackInst& in, const MT& M, const f3& nl,
                                            const f3& c2v, const f3& P, float& e, float J[6]) {
  if (nl.x == INVALID) return -1;
  if (c2v.x == 0 && c2v.y == 0 && c2v.z == 0) return -1;
  const f3 o2v = affine(M.invT, c2v);
  const f3 c1v = affine(M.A, o2v);
  const f3 pj = rotate(camK(p.fx, p.fy, p.cx, p.cy), c1v);
  const float ux = pj.x / pj.z, uy = pj.y / pj.z;
  if (ux < 1 || ux > (float)(p.W - 2) || uy < 1 || uy > (float)(p.H - 2) || isnan(ux) || isnan(uy)) return -2;
  const f3 o2n = rotate(M.invT, nl);
  f3 wrv, wrn;
  if (in.tiles && taps_all_miss(in.tiles, p.W, ux, uy)) {
    // every tap is a raycast miss of this object: vertex 0, normal
    // (INVALID, 0, 0), so the interpolation needs no loads
    wrv = bilin3_const(mk3(0.f, 0.f, 0.f), ux, uy);
    wrn = bilin3_const(mk3((float)INVALID, 0.f, 0.f), ux, uy);
  } else {
    wrv = bilin3(in.mvn, (size_t)p.W * p.H, p.W, ux, uy);
    wrn = bilin3(in.mvn + (size_t)3 * p.W * p.H, (size_t)p.W * p.H, p.W, ux, uy);
    hold(wrv); hold(wrn);
  }
  if (wrn.x == INVALID) return -3;
  const f3 o1rv = affine(M.B, wrv);
  const f3 o1rn = rotate(M.B, wrn);
  const f3 diff = o2v - o1rv;
  const float residual = dot(o1rn, diff);
  if (length(diff) > 0.1f) return -4;
  if (dot(o2n, o1rn) < 0.8f) return -5;
  const float sig = o1rn.x * o1rn.x * P.x + o1rn.y * o1rn.y * P.y + o1rn.z * o1rn.z * P.z;
  const float w = sqrtf((float)(1.0 / (double)sig));
  e = w * residual;
  const f3 jt = rotate(M.T, o1rn);
  const f3 j0 = ((-1.0f) * w) * ((-1.0f) * jt);
  const f3 j1 = w * cross(c2v, jt);
  J[0] = j0.x; J[1] = j0.y; J[2] = j0.z; J[3] = j1.x; J[4] = j1.y; J[5] = j1.z;
  return 1;
}
// The same row with its loads in the order the tests need them: the live
// vertex and the normal's x first, the model-map taps only for projections
// into the box of tiles with a surface (rows outside it interpolate the miss
// constants with the same arithmetic, so their results are unchanged), and
// the rest of the normal and the covariance only for rows that reach the
// normal test.  Most of an object's rows cover pixels far from it and end at
// the projection or at the taps.
template <class MT>
__device__ __forceinline__ int obj_icp_rest(const TrackLaunch& p, const TrackInst& in, const MT& M, size_t i,
                                            const f3& c2v, float nlx, float& e, float J[6]) {
  const float* nb = p.live + (size_t)PL_N * p.lstride;
  if (nlx == INVALID) return -1;
  if (c2v.x == 0 && c2v.y == 0 && c2v.z == 0) return -1;
  const f3 o2v = affine(M.invT, c2v);
  const f3 c1v = affine(M.A, o2v);
  const f3 pj = rotate(camK(p.fx, p.fy, p.cx, p.cy), c1v);
  const float ux = pj.x / pj.z, uy = pj.y / pj.z;
  if (ux < 1 || ux > (float)(p.W - 2) || uy < 1 || uy > (float)(p.H - 2) || isnan(ux) || isnan(uy)) return -2;
  f3 wrv, wrn;
  if (in.tiles && (taps_outside_box(M.tb, ux, uy) || taps_all_miss(in.tiles, p.W, ux, uy))) {
    wrv = bilin3_const(mk3(0.f, 0.f, 0.f), ux, uy);
    wrn = bilin3_const(mk3((float)INVALID, 0.f, 0.f), ux, uy);
  } else {
    wrv = bilin3(in.mvn, (size_t)p.W * p.H, p.W, ux, uy);
    wrn = bilin3(in.mvn + (size_t)3 * p.W * p.H, (size_t)p.W * p.H, p.W, ux, uy);
    hold(wrv); hold(wrn);
  }
  if (wrn.x == INVALID) return -3;
  const f3 o1rv = affine(M.B, wrv);
  const f3 o1rn = rotate(M.B, wrn);
  const f3 diff = o2v - o1rv;
  const float residual = dot(o1rn, diff);
  if (length(diff) > 0.1f) return -4;
  const f3 nl = mk3(nlx, ldx(nb + p.lstride, (uint32_t)i), ldx(nb + 2 * (size_t)p.lstride, (uint32_t)i));
  const f3 P = lv3(p, PL_COV, i);
  const f3 o2n = rotate(M.invT, nl);
  if (dot(o2n, o1rn) < 0.8f) return -5;
  const float sig = o1rn.x * o1rn.x * P.x + o1rn.y * o1rn.y * P.y + o1rn.z * o1rn.z * P.z;
  const float w = sqrtf((float)(1.0 / (double)sig));
  e = w * residual;
  const f3 jt = rotate(M.T, o1rn);
  const f3 j0 = ((-1.0f) * w) * ((-1.0f) * jt);
  const f3 j1 = w * cross(c2v, jt);
  J[0] = j0.x; J[1] = j0.y; J[2] = j0.z; J[3] = j1.x; J[4] = j1.y; J[5] = j1.z;
  return 1;
}
template <class MT>
__device__ __forceinline__ int obj_icp_row(const TrackLaunch& p, const TrackInst& in, const MT& M, int x, int y,
                                           size_t i, float& e, float J[6]) {
  const f3 c2v = lv3(p, PL_V, i);
  const float nlx = ldx(p.live + (size_t)PL_N * p.lstride, (uint32_t)i);
  hold(c2v); hold(nlx);
  return obj_icp_rest(p, in, M, i, c2v, nlx, e, J);
}

// Photometric rows: trackRGB (tracking.cpp:1023-1111) for the camera,
// track_obj_RGB_kernel (tracking.cpp:491-578) for objects, from the reference
// pixel's rendered depth dr, live vertex vl and intensity ir.
template <bool OBJ, class MT>
__device__ __forceinline__ int rgb_core(const TrackLaunch& p, const MT& M, int x, int y, float dr, const f3& vl,
                                        float ir, float& e, float J[6]) {
  const float s = (float)(1 << p.level);
  const float fx = p.fx / s, fy = p.fy / s, cx = p.cx / s, cy = p.cy / s;
  const f3 vr = dr > 0.0f ? dr * rotate(camKinv(fx, fy, cx, cy), mk3((float)x, (float)y, 1.0f)) : mk3(0.f, 0.f, 0.f);
  if (vr.z <= 0.f || vr.z == INVALID) return -1;
  if (length(vr - vl) > 0.1f) { if (vl.z > 0.f) return -3; }
  f3 lv, wv;
  if (OBJ) {
    const f3 o1v = affine(M.C, vr);
    lv = affine(M.T, o1v);
  } else {
    wv = affine(M.A, vr);
    lv = affine(M.invT, wv);
  }
  const f3 pj = rotate(camK(fx, fy, cx, cy), lv);
  const float ux = pj.x / pj.z, uy = pj.y / pj.z;
  if (OBJ) {
    if (ux < 1 || ux > (float)(p.w - 2) || uy < 1 || uy > (float)(p.h - 2)) return -2;
  } else {
    if (ux < 1 || ux > (float)(p.w - 2) || uy < 1 || uy > (float)(p.h - 2) || isnan(ux) || isnan(uy)) return -2;
  }
  const float li = bilin(p.live + (size_t)PL_I * p.lstride, p.w, ux, uy);
  const float res = ir - li;
  const float g_x = bilin(p.live + (size_t)PL_GX * p.lstride, p.w, ux, uy),
              g_y = bilin(p.live + (size_t)PL_GY * p.lstride, p.w, ux, uy);
  const float sc = 1.f / lv.z;
  f3 jt = sc * mk3(g_x * fx, g_y * fy, -(g_x * lv.x * fx + g_y * lv.y * fy) / lv.z);
  f3 jr;
  if (OBJ) {
    jr = cross(lv, jt);
  } else {
    jt = (-1.f) * rotate(M.T, jt);
    jr = cross(wv, jt);
  }
  const float ic = p.inv_cov_rgb;
  const f3 j0 = ic * jt, j1 = ic * jr;
  J[0] = j0.x; J[1] = j0.y; J[2] = j0.z; J[3] = j1.x; J[4] = j1.y; J[5] = j1.z;
  e = ic * res;
  return 1;
}
template <bool OBJ, class MT>
__device__ __forceinline__ int rgb_row(const TrackLaunch& p, const TrackInst& in, const MT& M, int x, int y,
                                       size_t i, float& e, float J[6]) {
  const float dr = ldx(in.rD + p.rd_off, (uint32_t)i);
  // an object's rendered depth is 0 outside its silhouette, where the row
  // ends at the validity test of the core: test it before the other loads
  if (OBJ && !(dr > 0.0f)) return -1;
  const f3 vl = rf3(p, PL_V, i);
  const float ir = ldx(p.ref + (size_t)PL_I * p.lstride, (uint32_t)i);
  hold(dr); hold(vl); hold(ir);
  return rgb_core<OBJ>(p, M, x, y, dr, vl, ir, e, J);
}

// RB = ROBUST_CAUCHY compiles the Cauchy weight alone; any other value the
// run-time choice of fn.  The switch over all four functions (FP64 log and
// pow among them) sets the row kernels' register count: 144 VGPRs for the
// camera rows against 106 with Cauchy alone.
static constexpr int ROBUST_CAUCHY = 3, ROBUST_ANY = -1;
template <int RB>
__device__ __forceinline__ void accum(float* s, int res, float e, const float* J, int fn, double k) {
  if (res < 1) {
    s[29] += res == -4 ? 1.f : 0.f;
    s[30] += res == -5 ? 1.f : 0.f;
    s[31] += res > -4 ? 1.f : 0.f;
    return;
  }
  float w, loss;
  if (RB == ROBUST_CAUCHY) robust_cauchy(e, k, w, loss);
  else robust(e, fn, k, w, loss);
  s[0] += loss;
  const float we = w * e;
#pragma unroll
  for (int a = 0; a < 6; a++) s[1 + a] += we * J[a];
  int c = 7;
#pragma unroll
  for (int a = 0; a < 6; a++) {
    const float wa = w * J[a];
#pragma unroll
    for (int b = a; b < 6; b++) s[c++] += wa * J[b];
  }
  s[28] += 1.f;
}


// The 32 sums of a workgroup into part (32 floats).  Wave butterfly: at each
// step a lane keeps half of its sums and adds the partner lane's copy of that
// half, so after 5 exchange steps (16+8+4+2+1 lane swaps) lane l holds sum
// l>>1 over 32 lanes, and one more swap adds the other 32.  Fixed order, so
// the partial is deterministic.  Called by every thread of the workgroup.
__device__ __forceinline__ void wg_reduce(float (&s)[32], float* __restrict__ part, float (*red)[32]) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  if (__ballot(s[28] != 0.f) == 0) {
    // No lane of this wave has a valid row (sum 28 counts them), so sums
    // 0-28 are +0 in every lane and the butterfly would give +0; only the
    // status counts 29-31 are reduced.  They are small integers, exact in
    // FP32 in any order, so the wave's column of `red` is the same.  (Most
    // of an object's rows lie away from the object and end invalid.)
    float c0 = s[29], c1 = s[30], c2 = s[31];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      c0 += __shfl_xor(c0, o); c1 += __shfl_xor(c1, o); c2 += __shfl_xor(c2, o);
    }
    if (lane < 32) red[tid >> 6][lane] = lane == 29 ? c0 : (lane == 30 ? c1 : (lane == 31 ? c2 : 0.f));
  } else {
#define MF_BFLY(HALF, OFF)                                                   \
  {                                                                          \
    const bool hi = (lane & OFF) != 0;                                       \
    _Pragma("unroll") for (int j = 0; j < HALF; j++) {                       \
      const float keep = hi ? s[j + HALF] : s[j];                            \
      const float send = hi ? s[j] : s[j + HALF];                            \
      s[j] = keep + __shfl_xor(send, OFF);                                   \
    }                                                                        \
  }
  MF_BFLY(16, 32)
  MF_BFLY(8, 16)
  MF_BFLY(4, 8)
  MF_BFLY(2, 4)
  MF_BFLY(1, 2)
#undef MF_BFLY
    const float v = s[0] + __shfl_xor(s[0], 1);
    if ((lane & 1) == 0) red[tid >> 6][lane >> 1] = v;
  }
  __syncthreads();
  if (tid < 32) {
    float a = 0.f;
#pragma unroll
    for (int w = 0; w < NW; w++) a += red[w][tid];
    part[tid] = a;
  }
  __syncthreads();   // red is reused by the caller's next instance
}

// Rows of one instance at one level over the workgroups [0, nactive): FP32
// per-thread sums, the wave butterfly, and this workgroup's 32 partial sums
// in `part` (32 floats).
template <bool ICP, bool RGB, bool OBJ, int RB, class MT>
__device__ __forceinline__ void track_rows(const TrackLaunch& p, const TrackInst& inst, const MT& M, int state,
                                           int wg, int nactive, float* __restrict__ part, float (*red)[32]) {
  const int tid = threadIdx.x;
  const TrackInst in = inst;   // descriptor held in registers across the row loop
  float s[32];
#pragma unroll
  for (int k = 0; k < 32; k++) s[k] = 0.f;
  const int n = p.w * p.h;
  const int nrows = (ICP ? n : 0) + (RGB ? n : 0);
  const size_t N0 = (size_t)p.W * p.H;
  // the row dump holds GN states 0 and 1
  TrackRow* dbg = (p.dbg && state < 2) ? p.dbg + ((size_t)state * p.n_levels + p.level) * 2 * N0 : nullptr;
  // rows [0, n) are ICP rows, [n, 2n) photometric rows of the same pixels
  // (the order of TrackData in the reference); n is a multiple of 64 at
  // every level of a 640x480 or 320x240 pyramid, so waves do not mix kinds
  for (int r = wg * TB + tid; r < nrows; r += nactive * TB) {
    asm volatile("" ::: "memory");   // keep the LDS matrix reads inside the loop
    const bool icp = ICP && (!RGB || r < n);
    const int i = (ICP && !icp) ? r - n : r;
    const int x = i % p.w, y = i / p.w;
    const size_t ro = (size_t)x + (size_t)y * p.W;   // row stride W at every level (tracking.cpp:49-50)
    float e = 0.f, J[6] = {0, 0, 0, 0, 0, 0};
    int res;
    if (icp) {
      if (OBJ) res = obj_icp_row(p, in, M, x, y, (size_t)i, e, J);
      else res = icp_row(p, in, M, x, y, (size_t)i, e, J);
    } else {
      res = rgb_row<OBJ>(p, in, M, x, y, (size_t)i, e, J);
    }
    accum<RB>(s, res, e, J, p.robust, p.robust_k);
    const size_t oo = icp ? ro : N0 + ro;
    if (in.codes && p.level == 0) in.codes[oo] = (int8_t)res;
    if (!icp && in.rgb_err && p.level == 0) in.rgb_err[ro] = e;
    if (dbg) { TrackRow& d = dbg[oo]; d.result = res; d.error = e; for (int q = 0; q < 6; q++) d.J[q] = J[q]; }
  }
  wg_reduce(s, part, red);
}


// Camera rows by pixel: each thread takes whole pixels and computes both
// rows of each (track_ICP_Kernel's and trackRGB's, the same arithmetic as
// icp_row and rgb_core<false>): the pixel's own inputs of both rows in one
// batch, then all its gathers in one batch.  The next pixel's own inputs are
// issued as soon as this pixel's have arrived, so they are in flight with
// this pixel's gathers: k pixels per thread wait on k + 1 memory round trips
// instead of 2k.  (Both pixels' loads in one batch
// took 172 VGPRs, 2 waves per SIMD; this form stays within 3.)  Pixels
// i = wg * TB + tid + k * nactive * TB; ICP row then photometric row of each
// pixel into the thread's FP32 sums.
struct PixelOwn { f3 nl, vl, P, rvl; float dr, ir; LabelTaps lt; };
__device__ __forceinline__ void pixel_own(const TrackLaunch& p, const TrackInst& in, int idx, int n, PixelOwn& o) {
  const uint32_t ii = idx < n ? (uint32_t)idx : 0u;   // past the end: pixel 0, loaded and unused
  o.nl = lv3(p, PL_N, ii); o.vl = lv3(p, PL_V, ii); o.P = lv3(p, PL_COV, ii);
  if (p.label) label_taps(p, (int)ii % p.w, (int)ii / p.w, o.lt);
  o.dr = ldx(in.rD + p.rd_off, ii); o.rvl = rf3(p, PL_V, ii); o.ir = ldx(p.ref + (size_t)PL_I * p.lstride, ii);
}
template <int RB, class MT>
__device__ __forceinline__ void camera_pixels(const TrackLaunch& p, const TrackInst& inst, const MT& M, int state,
                                              int wg, int nactive, float* __restrict__ part, float (*red)[32]) {
  const int tid = threadIdx.x;
  const TrackInst in = inst;
  float s[32];
#pragma unroll
  for (int k = 0; k < 32; k++) s[k] = 0.f;
  const int n = p.w * p.h;
  const size_t N0 = (size_t)p.W * p.H;
  TrackRow* dbg = (p.dbg && state < 2) ? p.dbg + ((size_t)state * p.n_levels + p.level) * 2 * N0 : nullptr;
  const int stride = nactive * TB;
  const float lsc = (float)(1 << p.level);
  const float fxl = p.fx / lsc, fyl = p.fy / lsc, cxl = p.cx / lsc, cyl = p.cy / lsc;
  const float* Ip = p.live + (size_t)PL_I * p.lstride;
  const float* Gx = p.live + (size_t)PL_GX * p.lstride;
  const float* Gy = p.live + (size_t)PL_GY * p.lstride;
  int idx = wg * TB + tid;
  PixelOwn cur;
  pixel_own(p, in, idx, n, cur);
  for (; idx < n; idx += stride) {
    asm volatile("" ::: "memory");   // keep the LDS matrix reads inside the loop
    hold(cur.nl); hold(cur.vl); hold(cur.P); hold(cur.dr); hold(cur.rvl); hold(cur.ir);
    if (p.label) {
#pragma unroll
      for (int k = 0; k < 4; k++) hold(cur.lt.l[k]);
    }
    // the next pixel's own inputs: in flight with this pixel's gathers
    PixelOwn nxt;
    pixel_own(p, in, idx + stride, n, nxt);
    // projections (codes 0: pending)
    const int x = idx % p.w, y = idx / p.w;
    int ci = 0, cr = 0;
    f3 pv = mk3(0.f, 0.f, 0.f), wv = pv, lvv = pv;
    uint32_t ri = 0, i0 = 0;
    float tfx = 0.f, tfy = 0.f;
    // track_ICP_Kernel (tracking.cpp:923-1020), up to the reference lookup
    if (p.label && taps_moving(p, cur.lt, M.mov)) ci = -6;
    else if (cur.nl.x == INVALID) ci = -1;
    else {
      pv = affine(M.T, cur.vl);
      const f3 pp = affine(M.B, pv);
      const float ux = pp.x / pp.z, uy = pp.y / pp.z;
      if (ux < 1 || ux > (float)(p.W - 2) || uy < 1 || uy > (float)(p.H - 2) || isnan(ux) || isnan(uy)) ci = -2;
      else ri = (unsigned)ux + (unsigned)uy * (unsigned)p.W;
    }
    // trackRGB (tracking.cpp:1023-1111), up to the bilinear taps
    const f3 vr = cur.dr > 0.0f ? cur.dr * rotate(camKinv(fxl, fyl, cxl, cyl), mk3((float)x, (float)y, 1.0f))
                                : mk3(0.f, 0.f, 0.f);
    if (vr.z <= 0.f || vr.z == INVALID) cr = -1;
    else if (length(vr - cur.rvl) > 0.1f && cur.rvl.z > 0.f) cr = -3;
    else {
      wv = affine(M.A, vr);
      lvv = affine(M.invT, wv);
      const f3 pj = rotate(camK(fxl, fyl, cxl, cyl), lvv);
      const float ux = pj.x / pj.z, uy = pj.y / pj.z;
      if (ux < 1 || ux > (float)(p.w - 2) || uy < 1 || uy > (float)(p.h - 2) || isnan(ux) || isnan(uy)) cr = -2;
      else {
        const float fpx = floorf(ux), fpy = floorf(uy);
        i0 = (uint32_t)(int)fpx + (uint32_t)(int)fpy * (uint32_t)p.w;
        tfx = ux - fpx; tfy = uy - fpy;
      }
    }
    // gathers of both pending rows
    const uint32_t r0 = ci == 0 ? ri : 0u;
    const f3 rn = mdl3(p, in, 3, r0), rv = mdl3(p, in, 0, r0);
    const uint32_t j0 = cr == 0 ? i0 : 0u;
    const uint32_t w = (uint32_t)p.w;
    float ta[12];
    ta[0] = ldx(Ip, j0); ta[1] = ldx(Ip, j0 + 1); ta[2] = ldx(Ip, j0 + w); ta[3] = ldx(Ip, j0 + w + 1);
    ta[4] = ldx(Gx, j0); ta[5] = ldx(Gx, j0 + 1); ta[6] = ldx(Gx, j0 + w); ta[7] = ldx(Gx, j0 + w + 1);
    ta[8] = ldx(Gy, j0); ta[9] = ldx(Gy, j0 + 1); ta[10] = ldx(Gy, j0 + w); ta[11] = ldx(Gy, j0 + w + 1);
    hold(rn); hold(rv);
#pragma unroll
    for (int k = 0; k < 12; k++) hold(ta[k]);
    // finish both rows
    const size_t ro = (size_t)x + (size_t)y * p.W;   // row stride W at every level (tracking.cpp:49-50)
    {
      float e = 0.f, J[6] = {0, 0, 0, 0, 0, 0};
      int res = ci;
      if (res == 0) {
        if (rn.x == INVALID) res = -3;
        else {
          const f3 diff = rv - pv;
          const f3 pn = rotate(M.T, cur.nl);
          if (length(diff) > 0.1f) res = -4;
          else if (dot(pn, rn) < 0.8f) res = -5;
          else {
            const f3 Pu = cur.P;
            const float sig = rn.x * rn.x * Pu.x + rn.y * rn.y * Pu.y + rn.z * rn.z * Pu.z;
            const float wgt = sqrtf((float)(1.0 / (double)sig));
            e = wgt * dot(rn, diff);
            const f3 j0v = wgt * rn, j1v = wgt * cross(pv, rn);
            J[0] = j0v.x; J[1] = j0v.y; J[2] = j0v.z; J[3] = j1v.x; J[4] = j1v.y; J[5] = j1v.z;
            res = 1;
          }
        }
      }
      accum<RB>(s, res, e, J, p.robust, p.robust_k);
      if (in.codes && p.level == 0) in.codes[ro] = (int8_t)res;
      if (dbg) { TrackRow& d = dbg[ro]; d.result = res; d.error = e; for (int q = 0; q < 6; q++) d.J[q] = J[q]; }
    }
    {
      float e = 0.f, J[6] = {0, 0, 0, 0, 0, 0};
      int res = cr;
      if (res == 0) {
        const float fx = tfx, fy = tfy;
        const float* t = ta;
        const float li = t[0] * (1.f - fx) * (1.f - fy) + t[1] * (1.f - fy) * fx + t[2] * (1.f - fx) * fy + t[3] * fx * fy;
        const float g_x = t[4] * (1.f - fx) * (1.f - fy) + t[5] * (1.f - fy) * fx + t[6] * (1.f - fx) * fy + t[7] * fx * fy;
        const float g_y = t[8] * (1.f - fx) * (1.f - fy) + t[9] * (1.f - fy) * fx + t[10] * (1.f - fx) * fy + t[11] * fx * fy;
        const float rres = cur.ir - li;
        const f3 lv = lvv;
        const float sc = 1.f / lv.z;
        f3 jt = sc * mk3(g_x * fxl, g_y * fyl, -(g_x * lv.x * fxl + g_y * lv.y * fyl) / lv.z);
        jt = (-1.f) * rotate(M.T, jt);
        const f3 jr = cross(wv, jt);
        const float ic = p.inv_cov_rgb;
        const f3 j0v = ic * jt, j1v = ic * jr;
        J[0] = j0v.x; J[1] = j0v.y; J[2] = j0v.z; J[3] = j1v.x; J[4] = j1v.y; J[5] = j1v.z;
        e = ic * rres;
        res = 1;
      }
      accum<RB>(s, res, e, J, p.robust, p.robust_k);
      if (in.codes && p.level == 0) in.codes[N0 + ro] = (int8_t)res;
      if (in.rgb_err && p.level == 0) in.rgb_err[ro] = e;
      if (dbg) { TrackRow& d = dbg[N0 + ro]; d.result = res; d.error = e; for (int q = 0; q < 6; q++) d.J[q] = J[q]; }
    }
    cur = nxt;
  }
  wg_reduce(s, part, red);
}

// Object rows by pixel: each thread takes whole pixels and computes the
// object's ICP row (track_Obj_ICP_Kernel) and photometric row
// (track_obj_RGB_kernel) of each, issuing both rows' leading loads (live
// vertex and normal x, rendered depth) in one batch, so the pixel costs one
// memory round trip where two rows on separate threads cost one each.  Most
// of an object's rows end after that batch (away from the object), so the
// next pixel's leading loads are issued as soon as this pixel's have
// arrived: they are in flight while this pixel's rows finish.  Pixels
// i = wg * TB + tid + k * nactive * TB; ICP row then photometric row of each
// pixel into the thread's FP32 sums.
struct ObjLead { f3 c2v; float nlx, dr; };
template <int RB, class MT>
__device__ __forceinline__ void object_pixels(const TrackLaunch& p, const TrackInst& inst, const MT& M, int state,
                                              int wg, int nactive, float* __restrict__ part, float (*red)[32]) {
  const int tid = threadIdx.x;
  const TrackInst in = inst;
  float s[32];
#pragma unroll
  for (int k = 0; k < 32; k++) s[k] = 0.f;
  const int n = p.w * p.h;
  const size_t N0 = (size_t)p.W * p.H;
  TrackRow* dbg = (p.dbg && state < 2) ? p.dbg + ((size_t)state * p.n_levels + p.level) * 2 * N0 : nullptr;
  const float* nb = p.live + (size_t)PL_N * p.lstride;
  // The rendered depth is vertex2depth of the object's model vertex map, 0
  // at every raycast miss, and each coarser level's pixel is a robust
  // half-sample of its 2x2 finer pixels (0 where all four are 0).  So it is
  // 0 at pixels whose level-0 footprint lies outside the box of model-map
  // tiles with a surface (M.tb, in MAP_TILE tiles; empty: x0 > x1): their
  // photometric row ends at the depth test, and the load is skipped.
  const bool boxed = in.tbox != nullptr;
  const int l = p.level;
  const int stride = nactive * TB;
  auto lead = [&](int i, ObjLead& o) {
    const uint32_t ii = i < n ? (uint32_t)i : 0u;   // past the end: pixel 0, loaded and unused
    const int x = (int)ii % p.w, y = (int)ii / p.w;
    const int fx0 = x << l, fy0 = y << l, fx1 = fx0 + (1 << l) - 1, fy1 = fy0 + (1 << l) - 1;
    const bool rgb_out = boxed && (fx1 / MAP_TILE < M.tb[0] || fx0 / MAP_TILE > M.tb[2] ||
                                   fy1 / MAP_TILE < M.tb[1] || fy0 / MAP_TILE > M.tb[3]);
    o.c2v = lv3(p, PL_V, ii);
    o.nlx = ldx(nb, ii);
    o.dr = rgb_out ? 0.0f : ldx(in.rD + p.rd_off, ii);
  };
  int i = wg * TB + tid;
  ObjLead cur;
  lead(i, cur);
  for (; i < n; i += stride) {
    asm volatile("" ::: "memory");   // keep the LDS matrix reads inside the loop
    hold(cur.c2v); hold(cur.nlx); hold(cur.dr);
    ObjLead nxt;
    lead(i + stride, nxt);
    const int x = i % p.w, y = i / p.w;
    const size_t ro = (size_t)x + (size_t)y * p.W;
    {
      float e = 0.f, J[6] = {0, 0, 0, 0, 0, 0};
      const int res = obj_icp_rest(p, in, M, (size_t)i, cur.c2v, cur.nlx, e, J);
      accum<RB>(s, res, e, J, p.robust, p.robust_k);
      if (in.codes && p.level == 0) in.codes[ro] = (int8_t)res;
      if (dbg) { TrackRow& d = dbg[ro]; d.result = res; d.error = e; for (int q = 0; q < 6; q++) d.J[q] = J[q]; }
    }
    {
      float e = 0.f, J[6] = {0, 0, 0, 0, 0, 0};
      int res = -1;
      if (cur.dr > 0.0f) {
        const f3 vl = rf3(p, PL_V, (size_t)i);
        const float ir = ldx(p.ref + (size_t)PL_I * p.lstride, (uint32_t)i);
        hold(vl); hold(ir);
        res = rgb_core<true>(p, M, x, y, cur.dr, vl, ir, e, J);
      }
      accum<RB>(s, res, e, J, p.robust, p.robust_k);
      if (in.codes && p.level == 0) in.codes[N0 + ro] = (int8_t)res;
      if (in.rgb_err && p.level == 0) in.rgb_err[ro] = e;
      if (dbg) { TrackRow& d = dbg[N0 + ro]; d.result = res; d.error = e; for (int q = 0; q < 6; q++) d.J[q] = J[q]; }
    }
    cur = nxt;
  }
  wg_reduce(s, part, red);
}

// ---- FP64 6x6 solve (commons.h:672-696 semantics: TooN GR_SVD backsub with
// condition 1e6).  C is the symmetric JtWJ of the reduction, b its JtWe.
// Wave 0 runs Gauss-Jordan on [C | I | b], lane j holding column j in
// registers and taking the pivot column from lane k with v_readlane; when
// every pivot is positive and tr(C)*tr(C^-1) < 1e6 (so lambda_max/lambda_min
// < 1e6 and backsub would truncate nothing), x = C^-1 b.  Otherwise the wave
// computes the pseudo-inverse from a parallel Jacobi eigendecomposition with
// the reference's truncation rule.
__device__ __forceinline__ int tri(int i, int j) {   // packed upper-triangular index of the 21 JtWJ sums
  if (i > j) { const int t = i; i = j; j = t; }
  return 6 + 6 * i - (i * (i - 1)) / 2 + (j - i);
}
// value of v in lane l (l uniform)
__device__ __forceinline__ float rlf(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rld(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Pseudo-inverse x = sum_i [|l_i| * 1e6 > max |l|] (v_i . b / l_i) v_i of the
// symmetric C by cyclic Jacobi, on the 64 lanes of one wave: lane j (< 6)
// holds column j of A and of the eigenvector matrix V.  Each step applies
// three disjoint rotations A <- J^T A J, V <- V J; a round-robin schedule
// covers the 15 index pairs in 5 steps.  Disjoint rotations leave each
// other's (p,p), (q,q), (p,q) entries alone, so each angle is the one the
// serial order would compute at that point.  Column updates exchange a column
// with the partner lane, row updates are lane-local.  Sweeps stop when the
// off-diagonal mass is under 1e-30 of the diagonal's (eigenvalues at full
// FP64 accuracy) or after 40 sweeps.  x is uniform on return.
__device__ void jacobi_pinv(const double* s27, int lane, double (&x)[6]) {
  constexpr int RR[5][6] = {{0, 1, 2, 3, 4, 5}, {0, 2, 1, 4, 3, 5}, {0, 3, 1, 5, 2, 4},
                            {0, 4, 1, 3, 2, 5}, {0, 5, 1, 2, 3, 4}};
  const int j = lane < 6 ? lane : 5;
  double a[6], v[6];
#pragma unroll
  for (int i = 0; i < 6; i++) { a[i] = s27[tri(i, j)]; v[i] = (i == j) ? 1.0 : 0.0; }
  for (int sweep = 0; sweep < 40; sweep++) {
    double offl = 0, dl = 0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
      if (i < j) offl += a[i] * a[i];
      if (i == j) dl = a[i] * a[i];
    }
    double off = 0, dd = 0;
#pragma unroll
    for (int l = 0; l < 6; l++) { off += rld(offl, l); dd += rld(dl, l); }
    if (off <= 1e-30 * dd) break;
#pragma unroll
    for (int st = 0; st < 5; st++) {
      // this lane's pair: its index k, partner lane and role (p or q)
      int kk = 0, partner = j;
      bool isp = false;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int pp = RR[st][2 * k], qq = RR[st][2 * k + 1];
        if (j == pp) { kk = k; partner = qq; isp = true; }
        if (j == qq) { kk = k; partner = pp; isp = false; }
      }
      double app = 0, aqq = 0, apq = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int pp = RR[st][2 * k], qq = RR[st][2 * k + 1];
        const double e_pp = rld(a[pp], pp), e_qq = rld(a[qq], qq), e_pq = rld(a[pp], qq);
        if (kk == k) { app = e_pp; aqq = e_qq; apq = e_pq; }
      }
      double c = 1.0, sn = 0.0;
      if (apq != 0) {
        const double theta = (aqq - app) / (2 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
        c = 1 / sqrt(t * t + 1);
        sn = t * c;
      }
      // columns: A J and V J
#pragma unroll
      for (int i = 0; i < 6; i++) {
        const double pa = __shfl(a[i], partner), pv = __shfl(v[i], partner);
        a[i] = isp ? c * a[i] - sn * pa : sn * pa + c * a[i];
        v[i] = isp ? c * v[i] - sn * pv : sn * pv + c * v[i];
      }
      // rows: J^T (A J), every pair's rotation on this lane's column
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int pp = RR[st][2 * k], qq = RR[st][2 * k + 1];
        const double ck = rld(c, pp), sk = rld(sn, pp);
        const double ap = a[pp], aq = a[qq];
        a[pp] = ck * ap - sk * aq;
        a[qq] = sk * ap + ck * aq;
      }
    }
  }
  double lam = a[0];
#pragma unroll
  for (int i = 1; i < 6; i++) if (i == j) lam = a[i];
  double mx = 0;
#pragma unroll
  for (int l = 0; l < 6; l++) mx = fmax(mx, fabs(rld(lam, l)));
  // backsub keeps 1/lambda unless |lambda| * 1e6 <= max: a NaN eigenvalue
  // (a NaN system) is kept, so NaN sums give a NaN step as in the reference
  double pr = 0;
  if (!(fabs(lam) * 1e6 <= mx)) {
#pragma unroll
    for (int k = 0; k < 6; k++) pr += v[k] * s27[k];
    pr /= lam;
  }
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const double y = pr * v[k];
    double t = 0;
#pragma unroll
    for (int l = 0; l < 6; l++) t += rld(y, l);
    x[k] = t;
  }
}
// s27 = sums[1..27] (b then the packed JtWJ) in LDS; called by the 64 lanes of
// one wave, x uniform on return.  Returns 1 (Gauss-Jordan, certified) or 2
// (Jacobi pseudo-inverse).
__device__ int solve6(const double* s27, int lane, double (&x)[6]) {
  const int j = lane < 13 ? lane : 12;
  double m[6];
#pragma unroll
  for (int i = 0; i < 6; i++) m[i] = j < 6 ? s27[tri(i, j)] : (j < 12 ? (j - 6 == i ? 1.0 : 0.0) : s27[i]);
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    double ck[6];
#pragma unroll
    for (int i = 0; i < 6; i++) ck[i] = rld(m[i], k);
    const double pk = ck[k];
    ok = ok && (pk > 0.0);
    const double r = m[k] / pk;
#pragma unroll
    for (int i = 0; i < 6; i++) m[i] = (i == k) ? r : m[i] - ck[i] * r;
  }
  double trC = 0, trI = 0;
#pragma unroll
  for (int i = 0; i < 6; i++) { trC += s27[tri(i, i)]; trI += rld(m[i], 6 + i); }
  if (ok && trI > 0.0 && trC * trI < 1e6) {
#pragma unroll
    for (int i = 0; i < 6; i++) x[i] = rld(m[i], 12);
    return 1;
  }
  jacobi_pinv(s27, lane, x);
  return 2;
}
// inverse(Matrix4) of mf_math.h, column-parallel over lanes: lane l holds
// column (l & 7) of [A | I] (rows a[0..3]); on return lanes with (l & 4) hold
// the columns of the inverse.  Every element sees the operations of the
// serial form in the same order, so the result is bit-identical to it.
__device__ __forceinline__ void inverse_cols(float (&a)[4], int lane) {
  const int cc = lane & 7;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float ci[4];
#pragma unroll
    for (int r = 0; r < 4; r++) ci[r] = rlf(a[r], i);
    int am = i;
    float mx = fabsf(ci[i]);
#pragma unroll
    for (int ii = i + 1; ii < 4; ii++) { const float v = fabsf(ci[ii]); if (v > mx) { mx = v; am = ii; } }
    float piv = ci[i];
#pragma unroll
    for (int r = i + 1; r < 4; r++)
      if (r == am) { piv = ci[r]; ci[r] = ci[i]; const float t = a[i]; a[i] = a[r]; a[r] = t; }
    const float inv = 1.0f / piv;
    if (cc > i) a[i] = a[i] * inv;
#pragma unroll
    for (int u = i + 1; u < 4; u++)
      if (cc > i) a[u] = a[u] - ci[u] * a[i];
  }
#pragma unroll
  for (int i = 3; i > 0; i--) {
    float f[4];
#pragma unroll
    for (int u = 0; u < i; u++) f[u] = rlf(a[u], i);
#pragma unroll
    for (int u = 0; u < i; u++)
      if (cc >= 4) a[u] = a[u] - f[u] * a[i];
  }
}
// 1/n! for n = 0..20 (exact to the double's rounding of each quotient).
struct InvFact {
  double v[21];
  constexpr InvFact() : v() {
    double f = 1.0;
    for (int n = 0; n < 21; n++) {
      if (n > 0) f = f * n;
      v[n] = 1.0 / f;
    }
  }
};
__device__ constexpr InvFact kInvFact{};
// sum_{k=0..8} (-t)^k / (2k + o)!, Horner from k = 8.
__device__ __forceinline__ double alt_series(double t, int o) {
  double s = kInvFact.v[16 + o];
#pragma unroll
  for (int k = 7; k >= 0; k--) s = kInvFact.v[2 * k + o] - t * s;
  return s;
}
// SE3 exponential, x[0..2] translation, x[3..5] rotation (TooN SE3::exp as
// used at tracking.cpp:1361): Rodrigues + left Jacobian in FP64.
__device__ Mat4 se3_exp(const double x[6]) {
  const double w0 = x[3], w1 = x[4], w2 = x[5];
  const double th2 = w0 * w0 + w1 * w1 + w2 * w2;
  double A, B, C;
  if (th2 < 1e-8) { A = 1.0 - th2 / 6.0; B = 0.5 - th2 / 24.0; C = 1.0 / 6.0 - th2 / 120.0; }
  else if (th2 < 1e-6) { C = (1.0 / 6.0) * (1.0 - th2 / 20.0); A = 1.0 - th2 * C; B = 0.5 - th2 / 24.0; }
  else if (th2 < 0.25) {
    // the same three functions by their series in th2: sin(th)/th,
    // (1 - cos(th))/th2 and (1 - A)/th2 are sum_k (-th2)^k / (2k+1)!, (2k+2)!
    // and (2k+3)!; below 0.25 the terms past k = 8 are under 1e-22 of the
    // sums, which keeps sin, cos and three divisions off the solve's serial
    // path (and the cancellation in 1 - cos(th) out of B and C)
    A = alt_series(th2, 1); B = alt_series(th2, 2); C = alt_series(th2, 3);
  }
  else { const double th = sqrt(th2); A = sin(th) / th; B = (1.0 - cos(th)) / th2; C = (1.0 - A) / th2; }
  const double W[3][3] = {{0, -w2, w1}, {w2, 0, -w0}, {-w1, w0, 0}};
  double W2[3][3];
  #pragma unroll
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) { double s = 0; for (int k = 0; k < 3; k++) s += W[i][k] * W[k][j]; W2[i][j] = s; }
  Mat4 M = ident();
  #pragma unroll
  for (int i = 0; i < 3; i++) {
    #pragma unroll
    for (int j = 0; j < 3; j++) M.m[i][j] = (float)((i == j) + A * W[i][j] + B * W2[i][j]);
    double s = x[i];
    #pragma unroll
    for (int k = 0; k < 3; k++) s += (B * W[i][k] + C * W2[i][k]) * x[k];
    M.m[i][3] = (float)s;
  }
  return M;
}

// FP64 sum of `nparts` workgroup partials (32 floats each) into sums (LDS)
// and, when `out` is set, into out (device), by an NT-thread workgroup: U
// float4 loads in flight per thread and round, fixed order (deterministic);
// thread t always reads sums 4*(t%8)..4*(t%8)+3.  The caller issues round 0
// (load_round with q0 = 0) ahead of its own state loads, so both latencies
// overlap.
template <int NT, int U>
__device__ __forceinline__ void load_round(const float* __restrict__ part, int nparts, int q0, float4 (&v)[U]) {
  const float4* P = reinterpret_cast<const float4*>(part);
  const int nq = nparts * 8;
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int q = q0 + u * NT + (int)threadIdx.x;
    v[u] = q < nq ? P[q] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
}
template <int NT, int U>
__device__ __forceinline__ void accumulate_sums(const float* __restrict__ part, int nparts, float4 (&v)[U],
                                                double* sums, double* out) {
  __shared__ double red[NT / 64][8][4];
  const int tid = threadIdx.x;
  const int nq = nparts * 8;
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (int q0 = 0;;) {
#pragma unroll
    for (int u = 0; u < U; u++) { a0 += (double)v[u].x; a1 += (double)v[u].y; a2 += (double)v[u].z; a3 += (double)v[u].w; }
    q0 += U * NT;
    if (q0 >= nq) break;
    load_round<NT, U>(part, nparts, q0, v);
  }
  // lanes with equal (lane % 8) hold the same four sums
#pragma unroll
  for (int o = 8; o < 64; o <<= 1) {
    a0 += __shfl_xor(a0, o); a1 += __shfl_xor(a1, o); a2 += __shfl_xor(a2, o); a3 += __shfl_xor(a3, o);
  }
  const int lane = tid & 63;
  if (lane < 8) { red[tid >> 6][lane][0] = a0; red[tid >> 6][lane][1] = a1; red[tid >> 6][lane][2] = a2; red[tid >> 6][lane][3] = a3; }
  __syncthreads();
  if (tid < 32) {
    double t = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) t += red[w][tid >> 2][tid & 3];
    sums[tid] = t;
    if (out) out[tid] = t;
  }
  __syncthreads();
}

// Per-run constants of an instance that the solve needs, read with the
// iteration state at kernel start.
struct SolvePre {
  Mat4 Twc2;
  Mat4* out_pose;
  int object;
};
__device__ __forceinline__ void load_pre(SolvePre& pre, const GNState& S) {
  pre.Twc2 = S.Twc2; pre.out_pose = S.out_pose; pre.object = S.object;
}

// One GN step of trackLiveFrame / trackEachObject (tracking.cpp:300-334,
// 664-688) for the rows computed with X's pose at (X.pend_level,
// X.pend_iter), from the FP64 sums in LDS, by the 64 lanes of wave 0: lanes
// 0-3 hold the columns of T and previous_pose, lanes 4-7 those of the
// inverse; only the `writer` workgroup publishes the outputs (pose, trace).
__device__ void gn_update(GNIter& X, GNState& S, const SolvePre& pre, const double* sums, int state,
                          bool writer, const TrackSync& sy, uint64_t t0) {
  const int lane = threadIdx.x;
  const uint64_t t1 = wall_clock64();
  const int level = X.pend_level, iter = X.pend_iter;
  const float s0 = (float)sums[0], s28 = (float)sums[28];
  const float err = s0 / s28;
  const float prev_err = (iter == 0) ? INFINITY : X.prev_err;
  // 3 = no inliers (objects), 2 = error grew (step back), 0 = solve
  const int mode = (pre.object && s28 == 0) ? 3 : (err > prev_err ? 2 : 0);
  const int c = lane & 3;
  float t[4], pv[4];
#pragma unroll
  for (int r = 0; r < 4; r++) { t[r] = X.T.m[r][c]; pv[r] = X.prev.m[r][c]; }
  int flag = mode, done = X.done_level, solver = 0;
  float perr = X.prev_err;
  bool newT = false;
  if (mode == 3) {
    done = level;
  } else if (mode == 2) {
#pragma unroll
    for (int r = 0; r < 4; r++) t[r] = pv[r];
    newT = true;
    done = level;
  } else {
    perr = err;
    double x[6];
    solver = solve6(sums + 1, lane, x);
    const Mat4 upd = se3_exp(x);
    const double nx = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
    // icp_threshold is a float constant (tracking.h:387) compared with the
    // double norm(x) (tracking.cpp:1365): its value is (double)1e-5f
    if (nx < (double)1e-5f && (double)sqrtf(s0 / s28) < 2e-3) {
      done = level;
      flag = 1;
    } else {
      // previous_pose = T; T = upd * previous_pose (commons.h:657-662 operand order)
#pragma unroll
      for (int r = 0; r < 4; r++) pv[r] = t[r];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        float s = upd.m[r][0] * pv[0];
        s = s + upd.m[r][1] * pv[1];
        s = s + upd.m[r][2] * pv[2];
        s = s + upd.m[r][3] * pv[3];
        t[r] = s;
      }
      newT = true;
    }
  }
  if (newT) {
    float a[4];
#pragma unroll
    for (int r = 0; r < 4; r++) a[r] = (lane & 4) ? (r == c ? 1.f : 0.f) : t[r];
    inverse_cols(a, lane);
    if (lane >= 4 && lane < 8) {
#pragma unroll
      for (int r = 0; r < 4; r++) X.invT.m[r][c] = a[r];
    }
  }
  if (lane < 4) {
#pragma unroll
    for (int r = 0; r < 4; r++) { X.T.m[r][c] = t[r]; X.prev.m[r][c] = pv[r]; }
  }
  if (lane == 0) { X.prev_err = perr; X.done_level = done; X.pend_level = -1; }
  if (!writer) return;
  if (lane < 4) {
    // camera: T_w_l; object: volume_pose_ = T_w_c2 * T_c2_o2 (tracking.cpp:695)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      S.T_final.m[r][c] = t[r];
      float o = t[r];
      if (pre.object) {
        o = pre.Twc2.m[r][0] * t[0];
        o = o + pre.Twc2.m[r][1] * t[1];
        o = o + pre.Twc2.m[r][2] * t[2];
        o = o + pre.Twc2.m[r][3] * t[3];
      }
      pre.out_pose->m[r][c] = o;
    }
  }
  if (sy.trace) {
    const uint64_t t2 = wall_clock64();
    int ti = 0;
    if (lane == 0) ti = atomicAdd(sy.trace_n, 1);
    ti = __shfl(ti, 0);
    if (ti < sy.trace_cap) {
      TraceRec& rec = sy.trace[ti];
      if (lane < 4) {
#pragma unroll
        for (int r = 0; r < 4; r++) rec.pose[4 * r + c] = t[r];
      }
      if (lane == 0) {
        rec.inst = state; rec.level = level; rec.iter = iter; rec.flag = flag; rec.err = err; rec.s28 = s28;
        rec.t_accum = (unsigned)(t1 - t0); rec.t_solve = (unsigned)(t2 - t1);
        rec.solver = solver;
      }
    }
  }
}

template <int NT, int U>
__device__ __forceinline__ void gn_solve(GNIter& X, GNState& S, const SolvePre& pre, int state,
                                         const float* __restrict__ part, int nparts, float4 (&v)[U], bool writer,
                                         const TrackSync& sy) {
  __shared__ double sums[32];
  const uint64_t t0 = wall_clock64();
  accumulate_sums<NT, U>(part, nparts, v, sums, writer ? S.sums : nullptr);
  if (threadIdx.x < 64) gn_update(X, S, pre, sums, state, writer, sy, t0);
  __syncthreads();
}

// Tracking state is double-buffered by launch index q: a launch reads
// GNState::it[q & 1] and the solve step of iteration q writes it[(q+1) & 1],
// so no workgroup ever reads a slot that another workgroup of the same
// launch writes.

// Moving flags of the labels as one mask (wave 0; tid < 64).
__device__ __forceinline__ void stage_moving(const TrackLaunch& p, Mats& M, int tid) {
  const bool mv = p.label && tid < p.n_labels && p.moving[tid] != 0;
  const uint64_t b = __ballot(mv);
  if (tid == 0) M.mov = b;
}

// Chunk of rows taken by workgroup b of n, XCD-aware.  Workgroups are
// dispatched to the 8 XCDs round-robin by linear id and each XCD has its own
// L2; with n a multiple of 8 the linear id of (b, instance) is b modulo 8, and
// the workgroups of one XCD take consecutive chunks, so the image rows that
// neighbouring chunks share (bilinear taps, projected gathers) are fetched
// into one L2.  The partial of a chunk is written at the chunk's index, so
// the sums and their order do not depend on the mapping.
__device__ __forceinline__ int xcd_chunk(int b, int n) {
  if (n & 7) return b;
  return (b & 7) * (n >> 3) + (b >> 3);
}

// Two-kernel form, rows: GN iteration q of every instance (grid.y =
// instance), one partial of 32 FP32 sums per workgroup.
template <bool ICP, bool RGB, bool OBJ, int RB>
__global__ __launch_bounds__(TB) void k_track(TrackLaunch p, const TrackInst* __restrict__ insts,
                                              const GNState* __restrict__ st, float* __restrict__ partial,
                                              int nblk, int q) {
  __shared__ float red[NW][32];
  __shared__ Mats Msh;
  __shared__ int skip;
  const int tid = threadIdx.x;
  const int j = blockIdx.y;
  const int state = insts[j].state;
  const GNState& S = st[state];
  if (tid == 0) {
    const GNIter& X = S.it[q & 1];
    // a level starts afresh (tracking.cpp:246); later iterations are skipped
    // once this level's loop has broken (uniform for the instance)
    skip = S.off || (p.iter > 0 && X.done_level == p.level);
    if (!skip) {
      Msh.T = X.T; Msh.invT = X.invT; Msh.A = S.A; Msh.B = S.B; Msh.C = S.C;
      if (OBJ && insts[j].tbox) {
        const int* tb = insts[j].tbox;
        Msh.tb[0] = tb[0]; Msh.tb[1] = tb[1]; Msh.tb[2] = tb[2]; Msh.tb[3] = tb[3];
      }
    }
  }
  if (tid < 64) stage_moving(p, Msh, tid);
  __syncthreads();
  if (skip) return;
  // object rows cluster where the object is: they keep the round-robin
  // spread over the XCDs, which balances them better than L2 locality pays
  const int c = OBJ ? (int)blockIdx.x : xcd_chunk(blockIdx.x, nblk);
  if constexpr (ICP && RGB && OBJ) {
    if (p.paired) { object_pixels<RB>(p, insts[j], Msh, state, c, nblk, partial + ((size_t)j * nblk + c) * 32, red); return; }
  }
  track_rows<ICP, RGB, OBJ, RB>(p, insts[j], Msh, state, c, nblk, partial + ((size_t)j * nblk + c) * 32, red);
}

// Two-kernel form, camera rows by pixel (camera_pixels, one pixel per loop
// iteration): the paired variant of k_track<true, true, false, RB>.
template <int RB>
__global__ __launch_bounds__(TB) void k_track_px(TrackLaunch p, const TrackInst* __restrict__ insts,
                                                 const GNState* __restrict__ st, float* __restrict__ partial,
                                                 int nblk, int q) {
  __shared__ float red[NW][32];
  __shared__ Mats Msh;
  __shared__ int skip;
  const int tid = threadIdx.x;
  const int j = blockIdx.y;
  const int state = insts[j].state;
  const GNState& S = st[state];
  if (tid == 0) {
    const GNIter& X = S.it[q & 1];
    skip = S.off || (p.iter > 0 && X.done_level == p.level);
    if (!skip) { Msh.T = X.T; Msh.invT = X.invT; Msh.A = S.A; Msh.B = S.B; Msh.C = S.C; }
  }
  if (tid < 64) stage_moving(p, Msh, tid);
  __syncthreads();
  if (skip) return;
  const int c = xcd_chunk(blockIdx.x, nblk);
  camera_pixels<RB>(p, insts[j], Msh, state, c, nblk, partial + ((size_t)j * nblk + c) * 32, red);
}

// Two-kernel form, rows of every moving object in one grid over the pixels
// (trackEachObject's rows, tracking.cpp:392-578, for all objects at once):
// each thread takes PPT pixels TB apart, loads their live vertex, normal and
// covariance and their reference vertex and intensity once, then for each
// object computes the object's ICP row of the live pixel and photometric row
// of the reference pixel from those values, and writes the object's partial
// of this workgroup at partial[(object * nblk + workgroup) * 32].  Only the
// rendered depth and, near the object's silhouette, the model-map taps are
// per-object loads.  Per-object pose matrices sit in dynamic LDS.
template <bool ICP, bool RGB, int PPT, int RB>
__global__ __launch_bounds__(TB) void k_track_objs(TrackLaunch p, const TrackInst* __restrict__ insts, int ninst,
                                                   const GNState* __restrict__ st, float* __restrict__ partial,
                                                   int nblk, int q) {
  extern __shared__ __align__(16) unsigned char mf_dyn[];
  Mats* Mo = reinterpret_cast<Mats*>(mf_dyn);
  __shared__ float red[NW][32];
  __shared__ int skip[MAX_INST];
  const int tid = threadIdx.x;
  for (int j = tid; j < ninst; j += TB) {
    const GNState& S = st[insts[j].state];
    const GNIter& X = S.it[q & 1];
    // a level starts afresh (tracking.cpp:246); later iterations are skipped
    // once this object's loop at this level has broken
    const int sk = S.off || (p.iter > 0 && X.done_level == p.level);
    skip[j] = sk;
    if (!sk) { Mo[j].T = X.T; Mo[j].invT = X.invT; Mo[j].A = S.A; Mo[j].B = S.B; Mo[j].C = S.C; Mo[j].mov = 0; }
  }
  __syncthreads();
  const int n = p.w * p.h;
  const size_t N0 = (size_t)p.W * p.H;
  int pix[PPT];
  f3 nl[PPT], vl[PPT], cv[PPT], rv[PPT];
  float ir[PPT];
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    const int i = blockIdx.x * TB * PPT + k * TB + tid;
    pix[k] = i < n ? i : -1;
    const size_t ii = i < n ? (size_t)i : 0;
    if (ICP) { nl[k] = lv3(p, PL_N, ii); vl[k] = lv3(p, PL_V, ii); cv[k] = lv3(p, PL_COV, ii); }
    if (RGB) { rv[k] = rf3(p, PL_V, ii); ir[k] = ldx(p.ref + (size_t)PL_I * p.lstride, (uint32_t)ii); }
  }
  for (int j = 0; j < ninst; j++) {
    if (skip[j]) continue;   // uniform over the workgroup
    const TrackInst in = insts[j];
    const Mats& M = Mo[j];
    TrackRow* dbg = (p.dbg && in.state < 2) ? p.dbg + ((size_t)in.state * p.n_levels + p.level) * 2 * N0 : nullptr;
    float s[32];
#pragma unroll
    for (int k = 0; k < 32; k++) s[k] = 0.f;
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      asm volatile("" ::: "memory");   // keep the LDS matrix reads inside the loop
      const int i = pix[k];
      if (i < 0) continue;
      const int x = i % p.w, y = i / p.w;
      const size_t ro = (size_t)x + (size_t)y * p.W;   // row stride W at every level (tracking.cpp:49-50)
      if (ICP) {
        float e = 0.f, J[6] = {0, 0, 0, 0, 0, 0};
        const int res = obj_icp_core(p, in, M, nl[k], vl[k], cv[k], e, J);
        accum<RB>(s, res, e, J, p.robust, p.robust_k);
        if (in.codes && p.level == 0) in.codes[ro] = (int8_t)res;
        if (dbg) { TrackRow& d = dbg[ro]; d.result = res; d.error = e; for (int u = 0; u < 6; u++) d.J[u] = J[u]; }
      }
      if (RGB) {
        float e = 0.f, J[6] = {0, 0, 0, 0, 0, 0};
        const float dr = ldx(in.rD + p.rd_off, (uint32_t)i);
        const int res = (dr > 0.0f) ? rgb_core<true>(p, M, x, y, dr, rv[k], ir[k], e, J) : -1;
        accum<RB>(s, res, e, J, p.robust, p.robust_k);
        if (in.codes && p.level == 0) in.codes[N0 + ro] = (int8_t)res;
        if (in.rgb_err && p.level == 0) in.rgb_err[ro] = e;
        if (dbg) { TrackRow& d = dbg[N0 + ro]; d.result = res; d.error = e; for (int u = 0; u < 6; u++) d.J[u] = J[u]; }
      }
    }
    wg_reduce(s, partial + ((size_t)j * nblk + blockIdx.x) * 32, red);
  }
}

// Two-kernel form, solve: the iteration's FP64 accumulation, solve and GN
// control, one workgroup per instance, after k_track (the kernel boundary
// publishes the partials; a device-scope fence in every track workgroup costs
// far more on gfx950).  SOLVE_U float4 loads per thread cover the 600
// partials of a 640x480 level in one round.
static constexpr int SOLVE_NT = 512;    // threads of the solve workgroup (partials in flight)
static constexpr int SOLVE_U = 10;
__global__ __launch_bounds__(SOLVE_NT) void k_solve(int level, int iter, const TrackInst* __restrict__ insts,
                                              GNState* __restrict__ st, const float* __restrict__ partial,
                                              int nblk, int q, TrackSync sy) {
  __shared__ GNIter X;
  __shared__ SolvePre pre;
  __shared__ int skip;
  const int tid = threadIdx.x;
  const int j = blockIdx.x;
  const float* part = partial + (size_t)j * nblk * 32;
  float4 v[SOLVE_U];
  load_round<SOLVE_NT, SOLVE_U>(part, nblk, 0, v);
  const int state = insts[j].state;
  GNState& S = st[state];
  if (tid == 0) {
    X = S.it[q & 1];
    load_pre(pre, S);
    skip = S.off || (iter > 0 && X.done_level == level);
    if (skip) S.it[(q + 1) & 1] = X;
    X.pend_level = level; X.pend_iter = iter;
  }
  __syncthreads();
  if (skip) return;
  gn_solve<SOLVE_NT, SOLVE_U>(X, S, pre, state, part, nblk, v, true, sy);
  if (tid == 0) S.it[(q + 1) & 1] = X;
}

// Fused form, one launch per iteration: every workgroup first finishes
// iteration q-1 itself from it[q & 1] (rows pending) and iteration q-1's
// partials (the same solve on the same data in every workgroup, so all reach
// bit-identical states), then computes iteration q's rows; workgroup 0
// stores the state, with these rows pending, in it[(q+1) & 1].  With
// rows == 0 the launch only finishes iteration q-1.  Each workgroup reads all
// of the previous partials, so this form pays nblk^2 partial reads: it suits
// levels with few workgroups.
template <bool ICP, bool RGB, bool OBJ, int RB>
__global__ __launch_bounds__(TB) void k_gn_iter(TrackLaunch p, const TrackInst* __restrict__ insts,
                                                GNState* __restrict__ st, const float* __restrict__ part_prev,
                                                int nblk_prev, float* __restrict__ part_cur, int nblk, int q,
                                                int rows, TrackSync sy) {
  __shared__ float red[NW][32];
  __shared__ Mats Msh;
  __shared__ GNIter X;
  __shared__ SolvePre pre;
  __shared__ int skip;
  const int tid = threadIdx.x;
  const int j = blockIdx.y;
  const float* prev = part_prev + (size_t)j * nblk_prev * 32;
  float4 v[8];
  load_round<TB, 8>(prev, nblk_prev, 0, v);
  const int state = insts[j].state;
  GNState& S = st[state];
  const bool writer = blockIdx.x == 0;
  if (tid == 0) { X = S.it[q & 1]; load_pre(pre, S); }
  __syncthreads();
  if (X.pend_level >= 0) gn_solve<TB, 8>(X, S, pre, state, prev, nblk_prev, v, writer, sy);
  if (tid == 0) {
    skip = S.off || !rows || (p.iter > 0 && X.done_level == p.level);
    if (!skip) {
      Msh.T = X.T; Msh.invT = X.invT; Msh.A = S.A; Msh.B = S.B; Msh.C = S.C;
      if (OBJ && insts[j].tbox) {
        const int* tb = insts[j].tbox;
        Msh.tb[0] = tb[0]; Msh.tb[1] = tb[1]; Msh.tb[2] = tb[2]; Msh.tb[3] = tb[3];
      }
      X.pend_level = p.level; X.pend_iter = p.iter;
    }
    if (writer) S.it[(q + 1) & 1] = X;
  }
  if (tid < 64) stage_moving(p, Msh, tid);
  __syncthreads();
  if (skip) return;
  float* part = part_cur + ((size_t)j * nblk + blockIdx.x) * 32;
  if constexpr (ICP && RGB && !OBJ) {
    if (p.paired) { camera_pixels<RB>(p, insts[j], Msh, state, blockIdx.x, nblk, part, red); return; }
  }
  if constexpr (ICP && RGB && OBJ) {
    if (p.paired) { object_pixels<RB>(p, insts[j], Msh, state, blockIdx.x, nblk, part, red); return; }
  }
  track_rows<ICP, RGB, OBJ, RB>(p, insts[j], Msh, state, blockIdx.x, nblk, part, red);
}

// Residual evaluation at a given pose (forward/backward of the tracking
// residual): the state's pose is set to T, rows are computed by k_track, and
// k_eval_sums writes the instance's 32 FP64 sums to S.sums.
__global__ void k_eval_set(GNState* st, int state, Mat4 T) {
  st[state].off = 0;
  GNIter& X = st[state].it[0];
  X.T = T; X.invT = inverse(T); X.prev = T;
  X.done_level = -1; X.pend_level = -1; X.pend_iter = -1;
}
__global__ __launch_bounds__(SOLVE_NT) void k_eval_sums(GNState* st, int state, const float* __restrict__ partial,
                                                        int nblk) {
  __shared__ double sums[32];
  float4 v[SOLVE_U];
  load_round<SOLVE_NT, SOLVE_U>(partial, nblk, 0, v);
  accumulate_sums<SOLVE_NT, SOLVE_U>(partial, nblk, v, sums, st[state].sums);
}

__device__ void gn_iter_init(GNIter& X, const Mat4& T) {
  X.T = T; X.invT = inverse(T); X.prev = T;
  X.prev_err = INFINITY; X.done_level = -1; X.pend_level = -1; X.pend_iter = -1;
}
__device__ void gn_init_camera(GNState& S, Mat4* T_w_l, const Mat4& Twr, float fx, float fy, float cx, float cy) {
  gn_iter_init(S.it[0], *T_w_l);
  S.T_final = *T_w_l;
  S.A = Twr;
  S.B = mul(camK(fx, fy, cx, cy), inverse(Twr));
  S.C = ident(); S.T_wo1 = ident();
  S.object = 0; S.off = 0;
  S.out_pose = T_w_l; S.Twc2 = ident();
}
// Tracking frame start: T_w_r = camera pose (kfusion.cpp:243), its inverse for
// the rendered reference, and the camera's GN state.
__global__ void k_camera_begin(Mat4* mats, GNState* st, float fx, float fy, float cx, float cy) {
  const Mat4 P = mats[0];
  mats[1] = P;
  mats[2] = inverse(P);
  gn_init_camera(st[0], mats + 0, P, fx, fy, cx, cy);
}
__global__ void k_gn_begin_camera(GNState* st, Mat4* T_w_l, const Mat4* T_w_r, const int* gate, float fx, float fy,
                                  float cx, float cy) {
  gn_init_camera(*st, T_w_l, *T_w_r, fx, fy, cx, cy);
  st->off = (gate && *gate == 0) ? 1 : 0;
}
// trackEachObject setup (tracking.cpp:587-596): T_c1_o1 and the initial
// T_c2_o2, one thread per object.
__global__ void k_gn_begin_objects(GNState* st, int first, int n, const int* ids, Mat4* poses, const Mat4* T_w_c2,
                                   const Mat4* T_w_c1, int publish, const int* gate) {
  const int j = threadIdx.x;
  if (j >= n) return;
  GNState& S = st[first + j];
  Mat4* pose = poses + ids[j];
  const Mat4 To = *pose;
  S.T_wo1 = To;
  S.A = mul(inverse(*T_w_c1), To);
  const Mat4 T = mul(inverse(*T_w_c2), To);
  gn_iter_init(S.it[0], T);
  S.T_final = T;
  S.B = inverse(To);
  S.C = inverse(S.A);
  S.object = 1;
  S.off = (gate && *gate == 0) ? 1 : 0;
  S.out_pose = pose; S.Twc2 = *T_w_c2;
  if (publish && !S.off) *pose = mul(*T_w_c2, T);
}

// ---- motion residuals (Tracking::compute_residuals, tracking.cpp:1535-1747),
// the per-object inlier counts of check_static_state (rendering.cpp:907-947)
// and the moving decision and gating of Kfusion::tracking (kfusion.cpp:
// 286-471).

// Instance sets (split_labels, kfusion.cpp:1527-1553) are flags by label,
// flags[l] = 1 for each label l in [0, MAX_INST) of an image, written by the
// raycast that writes the image (k_raycast's `present`).
__device__ int count_flags(const int* f, int nobj) {
  int c = 0;
  for (int j = 0; j < nobj; j++) c += f[j] != 0;
  return c;
}
// Frame start: the motion state cleared, then gate = more than one instance
// in the last raycast (kfusion.cpp:292); under it the in-view set is cleared
// before the raycast of every object (:294).
__global__ void k_motion_gate(MotionDev* md, int nobj) {
  for (int j = threadIdx.x; j < MAX_INST; j += blockDim.x) {
    md->present[j] = 0; md->move_set[j] = 0; md->counts[2 * j] = 0; md->counts[2 * j + 1] = 0;
  }
  const int g = count_flags(md->present_main, nobj) > 1 ? 1 : 0;
  if (threadIdx.x == 0) { md->gate = g; md->move_any = 0; md->retrack = 0; }
  if (g)
    for (int j = threadIdx.x; j < MAX_INST; j += blockDim.x) md->in_view[j] = 0;
}
__global__ void k_motion_inview(MotionDev* md, int nobj, const uint8_t* moving, int manual) {
  if (!md->gate) return;
  const int rn = count_flags(md->in_view, nobj);
  md->rendered_n = rn;
  if (!manual) return;
  int any = 0;
  for (int j = 1; j < nobj; j++) { md->move_set[j] = moving[j]; any |= moving[j]; }
  md->move_any = any;
  md->retrack = (any && rn != 0) ? 1 : 0;
}
__global__ void k_motion_classify(MotionDev* md, int nobj, uint8_t* moving) {
  if (!md->gate) return;
  int any = 0;
  for (int j = 1; j < nobj; j++) {
    const int px = md->counts[2 * j], inl = md->counts[2 * j + 1];
    if (px == 0) continue;
    const float ratio = (float)inl / (float)px;
    const uint8_t mv = (ratio < 0.8f && ratio > 0.05f) ? 1 : 0;
    moving[j] = mv; md->move_set[j] = mv; any |= mv;
  }
  md->move_any = any;
  md->retrack = (any && md->rendered_n != 0) ? 1 : 0;
}
__global__ void k_motion_after(MotionDev* md, int nobj, int use_rgb) {
  if (md->move_any) md->rendered_n = count_flags(md->present, nobj);
  md->cond2 = ((md->rendered_n > 1 && use_rgb) || md->move_any) ? 1 : 0;
  md->copy_static = (md->gate || md->cond2) ? 1 : 0;
}
// virtual_camera_pose_ = T_wo1 * inverse(T_c2_o2) (tracking.cpp:696)
__global__ void k_virtual_cams(const GNState* st, int first, int n, const int* ids, Mat4* vcam) {
  const int j = threadIdx.x;
  if (j >= n) return;
  const GNState& S = st[first + j];
  if (S.off) return;
  vcam[ids[j] - 3] = mul(S.T_wo1, inverse(S.T_final));
}
__global__ void k_copy_codes(const int8_t* __restrict__ src, int8_t* __restrict__ dst, size_t n,
                             const float* __restrict__ esrc, float* __restrict__ edst, size_t ne, const int* gate) {
  if (!*gate) return;
  const size_t n4 = n / 4;
  const int* s4 = reinterpret_cast<const int*>(src);
  int* d4 = reinterpret_cast<int*>(dst);
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  for (size_t i = t; i < n4; i += st) d4[i] = s4[i];
  if (t < n - 4 * n4) dst[4 * n4 + t] = src[4 * n4 + t];
  for (size_t i = t; i < ne; i += st) edst[i] = esrc[i];
}
__device__ __forceinline__ void motion_rows(const MotionLaunch& m) {
  if (m.gate && *m.gate == 0) return;
  __shared__ int hist[2 * MAX_INST];
  const bool count = m.label_bf != nullptr;
  if (count) {
    for (int j = threadIdx.x; j < 2 * MAX_INST; j += blockDim.x) hist[j] = 0;
    __syncthreads();
  }
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x < m.W && y < m.H) {
    const size_t n = (size_t)m.W * m.H;
    const size_t i = (size_t)x + (size_t)y * m.W;
    const Mat4 Tl = *m.T_w_l, iTr = *m.inv_T_w_r;
    const Mat4 K = camK(m.fx, m.fy, m.cx, m.cy);
    const float* L = m.live;
    int ri = 0, gi = 0;
    float ge = 0.f;   // photometric row error (TrackData.error of a result-1 row)
    if (m.use_icp) {
      // compute_icp_residual_kernel (tracking.cpp:1617-1673)
      const f3 nl = mk3(L[(size_t)PL_N * n + i], L[(size_t)(PL_N + 1) * n + i], L[(size_t)(PL_N + 2) * n + i]);
      if (nl.x == INVALID) ri = -1;
      else {
        const f3 vl = mk3(L[(size_t)PL_V * n + i], L[(size_t)(PL_V + 1) * n + i], L[(size_t)(PL_V + 2) * n + i]);
        const f3 wv = affine(Tl, vl);
        const f3 rv = affine(iTr, wv);
        const f3 pj = rotate(K, rv);
        const float ux = pj.x / pj.z, uy = pj.y / pj.z;
        if (ux < 1 || ux > (float)(m.W - 2) || uy < 1 || uy > (float)(m.H - 2) || isnan(ux) || isnan(uy)) ri = -2;
        else {
          const size_t q = (unsigned)ux + (size_t)((unsigned)uy) * m.W;
          const f3 rn = mk3(m.mvn[3 * n + q], m.mvn[4 * n + q], m.mvn[5 * n + q]);
          if (rn.x == INVALID) ri = -3;
          else {
            const f3 diff = mk3(m.mvn[q], m.mvn[n + q], m.mvn[2 * n + q]) - wv;
            const f3 wn = rotate(Tl, nl);
            if (length(diff) > 0.1f) ri = -4;
            else if (dot(wn, rn) < 0.8f) ri = -5;
            else {
              const f3 P = mk3(L[(size_t)PL_COV * n + i], L[(size_t)(PL_COV + 1) * n + i], L[(size_t)(PL_COV + 2) * n + i]);
              const double sig = (double)(rn.x * rn.x * P.x + rn.y * rn.y * P.y + rn.z * rn.z * P.z);
              const double err = sqrt(1.0 / sig) * (double)dot(rn, diff);
              ri = fabs(err) > (double)m.threshold ? -4 : 1;
            }
          }
        }
      }
    }
    if (m.use_rgb) {
      // compute_rgb_residual_kernel (tracking.cpp:1676-1747); the reference
      // vertex is vertex2depth of the map before fusion at T_w_r,
      // back-projected (tracking.cpp:1546-1547)
      const f3 vb = mk3(m.vbf[i], m.vbf[n + i], m.vbf[2 * n + i]);
      const float d = vb.z > 0.0f ? affine(iTr, vb).z : 0.0f;
      const f3 vr = d > 0.0f ? d * rotate(camKinv(m.fx, m.fy, m.cx, m.cy), mk3((float)x, (float)y, 1.0f))
                             : mk3(0.f, 0.f, 0.f);
      const f3 vl = mk3(L[(size_t)PL_V * n + i], L[(size_t)(PL_V + 1) * n + i], L[(size_t)(PL_V + 2) * n + i]);
      if (vr.z <= 0.f || vr.z == INVALID) gi = -1;
      else if (length(vr - vl) > 0.1f && vl.z > 0.f) gi = -5;
      else {
        const f3 wv = affine(Tl, vr);
        const f3 rv = affine(iTr, wv);
        const f3 pj = rotate(K, rv);
        const float ux = pj.x / pj.z, uy = pj.y / pj.z;
        if (ux < 1 || ux > (float)(m.W - 2) || uy < 1 || uy > (float)(m.H - 2) || isnan(ux) || isnan(uy)) gi = -2;
        else {
          const float residual = L[(size_t)PL_I * n + i] - bilin(m.ref + (size_t)PL_I * n, m.W, ux, uy);
          const double err = (1.0 / (double)m.sigma_bright) * (double)residual;
          gi = fabs(err) > (double)m.threshold ? -4 : 1;
          ge = gi == 1 ? (float)err : 0.f;
        }
      }
    }
    m.codes[i] = (int8_t)ri;
    m.codes[n + i] = (int8_t)gi;
    if (m.rgb_err) m.rgb_err[i] = ge;
    if (count) {
      const int l = m.label_bf[i];
      if (l > 0 && l < MAX_INST) {
        bool inl;
        if (m.use_icp && m.use_rgb) inl = ri > -3 && gi > -3;
        else if (m.use_icp) inl = ri > -3;
        else inl = gi > -3;
        atomicAdd(&hist[2 * l], 1);
        if (inl) atomicAdd(&hist[2 * l + 1], 1);
      }
    }
  }
  if (count) {
    __syncthreads();
    for (int j = threadIdx.x; j < 2 * MAX_INST; j += blockDim.x)
      if (hist[j]) atomicAdd(&m.counts[j], hist[j]);
  }
}
__global__ __launch_bounds__(256) void k_motion_rows(MotionLaunch m) { motion_rows(m); }
// Several instances' rows in one launch (grid.z = instance): `base` with the
// instance's own maps, pose and outputs.  Their outputs are disjoint, so this
// equals one launch per instance.
__global__ __launch_bounds__(256) void k_motion_rows_n(MotionLaunch base, const MotionObj* __restrict__ objs) {
  MotionLaunch m = base;
  const MotionObj o = objs[blockIdx.z];
  m.mvn = o.mvn; m.vbf = o.vbf; m.T_w_l = o.T_w_l; m.codes = o.codes; m.rgb_err = o.rgb_err;
  motion_rows(m);
}

// ------------------------------------------------------------ launchers
int track_blocks(int nrows, int rows_per_thread) {
  if (rows_per_thread < 1) rows_per_thread = 1;
  const int nb = (nrows + TB * rows_per_thread - 1) / (TB * rows_per_thread);
  return nb < 1 ? 1 : nb;
}
void launch_track_rows(hipStream_t s, const TrackLaunch& p, const TrackInst* inst, int ninst, GNState* st,
                       float* partial, int nblk, int q) {
  dim3 grd(nblk, ninst), blk(TB);
  if (p.paired && !p.object && p.use_icp && p.use_rgb) {
    if (p.robust == ROBUST_CAUCHY) k_track_px<ROBUST_CAUCHY><<<grd, blk, 0, s>>>(p, inst, st, partial, nblk, q);
    else k_track_px<ROBUST_ANY><<<grd, blk, 0, s>>>(p, inst, st, partial, nblk, q);
    return;
  }
#define MF_TRACK(RB)                                                                                   \
  if (p.object) {                                                                                      \
    if (p.use_icp && p.use_rgb) k_track<true, true, true, RB><<<grd, blk, 0, s>>>(p, inst, st, partial, nblk, q); \
    else if (p.use_icp) k_track<true, false, true, RB><<<grd, blk, 0, s>>>(p, inst, st, partial, nblk, q);        \
    else k_track<false, true, true, RB><<<grd, blk, 0, s>>>(p, inst, st, partial, nblk, q);                       \
  } else {                                                                                             \
    if (p.use_icp && p.use_rgb) k_track<true, true, false, RB><<<grd, blk, 0, s>>>(p, inst, st, partial, nblk, q); \
    else if (p.use_icp) k_track<true, false, false, RB><<<grd, blk, 0, s>>>(p, inst, st, partial, nblk, q);        \
    else k_track<false, true, false, RB><<<grd, blk, 0, s>>>(p, inst, st, partial, nblk, q);                       \
  }
  if (p.robust == ROBUST_CAUCHY) { MF_TRACK(ROBUST_CAUCHY) } else { MF_TRACK(ROBUST_ANY) }
#undef MF_TRACK
}
int track_obj_blocks(int npix, int pixels_per_thread) {
  const int ppt = pixels_per_thread < 1 ? 1 : (pixels_per_thread > 2 ? 2 : pixels_per_thread);
  return track_blocks(npix, ppt);
}
void launch_track_rows_objs(hipStream_t s, const TrackLaunch& p, const TrackInst* inst, int ninst, GNState* st,
                            float* partial, int nblk, int pixels_per_thread, int q) {
  const size_t lds = (size_t)ninst * sizeof(Mats);
  dim3 grd(nblk), blk(TB);
#define MF_OBJS(PPT, RB)                                                                                                  \
  if (p.use_icp && p.use_rgb) k_track_objs<true, true, PPT, RB><<<grd, blk, lds, s>>>(p, inst, ninst, st, partial, nblk, q); \
  else if (p.use_icp) k_track_objs<true, false, PPT, RB><<<grd, blk, lds, s>>>(p, inst, ninst, st, partial, nblk, q);        \
  else k_track_objs<false, true, PPT, RB><<<grd, blk, lds, s>>>(p, inst, ninst, st, partial, nblk, q);
  const bool cauchy = p.robust == ROBUST_CAUCHY;
  if (pixels_per_thread >= 2) {
    if (cauchy) { MF_OBJS(2, ROBUST_CAUCHY) } else { MF_OBJS(2, ROBUST_ANY) }
  } else {
    if (cauchy) { MF_OBJS(1, ROBUST_CAUCHY) } else { MF_OBJS(1, ROBUST_ANY) }
  }
#undef MF_OBJS
}
void launch_track_solve(hipStream_t s, const TrackLaunch& p, const TrackInst* inst, int ninst, GNState* st,
                        const float* partial, int nblk, int q, const TrackSync& sy) {
  k_solve<<<ninst, SOLVE_NT, 0, s>>>(p.level, p.iter, inst, st, partial, nblk, q, sy);
}
void launch_gn_fused(hipStream_t s, const TrackLaunch& p, const TrackInst* inst, int ninst, GNState* st,
                     const float* part_prev, int nblk_prev, float* part_cur, int nblk, int q, int rows,
                     const TrackSync& sy) {
  dim3 grd(rows ? nblk : 1, ninst), blk(TB);
#define MF_FUSED(RB) \
  if (p.object) { \
    if (p.use_icp && p.use_rgb) k_gn_iter<true, true, true, RB><<<grd, blk, 0, s>>>(p, inst, st, part_prev, nblk_prev, part_cur, nblk, q, rows, sy); \
    else if (p.use_icp) k_gn_iter<true, false, true, RB><<<grd, blk, 0, s>>>(p, inst, st, part_prev, nblk_prev, part_cur, nblk, q, rows, sy); \
    else k_gn_iter<false, true, true, RB><<<grd, blk, 0, s>>>(p, inst, st, part_prev, nblk_prev, part_cur, nblk, q, rows, sy); \
  } else { \
    if (p.use_icp && p.use_rgb) k_gn_iter<true, true, false, RB><<<grd, blk, 0, s>>>(p, inst, st, part_prev, nblk_prev, part_cur, nblk, q, rows, sy); \
    else if (p.use_icp) k_gn_iter<true, false, false, RB><<<grd, blk, 0, s>>>(p, inst, st, part_prev, nblk_prev, part_cur, nblk, q, rows, sy); \
    else k_gn_iter<false, true, false, RB><<<grd, blk, 0, s>>>(p, inst, st, part_prev, nblk_prev, part_cur, nblk, q, rows, sy); \
  }
  if (p.robust == ROBUST_CAUCHY) { MF_FUSED(ROBUST_CAUCHY) } else { MF_FUSED(ROBUST_ANY) }
#undef MF_FUSED
}
void launch_eval(hipStream_t s, const TrackLaunch& p, const TrackInst* inst, GNState* st, int state, const Mat4& T,
                 float* partial, int nblk) {
  k_eval_set<<<1, 1, 0, s>>>(st, state, T);
  launch_track_rows(s, p, inst, 1, st, partial, nblk, 0);
  k_eval_sums<<<1, SOLVE_NT, 0, s>>>(st, state, partial, nblk);
}
// Tile flags and tile box of model maps written by the caller (the main
// raycast computes them for the maps it writes): one wave per 8x8 tile of
// the 6 planes at mvn, flag 1 when any pixel is not a raycast miss (vertex 0,
// normal (INVALID, 0, 0)); tbox must hold the empty box beforehand.
__global__ __launch_bounds__(64) void k_map_tiles(const float* __restrict__ mvn, uint8_t* __restrict__ tiles,
                                                  int* __restrict__ tbox, int W, int H) {
  const int ntx = (W + MAP_TILE - 1) / MAP_TILE;
  const int tx = blockIdx.x % ntx, ty = blockIdx.x / ntx;
  const int x = tx * MAP_TILE + (threadIdx.x & 7), y = ty * MAP_TILE + (threadIdx.x >> 3);
  bool hit = false;
  if (x < W && y < H) {
    const size_t n = (size_t)W * H, i = (size_t)x + (size_t)y * W;
    const unsigned inv = __float_as_uint((float)INVALID);
    hit = (__float_as_uint(mvn[i]) | __float_as_uint(mvn[n + i]) | __float_as_uint(mvn[2 * n + i]) |
           (__float_as_uint(mvn[3 * n + i]) ^ inv) | __float_as_uint(mvn[4 * n + i]) |
           __float_as_uint(mvn[5 * n + i])) != 0u;
  }
  const uint64_t b = __ballot(hit);
  if (threadIdx.x == 0) {
    tiles[blockIdx.x] = b != 0 ? 1 : 0;
    if (b != 0) { atomicMin(&tbox[0], tx); atomicMin(&tbox[1], ty); atomicMax(&tbox[2], tx); atomicMax(&tbox[3], ty); }
  }
}
void launch_map_tiles(hipStream_t s, const float* mvn, uint8_t* tiles, int* tbox, int W, int H) {
  const int nt = ((W + MAP_TILE - 1) / MAP_TILE) * ((H + MAP_TILE - 1) / MAP_TILE);
  k_map_tiles<<<nt, 64, 0, s>>>(mvn, tiles, tbox, W, H);
}

// The 6x6 solve on caller-given systems (27 doubles each: JtWe, then the
// packed upper JtWJ), one wave per system; the tracking solve runs the same
// solve6 on its FP64 sums.
__global__ __launch_bounds__(64) void k_solve6(const double* __restrict__ s27, int n, double* __restrict__ x,
                                               int* __restrict__ solver) {
  const int b = blockIdx.x;
  if (b >= n) return;
  double xx[6];
  const int sv = solve6(s27 + (size_t)b * 27, threadIdx.x, xx);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 6; i++) x[(size_t)b * 6 + i] = xx[i];
    solver[b] = sv;
  }
}
void launch_solve6(hipStream_t s, const double* s27, int n, double* x, int* solver) {
  if (n > 0) k_solve6<<<n, 64, 0, s>>>(s27, n, x, solver);
}

__global__ void k_pose_from_table(Mat4* dst, const Mat4* table, const int* index) { *dst = table[*index]; }
__global__ void k_set_int(int* dst, int v) { *dst = v; }
void launch_set_int(hipStream_t s, int* dst, int v) { k_set_int<<<1, 1, 0, s>>>(dst, v); }
void launch_pose_from_table(hipStream_t s, Mat4* dst, const Mat4* table, const int* index) {
  k_pose_from_table<<<1, 1, 0, s>>>(dst, table, index);
}
void launch_motion_rows(hipStream_t s, const MotionLaunch& m) {
  k_motion_rows<<<dim3((m.W + 255) / 256, m.H), 256, 0, s>>>(m);
}
void launch_motion_rows_n(hipStream_t s, const MotionLaunch& base, const MotionObj* objs, int n) {
  if (n > 0) k_motion_rows_n<<<dim3((base.W + 255) / 256, base.H, n), 256, 0, s>>>(base, objs);
}
void launch_motion_gate(hipStream_t s, MotionDev* md, int nobj) { k_motion_gate<<<1, 64, 0, s>>>(md, nobj); }
void launch_motion_inview(hipStream_t s, MotionDev* md, int nobj, const uint8_t* moving, int manual) {
  k_motion_inview<<<1, 1, 0, s>>>(md, nobj, moving, manual);
}
void launch_motion_classify(hipStream_t s, MotionDev* md, int nobj, uint8_t* moving) {
  k_motion_classify<<<1, 1, 0, s>>>(md, nobj, moving);
}
void launch_motion_after(hipStream_t s, MotionDev* md, int nobj, int use_rgb) {
  k_motion_after<<<1, 1, 0, s>>>(md, nobj, use_rgb);
}
void launch_virtual_cams(hipStream_t s, const GNState* st, int first, int n, const int* ids, Mat4* vcam) {
  k_virtual_cams<<<1, 64, 0, s>>>(st, first, n, ids, vcam);
}
void launch_copy_codes(hipStream_t s, const int8_t* src, int8_t* dst, size_t n2, const float* esrc, float* edst,
                       size_t ne, const int* gate) {
  k_copy_codes<<<(unsigned)std::min<size_t>(1024, (n2 / 4 + 256) / 256), 256, 0, s>>>(src, dst, n2, esrc, edst, ne,
                                                                                       gate);
}
void launch_camera_begin(hipStream_t s, Mat4* mats, GNState* st, float fx, float fy, float cx, float cy) {
  k_camera_begin<<<1, 1, 0, s>>>(mats, st, fx, fy, cx, cy);
}
void launch_gn_begin_camera(hipStream_t s, GNState* st, Mat4* T_w_l, const Mat4* T_w_r, const int* gate, float fx,
                            float fy, float cx, float cy) {
  k_gn_begin_camera<<<1, 1, 0, s>>>(st, T_w_l, T_w_r, gate, fx, fy, cx, cy);
}
void launch_gn_begin_objects(hipStream_t s, GNState* st, int first, int n, const int* ids, Mat4* poses,
                             const Mat4* T_w_c2, const Mat4* T_w_c1, int publish, const int* gate) {
  k_gn_begin_objects<<<1, 64, 0, s>>>(st, first, n, ids, poses, T_w_c2, T_w_c1, publish, gate);
}

}  // namespace mf
