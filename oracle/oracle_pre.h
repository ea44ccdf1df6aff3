/*
 * CPU ORACLE (test infrastructure) — preprocessing and pyramid kernels.
 * Restates apps/kfusion/src/cpp/preprocessing/preprocessing.{h,cpp} and
 * Tracking::calc_icp_cov (tracking/tracking.cpp:845-871).
 */
#pragma once
#include "oracle_core.h"

namespace orc {

// Spatial Gaussian of the bilateral filter: g[i] = expf(-(x*x)/(2*delta^2)),
// x = i-2, delta = 4 (kfusion.cpp:80-88, tracking.cpp:77-85,
// constant_parameters.h:21).  -(x*x) is an int, divided by a float.
static inline void gaussian5(R g[5]) {
  const R delta = R(4.0);
  for (int i = 0; i < 5; i++) { int x = i - 2; g[i] = std::exp(R(-(x * x)) / (R(2) * delta * delta)); }
}

// mm2metersKernel (preprocessing.cpp:132-159).  Returns false on a bad ratio
// (the reference calls exit(1)).
static inline bool ratio_ok(int iw, int ih, int ow, int oh) {
  if (iw < ow || ih < oh) return false;
  if (iw % ow != 0 || ih % oh != 0) return false;
  if (iw / ow != ih / oh) return false;
  return true;
}
static inline bool mm2meters(R* out, int ow, int oh, const uint16_t* in, int iw, int ih) {
  if (!ratio_ok(iw, ih, ow, oh)) return false;
  int ratio = iw / ow;
  for (int y = 0; y < oh; y++)
    for (int x = 0; x < ow; x++) out[x + ow * y] = R(in[x * ratio + iw * y * ratio]) / R(1000.0);
  return true;
}

// rgb2intensity (grey + rgb/255 overload, preprocessing.cpp:73-102);
// rgb2gs (commons.h:150-154): (0.299R + 0.587G + 0.114B)/255.
static inline bool rgb2intensity(R* grey, V3* rgb, int ow, int oh, const uint8_t* in, int iw, int ih) {
  if (!ratio_ok(iw, ih, ow, oh)) return false;
  int ratio = iw / ow;
  for (int y = 0; y < oh; y++)
    for (int x = 0; x < ow; x++) {
      const uint8_t* p = in + 3 * (x * ratio + iw * y * ratio);
      if (rgb) {
        rgb[x + ow * y].x = R(p[0]) / R(255.0);
        rgb[x + ow * y].y = R(p[1]) / R(255.0);
        rgb[x + ow * y].z = R(p[2]) / R(255.0);
      }
      grey[x + ow * y] = (R(0.299) * R(p[0]) + R(0.587) * R(p[1]) + R(0.114) * R(p[2])) / R(255.0);
    }
  return true;
}

// bilateralFilterKernel (preprocessing.cpp:11-50).  Out-of-place only: the
// in-place use at kfusion.cpp:218-221 has a read/write race, see the design notes.
static inline void bilateral(R* out, const R* in, int w, int h, R e_d, int r) {
  R g[5]; gaussian5(g);
  const R e2 = e_d * e_d * R(2);
#pragma omp parallel for
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      int pos = x + y * w;
      if (in[pos] == 0) { out[pos] = 0; continue; }
      R sum = 0, t = 0;
      const R c = in[pos];
      for (int i = -r; i <= r; ++i)
        for (int j = -r; j <= r; ++j) {
          int cx = std::min(std::max(x + i, 0), w - 1);
          int cy = std::min(std::max(y + j, 0), h - 1);
          const R p = in[cx + cy * w];
          if (p > 0) {
            const R d = p - c;
            const R mod = d * d;
            // expf evaluated as the correctly rounded float of a double exp
            const R f = g[i + r] * g[j + r] * (R)std::exp((double)(-mod / e2));
            t = t + f * p;
            sum = sum + f;
          }
        }
      out[pos] = t / sum;
    }
}

// halfSampleRobustImageKernel (preprocessing.cpp:161-196); callers use
// e_d = e_delta*3 and r = 1 (tracking.cpp:789-792).
static inline void half_sample(R* out, const R* in, int iw, int ih, R e_d, int r) {
  int ow = iw / 2, oh = ih / 2;
#pragma omp parallel for
  for (int y = 0; y < oh; y++)
    for (int x = 0; x < ow; x++) {
      int cxp = 2 * x, cyp = 2 * y;
      R sum = 0, t = 0;
      const R c = in[cxp + cyp * iw];
      for (int i = -r + 1; i <= r; ++i)
        for (int j = -r + 1; j <= r; ++j) {
          int ux = std::min(std::max(cxp + j, 0), 2 * ow - 1);
          int uy = std::min(std::max(cyp + i, 0), 2 * oh - 1);
          R cur = in[ux + uy * iw];
          if (std::fabs(cur - c) < e_d) { sum = sum + R(1); t = t + cur; }
        }
      out[x + y * ow] = t / sum;
    }
}

// depth2vertexKernel (preprocessing.cpp:52-71): v = d * rotate(invK, (x,y,1)).
static inline void depth2vertex(V3* v, const R* d, int w, int h, const M4& invK) {
#pragma omp parallel for
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      R dd = d[x + y * w];
      if (dd > 0) v[x + y * w] = dd * rotate(invK, v3(R(x), R(y), R(1)));
      else v[x + y * w] = v3(0, 0, 0);
    }
}

// vertex2normalKernel<SDF, NegY> (preprocessing.h:39-81).  INVALID writes
// (-2, 0, 0); the reference only writes .x and leaves y/z stale (every
// consumer tests .x first).
static inline void vertex2normal(V3* out, const V3* in, int w, int h, bool negY) {
#pragma omp parallel for
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      int lx = std::max(x - 1, 0), rx = std::min(x + 1, w - 1);
      int yu, yd;
      if (negY) { yu = std::max(y - 1, 0); yd = std::min(y + 1, h - 1); }
      else { yd = std::max(y - 1, 0); yu = std::min(y + 1, h - 1); }
      V3 left = in[lx + w * y], right = in[rx + w * y], up = in[x + w * yu], down = in[x + w * yd];
      if (left.z == 0 || right.z == 0 || up.z == 0 || down.z == 0) {
        out[x + y * w] = v3(R(INVALID), 0, 0);
        continue;
      }
      V3 dxv = right - left, dyv = up - down;
      out[x + y * w] = normalize(cross(dyv, dxv));
    }
}

// gradientsKernel (preprocessing.cpp:235-246): interior only; the border
// keeps its zero initialisation.
static inline void gradients(R* gx, R* gy, const R* in, int w, int h) {
  for (int i = 0; i < w * h; i++) { gx[i] = 0; gy[i] = 0; }
  for (int y = 1; y < h - 1; y++)
    for (int x = 1; x < w - 1; x++) {
      gx[x + w * y] = (in[x + 1 + w * y] - in[x - 1 + w * y]) * R(0.5);
      gy[x + w * y] = (in[x + w * (y + 1)] - in[x + w * (y - 1)]) * R(0.5);
    }
}

// vertex2depth (preprocessing.cpp:292-309): z of inverse(T_w_c)*v, 0 if v.z<=0.
static inline void vertex2depth(R* d, const V3* v, int w, int h, const M4& T_w_c) {
  M4 inv = inverse(T_w_c);
  for (int i = 0; i < w * h; i++) d[i] = (v[i].z > 0) ? affine(inv, v[i]).z : R(0);
}

// Tracking::calc_icp_cov (tracking.cpp:845-871)
static inline void icp_cov(V3* cov, const R* depth, int w, int h, R focal, R sxy, R sdisp, R baseline) {
  for (int i = 0; i < w * h; i++) {
    R d = depth[i];
    V3 c;
    c.x = (d / focal) * sxy;
    c.y = (d / focal) * sxy;
    c.z = (d * d * sdisp) / (focal * baseline);
    c.x = c.x * c.x; c.y = c.y * c.y; c.z = c.z * c.z;
    cov[i] = c;
  }
}

}  // namespace orc
