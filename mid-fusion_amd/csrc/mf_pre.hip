// Depth/RGB preprocessing and image pyramids on gfx950.
//
// Reference: apps/kfusion/src/cpp/preprocessing/preprocessing.{h,cpp},
// Tracking::buildPyramid / calc_icp_cov (tracking/tracking.cpp:776-871),
// vertex2depth + half-sample of the rendered reference (tracking.cpp:218-238).
// All kernels are HBM-bound per-pixel maps; the 5x5 bilateral stencil is
// staged through LDS.
#include "mf_kernels.h"

namespace mf {

// mm2metersKernel (preprocessing.cpp:132-159) + rgb2intensity
// (preprocessing.cpp:73-102, rgb2gs commons.h:150-154), fused: one pass over
// the input frame writes depth (m), grey and rgb/255 planes.
__global__ void k_convert(const uint16_t* __restrict__ din, const uint8_t* __restrict__ rgb, int iw, int ratio,
                          int ow, int oh, float* __restrict__ depth, float* __restrict__ grey,
                          float* __restrict__ r, float* __restrict__ g, float* __restrict__ b) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= ow || y >= oh) return;
  const size_t src = (size_t)x * ratio + (size_t)iw * y * ratio;
  const size_t o = (size_t)x + (size_t)ow * y;
  depth[o] = (float)din[src] / 1000.0f;
  if (rgb) {
    const float R = rgb[3 * src], G = rgb[3 * src + 1], B = rgb[3 * src + 2];
    r[o] = R / 255.0f; g[o] = G / 255.0f; b[o] = B / 255.0f;
    grey[o] = (0.299f * R + 0.587f * G + 0.114f * B) / 255.0f;
  }
}

// bilateralFilterKernel (preprocessing.cpp:11-50), r = 2, applied to up to two
// images per launch (depth and grey at pyramid level 0, tracking.cpp:785-786).
// 16x16 tile + 2-pixel clamped halo in LDS.
#define BT 16
__global__ __launch_bounds__(256) void k_bilateral(const float* __restrict__ in0, float* __restrict__ out0,
                                                   const float* __restrict__ in1, float* __restrict__ out1, int w,
                                                   int h, float e_d, const float* __restrict__ gauss_g) {
  __shared__ float tile[BT + 4][BT + 4 + 1];
  __shared__ float g[5];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int x0 = blockIdx.x * BT, y0 = blockIdx.y * BT;
  const int tid = ty * BT + tx;
  if (tid < 5) g[tid] = gauss_g[tid];
  const float e2 = e_d * e_d * 2.0f;
  for (int img = 0; img < 2; img++) {
    const float* in = img ? in1 : in0;
    float* out = img ? out1 : out0;
    if (!in) break;
    for (int i = tid; i < (BT + 4) * (BT + 4); i += BT * BT) {
      const int ly = i / (BT + 4), lx = i % (BT + 4);
      const int gx = min(max(x0 + lx - 2, 0), w - 1), gy = min(max(y0 + ly - 2, 0), h - 1);
      tile[ly][lx] = in[(size_t)gx + (size_t)gy * w];
    }
    __syncthreads();
    const int x = x0 + tx, y = y0 + ty;
    if (x < w && y < h) {
      const float c = tile[ty + 2][tx + 2];
      float res = 0.0f;
      if (c != 0.0f) {
        float sum = 0.0f, t = 0.0f;
        for (int i = -2; i <= 2; ++i)       // x offset (outer loop in the reference)
          for (int j = -2; j <= 2; ++j) {   // y offset
            // clamped coordinates: the halo already holds clamped samples,
            // but the clamp must act on the absolute index
            const int cx = min(max(x + i, 0), w - 1) - x0 + 2;
            const int cy = min(max(y + j, 0), h - 1) - y0 + 2;
            const float p = tile[cy][cx];
            if (p > 0.0f) {
              const float d = p - c;
              const float mod = d * d;
              // expf as the correctly rounded float of a double exp, so the
              // host reference and the device agree bit for bit
              const float f = g[i + 2] * g[j + 2] * (float)exp((double)(-mod / e2));
              t = t + f * p;
              sum = sum + f;
            }
          }
        res = t / sum;
      }
      out[(size_t)x + (size_t)y * w] = res;
    }
    __syncthreads();
  }
}

// halfSampleRobustImageKernel (preprocessing.cpp:161-196), r = 1,
// e_d = e_delta*3 (tracking.cpp:789-792); up to two images per launch.
__global__ void k_half(const float* __restrict__ in0, float* __restrict__ out0, const float* __restrict__ in1,
                       float* __restrict__ out1, int iw, int ih, float e_d) {
  const int ow = iw / 2, oh = ih / 2;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= ow || y >= oh) return;
  for (int img = 0; img < 2; img++) {
    const float* in = img ? in1 : in0;
    float* out = img ? out1 : out0;
    if (!in) break;
    const int cx = 2 * x, cy = 2 * y;
    const float c = in[cx + (size_t)cy * iw];
    float sum = 0.0f, t = 0.0f;
    for (int i = 0; i <= 1; ++i)
      for (int j = 0; j <= 1; ++j) {
        const int ux = min(cx + j, 2 * ow - 1), uy = min(cy + i, 2 * oh - 1);
        const float cur = in[ux + (size_t)uy * iw];
        if (fabsf(cur - c) < e_d) { sum = sum + 1.0f; t = t + cur; }
      }
    out[x + (size_t)y * ow] = t / sum;
  }
}

// depth2vertexKernel (preprocessing.cpp:52-71) as a pure function of a pixel.
__device__ __forceinline__ f3 d2v(float d, int x, int y, const Mat4& invK) {
  return d * rotate(invK, mk3((float)x, (float)y, 1.0f));
}

// Per-level live maps, fused: depth2vertex, vertex2normal<SDF,NegY>
// (preprocessing.h:39-81), gradientsKernel (preprocessing.cpp:235-246,
// border left 0) and calc_icp_cov (tracking.cpp:845-871).
__global__ void k_level(const float* __restrict__ D, const float* __restrict__ I, int w, int h, Mat4 invK, int negY,
                        float focal, float sxy, float sdisp, float baseline, float* __restrict__ vx,
                        float* __restrict__ vy, float* __restrict__ vz, float* __restrict__ nx,
                        float* __restrict__ ny, float* __restrict__ nz, float* __restrict__ gx,
                        float* __restrict__ gy, float* __restrict__ cx, float* __restrict__ cy,
                        float* __restrict__ cz) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w || y >= h) return;
  const size_t i = (size_t)x + (size_t)y * w;
  const float d = D[i];
  const f3 v = d > 0.0f ? d2v(d, x, y, invK) : mk3(0.f, 0.f, 0.f);
  vx[i] = v.x; vy[i] = v.y; vz[i] = v.z;
  // neighbours
  const int lx = max(x - 1, 0), rx = min(x + 1, w - 1);
  int yu, yd;
  if (negY) { yu = max(y - 1, 0); yd = min(y + 1, h - 1); }
  else { yd = max(y - 1, 0); yu = min(y + 1, h - 1); }
  auto vert = [&](int px, int py) -> f3 {
    const float dd = D[(size_t)px + (size_t)py * w];
    return dd > 0.0f ? d2v(dd, px, py, invK) : mk3(0.f, 0.f, 0.f);
  };
  const f3 left = vert(lx, y), right = vert(rx, y), up = vert(x, yu), down = vert(x, yd);
  f3 n;
  if (left.z == 0 || right.z == 0 || up.z == 0 || down.z == 0) n = mk3((float)INVALID, 0.f, 0.f);
  else n = normalize(cross(up - down, right - left));
  nx[i] = n.x; ny[i] = n.y; nz[i] = n.z;
  if (gx) {
    float a = 0.f, bq = 0.f;
    if (x >= 1 && x < w - 1 && y >= 1 && y < h - 1) {
      a = (I[i + 1] - I[i - 1]) * 0.5f;
      bq = (I[i + w] - I[i - w]) * 0.5f;
    }
    gx[i] = a; gy[i] = bq;
  }
  if (cx) {
    float c0 = (d / focal) * sxy, c1 = (d / focal) * sxy, c2 = (d * d * sdisp) / (focal * baseline);
    cx[i] = c0 * c0; cy[i] = c1 * c1; cz[i] = c2 * c2;
  }
}

// vertex2depth (preprocessing.cpp:292-309) of the model vertex map at the
// reference pose: z of inverse(T_w_r)*v, 0 where v.z <= 0.  `invT` is read
// from device memory (poses stay on the device).
__global__ void k_vertex2depth(const float* __restrict__ vx, const float* __restrict__ vy,
                               const float* __restrict__ vz, int n, const Mat4* __restrict__ invT,
                               float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Mat4 M = *invT;
  const f3 v = mk3(vx[i], vy[i], vz[i]);
  out[i] = (v.z > 0.0f) ? affine(M, v).z : 0.0f;
}

// Rendered reference depth pyramids of `gridDim.z` tracking instances in one
// launch (tracking.cpp:218-238): each workgroup renders a 32x16 tile of level
// 0, vertex2depth of the instance's model vertex map at the reference pose
// (preprocessing.cpp:292-309), and half-samples it level by level in LDS with
// the robust rule (the 2x2 values within e_d of the top-left one, averaged).
// Level l's tile is (32 >> l) x (16 >> l) pixels, and the 2x2 sources of
// every level-l pixel (x < w_l = w_{l-1} / 2) lie inside the image and inside
// the tile, so the levels come out as separate half-sampling passes over
// the whole image would make them.  n_levels <= 4.
static constexpr int RR_TX = 32, RR_TY = 16;
struct RefOffsets { int off[MAX_LEVELS]; };
__global__ __launch_bounds__(256) void k_render_ref(const TrackInst* __restrict__ insts, int W, int H, int n_levels,
                                                   const Mat4* __restrict__ invT, RefOffsets ro, float e_d) {
  __shared__ float tile[RR_TX * RR_TY + (RR_TX / 2) * (RR_TY / 2) + (RR_TX / 4) * (RR_TY / 4) + (RR_TX / 8) * (RR_TY / 8)];
  const int tid = threadIdx.x;
  const TrackInst& ti = insts[blockIdx.z];
  float* rD = const_cast<float*>(ti.rD);
  const Mat4 M = *invT;
  const int X0 = blockIdx.x * RR_TX, Y0 = blockIdx.y * RR_TY;
  const size_t n0 = (size_t)W * H;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int lx = tid & (RR_TX - 1), ly = (tid >> 5) + (RR_TY / 2) * k;
    const int x = X0 + lx, y = Y0 + ly;
    float d = 0.0f;
    if (x < W && y < H) {
      const size_t i = (size_t)x + (size_t)y * W;
      const f3 v = mk3(ti.mvn[i], ti.mvn[n0 + i], ti.mvn[2 * n0 + i]);
      d = (v.z > 0.0f) ? affine(M, v).z : 0.0f;
      rD[i] = d;
    }
    tile[ly * RR_TX + lx] = d;
  }
  __syncthreads();
  int iw = W, ih = H, src = 0, dst = RR_TX * RR_TY;
  for (int l = 1; l < n_levels; l++) {
    const int ow = iw / 2, oh = ih / 2;
    const int sw = RR_TX >> (l - 1), tw = RR_TX >> l, th = RR_TY >> l;
    if (tid < tw * th) {
      const int lx = tid % tw, ly = tid / tw;
      const int x = (X0 >> l) + lx, y = (Y0 >> l) + ly;
      if (x < ow && y < oh) {
        const float* in = tile + src + (2 * ly) * sw + 2 * lx;
        const float c = in[0];
        float sum = 0.0f, t = 0.0f;
        for (int i = 0; i <= 1; ++i)
          for (int j = 0; j <= 1; ++j) {
            const float cur = in[i * sw + j];
            if (fabsf(cur - c) < e_d) { sum = sum + 1.0f; t = t + cur; }
          }
        const float r = t / sum;
        rD[ro.off[l] + x + (size_t)y * ow] = r;
        tile[dst + ly * tw + lx] = r;
      }
    }
    __syncthreads();
    iw = ow; ih = oh;
    src = dst;
    dst += tw * th;
  }
}

// ------------------------------------------------------------ launchers
void launch_render_reference(hipStream_t s, const TrackInst* insts, int n, int w, int h, int n_levels,
                             const Mat4* invT, const int* rd_off) {
  RefOffsets ro;
  for (int l = 0; l < MAX_LEVELS; l++) ro.off[l] = rd_off[l];
  const int nl = n_levels < MAX_LEVELS ? n_levels : MAX_LEVELS;
  k_render_ref<<<dim3((w + RR_TX - 1) / RR_TX, (h + RR_TY - 1) / RR_TY, n), 256, 0, s>>>(insts, w, h, nl, invT, ro,
                                                                                         0.1f * 3);
}
void launch_convert(hipStream_t s, const uint16_t* din, const uint8_t* rgb, int iw, int ratio, int ow, int oh,
                    float* depth, float* grey, float* r, float* g, float* b) {
  dim3 blk(256), grd((ow + 255) / 256, oh);
  k_convert<<<grd, blk, 0, s>>>(din, rgb, iw, ratio, ow, oh, depth, grey, r, g, b);
}
void launch_bilateral(hipStream_t s, const float* in0, float* out0, const float* in1, float* out1, int w, int h,
                      float e_d, const float* gauss) {
  dim3 blk(BT, BT), grd((w + BT - 1) / BT, (h + BT - 1) / BT);
  k_bilateral<<<grd, blk, 0, s>>>(in0, out0, in1, out1, w, h, e_d, gauss);
}
void launch_half(hipStream_t s, const float* in0, float* out0, const float* in1, float* out1, int iw, int ih,
                 float e_d) {
  const int ow = iw / 2, oh = ih / 2;
  dim3 blk(256), grd((ow + 255) / 256, oh);
  k_half<<<grd, blk, 0, s>>>(in0, out0, in1, out1, iw, ih, e_d);
}
void launch_level(hipStream_t s, const Level& L, const Mat4& invK, int negY, float focal, float sxy, float sdisp,
                  float baseline, bool grads, bool cov) {
  dim3 blk(256), grd((L.w + 255) / 256, L.h);
  k_level<<<grd, blk, 0, s>>>(L.D, L.I, L.w, L.h, invK, negY, focal, sxy, sdisp, baseline, L.V[0], L.V[1], L.V[2],
                              L.N[0], L.N[1], L.N[2], grads ? L.gx : nullptr, grads ? L.gy : nullptr,
                              cov ? L.cov[0] : nullptr, cov ? L.cov[1] : nullptr, cov ? L.cov[2] : nullptr);
}
void launch_vertex2depth(hipStream_t s, const float* const v[3], int n, const Mat4* invT, float* out) {
  k_vertex2depth<<<(n + 255) / 256, 256, 0, s>>>(v[0], v[1], v[2], n, invT, out);
}

}  // namespace mf
