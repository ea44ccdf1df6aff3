// Display rendering of the per-frame state on gfx950: tracking-result,
// depth and intensity images (rendering/rendering.cpp:183-237, 368-487,
// 648-664).  One pixel per thread, uchar4 (x, y, z, w) out as one 32-bit
// store.
#include "mf_kernels.h"

namespace mf {

__device__ __forceinline__ uint32_t rgba(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
  return (r & 0xffu) | ((g & 0xffu) << 8) | ((b & 0xffu) << 16) | ((a & 0xffu) << 24);
}
// float -> unsigned char as the reference's host conversion does it: to a
// 32-bit integer (truncation), then the low byte
__device__ __forceinline__ uint32_t to_u8(float v) { return (uint32_t)(int)v & 0xffu; }

// renderTrackKernel (mode 0, ICP rows), renderRGBTrackKernel (mode 1,
// photometric rows and their error), render_RGBD_TrackKernel (mode 2)
__global__ void k_render_track(const int8_t* __restrict__ codes, const float* __restrict__ err, int n, int mode,
                               uint32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ri = codes[i], gi = codes[n + i];
  uint32_t c;
  if (mode == 0) {
    switch (ri) {
      case 1: c = rgba(128, 128, 128, 0); break;
      case -1: c = rgba(0, 0, 0, 0); break;
      case -2: c = rgba(255, 0, 0, 0); break;
      case -3: c = rgba(0, 255, 0, 0); break;
      case -4: c = rgba(0, 0, 255, 0); break;
      case -5: c = rgba(255, 255, 0, 0); break;
      case -6: c = rgba(255, 175, 175, 0); break;
      default: c = rgba(255, 200, 128, 0); break;
    }
  } else if (mode == 1) {
    switch (gi) {
      case 1: { const uint32_t e = to_u8(fabsf(err[i] * 255.0f)); c = rgba(e, e, e, 255); break; }
      case -1: c = rgba(0, 255, 0, 0); break;
      case -2: c = rgba(0, 0, 0, 0); break;
      case -3: c = rgba(255, 0, 0, 0); break;
      case -4: c = rgba(0, 0, 255, 0); break;
      case -5: c = rgba(255, 255, 0, 0); break;
      default: c = rgba(255, 128, 128, 0); break;
    }
  } else {
    if (ri == 1 && gi == 1) { const uint32_t e = to_u8(fabsf(err[i] * 255.0f)); c = rgba(e, e, e, 255); }
    else if ((ri < -3 && ri > -6) || (gi < -3 && gi > -6)) c = rgba(255, 0, 0, 0);
    else if (ri == -2 || gi == -2) c = rgba(255, 255, 0, 0);
    else if (ri == -3 || gi == -3) c = rgba(255, 255, 0, 0);
    else if (ri == -1 || gi == -1) c = rgba(0, 255, 0, 0);
    else if (ri == -6 || gi == -6) c = rgba(255, 175, 175, 0);
    else c = rgba(128, 128, 128, 0);
  }
  out[i] = c;
}

// gs2rgb (commons.h:87-140): hue ramp of h in [0, 1], in double
__device__ uint32_t gs2rgb(double h) {
  const double v = 0.75, m = 0.25, sv = 0.6667;
  double r = 0, g = 0, b = 0;
  h *= 6.0;
  const int sextant = (int)h;
  const double fract = h - sextant;
  const double vsf = v * sv * fract;
  const double mid1 = m + vsf, mid2 = v - vsf;
  switch (sextant) {
    case 0: r = v; g = mid1; b = m; break;
    case 1: r = mid2; g = v; b = m; break;
    case 2: r = m; g = v; b = mid1; break;
    case 3: r = m; g = mid2; b = v; break;
    case 4: r = mid1; g = m; b = v; break;
    case 5: r = v; g = m; b = mid2; break;
    default: break;
  }
  return rgba((uint32_t)(int)(r * 255.0), (uint32_t)(int)(g * 255.0), (uint32_t)(int)(b * 255.0), 0);
}

// renderDepthKernel (rendering.cpp:368-394)
__global__ void k_render_depth(const float* __restrict__ depth, int n, float nearp, float farp,
                               uint32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float rangeScale = 1 / (farp - nearp);
  const float d = depth[i];
  uint32_t c;
  if (d < nearp) c = rgba(255, 255, 255, 0);
  else if (d > farp) c = rgba(0, 0, 0, 0);
  else c = gs2rgb((double)((d - nearp) * rangeScale));
  out[i] = c;
}

// renderIntensityKernel (rendering.cpp:648-664)
__global__ void k_render_intensity(const float* __restrict__ grey, int n, uint32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t gs = to_u8(255.f * grey[i]);
  out[i] = rgba(gs, gs, gs, 255);
}

void launch_render_track(hipStream_t s, const int8_t* codes, const float* err, int n, int mode, uint32_t* out) {
  k_render_track<<<(n + 255) / 256, 256, 0, s>>>(codes, err, n, mode, out);
}
void launch_render_depth(hipStream_t s, const float* depth, int n, float nearp, float farp, uint32_t* out) {
  k_render_depth<<<(n + 255) / 256, 256, 0, s>>>(depth, n, nearp, farp, out);
}
void launch_render_intensity(hipStream_t s, const float* grey, int n, uint32_t* out) {
  k_render_intensity<<<(n + 255) / 256, 256, 0, s>>>(grey, n, out);
}

}  // namespace mf
