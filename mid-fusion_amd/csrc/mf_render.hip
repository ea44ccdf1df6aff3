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

// ---- instance_seg::generalize_label (segmentation/segmentation.cpp:10-47)
// on 0/non-zero instance masks: the instance dilated by the 5x5 elliptic
// element (cv::getStructuringElement(MORPH_ELLIPSE, 5x5): rows 0 and 4 hold
// the centre column only; pixels outside the image do not take part), OR the
// model mask if one is given; -1 outside that, 0 on it but off the instance
// (the background layer), the foreground value on the instance.
__global__ void k_mask_count(const uint8_t* __restrict__ mask, int n, int* __restrict__ count) {
  int c = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c += mask[i] != 0;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}
__global__ void k_generalize(const uint8_t* __restrict__ mask, const uint8_t* __restrict__ global, int w, int h,
                             int rendered, const int* __restrict__ count, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w || y >= h) return;
  float fg = 1.0f;
  if (rendered)   // render_mask_size_ratio_ 0.03, large 0.5, small 0.1 (segmentation.h:110-113)
    fg = ((float)*count < 0.03f * (float)w * (float)h) ? 0.1f : 0.5f;
  bool dil = false;
  for (int dy = -2; dy <= 2 && !dil; dy++) {
    const int yy = y + dy;
    if (yy < 0 || yy >= h) continue;
    const int r = (dy == -2 || dy == 2) ? 0 : 2;
    for (int dx = -r; dx <= r; dx++) {
      const int xx = x + dx;
      if (xx >= 0 && xx < w && mask[(size_t)yy * w + xx]) { dil = true; break; }
    }
  }
  const size_t i = (size_t)y * w + x;
  if (global && global[i]) dil = true;
  out[i] = !dil ? -1.0f : (mask[i] ? fg : 0.0f);
}
void launch_generalize(hipStream_t s, const uint8_t* mask, const uint8_t* global, int w, int h, int rendered,
                       int* count, float* out) {
  if (rendered) {
    launch_set_int(s, count, 0);
    k_mask_count<<<256, 256, 0, s>>>(mask, w * h, count);
  }
  k_generalize<<<dim3((w + 255) / 256, h), 256, 0, s>>>(mask, global, w, h, rendered, count, out);
}

// Integration masks of several objects in one launch (grid.y = object):
// float4 copies when every pointer is 16-byte aligned, else floats.
__global__ void k_copy_masks(MaskCopies mc, size_t n, int vec) {
  const float* __restrict__ src = mc.src[blockIdx.y];
  float* __restrict__ dst = mc.dst[blockIdx.y];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (vec) {
    const float4* __restrict__ s4 = reinterpret_cast<const float4*>(src);
    float4* __restrict__ d4 = reinterpret_cast<float4*>(dst);
    for (size_t i = t; i < n / 4; i += stride) d4[i] = s4[i];
    for (size_t i = (n / 4) * 4 + t; i < n; i += stride) dst[i] = src[i];
  } else {
    for (size_t i = t; i < n; i += stride) dst[i] = src[i];
  }
}
void launch_copy_masks(hipStream_t s, const MaskCopies& mc, int m, size_t n) {
  if (m <= 0) return;
  int vec = 1;
  for (int i = 0; i < m; i++)
    if (((uintptr_t)mc.src[i] | (uintptr_t)mc.dst[i]) & 15) vec = 0;
  const size_t units = vec ? (n + 3) / 4 : n;
  const unsigned gx = (unsigned)std::min<size_t>(1024, (units + 255) / 256);
  k_copy_masks<<<dim3(gx, m), 256, 0, s>>>(mc, n, vec);
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
