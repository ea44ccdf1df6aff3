// Multi-object raycasting on gfx950.
//
// Reference: raycastObjectList (rendering/rendering.cpp:9-181), ray_iterator
// (lib/octree.hpp:919-1130), raycast() (kfusion/rendering_impl.hpp:80-140),
// Octree::interp / grad / get_fine (lib/octree.hpp:359-379, 479-501, 593-680)
// and gather_points (lib/interpolation/interp_gather.hpp:107-237).
//
// One thread per pixel walks every object: ESVO traversal over the per-level
// occupancy bytes and the Morton block table gives the entry distance of the
// first allocated block (the same float operations as ray_iterator, so the
// march starts at the same t), then the SDF march reads voxels through the
// block table (missing voxels read as tsdf 1, fg 0).
#include "mf_kernels.h"

namespace mf {

__device__ __forceinline__ int32_t vslot(const VolDev& v, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= v.res || y >= v.res || z >= v.res) return -1;
  return gload(v.slot + morton3_10((uint32_t)x >> 3, (uint32_t)y >> 3, (uint32_t)z >> 3));
}

// One-entry block cache: consecutive voxel reads of a ray (march steps, the
// 8 interpolation corners, the gradient stencil) mostly stay in one 8^3
// block, so the block-table probe is skipped and each read is one load.
// Block coordinates are packed 10 bits each (res <= 8192, checked at
// volume creation).
struct VCache {
  int key = -1;
  int32_t slot = -1;
};
__device__ __forceinline__ int32_t cslot(const VolDev& v, VCache& c, int x, int y, int z) {
  if ((unsigned)x >= (unsigned)v.res || (unsigned)y >= (unsigned)v.res || (unsigned)z >= (unsigned)v.res) return -1;
  const int key = ((x >> 3) << 20) | ((y >> 3) << 10) | (z >> 3);
  if (key != c.key) {
    c.key = key;
    c.slot = gload(v.slot + morton3_10((uint32_t)x >> 3, (uint32_t)y >> 3, (uint32_t)z >> 3));
  }
  return c.slot;
}
__device__ __forceinline__ size_t vox(int32_t s, int x, int y, int z) {
  return (size_t)s * BLOCK_VOX + (x & 7) + 8 * (y & 7) + 64 * (z & 7);
}
// Field value at voxel (x, y, z) of block slot s; missing voxels read as
// voxel_traits<SDF>::empty(): tsdf 1, fg 0.
// TSDF reads take a 32-bit index from the pool base (slot < 2^21 blocks, so
// the byte offset fits), which saves the per-voxel 64-bit address arithmetic.
template <bool FG>
__device__ __forceinline__ float field_at(const VolDev& v, int32_t s, int x, int y, int z) {
  if (s < 0) return FG ? 0.f : 1.f;
  if (FG) return (float)gload(v.fg + vox(s, x, y, z));
  return ldx(v.tsdf, (uint32_t)s * (uint32_t)BLOCK_VOX + (uint32_t)((x & 7) + 8 * (y & 7) + 64 * (z & 7)));
}
__device__ __forceinline__ float tsdf_at(const VolDev& v, VCache& c, int x, int y, int z) {
  return field_at<false>(v, cslot(v, c, x, y, z), x, y, z);
}

// Octree::interp (octree.hpp:479-501): voxel values at integer corners, base
// clamped at 0, fraction of the unclamped position.  The 8 corner loads are
// issued together: when the corners share one 8^3 block (no corner
// coordinate crosses a multiple of 8) after one block-table probe, otherwise
// after 8 independent probes.
template <bool FG>
__device__ __forceinline__ float interp(const VolDev& v, VCache& c, f3 pos) {
  const float flx = floorf(pos.x), fly = floorf(pos.y), flz = floorf(pos.z);
  const int bx = (int)flx, by = (int)fly, bz = (int)flz;
  const float fx = pos.x - flx, fy = pos.y - fly, fz = pos.z - flz;
  const int lx = max(bx, 0), ly = max(by, 0), lz = max(bz, 0);
  float p[8];
  if ((lx & 7) != 7 && (ly & 7) != 7 && (lz & 7) != 7) {
    // res is a multiple of 8, so the +1 corners are inside exactly when the
    // base corner is
    const int32_t sl = cslot(v, c, lx, ly, lz);
#pragma unroll
    for (int i = 0; i < 8; i++) p[i] = field_at<FG>(v, sl, lx + (i & 1), ly + ((i >> 1) & 1), lz + ((i >> 2) & 1));
  } else {
    // the corners span 2 blocks along some axis: each corner's block-table
    // index from the two spread block coordinates per axis, and the 8 table
    // loads issued together (an out-of-range corner, only at the volume's
    // far faces, reads a clamped entry and takes -1 instead)
    const int nb = v.res >> 3;
    uint32_t e[3][2];
    bool in[3][2];
    const int L[3] = {lx, ly, lz};
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int d = 0; d < 2; d++) {
        const int c = L[a] + d;
        in[a][d] = c < v.res;
        e[a][d] = expand10((uint32_t)min(c >> 3, nb - 1)) << a;
      }
    int32_t sl[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int dx = i & 1, dy = (i >> 1) & 1, dz = (i >> 2) & 1;
      const int32_t s = gload(v.slot + (e[0][dx] | e[1][dy] | e[2][dz]));
      sl[i] = (in[0][dx] && in[1][dy] && in[2][dz]) ? s : -1;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) p[i] = field_at<FG>(v, sl[i], lx + (i & 1), ly + ((i >> 1) & 1), lz + ((i >> 2) & 1));
  }
  return (((p[0] * (1 - fx) + p[1] * fx) * (1 - fy) + (p[2] * (1 - fx) + p[3] * fx) * fy) * (1 - fz) +
          ((p[4] * (1 - fx) + p[5] * fx) * (1 - fy) + (p[6] * (1 - fx) + p[7] * fx) * fy) * fz);
}

// Octree::grad(pos, select) (octree.hpp:593-680), tsdf field: trilinearly
// weighted central differences over the clamped 4^3 stencil.  (Issuing the
// stencil's loads in batches per axis measured no faster: the raycast is
// bound by its divergent traversal, and the batch cost registers.)
__device__ f3 grad(const VolDev& v, VCache& cc, f3 pos) {
  const float fl[3] = {floorf(pos.x), floorf(pos.y), floorf(pos.z)};
  const int b[3] = {(int)fl[0], (int)fl[1], (int)fl[2]};
  const float f[3] = {pos.x - fl[0], pos.y - fl[1], pos.z - fl[2]};
  int ll[3], lu[3], ul[3], uu[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    ll[a] = max(b[a] - 1, 0); lu[a] = max(b[a], 0);
    ul[a] = min(b[a] + 1, v.res - 1); uu[a] = min(b[a] + 2, v.res - 1);
  }
  float g[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float D[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int id[3] = {q & 1, (q >> 1) & 1, (q >> 2) & 1};
      int hi[3], lo[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if (c == a) { hi[c] = id[c] ? uu[c] : ul[c]; lo[c] = id[c] ? lu[c] : ll[c]; }
        else { hi[c] = lo[c] = id[c] ? ul[c] : lu[c]; }
      }
      D[q] = tsdf_at(v, cc, hi[0], hi[1], hi[2]) - tsdf_at(v, cc, lo[0], lo[1], lo[2]);
    }
    g[a] = (((D[0] * (1 - f[0]) + D[1] * f[0]) * (1 - f[1]) + (D[2] * (1 - f[0]) + D[3] * f[0]) * f[1]) * (1 - f[2]) +
            ((D[4] * (1 - f[0]) + D[5] * f[0]) * (1 - f[1]) + (D[6] * (1 - f[0]) + D[7] * f[0]) * f[1]) * f[2]);
  }
  const float s = v.size / (float)v.res;
  return mk3(g[0] * s * 0.5f, g[1] * s * 0.5f, g[2] * s * 0.5f);
}

// Byte offset of internal-node level lv >= 1 in VolDev::nodes: levels are
// stored back to back from level 1, level l holding 8^l occupancy bytes, so
// the offset is sum_{k=1}^{lv-1} 8^k = 8 + 64 + ..., the octal digits 1 at
// positions 1..lv-1: the bit pattern 0x49249248 below bit 3*lv.
__device__ __forceinline__ uint32_t node_offset(int lv) { return 0x49249248u & ((1u << (3 * lv)) - 1u); }

// Child occupancy of an octree node: level `lv` (1 = children of the root),
// Morton prefix `pre` of the child.  One byte load on either kind of level,
// so that a wave whose lanes test leaf and internal children in the same
// step waits for one memory round trip, not two: at the leaf level the most
// significant byte of the block-table entry (its sign: -1 = missing), above
// it the occupancy byte.
__device__ __forceinline__ bool child_exists(const VolDev& v, int lv, uint32_t pre) {
  const bool leaf = lv == v.leaf_level;
  const uint8_t* a = leaf ? reinterpret_cast<const uint8_t*>(v.slot) + 4 * (size_t)pre + 3
                          : v.nodes + node_offset(lv) + pre;
  const uint8_t b = gload(a);
  return leaf ? (b & 0x80u) == 0 : b != 0;
}

// ray_iterator::next() first call (octree.hpp:923-1093): entry distance, in
// voxel units, of the first allocated leaf block along the ray; -1 if none.
// Same float operations as the reference traversal over [1,2]^3.  The
// traversal stack lives in LDS, one column per thread (`stride` apart): entry
// scale - min_scale - 1 holds the parent's Morton prefix (bit 31: pushed) and
// its t_max, for the leaf_level + 1 scales a push or pop can reach.
static constexpr int RT = 256;          // raycast threads per workgroup
static constexpr int RT_TILE = 16;      // pixel tile side per workgroup (RT = RT_TILE^2)
static_assert(RT_TILE == 2 * MAP_TILE, "each raycast wave covers one model-map tile");
static constexpr int MAX_LG = 10;       // leaf level <= 10: res <= 8192
template <bool STATS>
__device__ float ray_first_block(const VolDev& v, f3 o, f3 d_in, float nearp, float farp, float eps,
                                 uint32_t* __restrict__ st_par, float* __restrict__ st_tmax, unsigned& nnodes) {
  constexpr int DEPTH = 23;
  const int lg = v.leaf_level;                   // log2(res/8)
  const int min_scale = DEPTH - lg;
  float d[3] = {d_in.x, d_in.y, d_in.z};
  for (int a = 0; a < 3; a++) if (fabsf(d[a]) < eps) d[a] = copysignf(eps, d[a]);
  const float oo[3] = {o.x, o.y, o.z};
  float t_coef[3], t_bias[3], pos[3] = {1.f, 1.f, 1.f};
  for (int a = 0; a < 3; a++) {
    const float so = oo[a] / v.size + 1.f;
    t_coef[a] = -1.f / fabsf(d[a]);
    t_bias[a] = t_coef[a] * so;
  }
  int octant_mask = 7;
  for (int a = 0; a < 3; a++)
    if (d[a] > 0.0f) { octant_mask ^= (1 << a); t_bias[a] = 3.0f * t_coef[a] - t_bias[a]; }
  float t_min = fmax_r(fmax_r(2.0f * t_coef[0] - t_bias[0], 2.0f * t_coef[1] - t_bias[1]), 2.0f * t_coef[2] - t_bias[2]);
  float t_max = fmin_r(fmin_r(t_coef[0] - t_bias[0], t_coef[1] - t_bias[1]), t_coef[2] - t_bias[2]);
  float h = t_max;
  t_min = fmax_r(t_min, nearp / v.size);
  t_max = fmin_r(t_max, farp / v.size);
  int idx = 0;
  for (int a = 0; a < 3; a++)
    if (1.5f * t_coef[a] - t_bias[a] > t_min) { idx ^= (1 << a); pos[a] = 1.5f; }
  int scale = DEPTH - 1;
  float scale_exp2 = 0.5f;
  uint32_t parent = 0;          // Morton prefix of the parent node (root: 0)
  bool parent_ok = true;
  for (int i = 0; i <= lg; i++) st_par[i * RT] = 0u;
  float t_corner[3] = {0, 0, 0}, tc_max = 0.f;
  const float vs = v.size / (float)v.res;
  for (int guard = 0; scale < DEPTH; guard++) {
    if (!parent_ok || guard >= 8192) return -1.f;   // bounded walk
    if (STATS) nnodes++;
    for (int a = 0; a < 3; a++) t_corner[a] = pos[a] * t_coef[a] - t_bias[a];
    tc_max = fmin_r(fmin_r(t_corner[0], t_corner[1]), t_corner[2]);
    const int cidx = idx ^ octant_mask ^ 7;
    const int clevel = DEPTH - scale;            // level of the child examined
    const uint32_t child = (parent << 3) | (uint32_t)cidx;
    const bool has_child = child_exists(v, clevel, child);
    if (scale == min_scale && has_child) return t_min * v.size / vs;
    if (has_child && t_min <= t_max) {
      // descend
      const float tv_max = fmin_r(t_max, tc_max);
      const float half = scale_exp2 * 0.5f;
      float t_center[3];
      for (int a = 0; a < 3; a++) t_center[a] = half * t_coef[a] + t_corner[a];
      if (tc_max < h) {
        const int e = (scale - min_scale - 1) * RT;   // scale > min_scale here
        st_par[e] = parent | (parent_ok ? 0x80000000u : 0u);
        st_tmax[e] = t_max;
      }
      h = tc_max;
      parent = child;
      scale--;
      scale_exp2 = half;
      idx = (t_center[0] > t_min) | ((t_center[1] > t_min) << 1) | ((t_center[2] > t_min) << 2);
      // pos + scale_exp2 * bit, as a select: pos + 0 is pos (pos in [1, 2))
      for (int a = 0; a < 3; a++) pos[a] = (idx & (1 << a)) ? pos[a] + scale_exp2 : pos[a];
      t_max = tv_max;
      continue;
    }
    // advance
    const int step_mask = (t_corner[0] <= tc_max) | ((t_corner[1] <= tc_max) << 1) | ((t_corner[2] <= tc_max) << 2);
    for (int a = 0; a < 3; a++) pos[a] = (step_mask & (1 << a)) ? pos[a] - scale_exp2 : pos[a];
    t_min = tc_max;
    idx ^= step_mask;
    if ((idx & step_mask) != 0) {
      unsigned diff = 0;
      for (int a = 0; a < 3; a++)
        if (step_mask & (1 << a)) diff |= (unsigned)(__float_as_int(pos[a]) ^ __float_as_int(pos[a] + scale_exp2));
      scale = (__float_as_int((float)diff) >> 23) - 127;
      scale_exp2 = __int_as_float((scale - DEPTH + 127) << 23);
      if (scale <= min_scale || scale > DEPTH) return -1.f;
      if (scale == DEPTH) return -1.f;            // popped past the root
      {
        const int e = (scale - min_scale - 1) * RT;
        const uint32_t sp = st_par[e];
        parent = sp & 0x7fffffffu;
        parent_ok = (sp >> 31) != 0;
        t_max = st_tmax[e];
      }
      int sh[3];
      for (int a = 0; a < 3; a++) { sh[a] = __float_as_int(pos[a]) >> scale; pos[a] = __int_as_float(sh[a] << scale); }
      idx = (sh[0] & 1) | ((sh[1] & 1) << 1) | ((sh[2] & 1) << 2);
      h = 0.0f;
    }
  }
  return -1.f;
}

// raycast(volume, origin, dir, tnear, tfar, mu, step, largestep)
// (kfusion/rendering_impl.hpp:80-140); returns t (0: no hit).
template <bool STATS>
__device__ float march(const VolDev& v, f3 origin, f3 dir, float nearp, float farp, float mu, float step,
                       float largestep, f3& hit, unsigned& nsteps, unsigned& ninterp) {
  const float size = v.size;
  const float inv_vs = (float)v.res / v.size;
  const f3 invR = mk3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
  const f3 tbot = cmul((-1.f) * invR, origin);
  const f3 ttop = cmul(invR, mk3(size - origin.x, size - origin.y, size - origin.z));
  const f3 tmn = mk3(fmin_r(ttop.x, tbot.x), fmin_r(ttop.y, tbot.y), fmin_r(ttop.z, tbot.z));
  const f3 tmx = mk3(fmax_r(ttop.x, tbot.x), fmax_r(ttop.y, tbot.y), fmax_r(ttop.z, tbot.z));
  const float largest_tmin = fmax_r(fmax_r(tmn.x, tmn.y), fmax_r(tmn.x, tmn.z));
  const float smallest_tmax = fmin_r(fmin_r(tmx.x, tmx.y), fmin_r(tmx.x, tmx.z));
  const float tnear = fmax_r(largest_tmin, nearp);
  const float tfar = fmin_r(smallest_tmax, farp);
  hit = mk3(0.f, 0.f, 0.f);
  if (!(tnear < tfar)) return 0.f;
  float t = tnear, stepsize = largestep;
  f3 p = origin + dir * t;
  VCache c;
  float f_t = interp<false>(v, c, p * inv_vs);
  if (STATS) ninterp++;
  float f_tt = 0;
  if (f_t > 0) {
    for (; t < tfar; t += stepsize) {
      const f3 q = p * inv_vs;
      if (STATS) nsteps++;
      f_tt = tsdf_at(v, c, (int)q.x, (int)q.y, (int)q.z);
      if (f_tt < 0.1f && f_tt >= -0.5f) { f_tt = interp<false>(v, c, q); if (STATS) ninterp++; }   // (double)f <= 0.1
      if (f_tt < 0) break;
      stepsize = fmax_r(f_tt * mu, step);
      p = p + stepsize * dir;
      f_t = f_tt;
    }
    if (f_tt < 0) {
      t = t + stepsize * f_tt / (f_t - f_tt);
      hit = origin + dir * t;
      return t;
    }
  }
  return 0.f;
}

// False only when no ray through the pixel tile [x0, x0 + n) x [y0, y0 + n)
// can meet the object's volume cube [0, size]^3: every cube corner is in
// front of the camera and the tile lies more than 2 pixels outside the
// bounding box of their projections (the projection of a convex body in
// front of the camera is the convex hull of its corners' projections).
// Rays through pixel (x, y) follow view * (x, y, 1), so the corners are
// projected with inverse(view) = K * inverse(T_w_c).  A ray that misses the
// cube is clipped to nothing by march(), so skipping it changes no result.
__device__ bool tile_meets_cube(const ObjParams& P, float size, int x0, int y0, int n) {
  const Mat4 Vi = inverse(P.view);
  float umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
  for (int c = 0; c < 8; c++) {
    const f3 q = affine(Vi, affine(P.pose, mk3((c & 1) ? size : 0.f, (c & 2) ? size : 0.f, (c & 4) ? size : 0.f)));
    if (!(q.z > 1e-3f)) return true;   // corner behind or at the camera plane: no bound
    const float u = q.x / q.z, v = q.y / q.z;
    umin = fminf(umin, u); umax = fmaxf(umax, u); vmin = fminf(vmin, v); vmax = fmaxf(vmax, v);
  }
  const float m = 2.f;
  return !((float)(x0 + n - 1) < umin - m || (float)x0 > umax + m || (float)(y0 + n - 1) < vmin - m ||
           (float)y0 > vmax + m);
}

template <bool STATS>
__global__ __launch_bounds__(RT) void k_raycast(const RayObj* __restrict__ objs, int nobj, int w, int h, float mu,
                                                 float nearp, float farp, float eps, float* __restrict__ vx,
                                                 float* __restrict__ vy, float* __restrict__ vz,
                                                 float* __restrict__ nx, float* __restrict__ ny,
                                                 float* __restrict__ nz, int32_t* __restrict__ label,
                                                 unsigned long long* __restrict__ stats,
                                                 const int* __restrict__ in_view, const int* __restrict__ gate,
                                                 int* __restrict__ present) {
  if (gate && *gate == 0) return;
  __shared__ unsigned cnt[5];
  __shared__ uint32_t st_par[MAX_LG + 1][RT];
  __shared__ float st_tmax[MAX_LG + 1][RT];
  __shared__ uint64_t tile_objs;   // bit oi: this tile's rays may meet object oi's volume
  // fg of each volume at the hit point 0: every miss (hit 0) reads it
  // (rendering.cpp:91-92), so it is looked up once per workgroup instead of
  // once per ray and object
  __shared__ float fg_miss[MAX_INST];
  if (threadIdx.x < 5) cnt[threadIdx.x] = 0;
  if (threadIdx.x < 64) {
    const int oi = threadIdx.x;
    const bool meets = oi >= nobj ||
                       tile_meets_cube(*objs[oi].prm, objs[oi].v.size, blockIdx.x * RT_TILE, blockIdx.y * RT_TILE, RT_TILE);
    const uint64_t b = __ballot(meets);
    if (oi == 0) tile_objs = b;
  } else if (threadIdx.x < 64 + nobj) {
    const int oi = threadIdx.x - 64;
    VCache c;
    fg_miss[oi] = interp<true>(objs[oi].v, c, mk3(0.f, 0.f, 0.f));
  }
  __syncthreads();
  // 16x16 pixel tile per workgroup, 8x8 per wave: neighbouring rays visit
  // the same blocks, so a wave's gathers share cache lines and its lanes
  // march similar distances
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int x = blockIdx.x * RT_TILE + (wv & 1) * 8 + (lane & 7);
  const int y = blockIdx.y * RT_TILE + (wv >> 1) * 8 + (lane >> 3);
  unsigned nsteps = 0, ninterp = 0, nhits = 0, npairs = 0, nnodes = 0;
  int surf = -1;   // object whose maps get a surface pixel here (not the miss constants)
  int plabel = -1; // the pixel's label, for the instance set
  if (x < w && y < h) {
  const size_t pix = (size_t)x + (size_t)y * w;
  float nearest = 1.0e3f;
  float high_fg = 0;   // fg values are floats widened to double: float compares are exact
  // every object either takes the pixel or counts as a miss, so "missed by
  // every object" is "no take": win < 0
  int win = -1;                  // object of the last take, and its hit (object frame, metres)
  f3 win_hit = mk3(0.f, 0.f, 0.f);
  // should_raycast_all when the in-view set is empty (rendering.cpp:16, 27-33)
  bool all = true;
  if (in_view)
    for (int oi = 0; oi < nobj; oi++) if (in_view[objs[oi].label]) all = false;
  for (int oi = 0; oi < nobj; oi++) {
    const RayObj& O = objs[oi];
    if (!all && !in_view[O.label]) continue;
    const VolDev v = O.v;
    f3 hit = mk3(0.f, 0.f, 0.f);
    float hw = 0.f;
    if (STATS) npairs++;
    // rays of a tile outside the object's projected volume: the miss outcome
    // (no hit, hit point 0) without the traversal
    if ((tile_objs >> oi) & 1) {
      const ObjParams& P = *O.prm;
      const float step = v.size / (float)v.res;
      const float largestep = step * 8;
      const f3 wo = translation(P.view);
      const f3 oo = affine(P.inv_pose, wo);
      const f3 wd = rotate(P.view, mk3((float)x, (float)y, 1.f));
      const f3 od = normalize(rotate(P.inv_pose, wd));
      const float nfar = fmin_r(nearest, farp);
      const float tmin = ray_first_block<STATS>(v, oo, od, nearp, nfar, eps, &st_par[0][threadIdx.x],
                                                &st_tmax[0][threadIdx.x], nnodes);
      if (tmin > 0.f)
        hw = march<STATS>(v, oo, od, tmin * v.size / (float)v.res, nfar, mu, step, largestep, hit, nsteps, ninterp);
    }
    const float hd = (O.label == 0) ? hw + 0.05f : hw;
    if (hd > nearest) continue;
    const float inv_vs = (float)v.res / v.size;
    VCache c;
    // a miss leaves hit 0 (a hit has t >= the near plane > 0)
    const float fg = hw == 0.f ? fg_miss[oi] : interp<true>(v, c, hit * inv_vs);
    if (STATS) ninterp++;
    if (fg <= 0.5f) continue;
    bool take = false;
    if (hd == nearest) { if (fg > high_fg) take = true; }
    else if (hd < nearest) { nearest = hd; take = true; }
    if (!take) continue;
    high_fg = fg;
    win = oi;
    win_hit = hit;
    if (STATS) nhits++;
  }
  f3 out_v = mk3(0.f, 0.f, 0.f), out_n = mk3((float)INVALID, 0.f, 0.f);
  int out_l = INVALID;
  if (win >= 0) {
    // vertex and normal of the last take (the values a later take would
    // overwrite are never needed, so the gradient runs once per pixel)
    const RayObj& O = objs[win];
    const VolDev v = O.v;
    const ObjParams& P = *O.prm;
    out_l = O.label;
    out_v = affine(P.pose, win_hit);
    if (nx) {
      VCache c;
      const f3 g = grad(v, c, win_hit * ((float)v.res / v.size));
      out_n = (length(g) == 0) ? mk3((float)INVALID, 0.f, 0.f) : normalize(rotate(P.pose, g));
    }
  }
  if (vx) { vx[pix] = out_v.x; vy[pix] = out_v.y; vz[pix] = out_v.z; }
  if (nx) { nx[pix] = out_n.x; ny[pix] = out_n.y; nz[pix] = out_n.z; }
  label[pix] = out_l;
  plabel = out_l;
  for (int oi = 0; oi < nobj; oi++) {
    const RayObj& O = objs[oi];
    const bool mine = (out_l == O.label);
    if (O.mvx) { O.mvx[pix] = mine ? out_v.x : 0.f; O.mvy[pix] = mine ? out_v.y : 0.f; O.mvz[pix] = mine ? out_v.z : 0.f; }
    if (O.mnx) { O.mnx[pix] = mine ? out_n.x : (float)INVALID; O.mny[pix] = mine ? out_n.y : 0.f; O.mnz[pix] = mine ? out_n.z : 0.f; }
    if (mine && O.tiles &&
        (__float_as_uint(out_v.x) | __float_as_uint(out_v.y) | __float_as_uint(out_v.z) | __float_as_uint(out_n.y) |
         __float_as_uint(out_n.z) | (__float_as_uint(out_n.x) ^ __float_as_uint((float)INVALID))) != 0u)
      surf = oi;
  }
  }
  // The instance set of the label image: one same-value store per distinct
  // label of the wave.
  if (present) {
    bool todo = plabel >= 0 && plabel < MAX_INST;
    for (uint64_t b = __ballot(todo); b != 0; b = __ballot(todo)) {
      const int first = __ffsll((unsigned long long)b) - 1;
      const int l0 = __shfl(plabel, first);
      if (plabel == l0) todo = false;
      if (lane == first) present[l0] = 1;
    }
  }
  // Tile flags of the objects' model maps (TrackInst::tiles; each wave
  // covers one 8x8 tile): 1 when a pixel of the tile got anything but the
  // miss constants, and the box of those tiles, which the object rows use to
  // skip model-map loads.
  {
    const int ntx = (w + MAP_TILE - 1) / MAP_TILE, nty = (h + MAP_TILE - 1) / MAP_TILE;
    const int ttx = (int)blockIdx.x * 2 + (wv & 1), tty = (int)blockIdx.y * 2 + (wv >> 1);
    for (int oi = 0; oi < nobj; oi++) {
      const RayObj& O = objs[oi];
      if (!O.tiles) continue;
      const uint64_t b = __ballot(surf == oi);
      if (lane == 0 && ttx < ntx && tty < nty) {
        O.tiles[tty * ntx + ttx] = b != 0 ? 1 : 0;
        if (b != 0) {
          atomicMin(&O.tbox[0], ttx); atomicMin(&O.tbox[1], tty);
          atomicMax(&O.tbox[2], ttx); atomicMax(&O.tbox[3], tty);
        }
      }
    }
  }
  if (!STATS) return;
  atomicAdd(&cnt[0], nsteps); atomicAdd(&cnt[1], ninterp); atomicAdd(&cnt[2], nhits); atomicAdd(&cnt[3], npairs);
  atomicAdd(&cnt[4], nnodes);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long* sl = stats + stat_slot();
    atomicAdd(&sl[2 * STAT_SLOTS], (unsigned long long)cnt[0]); atomicAdd(&sl[3 * STAT_SLOTS], (unsigned long long)cnt[1]);
    atomicAdd(&sl[4 * STAT_SLOTS], (unsigned long long)cnt[2]); atomicAdd(&sl[7 * STAT_SLOTS], (unsigned long long)cnt[3]);
    atomicAdd(&sl[8 * STAT_SLOTS], (unsigned long long)cnt[4]);
  }
}

// ---- renderVolume_many_Kernel (rendering/rendering.cpp:489-646)

// Interpolated colour channel at `pos` (voxel units); missing voxels are
// voxel_traits<SDF>::empty(), colour 0 (volume_traits.hpp:36).
__device__ __forceinline__ float interp_col(const VolDev& v, VCache& c, f3 pos, const float* __restrict__ field) {
  const float flx = floorf(pos.x), fly = floorf(pos.y), flz = floorf(pos.z);
  const int bx = (int)flx, by = (int)fly, bz = (int)flz;
  const float fx = pos.x - flx, fy = pos.y - fly, fz = pos.z - flz;
  const int lx = max(bx, 0), ly = max(by, 0), lz = max(bz, 0);
  float p[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int X = lx + (i & 1), Y = ly + ((i >> 1) & 1), Z = lz + ((i >> 2) & 1);
    const int32_t s = cslot(v, c, X, Y, Z);
    p[i] = s < 0 ? 0.f : field[vox(s, X, Y, Z)];
  }
  return (((p[0] * (1 - fx) + p[1] * fx) * (1 - fy) + (p[2] * (1 - fx) + p[3] * fx) * fy) * (1 - fz) +
          ((p[4] * (1 - fx) + p[5] * fx) * (1 - fy) + (p[6] * (1 - fx) + p[7] * fx) * fy) * fz);
}
__device__ __forceinline__ float clamp01(float f) { return fmaxf(0.f, fminf(f, 1.f)); }
__device__ __forceinline__ uint32_t col_u8(float v) { return (uint32_t)(int)v & 0xffu; }

// Shading of one pixel from a world hit, its normal and label: Lambert term
// towards the camera plus ambient, times the interpolated voxel colour (or the
// label's palette colour, white for the background).
__device__ uint32_t shade(const RayObj* __restrict__ objs, int nobj, f3 w_origin, f3 w_hit, f3 w_n, int w_label,
                          int render_color, const uint32_t* __restrict__ palette, int npal) {
  const float amb = 0.1f;   // ambient (constant_parameters.h:24)
  f3 col = mk3(0.f, 0.f, 0.f);
  for (int oi = 0; oi < nobj; oi++) {
    const RayObj& O = objs[oi];
    if (w_label != O.label) continue;
    const VolDev v = O.v;
    const ObjParams& P = *O.prm;
    const f3 o_hit = affine(P.inv_pose, w_hit);
    const f3 diff = normalize(w_origin - w_hit);
    const float dir = fmaxf(dot(normalize(w_n), diff), 0.f);
    if (render_color) {
      const f3 q = o_hit * ((float)v.res / v.size);
      VCache c;
      const float r = interp_col(v, c, q, v.r), g = interp_col(v, c, q, v.g), b = interp_col(v, c, q, v.b);
      col = mk3(clamp01(dir * r + amb), clamp01(dir * g + amb), clamp01(dir * b + amb)) * 255.f;
    } else if (w_label == 0) {
      col = mk3(clamp01(dir + amb), clamp01(dir + amb), clamp01(dir + amb)) * 255.f;
    } else {
      const uint32_t pc = (w_label >= 0 && w_label < npal) ? palette[w_label] : 0u;
      const f3 rgb = mk3((float)(pc & 0xffu) / 255, (float)((pc >> 8) & 0xffu) / 255, (float)((pc >> 16) & 0xffu) / 255);
      col = mk3(clamp01(dir * rgb.x + amb), clamp01(dir * rgb.y + amb), clamp01(dir * rgb.z + amb)) * 255.f;
    }
  }
  return col_u8(col.x) | (col_u8(col.y) << 8) | (col_u8(col.z) << 16);
}

// From the last raycast's maps (the view is the raycast pose, same size).
__global__ void k_render_volume_maps(const RayObj* __restrict__ objs, int nobj, int n, const Mat4* __restrict__ T_w_c,
                                     const float* __restrict__ vx, const float* __restrict__ vy,
                                     const float* __restrict__ vz, const float* __restrict__ nx,
                                     const float* __restrict__ ny, const float* __restrict__ nz,
                                     const int32_t* __restrict__ label, int render_color,
                                     const uint32_t* __restrict__ palette, int npal, uint32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = shade(objs, nobj, translation(*T_w_c), mk3(vx[i], vy[i], vz[i]), mk3(nx[i], ny[i], nz[i]), label[i],
                 render_color, palette, npal);
}

// Raycasting at another view or size: every object, the take rule of
// renderVolume_many_Kernel (a later hit at the same distance needs a higher
// foreground probability; a take needs hit.w > 0 and sets the label only
// where the gradient is non-zero), then shading.
__global__ __launch_bounds__(RT) void k_render_volume_ray(const RayObj* __restrict__ objs, int nobj, int w, int h,
                                                          const Mat4* __restrict__ T_w_c, float mu, float nearp,
                                                          float farp, float largestep, float eps, int render_color,
                                                          const uint32_t* __restrict__ palette, int npal,
                                                          uint32_t* __restrict__ out) {
  __shared__ uint32_t st_par[MAX_LG + 1][RT];
  __shared__ float st_tmax[MAX_LG + 1][RT];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int x = blockIdx.x * RT_TILE + (wv & 1) * 8 + (lane & 7);
  const int y = blockIdx.y * RT_TILE + (wv >> 1) * 8 + (lane >> 3);
  if (x >= w || y >= h) return;
  const f3 w_origin = translation(*T_w_c);
  float nearest = 1.0e10f;
  float high_fg = 0.f;   // floats widened to double in the reference: float compares are exact
  int miss = 0;
  f3 w_hit = mk3(0.f, 0.f, 0.f), w_n = mk3(0.f, 0.f, 0.f);
  int w_label = 0;
  unsigned dummy = 0;
  for (int oi = 0; oi < nobj; oi++) {
    const RayObj& O = objs[oi];
    const VolDev v = O.v;
    const ObjParams& P = *O.prm;
    const float step = v.size / (float)v.res;
    const f3 oo = affine(P.inv_pose, w_origin);
    const f3 wd = rotate(P.view, mk3((float)x, (float)y, 1.f));
    const f3 od = normalize(rotate(P.inv_pose, wd));
    const float nfar = fminf(nearest, farp);
    const float tmin = ray_first_block<false>(v, oo, od, nearp, nfar, eps, &st_par[0][threadIdx.x],
                                              &st_tmax[0][threadIdx.x], dummy);
    f3 hit = mk3(0.f, 0.f, 0.f);
    float hw = 0.f;
    if (tmin > 0.f) hw = march<false>(v, oo, od, tmin * v.size / (float)v.res, nfar, mu, step, largestep, hit, dummy, dummy);
    const float hd = (O.label == 0) ? hw + 0.05f : hw;
    if (hd > nearest) { miss++; continue; }
    const float inv_vs = (float)v.res / v.size;
    VCache c;
    const float fg = interp<true>(v, c, hit * inv_vs);
    if (fg <= 0.5f) { miss++; continue; }
    if (hd == nearest) {
      if (fg > high_fg) high_fg = fg;
      else { miss++; continue; }
    }
    if (hw > 0.f) {
      nearest = hd;
      const f3 g = grad(v, c, hit * inv_vs);
      if (length(g) > 0.f) {
        w_n = rotate(P.pose, g);
        w_hit = affine(P.pose, hit);
        w_label = O.label;
      }
    } else {
      miss++;
    }
  }
  const size_t pix = (size_t)x + (size_t)y * w;
  out[pix] = (miss == nobj) ? 0u : shade(objs, nobj, w_origin, w_hit, w_n, w_label, render_color, palette, npal);
}

void launch_render_volume_maps(hipStream_t s, const RayObj* objs, int nobj, int n, const Mat4* T_w_c,
                               float* const v[3], float* const nrm[3], const int32_t* label, int render_color,
                               const uint32_t* palette, int npal, uint32_t* out) {
  k_render_volume_maps<<<(n + 255) / 256, 256, 0, s>>>(objs, nobj, n, T_w_c, v[0], v[1], v[2], nrm[0], nrm[1], nrm[2],
                                                       label, render_color, palette, npal, out);
}
void launch_render_volume_ray(hipStream_t s, const RayObj* objs, int nobj, int w, int h, const Mat4* T_w_c, float mu,
                              float nearp, float farp, float largestep, float eps, int render_color,
                              const uint32_t* palette, int npal, uint32_t* out) {
  dim3 blk(RT), grd((w + RT_TILE - 1) / RT_TILE, (h + RT_TILE - 1) / RT_TILE);
  k_render_volume_ray<<<grd, blk, 0, s>>>(objs, nobj, w, h, T_w_c, mu, nearp, farp, largestep, eps, render_color,
                                          palette, npal, out);
}

void launch_raycast(hipStream_t s, const RayObj* objs, int nobj, int w, int h, float mu, float nearp, float farp,
                    float eps, float* vx, float* vy, float* vz, float* nx, float* ny, float* nz, int32_t* label,
                    unsigned long long* stats, const int* in_view, const int* gate, int* present) {
  dim3 blk(RT), grd((w + RT_TILE - 1) / RT_TILE, (h + RT_TILE - 1) / RT_TILE);
  if (stats)
    k_raycast<true><<<grd, blk, 0, s>>>(objs, nobj, w, h, mu, nearp, farp, eps, vx, vy, vz, nx, ny, nz, label, stats,
                                        in_view, gate, present);
  else
    k_raycast<false><<<grd, blk, 0, s>>>(objs, nobj, w, h, mu, nearp, farp, eps, vx, vy, vz, nx, ny, nz, label,
                                         nullptr, in_view, gate, present);
}

}  // namespace mf
