// Voxel-block allocation and projective TSDF integration on gfx950.
//
// Reference: buildVolumeAllocationList (kfusion/alloc_impl.hpp:121-191),
// Octree::alloc_update (lib/octree.hpp:761-863), projective_functor
// (lib/functors/projective_functor.hpp:54-156), in_frustum
// (lib/algorithms/filter.hpp:42-51), sdf_update (kfusion/mapping_impl.hpp:
// 15-72), Object::integrate_volume_kernel / refine_mask_use_motion
// (object/object.cpp:128-195, 255-315).
//
// The octree is replaced by a Morton-indexed block table: slot[morton(bx,by,
// bz)] holds the pool slot of an allocated 8^3 block, so a band sample's
// lookup is one 4-byte load and allocation is a request flag plus a commit
// pass in Morton order.  Internal-node occupancy bytes per level keep what
// the ESVO raycast traversal needs.  Allocated key sets are bit-identical to
// the reference's keyops::encode keys.
#include "mf_kernels.h"

namespace mf {

__device__ __forceinline__ float refined_mask(const float* mask, const int8_t* codes, size_t i, size_t n, int refine,
                                              int use_icp, int use_rgb) {
  float m = mask[i];
  if (refine && m >= 0.f && codes) {
    if (use_icp && (codes[i] == -4 || codes[i] == -6)) return -1.f;
    if (use_rgb && codes[n + i] < -3) return -1.f;
  }
  return m;
}

__device__ void obj_params(const Mat4& Twc, const Mat4& Two, float fx, float fy, float cx, float cy, ObjParams* out) {
  const Mat4 K = camK(fx, fy, cx, cy);
  ObjParams p;
  p.pose = Two;
  p.inv_pose = inverse(Two);
  p.Tco = mul(inverse(Twc), Two);
  p.KT = mul(K, p.Tco);
  p.kPose = mul(Twc, inverse(K));
  p.o_cam = affine(p.inv_pose, translation(Twc));
  p.view = mul(Twc, camKinv(fx, fy, cx, cy));
  *out = p;
}
// Matrices of every object of a stage (one thread each), the start of this
// frame's allocation bookkeeping with OBJ_ALLOC_RESET, and two small jobs of
// the launch that precedes a raycast: *copy_to = *T_w_c and clear[0 ..
// MAX_INST) = 0 (each optional).
__global__ void k_obj_params(const Mat4* T_w_c, const Mat4* mats, const MapObj* objs, int n, float fx, float fy,
                             float cx, float cy, int flags, int* act_count, Mat4* copy_to, int* clear) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if ((flags & OBJ_ALLOC_RESET) && j == 0) *act_count = 0;
  if (copy_to && j == 0) *copy_to = *T_w_c;
  if (clear && blockIdx.x == 0)
    for (int k = threadIdx.x; k < MAX_INST; k += blockDim.x) clear[k] = 0;
  if (j >= n) return;
  const MapObj& o = objs[j];
  obj_params(*T_w_c, mats[3 + o.id], fx, fy, cx, cy, o.prm);
  if (flags & OBJ_ALLOC_RESET) o.v.counters[3] = o.v.counters[0];
  if ((flags & OBJ_TBOX_RESET) && o.tbox) {
    o.tbox[0] = o.tbox[1] = 0x7fffffff;
    o.tbox[2] = o.tbox[3] = -1;
  }
}

// Band ray-march of buildVolumeAllocationList: every sample in a missing
// block raises that block's request flag; a sample in an allocated block sets
// the block active.
__global__ void k_alloc(const MapObj* __restrict__ objs, const float* __restrict__ depth, int refine, int use_icp,
                        int use_rgb, int w, int h, float band, unsigned long long* __restrict__ stats) {
  const MapObj& O = objs[blockIdx.z];
  const VolDev& v = O.v;
  const ObjParams* __restrict__ prm = O.prm;
  const float* __restrict__ mask = O.mask;
  const int8_t* __restrict__ codes = O.codes;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  const size_t i = (size_t)x + (size_t)y * w;
  const bool valid = x < w && y < h && depth[i] != 0.0f &&
                     refined_mask(mask, codes, i, (size_t)w * h, refine, use_icp, use_rgb) >= 0.f;
  const float inv_vox = 1.0f / v.voxel;
  const int num_steps = (int)ceilf(band * inv_vox);
  if (stats) {
    const int nvalid = __syncthreads_count(valid);
    if (threadIdx.x == 0 && nvalid) {
      unsigned long long* sl = stats + stat_slot();
      atomicAdd(&sl[6 * STAT_SLOTS], (unsigned long long)nvalid);
      atomicAdd(&sl[5 * STAT_SLOTS], (unsigned long long)nvalid * num_steps);
    }
  }
  if (!valid) return;
  const float d = depth[i];
  const f3 wv = affine(prm->kPose, mk3(((float)x + 0.5f) * d, ((float)y + 0.5f) * d, d));
  const f3 ov = affine(prm->inv_pose, wv);
  const f3 dir = normalize(prm->o_cam - ov);
  const f3 origin = ov - (band * 0.5f) * dir;
  const f3 step = (dir * band) / (float)num_steps;
  const float size = (float)v.res;
  // Samples in batches of ALLOC_BATCH: the batch's block-table loads are
  // issued together and its flag stores follow them.  (One sample at a time,
  // each byte store may alias the next table load, so every sample waited a
  // full memory round trip.)  A sample in the same block as the sample
  // before it repeats that sample's load and store, so it is skipped.  The
  // flags are set-to-1 stores, so the set of requested and active blocks is
  // the one the sample-by-sample loop leaves.  Block coordinates are below
  // 1024 (res <= 8192), so the 30-bit Morton code is the table index.
  constexpr int ALLOC_BATCH = 8;
  // descriptor pointers held in registers: a byte store may alias the
  // descriptor, which would otherwise be re-read before every store
  const int32_t* const slot = v.slot;
  __attribute__((address_space(1))) uint8_t* const req = (__attribute__((address_space(1))) uint8_t*)v.req;
  __attribute__((address_space(1))) uint8_t* const act = (__attribute__((address_space(1))) uint8_t*)v.blk_active;
  f3 p = origin;
  uint32_t last = 0xffffffffu;
  for (int s0 = 0; s0 < num_steps; s0 += ALLOC_BATCH) {
    uint32_t m[ALLOC_BATCH];
    bool use[ALLOC_BATCH];
#pragma unroll
    for (int k = 0; k < ALLOC_BATCH; k++) {
      use[k] = false;
      m[k] = 0;
      if (s0 + k < num_steps) {
        const float sx = floorf(p.x * inv_vox), sy = floorf(p.y * inv_vox), sz = floorf(p.z * inv_vox);
        if (sx < size && sy < size && sz < size && sx >= 0 && sy >= 0 && sz >= 0) {
          const uint32_t vx = (uint32_t)(int)sx, vy = (uint32_t)(int)sy, vz = (uint32_t)(int)sz;
          const uint32_t mk = morton3_10(vx >> 3, vy >> 3, vz >> 3);
          use[k] = mk != last;
          m[k] = mk;
          last = mk;
        }
        p = p + step;
      }
    }
    int32_t sl[ALLOC_BATCH];
#pragma unroll
    for (int k = 0; k < ALLOC_BATCH; k++) sl[k] = use[k] ? ldx(slot, m[k]) : 0;
    // one wait for the batch: the stores below then carry no load waits
#pragma unroll
    for (int k = 0; k < ALLOC_BATCH; k++) asm volatile("" ::"v"(sl[k]));
#pragma unroll
    for (int k = 0; k < ALLOC_BATCH; k++) {
      if (!use[k]) continue;
      if (sl[k] < 0) req[m[k]] = 1;
      else act[sl[k]] = 1;
    }
  }
}

// Commit: requested blocks receive pool slots (Octree::updateLevel: new leaf
// blocks start active, octree.hpp:846-851); their ancestors are marked in the
// per-level node-occupancy bytes.
// Wave-aggregated append: one atomic per wave on `counter`, ranks by ballot.
// Must be reached by every lane of the wave.
__device__ __forceinline__ int wave_append(int* counter, bool pred) {
  const unsigned long long m = __ballot(pred);
  if (!m) return -1;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader);
  return pred ? base + __popcll(m & ((1ull << lane) - 1)) : -1;
}

// Octree::alloc_update commit over the concatenated Morton tables of the
// launch's objects (prefix `wl.table_pre`): every requested block gets a pool
// slot, its path's internal-node occupancy is set, and (object, slot) joins
// this frame's new-block list.
__global__ void k_alloc_commit(const MapObj* __restrict__ objs, int n, MapWork wl) {
  const long total = wl.table_pre[n];
  const long stride = (long)gridDim.x * blockDim.x;
  for (long a0 = (long)blockIdx.x * blockDim.x; a0 < total; a0 += stride) {
    const long a = a0 + threadIdx.x;
    int j = 0;
    while (j + 1 < n && wl.table_pre[j + 1] <= a) j++;
    const VolDev& v = objs[j].v;
    const long m = a - wl.table_pre[j];
    bool fresh = false;
    int s = -1;
    if (a < total && v.req[m]) {
      v.req[m] = 0;
      s = atomicAdd(&v.counters[0], 1);
      if (s >= v.cap) { v.counters[1] = 1; atomicSub(&v.counters[0], 1); }
      else fresh = true;
    }
    if (fresh) {
      v.slot[m] = s;
      v.blk_morton[s] = (int32_t)m;
      v.blk_active[s] = 1;
      for (int l = v.leaf_level - 1; l >= 1; l--) v.nodes[v.node_off[l] + (m >> (3 * (v.leaf_level - l)))] = 1;
    }
  }
}

// Voxels of the blocks with the given keys (keyops::encode of a leaf block:
// the block's Morton index shifted left by 9, level in the low bits), one
// 512-thread workgroup per key: out[k][voxel x + 8y + 64z][8] as double
// {tsdf, w, r, g, b, c, fg, fgw}.  A key without an allocated block gives
// voxel_traits<SDF>::initValue() (tsdf 1, everything else 0) and found 0.
__global__ __launch_bounds__(512) void k_export_voxels(VolDev v, const uint64_t* __restrict__ keys, int n,
                                                       double* __restrict__ out, int* __restrict__ found) {
  const uint64_t nb3 = (uint64_t)v.nb * v.nb * v.nb;
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    const uint64_t m = keys[k] >> 9;
    const int32_t s = m < nb3 ? v.slot[m] : -1;
    const int i = threadIdx.x;
    double* d = out + ((size_t)k * BLOCK_VOX + i) * 8;
    if (s < 0) {
      d[0] = 1.0; d[1] = 0.0; d[2] = 0.0; d[3] = 0.0; d[4] = 0.0; d[5] = 0.0; d[6] = 0.0; d[7] = 0.0;
    } else {
      const size_t vi = (size_t)s * BLOCK_VOX + i;
      d[0] = v.tsdf[vi]; d[1] = v.wt[vi]; d[2] = v.r[vi]; d[3] = v.g[vi]; d[4] = v.b[vi];
      d[5] = v.c[vi]; d[6] = v.fg[vi]; d[7] = v.fgw[vi];
    }
    if (i == 0) found[k] = s >= 0 ? 1 : 0;
  }
}
void launch_export_voxels(hipStream_t st, const VolDev& v, const uint64_t* keys, int n, double* out, int* found) {
  if (n > 0) k_export_voxels<<<n < 4096 ? n : 4096, BLOCK_VOX, 0, st>>>(v, keys, n, out, found);
}

__device__ __forceinline__ void block_coords(int32_t m, int& x, int& y, int& z) {
  x = (int)compact21((uint64_t)m) * 8; y = (int)compact21((uint64_t)m >> 1) * 8; z = (int)compact21((uint64_t)m >> 2) * 8;
}

// projective_functor::build_active_list: active || in_frustum(block corner),
// over the concatenated block pools of the launch's objects (the prefix of
// their allocated counts, computed by every workgroup); entries (object,
// slot, Morton index of the block) of one list for the whole launch, counted
// in wl.act_count (zeroed by k_obj_params at the start of the mapping pass).
__global__ void k_active_list(const MapObj* __restrict__ objs, int n, int w, int h, MapWork wl) {
  __shared__ long pre[MAX_INST + 1];
  if (threadIdx.x == 0) {
    long t = 0;
    pre[0] = 0;
    for (int j = 0; j < n; j++) { t += min(objs[j].v.counters[0], (int)objs[j].v.cap); pre[j + 1] = t; }
  }
  __syncthreads();
  const long total = pre[n];
  const long stride = (long)gridDim.x * blockDim.x;
  for (long a0 = (long)blockIdx.x * blockDim.x; a0 < total; a0 += stride) {
    const long a = a0 + threadIdx.x;
    int j = 0;
    while (j + 1 < n && pre[j + 1] <= a) j++;
    bool keep = false;
    const int s = (int)(a - pre[j]);
    int32_t mor = 0;
    if (a < total) {
      const VolDev& v = objs[j].v;
      mor = v.blk_morton[s];
      keep = v.blk_active[s] != 0;
      if (!keep) {
        int bx, by, bz;
        block_coords(mor, bx, by, bz);
        const f3 bc = mk3((float)bx * v.voxel, (float)by * v.voxel, (float)bz * v.voxel);
        const f3 vc = affine(objs[j].prm->KT, bc);
        const int px = (int)(vc.x / vc.z), py = (int)(vc.y / vc.z);
        keep = px >= 0 && px < w && py >= 0 && py < h;
      }
    }
    const int slot = wave_append(wl.act_count, keep);
    if (keep) wl.act_list[slot] = make_int4(j, s, mor, 0);
  }
}

// update_block + sdf_update: one 512-thread workgroup per active block, one
// voxel per thread (x fastest, the block's storage order).  The row start
// and x-increment are formed exactly as the reference forms them.
//
// Memory round trips per block: the list entry (object, slot, Morton index;
// prefetched one block ahead), one batch of pixel gathers (depth, mask and,
// for the motion-refined mask, both track codes) and one batch of voxel
// fields with the colour gathers; the stores come last.  Each object's
// descriptor fields are staged in LDS once per workgroup (the voxel stores
// may alias the descriptor as far as the compiler knows, which would
// otherwise force re-reads).  A block allocated this frame (slot at or past
// the frame-start count) holds voxel_traits<SDF>::initValue()
// (continuous/volume_traits.hpp:38) before its first update, so its voxels
// take those values without loading them and store them whether they update
// or not.
struct IntegObj {
  Mat4 Tco;
  float *tsdf, *wt, *r, *g, *b;
  int32_t *c, *fgw;
  double* fg;
  uint8_t* blk_active;
  const float* mask;
  const int8_t* codes;
  float voxel;
  int start_count;
};
__global__ __launch_bounds__(512) void k_integrate(const MapObj* __restrict__ objs, int n, float fx, float fy,
                                                   float cx, float cy, const float* __restrict__ depth,
                                                   const float* __restrict__ rr, const float* __restrict__ gg,
                                                   const float* __restrict__ bb, int refine, int use_icp, int use_rgb,
                                                   int w, int h, float mu, int maxweight, MapWork wl,
                                                   unsigned long long* __restrict__ stats) {
  __shared__ IntegObj iobj[MAX_INST];
  const int tid = threadIdx.x;
  const int nact = *wl.act_count;
  if ((int)blockIdx.x >= nact) return;   // uniform: no block for this workgroup
  for (int j = tid; j < n; j += (int)blockDim.x) {
    const MapObj& O = objs[j];
    IntegObj& d = iobj[j];
    d.Tco = O.prm->Tco;
    d.tsdf = O.v.tsdf; d.wt = O.v.wt; d.r = O.v.r; d.g = O.v.g; d.b = O.v.b; d.c = O.v.c; d.fgw = O.v.fgw;
    d.fg = O.v.fg; d.blk_active = O.v.blk_active; d.mask = O.mask; d.codes = O.codes;
    d.voxel = O.v.voxel; d.start_count = O.v.counters[3];
  }
  const int lx = tid & 7, ly = (tid >> 3) & 7, lz = tid >> 6;
  const Mat4 K = camK(fx, fy, cx, cy);
  const size_t npx = (size_t)w * h;
  const int G = (int)gridDim.x;
  int4 next = wl.act_list[blockIdx.x];
  __syncthreads();
  for (int a = blockIdx.x; a < nact; a += G) {
    const int4 e = next;
    if (a + G < nact) next = wl.act_list[a + G];
    const int s = e.y;
    const IntegObj& D = iobj[e.x];
    const Mat4 Tco = D.Tco;
    const float vs = D.voxel;
    const bool fresh = s >= D.start_count;
    const int32_t morton = e.z;
    float* const tsdf = D.tsdf;
    float* const wt = D.wt;
    float* const vr = D.r;
    float* const vg = D.g;
    float* const vb = D.b;
    int32_t* const vc = D.c;
    int32_t* const vfgw = D.fgw;
    double* const vfg = D.fg;
    uint8_t* const blk_active = D.blk_active;
    const float* const mask = D.mask;
    const int8_t* const codes = D.codes;
    const f3 delta = rotate(Tco, mk3(vs, 0.f, 0.f));
    const f3 cdelta = rotate(K, delta);
    int bx, by, bz;
    block_coords(morton, bx, by, bz);
    const int y = by + ly, z = bz + lz;
    const f3 start = affine(Tco, mk3((float)bx * vs, (float)y * vs, (float)z * vs));
    const f3 cstart = affine(K, start);
    const f3 cv = cstart + (float)lx * cdelta;
    const f3 pos = start + (float)lx * delta;
    bool vis = false;
    size_t pi = 0;
    if (!(pos.z < 0.0001f)) {
      const float invd = 1.f / cv.z;
      const float pxf = cv.x * invd + 0.5f, pyf = cv.y * invd + 0.5f;
      if (!(pxf < 0.5f || pxf > (float)w - 1.5f || pyf < 0.5f || pyf > (float)h - 1.5f)) {
        vis = true;
        pi = (size_t)(unsigned)pxf + (size_t)w * (unsigned)pyf;
      }
    }
    // pixel gathers, one batch (refine_mask_use_motion reads both codes)
    const bool use_codes = refine && codes;
    float ds = 0.f, m = -1.f;
    int8_t ci = 0, cr = 0;
    if (vis) {
      ds = depth[pi];
      m = mask[pi];
      if (use_codes) { ci = codes[pi]; cr = codes[npx + pi]; }
    }
    if (use_codes && m >= 0.f) {
      if (use_icp && (ci == -4 || ci == -6)) m = -1.f;
      else if (use_rgb && cr < -3) m = -1.f;
    }
    bool upd = false;
    float sdf = 0.f;
    if (vis && !(ds <= 0.f || m < 0.f)) {
      const float qx = pos.x / pos.z, qy = pos.y / pos.z;
      const float diff = (ds - pos.z) * sqrtf(1.f + qx * qx + qy * qy);
      if (diff > -mu) {
        upd = true;
        sdf = fmin_r(1.f, diff / mu);
      }
    }
    // voxel fields and colour gathers, one batch
    const size_t vi = (size_t)s * BLOCK_VOX + tid;
    float t0 = 1.f, w0 = 0.f, r0 = 0.f, g0 = 0.f, b0 = 0.f, pr = 0.f, pg = 0.f, pb = 0.f;
    int c0 = 0, fw = 0;
    double fg0 = 0.0;
    if (upd) {
      if (!fresh) {
        t0 = tsdf[vi]; w0 = wt[vi]; r0 = vr[vi]; g0 = vg[vi]; b0 = vb[vi];
        c0 = vc[vi]; fw = vfgw[vi]; fg0 = vfg[vi];
      }
      pr = rr[pi]; pg = gg[pi]; pb = bb[pi];
    }
    if (upd) {
      tsdf[vi] = fmax_r(-1.f, fmin_r((w0 * t0 + sdf) / (w0 + 1), 1.f));
      wt[vi] = fmin_r(w0 + 1, (float)maxweight);
      const float fc = (float)c0, fc1 = (float)(c0 + 1);
      vr[vi] = (pr + r0 * fc) / fc1;
      vg[vi] = (pg + g0 * fc) / fc1;
      vb[vi] = (pb + b0 * fc) / fc1;
      vc[vi] = c0 + 1;
      vfg[vi] = ((double)m + fg0 * (double)fw) / (double)(fw + 1);
      vfgw[vi] = fw + 1;
    } else if (fresh) {
      tsdf[vi] = 1.f; wt[vi] = 0.f; vr[vi] = 0.f; vg[vi] = 0.f; vb[vi] = 0.f; vc[vi] = 0; vfg[vi] = 0.0; vfgw[vi] = 0;
    }
    const int any = __syncthreads_or(vis ? 1 : 0);
    if (tid == 0) blk_active[s] = any ? 1 : 0;
    if (stats) {
      const int nupd = __syncthreads_count(upd ? 1 : 0);
      if (tid == 0) {
        unsigned long long* sl = stats + stat_slot();
        atomicAdd(&sl[0], (unsigned long long)nupd);
        atomicAdd(&sl[STAT_SLOTS], 1ULL);
      }
    }
  }
}

// ------------------------------------------------------------ launchers
// Every launcher covers the `n` objects of `objs` (device table) in one
// launch; a one-object-per-launch caller passes n = 1 and a table offset.
void launch_obj_params(hipStream_t s, const Mat4* T_w_c, const Mat4* mats, const MapObj* objs, int n, float fx,
                       float fy, float cx, float cy, int flags, const MapWork& wl, Mat4* copy_to, int* clear) {
  k_obj_params<<<(n + 63) / 64, 64, 0, s>>>(T_w_c, mats, objs, n, fx, fy, cx, cy, flags, wl.act_count, copy_to,
                                            clear);
}
void launch_alloc(hipStream_t s, const MapObj* objs, int n, const float* depth, int refine, int use_icp,
                  int use_rgb, int w, int h, float band, unsigned long long* stats) {
  dim3 blk(256), grd((w + 255) / 256, h, n);
  k_alloc<<<grd, blk, 0, s>>>(objs, depth, refine, use_icp, use_rgb, w, h, band, stats);
}
void launch_alloc_commit(hipStream_t s, const MapObj* objs, int n, const MapWork& wl) {
  long g = (wl.table_pre[n] + 255) / 256;
  if (g > 4096) g = 4096;
  k_alloc_commit<<<(unsigned)g, 256, 0, s>>>(objs, n, wl);
}
void launch_integrate(hipStream_t s, const MapObj* objs, int n, long sum_cap, float fx, float fy, float cx,
                      float cy, const float* depth, const float* r, const float* g, const float* b, int refine,
                      int use_icp, int use_rgb, int w, int h, float mu, int maxweight, const MapWork& wl,
                      unsigned long long* stats) {
  long ga = (sum_cap + 255) / 256;
  if (ga > 2048) ga = 2048;
  k_active_list<<<(unsigned)ga, 256, 0, s>>>(objs, n, w, h, wl);
  k_integrate<<<4096, 512, 0, s>>>(objs, n, fx, fy, cx, cy, depth, r, g, b, refine, use_icp, use_rgb, w, h, mu,
                                   maxweight, wl, stats);
}

}  // namespace mf
