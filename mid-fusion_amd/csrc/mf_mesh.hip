// Marching cubes over an object's allocated voxel blocks on gfx950
// (algorithms::marching_cube, lib/algorithms/meshing.hpp:157-208, with
// meshing::compute_index :88-146 and interp_vertexes :21-61).  One 512-thread
// workgroup per allocated block, one voxel per thread; triangles are
// appended with one atomic per wave.  The triangle table is this library's
// own (mf_mc_table.h, generated from the cube geometry), so triangles share
// the reference's edge vertices but may split a cube's polygons differently.
#include "mf_kernels.h"
#include "mf_mc_table.h"

namespace mf {

namespace {

struct Corner {
  float tsdf, wt;
  double fg;
  int fgw;
};

// get_fine: missing voxels are voxel_traits<SDF>::initValue() (weight 0,
// foreground 0, witness 0)
__device__ __forceinline__ Corner corner(const VolDev& v, int x, int y, int z) {
  Corner c = {1.f, 0.f, 0.0, 0};
  if ((unsigned)x >= (unsigned)v.res || (unsigned)y >= (unsigned)v.res || (unsigned)z >= (unsigned)v.res) return c;
  const int32_t s = v.slot[morton3((uint32_t)x >> 3, (uint32_t)y >> 3, (uint32_t)z >> 3)];
  if (s < 0) return c;
  const size_t i = (size_t)s * BLOCK_VOX + (x & 7) + 8 * (y & 7) + 64 * (z & 7);
  c.tsdf = v.tsdf[i]; c.wt = v.wt[i]; c.fg = v.fg[i]; c.fgw = v.fgw[i];
  return c;
}

__constant__ int kCornerOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1},
                                     {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
__constant__ int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {4, 5}, {5, 6},
                                 {6, 7}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// compute_index: 0 unless every corner is observed (weight != 0) and
// foreground (fg > 0.5 with more than 3 witnesses); bit k = corner k inside
// (tsdf < 0)
__device__ int cube_index(const Corner (&p)[8]) {
  for (int k = 0; k < 8; k++)
    if (p[k].wt == 0.f) return 0;
  for (int k = 0; k < 8; k++)
    if (p[k].fg <= 0.5 || p[k].fgw <= 3) return 0;
  int idx = 0;
  for (int k = 0; k < 8; k++)
    if (p[k].tsdf < 0.f) idx |= 1 << k;
  return idx;
}

// compute_intersection: s + (0 - v1) * (d - s) / (v2 - v1), metres
__device__ f3 edge_vertex(const Corner (&p)[8], int x, int y, int z, int e, float vs) {
  const int a = kEdge[e][0], b = kEdge[e][1];
  const f3 s = mk3((float)(uint32_t)(x + kCornerOff[a][0]) * vs, (float)(uint32_t)(y + kCornerOff[a][1]) * vs,
                   (float)(uint32_t)(z + kCornerOff[a][2]) * vs);
  const f3 d = mk3((float)(uint32_t)(x + kCornerOff[b][0]) * vs, (float)(uint32_t)(y + kCornerOff[b][1]) * vs,
                   (float)(uint32_t)(z + kCornerOff[b][2]) * vs);
  const float v1 = p[a].tsdf, v2 = p[b].tsdf;
  const float t = -v1;
  return s + (t * (d - s)) / (v2 - v1);
}

// checkVertex (meshing.hpp:148-150): a vertex at or below 0, or beyond the
// volume size, drops its triangle.  The size is compared in metres; the
// reference truncates it to an int (meshing.hpp:163), which would drop every
// triangle of a volume smaller than 1 m.
__device__ __forceinline__ bool bad_vertex(f3 v, float dim) {
  return v.x <= 0 || v.y <= 0 || v.z <= 0 || v.x > dim || v.y > dim || v.z > dim;
}

}  // namespace

// Mesh of one volume: count == null counts triangles into *total; otherwise
// appends up to `cap` triangles (9 floats each) and counts all into *total.
__global__ __launch_bounds__(512) void k_mesh(VolDev v, int nblocks, float* __restrict__ out, long cap,
                                              unsigned long long* __restrict__ total) {
  const int blk = blockIdx.x;
  if (blk >= nblocks) return;
  const int32_t m = v.blk_morton[blk];
  const int bx = (int)compact21((uint64_t)m) * 8, by = (int)compact21((uint64_t)m >> 1) * 8,
            bz = (int)compact21((uint64_t)m >> 2) * 8;
  const int lx = threadIdx.x & 7, ly = (threadIdx.x >> 3) & 7, lz = threadIdx.x >> 6;
  const int x = bx + lx, y = by + ly, z = bz + lz;
  const int size = v.res;
  const int xb = min(max(bx + 8, 0), size - 1), yb = min(max(by + 8, 0), size - 1), zb = min(max(bz + 8, 0), size - 1);
  int ntri = 0;
  Corner p[8];
  int idx = 0;
  if (x < xb && y < yb && z < zb) {
#pragma unroll
    for (int k = 0; k < 8; k++) p[k] = corner(v, x + kCornerOff[k][0], y + kCornerOff[k][1], z + kCornerOff[k][2]);
    idx = cube_index(p);
  }
  const float vs = v.size / (float)size;
  const float dim = v.size;
  // triangles kept after the vertex check
  f3 tv[5][3];
  for (int t = 0; t < 5 && idx != 0 && idx != 255; t++) {
    const int e0 = kMcTri[idx][3 * t];
    if (e0 < 0) break;
    const f3 a = edge_vertex(p, x, y, z, e0, vs);
    const f3 b = edge_vertex(p, x, y, z, kMcTri[idx][3 * t + 1], vs);
    const f3 c = edge_vertex(p, x, y, z, kMcTri[idx][3 * t + 2], vs);
    if (bad_vertex(a, dim) || bad_vertex(b, dim) || bad_vertex(c, dim)) continue;
    tv[ntri][0] = a; tv[ntri][1] = b; tv[ntri][2] = c;
    ntri++;
  }
  // one atomic per wave: lane offsets by prefix sum of the wave's counts
  const int lane = threadIdx.x & 63;
  int pre = ntri;
  for (int off = 1; off < 64; off <<= 1) {
    const int o = __shfl_up(pre, off);
    if (lane >= off) pre += o;
  }
  const int wave_total = __shfl(pre, 63);
  unsigned long long base = 0;
  if (lane == 63 && wave_total) base = atomicAdd(total, (unsigned long long)wave_total);
  base = __shfl(base, 63);
  if (!out || ntri == 0) return;
  const unsigned long long first = base + (unsigned long long)(pre - ntri);
  for (int t = 0; t < ntri; t++) {
    const unsigned long long q = first + t;
    if ((long long)q >= cap) break;
    float* o = out + 9 * q;
    for (int k = 0; k < 3; k++) { o[3 * k] = tv[t][k].x; o[3 * k + 1] = tv[t][k].y; o[3 * k + 2] = tv[t][k].z; }
  }
}

void launch_mesh(hipStream_t s, const VolDev& v, int nblocks, float* out, long cap, unsigned long long* total) {
  if (nblocks <= 0) return;
  k_mesh<<<nblocks, 512, 0, s>>>(v, nblocks, out, cap, total);
}

}  // namespace mf
