// Internal device-side data layout of the MID-Fusion dense path.
//
// HBM layout (all SoA, 4-byte planes unless stated):
//   frame:    depth, grey, rgb r/g/b                    [W*H]
//   pyramid:  per level l: D, I, V x/y/z, N x/y/z, gx, gy, cov x/y/z  [W*H/4^l]
//   model:    raycast vertex x/y/z, normal x/y/z, label [W*H]
//   volume:   Morton-indexed block table slot[(res/8)^3] (int32, -1 missing),
//             internal-node occupancy bytes per octree level, block pool with
//             per-field voxel planes [capacity*512] (tsdf, w, r, g, b, c,
//             fgw as 4 B; fg as 8 B).
#pragma once
#include "mf_math.h"
#include <vector>
#include <string>

namespace mf {

static constexpr int MAX_LEVELS = 4;
static constexpr int MAX_INST = 64;     // camera + objects per context
static constexpr int BLOCK_VOX = 512;   // 8^3 voxels per block (octree_defines.h:38)
// blocks per volume: TSDF byte offsets cap * 512 * 4 stay below 2^32
static constexpr long MAX_BLOCKS = 1L << 21;

// Device view of one object volume (passed by value to kernels).
struct VolDev {
  int res, nb, max_level, leaf_level;
  float size, voxel;
  long cap;
  int32_t* slot;        // [nb^3] block slot by Morton index, -1 missing
  uint8_t* req;         // [nb^3] allocation request flags
  uint8_t* nodes;       // internal-node occupancy, levels 1..leaf-1
  long node_off[24];    // byte offset of level l in `nodes`
  int32_t* blk_morton;  // [cap] Morton index of each slot
  uint8_t* blk_active;  // [cap] VoxelBlock::active_
  float *tsdf, *wt, *r, *g, *b;
  int32_t *c, *fgw;
  double* fg;
  int32_t* counters;    // [0] allocated blocks, [1] overflow flag, [3] allocated blocks at frame start
};

// Matrices derived from the poses, computed on the device once per stage.
struct ObjParams {
  Mat4 pose, inv_pose;      // T_w_o and its inverse
  Mat4 Tco, KT;             // inverse(T_w_c)*T_w_o, K*Tco (integration)
  Mat4 kPose;               // T_w_c * inverse(K) (allocation)
  f3 o_cam;                 // camera centre in the object frame
  Mat4 view;                // T_w_c * getInverseCameraMatrix(k) (raycast)
};

// Per-iteration GN state (double-buffered in GNState::it by launch parity).
struct GNIter {
  Mat4 T;          // camera: T_w_l; object: T_c2_o2
  Mat4 invT;       // inverse(T)
  Mat4 prev;       // previous_pose (carried across levels)
  float prev_err;
  int done_level;  // pyramid level at which the GN loop broke, -1 if none
  int pend_level;  // rows computed with T whose solve is pending, -1 if none
  int pend_iter;
};

// Gauss-Newton state of one tracked instance (device resident).
struct GNState {
  GNIter it[2];    // by launch parity
  Mat4 A;          // camera: T_w_r;            object: T_c1_o1
  Mat4 B;          // camera: K * inverse(T_w_r); object: inverse(T_w_o1)
  Mat4 C;          // object: inverse(T_c1_o1)
  Mat4 T_wo1;      // object: pose at the previous frame
  int object;      // 1: object instance (zero-inlier break, pose composed with T_w_c2)
  Mat4* out_pose;  // written after every update: T_w_l, or volume_pose = T_w_c2 * T_c2_o2
  Mat4 Twc2;       // object: the camera pose T_w_c2 of this frame (fixed during the run)
  Mat4 T_final;    // T after the last solve (object: T_c2_o2)
  int off;         // 1: this run is skipped entirely (no rows, no solves)
  double sums[32]; // last reduction (FP64)
};

// Per-instance inputs of one tracking pass.
struct TrackInst {
  const float* mvn;     // model vertex (planes 0-2, world frame) and normal (3-5) maps, full res
  const float* rD;      // rendered reference depth pyramid, levels back to back
  int8_t* codes;        // level-0 result codes out [2*W*H] (ICP, RGB) or null
  float* rgb_err;       // level-0 photometric row errors out [W*H] or null
  uint8_t* tiles;       // objects: per 8x8 pixel tile of the model maps, 0 when every
                        // pixel is a raycast miss (vertex 0, normal (INVALID, 0, 0)); or null
  const int* tbox;      // with tiles: {x0, y0, x1, y1}, the box (in tiles) of the tiles flagged 1
  int state;            // index into GNState array
  int pad;
};
static constexpr int MAP_TILE = 8;   // model-map tile side of TrackInst::tiles

// Pyramid buffers of one frame (live or reference) at one level: 13 planes
// of w*h floats in one allocation, in the order of the PL_* indices.
enum { PL_D = 0, PL_I = 1, PL_V = 2, PL_N = 5, PL_GX = 8, PL_GY = 9, PL_COV = 10, PL_COUNT = 13 };
struct Level {
  int w, h;
  float *D, *I, *V[3], *N[3], *gx, *gy, *cov[3];
};

// t_*: 100 MHz ticks; solver: 0 none, 1 Gauss-Jordan (certified), 2 Jacobi pseudo-inverse
struct TraceRec { int inst, level, iter, flag; float err, s28; float pose[16]; unsigned t_accum, t_solve; int solver, pad; };

}  // namespace mf
