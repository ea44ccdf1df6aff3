// Host-side launchers of the gfx950 kernels (one translation unit per stage).
#pragma once
#include "mf_internal.h"

namespace mf {

// Load through the global address space.  Pointers read out of a descriptor
// are generic to the compiler, which would otherwise issue flat loads (which
// count against LDS waits as well).
template <class T>
__device__ __forceinline__ T gload(const T* a) {
  return *(const __attribute__((address_space(1))) T*)a;
}
// base[idx] for a wave-uniform base and a 32-bit index: addressed as scalar
// base + 32-bit vector byte offset, half the vector registers of a 64-bit
// per-lane address.
template <class T>
__device__ __forceinline__ T ldx(const T* base, uint32_t idx) {
  return gload((const T*)((const char*)base + (size_t)(idx * (uint32_t)sizeof(T))));
}

// Voxels of the leaf blocks with the given keys into out[n][512][8] (double
// {tsdf, w, r, g, b, c, fg, fgw}) and found[n] (mf_map.hip).
void launch_export_voxels(hipStream_t st, const VolDev& v, const uint64_t* keys, int n, double* out, int* found);

// preprocessing (mf_pre.hip)
void launch_convert(hipStream_t s, const uint16_t* din, const uint8_t* rgb, int iw, int ratio, int ow, int oh,
                    float* depth, float* grey, float* r, float* g, float* b);
void launch_bilateral(hipStream_t s, const float* in0, float* out0, const float* in1, float* out1, int w, int h,
                      float e_d, const float* gauss);
void launch_half(hipStream_t s, const float* in0, float* out0, const float* in1, float* out1, int iw, int ih,
                 float e_d);
void launch_level(hipStream_t s, const Level& L, const Mat4& invK, int negY, float focal, float sxy, float sdisp,
                  float baseline, bool grads, bool cov);
void launch_vertex2depth(hipStream_t s, const float* const v[3], int n, const Mat4* invT, float* out);
// Rendered reference depth pyramids (vertex2depth + robust half-sampling) of
// `n` tracking instances, every level in one launch; rd_off[l] is the offset
// of level l in each instance's pyramid.
void launch_render_reference(hipStream_t s, const TrackInst* insts, int n, int w, int h, int n_levels,
                             const Mat4* invT, const int* rd_off);

// tracking (mf_track.hip)
struct TrackLaunch {
  int level, iter, w, h, W, H;         // level size and full size
  int use_icp, use_rgb, object;        // instance kind
  int robust; double robust_k;
  float fx, fy, cx, cy;                // full-resolution intrinsics
  float inv_cov_rgb;                   // 1/sigma_bright
  const float* live;                   // live pyramid level planes (PL_*), stride lstride
  const float* ref;                    // reference (previous frame) level planes
  int lstride;                         // w*h
  int rd_off;                          // offset of this level in a rendered-depth pyramid
  const int32_t* label;                // camera re-track: outlier mask from raycast labels
  const uint8_t* moving;               // per-object moving flags (label-indexed)
  int n_labels;
  TrackRow* dbg;                       // optional row dump: [state][level][2*W*H], or null
  int n_levels;
  int paired;                          // camera ICP + photometric rows: each thread takes whole pixels
                                       // (both rows), two at a time (rows per thread even)
};
struct TrackSync {             // accumulation + solve of one launch
  TraceRec* trace;              // optional per-iteration trace
  int* trace_n;
  int trace_cap;
};
int track_blocks(int nrows, int rows_per_thread);

// GN iteration q of `ninst` instances, two-kernel form: rows into `partial`
// (nblk per instance), then one solve workgroup per instance.
void launch_track_rows(hipStream_t s, const TrackLaunch& p, const TrackInst* inst, int ninst, GNState* st,
                       float* partial, int nblk, int q);
// GN iteration q's rows of `ninst` objects in one grid over the pixels
// (track_obj_blocks(w*h, pixels_per_thread) workgroups, each writing one
// partial per object); pixels_per_thread is 1 or 2.
int track_obj_blocks(int npix, int pixels_per_thread);
void launch_track_rows_objs(hipStream_t s, const TrackLaunch& p, const TrackInst* inst, int ninst, GNState* st,
                            float* partial, int nblk, int pixels_per_thread, int q);
void launch_track_solve(hipStream_t s, const TrackLaunch& p, const TrackInst* inst, int ninst, GNState* st,
                        const float* partial, int nblk, int q, const TrackSync& sy);
// GN iteration q, fused form: finishes iteration q-1 (part_prev, nblk_prev
// per instance) in every workgroup, then iteration q's rows into part_cur;
// rows == 0 only finishes iteration q-1.
void launch_gn_fused(hipStream_t s, const TrackLaunch& p, const TrackInst* inst, int ninst, GNState* st,
                     const float* part_prev, int nblk_prev, float* part_cur, int nblk, int q, int rows,
                     const TrackSync& sy);
// Tile flags and tile box (TrackInst::tiles, tbox) of caller-written model
// maps (6 planes of W*H at mvn); tbox must hold the empty box.
void launch_map_tiles(hipStream_t s, const float* mvn, uint8_t* tiles, int* tbox, int W, int H);
// The tracking 6x6 solve on n systems (27 doubles each) into x[6n] and
// solver[n] (1 certified Gauss-Jordan, 2 Jacobi pseudo-inverse).
void launch_solve6(hipStream_t s, const double* s27, int n, double* x, int* solver);
// Residual sums of one instance (`inst` points at its descriptor) at pose T:
// GN state `state` gets T, then rows and the FP64 sums (into GNState::sums).
void launch_eval(hipStream_t s, const TrackLaunch& p, const TrackInst* inst, GNState* st, int state, const Mat4& T,
                 float* partial, int nblk);
// Motion gating state on the device (Kfusion::tracking, kfusion.cpp:286-471),
// so that every decision of a frame stays stream-ordered.
struct MotionDev {
  int gate;          // the last raycast shows more than one instance (kfusion.cpp:292)
  int move_any;      // the moving set of this frame is not empty
  int retrack;       // move_any and the raycast before fusion shows an instance (kfusion.cpp:358-359)
  int cond2;         // moving objects take motion rows (kfusion.cpp:445-446)
  int copy_static;   // static objects take the camera's rows (kfusion.cpp:399-402, 450-454)
  int rendered_n;    // instances in the last raycast before fusion (persists across frames)
  int pad[2];
  int present[MAX_INST];       // labels present in the second raycast before fusion
  int present_main[MAX_INST];  // labels present in the model maps' label image
  int in_view[MAX_INST];       // objects_in_view_ by label
  int counts[2 * MAX_INST];    // {pixels, inliers} by label of the motion check
  uint8_t move_set[MAX_INST];  // move_obj_set by label
};
// Motion residuals of one instance over the full-resolution live frame
// (Tracking::compute_residuals, tracking.cpp:1535-1747).
struct MotionLaunch {
  int W, H;
  const float* live;        // live level-0 planes (PL_*)
  const float* ref;         // reference level-0 planes (PL_I read)
  const float* mvn;         // model vertex (0-2) and normal (3-5) planes
  const float* vbf;         // vertex planes of a raycast before fusion
  const Mat4* T_w_l;        // pose of the residual
  const Mat4* inv_T_w_r;    // inverse(T_w_r)
  float fx, fy, cx, cy;
  int use_icp, use_rgb;
  float sigma_bright, threshold;
  int8_t* codes;            // out: ICP codes then RGB codes [2*W*H]
  float* rgb_err;           // out (optional): photometric row errors [W*H]
  const int* gate;          // run only if *gate != 0 (null: always)
  const int32_t* label_bf;  // optional: labels of the raycast before fusion, for
  int* counts;              //   per-label {pixels, inliers} counts [2*MAX_INST]
};
void launch_motion_rows(hipStream_t s, const MotionLaunch& m);
// n instances' rows in one launch: `base` (no per-label counts) with each
// instance's maps, pose and outputs from a device table; outputs disjoint.
struct MotionObj {
  const float* mvn; const float* vbf; const Mat4* T_w_l; int8_t* codes; float* rgb_err;
};
void launch_motion_rows_n(hipStream_t s, const MotionLaunch& base, const MotionObj* objs, int n);
// Display images (uchar4 as one 32-bit word per pixel).
void launch_render_track(hipStream_t s, const int8_t* codes, const float* err, int n, int mode, uint32_t* out);
void launch_render_depth(hipStream_t s, const float* depth, int n, float nearp, float farp, uint32_t* out);
void launch_render_intensity(hipStream_t s, const float* grey, int n, uint32_t* out);
// Marching cubes over the first `nblocks` pool slots of a volume: counts
// triangles into *total and, with `out`, writes up to `cap` of them.
void launch_mesh(hipStream_t s, const VolDev& v, int nblocks, float* out, long cap, unsigned long long* total);
// generalize_label of a w*h 0/non-zero instance mask (global: optional model
// mask) into an integration mask; `count` is device scratch for the
// rendered-mask size.
void launch_generalize(hipStream_t s, const uint8_t* mask, const uint8_t* global, int w, int h, int rendered,
                       int* count, float* out);
// *dst = table[*index] (ground-truth camera pose of the frame)
void launch_pose_from_table(hipStream_t s, Mat4* dst, const Mat4* table, const int* index);
void launch_set_int(hipStream_t s, int* dst, int v);
// Frame-start clear of the motion state, then the gate from the labels
// present in the last raycast (present_main, kept by whatever writes the
// label image) and the in-view set cleared under it.
void launch_motion_gate(hipStream_t s, MotionDev* md, int nobj);
// After the raycast before fusion (which flagged its labels in in_view): the
// rendered count; with `manual`, the moving set is the caller's moving flags.
void launch_motion_inview(hipStream_t s, MotionDev* md, int nobj, const uint8_t* moving, int manual);
// check_static_state decision from the counts (kfusion.cpp:314-354): updates
// the moving flags and the moving set.
void launch_motion_classify(hipStream_t s, MotionDev* md, int nobj, uint8_t* moving);
// Object phase: rendered count after the second raycast before fusion (when
// objects move; that raycast flags its labels in `present`, zero since the
// frame-start clear), then cond2 and copy_static.
void launch_motion_after(hipStream_t s, MotionDev* md, int nobj, int use_rgb);
// virtual_camera_pose_ of tracked objects: vcam[ids[j] - 3] for instances
// [first, first + n), unless the run was skipped.
void launch_virtual_cams(hipStream_t s, const GNState* st, int first, int n, const int* ids, Mat4* vcam);
// dst <- src (2*n codes) and edst <- esrc (ne photometric errors) when *gate != 0
// Copies of up to MAX_INST float images of n elements in one launch.
struct MaskCopies {
  const float* src[MAX_INST];
  float* dst[MAX_INST];
};
void launch_copy_masks(hipStream_t s, const MaskCopies& mc, int m, size_t n);
void launch_copy_codes(hipStream_t s, const int8_t* src, int8_t* dst, size_t n2, const float* esrc, float* edst,
                       size_t ne, const int* gate);
// Frame start for the camera: T_w_r = T_w_c, inverse(T_w_r), GN state 0.
void launch_camera_begin(hipStream_t s, Mat4* mats, GNState* st, float fx, float fy, float cx, float cy);
// GN state of a camera pass starting from the current camera pose.
void launch_gn_begin_camera(hipStream_t s, GNState* st, Mat4* T_w_l, const Mat4* T_w_r, const int* gate, float fx, float fy,
                            float cx, float cy);
// GN states of `n` objects: states[first+j] <- object poses poses[ids[j]].
// With a gate, the runs are skipped unless *gate != 0.
void launch_gn_begin_objects(hipStream_t s, GNState* st, int first, int n, const int* ids, Mat4* poses,
                             const Mat4* T_w_c2, const Mat4* T_w_c1, int publish = 1, const int* gate = nullptr);

// mapping (mf_map.hip)
// One object of a mapping launch (device table, indexed by the launch's
// object dimension).
struct MapObj {
  VolDev v;
  ObjParams* prm;
  const float* mask;      // generalised integration mask
  const int8_t* codes;    // the instance's last level-0 track codes (refine_mask_use_motion)
  int* tbox;              // the object's model-map tile box (TrackInst::tbox)
  int id;                 // object id: pose in mats[3 + id]
  int pad;
};
// Per-launch work lists of the mapping kernels.
struct MapWork {
  long table_pre[MAX_INST + 1];   // prefix of the objects' Morton-table sizes (host computed)
  int* act_count;                 // active blocks this frame
  int4* act_list;                 // (object, slot, Morton index, 0) of each active block
};
// flags: OBJ_ALLOC_RESET starts a mapping pass (the active-list count, and the
// volumes' frame-start block counts that mark this pass's new blocks for
// k_integrate to initialise), OBJ_TBOX_RESET empties the objects' model-map
// tile boxes ahead of the raycast that fills them.
// Optionally also *copy_to = *T_w_c and clear[0 .. MAX_INST) = 0.
enum { OBJ_ALLOC_RESET = 1, OBJ_TBOX_RESET = 2 };
void launch_obj_params(hipStream_t s, const Mat4* T_w_c, const Mat4* mats, const MapObj* objs, int n, float fx,
                       float fy, float cx, float cy, int flags, const MapWork& wl, Mat4* copy_to = nullptr,
                       int* clear = nullptr);
// Work counters accumulated on the device (algorithmic-byte accounting):
// [0] voxel updates, [1] blocks integrated, [2] march steps, [3] trilinear
// evaluations, [4] raycast hits, [5] allocation band samples, [6] pixels
// allocated from, [7] ray-object pairs, [8] octree nodes visited by the
// first-block traversal.
static constexpr int N_STATS = 9;
// Work counters are spread over STAT_SLOTS addresses per counter (stats[k *
// STAT_SLOTS + slot]) so that workgroups do not all hit one atomic address;
// the host sums the slots.  A null stats pointer disables counting.
static constexpr int STAT_SLOTS = 64;
__device__ __forceinline__ int stat_slot() { return (blockIdx.x + 7 * blockIdx.y + 13 * blockIdx.z) & (STAT_SLOTS - 1); }
void launch_alloc(hipStream_t s, const MapObj* objs, int n, const float* depth, int refine, int use_icp,
                  int use_rgb, int w, int h, float band, unsigned long long* stats);
void launch_alloc_commit(hipStream_t s, const MapObj* objs, int n, const MapWork& wl);
void launch_integrate(hipStream_t s, const MapObj* objs, int n, long sum_cap, float fx, float fy, float cx,
                      float cy, const float* depth, const float* r, const float* g, const float* b, int refine,
                      int use_icp, int use_rgb, int w, int h, float mu, int maxweight, const MapWork& wl,
                      unsigned long long* stats);

// raycasting (mf_raycast.hip)
struct RayObj {
  VolDev v;
  const ObjParams* prm;
  int label;
  float* mvx; float* mvy; float* mvz; float* mnx; float* mny; float* mnz;   // per-object output maps
  uint8_t* tiles;         // set with the maps: their 8x8 tile flags and tile box (TrackInst::tiles,
  int* tbox;              // tbox), the box reset beforehand by launch_obj_params(..., OBJ_TBOX_RESET)
};
// raycastObjectList over `objs`: objects whose label is not flagged in
// `in_view` are skipped unless no object is flagged (null: every object); the
// per-object maps are written where the object's pointers are set; with
// nx == nullptr no normal is computed; with a gate, nothing runs unless
// *gate != 0.
// renderVolume_many_Kernel from the last raycast's maps, or by raycasting at
// the view pose T_w_c (prm of every object computed for that view).
void launch_render_volume_maps(hipStream_t s, const RayObj* objs, int nobj, int n, const Mat4* T_w_c,
                               float* const v[3], float* const nrm[3], const int32_t* label, int render_color,
                               const uint32_t* palette, int npal, uint32_t* out);
void launch_render_volume_ray(hipStream_t s, const RayObj* objs, int nobj, int w, int h, const Mat4* T_w_c, float mu,
                              float nearp, float farp, float largestep, float eps, int render_color,
                              const uint32_t* palette, int npal, uint32_t* out);
// present (optional): present[l] = 1 for every label l in [0, MAX_INST) the
// raycast writes (split_labels' instance set, kfusion.cpp:1527-1553).
void launch_raycast(hipStream_t s, const RayObj* objs, int nobj, int w, int h, float mu, float nearp, float farp,
                    float eps, float* vx, float* vy, float* vz, float* nx, float* ny, float* nz, int32_t* label,
                    unsigned long long* stats, const int* in_view, const int* gate, int* present = nullptr);

}  // namespace mf
