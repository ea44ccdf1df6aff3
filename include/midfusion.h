/*
 * midfusion.h — C ABI of the MI355X (gfx950) implementation of MID-Fusion's
 * per-frame dense path: depth preprocessing, ICP + photometric Gauss-Newton
 * tracking of the camera and of moving objects, octree TSDF block
 * allocation, projective integration and raycasting.
 *
 * The entry points mirror the reference pipeline facade `Kfusion`
 * (apps/kfusion/include/kernels.h:22-361) and the per-object calls of
 * `Object` / `Tracking` (object/object.h:33-154, tracking/tracking.h:34-409).
 * Plain pointers and sizes only.  Host pointers unless a name ends in
 * `_device`.  Every call is stream-ordered on the context's HIP stream and
 * returns 0 (or a non-negative count/id) on success, a negative MF_E_* code
 * on failure; nothing calls exit().
 */
#ifndef MIDFUSION_H
#define MIDFUSION_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MF_OK 0
#define MF_E_ARG (-1)       /* bad argument / image ratio (preprocessing.cpp:76-88 exits) */
#define MF_E_OOM (-2)       /* device allocation failed */
#define MF_E_HIP (-3)       /* HIP runtime error */
#define MF_E_CAPACITY (-4)  /* block pool of a volume is full */
#define MF_E_EOF (-5)       /* end of a .raw sequence, or a frame that does not parse (interface.h:297-303) */

#define MF_ROBUST_NONE 0
#define MF_ROBUST_HUBER 1
#define MF_ROBUST_TUKEY 2
#define MF_ROBUST_CAUCHY 3   /* reference default, k = 2.3849 (tracking.h:342-345) */

typedef struct mf_ctx mf_ctx;

/* Configuration: the `Configuration` fields the dense path reads
 * (include/config.h:15-63, default_parameters.h:22-57). */
typedef struct mf_config {
  int width, height;    /* computation size */
  int n_levels;         /* pyramid levels, <= 4 (reference: 3) */
  int iters[4];         /* GN iterations per level, [0] = finest (reference {10,5,4}) */
  int use_icp, use_rgb; /* Tracking(..., using_ICP, using_RGB) */
  int robust;           /* MF_ROBUST_* */
  double robust_k;      /* robust-function constant */
  float mu;             /* TSDF truncation (default_parameters.h:25) */
  int track_rate, integ_rate;
  int init_frame;       /* segment_startFrame_: the frame that only seeds the reference */
  int batch_objects;    /* 1: all moving objects share one launch per GN iteration */
} mf_config;

/* Kfusion::Kfusion (kernels.h:115-170): allocates every frame buffer and the
 * background object (id 0, identity pose, volume res^3 of side size_m). */
mf_ctx* mf_create(const mf_config* cfg, const float init_pose[16], float bg_size_m, int bg_res, long bg_capacity);
void mf_destroy(mf_ctx* ctx);
/* Last error message of the context (static storage of the context). */
const char* mf_last_error(mf_ctx* ctx);

/* Object::Object (object/object.cpp:19-46); res a power of two in [16, 8192].  Returns the instance id
 * (= instance_label_), capacity = max voxel blocks (0: min(blocks in the volume, 262144); at most
 * 2^21, else MF_E_ARG; the same limit holds for mf_create's bg_capacity). */
int mf_object_create(mf_ctx* ctx, float size_m, int res, const float T_w_o[16], long capacity);
/* Object::set_static_state (object.cpp:317-320): moving objects are tracked
 * by trackEachObject and masked out of the camera re-track. */
int mf_object_set_moving(mf_ctx* ctx, int id, int moving);
/* Motion detection (Kfusion::tracking, kfusion.cpp:286-354): on, after the
 * camera's first pass, the motion residuals (Tracking::compute_residuals,
 * tracking.cpp:1535-1747) against the maps before fusion classify every
 * object in view by its inlier ratio (moving iff 0.05 < ratio < 0.8,
 * check_static_state rendering.cpp:907-947), which replaces the caller's
 * moving flags for those objects; one host readback per tracked frame.  Off
 * (default), the moving set of a frame is the caller's flags. */
int mf_set_motion_detection(mf_ctx* ctx, int on);
/* Last motion check, per object id < cap: {pixels in view, inliers} and the
 * current moving flag.  Returns the object count. */
int mf_get_motion(mf_ctx* ctx, int64_t* pixels, int64_t* inliers, int32_t* moving, int cap);
/* Gating state of the last tracking call: gate_rendered = {more than one
 * instance in the last raycast, instances in the last raycast before fusion};
 * in_view = objects_in_view_ flags and move_set = move_obj_set flags, by id. */
int mf_get_motion_state(mf_ctx* ctx, int32_t gate_rendered[2], int32_t* in_view, int32_t* move_set, int cap);
/* Residual codes of the last motion check (2*W*H int32, ICP rows then RGB
 * rows, the TrackData.result values of compute_residuals). */
int mf_get_motion_codes(mf_ctx* ctx, int32_t* codes);
/* Tracking::compute_residuals (tracking.cpp:1535-1747) at pose T_w_l against
 * the given model vertex and normal maps and vertex map before fusion (AoS
 * float3, W*H each), with this frame's live and reference images and T_w_r:
 * 2*W*H result codes.  Valid between mf_track_camera and mf_track_objects. */
int mf_motion_residuals(mf_ctx* ctx, const float k[4], const float T_w_l[16], const float* vertex,
                        const float* normal, const float* vertex_bf, int32_t* codes);
/* Generalised integration mask of an instance (instance_seg::generalize_label
 * output, segmentation.cpp:10-47): fg value >= 0 inside, -1 outside.  NULL
 * clears it (the instance is then not integrated). */
int mf_object_set_mask(mf_ctx* ctx, int id, const float* mask);
/* instance_seg::generalize_label (segmentation.cpp:10-47) on the device: from
 * a W*H instance mask (0 / non-zero, as Mask R-CNN gives it) the integration
 * mask: the foreground value on the instance (1; for a mask rendered from
 * the volumes, 0.1 if it covers under 3% of the image, else 0.5), 0 on the
 * ring of its 5x5 elliptic dilation (and on the optional model mask, OR-ed
 * into the dilation), -1 elsewhere. */
int mf_object_set_instance_mask(mf_ctx* ctx, int id, const uint8_t* mask, int rendered, const uint8_t* model_mask);
int mf_object_set_instance_mask_device(mf_ctx* ctx, int id, const uint8_t* d_mask, int rendered,
                                       const uint8_t* d_model_mask);
/* The instance's current integration mask (W*H floats). */
int mf_object_get_mask(mf_ctx* ctx, int id, float* mask);
int mf_object_set_mask_device(mf_ctx* ctx, int id, const float* d_mask);
/* mf_object_set_mask_device for n objects at once (ids[i] gets d_masks[i];
 * a null mask clears that object's mask), in one copy launch: the per-frame
 * masks of all instances (Object::mask_ of every object, set per frame from
 * the segmentation, segmentation.cpp:10-47).  Ids may not repeat. */
int mf_objects_set_masks_device(mf_ctx* ctx, int n, const int32_t* ids, const float* const* d_masks);

/* Kfusion::preprocessing(depth, rgb, inSize, filter) (kfusion.cpp:210-235):
 * mm->metres, grey + rgb/255, optional bilateral filter of the depth. */
int mf_preprocess(mf_ctx* ctx, const uint16_t* depth_mm, const uint8_t* rgb8, int in_w, int in_h, int filter);
int mf_preprocess_device(mf_ctx* ctx, const uint16_t* d_depth_mm, const uint8_t* d_rgb8, int in_w, int in_h,
                         int filter);
/* Kfusion::tracking(k, rate, frame) (kfusion.cpp:237-488).  Returns 1 if
 * tracking ran, 0 if skipped (rate / seeding frame). */
int mf_track(mf_ctx* ctx, const float k[4], unsigned frame);
/* The two phases of mf_track: the camera phase builds the pyramid and runs
 * Tracking::trackLiveFrame (plus the masked re-track when objects move,
 * kfusion.cpp:237-392); the object phase runs Tracking::trackEachObject for
 * every moving object at the current camera pose and hands the live pyramid
 * over as the next reference (kfusion.cpp:395-471, tracking.cpp:580-708,
 * 1401-1414).  mf_track_objects does nothing unless a camera phase that
 * tracked is pending. */
int mf_track_camera(mf_ctx* ctx, const float k[4], unsigned frame);
int mf_track_objects(mf_ctx* ctx, const float k[4], unsigned frame);
/* Tracking::trackEachObject(ObjectPointer&, k, T_w_c2, T_w_c1, I, D)
 * (tracking.h:63-68, tracking.cpp:580-708) for object id > 0 alone, between
 * mf_track_camera and mf_track_objects: GN from the object's pose at the
 * frame's start (T_c1_o1 = T_w_c1^-1 T_w_o1, T_c2_o2 = T_w_c2^-1 T_w_o1) at
 * the current camera pose T_w_c2, updating volume_pose_ (mf_get_pose), the
 * virtual camera pose and the object's level-0 trackresult_ codes.  It runs
 * whatever the moving flags and the motion gate say.  mf_track_objects then
 * keeps this result (it does not track the object again) and completes the
 * frame.  MF_E_ARG if no camera phase is pending or the object was already
 * tracked in this frame. */
int mf_track_object(mf_ctx* ctx, int id, const float k[4]);
/* Kfusion::integration(k, rate, mu, frame) (kfusion.cpp:573-677). */
int mf_integrate(mf_ctx* ctx, const float k[4], unsigned frame);
/* Kfusion::raycasting(k, mu, frame) (kfusion.cpp:539-569). */
int mf_raycast(mf_ctx* ctx, const float k[4], unsigned frame);
/* Kfusion::synchroniseDevices (kfusion.cpp:1697-1699).  Also reports a
 * pending capacity overflow as MF_E_CAPACITY. */
int mf_sync(mf_ctx* ctx);

/* Kfusion::getPose (kernels.h:282-284) for id < 0; Object::volume_pose_
 * otherwise.  Row-major 4x4. */
int mf_get_pose(mf_ctx* ctx, int id, float T[16]);
int mf_set_pose(mf_ctx* ctx, int id, const float T[16]);

/* Raycast output (Kfusion::vertex/normal + raycast label image): AoS float3
 * per pixel; any pointer may be NULL. */
int mf_get_maps(mf_ctx* ctx, float* vertex, float* normal, int32_t* label);
int mf_set_maps(mf_ctx* ctx, const float* vertex, const float* normal, const int32_t* label);
/* Object::m_vertex / m_normal. */
int mf_get_object_maps(mf_ctx* ctx, int id, float* vertex, float* normal);
int mf_set_object_maps(mf_ctx* ctx, int id, const float* vertex, const float* normal);

/* Allocated voxel-block keys of an instance in the reference encoding
 * (keyops::encode, octant_ops.hpp:51-55), ascending.  Returns the count. */
long mf_object_keys(mf_ctx* ctx, int id, uint64_t* out, long cap);
/* Voxels of the block holding voxel (x,y,z): 512 x {tsdf, w, r, g, b, c, fg,
 * fgw} as double, index x + 8y + 64z.  Returns 1, or 0 if unallocated. */
int mf_object_block(mf_ctx* ctx, int id, int x, int y, int z, double* out);
/* Bulk form of mf_object_block: voxels of the n blocks with the given keys
 * (mf_object_keys encoding), out[n][512][8] as double {tsdf, w, r, g, b, c,
 * fg, fgw}, index x + 8y + 64z; a key without an allocated block reads as
 * voxel_traits<SDF>::initValue() (tsdf 1, the rest 0).  Returns how many keys
 * were allocated.  (Octree contents for checks against the reference's
 * VoxelBlock data, node.hpp:83-157.) */
long mf_object_voxels(mf_ctx* ctx, int id, const uint64_t* keys, long n, double* out);
/* mf_object_voxels for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc);
 * no counterpart in the reference.  d_keys (n keys) and d_out (n * 512 * 8
 * doubles, same layout) are device pointers; n <= 65536, else MF_E_ARG.  It
 * calls no hipMalloc or hipFree: the per-key found flags and their count live
 * in a fixed 256-KiB device array of the library, one per GPU, held until the
 * process ends and shared by every context on that GPU, so calls that
 * overlap on one GPU race.  Blocking; returns how many keys were allocated. */
long mf_object_voxels_device(mf_ctx* ctx, int id, const uint64_t* d_keys, long n, double* d_out);
/* Number of internal octree nodes implied by the allocated blocks. */
long mf_object_node_count(mf_ctx* ctx, int id);

/* Level-0 tracking result codes of an instance's last pass (TrackData.result
 * of Object::trackresult_, ICP rows then RGB rows: 2*W*H int32). */
int mf_get_trackresult(mf_ctx* ctx, int id, int32_t* codes);
/* Overwrite an instance's trackresult_ codes (2*W*H int32), e.g. to integrate
 * from codes computed elsewhere. */
int mf_set_trackresult(mf_ctx* ctx, int id, const int32_t* codes);
/* Per-iteration GN trace since the last clear (the first clear turns tracing
 * on): rows of 25 doubles {GN state, level, iteration, error, inliers, flag
 * (0 step, 1 converged, 2 step-back, 3 no inliers), pose[16], FP64
 * accumulation us, solve + control us, solver (0 none, 1 certified
 * Gauss-Jordan, 2 Jacobi pseudo-inverse)}; returns the row count. */
long mf_get_trace(mf_ctx* ctx, double* out, long cap);
void mf_clear_trace(mf_ctx* ctx);
/* Row dump for parity checks: when on, every tracking launch also writes its
 * TrackData rows (commons.h:624-628; ICP rows then RGB rows, row stride W at
 * every level) for GN state 0 (camera) and 1 (first moving object).
 * mf_get_rows returns the rows of the last launch at that level. */
int mf_debug_rows(mf_ctx* ctx, int on);
int mf_get_rows(mf_ctx* ctx, int state, int level, int32_t* result, float* error, float* J);
/* Pyramid buffers of the live frame: which = 0 depth, 1 grey, 2 vertex (3/px),
 * 3 normal (3/px), 4 grad x, 5 grad y, 6 icp covariance (3/px), 7 the
 * camera's rendered reference depth (vertex2depth pyramid, r_D_render). */
long mf_get_level(mf_ctx* ctx, int level, int which, float* out);
/* Preprocessed frame: depth (m), grey, rgb (3/px). */
int mf_get_frame(mf_ctx* ctx, float* depth, float* grey, float* rgb);

/* Forward/backward of the ICP + photometric tracking residual at a given
 * pose (the residual kernels and Tracking::reduceKernel, tracking.cpp:
 * 392-578, 923-1111, 1191-1352).  For the camera (id < 0) T = T_w_l; for
 * moving object id > 0, T = T_c2_o2 with the object's pose at the frame's
 * start.  out = the 32 FP64 sums at pyramid level `level`: [0] the robust
 * loss L (forward), [1..6] J^T W e and [7..27] the upper triangle of J^T W J,
 * [28] inliers, [29..31] outlier counts.  The rows' J is minus the derivative
 * of the residual for a left perturbation T <- exp(xi) T, xi = (v, w), so
 * dL/dxi = -out[1..6] (backward) wherever the rows' correspondences are
 * locally fixed.  Valid between mf_track_camera and mf_track_objects of a
 * frame; the robust function is the context's. */
int mf_residual_eval(mf_ctx* ctx, int id, const float k[4], int level, const float T[16], double out[32]);

/* The tracking solve (solve / makeJTJ, commons.h:672-696: TooN GR_SVD backsub
 * with condition 1e6) on n caller-given systems of 27 doubles each {JtWe[6],
 * upper triangle of JtWJ row-major[21]}, computed by the device code the GN
 * iterations use: x[6n] the solutions, solver[n] 1 (certified Gauss-Jordan) or
 * 2 (Jacobi pseudo-inverse).  Blocking; n <= 2^20. */
int mf_solve6(mf_ctx* ctx, const double* s27, long n, double* x, int32_t* solver);

/* Kernel timing: when enabled every launch is bracketed by HIP events on the
 * context stream.  out[i] = {total_ms, launches} for each kernel class in
 * mf_kernel_names order; returns the number of classes. */
int mf_enable_timing(mf_ctx* ctx, int on);
int mf_get_timing(mf_ctx* ctx, double* out, int cap);
const char* mf_kernel_name(int i);
/* Device work counters since the last reset, counted while kernel timing is
 * enabled, for algorithmic-byte accounting: {voxel updates, blocks integrated, march steps, trilinear
 * evaluations, raycast hits, allocation band samples, pixels allocated from,
 * ray-object pairs, octree nodes visited by the raycast's first-block
 * traversal}.  Returns the number of counters. */
int mf_get_stats(mf_ctx* ctx, uint64_t* out, int cap);
int mf_reset_stats(mf_ctx* ctx);
/* Launches a plain float4 device copy of `bytes` bytes `reps` times on the
 * context stream (no counterpart in the reference: the calibration point for
 * hardware traffic counters). */
int mf_copy_reference(mf_ctx* ctx, size_t bytes, int reps);
/* Replays the camera's tracking-row kernel of pyramid level `level` `reps`
 * times back to back on a scratch copy of the camera's GN state (run even
 * when the last frame's re-track left the camera's state off) and returns
 * its average launch time (HIP events on the context stream) and algorithmic
 * bytes per launch (no counterpart in the reference: roofline measurement).
 * Call it between frames; it overwrites the camera's track codes. */
int mf_replay_track_rows(mf_ctx* ctx, const float k[4], int level, int reps, double* avg_us, double* alg_bytes);
/* Device bytes held by the context. */
long mf_device_bytes(mf_ctx* ctx);

/* Launch-shape knobs (no counterpart in the reference, whose loops run on the
 * host CPU).  Results do not depend on them beyond FP32 summation order.
 *   MF_TUNE_TRACK_ROWS_PER_THREAD_L0 + l: residual rows (ICP or photometric)
 *   each tracking thread sums before the workgroup reduction at pyramid
 *   level l, 1..64 (defaults 4, 1, 1, 1). */
/*   MF_TUNE_TRACK_FUSED_L0 + l: 1 runs each GN iteration at level l as one
 *   launch whose workgroups first finish the previous iteration themselves;
 *   0 as a row launch and a one-workgroup solve launch; -1 (default) fused
 *   at levels l >= 1 when the launch has at most 1200 row workgroups. */
/*   MF_TUNE_OBJ_PIXELS_PER_THREAD_L0 + l: moving objects' rows at level l in
 *   one grid over the pixels, each thread loading its pixels' frame inputs
 *   once for every object, with 1 or 2 pixels per thread; 0 (default) gives
 *   each object its own row grid in the same launch. */
/*   MF_TUNE_OBJ_ROWS_PER_THREAD_L0 + l: rows per thread of the moving
 *   objects' row grids at level l, 1..64, or 0 (default): {8, 8, 4, 4} when
 *   two or more objects share the launch, else the camera's. */
enum { MF_TUNE_TRACK_ROWS_PER_THREAD_L0 = 0, MF_TUNE_TRACK_FUSED_L0 = 8, MF_TUNE_OBJ_PIXELS_PER_THREAD_L0 = 16,
       MF_TUNE_OBJ_ROWS_PER_THREAD_L0 = 24 };
int mf_set_tuning(mf_ctx* ctx, int knob, int value);
int mf_get_tuning(mf_ctx* ctx, int knob, int* value);
/* Launch-shape autotuner (no counterpart in the reference): for every
 * pyramid level, times bursts of `reps` GN iterations of the camera (row
 * launch + solve launch) on the last tracked frame for each of the
 * MF_TUNE_CANDIDATES rows-per-thread values {1, 2, 3, 4, 8, 16} (best of
 * three bursts each), and sets MF_TUNE_TRACK_ROWS_PER_THREAD_L0 + level to the
 * fastest, keeping the current value unless another is more than 1 % faster.  The iterations
 * run on a scratch copy of the camera's GN state: poses, result codes and
 * maps are left as they were.  us[level * MF_TUNE_CANDIDATES + i] (or null)
 * gets the measured microseconds per iteration, chosen[level] (or null) the
 * value set.  Blocking; call between frames, after at least one tracked
 * frame.  Later results change only by FP32 summation order. */
enum { MF_TUNE_CANDIDATES = 6 };
int mf_autotune(mf_ctx* ctx, const float k[4], int reps, int32_t* chosen, double* us);

/* ---- On-disk formats on either side of the path ---------------------- */
typedef struct mf_raw mf_raw;
/* RawDepthReader (interface.h:207-333): a SLAMBench .raw sequence, frames of
 * {uint32 w, uint32 h, uint16 depth[w*h], uint32 w, uint32 h, uint8
 * rgb[3*w*h]}.  NULL if the file cannot be opened or its first frame header
 * read ("Invalid Raw file"). */
mf_raw* mf_raw_open(const char* path);
/* Frame size, whole frames in the file, and RawDepthReader::getK's default
 * intrinsics (531.15, 531.15, 320, 240) (interface.h:327-329). */
int mf_raw_info(mf_raw* raw, int* width, int* height, long* frames, float k[4]);
/* readNextDepthFrame (interface.h:236-304) for frame `frame`: depth (w*h
 * uint16, mm) and rgb (3*w*h bytes), either may be NULL.  MF_E_EOF past the
 * end or on a frame whose headers do not repeat frame 0's size. */
int mf_raw_read(mf_raw* raw, long frame, uint16_t* depth, uint8_t* rgb);
void mf_raw_close(mf_raw* raw);
/* mf_preprocess of frame `frame` of a .raw sequence: the file is read into
 * pinned host staging (two buffers, so the read of one frame overlaps the
 * copy of the previous one) and copied on the context stream. */
int mf_preprocess_raw(mf_ctx* ctx, mf_raw* raw, long frame, int filter);
/* parseGTFile (vtk-io.h:304-330): TUM ground truth; every line of exactly 8
 * space-separated fields "t tx ty tz qx qy qz qw" gives a row-major 4x4 pose
 * (toMatrix4(quat, trans), math_utils.h:137-146) in poses[16*i] (up to cap).
 * Returns the number of poses, MF_E_ARG if the file cannot be opened or a
 * field does not parse. */
long mf_parse_gt(const char* path, float* poses, long cap);
/* Kfusion::tracking with ground-truth poses (kfusion.cpp:262-268): from now
 * on the camera pose of frame f is poses[f] with its translation moved by
 * (initial pose - poses[0]) translation, instead of the camera's GN run.
 * n = 0 turns it off.  mf_track_camera fails with MF_E_ARG for f >= n. */
int mf_set_gt_poses(mf_ctx* ctx, const float* poses, long n);
/* Kfusion::save_poses (kfusion.cpp:1591-1626): appends the TUM line of the
 * camera to "<prefix>_camera" and of each object to "<prefix>_<id>"; the
 * first call for a file truncates it.  Each line carries the instance's own
 * translation (the reference writes the camera's for every instance,
 * kfusion.cpp:1617). */
int mf_save_poses(mf_ctx* ctx, const char* prefix, int frame);
/* One TUM line "frame\ttx\tty\ttz\tqx\tqy\tqz\tqw\t\n" for pose T, with
 * getQuaternion's rotation quaternion (preprocessing.cpp:312-320).  Returns
 * the line length. */
int mf_format_tum(const float T[16], int frame, char* out, int cap);

/* ---- Display rendering and label projection ----------------------------
 * Images are W*H (or w*h) uchar4 (x, y, z, w bytes); every call blocks until
 * the image is on the host. */
/* Kfusion::renderTrack (kfusion.cpp:1332-1425) of instance id's
 * trackresult_: mode 0 renderTrackKernel (ICP rows), 1 renderRGBTrackKernel
 * (photometric rows), 2 render_RGBD_TrackKernel (rendering.cpp:183-237,
 * 395-487). */
int mf_render_track(mf_ctx* ctx, int id, int mode, uint8_t* out_rgba);
/* renderDepthKernel of the frame depth, near 0.1 m, far 4 m
 * (rendering.cpp:368-394, kfusion.cpp:1435-1441). */
int mf_render_depth(mf_ctx* ctx, uint8_t* out_rgba);
/* renderIntensityKernel of the frame intensity (rendering.cpp:648-664). */
int mf_render_intensity(mf_ctx* ctx, uint8_t* out_rgba);
/* volume2MaskKernel (rendering.cpp:240-343; Kfusion::volume2mask,
 * kfusion.cpp:679-686): instance labels of every object raycast at the
 * camera pose (INVALID where every object misses), W*H int32. */
int mf_volume2mask(mf_ctx* ctx, const float k[4], int32_t* labels);
/* Kfusion::renderVolume (kfusion.cpp:1284-1328, renderVolume_many_Kernel
 * rendering.cpp:489-646): shaded w*h image at view_pose.  The last raycast's
 * maps are shaded when view_pose is that raycast's pose and w, h the
 * computation size; otherwise every object is raycast again (far plane 8 m,
 * march step `largestep`, 0.75*mu in the benchmark).  render_color: voxel
 * colours, else palette colours by label (rgba bytes, n_palette <= 256; the
 * reference draws them at random, rendering.cpp:666-676) and white for the
 * background.  Returns 1 if it raycast, 0 if it used the maps. */
int mf_render_volume(mf_ctx* ctx, const float k[4], const float view_pose[16], int w, int h, float largestep,
                     int render_color, const uint8_t* palette_rgba, int n_palette, uint8_t* out_rgba);
/* Photometric row errors (TrackData.error of the rows in trackresult_ from W*H
 * on) of instance id: read, or overwrite (e.g. to render rows computed
 * elsewhere). */
int mf_get_track_errors(mf_ctx* ctx, int id, float* rgb_err);
/* algorithms::marching_cube (lib/algorithms/meshing.hpp:157-208) of object
 * id's volume, as Kfusion::dump_mesh calls it (tsdf field, inside = tsdf < 0;
 * kfusion.cpp:1555-1588): triangles as 9 floats (object frame, metres), up
 * to cap written; returns the number of triangles.  Cubes are meshed only
 * where all 8 corners are observed foreground (weight != 0, fg > 0.5, more
 * than 3 witnesses); triangles with a vertex at or below 0 or beyond the
 * volume size are dropped.  The triangle table is this
 * library's own: every cube yields the reference's edge vertices, but its
 * polygons may be split into other triangles. */
long mf_object_mesh(mf_ctx* ctx, int id, float* triangles, long cap);
/* Kfusion::dump_mesh: "<prefix>_<id>.vtk" per object in the writeVtkMesh
 * format (commons.h:945-1011). */
int mf_dump_mesh(mf_ctx* ctx, const char* prefix);
int mf_set_track_errors(mf_ctx* ctx, int id, const float* rgb_err);

#ifdef __cplusplus
}
#endif
#endif
