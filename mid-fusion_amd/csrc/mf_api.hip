// C-ABI implementation: the Kfusion per-frame facade on one MI355X.
//
// Reference sequencing: Kfusion::preprocessing / tracking / integration /
// raycasting (apps/kfusion/src/cpp/kfusion.cpp:165-677), Tracking
// (tracking/tracking.cpp:142-357, 580-708, 1401-1414), Object
// (object/object.cpp:19-46, 128-195).  Poses, GN state and every frame
// buffer stay in HBM; a frame is a stream-ordered sequence of launches with
// no host synchronisation until mf_sync / a getter.
#include "mf_kernels.h"
#include "../../include/midfusion.h"
#include "mf_io.h"
#include <cstdio>
#include <string>
#include <cstring>
#include <cmath>
#include <algorithm>
#include <map>
#include <mutex>

using namespace mf;

namespace {

enum KClass {
  KC_CONVERT, KC_BILATERAL, KC_HALF, KC_LEVEL, KC_V2D, KC_TRACK, KC_TRACK_OBJ, KC_SOLVE, KC_FUSED, KC_GNSMALL, KC_PARAMS, KC_ALLOC,
  KC_COMMIT, KC_INTEGRATE, KC_RAYCAST, KC_MOTION, KC_COUNT
};
const char* kNames[KC_COUNT] = {"convert", "bilateral", "half_sample", "level_maps", "vertex2depth", "track_rows",
                                "track_rows_obj", "track_solve", "track_fused", "gn_control", "obj_params", "alloc_band", "alloc_commit",
                                "integrate", "raycast", "motion"};

struct Object {
  int id = 0;
  float size = 0;
  int res = 0;
  VolDev v;
  bool moving = false;
  bool has_mask = false;
  float* mask = nullptr;
  float* mv[3] = {nullptr, nullptr, nullptr};
  float* mn[3] = {nullptr, nullptr, nullptr};
  int8_t* codes = nullptr;       // level-0 result codes of this instance's last pass
  float* rgb_err = nullptr;      // level-0 photometric row errors of that pass
  float* mvbf[3] = {nullptr, nullptr, nullptr};   // vertex map of the last raycast before fusion
  float* rD[MAX_LEVELS] = {};    // rendered reference depth pyramid
  uint8_t* tiles = nullptr;      // model-map tile flags (TrackInst::tiles)
  int* tbox = nullptr;           // their box (TrackInst::tbox), filled by the main raycast
  ObjParams* prm = nullptr;
  bool tracked = false;          // tracked by mf_track_object in the pending object phase
};

}  // namespace

struct mf_ctx {
  mf_config cfg;
  hipStream_t stream = nullptr;
  int W = 0, H = 0;
  size_t N0 = 0;
  std::vector<void*> allocs;
  long bytes = 0;
  std::string err;
  // frame
  uint16_t* in_depth = nullptr;
  uint8_t* in_rgb = nullptr;
  size_t in_cap = 0;
  float *depth = nullptr, *grey = nullptr, *rgb[3] = {};
  float* gauss = nullptr;
  Level live[MAX_LEVELS], ref[MAX_LEVELS];
  // model maps (raycast output)
  float *vtx[3] = {}, *nrm[3] = {};
  int32_t* label = nullptr;
  // poses on the device: [0] camera, [1] T_w_r, [2] inverse(T_w_r), [3 + id] objects
  Mat4* mats = nullptr;
  GNState* gn = nullptr;
  TrackInst* inst = nullptr;     // device descriptors [MAX_INST]
  float* partial = nullptr;
  int nblk0 = 0;
  int track_rpt[MAX_LEVELS] = {4, 1, 1, 1};   // tracking rows per thread, per pyramid level
  int track_fused[MAX_LEVELS] = {-1, -1, -1, -1};   // 1: one fused launch per iteration at that level, 0: two, -1: fused_level
  // objects' rows in one grid over the pixels with this many pixels per
  // thread (1, 2), or 0: one row grid per object (grid.y), per pyramid level.
  // 0 everywhere measured fastest on MI355X at 1, 8 and 16 objects.
  int track_obj_ppt[MAX_LEVELS] = {0, 0, 0, 0};
  // rows per thread of the objects' row grids per pyramid level (0: automatic,
  // obj_rows_per_thread)
  int track_obj_rpt[MAX_LEVELS] = {0, 0, 0, 0};
  uint8_t* moving = nullptr;     // [MAX_INST] label-indexed flags
  TraceRec* trace = nullptr;
  int* trace_n = nullptr;
  bool trace_on = false;
  int* obj_ids = nullptr;        // [MAX_INST] ids of the moving objects being tracked
  int trace_cap = 4096;
  RayObj* rayobjs = nullptr;
  MapObj* mapobjs = nullptr;     // device [2*MAX_INST]: integration set, then every object
  MapWork work;                  // mapping work lists (sized for every object's pool)
  long work_cap = 0;
  std::vector<MapObj> mapobjs_host = std::vector<MapObj>(2 * MAX_INST);
  unsigned long long* stats = nullptr;   // device work counters [N_STATS]
  TrackInst inst_host[MAX_INST];   // mirrors of the device descriptor tables
  RayObj ray_host[MAX_INST];
  int ray_n = -1;
  // motion detection and gating (kfusion.cpp:286-471)
  bool motion_auto = false;
  MotionDev* md = nullptr;
  float* vbf[3] = {};                 // raycast before fusion: vertex planes
  int32_t* label_bf = nullptr;
  int8_t* motion_codes = nullptr;     // [2*N0] residual codes of the last motion check
  Mat4* vcam = nullptr;               // [MAX_INST] virtual camera poses by object id
  RayObj* rayobjs_bf = nullptr;       // raycast table writing the objects' maps before fusion
  RayObj ray_bf_host[MAX_INST];
  // .raw input: two pinned host staging buffers, each released by an event
  // once its copy has run
  uint8_t* raw_pin[2] = {nullptr, nullptr};
  hipEvent_t raw_ev[2] = {nullptr, nullptr};
  bool raw_ev_used[2] = {false, false};
  size_t raw_cap = 0;
  int raw_next = 0;
  // ground-truth camera poses (re-based to the initial position), device copy
  // and the device frame index the graph reads
  std::vector<Mat4> gt;
  Mat4* gt_dev = nullptr;
  long gt_cap = 0;
  int* gt_frame = nullptr;
  float init_t[3] = {0.f, 0.f, 0.f};
  std::vector<bool> pose_saved;     // per instance: its pose file was started
  // display rendering: the pose of the last raycast (raycastPose), the view
  // pose and palette of a render call, and the output image
  Mat4* ray_pose = nullptr;
  Mat4* view_dev = nullptr;
  uint32_t* palette_dev = nullptr;
  uint32_t* render_out = nullptr;
  int32_t* label_scratch = nullptr;
  size_t render_cap = 0;
  MotionObj* motion_tab = nullptr;    // moving objects' motion-residual inputs and outputs (one launch)
  std::vector<MotionObj> motion_host;
  RayObj* rayobjs_lab = nullptr;      // raycast table writing no per-object maps
  RayObj ray_lab_host[MAX_INST];
  unsigned long long* mesh_total = nullptr;   // marching-cubes triangle counter
  uint8_t* mask_stage = nullptr;      // instance / model masks from the host
  int* mask_count = nullptr;          // rendered-mask pixel count
  float* eval_maps = nullptr;         // mf_motion_residuals inputs (9 planes), allocated on first use
  Mat4* eval_T = nullptr;
  Mat4* tune_pose = nullptr;     // autotuner: pose output of its scratch GN state
  int8_t* eval_codes = nullptr;
  std::vector<int> ids_host;
  std::vector<Object*> objs;
  int* scratch = nullptr;        // this device's no-allocation scratch (address of `scratch` on the device)
  float ray_eps = -1.f;
  bool have_ref = false;
  bool pending_objects = false;   // camera phase done, object phase due
  // hipGraph cache: each stage's launch sequence is captured once per
  // distinct set of kernel arguments (key) and replayed afterwards
  bool graphs = true;
  std::map<std::vector<uint64_t>, hipGraphExec_t> gcache;
  TrackRow* dbg_rows = nullptr;   // row dump [2 states][levels][2*W*H] when enabled
  // timing
  bool timing = false;
  struct Ev { hipEvent_t a, b; int cls; };
  std::vector<Ev> evs;
  std::vector<hipEvent_t> pool;
  double acc_ms[KC_COUNT] = {};
  long acc_n[KC_COUNT] = {};
};

namespace {

template <typename T>
T* dalloc(mf_ctx* c, size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
  hipMemsetAsync(p, 0, n * sizeof(T), c->stream);
  c->allocs.push_back(p);
  c->bytes += (long)(n * sizeof(T));
  return (T*)p;
}

hipEvent_t ev_get(mf_ctx* c) {
  if (!c->pool.empty()) { hipEvent_t e = c->pool.back(); c->pool.pop_back(); return e; }
  hipEvent_t e; hipEventCreate(&e); return e;
}
struct Timed {
  mf_ctx* c; int cls; hipEvent_t a = nullptr, b = nullptr;
  Timed(mf_ctx* cc, int k) : c(cc), cls(k) { if (c->timing) { a = ev_get(c); b = ev_get(c); hipEventRecord(a, c->stream); } }
  ~Timed() { if (c->timing) { hipEventRecord(b, c->stream); c->evs.push_back({a, b, cls}); } }
};
void drain_timing(mf_ctx* c) {
  if (c->evs.empty()) return;
  hipStreamSynchronize(c->stream);
  for (auto& e : c->evs) {
    float ms = 0; hipEventElapsedTime(&ms, e.a, e.b);
    c->acc_ms[e.cls] += ms; c->acc_n[e.cls]++;
    c->pool.push_back(e.a); c->pool.push_back(e.b);
  }
  c->evs.clear();
}

Mat4 m4(const float* T) { Mat4 r; for (int i = 0; i < 16; i++) r.m[i / 4][i % 4] = T[i]; return r; }
int fail(mf_ctx* c, int code, const char* msg) { c->err = msg; return code; }

bool alloc_level(mf_ctx* c, Level& L, int w, int h) {
  const size_t n = (size_t)w * h;
  L.w = w; L.h = h;
  float* b = dalloc<float>(c, n * PL_COUNT);
  if (!b) return false;
  L.D = b + PL_D * n; L.I = b + PL_I * n; L.gx = b + PL_GX * n; L.gy = b + PL_GY * n;
  for (int k = 0; k < 3; k++) { L.V[k] = b + (PL_V + k) * n; L.N[k] = b + (PL_N + k) * n; L.cov[k] = b + (PL_COV + k) * n; }
  return true;
}
// Offset of pyramid level l in a rendered-depth pyramid stored back to back
// (one guard element after each level).
size_t rd_offset(const mf_ctx* c, int l) {
  size_t off = 0;
  for (int k = 0; k < l; k++) off += (c->N0 >> (2 * k)) + 1;
  return off;
}

int make_object(mf_ctx* c, float size, int res, const float* T, long cap) {
  if (res < 16 || res > 8192 || (res & (res - 1)) || size <= 0)
    return fail(c, MF_E_ARG, "object resolution must be a power of two in [16, 8192]");
  if ((int)c->objs.size() >= MAX_INST - 3) return fail(c, MF_E_ARG, "too many objects");
  // a failed call frees what it allocated and leaves the context as it was
  const size_t mark = c->allocs.size();
  const long bytes_mark = c->bytes;
  Object* o = new Object();
  auto rollback = [&](int code, const char* msg) {
    hipStreamSynchronize(c->stream);
    for (size_t i = mark; i < c->allocs.size(); i++) hipFree(c->allocs[i]);
    c->allocs.resize(mark);
    c->bytes = bytes_mark;
    delete o;
    return fail(c, code, msg);
  };
  o->id = (int)c->objs.size();
  o->size = size; o->res = res;
  VolDev& v = o->v;
  v.res = res; v.size = size; v.voxel = size / (float)res;
  v.max_level = 0; while ((1 << v.max_level) < res) v.max_level++;
  v.leaf_level = v.max_level - 3;
  v.nb = res / 8;
  const long nb3 = (long)v.nb * v.nb * v.nb;
  if (cap <= 0) cap = std::min(nb3, (long)262144);
  cap = std::min(cap, nb3);
  // the raycast addresses TSDF voxels by a 32-bit byte offset from the pool
  if (cap > MAX_BLOCKS) return rollback(MF_E_ARG, "capacity above 2^21 voxel blocks per volume");
  v.cap = cap;
  v.slot = dalloc<int32_t>(c, nb3);
  if (v.slot) hipMemsetAsync(v.slot, 0xff, nb3 * sizeof(int32_t), c->stream);
  v.req = dalloc<uint8_t>(c, nb3);
  long off = 0;
  for (int l = 0; l < 24; l++) v.node_off[l] = 0;
  for (int l = 1; l < v.leaf_level; l++) { v.node_off[l] = off; off += 1L << (3 * l); }
  v.nodes = dalloc<uint8_t>(c, off);
  v.blk_morton = dalloc<int32_t>(c, cap);
  v.blk_active = dalloc<uint8_t>(c, cap);
  const size_t nv = (size_t)cap * BLOCK_VOX;
  v.tsdf = dalloc<float>(c, nv); v.wt = dalloc<float>(c, nv);
  v.r = dalloc<float>(c, nv); v.g = dalloc<float>(c, nv); v.b = dalloc<float>(c, nv);
  v.c = dalloc<int32_t>(c, nv); v.fgw = dalloc<int32_t>(c, nv); v.fg = dalloc<double>(c, nv);
  v.counters = dalloc<int32_t>(c, 8);
  o->mask = dalloc<float>(c, c->N0);
  {
    float* b = dalloc<float>(c, 6 * c->N0);   // vertex planes then normal planes
    for (int k = 0; k < 3; k++) { o->mv[k] = b ? b + k * c->N0 : nullptr; o->mn[k] = b ? b + (3 + k) * c->N0 : nullptr; }
  }
  o->codes = dalloc<int8_t>(c, 2 * c->N0);
  o->rgb_err = dalloc<float>(c, c->N0);
  o->tiles = dalloc<uint8_t>(c, (size_t)((c->W + MAP_TILE - 1) / MAP_TILE) * ((c->H + MAP_TILE - 1) / MAP_TILE));
  o->tbox = dalloc<int>(c, 4);
  {
    float* b = dalloc<float>(c, 3 * c->N0);
    for (int k = 0; k < 3; k++) o->mvbf[k] = b ? b + k * c->N0 : nullptr;
  }
  {
    float* b = dalloc<float>(c, rd_offset(c, c->cfg.n_levels));
    for (int l = 0; l < c->cfg.n_levels; l++) o->rD[l] = b ? b + rd_offset(c, l) : nullptr;
  }
  o->prm = dalloc<ObjParams>(c, 1);
  if (!v.slot || !v.req || !v.nodes || !v.tsdf || !v.fg || !v.c || !v.fgw || !o->mask ||
      !o->codes || !o->rgb_err || !o->prm || !o->mv[2] || !o->mn[2] || !o->mvbf[2] || !o->tiles || !o->tbox || !v.blk_morton ||
      !v.blk_active || !v.wt || !v.r || !v.g || !v.b || !v.counters || !o->rD[0])
    return rollback(MF_E_OOM, "device allocation failed");
  // INVALID normals in the model maps until the first raycast
  std::vector<float> inv(c->N0, (float)INVALID);
  hipMemcpyAsync(o->mn[0], inv.data(), c->N0 * sizeof(float), hipMemcpyHostToDevice, c->stream);
  // every tile flag 0 (zeroed) and an empty tile box until then
  static const int empty_box[4] = {0x7fffffff, 0x7fffffff, -1, -1};
  hipMemcpyAsync(o->tbox, empty_box, sizeof empty_box, hipMemcpyHostToDevice, c->stream);
  const Mat4 M = m4(T);
  hipMemcpyAsync(c->mats + 3 + o->id, &M, sizeof(Mat4), hipMemcpyHostToDevice, c->stream);
  // virtual_camera_pose_ starts at the current camera pose
  // (kfusion.cpp:1058-1061); a new object is in view (kfusion.cpp:122, 1079)
  hipMemcpyAsync(c->vcam + o->id, c->mats + 0, sizeof(Mat4), hipMemcpyDeviceToDevice, c->stream);
  const int one = 1;
  hipMemcpyAsync(&c->md->in_view[o->id], &one, sizeof one, hipMemcpyHostToDevice, c->stream);
  hipStreamSynchronize(c->stream);
  // mapping work lists cover every object's pool (grown with each object;
  // the previous lists stay allocated until the context is destroyed)
  const long work_cap = c->work_cap + cap;
  int4* act_list = dalloc<int4>(c, work_cap);
  if (!act_list) return rollback(MF_E_OOM, "device allocation failed");
  c->work_cap = work_cap;
  c->work.act_list = act_list;
  c->objs.push_back(o);
  return o->id;
}

TrackLaunch base_launch(mf_ctx* c, const float k[4], int level, int object) {
  TrackLaunch p;
  memset(&p, 0, sizeof p);
  p.level = level; p.w = c->live[level].w; p.h = c->live[level].h; p.W = c->W; p.H = c->H;
  p.use_icp = c->cfg.use_icp; p.use_rgb = c->cfg.use_rgb; p.object = object;
  p.robust = c->cfg.robust; p.robust_k = c->cfg.robust_k;
  p.fx = k[0]; p.fy = k[1]; p.cx = k[2]; p.cy = k[3];
  // sigma_bright_ = sqrtf(100.0)/255.0 (tracking.cpp:764,769); inv_cov = 1.0/sigma
  const float sigma_bright = (float)(sqrtf(100.0f) / 255.0);
  p.inv_cov_rgb = (float)(1.0 / (double)sigma_bright);
  p.live = c->live[level].D; p.ref = c->ref[level].D; p.lstride = c->live[level].w * c->live[level].h;
  p.rd_off = (int)rd_offset(c, level);
  p.label = nullptr; p.moving = c->moving; p.n_labels = MAX_INST;
  p.dbg = c->dbg_rows; p.n_levels = c->cfg.n_levels;
  return p;
}

// ICP + photometric rows run by pixel (both rows of a pixel on one thread),
// camera and objects, when the level's rows per thread is even.
void set_shape(TrackLaunch& p, int rows_per_thread) {
  p.paired = (p.use_icp && p.use_rgb && rows_per_thread % 2 == 0) ? 1 : 0;
}

// Rendered reference depth pyramids (tracking.cpp:218-238): vertex2depth at
// T_w_r (inverse in mats[2]) then robust half-sampling, for the tracking
// instances [first, first + n) of the descriptor table c->inst.
void render_reference(mf_ctx* c, int first, int n) {
  int off[MAX_LEVELS];
  for (int l = 0; l < MAX_LEVELS; l++) off[l] = (int)rd_offset(c, l);
  Timed t(c, KC_V2D);
  launch_render_reference(c->stream, c->inst + first, n, c->W, c->H, c->cfg.n_levels, c->mats + 2, off);
}

// Rows per thread of the objects' row grids at level l in a launch of n
// objects: the knob, or when it is 0, {8, 8, 4, 4} for two or more objects
// and the camera's for one.  Most of an object's rows lie away from it and
// end at their first loads, so with several objects the launch fills the
// chip with fewer, longer workgroups (fewer partials to reduce, and at the
// fused coarse levels every workgroup reads all of the previous iteration's
// partials); one object alone needs the camera's shape to fill it.  Measured
// on MI355X at 640x480: 8 objects 532 -> 548 frames/s with {8, 4, 4}, then
// 570 -> 575 with 8 at level 1; 16 objects 427 -> 450, 1280x720 with 16
// objects 215 -> 236 with {8, 4, 4}; {8, 4, 4} for a single object lost 1 %
// (5 % in the one-object-per-launch form at 8 objects).
int obj_rows_per_thread(const mf_ctx* c, int l, int n) {
  static constexpr int kBatched[MAX_LEVELS] = {8, 8, 4, 4};
  if (c->track_obj_rpt[l] > 0) return c->track_obj_rpt[l];
  return n >= 2 ? kBatched[l] : c->track_rpt[l];
}

// Whether the GN iterations of level l run as one fused launch each (its
// workgroups first finish the previous iteration themselves from all of its
// partials) for a launch of n instances with nblk row workgroups each: the
// knob, or when it is -1, at the coarse levels (l >= 1) when the launch has
// at most 1200 workgroups.  The fused form saves a launch per iteration but
// every workgroup reads all nblk partials and runs the solve step; measured
// on MI355X: 640x480 with 1 object 794 -> 810 frames/s, 8 objects 566 -> 570
// (one object per launch 313 -> 321); level 1 of 1280x720 (1800
// workgroups) fused 238 -> 220, and at 640x480 the level-0 launch fused is
// no faster.
bool fused_level(const mf_ctx* c, int l, int nblk, int n) {
  if (c->track_fused[l] >= 0) return c->track_fused[l] != 0;
  return l >= 1 && (long)nblk * n <= 1200;
}

// One coarse-to-fine GN run over `n` instances sharing a launch
// (trackLiveFrame / trackEachObject loops, tracking.cpp:242-335, 624-690):
// one fused track+reduce+finalize launch per (level, iteration).
void gn_run(mf_ctx* c, const float k[4], int object, int first_inst, int n, const int32_t* label_mask,
            const uint8_t* mask_labels = nullptr) {
  TrackSync sy;
  sy.trace = c->trace_on ? c->trace : nullptr; sy.trace_n = c->trace_n;
  sy.trace_cap = c->trace_cap;
  const int kinds = (c->cfg.use_icp ? 1 : 0) + (c->cfg.use_rgb ? 1 : 0);
  // launch q of the run reads state slot q&1; fused launches read partials
  // [(q+1)&1] (iteration q-1) and write partials[q&1]
  const size_t half = (size_t)MAX_INST * c->nblk0 * 32;
  int q = 0, nblk_prev = 1;
  bool pending = false;   // a fused launch left its rows' solve to the next launch
  TrackLaunch p = base_launch(c, k, c->cfg.n_levels - 1, object);
  auto finish = [&]() {
    if (!pending) return;
    Timed t(c, KC_FUSED);
    launch_gn_fused(c->stream, p, c->inst + first_inst, n, c->gn, c->partial + ((q + 1) & 1) * half, nblk_prev,
                    c->partial + (q & 1) * half, 1, q, 0, sy);
    ++q;
    pending = false;
  };
  for (int l = c->cfg.n_levels - 1; l >= 0; --l) {
    p = base_launch(c, k, l, object);
    p.label = label_mask;
    if (mask_labels) p.moving = mask_labels;
    const int rpt = object ? obj_rows_per_thread(c, l, n) : c->track_rpt[l];
    set_shape(p, rpt);
    const bool fused = fused_level(c, l, track_blocks(p.w * p.h * kinds, rpt), n);
    const bool objs_grid = object && c->cfg.batch_objects && !fused && c->track_obj_ppt[l] > 0;
    const int nblk = objs_grid ? track_obj_blocks(p.w * p.h, c->track_obj_ppt[l]) : track_blocks(p.w * p.h * kinds, rpt);
    if (!fused) finish();
    for (int it = 0; it < c->cfg.iters[l]; ++it, ++q) {
      p.iter = it;
      if (fused) {
        Timed t(c, KC_FUSED);
        launch_gn_fused(c->stream, p, c->inst + first_inst, n, c->gn, c->partial + ((q + 1) & 1) * half, nblk_prev,
                        c->partial + (q & 1) * half, nblk, q, 1, sy);
        pending = true;
      } else {
        { Timed t2(c, object ? KC_TRACK_OBJ : KC_TRACK);
          if (objs_grid)
            launch_track_rows_objs(c->stream, p, c->inst + first_inst, n, c->gn, c->partial + (q & 1) * half, nblk,
                                   c->track_obj_ppt[l], q);
          else
            launch_track_rows(c->stream, p, c->inst + first_inst, n, c->gn, c->partial + (q & 1) * half, nblk, q); }
        { Timed t2(c, KC_SOLVE);
          launch_track_solve(c->stream, p, c->inst + first_inst, n, c->gn, c->partial + (q & 1) * half, nblk, q, sy); }
      }
      nblk_prev = nblk;
    }
  }
  finish();
  // the run's result is the state of slot q&1 (poses are already published)
}

void set_inst(mf_ctx* c, int slot, float* const mv[3], float* const mn[3], float* const rD[MAX_LEVELS],
              int8_t* codes, float* rgb_err, int state, uint8_t* tiles = nullptr, const int* tbox = nullptr) {
  TrackInst ti;
  memset(&ti, 0, sizeof ti);
  ti.mvn = mv[0];      // planes: mv[0..2], mn[0..2] back to back
  ti.rD = rD[0];
  (void)mn;
  ti.codes = codes; ti.rgb_err = rgb_err; ti.state = state; ti.tiles = tiles; ti.tbox = tbox;
  // descriptors only change with the object set: upload on change, ordered
  // after every pending launch that may still read the old table
  if (memcmp(&ti, &c->inst_host[slot], sizeof ti) != 0) {
    c->inst_host[slot] = ti;
    hipMemcpyAsync(c->inst + slot, &c->inst_host[slot], sizeof ti, hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
  }
}

void set_ids(mf_ctx* c, const std::vector<int>& ids) {
  if (ids == c->ids_host) return;
  c->ids_host = ids;
  hipMemcpyAsync(c->obj_ids, c->ids_host.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
  hipStreamSynchronize(c->stream);
}

// Runs `enqueue` (kernel launches only: no copies, memsets or host syncs) on
// the context stream, through a captured graph keyed by every value its
// kernel arguments depend on.  Kernel timing needs per-launch events, so it
// bypasses the graphs.
uint64_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
template <class F>
int run_graph(mf_ctx* c, const std::vector<uint64_t>& key, F&& enqueue) {
  if (!c->graphs || c->timing) { enqueue(); return 0; }
  auto it = c->gcache.find(key);
  if (it == c->gcache.end()) {
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed) != hipSuccess)
      return fail(c, MF_E_HIP, "stream capture failed to start");
    enqueue();
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(c->stream, &g) != hipSuccess || !g) return fail(c, MF_E_HIP, "stream capture failed");
    hipGraphExec_t e = nullptr;
    const hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (r != hipSuccess) return fail(c, MF_E_HIP, "graph instantiation failed");
    it = c->gcache.emplace(key, e).first;
  }
  if (hipGraphLaunch(it->second, c->stream) != hipSuccess) return fail(c, MF_E_HIP, "graph launch failed");
  return 0;
}
std::vector<uint64_t> base_key(mf_ctx* c, int stage, const float k[4]) {
  std::vector<uint64_t> key = {(uint64_t)stage, (uint64_t)(uintptr_t)c->live[0].D, (uint64_t)(uintptr_t)c->ref[0].D,
                               fbits(k[0]), fbits(k[1]), fbits(k[2]), fbits(k[3]), (uint64_t)c->trace_on,
                               (uint64_t)(uintptr_t)c->dbg_rows};
  for (int l = 0; l < MAX_LEVELS; l++)
    key.push_back((uint64_t)c->track_rpt[l] | ((uint64_t)(c->track_fused[l] + 1) << 32) | ((uint64_t)c->track_obj_ppt[l] << 40) |
                  ((uint64_t)c->track_obj_rpt[l] << 48));
  return key;
}

// Mapping descriptor tables: entries [0, n_int) are the objects integrated
// this frame (those with a mask), [MAX_INST, MAX_INST + n_all) every object.
// Uploaded only when they change (ordered after pending launches).
void set_mapobjs(mf_ctx* c, int& n_int, int& n_all, long& sum_nb3, long& sum_cap) {
  std::vector<MapObj> t(2 * MAX_INST);
  memset(t.data(), 0, t.size() * sizeof(MapObj));
  n_int = 0; n_all = 0; sum_nb3 = 0; sum_cap = 0;
  for (Object* o : c->objs) {
    MapObj m;
    memset(&m, 0, sizeof m);
    m.v = o->v; m.prm = o->prm; m.mask = o->mask; m.id = o->id; m.tbox = o->tbox;
    // refine_mask_use_motion with the instance's trackresult_ (static objects
    // hold copies of the camera's rows, kfusion.cpp:399-402, 450-454)
    m.codes = o->codes;
    t[MAX_INST + n_all++] = m;
    if (o->has_mask) {
      t[n_int++] = m;
      sum_nb3 += (long)o->v.nb * o->v.nb * o->v.nb;
      sum_cap += (long)o->v.cap;
    }
  }
  if (memcmp(t.data(), c->mapobjs_host.data(), t.size() * sizeof(MapObj)) != 0) {
    c->mapobjs_host = t;
    hipMemcpyAsync(c->mapobjs, c->mapobjs_host.data(), t.size() * sizeof(MapObj), hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
  }
}

void download_planes(mf_ctx* c, float* const src[3], float* dst_aos, size_t n) {
  std::vector<float> tmp(3 * n);
  for (int k = 0; k < 3; k++) hipMemcpyAsync(tmp.data() + k * n, src[k], n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  for (size_t i = 0; i < n; i++) for (int k = 0; k < 3; k++) dst_aos[3 * i + k] = tmp[k * n + i];
}
void upload_planes(mf_ctx* c, float* const dst[3], const float* src_aos, size_t n) {
  std::vector<float> tmp(3 * n);
  for (size_t i = 0; i < n; i++) for (int k = 0; k < 3; k++) tmp[k * n + i] = src_aos[3 * i + k];
  for (int k = 0; k < 3; k++) hipMemcpyAsync(dst[k], tmp.data() + k * n, n * sizeof(float), hipMemcpyHostToDevice, c->stream);
  hipStreamSynchronize(c->stream);
}

}  // namespace

// ---- Scratch of the no-allocation voxel export (mf_object_voxels_device):
// word 0 counts the keys found, words [64, 64 + n) hold the per-key found
// flags.  A file-scope device array: every GPU the library runs on holds its
// own copy until the process ends, shared by every context and call on that
// GPU, so calls that overlap on one GPU race.  Host code never passes the
// array's name to a kernel: it takes the device address from
// hipGetSymbolAddress once per device (scratch_setup, from mf_create).
static constexpr long NOALLOC_MAX_KEYS = 65536;
static __device__ __align__(256) int scratch[64 + NOALLOC_MAX_KEYS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int* g_scratch[MAX_DEVICES] = {};

// Address of this device's `scratch`, looked up once per device with that
// device current.
int scratch_setup(mf_ctx* c) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
    return fail(c, MF_E_HIP, "no current HIP device for the no-allocation scratch");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(c, MF_E_HIP, "hipGetSymbolAddress of the no-allocation scratch failed");
    g_scratch[dev] = static_cast<int*>(p);
  }
  c->scratch = g_scratch[dev];
  return 0;
}

__global__ void k_fill_i32(int* p, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
// *total += number of non-zero flags (one workgroup)
__global__ __launch_bounds__(1024) void k_count_flags(const int* __restrict__ f, int n, int* total) {
  __shared__ int part[1024 / 64];
  int t = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) t += f[i] != 0;
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) a += part[w];
    *total += a;
  }
}
}  // namespace

extern "C" {

mf_ctx* mf_create(const mf_config* cfg, const float init_pose[16], float bg_size_m, int bg_res, long bg_capacity) {
  if (!cfg || cfg->width <= 0 || cfg->height <= 0 || cfg->n_levels < 1 || cfg->n_levels > MAX_LEVELS) return nullptr;
  mf_ctx* c = new mf_ctx();
  memset(c->inst_host, 0, sizeof c->inst_host);
  memset(c->ray_host, 0, sizeof c->ray_host);
  memset(c->ray_bf_host, 0, sizeof c->ray_bf_host);
  memset(c->ray_lab_host, 0, sizeof c->ray_lab_host);
  c->cfg = *cfg;
  if (c->cfg.track_rate <= 0) c->cfg.track_rate = 1;
  if (c->cfg.integ_rate <= 0) c->cfg.integ_rate = 1;
  c->W = cfg->width; c->H = cfg->height; c->N0 = (size_t)c->W * c->H;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return nullptr; }
  // a failed lookup leaves c->scratch null: mf_object_voxels_device then fails
  scratch_setup(c);
  bool ok = true;
  c->depth = dalloc<float>(c, c->N0); c->grey = dalloc<float>(c, c->N0);
  for (int k = 0; k < 3; k++) c->rgb[k] = dalloc<float>(c, c->N0);
  for (int l = 0; l < cfg->n_levels; l++) {
    ok = ok && alloc_level(c, c->live[l], c->W >> l, c->H >> l);
    ok = ok && alloc_level(c, c->ref[l], c->W >> l, c->H >> l);
  }
  {
    float* b = dalloc<float>(c, 6 * c->N0);   // vertex planes then normal planes
    for (int k = 0; k < 3; k++) { c->vtx[k] = b ? b + k * c->N0 : nullptr; c->nrm[k] = b ? b + (3 + k) * c->N0 : nullptr; }
  }
  c->label = dalloc<int32_t>(c, c->N0);
  {
    float* b = dalloc<float>(c, 3 * c->N0);
    for (int k = 0; k < 3; k++) c->vbf[k] = b ? b + k * c->N0 : nullptr;
  }
  c->label_bf = dalloc<int32_t>(c, c->N0);
  c->motion_codes = dalloc<int8_t>(c, 2 * c->N0);
  c->md = dalloc<MotionDev>(c, 1);
  c->vcam = dalloc<Mat4>(c, MAX_INST);
  c->rayobjs_bf = dalloc<RayObj>(c, MAX_INST);
  c->rayobjs_lab = dalloc<RayObj>(c, MAX_INST);
  c->motion_tab = dalloc<MotionObj>(c, MAX_INST);
  c->ray_pose = dalloc<Mat4>(c, 1);
  c->view_dev = dalloc<Mat4>(c, 1);
  c->palette_dev = dalloc<uint32_t>(c, 256);
  c->mats = dalloc<Mat4>(c, 3 + MAX_INST);
  c->gn = dalloc<GNState>(c, MAX_INST);
  c->inst = dalloc<TrackInst>(c, MAX_INST);
  c->nblk0 = track_blocks(2 * (int)c->N0, 1);
  c->partial = dalloc<float>(c, 2 * (size_t)MAX_INST * c->nblk0 * 32);   // two launch parities
  c->moving = dalloc<uint8_t>(c, MAX_INST);
  c->trace = dalloc<TraceRec>(c, c->trace_cap);
  c->trace_n = dalloc<int>(c, 1);
  c->rayobjs = dalloc<RayObj>(c, MAX_INST);
  c->mapobjs = dalloc<MapObj>(c, 2 * MAX_INST);
  c->gauss = dalloc<float>(c, 5);
  c->stats = dalloc<unsigned long long>(c, (size_t)N_STATS * STAT_SLOTS);
  c->obj_ids = dalloc<int>(c, MAX_INST);
  c->work.act_count = dalloc<int>(c, 1);
  ok = ok && c->work.act_count;
  ok = ok && c->vbf[2] && c->label_bf && c->motion_codes && c->md && c->vcam && c->rayobjs_bf && c->rayobjs_lab &&
       c->motion_tab && c->ray_pose && c->view_dev && c->palette_dev;
  ok = ok && c->stats && c->obj_ids && c->depth && c->grey && c->rgb[2] && c->vtx[2] && c->nrm[2] && c->label && c->mats && c->gn &&
       c->inst && c->partial && c->moving && c->trace && c->trace_n && c->rayobjs && c->mapobjs && c->gauss;
  if (!ok) { mf_destroy(c); return nullptr; }
  // Gaussian of the bilateral filter (kfusion.cpp:80-88): expf(-(x*x)/(2*delta^2)), delta = 4
  float g[5];
  for (int i = 0; i < 5; i++) { int x = i - 2; g[i] = expf((float)(-(x * x)) / (2.0f * 4.0f * 4.0f)); }
  hipMemcpyAsync(c->gauss, g, sizeof g, hipMemcpyHostToDevice, c->stream);
  const Mat4 P = m4(init_pose);
  c->init_t[0] = init_pose[3]; c->init_t[1] = init_pose[7]; c->init_t[2] = init_pose[11];
  hipMemcpyAsync(c->mats, &P, sizeof P, hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(c->mats + 1, &P, sizeof P, hipMemcpyHostToDevice, c->stream);
  // model normals INVALID until the first raycast
  std::vector<float> inv(c->N0, (float)INVALID);
  hipMemcpyAsync(c->nrm[0], inv.data(), c->N0 * sizeof(float), hipMemcpyHostToDevice, c->stream);
  std::vector<int32_t> il(c->N0, INVALID);
  hipMemcpyAsync(c->label, il.data(), c->N0 * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
  hipStreamSynchronize(c->stream);
  // background object, identity pose (kernels.h:146-154)
  float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  if (make_object(c, bg_size_m, bg_res, I, bg_capacity) != 0) { mf_destroy(c); return nullptr; }
  return c;
}

void mf_destroy(mf_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  for (auto& e : c->evs) { hipEventDestroy(e.a); hipEventDestroy(e.b); }
  for (auto e : c->pool) hipEventDestroy(e);
  for (auto& g : c->gcache) hipGraphExecDestroy(g.second);
  for (void* p : c->allocs) hipFree(p);
  for (int b = 0; b < 2; b++) {
    if (c->raw_pin[b]) hipHostFree(c->raw_pin[b]);
    if (c->raw_ev[b]) hipEventDestroy(c->raw_ev[b]);
  }
  for (Object* o : c->objs) delete o;
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

const char* mf_last_error(mf_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mf_object_create(mf_ctx* c, float size_m, int res, const float T_w_o[16], long capacity) {
  return make_object(c, size_m, res, T_w_o, capacity);
}

int mf_object_set_moving(mf_ctx* c, int id, int moving) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  c->objs[id]->moving = id > 0 && moving != 0;
  uint8_t f = c->objs[id]->moving ? 1 : 0;
  hipMemcpyAsync(c->moving + id, &f, 1, hipMemcpyHostToDevice, c->stream);
  hipStreamSynchronize(c->stream);
  return 0;
}

int mf_set_motion_detection(mf_ctx* c, int on) {
  c->motion_auto = on != 0;
  return 0;
}
int mf_get_motion(mf_ctx* c, int64_t* pixels, int64_t* inliers, int32_t* moving, int cap) {
  MotionDev m;
  uint8_t mv[MAX_INST];
  hipMemcpyAsync(&m, c->md, sizeof m, hipMemcpyDeviceToHost, c->stream);
  hipMemcpyAsync(mv, c->moving, MAX_INST, hipMemcpyDeviceToHost, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  const int n = (int)c->objs.size();
  for (int j = 0; j < n && j < cap; j++) {
    if (pixels) pixels[j] = m.counts[2 * j];
    if (inliers) inliers[j] = m.counts[2 * j + 1];
    if (moving) moving[j] = mv[j];
  }
  return n;
}
int mf_get_motion_state(mf_ctx* c, int32_t gate_rendered[2], int32_t* in_view, int32_t* move_set, int cap) {
  MotionDev m;
  hipMemcpyAsync(&m, c->md, sizeof m, hipMemcpyDeviceToHost, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  const int n = (int)c->objs.size();
  if (gate_rendered) { gate_rendered[0] = m.gate; gate_rendered[1] = m.rendered_n; }
  for (int j = 0; j < n && j < cap; j++) {
    if (in_view) in_view[j] = m.in_view[j];
    if (move_set) move_set[j] = m.move_set[j];
  }
  return n;
}
int mf_get_motion_codes(mf_ctx* c, int32_t* codes) {
  std::vector<int8_t> t(2 * c->N0);
  hipMemcpyAsync(t.data(), c->motion_codes, t.size(), hipMemcpyDeviceToHost, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  for (size_t i = 0; i < t.size(); i++) codes[i] = t[i];
  return 0;
}

int mf_object_set_mask(mf_ctx* c, int id, const float* mask) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  Object* o = c->objs[id];
  if (!mask) { o->has_mask = false; return 0; }
  hipMemcpyAsync(o->mask, mask, c->N0 * sizeof(float), hipMemcpyHostToDevice, c->stream);
  hipStreamSynchronize(c->stream);
  o->has_mask = true;
  return 0;
}
// instance_seg::generalize_label on the device (segmentation.cpp:10-47)
int mf_object_set_instance_mask_device(mf_ctx* c, int id, const uint8_t* d_mask, int rendered,
                                       const uint8_t* d_model_mask) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  Object* o = c->objs[id];
  if (!d_mask) { o->has_mask = false; return 0; }
  if (!c->mask_count) {
    c->mask_count = dalloc<int>(c, 1);
    if (!c->mask_count) return fail(c, MF_E_OOM, "mask scratch");
  }
  launch_generalize(c->stream, d_mask, d_model_mask, c->W, c->H, rendered ? 1 : 0, c->mask_count, o->mask);
  o->has_mask = true;
  return 0;
}
int mf_object_set_instance_mask(mf_ctx* c, int id, const uint8_t* mask, int rendered, const uint8_t* model_mask) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  if (!mask) { c->objs[id]->has_mask = false; return 0; }
  if (!c->mask_stage) {
    c->mask_stage = dalloc<uint8_t>(c, 2 * c->N0);
    if (!c->mask_stage) return fail(c, MF_E_OOM, "mask staging");
  }
  hipMemcpyAsync(c->mask_stage, mask, c->N0, hipMemcpyHostToDevice, c->stream);
  if (model_mask) hipMemcpyAsync(c->mask_stage + c->N0, model_mask, c->N0, hipMemcpyHostToDevice, c->stream);
  const int r = mf_object_set_instance_mask_device(c, id, c->mask_stage, rendered,
                                                   model_mask ? c->mask_stage + c->N0 : nullptr);
  // the staging buffer is reused by the next call: wait for this one
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  return r;
}

int mf_object_get_mask(mf_ctx* c, int id, float* mask) {
  if (id < 0 || id >= (int)c->objs.size() || !mask) return fail(c, MF_E_ARG, "bad object id");
  hipMemcpyAsync(mask, c->objs[id]->mask, c->N0 * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  return 0;
}

int mf_object_set_mask_device(mf_ctx* c, int id, const float* d_mask) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  Object* o = c->objs[id];
  if (!d_mask) { o->has_mask = false; return 0; }
  hipMemcpyAsync(o->mask, d_mask, c->N0 * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
  o->has_mask = true;
  return 0;
}

int mf_objects_set_masks_device(mf_ctx* c, int n, const int32_t* ids, const float* const* d_masks) {
  if (n < 0 || n > MAX_INST || (n > 0 && (!ids || !d_masks))) return fail(c, MF_E_ARG, "bad mask list");
  for (int i = 0; i < n; i++)
    if (ids[i] < 0 || ids[i] >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  MaskCopies mc;
  int m = 0;
  for (int i = 0; i < n; i++) {
    Object* o = c->objs[ids[i]];
    o->has_mask = d_masks[i] != nullptr;
    if (!d_masks[i]) continue;
    mc.src[m] = d_masks[i];
    mc.dst[m] = o->mask;
    ++m;
  }
  launch_copy_masks(c->stream, mc, m, c->N0);
  return 0;
}

int mf_preprocess_device(mf_ctx* c, const uint16_t* d_depth, const uint8_t* d_rgb, int iw, int ih, int filter) {
  const int W = c->W, H = c->H;
  if (iw < W || ih < H || iw % W || ih % H || iw / W != ih / H) return fail(c, MF_E_ARG, "Invalid ratio.");
  { Timed t(c, KC_CONVERT); launch_convert(c->stream, d_depth, d_rgb, iw, iw / W, W, H, c->depth, c->grey, c->rgb[0], c->rgb[1], c->rgb[2]); }
  if (filter) {
    // bilateral of the integration depth (kfusion.cpp:218-221), out of place
    float* tmp = c->ref[0].gx;   // scratch plane of the same size, rewritten every frame
    hipMemcpyAsync(tmp, c->depth, c->N0 * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
    Timed t(c, KC_BILATERAL);
    launch_bilateral(c->stream, tmp, c->depth, nullptr, nullptr, W, H, 0.1f, c->gauss);
  }
  return 0;
}

// Device input buffers for host frames of n pixels.
bool ensure_inputs(mf_ctx* c, size_t n) {
  if (n <= c->in_cap) return true;
  c->in_depth = dalloc<uint16_t>(c, n);
  c->in_rgb = dalloc<uint8_t>(c, 3 * n);
  if (!c->in_depth || !c->in_rgb) return false;
  c->in_cap = n;
  return true;
}

int mf_preprocess(mf_ctx* c, const uint16_t* depth, const uint8_t* rgb, int iw, int ih, int filter) {
  const size_t n = (size_t)iw * ih;
  if (!ensure_inputs(c, n)) return fail(c, MF_E_OOM, "input staging");
  hipMemcpyAsync(c->in_depth, depth, n * sizeof(uint16_t), hipMemcpyHostToDevice, c->stream);
  if (rgb) hipMemcpyAsync(c->in_rgb, rgb, 3 * n, hipMemcpyHostToDevice, c->stream);
  return mf_preprocess_device(c, c->in_depth, rgb ? c->in_rgb : nullptr, iw, ih, filter);
}

int mf_preprocess_raw(mf_ctx* c, mf_raw* raw, long frame, int filter) {
  int w = 0, h = 0;
  long frames = 0;
  if (!raw || mf_raw_info(raw, &w, &h, &frames, nullptr) != MF_OK) return fail(c, MF_E_ARG, "no raw sequence");
  const size_t n = (size_t)w * h;
  if (!ensure_inputs(c, n)) return fail(c, MF_E_OOM, "input staging");
  if (5 * n > c->raw_cap) {
    for (int b = 0; b < 2; b++) {
      if (c->raw_ev_used[b]) hipEventSynchronize(c->raw_ev[b]);
      if (c->raw_pin[b]) hipHostFree(c->raw_pin[b]);
      c->raw_pin[b] = nullptr;
      c->raw_ev_used[b] = false;
      if (hipHostMalloc((void**)&c->raw_pin[b], 5 * n, hipHostMallocDefault) != hipSuccess)
        return fail(c, MF_E_OOM, "pinned staging");
      if (!c->raw_ev[b] && hipEventCreateWithFlags(&c->raw_ev[b], hipEventDisableTiming) != hipSuccess)
        return fail(c, MF_E_HIP, "staging event");
    }
    c->raw_cap = 5 * n;
  }
  const int b = c->raw_next;
  c->raw_next ^= 1;
  if (c->raw_ev_used[b]) hipEventSynchronize(c->raw_ev[b]);   // the copy that last read this buffer has run
  uint8_t* p = c->raw_pin[b];
  if (mf_raw_read(raw, frame, reinterpret_cast<uint16_t*>(p), p + 2 * n) != MF_OK)
    return fail(c, MF_E_EOF, "end of the raw sequence");
  hipMemcpyAsync(c->in_depth, p, 2 * n, hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(c->in_rgb, p + 2 * n, 3 * n, hipMemcpyHostToDevice, c->stream);
  hipEventRecord(c->raw_ev[b], c->stream);
  c->raw_ev_used[b] = true;
  return mf_preprocess_device(c, c->in_depth, c->in_rgb, w, h, filter);
}

int mf_set_gt_poses(mf_ctx* c, const float* poses, long n) {
  if (n < 0 || (n > 0 && !poses)) return fail(c, MF_E_ARG, "bad ground-truth poses");
  c->gt.clear();
  for (long f = 0; f < n; f++) {
    // camera_pose = poses[frame] with its translation re-based to the
    // initial position (kfusion.cpp:263-266), in float
    Mat4 P = m4(poses + 16 * f);
    for (int r = 0; r < 3; r++) P.m[r][3] = (P.m[r][3] - poses[4 * r + 3]) + c->init_t[r];
    c->gt.push_back(P);
  }
  if (n > c->gt_cap) {
    c->gt_dev = dalloc<Mat4>(c, n);
    if (!c->gt_frame) c->gt_frame = dalloc<int>(c, 1);
    if (!c->gt_dev || !c->gt_frame) return fail(c, MF_E_OOM, "ground-truth poses");
    c->gt_cap = n;
  }
  if (n > 0) hipMemcpyAsync(c->gt_dev, c->gt.data(), n * sizeof(Mat4), hipMemcpyHostToDevice, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  return 0;
}

int mf_save_poses(mf_ctx* c, const char* prefix, int frame) {
  if (!prefix) return fail(c, MF_E_ARG, "no file prefix");
  const int n = (int)c->objs.size();
  std::vector<Mat4> M(3 + n);
  hipMemcpyAsync(M.data(), c->mats, M.size() * sizeof(Mat4), hipMemcpyDeviceToHost, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  if ((int)c->pose_saved.size() < n) c->pose_saved.resize(n, false);
  for (int id = 0; id < n; id++) {
    const Mat4& P = (id == 0) ? M[0] : M[3 + id];
    float T[16];
    for (int i = 0; i < 16; i++) T[i] = P.m[i / 4][i % 4];
    char line[256];
    if (mf::tum_line(line, sizeof line, frame, T) < 0) return fail(c, MF_E_ARG, "pose line");
    const std::string path = std::string(prefix) + (id == 0 ? std::string("_camera") : "_" + std::to_string(id));
    FILE* f = fopen(path.c_str(), c->pose_saved[id] ? "a" : "w");
    if (!f) return fail(c, MF_E_ARG, "cannot open pose file");
    fputs(line, f);
    fclose(f);
    c->pose_saved[id] = true;
  }
  return 0;
}

// Raycast descriptor tables: the model-map table (each object's maps) and the
// table of the raycasts before fusion (the objects' vertex maps before fusion
// only).  They change only with the object set.
void upload_ray_tables(mf_ctx* c) {
  const int n = (int)c->objs.size();
  if (c->ray_n == n) return;
  for (int i = 0; i < n; i++) {
    Object* o = c->objs[i];
    RayObj& r = c->ray_host[i];
    r.v = o->v; r.prm = o->prm; r.label = o->id;
    r.mvx = o->mv[0]; r.mvy = o->mv[1]; r.mvz = o->mv[2];
    r.mnx = o->mn[0]; r.mny = o->mn[1]; r.mnz = o->mn[2];
    // moving objects track against these maps: the raycast flags their tiles
    r.tiles = i > 0 ? o->tiles : nullptr; r.tbox = i > 0 ? o->tbox : nullptr;
    RayObj& lb = c->ray_lab_host[i];
    lb = r;
    lb.mvx = lb.mvy = lb.mvz = lb.mnx = lb.mny = lb.mnz = nullptr;
    lb.tiles = nullptr; lb.tbox = nullptr;
    RayObj& b = c->ray_bf_host[i];
    b = r;
    b.mnx = b.mny = b.mnz = nullptr;
    b.tiles = nullptr; b.tbox = nullptr;
    if (i == 0) { b.mvx = b.mvy = b.mvz = nullptr; }
    else { b.mvx = o->mvbf[0]; b.mvy = o->mvbf[1]; b.mvz = o->mvbf[2]; }
  }
  hipMemcpyAsync(c->rayobjs, c->ray_host, n * sizeof(RayObj), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(c->rayobjs_bf, c->ray_bf_host, n * sizeof(RayObj), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(c->rayobjs_lab, c->ray_lab_host, n * sizeof(RayObj), hipMemcpyHostToDevice, c->stream);
  hipStreamSynchronize(c->stream);
  c->ray_n = n;
}

// Raycast of every object at the current camera pose and the objects' current
// poses into the maps before fusion (kfusion.cpp:295-297, 432-434).  Only the
// vertices and labels of these maps are read downstream (compute_residuals,
// split_labels), so no normals are computed; the raycast flags the labels it
// writes in `present` (the instance set split_labels takes, kfusion.cpp:299).
void raycast_before_fusion(mf_ctx* c, const float k[4], const int* in_view, const int* gate, int* present) {
  hipStream_t s = c->stream;
  const int n = (int)c->objs.size();
  { Timed t(c, KC_PARAMS); launch_obj_params(s, c->mats + 0, c->mats, c->mapobjs + MAX_INST, n, k[0], k[1], k[2], k[3], 0, c->work); }
  Timed t(c, KC_RAYCAST);
  launch_raycast(s, c->rayobjs_bf, n, c->W, c->H, c->cfg.mu, 0.1f, 4.0f, c->ray_eps, c->vbf[0], c->vbf[1], c->vbf[2],
                 nullptr, nullptr, nullptr, c->label_bf, c->timing ? c->stats : nullptr, in_view, gate, present);
}

// compute_residuals launch of one instance against model maps `mvn` (6
// planes) and the vertex map before fusion `vbf` (3 planes).
MotionLaunch motion_launch(mf_ctx* c, const float k[4], const Mat4* T_w_l, const float* mvn, const float* vbf,
                           int8_t* codes, const int* gate) {
  MotionLaunch m;
  memset(&m, 0, sizeof m);
  m.W = c->W; m.H = c->H;
  m.live = c->live[0].D; m.ref = c->ref[0].D;
  m.mvn = mvn; m.vbf = vbf;
  m.T_w_l = T_w_l; m.inv_T_w_r = c->mats + 2;
  m.fx = k[0]; m.fy = k[1]; m.cx = k[2]; m.cy = k[3];
  m.use_icp = c->cfg.use_icp; m.use_rgb = c->cfg.use_rgb;
  m.sigma_bright = (float)(sqrtf(100.0f) / 255.0);   // tracking.cpp:764, 769
  m.threshold = 1.5f;                                 // residual_threshold_ (kernels.h:95)
  m.codes = codes; m.gate = gate;
  return m;
}

int mf_track_camera(mf_ctx* c, const float k[4], unsigned frame) {
  c->pending_objects = false;
  for (Object* o : c->objs) o->tracked = false;
  if (frame % c->cfg.track_rate != 0) return 0;
  hipStream_t s = c->stream;
  const int L = c->cfg.n_levels;
  // set_params_frame + buildPyramid (tracking.cpp:142-171, 776-818)
  auto pyramid = [&]() {
    { Timed t(c, KC_BILATERAL); launch_bilateral(s, c->grey, c->live[0].I, c->depth, c->live[0].D, c->W, c->H, 0.1f, c->gauss); }
    for (int l = 0; l < L; l++) {
      if (l > 0) { Timed t(c, KC_HALF); launch_half(s, c->live[l - 1].D, c->live[l].D, c->live[l - 1].I, c->live[l].I, c->live[l - 1].w, c->live[l - 1].h, 0.1f * 3); }
      const float sc = (float)(1 << l);
      const Mat4 invK = camKinv(k[0] / sc, k[1] / sc, k[2] / sc, k[3] / sc);
      Timed t(c, KC_LEVEL);
      launch_level(s, c->live[l], invK, k[1] < 0 ? 1 : 0, k[0] / sc, 1.0f, 1.0f, 0.075f, c->cfg.use_rgb != 0, c->cfg.use_icp != 0);
    }
  };
  const bool gt = !c->gt.empty();
  if (gt && (long)frame >= (long)c->gt.size()) return fail(c, MF_E_ARG, "no ground-truth pose for this frame");
  if ((int)frame == c->cfg.init_frame || !c->have_ref) {
    if (gt) hipMemcpyAsync(c->mats + 0, &c->gt[frame], sizeof(Mat4), hipMemcpyHostToDevice, s);
    pyramid();
    for (int l = 0; l < L; l++) std::swap(c->live[l], c->ref[l]);   // setRefImgFromCurr
    c->have_ref = true;
    return 0;
  }
  Object* bg = c->objs[0];
  // camera: trackLiveFrame against the global model maps
  float* camD[MAX_LEVELS];
  for (int l = 0; l < MAX_LEVELS; l++) camD[l] = bg->rD[l];
  set_inst(c, 0, c->vtx, c->nrm, camD, bg->codes, bg->rgb_err, 0);
  const int nobj = (int)c->objs.size();
  const bool multi = nobj >= 2;
  const bool autom = multi && c->motion_auto;
  if (multi) {
    if (c->ray_eps < 0) c->ray_eps = exp2f(-(float)c->objs[0]->v.max_level);
    upload_ray_tables(c);
    int n_int, n_all;
    long sum_nb3, sum_cap;
    set_mapobjs(c, n_int, n_all, sum_nb3, sum_cap);
  }
  MotionDev* md = c->md;
  // first pass, then (with objects) the raycast before fusion, the in-view
  // set and the moving set (kfusion.cpp:254-354)
  auto first_pass = [&]() {
    pyramid();
    { Timed t(c, KC_GNSMALL); launch_camera_begin(s, c->mats, c->gn, k[0], k[1], k[2], k[3]); }
    if (gt) {
      // the ground-truth pose replaces the camera's GN run (kfusion.cpp:262-268)
      Timed t(c, KC_GNSMALL);
      launch_pose_from_table(s, c->mats + 0, c->gt_dev, c->gt_frame);
    } else {
      if (c->cfg.use_rgb) render_reference(c, 0, 1);
      gn_run(c, k, 0, 0, 1, nullptr);
    }
    if (!multi) return;
    { Timed t(c, KC_MOTION); launch_motion_gate(s, md, nobj); }
    // the in-view set: the labels of this raycast (cleared by the gate)
    raycast_before_fusion(c, k, nullptr, &md->gate, md->in_view);
    { Timed t(c, KC_MOTION); launch_motion_inview(s, md, nobj, c->moving, autom ? 0 : 1); }
    if (autom) {
      MotionLaunch m = motion_launch(c, k, c->mats + 0, c->vtx[0], c->vbf[0], c->motion_codes, &md->gate);
      m.label_bf = c->label_bf; m.counts = md->counts;
      { Timed t(c, KC_MOTION); launch_motion_rows(s, m); }
      { Timed t(c, KC_MOTION); launch_motion_classify(s, md, nobj, c->moving); }
    }
  };
  // re-track the camera with the moving objects' raycast pixels masked
  // (kfusion.cpp:357-392, trackLiveFrame mask overload)
  auto retrack = [&]() {
    { Timed t(c, KC_GNSMALL); launch_gn_begin_camera(s, c->gn + 0, c->mats + 0, c->mats + 1, &md->retrack, k[0], k[1], k[2], k[3]); }
    gn_run(c, k, 0, 0, 1, c->label, md->move_set);
  };
  if (gt) launch_set_int(s, c->gt_frame, (int)frame);   // by value, ahead of the stage's graph
  std::vector<uint64_t> key = base_key(c, 1, k);
  key.push_back((uint64_t)nobj | ((uint64_t)autom << 32) | ((uint64_t)gt << 33));
  if (!autom) {
    bool any = false;
    for (int i = 1; i < nobj; i++) any = any || c->objs[i]->moving;
    key.push_back(any ? 1 : 0);
    const int r = run_graph(c, key, [&]() { first_pass(); if (any) retrack(); });
    if (r < 0) return r;
  } else {
    if (run_graph(c, key, first_pass) < 0) return -1;
    // the moving decision selects the launches that follow: read it back
    uint8_t mv[MAX_INST], ms[MAX_INST];
    hipMemcpyAsync(mv, c->moving, MAX_INST, hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(ms, md->move_set, MAX_INST, hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
    bool any = false;
    for (int i = 1; i < nobj; i++) { c->objs[i]->moving = mv[i] != 0; any = any || ms[i] != 0; }
    if (any) {
      std::vector<uint64_t> k2 = base_key(c, 5, k);
      if (run_graph(c, k2, retrack) < 0) return -1;
    }
  }
  c->pending_objects = true;
  return 1;
}

// Tracking::trackEachObject (tracking.cpp:580-708) of one object: GN state
// and descriptor slot OBJ_SLOT, ids entry OBJ_SLOT of obj_ids.
static constexpr int OBJ_SLOT = MAX_INST - 2;

int mf_track_object(mf_ctx* c, int id, const float k[4]) {
  if (!c->pending_objects) return fail(c, MF_E_ARG, "call between mf_track_camera and mf_track_objects");
  if (id <= 0 || id >= (int)c->objs.size() || !k) return fail(c, MF_E_ARG, "bad object id");
  Object* o = c->objs[id];
  if (o->tracked) return fail(c, MF_E_ARG, "object already tracked in this frame");
  hipStream_t s = c->stream;
  set_inst(c, OBJ_SLOT, o->mv, o->mn, o->rD, o->codes, o->rgb_err, OBJ_SLOT, o->tiles, o->tbox);
  const int oid = 3 + id;
  hipMemcpyAsync(c->obj_ids + OBJ_SLOT, &oid, sizeof oid, hipMemcpyHostToDevice, s);
  if (hipStreamSynchronize(s) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  std::vector<uint64_t> key = base_key(c, 6, k);
  key.push_back((uint64_t)id);
  const int r = run_graph(c, key, [&]() {
    { Timed t(c, KC_GNSMALL); launch_gn_begin_objects(s, c->gn, OBJ_SLOT, 1, c->obj_ids + OBJ_SLOT, c->mats, c->mats + 0, c->mats + 1, 1, nullptr); }
    if (c->cfg.use_rgb) render_reference(c, OBJ_SLOT, 1);
    gn_run(c, k, 1, OBJ_SLOT, 1, nullptr);
    { Timed t(c, KC_GNSMALL); launch_virtual_cams(s, c->gn, OBJ_SLOT, 1, c->obj_ids + OBJ_SLOT, c->vcam); }
  });
  if (r < 0) return r;
  o->tracked = true;
  return 0;
}

int mf_track_objects(mf_ctx* c, const float k[4], unsigned frame) {
  (void)frame;
  if (!c->pending_objects) return 0;
  c->pending_objects = false;
  hipStream_t s = c->stream;
  const int L = c->cfg.n_levels;
  const int nobj = (int)c->objs.size();
  std::vector<Object*> mov, stat, gnmov;
  for (int i = 1; i < nobj; i++) (c->objs[i]->moving ? mov : stat).push_back(c->objs[i]);
  // objects already tracked by mf_track_object keep that result
  for (Object* o : mov) if (!o->tracked) gnmov.push_back(o);
  if (nobj >= 2) {
    MotionDev* md = c->md;
    Object* bg = c->objs[0];
    std::vector<int> ids(gnmov.size());
    for (size_t j = 0; j < gnmov.size(); j++) {
      Object* o = gnmov[j];
      ids[j] = 3 + o->id;
      set_inst(c, 1 + (int)j, o->mv, o->mn, o->rD, o->codes, o->rgb_err, 1 + (int)j, o->tiles, o->tbox);
    }
    if (!gnmov.empty()) set_ids(c, ids);
    // the moving objects' motion-residual table (changes with the moving set)
    std::vector<MotionObj> mtab;
    for (Object* o : mov) mtab.push_back({o->mv[0], o->mvbf[0], c->vcam + o->id, o->codes, o->rgb_err});
    if (mtab.size() != c->motion_host.size() ||
        (!mtab.empty() && memcmp(mtab.data(), c->motion_host.data(), mtab.size() * sizeof(MotionObj)) != 0)) {
      c->motion_host = mtab;
      if (!mtab.empty()) {
        hipMemcpyAsync(c->motion_tab, c->motion_host.data(), mtab.size() * sizeof(MotionObj), hipMemcpyHostToDevice, s);
        hipStreamSynchronize(s);
      }
    }
    std::vector<uint64_t> key = base_key(c, 2, k);
    key.push_back((uint64_t)c->cfg.batch_objects);
    key.push_back((uint64_t)nobj);
    for (int id : ids) key.push_back((uint64_t)id);
    for (Object* o : mov) key.push_back((uint64_t)o->id | (1ull << 32));
    const int ng = (int)gnmov.size();
    const int r = run_graph(c, key, [&]() {
      if (!mov.empty()) {
        // trackEachObject for every moving object when the last raycast shows
        // more than one instance (kfusion.cpp:292, 395-414)
        if (ng > 0) {
          { Timed t(c, KC_GNSMALL); launch_gn_begin_objects(s, c->gn, 1, ng, c->obj_ids, c->mats, c->mats + 0, c->mats + 1, 1, &md->gate); }
          if (c->cfg.use_rgb) {
            if (c->cfg.batch_objects) render_reference(c, 1, ng);
            else for (int j = 0; j < ng; j++) render_reference(c, 1 + j, 1);
          }
          if (c->cfg.batch_objects) gn_run(c, k, 1, 1, ng, nullptr);
          else for (int j = 0; j < ng; j++) gn_run(c, k, 1, 1 + j, 1, nullptr);
          { Timed t(c, KC_GNSMALL); launch_virtual_cams(s, c->gn, 1, ng, c->obj_ids, c->vcam); }
        }
        // objects moving this frame: the camera's rows become motion residuals
        // against the maps before fusion, and a second raycast before fusion
        // renders the objects at their new poses (kfusion.cpp:416-441)
        { Timed t(c, KC_MOTION);
          MotionLaunch m = motion_launch(c, k, c->mats + 0, c->vtx[0], c->vbf[0], bg->codes, &md->move_any);
          m.rgb_err = bg->rgb_err;
          launch_motion_rows(s, m); }
        raycast_before_fusion(c, k, md->in_view, &md->move_any, md->present);
        { Timed t(c, KC_MOTION); launch_motion_after(s, md, nobj, c->cfg.use_rgb); }
        // each moving object's rows: motion residuals at its virtual camera
        // pose (kfusion.cpp:445-467), one launch for all of them
        { Timed t(c, KC_MOTION);
          launch_motion_rows_n(s, motion_launch(c, k, nullptr, nullptr, nullptr, nullptr, &md->cond2), c->motion_tab,
                               (int)mov.size()); }
      } else {
        Timed t(c, KC_MOTION);
        launch_motion_after(s, md, nobj, c->cfg.use_rgb);
      }
      // static objects carry the camera's rows (kfusion.cpp:399-402, 450-454)
      for (Object* o : stat) {
        Timed t(c, KC_MOTION);
        launch_copy_codes(s, bg->codes, o->codes, 2 * c->N0, bg->rgb_err, o->rgb_err, c->N0, &md->copy_static);
      }
    });
    if (r < 0) return r;
  }
  for (Object* o : c->objs) o->tracked = false;
  for (int l = 0; l < L; l++) std::swap(c->live[l], c->ref[l]);   // setRefImgFromCurr
  return 1;
}

int mf_track(mf_ctx* c, const float k[4], unsigned frame) {
  const int r = mf_track_camera(c, k, frame);
  if (r <= 0) return r;
  return mf_track_objects(c, k, frame);
}

int mf_integrate(mf_ctx* c, const float k[4], unsigned frame) {
  if (!((frame % c->cfg.integ_rate) == 0 || frame <= 3)) return 0;
  hipStream_t s = c->stream;
  const float band = 2.0f * c->cfg.mu;
  const int refine = frame > 1 ? 1 : 0;
  int n_int, n_all;
  long sum_nb3, sum_cap;
  set_mapobjs(c, n_int, n_all, sum_nb3, sum_cap);
  if (n_int == 0) return 1;
  std::vector<uint64_t> key = base_key(c, 3, k);
  key.push_back((uint64_t)refine);
  key.push_back((uint64_t)c->cfg.batch_objects);
  for (Object* o : c->objs) key.push_back((uint64_t)o->has_mask | ((uint64_t)o->moving << 1));
  unsigned long long* stats = c->timing ? c->stats : nullptr;
  // all integrated objects per launch, or one object per launch
  const int per = c->cfg.batch_objects ? n_int : 1;
  return run_graph(c, key, [&]() {
    for (int j0 = 0; j0 < n_int; j0 += per) {
      const MapObj* t = c->mapobjs + j0;
      MapWork wl = c->work;
      long cap = 0;
      wl.table_pre[0] = 0;
      for (int j = 0; j < per; j++) {
        const VolDev& v = c->mapobjs_host[j0 + j].v;
        wl.table_pre[j + 1] = wl.table_pre[j] + (long)v.nb * v.nb * v.nb;
        cap += v.cap;
      }
      { Timed tt(c, KC_PARAMS); launch_obj_params(s, c->mats + 0, c->mats, t, per, k[0], k[1], k[2], k[3], OBJ_ALLOC_RESET, wl); }
      { Timed tt(c, KC_ALLOC); launch_alloc(s, t, per, c->depth, refine, c->cfg.use_icp, c->cfg.use_rgb, c->W, c->H, band, stats); }
      { Timed tt(c, KC_COMMIT); launch_alloc_commit(s, t, per, wl); }
      { Timed tt(c, KC_INTEGRATE);
        launch_integrate(s, t, per, cap, k[0], k[1], k[2], k[3], c->depth, c->rgb[0], c->rgb[1], c->rgb[2],
                         refine, c->cfg.use_icp, c->cfg.use_rgb, c->W, c->H, c->cfg.mu, 100, wl, stats); }
    }
  }) < 0 ? -1 : 1;
}

int mf_raycast(mf_ctx* c, const float k[4], unsigned frame) {
  (void)frame;
  hipStream_t s = c->stream;
  if (c->ray_eps < 0) c->ray_eps = exp2f(-(float)c->objs[0]->v.max_level);   // ray_iterator's static epsilon
  const int n = (int)c->objs.size();
  upload_ray_tables(c);
  std::vector<uint64_t> key = base_key(c, 4, k);
  key.push_back((uint64_t)n);
  key.push_back(fbits(c->ray_eps));
  int n_int, n_all;
  long sum_nb3, sum_cap;
  set_mapobjs(c, n_int, n_all, sum_nb3, sum_cap);
  return run_graph(c, key, [&]() {
    // the objects' model-map tile boxes and the label image's instance set
    // restart empty (this raycast fills them); raycastPose = camera_pose
    // (kfusion.cpp:554), nothing in the raycast reads it
    { Timed t(c, KC_PARAMS);
      launch_obj_params(s, c->mats + 0, c->mats, c->mapobjs + MAX_INST, n_all, k[0], k[1], k[2], k[3], OBJ_TBOX_RESET,
                        c->work, c->ray_pose, c->md->present_main); }
    Timed t(c, KC_RAYCAST);
    // objects in view only (kfusion.cpp:557-558)
    launch_raycast(s, c->rayobjs, n, c->W, c->H, c->cfg.mu, 0.1f, 4.0f, c->ray_eps, c->vtx[0], c->vtx[1],
                   c->vtx[2], c->nrm[0], c->nrm[1], c->nrm[2], c->label, c->timing ? c->stats : nullptr,
                   c->md->in_view, nullptr, c->md->present_main);
  }) < 0 ? -1 : 1;
}

// ---- meshing (algorithms::marching_cube, meshing.hpp:157-208; Kfusion::
// dump_mesh, kfusion.cpp:1555-1588)

long mf_object_mesh(mf_ctx* c, int id, float* tri, long cap) {
  if (id < 0 || id >= (int)c->objs.size() || cap < 0 || (cap && !tri)) return fail(c, MF_E_ARG, "bad mesh request");
  hipStream_t s = c->stream;
  const VolDev& v = c->objs[id]->v;
  int cnt[4];
  hipMemcpyAsync(cnt, v.counters, sizeof cnt, hipMemcpyDeviceToHost, s);
  if (!c->mesh_total) c->mesh_total = dalloc<unsigned long long>(c, 1);
  if (!c->mesh_total) return fail(c, MF_E_OOM, "mesh counter");
  if (hipStreamSynchronize(s) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  const int nblocks = (int)std::min<long>(cnt[0], v.cap);
  unsigned long long total = 0;
  hipMemsetAsync(c->mesh_total, 0, sizeof(unsigned long long), s);
  launch_mesh(s, v, nblocks, nullptr, 0, c->mesh_total);
  hipMemcpyAsync(&total, c->mesh_total, sizeof total, hipMemcpyDeviceToHost, s);
  if (hipStreamSynchronize(s) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  const long want = std::min<long>((long)total, cap);
  if (want > 0) {
    float* d = nullptr;
    if (hipMalloc((void**)&d, (size_t)want * 9 * sizeof(float)) != hipSuccess) return fail(c, MF_E_OOM, "mesh buffer");
    hipMemsetAsync(c->mesh_total, 0, sizeof(unsigned long long), s);
    launch_mesh(s, v, nblocks, d, want, c->mesh_total);
    hipMemcpyAsync(tri, d, (size_t)want * 9 * sizeof(float), hipMemcpyDeviceToHost, s);
    const hipError_t e = hipStreamSynchronize(s);
    hipFree(d);
    if (e != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(e));
  }
  return (long)total;
}

int mf_dump_mesh(mf_ctx* c, const char* prefix) {
  if (!prefix) return fail(c, MF_E_ARG, "no file prefix");
  for (Object* o : c->objs) {
    const long n = mf_object_mesh(c, o->id, nullptr, 0);
    if (n < 0) return (int)n;
    std::vector<float> t((size_t)std::max(n, 1L) * 9);
    if (n > 0 && mf_object_mesh(c, o->id, t.data(), n) < 0) return -1;
    // writeVtkMesh (commons.h:945-1011), object frame, metres
    const std::string path = std::string(prefix) + "_" + std::to_string(o->id) + ".vtk";
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return fail(c, MF_E_ARG, "cannot open mesh file");
    fprintf(f, "# vtk DataFile Version 1.0\nvtk mesh generated from KFusion\nASCII\nDATASET POLYDATA\n");
    fprintf(f, "POINTS %ld FLOAT\n", 3 * n);
    for (long i = 0; i < 3 * n; i++) fprintf(f, "%g %g %g\n", t[3 * i], t[3 * i + 1], t[3 * i + 2]);
    fprintf(f, "POLYGONS %ld %ld\n", n, 4 * n);
    for (long i = 0; i < n; i++) fprintf(f, "3 %ld %ld %ld\n", 3 * i, 3 * i + 1, 3 * i + 2);
    fprintf(f, "\n");
    fclose(f);
  }
  return 0;
}

// ---- display rendering (rendering/rendering.cpp, kfusion.cpp:1284-1475)

bool ensure_render(mf_ctx* c, size_t n) {
  if (n <= c->render_cap) return true;
  c->render_out = dalloc<uint32_t>(c, n);
  c->label_scratch = dalloc<int32_t>(c, n);
  if (!c->render_out || !c->label_scratch) return false;
  c->render_cap = n;
  return true;
}
int download_image(mf_ctx* c, void* out, size_t bytes, const void* src) {
  hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  return 0;
}

int mf_render_track(mf_ctx* c, int id, int mode, uint8_t* out) {
  if (id < 0 || id >= (int)c->objs.size() || mode < 0 || mode > 2 || !out) return fail(c, MF_E_ARG, "bad render request");
  if (!ensure_render(c, c->N0)) return fail(c, MF_E_OOM, "render buffer");
  const Object* o = c->objs[id];
  launch_render_track(c->stream, o->codes, o->rgb_err, (int)c->N0, mode, c->render_out);
  return download_image(c, out, 4 * c->N0, c->render_out);
}

int mf_render_depth(mf_ctx* c, uint8_t* out) {
  if (!out) return fail(c, MF_E_ARG, "no output");
  if (!ensure_render(c, c->N0)) return fail(c, MF_E_OOM, "render buffer");
  launch_render_depth(c->stream, c->depth, (int)c->N0, 0.1f, 4.0f, c->render_out);   // nearPlane, farPlane
  return download_image(c, out, 4 * c->N0, c->render_out);
}

int mf_render_intensity(mf_ctx* c, uint8_t* out) {
  if (!out) return fail(c, MF_E_ARG, "no output");
  if (!ensure_render(c, c->N0)) return fail(c, MF_E_OOM, "render buffer");
  launch_render_intensity(c->stream, c->grey, (int)c->N0, c->render_out);
  return download_image(c, out, 4 * c->N0, c->render_out);
}

int mf_volume2mask(mf_ctx* c, const float k[4], int32_t* labels) {
  if (!labels) return fail(c, MF_E_ARG, "no output");
  if (!ensure_render(c, c->N0)) return fail(c, MF_E_OOM, "render buffer");
  if (c->ray_eps < 0) c->ray_eps = exp2f(-(float)c->objs[0]->v.max_level);
  upload_ray_tables(c);
  int n_int, n_all;
  long sum_nb3, sum_cap;
  set_mapobjs(c, n_int, n_all, sum_nb3, sum_cap);
  hipStream_t s = c->stream;
  launch_obj_params(s, c->mats + 0, c->mats, c->mapobjs + MAX_INST, n_all, k[0], k[1], k[2], k[3], 0, c->work);
  // every object, labels only, at the camera pose (rendering.cpp:240-343)
  launch_raycast(s, c->rayobjs_lab, n_all, c->W, c->H, c->cfg.mu, 0.1f, 4.0f, c->ray_eps, nullptr, nullptr, nullptr,
                 nullptr, nullptr, nullptr, c->label_scratch, nullptr, nullptr, nullptr);
  return download_image(c, labels, 4 * c->N0, c->label_scratch);
}

int mf_render_volume(mf_ctx* c, const float k[4], const float view_pose[16], int w, int h, float largestep,
                     int render_color, const uint8_t* palette, int n_palette, uint8_t* out) {
  if (!k || !view_pose || !out || w <= 0 || h <= 0 || n_palette < 0 || n_palette > 256 || (n_palette && !palette))
    return fail(c, MF_E_ARG, "bad render request");
  const size_t n = (size_t)w * h;
  if (!ensure_render(c, n)) return fail(c, MF_E_OOM, "render buffer");
  hipStream_t s = c->stream;
  if (c->ray_eps < 0) c->ray_eps = exp2f(-(float)c->objs[0]->v.max_level);
  // raycast again unless the view is the last raycast's pose at the
  // computation size (kfusion.cpp:1305-1307)
  Mat4 rp;
  hipMemcpyAsync(&rp, c->ray_pose, sizeof rp, hipMemcpyDeviceToHost, s);
  if (hipStreamSynchronize(s) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  const Mat4 V = m4(view_pose);
  bool same = true;
  for (int i = 0; i < 16; i++) same = same && (V.m[i / 4][i % 4] == rp.m[i / 4][i % 4]);
  const bool doray = !same || w != c->W || h != c->H;
  upload_ray_tables(c);
  int n_int, n_all;
  long sum_nb3, sum_cap;
  set_mapobjs(c, n_int, n_all, sum_nb3, sum_cap);
  std::vector<uint32_t> pal(256, 0u);
  for (int i = 0; i < n_palette; i++)
    pal[i] = palette[4 * i] | (palette[4 * i + 1] << 8) | (palette[4 * i + 2] << 16) | ((uint32_t)palette[4 * i + 3] << 24);
  hipMemcpyAsync(c->palette_dev, pal.data(), pal.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s);
  hipMemcpyAsync(c->view_dev, &V, sizeof V, hipMemcpyHostToDevice, s);
  launch_obj_params(s, c->view_dev, c->mats, c->mapobjs + MAX_INST, n_all, k[0], k[1], k[2], k[3], 0, c->work);
  if (doray)
    launch_render_volume_ray(s, c->rayobjs_lab, n_all, w, h, c->view_dev, c->cfg.mu, 0.1f, 2.0f * 4.0f, largestep,
                             c->ray_eps, render_color, c->palette_dev, n_palette, c->render_out);
  else
    launch_render_volume_maps(s, c->rayobjs_lab, n_all, (int)n, c->view_dev, c->vtx, c->nrm, c->label, render_color,
                              c->palette_dev, n_palette, c->render_out);
  const int r = download_image(c, out, 4 * n, c->render_out);
  return r < 0 ? r : (doray ? 1 : 0);
}

int mf_get_track_errors(mf_ctx* c, int id, float* rgb_err) {
  if (id < 0 || id >= (int)c->objs.size() || !rgb_err) return fail(c, MF_E_ARG, "bad object id");
  return download_image(c, rgb_err, sizeof(float) * c->N0, c->objs[id]->rgb_err);
}
int mf_set_track_errors(mf_ctx* c, int id, const float* rgb_err) {
  if (id < 0 || id >= (int)c->objs.size() || !rgb_err) return fail(c, MF_E_ARG, "bad object id");
  hipMemcpyAsync(c->objs[id]->rgb_err, rgb_err, sizeof(float) * c->N0, hipMemcpyHostToDevice, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  return 0;
}

int mf_sync(mf_ctx* c) {
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(e));
  for (Object* o : c->objs) {
    int32_t cnt[4];
    hipMemcpy(cnt, o->v.counters, sizeof cnt, hipMemcpyDeviceToHost);
    if (cnt[1]) return fail(c, MF_E_CAPACITY, "voxel block pool full");
  }
  return 0;
}

int mf_get_pose(mf_ctx* c, int id, float T[16]) {
  if (id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  Mat4 M;
  hipMemcpyAsync(&M, c->mats + (id < 0 ? 0 : 3 + id), sizeof M, hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  for (int i = 0; i < 16; i++) T[i] = M.m[i / 4][i % 4];
  return 0;
}
int mf_set_pose(mf_ctx* c, int id, const float T[16]) {
  if (id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  const Mat4 M = m4(T);
  hipMemcpyAsync(c->mats + (id < 0 ? 0 : 3 + id), &M, sizeof M, hipMemcpyHostToDevice, c->stream);
  hipStreamSynchronize(c->stream);
  return 0;
}

int mf_get_maps(mf_ctx* c, float* vertex, float* normal, int32_t* label) {
  if (vertex) download_planes(c, c->vtx, vertex, c->N0);
  if (normal) download_planes(c, c->nrm, normal, c->N0);
  if (label) { hipMemcpyAsync(label, c->label, c->N0 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream); hipStreamSynchronize(c->stream); }
  return 0;
}
int mf_set_maps(mf_ctx* c, const float* vertex, const float* normal, const int32_t* label) {
  if (vertex) upload_planes(c, c->vtx, vertex, c->N0);
  if (normal) upload_planes(c, c->nrm, normal, c->N0);
  if (label) {
    hipMemcpyAsync(c->label, label, c->N0 * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
    // the label image's instance set, which the frame-start gate reads
    int present[MAX_INST] = {};
    for (size_t i = 0; i < c->N0; i++)
      if (label[i] >= 0 && label[i] < MAX_INST) present[label[i]] = 1;
    hipMemcpyAsync(c->md->present_main, present, sizeof present, hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
  }
  return 0;
}
int mf_get_object_maps(mf_ctx* c, int id, float* vertex, float* normal) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  if (vertex) download_planes(c, c->objs[id]->mv, vertex, c->N0);
  if (normal) download_planes(c, c->objs[id]->mn, normal, c->N0);
  return 0;
}
int mf_set_object_maps(mf_ctx* c, int id, const float* vertex, const float* normal) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  Object* o = c->objs[id];
  if (vertex) upload_planes(c, o->mv, vertex, c->N0);
  if (normal) upload_planes(c, o->mn, normal, c->N0);
  // their tile flags and box, which the main raycast computes for its maps
  static const int empty_box[4] = {0x7fffffff, 0x7fffffff, -1, -1};
  hipMemcpyAsync(o->tbox, empty_box, sizeof empty_box, hipMemcpyHostToDevice, c->stream);
  launch_map_tiles(c->stream, o->mv[0], o->tiles, o->tbox, c->W, c->H);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  return 0;
}

long mf_object_keys(mf_ctx* c, int id, uint64_t* out, long cap) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  const VolDev& v = c->objs[id]->v;
  int32_t cnt[4];
  hipMemcpyAsync(cnt, v.counters, sizeof cnt, hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  const int n = std::min(cnt[0], (int)v.cap);
  std::vector<int32_t> m(std::max(n, 1));
  hipMemcpy(m.data(), v.blk_morton, n * sizeof(int32_t), hipMemcpyDeviceToHost);
  std::vector<uint64_t> keys(n);
  for (int i = 0; i < n; i++) keys[i] = ((uint64_t)(uint32_t)m[i] << 9) | (uint64_t)v.leaf_level;
  std::sort(keys.begin(), keys.end());
  for (long i = 0; i < n && i < cap; i++) out[i] = keys[i];
  return n;
}

int mf_object_block(mf_ctx* c, int id, int x, int y, int z, double* out) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  const VolDev& v = c->objs[id]->v;
  if (x < 0 || y < 0 || z < 0 || x >= v.res || y >= v.res || z >= v.res) return 0;
  const uint64_t m = morton3((uint32_t)x >> 3, (uint32_t)y >> 3, (uint32_t)z >> 3);
  int32_t s = -1;
  hipMemcpyAsync(&s, v.slot + m, sizeof s, hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  if (s < 0) return 0;
  std::vector<float> f(5 * BLOCK_VOX);
  std::vector<int32_t> iv(2 * BLOCK_VOX);
  std::vector<double> fg(BLOCK_VOX);
  const size_t o = (size_t)s * BLOCK_VOX, bs = BLOCK_VOX * sizeof(float);
  hipMemcpy(f.data(), v.tsdf + o, bs, hipMemcpyDeviceToHost);
  hipMemcpy(f.data() + BLOCK_VOX, v.wt + o, bs, hipMemcpyDeviceToHost);
  hipMemcpy(f.data() + 2 * BLOCK_VOX, v.r + o, bs, hipMemcpyDeviceToHost);
  hipMemcpy(f.data() + 3 * BLOCK_VOX, v.g + o, bs, hipMemcpyDeviceToHost);
  hipMemcpy(f.data() + 4 * BLOCK_VOX, v.b + o, bs, hipMemcpyDeviceToHost);
  hipMemcpy(iv.data(), v.c + o, BLOCK_VOX * sizeof(int32_t), hipMemcpyDeviceToHost);
  hipMemcpy(iv.data() + BLOCK_VOX, v.fgw + o, BLOCK_VOX * sizeof(int32_t), hipMemcpyDeviceToHost);
  hipMemcpy(fg.data(), v.fg + o, BLOCK_VOX * sizeof(double), hipMemcpyDeviceToHost);
  for (int i = 0; i < BLOCK_VOX; i++) {
    double* d = out + 8 * i;
    d[0] = f[i]; d[1] = f[BLOCK_VOX + i]; d[2] = f[2 * BLOCK_VOX + i]; d[3] = f[3 * BLOCK_VOX + i];
    d[4] = f[4 * BLOCK_VOX + i]; d[5] = iv[i]; d[6] = fg[i]; d[7] = iv[BLOCK_VOX + i];
  }
  return 1;
}

long mf_object_voxels(mf_ctx* c, int id, const uint64_t* keys, long n, double* out) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  if (n < 0 || (n > 0 && (!keys || !out))) return fail(c, MF_E_ARG, "bad voxel export arguments");
  const VolDev& v = c->objs[id]->v;
  hipStream_t s = c->stream;
  const long chunk = 4096;   // keys per round: 128 MB of doubles on the device
  const long cn = std::min(n, chunk);
  uint64_t* d_keys = nullptr;
  double* d_out = nullptr;
  int* d_found = nullptr;
  if (cn > 0 && (hipMalloc(&d_keys, cn * sizeof(uint64_t)) != hipSuccess ||
                 hipMalloc(&d_out, cn * BLOCK_VOX * 8 * sizeof(double)) != hipSuccess ||
                 hipMalloc(&d_found, cn * sizeof(int)) != hipSuccess)) {
    hipFree(d_keys); hipFree(d_out); hipFree(d_found);
    return fail(c, MF_E_OOM, "voxel export buffers");
  }
  long nfound = 0;
  std::vector<int> f(cn);
  for (long k0 = 0; k0 < n; k0 += chunk) {
    const long k = std::min(chunk, n - k0);
    hipMemcpyAsync(d_keys, keys + k0, k * sizeof(uint64_t), hipMemcpyHostToDevice, s);
    launch_export_voxels(s, v, d_keys, (int)k, d_out, d_found);
    hipMemcpyAsync(out + k0 * BLOCK_VOX * 8, d_out, k * BLOCK_VOX * 8 * sizeof(double), hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(f.data(), d_found, k * sizeof(int), hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) {
      hipFree(d_keys); hipFree(d_out); hipFree(d_found);
      return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
    }
    for (long i = 0; i < k; i++) nfound += f[i];
  }
  hipFree(d_keys); hipFree(d_out); hipFree(d_found);
  return nfound;
}

// mf_object_voxels on device buffers with no allocation: the found flags and
// their count live in this device's `scratch`.
long mf_object_voxels_device(mf_ctx* c, int id, const uint64_t* d_keys, long n, double* d_out) {
  if (!c) return MF_E_ARG;
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  if (n < 0 || n > NOALLOC_MAX_KEYS)
    return fail(c, MF_E_ARG, "mf_object_voxels_device takes at most 65536 keys per call");
  if (n > 0 && (!d_keys || !d_out)) return fail(c, MF_E_ARG, "bad voxel export arguments");
  if (!c->scratch) return fail(c, MF_E_HIP, "no-allocation scratch unavailable on this device");
  hipStream_t s = c->stream;
  int* count = c->scratch;
  int* found = c->scratch + 64;
  k_fill_i32<<<1, 64, 0, s>>>(count, 1, 0);
  if (n > 0) {
    launch_export_voxels(s, c->objs[id]->v, d_keys, (int)n, d_out, found);
    k_count_flags<<<1, 1024, 0, s>>>(found, (int)n, count);
  }
  int total = 0;
  hipMemcpyAsync(&total, count, sizeof total, hipMemcpyDeviceToHost, s);
  const hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(e));
  return total;
}

long mf_object_node_count(mf_ctx* c, int id) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  const VolDev& v = c->objs[id]->v;
  long tot = 0;
  for (int l = 1; l < v.leaf_level; l++) tot += 1L << (3 * l);
  std::vector<uint8_t> b(std::max(tot, 1L));
  hipStreamSynchronize(c->stream);
  hipMemcpy(b.data(), v.nodes, tot, hipMemcpyDeviceToHost);
  long n = 0;
  for (long i = 0; i < tot; i++) n += b[i] != 0;
  return n;
}

int mf_get_trackresult(mf_ctx* c, int id, int32_t* codes) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  Object* o = c->objs[id];
  const int8_t* src = o->codes;
  std::vector<int8_t> t(2 * c->N0);
  hipMemcpyAsync(t.data(), src, t.size(), hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  for (size_t i = 0; i < t.size(); i++) codes[i] = t[i];
  return 0;
}

int mf_set_trackresult(mf_ctx* c, int id, const int32_t* codes) {
  if (id < 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad object id");
  std::vector<int8_t> t(2 * c->N0);
  for (size_t i = 0; i < t.size(); i++) t[i] = (int8_t)codes[i];
  hipMemcpyAsync(c->objs[id]->codes, t.data(), t.size(), hipMemcpyHostToDevice, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  return 0;
}

int mf_motion_residuals(mf_ctx* c, const float k[4], const float T16[16], const float* vertex, const float* normal,
                        const float* vertex_bf, int32_t* codes) {
  if (!c->pending_objects) return fail(c, MF_E_ARG, "call between mf_track_camera and mf_track_objects");
  if (!c->eval_maps) {
    c->eval_maps = dalloc<float>(c, 9 * c->N0);
    c->eval_T = dalloc<Mat4>(c, 1);
    c->eval_codes = dalloc<int8_t>(c, 2 * c->N0);
    if (!c->eval_maps || !c->eval_T || !c->eval_codes) return fail(c, MF_E_OOM, "device allocation failed");
  }
  float* mv[3]; float* mn[3]; float* vb[3];
  for (int k2 = 0; k2 < 3; k2++) {
    mv[k2] = c->eval_maps + k2 * c->N0; mn[k2] = c->eval_maps + (3 + k2) * c->N0; vb[k2] = c->eval_maps + (6 + k2) * c->N0;
  }
  upload_planes(c, mv, vertex, c->N0);
  upload_planes(c, mn, normal, c->N0);
  upload_planes(c, vb, vertex_bf, c->N0);
  const Mat4 T = m4(T16);
  hipMemcpyAsync(c->eval_T, &T, sizeof T, hipMemcpyHostToDevice, c->stream);
  launch_motion_rows(c->stream, motion_launch(c, k, c->eval_T, c->eval_maps, vb[0], c->eval_codes, nullptr));
  std::vector<int8_t> t(2 * c->N0);
  hipMemcpyAsync(t.data(), c->eval_codes, t.size(), hipMemcpyDeviceToHost, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  for (size_t i = 0; i < t.size(); i++) codes[i] = t[i];
  return 0;
}

long mf_get_trace(mf_ctx* c, double* out, long cap) {
  int n = 0;
  hipMemcpyAsync(&n, c->trace_n, sizeof n, hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  n = std::min(n, c->trace_cap);
  std::vector<TraceRec> r(std::max(n, 1));
  hipMemcpy(r.data(), c->trace, n * sizeof(TraceRec), hipMemcpyDeviceToHost);
  for (long i = 0; i < n && i < cap; i++) {
    double* d = out + 25 * i;
    d[0] = r[i].inst; d[1] = r[i].level; d[2] = r[i].iter; d[3] = r[i].err; d[4] = r[i].s28; d[5] = r[i].flag;
    for (int j = 0; j < 16; j++) d[6 + j] = r[i].pose[j];
    d[22] = r[i].t_accum * 0.01; d[23] = r[i].t_solve * 0.01;   // microseconds
    d[24] = r[i].solver;
  }
  return n;
}
void mf_clear_trace(mf_ctx* c) {
  hipMemsetAsync(c->trace_n, 0, sizeof(int), c->stream);
  hipStreamSynchronize(c->stream);
  c->trace_on = true;
}

long mf_get_level(mf_ctx* c, int level, int which, float* out) {
  if (level < 0 || level >= c->cfg.n_levels) return fail(c, MF_E_ARG, "bad level");
  // after tracking the live pyramid has been handed over to the reference slot
  const Level& L = c->ref[level];
  const size_t n = (size_t)L.w * L.h;
  hipStreamSynchronize(c->stream);
  switch (which) {
    case 0: hipMemcpy(out, L.D, n * 4, hipMemcpyDeviceToHost); break;
    case 1: hipMemcpy(out, L.I, n * 4, hipMemcpyDeviceToHost); break;
    case 2: download_planes(c, L.V, out, n); break;
    case 3: download_planes(c, L.N, out, n); break;
    case 4: hipMemcpy(out, L.gx, n * 4, hipMemcpyDeviceToHost); break;
    case 5: hipMemcpy(out, L.gy, n * 4, hipMemcpyDeviceToHost); break;
    case 6: download_planes(c, L.cov, out, n); break;
    case 7: hipMemcpy(out, c->objs[0]->rD[level], n * 4, hipMemcpyDeviceToHost); break;   // camera's rendered reference depth
    default: return fail(c, MF_E_ARG, "bad buffer");
  }
  return (long)n;
}

int mf_get_frame(mf_ctx* c, float* depth, float* grey, float* rgb) {
  hipStreamSynchronize(c->stream);
  if (depth) hipMemcpy(depth, c->depth, c->N0 * 4, hipMemcpyDeviceToHost);
  if (grey) hipMemcpy(grey, c->grey, c->N0 * 4, hipMemcpyDeviceToHost);
  if (rgb) download_planes(c, c->rgb, rgb, c->N0);
  return 0;
}

int mf_debug_rows(mf_ctx* c, int on) {
  if (on && !c->dbg_rows) {
    c->dbg_rows = dalloc<TrackRow>(c, (size_t)2 * c->cfg.n_levels * 2 * c->N0);
    if (!c->dbg_rows) return fail(c, MF_E_OOM, "row dump");
  }
  if (!on) c->dbg_rows = nullptr;   // buffer stays owned by the context
  return 0;
}
int mf_get_rows(mf_ctx* c, int state, int level, int32_t* result, float* error, float* J) {
  if (!c->dbg_rows || state < 0 || state > 1 || level < 0 || level >= c->cfg.n_levels)
    return fail(c, MF_E_ARG, "row dump not enabled / bad index");
  const size_t n = 2 * c->N0;
  std::vector<TrackRow> r(n);
  hipStreamSynchronize(c->stream);
  hipMemcpy(r.data(), c->dbg_rows + ((size_t)state * c->cfg.n_levels + level) * n, n * sizeof(TrackRow),
            hipMemcpyDeviceToHost);
  for (size_t i = 0; i < n; i++) {
    result[i] = r[i].result; error[i] = r[i].error;
    for (int q = 0; q < 6; q++) J[6 * i + q] = r[i].J[q];
  }
  return 0;
}

int mf_enable_timing(mf_ctx* c, int on) {
  drain_timing(c);
  c->timing = on != 0;
  for (int i = 0; i < KC_COUNT; i++) { c->acc_ms[i] = 0; c->acc_n[i] = 0; }
  return 0;
}
int mf_get_timing(mf_ctx* c, double* out, int cap) {
  drain_timing(c);
  for (int i = 0; i < KC_COUNT && i < cap; i++) { out[2 * i] = c->acc_ms[i]; out[2 * i + 1] = (double)c->acc_n[i]; }
  return KC_COUNT;
}
int mf_get_stats(mf_ctx* c, uint64_t* out, int cap) {
  std::vector<unsigned long long> v((size_t)N_STATS * STAT_SLOTS);
  hipMemcpyAsync(v.data(), c->stats, v.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  for (int i = 0; i < N_STATS && i < cap; i++) {
    uint64_t t = 0;
    for (int k = 0; k < STAT_SLOTS; k++) t += v[(size_t)i * STAT_SLOTS + k];
    out[i] = t;
  }
  return N_STATS;
}
int mf_reset_stats(mf_ctx* c) {
  hipMemsetAsync(c->stats, 0, (size_t)N_STATS * STAT_SLOTS * sizeof(unsigned long long), c->stream);
  return 0;
}
const char* mf_kernel_name(int i) { return (i >= 0 && i < KC_COUNT) ? kNames[i] : ""; }
long mf_device_bytes(mf_ctx* c) { return c->bytes; }

int mf_set_tuning(mf_ctx* c, int knob, int value) {
  if (!c) return MF_E_ARG;
  if (knob >= MF_TUNE_OBJ_ROWS_PER_THREAD_L0 && knob < MF_TUNE_OBJ_ROWS_PER_THREAD_L0 + MAX_LEVELS) {
    if (value < 0 || value > 64) return fail(c, MF_E_ARG, "object rows per thread must be in [0, 64]");
    c->track_obj_rpt[knob - MF_TUNE_OBJ_ROWS_PER_THREAD_L0] = value;
    return 0;
  }
  if (knob >= MF_TUNE_OBJ_PIXELS_PER_THREAD_L0 && knob < MF_TUNE_OBJ_PIXELS_PER_THREAD_L0 + MAX_LEVELS) {
    if (value < 0 || value > 2) return fail(c, MF_E_ARG, "object pixels per thread must be 0, 1 or 2");
    c->track_obj_ppt[knob - MF_TUNE_OBJ_PIXELS_PER_THREAD_L0] = value;
    return 0;
  }
  if (knob >= MF_TUNE_TRACK_FUSED_L0 && knob < MF_TUNE_TRACK_FUSED_L0 + MAX_LEVELS) {
    c->track_fused[knob - MF_TUNE_TRACK_FUSED_L0] = value < 0 ? -1 : (value ? 1 : 0);
    return 0;
  }
  if (knob >= MF_TUNE_TRACK_ROWS_PER_THREAD_L0 && knob < MF_TUNE_TRACK_ROWS_PER_THREAD_L0 + MAX_LEVELS) {
    if (value < 1 || value > 64) return fail(c, MF_E_ARG, "rows per thread must be in [1, 64]");
    c->track_rpt[knob - MF_TUNE_TRACK_ROWS_PER_THREAD_L0] = value;
    return 0;
  }
  return fail(c, MF_E_ARG, "unknown tuning knob");
}
int mf_get_tuning(mf_ctx* c, int knob, int* value) {
  if (!c || !value) return MF_E_ARG;
  if (knob >= MF_TUNE_OBJ_ROWS_PER_THREAD_L0 && knob < MF_TUNE_OBJ_ROWS_PER_THREAD_L0 + MAX_LEVELS) {
    *value = c->track_obj_rpt[knob - MF_TUNE_OBJ_ROWS_PER_THREAD_L0];
    return 0;
  }
  if (knob >= MF_TUNE_OBJ_PIXELS_PER_THREAD_L0 && knob < MF_TUNE_OBJ_PIXELS_PER_THREAD_L0 + MAX_LEVELS) {
    *value = c->track_obj_ppt[knob - MF_TUNE_OBJ_PIXELS_PER_THREAD_L0];
    return 0;
  }
  if (knob >= MF_TUNE_TRACK_FUSED_L0 && knob < MF_TUNE_TRACK_FUSED_L0 + MAX_LEVELS) {
    *value = c->track_fused[knob - MF_TUNE_TRACK_FUSED_L0];
    return 0;
  }
  if (knob >= MF_TUNE_TRACK_ROWS_PER_THREAD_L0 && knob < MF_TUNE_TRACK_ROWS_PER_THREAD_L0 + MAX_LEVELS) {
    *value = c->track_rpt[knob - MF_TUNE_TRACK_ROWS_PER_THREAD_L0];
    return 0;
  }
  return fail(c, MF_E_ARG, "unknown tuning knob");
}

// Plain device-to-device copy kernel: the reference point for scaling HBM
// traffic counters (FETCH_SIZE / WRITE_SIZE) to bytes.
__global__ void k_copy_ref(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}
// Forward/backward of the tracking residual at a given pose (see the header).
int mf_residual_eval(mf_ctx* c, int id, const float k[4], int level, const float T16[16], double out[32]) {
  if (!c->pending_objects) return fail(c, MF_E_ARG, "call between mf_track_camera and mf_track_objects");
  if (level < 0 || level >= c->cfg.n_levels) return fail(c, MF_E_ARG, "bad level");
  if (id == 0 || id >= (int)c->objs.size()) return fail(c, MF_E_ARG, "bad instance id");
  hipStream_t s = c->stream;
  const int object = id > 0 ? 1 : 0;
  int slot = 0, state = 0;
  if (object) {
    // a spare descriptor / state slot: the object's GN constants at the
    // frame's start (T_c1_o1 etc.), without publishing a pose
    Object* o = c->objs[id];
    slot = state = MAX_INST - 1;
    set_inst(c, slot, o->mv, o->mn, o->rD, nullptr, nullptr, state);
    const int oid = 3 + id;
    hipMemcpyAsync(c->obj_ids + MAX_INST - 1, &oid, sizeof oid, hipMemcpyHostToDevice, s);
    hipStreamSynchronize(s);
    launch_gn_begin_objects(s, c->gn, state, 1, c->obj_ids + MAX_INST - 1, c->mats, c->mats + 0, c->mats + 1, 0,
                            nullptr);
    if (c->cfg.use_rgb) render_reference(c, slot, 1);
  }
  TrackLaunch p = base_launch(c, k, level, object);
  p.iter = 0;
  set_shape(p, c->track_rpt[level]);
  const int kinds = (c->cfg.use_icp ? 1 : 0) + (c->cfg.use_rgb ? 1 : 0);
  const int nblk = track_blocks(p.w * p.h * kinds, c->track_rpt[level]);
  launch_eval(s, p, c->inst + slot, c->gn, state, m4(T16), c->partial, nblk);
  hipMemcpyAsync(out, &c->gn[state].sums[0], 32 * sizeof(double), hipMemcpyDeviceToHost, s);
  if (hipStreamSynchronize(s) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  return 0;
}

int mf_solve6(mf_ctx* c, const double* s27, long n, double* x, int32_t* solver) {
  if (n < 0 || (n > 0 && (!s27 || !x || !solver))) return fail(c, MF_E_ARG, "bad solve6 arguments");
  if (n == 0) return 0;
  if (n > (1 << 20)) return fail(c, MF_E_CAPACITY, "at most 2^20 systems per call");
  hipStream_t s = c->stream;
  double *d_s = nullptr, *d_x = nullptr;
  int* d_v = nullptr;
  if (hipMalloc(&d_s, n * 27 * sizeof(double)) != hipSuccess || hipMalloc(&d_x, n * 6 * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_v, n * sizeof(int)) != hipSuccess) {
    hipFree(d_s); hipFree(d_x); hipFree(d_v);
    return fail(c, MF_E_OOM, "solve6 buffers");
  }
  hipMemcpyAsync(d_s, s27, n * 27 * sizeof(double), hipMemcpyHostToDevice, s);
  launch_solve6(s, d_s, (int)n, d_x, d_v);
  hipMemcpyAsync(x, d_x, n * 6 * sizeof(double), hipMemcpyDeviceToHost, s);
  hipMemcpyAsync(solver, d_v, n * sizeof(int), hipMemcpyDeviceToHost, s);
  const hipError_t e = hipStreamSynchronize(s);
  hipFree(d_s); hipFree(d_x); hipFree(d_v);
  if (e != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(e));
  return 0;
}

// Replays the camera's row kernel of one pyramid level `reps` times back to
// back (iteration 0 at the current camera state), bracketed by two events on
// the context stream, and returns the average launch time and the
// launch's algorithmic bytes.  Overwrites the camera's level-0 track codes and
// the row partials; meant for measurement between frames.
// Scratch GN state and descriptor in the spare slot (MAX_INST - 1, shared
// with the residual evaluation): a copy of the camera's state, switched on
// (the last frame's re-track may have left the camera's state off), whose
// pose output goes to tune_pose; the descriptor takes the camera's inputs,
// and its code and error outputs when `outputs` is set.
static int scratch_camera(mf_ctx* c, bool outputs) {
  hipStream_t s = c->stream;
  if (!c->tune_pose && !(c->tune_pose = dalloc<Mat4>(c, 1))) return fail(c, MF_E_OOM, "device allocation failed");
  const int spare = MAX_INST - 1;
  GNState* sg = c->gn + spare;
  hipMemcpyAsync(sg, c->gn, sizeof(GNState), hipMemcpyDeviceToDevice, s);
  Mat4* tp = c->tune_pose;
  const int zero = 0;
  hipMemcpyAsync(&sg->out_pose, &tp, sizeof tp, hipMemcpyHostToDevice, s);
  hipMemcpyAsync(&sg->off, &zero, sizeof zero, hipMemcpyHostToDevice, s);
  if (hipStreamSynchronize(s) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  Object* bg = c->objs[0];
  set_inst(c, spare, c->vtx, c->nrm, bg->rD, outputs ? bg->codes : nullptr, outputs ? bg->rgb_err : nullptr, spare);
  return spare;
}

int mf_replay_track_rows(mf_ctx* c, const float k[4], int level, int reps, double* avg_us, double* alg_bytes) {
  if (level < 0 || level >= c->cfg.n_levels || reps < 1) return fail(c, MF_E_ARG, "bad level or repetition count");
  if (!c->have_ref) return fail(c, MF_E_ARG, "no tracked frame yet");
  // the camera's launch on a scratch copy of its state, with the camera's
  // code and error outputs (the same stores as in a frame)
  const int spare = scratch_camera(c, true);
  if (spare < 0) return spare;
  TrackLaunch p = base_launch(c, k, level, 0);
  p.iter = 0;
  set_shape(p, c->track_rpt[level]);
  // the reference pyramid is the last tracked frame, the live one the frame
  // before it: after a frame's hand-over both hold valid planes
  const int kinds = (c->cfg.use_icp ? 1 : 0) + (c->cfg.use_rgb ? 1 : 0);
  const int nblk = track_blocks(p.w * p.h * kinds, c->track_rpt[level]);
  hipEvent_t a = ev_get(c), b = ev_get(c);
  launch_track_rows(c->stream, p, c->inst + spare, 1, c->gn, c->partial, nblk, 0);   // warm
  hipEventRecord(a, c->stream);
  for (int r = 0; r < reps; r++) launch_track_rows(c->stream, p, c->inst + spare, 1, c->gn, c->partial, nblk, 0);
  hipEventRecord(b, c->stream);
  if (hipEventSynchronize(b) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
  float ms = 0;
  hipEventElapsedTime(&ms, a, b);
  c->pool.push_back(a); c->pool.push_back(b);
  if (avg_us) *avg_us = 1000.0 * ms / reps;
  // live v, n, cov + model v, n + mask (61 B) and the photometric inputs
  // (41 B) per pixel, one 128-B partial per workgroup
  if (alg_bytes) *alg_bytes = (double)p.w * p.h * ((c->cfg.use_icp ? 61 : 0) + (c->cfg.use_rgb ? 41 : 0)) + 128.0 * nblk;
  return 0;
}

int mf_autotune(mf_ctx* c, const float k[4], int reps, int32_t* chosen, double* us) {
  static const int cand[MF_TUNE_CANDIDATES] = {1, 2, 3, 4, 8, 16};
  if (!k || reps < 1 || reps > 10000) return fail(c, MF_E_ARG, "bad arguments");
  if (!c->have_ref) return fail(c, MF_E_ARG, "no tracked frame yet");
  hipStream_t s = c->stream;
  const int spare = scratch_camera(c, false);
  if (spare < 0) return spare;
  TrackSync sy;
  memset(&sy, 0, sizeof sy);
  const int kinds = (c->cfg.use_icp ? 1 : 0) + (c->cfg.use_rgb ? 1 : 0);
  hipEvent_t a = ev_get(c), b = ev_get(c);
  int q = 0;
  for (int l = 0; l < c->cfg.n_levels; l++) {
    TrackLaunch p = base_launch(c, k, l, 0);
    p.iter = 0;          // rows and solves of iteration 0 never skip
    p.dbg = nullptr;
    // each candidate's time is the best of three bursts; the current setting
    // is replaced only by one more than 1 % faster (near-equal candidates
    // would otherwise flip between calls)
    double t_cur = -1;
    double tc[MF_TUNE_CANDIDATES];
    for (int ci = 0; ci < MF_TUNE_CANDIDATES; ci++) {
      const int nblk = track_blocks(p.w * p.h * kinds, cand[ci]);
      set_shape(p, cand[ci]);
      auto iter = [&]() {
        launch_track_rows(s, p, c->inst + spare, 1, c->gn, c->partial, nblk, q);
        launch_track_solve(s, p, c->inst + spare, 1, c->gn, c->partial, nblk, q, sy);
        ++q;
      };
      iter();   // warm
      double t = -1;
      for (int burst = 0; burst < 3; burst++) {
        hipEventRecord(a, s);
        for (int r = 0; r < reps; r++) iter();
        hipEventRecord(b, s);
        if (hipEventSynchronize(b) != hipSuccess) return fail(c, MF_E_HIP, hipGetErrorString(hipGetLastError()));
        float ms = 0;
        hipEventElapsedTime(&ms, a, b);
        const double tb = 1000.0 * ms / reps;
        if (t < 0 || tb < t) t = tb;
      }
      tc[ci] = t;
      if (us) us[l * MF_TUNE_CANDIDATES + ci] = t;
      if (cand[ci] == c->track_rpt[l]) t_cur = t;
    }
    int arg = c->track_rpt[l];
    double best = t_cur;
    for (int ci = 0; ci < MF_TUNE_CANDIDATES; ci++)
      if (best < 0 || tc[ci] < 0.99 * best) { best = tc[ci]; arg = cand[ci]; }
    c->track_rpt[l] = arg;
    if (chosen) chosen[l] = arg;
  }
  c->pool.push_back(a); c->pool.push_back(b);
  return 0;
}

int mf_copy_reference(mf_ctx* c, size_t bytes, int reps) {
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    if (a) hipFree(a);
    return fail(c, MF_E_OOM, "copy buffers");
  }
  hipMemsetAsync(a, 1, bytes, c->stream);
  for (int r = 0; r < reps; r++) k_copy_ref<<<4096, 256, 0, c->stream>>>((const float4*)a, (float4*)b, bytes / 16);
  hipStreamSynchronize(c->stream);
  hipFree(a); hipFree(b);
  return 0;
}

}  // extern "C"
