// Test driver of the C++ Function interface (include/midfusion_function.hpp):
// Gauss-Newton to convergence over the GPU forward/backward of the tracking
// residual.
//
//   function_gn <dir>
//
// <dir>/params.txt: W H  fx fy cx cy  bg_size bg_res  huber_delta  n_obj
//                   init_pose[16]  then per object: size res T_w_o[16]
//                   then the GN start poses: camera T_w_l[16], per object
//                   T_c2_o2[16]; then max_iters, iterations at levels 0 1 2
//                   (0: max_iters), step_tol, loss_rtol
// <dir>/frames.raw: frames 0 and 1 (SLAMBench .raw layout)
// <dir>/masks.bin:  float32 integration masks [frame][instance][H*W]
//
// Frame 0 seeds the tracker and is integrated and raycast; frame 1 is
// preprocessed and its camera phase run, then each instance's pose is
// solved from its start pose by midfusion::gauss_newton.  Prints one line per
// instance: "<id> <iterations L0 L1 L2> <stop reasons L0 L1 L2> <loss L0>
// <pose[16]>".
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/midfusion_function.hpp"

static int die(mf_ctx* c, const char* what) {
  std::fprintf(stderr, "%s: %s\n", what, c ? mf_last_error(c) : "");
  return 1;
}

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: function_gn <dir>\n"); return 2; }
  const std::string dir = argv[1];
  FILE* f = std::fopen((dir + "/params.txt").c_str(), "r");
  if (!f) return die(nullptr, "params.txt");
  int W, H, bg_res, n_obj;
  float k[4], bg_size, init[16];
  double delta;
  if (std::fscanf(f, "%d %d %f %f %f %f %f %d %lf %d", &W, &H, &k[0], &k[1], &k[2], &k[3], &bg_size, &bg_res, &delta,
                  &n_obj) != 10)
    return die(nullptr, "params.txt header");
  for (float& v : init) if (std::fscanf(f, "%f", &v) != 1) return die(nullptr, "init pose");
  std::vector<float> size(n_obj), pose(16 * n_obj), start(16 * (1 + n_obj));
  std::vector<int> res(n_obj);
  for (int j = 0; j < n_obj; j++) {
    if (std::fscanf(f, "%f %d", &size[j], &res[j]) != 2) return die(nullptr, "object");
    for (int i = 0; i < 16; i++) if (std::fscanf(f, "%f", &pose[16 * j + i]) != 1) return die(nullptr, "object pose");
  }
  for (float& v : start) if (std::fscanf(f, "%f", &v) != 1) return die(nullptr, "start pose");
  int max_iters, level_iters[3];
  double step_tol, loss_rtol;
  if (std::fscanf(f, "%d %d %d %d %lf %lf", &max_iters, &level_iters[0], &level_iters[1], &level_iters[2], &step_tol,
                  &loss_rtol) != 6)
    return die(nullptr, "options");
  std::fclose(f);

  mf_config cfg = {};
  cfg.width = W; cfg.height = H; cfg.n_levels = 3;
  cfg.iters[0] = 10; cfg.iters[1] = 5; cfg.iters[2] = 4;
  cfg.use_icp = 1; cfg.use_rgb = 1;
  cfg.robust = MF_ROBUST_HUBER; cfg.robust_k = delta;
  cfg.mu = 0.1f; cfg.track_rate = 1; cfg.integ_rate = 1; cfg.init_frame = 0; cfg.batch_objects = 1;
  mf_ctx* c = mf_create(&cfg, init, bg_size, bg_res, 0);
  if (!c) return die(nullptr, "mf_create");
  for (int j = 0; j < n_obj; j++) {
    if (mf_object_create(c, size[j], res[j], &pose[16 * j], 0) != j + 1) return die(c, "mf_object_create");
    mf_object_set_moving(c, j + 1, 1);
  }
  mf_raw* raw = mf_raw_open((dir + "/frames.raw").c_str());
  if (!raw) return die(c, "frames.raw");
  std::vector<float> masks((size_t)2 * (1 + n_obj) * W * H);
  FILE* fm = std::fopen((dir + "/masks.bin").c_str(), "rb");
  if (!fm || std::fread(masks.data(), sizeof(float), masks.size(), fm) != masks.size()) return die(c, "masks.bin");
  std::fclose(fm);
  for (unsigned frame = 0; frame < 2; frame++) {
    for (int i = 0; i <= n_obj; i++)
      if (mf_object_set_mask(c, i, &masks[((size_t)frame * (1 + n_obj) + i) * W * H]) != 0) return die(c, "mask");
    if (mf_preprocess_raw(c, raw, frame, 0) != 0) return die(c, "mf_preprocess_raw");
    if (frame == 0) {
      if (mf_track(c, k, frame) < 0 || mf_integrate(c, k, frame) < 0 || mf_raycast(c, k, frame) < 0)
        return die(c, "frame 0");
    } else if (mf_track_camera(c, k, frame) < 0) {
      return die(c, "mf_track_camera");
    }
  }
  for (int i = 0; i <= n_obj; i++) {
    float T[16];
    for (int q = 0; q < 16; q++) T[q] = start[16 * i + q];
    midfusion::GaussNewtonOptions opt;
    opt.max_iters = max_iters; opt.step_tol = step_tol; opt.loss_rtol = loss_rtol;
    for (int l = 0; l < 3; l++) opt.level_iters[l] = level_iters[l];
    midfusion::GaussNewtonReport rep;
    if (midfusion::gauss_newton(c, i == 0 ? -1 : i, k, 3, T, opt, &rep) != 0) return die(c, "gauss_newton");
    std::printf("%d %d %d %d %d %d %d %.12g", i == 0 ? -1 : i, rep.iterations[0], rep.iterations[1], rep.iterations[2],
                rep.stop[0], rep.stop[1], rep.stop[2], rep.loss[0]);
    for (float v : T) std::printf(" %.9g", v);
    std::printf("\n");
  }
  mf_raw_close(raw);
  mf_destroy(c);
  return 0;
}
