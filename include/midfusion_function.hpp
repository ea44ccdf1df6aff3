/*
 * midfusion_function.hpp — forward/backward Function interface to the
 * ICP + photometric tracking residual of one instance (camera or moving
 * object), for C++ callers that differentiate through the tracking loss or
 * run their own optimiser over it.  Header-only, over the C ABI of
 * midfusion.h: the residual rows, their robust weights and the reduction run
 * on the GPU (mf_residual_eval); this header holds the host-side algebra.
 *
 * The loss is the reference's robust tracking loss L(T) = sum rho(e_i) over
 * the ICP and photometric rows of one pyramid level (Tracking::reduceKernel,
 * tracking.cpp:1191-1352, rows tracking.cpp:392-578, 923-1111), at the pose
 * T: T_w_l for the camera, T_c2_o2 for object id > 0.  The rows' J is minus
 * the derivative of the residual for a left perturbation T <- exp(xi) T,
 * xi = (v, w) (translation first, as TooN's SE3::exp, tracking.cpp:1361), so
 *   forward:  L(T)
 *   backward: dL/dxi = -J^T W e,  and the Gauss-Newton Hessian J^T W J,
 * valid wherever the rows' correspondences are locally fixed.
 */
#ifndef MIDFUSION_FUNCTION_HPP
#define MIDFUSION_FUNCTION_HPP

#include <cmath>
#include <cstring>

#include "midfusion.h"

namespace midfusion {

// Row-major 4x4 of exp(xi) for xi = (v, w): Rodrigues rotation and the left
// Jacobian V applied to v (the SE3 exponential the tracking step uses).
inline void se3_exp(const double xi[6], double T[16]) {
  const double w0 = xi[3], w1 = xi[4], w2 = xi[5];
  const double th2 = w0 * w0 + w1 * w1 + w2 * w2;
  double A, B, C;
  if (th2 < 1e-8) {
    A = 1.0 - th2 / 6.0; B = 0.5 - th2 / 24.0; C = 1.0 / 6.0 - th2 / 120.0;
  } else {
    const double th = std::sqrt(th2);
    A = std::sin(th) / th; B = (1.0 - std::cos(th)) / th2; C = (1.0 - A) / th2;
  }
  const double W[3][3] = {{0, -w2, w1}, {w2, 0, -w0}, {-w1, w0, 0}};
  double W2[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0;
      for (int k = 0; k < 3; k++) s += W[i][k] * W[k][j];
      W2[i][j] = s;
    }
  for (int i = 0; i < 3; i++) {
    double t = xi[i];
    for (int j = 0; j < 3; j++) {
      T[4 * i + j] = (i == j ? 1.0 : 0.0) + A * W[i][j] + B * W2[i][j];
      t += (B * W[i][j] + C * W2[i][j]) * xi[j];
    }
    T[4 * i + 3] = t;
  }
  T[12] = T[13] = T[14] = 0.0;
  T[15] = 1.0;
}

// x = H^-1 b for the symmetric 6x6 H by Cholesky; false if H is not
// positive definite.
inline bool solve_spd6(const double H[36], const double b[6], double x[6]) {
  double L[36] = {};
  for (int i = 0; i < 6; i++)
    for (int j = 0; j <= i; j++) {
      double s = H[6 * i + j];
      for (int k = 0; k < j; k++) s -= L[6 * i + k] * L[6 * j + k];
      if (i == j) {
        if (!(s > 0.0)) return false;
        L[6 * i + i] = std::sqrt(s);
      } else {
        L[6 * i + j] = s / L[6 * j + j];
      }
    }
  double y[6];
  for (int i = 0; i < 6; i++) {
    double s = b[i];
    for (int k = 0; k < i; k++) s -= L[6 * i + k] * y[k];
    y[i] = s / L[6 * i + i];
  }
  for (int i = 5; i >= 0; i--) {
    double s = y[i];
    for (int k = i + 1; k < 6; k++) s -= L[6 * k + i] * x[k];
    x[i] = s / L[6 * i + i];
  }
  return true;
}

// The tracking loss of one instance at one pyramid level as a Function.
// Valid between mf_track_camera and mf_track_objects of a frame (the live
// and reference images of that frame); the robust function and its constant
// are the context's (mf_config::robust, robust_k).
class ResidualFunction {
 public:
  // id < 0: the camera; id > 0: moving object id.
  ResidualFunction(mf_ctx* ctx, int id, const float k[4], int level) : ctx_(ctx), id_(id), level_(level) {
    std::memcpy(k_, k, sizeof k_);
  }

  // Forward at pose T (row-major 4x4): evaluates the rows and their sums on
  // the GPU; returns 0 or a negative MF_E_* code.  *loss (if given) = L(T).
  int forward(const float T[16], double* loss = nullptr) {
    const int r = mf_residual_eval(ctx_, id_, k_, level_, T, sums_);
    have_ = r == 0;
    if (have_ && loss) *loss = sums_[0];
    return r;
  }
  // Backward of the last forward: dL/dxi for T <- exp(xi) T.
  void backward(double grad[6]) const {
    for (int a = 0; a < 6; a++) grad[a] = -sums_[1 + a];
  }
  // The Gauss-Newton Hessian J^T W J of the last forward (row-major 6x6).
  void hessian(double H[36]) const {
    int c = 7;
    for (int a = 0; a < 6; a++)
      for (int b = a; b < 6; b++) {
        H[6 * a + b] = H[6 * b + a] = sums_[c];
        c++;
      }
  }
  double loss() const { return sums_[0]; }
  // Rows with a valid residual (the reduction's sum 28).
  double inliers() const { return sums_[28]; }
  // The 32 reduction sums of the last forward ([0] loss, [1..6] J^T W e,
  // [7..27] the upper triangle of J^T W J, [28] inliers, [29..31] outliers).
  const double* sums() const { return sums_; }
  bool valid() const { return have_; }
  int level() const { return level_; }

 private:
  mf_ctx* ctx_;
  int id_;
  float k_[4];
  int level_;
  double sums_[32] = {};
  bool have_ = false;
};

struct GaussNewtonOptions {
  int max_iters = 10;      // iterations per pyramid level, unless level_iters[level] > 0
  int level_iters[4] = {0, 0, 0, 0};
  double step_tol = 1e-5;  // converged when |x| < step_tol
  // a step is undone and the level ends when it raises the loss by more than
  // loss_rtol * loss: 0 is the reference's test (any increase); the FP32
  // rows carry a relative loss error near 1e-4, so a loop run to
  // convergence needs a tolerance above that
  double loss_rtol = 0.0;
};

struct GaussNewtonReport {
  int iterations[4] = {0, 0, 0, 0};   // executed iterations per pyramid level
  int stop[4] = {0, 0, 0, 0};         // per level: 0 iteration cap, 1 converged, 2 loss increase, 3 no rows, 4 singular
  double loss[4] = {0, 0, 0, 0};      // final loss per level
  int status = 0;                     // 0, or the MF_E_* code of a failed forward
};

// Coarse-to-fine Gauss-Newton over ResidualFunctions (levels n_levels-1 ..
// 0): at each level, T <- exp(x) T with x = (J^T W J)^-1 J^T W e until
// |x| < step_tol or max_iters; a step that raises the loss (beyond
// loss_rtol) is undone and ends the level, as the reference's step-back
// (tracking.cpp:304-323).  T is updated in place.
inline int gauss_newton(mf_ctx* ctx, int id, const float k[4], int n_levels, float T[16],
                        const GaussNewtonOptions& opt = GaussNewtonOptions(), GaussNewtonReport* rep = nullptr) {
  GaussNewtonReport local;
  GaussNewtonReport& R = rep ? *rep : local;
  for (int level = n_levels - 1; level >= 0; --level) {
    ResidualFunction f(ctx, id, k, level);
    float prev[16];
    double prev_loss = INFINITY;
    int it = 0, stop = 0;
    const int n_iters = opt.level_iters[level] > 0 ? opt.level_iters[level] : opt.max_iters;
    for (; it < n_iters; ++it) {
      double loss = 0;
      const int r = f.forward(T, &loss);
      if (r != 0) { R.status = r; return r; }
      if (!(f.inliers() > 0)) { stop = 3; break; }
      if (loss > prev_loss * (1.0 + opt.loss_rtol)) {   // undo the last step and end the level
        std::memcpy(T, prev, sizeof prev);
        R.loss[level] = prev_loss;
        stop = 2;
        break;
      }
      R.loss[level] = loss;
      double H[36], g[6], x[6];
      f.hessian(H);
      f.backward(g);
      for (int a = 0; a < 6; a++) g[a] = -g[a];   // J^T W e
      if (!solve_spd6(H, g, x)) { stop = 4; break; }
      double E[16], nx = 0;
      for (int a = 0; a < 6; a++) nx += x[a] * x[a];
      nx = std::sqrt(nx);
      se3_exp(x, E);
      std::memcpy(prev, T, sizeof prev);
      prev_loss = loss;
      float Tn[16];
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
          double s = 0;
          for (int q = 0; q < 4; q++) s += E[4 * i + q] * (double)T[4 * q + j];
          Tn[4 * i + j] = (float)s;
        }
      std::memcpy(T, Tn, sizeof Tn);
      if (nx < opt.step_tol) { ++it; stop = 1; break; }
    }
    R.iterations[level] = it;
    R.stop[level] = stop;
  }
  return 0;
}

}  // namespace midfusion

#endif
