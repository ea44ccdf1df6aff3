// On-disk formats on either side of the dense path: SLAMBench .raw frames,
// TUM trajectories and TUM ground truth (host code, no device work).
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/midfusion.h"
#include "mf_io.h"

namespace mf {

bool raw_open(RawFile& r, const char* path) {
  r.f = fopen(path, "rb");
  if (!r.f) return false;
  uint32_t sz[2];
  if (fread(sz, sizeof sz, 1, r.f) != 1 || sz[0] == 0 || sz[1] == 0) {
    fclose(r.f);
    r.f = nullptr;
    return false;
  }
  r.w = sz[0];
  r.h = sz[1];
  const long n = (long)r.w * r.h;
  r.frame_bytes = 4 * (long)sizeof(uint32_t) + n * (long)sizeof(uint16_t) + n * 3;
  fseeko(r.f, 0, SEEK_END);
  const off_t len = ftello(r.f);
  r.frames = (long)(len / r.frame_bytes);
  fseeko(r.f, 0, SEEK_SET);
  return true;
}

int raw_read(RawFile& r, long frame, uint16_t* depth, uint8_t* rgb) {
  if (!r.f || frame < 0 || frame >= r.frames) return -1;
  const long n = (long)r.w * r.h;
  if (fseeko(r.f, (off_t)frame * r.frame_bytes, SEEK_SET) != 0) return -1;
  uint32_t sz[2];
  if (fread(sz, sizeof sz, 1, r.f) != 1 || sz[0] != r.w || sz[1] != r.h) return -1;
  if (depth) {
    if (fread(depth, sizeof(uint16_t), n, r.f) != (size_t)n) return -1;
  } else if (fseeko(r.f, n * (off_t)sizeof(uint16_t), SEEK_CUR) != 0) {
    return -1;
  }
  if (fread(sz, sizeof sz, 1, r.f) != 1 || sz[0] != r.w || sz[1] != r.h) return -1;
  if (rgb && fread(rgb, 3, n, r.f) != (size_t)n) return -1;
  return 0;
}

void raw_close(RawFile& r) {
  if (r.f) fclose(r.f);
  r.f = nullptr;
}

void rot_quat(const float T[16], float q[4]) {
  auto m = [&](int i, int j) { return T[4 * i + j]; };
  float t = m(0, 0) + m(1, 1) + m(2, 2);
  float x, y, z, w;
  if (t > 0.0f) {
    t = std::sqrt(t + 1.0f);
    w = 0.5f * t;
    t = 0.5f / t;
    x = (m(2, 1) - m(1, 2)) * t;
    y = (m(0, 2) - m(2, 0)) * t;
    z = (m(1, 0) - m(0, 1)) * t;
  } else {
    int i = 0;
    if (m(1, 1) > m(0, 0)) i = 1;
    if (m(2, 2) > m(i, i)) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    float v[3];
    t = std::sqrt(m(i, i) - m(j, j) - m(k, k) + 1.0f);
    v[i] = 0.5f * t;
    t = 0.5f / t;
    w = (m(k, j) - m(j, k)) * t;
    v[j] = (m(j, i) + m(i, j)) * t;
    v[k] = (m(k, i) + m(i, k)) * t;
    x = v[0]; y = v[1]; z = v[2];
  }
  q[0] = x; q[1] = y; q[2] = z; q[3] = w;
}

int tum_line(char* out, int cap, int frame, const float T[16]) {
  float q[4];
  rot_quat(T, q);
  const int n = snprintf(out, (size_t)cap, "%d\t%g\t%g\t%g\t%g\t%g\t%g\t%g\t\n", frame, (double)T[3], (double)T[7],
                         (double)T[11], (double)q[0], (double)q[1], (double)q[2], (double)q[3]);
  return (n < 0 || n >= cap) ? -1 : n;
}

void gt_pose(const float quat[4], const float t[3], float T[16]) {
  // toMatrix4 (utils/math_utils.h:137-146): normalise to |q|^2 = 2, outer
  // product; the diagonal terms are evaluated in double (1.0 - a - b) and
  // rounded once
  const float n = quat[0] * quat[0] + quat[1] * quat[1] + quat[2] * quat[2] + quat[3] * quat[3];
  const float s = sqrtf((float)(2.0 / (double)n));
  const float a[4] = {quat[0] * s, quat[1] * s, quat[2] * s, quat[3] * s};
  float M[4][4];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) M[i][j] = a[i] * a[j];
  const float R[16] = {(float)(1.0 - (double)M[1][1] - (double)M[2][2]), M[0][1] - M[2][3], M[0][2] + M[1][3], t[0],
                       M[0][1] + M[2][3], (float)(1.0 - (double)M[0][0] - (double)M[2][2]), M[1][2] - M[0][3], t[1],
                       M[0][2] - M[1][3], M[1][2] + M[0][3], (float)(1.0 - (double)M[0][0] - (double)M[1][1]), t[2],
                       0.f, 0.f, 0.f, 1.f};
  memcpy(T, R, sizeof R);
}

}  // namespace mf

struct mf_raw {
  mf::RawFile r;
};

extern "C" {

mf_raw* mf_raw_open(const char* path) {
  if (!path) return nullptr;
  mf_raw* h = new mf_raw();
  if (!mf::raw_open(h->r, path)) {
    delete h;
    return nullptr;
  }
  return h;
}

int mf_raw_info(mf_raw* h, int* width, int* height, long* frames, float k[4]) {
  if (!h) return MF_E_ARG;
  if (width) *width = (int)h->r.w;
  if (height) *height = (int)h->r.h;
  if (frames) *frames = h->r.frames;
  if (k) {   // RawDepthReader::getK (interface.h:327-329)
    k[0] = 531.15f; k[1] = 531.15f; k[2] = 640 / 2; k[3] = 480 / 2;
  }
  return MF_OK;
}

int mf_raw_read(mf_raw* h, long frame, uint16_t* depth, uint8_t* rgb) {
  if (!h) return MF_E_ARG;
  return mf::raw_read(h->r, frame, depth, rgb) == 0 ? MF_OK : MF_E_EOF;
}

void mf_raw_close(mf_raw* h) {
  if (!h) return;
  mf::raw_close(h->r);
  delete h;
}

long mf_parse_gt(const char* path, float* poses, long cap) {
  std::ifstream file(path ? path : "");
  if (!file.is_open()) return MF_E_ARG;
  long n = 0;
  std::string line;
  while (std::getline(file, line)) {
    // split(line, ' ') (vtk-io.h:42-48): every single space separates a field
    std::vector<std::string> f;
    std::stringstream ss(line);
    std::string item;
    while (std::getline(ss, item, ' ')) f.push_back(item);
    if (f.size() != 8) continue;
    float v[7];
    try {
      for (int i = 0; i < 7; i++) v[i] = std::stof(f[1 + i]);
    } catch (...) {
      return MF_E_ARG;   // std::stof throws where the reference would abort
    }
    if (poses && n < cap) {
      const float t[3] = {v[0], v[1], v[2]}, q[4] = {v[3], v[4], v[5], v[6]};
      mf::gt_pose(q, t, poses + 16 * n);
    }
    n++;
  }
  return n;
}

int mf_format_tum(const float T[16], int frame, char* out, int cap) {
  if (!T || !out || cap <= 0) return MF_E_ARG;
  const int n = mf::tum_line(out, cap, frame, T);
  return n < 0 ? MF_E_ARG : n;
}

}  // extern "C"
