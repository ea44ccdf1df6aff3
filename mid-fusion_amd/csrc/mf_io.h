// On-disk formats on either side of the dense path (host code).
//
// * SLAMBench .raw sequences, read as RawDepthReader does
//   (apps/kfusion/include/interface.h:207-333): every frame is
//   {uint32 w, uint32 h, uint16 depth[w*h], uint32 w, uint32 h,
//   uint8 rgb[3*w*h]} and frame i starts at i times the size of frame 0.
// * TUM trajectories as Kfusion::save_poses writes them
//   (apps/kfusion/src/cpp/kfusion.cpp:1591-1626), with the rotation's
//   quaternion from getQuaternion (preprocessing.cpp:312-320).
// * TUM ground truth as parseGTFile reads it (include/vtk-io.h:304-330),
//   poses from toMatrix4(quat, trans) (utils/math_utils.h:137-146).
#pragma once
#include <cstdint>
#include <cstdio>

namespace mf {

struct RawFile {
  FILE* f = nullptr;
  uint32_t w = 0, h = 0;    // size of frame 0
  long frame_bytes = 0;     // 16 + 5*w*h
  long frames = 0;          // whole frames in the file
};
// Opens a .raw sequence; false if the first frame header cannot be read.
bool raw_open(RawFile& r, const char* path);
// Frame `frame` into depth (w*h uint16) and rgb (3*w*h bytes); either may be
// null to skip it.  0 on success, -1 at or past the end of the file or on a
// frame whose header does not repeat frame 0's size.
int raw_read(RawFile& r, long frame, uint16_t* depth, uint8_t* rgb);
void raw_close(RawFile& r);

// Quaternion (x, y, z, w) of the rotation block of a row-major 4x4 pose, by
// the trace / largest-diagonal method of Shepperd (J. Guidance and Control
// 1(3), 1978) in float arithmetic.
void rot_quat(const float T[16], float q[4]);
// One trajectory line "frame\ttx\tty\ttz\tqx\tqy\tqz\tqw\t\n" with the
// iostream default formatting of floats (6 significant digits).  Returns the
// length written (excluding the terminating zero), or -1 if cap is too small.
int tum_line(char* out, int cap, int frame, const float T[16]);
// Pose of a ground-truth line: quaternion (x, y, z, w) and translation.
void gt_pose(const float quat[4], const float t[3], float T[16]);

}  // namespace mf
