// kf_recording.cpp — the reference's recorded-stream format (Parameters.txt +
// raw recorded<N> uint16 frames), read and written on the host.
//
// Restates CameraFromFile::Init / GetDepthFrame (Source/src/CameraSupport/
// CameraFromFile.cpp:22-92) and WriteToFile (Source/src/CameraSupport/
// WriteToFile.cpp:16-131) with status codes instead of console messages.
#include "kf_recording.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <string>

namespace {

std::string path_of(const char* dir, const std::string& file) {
  std::string d(dir);
  if (!d.empty() && d.back() != '/') d += '/';
  return d + file;
}

void copy_line(std::ifstream& in, char* dst, size_t cap) {
  std::string s;
  if (std::getline(in, s)) {
    while (!s.empty() && (s.back() == '\r' || s.back() == '\n')) s.pop_back();
    snprintf(dst, cap, "%s", s.c_str());
  } else {
    dst[0] = 0;
  }
}

}  // namespace

extern "C" int32_t kf_recording_read_info(const char* dir, kf_recording_info* out) {
  if (!dir || !out) return 1;
  memset(out, 0, sizeof(*out));
  std::ifstream in(path_of(dir, "Parameters.txt"));
  if (!in.is_open()) return 2;
  // The reference reads the first nine numbers with operator>> (CameraFromFile.cpp:34-38).
  if (!(in >> out->width >> out->height >> out->fx >> out->fy >> out->cx >> out->cy >> out->z_offset_mm >>
        out->min_depth_mm >> out->max_depth_mm))
    return 2;
  // Frame count, name, date and camera type follow (written by WriteToFile.cpp:20-44).
  if (!(in >> out->frames)) out->frames = 0;
  std::string rest;
  std::getline(in, rest);  // end of the frame-count line
  copy_line(in, out->name, sizeof(out->name));
  copy_line(in, out->date, sizeof(out->date));
  copy_line(in, out->camera_type, sizeof(out->camera_type));
  if (out->width <= 0 || out->height <= 0) return 2;
  return 0;
}

extern "C" int32_t kf_recording_write_info(const char* dir, const kf_recording_info* info) {
  if (!dir || !info || info->width <= 0 || info->height <= 0) return 1;
  std::ofstream out(path_of(dir, "Parameters.txt"));
  if (!out.is_open()) return 2;
  out.precision(9);
  out << info->width << "\n" << info->height << "\n";
  out << info->fx << " " << info->fy << "\n" << info->cx << " " << info->cy << "\n";
  out << info->z_offset_mm << "\n" << info->min_depth_mm << "\n" << info->max_depth_mm << "\n";
  out << info->frames << "\n" << info->name << "\n" << info->date << "\n" << info->camera_type << "\n";
  return out.good() ? 0 : 2;
}

extern "C" int32_t kf_recording_read_depth(const char* dir, int32_t frame, int32_t width, int32_t height,
                                           uint16_t* dst) {
  if (!dir || !dst || frame < 1 || width <= 0 || height <= 0) return 1;
  char name[32];
  snprintf(name, sizeof(name), "recorded%d", frame);
  std::ifstream in(path_of(dir, name), std::ios::binary);
  if (!in.is_open()) return 2;
  const std::streamsize bytes = (std::streamsize)width * height * 2;
  in.read(reinterpret_cast<char*>(dst), bytes);
  return in.gcount() == bytes ? 0 : 2;
}

extern "C" int32_t kf_recording_write_depth(const char* dir, int32_t frame, int32_t width, int32_t height,
                                            const uint16_t* depth) {
  if (!dir || !depth || frame < 1 || width <= 0 || height <= 0) return 1;
  char name[32];
  snprintf(name, sizeof(name), "recorded%d", frame);
  std::ofstream out(path_of(dir, name), std::ios::binary);
  if (!out.is_open()) return 2;
  out.write(reinterpret_cast<const char*>(depth), (std::streamsize)width * height * 2);
  return out.good() ? 0 : 2;
}

extern "C" int32_t kf_recording_intrinsics(const kf_recording_info* info, kf_intrinsics* out) {
  if (!info || !out) return 1;
  out->fx = fabsf(info->fx);
  out->fy = fabsf(info->fy);
  out->cx = info->cx;
  out->cy = info->cy;
  out->width = info->width;
  out->height = info->height;
  return out->fx > 0.0f && out->fy > 0.0f ? 0 : 1;
}
