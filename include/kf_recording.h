/*
 * kf_recording.h — recorded depth streams in the reference's on-disk format
 * (host C ABI, in libkfsynth with the other frame sources).
 *
 * Replaces CameraFromFile (Source/src/CameraSupport/CameraFromFile.cpp:22-92) and
 * WriteToFile (Source/src/CameraSupport/WriteToFile.cpp:16-174); format as in
 * Source/data/ParametersFileFormat.txt:1-11:
 *   <dir>/Parameters.txt   text, one item per line: width, height, "fx fy",
 *                          "cx cy", z offset (mm), min depth (mm), max depth (mm),
 *                          frame count, name, date, camera type
 *   <dir>/recorded<N>      raw little-endian uint16 millimetre depth, width*height,
 *                          row-major, N = 1, 2, ... (the reference counts from 1)
 * All functions return 0 on success, nonzero on error (1 = bad argument, 2 = I/O).
 */
#ifndef KF_RECORDING_H_
#define KF_RECORDING_H_

#include <stdint.h>

#include "kinfu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kf_recording_info {
  int32_t width, height;
  float fx, fy, cx, cy; /* as stored (the reference may store negative focal lengths) */
  float z_offset_mm;
  float min_depth_mm, max_depth_mm;
  int32_t frames;
  char name[128];
  char date[64];
  char camera_type[64];
} kf_recording_info;

int32_t kf_recording_read_info(const char* dir, kf_recording_info* out);
int32_t kf_recording_write_info(const char* dir, const kf_recording_info* info);
/* frame is 1-based, like the reference's recorded<N> files. */
int32_t kf_recording_read_depth(const char* dir, int32_t frame, int32_t width, int32_t height, uint16_t* out);
int32_t kf_recording_write_depth(const char* dir, int32_t frame, int32_t width, int32_t height,
                                 const uint16_t* depth);
/* Intrinsics for this library (z-forward pinhole: |fx|, |fy|, cx, cy). */
int32_t kf_recording_intrinsics(const kf_recording_info* info, kf_intrinsics* out);

#ifdef __cplusplus
}
#endif

#endif /* KF_RECORDING_H_ */
