// kinfu.hpp — header-only C++ wrappers over the kinfu C ABI with the reference's
// class and method names, so a driver written against OpenFusion's reconstruction
// API keeps its shape:
//
//   OpenFusion                                  here
//   General::Instance()->InitVoxelArray(..)     kinfu::General::Instance()->InitVoxelArray(vol)
//     (Source/include/General.h:12-19)
//   General::UpdateTransforms(..)               kinfu::General::Instance()->UpdateTransforms(d_pose, T_wc)
//   CameraHandler::updateCameraData(..)         kinfu::CameraHandler::updateCameraData(k, p, live, d_depth)
//     (Source/include/CameraHandler.h:24)
//   VolumeIntegrator(params).Integrate(..)      kinfu::VolumeIntegrator(k, p).Integrate(vol, d_pose, live)
//     (Source/include/VolumeIntegrator.h:14-15)
//   Raycaster(params).Raycast(..)               kinfu::Raycaster(k, p).Raycast(vol, d_pose, model)
//     (Source/include/Raycaster.h:11-12)
//   ICP(grid, block, params).Iterate_ICP(..)    kinfu::ICP(k, p).Iterate_ICP(live, model, d_pose)
//     (Source/include/ICP.h:53-68)
//
// Differences: buffers are kf_volume / kf_frame_maps / device kf_pose instead of
// raw tsdf/weight/float[16] pointers; errors throw kinfu::Error (the reference
// calls exit(-1)); every call takes an optional stream.
#ifndef KINFU_HPP_
#define KINFU_HPP_

#include <stdexcept>
#include <string>

#include "kinfu.h"

namespace kinfu {

class Error : public std::runtime_error {
 public:
  Error(kf_status s, const std::string& what)
      : std::runtime_error(what + ": " + kf_last_error()), status(s) {}
  kf_status status;
};

inline void check(kf_status s, const char* what) {
  if (s != KF_OK) throw Error(s, what);
}

class General {
 public:
  static General* Instance() {
    static General g;
    return &g;
  }
  void InitVoxelArray(kf_volume* vol, kf_stream s = nullptr) { check(kf_volume_reset(vol, s), "InitVoxelArray"); }
  // Sets the device pose from a host camera->world transform; the inverse and the
  // FP32 copies (the reference's d_invCurrentTransform / d_cameraPosition) follow.
  void UpdateTransforms(kf_pose* d_pose, const double T_wc[16], kf_stream s = nullptr) {
    check(kf_pose_upload(d_pose, T_wc, s), "UpdateTransforms");
  }

 private:
  General() = default;
};

class CameraHandler {
 public:
  static void updateCameraData(const kf_intrinsics& k, const kf_preprocess_params& p, kf_frame_maps& live,
                               const uint16_t* d_depth, kf_stream s = nullptr) {
    check(kf_preprocess(&k, &p, d_depth, &live, s), "updateCameraData");
  }
};

class VolumeIntegrator {
 public:
  VolumeIntegrator(const kf_intrinsics& k, const kf_integrate_params& p) : k_(k), p_(p) {}
  void Integrate(kf_volume* vol, const kf_pose* d_pose, const kf_frame_maps& depthData, kf_stream s = nullptr,
                 unsigned long long* d_updated = nullptr) {
    check(kf_integrate(vol, &k_, &p_, depthData.raylen, d_pose, d_updated, nullptr, s), "Integrate");
  }

 private:
  kf_intrinsics k_;
  kf_integrate_params p_;
};

class Raycaster {
 public:
  Raycaster(const kf_intrinsics& k, const kf_raycast_params& p) : k_(k), p_(p) {}
  void Raycast(const kf_volume* vol, const kf_pose* d_pose, kf_frame_maps& model, kf_stream s = nullptr) {
    check(kf_raycast(vol, &k_, &p_, d_pose, &model, nullptr, s), "Raycast");
  }

 private:
  kf_intrinsics k_;
  kf_raycast_params p_;
};

class ICP {
 public:
  ICP(const kf_intrinsics& k, const kf_icp_params& p) { check(kf_tracker_create(&k, &p, &t_), "ICP"); }
  ~ICP() { kf_tracker_destroy(t_); }
  ICP(const ICP&) = delete;
  ICP& operator=(const ICP&) = delete;

  // d_currentTransform: in = initial estimate, out = tracked pose (device).
  void Iterate_ICP(const kf_frame_maps& live, const kf_frame_maps& model, kf_pose* d_currentTransform,
                   kf_stream s = nullptr, kf_icp_stats* d_stats = nullptr) {
    check(kf_icp_track(t_, &live, &model, d_currentTransform, d_currentTransform, d_stats, s), "Iterate_ICP");
  }
  // One association + reduction pass (30 FP64 sums on the device).
  void FindCorresponding(int level, const kf_frame_maps& live, const kf_frame_maps& model, const kf_pose* d_est,
                         double* d_sums, uint8_t* d_mask = nullptr, kf_stream s = nullptr) {
    check(kf_icp_reduce(t_, level, &live, &model, d_est, d_sums, d_mask, s), "FindCorresponding");
  }
  kf_tracker* handle() { return t_; }

 private:
  kf_tracker* t_ = nullptr;
};

}  // namespace kinfu

#endif  // KINFU_HPP_
