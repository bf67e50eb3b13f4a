// kf_pipeline.cpp — per-frame driver: preprocess -> track -> integrate -> raycast.
//
// Frame order and buffer ownership follow RunApp::display / initCudaPointers
// (Source/src/MainApp.cpp:774-970, 591-647), rotated into the standard
// KinectFusion order.  Differences from the reference driver, by design:
//  * the initial guess is the constant-velocity prediction T1 (T0^-1 T1) instead of
//    the previous pose (Source/src/ICP.cpp:94);
//  * nothing returns to the host inside a frame: poses, stats and the ICP solve
//    stay on the device, and frames after the first tracked one replay a hipGraph.
#include <hip/hip_runtime.h>
#include <string.h>

#include "kf_internal.hpp"
#include "kf_pose_math.hpp"

namespace {

enum PoseSlot { kCur = 0, kPrev1 = 1, kPrev2 = 2, kGuess = 3, kSlots = 4 };

}  // namespace

struct kf_pipeline {
  kf_pipeline_desc d;
  hipStream_t stream = nullptr;
  bool own_stream = false;  // created here when the caller passes NULL (graphs cannot capture the null stream)
  kf_volume* vol = nullptr;
  kf_frame_maps live = {}, model = {};
  kf_tracker* tracker = nullptr;
  kf_pose* d_poses = nullptr;        // kSlots poses
  kf_icp_stats* d_stats = nullptr;   // 2: plain, prior
  uint16_t* d_depth = nullptr;       // staging input buffer
  int frame = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool stage_timing = false;
  hipEvent_t ev[KF_STAGES + 1] = {};
  float stage_ms[KF_STAGES] = {};
  unsigned long long* d_counters = nullptr;  // [updated voxels, raycast steps] (stage-timing mode)
  double* d_traj = nullptr;                  // KF_TRAJ_CAPACITY poses (16 doubles each)
  int* d_traj_count = nullptr;
};

namespace kf {
namespace {

// End of frame: log the frame's pose and shift the history (k-2 <- k-1 <- k).
__global__ void pose_shift(kf_pose* poses, double* traj, int* count) {
  if (threadIdx.x < 16) {
    const int c = *count;
    if (c < KF_TRAJ_CAPACITY) traj[(size_t)c * 16 + threadIdx.x] = poses[kCur].T_wc[threadIdx.x];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  *count = *count + 1;
  poses[kPrev2] = poses[kPrev1];
  poses[kPrev1] = poses[kCur];
}

// One tracked frame.  In stage-timing mode (eager only) events bracket each stage
// and the sweep kernel alone, and the sweep / raycast counters are enabled.
kf_status record_track_frame(kf_pipeline* p, bool timed) {
  hipStream_t s = p->stream;
  const kf_pipeline_desc& d = p->d;
  unsigned long long* upd = timed ? &p->d_counters[0] : nullptr;
  unsigned long long* steps = timed ? &p->d_counters[1] : nullptr;
  if (timed) {
    KF_HIP(hipMemsetAsync(p->d_counters, 0, 2 * sizeof(unsigned long long), s));
    KF_HIP(hipEventRecord(p->ev[0], s));
  }
  KF_TRY(launch_preprocess(d.intr, d.pre, p->d_depth, p->live, s));
  if (timed) KF_HIP(hipEventRecord(p->ev[1], s));
  KF_TRY(launch_pose_constant_velocity(&p->d_poses[kPrev2], &p->d_poses[kPrev1], &p->d_poses[kGuess], s));
  KF_TRY(launch_icp_track(p->tracker, p->live, p->model, &p->d_poses[kGuess], &p->d_poses[kCur],
                          &p->d_stats[0], s));
  if (timed) KF_HIP(hipEventRecord(p->ev[2], s));
  KF_TRY(launch_integrate(p->vol, d.intr, d.integ, p->live.raylen, &p->d_poses[kCur], upd, nullptr, s,
                          timed ? p->ev[3] : nullptr));
  if (timed) KF_HIP(hipEventRecord(p->ev[4], s));
  KF_TRY(launch_raycast(p->vol, d.intr, d.ray, &p->d_poses[kCur], p->model, steps, s));
  hipLaunchKernelGGL(pose_shift, dim3(1), dim3(64), 0, s, p->d_poses, p->d_traj, p->d_traj_count);
  KF_TRY(check_launch("pose_shift"));
  if (timed) KF_HIP(hipEventRecord(p->ev[5], s));
  return KF_OK;
}

kf_status bootstrap_frame(kf_pipeline* p) {
  hipStream_t s = p->stream;
  const kf_pipeline_desc& d = p->d;
  KF_TRY(launch_preprocess(d.intr, d.pre, p->d_depth, p->live, s));
  KF_TRY(launch_integrate(p->vol, d.intr, d.integ, p->live.raylen, &p->d_poses[kCur], nullptr, nullptr, s));
  KF_TRY(launch_raycast(p->vol, d.intr, d.ray, &p->d_poses[kCur], p->model, nullptr, s));
  hipLaunchKernelGGL(pose_shift, dim3(1), dim3(64), 0, s, p->d_poses, p->d_traj, p->d_traj_count);
  return check_launch("pose_shift");
}

void destroy_graph(kf_pipeline* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
}

}  // namespace
}  // namespace kf

using namespace kf;

extern "C" kf_status kf_pipeline_default_desc(kf_pipeline_desc* d) {
  KF_CHECK_ARG(d != nullptr, "desc is NULL");
  memset(d, 0, sizeof(*d));
  d->intr = kf_intrinsics{525.0f, 525.0f, 319.5f, 239.5f, 640, 480};
  d->volume.n = 512;
  d->volume.extent_m = 4.0f;
  d->volume.origin_m[0] = d->volume.origin_m[1] = d->volume.origin_m[2] = -2.0f;
  d->volume.trunc_m = 0.02f;
  d->volume.max_weight = 128.0f;
  d->volume.format = KF_VOXEL_F32;
  d->pre.depth_scale = 0.001f;
  d->pre.min_depth_m = 0.3f;
  d->pre.max_depth_m = 4.0f;
  d->pre.bilateral_radius = 3;
  d->pre.sigma_space_px = 3.0f;
  d->pre.sigma_range_m = 0.03f;
  d->pre.pyramid_gate_m = 0.09f;
  d->integ.min_depth_m = 0.3f;
  d->integ.max_depth_m = 4.0f;
  d->ray.min_depth_m = 0.3f;
  d->ray.step_far = 0.8f;
  d->ray.step_near = 0.5f;
  d->ray.max_steps = 2048;
  d->icp.levels = 3;
  d->icp.iterations[0] = 10;
  d->icp.iterations[1] = 5;
  d->icp.iterations[2] = 4;
  d->icp.dist_thresh_m = 0.1f;
  d->icp.normal_thresh_deg = 20.0f;
  d->icp.compensated = 1;
  d->icp_prior = d->icp;
  d->icp_prior.iterations[0] = 4;
  d->icp_prior.iterations[1] = 3;
  d->icp_prior.iterations[2] = 3;
  d->fallback_residual_ratio = 1.5f;
  d->fallback_min_inlier_frac = 0.4f;
  d->use_graph = 1;
  return KF_OK;
}

extern "C" kf_status kf_pipeline_destroy(kf_pipeline* p) {
  if (!p) return KF_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  destroy_graph(p);
  if (p->own_stream) (void)hipStreamDestroy(p->stream);
  for (hipEvent_t& e : p->ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipFree(p->d_counters);
  (void)hipFree(p->d_traj);
  (void)hipFree(p->d_traj_count);
  kf_volume_destroy(p->vol);
  kf_tracker_destroy(p->tracker);
  if (p->live.level[0].vx) kf_maps_free(&p->live);
  if (p->model.level[0].vx) kf_maps_free(&p->model);
  (void)hipFree(p->d_poses);
  (void)hipFree(p->d_stats);
  (void)hipFree(p->d_depth);
  delete p;
  return KF_OK;
}

extern "C" kf_status kf_pipeline_create(const kf_pipeline_desc* d, kf_stream s, kf_pipeline** out) {
  KF_CHECK_ARG(d && out, "NULL argument");
  KF_CHECK_ARG(d->icp.levels == d->icp_prior.levels, "plain and prior ICP must use the same levels");
  kf_pipeline* p = new kf_pipeline();
  p->d = *d;
  p->stream = S(s);
  const int levels = d->icp.levels;
  kf_status st = KF_OK;
  if (!p->stream) {
    hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete p;
      return hip_fail(e, "hipStreamCreate");
    }
    p->own_stream = true;
  }
  st = kf_volume_create(&d->volume, &p->vol);
  if (st == KF_OK) st = kf_tracker_create(&d->intr, &d->icp, &p->tracker);
  if (st == KF_OK) st = kf_maps_alloc(&d->intr, levels, 1, &p->live);
  if (st == KF_OK) st = kf_maps_alloc(&d->intr, levels, 0, &p->model);
  if (st == KF_OK) {
    hipError_t e = hipMalloc(&p->d_poses, kSlots * sizeof(kf_pose));
    if (e == hipSuccess) e = hipMalloc(&p->d_stats, 2 * sizeof(kf_icp_stats));
    if (e == hipSuccess) e = hipMalloc(&p->d_depth, (size_t)d->intr.width * d->intr.height * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(&p->d_counters, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&p->d_traj, (size_t)KF_TRAJ_CAPACITY * 16 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&p->d_traj_count, sizeof(int));
    if (e == hipSuccess) e = hipMemset(p->d_traj_count, 0, sizeof(int));
    if (e == hipSuccess) e = hipMemset(p->d_counters, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(p->d_stats, 0, 2 * sizeof(kf_icp_stats));
    for (int i = 0; i < KF_STAGES + 1 && e == hipSuccess; ++i) e = hipEventCreate(&p->ev[i]);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) st = hip_fail(e, "pipeline allocation");
  }
  if (st != KF_OK) {
    kf_pipeline_destroy(p);
    return st;
  }
  *out = p;
  return KF_OK;
}

extern "C" kf_status kf_pipeline_reset(kf_pipeline* p, const double T0[16]) {
  KF_CHECK_ARG(p && T0, "NULL argument");
  KF_HIP(hipStreamSynchronize(p->stream));
  KF_TRY(launch_volume_reset(p->vol, p->stream));
  for (int i = 0; i < kSlots; ++i) KF_TRY(kf_pose_upload(&p->d_poses[i], T0, p->stream));
  KF_HIP(hipMemsetAsync(p->d_stats, 0, 2 * sizeof(kf_icp_stats), p->stream));
  KF_HIP(hipMemsetAsync(p->d_traj_count, 0, sizeof(int), p->stream));
  KF_HIP(hipStreamSynchronize(p->stream));
  p->frame = 0;
  return KF_OK;
}

extern "C" kf_status kf_pipeline_process(kf_pipeline* p, const uint16_t* d_depth) {
  KF_CHECK_ARG(p && d_depth, "NULL argument");
  hipStream_t s = p->stream;
  const size_t bytes = (size_t)p->d.intr.width * p->d.intr.height * sizeof(uint16_t);
  KF_HIP(hipMemcpyAsync(p->d_depth, d_depth, bytes, hipMemcpyDeviceToDevice, s));
  if (p->frame == 0) {
    KF_TRY(bootstrap_frame(p));
  } else if (p->stage_timing || !p->d.use_graph) {
    KF_TRY(record_track_frame(p, p->stage_timing));
    if (p->stage_timing) {
      KF_HIP(hipEventSynchronize(p->ev[KF_STAGES]));
      for (int i = 0; i < KF_STAGES; ++i)
        KF_HIP(hipEventElapsedTime(&p->stage_ms[i], p->ev[i], p->ev[i + 1]));
    }
  } else {
    if (!p->exec) {
      KF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      kf_status st = record_track_frame(p, false);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s, &g);
      if (st != KF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
      p->graph = g;
      e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        destroy_graph(p);
        return hip_fail(e, "hipGraphInstantiate");
      }
    }
    KF_HIP(hipGraphLaunch(p->exec, s));
  }
  ++p->frame;
  return KF_OK;
}

extern "C" kf_status kf_pipeline_get_pose(kf_pipeline* p, double T[16]) {
  KF_CHECK_ARG(p && T, "NULL argument");
  return kf_pose_download(T, &p->d_poses[kCur], p->stream);
}

extern "C" kf_status kf_pipeline_get_trajectory(kf_pipeline* p, double* T_out, int32_t max_frames,
                                                int32_t* frames) {
  KF_CHECK_ARG(p && frames && (T_out || max_frames == 0), "NULL argument");
  int c = 0;
  KF_HIP(hipMemcpyAsync(&c, p->d_traj_count, sizeof(int), hipMemcpyDeviceToHost, p->stream));
  KF_HIP(hipStreamSynchronize(p->stream));
  *frames = c;
  int n = c < KF_TRAJ_CAPACITY ? c : KF_TRAJ_CAPACITY;
  if (n > max_frames) n = max_frames;
  if (n > 0) {
    KF_HIP(hipMemcpyAsync(T_out, p->d_traj, (size_t)n * 16 * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    KF_HIP(hipStreamSynchronize(p->stream));
  }
  return KF_OK;
}

extern "C" kf_status kf_pipeline_get_stats(kf_pipeline* p, kf_frame_stats* st) {
  KF_CHECK_ARG(p && st, "NULL argument");
  memset(st, 0, sizeof(*st));
  kf_icp_stats h[2];
  unsigned long long c[2];
  KF_HIP(hipMemcpyAsync(h, p->d_stats, sizeof(h), hipMemcpyDeviceToHost, p->stream));
  KF_HIP(hipMemcpyAsync(c, p->d_counters, sizeof(c), hipMemcpyDeviceToHost, p->stream));
  KF_HIP(hipStreamSynchronize(p->stream));
  st->frame = p->frame - 1;
  st->icp = h[0];
  st->icp_prior = h[1];
  st->updated_voxels = c[0];
  st->raycast_steps = c[1];
  return KF_OK;
}

extern "C" kf_status kf_pipeline_get_volume(kf_pipeline* p, kf_volume** v) {
  KF_CHECK_ARG(p && v, "NULL argument");
  *v = p->vol;
  return KF_OK;
}

extern "C" kf_status kf_pipeline_get_maps(kf_pipeline* p, kf_frame_maps* live, kf_frame_maps* model) {
  KF_CHECK_ARG(p, "NULL argument");
  if (live) *live = p->live;
  if (model) *model = p->model;
  return KF_OK;
}

extern "C" kf_status kf_pipeline_get_stream(kf_pipeline* p, kf_stream* s) {
  KF_CHECK_ARG(p && s, "NULL argument");
  *s = p->stream;
  return KF_OK;
}

extern "C" kf_status kf_pipeline_set_stage_timing(kf_pipeline* p, int32_t on) {
  KF_CHECK_ARG(p, "NULL argument");
  p->stage_timing = on != 0;
  return KF_OK;
}

extern "C" kf_status kf_pipeline_get_stage_ms(kf_pipeline* p, float ms[KF_STAGES]) {
  KF_CHECK_ARG(p && ms, "NULL argument");
  for (int i = 0; i < KF_STAGES; ++i) ms[i] = p->stage_ms[i];
  return KF_OK;
}
