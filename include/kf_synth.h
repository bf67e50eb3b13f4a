/*
 * kf_synth.h — seeded synthetic depth streams (host C ABI, libkfsynth).
 *
 * The reference has no synthetic input: frames come from cameras or recorded
 * files (Source/src/CameraSupport/CameraFromFile.cpp:22-92).  This generator
 * stands in for that frame source: analytic scenes (box room interior, spheres,
 * axis-aligned boxes), exact ray casting, the Kinect axial noise model
 * sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m from a counter-based RNG keyed by
 * (seed, frame, pixel), range cut-off and uint16 millimetre quantisation, plus
 * smooth camera trajectories with per-frame motion limits.  Everything is
 * deterministic: the same seeds give the same bytes on every machine that runs
 * this library.
 *
 * World frame: y up, room centred at the origin.  Camera frame: x right, y down,
 * z forward.  Poses are row-major 4x4 camera->world (T_wc), FP64.
 */
#ifndef KF_SYNTH_H_
#define KF_SYNTH_H_

#include <stdint.h>

#include "kinfu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define KF_SCENE_MAX_PRIMS 64

typedef enum kf_prim_kind { KF_PRIM_SPHERE = 0, KF_PRIM_BOX = 1 } kf_prim_kind;

typedef struct kf_primitive {
  int32_t kind;
  int32_t pad_;
  double center[3];
  double radius;  /* sphere */
  double half[3]; /* box half extents */
} kf_primitive;

typedef struct kf_scene {
  double room[3]; /* interior size (x, y, z); the room spans [-room/2, room/2] */
  int32_t count;
  int32_t pad_;
  kf_primitive prim[KF_SCENE_MAX_PRIMS];
} kf_scene;

typedef struct kf_scene_desc {
  double room[3];
  int32_t n_spheres;     /* fixed counts ...                                  */
  int32_t n_boxes;
  int32_t n_random_min;  /* ... or, when n_random_max > 0, a random count in   */
  int32_t n_random_max;  /*     [min, max] of random kinds                     */
  double sphere_radius[2];
  double box_half[2];
  double clearance_m;    /* min distance from every camera centre to a surface */
  uint64_t seed;
} kf_scene_desc;

typedef enum kf_trajectory_kind {
  KF_TRAJ_ARC = 0,    /* circle around the room centre looking inwards: exactly
                         max_trans_m chord and max_rot_deg per frame            */
  KF_TRAJ_SPLINE = 1, /* smooth random path, per-frame motion <= the limits     */
  KF_TRAJ_BURST = 2   /* smooth random path with fast bursts up to the limits   */
} kf_trajectory_kind;

typedef struct kf_trajectory_desc {
  int32_t kind;
  int32_t frames;
  double max_trans_m;
  double max_rot_deg;
  double room[3];
  double margin_m; /* keep camera centres this far inside the walls */
  uint64_t seed;
} kf_trajectory_desc;

/* All functions return 0 on success, nonzero on invalid arguments. */
int32_t kf_synth_trajectory(const kf_trajectory_desc* d, double* poses /* frames*16 */);
int32_t kf_synth_scene(const kf_scene_desc* d, const double* poses, int32_t frames, kf_scene* out);
/* Renders one frame.  out_mm (nullable): noisy, cut off at max_range_m, uint16 mm
 * (noise_scale multiplies sigma; 0 = noise-free).  out_true_m (nullable): exact
 * camera-z depth of the first surface in metres (0 where nothing is hit). */
int32_t kf_synth_render(const kf_scene* s, const kf_intrinsics* k, const double T_wc[16],
                        uint64_t noise_seed, int32_t frame, double noise_scale, double max_range_m,
                        uint16_t* out_mm, float* out_true_m);
/* Renders frames [0, frames) of a trajectory into out_mm (frames * width * height). */
int32_t kf_synth_render_batch(const kf_scene* s, const kf_intrinsics* k, const double* poses,
                              int32_t frames, uint64_t noise_seed, double noise_scale,
                              double max_range_m, uint16_t* out_mm);

#ifdef __cplusplus
}
#endif

#endif /* KF_SYNTH_H_ */
