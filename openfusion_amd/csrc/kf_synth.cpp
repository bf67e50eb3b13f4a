// kf_synth.cpp — seeded synthetic scenes, trajectories and depth frames (host).
//
// Stands in for the reference's frame sources (CameraAbstract::GetDepthFrame,
// Source/include/CameraAbstract.h:76; CameraFromFile, Source/src/CameraSupport/
// CameraFromFile.cpp:68-92): it produces uint16 millimetre depth frames of an
// analytic scene plus ground-truth poses.
#include "kf_synth.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
uint64_t hash3(uint64_t a, uint64_t b, uint64_t c) {
  uint64_t s = a * 0xD1B54A32D192ED03ull ^ b * 0x9E3779B97F4A7C15ull ^ c * 0xABC98388FB8FAC03ull;
  splitmix(s);
  return splitmix(s);
}
double unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed ^ 0x5DEECE66Dull) {}
  double u() { return unit(splitmix(s)); }
  double range(double a, double b) { return a + (b - a) * u(); }
};

struct V3 {
  double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double norm(V3 a) { return sqrt(dot(a, a)); }
V3 normalize(V3 a) { return a * (1.0 / norm(a)); }

// Rotation matrices (row-major 3x3).
void mat3_mul(const double* A, const double* B, double* C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
}
void rot_x(double a, double* R) {
  const double c = cos(a), s = sin(a);
  const double M[9] = {1, 0, 0, 0, c, -s, 0, s, c};
  memcpy(R, M, sizeof(M));
}
void rot_y(double a, double* R) {
  const double c = cos(a), s = sin(a);
  const double M[9] = {c, 0, s, 0, 1, 0, -s, 0, c};
  memcpy(R, M, sizeof(M));
}
void rot_z(double a, double* R) {
  const double c = cos(a), s = sin(a);
  const double M[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
  memcpy(R, M, sizeof(M));
}

void set_pose(double* T, const double* R, V3 p) {
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T[r * 4 + c] = R[r * 3 + c];
  }
  T[3] = p.x;
  T[7] = p.y;
  T[11] = p.z;
  T[12] = T[13] = T[14] = 0.0;
  T[15] = 1.0;
}

// Rotation angle (rad) of R_a^T R_b.
double rel_angle(const double* Ta, const double* Tb) {
  double tr = 0.0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) tr += Ta[r * 4 + c] * Tb[r * 4 + c];
  double ca = std::min(1.0, std::max(-1.0, 0.5 * (tr - 1.0)));
  return acos(ca);
}
double rel_trans(const double* Ta, const double* Tb) {
  V3 d = {Tb[3] - Ta[3], Tb[7] - Ta[7], Tb[11] - Ta[11]};
  return norm(d);
}

// Smooth random path: a wobbling orbit around the room centre.  The parameter s
// is the orbit angle; radius, height, look-at target and roll are sums of random
// sinusoids of s, and the camera looks at the target near the centre, so the far
// side of the room and its primitives stay in view (a camera staring at a bare wall
// sees a plane, which leaves ICP's translation along it unobservable).
struct Path {
  double r0, dir, th0;
  double ra[3], rw[3], rp[3];   // radius wobble
  double ya[3], yw[3], yp[3];   // height
  double ta[3][3], tw[3][3], tp[3][3];  // look-at target wobble (x, y, z)
  double qa[2], qw[2], qp[2];   // roll

  void init(Rng& g, const double* room, double margin) {
    const double half = 0.5 * std::min(room[0], room[2]);
    r0 = std::max(0.3, half - margin - 0.1 * half);
    dir = g.u() < 0.5 ? -1.0 : 1.0;
    th0 = g.range(0, 2 * kPi);
    const double hy = 0.5 * room[1];
    for (int i = 0; i < 3; ++i) {
      ra[i] = 0.1 * r0 * g.range(0.3, 1.0) / (i + 1);
      rw[i] = g.range(0.7, 1.5) * (i + 1);
      rp[i] = g.range(0, 2 * kPi);
      ya[i] = 0.25 * hy * g.range(0.3, 1.0) / (i + 1);
      yw[i] = g.range(0.7, 1.5) * (i + 1);
      yp[i] = g.range(0, 2 * kPi);
      for (int k = 0; k < 3; ++k) {
        ta[k][i] = 0.15 * r0 * g.range(0.3, 1.0) / (i + 1);
        tw[k][i] = g.range(0.7, 2.0) * (i + 1);
        tp[k][i] = g.range(0, 2 * kPi);
      }
    }
    for (int i = 0; i < 2; ++i) {
      qa[i] = g.range(0.02, 0.06) / (i + 1);
      qw[i] = g.range(0.7, 1.6) * (i + 1);
      qp[i] = g.range(0, 2 * kPi);
    }
  }
  void eval(double s, double* T) const {
    double r = r0, y = 0.0, roll = 0.0;
    double tg[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; ++i) {
      r += ra[i] * sin(rw[i] * s + rp[i]);
      y += ya[i] * sin(yw[i] * s + yp[i]);
      for (int k = 0; k < 3; ++k) tg[k] += ta[k][i] * sin(tw[k][i] * s + tp[k][i]);
    }
    for (int i = 0; i < 2; ++i) roll += qa[i] * sin(qw[i] * s + qp[i]);
    const double th = th0 + dir * s;
    const V3 p = {r * sin(th), y, r * cos(th)};
    const V3 fwd = normalize(V3{tg[0], tg[1], tg[2]} - p);
    const V3 right0 = normalize(cross(V3{0, -1, 0}, fwd));
    const V3 down0 = cross(fwd, right0);
    // roll about the viewing direction
    const double cr = cos(roll), sr = sin(roll);
    const V3 right = right0 * cr + down0 * sr;
    const V3 down = down0 * cr - right0 * sr;
    const double R[9] = {right.x, down.x, fwd.x, right.y, down.y, fwd.y, right.z, down.z, fwd.z};
    set_pose(T, R, p);
  }
};

// Per-frame stepping along a path under translation/rotation limits.
void step_path(const Path& path, int frames, double lt, double lr, const std::vector<double>& speed,
               double* poses) {
  double s = 0.0;
  path.eval(s, poses);
  for (int k = 1; k < frames; ++k) {
    double* Tp = poses + (k - 1) * 16;
    double* Tk = poses + k * 16;
    const double h = 1e-4;
    double Ta[16], Tb[16];
    path.eval(s, Ta);
    path.eval(s + h, Tb);
    const double vt = rel_trans(Ta, Tb) / h, vr = rel_angle(Ta, Tb) / h;
    double ds = speed[k] * std::min(lt / std::max(vt, 1e-9), lr / std::max(vr, 1e-9));
    for (int it = 0; it < 60; ++it) {
      path.eval(s + ds, Tk);
      if (rel_trans(Tp, Tk) <= lt && rel_angle(Tp, Tk) <= lr) break;
      ds *= 0.9;
    }
    s += ds;
    path.eval(s, Tk);
  }
}

bool inside_room(const kf_scene* sc, V3 p) {
  return fabs(p.x) < sc->room[0] * 0.5 && fabs(p.y) < sc->room[1] * 0.5 && fabs(p.z) < sc->room[2] * 0.5;
}

double prim_distance(const kf_primitive& q, V3 p) {
  V3 c = {q.center[0], q.center[1], q.center[2]};
  if (q.kind == KF_PRIM_SPHERE) return norm(p - c) - q.radius;
  V3 d = {fabs(p.x - c.x) - q.half[0], fabs(p.y - c.y) - q.half[1], fabs(p.z - c.z) - q.half[2]};
  V3 o = {std::max(d.x, 0.0), std::max(d.y, 0.0), std::max(d.z, 0.0)};
  const double inside = std::min(std::max(d.x, std::max(d.y, d.z)), 0.0);
  return norm(o) + inside;
}

// Ray (o + t d) hits: returns the smallest t > 0 or +inf.
double hit_sphere(const kf_primitive& q, V3 o, V3 d) {
  V3 c = {q.center[0], q.center[1], q.center[2]};
  V3 oc = o - c;
  const double a = dot(d, d), b = dot(oc, d), cc = dot(oc, oc) - q.radius * q.radius;
  const double disc = b * b - a * cc;
  if (disc < 0.0) return INFINITY;
  const double sq = sqrt(disc);
  double t = (-b - sq) / a;
  if (t > 0.0) return t;
  t = (-b + sq) / a;
  return t > 0.0 ? t : INFINITY;
}
double hit_box(const kf_primitive& q, V3 o, V3 d) {
  const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  double tn = -INFINITY, tf = INFINITY;
  for (int k = 0; k < 3; ++k) {
    const double lo = q.center[k] - q.half[k], hi = q.center[k] + q.half[k];
    if (dd[k] == 0.0) {
      if (oo[k] < lo || oo[k] > hi) return INFINITY;
      continue;
    }
    double t1 = (lo - oo[k]) / dd[k], t2 = (hi - oo[k]) / dd[k];
    if (t1 > t2) std::swap(t1, t2);
    tn = std::max(tn, t1);
    tf = std::min(tf, t2);
  }
  if (tn > tf || tf <= 0.0) return INFINITY;
  return tn > 0.0 ? tn : INFINITY;
}
double hit_room(const kf_scene* sc, V3 o, V3 d) {
  const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  double t = INFINITY;
  for (int k = 0; k < 3; ++k) {
    if (dd[k] == 0.0) continue;
    const double wall = dd[k] > 0.0 ? sc->room[k] * 0.5 : -sc->room[k] * 0.5;
    t = std::min(t, (wall - oo[k]) / dd[k]);
  }
  return t;
}

double gauss(uint64_t seed, int32_t frame, int64_t pixel) {
  const uint64_t h1 = hash3(seed, (uint64_t)(uint32_t)frame, (uint64_t)pixel * 2 + 0);
  const uint64_t h2 = hash3(seed, (uint64_t)(uint32_t)frame, (uint64_t)pixel * 2 + 1);
  const double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);  // (0, 1]
  const double u2 = unit(h2);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
}

}  // namespace

extern "C" int32_t kf_synth_trajectory(const kf_trajectory_desc* d, double* poses) {
  if (!d || !poses || d->frames < 1 || d->max_trans_m <= 0.0 || d->max_rot_deg <= 0.0) return 1;
  Rng g(d->seed);
  const double lt = d->max_trans_m, lr = d->max_rot_deg * kPi / 180.0;
  if (d->kind == KF_TRAJ_ARC) {
    const double r = lt / (2.0 * sin(lr * 0.5));  // chord per frame == lt exactly
    const double th0 = g.range(0, 2 * kPi);
    for (int k = 0; k < d->frames; ++k) {
      const double th = th0 + k * lr;
      V3 p = {r * sin(th), 0.0, r * cos(th)};
      V3 fwd = normalize(V3{0, 0, 0} - p);
      V3 down = {0, -1, 0};
      V3 right = normalize(cross(down, fwd));
      down = cross(fwd, right);
      const double R[9] = {right.x, down.x, fwd.x, right.y, down.y, fwd.y, right.z, down.z, fwd.z};
      set_pose(poses + k * 16, R, p);
    }
    return 0;
  }
  if (d->kind != KF_TRAJ_SPLINE && d->kind != KF_TRAJ_BURST) return 1;
  Path path;
  path.init(g, d->room, d->margin_m);
  std::vector<double> speed(d->frames, 0.9);
  if (d->kind == KF_TRAJ_SPLINE) {
    const double ph = g.range(0, 2 * kPi);
    for (int k = 0; k < d->frames; ++k) speed[k] = 0.75 + 0.2 * sin(0.05 * k + ph);
  } else {
    const int bursts = std::max(1, d->frames / 60);
    std::vector<double> at(bursts);
    for (int b = 0; b < bursts; ++b) at[b] = g.range(0.0, (double)d->frames);
    for (int k = 0; k < d->frames; ++k) {
      double v = 0.35;
      for (int b = 0; b < bursts; ++b) {
        const double x = (k - at[b]) / 6.0;
        v += 0.65 * exp(-0.5 * x * x);
      }
      speed[k] = std::min(1.0, v);
    }
  }
  step_path(path, d->frames, lt, lr, speed, poses);
  return 0;
}

extern "C" int32_t kf_synth_scene(const kf_scene_desc* d, const double* poses, int32_t frames, kf_scene* out) {
  if (!d || !out || (frames > 0 && !poses)) return 1;
  memset(out, 0, sizeof(*out));
  for (int k = 0; k < 3; ++k) out->room[k] = d->room[k];
  Rng g(d->seed);
  std::vector<int> kinds;
  if (d->n_random_max > 0) {
    const int n = d->n_random_min + (int)(g.u() * (d->n_random_max - d->n_random_min + 1));
    for (int i = 0; i < n; ++i) kinds.push_back(g.u() < 0.5 ? KF_PRIM_SPHERE : KF_PRIM_BOX);
  } else {
    for (int i = 0; i < d->n_spheres; ++i) kinds.push_back(KF_PRIM_SPHERE);
    for (int i = 0; i < d->n_boxes; ++i) kinds.push_back(KF_PRIM_BOX);
  }
  if ((int)kinds.size() > KF_SCENE_MAX_PRIMS) return 1;
  for (int kind : kinds) {
    for (int attempt = 0; attempt < 2000; ++attempt) {
      kf_primitive q;
      memset(&q, 0, sizeof(q));
      q.kind = kind;
      double ext[3];
      if (kind == KF_PRIM_SPHERE) {
        q.radius = g.range(d->sphere_radius[0], d->sphere_radius[1]);
        ext[0] = ext[1] = ext[2] = q.radius;
      } else {
        for (int k = 0; k < 3; ++k) ext[k] = q.half[k] = g.range(d->box_half[0], d->box_half[1]);
      }
      bool fits = true;
      for (int k = 0; k < 3; ++k) {
        const double lim = d->room[k] * 0.5 - ext[k];
        if (lim <= 0.0) fits = false;
        q.center[k] = fits ? g.range(-lim, lim) : 0.0;
      }
      if (!fits) continue;
      bool clear = true;
      for (int f = 0; f < frames && clear; ++f) {
        const double* T = poses + f * 16;
        if (prim_distance(q, V3{T[3], T[7], T[11]}) < d->clearance_m) clear = false;
      }
      if (!clear) continue;
      out->prim[out->count++] = q;
      break;
    }
  }
  return 0;
}

extern "C" int32_t kf_synth_render(const kf_scene* sc, const kf_intrinsics* k, const double* T,
                                   uint64_t noise_seed, int32_t frame, double noise_scale,
                                   double max_range, uint16_t* out_mm, float* out_true) {
  if (!sc || !k || !T || k->width <= 0 || k->height <= 0) return 1;
  const int W = k->width, H = k->height;
  const V3 o = {T[3], T[7], T[11]};
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      const double xn = ((double)x - k->cx) / k->fx, yn = ((double)y - k->cy) / k->fy;
      // Direction with unit camera-z, so the ray parameter is the depth.
      const V3 d = {T[0] * xn + T[1] * yn + T[2], T[4] * xn + T[5] * yn + T[6], T[8] * xn + T[9] * yn + T[10]};
      double z = inside_room(sc, o) ? hit_room(sc, o, d) : INFINITY;
      for (int i = 0; i < sc->count; ++i) {
        const kf_primitive& q = sc->prim[i];
        z = std::min(z, q.kind == KF_PRIM_SPHERE ? hit_sphere(q, o, d) : hit_box(q, o, d));
      }
      if (!(z < INFINITY)) z = 0.0;
      const int64_t pix = (int64_t)y * W + x;
      if (out_true) out_true[pix] = (float)z;
      if (out_mm) {
        uint16_t mm = 0;
        if (z > 0.0) {
          const double sig = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
          const double zn = z + noise_scale * sig * gauss(noise_seed, frame, pix);
          if (zn > 0.0 && zn <= max_range) mm = (uint16_t)std::min(65535.0, floor(zn * 1000.0 + 0.5));
        }
        out_mm[pix] = mm;
      }
    }
  }
  return 0;
}

extern "C" int32_t kf_synth_render_batch(const kf_scene* sc, const kf_intrinsics* k, const double* poses,
                                         int32_t frames, uint64_t noise_seed, double noise_scale,
                                         double max_range, uint16_t* out) {
  if (!poses || !out || frames < 0) return 1;
  const size_t P = (size_t)k->width * k->height;
  for (int f = 0; f < frames; ++f) {
    const int32_t r = kf_synth_render(sc, k, poses + f * 16, noise_seed, f, noise_scale, max_range,
                                      out + f * P, nullptr);
    if (r) return r;
  }
  return 0;
}
