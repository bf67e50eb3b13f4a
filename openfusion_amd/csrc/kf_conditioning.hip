// kf_conditioning.hip — camera-side depth conditioning ahead of preprocessing.
//
//  * depth_median: 5x5 median of the valid (non-null) neighbours, null when at
//    most 12 of the 25 are valid, mean of the two middle values (integer division)
//    for an even count — d_MedianFilter / getMedian
//    (Source/src/CameraSupport/CameraAbstract.cu:76-124).  Pixels outside the image
//    are skipped; the reference wraps their indices through unsigned arithmetic
//    (CameraAbstract.cu:111), which is not reproduced.
//  * depth_undistort: Brown-Conrady remap (k1, k2 radial; p1, p2 tangential) with
//    bilinear sampling — d_undistort (CameraAbstract.cu:160-221).  Differences, by
//    design: both distorted coordinates are computed from the undistorted ones (the
//    reference feeds the already distorted x into y's formula, :182-183), a sample
//    whose four source pixels are not all valid is null (interpolating into
//    missing depth creates flying pixels), and out-of-image samples are null
//    instead of left unwritten.  The result is rounded to the nearest millimetre.
#include "kf_device.hpp"
#include "kf_internal.hpp"

namespace kf {
namespace {

constexpr int kMedR = 2, kMedN = 25;

__global__ __launch_bounds__(256) void depth_median(const uint16_t* in, uint16_t* out, int W, int H,
                                                    uint16_t null_v) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= W || y >= H) return;
  uint16_t v[kMedN];
  int n = 0;
  for (int j = -kMedR; j <= kMedR; ++j) {
    const int yy = y + j;
    if (yy < 0 || yy >= H) continue;
    for (int i = -kMedR; i <= kMedR; ++i) {
      const int xx = x + i;
      if (xx < 0 || xx >= W) continue;
      const uint16_t d = in[yy * W + xx];
      if (d != null_v) v[n++] = d;
    }
  }
  uint16_t r = null_v;
  if (n > kMedN / 2) {
    for (int a = 1; a < n; ++a) {  // insertion sort (n <= 25)
      const uint16_t t = v[a];
      int b = a - 1;
      while (b >= 0 && v[b] > t) {
        v[b + 1] = v[b];
        --b;
      }
      v[b + 1] = t;
    }
    r = (n & 1) ? v[n / 2] : (uint16_t)(((uint32_t)v[n / 2] + (uint32_t)v[n / 2 - 1]) / 2u);
  }
  out[y * W + x] = r;
}

__global__ __launch_bounds__(256) void depth_undistort(const uint16_t* in, uint16_t* out, int W, int H, float fx,
                                                       float fy, float cx, float cy, float k1, float k2, float p1,
                                                       float p2) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= W || y >= H) return;
  const float xu = ((float)x - cx) / fx;
  const float yu = ((float)y - cy) / fy;
  const float r2 = xu * xu + yu * yu;
  const float radial = 1.0f + (k1 * r2 + (k2 * r2) * r2);
  const float xd = xu * radial + ((2.0f * p1) * xu * yu + p2 * (r2 + (2.0f * xu) * xu));
  const float yd = yu * radial + ((2.0f * p2) * xu * yu + p1 * (r2 + (2.0f * yu) * yu));
  const float u = fx * xd + cx;
  const float v = fy * yd + cy;
  uint16_t r = 0;
  const float fu = floorf(u), fv = floorf(v);
  if (fu >= 0.0f && fv >= 0.0f && fu <= (float)(W - 2) && fv <= (float)(H - 2)) {
    const int iu = (int)fu, iv = (int)fv;
    const float tu = u - fu, tv = v - fv;
    const uint16_t a = in[iv * W + iu], b = in[iv * W + iu + 1];
    const uint16_t c = in[(iv + 1) * W + iu], d = in[(iv + 1) * W + iu + 1];
    if (a && b && c && d) {
      const float top = (float)a + ((float)b - (float)a) * tu;
      const float bot = (float)c + ((float)d - (float)c) * tu;
      r = (uint16_t)(top + (bot - top) * tv + 0.5f);
    }
  }
  out[y * W + x] = r;
}

}  // namespace
}  // namespace kf

using namespace kf;

extern "C" kf_status kf_depth_median(const uint16_t* d_in, uint16_t* d_out, int32_t width, int32_t height,
                                     uint16_t null_value, kf_stream s) {
  KF_CHECK_ARG(d_in && d_out && d_in != d_out, "bad buffers (in-place is not supported)");
  KF_CHECK_ARG(width > 0 && height > 0, "bad size");
  hipLaunchKernelGGL(depth_median, dim3(div_up(width, 16), div_up(height, 16)), dim3(256), 0, S(s), d_in, d_out,
                     width, height, null_value);
  return check_launch("depth_median");
}

extern "C" kf_status kf_depth_undistort(const uint16_t* d_in, uint16_t* d_out, const kf_intrinsics* k,
                                        const float dist[4], kf_stream s) {
  KF_CHECK_ARG(d_in && d_out && d_in != d_out && k && dist, "bad arguments (in-place is not supported)");
  KF_CHECK_ARG(k->width >= 2 && k->height >= 2 && k->fx > 0.0f && k->fy > 0.0f, "bad intrinsics");
  hipLaunchKernelGGL(depth_undistort, dim3(div_up(k->width, 16), div_up(k->height, 16)), dim3(256), 0, S(s), d_in,
                     d_out, k->width, k->height, k->fx, k->fy, k->cx, k->cy, dist[0], dist[1], dist[2], dist[3]);
  return check_launch("depth_undistort");
}
