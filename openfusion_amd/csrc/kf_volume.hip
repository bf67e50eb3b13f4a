// kf_volume.hip — TSDF volume initialisation and the projective integration sweep.
//
// Replaces d_initVoxelArray (Source/src/General.cu:27-50) and d_integrateVolume /
// VolumeIntegrator::Integrate (Source/src/VolumeIntegrator.cu:86-201).
//
// Layout: voxel (x,y,z) at index (x*n + y)*n + z (z fastest), one 8-byte {f, w}
// (FP32) or one 4-byte {half f, u16 w} (FP16) word per voxel, f = tsdf/T.
//
// The sweep is HBM-bound: it reads and writes only voxels that are updated.
//   1. column_spans: one thread per (x,y) column computes, in FP64, a conservative
//      z-interval outside which no voxel can pass the exact per-voxel test (the
//      column is a line in camera space, the frustum is convex).
//   2. integrate: one wavefront per column (waves stride over the columns) walks
//      its span 64 voxels at a time (one voxel per lane, coalesced), evaluates the
//      exact FP32 test per voxel (same expressions as the reference's per-voxel
//      body, in metres and with a z-forward camera), loads a voxel only if it is
//      updated and stores it only if its bits change.
#include "kf_device.hpp"
#include "kf_internal.hpp"

namespace kf {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kSweepMaxBlocks = 4096;  // 16384 waves: two full waves of residency on 256 CUs

struct SweepArgs {
  void* vox;
  uint32_t* spans;
  const float* raylen;
  const kf_pose* pose;
  unsigned long long* updated;
  uint32_t* mask;
  OccMap occ;  // brick occupancy map (occ.words NULL: none)
  int n;
  float h, ox, oy, oz;
  float fx, fy, cx, cy;
  int W, H;
  float dmin, dmax;
  float trunc, inv_trunc, maxw;
  float band;  // T + max(0.01 T, 0.1 mm): half-width of the band where the square root is taken
};

__global__ __launch_bounds__(256) void volume_fill(uint4* p, size_t n16, uint4 word) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride) p[i] = word;
}

// Conservative frustum span per column (FP64).  Margins: 0.05 px in the image
// plane and 1e-4 m in depth, each far above the FP32 error of the per-voxel path
// (< 3e-3 px, < 2e-6 m), plus one voxel on each side.
__global__ __launch_bounds__(256) void column_spans(SweepArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n * a.n) return;
  const int ix = c / a.n, iy = c - ix * a.n;
  const kf_posef P = a.pose->cw;
  const float wxf = a.ox + ((float)ix + 0.5f) * a.h;
  const float wyf = a.oy + ((float)iy + 0.5f) * a.h;
  const double wx = wxf, wy = wyf;
  const double wz0 = (double)a.oz + 0.5 * (double)a.h;
  const double h = a.h;
  double A[3], D[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    A[r] = (double)P.r[3 * r] * wx + (double)P.r[3 * r + 1] * wy + (double)P.r[3 * r + 2] * wz0 +
           (double)P.t[r];
    D[r] = (double)P.r[3 * r + 2] * h;
  }
  const double eps_px = 0.05, eps_z = 1e-4;
  double lo = 0.0, hi = (double)(a.n - 1);
  bool empty = false;
  auto clip = [&](double c0, double c1) {
    if (c1 > 0.0) {
      lo = fmax(lo, -c0 / c1);
    } else if (c1 < 0.0) {
      hi = fmin(hi, -c0 / c1);
    } else if (c0 < 0.0) {
      empty = true;
    }
  };
  clip(A[2] - ((double)a.dmin - eps_z), D[2]);
  clip(((double)a.dmax + eps_z) - A[2], -D[2]);
  const double ulo = 1.5 - eps_px, uhi = (double)a.W - 2.5 + eps_px;
  const double vlo = 1.5 - eps_px, vhi = (double)a.H - 2.5 + eps_px;
  const double fx = a.fx, fy = a.fy, cx = a.cx, cy = a.cy;
  clip(fx * A[0] + (cx - ulo) * A[2], fx * D[0] + (cx - ulo) * D[2]);
  clip((uhi - cx) * A[2] - fx * A[0], (uhi - cx) * D[2] - fx * D[0]);
  clip(fy * A[1] + (cy - vlo) * A[2], fy * D[1] + (cy - vlo) * D[2]);
  clip((vhi - cy) * A[2] - fy * A[1], (vhi - cy) * D[2] - fy * D[1]);
  uint32_t span = 0x0000ffffu;  // z1 < z0: empty
  if (!empty && lo <= hi + 2.0 && lo <= (double)a.n && hi >= -1.0) {
    const int z0 = (int)fmax(0.0, floor(lo) - 1.0);
    const int z1 = (int)fmin((double)(a.n - 1), ceil(hi) + 1.0);
    if (z0 <= z1) span = (uint32_t)z0 | ((uint32_t)z1 << 16);
  }
  a.spans[c] = span;
}

// Exact per-voxel test; returns true and the new normalised sample f if updated.
__device__ __forceinline__ bool voxel_sample(const SweepArgs& a, float ax, float ay, float az,
                                             const float* r, int k, float& f) {
  const float wz = a.oz + ((float)k + 0.5f) * a.h;
  const float px = ax + r[2] * wz;
  const float py = ay + r[5] * wz;
  const float pz = az + r[8] * wz;
  if (!(pz >= a.dmin && pz <= a.dmax)) return false;
  const float iz = 1.0f / pz;
  const float u = (a.fx * px + a.cx * pz) * iz;
  const float v = (a.fy * py + a.cy * pz) * iz;
  const float ru = round_half_up(u);
  const float rv = round_half_up(v);
  if (!(ru >= 2.0f && ru < (float)(a.W - 2) && rv >= 2.0f && rv < (float)(a.H - 2))) return false;
  const float rl = a.raylen[(int)rv * a.W + (int)ru];
  if (!(rl > 0.0f)) return false;
  const float d2 = (px * px + py * py) + pz * pz;
  // Outside |rl - dist| <= band the outcome follows from d2 alone: the margin
  // band - T >= max(0.01 T, 0.1 mm) is far above the FP32 error of d2, lo^2, hi^2,
  // dist and sdf (< 2 um for rays up to 10 m), so free space yields exactly f = 1
  // (sdf / T > 1 clamps to 1) and the far side is exactly "not updated" (sdf < -T),
  // as the full expression gives.
  const float lo = rl - a.band, hi = rl + a.band;
  if (lo > 0.0f && d2 < lo * lo) {
    f = 1.0f;
    return true;
  }
  if (d2 > hi * hi) return false;
  const float dist = sqrtf(d2);
  const float sdf = rl - dist;
  if (sdf < -a.trunc) return false;
  f = fmaxf(fminf(sdf * a.inv_trunc, 1.0f), -1.0f);
  return true;
}

template <int FMT>
__global__ __launch_bounds__(256) void integrate(SweepArgs a) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
  const kf_posef P = a.pose->cw;
  const int ncol = a.n * a.n;
  int count = 0;
  // Waves stride over the columns; neighbouring waves take neighbouring columns.
  for (int c = blockIdx.x * kWavesPerBlock + wave; c < ncol; c += gridDim.x * kWavesPerBlock) {
    const uint32_t span = a.spans[c];
    const int z0 = (int)(span & 0xffffu), z1 = (int)(span >> 16);
    if (z1 < z0) continue;
    const int ix = c / a.n, iy = c - ix * a.n;
    const float wx = a.ox + ((float)ix + 0.5f) * a.h;
    const float wy = a.oy + ((float)iy + 0.5f) * a.h;
    const float ax = (P.r[0] * wx + P.r[1] * wy) + P.t[0];
    const float ay = (P.r[3] * wx + P.r[4] * wy) + P.t[1];
    const float az = (P.r[6] * wx + P.r[7] * wy) + P.t[2];
    const size_t col = (size_t)c * (size_t)a.n;
    unsigned long long occ_bits = 0ull;  // z-bricks receiving f != 1 from this lane
    // One voxel per lane per step: a wave covers 64 consecutive z (256 B of FP16
    // or 512 B of FP32 voxels, fully coalesced), so few lanes idle at span ends.
    for (int zb = z0; zb <= z1; zb += kWave) {
      const int z = zb + lane;
      float fs = 0.0f;
      const bool upd = z <= z1 && voxel_sample(a, ax, ay, az, P.r, z, fs);
      if (upd) {
        ++count;
        const size_t idx = col + (size_t)z;
        if (FMT == KF_VOXEL_F16) {
          uint32_t* p = static_cast<uint32_t*>(a.vox) + idx;
          const uint32_t w = *p;
          const float fo = half_bits_to_float(w & 0xffffu);
          const float wo = (float)(w >> 16);
          // Free space (f_old = f = 1) averages to exactly 1: (w + 1) / (w + 1).
          const float fn = (fo == 1.0f && fs == 1.0f) ? 1.0f : (fo * wo + fs) / (wo + 1.0f);
          const float wn = fminf(wo + 1.0f, a.maxw);
          const uint32_t nw = float_to_half_bits(fn) | ((uint32_t)wn << 16);
          if (nw != w) *p = nw;  // saturated free space (f = 1, w = max) is left as it is
          if (a.occ.words && (nw & 0xffffu) != 0x3C00u) occ_bits |= occ_zbits(z);
        } else {
          float2* p = static_cast<float2*>(a.vox) + idx;
          const float2 v = *p;
          const float fn = (v.x == 1.0f && fs == 1.0f) ? 1.0f : (v.x * v.y + fs) / (v.y + 1.0f);
          const float wn = fminf(v.y + 1.0f, a.maxw);
          if (__float_as_uint(fn) != __float_as_uint(v.x) || __float_as_uint(wn) != __float_as_uint(v.y))
            *p = make_float2(fn, wn);
          if (a.occ.words && fn != 1.0f) occ_bits |= occ_zbits(z);
        }
      }
      if (a.mask) {  // 64 consecutive voxels -> up to three 32-bit mask words
        const unsigned long long m = __ballot(upd);
        if (m) {
          const size_t base = col + (size_t)zb;
          const int sh = (int)(base & 31);
          const size_t w0 = base >> 5;
          if (lane == 0) {
            const uint32_t v0 = (uint32_t)(m << sh);
            if (v0) atomicOr(&a.mask[w0], v0);
          } else if (lane == 1) {
            const uint32_t v1 = (uint32_t)(m >> (32 - sh));
            if (v1) atomicOr(&a.mask[w0 + 1], v1);
          } else if (lane == 2 && sh > 0) {
            const uint32_t v2 = (uint32_t)(m >> (64 - sh));
            if (v2) atomicOr(&a.mask[w0 + 2], v2);
          }
        }
      }
    }
    if (a.occ.words) occ_mark_column(a.occ, ix, iy, occ_bits, lane);
  }
  if (a.updated) {  // one device atomic per block (uniform branch)
    __shared__ int red[kWavesPerBlock];
    count = wave_sum_i_lane63(count);
    if (lane == 63) red[wave] = count;
    __syncthreads();
    if (threadIdx.x == 0) {
      const int tot = (red[0] + red[1]) + (red[2] + red[3]);
      if (tot) atomicAdd(a.updated, (unsigned long long)tot);
    }
  }
}

// Whole-volume occupancy statistics: one pass over every voxel with 16-byte
// loads, block reduction in LDS, one FP64 atomic per block and statistic.
template <int FMT>
__global__ __launch_bounds__(256) void volume_stats(const uint4* p, size_t n16, double* out) {
  double obs = 0.0, sw = 0.0, sf = 0.0, neg = 0.0;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride) {
    const uint4 q = p[i];
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
    if (FMT == KF_VOXEL_F16) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float wt = (float)(w[j] >> 16);
        if (wt > 0.0f) {
          const float f = half_bits_to_float(w[j] & 0xffffu);
          obs += 1.0;
          sw += wt;
          sf += f;
          neg += f < 0.0f ? 1.0 : 0.0;
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const float f = __uint_as_float(w[2 * j]), wt = __uint_as_float(w[2 * j + 1]);
        if (wt > 0.0f) {
          obs += 1.0;
          sw += wt;
          sf += f;
          neg += f < 0.0f ? 1.0 : 0.0;
        }
      }
    }
  }
  __shared__ double red[4][256];
  red[0][threadIdx.x] = obs;
  red[1][threadIdx.x] = sw;
  red[2][threadIdx.x] = sf;
  red[3][threadIdx.x] = neg;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 4) atomicAdd(&out[threadIdx.x], red[threadIdx.x][0]);
}

// Occupancy from scratch: one wavefront per voxel column (waves stride over the
// columns) reads the whole column with 16-byte loads and marks the z-bricks that
// hold f != 1.
template <int FMT>
__global__ __launch_bounds__(256) void occupancy_rebuild(const void* vox, int n, OccMap occ) {
  constexpr int V = (FMT == KF_VOXEL_F16) ? 4 : 2;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
  const int ncol = n * n;
  for (int c = blockIdx.x * kWavesPerBlock + wave; c < ncol; c += gridDim.x * kWavesPerBlock) {
    const int ix = c / n, iy = c - ix * n;
    const size_t col = (size_t)c * (size_t)n;
    unsigned long long bits = 0ull;
    for (int zb = lane * V; zb < n; zb += kWave * V) {
      const uint4 q = *reinterpret_cast<const uint4*>(static_cast<const uint32_t*>(vox) + (FMT == KF_VOXEL_F16 ? col + zb : 2 * (col + zb)));
      if (FMT == KF_VOXEL_F16) {
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if ((w[j] & 0xffffu) != 0x3C00u) bits |= occ_zbits(zb + j);
      } else {
        if (__uint_as_float(q.x) != 1.0f) bits |= occ_zbits(zb);
        if (__uint_as_float(q.z) != 1.0f) bits |= occ_zbits(zb + 1);
      }
    }
    occ_mark_column(occ, ix, iy, bits, lane);
  }
}

__global__ __launch_bounds__(256) void fill_u64(unsigned long long* p, int count, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

__global__ void zero_doubles(double* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0;
}

}  // namespace

kf_status launch_volume_stats(const kf_volume* v, double* d_out, hipStream_t s) {
  KF_CHECK_ARG(d_out != nullptr, "out is NULL");
  hipLaunchKernelGGL(zero_doubles, dim3(1), dim3(64), 0, s, d_out, 4);
  KF_TRY(check_launch("zero_doubles"));
  const size_t n16 = v->bytes / 16;
  const int blocks = (int)std::min<size_t>((n16 + 255) / 256, 4096);
  if (v->desc.format == KF_VOXEL_F16)
    hipLaunchKernelGGL(volume_stats<KF_VOXEL_F16>, dim3(blocks), dim3(256), 0, s,
                       static_cast<const uint4*>(v->d_voxels), n16, d_out);
  else
    hipLaunchKernelGGL(volume_stats<KF_VOXEL_F32>, dim3(blocks), dim3(256), 0, s,
                       static_cast<const uint4*>(v->d_voxels), n16, d_out);
  return check_launch("volume_stats");
}

OccMap occ_map(const kf_volume* v) {
  return OccMap{v->d_occ, v->d_occ2, v->occ_nb, v->occ_nsb};
}

// Both levels to zero (a fill kernel, so that graph capture stays valid).
kf_status clear_occupancy(const kf_volume* v, hipStream_t s) {
  const int words = v->occ_nb * v->occ_nb + (v->occ_nsb * v->occ_nsb + 1) / 2;  // the whole allocation
  hipLaunchKernelGGL(fill_u64, dim3(div_up(words, 256)), dim3(256), 0, s, v->d_occ, words, 0ull);
  return check_launch("fill_u64(occupancy)");
}

kf_status launch_volume_reset(kf_volume* v, hipStream_t s) {
  uint4 word;
  if (v->desc.format == KF_VOXEL_F16) {
    const uint32_t w = 0x3C00u;  // half(1.0), weight 0
    word = make_uint4(w, w, w, w);
  } else {
    const uint32_t one = 0x3F800000u;  // 1.0f, weight 0.0f
    word = make_uint4(one, 0u, one, 0u);
  }
  const size_t n16 = v->bytes / 16;
  const int blocks = (int)std::min<size_t>((n16 + 255) / 256, 8192);
  hipLaunchKernelGGL(volume_fill, dim3(blocks), dim3(256), 0, s, static_cast<uint4*>(v->d_voxels), n16,
                     word);
  KF_TRY(check_launch("volume_fill"));
  if (v->d_occ) KF_TRY(clear_occupancy(v, s));  // every voxel is free space: all bricks clear
  v->occ_stale = false;
  return KF_OK;
}

kf_status ensure_occupancy(const kf_volume* v, hipStream_t s) {
  if (!v->d_occ || !v->occ_stale) return KF_OK;
  KF_TRY(clear_occupancy(v, s));
  const int n = v->desc.n;
  const int blocks = std::min(div_up(n * n, kWavesPerBlock), kSweepMaxBlocks);
  if (v->desc.format == KF_VOXEL_F16)
    hipLaunchKernelGGL(occupancy_rebuild<KF_VOXEL_F16>, dim3(blocks), dim3(256), 0, s,
                       (const void*)v->d_voxels, n, occ_map(v));
  else
    hipLaunchKernelGGL(occupancy_rebuild<KF_VOXEL_F32>, dim3(blocks), dim3(256), 0, s,
                       (const void*)v->d_voxels, n, occ_map(v));
  KF_TRY(check_launch("occupancy_rebuild"));
  v->occ_stale = false;
  return KF_OK;
}

kf_status launch_integrate(kf_volume* v, const kf_intrinsics& k0, const kf_integrate_params& p,
                           const float* d_raylen, const kf_pose* d_pose, unsigned long long* d_updated,
                           uint32_t* d_mask, hipStream_t s, hipEvent_t between) {
  KF_CHECK_ARG(d_raylen && d_pose, "raylen/pose is NULL");
  KF_CHECK_ARG(p.min_depth_m >= 0.001f && p.max_depth_m > p.min_depth_m, "bad depth range");
  KF_CHECK_ARG(k0.width > 4 && k0.height > 4, "image too small");
  SweepArgs a;
  const kf_volume_desc& d = v->desc;
  a.vox = v->d_voxels;
  a.spans = v->d_spans;
  a.raylen = d_raylen;
  a.pose = d_pose;
  a.updated = d_updated;
  a.mask = d_mask;
  a.occ = occ_map(v);
  a.n = d.n;
  a.h = d.extent_m / (float)d.n;
  a.ox = d.origin_m[0];
  a.oy = d.origin_m[1];
  a.oz = d.origin_m[2];
  a.fx = k0.fx;
  a.fy = k0.fy;
  a.cx = k0.cx;
  a.cy = k0.cy;
  a.W = k0.width;
  a.H = k0.height;
  a.dmin = p.min_depth_m;
  a.dmax = p.max_depth_m;
  a.trunc = d.trunc_m;
  a.inv_trunc = 1.0f / d.trunc_m;
  a.band = d.trunc_m + std::max(0.01f * d.trunc_m, 1e-4f);
  a.maxw = d.max_weight;
  const int cols = d.n * d.n;
  hipLaunchKernelGGL(column_spans, dim3(div_up(cols, 256)), dim3(256), 0, s, a);
  KF_TRY(check_launch("column_spans"));
  if (between) KF_HIP(hipEventRecord(between, s));
  const int blocks = std::min(div_up(cols, kWavesPerBlock), kSweepMaxBlocks);
  if (d.format == KF_VOXEL_F16)
    hipLaunchKernelGGL(integrate<KF_VOXEL_F16>, dim3(blocks), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(integrate<KF_VOXEL_F32>, dim3(blocks), dim3(256), 0, s, a);
  return check_launch("integrate");
}

}  // namespace kf
