// kf_prior.hip — the learned pose prior: CNN inference and training on gfx950.
//
// No counterpart in the reference (its ICP starts from the previous pose,
// Source/src/ICP.cpp:94).  Network: four 3x3 stride-2 pad-1 convolutions
// (2->32->64->64->128, ReLU) on two 160x120 depth maps, global average pooling,
// concatenation with the previous frame's twist, FC 134->256 (ReLU), FC 256->6,
// plus the previous twist (skip connection: the layers predict the change of motion).
//
// Layout: activations NHWC (pixel-major, channel fastest), FP32; GEMM operands
// are BF16 (round to nearest even), accumulation FP32.  A convolution is
// im2col (rows = output pixels, columns = (ky, kx, cin)) followed by one GEMM
// against the weight matrix (rows = cout); its backward is two GEMMs (weight
// gradient with deterministic split-K over pixels, column gradient) and a
// gather col2im.  The GEMM is one 16x16 output tile per wavefront on
// v_mfma_f32_16x16x16bf16_1k, four wavefronts per 256-thread block.
#include "kf_device.hpp"
#include "kf_internal.hpp"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace kf {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kH0 = KF_PRIOR_HEIGHT, kW0 = KF_PRIOR_WIDTH;
constexpr int kConvs = 4;
constexpr int kFeat = 128, kTwist = 6, kFc1In = kFeat + kTwist, kHidden = 256;

struct ConvShape {
  int cin, cout, hin, win, hout, wout, kc;
};
constexpr ConvShape kConv[kConvs] = {
    {2, 32, 120, 160, 60, 80, 18},
    {32, 64, 60, 80, 30, 40, 288},
    {64, 64, 30, 40, 15, 20, 576},
    {64, 128, 15, 20, 8, 10, 576},
};
constexpr int kPool = 8 * 10;

// Parameter blob offsets (floats).
struct ParamLayout {
  long w[6], b[6], total;
  ParamLayout() {
    long o = 0;
    for (int l = 0; l < kConvs; ++l) {
      w[l] = o;
      o += (long)kConv[l].cout * kConv[l].kc;
      b[l] = o;
      o += kConv[l].cout;
    }
    w[4] = o;
    o += (long)kHidden * kFc1In;
    b[4] = o;
    o += kHidden;
    w[5] = o;
    o += (long)kTwist * kHidden;
    b[5] = o;
    o += kTwist;
    total = o;
  }
};
const ParamLayout& layout() {
  static const ParamLayout L;
  return L;
}

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// ---------------------------------------------------------------------------
// GEMM: C[m][n] (+)= sum_k A(m,k) B(k,n); A(m,k) = A[m*sam + k*sak], B(k,n) =
// B[k*sbk + n*sbn] (BF16 bits).  Split s of gridDim.z covers k in
// [s*kps, (s+1)*kps) and, when Cpart is set, writes its own partial plane.
struct GemmArgs {
  const uint16_t* A;
  long sam, sak;
  const uint16_t* B;
  long sbk, sbn;
  float* C;
  long ldc;
  float* Cpart;
  int M, N, K, kps, accumulate;
  // Epilogue (no split-K): v = acc [+ bias[n], ReLU] [+ addend[m * N + n]]; C = v and,
  // with Cb, the BF16 copy (RNE) at the same position.
  const float* bias;
  int relu;
  const float* addend;
  uint16_t* Cb;
  // Convolution operand (convX set): A(m, k) is the im2col element of a 3x3 /
  // stride-2 / pad-1 convolution read from the BF16 NHWC input convX instead.
  const uint16_t* convX;
  int conv_cin, conv_hin, conv_win, conv_hout, conv_wout;
};

// Each 256-thread block computes a 32x32 tile of C (four 16x16 wave tiles on
// v_mfma_f32_16x16x16bf16_1k).  K is staged through LDS in chunks of kKc: the block
// loads the 32 x chunk slices of A and B with coalesced accesses (adjacent threads
// take adjacent addresses along whichever of K or M/N is unit-stride), then runs
// the chunk's MFMAs from LDS.  The MFMA sequence over k (steps of 16, zeros beyond
// K) is the same as a direct k loop, so results do not depend on the chunking.
constexpr int kKc = 576;        // whole K of every convolution layer fits one chunk
constexpr int kPitch = kKc + 8;  // LDS row pitch (elements): rows start on different banks
constexpr int kVecPerThread = (32 * kKc / 8 + 255) / 256;  // 16-byte vectors per thread and operand

// Stages a 32 x kpad operand slice into T[32][kPitch] (row = m or n, column = k)
// from X(row, k) = X[row * s_row + k * s_k], zero outside rows < nrows, k < kn.
__device__ __forceinline__ void stage(uint16_t (*T)[kPitch], const uint16_t* X, long s_row, long s_k, int row0,
                                      int nrows, long k0, int kn, int kpad) {
  const int tid = threadIdx.x;
  const bool vec = s_k == 1 && (s_row & 7) == 0 && (k0 & 7) == 0 && (kn & 7) == 0 &&
                   ((reinterpret_cast<uintptr_t>(X) & 15) == 0);
  if (vec) {  // 16-byte loads of 8 consecutive k, all issued before any LDS store
    const int vpr = kpad >> 3;  // vectors per row (the last one may be k padding)
    const int total = 32 * vpr;
    uint4 v[kVecPerThread];
#pragma unroll
    for (int u = 0; u < kVecPerThread; ++u) {
      const int e = tid + 256 * u;
      const int r = e / vpr, c = e - (e / vpr) * vpr;
      const bool in = e < total && row0 + r < nrows && 8 * c < kn;
      v[u] = in ? *reinterpret_cast<const uint4*>(X + (long)(row0 + r) * s_row + k0 + 8 * c) : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < kVecPerThread; ++u) {
      const int e = tid + 256 * u;
      if (e < total) {
        const int r = e / vpr, c = e - (e / vpr) * vpr;
        *reinterpret_cast<uint4*>(&T[r][8 * c]) = v[u];
      }
    }
  } else if (s_k == 1) {  // K contiguous: thread e takes (row, k) = (e / kpad, e % kpad)
    const int total = 32 * kpad;
    for (int e0 = tid; e0 < total; e0 += 256 * 4) {
      uint16_t v[4];
      int rr[4], kk[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + 256 * u;
        rr[u] = e / kpad;
        kk[u] = e - rr[u] * kpad;
        const bool in = e < total && row0 + rr[u] < nrows && kk[u] < kn;
        v[u] = in ? X[(long)(row0 + rr[u]) * s_row + (k0 + kk[u])] : (uint16_t)0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (e0 + 256 * u < total) T[rr[u]][kk[u]] = v[u];
    }
  } else {  // rows contiguous (or general): thread e takes (k, row) = (e / 32, e % 32)
    const int total = 32 * kpad;
    for (int e0 = tid; e0 < total; e0 += 256 * 4) {
      uint16_t v[4];
      int rr[4], kk[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + 256 * u;
        kk[u] = e >> 5;
        rr[u] = e & 31;
        const bool in = e < total && row0 + rr[u] < nrows && kk[u] < kn;
        v[u] = in ? X[(long)(row0 + rr[u]) * s_row + (k0 + kk[u]) * s_k] : (uint16_t)0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (e0 + 256 * u < total) T[rr[u]][kk[u]] = v[u];
    }
  }
}

// Stages A rows [row0, row0 + 32) of a convolution directly from its BF16 NHWC
// input: k = (ky * 3 + kx) * cin + ci with cin a multiple of 8, so 8 consecutive k
// are 8 channels of one tap (one 16-byte load, zero outside the image); the values
// are those an im2col of the same input would hold.
__device__ __forceinline__ void stage_conv(uint16_t (*T)[kPitch], const GemmArgs& g, int row0, int kn, int kpad) {
  const int tid = threadIdx.x;
  const int vpr = kpad >> 3;
  const int total = 32 * vpr;
  const int pix = g.conv_hout * g.conv_wout;
  uint4 v[kVecPerThread];
#pragma unroll
  for (int u = 0; u < kVecPerThread; ++u) {
    const int e = tid + 256 * u;
    const int r = e / vpr, c = e - (e / vpr) * vpr;
    const int m = row0 + r;
    bool in = e < total && m < g.M && 8 * c < kn;
    long off = 0;
    if (in) {
      const int n = m / pix, q = m - n * pix;
      const int oy = q / g.conv_wout, ox = q - oy * g.conv_wout;
      const int tap = (8 * c) / g.conv_cin, ci = 8 * c - tap * g.conv_cin;
      const int ky = tap / 3, kx = tap - 3 * ky;
      const int iy = 2 * oy + ky - 1, ix = 2 * ox + kx - 1;
      in = iy >= 0 && iy < g.conv_hin && ix >= 0 && ix < g.conv_win;
      off = (((long)n * g.conv_hin + iy) * g.conv_win + ix) * g.conv_cin + ci;
    }
    v[u] = in ? *reinterpret_cast<const uint4*>(g.convX + off) : make_uint4(0, 0, 0, 0);
  }
#pragma unroll
  for (int u = 0; u < kVecPerThread; ++u) {
    const int e = tid + 256 * u;
    if (e < total) {
      const int r = e / vpr, c = e - (e / vpr) * vpr;
      *reinterpret_cast<uint4*>(&T[r][8 * c]) = v[u];
    }
  }
}

__global__ __launch_bounds__(256) void gemm_bf16(GemmArgs g) {
  __shared__ __align__(16) uint16_t As[32][kPitch];
  __shared__ __align__(16) uint16_t Bt[32][kPitch];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int mt = blockIdx.y * 32, nt = blockIdx.x * 32;
  const int wm = (wave >> 1) * 16, wn = (wave & 1) * 16;
  const int kbeg = blockIdx.z * g.kps;
  const int kend = min(g.K, kbeg + g.kps);
  const int r = lane & 15, kq = (lane >> 4) * 4;
  v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int kc0 = kbeg; kc0 < kend; kc0 += kKc) {
    const int kn = min(kKc, kend - kc0);
    const int kpad = (kn + 15) & ~15;
    if (g.convX)  // uniform; the whole K is one chunk for every convolution
      stage_conv(As, g, mt, kn, kpad);
    else
      stage(As, g.A, g.sam, g.sak, mt, g.M, kc0, kn, kpad);
    stage(Bt, g.B, g.sbn, g.sbk, nt, g.N, kc0, kn, kpad);
    __syncthreads();
    for (int kk = 0; kk < kpad; kk += 16) {
      v4s a, b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a[j] = (short)As[wm + r][kk + kq + j];
        b[j] = (short)Bt[wn + r][kk + kq + j];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int m0 = mt + wm, n0 = nt + wn;
  if (m0 >= g.M || n0 >= g.N) return;  // wave-uniform
  const int cn = n0 + (lane & 15);
  if (cn >= g.N) return;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int cm = m0 + (lane >> 4) * 4 + j;
    if (cm >= g.M) continue;
    if (g.Cpart) {
      g.Cpart[((long)blockIdx.z * g.M + cm) * g.N + cn] = acc[j];
    } else {
      float* c = g.C + (long)cm * g.ldc + cn;
      float v = g.accumulate ? *c + acc[j] : acc[j];
      if (g.bias) {
        v = v + g.bias[cn];
        if (g.relu) v = fmaxf(v, 0.0f);
      }
      if (g.addend) v = v + g.addend[(long)cm * g.N + cn];
      *c = v;
      if (g.Cb) g.Cb[(long)cm * g.ldc + cn] = bf16_rne(v);
    }
  }
}

__global__ void splitk_reduce(const float* part, int splits, long MN, int N, float* C, long ldc) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MN) return;
  float s = 0.0f;
  for (int z = 0; z < splits; ++z) s = s + part[(long)z * MN + i];
  C[(i / N) * ldc + (i % N)] = s;
}

// ---------------------------------------------------------------------------
// im2col.  Layer 1 reads the two depth planes of each pair directly.
__global__ void im2col_pairs(const float* frames, const int32_t* pairs, int batch, uint16_t* col, int kcp) {
  const ConvShape c = kConv[0];
  const long M = (long)batch * c.hout * c.wout;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * kcp) return;
  const long m = i / kcp;
  const int k = (int)(i - m * kcp);
  float v = 0.0f;
  if (k < c.kc) {
    const int n = (int)(m / (c.hout * c.wout));
    const int p = (int)(m - (long)n * c.hout * c.wout);
    const int oy = p / c.wout, ox = p - oy * c.wout;
    const int ky = k / (3 * c.cin), kx = (k / c.cin) % 3, ci = k % c.cin;
    const int iy = 2 * oy + ky - 1, ix = 2 * ox + kx - 1;
    if (iy >= 0 && iy < c.hin && ix >= 0 && ix < c.win)
      v = frames[((long)pairs[2 * n + ci] * kH0 + iy) * kW0 + ix];
  }
  col[i] = bf16_rne(v);
}

__global__ void im2col_nhwc(const float* X, int layer, int batch, uint16_t* col) {
  const ConvShape c = kConv[layer];
  const long M = (long)batch * c.hout * c.wout;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * c.kc) return;
  const long m = i / c.kc;
  const int k = (int)(i - m * c.kc);
  const int n = (int)(m / (c.hout * c.wout));
  const int p = (int)(m - (long)n * c.hout * c.wout);
  const int oy = p / c.wout, ox = p - oy * c.wout;
  const int ky = k / (3 * c.cin), kx = (k / c.cin) % 3, ci = k % c.cin;
  const int iy = 2 * oy + ky - 1, ix = 2 * ox + kx - 1;
  float v = 0.0f;
  if (iy >= 0 && iy < c.hin && ix >= 0 && ix < c.win) v = X[(((long)n * c.hin + iy) * c.win + ix) * c.cin + ci];
  col[i] = bf16_rne(v);
}

// dX[n][iy][ix][ci] = sum over (ky,kx) of dcol[(n,oy,ox)][(ky,kx,ci)], times the
// ReLU mask of the forward activation X (the previous layer's output).
__global__ void col2im_relu(const float* dcol, const float* X, int layer, int batch, float* dX) {
  const ConvShape c = kConv[layer];
  const long total = (long)batch * c.hin * c.win * c.cin;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int ci = (int)(i % c.cin);
  const long q = i / c.cin;
  const int ix = (int)(q % c.win);
  const long q2 = q / c.win;
  const int iy = (int)(q2 % c.hin);
  const int n = (int)(q2 / c.hin);
  float s = 0.0f;
  for (int ky = 0; ky < 3; ++ky) {
    const int ty = iy + 1 - ky;
    if (ty < 0 || (ty & 1)) continue;
    const int oy = ty >> 1;
    if (oy >= c.hout) continue;
    for (int kx = 0; kx < 3; ++kx) {
      const int tx = ix + 1 - kx;
      if (tx < 0 || (tx & 1)) continue;
      const int ox = tx >> 1;
      if (ox >= c.wout) continue;
      const long m = ((long)n * c.hout + oy) * c.wout + ox;
      s = s + dcol[m * c.kc + (ky * 3 + kx) * c.cin + ci];
    }
  }
  dX[i] = X[i] > 0.0f ? s : 0.0f;
}

__global__ void to_bf16(const float* x, uint16_t* y, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = bf16_rne(x[i]);
}

// feat[n][0..127] = mean over the 80 pooled pixels of Y4; [128..133] = prev twist.
__global__ void pool_concat(const float* Y4, const float* prev, int batch, float* X5, uint16_t* X5b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch * kFc1In) return;
  const int n = i / kFc1In, c = i - n * kFc1In;
  float v;
  if (c < kFeat) {
    float s = 0.0f;
    for (int p = 0; p < kPool; ++p) s = s + Y4[((long)n * kPool + p) * kFeat + c];
    v = s * (1.0f / (float)kPool);
  } else {
    v = prev[n * kTwist + (c - kFeat)];
  }
  X5[i] = v;
  X5b[i] = bf16_rne(v);
}

// dY4[n][p][c] = dX5[n][c] / 80 where Y4 > 0.
__global__ void pool_backward(const float* dX5, const float* Y4, int batch, float* dY4) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)batch * kPool * kFeat) return;
  const int c = (int)(i % kFeat);
  const int n = (int)(i / ((long)kPool * kFeat));
  dY4[i] = Y4[i] > 0.0f ? dX5[n * kFc1In + c] * (1.0f / (float)kPool) : 0.0f;
}

// Loss mean((O - T)^2) and dO = 2 (O - T) / (6 B); one block.
__global__ void mse_loss(const float* O, const float* T, int batch, float* dO, float* loss) {
  __shared__ float red[256];
  const int n = batch * kTwist;
  const float inv = 1.0f / (float)n;
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float d = O[i] - T[i];
    dO[i] = 2.0f * d * inv;
    s = s + d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) *loss = red[0] * inv;
}

// db[c] = sum_m dY[m][c]; one block per channel, fixed reduction order.
__global__ void bias_grad(const float* dY, long M, int C, float* db) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float s = 0.0f;
  for (long m = threadIdx.x; m < M; m += blockDim.x) s = s + dY[m * C + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = red[0];
}

__global__ void relu_mask(float* dY, const float* Y, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(Y[i] > 0.0f)) dY[i] = 0.0f;
}

__global__ void adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                     float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline unsigned blocks_for(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace
}  // namespace kf

struct kf_prior {
  int max_batch = 0;
  long nparams = 0;
  float* params = nullptr;   // FP32 master
  uint16_t* wbf = nullptr;   // BF16 copy
  float* grad = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  long step = 0;
  // activations (FP32, NHWC) and BF16 im2col buffers, sized for max_batch
  float* Y[kf::kConvs] = {};
  uint16_t* col[kf::kConvs] = {};
  float* X5 = nullptr;
  uint16_t* X5b = nullptr;
  float* Hd = nullptr;
  uint16_t* Hb = nullptr;
  float* O = nullptr;
  float* dO = nullptr;
  uint16_t* dOb = nullptr;
  float* dH = nullptr;
  uint16_t* dHb = nullptr;
  float* dX5 = nullptr;
  float* dY = nullptr;    // largest activation gradient
  float* dY2 = nullptr;   // second buffer (ping-pong)
  uint16_t* dYb = nullptr;
  float* dcol = nullptr;
  float* part = nullptr;  // split-K partial planes
  long part_floats = 0;
  float* loss_tmp = nullptr;
};

namespace kf {
namespace {

constexpr int kSplitMaxK = 4096;  // pixels per split in weight-gradient GEMMs

struct Epilogue {
  const float* bias = nullptr;
  int relu = 0;
  const float* addend = nullptr;
  uint16_t* Cb = nullptr;
};

kf_status gemm(hipStream_t s, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, float* C,
               long ldc, int M, int N, int K, bool accumulate, float* part = nullptr, long part_floats = 0,
               const Epilogue& ep = Epilogue(), const ConvShape* conv = nullptr, const uint16_t* convX = nullptr) {
  GemmArgs g;
  g.convX = convX;
  g.conv_cin = conv ? conv->cin : 0;
  g.conv_hin = conv ? conv->hin : 0;
  g.conv_win = conv ? conv->win : 0;
  g.conv_hout = conv ? conv->hout : 0;
  g.conv_wout = conv ? conv->wout : 0;
  if (convX) KF_CHECK_ARG(conv->cin % 8 == 0 && K == conv->kc && K <= kKc, "implicit convolution shape");
  g.bias = ep.bias;
  g.relu = ep.relu;
  g.addend = ep.addend;
  g.Cb = ep.Cb;
  g.A = A;
  g.sam = sam;
  g.sak = sak;
  g.B = B;
  g.sbk = sbk;
  g.sbn = sbn;
  g.C = C;
  g.ldc = ldc;
  g.M = M;
  g.N = N;
  g.K = K;
  g.accumulate = accumulate ? 1 : 0;
  int splits = 1;
  if (part && K > kSplitMaxK) {
    splits = (K + kSplitMaxK - 1) / kSplitMaxK;
    while ((long)splits * M * N > part_floats && splits > 1) splits = (splits + 1) / 2;
  }
  g.kps = splits > 1 ? (((K + splits - 1) / splits + 15) / 16) * 16 : ((K + 15) / 16) * 16;
  if (splits > 1) splits = (K + g.kps - 1) / g.kps;
  g.Cpart = splits > 1 ? part : nullptr;
  dim3 grid((N + 31) / 32, (M + 31) / 32, splits);
  hipLaunchKernelGGL(gemm_bf16, grid, dim3(256), 0, s, g);
  KF_TRY(check_launch("gemm_bf16"));
  if (splits > 1) {
    KF_CHECK_ARG(!accumulate && !ep.bias && !ep.addend && !ep.Cb, "split-K GEMM has no epilogue");
    const long MN = (long)M * N;
    hipLaunchKernelGGL(splitk_reduce, dim3(blocks_for(MN)), dim3(256), 0, s, (const float*)part, splits, MN, N, C,
                       ldc);
    KF_TRY(check_launch("splitk_reduce"));
  }
  return KF_OK;
}

const ConvShape& conv(int l) { return kConv[l]; }
int kcp(int l) { return l == 0 ? 32 : kConv[l].kc; }  // im2col row length (layer 1 padded to 32)

// BF16 copy of the weights; every call that changes the FP32 parameters refreshes
// it, so inference (forward) never converts.
kf_status sync_weights(kf_prior* P, hipStream_t s) {
  hipLaunchKernelGGL(to_bf16, dim3(blocks_for(P->nparams)), dim3(256), 0, s, (const float*)P->params, P->wbf,
                     P->nparams);
  return check_launch("to_bf16(weights)");
}

// out: where the twists go (P->O when NULL; training reads P->O).  cols: the
// im2col matrices of every layer are materialised (training: the weight gradients
// read them); without, layers 2-4 read their input directly in the GEMM (the
// previous layer's epilogue leaves its BF16 copy in the next layer's im2col
// buffer), with the same A values and hence the same outputs.
kf_status forward(kf_prior* P, const float* frames, const int32_t* pairs, const float* prev, int B, hipStream_t s,
                  float* out = nullptr, bool cols = true) {
  const ParamLayout& L = layout();
  for (int l = 0; l < kConvs; ++l) {
    const ConvShape& c = conv(l);
    const long M = (long)B * c.hout * c.wout;
    if (l == 0) {
      hipLaunchKernelGGL(im2col_pairs, dim3(blocks_for(M * kcp(0))), dim3(256), 0, s, frames, pairs, B, P->col[0],
                         kcp(0));
      KF_TRY(check_launch("im2col_pairs"));
    } else if (cols) {
      hipLaunchKernelGGL(im2col_nhwc, dim3(blocks_for(M * c.kc)), dim3(256), 0, s, (const float*)P->Y[l - 1], l, B,
                         P->col[l]);
      KF_TRY(check_launch("im2col_nhwc"));
    }
    // Y[m][co] = sum_k col[m][k] W[co][k]; the layer-1 padding columns are zero in col.
    // Y = ReLU(col W^T + b): bias and ReLU in the GEMM epilogue
    Epilogue ep;
    ep.bias = (const float*)P->params + L.b[l];
    ep.relu = 1;
    if (!cols && l + 1 < kConvs) ep.Cb = P->col[l + 1];  // BF16 NHWC copy, the next layer's input
    const bool implicit = !cols && l > 0;
    KF_TRY(gemm(s, P->col[l], kcp(l), 1, P->wbf + L.w[l], 1, c.kc, P->Y[l], c.cout, (int)M, c.cout, c.kc, false,
                nullptr, 0, ep, implicit ? &c : nullptr, implicit ? P->col[l] : nullptr));
  }
  hipLaunchKernelGGL(pool_concat, dim3(blocks_for((long)B * kFc1In)), dim3(256), 0, s, (const float*)P->Y[3], prev, B,
                     P->X5, P->X5b);
  KF_TRY(check_launch("pool_concat"));
  {  // FC1: H = ReLU(X5 W5^T + b5), also as BF16 for FC2
    Epilogue ep;
    ep.bias = (const float*)P->params + L.b[4];
    ep.relu = 1;
    ep.Cb = P->Hb;
    KF_TRY(gemm(s, P->X5b, kFc1In, 1, P->wbf + L.w[4], 1, kFc1In, P->Hd, kHidden, B, kHidden, kFc1In, false, nullptr,
                0, ep));
  }
  {  // FC2 plus the skip connection: O = (H W6^T + b6) + previous twist
    Epilogue ep;
    ep.bias = (const float*)P->params + L.b[5];
    ep.addend = prev;
    KF_TRY(gemm(s, P->Hb, kHidden, 1, P->wbf + L.w[5], 1, kHidden, out ? out : P->O, kTwist, B, kTwist, kHidden,
                false, nullptr, 0, ep));
  }
  return KF_OK;
}

kf_status backward(kf_prior* P, const float* target, int B, float* loss, hipStream_t s) {
  const ParamLayout& L = layout();
  float* G = P->grad;
  hipLaunchKernelGGL(mse_loss, dim3(1), dim3(256), 0, s, (const float*)P->O, target, B, P->dO, loss);
  hipLaunchKernelGGL(to_bf16, dim3(blocks_for((long)B * kTwist)), dim3(256), 0, s, (const float*)P->dO, P->dOb,
                     (long)B * kTwist);
  KF_TRY(check_launch("mse_loss"));
  // FC2: dW6[j][k] = sum_n dO[n][j] H[n][k]; db6; dH = dO W6, masked by H > 0.
  KF_TRY(gemm(s, P->dOb, 1, kTwist, P->Hb, kHidden, 1, G + L.w[5], kHidden, kTwist, kHidden, B, false));
  hipLaunchKernelGGL(bias_grad, dim3(kTwist), dim3(256), 0, s, (const float*)P->dO, (long)B, kTwist, G + L.b[5]);
  KF_TRY(gemm(s, P->dOb, kTwist, 1, P->wbf + L.w[5], kHidden, 1, P->dH, kHidden, B, kHidden, kTwist, false));
  hipLaunchKernelGGL(relu_mask, dim3(blocks_for((long)B * kHidden)), dim3(256), 0, s, P->dH, (const float*)P->Hd,
                     (long)B * kHidden);
  hipLaunchKernelGGL(to_bf16, dim3(blocks_for((long)B * kHidden)), dim3(256), 0, s, (const float*)P->dH, P->dHb,
                     (long)B * kHidden);
  // FC1: dW5[j][k] = sum_n dH[n][j] X5[n][k]; db5; dX5 (features only) = dH W5.
  KF_TRY(gemm(s, P->dHb, 1, kHidden, P->X5b, kFc1In, 1, G + L.w[4], kFc1In, kHidden, kFc1In, B, false));
  hipLaunchKernelGGL(bias_grad, dim3(kHidden), dim3(256), 0, s, (const float*)P->dH, (long)B, kHidden, G + L.b[4]);
  KF_TRY(gemm(s, P->dHb, kHidden, 1, P->wbf + L.w[4], kFc1In, 1, P->dX5, kFc1In, B, kFeat, kHidden, false));
  // Pool: dY4.
  float* dY = P->dY;
  float* dYn = P->dY2;
  hipLaunchKernelGGL(pool_backward, dim3(blocks_for((long)B * kPool * kFeat)), dim3(256), 0, s,
                     (const float*)P->dX5, (const float*)P->Y[3], B, dY);
  KF_TRY(check_launch("pool_backward"));
  for (int l = kConvs - 1; l >= 0; --l) {
    const ConvShape& c = conv(l);
    const long M = (long)B * c.hout * c.wout;
    hipLaunchKernelGGL(to_bf16, dim3(blocks_for(M * c.cout)), dim3(256), 0, s, (const float*)dY, P->dYb,
                       M * c.cout);
    // dW[co][k] = sum_m dY[m][co] col[m][k]   (split-K over pixels, deterministic)
    KF_TRY(gemm(s, P->dYb, 1, c.cout, P->col[l], kcp(l), 1, G + L.w[l], c.kc, c.cout, c.kc, (int)M, false, P->part,
                P->part_floats));
    hipLaunchKernelGGL(bias_grad, dim3(c.cout), dim3(256), 0, s, (const float*)dY, M, c.cout, G + L.b[l]);
    KF_TRY(check_launch("bias_grad"));
    if (l == 0) break;
    // dcol[m][k] = sum_co dY[m][co] W[co][k]; dX = col2im(dcol) masked by Y[l-1] > 0.
    KF_TRY(gemm(s, P->dYb, c.cout, 1, P->wbf + L.w[l], c.kc, 1, P->dcol, c.kc, (int)M, c.kc, c.cout, false));
    const long nin = (long)B * c.hin * c.win * c.cin;
    hipLaunchKernelGGL(col2im_relu, dim3(blocks_for(nin)), dim3(256), 0, s, (const float*)P->dcol,
                       (const float*)P->Y[l - 1], l, B, dYn);
    KF_TRY(check_launch("col2im_relu"));
    std::swap(dY, dYn);
  }
  return KF_OK;
}

kf_status check_batch(const kf_prior* P, const float* frames, const int32_t* pairs, const float* prev, int B) {
  KF_CHECK_ARG(P && frames && pairs && prev, "NULL argument");
  KF_CHECK_ARG(B >= 1 && B <= P->max_batch, "batch must be 1..max_batch");
  return KF_OK;
}

}  // namespace
}  // namespace kf

using namespace kf;

extern "C" int64_t kf_prior_param_count(void) { return layout().total; }

extern "C" kf_status kf_prior_destroy(kf_prior* P) {
  if (!P) return KF_OK;
  void* ptrs[] = {P->params, P->wbf, P->grad, P->adam_m, P->adam_v, P->X5, P->X5b, P->Hd, P->Hb, P->O, P->dO,
                  P->dOb, P->dH, P->dHb, P->dX5, P->dY, P->dY2, P->dYb, P->dcol, P->part, P->loss_tmp};
  for (void* q : ptrs) (void)hipFree(q);
  for (int l = 0; l < kConvs; ++l) {
    (void)hipFree(P->Y[l]);
    (void)hipFree(P->col[l]);
  }
  delete P;
  return KF_OK;
}

extern "C" kf_status kf_prior_create(int32_t max_batch, kf_prior** out) {
  KF_CHECK_ARG(out != nullptr, "out is NULL");
  KF_CHECK_ARG(max_batch >= 1 && max_batch <= 4096, "max_batch must be 1..4096");
  kf_prior* P = new kf_prior();
  P->max_batch = max_batch;
  P->nparams = layout().total;
  const long B = max_batch;
  hipError_t e = hipSuccess;
  auto al = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  al((void**)&P->params, P->nparams * 4);
  al((void**)&P->wbf, P->nparams * 2);
  al((void**)&P->grad, P->nparams * 4);
  al((void**)&P->adam_m, P->nparams * 4);
  al((void**)&P->adam_v, P->nparams * 4);
  long maxact = 0, maxdcol = 0;
  for (int l = 0; l < kConvs; ++l) {
    const ConvShape& c = kConv[l];
    const long M = B * c.hout * c.wout;
    al((void**)&P->Y[l], M * c.cout * 4);
    al((void**)&P->col[l], M * kcp(l) * 2);
    maxact = std::max(maxact, M * c.cout);
    maxact = std::max(maxact, B * c.hin * c.win * c.cin);
    if (l > 0) maxdcol = std::max(maxdcol, M * c.kc);
  }
  al((void**)&P->X5, B * kFc1In * 4);
  al((void**)&P->X5b, B * kFc1In * 2);
  al((void**)&P->Hd, B * kHidden * 4);
  al((void**)&P->Hb, B * kHidden * 2);
  al((void**)&P->O, B * kTwist * 4);
  al((void**)&P->dO, B * kTwist * 4);
  al((void**)&P->dOb, B * kTwist * 2);
  al((void**)&P->dH, B * kHidden * 4);
  al((void**)&P->dHb, B * kHidden * 2);
  al((void**)&P->dX5, B * kFc1In * 4);
  al((void**)&P->dY, maxact * 4);
  al((void**)&P->dY2, maxact * 4);
  al((void**)&P->dYb, maxact * 2);
  al((void**)&P->dcol, maxdcol * 4);
  P->part_floats = 64L * 128 * 576;  // up to 64 splits of the largest weight matrix
  al((void**)&P->part, P->part_floats * 4);
  al((void**)&P->loss_tmp, 16);
  if (e == hipSuccess) e = hipMemset(P->adam_m, 0, P->nparams * 4);
  if (e == hipSuccess) e = hipMemset(P->adam_v, 0, P->nparams * 4);
  if (e == hipSuccess) e = hipMemset(P->params, 0, P->nparams * 4);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    kf_prior_destroy(P);
    return hip_fail(e, "prior allocation");
  }
  *out = P;
  return KF_OK;
}

extern "C" kf_status kf_prior_set_params(kf_prior* P, const float* h, int64_t count) {
  KF_CHECK_ARG(P && h, "NULL argument");
  KF_CHECK_ARG(count == P->nparams, "parameter count mismatch");
  KF_HIP(hipMemcpy(P->params, h, count * 4, hipMemcpyHostToDevice));
  KF_TRY(sync_weights(P, nullptr));
  KF_HIP(hipDeviceSynchronize());
  return KF_OK;
}

extern "C" kf_status kf_prior_get_params(kf_prior* P, float* h, int64_t count) {
  KF_CHECK_ARG(P && h, "NULL argument");
  KF_CHECK_ARG(count == P->nparams, "parameter count mismatch");
  KF_HIP(hipDeviceSynchronize());
  KF_HIP(hipMemcpy(h, P->params, count * 4, hipMemcpyDeviceToHost));
  return KF_OK;
}

extern "C" kf_status kf_prior_init_params(kf_prior* P, uint64_t seed) {
  KF_CHECK_ARG(P != nullptr, "NULL argument");
  const ParamLayout& L = layout();
  std::vector<float> h(P->nparams, 0.0f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x1234567ull;
  auto uni = [&]() {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0);
  };
  const long fan_in[6] = {kConv[0].kc, kConv[1].kc, kConv[2].kc, kConv[3].kc, kFc1In, kHidden};
  for (int l = 0; l < 6; ++l) {
    const long n = L.b[l] - L.w[l];
    const double a = sqrt(6.0 / (double)fan_in[l]) * (l == 5 ? 0.1 : 1.0);
    for (long i = 0; i < n; ++i) h[L.w[l] + i] = (float)((2.0 * uni() - 1.0) * a);
  }
  KF_TRY(kf_prior_set_params(P, h.data(), P->nparams));
  KF_HIP(hipMemset(P->adam_m, 0, P->nparams * 4));
  KF_HIP(hipMemset(P->adam_v, 0, P->nparams * 4));
  P->step = 0;
  return KF_OK;
}

extern "C" kf_status kf_prior_forward(kf_prior* P, const float* frames, const int32_t* pairs, const float* prev,
                                      int32_t B, float* d_out, kf_stream s) {
  KF_TRY(check_batch(P, frames, pairs, prev, B));
  KF_CHECK_ARG(d_out != nullptr, "out is NULL");
  return forward(P, frames, pairs, prev, B, S(s), d_out, false);
}

extern "C" kf_status kf_prior_gradients(kf_prior* P, const float* frames, const int32_t* pairs, const float* prev,
                                        const float* target, int32_t B, float* d_grad, float* d_loss, kf_stream s) {
  KF_TRY(check_batch(P, frames, pairs, prev, B));
  KF_CHECK_ARG(target && d_grad, "NULL argument");
  KF_TRY(forward(P, frames, pairs, prev, B, S(s)));
  KF_TRY(backward(P, target, B, d_loss, S(s)));
  KF_HIP(hipMemcpyAsync(d_grad, P->grad, (size_t)P->nparams * 4, hipMemcpyDeviceToDevice, S(s)));
  return KF_OK;
}

extern "C" kf_status kf_prior_train_step(kf_prior* P, const float* frames, const int32_t* pairs, const float* prev,
                                         const float* target, int32_t B, float lr, float* d_loss, kf_stream s) {
  KF_TRY(check_batch(P, frames, pairs, prev, B));
  KF_CHECK_ARG(target != nullptr, "target is NULL");
  KF_TRY(forward(P, frames, pairs, prev, B, S(s)));
  KF_TRY(backward(P, target, B, d_loss, S(s)));
  P->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - (float)pow(0.9, (double)P->step), c2 = 1.0f - (float)pow(0.999, (double)P->step);
  hipLaunchKernelGGL(adam, dim3(blocks_for(P->nparams)), dim3(256), 0, S(s), P->params, (const float*)P->grad,
                     P->adam_m, P->adam_v, P->nparams, lr, b1, b2, 1e-8f, c1, c2);
  KF_TRY(check_launch("adam"));
  return sync_weights(P, S(s));
}
