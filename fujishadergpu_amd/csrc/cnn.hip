// Small BF16 CNN on MFMA for the learned openness surrogate: 3x3 "same"
// convolutions (zero padding) on NHWC bfloat16 activations with float32
// accumulation, bias + GELU epilogues, the backward passes and Adam.
//
// No reference counterpart: the reference has no learned component.  The
// surrogate's structure (6 conv3x3 layers, 32 channels, GELU, BF16 weights and
// activations, FP32 accumulation) is fixed by the build's own specification.
//
// Convolution as implicit GEMM on v_mfma_f32_32x32x16_bf16 (wave64):
//   M = 32 consecutive pixels of one image row, N = 32 output channels,
//   K = 9 taps x Cin (tap-major, channel-minor), padded to a multiple of 16.
// Lane l holds A row (pixel) l%32, k (l/32)*8..+7 -- one tap's 8 consecutive
// channels, one 16-byte load -- and B column (channel) l%32 over the same k;
// the accumulator holds column l%32, rows 8i + (l/32)*4 + j at acc[4i + j].
// Weights are staged in LDS once per workgroup; each wave walks 32-pixel
// tiles grid-stride.
#include <algorithm>
#include <cmath>

#include "fsg_common.hpp"

namespace fsg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16;

constexpr int CNN_C = 32;          // hidden width
constexpr int CNN_MAX_KPAD = 288;  // 9 taps x 32 channels

__device__ __forceinline__ float bf2f(u16 v) { return __uint_as_float((unsigned)v << 16); }

// float -> bfloat16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ u16 f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u16)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}

__device__ __forceinline__ float gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

enum { EPI_GELU = 0, EPI_LINEAR = 1, EPI_GELU_GRAD = 2 };

struct ConvArgs {
  const u16* x;        // [B][H][W][Cin] bf16
  const u16* wpack;    // [32][kpad] bf16: row = output channel, k = tap*Cin + ci
  const float* bias;   // [32] (EPI_GELU, EPI_LINEAR)
  const float* pre_in; // [B][H][W][32] f32 pre-activations (EPI_GELU_GRAD)
  float* pre_out;      // optional [B][H][W][32] f32 pre-activations (EPI_GELU)
  u16* y;              // [B][H][W][32] bf16 (EPI_GELU, EPI_GELU_GRAD)
  float* yf;           // [B][H][W][cout] f32 (EPI_LINEAR)
  int B, H, W, cin, kpad, cout;
};

template <int EPI>
__global__ void __launch_bounds__(256) k_conv3x3(ConvArgs A) {
  __shared__ __attribute__((aligned(16))) u16 wl[CNN_C * CNN_MAX_KPAD];
  const int t = threadIdx.x;
  const int nw = CNN_C * A.kpad;
  for (int e = t * 8; e < nw; e += 256 * 8)
    *(uint4*)(wl + e) = *(const uint4*)(A.wpack + e);
  __syncthreads();
  const int lane = t & 63, wave = t >> 6;
  const int n = lane & 31, half = lane >> 5;
  const int tiles_x = (A.W + 31) >> 5;
  const int n_tiles = A.B * A.H * tiles_x;
  const int ksteps = A.kpad >> 4;
  const float bias = (EPI != EPI_GELU_GRAD) ? A.bias[n] : 0.0f;
  for (int tile = blockIdx.x * 4 + wave; tile < n_tiles; tile += gridDim.x * 4) {
    const int tx = tile % tiles_x;
    const int row = tile / tiles_x;  // b * H + y
    const int b = row / A.H, y = row - b * A.H;
    const int x0 = tx * 32;
    const int px = x0 + n;           // this lane's A-row pixel
    f32x16 acc = {};
    for (int ks = 0; ks < ksteps; ks++) {
      const int k = ks * 16 + half * 8;
      const int tap = k / A.cin, ci = k - tap * A.cin;
      bf16x8 a;
      const int yy = y + tap / 3 - 1, xx = px + tap % 3 - 1;
      if (tap < 9 && yy >= 0 && yy < A.H && xx >= 0 && xx < A.W) {
        const uint4 v = *(const uint4*)(A.x + ((((int64_t)b * A.H + yy) * A.W + xx) * A.cin + ci));
        a = __builtin_bit_cast(bf16x8, v);
      } else {
        a = __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
      }
      const bf16x8 w = __builtin_bit_cast(bf16x8, *(const uint4*)(wl + n * A.kpad + k));
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, w, acc, 0, 0, 0);
    }
    // epilogue: lane holds channel n for pixels x0 + 8i + half*4 + j
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int xo = x0 + 8 * i + half * 4 + j;
        if (xo >= A.W) continue;
        const int64_t p = ((int64_t)b * A.H + y) * A.W + xo;
        const float v = acc[4 * i + j];
        if (EPI == EPI_GELU) {
          const float z = v + bias;
          if (A.pre_out) A.pre_out[p * CNN_C + n] = z;
          A.y[p * CNN_C + n] = f2bf(gelu(z));
        } else if (EPI == EPI_LINEAR) {
          if (n < A.cout) A.yf[p * A.cout + n] = v + bias;
        } else {
          A.y[p * CNN_C + n] = f2bf(v * gelu_grad(A.pre_in[p * CNN_C + n]));
        }
      }
    }
  }
}

// dW[co][ci][tap] += sum_p dz[p][co] * x[p + off(tap)][ci], db[co] += sum_p dz[p][co]
// (the master-weight layout, w[co][ci][3][3]).
// MFMA with M = output channel, N = input channel, K = 16 pixels per step.
// A workgroup stages one 64-pixel row segment of dz and the three x rows
// around it (zero outside the image) in LDS; wave w owns taps w, w+4, w+8.
struct WgradArgs {
  const u16* x;   // [B][H][W][cin]
  const u16* dz;  // [B][H][W][32]
  float* dw;      // [cout][cin][9] f32 (accumulated), cout = 32
  float* db;      // [32] f32 (accumulated)
  int B, H, W, cin;
};

constexpr int WG_SEG = 64;

__global__ void __launch_bounds__(256) k_conv3x3_wgrad(WgradArgs A) {
  __shared__ __attribute__((aligned(16))) u16 sdz[WG_SEG * CNN_C];          // [px][co]
  __shared__ __attribute__((aligned(16))) u16 sx[3 * (WG_SEG + 2) * CNN_C]; // [r][px][ci]
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int n = lane & 31, half = lane >> 5;  // n: A row = co, B column = ci
  const int segs_x = (A.W + WG_SEG - 1) / WG_SEG;
  const int n_segs = A.B * A.H * segs_x;
  f32x16 acc[3] = {{}, {}, {}};
  float bsum = 0.0f;
  for (int sg = blockIdx.x; sg < n_segs; sg += gridDim.x) {
    const int sx0 = (sg % segs_x) * WG_SEG;
    const int row = sg / segs_x;
    const int b = row / A.H, y = row - b * A.H;
    __syncthreads();
    for (int e = t; e < WG_SEG * CNN_C; e += 256) {
      const int p = e / CNN_C, c = e - p * CNN_C;
      const int xx = sx0 + p;
      sdz[e] = (xx < A.W) ? A.dz[(((int64_t)b * A.H + y) * A.W + xx) * CNN_C + c] : (u16)0;
    }
    for (int e = t; e < 3 * (WG_SEG + 2) * A.cin; e += 256) {
      const int r = e / ((WG_SEG + 2) * A.cin);
      const int rem = e - r * (WG_SEG + 2) * A.cin;
      const int p = rem / A.cin, c = rem - p * A.cin;
      const int yy = y + r - 1, xx = sx0 + p - 1;
      sx[(r * (WG_SEG + 2) + p) * CNN_C + c] =
          (yy >= 0 && yy < A.H && xx >= 0 && xx < A.W)
              ? A.x[(((int64_t)b * A.H + yy) * A.W + xx) * A.cin + c]
              : (u16)0;
    }
    __syncthreads();
    if (t < CNN_C)
      for (int p = 0; p < WG_SEG; p++) bsum += bf2f(sdz[p * CNN_C + t]);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int tap = wave + 4 * q;
      if (tap >= 9) continue;
      const int dy = tap / 3, dx = tap % 3;  // staged row dy, column offset dx
      for (int ks = 0; ks < WG_SEG / 16; ks++) {
        bf16x8 a, bb;
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const int p = ks * 16 + half * 8 + e;
          a[e] = __builtin_bit_cast(__bf16, sdz[p * CNN_C + n]);
          bb[e] = (n < A.cin) ? __builtin_bit_cast(__bf16, sx[(dy * (WG_SEG + 2) + p + dx) * CNN_C + n])
                              : __builtin_bit_cast(__bf16, (u16)0);
        }
        acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc[q], 0, 0, 0);
      }
    }
  }
  // accumulate: lane holds column ci = n, rows co = 8i + half*4 + j
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int tap = wave + 4 * q;
    if (tap >= 9 || n >= A.cin) continue;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int co = 8 * i + half * 4 + j;
        atomicAdd(&A.dw[(co * A.cin + n) * 9 + tap], acc[q][4 * i + j]);
      }
  }
  if (t < CNN_C) atomicAdd(&A.db[t], bsum);
}

// Pack float32 master weights w[co][ci][3][3] into the forward layout
// wf[32][kpad] (k = tap*cin + ci) and the flipped, transposed backward layout
// wt[32][kpad_t] (row = ci, k = tap*cout_pad + co with tap' = 8 - tap), both
// bfloat16 with zero padding.
__global__ void k_cnn_pack(const float* __restrict__ w, int cout, int cin, u16* __restrict__ wf,
                           int kpad, u16* __restrict__ wt, int kpad_t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < CNN_C * kpad) {
    const int co = i / kpad, k = i - co * kpad;
    const int tap = k / cin, ci = k - tap * cin;
    float v = 0.0f;
    if (co < cout && tap < 9) v = w[(co * cin + ci) * 9 + tap];
    wf[i] = f2bf(v);
  }
  if (wt && i < CNN_C * kpad_t) {
    const int ci = i / kpad_t, k = i - ci * kpad_t;
    const int tap = k / CNN_C, co = k - tap * CNN_C;
    float v = 0.0f;
    if (ci < cin && co < cout && tap < 9) v = w[(co * cin + ci) * 9 + (8 - tap)];
    wt[i] = f2bf(v);
  }
}

// Adam (Kingma & Ba 2015) with bias correction, float32 master weights.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, u16* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

// Detrended elevation pyramid: channel k of pixel p = (mean_k[p] - z[p]) *
// inv[k] for k < 4, channels 4..7 zero, bfloat16 NHWC [h][w][8].
struct FeatArgs {
  const float* z;
  int64_t ldz;
  const float* mean[4];
  int64_t ldm;
  float inv[4];
  u16* out;
  int h, w;
};

__global__ void k_cnn_features(FeatArgs A) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (c >= A.w || r >= A.h) return;
  const float z = A.z[(int64_t)r * A.ldz + c];
  u16 v[8];
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = f2bf((A.mean[k][(int64_t)r * A.ldm + c] - z) * A.inv[k]);
#pragma unroll
  for (int k = 4; k < 8; k++) v[k] = 0;
  uint4 pk;
  pk.x = v[0] | ((unsigned)v[1] << 16);
  pk.y = v[2] | ((unsigned)v[3] << 16);
  pk.z = 0;
  pk.w = 0;
  *(uint4*)(A.out + ((int64_t)r * A.w + c) * 8) = pk;
}

// out[i][j] = bilinear sample of (src + add) at ((i + .5) / f - .5, (j + .5) / f - .5),
// coordinates clamped to the source grid (pixel-centre alignment).
__global__ void k_cnn_upsample(const float* __restrict__ src, const float* __restrict__ add,
                               int64_t lds, int hs, int ws, int f, float* __restrict__ out,
                               int64_t ldo, int h, int w) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= w) return;
  const float inv = 1.0f / (float)f;
  for (int i = blockIdx.y; i < h; i += gridDim.y) {
  float x = ((float)j + 0.5f) * inv - 0.5f;
  x = fminf(fmaxf(x, 0.0f), (float)(ws - 1));
  const int x0 = min((int)x, ws - 1), x1 = min(x0 + 1, ws - 1);
  const float tx = x - (float)x0;
  float y = ((float)i + 0.5f) * inv - 0.5f;
  y = fminf(fmaxf(y, 0.0f), (float)(hs - 1));
  const int y0 = min((int)y, hs - 1), y1 = min(y0 + 1, hs - 1);
  const float ty = y - (float)y0;
  auto s = [&](int a, int b) {
    const int64_t o = (int64_t)a * lds + b;
    return src[o] + (add ? add[o] : 0.0f);
  };
  const float top = s(y0, x0) * (1.0f - tx) + s(y0, x1) * tx;
  const float bot = s(y1, x0) * (1.0f - tx) + s(y1, x1) * tx;
  out[(int64_t)i * ldo + j] = top * (1.0f - ty) + bot * ty;
  }
}

// Masked squared error of pred = yf + base against label (NaN = no label):
// loss += sum d^2, dz[p][0] = bf16(scale * 2 d), dz[p][1..31] = 0.
__global__ void k_cnn_mse(const float* __restrict__ yf, const float* __restrict__ base,
                          const float* __restrict__ label, int64_t n, float scale,
                          u16* __restrict__ dz, float* __restrict__ loss) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float d2 = 0.0f;
  if (p < n) {
    const float lb = label[p];
    float g = 0.0f;
    if (!isnan(lb)) {
      const float d = yf[p] + base[p] - lb;
      d2 = d * d;
      g = scale * 2.0f * d;
    }
    uint4* row = (uint4*)(dz + p * CNN_C);
    row[0] = make_uint4(f2bf(g), 0, 0, 0);
    row[1] = make_uint4(0, 0, 0, 0);
    row[2] = make_uint4(0, 0, 0, 0);
    row[3] = make_uint4(0, 0, 0, 0);
  }
  // wave sum, one atomic per wave
  for (int o = 32; o > 0; o >>= 1) d2 += __shfl_down(d2, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, d2);
}

}  // namespace fsg

using namespace fsg;

extern "C" int fsg_cnn_features(const float* z, int64_t ldz, const float* const* means,
                                int64_t ldm, const float* inv, int32_t h, int32_t w, void* out,
                                void* stream) {
  if (!z || !means || !inv || !out || h <= 0 || w <= 0 || ldz < w || ldm < w)
    return fail(FSG_E_INVALID, "fsg_cnn_features: bad arguments");
  FeatArgs A;
  A.z = z;
  A.ldz = ldz;
  for (int k = 0; k < 4; k++) {
    if (!means[k]) return fail(FSG_E_INVALID, "fsg_cnn_features: mean NULL");
    A.mean[k] = means[k];
    A.inv[k] = inv[k];
  }
  A.ldm = ldm;
  A.out = (u16*)out;
  A.h = h;
  A.w = w;
  hipLaunchKernelGGL(k_cnn_features, dim3((w + 255) / 256, h), dim3(256), 0, as_stream(stream), A);
  FSG_CHECK_LAUNCH("k_cnn_features");
  return FSG_OK;
}

extern "C" int fsg_cnn_upsample(const float* src, const float* add, int64_t lds, int32_t hs,
                                int32_t ws, int32_t factor, float* out, int64_t ldo, int32_t h,
                                int32_t w, void* stream) {
  if (!src || !out || hs <= 0 || ws <= 0 || factor <= 0 || h <= 0 || w <= 0 || lds < ws ||
      ldo < w)
    return fail(FSG_E_INVALID, "fsg_cnn_upsample: bad arguments");
  hipLaunchKernelGGL(k_cnn_upsample, dim3((w + 255) / 256, std::min(h, 16384)), dim3(256), 0,
                     as_stream(stream),
                     src, add, lds, hs, ws, factor, out, ldo, h, w);
  FSG_CHECK_LAUNCH("k_cnn_upsample");
  return FSG_OK;
}

extern "C" int fsg_cnn_mse(const float* yf, const float* base, const float* label, int64_t n,
                           float scale, void* dz, float* loss, void* stream) {
  if (!yf || !base || !label || !dz || !loss || n <= 0)
    return fail(FSG_E_INVALID, "fsg_cnn_mse: bad arguments");
  hipLaunchKernelGGL(k_cnn_mse, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     as_stream(stream), yf, base, label, n, scale, (u16*)dz, loss);
  FSG_CHECK_LAUNCH("k_cnn_mse");
  return FSG_OK;
}

static int conv_check(const fsg_cnn_conv* c) {
  if (!c || !c->x || !c->wpack) return fail(FSG_E_INVALID, "fsg_cnn_conv3x3: NULL argument");
  if (c->batch <= 0 || c->h <= 0 || c->w <= 0)
    return fail(FSG_E_INVALID, "fsg_cnn_conv3x3: empty input");
  if (c->cin != 8 && c->cin != 32) return fail(FSG_E_INVALID, "fsg_cnn_conv3x3: cin must be 8 or 32");
  if (c->kpad % 16 || c->kpad < 9 * c->cin || c->kpad > CNN_MAX_KPAD + 16)
    return fail(FSG_E_INVALID, "fsg_cnn_conv3x3: kpad must be a multiple of 16 >= 9*cin");
  if (c->kpad > CNN_MAX_KPAD) return fail(FSG_E_INVALID, "fsg_cnn_conv3x3: kpad too large");
  if (c->cout <= 0 || c->cout > 32) return fail(FSG_E_INVALID, "fsg_cnn_conv3x3: cout out of range");
  return FSG_OK;
}

extern "C" int fsg_cnn_conv3x3(const fsg_cnn_conv* c, void* stream) {
  int rc = conv_check(c);
  if (rc) return rc;
  ConvArgs A{(const u16*)c->x, (const u16*)c->wpack, c->bias, c->pre_in, c->pre_out,
             (u16*)c->y, c->yf, c->batch, c->h, c->w, c->cin, c->kpad, c->cout};
  if (c->epilogue == FSG_CNN_GELU && (!c->bias || !c->y))
    return fail(FSG_E_INVALID, "fsg_cnn_conv3x3: GELU epilogue needs bias and y");
  if (c->epilogue == FSG_CNN_LINEAR && (!c->bias || !c->yf))
    return fail(FSG_E_INVALID, "fsg_cnn_conv3x3: linear epilogue needs bias and yf");
  if (c->epilogue == FSG_CNN_GELU_GRAD && (!c->pre_in || !c->y))
    return fail(FSG_E_INVALID, "fsg_cnn_conv3x3: GELU-gradient epilogue needs pre_in and y");
  const int64_t tiles = (int64_t)c->batch * c->h * ((c->w + 31) / 32);
  const int grid = (int)std::min<int64_t>((tiles + 3) / 4, 256 * 16);
  hipStream_t s = as_stream(stream);
  TimedScope ts("k_conv3x3", s);
  switch (c->epilogue) {
    case FSG_CNN_GELU:
      hipLaunchKernelGGL(k_conv3x3<EPI_GELU>, dim3(grid), dim3(256), 0, s, A);
      break;
    case FSG_CNN_LINEAR:
      hipLaunchKernelGGL(k_conv3x3<EPI_LINEAR>, dim3(grid), dim3(256), 0, s, A);
      break;
    case FSG_CNN_GELU_GRAD:
      hipLaunchKernelGGL(k_conv3x3<EPI_GELU_GRAD>, dim3(grid), dim3(256), 0, s, A);
      break;
    default:
      return fail(FSG_E_INVALID, "fsg_cnn_conv3x3: bad epilogue");
  }
  FSG_CHECK_LAUNCH("k_conv3x3");
  return FSG_OK;
}

extern "C" int fsg_cnn_conv3x3_wgrad(const void* x, const void* dz, int32_t batch, int32_t h,
                                     int32_t w, int32_t cin, float* dw, float* db,
                                     void* stream) {
  if (!x || !dz || !dw || !db) return fail(FSG_E_INVALID, "fsg_cnn_conv3x3_wgrad: NULL argument");
  if (batch <= 0 || h <= 0 || w <= 0 || (cin != 8 && cin != 32))
    return fail(FSG_E_INVALID, "fsg_cnn_conv3x3_wgrad: bad shape");
  WgradArgs A{(const u16*)x, (const u16*)dz, dw, db, batch, h, w, cin};
  const int64_t segs = (int64_t)batch * h * ((w + WG_SEG - 1) / WG_SEG);
  const int grid = (int)std::min<int64_t>(segs, 1024);
  hipStream_t s = as_stream(stream);
  TimedScope ts("k_conv3x3_wgrad", s);
  hipLaunchKernelGGL(k_conv3x3_wgrad, dim3(grid), dim3(256), 0, s, A);
  FSG_CHECK_LAUNCH("k_conv3x3_wgrad");
  return FSG_OK;
}

extern "C" int fsg_cnn_pack(const float* w, int32_t cout, int32_t cin, void* wf, int32_t kpad,
                            void* wt, int32_t kpad_t, void* stream) {
  if (!w || !wf || cout <= 0 || cout > 32 || cin <= 0 || cin > 32 || kpad % 16 ||
      kpad < 9 * cin || (wt && (kpad_t % 16 || kpad_t < 9 * 32)))
    return fail(FSG_E_INVALID, "fsg_cnn_pack: bad arguments");
  const int n = CNN_C * std::max(kpad, wt ? kpad_t : 0);
  hipLaunchKernelGGL(k_cnn_pack, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), w, cout,
                     cin, (u16*)wf, kpad, (u16*)wt, kpad_t);
  FSG_CHECK_LAUNCH("k_cnn_pack");
  return FSG_OK;
}

extern "C" int fsg_cnn_adam(float* p, const float* g, float* m, float* v, int32_t n, float lr,
                            float beta1, float beta2, float eps, int32_t step, void* stream) {
  if (!p || !g || !m || !v || n <= 0 || step <= 0)
    return fail(FSG_E_INVALID, "fsg_cnn_adam: bad arguments");
  const float c1 = 1.0f - std::pow(beta1, (float)step), c2 = 1.0f - std::pow(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), p, g, m, v,
                     n, lr, beta1, beta2, eps, c1, c2);
  FSG_CHECK_LAUNCH("k_adam");
  return FSG_OK;
}

extern "C" int fsg_cnn_to_bf16(const float* a, void* b, int64_t n, void* stream) {
  if (!a || !b || n < 0) return fail(FSG_E_INVALID, "fsg_cnn_to_bf16: bad arguments");
  if (n == 0) return FSG_OK;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     as_stream(stream), a, (u16*)b, n);
  FSG_CHECK_LAUNCH("k_f32_to_bf16");
  return FSG_OK;
}
