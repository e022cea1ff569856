// Small BF16 CNN on MFMA for the learned openness surrogate: 3x3 "same"
// convolutions (zero padding) on NHWC bfloat16 activations with float32
// accumulation, bias + GELU epilogues, the backward passes and Adam.
//
// No reference counterpart: the reference has no learned component.  The
// surrogate's structure (6 conv3x3 layers, 32 channels, GELU, BF16 weights and
// activations, FP32 accumulation) is fixed by the build's own specification.
//
// Convolution as implicit GEMM on v_mfma_f32_32x32x16_bf16 (wave64):
//   M = 32 output channels, N = 32 consecutive pixels of one image row,
//   K = 9 taps x Cin (tap-major, channel-minor), padded to a multiple of 16.
// Lane l holds A row (channel) l%32, k (l/32)*8..+7 and B column (pixel) l%32
// over the same k -- one tap's 8 consecutive channels, one 16-byte load; the
// accumulator holds pixel l%32 and channels 8i + (l/32)*4 + j at acc[4i + j],
// so each lane stores 4 x 4 consecutive channels of one pixel.
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

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest even: gfx950's v_cvt_pk_bf16_f32
__device__ __forceinline__ u16 f2bf(float f) { return __builtin_bit_cast(u16, (__bf16)f); }
// two floats -> packed bfloat16 pair (a in the low half), one instruction
__device__ __forceinline__ unsigned pk_bf16(float a, float b) {
  const f32x2 v = {a, b};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
}

// GELU in its tanh form (Hendrycks & Gimpel 2016):
//   gelu(x) = 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))) = x * sigmoid(u),
//   u = 2 sqrt(2/pi) (x + 0.044715 x^3);
// one v_exp_f32 and one v_rcp_f32 per value (exp overflow gives -0 for x << 0).
constexpr float GELU_K0 = 1.5957691216057308f;   // 2 sqrt(2/pi)
constexpr float GELU_K1 = 0.07135481627260025f;  // 2 sqrt(2/pi) 0.044715
constexpr float LOG2E = 1.4426950408889634f;

__device__ __forceinline__ float sigmoid_of_u(float x) {
  // exp(-u) = exp2(x * (x^2 * (-K1 log2 e) + (-K0 log2 e)))
  const float m = x * fmaf(x * x, -GELU_K1 * LOG2E, -GELU_K0 * LOG2E);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(m));
}
__device__ __forceinline__ float gelu(float x) { return x * sigmoid_of_u(x); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float s = sigmoid_of_u(x);
  return s + x * s * (1.0f - s) * fmaf(3.0f * GELU_K1, x * x, GELU_K0);
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
  // C = W (32 out channels x K) * X (K x 32 pixels): lane holds pixel n and
  // channels 8i + half*4 + j at acc[4i + j]
  float bias[16];
#pragma unroll
  for (int e = 0; e < 16; e++)
    bias[e] = (EPI != EPI_GELU_GRAD && (8 * (e >> 2) + half * 4 + (e & 3)) < A.cout)
                  ? A.bias[8 * (e >> 2) + half * 4 + (e & 3)]
                  : 0.0f;
  for (int tile = blockIdx.x * 4 + wave; tile < n_tiles; tile += gridDim.x * 4) {
    const int tx = tile % tiles_x;
    const int row = tile / tiles_x;  // b * H + y
    const int b = row / A.H, y = row - b * A.H;
    const int x0 = tx * 32;
    const int px = x0 + n;           // this lane's pixel (B column)
    f32x16 acc;   // starts at the bias: z = bias + sum_k w x
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = bias[e];
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
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, a, acc, 0, 0, 0);
    }
    if (px >= A.W) continue;
    const int64_t p = ((int64_t)b * A.H + y) * A.W + px;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int c0 = 8 * i + half * 4;  // channels c0 .. c0 + 3
      if (EPI == EPI_GELU) {
        float z[4];
#pragma unroll
        for (int j = 0; j < 4; j++) z[j] = acc[4 * i + j];
        if (A.pre_out) *(float4*)(A.pre_out + p * CNN_C + c0) = make_float4(z[0], z[1], z[2], z[3]);
        *(uint2*)(A.y + p * CNN_C + c0) =
            make_uint2(pk_bf16(gelu(z[0]), gelu(z[1])), pk_bf16(gelu(z[2]), gelu(z[3])));
      } else if (EPI == EPI_LINEAR) {
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (c0 + j < A.cout) A.yf[p * A.cout + c0 + j] = acc[4 * i + j];
      } else {
        const float4 pz = *(const float4*)(A.pre_in + p * CNN_C + c0);
        *(uint2*)(A.y + p * CNN_C + c0) =
            make_uint2(pk_bf16(acc[4 * i + 0] * gelu_grad(pz.x), acc[4 * i + 1] * gelu_grad(pz.y)),
                       pk_bf16(acc[4 * i + 2] * gelu_grad(pz.z), acc[4 * i + 3] * gelu_grad(pz.w)));
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

// Masked squared error of the network output yf [n][cout]:
//   angle term     d = yf[p][0] + base[p] - label[p]   (label NaN: no label)
//   detail term    e = yf[p][1] - label_s[p], weighted by weight_s[p]
//                  (cout >= 2 and label_s given; label_s NaN: no label)
// loss[0] += sum d^2, loss[1] += sum w e^2; dz[p] = bf16 of (2 scale d,
// 2 scale w e, 0, ..., 0) over 32 channels.
__global__ void k_cnn_mse(const float* __restrict__ yf, int cout, const float* __restrict__ base,
                          const float* __restrict__ label, const float* __restrict__ label_s,
                          const float* __restrict__ weight_s, int64_t n, float scale,
                          u16* __restrict__ dz, float* __restrict__ loss) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float d2 = 0.0f, e2 = 0.0f;
  if (p < n) {
    const float lb = label[p];
    float g0 = 0.0f, g1 = 0.0f;
    if (!isnan(lb)) {
      const float d = yf[p * cout] + base[p] - lb;
      d2 = d * d;
      g0 = scale * 2.0f * d;
    }
    if (label_s) {
      const float ls = label_s[p];
      if (!isnan(ls)) {
        const float w = weight_s[p];
        const float e = yf[p * cout + 1] - ls;
        e2 = w * e * e;
        g1 = scale * 2.0f * w * e;
      }
    }
    uint4* row = (uint4*)(dz + p * CNN_C);
    row[0] = make_uint4(pk_bf16(g0, g1), 0, 0, 0);
    row[1] = make_uint4(0, 0, 0, 0);
    row[2] = make_uint4(0, 0, 0, 0);
    row[3] = make_uint4(0, 0, 0, 0);
  }
  // wave sums, one atomic per wave and term
  for (int o = 32; o > 0; o >>= 1) {
    d2 += __shfl_down(d2, o);
    e2 += __shfl_down(e2, o);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(loss, d2);
    if (label_s) atomicAdd(loss + 1, e2);
  }
}

// ------------------------------------------------------------ fused forward
//
// All six layers in one launch.  A workgroup owns a strip of FZ_CW = 96
// computed columns (FZ_SW = 84 output columns plus the 6-column receptive-
// field halo on each side) and a block of SH output rows, and streams down
// the rows: every layer keeps a ring of 4 rows of its input in LDS (bf16, 32
// channels = 64 bytes per pixel, 16-byte chunks XOR-swizzled by pixel so that
// 8 consecutive pixels of one chunk hit 8 different bank groups).  Layer l at
// step t produces row y0 - 6 + t - 2(l + 1) of its output: each layer reads
// only rows written in earlier steps, so one barrier per step separates all
// producers from all consumers.
//
// Waves are specialised and hold their layers' weights in VGPRs (the MFMA A
// operand, 16 bytes per lane and k-step), balanced so that every SIMD carries
// about the same GELU (VALU) and MFMA load:
//   waves 0..3  layers 1..4 (32 -> 32), all three 32-pixel tiles of the row,
//               the GELU epilogue of tile i interleaved with the MFMAs of i+1;
//   wave  4     loads input rows (8 channels) and runs layer 0 (8 -> 32) on tile 0;
//   waves 5, 6  layer 0 on tiles 1, 2 and layer 5 (32 -> cout) on tiles 0, 1;
//   wave  7     layer 5 on tile 2.
// Activations outside the image are zero in every layer (the "same" zero
// padding of the layer-by-layer kernels), so both paths issue identical MFMA
// sequences on identical operands and agree bit for bit.
constexpr int FZ_TILES = 3;
constexpr int FZ_CW = 32 * FZ_TILES;     // computed columns per strip
constexpr int FZ_SW = FZ_CW - 12;        // output columns per strip
constexpr int FZ_PW = FZ_CW + 2;         // ring row: one zero pixel each side
constexpr int FZ_ROW32 = FZ_PW * 64;     // bytes per 32-channel ring row
constexpr int FZ_ROW8 = FZ_PW * 16;      // bytes per 8-channel ring row
constexpr int FZ_RING0 = 4 * FZ_ROW8;
constexpr int FZ_RING32 = 4 * FZ_ROW32;
constexpr int FZ_LDS = FZ_RING0 + 5 * FZ_RING32;

struct FusedArgs {
  const u16* x;         // [B][H][W][8] bf16
  const u16* w[6];      // [32][80] (layer 0), [32][288] (layers 1..5) bf16
  const float* bias[6]; // [32] each
  float* yf;            // [B][H][W][cout] f32
  int B, H, W, cout, SH, n_strips, n_rb;
};

__device__ __forceinline__ int fz_off32(int pc, int chunk) {
  return pc * 64 + ((chunk ^ ((pc >> 1) & 3)) << 4);
}
__device__ __forceinline__ int fz_slot(int r, int y0) { return (r - y0 + 8) & 3; }

// One 32-channel 3x3 layer over the three tiles of row r (MFMA accumulators
// per tile, ds_read_b128 operands).  Returns the accumulators of tile nt.
__device__ __forceinline__ f32x16 fz_mma32(const char* rin, const bf16x8 (&w)[18],
                                           const f32x16& init, int r, int y0, int nt, int n,
                                           int half) {
  const char* rows[3] = {rin + fz_slot(r - 1, y0) * FZ_ROW32, rin + fz_slot(r, y0) * FZ_ROW32,
                         rin + fz_slot(r + 1, y0) * FZ_ROW32};
  f32x16 acc = init;
#pragma unroll
  for (int ks = 0; ks < 18; ks++) {
    const int tap = ks >> 1, dy = tap / 3, dx = tap % 3;
    const int pc = nt * 32 + n + dx;
    const bf16x8 b = *(const bf16x8*)(rows[dy] + fz_off32(pc, (ks & 1) * 2 + half));
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[ks], b, acc, 0, 0, 0);
  }
  return acc;
}

// GELU of accumulator values e, e + 1 (channels of one 4-group) as a packed pair
__device__ __forceinline__ unsigned fz_gelu2(const f32x16& acc, int e) {
  return pk_bf16(gelu(acc[e]), gelu(acc[e + 1]));
}

// GELU epilogue into a 32-channel ring row (zeros outside the image); the
// accumulators already hold bias + sum.
__device__ __forceinline__ void fz_store_gelu(char* dst, const f32x16& acc, bool ok, int pc,
                                              int half) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    uint2 v = make_uint2(0, 0);
    if (ok) v = make_uint2(fz_gelu2(acc, 4 * i), fz_gelu2(acc, 4 * i + 2));
    *(uint2*)(dst + fz_off32(pc, i) + half * 8) = v;
  }
}

// Layer 0 (8 -> 32) on tile nt of row r: cin 8 puts one tap in each half-wave.
__device__ __forceinline__ void fz_layer0(const char* lds, const bf16x8 (&w0)[5],
                                          const f32x16& bias, int r, int y0, int H, int W,
                                          int xs, int nt, int n, int half) {
  const bool row_ok = r >= 0 && r < H;
  f32x16 acc = bias;
  if (row_ok) {
#pragma unroll
    for (int ks = 0; ks < 5; ks++) {
      const int tap = 2 * ks + half;
      bf16x8 bv = __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
      if (tap < 9) {
        const int dy = tap / 3, dx = tap % 3;
        bv = *(const bf16x8*)(lds + fz_slot(r + dy - 1, y0) * FZ_ROW8 + (nt * 32 + n + dx) * 16);
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w0[ks], bv, acc, 0, 0, 0);
    }
  }
  const int x = xs + nt * 32 + n;
  fz_store_gelu((char*)lds + FZ_RING0 + fz_slot(r, y0) * FZ_ROW32, acc,
                row_ok && x >= 0 && x < W, nt * 32 + n + 1, half);
}

// Layer 5 (32 -> cout) on tile nt of output row r, stored to global memory.
__device__ __forceinline__ void fz_layer5(const FusedArgs& A, const char* lds,
                                          const bf16x8 (&w)[18], const f32x16& bias, int r,
                                          int y0, int b, int xs, int nt, int n, int half) {
  const f32x16 acc = fz_mma32(lds + FZ_RING0 + 4 * FZ_RING32, w, bias, r, y0, nt, n, half);
  const int c = nt * 32 + n, x = xs + c;
  if (half == 0 && c >= 6 && c < 6 + FZ_SW && x < A.W) {
    float* o = A.yf + (((int64_t)b * A.H + r) * A.W + x) * A.cout;
    if (A.cout == 2) {
      *(float2*)o = make_float2(acc[0], acc[1]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (j < A.cout) o[j] = acc[j];
    }
  }
}

__global__ void __launch_bounds__(512) k_cnn_fused(FusedArgs A) {
  extern __shared__ __attribute__((aligned(16))) char fz_lds[];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int n = lane & 31, half = lane >> 5;
  int blk = blockIdx.x;
  const int rb = blk % A.n_rb;
  blk /= A.n_rb;
  const int strip = blk % A.n_strips;
  const int b = blk / A.n_strips;
  const int y0 = rb * A.SH;
  const int y_end = min(y0 + A.SH, A.H);
  const int xs = strip * FZ_SW - 6;   // image column of computed column 0

  // zero the pad pixels of every ring row (never written afterwards)
  for (int e = tid; e < 4 * 6 * 2; e += 512) {
    const int ring = e / 8, rest = e % 8, slot = rest >> 1, side = rest & 1;
    if (ring == 0) {
      *(uint4*)(fz_lds + slot * FZ_ROW8 + side * (FZ_PW - 1) * 16) = make_uint4(0, 0, 0, 0);
    } else {
      char* row = fz_lds + FZ_RING0 + (ring - 1) * FZ_RING32 + slot * FZ_ROW32 +
                  side * (FZ_PW - 1) * 64;
      for (int q = 0; q < 4; q++) *(uint4*)(row + q * 16) = make_uint4(0, 0, 0, 0);
    }
  }

  // weights: w = layer wave + 1 (waves 0..3) or layer 5 (waves 5..7); w0 = layer 0
  const int big = wave < 4 ? wave + 1 : 5;
  bf16x8 w[18], w0[5];
  f32x16 bias, bias0;
#pragma unroll
  for (int ks = 0; ks < 18; ks++)
    w[ks] = __builtin_bit_cast(bf16x8, *(const uint4*)(A.w[big] + n * 288 + ks * 16 + half * 8));
#pragma unroll
  for (int ks = 0; ks < 5; ks++)
    w0[ks] = __builtin_bit_cast(bf16x8, *(const uint4*)(A.w[0] + n * 80 + ks * 16 + half * 8));
#pragma unroll
  for (int e = 0; e < 16; e++) {
    bias[e] = A.bias[big][8 * (e >> 2) + half * 4 + (e & 3)];
    bias0[e] = A.bias[0][8 * (e >> 2) + half * 4 + (e & 3)];
  }
  __syncthreads();

  const int T = A.SH + 18;
  for (int t = 0; t < T; t++) {
    const int r1 = y0 - 8 + t;                          // layer 0 output row
    const bool l0 = r1 >= y0 - 5 && r1 < y0 + A.SH + 5;
    const int r6 = y0 + t - 18;                         // layer 5 output row
    const bool l5 = r6 >= y0 && r6 < y_end;
    if (wave < 4) {
      // layer l = wave + 1 computes row r of a_{l+1}; rows needed [y0 - (5 - l), y0 + SH + 5 - l)
      const int l = wave + 1;
      const int r = y0 - 6 + t - 2 * (l + 1);
      if (r >= y0 - (5 - l) && r < y0 + A.SH + 5 - l) {
        const bool row_ok = r >= 0 && r < A.H;
        const char* rin = fz_lds + FZ_RING0 + (l - 1) * FZ_RING32;
        char* dst = fz_lds + FZ_RING0 + l * FZ_RING32 + fz_slot(r, y0) * FZ_ROW32;
        if (!row_ok) {
          for (int nt = 0; nt < FZ_TILES; nt++)
            fz_store_gelu(dst, bias, false, nt * 32 + n + 1, half);
        } else {
          const char* rows[3] = {rin + fz_slot(r - 1, y0) * FZ_ROW32,
                                 rin + fz_slot(r, y0) * FZ_ROW32,
                                 rin + fz_slot(r + 1, y0) * FZ_ROW32};
          f32x16 prev = fz_mma32(rin, w, bias, r, y0, 0, n, half);
#pragma unroll
          for (int nt = 1; nt <= FZ_TILES; nt++) {
            // MFMAs of tile nt with the epilogue of tile nt - 1 in between
            const int pp = (nt - 1) * 32 + n + 1;
            const int xp = xs + (nt - 1) * 32 + n;
            const bool okp = xp >= 0 && xp < A.W;
            f32x16 acc = bias;
            unsigned pk[8];
#pragma unroll
            for (int ks = 0; ks < 18; ks++) {
              if (nt < FZ_TILES) {
                const int tap = ks >> 1, dy = tap / 3, dx = tap % 3;
                const bf16x8 bv = *(const bf16x8*)(rows[dy] +
                                                   fz_off32(nt * 32 + n + dx, (ks & 1) * 2 + half));
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[ks], bv, acc, 0, 0, 0);
              }
              if (ks < 8) pk[ks] = okp ? fz_gelu2(prev, 2 * ks) : 0u;
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
              *(uint2*)(dst + fz_off32(pp, i) + half * 8) = make_uint2(pk[2 * i], pk[2 * i + 1]);
            prev = acc;
          }
        }
      }
    } else if (wave == 4) {
      // input row r0 into registers (global loads in flight during layer 0)
      const int r0 = y0 - 6 + t;
      const bool load = r0 < y0 + A.SH + 6;
      uint4 in0 = make_uint4(0, 0, 0, 0), in1 = make_uint4(0, 0, 0, 0);
      if (load && r0 >= 0 && r0 < A.H) {
        const u16* src = A.x + ((int64_t)b * A.H + r0) * A.W * 8;
        const int xa = xs + lane, xb = xs + 64 + lane;
        if (xa >= 0 && xa < A.W) in0 = *(const uint4*)(src + (int64_t)xa * 8);
        if (lane < FZ_CW - 64 && xb >= 0 && xb < A.W) in1 = *(const uint4*)(src + (int64_t)xb * 8);
      }
      if (l0) fz_layer0(fz_lds, w0, bias0, r1, y0, A.H, A.W, xs, 0, n, half);
      // write after the barrier of the previous step: the slot of r0 held row
      // r0 - 4, last read by layer 0 in the previous step
      if (load) {
        char* row = fz_lds + fz_slot(r0, y0) * FZ_ROW8;
        *(uint4*)(row + (lane + 1) * 16) = in0;
        if (lane < FZ_CW - 64) *(uint4*)(row + (64 + lane + 1) * 16) = in1;
      }
    } else {
      const int nt = wave - 5;   // 0, 1, 2
      if (nt < 2 && l0) fz_layer0(fz_lds, w0, bias0, r1, y0, A.H, A.W, xs, nt + 1, n, half);
      if (l5) fz_layer5(A, fz_lds, w, bias, r6, y0, b, xs, nt, n, half);
    }
    __syncthreads();
  }
}

// Bilinear sample (pixel-centre aligned, clamped) of a grid field at full-
// resolution pixel (i, j): the weights and the four source offsets.
struct Bilerp {
  int64_t o00, o01, o10, o11;
  float tx, ty;
};
__device__ __forceinline__ Bilerp bilerp_at(int i, int j, float inv, int hs, int ws, int64_t ld) {
  float x = ((float)j + 0.5f) * inv - 0.5f;
  x = fminf(fmaxf(x, 0.0f), (float)(ws - 1));
  const int x0 = min((int)x, ws - 1), x1 = min(x0 + 1, ws - 1);
  float y = ((float)i + 0.5f) * inv - 0.5f;
  y = fminf(fmaxf(y, 0.0f), (float)(hs - 1));
  const int y0 = min((int)y, hs - 1), y1 = min(y0 + 1, hs - 1);
  Bilerp b;
  b.tx = x - (float)x0;
  b.ty = y - (float)y0;
  b.o00 = (int64_t)y0 * ld + x0;
  b.o01 = (int64_t)y0 * ld + x1;
  b.o10 = (int64_t)y1 * ld + x0;
  b.o11 = (int64_t)y1 * ld + x1;
  return b;
}
__device__ __forceinline__ float bilerp(const Bilerp& b, float v00, float v01, float v10,
                                        float v11) {
  const float top = v00 * (1.0f - b.tx) + v01 * b.tx;
  const float bot = v10 * (1.0f - b.tx) + v11 * b.tx;
  return top * (1.0f - b.ty) + bot * b.ty;
}

// Full-resolution surrogate angle from the network-grid output:
//   out = bilerp(yf[.][0] + base) + s_scale * yf[cell][1] * (z - bilerp(zg))
// (the second term only when cout >= 2 and z is given).  yf is [hs][ld][cout],
// base and zg [hs][ld]; cell = (i / f, j / f).  Four output pixels per lane.
// When f is a multiple of 8, the four pixels j0 .. j0 + 3 (j0 % 4 == 0) share
// their bilinear columns and their cell, so the column pair is interpolated
// along y once and each pixel costs one interpolation along x.
__global__ void k_cnn_compose(const float* __restrict__ yf, int cout,
                              const float* __restrict__ base, const float* __restrict__ zg,
                              int64_t ldg, int hs, int ws, int f, const float* __restrict__ z,
                              int64_t ldz, float s_scale, float* __restrict__ out, int64_t ldo,
                              int h, int w) {
  const int j0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (j0 >= w) return;
  const float inv = 1.0f / (float)f;
  const bool grouped = (f & 7) == 0;
  for (int i = blockIdx.y; i < h; i += gridDim.y) {
    float r[4];
    if (grouped) {
      const Bilerp b = bilerp_at(i, j0, inv, hs, ws, ldg);
      const float a0 = (yf[b.o00 * cout] + base[b.o00]) * (1.0f - b.ty) +
                       (yf[b.o10 * cout] + base[b.o10]) * b.ty;
      const float a1 = (yf[b.o01 * cout] + base[b.o01]) * (1.0f - b.ty) +
                       (yf[b.o11 * cout] + base[b.o11]) * b.ty;
      float z0 = 0.0f, z1 = 0.0f, sc = 0.0f;
      float zz[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      if (z) {
        z0 = zg[b.o00] * (1.0f - b.ty) + zg[b.o10] * b.ty;
        z1 = zg[b.o01] * (1.0f - b.ty) + zg[b.o11] * b.ty;
        const int64_t c = (int64_t)min(i / f, hs - 1) * ldg + min(j0 / f, ws - 1);
        sc = s_scale * yf[c * cout + 1];
        const float* zr = z + (int64_t)i * ldz + j0;
        if (j0 + 3 < w && ((reinterpret_cast<uintptr_t>(zr) & 15) == 0)) {
          // streamed once: no reuse, keep it out of the caches' way
          const f32x4 v = __builtin_nontemporal_load((const f32x4*)zr);
          zz[0] = v[0]; zz[1] = v[1]; zz[2] = v[2]; zz[3] = v[3];
        } else {
          for (int q = 0; q < 4; q++) zz[q] = zr[min(q, w - 1 - j0)];
        }
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        float x = ((float)(j0 + q) + 0.5f) * inv - 0.5f;
        x = fminf(fmaxf(x, 0.0f), (float)(ws - 1));
        const float tx = x - floorf(x);
        float v = a0 + tx * (a1 - a0);
        if (z) v += sc * (zz[q] - (z0 + tx * (z1 - z0)));
        r[q] = v;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int j = min(j0 + q, w - 1);
        const Bilerp b = bilerp_at(i, j, inv, hs, ws, ldg);
        float v = bilerp(b, yf[b.o00 * cout] + base[b.o00], yf[b.o01 * cout] + base[b.o01],
                         yf[b.o10 * cout] + base[b.o10], yf[b.o11 * cout] + base[b.o11]);
        if (z) {
          const float zb = bilerp(b, zg[b.o00], zg[b.o01], zg[b.o10], zg[b.o11]);
          const int64_t c = (int64_t)min(i / f, hs - 1) * ldg + min(j / f, ws - 1);
          v += s_scale * yf[c * cout + 1] * (z[(int64_t)i * ldz + j] - zb);
        }
        r[q] = v;
      }
    }
    float* o = out + (int64_t)i * ldo + j0;
    if (j0 + 3 < w && ((reinterpret_cast<uintptr_t>(o) & 15) == 0)) {
      const f32x4 v = {r[0], r[1], r[2], r[3]};
      __builtin_nontemporal_store(v, (f32x4*)o);
    } else {
      for (int q = 0; q < 4 && j0 + q < w; q++) o[q] = r[q];
    }
  }
}

// Training label of the detail channel for each labelled grid cell: the
// least-squares slope s* of r = exact - bilerp(lab) on dz = z - bilerp(zg)
// over the cell's f x f pixels, label_s = s_unit * s*, weight_s = sum dz^2 /
// (f^2 s_unit^2).  Cells without a label, with a pixel outside the exact
// rectangle [ey0, ey0 + eh) x [ex0, ex0 + ew), with a NaN in the bilinear
// stencil of lab, or with sum dz^2 == 0 get label_s = NaN, weight_s = 0.
__global__ void k_cnn_detail_label(const float* __restrict__ exact, int64_t lde, int ey0,
                                   int ex0, int eh, int ew, const float* __restrict__ lab,
                                   const float* __restrict__ zg, int64_t ldg, int hs, int ws,
                                   int f, const float* __restrict__ z, int64_t ldz,
                                   float s_unit, float* __restrict__ label_s,
                                   float* __restrict__ weight_s) {
  const int cx = blockIdx.x * blockDim.x + threadIdx.x, cy = blockIdx.y;
  if (cx >= ws || cy >= hs) return;
  const int64_t co = (int64_t)cy * ldg + cx;
  float ls = __int_as_float(0x7fc00000), wt = 0.0f;
  const int i0 = cy * f, j0 = cx * f;
  if (!isnan(lab[co]) && i0 >= ey0 && i0 + f <= ey0 + eh && j0 >= ex0 && j0 + f <= ex0 + ew) {
    const float inv = 1.0f / (float)f;
    double srd = 0.0, sdd = 0.0;
    bool ok = true;
    for (int i = i0; i < i0 + f && ok; i++) {
      for (int j = j0; j < j0 + f; j++) {
        const Bilerp b = bilerp_at(i, j, inv, hs, ws, ldg);
        const float bl = bilerp(b, lab[b.o00], lab[b.o01], lab[b.o10], lab[b.o11]);
        if (isnan(bl)) {
          ok = false;
          break;
        }
        const float zb = bilerp(b, zg[b.o00], zg[b.o01], zg[b.o10], zg[b.o11]);
        const double r = (double)exact[(int64_t)(i - ey0) * lde + (j - ex0)] - (double)bl;
        const double d = (double)z[(int64_t)i * ldz + j] - (double)zb;
        srd += r * d;
        sdd += d * d;
      }
    }
    if (ok && sdd > 0.0) {
      ls = (float)(s_unit * srd / sdd);
      wt = (float)(sdd / ((double)f * f * (double)s_unit * s_unit));
    }
  }
  label_s[co] = ls;
  weight_s[co] = wt;
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

extern "C" int fsg_cnn_mse(const float* yf, int32_t cout, const float* base, const float* label,
                           const float* label_s, const float* weight_s, int64_t n, float scale,
                           void* dz, float* loss, void* stream) {
  if (!yf || !base || !label || !dz || !loss || n <= 0 || cout < 1 || cout > 32)
    return fail(FSG_E_INVALID, "fsg_cnn_mse: bad arguments");
  if (label_s && (cout < 2 || !weight_s))
    return fail(FSG_E_INVALID, "fsg_cnn_mse: the detail term needs cout >= 2 and weight_s");
  hipLaunchKernelGGL(k_cnn_mse, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     as_stream(stream), yf, cout, base, label, label_s, weight_s, n, scale,
                     (u16*)dz, loss);
  FSG_CHECK_LAUNCH("k_cnn_mse");
  return FSG_OK;
}

extern "C" int fsg_cnn_forward(const fsg_cnn_net* net, const void* x, int32_t batch, int32_t h,
                               int32_t w, float* yf, int32_t rows_per_block, void* stream) {
  if (!net || !x || !yf || batch <= 0 || h <= 0 || w <= 0)
    return fail(FSG_E_INVALID, "fsg_cnn_forward: bad arguments");
  if (net->cin != 8 || net->cout < 1 || net->cout > 4)
    return fail(FSG_E_INVALID, "fsg_cnn_forward: the fused network is 8 -> 32 x 5 -> cout <= 4");
  FusedArgs A;
  for (int l = 0; l < 6; l++) {
    if (!net->wpack[l] || !net->bias[l]) return fail(FSG_E_INVALID, "fsg_cnn_forward: NULL layer");
    A.w[l] = (const u16*)net->wpack[l];
    A.bias[l] = net->bias[l];
  }
  A.x = (const u16*)x;
  A.yf = yf;
  A.B = batch;
  A.H = h;
  A.W = w;
  A.cout = net->cout;
  int sh = rows_per_block > 0 ? rows_per_block : 256;
  A.n_rb = (h + sh - 1) / sh;
  A.SH = (h + A.n_rb - 1) / A.n_rb;   // balanced blocks
  A.n_strips = (w + FZ_SW - 1) / FZ_SW;
  const int64_t blocks = (int64_t)batch * A.n_strips * A.n_rb;
  if (blocks > 0x7fffffff) return fail(FSG_E_INVALID, "fsg_cnn_forward: too many blocks");
  hipStream_t s = as_stream(stream);
  TimedScope ts("k_cnn_fused", s);
  hipLaunchKernelGGL(k_cnn_fused, dim3((unsigned)blocks), dim3(512), FZ_LDS, s, A);
  FSG_CHECK_LAUNCH("k_cnn_fused");
  return FSG_OK;
}

extern "C" int fsg_cnn_compose(const float* yf, int32_t cout, const float* base, const float* zg,
                               int64_t ldg, int32_t hs, int32_t ws, int32_t factor,
                               const float* z, int64_t ldz, float s_scale, float* out,
                               int64_t ldo, int32_t h, int32_t w, void* stream) {
  if (!yf || !base || !out || cout < 1 || cout > 32 || hs <= 0 || ws <= 0 || factor <= 0 ||
      h <= 0 || w <= 0 || ldg < ws || ldo < w)
    return fail(FSG_E_INVALID, "fsg_cnn_compose: bad arguments");
  if (z && (cout < 2 || !zg || ldz < w))
    return fail(FSG_E_INVALID, "fsg_cnn_compose: the detail term needs cout >= 2, zg and z");
  if ((int64_t)hs * factor < h || (int64_t)ws * factor < w)
    return fail(FSG_E_INVALID, "fsg_cnn_compose: output larger than the grid covers");
  hipStream_t s = as_stream(stream);
  TimedScope ts("k_cnn_compose", s);
  hipLaunchKernelGGL(k_cnn_compose, dim3((w + 1023) / 1024, std::min(h, 16384)), dim3(256), 0,
                     s, yf, cout, base, zg, ldg, hs, ws, factor, z, ldz, s_scale, out, ldo, h,
                     w);
  FSG_CHECK_LAUNCH("k_cnn_compose");
  return FSG_OK;
}

extern "C" int fsg_cnn_detail_label(const float* exact, int64_t lde, int32_t ey0, int32_t ex0,
                                    int32_t eh, int32_t ew, const float* lab, const float* zg,
                                    int64_t ldg, int32_t hs, int32_t ws, int32_t factor,
                                    const float* z, int64_t ldz, float s_unit, float* label_s,
                                    float* weight_s, void* stream) {
  if (!exact || !lab || !zg || !z || !label_s || !weight_s || hs <= 0 || ws <= 0 ||
      factor <= 0 || ldg < ws || lde < ew || ldz < (int64_t)ws * factor || s_unit <= 0.0f ||
      ey0 < 0 || ex0 < 0 || eh < 0 || ew < 0)
    return fail(FSG_E_INVALID, "fsg_cnn_detail_label: bad arguments");
  hipLaunchKernelGGL(k_cnn_detail_label, dim3((ws + 63) / 64, hs), dim3(64), 0,
                     as_stream(stream), exact, lde, ey0, ex0, eh, ew, lab, zg, ldg, hs, ws,
                     factor, z, ldz, s_unit, label_s, weight_s);
  FSG_CHECK_LAUNCH("k_cnn_detail_label");
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
