// The prepare step on a raster resident in HBM: NoData inference and the
// push-pull void fill of preprocess_dem_to_cog (io/dem_preprocess.py:462-727),
// without the file reads and the COG write.
//
//   fsg_prepare_sample      NEAREST read of the raster at (oh, ow)
//   fsg_value_mode          mode / min / max of the valid finite values
//                           (np.unique + argmax of _detect_sentinel_nodata and
//                           _detect_border_nodata): radix sort, run-length
//                           encode, first arg-max
//   fsg_count_equal         count of cells equal to a value
//   fsg_coarse_average      _nan_aware_coarse_average (:135-188)
//   fsg_pushpull_fill       pushpull_fill (algorithms/_pyramid_fill.py:24-87)
//   fsg_edge_connected      _edge_connected_mask (utils/nodata_handler.py:13-25)
//   fsg_prepare_fill        _apply_band_fill + _sample_coarse (:729-826) over
//                           the whole raster
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fsg_common.hpp"

namespace fsg {

// ------------------------------------------------------------ NEAREST read
// GDAL RasterIO NEAREST: source index floor((i + 0.5) * n_src / n_dst).
__global__ void k_nearest_sample(const float* __restrict__ src, int64_t lds, int h, int w,
                                 int oh, int ow, float* __restrict__ out, int64_t ldo) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ow) return;
  const int sj = min((int)floor(((double)j + 0.5) * ((double)w / (double)ow)), w - 1);
  for (int i = blockIdx.y; i < oh; i += gridDim.y) {
    const int si = min((int)floor(((double)i + 0.5) * ((double)h / (double)oh)), h - 1);
    out[(int64_t)i * ldo + j] = src[(int64_t)si * lds + sj];
  }
}

// --------------------------------------------------------------- value mode
struct ModeWs {
  float* keys;
  float* sorted;
  float* uniq;
  int* counts;
  unsigned long long* n_valid;
  int* n_runs;
  hipcub::KeyValuePair<int, int>* best;
  void* temp;
  size_t temp_bytes;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// invalid (non-finite or the declared value) -> a positive NaN, sorted last;
// -0 -> +0 so the zero run has one representative; counts zeroed
__global__ void k_mode_prep(const float* __restrict__ v, int64_t n, int has_decl, float decl,
                            float* __restrict__ keys, int* __restrict__ counts,
                            unsigned long long* __restrict__ n_valid) {
  __shared__ unsigned long long part[4];
  unsigned long long c = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    float x = v[e];
    const bool ok = isfinite(x) && !(has_decl && x == decl);
    if (x == 0.0f) x = 0.0f;
    keys[e] = ok ? x : __int_as_float(0x7fffffff);
    counts[e] = 0;
    c += ok ? 1ull : 0ull;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += part[k];
    if (t) atomicAdd(n_valid, t);
  }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void k_mode_finish(const float* __restrict__ sorted, const float* __restrict__ uniq,
                              const int* __restrict__ n_runs,
                              const hipcub::KeyValuePair<int, int>* __restrict__ best,
                              const unsigned long long* __restrict__ n_valid, int64_t n,
                              fsg_mode_stats* __restrict__ st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const unsigned long long nv = *n_valid;
  st->n_valid = (int64_t)nv;
  if (nv == 0) {
    st->mode = 0.0;
    st->count = 0;
    st->vmin = 0.0;
    st->vmax = 0.0;
    st->n_distinct = 0;
    return;
  }
  st->mode = (double)uniq[best->key];
  st->count = best->value;
  st->vmin = (double)sorted[0];
  st->vmax = (double)sorted[nv - 1];
  st->n_distinct = (int64_t)*n_runs - (n - (int64_t)nv);   // less the NaN runs of 1
}

static int mode_layout(int64_t n, void* base, size_t bytes, ModeWs* w, size_t* need) {
  size_t t_sort = 0, t_rle = 0, t_arg = 0;
  const int ni = (int)n;
  // size queries only (no launch, no device access)
  if (hipcub::DeviceRadixSort::SortKeys(nullptr, t_sort, (const float*)nullptr, (float*)nullptr,
                                        ni) != hipSuccess ||
      hipcub::DeviceRunLengthEncode::Encode(nullptr, t_rle, (const float*)nullptr,
                                            (float*)nullptr, (int*)nullptr, (int*)nullptr,
                                            ni) != hipSuccess ||
      hipcub::DeviceReduce::ArgMax(nullptr, t_arg, (const int*)nullptr,
                                   (hipcub::KeyValuePair<int, int>*)nullptr, ni) != hipSuccess) {
    if (need) *need = 0;
    return 2;
  }
  const size_t temp = align256(std::max(t_sort, std::max(t_rle, t_arg)));
  const size_t fn = align256(sizeof(float) * (size_t)n);
  const size_t in = align256(sizeof(int) * (size_t)n);
  const size_t total = 3 * fn + in + 3 * 256 + temp;
  if (need) *need = total;
  if (!w) return 0;
  if (!base || bytes < total) return 1;
  char* p = (char*)base;
  w->keys = (float*)p; p += fn;
  w->sorted = (float*)p; p += fn;
  w->uniq = (float*)p; p += fn;
  w->counts = (int*)p; p += in;
  w->n_valid = (unsigned long long*)p; p += 256;
  w->n_runs = (int*)p; p += 256;
  w->best = (hipcub::KeyValuePair<int, int>*)p; p += 256;
  w->temp = p;
  w->temp_bytes = temp;
  return 0;
}

// ------------------------------------------------------------ count equal
__global__ void k_count_equal(const float* __restrict__ v, int64_t n, float x,
                              unsigned long long* __restrict__ out) {
  __shared__ unsigned long long part[4];
  unsigned long long c = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    c += v[e] == x ? 1ull : 0ull;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += part[k];
    if (t) atomicAdd(out, t);
  }
}

// ----------------------------------------------------------- coarse average
struct NodataList {
  int n;            // finite values compared with ==
  float v[FSG_PREPARE_MAX_NODATA];
};

__device__ __forceinline__ bool is_nodata(float x, const NodataList& L) {
  if (!isfinite(x)) return true;
  for (int k = 0; k < L.n; k++)
    if (x == L.v[k]) return true;
  return false;
}

// a coarse cell: numpy's nansum over axes (1, 3) of the (ch, k, cw, k) view
// adds, for each sub-row, the sub-row's float32 running sum (from 0) to the
// cell's float32 running total; the mean divides in float64 by the count
__global__ void k_coarse_average(const float* __restrict__ s, int64_t lds, int ch, int cw, int k,
                                 NodataList L, float* __restrict__ coarse, int64_t ldc,
                                 uint8_t* __restrict__ cvalid,
                                 unsigned long long* __restrict__ n_void) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  bool void_cell = false;
  if (x < cw && y < ch) {
    float tot = 0.0f;
    int cnt = 0;
    for (int a = 0; a < k; a++) {
      const float* row = s + (int64_t)(y * k + a) * lds + (int64_t)x * k;
      float seg = 0.0f;
      for (int b = 0; b < k; b++) {
        const float v = row[b];
        const bool m = is_nodata(v, L);
        seg = seg + (m ? 0.0f : v);
        cnt += m ? 0 : 1;
      }
      tot = tot + seg;
    }
    const float c = cnt == 0 ? 0.0f : (float)((double)tot / (double)cnt);
    const bool ok = cnt > 0 && isfinite(c);
    coarse[(int64_t)y * ldc + x] = c;
    cvalid[(int64_t)y * cw + x] = ok ? 1 : 0;
    void_cell = !ok;
  }
  const unsigned long long b = __ballot(void_cell);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(n_void, (unsigned long long)__popcll(b));
}

// ---------------------------------------------------------------- push-pull
// scipy.ndimage.zoom / map_coordinates, order 1, 'nearest': per axis the
// coordinate clamped to [0, n - 1], start = floor, t = c - start; the value
// accumulates coeff = (v * w_y) * w_x over the 2 x 2 footprint (row-major) in
// float64 and is rounded to float32 once.
struct Lin {
  int i0, i1;
  double w0, w1;
};

__device__ __forceinline__ Lin lin_at(double c, int n) {
  c = fmin(fmax(c, 0.0), (double)(n - 1));
  const int s = (int)floor(c);
  const double t = c - (double)s;
  return Lin{s, min(s + 1, n - 1), 1.0 - t, t};
}

__device__ __forceinline__ float bilin(const float* __restrict__ a, int64_t ld, const Lin& Y,
                                       const Lin& X) {
  double t = 0.0;
  t = t + ((double)a[(int64_t)Y.i0 * ld + X.i0] * Y.w0) * X.w0;
  t = t + ((double)a[(int64_t)Y.i0 * ld + X.i1] * Y.w0) * X.w1;
  t = t + ((double)a[(int64_t)Y.i1 * ld + X.i0] * Y.w1) * X.w0;
  t = t + ((double)a[(int64_t)Y.i1 * ld + X.i1] * Y.w1) * X.w1;
  return (float)t;
}

constexpr float PP_EPS = 1e-6f;

__global__ void k_pp_init(const float* __restrict__ coarse, int64_t ldc,
                          const uint8_t* __restrict__ valid, int h, int w,
                          float* __restrict__ vw, float* __restrict__ wt) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= w) return;
  for (int y = blockIdx.y; y < h; y += gridDim.y) {
    const bool ok = valid[(int64_t)y * w + x] != 0;
    vw[(int64_t)y * w + x] = ok ? coarse[(int64_t)y * ldc + x] : 0.0f;
    wt[(int64_t)y * w + x] = ok ? 1.0f : 0.0f;
  }
}

// minimum of non-negative floats: their bit patterns order as unsigned ints
__global__ void k_min_nonneg(const float* __restrict__ a, int64_t n, unsigned* __restrict__ out) {
  unsigned m = 0x7f800000u;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    m = min(m, __float_as_uint(a[e]));
  for (int o = 32; o > 0; o >>= 1) m = min(m, (unsigned)__shfl_down((int)m, o));
  __shared__ unsigned part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) m = min(m, part[k]);
    atomicMin(out, m);   // one atomic per block
  }
}

__global__ void k_set_u32n(unsigned* p, int n, unsigned v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// push: num = zoom(vw), den = zoom(w) to (ho, wo); wv = min(den, 1);
// mean = den > eps ? num / max(den, eps) : 0; vw' = mean * wv; w' = wv
__global__ void k_pp_push(const float* __restrict__ vw, const float* __restrict__ wt, int hi,
                          int wi, float* __restrict__ vwo, float* __restrict__ wto, int ho,
                          int wo, double zy, double zx) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= wo) return;
  const Lin X = lin_at((double)x * zx, wi);
  for (int y = blockIdx.y; y < ho; y += gridDim.y) {
    const Lin Y = lin_at((double)y * zy, hi);
    const float num = bilin(vw, wi, Y, X);
    const float den = bilin(wt, wi, Y, X);
    const float wv = fminf(den, 1.0f);
    const float mean = den > PP_EPS ? num / fmaxf(den, PP_EPS) : 0.0f;
    vwo[(int64_t)y * wo + x] = mean * wv;
    wto[(int64_t)y * wo + x] = wv;
  }
}

__global__ void k_pp_top(const float* __restrict__ vw, const float* __restrict__ wt, int64_t n,
                         float* __restrict__ filled) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float w = wt[e];
  filled[e] = w > PP_EPS ? vw[e] / fmaxf(w, PP_EPS) : 0.0f;
}

// pull: up = zoom(filled) to (zh, zw), edge-padded / cropped to (th, tw);
// filled' = w > eps ? vw / max(w, eps) : up
__global__ void k_pp_pull(const float* __restrict__ fin, int fh, int fw, int zh, int zw,
                          double zy, double zx, const float* __restrict__ vw,
                          const float* __restrict__ wt, int th, int tw,
                          float* __restrict__ fout) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= tw) return;
  const Lin X = lin_at((double)min(x, zw - 1) * zx, fw);
  for (int y = blockIdx.y; y < th; y += gridDim.y) {
    const int64_t e = (int64_t)y * tw + x;
    const float w = wt[e];
    float v;
    if (w > PP_EPS) {
      v = vw[e] / fmaxf(w, PP_EPS);
    } else {
      const Lin Y = lin_at((double)min(y, zh - 1) * zy, fh);
      v = bilin(fin, fw, Y, X);
    }
    fout[e] = v;
  }
}

__global__ void k_pp_final(const float* __restrict__ coarse, int64_t ldc,
                           const uint8_t* __restrict__ valid, const float* __restrict__ filled,
                           int h, int w, float* __restrict__ out, int64_t ldo) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= w) return;
  for (int y = blockIdx.y; y < h; y += gridDim.y)
    out[(int64_t)y * ldo + x] =
        valid[(int64_t)y * w + x] ? coarse[(int64_t)y * ldc + x] : filled[(int64_t)y * w + x];
}

__global__ void k_fill_f32(float* p, int64_t n, float v) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}

// ------------------------------------------------------------ edge-connected
constexpr int EC_T = 32;   // tile edge; 256 threads, 4 cells each

__global__ void k_ec_seed(const uint8_t* __restrict__ mask, int h, int w,
                          uint8_t* __restrict__ ext) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= w) return;
  for (int y = blockIdx.y; y < h; y += gridDim.y) {
    const bool border = y == 0 || y == h - 1 || x == 0 || x == w - 1;
    ext[(int64_t)y * w + x] = (border && mask[(int64_t)y * w + x]) ? 1 : 0;
  }
}

// One sweep: each 32 x 32 tile propagates 'exterior' through its masked cells
// (4-connected) to a fixed point in LDS, seeded by the current exterior of
// the tile and its one-cell ring; *changed is set if the tile gained a cell.
// Reads of a neighbour tile's ring may see older values (cells only ever turn
// on), which the next sweep picks up; a sweep that changes nothing anywhere
// means every tile saw final rings, so the result is the fixed point.
__global__ void __launch_bounds__(256) k_ec_sweep(const uint8_t* __restrict__ mask, int h, int w,
                                                  uint8_t* __restrict__ ext,
                                                  int* __restrict__ changed) {
  __shared__ uint8_t E[EC_T + 2][EC_T + 2];
  __shared__ uint8_t M[EC_T][EC_T];
  const int tx0 = blockIdx.x * EC_T, ty0 = blockIdx.y * EC_T;
  for (int e = threadIdx.x; e < (EC_T + 2) * (EC_T + 2); e += blockDim.x) {
    const int ly = e / (EC_T + 2), lx = e % (EC_T + 2);
    const int y = ty0 + ly - 1, x = tx0 + lx - 1;
    E[ly][lx] = (y >= 0 && y < h && x >= 0 && x < w) ? ext[(int64_t)y * w + x] : 0;
  }
  for (int e = threadIdx.x; e < EC_T * EC_T; e += blockDim.x) {
    const int ly = e / EC_T, lx = e % EC_T;
    const int y = ty0 + ly, x = tx0 + lx;
    M[ly][lx] = (y < h && x < w) ? mask[(int64_t)y * w + x] : 0;
  }
  __syncthreads();
  bool gained = false;
  for (;;) {
    bool ch = false;
    for (int e = threadIdx.x; e < EC_T * EC_T; e += blockDim.x) {
      const int ly = e / EC_T, lx = e % EC_T;
      if (M[ly][lx] && !E[ly + 1][lx + 1] &&
          (E[ly][lx + 1] | E[ly + 2][lx + 1] | E[ly + 1][lx] | E[ly + 1][lx + 2])) {
        E[ly + 1][lx + 1] = 1;
        ch = true;
      }
    }
    gained |= ch;
    if (!__syncthreads_or(ch)) break;
  }
  for (int e = threadIdx.x; e < EC_T * EC_T; e += blockDim.x) {
    const int ly = e / EC_T, lx = e % EC_T;
    const int y = ty0 + ly, x = tx0 + lx;
    if (y < h && x < w && E[ly + 1][lx + 1]) ext[(int64_t)y * w + x] = 1;
  }
  if (__syncthreads_or(gained) && threadIdx.x == 0) atomicOr(changed, 1);
}

__global__ void k_set_i32(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ------------------------------------------------------------------ the fill
struct FillArgs {
  const float* src;
  int64_t lds;
  int h, w;
  NodataList L;
  int do_fill, enclosed;
  const float* surface;    // ch x cw, row stride cw
  const uint8_t* ext;      // ch x cw or NULL
  int ch, cw;
  float fy, fx;            // f32(ch / H), f32(cw / W)
  float* out;
  int64_t ldo;
};

__device__ __forceinline__ float bilin_u8(const uint8_t* __restrict__ a, int64_t ld, const Lin& Y,
                                          const Lin& X) {
  double t = 0.0;
  t = t + ((double)a[(int64_t)Y.i0 * ld + X.i0] * Y.w0) * X.w0;
  t = t + ((double)a[(int64_t)Y.i0 * ld + X.i1] * Y.w0) * X.w1;
  t = t + ((double)a[(int64_t)Y.i1 * ld + X.i0] * Y.w1) * X.w0;
  t = t + ((double)a[(int64_t)Y.i1 * ld + X.i1] * Y.w1) * X.w1;
  return (float)t;
}

__global__ void k_prepare_fill(FillArgs A) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= A.w) return;
  const float cc = ((float)j + 0.5f) * A.fx - 0.5f;
  const Lin X = lin_at((double)cc, A.cw);
  for (int i = blockIdx.y; i < A.h; i += gridDim.y) {
    const float v = A.src[(int64_t)i * A.lds + j];
    float o = v;
    if (is_nodata(v, A.L)) {
      o = __builtin_nanf("");
      if (A.do_fill) {
        const float rr = ((float)i + 0.5f) * A.fy - 0.5f;
        const Lin Y = lin_at((double)rr, A.ch);
        const bool exterior = A.enclosed && A.ext && bilin_u8(A.ext, A.cw, Y, X) > 0.5f;
        if (!exterior) o = bilin(A.surface, A.cw, Y, X);
      }
    }
    A.out[(int64_t)i * A.ldo + j] = o;
  }
}

// four pixels per lane (float4) when rows and pointers are 16-byte aligned
__global__ void k_prepare_fill4(FillArgs A) {
  const int j0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (j0 >= A.w) return;
  for (int i = blockIdx.y; i < A.h; i += gridDim.y) {
    const float4 v = *(const float4*)(A.src + (int64_t)i * A.lds + j0);
    float o[4] = {v.x, v.y, v.z, v.w};
    bool m[4];
    bool any = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      m[k] = is_nodata(o[k], A.L);
      any = any || m[k];
    }
    if (any) {
      const float rr = ((float)i + 0.5f) * A.fy - 0.5f;
      const Lin Y = lin_at((double)rr, A.ch);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (!m[k]) continue;
        o[k] = __builtin_nanf("");
        if (A.do_fill) {
          const float cc = ((float)(j0 + k) + 0.5f) * A.fx - 0.5f;
          const Lin X = lin_at((double)cc, A.cw);
          const bool exterior = A.enclosed && A.ext && bilin_u8(A.ext, A.cw, Y, X) > 0.5f;
          if (!exterior) o[k] = bilin(A.surface, A.cw, Y, X);
        }
      }
    }
    *(float4*)(A.out + (int64_t)i * A.ldo + j0) = make_float4(o[0], o[1], o[2], o[3]);
  }
}

static dim3 grid_rows(int w, int h, int bx = 256) {
  return dim3((unsigned)((w + bx - 1) / bx), (unsigned)std::max(1, std::min(h, 16384)));
}

// device -> host read of one word on stream s (synchronising s)
static hipError_t read_word(void* dst, const void* src, size_t bytes, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

static int nodata_list(const float* vals, int n, NodataList* L, const char* who) {
  if (n < 0 || n > FSG_PREPARE_MAX_NODATA || (n > 0 && !vals))
    return fail(FSG_E_INVALID, std::string(who) + ": bad NoData list (at most " +
                                   std::to_string(FSG_PREPARE_MAX_NODATA) + " values)");
  L->n = 0;
  for (int k = 0; k < n; k++)
    if (std::isfinite(vals[k])) L->v[L->n++] = vals[k];
  return 0;
}

}  // namespace fsg

using namespace fsg;

extern "C" int fsg_prepare_sample(const fsg_window* src, int32_t oh, int32_t ow, float* out,
                                  int64_t ldo, void* stream) {
  int rc = validate_window(src, "fsg_prepare_sample");
  if (rc) return rc;
  if (oh <= 0 || ow <= 0 || !out || ldo < ow)
    return fail(FSG_E_INVALID, "fsg_prepare_sample: bad output shape");
  hipLaunchKernelGGL(k_nearest_sample, grid_rows(ow, oh), dim3(256), 0, as_stream(stream),
                     src->data, src->ld, src->h, src->w, oh, ow, out, ldo);
  FSG_CHECK_LAUNCH("k_nearest_sample");
  return FSG_OK;
}

extern "C" size_t fsg_value_mode_workspace_size(int64_t n) {
  if (n <= 0 || n > 0x7fffffff) return 0;
  size_t need = 0;
  mode_layout(n, nullptr, 0, nullptr, &need);
  return need;
}

extern "C" int fsg_value_mode(const float* values, int64_t n, int32_t has_declared,
                              float declared, fsg_mode_stats* stats, void* workspace,
                              size_t workspace_bytes, void* stream) {
  if (!values || n <= 0 || n > 0x7fffffff || !stats)
    return fail(FSG_E_INVALID, "fsg_value_mode: bad arguments (0 < n < 2^31)");
  ModeWs W{};
  const int lay = mode_layout(n, workspace, workspace_bytes, &W, nullptr);
  if (lay == 2) return fail(FSG_E_HIP, "fsg_value_mode: hipCUB size query failed");
  if (lay) return fail(FSG_E_WORKSPACE, "fsg_value_mode: workspace too small");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, W.n_valid, 1);
  FSG_CHECK_LAUNCH("k_zero_u64");
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_mode_prep, dim3(blocks), dim3(256), 0, s, values, n, has_declared,
                     declared, W.keys, W.counts, W.n_valid);
  FSG_CHECK_LAUNCH("k_mode_prep");
  // invalid keys sort last as NaN, each its own run of 1 (NaN != NaN), so the
  // first arg-max over all n run counts is the mode of the valid values
  // whenever there is one; no host read is needed before the finish kernel
  size_t tb = W.temp_bytes;
  const int ni = (int)n;
  if (check_hip(hipcub::DeviceRadixSort::SortKeys(W.temp, tb, W.keys, W.sorted, ni, 0, 32, s),
                "fsg_value_mode: sort"))
    return FSG_E_HIP;
  tb = W.temp_bytes;
  if (check_hip(hipcub::DeviceRunLengthEncode::Encode(W.temp, tb, W.sorted, W.uniq, W.counts,
                                                      W.n_runs, ni, s),
                "fsg_value_mode: run-length encode"))
    return FSG_E_HIP;
  tb = W.temp_bytes;
  if (check_hip(hipcub::DeviceReduce::ArgMax(W.temp, tb, W.counts, W.best, ni, s),
                "fsg_value_mode: arg-max"))
    return FSG_E_HIP;
  hipLaunchKernelGGL(k_mode_finish, dim3(1), dim3(64), 0, s, W.sorted, W.uniq, W.n_runs, W.best,
                     W.n_valid, n, stats);
  FSG_CHECK_LAUNCH("k_mode_finish");
  return FSG_OK;
}

extern "C" int fsg_count_equal(const float* values, int64_t n, float value,
                               uint64_t* count, void* stream) {
  if (!values || n < 0 || !count) return fail(FSG_E_INVALID, "fsg_count_equal: bad arguments");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, (unsigned long long*)count, 1);
  FSG_CHECK_LAUNCH("k_zero_u64");
  if (n == 0) return FSG_OK;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_count_equal, dim3(blocks), dim3(256), 0, s, values, n, value,
                     (unsigned long long*)count);
  FSG_CHECK_LAUNCH("k_count_equal");
  return FSG_OK;
}

extern "C" int fsg_coarse_average(const float* sample, int64_t lds, int32_t ch, int32_t cw,
                                  int32_t k, const float* nodata, int32_t n_nodata, float* coarse,
                                  int64_t ldc, uint8_t* cvalid, uint64_t* n_void, void* stream) {
  if (!sample || !coarse || !cvalid || !n_void || ch <= 0 || cw <= 0 || ch > 65535 || k <= 0 ||
      lds < (int64_t)cw * k || ldc < cw)
    return fail(FSG_E_INVALID, "fsg_coarse_average: bad arguments");
  NodataList L{};
  if (int rc = nodata_list(nodata, n_nodata, &L, "fsg_coarse_average")) return rc;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, (unsigned long long*)n_void, 1);
  FSG_CHECK_LAUNCH("k_zero_u64");
  hipLaunchKernelGGL(k_coarse_average, dim3((cw + 255) / 256, ch), dim3(256), 0, s, sample, lds,
                     ch, cw, k, L, coarse, ldc, cvalid, (unsigned long long*)n_void);
  FSG_CHECK_LAUNCH("k_coarse_average");
  return FSG_OK;
}

// push-pull level shapes are decided on the host from each level's min weight
extern "C" size_t fsg_pushpull_workspace_size(int32_t h, int32_t w) {
  if (h <= 0 || w <= 0) return 0;
  // vw and w of every level (a geometric series below 2 h w each), two
  // filled buffers of the largest level, the min word
  size_t total = 0;
  int ch = h, cw = w;
  for (;;) {
    total += 2 * align256(sizeof(float) * (size_t)ch * cw);
    if (std::max(ch, cw) <= 1) break;
    ch = std::max(1, ch / 2);
    cw = std::max(1, cw / 2);
  }
  total += 2 * align256(sizeof(float) * (size_t)h * w) + 256;   // filled x 2, 64 minima
  return total;
}

extern "C" int fsg_pushpull_fill(const float* coarse, int64_t ldc, const uint8_t* valid,
                                 int32_t h, int32_t w, int32_t n_valid_cells, float* out,
                                 int64_t ldo, void* workspace, size_t workspace_bytes,
                                 void* stream) {
  if (!coarse || !valid || !out || h <= 0 || w <= 0 || ldc < w || ldo < w)
    return fail(FSG_E_INVALID, "fsg_pushpull_fill: bad arguments");
  const size_t need = fsg_pushpull_workspace_size(h, w);
  if (!workspace || workspace_bytes < need)
    return fail(FSG_E_WORKSPACE, "fsg_pushpull_fill: workspace too small");
  hipStream_t s = as_stream(stream);
  const int64_t n0 = (int64_t)h * w;
  if (n_valid_cells == 0) {   // no reference data: zeros
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, s,
                       (float*)workspace, n0, 0.0f);
    FSG_CHECK_LAUNCH("k_fill_f32");
    hipLaunchKernelGGL(k_pp_final, grid_rows(w, h), dim3(256), 0, s, (const float*)workspace,
                       (int64_t)w, valid, (const float*)workspace, h, w, out, ldo);
    FSG_CHECK_LAUNCH("k_pp_final");
    return FSG_OK;
  }
  char* p = (char*)workspace;
  std::vector<float*> vws, wts;
  std::vector<int> hs, wsz;
  auto take = [&](size_t b) {
    char* q = p;
    p += align256(b);
    return (float*)q;
  };
  hs.push_back(h);
  wsz.push_back(w);
  vws.push_back(take(sizeof(float) * n0));
  wts.push_back(take(sizeof(float) * n0));
  hipLaunchKernelGGL(k_pp_init, grid_rows(w, h), dim3(256), 0, s, coarse, ldc, valid, h, w,
                     vws[0], wts[0]);
  FSG_CHECK_LAUNCH("k_pp_init");
  if (n_valid_cells == h * w) {   // nothing to fill
    hipLaunchKernelGGL(k_pp_final, grid_rows(w, h), dim3(256), 0, s, coarse, ldc, valid, vws[0],
                       h, w, out, ldo);
    FSG_CHECK_LAUNCH("k_pp_final");
    return FSG_OK;
  }
  // level storage follows in the same order the size function counted it
  std::vector<char*> slots;
  {
    int ch = h, cw = w;
    for (;;) {
      if (std::max(ch, cw) <= 1) break;
      ch = std::max(1, ch / 2);
      cw = std::max(1, cw / 2);
      slots.push_back(p);
      p += 2 * align256(sizeof(float) * (size_t)ch * cw);
    }
  }
  float* fa = take(sizeof(float) * n0);
  float* fb = take(sizeof(float) * n0);
  unsigned* minw = (unsigned*)p;   // one word per level (at most 64 levels)
  const int n_lv = (int)slots.size() + 1;
  // every level down to 1 x 1 is pushed, with each level's min weight; the
  // reference stops at the first level whose min weight exceeds 1e-6, so the
  // levels past it are dropped after one read of the minima
  hipLaunchKernelGGL(k_set_u32n, dim3(1), dim3(64), 0, s, minw, n_lv, 0x7f800000u);
  FSG_CHECK_LAUNCH("k_set_u32n");
  for (int l = 0;; l++) {
    const int ch = hs.back(), cw = wsz.back();
    const int64_t n = (int64_t)ch * cw;
    hipLaunchKernelGGL(k_min_nonneg, dim3((unsigned)std::min<int64_t>((n + 1023) / 1024, 512)),
                       dim3(256), 0, s, wts.back(), n, minw + l);
    FSG_CHECK_LAUNCH("k_min_nonneg");
    if (std::max(ch, cw) <= 1) break;
    const int nh = std::max(1, ch / 2), nw = std::max(1, cw / 2);
    // scipy zoom: output shape round(n * (m / n)), coordinate scale (n - 1) / (m' - 1)
    const int zh = (int)std::nearbyint((double)ch * ((double)nh / (double)ch));
    const int zw = (int)std::nearbyint((double)cw * ((double)nw / (double)cw));
    if (zh != nh || zw != nw)
      return fail(FSG_E_UNSUPPORTED, "fsg_pushpull_fill: zoom shape differs from the level");
    const double zy = zh > 1 ? (double)(ch - 1) / (double)(zh - 1) : 1.0;
    const double zx = zw > 1 ? (double)(cw - 1) / (double)(zw - 1) : 1.0;
    float* nvw = (float*)slots[l];
    float* nwt = (float*)(slots[l] + align256(sizeof(float) * (size_t)nh * nw));
    hipLaunchKernelGGL(k_pp_push, grid_rows(nw, nh), dim3(256), 0, s, vws.back(), wts.back(), ch,
                       cw, nvw, nwt, nh, nw, zy, zx);
    FSG_CHECK_LAUNCH("k_pp_push");
    vws.push_back(nvw);
    wts.push_back(nwt);
    hs.push_back(nh);
    wsz.push_back(nw);
  }
  std::vector<unsigned> mins(n_lv);
  if (hipError_t e = read_word(mins.data(), minw, sizeof(unsigned) * n_lv, s))
    return check_hip(e, "fsg_pushpull_fill: reading the level minima");
  // levels kept: 0 .. L where L is the first level that stops the push
  int L = 0;
  while (L + 1 < n_lv) {
    float mf;
    std::memcpy(&mf, &mins[L], sizeof(mf));
    if (!((double)mf <= 1e-6)) break;   // float(ws.min()) <= 1e-6 in float64
    L++;
  }
  vws.resize(L + 1);
  wts.resize(L + 1);
  hs.resize(L + 1);
  wsz.resize(L + 1);
  // pull from the top level down
  int fh = hs.back(), fw = wsz.back();
  float* cur = fa;
  float* nxt = fb;
  hipLaunchKernelGGL(k_pp_top, dim3((unsigned)(((int64_t)fh * fw + 255) / 256)), dim3(256), 0, s,
                     vws.back(), wts.back(), (int64_t)fh * fw, cur);
  FSG_CHECK_LAUNCH("k_pp_top");
  for (int l = (int)vws.size() - 2; l >= 0; l--) {
    const int th = hs[l], tw = wsz[l];
    const int zh = (int)std::nearbyint((double)fh * ((double)th / (double)fh));
    const int zw = (int)std::nearbyint((double)fw * ((double)tw / (double)fw));
    const double zy = zh > 1 ? (double)(fh - 1) / (double)(zh - 1) : 1.0;
    const double zx = zw > 1 ? (double)(fw - 1) / (double)(zw - 1) : 1.0;
    hipLaunchKernelGGL(k_pp_pull, grid_rows(tw, th), dim3(256), 0, s, cur, fh, fw,
                       std::max(1, zh), std::max(1, zw), zy, zx, vws[l], wts[l], th, tw, nxt);
    FSG_CHECK_LAUNCH("k_pp_pull");
    std::swap(cur, nxt);
    fh = th;
    fw = tw;
  }
  hipLaunchKernelGGL(k_pp_final, grid_rows(w, h), dim3(256), 0, s, coarse, ldc, valid, cur, h, w,
                     out, ldo);
  FSG_CHECK_LAUNCH("k_pp_final");
  return FSG_OK;
}

extern "C" int fsg_edge_connected(const uint8_t* mask, int32_t h, int32_t w, uint8_t* ext,
                                  int32_t* flag, int32_t* sweeps, void* stream) {
  if (!mask || !ext || !flag || h <= 0 || w <= 0)
    return fail(FSG_E_INVALID, "fsg_edge_connected: bad arguments");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_ec_seed, grid_rows(w, h), dim3(256), 0, s, mask, h, w, ext);
  FSG_CHECK_LAUNCH("k_ec_seed");
  const dim3 grid((w + EC_T - 1) / EC_T, (h + EC_T - 1) / EC_T);
  // every sweep that changes something turns on at least one cell
  const int64_t max_sweeps = (int64_t)h * w + 2;
  int64_t n = 0;
  for (;;) {
    // four sweeps per host read; the flag read is the last sweep's
    for (int b = 0; b < 4; b++) {
      hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(64), 0, s, flag, 0);
      FSG_CHECK_LAUNCH("k_set_i32");
      hipLaunchKernelGGL(k_ec_sweep, grid, dim3(256), 0, s, mask, h, w, ext, flag);
      FSG_CHECK_LAUNCH("k_ec_sweep");
      n++;
    }
    int f = 0;
    if (hipError_t e = read_word(&f, flag, sizeof(f), s))
      return check_hip(e, "fsg_edge_connected: reading the sweep flag");
    if (!f) break;
    if (n > max_sweeps) return fail(FSG_E_HIP, "fsg_edge_connected: no fixed point");
  }
  if (sweeps) *sweeps = (int32_t)std::min<int64_t>(n, 0x7fffffff);
  return FSG_OK;
}

extern "C" int fsg_prepare_fill(const fsg_window* src, const float* nodata, int32_t n_nodata,
                                int32_t fill_mode, int32_t do_fill, const float* surface,
                                const uint8_t* exterior, int32_t ch, int32_t cw, float* out,
                                int64_t ldo, void* stream) {
  int rc = validate_window(src, "fsg_prepare_fill");
  if (rc) return rc;
  if (!out || ldo < src->w || fill_mode < FSG_FILL_NONE || fill_mode > FSG_FILL_ALL)
    return fail(FSG_E_INVALID, "fsg_prepare_fill: bad output or fill mode");
  if (do_fill && (!surface || ch <= 0 || cw <= 0 ||
                  (fill_mode == FSG_FILL_ENCLOSED && !exterior)))
    return fail(FSG_E_INVALID, "fsg_prepare_fill: the fill needs the surface (and exterior)");
  FillArgs A{};
  if ((rc = nodata_list(nodata, n_nodata, &A.L, "fsg_prepare_fill"))) return rc;
  A.src = src->data;
  A.lds = src->ld;
  A.h = src->h;
  A.w = src->w;
  A.do_fill = do_fill && fill_mode != FSG_FILL_NONE;
  A.enclosed = fill_mode == FSG_FILL_ENCLOSED;
  A.surface = surface;
  A.ext = exterior;
  A.ch = ch;
  A.cw = cw;
  A.fy = (float)((double)ch / (double)src->h);
  A.fx = (float)((double)cw / (double)src->w);
  A.out = out;
  A.ldo = ldo;
  TimedScope ts("k_prepare_fill", as_stream(stream));
  const bool vec = src->w % 4 == 0 && src->ld % 4 == 0 && ldo % 4 == 0 &&
                   ((uintptr_t)src->data & 15) == 0 && ((uintptr_t)out & 15) == 0;
  if (vec) {
    hipLaunchKernelGGL(k_prepare_fill4, grid_rows(src->w / 4, src->h), dim3(256), 0,
                       as_stream(stream), A);
    FSG_CHECK_LAUNCH("k_prepare_fill4");
  } else {
    hipLaunchKernelGGL(k_prepare_fill, grid_rows(src->w, src->h), dim3(256), 0,
                       as_stream(stream), A);
    FSG_CHECK_LAUNCH("k_prepare_fill");
  }
  return FSG_OK;
}
