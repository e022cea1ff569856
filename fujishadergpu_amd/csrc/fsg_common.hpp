// Shared host/device helpers of libfsg (MI355X / gfx950).
//
// Arithmetic policy: the library is built with -ffp-contract=off so every
// float32/float64 expression rounds after each operation in source order,
// as the reference's one-ufunc-per-operation CuPy code does; '/' and sqrt are
// correctly rounded (hipcc default).  Transcendentals come from ROCm's device
// math library.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string>

#include "../../include/fsg.h"

namespace fsg {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);
void count_launch(int n = 1);

#define FSG_CHECK_LAUNCH(name)                                              \
  do {                                                                      \
    ::fsg::count_launch();                                                  \
    hipError_t _e = hipGetLastError();                                      \
    if (_e != hipSuccess) return ::fsg::check_hip(_e, name);                \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Optional per-kernel timing: when enabled (fsg_timing_enable), HIP events are
// recorded on the launch stream around named kernels; fsg_timing_read sums them.
bool timing_on();
void timing_start(const char* name, hipStream_t s);
void timing_stop(const char* name, hipStream_t s);
struct TimedScope {
  const char* name;
  hipStream_t s;
  bool on;
  TimedScope(const char* n, hipStream_t st) : name(n), s(st), on(timing_on()) {
    if (on) timing_start(name, s);
  }
  ~TimedScope() {
    if (on) timing_stop(name, s);
  }
};

// Descriptor i of an array the host wrote before the launch and no
// kernel writes, read through the constant address space: uniform fields take
// scalar loads, and no store of the kernel makes the compiler read them again
// (through a plain pointer every output store could alias them).
template <class T>
__device__ __forceinline__ const T& desc(const T* p, int i) {
  return *(const T*)((const __attribute__((address_space(4))) T*)p + i);
}

int validate_window(const fsg_window* w, const char* who);
int validate_region(const fsg_window* w, const fsg_region* r, const char* who);
int validate_output(const fsg_output* o, const fsg_region* r, const char* who);

// ------------------------------------------------------ boundary indexing
// scipy.ndimage 'reflect' (half-sample symmetric: d c b a | a b c d | d c b a)
__host__ __device__ inline int refl(int i, int n) {
  if (n <= 1) return 0;
  const int period = 2 * n;
  i %= period;
  if (i < 0) i += period;
  return i < n ? i : period - 1 - i;
}
// scipy.ndimage 'nearest' (edge value repeated)
__host__ __device__ inline int clampi(int i, int n) {
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// --------------------------------------------------------- output epilogue
// Device-side copy of fsg_output with the statistics resolved to floats.
struct OutDev {
  void* data;
  int64_t ld;
  int dtype;
  int clip01;
  float a, b;
  float dn_min, dn_max;
  int post;
  const double* dev_stats;
  double host_stats0, host_stats1;
};

inline OutDev make_out(const fsg_output* o) {
  OutDev d{};
  if (!o) return d;
  d.data = o->data;
  d.ld = o->ld;
  d.dtype = o->dtype;
  d.clip01 = o->clip01;
  d.a = o->a_coef;
  d.b = o->b_coef;
  d.dn_min = (float)o->dn_min;
  d.dn_max = (float)o->dn_max;
  d.post = o->post_kind;
  d.dev_stats = o->dev_stats;
  d.host_stats0 = o->host_stats[0];
  d.host_stats1 = o->host_stats[1];
  return d;
}

// Post-processing constants (uniform per launch).
struct PostK {
  int kind;     // 0 none, 1 div, 2 stretch (active)
  float lo, sc; // f32(lo), f32(scale)
  int zero;     // div with scale <= 0 -> zeros
};

__device__ inline PostK load_post(const OutDev& o) {
  PostK p{0, 0.f, 1.f, 0};
  if (o.data == nullptr || o.post == FSG_POST_NONE) return p;
  double s0 = o.dev_stats ? o.dev_stats[0] : o.host_stats0;
  double s1 = o.dev_stats ? o.dev_stats[1] : o.host_stats1;
  if (o.post == FSG_POST_DIV) {
    p.kind = 1;
    p.sc = (float)s0;
    p.zero = !(s0 > 0.0);
  } else if (o.post == FSG_POST_STRETCH) {
    if (s1 > 1e-12) {
      p.kind = 2;
      p.lo = (float)s0;
      p.sc = (float)s1;
    }
  }
  return p;
}

__device__ inline float apply_post(const PostK& p, float v) {
  if (p.kind == 1) {
    if (isnan(v)) return v;
    return p.zero ? 0.0f : v / p.sc;
  }
  if (p.kind == 2) {
    float t = (v - p.lo) / p.sc;
    return isnan(t) ? t : fmaxf(t, 0.0f);
  }
  return v;
}

// DN = clip(rint(f32(a)*v + f32(b)), dn_min, dn_max); non-finite -> 0.
__device__ inline float quantize_dn(const OutDev& o, float v) {
  if (!isfinite(v)) return 0.0f;
  float t = o.a * v;   // rounded f32 multiply (no contraction)
  t = t + o.b;         // rounded f32 add
  t = rintf(t);        // ties to even
  return fminf(fmaxf(t, o.dn_min), o.dn_max);
}

// Write one value (already post-processed) at region coordinates (r, c).
__device__ inline void store_out(const OutDev& o, int r, int c, float v) {
  if (o.clip01 && !isnan(v)) v = fminf(fmaxf(v, 0.0f), 1.0f);
  const int64_t off = (int64_t)r * o.ld + c;
  if (o.dtype == FSG_OUT_F32) {
    ((float*)o.data)[off] = v;
  } else if (o.dtype == FSG_OUT_I16) {
    ((int16_t*)o.data)[off] = (int16_t)(int)quantize_dn(o, v);
  } else {
    ((uint8_t*)o.data)[off] = (uint8_t)(int)quantize_dn(o, v);
  }
}

// ---------------------------------------------- numpy-compatible host math
// numpy pairwise summation (float64, contiguous), used for kernel weights.
double np_pairwise_sum(const double* a, int64_t n);
// scipy.ndimage._gaussian_kernel1d(sigma, 0, radius) (symmetric, normalised).
void gaussian_weights(double sigma, int radius, double* w /* 2*radius+1 */);
inline int gaussian_radius(double sigma, double truncate = 4.0) {
  return (int)(truncate * sigma + 0.5);
}

// ---------------------------------------------- shared device-side types
// A window as the kernels see it (row-major float32 view of the raster).
struct WinDev {
  const float* p;
  int64_t ld;
  int h, w;
};

// Gaussian sigma=1, truncate=4 weights: g[0] centre, g[d] distance d.
struct Gauss9 {
  double g[5];
};

inline Gauss9 make_gauss9() {
  double w[9];
  gaussian_weights(1.0, 4, w);
  Gauss9 G;
  for (int d = 0; d <= 4; d++) G.g[d] = w[4 + d];
  return G;
}

}  // namespace fsg
