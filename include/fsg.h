/*
 * fsg.h -- C ABI of the MI355X-native FujiShaderGPU raster hot path.
 *
 * Every entry point is extern "C", takes plain pointers, sizes and POD
 * structs, and never allocates or frees caller memory.  Device buffers are
 * owned by the caller (e.g. torch tensors, via data_ptr()).  All work is
 * enqueued on the caller's HIP stream (`stream`, a hipStream_t passed as
 * void*; NULL = the default stream) and is stream-ordered; nothing blocks the
 * host except where a function says so.
 *
 * Return codes: FSG_OK (0) or a positive FSG_E_* code; fsg_last_error()
 * returns a thread-local message for the last failure.  No C++ exception
 * crosses this boundary.
 *
 * Geometry convention: a *window* is the array a per-tile call receives (tile
 * core + halo, clipped at the raster edge), exactly as the reference's
 * TileAlgorithm.process(dem_gpu) receives it.  A *region* is the sub-rectangle
 * of the window whose output is written; the reference always computes the
 * whole window and crops the core afterwards
 * (core/tile_processor.py:944-971), so region == whole window reproduces
 * process() and region == core fuses the crop.  Results inside the region are
 * identical either way: every window-edge rule (one-sided gradients, reflect /
 * nearest padding, out-of-window rays) is evaluated against the window.
 */
#ifndef FSG_H
#define FSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSG_ABI_VERSION 20

enum {
  FSG_OK = 0,
  FSG_E_INVALID = 1,     /* bad argument (shape, pointer, parameter) */
  FSG_E_HIP = 2,         /* a HIP runtime call failed */
  FSG_E_UNSUPPORTED = 3, /* valid request this build does not implement */
  FSG_E_WORKSPACE = 4    /* workspace too small (see *_workspace_size) */
};

/* Output encodings (io/output_encoding.py:75 SUPPORTED_OUTPUT_DTYPES). */
enum { FSG_OUT_F32 = 0, FSG_OUT_I16 = 1, FSG_OUT_U8 = 2 };

/* Post-processing applied to the float result before encoding. */
enum {
  FSG_POST_NONE = 0,
  FSG_POST_DIV = 1,     /* v / f32(scale)          (_normalization.py:35-41)   */
  FSG_POST_STRETCH = 2  /* max((v - f32(lo)) / f32(scale), 0), no-op unless
                           scale > 1e-12          (tile/dask_bridge.py:173-187) */
};

typedef struct fsg_window {
  const float* data; /* device pointer to window element (0, 0) */
  int64_t ld;        /* row stride in elements (>= w) */
  int32_t h, w;      /* window rows, columns */
} fsg_window;

typedef struct fsg_region {
  int32_t r0, c0, h, w; /* output rectangle, in window coordinates */
} fsg_region;

/*
 * Output descriptor: post-processing + encoding + NoData policy.
 *   float32: NaN stays NaN (NoData), optional [0,1] re-clip
 *            (core/tile_processor.py:606-624 _format_algorithm_output).
 *   int16/uint8: DN = clip(rint(f32(a)*v + f32(b)), dn_min, dn_max) with two
 *            rounded f32 operations and ties-to-even; non-finite v -> 0
 *            (io/output_encoding.py:178-190, core/dask_processor.py:997-1008).
 * post_kind uses host_stats, or dev_stats (device doubles) when non-NULL, so a
 * statistic computed on the GPU is consumed without a host round trip.
 */
typedef struct fsg_output {
  void* data;     /* device pointer to region element (0, 0) */
  int64_t ld;     /* row stride in elements of the output dtype */
  int32_t dtype;  /* FSG_OUT_* */
  int32_t clip01; /* 1: clip finite values to [0, 1] before encoding */
  float a_coef, b_coef;
  int32_t dn_min, dn_max;
  int32_t post_kind;        /* FSG_POST_* */
  int32_t pad0;
  double host_stats[2];     /* DIV: {scale}; STRETCH: {lo, scale} */
  const double* dev_stats;  /* optional device copy of host_stats */
} fsg_output;

/* ------------------------------------------------------------------ */
/* Library                                                             */
/* ------------------------------------------------------------------ */

int fsg_abi_version(void);
const char* fsg_last_error(void);
/* Number of kernels launched by this thread since load (testing aid). */
int64_t fsg_launch_count(void);
/*
 * Kernel timing for measurement: while enabled, HIP events are recorded on
 * the launch stream around the named hot kernels ("k_usm_main",
 * "k_hillshade_slope", "k_curvature", "k_openness", "k_ao_raw", ...).
 * fsg_timing_read synchronises those events, returns the summed milliseconds
 * and launch count for one name, and forgets them.  No reference counterpart
 * (the reference has no kernels of its own to time).
 */
int fsg_timing_enable(int32_t on);
int fsg_timing_read(const char* name, double* total_ms, int64_t* count);
/* The same, plus the length of the union of the launches' [start, end]
   intervals: launches on several streams may run at the same time, and then
   the union, not the sum, is the time the GPU spent on them. */
int fsg_timing_read_union(const char* name, double* total_ms, int64_t* count,
                          double* union_ms);

/* ------------------------------------------------------------------ */
/* Input preparation                                                   */
/* ------------------------------------------------------------------ */

/*
 * NoData sentinel -> NaN, in place, and per-tile NoData counts.
 * Replaces _build_nodata_mask + dem_tile[mask] = NaN
 * (core/tile_processor.py:185-196, :854-885): a cell is NoData when it is NaN
 * or |f64(v) - nodata| <= 1e-6 (numpy isclose, rtol=0, atol=1e-6).
 * has_nodata = 0 converts nothing (NaN cells still count).
 * counts (device int64[n_tiles], optional) receives the NoData count of each
 * window listed in windows (host array of int32[4] = {r0, c0, h, w} in raster
 * coordinates); counts are accumulated, so zero them first.
 */
int fsg_nodata_to_nan(float* raster, int64_t ld, int32_t h, int32_t w,
                      int32_t has_nodata, double nodata,
                      const int32_t* windows, int32_t n_windows,
                      int64_t* counts, void* stream);

/* ------------------------------------------------------------------ */
/* Gradient stencils                                                   */
/* ------------------------------------------------------------------ */

typedef struct fsg_shade_params {
  /* Derivative steps: |pixel_scale| (or pixel_size when the scale is None or
     below 1e-9), exactly as handle_nan_for_gradient resolves them
     (algorithms/_nan_utils.py:65-73). */
  double step_x, step_y;
  double sign_x, sign_y; /* +1/-1 east/north orientation (_impl_hillshade.py:39-42) */
  double z_factor;       /* elevation scale before the gradient (hillshade) */
  double azimuth_deg, altitude_deg;
  int32_t slope_unit;    /* 0 degree, 1 percent, 2 radian (_impl_slope.py:27-33) */
  int32_t nan_hint;      /* 0: window known NaN-free (no fill path); else NaN-aware */
  /* NULL, or a device array with the signed x step of every window row
     (geographic grids with per-row latitude, metres per pixel); it replaces
     step_x row by row (|value| here).  No reference counterpart: the
     reference uses one latitude per tile (core/tile_processor.py:910-932). */
  const double* step_x_rows;
} fsg_shade_params;

/*
 * Hillshade and/or slope from the central-difference gradient with
 * edge_order=2 window edges and the NaN-aware sigma=1 fill.
 * Replaces compute_hillshade_block (_impl_hillshade.py:20-54) and
 * compute_slope_block (_impl_slope.py:19-35); both outputs come from one
 * read of the window.  Either output may be NULL.  The window needs h, w >= 3
 * (numpy/cupy gradient edge_order=2 requirement).
 */
int fsg_hillshade_slope(const fsg_window* win, const fsg_region* reg,
                        const fsg_shade_params* p,
                        const fsg_output* hillshade, const fsg_output* slope,
                        void* stream);

typedef struct fsg_curvature_params {
  /* Signed steps (pixel_scale or pixel_size, with the 1e-9 fallback) as
     compute_curvature_block resolves them (_impl_curvature.py:23-28); the
     first derivatives use |step|, the second ones the signed step. */
  double step_x, step_y;
  int32_t curvature_type; /* 0 mean, 1 gaussian, 2 planform, 3 profile */
  int32_t nan_hint;       /* 0: window known NaN-free (no fill path); else NaN-aware */
  /* NULL, or the signed x step of every window row (device array), as in
     fsg_shade_params; first derivatives use |value|, second ones the value. */
  const double* step_x_rows;
} fsg_curvature_params;

/* Replaces compute_curvature_block (_impl_curvature.py:19-57). */
int fsg_curvature(const fsg_window* win, const fsg_region* reg,
                  const fsg_curvature_params* p, const fsg_output* out,
                  void* stream);

/*
 * Smallest interior (in pixels: window rows and columns where every
 * derivative is the central one, within the region) that fsg_hillshade_slope
 * and fsg_curvature give to the row-walking interior kernels, with one more
 * launch for the border strips; smaller ones take the general kernel alone,
 * one launch.  A process-wide setting with no reference counterpart, for
 * equality tests and tuning: results are identical either way.  px < 0
 * restores the default (2^23); returns the previous value.  Windows with
 * NaN take the row-walking kernels from min(px, 2^20) pixels.
 */
int64_t fsg_stencil_set_split_min(int64_t px);

/* ------------------------------------------------------------------ */
/* Ray marches                                                          */
/* ------------------------------------------------------------------ */

/*
 * Openness ray table, built by the caller exactly as
 * compute_openness_vectorized builds it (_impl_openness.py:58-109):
 * for direction k (0..n_dirs-1) the samples k*n_dist .. k*n_dist+n_dist-1 hold
 * integer offsets (dx = column, dy = row) and the physical distance
 * max(hypot(dx*|sx|, dy*|sy|), 1e-9).  Samples with dx == dy == 0 must be
 * marked skip = 1.
 */
typedef struct fsg_ray_sample {
  int32_t dx, dy;
  int32_t skip, pad0;
  double phys;
} fsg_ray_sample;

typedef struct fsg_openness_params {
  int32_t openness_type; /* 0 positive, 1 negative */
  int32_t n_dirs, n_dist;
  int32_t raw_degrees;   /* 0: display value (reference); 1: mean angle in degrees */
  const fsg_ray_sample* samples; /* host pointer, n_dirs*n_dist entries */
} fsg_openness_params;

/* Replaces compute_openness_vectorized (_impl_openness.py:31-132).  With
   raw_degrees the output is the mean openness angle in degrees, before the
   reference's / (pi/2), [0, 1] clip and gamma (the surrogate's target). */
int fsg_openness(const fsg_window* win, const fsg_region* reg,
                 const fsg_openness_params* p, const fsg_output* out,
                 void* stream);

/* The same openness at n window pixels (device int32 [n][2] = row, col) into
   float32 out[n] (the surrogate's sampled exact check).  A point outside the
   window (row not in [0, h) or col not in [0, w)) reads nothing and gets NaN. */
int fsg_openness_points(const fsg_window* win, const int32_t* points, int32_t n,
                        const fsg_openness_params* p, float* out, void* stream);

/* No reference counterpart: the learned surrogate's judge.  The window is
   cut into tile_h x tile_w tiles (row-major, ragged at the far edges); tile t
   gets n_t = max(1, ceil(sample_frac * th * tw / seg_len)) row segments of
   seg_len (1..64, at most every tile's width) pixels, segment k at
     h = splitmix64(splitmix64(seed ^ splitmix64(t)) + k),
     row = y0 + (h >> 32) % th,  col0 = x0 + (h & 0xffffffff) % (tw - seg_len + 1)
   (splitmix64: Steele, Lea & Flood 2014).  tile_mae[t] (device f32) receives
   the mean |exact openness - approx| over the segments' pixels where both are
   numbers (NaN when none).  approx is [h][lda] in window coordinates and is
   compared with the openness of p (normally raw_degrees).  Workspace:
   fsg_openness_judge_workspace_size bytes. */
size_t fsg_openness_judge_workspace_size(const fsg_window* win, int32_t tile_h, int32_t tile_w);
int fsg_openness_judge(const fsg_window* win, const fsg_openness_params* p, const float* approx,
                       int64_t lda, int32_t tile_h, int32_t tile_w, double sample_frac,
                       uint64_t seed, int32_t seg_len, float* tile_mae, void* workspace,
                       size_t workspace_bytes, void* stream);

/*
 * Ambient-occlusion sample table (_impl_ambient_occlusion.py:54-81): ring
 * r (0..3) sample i holds the np.round offsets, skip when both are 0, the
 * physical distance and the ring's distance factor 1 - 0.3*r_factor.
 */
typedef struct fsg_ao_sample {
  int32_t dx, dy;
  int32_t skip, pad0;
  double phys;
  double dist_factor;
} fsg_ao_sample;

typedef struct fsg_ao_params {
  int32_t n_samples; /* total table entries (4 rings x num_samples) */
  int32_t pad_d;     /* D = max(1, int(round(radius))) edge padding */
  double intensity;
  const fsg_ao_sample* samples; /* host pointer */
} fsg_ao_params;

/* Replaces compute_ambient_occlusion_block (_impl_ambient_occlusion.py:32-121). */
size_t fsg_ambient_occlusion_workspace_size(const fsg_window* win,
                                            const fsg_region* reg);
int fsg_ambient_occlusion(const fsg_window* win, const fsg_region* reg,
                          const fsg_ao_params* p, const fsg_output* out,
                          void* workspace, size_t workspace_bytes,
                          void* stream);

/* ------------------------------------------------------------------ */
/* TopoUSM Fast                                                         */
/* ------------------------------------------------------------------ */

#define FSG_MAX_RADII 16

typedef struct fsg_topousm_params {
  int32_t n_radii;
  int32_t radii[FSG_MAX_RADII];     /* small (in-window) radii, in order */
  float weights[FSG_MAX_RADII];     /* f32 weights (cp.array(weights, f32)) */
  int32_t ds[FSG_MAX_RADII];        /* _radius_to_downsample_factor per radius */
  /* Large-radius overview field (tile/dask_bridge.py:237-249); field == NULL
     disables it.  field is a device f32 grid of field_h x field_w. */
  const float* field;
  int32_t field_h, field_w;
  int32_t full_h, full_w;           /* raster shape the field spans */
  int32_t win_r0, win_c0;           /* window origin in raster coordinates */
  float w_large;                    /* f32(sum of the large weights) */
  /* 0: the caller knows the window holds no NaN (enables the plain-mean
     variant); any other value: NaN-aware means (always correct). */
  int32_t nan_hint;
  /* Optional precomputed coarse means, per decimated radius (NULL: built
     here from the window).  mean[i] points at cell (0, 0) of this window's
     ds-grid; ceil(h/ds) x ceil(w/ds) cells are read with leading dimension
     mean_ld[i].  The caller vouches that the values equal the window's own
     coarse means, e.g. views of fsg_topousm_coarse_means over the raster
     when the raster holds no NaN, the window origin and shape are multiples
     of ds and the region's samples never reach a window edge that is not a
     raster edge. */
  const float* mean[FSG_MAX_RADII];
  int64_t mean_ld[FSG_MAX_RADII];
} fsg_topousm_params;

/*
 * Replaces compute_topousm_fast_efficient_block (_impl_topousm_fast.py:49-100)
 * plus the large-radius part and normalisation of _direct_topousm_fast
 * (tile/dask_bridge.py:225-271).  Normalisation is out->post_kind = DIV.
 * out2 (optional) receives the same post-processed values (out's post_kind
 * and statistics apply to both) in a second encoding, e.g. the float32 and the
 * uint8 product of one pass.
 */
size_t fsg_topousm_workspace_size(const fsg_window* win,
                                  const fsg_topousm_params* p);
int fsg_topousm(const fsg_window* win, const fsg_region* reg,
                const fsg_topousm_params* p, const fsg_output* out,
                const fsg_output* out2, /* optional second encoding */
                void* workspace, size_t workspace_bytes, void* stream);

/*
 * Which main-pass kernels fsg_topousm uses (a process-wide setting, for
 * benchmarks and equality tests; no reference counterpart): 0 (default) the
 * single-wavefront kernel wherever it applies (NaN-free windows, terms in
 * ascending radius order, 1-3 box radii up to 32, at most 3 bilinear
 * sources), 1 the workgroup kernels everywhere, 8..1024 the single-wavefront
 * kernel with that many rows per block (launch-shape sweeps).  Results are
 * identical.
 */
int fsg_topousm_set_variant(int32_t variant);

/*
 * fsg_topousm without a workspace argument and without any allocation, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, hosts that forbid hipMalloc).  No counterpart in the
 * reference.  The coarse grids live in a library-owned device array of
 * FSG_TOPOUSM_SCRATCH_BYTES (64 MiB) per GPU, held until the process ends;
 * a window whose fsg_topousm_workspace_size exceeds it is rejected with
 * FSG_E_WORKSPACE.  Every call on one GPU shares that array (each call
 * re-initialises what it reads), so calls that overlap on one GPU -- from
 * two streams or two host threads -- race; serialise them on one stream.
 * Results equal fsg_topousm's bit for bit.
 */
#define FSG_TOPOUSM_SCRATCH_BYTES (64u << 20)
int fsg_topousm_noalloc(const fsg_window* win, const fsg_region* reg,
                        const fsg_topousm_params* p, const fsg_output* out,
                        const fsg_output* out2, void* stream);

/*
 * The coarse means fsg_topousm builds per window for its decimated radii,
 * for a whole window (typically the raster): the NaN-aware ds x ds cell means
 * with the enclosed-void fill of _downsample_nan_aware (_nan_utils.py:604-668),
 * then per radius the NaN-aware box mean of radius radii_coarse[i]
 * (max(1, round(r / ds)); <= 1 selects the sigma=1 Gaussian) with 'reflect'
 * edges, as compute_topousm_fast_efficient_block does
 * (_impl_topousm_fast.py:72-91).  outs[i] are ceil(h/ds) x ceil(w/ds) grids
 * with leading dimension ld_out; ds is a power of two <= 16, or 1 when the
 * window already is the cell grid (no decimation: box means only).
 */
size_t fsg_topousm_coarse_means_workspace_size(const fsg_window* win, int32_t ds);
int fsg_topousm_coarse_means(const fsg_window* win, int32_t ds, int32_t n_radii,
                             const int32_t* radii_coarse, float* const* outs,
                             int64_t ld_out, void* workspace, size_t workspace_bytes,
                             void* stream);

/*
 * The same coarse means for many NaN-free windows in a few launches: every
 * decimation in one launch (two jobs on one window whose factors differ by 2x
 * or 4x share one read of it), then every (job, radius) box pass 0 in one
 * launch and every pass 1 in one more.  Outputs equal
 * fsg_topousm_coarse_means' per job bit for bit.  The caller vouches that each
 * window holds no NaN (the enclosed-void fill is not run); ds is 4, 8 or 16,
 * or 1 when the window already is the decimated grid (box passes only), and
 * every coarse radius >= 2 (FSG_E_UNSUPPORTED otherwise: those take
 * fsg_topousm_coarse_means).  Replaces, for the windows of the stratified
 * statistics pre-pass (algorithms/_norm_stats.py:176-350) and the raster-wide
 * grids, the per-window decimate + uniform_filter sequence of
 * compute_topousm_fast_efficient_block (_impl_topousm_fast.py:72-91).
 */
#define FSG_COARSE_JOB_MAX_RADII 4
typedef struct fsg_coarse_job {
  fsg_window win;
  int32_t ds;
  int32_t n_radii;
  int32_t radii[FSG_COARSE_JOB_MAX_RADII];   /* coarse radii (max(1, round(r / ds))) */
  float* outs[FSG_COARSE_JOB_MAX_RADII];     /* ceil(h/ds) x ceil(w/ds) grids */
  int64_t ld_out;
} fsg_coarse_job;
size_t fsg_topousm_coarse_means_batch_workspace_size(const fsg_coarse_job* jobs, int32_t n);
int fsg_topousm_coarse_means_batch(const fsg_coarse_job* jobs, int32_t n, void* workspace,
                                   size_t workspace_bytes, void* stream);

/*
 * The large-radius field of fsg_topousm_coarse_field from box means already
 * computed (e.g. by fsg_topousm_coarse_means_batch over a NaN-free overview,
 * ds = 1): out = sum_i f32(weights[i]) * means[i] in radius order, with
 * fsg_topousm_coarse_field's arithmetic (_impl_topousm_fast.py:133-155).
 * n <= 8 grids of h x w.
 */
int fsg_topousm_field_from_means(const float* const* means, const int64_t* ld_means,
                                 const float* weights, int32_t n, int32_t h, int32_t w,
                                 float* out, int64_t ld_out, void* stream);

/*
 * One read of a raster for what the tile pipeline needs of it before its
 * first tile: the number of NaN pixels (added to *nan_count, device memory,
 * which the caller zeroes) -- the all-zero test of the per-window NoData
 * counts (core/tile_processor.py:864-874) -- and NaN-aware block means at up
 * to three factors 4, 8 or 16: the integer-decimation overview
 * (core/tile_processor.py:575-590) and the decimated grids of
 * _downsample_nan_aware (_nan_utils.py:620-643) before their void fill.
 * outs[k].out is ceil(h/f) x ceil(w/f) with leading dimension ld_out; every
 * cell equals fsg_block_mean's.
 */
typedef struct fsg_scan_out {
  int32_t factor;
  float* out;
  int64_t ld_out;
} fsg_scan_out;
int fsg_raster_scan(const fsg_window* raster, int32_t n_out, const fsg_scan_out* outs,
                    uint64_t* nan_count, void* stream);

/*
 * NaN-aware block mean (area average for an integer factor), the restatement
 * of the overview read with Resampling.average
 * (core/tile_processor.py:575-590) and of _downsample_nan_aware's cell mean
 * (_nan_utils.py:620-643).  out is ceil(h/f) x ceil(w/f); ragged cells
 * average their in-raster cells; all-NaN cells become NaN.
 */
int fsg_block_mean(const fsg_window* src, int32_t factor, float* out,
                   int64_t ld_out, void* stream);

/*
 * The overview read of the large-radius paths at any decimation: rasterio's
 * src.read(out_shape=(oh, ow), resampling=Resampling.average, masked=True)
 * (core/tile_processor.py:575-590, read_overview_coarse_dem
 * _nan_utils.py:836-871), restated as an area-weighted average over the
 * full-resolution raster: output cell (i, j) covers source rows
 * [i*h/oh, (i+1)*h/oh) and columns [j*w/ow, (j+1)*w/ow), every source pixel
 * weighs its overlap length on each axis, NaN pixels are left out, and a cell
 * with no finite pixel is NaN.  Sums and the division in float64, rounded
 * once to float32.  For an integer ratio that is the float64 mean of each
 * cell's finite pixels (fsg_block_mean rounds its sum first, so the two can
 * differ by one float32 ulp where a cell's finite count is not a power of
 * two).  Needs 1 <= oh <= h and 1 <= ow <= w.
 */
int fsg_overview_average(const fsg_window* src, int32_t oh, int32_t ow, float* out,
                         int64_t ld_out, void* stream);

/*
 * Large-radius coarse field sum_i w_i * mean_{r_i}(coarse)
 * (_impl_topousm_fast.py:133-155).  radii are already scaled into the coarse
 * grid (max(1, round(r / decimation))).  out must not alias coarse.
 */
size_t fsg_topousm_field_workspace_size(int32_t h, int32_t w);
int fsg_topousm_coarse_field(const fsg_window* coarse, int32_t n_radii,
                             const int32_t* radii_coarse,
                             const float* weights, float* out, int64_t ld_out,
                             void* workspace, size_t workspace_bytes,
                             void* stream);

/* ------------------------------------------------------------------ */
/* Statistics                                                           */
/* ------------------------------------------------------------------ */

/*
 * Order statistics of the finite values of a device f32 array, exactly:
 * radix selection on the order-preserving bit pattern, then numpy's
 * method='linear' interpolation in FP64 (cp.percentile as called by
 * topousm_fast_stat_func, _normalization.py:22-32, and
 * robust_unsigned_stretch_stat_func, _global_stats.py:181-203).
 * flags: FSG_PCT_ABS selects on |v| (topousm_fast_stat_func); FSG_PCT_FINITE
 * keeps only finite values (robust_unsigned_stretch_stat_func uses isfinite,
 * topousm_fast_stat_func only drops NaN).  q (device double[nq], nq <= 2) are
 * percentages in [0, 100]; out (device double[nq]) receives the percentiles;
 * n_valid (device int64, optional) the selected count; moments (device
 * double[2], optional) the FP64 mean and population std of the selected
 * (signed) values -- the std fallback of topousm_fast_stat_func.
 */
enum { FSG_PCT_ABS = 1, FSG_PCT_FINITE = 2 };
size_t fsg_percentile_workspace_size(int64_t n);
int fsg_percentile(const float* values, int64_t n, int32_t flags,
                   const double* q, int32_t nq, double* out, int64_t* n_valid,
                   double* moments, void* workspace, size_t workspace_bytes,
                   void* stream);

/*
 * Finalise device statistics in place for the normalisation epilogue:
 *   kind 0 (topousm): stats[0] = p99 > 1e-9 ? p99 : (std > 1e-9 ? std : 1.0)
 *   kind 1 (stretch): (lo, scale) = (p1, p99 - p1); scale 0 (stretch off) when
 *                     not > 1e-9, as the pre-pass drops such stats
 *                     (_norm_stats.py:288-289)
 * pct holds the percentiles from fsg_percentile, moments its moments.
 */
int fsg_finalize_stats(int32_t kind, const double* pct, const double* moments,
                       const int64_t* n_valid, double* stats, void* stream);

/*
 * Post-process + encode a float32 array region into an output: the
 * drop-in for quantize_array / _quantize_block_cp
 * (io/output_encoding.py:178-190, core/dask_processor.py:997-1008), the
 * display stretch (_apply_display_stretch_block, tile/dask_bridge.py:173-187)
 * and the normalisation divide (_global_stats.py:123-153).
 */
int fsg_encode(const fsg_window* src, const fsg_region* reg,
               const fsg_output* out, void* stream);

/* Gather the region of a window into a dense f32 pool (stats pre-pass). */
int fsg_copy_region(const fsg_window* win, const fsg_region* reg, float* dst,
                    void* stream);

/* ------------------------------------------------------------------ */
/* Spatial mode (multi-radius hillshade, slope, curvature)              */
/* ------------------------------------------------------------------ */

/*
 * Building blocks of the spatial-mode path, which the reference's tile
 * backend runs through Dask as one chunk (algorithms/tile/dask_bridge.py:
 * 395-405, multiscale_response_fields _nan_utils.py:455-524):
 *
 * fsg_pad_symmetric: the window padded by pad on every side as dask's
 *   map_overlap boundary='reflect' does ([d c b a | a b c d | d c b a]);
 *   out is (h + 2 pad) x (w + 2 pad), pad <= min(h, w).
 * fsg_gaussian_nan: handle_nan_with_gaussian (_nan_utils.py:18-31) of that
 *   padded window without materialising it -- scipy gaussian_filter(sigma,
 *   mode='nearest', truncate=4) of where(nan, 0, x) and of valid, axis 0 then
 *   axis 1 with float32 intermediates, where(sw > 0, sv / sw, 0).  Workspace:
 *   fsg_gaussian_nan_workspace_size(h, w, pad).
 * fsg_downsample_nan_aware: _downsample_nan_aware (_nan_utils.py:604-668),
 *   cell means with the enclosed-void fill, factor a power of two <= 16, out
 *   ceil(h/f) x ceil(w/f).  Workspace: fsg_downsample_workspace_size.
 * fsg_upsample_to_shape: _upsample_to_shape (_nan_utils.py:671-698).
 * fsg_sample_coarse: _bilinear_sample_coarse (_nan_utils.py:255-281) of a
 *   global coarse response at the window rows r0 .. r0 + h - 1 and columns
 *   c0 .. c0 + w - 1 of a full_h x full_w raster; NaN where nan_src (the
 *   window, optional) is NaN, as coarse_large_radius_response does.
 * fsg_combine_fields: _combine_multiscale_dask (_nan_utils.py:182-213) and
 *   hillshade's inline variant (_impl_hillshade.py:110-131) over n <= 16 h x w
 *   fields with leading dimension ld, written through an output descriptor.
 */
enum {
  FSG_AGG_WEIGHTED = 0, /* f0 * w0 + f1 * w1 + ... (float32, in order) */
  FSG_AGG_MEAN = 1,     /* float32 mean */
  FSG_AGG_SUM = 2,
  FSG_AGG_MAX = 3,
  FSG_AGG_MIN = 4
};

int fsg_pad_symmetric(const fsg_window* src, int32_t pad, float* out, int64_t ldo,
                      void* stream);
/*
 * fsg_pad_reflect: the window padded by its own amount on each side with the
 * same symmetric rule -- the Dask backend's map_overlap(boundary='reflect')
 * at the raster edges (core/dask_processor.py:1473-1488 runs each algorithm's
 * map_overlap over the whole raster), for windows that touch them.  Each pad
 * is at most the window's extent on its axis; out is (h + top + bottom) x
 * (w + left + right).
 */
int fsg_pad_reflect(const fsg_window* src, int32_t top, int32_t bottom, int32_t left,
                    int32_t right, float* out, int64_t ldo, void* stream);
size_t fsg_gaussian_nan_workspace_size(int32_t h, int32_t w, int32_t pad);
int fsg_gaussian_nan(const fsg_window* src, int32_t pad, double sigma, float* out, int64_t ldo,
                     void* workspace, size_t workspace_bytes, void* stream);
size_t fsg_downsample_workspace_size(int32_t h, int32_t w, int32_t factor);
int fsg_downsample_nan_aware(const fsg_window* src, int32_t factor, float* out, int64_t ldo,
                             void* workspace, size_t workspace_bytes, void* stream);
int fsg_upsample_to_shape(const float* src, int64_t lds, int32_t hs, int32_t ws, float* out,
                          int64_t ldo, int32_t h, int32_t w, void* stream);
int fsg_sample_coarse(const float* coarse, int64_t ldc, int32_t hc, int32_t wc, int32_t r0,
                      int32_t c0, int32_t full_h, int32_t full_w, const float* nan_src,
                      int64_t ld_nan, float* out, int64_t ldo, int32_t h, int32_t w,
                      void* stream);
int fsg_combine_fields(const float* const* fields, int32_t n, int64_t ld, const float* weights,
                       int32_t agg, int32_t h, int32_t w, const fsg_output* out, void* stream);

/* ------------------------------------------------------------------ */
/* Openness surrogate: BF16 CNN on MFMA                                 */
/* ------------------------------------------------------------------ */

/*
 * No reference counterpart (the reference has no learned component): the
 * building blocks of the learned openness surrogate -- 3x3 "same" convolutions
 * with zero padding on NHWC bfloat16 activations, float32 accumulation on
 * MFMA, bias + GELU epilogues, the backward passes and Adam.
 *
 * fsg_cnn_conv3x3: y = epi(conv(x, w) + bias).  x is [batch][h][w][cin]
 * bf16 with cin 8 or 32; wpack is [32][kpad] bf16 from fsg_cnn_pack (k =
 * tap * cin + ci, kpad = 9 * cin rounded up to 16).  Epilogues:
 *   FSG_CNN_GELU       y = bf16(gelu(z)), z = acc + bias; pre_out (optional,
 *                      f32 [..][32]) receives z for the backward pass;
 *   FSG_CNN_LINEAR     yf = acc + bias, f32 [..][cout] (cout <= 32);
 *   FSG_CNN_GELU_GRAD  y = bf16(acc * gelu'(pre_in)) -- with the flipped
 *                      weights from fsg_cnn_pack, the gradient of the previous
 *                      layer's pre-activation (backward through conv + GELU).
 * fsg_cnn_conv3x3_wgrad accumulates dw [32][cin][9] and db [32] (f32; zero
 * them first) from the layer input x and the pre-activation gradient dz
 * [..][32] bf16.  fsg_cnn_pack writes the forward and (wt != NULL) backward
 * bf16 layouts of float32 weights w [cout][cin][3][3].  fsg_cnn_adam is one
 * bias-corrected Adam step on float32 parameters (step counts from 1).
 */
enum { FSG_CNN_GELU = 0, FSG_CNN_LINEAR = 1, FSG_CNN_GELU_GRAD = 2 };

typedef struct fsg_cnn_conv {
  const void* x;
  const void* wpack;
  const float* bias;
  const float* pre_in;
  float* pre_out;
  void* y;
  float* yf;
  int32_t batch, h, w, cin, kpad, cout, epilogue, pad0;
} fsg_cnn_conv;

int fsg_cnn_conv3x3(const fsg_cnn_conv* c, void* stream);
int fsg_cnn_conv3x3_wgrad(const void* x, const void* dz, int32_t batch, int32_t h, int32_t w,
                          int32_t cin, float* dw, float* db, void* stream);
int fsg_cnn_pack(const float* w, int32_t cout, int32_t cin, void* wf, int32_t kpad, void* wt,
                 int32_t kpad_t, void* stream);
int fsg_cnn_adam(float* p, const float* g, float* m, float* v, int32_t n, float lr,
                 float beta1, float beta2, float eps, int32_t step, void* stream);
int fsg_cnn_to_bf16(const float* a, void* b, int64_t n, void* stream);

/* fsg_cnn_forward: the whole 6-layer network (8 -> 32 x 5 -> cout, GELU after
 * layers 0..4, linear layer 5) in one launch, identical bit for bit to the
 * layer-by-layer fsg_cnn_conv3x3 chain.  x [batch][h][w][8] bf16, yf
 * [batch][h][w][cout] f32 (cout <= 4); wpack[l] from fsg_cnn_pack (kpad 80
 * for layer 0, 288 for the others); rows_per_block <= 0 picks the default
 * (256 rows, balanced). */
typedef struct fsg_cnn_net {
  const void* wpack[6];
  const float* bias[6];
  int32_t cin, cout;
} fsg_cnn_net;

int fsg_cnn_forward(const fsg_cnn_net* net, const void* x, int32_t batch, int32_t h, int32_t w,
                    float* yf, int32_t rows_per_block, void* stream);

/* Surrogate plumbing.  fsg_cnn_features: the 4-level detrended elevation
 * pyramid, channel k = bf16((means[k] - z) * inv[k]) for k < 4 and zero
 * channels 4..7, NHWC [h][w][8].  fsg_cnn_upsample: bilinear, pixel-centre
 * aligned (sample at ((i + .5) / factor - .5, (j + .5) / factor - .5),
 * clamped) of src (+ add, optional) into out [h][w].
 *
 * fsg_cnn_mse: loss of the network output yf [n][cout] -- the angle term
 * d = yf[p][0] + base[p] - label[p] (label NaN: no label) and, when label_s
 * is given (cout >= 2), the detail term e = yf[p][1] - label_s[p] weighted by
 * weight_s[p]; loss[0] += sum d^2, loss[1] += sum w e^2, dz[p] = bf16 of
 * (2 scale d, 2 scale w e, 0, ..., 0) over 32 channels.
 *
 * fsg_cnn_compose: the full-resolution angle out [h][w] = bilinear(yf[.][0]
 * + base) + s_scale * yf[cell][1] * (z - bilinear(zg)), cell = (i / factor,
 * j / factor); yf [hs][ldg][cout], base and zg [hs][ldg]; z == NULL drops the
 * detail term.
 *
 * fsg_cnn_detail_label: per grid cell of lab/zg [hs][ldg], the least-squares
 * slope s* of exact - bilinear(lab) on z - bilinear(zg) over the cell's
 * factor^2 pixels; label_s = s_unit * s*, weight_s = sum dz^2 / (factor^2
 * s_unit^2).  exact covers full-resolution rows [ey0, ey0 + eh) and columns
 * [ex0, ex0 + ew) of the grid's pixel frame, z covers the whole frame; cells
 * without a full label get label_s = NaN, weight_s = 0. */
int fsg_cnn_features(const float* z, int64_t ldz, const float* const* means, int64_t ldm,
                     const float* inv, int32_t h, int32_t w, void* out, void* stream);
int fsg_cnn_upsample(const float* src, const float* add, int64_t lds, int32_t hs, int32_t ws,
                     int32_t factor, float* out, int64_t ldo, int32_t h, int32_t w,
                     void* stream);
int fsg_cnn_mse(const float* yf, int32_t cout, const float* base, const float* label,
                const float* label_s, const float* weight_s, int64_t n, float scale, void* dz,
                float* loss, void* stream);
int fsg_cnn_compose(const float* yf, int32_t cout, const float* base, const float* zg,
                    int64_t ldg, int32_t hs, int32_t ws, int32_t factor, const float* z,
                    int64_t ldz, float s_scale, float* out, int64_t ldo, int32_t h, int32_t w,
                    void* stream);
int fsg_cnn_detail_label(const float* exact, int64_t lde, int32_t ey0, int32_t ex0, int32_t eh,
                         int32_t ew, const float* lab, const float* zg, int64_t ldg, int32_t hs,
                         int32_t ws, int32_t factor, const float* z, int64_t ldz, float s_unit,
                         float* label_s, float* weight_s, void* stream);

/* ------------------------------------------------------------------ */
/* Batched fused launch                                                 */
/* ------------------------------------------------------------------ */

/*
 * One tile of a batch: an independent window, the region computed, and the
 * outputs wanted from it (an output with data == NULL is not computed;
 * topousm needs n_radii > 0 or a field).  The parameters mean what they mean
 * for fsg_hillshade_slope, fsg_curvature, fsg_topousm and fsg_openness (the
 * openness tiles of one batch share one ray table).
 */
typedef struct fsg_batch_tile {
  fsg_window win;
  fsg_region reg;
  fsg_output hillshade, slope, curvature, topousm, topousm2;
  fsg_shade_params shade;
  fsg_curvature_params curv;
  fsg_topousm_params usm;
  fsg_output openness;
  fsg_openness_params open;
} fsg_batch_tile;

/* Launch shape of the persistent kernel (what the tuner searches). */
typedef struct fsg_batch_config {
  int32_t wg_per_cu;  /* persistent workgroups per CU (0: 3); with openness tiles,
                         5..8 select kernels capped at that many waves per SIMD */
  int32_t order;      /* 0: TopoUSM, openness, stencil blocks; 1: stencil, openness, TopoUSM */
  int32_t n_cu;       /* CUs to size the grid for (0: the device's) */
  int32_t chunk;      /* work items claimed per queue access (0: 8) */
} fsg_batch_config;

typedef struct fsg_batch fsg_batch;

/*
 * The batch replaces the reference's per-tile, per-algorithm loop
 * (core/tile_processor.py:1635-1683 -> run_tile_algorithm,
 * core/tile_compute.py:93-129) for hillshade, slope, curvature, openness and
 * topousm_fast: every tile's work goes through ONE persistent kernel that
 * drains a device queue of work items, after the per-tile coarse-grid
 * preparation.  Outputs equal the per-tile entry points' bit for bit.
 *   - Without openness tiles the items are TopoUSM main-pass blocks and
 *     stencil blocks of NaN-free tiles; stencil tiles with NaN
 *     (nan_hint != 0) take their own launches.
 *   - With openness tiles the items are openness blocks (64 x 16 pixels)
 *     and stencil blocks, NaN-aware ones included; every TopoUSM tile then
 *     takes its own fsg_topousm launches inside the same sequence.
 * TopoUSM tiles off the fast path, or with NaN (nan_hint != 0), or with a
 * box-radius count other than the batch's most common one also run their own
 * fsg_topousm launches.
 *
 * fsg_batch_create plans the batch (host work, one device allocation for the
 * coarse grids and descriptors, a synchronous upload); the tiles' buffers must
 * outlive it.  fsg_batch_run enqueues one pass on a stream; fsg_batch_capture
 * records one pass into a hipGraph (non-default stream; per-kernel timing
 * off) and fsg_batch_replay launches it.  fsg_batch_info fills int64[12]:
 * TopoUSM items, stencil items, fused TopoUSM tiles, fallback tiles, stencil
 * tiles, LDS bytes, box radii of the fused kernel, device bytes held,
 * openness items, openness tiles, stencil tiles whose interior is walked by
 * rows, and the row-walking items among the stencil items.
 */
int fsg_batch_create(const fsg_batch_tile* tiles, int32_t n, fsg_batch** out);
int fsg_batch_run(fsg_batch* b, const fsg_batch_config* cfg, void* stream);
int fsg_batch_capture(fsg_batch* b, const fsg_batch_config* cfg, void* stream);
int fsg_batch_replay(fsg_batch* b, void* stream);
int fsg_batch_info(const fsg_batch* b, int64_t* info);
void fsg_batch_destroy(fsg_batch* b);

/* ------------------------------------------------------------------ */
/* prepare: NoData inference and push-pull void fill                    */
/* ------------------------------------------------------------------ */

/*
 * preprocess_dem_to_cog (io/dem_preprocess.py:462-727) for a raster already
 * in device memory; the file reads and the COG write stay with the caller.
 *
 * fsg_prepare_sample: src.read(out_shape=(oh, ow), resampling=nearest)
 *   (io/dem_preprocess.py:536-542): source index floor((i + 0.5) n_src / n_dst)
 *   per axis.
 * fsg_value_mode: the np.unique(..., return_counts=True) + argmax step of
 *   _detect_sentinel_nodata (io/dem_preprocess.py:241-275) and
 *   _detect_border_nodata (:191-238) over the values that are finite and not
 *   the declared NoData: the most frequent value (the smallest on ties), its
 *   count, the number of such values, the smallest and largest.  stats is
 *   device memory, written when the stream reaches it.  Workspace:
 *   fsg_value_mode_workspace_size(n); n < 2^31.
 * fsg_count_equal: number of the n values equal to value (the
 *   np.mean(sample == cand) of _detect_border_nodata); count is device memory.
 * fsg_coarse_average: _nan_aware_coarse_average (io/dem_preprocess.py:135-188)
 *   of the (ch k) x (cw k) sample: a cell is the mean of its sub-cells that
 *   are finite and equal to none of the n_nodata values (<= 16), summed in
 *   numpy's order; cvalid (ch x cw bytes) is 1 for cells with a finite mean;
 *   n_void (device) counts the others.
 * fsg_pushpull_fill: pushpull_fill (algorithms/_pyramid_fill.py:24-87) of the
 *   coarse grid: scipy zoom (order 1, corner-aligned) pyramid, valid cells
 *   kept; n_valid_cells = h w - n_void.  Synchronises the stream once per
 *   level.  Workspace: fsg_pushpull_workspace_size(h, w).
 * fsg_edge_connected: _edge_connected_mask (utils/nodata_handler.py:13-25):
 *   cells of mask (1 = void) 4-connected to the grid edge.  flag is one
 *   device int32; synchronises the stream once per sweep; sweeps (optional,
 *   host) receives the number of sweeps.
 * fsg_prepare_fill: _apply_band_fill and _sample_coarse (io/dem_preprocess.py:
 *   729-826) over the whole raster: NoData (non-finite or equal to one of the
 *   n_nodata values) -> NaN, then with do_fill the bilinear sample of the
 *   coarse surface at the pixel centre (global coarse coordinates) where the
 *   sampled exterior is <= 0.5 (FSG_FILL_ENCLOSED) or everywhere
 *   (FSG_FILL_ALL).  out may alias src.data.
 */
enum { FSG_FILL_NONE = 0, FSG_FILL_ENCLOSED = 1, FSG_FILL_ALL = 2 };
#define FSG_PREPARE_MAX_NODATA 16

typedef struct fsg_mode_stats {
  double mode;        /* most frequent valid value */
  int64_t count;      /* its count */
  int64_t n_valid;    /* finite values that are not the declared NoData */
  double vmin, vmax;  /* smallest and largest of them */
  int64_t n_distinct; /* distinct values among them */
} fsg_mode_stats;

int fsg_prepare_sample(const fsg_window* src, int32_t oh, int32_t ow, float* out, int64_t ldo,
                       void* stream);
size_t fsg_value_mode_workspace_size(int64_t n);
int fsg_value_mode(const float* values, int64_t n, int32_t has_declared, float declared,
                   fsg_mode_stats* stats, void* workspace, size_t workspace_bytes, void* stream);
int fsg_count_equal(const float* values, int64_t n, float value, uint64_t* count, void* stream);
int fsg_coarse_average(const float* sample, int64_t lds, int32_t ch, int32_t cw, int32_t k,
                       const float* nodata, int32_t n_nodata, float* coarse, int64_t ldc,
                       uint8_t* cvalid, uint64_t* n_void, void* stream);
size_t fsg_pushpull_workspace_size(int32_t h, int32_t w);
int fsg_pushpull_fill(const float* coarse, int64_t ldc, const uint8_t* valid, int32_t h,
                      int32_t w, int32_t n_valid_cells, float* out, int64_t ldo, void* workspace,
                      size_t workspace_bytes, void* stream);
int fsg_edge_connected(const uint8_t* mask, int32_t h, int32_t w, uint8_t* ext, int32_t* flag,
                       int32_t* sweeps, void* stream);
int fsg_prepare_fill(const fsg_window* src, const float* nodata, int32_t n_nodata,
                     int32_t fill_mode, int32_t do_fill, const float* surface,
                     const uint8_t* exterior, int32_t ch, int32_t cw, float* out, int64_t ldo,
                     void* stream);

/* ------------------------------------------------------------------ */
/* Cloud-optimized GeoTIFF output and windowed input                    */
/* ------------------------------------------------------------------ */

/*
 * The on-disk side of the tile pipeline: the output COG the reference builds
 * with GDAL (io/cog_builder.py:658-716; options core/dask_processor.py:201-228:
 * COMPRESS=ZSTD, LEVEL=1, BLOCKSIZE=512, PREDICTOR=3 for float32 and 2 for
 * int16 (none for uint8), OVERVIEW_RESAMPLING=AVERAGE, OVERVIEW_COUNT=8,
 * BIGTIFF=YES) and the windowed read of read_tile_window (core/tile_io.py:
 * 49-51).
 *
 * fsg_cog_write: a BigTIFF COG of the device raster r (float32, int16 or
 *   uint8, row stride ld): every IFD first, then the tile data from the
 *   smallest overview to full resolution.  Overviews are 2 x 2 AVERAGE of the
 *   valid pixels (NaN and the NoData value excluded), level by level, while
 *   the level exceeds one block and up to overview_count; the GPU builds them
 *   and applies the predictor, host threads compress with ZSTD.  GeoTIFF tags
 *   (north-up geotransform x0, a, 0, y0, 0, e; EPSG code) and GDAL_NODATA are
 *   written when given.  opt may be NULL (512, level 1, 8 overviews, all
 *   host threads).  The call synchronises the stream.
 * fsg_cog_open / fsg_cog_info / fsg_cog_read / fsg_cog_close: a tiled
 *   single-band float32 / int16 / uint8 TIFF or BigTIFF (ZSTD, DEFLATE or
 *   none; predictor 1, 2 or 3); level 0 is full resolution, 1.. the overviews
 *   in file order.  fsg_cog_read writes the window rows r0 .. r0 + h - 1 and
 *   columns c0 .. c0 + w - 1 of a level to device memory out (row stride
 *   ldo, the file's sample type) and synchronises the stream.
 */
typedef struct fsg_cog_raster {
  const void* data;      /* device */
  int64_t ld;
  int32_t h, w;
  int32_t dtype;         /* FSG_OUT_F32 / FSG_OUT_I16 / FSG_OUT_U8 */
  int32_t has_geotransform;
  double geotransform[6];
  int32_t epsg;          /* 0: none */
  int32_t epsg_geographic;
  int32_t has_nodata;
  int32_t pad0;
  double nodata;
} fsg_cog_raster;

typedef struct fsg_cog_options {
  int32_t block;          /* tile edge, multiple of 16, <= 4096 */
  int32_t zstd_level;
  int32_t overview_count;
  int32_t threads;        /* 0: all host threads */
} fsg_cog_options;

typedef struct fsg_cog_report {
  int32_t n_levels;
  int32_t pad0;
  uint64_t data_start;
  uint64_t compressed_bytes;
  uint64_t raw_bytes;
} fsg_cog_report;

typedef struct fsg_cog_info_t {
  int32_t n_levels, width, height, tile_w, tile_h, dtype, compression, predictor;
  int32_t has_geotransform;
  int32_t pad0;
  double geotransform[6];
  int32_t epsg, epsg_geographic, has_nodata, pad1;
  double nodata;
} fsg_cog_info_t;

typedef struct fsg_cog fsg_cog;

int fsg_cog_write(const char* path, const fsg_cog_raster* r, const fsg_cog_options* opt,
                  fsg_cog_report* report, void* stream);
int fsg_cog_open(const char* path, fsg_cog** out);
int fsg_cog_info(const fsg_cog* c, int32_t level, fsg_cog_info_t* info);
int fsg_cog_read(fsg_cog* c, int32_t level, int32_t r0, int32_t c0, int32_t h, int32_t w,
                 void* out, int64_t ldo, int32_t threads, void* stream);
void fsg_cog_close(fsg_cog* c);

#ifdef __cplusplus
}
#endif
#endif /* FSG_H */
