// Cloud-optimized GeoTIFF output and windowed input for rasters in device
// memory: the on-disk side of the tile pipeline (core/tile_io.py:49-72,
// io/cog_builder.py:658-716 with the COG options of
// core/dask_processor.py:201-228: ZSTD level 1, 512 x 512 tiles, PREDICTOR 3
// for float32 and 2 for int16, none for uint8, AVERAGE overviews, BIGTIFF).
//
// Writer: the overview pyramid (2 x 2 AVERAGE of valid pixels, level by
// level) and the predictor encoding of every tile run on the GPU; ZSTD
// compression runs on host threads (libzstd); the file is BigTIFF with every
// IFD ahead of the tile data and the tile data ordered from the smallest
// overview to full resolution, as the COG layout asks.
// Reader: the IFDs are parsed once (BigTIFF or classic, tiled, ZSTD / DEFLATE
// / none, predictor 1 / 2 / 3); a window read decompresses the tiles it
// touches on host threads, and one GPU kernel undoes the predictor and places
// the window in device memory.
#include <fcntl.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "fsg_common.hpp"
#include "tiff_ifd.hpp"

// libzstd's public API (zstd.h is not installed; the shared library is)
extern "C" {
typedef struct ZSTD_CCtx_s ZSTD_CCtx;
typedef struct ZSTD_DCtx_s ZSTD_DCtx;
ZSTD_CCtx* ZSTD_createCCtx(void);
size_t ZSTD_freeCCtx(ZSTD_CCtx* cctx);
size_t ZSTD_compressCCtx(ZSTD_CCtx* cctx, void* dst, size_t dstCapacity, const void* src,
                         size_t srcSize, int compressionLevel);
size_t ZSTD_compressBound(size_t srcSize);
unsigned ZSTD_isError(size_t code);
const char* ZSTD_getErrorName(size_t code);
ZSTD_DCtx* ZSTD_createDCtx(void);
size_t ZSTD_freeDCtx(ZSTD_DCtx* dctx);
size_t ZSTD_decompressDCtx(ZSTD_DCtx* dctx, void* dst, size_t dstCapacity, const void* src,
                           size_t srcSize);
}

namespace fsg {


static int elem_bytes(int dtype) { return dtype == FSG_OUT_F32 ? 4 : (dtype == FSG_OUT_I16 ? 2 : 1); }

// ------------------------------------------------------------------ kernels
template <typename T>
__device__ __forceinline__ bool cog_valid(T v, int has_nd, double nd) {
  if constexpr (std::is_same<T, float>::value) {
    if (isnan(v)) return false;
  }
  return !(has_nd && (double)v == nd);
}

// AVERAGE overview: the mean of the valid pixels of each 2 x 2 source block
// (clipped at the far edges), accumulated in float64; integer types round to
// nearest; a block with no valid pixel takes the NoData value (NaN for float
// without one, 0 for integers without one).
template <typename T>
__global__ void k_ovr_avg(const T* __restrict__ src, int64_t lds, int h, int w,
                          T* __restrict__ dst, int64_t ldd, int hd, int wd, int has_nd,
                          double nd) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= wd) return;
  for (int y = blockIdx.y; y < hd; y += gridDim.y) {
    double s = 0.0;
    int n = 0;
    for (int dy = 0; dy < 2; dy++) {
      const int sy = 2 * y + dy;
      if (sy >= h) break;
      for (int dx = 0; dx < 2; dx++) {
        const int sx = 2 * x + dx;
        if (sx >= w) break;
        const T v = src[(int64_t)sy * lds + sx];
        if (cog_valid(v, has_nd, nd)) {
          s += (double)v;
          n++;
        }
      }
    }
    T o;
    if constexpr (std::is_same<T, float>::value) {
      o = n ? (float)(s / (double)n) : (has_nd ? (float)nd : __builtin_nanf(""));
    } else {
      o = n ? (T)floor(s / (double)n + 0.5) : (T)(has_nd ? nd : 0.0);
    }
    dst[(int64_t)y * ldd + x] = o;
  }
}

// Tile encoding, one workgroup per tile row of `block` samples (block <= 4096):
// predictor 3 (Adobe floating-point predictor: the row's bytes regrouped by
// significance, most significant plane first, then byte-wise horizontal
// differencing), predictor 2 (sample-wise horizontal differencing) or none.
// Pixels outside the level are `pad`.
template <typename T>
__global__ void k_tile_encode(const T* __restrict__ lvl, int64_t ld, int h, int w, int block,
                              int ntx, int t0, int predictor, T pad, uint8_t* __restrict__ out) {
  extern __shared__ uint8_t rowb[];
  const int t = t0 + blockIdx.y, r = blockIdx.x;
  const int ty = t / ntx, tx = t % ntx;
  const int y = ty * block + r;
  constexpr int B = sizeof(T);
  uint8_t* orow = out + ((size_t)blockIdx.y * block + r) * (size_t)block * B;
  for (int i = threadIdx.x; i < block; i += blockDim.x) {
    const int x = tx * block + i;
    const T v = (y < h && x < w) ? lvl[(int64_t)y * ld + x] : pad;
    bool planes = false;
    if constexpr (B == 4) {
      if (predictor == 3) {
        uint32_t b;
        memcpy(&b, &v, 4);
        for (int k = 0; k < 4; k++) rowb[k * block + i] = (uint8_t)(b >> (8 * (3 - k)));
        planes = true;
      }
    }
    if (!planes) memcpy(rowb + (size_t)i * B, &v, B);
  }
  __syncthreads();
  const int nb = block * B;
  if (predictor == 3) {
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
      orow[i] = (uint8_t)(rowb[i] - (i ? rowb[i - 1] : 0));
  } else if (predictor == 2) {
    for (int i = threadIdx.x; i < block; i += blockDim.x) {
      T cur, prev = 0;
      memcpy(&cur, rowb + (size_t)i * B, B);
      if (i) memcpy(&prev, rowb + (size_t)(i - 1) * B, B);
      const T d = (T)(cur - prev);
      memcpy(orow + (size_t)i * B, &d, B);
    }
  } else {
    for (int i = threadIdx.x; i < nb; i += blockDim.x) orow[i] = rowb[i];
  }
}

// Window decode: one workgroup per (tile row, staged tile).  Undoes the
// predictor with a workgroup prefix sum (mod 2^8 or 2^16) and writes the part
// of the row inside the window to out.
struct DecTile {
  int tile_y0, tile_x0;   // level coordinates of the tile's first pixel
};

template <typename T>
__global__ void __launch_bounds__(256) k_tile_decode(const uint8_t* __restrict__ raw,
                                                     const DecTile* __restrict__ tiles,
                                                     int tw, int th, int predictor, int r0,
                                                     int c0, int wh, int ww,
                                                     T* __restrict__ out, int64_t ldo) {
  extern __shared__ uint8_t smem[];
  constexpr int B = sizeof(T);
  const int r = blockIdx.x;
  const DecTile dt = tiles[blockIdx.y];
  const int y = dt.tile_y0 + r - r0;     // window row
  if (y < 0 || y >= wh) return;          // uniform per workgroup
  const uint8_t* irow = raw + ((size_t)blockIdx.y * th + r) * (size_t)tw * B;
  const int nb = tw * B;
  uint32_t* part = (uint32_t*)(smem + ((nb + 15) & ~15));
  if (predictor == 3 || predictor == 2) {
    // per-thread contiguous chunk sums, then an exclusive scan of the chunks
    const int unit = predictor == 3 ? 1 : B;           // scan element bytes
    const int n = nb / unit;
    const int per = (n + blockDim.x - 1) / blockDim.x;
    const int i0 = threadIdx.x * per, i1 = min(n, i0 + per);
    uint32_t s = 0;
    for (int i = i0; i < i1; i++) {
      uint32_t v = 0;
      memcpy(&v, irow + (size_t)i * unit, unit);
      s += v;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t acc = 0;
      for (int k = 0; k < (int)blockDim.x; k++) {
        const uint32_t v = part[k];
        part[k] = acc;
        acc += v;
      }
    }
    __syncthreads();
    uint32_t acc = part[threadIdx.x];
    for (int i = i0; i < i1; i++) {
      uint32_t v = 0;
      memcpy(&v, irow + (size_t)i * unit, unit);
      acc += v;
      if (unit == 1) {
        smem[i] = (uint8_t)acc;
      } else {
        const uint16_t a16 = (uint16_t)acc;
        memcpy(smem + (size_t)i * unit, &a16, unit);
      }
    }
  } else {
    for (int i = threadIdx.x; i < nb; i += blockDim.x) smem[i] = irow[i];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < tw; i += blockDim.x) {
    const int x = dt.tile_x0 + i - c0;
    if (x < 0 || x >= ww) continue;
    T v;
    bool planes = false;
    if constexpr (B == 4) {
      if (predictor == 3) {
        uint32_t b = 0;
        for (int k = 0; k < 4; k++) b |= (uint32_t)smem[k * tw + i] << (8 * (3 - k));
        memcpy(&v, &b, 4);
        planes = true;
      }
    }
    if (!planes) memcpy(&v, smem + (size_t)i * B, B);
    out[(int64_t)y * ldo + x] = v;
  }
}

// ------------------------------------------------------------ file helpers
struct Entry {
  uint16_t tag, type;
  uint64_t count;
  std::vector<uint8_t> data;   // little-endian values
};

template <typename V>
static Entry entry(uint16_t tag, uint16_t type, const std::vector<V>& vals) {
  Entry e{tag, type, vals.size(), {}};
  e.data.resize(sizeof(V) * vals.size());
  if (!vals.empty()) std::memcpy(e.data.data(), vals.data(), e.data.size());
  return e;
}

static Entry entry_ascii(uint16_t tag, const std::string& s) {
  Entry e{tag, TY_ASCII, s.size() + 1, {}};
  e.data.assign(s.begin(), s.end());
  e.data.push_back(0);
  return e;
}

// bytes an IFD takes with its out-of-line values (BigTIFF: inline <= 8 bytes)
static size_t ifd_bytes(const std::vector<Entry>& es) {
  size_t b = 8 + 20 * es.size() + 8;
  for (const Entry& e : es)
    if (e.data.size() > 8) b += (e.data.size() + 7) & ~(size_t)7;
  return b;
}

// serialise an IFD placed at `at`; next = offset of the following IFD or 0
static std::vector<uint8_t> ifd_serialise(std::vector<Entry> es, uint64_t at, uint64_t next) {
  std::sort(es.begin(), es.end(), [](const Entry& a, const Entry& b) { return a.tag < b.tag; });
  std::vector<uint8_t> out(ifd_bytes(es), 0);
  uint64_t n = es.size();
  std::memcpy(out.data(), &n, 8);
  size_t ext = 8 + 20 * es.size() + 8;
  for (size_t k = 0; k < es.size(); k++) {
    uint8_t* p = out.data() + 8 + 20 * k;
    std::memcpy(p, &es[k].tag, 2);
    std::memcpy(p + 2, &es[k].type, 2);
    std::memcpy(p + 4, &es[k].count, 8);
    if (es[k].data.size() <= 8) {
      std::memcpy(p + 12, es[k].data.data(), es[k].data.size());
    } else {
      const uint64_t off = at + ext;
      std::memcpy(p + 12, &off, 8);
      std::memcpy(out.data() + ext, es[k].data.data(), es[k].data.size());
      ext += (es[k].data.size() + 7) & ~(size_t)7;
    }
  }
  std::memcpy(out.data() + 8 + 20 * es.size(), &next, 8);
  return out;
}

static int write_all(int fd, const void* p, size_t n, uint64_t off) {
  const uint8_t* c = (const uint8_t*)p;
  while (n) {
    const ssize_t k = pwrite(fd, c, n, (off_t)off);
    if (k <= 0) return 1;
    c += k;
    n -= (size_t)k;
    off += (uint64_t)k;
  }
  return 0;
}

static int read_all(int fd, void* p, size_t n, uint64_t off) {
  uint8_t* c = (uint8_t*)p;
  while (n) {
    const ssize_t k = pread(fd, c, n, (off_t)off);
    if (k <= 0) return 1;
    c += k;
    n -= (size_t)k;
    off += (uint64_t)k;
  }
  return 0;
}

static int n_workers(int req) {
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  return std::max(1, std::min(req > 0 ? req : hw, 64));
}

// run f(i, worker) for i in [0, n) on up to `threads` host threads
template <typename F>
static void parallel_for(int n, int threads, F f) {
  std::atomic<int> next{0};
  auto work = [&](int wid) {
    for (int i = next++; i < n; i = next++) f(i, wid);
  };
  std::vector<std::thread> ts;
  for (int k = 1; k < std::min(threads, n); k++) ts.emplace_back(work, k);
  work(0);
  for (auto& t : ts) t.join();
}

static dim3 grid_r(int w, int h) {
  return dim3((unsigned)((w + 255) / 256), (unsigned)std::max(1, std::min(h, 16384)));
}

// --------------------------------------------------------------- the writer
struct Level {
  const void* data;
  int64_t ld;
  int h, w, ntx, nty;
};

template <typename T>
static int encode_and_write(int fd, const std::vector<Level>& L, const fsg_cog_options& o,
                            int dtype, int predictor, T pad, hipStream_t s, uint64_t data_start,
                            std::vector<std::vector<uint64_t>>& offs,
                            std::vector<std::vector<uint64_t>>& cnts, double* t_bytes) {
  const int block = o.block;
  const size_t tile_bytes = (size_t)block * block * sizeof(T);
  const int batch = (int)std::max<size_t>(1, (64u << 20) / tile_bytes);
  uint8_t* dstage = nullptr;
  uint8_t* hstage = nullptr;
  if (check_hip(hipMalloc(&dstage, tile_bytes * batch), "fsg_cog_write: device staging"))
    return FSG_E_HIP;
  if (check_hip(hipHostMalloc(&hstage, tile_bytes * batch, hipHostMallocDefault),
                "fsg_cog_write: pinned staging")) {
    (void)hipFree(dstage);
    return FSG_E_HIP;
  }
  const int threads = n_workers(o.threads);
  const size_t cap = ZSTD_compressBound(tile_bytes);
  // two sets of compressed tiles: one batch is written by a writer thread while
  // the next is compressed
  std::vector<std::vector<uint8_t>> comp[2];
  std::vector<size_t> csize[2];
  for (int k = 0; k < 2; k++) {
    comp[k].assign(batch, std::vector<uint8_t>(cap));
    csize[k].assign(batch, 0);
  }
  std::vector<ZSTD_CCtx*> ctx(threads, nullptr);
  for (auto& c : ctx) c = ZSTD_createCCtx();
  int rc = FSG_OK;
  for (auto* c : ctx)
    if (!c) rc = fail(FSG_E_HIP, "fsg_cog_write: ZSTD context");
  uint64_t pos = data_start;
  std::atomic<int> wbad{0};
  std::thread writer;
  auto join_writer = [&]() {
    if (writer.joinable()) writer.join();
  };
  int set = 0;
  for (int li = (int)L.size() - 1; li >= 0 && rc == FSG_OK; li--) {   // smallest level first
    const Level& lv = L[li];
    const int nt = lv.ntx * lv.nty;
    offs[li].assign(nt, 0);
    cnts[li].assign(nt, 0);
    for (int t0 = 0; t0 < nt && rc == FSG_OK; t0 += batch) {
      const int nb = std::min(batch, nt - t0);
      hipLaunchKernelGGL(k_tile_encode<T>, dim3(block, nb), dim3(256), block * sizeof(T) + 16, s,
                         (const T*)lv.data, lv.ld, lv.h, lv.w, block, lv.ntx, t0, predictor, pad,
                         dstage);
      if (hipError_t e = hipGetLastError()) { rc = check_hip(e, "k_tile_encode"); break; }
      count_launch();
      if (check_hip(hipMemcpyAsync(hstage, dstage, tile_bytes * nb, hipMemcpyDeviceToHost, s),
                    "fsg_cog_write: tile copy") ||
          check_hip(hipStreamSynchronize(s), "fsg_cog_write: tile copy")) {
        rc = FSG_E_HIP;
        break;
      }
      std::atomic<int> bad{0};
      auto& cs = comp[set];
      auto& sz = csize[set];
      parallel_for(nb, threads, [&](int i, int wid) {
        const size_t z = ZSTD_compressCCtx(ctx[wid], cs[i].data(), cap, hstage + tile_bytes * i,
                                           tile_bytes, o.zstd_level);
        if (ZSTD_isError(z)) bad = 1;
        else sz[i] = z;
      });
      if (bad) {
        rc = fail(FSG_E_HIP, "fsg_cog_write: ZSTD compression failed");
        break;
      }
      join_writer();   // the previous batch is on disk; its set is free again
      if (wbad) break;
      writer = std::thread([&, li, t0, nb, set]() {
        for (int i = 0; i < nb; i++) {
          if (write_all(fd, comp[set][i].data(), csize[set][i], pos)) {
            wbad = 1;
            return;
          }
          offs[li][t0 + i] = pos;
          cnts[li][t0 + i] = csize[set][i];
          pos += csize[set][i];
          *t_bytes += (double)csize[set][i];
        }
      });
      set ^= 1;
    }
  }
  join_writer();
  if (wbad && rc == FSG_OK) rc = fail(FSG_E_INVALID, "fsg_cog_write: write failed (disk full?)");
  for (auto& c : ctx) ZSTD_freeCCtx(c);
  (void)hipHostFree(hstage);
  (void)hipFree(dstage);
  (void)dtype;
  return rc;
}

template <typename T>
static int build_overviews(const fsg_cog_raster& R, int block, int max_ovr, hipStream_t s,
                           std::vector<Level>& L, std::vector<void*>& owned) {
  L.push_back({R.data, R.ld, R.h, R.w, (R.w + block - 1) / block, (R.h + block - 1) / block});
  while ((int)L.size() - 1 < max_ovr && std::max(L.back().h, L.back().w) > block) {
    const Level& p = L.back();
    const int h2 = (p.h + 1) / 2, w2 = (p.w + 1) / 2;
    void* buf = nullptr;
    if (check_hip(hipMalloc(&buf, sizeof(T) * (size_t)h2 * w2), "fsg_cog_write: overview"))
      return FSG_E_HIP;
    owned.push_back(buf);
    hipLaunchKernelGGL(k_ovr_avg<T>, grid_r(w2, h2), dim3(256), 0, s, (const T*)p.data, p.ld, p.h,
                       p.w, (T*)buf, (int64_t)w2, h2, w2, R.has_nodata, R.nodata);
    FSG_CHECK_LAUNCH("k_ovr_avg");
    L.push_back({buf, (int64_t)w2, h2, w2, (w2 + block - 1) / block, (h2 + block - 1) / block});
  }
  return FSG_OK;
}

static std::vector<Entry> level_entries(const fsg_cog_raster& R, const Level& lv, bool main,
                                        int dtype, int predictor, int block,
                                        const std::vector<uint64_t>& off,
                                        const std::vector<uint64_t>& cnt) {
  std::vector<Entry> es;
  const int bps = 8 * elem_bytes(dtype);
  const uint16_t fmt = dtype == FSG_OUT_F32 ? 3 : (dtype == FSG_OUT_I16 ? 2 : 1);
  if (!main) es.push_back(entry<uint32_t>(T_NEWSUBFILE, TY_LONG, {1u}));
  es.push_back(entry<uint32_t>(T_WIDTH, TY_LONG, {(uint32_t)lv.w}));
  es.push_back(entry<uint32_t>(T_HEIGHT, TY_LONG, {(uint32_t)lv.h}));
  es.push_back(entry<uint16_t>(T_BPS, TY_SHORT, {(uint16_t)bps}));
  es.push_back(entry<uint16_t>(T_COMPRESSION, TY_SHORT, {COMP_ZSTD}));
  es.push_back(entry<uint16_t>(T_PHOTOMETRIC, TY_SHORT, {1}));
  es.push_back(entry<uint16_t>(T_SPP, TY_SHORT, {1}));
  es.push_back(entry<uint16_t>(T_PLANAR, TY_SHORT, {1}));
  es.push_back(entry<uint16_t>(T_PREDICTOR, TY_SHORT, {(uint16_t)predictor}));
  es.push_back(entry<uint16_t>(T_TILEW, TY_SHORT, {(uint16_t)block}));
  es.push_back(entry<uint16_t>(T_TILEH, TY_SHORT, {(uint16_t)block}));
  es.push_back(entry<uint64_t>(T_TILEOFF, TY_LONG8, off));
  es.push_back(entry<uint64_t>(T_TILEBC, TY_LONG8, cnt));
  es.push_back(entry<uint16_t>(T_SAMPLEFMT, TY_SHORT, {fmt}));
  if (main && R.has_geotransform) {
    const double* g = R.geotransform;   // x0, a, b, y0, d, e (north-up: b = d = 0)
    es.push_back(entry<double>(T_PIXSCALE, TY_DOUBLE, {g[1], -g[5], 0.0}));
    es.push_back(entry<double>(T_TIEPOINT, TY_DOUBLE, {0.0, 0.0, 0.0, g[0], g[3], 0.0}));
    std::vector<uint16_t> k = {1, 1, 0, 0};
    auto key = [&](uint16_t id, uint16_t v) {
      k.push_back(id);
      k.push_back(0);
      k.push_back(1);
      k.push_back(v);
      k[3]++;
    };
    key(1024, R.epsg_geographic ? 2 : 1);   // GTModelTypeGeoKey
    key(1025, 1);                           // GTRasterTypeGeoKey: PixelIsArea
    if (R.epsg > 0) key(R.epsg_geographic ? 2048 : 3072, (uint16_t)R.epsg);
    es.push_back(entry<uint16_t>(T_GEOKEYS, TY_SHORT, k));
  }
  if (R.has_nodata) {
    char buf[64];
    if (std::isnan(R.nodata)) std::snprintf(buf, sizeof buf, "nan");
    else std::snprintf(buf, sizeof buf, "%.17g", R.nodata);
    es.push_back(entry_ascii(T_NODATA, buf));
  }
  return es;
}

template <typename T>
static int write_cog(const char* path, const fsg_cog_raster& R, const fsg_cog_options& o,
                     fsg_cog_report* rep, hipStream_t s) {
  const int predictor = R.dtype == FSG_OUT_F32 ? 3 : (R.dtype == FSG_OUT_I16 ? 2 : 1);
  std::vector<Level> L;
  std::vector<void*> owned;
  int rc = build_overviews<T>(R, o.block, o.overview_count, s, L, owned);
  auto cleanup = [&]() {
    for (void* p : owned) (void)hipFree(p);
  };
  if (rc) {
    cleanup();
    return rc;
  }
  std::vector<std::vector<uint64_t>> offs(L.size()), cnts(L.size());
  // IFD sizes are fixed by the tile counts, so the data starts right after them
  std::vector<size_t> isz(L.size());
  uint64_t at = 16;
  for (size_t li = 0; li < L.size(); li++) {
    std::vector<uint64_t> z((size_t)L[li].ntx * L[li].nty, 0);
    isz[li] = ifd_bytes(level_entries(R, L[li], li == 0, R.dtype, predictor, o.block, z, z));
    at += isz[li];
  }
  const uint64_t data_start = (at + 4095) & ~(uint64_t)4095;
  const int fd = open(path, O_CREAT | O_TRUNC | O_WRONLY, 0644);
  if (fd < 0) {
    cleanup();
    return fail(FSG_E_INVALID, std::string("fsg_cog_write: cannot create ") + path);
  }
  T pad;
  if constexpr (std::is_same<T, float>::value)
    pad = R.has_nodata ? (float)R.nodata : 0.0f;
  else
    pad = (T)(R.has_nodata ? R.nodata : 0.0);
  double tbytes = 0.0;
  rc = encode_and_write<T>(fd, L, o, R.dtype, predictor, pad, s, data_start, offs, cnts, &tbytes);
  if (rc == FSG_OK) {
    // header + IFDs (main first, each pointing to the next)
    std::vector<uint8_t> head(16, 0);
    head[0] = 'I';
    head[1] = 'I';
    head[2] = 43;
    head[4] = 8;
    const uint64_t first = 16;
    std::memcpy(head.data() + 8, &first, 8);
    std::vector<uint8_t> blob(head);
    uint64_t cur = 16;
    for (size_t li = 0; li < L.size(); li++) {
      const uint64_t next = li + 1 < L.size() ? cur + isz[li] : 0;
      auto es = level_entries(R, L[li], li == 0, R.dtype, predictor, o.block, offs[li], cnts[li]);
      auto b = ifd_serialise(es, cur, next);
      blob.insert(blob.end(), b.begin(), b.end());
      cur += isz[li];
    }
    if (write_all(fd, blob.data(), blob.size(), 0))
      rc = fail(FSG_E_INVALID, "fsg_cog_write: header write failed");
  }
  if (close(fd) != 0 && rc == FSG_OK) rc = fail(FSG_E_INVALID, "fsg_cog_write: close failed");
  if (rep && rc == FSG_OK) {
    rep->n_levels = (int32_t)L.size();
    rep->data_start = data_start;
    rep->compressed_bytes = (uint64_t)tbytes;
    uint64_t raw = 0;
    for (const Level& lv : L) raw += (uint64_t)lv.h * lv.w * sizeof(T);
    rep->raw_bytes = raw;
  }
  cleanup();
  return rc;
}

// --------------------------------------------------------------- the reader
using CogLevelR = TiffLevel;

}  // namespace fsg

struct fsg_cog {
  int fd = -1;
  bool big = false;
  std::vector<fsg::CogLevelR> lv;
  double gt[6] = {0, 1, 0, 0, 0, -1};
  int has_gt = 0, has_nodata = 0;
  double nodata = 0.0;
  int epsg = 0, geographic = 0;
};

namespace fsg {

template <typename T>
static int read_window(fsg_cog* c, const CogLevelR& L, int r0, int c0, int wh, int ww, T* out,
                       int64_t ldo, int threads, hipStream_t s) {
  const int tx0 = c0 / L.tw, tx1 = (c0 + ww - 1) / L.tw;
  const int ty0 = r0 / L.th, ty1 = (r0 + wh - 1) / L.th;
  const int ntx = (L.w + L.tw - 1) / L.tw;
  std::vector<DecTile> tiles;
  std::vector<int> ids;
  for (int ty = ty0; ty <= ty1; ty++)
    for (int tx = tx0; tx <= tx1; tx++) {
      tiles.push_back({ty * L.th, tx * L.tw});
      ids.push_back(ty * ntx + tx);
    }
  const size_t tile_bytes = (size_t)L.tw * L.th * sizeof(T);
  const int n = (int)tiles.size();
  // bytes a tile decompresses to: full tiles; a short last strip
  auto expect = [&](int i) {
    const int rows = L.strips ? std::min(L.th, L.h - tiles[i].tile_y0) : L.th;
    return (size_t)L.tw * rows * sizeof(T);
  };
  uint8_t* h = nullptr;
  uint8_t* d = nullptr;
  DecTile* dt = nullptr;
  if (check_hip(hipHostMalloc(&h, tile_bytes * n, hipHostMallocDefault), "fsg_cog_read: pinned"))
    return FSG_E_HIP;
  if (check_hip(hipMalloc(&d, tile_bytes * n), "fsg_cog_read: staging") ||
      check_hip(hipMalloc(&dt, sizeof(DecTile) * n), "fsg_cog_read: staging")) {
    (void)hipHostFree(h);
    if (d) (void)hipFree(d);
    return FSG_E_HIP;
  }
  std::atomic<int> bad{0};
  parallel_for(n, n_workers(threads), [&](int i, int) {
    const int id = ids[i];
    if (id >= (int)L.off.size() || id >= (int)L.cnt.size()) { bad = 1; return; }
    // the parse bounded every tile's bytes by the file size
    std::vector<uint8_t> comp;
    try {
      comp.resize(L.cnt[id]);
    } catch (const std::bad_alloc&) {
      bad = 1;
      return;
    }
    if (read_all(c->fd, comp.data(), comp.size(), L.off[id])) { bad = 1; return; }
    uint8_t* dst = h + tile_bytes * i;
    const size_t want = expect(i);
    if (L.comp == COMP_ZSTD) {
      ZSTD_DCtx* z = ZSTD_createDCtx();
      const size_t r = ZSTD_decompressDCtx(z, dst, tile_bytes, comp.data(), comp.size());
      ZSTD_freeDCtx(z);
      if (ZSTD_isError(r) || r != want) bad = 1;
    } else if (L.comp == COMP_DEFLATE || L.comp == COMP_ADOBE_DEFLATE) {
      uLongf dl = (uLongf)tile_bytes;
      if (uncompress(dst, &dl, comp.data(), (uLong)comp.size()) != Z_OK || dl != want)
        bad = 1;
    } else if (L.comp == COMP_NONE) {
      if (comp.size() < want) bad = 1;
      else std::memcpy(dst, comp.data(), want);
    } else {
      bad = 1;
    }
  });
  int rc = FSG_OK;
  if (bad) rc = fail(FSG_E_INVALID, "fsg_cog_read: a tile could not be read or decompressed");
  if (rc == FSG_OK &&
      (check_hip(hipMemcpyAsync(d, h, tile_bytes * n, hipMemcpyHostToDevice, s), "fsg_cog_read") ||
       check_hip(hipMemcpyAsync(dt, tiles.data(), sizeof(DecTile) * n, hipMemcpyHostToDevice, s),
                 "fsg_cog_read")))
    rc = FSG_E_HIP;
  if (rc == FSG_OK) {
    const size_t lds = (((size_t)L.tw * sizeof(T) + 15) & ~(size_t)15) + 256 * sizeof(uint32_t);
    hipLaunchKernelGGL(k_tile_decode<T>, dim3(L.th, n), dim3(256), lds, s, d, dt, L.tw, L.th,
                       L.pred, r0, c0, wh, ww, out, ldo);
    if (hipError_t e = hipGetLastError()) rc = check_hip(e, "k_tile_decode");
    else count_launch();
  }
  // the staging buffers are released once the stream is done with them
  if (check_hip(hipStreamSynchronize(s), "fsg_cog_read") && rc == FSG_OK) rc = FSG_E_HIP;
  (void)hipHostFree(h);
  (void)hipFree(d);
  (void)hipFree(dt);
  return rc;
}

}  // namespace fsg

using namespace fsg;

extern "C" int fsg_cog_write(const char* path, const fsg_cog_raster* r,
                             const fsg_cog_options* opt, fsg_cog_report* rep, void* stream) {
  if (!path || !r || !r->data || r->h <= 0 || r->w <= 0 || r->ld < r->w ||
      (r->dtype != FSG_OUT_F32 && r->dtype != FSG_OUT_I16 && r->dtype != FSG_OUT_U8))
    return fail(FSG_E_INVALID, "fsg_cog_write: bad raster");
  fsg_cog_options o = opt ? *opt : fsg_cog_options{512, 1, 8, 0};
  if (o.block < 16 || o.block > 4096 || o.block % 16 || o.zstd_level < 1 || o.zstd_level > 22 ||
      o.overview_count < 0)
    return fail(FSG_E_INVALID, "fsg_cog_write: bad options (block 16..4096, multiple of 16)");
  hipStream_t s = as_stream(stream);
  if (r->dtype == FSG_OUT_F32) return write_cog<float>(path, *r, o, rep, s);
  if (r->dtype == FSG_OUT_I16) return write_cog<int16_t>(path, *r, o, rep, s);
  return write_cog<uint8_t>(path, *r, o, rep, s);
}

extern "C" int fsg_cog_open(const char* path, fsg_cog** out) {
  if (!path || !out) return fail(FSG_E_INVALID, "fsg_cog_open: bad arguments");
  *out = nullptr;
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return fail(FSG_E_INVALID, std::string("fsg_cog_open: cannot open (") + path + ")");
  TiffFile F;
  std::string why;
  try {
    why = tiff_parse(fd, F);
  } catch (const std::bad_alloc&) {
    why = "out of host memory";
  }
  if (!why.empty()) {
    close(fd);
    return fail(FSG_E_INVALID, "fsg_cog_open: " + why + " (" + path + ")");
  }
  auto* c = new fsg_cog();
  c->fd = fd;
  c->big = F.big;
  c->lv = std::move(F.lv);
  for (int k = 0; k < 6; k++) c->gt[k] = F.gt[k];
  c->has_gt = F.has_gt;
  c->has_nodata = F.has_nodata;
  c->nodata = F.nodata;
  c->epsg = F.epsg;
  c->geographic = F.geographic;
  *out = c;
  return FSG_OK;
}

extern "C" int fsg_cog_info(const fsg_cog* c, int32_t level, fsg_cog_info_t* info) {
  if (!c || !info || level < 0 || level >= (int)c->lv.size())
    return fail(FSG_E_INVALID, "fsg_cog_info: bad handle or level");
  const CogLevelR& L = c->lv[level];
  info->n_levels = (int32_t)c->lv.size();
  info->width = L.w;
  info->height = L.h;
  info->tile_w = L.tw;
  info->tile_h = L.th;
  info->dtype = L.bps == 32 ? FSG_OUT_F32 : (L.bps == 16 ? FSG_OUT_I16 : FSG_OUT_U8);
  info->compression = L.comp;
  info->predictor = L.pred;
  info->has_geotransform = c->has_gt;
  for (int k = 0; k < 6; k++) info->geotransform[k] = c->gt[k];
  info->epsg = c->epsg;
  info->epsg_geographic = c->geographic;
  info->has_nodata = c->has_nodata;
  info->nodata = c->nodata;
  return FSG_OK;
}

extern "C" int fsg_cog_read(fsg_cog* c, int32_t level, int32_t r0, int32_t c0, int32_t h,
                            int32_t w, void* out, int64_t ldo, int32_t threads, void* stream) {
  if (!c || level < 0 || level >= (int)c->lv.size() || !out)
    return fail(FSG_E_INVALID, "fsg_cog_read: bad handle, level or output");
  const CogLevelR& L = c->lv[level];
  if (h <= 0 || w <= 0 || r0 < 0 || c0 < 0 || (int64_t)r0 + h > L.h ||
      (int64_t)c0 + w > L.w || ldo < w)
    return fail(FSG_E_INVALID, "fsg_cog_read: window outside the level");
  if (L.pred != 1 && L.pred != 2 && L.pred != 3)
    return fail(FSG_E_UNSUPPORTED, "fsg_cog_read: predictor not 1, 2 or 3");
  hipStream_t s = as_stream(stream);
  if (L.bps == 32) return read_window<float>(c, L, r0, c0, h, w, (float*)out, ldo, threads, s);
  if (L.bps == 16) return read_window<int16_t>(c, L, r0, c0, h, w, (int16_t*)out, ldo, threads, s);
  return read_window<uint8_t>(c, L, r0, c0, h, w, (uint8_t*)out, ldo, threads, s);
}

extern "C" void fsg_cog_close(fsg_cog* c) {
  if (!c) return;
  if (c->fd >= 0) close(c->fd);
  delete c;
}
