// TIFF / BigTIFF IFD parsing for the COG reader: host code only (no HIP), so
// the same source also builds with the host sanitizers for the malformed-file
// tests.  Reads the IFD chain of a little-endian classic TIFF or BigTIFF and
// keeps, per image, what the windowed reader needs: size, tile (or strip)
// geometry, sample type, compression, predictor and the tile tables, plus the
// first image's GeoTIFF transform, EPSG code and GDAL NoData.
//
// Every count, offset and size read from the file is checked against the file
// size and against fixed limits before anything is allocated or read from it:
// an untrusted file gives an error message, never an allocation of its
// choosing or a read outside it.
#pragma once
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace fsg {

// TIFF tags and codes (TIFF 6.0, BigTIFF, GeoTIFF 1.1, GDAL's NoData tag)
enum : uint16_t {
  T_NEWSUBFILE = 254, T_WIDTH = 256, T_HEIGHT = 257, T_BPS = 258, T_COMPRESSION = 259,
  T_PHOTOMETRIC = 262, T_SPP = 277, T_PLANAR = 284, T_PREDICTOR = 317, T_TILEW = 322,
  T_TILEH = 323, T_TILEOFF = 324, T_TILEBC = 325, T_SAMPLEFMT = 339, T_PIXSCALE = 33550,
  T_TIEPOINT = 33922, T_GEOKEYS = 34735, T_NODATA = 42113
};
enum : uint16_t { TY_BYTE = 1, TY_ASCII = 2, TY_SHORT = 3, TY_LONG = 4, TY_DOUBLE = 12,
                  TY_LONG8 = 16 };
constexpr uint16_t COMP_NONE = 1, COMP_DEFLATE = 8, COMP_ADOBE_DEFLATE = 32946,
                   COMP_ZSTD = 50000;

// limits on what a file may declare
constexpr uint64_t TIFF_MAX_ENTRIES = 4096;        // entries per IFD
constexpr int TIFF_MAX_IFDS = 64;                  // images in the chain
constexpr int64_t TIFF_MAX_SIDE = int64_t(1) << 30;  // image width / height
constexpr uint64_t TIFF_MAX_TILE_BYTES = uint64_t(1) << 30;  // one decoded tile / strip
constexpr uint64_t TIFF_MAX_TILES = uint64_t(1) << 26;

struct TiffLevel {
  int w = 0, h = 0, tw = 0, th = 0, bps = 0, fmt = 1, comp = 1, pred = 1, spp = 1;
  bool strips = false;   // strips: tiles of the full width, the last one short
  std::vector<uint64_t> off, cnt;
};

struct TiffFile {
  bool big = false;
  uint64_t size = 0;
  std::vector<TiffLevel> lv;
  double gt[6] = {0, 1, 0, 0, 0, -1};
  int has_gt = 0, has_nodata = 0;
  double nodata = 0.0;
  int epsg = 0, geographic = 0;
};

inline int tiff_read_all(int fd, void* p, size_t n, uint64_t off) {
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

inline uint64_t tiff_rd(const uint8_t* b, int n) {
  uint64_t v = 0;
  std::memcpy(&v, b, n);
  return v;
}

inline int tiff_type_size(int ty) {
  switch (ty) {
    case 1: case 2: case 6: case 7: return 1;
    case 3: case 8: return 2;
    case 4: case 9: case 11: case 13: return 4;
    case 5: case 10: case 12: case 16: case 17: case 18: return 8;
  }
  return 0;
}

// the bytes [off, off + n) lie inside the file
inline bool tiff_inside(const TiffFile& F, uint64_t off, uint64_t n) {
  return n <= F.size && off <= F.size - n;
}

// Values of one IFD entry as uint64 / double / string; "" or the reason the
// entry is refused.
inline std::string tiff_entry_values(int fd, const TiffFile& F, const uint8_t* ent,
                                     std::vector<uint64_t>* u, std::vector<double>* d,
                                     std::string* str) {
  const bool big = F.big;
  const int ty = (int)tiff_rd(ent + 2, 2);
  const uint64_t cnt = big ? tiff_rd(ent + 4, 8) : tiff_rd(ent + 4, 4);
  const int ts = tiff_type_size(ty);
  if (!ts) return "unknown field type";
  if (cnt > F.size / (uint64_t)ts + 1) return "value count beyond the file";
  const uint64_t bytes = cnt * (uint64_t)ts;
  const uint64_t inl = big ? 8 : 4;
  std::vector<uint8_t> v((size_t)bytes);
  if (bytes <= inl) {
    if (bytes) std::memcpy(v.data(), ent + (big ? 12 : 8), (size_t)bytes);
  } else {
    const uint64_t off = big ? tiff_rd(ent + 12, 8) : tiff_rd(ent + 8, 4);
    if (!tiff_inside(F, off, bytes)) return "value offset beyond the file";
    if (tiff_read_all(fd, v.data(), (size_t)bytes, off)) return "short read";
  }
  for (uint64_t k = 0; k < cnt; k++) {
    uint64_t x = 0;
    double y = 0;
    std::memcpy(&x, v.data() + k * ts, ts);
    if (ty == 12) std::memcpy(&y, v.data() + k * ts, 8);
    else if (ty == 11) { float f; std::memcpy(&f, v.data() + k * ts, 4); y = f; }
    else y = (double)x;
    if (u) u->push_back(x);
    if (d) d->push_back(y);
  }
  if (str && ty == 2) str->assign((const char*)v.data(), strnlen((const char*)v.data(), bytes));
  return "";
}

// an unsigned tag value as an int in [lo, hi], or -1
inline int64_t tiff_int(const std::vector<uint64_t>& u, int64_t lo, int64_t hi) {
  if (u.empty() || u[0] > (uint64_t)hi || (int64_t)u[0] < lo) return -1;
  return (int64_t)u[0];
}

// Parses the IFD chain of the file open at fd; "" or the reason it is refused.
inline std::string tiff_parse(int fd, TiffFile& F) {
  struct stat st;
  if (fstat(fd, &st) != 0 || st.st_size < 16) return "short file";
  F.size = (uint64_t)st.st_size;
  uint8_t hd[16];
  if (tiff_read_all(fd, hd, 16, 0)) return "short file";
  if (hd[0] != 'I' || hd[1] != 'I') return "not a little-endian TIFF";
  const int magic = (int)tiff_rd(hd + 2, 2);
  if (magic != 42 && magic != 43) return "not a TIFF";
  F.big = magic == 43;
  uint64_t ifd = F.big ? tiff_rd(hd + 8, 8) : tiff_rd(hd + 4, 4);
  const uint64_t es = F.big ? 20 : 12;
  const uint64_t nsz = F.big ? 8 : 2, lsz = F.big ? 8 : 4;
  std::vector<uint64_t> seen;
  for (int guard = 0; ifd; guard++) {
    if (guard >= TIFF_MAX_IFDS) return "too many images";
    if (std::find(seen.begin(), seen.end(), ifd) != seen.end()) return "IFD chain loops";
    seen.push_back(ifd);
    if (!tiff_inside(F, ifd, nsz)) return "bad IFD";
    uint8_t nb[8] = {0};
    if (tiff_read_all(fd, nb, (size_t)nsz, ifd)) return "bad IFD";
    const uint64_t n = tiff_rd(nb, (int)nsz);
    if (n > TIFF_MAX_ENTRIES) return "bad IFD entry count";
    const uint64_t tbytes = n * es + lsz;
    if (!tiff_inside(F, ifd + nsz, tbytes)) return "bad IFD";
    std::vector<uint8_t> all((size_t)tbytes);
    if (tiff_read_all(fd, all.data(), all.size(), ifd + nsz)) return "bad IFD";
    TiffLevel L;
    bool tiled = false, has_tw = false, has_th = false;
    int64_t rows_per_strip = 0;
    std::vector<uint64_t> soff, scnt;
    std::vector<double> scale, tie;
    for (uint64_t k = 0; k < n; k++) {
      const uint8_t* ent = all.data() + k * es;
      const int tag = (int)tiff_rd(ent, 2);
      std::vector<uint64_t> u;
      std::vector<double> d;
      std::string str;
      const std::string why = tiff_entry_values(fd, F, ent, &u, &d, &str);
      if (!why.empty()) return "bad tag value: " + why;
      if (u.empty() && str.empty()) continue;
      int64_t v;
      switch (tag) {
        case T_WIDTH:
          if ((v = tiff_int(u, 1, TIFF_MAX_SIDE)) < 0) return "bad width";
          L.w = (int)v;
          break;
        case T_HEIGHT:
          if ((v = tiff_int(u, 1, TIFF_MAX_SIDE)) < 0) return "bad height";
          L.h = (int)v;
          break;
        case T_BPS: L.bps = (int)tiff_int(u, 0, 64); break;
        case T_COMPRESSION: L.comp = (int)tiff_int(u, 0, 65535); break;
        case T_PREDICTOR: L.pred = (int)tiff_int(u, 0, 65535); break;
        case T_SPP: L.spp = (int)tiff_int(u, 0, 65535); break;
        case T_SAMPLEFMT: L.fmt = (int)tiff_int(u, 0, 65535); break;
        case T_TILEW:
          if ((v = tiff_int(u, 1, TIFF_MAX_SIDE)) < 0) return "bad tile width";
          L.tw = (int)v;
          has_tw = true;
          break;
        case T_TILEH:
          if ((v = tiff_int(u, 1, TIFF_MAX_SIDE)) < 0) return "bad tile height";
          L.th = (int)v;
          has_th = true;
          break;
        case T_TILEOFF: L.off = u; break;
        case T_TILEBC: L.cnt = u; break;
        case 273: soff = u; break;             // StripOffsets
        case 278: rows_per_strip = tiff_int(u, 0, TIFF_MAX_SIDE); break;
        case 279: scnt = u; break;             // StripByteCounts
        case T_PIXSCALE:
          if (F.lv.empty() && d.size() >= 2) scale = d;
          break;
        case T_TIEPOINT:
          if (F.lv.empty() && d.size() >= 6) tie = d;
          break;
        case T_GEOKEYS:
          if (F.lv.empty())
            for (size_t q = 4; q + 3 < u.size(); q += 4) {
              if (u[q] == 3072 || u[q] == 2048) {
                F.epsg = (int)std::min<uint64_t>(u[q + 3], 0x7fffffff);
                F.geographic = u[q] == 2048;
              }
            }
          break;
        case T_NODATA:
          if (F.lv.empty() && !str.empty()) {
            F.has_nodata = 1;
            F.nodata = std::strtod(str.c_str(), nullptr);
          }
          break;
        default: break;
      }
    }
    tiled = has_tw && has_th;
    if (!tiled && (has_tw || has_th)) return "tile width without tile height";
    if (!tiled && !soff.empty()) {
      L.strips = true;
      L.tw = L.w;
      L.th = rows_per_strip > 0 ? (int)std::min<int64_t>(rows_per_strip, L.h) : L.h;
      L.off = soff;
      L.cnt = scnt;
      tiled = L.w > 0 && L.h > 0;
    }
    if (!tiled || L.w <= 0 || L.h <= 0 || L.tw <= 0 || L.th <= 0 || L.spp != 1)
      return "only tiled or stripped single-band rasters are read";
    const bool ok_type = (L.bps == 32 && L.fmt == 3) || (L.bps == 16 && L.fmt == 2) ||
                         (L.bps == 8 && L.fmt == 1);
    if (!ok_type) return "sample type is not float32, int16 or uint8";
    if (L.pred == 3 && L.bps != 32) return "predictor 3 on a non-float raster";
    if (L.pred == 2 && L.bps == 32) return "predictor 2 on a float raster";
    if (L.pred != 1 && L.pred != 2 && L.pred != 3) return "unknown predictor";
    if (L.comp != COMP_NONE && L.comp != COMP_DEFLATE && L.comp != COMP_ADOBE_DEFLATE &&
        L.comp != COMP_ZSTD)
      return "compression is not none, DEFLATE or ZSTD";
    if ((uint64_t)L.tw * (uint64_t)L.th * (uint64_t)(L.bps / 8) > TIFF_MAX_TILE_BYTES)
      return "tile larger than 1 GiB";
    const uint64_t ntx = ((uint64_t)L.w + L.tw - 1) / L.tw;
    const uint64_t nty = ((uint64_t)L.h + L.th - 1) / L.th;
    if (ntx * nty > TIFF_MAX_TILES) return "too many tiles";
    const uint64_t nt = ntx * nty;
    if (L.off.size() != nt || L.cnt.size() != nt) return "tile tables do not match the size";
    for (uint64_t t = 0; t < nt; t++)
      if (!tiff_inside(F, L.off[t], L.cnt[t])) return "tile data beyond the file";
    if (F.lv.empty() && !scale.empty() && !tie.empty()) {
      // tiepoint (I, J, K, X, Y, Z) at pixel (I, J); north-up scale (sx, sy)
      F.gt[1] = scale[0];
      F.gt[5] = -scale[1];
      F.gt[0] = tie[3] - tie[0] * F.gt[1];
      F.gt[3] = tie[4] - tie[1] * F.gt[5];
      F.has_gt = 1;
    }
    F.lv.push_back(std::move(L));
    ifd = tiff_rd(all.data() + n * es, (int)lsz);
  }
  if (F.lv.empty()) return "no image";
  return "";
}

}  // namespace fsg
