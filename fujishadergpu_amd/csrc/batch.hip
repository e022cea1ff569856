// Batched fused launch: many independent tiles (ragged sizes, each its own
// window) and, per tile, hillshade + slope + curvature + TopoUSM, or
// hillshade + slope + curvature (NaN-aware included) + openness, in ONE
// persistent kernel that drains a device work queue, after one batched launch
// over the NaN-free stencil tiles' interiors (k_batch_rows).
//
// The reference runs one tile at a time and one algorithm at a time
// (core/tile_processor.py:1635-1683 calling run_tile_algorithm per tile);
// per-tile, per-algorithm launches are the "unfused" comparison here.
//
// Work items are TopoUSM main-pass blocks (SW x FAST_SH of a tile region,
// the heavier items), openness blocks (64 x 16 pixels, one sparse ray march
// per pixel: rays.hpp) and stencil blocks (64 x 16 output pixels).  A stencil
// item stages the block's filled window values with a 3-pixel halo in LDS
// once, tabulates the first derivatives and derives hillshade, slope and
// curvature from them (stencil_tile_block, the per-window kernels' block
// function too); both item kinds
// run the same device functions as the per-window kernels (stencil.hpp,
// usm.hpp), so fused and unfused outputs are identical bit for bit.  The row
// items (stencil4_wave) form the same first derivatives with the same
// correctly rounded divisions; curvature's second derivatives divide without
// the range guard, the same quotients for dividends in [2^-75, 2^75].
// Per-tile coarse-grid preparation (decimation, coarse box means) is
// enqueued before the persistent kernel; the whole sequence can be captured
// once in a hipGraph and replayed.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "rays.hpp"
#include "stencil.hpp"
#include "stencil_plan.hpp"
#include "usm.hpp"

namespace fsg {

// One tile's openness region: 64 x OPEN_ITEM_H pixel items.
constexpr int OPEN_ITEM_H = 16;
struct OpenTileD {
  WinDevR W;
  int r0, c0, rh, rw;
  int nbx;             // 64-column items per item row
  OutDev out;
};

// Item kinds a persistent kernel instance carries (template bit mask).
constexpr int BK_USM = 1, BK_ST = 2, BK_ST_NAN = 4, BK_OPEN = 8;

struct BatchDev {
  const UsmArgs* usm;
  const int* usm_pre;  // prefix of main-pass blocks per fused TopoUSM tile (n + 1)
  int n_usm_tiles, n_usm_items;
  const StencilTile* st;
  const int* st_pre;   // prefix of stencil blocks per tile (n + 1)
  int n_st_tiles, n_st_items;
  const OpenTileD* op;
  const int* op_pre;   // prefix of openness items per tile (n + 1)
  int n_op_tiles, n_op_items;
  int* queue;
  int order;           // 0: TopoUSM, openness, stencil items; 1: stencil, openness, TopoUSM
  int chunk;           // items claimed per queue access
  int pre_off;         // LDS byte offset of the staged prefix arrays
  Gauss9 G;
  OpenTab T;           // the batch's one openness ray table
};

// largest t with pre[t] <= item (pre[0] = 0, pre strictly covers the items)
__device__ __forceinline__ int find_tile(const int* pre, int n, int item) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One openness item: rows by*16 .. by*16 + 15, columns bx*64 .. bx*64 + 63 of
// the tile region, as four 64 x 4 passes of k_openness's block (the same
// per-block interior test, per-pixel functions and epilogue: the same bits).
__device__ __forceinline__ void openness_item(const OpenTileD& O, const OpenTab& T, int by,
                                              int bx) {
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  const PostK pp = load_post(O.out);
  for (int sub = 0; sub < OPEN_ITEM_H / 4; sub++) {
    const int rb = by * OPEN_ITEM_H + sub * 4;
    if (rb >= O.rh) break;  // item-uniform
    const int bi0 = O.r0 + rb, bj0 = O.c0 + bx * 64;
    const bool interior = openness_interior(O.W, T, bi0, bi0 + 3, bj0, bj0 + 63);
    const int r = rb + ly, c = bx * 64 + lx;
    if (r < O.rh && c < O.rw) {
      const int i = O.r0 + r, j = O.c0 + c;
      const float z = O.W.p[(int64_t)i * O.W.ld + j];
      const float v = isnan(z) ? __builtin_nanf("")
                               : openness_value(T, openness_angle_at(O.W, i, j, z, T, interior));
      store_out(O.out, r, c, apply_post(pp, v));
    }
  }
}

template <int NR, int KINDS>
__device__ __forceinline__ void batch_drain(const BatchDev& B) {
  extern __shared__ float lds[];
  __shared__ int s_item;
  const int n_stencil = B.n_st_items;
  const int total = B.n_usm_items + n_stencil + B.n_op_items;
  // the prefix arrays staged in LDS once: tile lookups stay on chip
  int* upre = (int*)((char*)lds + B.pre_off);
  int* spre = upre + B.n_usm_tiles + 1;
  int* opre = spre + B.n_st_tiles + 1;
  for (int e = threadIdx.x; e <= B.n_usm_tiles; e += 256) upre[e] = B.usm_pre[e];
  for (int e = threadIdx.x; e <= B.n_st_tiles; e += 256) spre[e] = B.st_pre[e];
  if (KINDS & BK_OPEN)
    for (int e = threadIdx.x; e <= B.n_op_tiles; e += 256) opre[e] = B.op_pre[e];
  // item groups in queue order: kinds 0 TopoUSM, 1 openness, 2 stencil
  const int first = B.order ? 2 : 0, last = B.order ? 0 : 2;
  const int n_first = B.order ? n_stencil : B.n_usm_items;
  // every workgroup leaves once the queue passes the item count
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(B.queue, B.chunk);
    __syncthreads();
    const int base = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (base >= total) break;
    const int end = min(base + B.chunk, total);
    for (int item = base; item < end; item++) {
      int kind, it = item;
      if (it < n_first) {
        kind = first;
      } else if (it - n_first < B.n_op_items) {
        kind = 1;
        it -= n_first;
      } else {
        kind = last;
        it -= n_first + B.n_op_items;
      }
      if ((KINDS & BK_USM) && kind == 0) {
        const int tl = __builtin_amdgcn_readfirstlane(find_tile(upre, B.n_usm_tiles, it));
        usm_fast_block<NR, false>(desc(B.usm, tl), it - upre[tl], lds);
      } else if ((KINDS & BK_OPEN) && kind == 1) {
        const int tl = __builtin_amdgcn_readfirstlane(find_tile(opre, B.n_op_tiles, it));
        const OpenTileD& O = desc(B.op, tl);
        const int blk = it - opre[tl];
        const int by = blk / O.nbx;
        openness_item(O, B.T, by, blk - by * O.nbx);
      } else if (kind == 2) {
        const int tl = __builtin_amdgcn_readfirstlane(find_tile(spre, B.n_st_tiles, it));
        const StencilTile& S = desc(B.st, tl);
        const int blk = it - spre[tl];
        const int by = blk / S.nbx, bx = blk - by * S.nbx;
        // the NaN-aware fill only in instances that carry NaN-aware tiles; in
        // the others those tiles take their own launches
        if ((KINDS & BK_ST_NAN) && S.nanaw)
          stencil_tile_block<true>(S, by, bx, B.G, lds);
        else
          stencil_tile_block<false>(S, by, bx, B.G, lds);
      }
    }
  }
}

template <int NR, int KINDS>
__global__ void __launch_bounds__(256) k_batch_fused(BatchDev B) {
  batch_drain<NR, KINDS>(B);
}

// The stencil + openness instance with a register budget of WPE waves per
// SIMD: the openness blocks, the bulk of such a batch, want more waves in
// flight than the stencil blocks' registers would leave.
template <int WPE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_batch_rays(BatchDev B) {
  batch_drain<0, BK_ST | BK_ST_NAN | BK_OPEN>(B);
}
template __global__ void k_batch_rays<5>(BatchDev);
template __global__ void k_batch_rays<6>(BatchDev);
template __global__ void k_batch_rays<8>(BatchDev);

// The NaN-free stencil tiles' interiors as row-walking items, one workgroup
// per item: four adjacent strips of one row segment, a strip per wave.  A
// launch of its own, ahead of the persistent kernel, so that the walk's
// registers alone set its occupancy (inside the persistent kernel the TopoUSM
// path's registers would halve it).  The items cover columns no other item
// writes.
__global__ void __launch_bounds__(256) k_batch_rows(const RowsTile4* rt, const int* pre, int n) {
  const int ir = blockIdx.x;
  int lo = 0, hi = n - 1;  // largest tile t with pre[t] <= ir
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (desc(pre, mid) <= ir) lo = mid; else hi = mid - 1;
  }
  const RowsTile4& R = desc(rt, lo);
  const int blk = ir - desc(pre, lo);
  const int ngx = (R.n_strips + 3) >> 2;
  const int sg = blk / ngx, xg = blk - sg * ngx;
  const int strip = xg * 4 + (threadIdx.x >> 6);
  if (R.T.xrows)
    stencil4_wave<true>(R, strip, sg);
  else
    stencil4_wave<false>(R, strip, sg);
}

// NaN-aware stencil tiles of a batch, one launch per tile.
__global__ void __launch_bounds__(256) k_stencil_nanaw(StencilTile S, Gauss9 G) {
  __shared__ float lds[ST_LDS_FLOATS_ROWS];
  stencil_tile_block<true>(S, blockIdx.y, blockIdx.x, G, lds);
}

__global__ void k_batch_reset(int* __restrict__ q) {
  if (threadIdx.x == 0) q[0] = 0;
}

}  // namespace fsg

using namespace fsg;

struct fsg_batch {
  std::vector<fsg_batch_tile> tiles;
  std::vector<size_t> ws_off;         // per tile: TopoUSM workspace offset, or SIZE_MAX
  std::vector<int> usm_fused;         // tile indices in the fused kernel
  std::vector<int> usm_wave;          // tile indices in the batched wavefront main pass
  WaveBatch wave;
  std::vector<int> usm_fallback;      // tile indices run with per-tile launches
  std::vector<StencilTile> st_fallback;  // NaN-aware stencil tiles (own launches)
  int nr = -1;
  int kinds = 0;                      // item kinds of the persistent kernel (BK_*)
  size_t lds = 0;
  int n_usm_items = 0, n_st_items = 0, n_st_tiles = 0;
  int n_rt_items = 0, n_rt_tiles = 0;
  int n_stencil_in = 0;               // input tiles with stencil work in the kernel
  RowsTile4* d_rt = nullptr;
  int* d_rt_pre = nullptr;
  int n_op_items = 0, n_op_tiles = 0, first_open = -1;
  OpenTab open_tab{};
  OpenTileD* d_op = nullptr;
  int* d_op_pre = nullptr;
  char* dev = nullptr;                // one allocation: workspaces, descriptors, queue
  size_t dev_bytes = 0;
  UsmArgs* d_usm = nullptr;
  int* d_usm_pre = nullptr;
  StencilTile* d_st = nullptr;
  int* d_st_pre = nullptr;
  int* d_queue = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // coarse preparation of the fused TopoUSM tiles: batched launches (decimations,
  // then per coarse-radius level the two box passes), or per tile when a tile's
  // preparation has no batched form
  struct PrepLaunch {
    int kind;     // 0 decimation, 1 box pass 0, 2 box pass 1
    const void* jobs;
    const int* pre;
    int n_jobs, n_blocks;
  };
  std::vector<int> prep_per_tile;
  std::vector<PrepLaunch> prep;
  char* d_prep = nullptr;
};

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

static bool wants_usm(const fsg_batch_tile& t) {
  return t.topousm.data != nullptr && (t.usm.n_radii > 0 || t.usm.field != nullptr);
}
static bool wants_stencil(const fsg_batch_tile& t) {
  return t.hillshade.data || t.slope.data || t.curvature.data;
}
static bool wants_open(const fsg_batch_tile& t) { return t.openness.data != nullptr; }

// Two ray tables are the same table (a batch holds one).
static bool same_open_table(const fsg_openness_params& a, const fsg_openness_params& b) {
  if (a.n_dirs != b.n_dirs || a.n_dist != b.n_dist || a.openness_type != b.openness_type ||
      a.raw_degrees != b.raw_degrees)
    return false;
  for (int e = 0; e < a.n_dirs * a.n_dist; e++) {
    const fsg_ray_sample &x = a.samples[e], &y = b.samples[e];
    if (x.dx != y.dx || x.dy != y.dy || x.phys != y.phys || x.skip != y.skip) return false;
  }
  return true;
}
static const fsg_output* opt(const fsg_output& o) { return o.data ? &o : nullptr; }

// Job tables of the batched coarse preparation, in one device allocation.
static int plan_prep(fsg_batch* b) {
  std::vector<DecJob> dec;
  std::vector<std::vector<BoxJob>> box0, box1;
  std::vector<int> with_prep = b->usm_fused;
  with_prep.insert(with_prep.end(), b->usm_wave.begin(), b->usm_wave.end());
  for (int i : with_prep) {
    const fsg_batch_tile& t = b->tiles[i];
    std::vector<DecJob> d1;
    std::vector<std::vector<BoxJob>> a0, a1;
    if (usm_prep_jobs(&t.win, &t.usm, b->dev + b->ws_off[i], d1, a0, a1) != FSG_OK) {
      b->prep_per_tile.push_back(i);
      continue;
    }
    dec.insert(dec.end(), d1.begin(), d1.end());
    for (size_t L = 0; L < a0.size(); L++) {
      if (box0.size() <= L) {
        box0.emplace_back();
        box1.emplace_back();
      }
      box0[L].insert(box0[L].end(), a0[L].begin(), a0[L].end());
      box1[L].insert(box1[L].end(), a1[L].begin(), a1[L].end());
    }
  }
  // host images of the tables: [jobs][prefix] per launch
  struct Part {
    int kind;
    std::vector<char> jobs;
    std::vector<int> pre;
  };
  std::vector<Part> parts;
  auto add_dec = [&](const std::vector<DecJob>& v) {
    if (v.empty()) return;
    Part P{0, std::vector<char>(sizeof(DecJob) * v.size()), {0}};
    std::memcpy(P.jobs.data(), v.data(), P.jobs.size());
    for (const DecJob& J : v) P.pre.push_back(P.pre.back() + dec_blocks(J));
    parts.push_back(std::move(P));
  };
  auto add_box = [&](int kind, const std::vector<BoxJob>& v) {
    if (v.empty()) return;
    Part P{kind, std::vector<char>(sizeof(BoxJob) * v.size()), {0}};
    std::memcpy(P.jobs.data(), v.data(), P.jobs.size());
    for (const BoxJob& J : v) P.pre.push_back(P.pre.back() + box_blocks(J));
    parts.push_back(std::move(P));
  };
  add_dec(dec);
  for (size_t L = 0; L < box0.size(); L++) {
    add_box(1, box0[L]);
    add_box(2, box1[L]);
  }
  if (parts.empty()) return FSG_OK;
  size_t bytes = 0;
  for (const Part& P : parts) bytes += al256(P.jobs.size()) + al256(sizeof(int) * P.pre.size());
  if (int rc = check_hip(hipMalloc((void**)&b->d_prep, bytes), "fsg_batch_create")) return rc;
  size_t off = 0;
  for (const Part& P : parts) {
    char* jobs = b->d_prep + off;
    off += al256(P.jobs.size());
    int* pre = (int*)(b->d_prep + off);
    off += al256(sizeof(int) * P.pre.size());
    int rc = check_hip(hipMemcpy(jobs, P.jobs.data(), P.jobs.size(), hipMemcpyHostToDevice),
                       "fsg_batch_create");
    if (!rc)
      rc = check_hip(hipMemcpy(pre, P.pre.data(), sizeof(int) * P.pre.size(),
                               hipMemcpyHostToDevice), "fsg_batch_create");
    if (rc) return rc;
    if (P.pre.back() <= 0) continue;
    b->prep.push_back(fsg_batch::PrepLaunch{P.kind, jobs, pre, (int)P.pre.size() - 1,
                                            P.pre.back()});
  }
  return FSG_OK;
}

// Interiors smaller than this keep the tile's blocks.
constexpr int64_t BATCH_ROWS_MIN_PX = int64_t(1) << 16;
// Block strips of a tile whose interior is walked by rows: the interior
// columns left at either side, and four border strips.
constexpr int BATCH_TILE_STRIPS = 6;

// A NaN-free stencil tile's interior as row items (appended to rt, rt_pre)
// and the rest of its region as up to six block strips; false: the tile
// keeps its blocks (NaN-aware windows, a hillshade z factor other than 1, a
// hillshade/slope gradient that differs from curvature's, interiors too small
// or too narrow for one 256-column wave).
static bool rows_item_plan(const StencilTile& S, std::vector<RowsTile4>& rt,
                           std::vector<int>& rt_pre, Rect* strips, int& n_strips) {
  const bool shade = S.want_hs || S.want_sl;
  if (S.nanaw || (S.want_hs && !S.shade.z_is_one) ||
      (shade && S.want_cv && !S.shade_from_tables))
    return false;
  const Rect in = interior(S.W.h, S.W.w, S.r0, S.c0, S.rh, S.rw, 2);
  if ((int64_t)in.h * in.w < BATCH_ROWS_MIN_PX) return false;
  RowsTile R{};
  R.W = S.W;
  R.r0 = S.r0;
  R.c0 = S.c0;
  R.ir0 = in.r0;
  R.ir1 = in.r0 + in.h;
  R.ic0 = in.c0;
  R.ic1 = in.c0 + in.w;
  if (S.want_cv) {
    R.dx = make_divk(S.curv.ax.two_h);
    R.dy = make_divk(S.curv.ay.two_h);
    R.sx = make_divk(S.curv.sx.two_h);
    R.sy = make_divk(S.curv.sy.two_h);
  } else {
    R.dx = make_divk(S.shade.ax.two_h);
    R.dy = make_divk(S.shade.ay.two_h);
  }
  R.shade = S.shade;
  R.curv = S.curv;
  R.want_hs = S.want_hs;
  R.want_sl = S.want_sl;
  R.want_cv = S.want_cv;
  R.o0 = S.hs;
  R.o1 = S.sl;
  R.o2 = S.cv;
  R.xrows = S.xrows;
  RowsTile4 S4;
  dim3 grd;
  int oc0 = 0, oc1 = 0;
  const OutDev outs[3] = {S.hs, S.sl, S.cv};
  if (!rows4_setup(R, in, outs, 3, SR4_LEFT_CURV, S4, grd, oc0, oc1, SR_SEG_MAX)) return false;
  n_strips = 0;
  const int ic1 = in.c0 + in.w;
  if (in.c0 < oc0) strips[n_strips++] = Rect{in.r0, in.c0, in.h, oc0 - in.c0};
  if (oc1 < ic1) strips[n_strips++] = Rect{in.r0, oc1, in.h, ic1 - oc1};
  n_strips += border_strips(S.r0, S.c0, S.rh, S.rw, in, strips + n_strips);
  rt.push_back(S4);
  rt_pre.push_back(rt_pre.back() + (int)grd.x * (int)grd.y);
  return true;
}

extern "C" int fsg_batch_create(const fsg_batch_tile* tiles, int32_t n, fsg_batch** out) {
  if (!out) return fail(FSG_E_INVALID, "fsg_batch_create: out NULL");
  *out = nullptr;
  if (n <= 0 || !tiles) return fail(FSG_E_INVALID, "fsg_batch_create: no tiles");
  fsg_batch* b = new fsg_batch();
  b->tiles.assign(tiles, tiles + n);
  b->ws_off.assign(n, SIZE_MAX);
  // validation (the per-tile entry points' rules) and TopoUSM workspaces
  size_t off = 0;
  for (int i = 0; i < n; i++) {
    const fsg_batch_tile& t = b->tiles[i];
    int rc = validate_window(&t.win, "fsg_batch_create");
    if (!rc) rc = validate_region(&t.win, &t.reg, "fsg_batch_create");
    for (const fsg_output* o : {opt(t.hillshade), opt(t.slope), opt(t.curvature), opt(t.topousm),
                                opt(t.topousm2), opt(t.openness)})
      if (!rc) rc = validate_output(o, &t.reg, "fsg_batch_create");
    if (!rc && wants_open(t)) {
      OpenTab T;
      rc = openness_table(&t.open, T, "fsg_batch_create");
      if (!rc && b->n_op_tiles == 0) b->open_tab = T;
      if (!rc && b->n_op_tiles > 0 && !same_open_table(t.open, b->tiles[b->first_open].open))
        rc = fail(FSG_E_UNSUPPORTED, "fsg_batch_create: the openness tiles of one batch share "
                                     "one ray table");
      if (!rc && b->n_op_tiles++ == 0) b->first_open = i;
    }
    if (!rc && wants_stencil(t)) {
      if (t.win.h < 3 || t.win.w < 3)
        rc = fail(FSG_E_INVALID, "fsg_batch_create: window smaller than 3 x 3");
      else if ((t.hillshade.data || t.slope.data) && !(t.shade.step_x > 0 && t.shade.step_y > 0))
        rc = fail(FSG_E_INVALID, "fsg_batch_create: shade steps must be positive");
      else if (t.curvature.data && (t.curv.step_x == 0.0 || t.curv.step_y == 0.0 ||
                                    t.curv.curvature_type < 0 || t.curv.curvature_type > 3))
        rc = fail(FSG_E_INVALID, "fsg_batch_create: bad curvature parameters");
    }
    if (rc) {
      delete b;
      return rc;
    }
    if (wants_usm(t)) {
      b->ws_off[i] = off;
      off += al256(std::max<size_t>(256, fsg_topousm_workspace_size(&t.win, &t.usm)));
    }
  }
  const size_t ws_total = off;
  const size_t usm_bytes = al256(sizeof(UsmArgs) * n), pre_bytes = al256(sizeof(int) * (n + 1));
  // a tile's stencil work: its blocks, or row items for the interior and up
  // to BATCH_TILE_STRIPS strips of blocks
  const size_t st_cap = (size_t)BATCH_TILE_STRIPS * n;
  const size_t st_bytes = al256(sizeof(StencilTile) * st_cap);
  const size_t st_pre_bytes = al256(sizeof(int) * (st_cap + 1));
  const size_t rt_bytes = al256(sizeof(RowsTile4) * n);
  const size_t op_bytes = al256(sizeof(OpenTileD) * n);
  b->dev_bytes = ws_total + usm_bytes + 3 * pre_bytes + st_bytes + st_pre_bytes + rt_bytes +
                 op_bytes + 256;
  if (int rc = check_hip(hipMalloc((void**)&b->dev, b->dev_bytes), "fsg_batch_create")) {
    delete b;
    return rc;
  }
  char* p = b->dev + ws_total;
  b->d_usm = (UsmArgs*)p;
  p += usm_bytes;
  b->d_usm_pre = (int*)p;
  p += pre_bytes;
  b->d_st = (StencilTile*)p;
  p += st_bytes;
  b->d_st_pre = (int*)p;
  p += st_pre_bytes;
  b->d_rt = (RowsTile4*)p;
  p += rt_bytes;
  b->d_rt_pre = (int*)p;
  p += pre_bytes;
  b->d_op = (OpenTileD*)p;
  p += op_bytes;
  b->d_op_pre = (int*)p;
  p += pre_bytes;
  b->d_queue = (int*)p;

  // The persistent kernel's item kinds: with openness tiles, stencil blocks
  // (NaN-aware included) and openness blocks, and the TopoUSM tiles take their
  // own launches; otherwise TopoUSM and NaN-free stencil blocks.
  b->kinds = b->n_op_tiles > 0 ? (BK_ST | BK_ST_NAN | BK_OPEN) : (BK_USM | BK_ST);
  // TopoUSM arguments (no launches): fused when on the fast path, NaN-free,
  // and with the batch's common number of box radii
  std::vector<UsmArgs> args(n);
  std::vector<int> blocks(n, 0), fast_ok(n, 0);
  std::map<int, int> nr_count;
  for (int i = 0; i < n; i++) {
    const fsg_batch_tile& t = b->tiles[i];
    if (!wants_usm(t)) continue;
    bool fast = false;
    int rc = usm_prepare(&t.win, &t.reg, &t.usm, &t.topousm, opt(t.topousm2),
                         b->dev + b->ws_off[i],
                         fsg_topousm_workspace_size(&t.win, &t.usm), nullptr, &args[i], &fast,
                         &blocks[i], /*launch_prep=*/false);
    if (rc) {
      fsg_batch_destroy(b);
      return rc;
    }
    if ((b->kinds & BK_USM) && fast && t.usm.nan_hint == 0 && blocks[i] > 0 &&
        fast_lds_bytes(args[i], false) <= 150 * 1024) {
      fast_ok[i] = 1;
      nr_count[args[i].nr]++;
    }
  }
  int best = -1, best_n = 0;
  for (auto& kv : nr_count)
    if (kv.second > best_n) best = kv.first, best_n = kv.second;
  b->nr = best;
  // the fused-path tiles the batched wavefront main pass takes leave the
  // persistent kernel (its single-wavefront blocks run the main pass faster
  // than the workgroup blocks, the same bits)
  std::vector<WaveTileIn> wave_in;
  std::vector<int> wave_idx;
  for (int i = 0; i < n; i++) {
    const fsg_batch_tile& t = b->tiles[i];
    if (!wants_usm(t) || !fast_ok[i] || args[i].nr != best) continue;
    wave_in.push_back(WaveTileIn{&t.win, &t.usm, b->dev + b->ws_off[i], &args[i]});
    wave_idx.push_back(i);
  }
  std::vector<char> taken;
  if (int rc = wave_batch_create(wave_in, taken, &b->wave)) {
    fsg_batch_destroy(b);
    return rc;
  }
  std::vector<char> in_wave(n, 0);
  for (size_t k = 0; k < wave_idx.size(); k++)
    if (taken[k]) {
      in_wave[wave_idx[k]] = 1;
      b->usm_wave.push_back(wave_idx[k]);
    }
  std::vector<UsmArgs> fused_args;
  std::vector<int> usm_pre{0};
  size_t lds = sizeof(float) * ST_LDS_FLOATS_ROWS;
  for (int i = 0; i < n; i++) {
    if (!wants_usm(b->tiles[i]) || in_wave[i]) continue;
    if (fast_ok[i] && args[i].nr == best) {
      b->usm_fused.push_back(i);
      fused_args.push_back(args[i]);
      usm_pre.push_back(usm_pre.back() + blocks[i]);
      lds = std::max(lds, fast_lds_bytes(args[i], false));
    } else {
      b->usm_fallback.push_back(i);
    }
  }
  b->n_usm_items = usm_pre.back();
  // stencil descriptors
  std::vector<StencilTile> st;
  std::vector<int> st_pre{0};
  std::vector<RowsTile4> rt;
  std::vector<int> rt_pre{0};
  for (int i = 0; i < n; i++) {
    const fsg_batch_tile& t = b->tiles[i];
    if (!wants_stencil(t) || t.reg.h == 0 || t.reg.w == 0) continue;
    StencilTile S{};
    S.W = WinDev{t.win.data, t.win.ld, t.win.h, t.win.w};
    S.r0 = t.reg.r0;
    S.c0 = t.reg.c0;
    S.rh = t.reg.h;
    S.rw = t.reg.w;
    S.nbx = (t.reg.w + ST_TW - 1) / ST_TW;
    S.want_hs = t.hillshade.data != nullptr;
    S.want_sl = t.slope.data != nullptr;
    S.want_cv = t.curvature.data != nullptr;
    S.nanaw = (S.want_hs || S.want_sl ? t.shade.nan_hint != 0 : 0) ||
              (S.want_cv ? t.curv.nan_hint != 0 : 0);
    if (S.want_hs || S.want_sl) S.shade = make_shadek(&t.shade, S.want_hs, S.want_sl);
    if (S.want_cv) S.curv = make_curvk(&t.curv);
    if ((S.want_hs || S.want_sl) && S.want_cv && t.shade.step_x_rows != t.curv.step_x_rows) {
      fsg_batch_destroy(b);
      return fail(FSG_E_INVALID, "fsg_batch_create: per-row x steps differ between the "
                                 "shade and curvature parameters of one tile");
    }
    S.xrows = (S.want_hs || S.want_sl) ? t.shade.step_x_rows : t.curv.step_x_rows;
    S.shade_from_tables = !S.want_cv || (same_axis(S.shade.ax, S.curv.ax) &&
                                         same_axis(S.shade.ay, S.curv.ay));
    S.hs = S.want_hs ? make_out(&t.hillshade) : OutDev{};
    S.sl = S.want_sl ? make_out(&t.slope) : OutDev{};
    S.cv = S.want_cv ? make_out(&t.curvature) : OutDev{};
    if (S.nanaw && !(b->kinds & BK_ST_NAN)) {
      b->st_fallback.push_back(S);
      continue;
    }
    b->n_stencil_in++;
    // NaN-free tiles with a big enough interior: row items over the interior
    // (the per-window interior kernels' walk), blocks over the border strips
    Rect strips[BATCH_TILE_STRIPS];
    int n_strips = 0;
    if (rows_item_plan(S, rt, rt_pre, strips, n_strips)) {
      for (int k = 0; k < n_strips; k++) {
        const Rect& q = strips[k];
        StencilTile T = S;
        T.r0 = q.r0;
        T.c0 = q.c0;
        T.rh = q.h;
        T.rw = q.w;
        T.nbx = (q.w + ST_TW - 1) / ST_TW;
        T.hs = out_shift(S.hs, q.r0 - S.r0, q.c0 - S.c0);
        T.sl = out_shift(S.sl, q.r0 - S.r0, q.c0 - S.c0);
        T.cv = out_shift(S.cv, q.r0 - S.r0, q.c0 - S.c0);
        st.push_back(T);
        st_pre.push_back(st_pre.back() + T.nbx * ((q.h + ST_TH - 1) / ST_TH));
      }
      continue;
    }
    st.push_back(S);
    st_pre.push_back(st_pre.back() + S.nbx * ((t.reg.h + ST_TH - 1) / ST_TH));
  }
  b->n_st_tiles = (int)st.size();
  b->n_st_items = st_pre.back();
  b->n_rt_tiles = (int)rt.size();
  b->n_rt_items = rt_pre.back();
  // openness descriptors
  std::vector<OpenTileD> op;
  std::vector<int> op_pre{0};
  for (int i = 0; i < n; i++) {
    const fsg_batch_tile& t = b->tiles[i];
    if (!wants_open(t) || t.reg.h == 0 || t.reg.w == 0) continue;
    OpenTileD O{};
    O.W = WinDevR{t.win.data, t.win.ld, t.win.h, t.win.w};
    O.r0 = t.reg.r0;
    O.c0 = t.reg.c0;
    O.rh = t.reg.h;
    O.rw = t.reg.w;
    O.nbx = (t.reg.w + 63) / 64;
    O.out = make_out(&t.openness);
    op.push_back(O);
    op_pre.push_back(op_pre.back() + O.nbx * ((t.reg.h + OPEN_ITEM_H - 1) / OPEN_ITEM_H));
  }
  b->n_op_tiles = (int)op.size();
  b->n_op_items = op_pre.back();
  b->lds = lds;
  if (st.size() > st_cap || rt.size() > (size_t)n) {
    fsg_batch_destroy(b);
    return fail(FSG_E_UNSUPPORTED, "fsg_batch_create: stencil descriptors exceed their buffer");
  }
  int rc = FSG_OK;
  if (!fused_args.empty())
    rc = check_hip(hipMemcpy(b->d_usm, fused_args.data(), sizeof(UsmArgs) * fused_args.size(),
                             hipMemcpyHostToDevice), "fsg_batch_create");
  if (!rc)
    rc = check_hip(hipMemcpy(b->d_usm_pre, usm_pre.data(), sizeof(int) * usm_pre.size(),
                             hipMemcpyHostToDevice), "fsg_batch_create");
  if (!rc && !st.empty())
    rc = check_hip(hipMemcpy(b->d_st, st.data(), sizeof(StencilTile) * st.size(),
                             hipMemcpyHostToDevice), "fsg_batch_create");
  if (!rc && !rt.empty())
    rc = check_hip(hipMemcpy(b->d_rt, rt.data(), sizeof(RowsTile4) * rt.size(),
                             hipMemcpyHostToDevice), "fsg_batch_create");
  if (!rc)
    rc = check_hip(hipMemcpy(b->d_rt_pre, rt_pre.data(), sizeof(int) * rt_pre.size(),
                             hipMemcpyHostToDevice), "fsg_batch_create");
  if (!rc)
    rc = check_hip(hipMemcpy(b->d_st_pre, st_pre.data(), sizeof(int) * st_pre.size(),
                             hipMemcpyHostToDevice), "fsg_batch_create");
  if (!rc && !op.empty())
    rc = check_hip(hipMemcpy(b->d_op, op.data(), sizeof(OpenTileD) * op.size(),
                             hipMemcpyHostToDevice), "fsg_batch_create");
  if (!rc)
    rc = check_hip(hipMemcpy(b->d_op_pre, op_pre.data(), sizeof(int) * op_pre.size(),
                             hipMemcpyHostToDevice), "fsg_batch_create");
  if (!rc) rc = plan_prep(b);
  if (rc) {
    fsg_batch_destroy(b);
    return rc;
  }
  *out = b;
  return FSG_OK;
}

template <int NR>
static void launch_fused(const BatchDev& D, int grid, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_fused<NR, BK_USM | BK_ST>), dim3(grid), dim3(256), lds, s, D);
}

static int batch_enqueue(fsg_batch* b, const fsg_batch_config* cfg, hipStream_t s) {
  // coarse grids of the fused TopoUSM tiles (decimation, coarse box means):
  // batched launches, and per-tile ones for tiles without a batched form
  for (const fsg_batch::PrepLaunch& L : b->prep) {
    const int rc = L.kind == 0
                       ? launch_dec_batch((const DecJob*)L.jobs, L.pre, L.n_jobs, L.n_blocks, s)
                       : launch_box_batch(L.kind - 1, (const BoxJob*)L.jobs, L.pre, L.n_jobs,
                                          L.n_blocks, s);
    if (rc) return rc;
  }
  for (int i : b->prep_per_tile) {
    const fsg_batch_tile& t = b->tiles[i];
    UsmArgs A;
    bool fast = false;
    int nb = 0;
    const int rc = usm_prepare(&t.win, &t.reg, &t.usm, &t.topousm, opt(t.topousm2),
                               b->dev + b->ws_off[i], fsg_topousm_workspace_size(&t.win, &t.usm),
                               s, &A, &fast, &nb, /*launch_prep=*/true);
    if (rc) return rc;
  }
  // tiles outside the fused kernel: their own complete launches
  for (int i : b->usm_fallback) {
    const fsg_batch_tile& t = b->tiles[i];
    const int rc = fsg_topousm(&t.win, &t.reg, &t.usm, &t.topousm, opt(t.topousm2),
                               b->dev + b->ws_off[i], fsg_topousm_workspace_size(&t.win, &t.usm),
                               s);
    if (rc) return rc;
  }
  if (int rc = wave_batch_launch(b->wave, s)) return rc;
  for (const StencilTile& S : b->st_fallback) {
    dim3 grd(S.nbx, (S.rh + ST_TH - 1) / ST_TH);
    hipLaunchKernelGGL(k_stencil_nanaw, grd, dim3(256), 0, s, S, make_gauss9());
    FSG_CHECK_LAUNCH("k_stencil_nanaw");
  }
  if (b->n_rt_items > 0) {
    hipLaunchKernelGGL(k_batch_rows, dim3(b->n_rt_items), dim3(256), 0, s, b->d_rt, b->d_rt_pre,
                       b->n_rt_tiles);
    FSG_CHECK_LAUNCH("k_batch_rows");
  }
  const int total = b->n_usm_items + b->n_st_items + b->n_op_items;
  if (total == 0) return FSG_OK;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, b->d_queue);
  FSG_CHECK_LAUNCH("k_batch_reset");
  int n_cu = cfg ? cfg->n_cu : 0;
  if (n_cu <= 0) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      n_cu = prop.multiProcessorCount;
    if (n_cu <= 0) n_cu = 256;
  }
  const int per_cu = (cfg && cfg->wg_per_cu > 0) ? cfg->wg_per_cu : 3;
  const int grid = std::max(1, std::min(total, n_cu * per_cu));
  BatchDev D;
  D.usm = b->d_usm;
  D.usm_pre = b->d_usm_pre;
  D.n_usm_tiles = (int)b->usm_fused.size();
  D.n_usm_items = b->n_usm_items;
  D.st = b->d_st;
  D.st_pre = b->d_st_pre;
  D.n_st_tiles = b->n_st_tiles;
  D.n_st_items = b->n_st_items;
  D.op = b->d_op;
  D.op_pre = b->d_op_pre;
  D.n_op_tiles = b->n_op_tiles;
  D.n_op_items = b->n_op_items;
  D.T = b->open_tab;
  D.queue = b->d_queue;
  D.order = cfg ? cfg->order : 0;
  D.chunk = (cfg && cfg->chunk > 0) ? cfg->chunk : 8;
  D.pre_off = (int)((b->lds + 15) & ~(size_t)15);
  const size_t lds_total =
      (size_t)D.pre_off +
      sizeof(int) * ((size_t)D.n_usm_tiles + D.n_st_tiles + D.n_op_tiles + 3);
  if (lds_total > 150 * 1024)
    return fail(FSG_E_UNSUPPORTED, "fsg_batch_run: too many tiles for the LDS tile index");
  D.G = make_gauss9();
  const int nr = b->usm_fused.empty() ? 0 : b->nr;
  TimedScope ts("k_batch_fused", s);
  if (b->kinds & BK_OPEN) {
    // more workgroups per CU than the full-register instance holds take a
    // register-capped instance (spilling some of the stencil path)
    if (per_cu <= 4)
      hipLaunchKernelGGL((k_batch_fused<0, BK_ST | BK_ST_NAN | BK_OPEN>), dim3(grid), dim3(256),
                         lds_total, s, D);
    else if (per_cu == 5)
      hipLaunchKernelGGL(k_batch_rays<5>, dim3(grid), dim3(256), lds_total, s, D);
    else if (per_cu == 6)
      hipLaunchKernelGGL(k_batch_rays<6>, dim3(grid), dim3(256), lds_total, s, D);
    else
      hipLaunchKernelGGL(k_batch_rays<8>, dim3(grid), dim3(256), lds_total, s, D);
    FSG_CHECK_LAUNCH("k_batch_fused");
    return FSG_OK;
  }
  switch (nr) {
    case 0: launch_fused<0>(D, grid, lds_total, s); break;
    case 1: launch_fused<1>(D, grid, lds_total, s); break;
    case 2: launch_fused<2>(D, grid, lds_total, s); break;
    case 3: launch_fused<3>(D, grid, lds_total, s); break;
    case 4: launch_fused<4>(D, grid, lds_total, s); break;
    case 5: launch_fused<5>(D, grid, lds_total, s); break;
    case 6: launch_fused<6>(D, grid, lds_total, s); break;
    default: return fail(FSG_E_UNSUPPORTED, "fsg_batch_run: more than 6 box radii");
  }
  FSG_CHECK_LAUNCH("k_batch_fused");
  return FSG_OK;
}

extern "C" int fsg_batch_run(fsg_batch* b, const fsg_batch_config* cfg, void* stream) {
  if (!b) return fail(FSG_E_INVALID, "fsg_batch_run: NULL batch");
  return batch_enqueue(b, cfg, as_stream(stream));
}

extern "C" int fsg_batch_capture(fsg_batch* b, const fsg_batch_config* cfg, void* stream) {
  if (!b) return fail(FSG_E_INVALID, "fsg_batch_capture: NULL batch");
  hipStream_t s = as_stream(stream);
  if (!s) return fail(FSG_E_INVALID, "fsg_batch_capture: capture needs a non-default stream");
  if (b->exec) {
    hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  int rc = check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "fsg_batch_capture");
  if (rc) return rc;
  const int erc = batch_enqueue(b, cfg, s);
  hipGraph_t g = nullptr;
  rc = check_hip(hipStreamEndCapture(s, &g), "fsg_batch_capture");
  if (erc) {
    if (g) hipGraphDestroy(g);
    return erc;
  }
  if (rc) return rc;
  b->graph = g;
  return check_hip(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "fsg_batch_capture");
}

extern "C" int fsg_batch_replay(fsg_batch* b, void* stream) {
  if (!b || !b->exec) return fail(FSG_E_INVALID, "fsg_batch_replay: nothing captured");
  return check_hip(hipGraphLaunch(b->exec, as_stream(stream)), "fsg_batch_replay");
}

extern "C" int fsg_batch_info(const fsg_batch* b, int64_t* info) {
  if (!b || !info) return fail(FSG_E_INVALID, "fsg_batch_info: NULL argument");
  info[0] = b->n_usm_items + b->wave.n_blocks;
  info[1] = b->n_st_items + b->n_rt_items;
  info[2] = (int64_t)(b->usm_fused.size() + b->usm_wave.size());
  info[3] = (int64_t)b->usm_fallback.size();
  info[4] = b->n_stencil_in + (int64_t)b->st_fallback.size();
  info[5] = (int64_t)b->lds;
  info[6] = b->nr;
  info[7] = (int64_t)b->dev_bytes;
  info[8] = b->n_op_items;
  info[9] = b->n_op_tiles;
  info[10] = b->n_rt_tiles;
  info[11] = b->n_rt_items;
  return FSG_OK;
}

extern "C" void fsg_batch_destroy(fsg_batch* b) {
  if (!b) return;
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->dev) hipFree(b->dev);
  if (b->d_prep) hipFree(b->d_prep);
  wave_batch_destroy(&b->wave);
  delete b;
}
