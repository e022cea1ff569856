// TopoUSM main-pass device code shared by the per-window launch
// (topousm.hip) and the batched fused launch (batch.hip): argument types,
// address helpers, and the fast-path body for one SW x FAST_SH block of the
// region.
#pragma once
#include <cmath>
#include <vector>

#include "fsg_common.hpp"

namespace fsg {

enum { TERM_BOX = 0, TERM_GAUSS = 1, TERM_COARSE = 2 };
constexpr int MAX_BOX = 6;
constexpr int MAX_SRC = 5;  // bilinear sources staged in LDS (coarse means + field)

// A bilinear source: a coarse mean grid sampled with the corner-aligned zoom
// (mode 0, window coordinates, float64 coordinates), or the overview field
// sampled at raster coordinates with float32 coordinates (mode 1), both with
// 'nearest' clamping (_nan_utils.py:255-281, :671-698).
struct SrcRef {
  const float* p;
  int64_t ld;
  int h, w;
  int mode;
  int oy, ox;           // raster origin of the window (mode 1)
  double zy, zx;        // zoom factors (mode 0)
  float fy, fx;         // f32 ratios (mode 1)
  int rows_max, cols_max, lds_off;
};

struct UsmArgs {
  WinDev W;
  int r0, c0, rh, rw;  // region in window coordinates
  int n_terms;
  int kind[FSG_MAX_RADII];
  int slot[FSG_MAX_RADII];  // box slot, or source index for coarse terms
  float wt[FSG_MAX_RADII];
  int nr;                   // number of full-resolution box radii
  int rad[MAX_BOX];
  double inv[MAX_BOX];      // 1 / (2r + 1)
  int R;                    // largest box radius (strip halo)
  int SW, SH;               // strip width, rows per workgroup
  int n_strips;
  int n_src;
  SrcRef src[MAX_SRC];
  int field_src;            // source index of the overview field, or -1
  float wl;                 // f32(w_large)
  int src_floats;           // LDS entries of all staging tiles (float64 in k_usm_fast)
  Gauss9 G;
  OutDev o1, o2;
};

__device__ __forceinline__ int reflf(int i, int n) {
  return (i >= 0 && i < n) ? i : refl(i, n);
}

// Row pointer of a block-uniform row index: the index is forced into a scalar
// register so loads become (scalar row base)[per-lane column offset].
__device__ __forceinline__ const float* urow(const float* base, int row, int64_t ld) {
  const int r = __builtin_amdgcn_readfirstlane(row);
  return base + (int64_t)r * ld;
}

__host__ __device__ inline void src_axis(const SrcRef& S, int row, int idx, int& i0, int& i1,
                                         double& t) {
  const int n = row ? S.h : S.w;
  double c;
  if (S.mode == 0) {
    c = (double)idx * (row ? S.zy : S.zx);
  } else {
    const float f = ((float)(idx + (row ? S.oy : S.ox)) + 0.5f) * (row ? S.fy : S.fx) - 0.5f;
    c = (double)f;
  }
  c = fmin(fmax(c, 0.0), (double)(n - 1));
  i0 = (int)floor(c);
  i1 = (i0 + 1 < n - 1) ? i0 + 1 : n - 1;
  t = c - (double)i0;
}

// Fast path, term order [distinct box radii ascending in slot order]
// [decimated radii][field] -- the order every sorted radius list produces.  A
// workgroup covers SW output columns x FAST_SH rows; each K-row chunk runs
// three phases:
//   1. vertical running sums (one LDS column per thread, all row loads of the
//      chunk in flight together) -> Vs, centre values -> Zs;
//   2. horizontal running sums over row segments and the box terms in order,
//      the partial sum written over the centre value in Zs;
//   3. one output row per step with lane = column: the decimated terms from
//      per-chunk float64 tables of horizontally interpolated coarse rows, the
//      field term, post-processing and both encodings, coalesced.
// Accumulation order is the term order, so results equal k_usm_main's.
constexpr int FAST_SH = 64;
constexpr int FAST_K = 8;

// scipy 'reflect' for an index at most n outside [0, n)
__device__ __forceinline__ int refl1(int i, int n) {
  return i < 0 ? -1 - i : (i >= n ? 2 * n - 1 - i : i);
}

template <int NR, bool NANAW>
__device__ __forceinline__ void usm_fast_block(const UsmArgs& A, int tile, float* lds) {
  constexpr int K = FAST_K;
  constexpr int KB = 4;            // rows per batch of vertical loads
  constexpr int NRA = NR > 0 ? NR : 1;
  const int NC = A.SW + 2 * A.R;
  const int LDC = NC + 1;
  const int LDZ = A.SW + 1;                          // output columns only
  float* Vs = lds;                                   // [NR][K][LDC]
  float* Ws = Vs + NR * K * LDC;                     // [NR][K][LDC] (NaN-aware only)
  float* Zs = NANAW ? (Ws + NR * K * LDC) : Ws;      // [K][LDZ] centre -> partial sum
  // float64 tables from a 16-byte boundary: per source, rows_max staged
  // coarse rows of horizontally interpolated values [row][SW]; then per source
  // K row entries (staged-row offsets, ty, 1 - ty)
  double* Hs_t = (double*)(lds + ((NR * K * LDC * (NANAW ? 2 : 1) + K * LDZ + 3) & ~3));
  int2* RowI = (int2*)(Hs_t + A.src_floats);
  double2* RowT = (double2*)(RowI + A.n_src * K);

  const int t = threadIdx.x;
  const int strip = tile % A.n_strips;
  const int seg = tile / A.n_strips;
  const int x0 = strip * A.SW;
  const int y0 = seg * FAST_SH;
  const int rows = min(FAST_SH, A.rh - y0);
  const int cols = min(A.SW, A.rw - x0);
  if (rows <= 0 || cols <= 0) return;
  const int wx0 = A.c0 + x0, wy0 = A.r0 + y0;
  const int H = A.W.h;
  const int64_t ld = A.W.ld;
  const float* Wp = A.W.p;
  const int nz = A.n_src - (A.field_src >= 0 ? 1 : 0);  // decimated terms

  // vertical running sums over rows wy0-1-r .. wy0-1+r (16 loads in flight)
  const bool has_col = t < NC;
  const int jw = has_col ? refl(wx0 - A.R + t, A.W.w) : 0;
  // block-uniform: every row this block reads (wy0 - R - 1 .. wy0 + FAST_SH - 1 + R)
  // lies inside the window, so no row index needs the reflection and row
  // addresses advance by ld (scalar adds instead of a reflect and a 64-bit
  // multiply per row)
  const bool rows_in = wy0 - A.R - 1 >= 0 && wy0 + FAST_SH - 1 + A.R <= H - 1;
  double S[NRA];
  int Cn[NRA];
#pragma unroll
  for (int i = 0; i < NR; i++) {
    double s = 0.0;
    int n = 0;
    const int r = A.rad[i];
    for (int d0 = -r; d0 <= r; d0 += 16) {
      float v[16];
      if (rows_in) {
        const float* rp = Wp + (int64_t)(wy0 - 1 + d0) * ld;
#pragma unroll
        for (int e = 0; e < 16; e++) {
          v[e] = rp[jw];
          if (d0 + e < r) rp += ld;   // rows past r repeat row r, as min(d0 + e, r) does
        }
      } else {
#pragma unroll
        for (int e = 0; e < 16; e++)
          v[e] = urow(Wp, refl1(wy0 - 1 + min(d0 + e, r), H), ld)[jw];
      }
#pragma unroll
      for (int e = 0; e < 16; e++) {
        if (d0 + e > r) break;
        if (NANAW) {
          if (!isnan(v[e])) { s += (double)v[e]; n++; }
        } else {
          s += (double)v[e];
        }
      }
    }
    S[i] = s;
    Cn[i] = n;
  }

  const int segs = 256 / K;
  const int L = A.SW / segs;
  const int n_chunks = (rows + K - 1) / K;
  const int tc = min(t, cols - 1);  // this lane's output column (clamped)

  for (int ch = 0; ch < n_chunks; ch++) {
    const int wyc = wy0 + ch * K;
    const int nrow = min(K, rows - ch * K);

    // ---------------- phase 1: vertical running sums of K rows
#pragma unroll
    for (int kb = 0; kb < K; kb += KB) {
      float vz[KB], vn[NRA][KB], vo[NRA][KB];
      if (rows_in) {
        const float* rz = Wp + (int64_t)(wyc + kb) * ld;
#pragma unroll
        for (int k = 0; k < KB; k++) {
          vz[k] = rz[jw];
#pragma unroll
          for (int i = 0; i < NR; i++) {
            vn[i][k] = (rz + (int64_t)A.rad[i] * ld)[jw];
            vo[i][k] = (rz - (int64_t)(A.rad[i] + 1) * ld)[jw];
          }
          rz += ld;
        }
      } else {
#pragma unroll
        for (int k = 0; k < KB; k++) {
          const int yw = wyc + kb + k;
          vz[k] = urow(Wp, min(yw, H - 1), ld)[jw];
#pragma unroll
          for (int i = 0; i < NR; i++) {
            vn[i][k] = urow(Wp, refl1(yw + A.rad[i], H), ld)[jw];
            vo[i][k] = urow(Wp, refl1(yw - A.rad[i] - 1, H), ld)[jw];
          }
        }
      }
      if (has_col) {
#pragma unroll
        for (int k = 0; k < KB; k++) {
#pragma unroll
          for (int i = 0; i < NR; i++) {
            if (NANAW) {
              if (!isnan(vn[i][k])) { S[i] += (double)vn[i][k]; Cn[i]++; }
              if (!isnan(vo[i][k])) { S[i] -= (double)vo[i][k]; Cn[i]--; }
              Ws[(i * K + kb + k) * LDC + t] = (float)((double)Cn[i] * A.inv[i]);
            } else {
              S[i] += (double)vn[i][k];
              S[i] -= (double)vo[i][k];
            }
            Vs[(i * K + kb + k) * LDC + t] = (float)(S[i] * A.inv[i]);
          }
          if (t >= A.R && t < A.R + A.SW) Zs[(kb + k) * LDZ + t - A.R] = vz[k];
        }
      }
    }
    // ---------------- staging: horizontal lerps of the coarse rows this
    // chunk samples, top/bot = a * (1 - tx) + b * tx per (coarse row, column)
    for (int s = 0; s < A.n_src; s++) {
      const SrcRef& SR = A.src[s];
      int a0, a1, b0, b1;
      double tt;
      src_axis(SR, 1, wyc, a0, a1, tt);
      src_axis(SR, 1, wyc + nrow - 1, b0, b1, tt);
      const int nr = min(b1 - a0 + 1, SR.rows_max);
      if (t < A.SW) {
        int i0, i1;
        double tx;
        src_axis(SR, 0, wx0 + tc, i0, i1, tx);
        const double mx = 1.0 - tx;
        double* hrow = Hs_t + SR.lds_off + t;
        const float* prow = SR.p + (int64_t)a0 * SR.ld;
        for (int rr = 0; rr < nr; rr++) {
          const double va = prow[i0], vb = prow[i1];
          hrow[rr * A.SW] = va * mx + vb * tx;
          prow += SR.ld;
        }
      }
      if (t < K) {
        int i0, i1;
        double ty;
        src_axis(SR, 1, wyc + min(t, nrow - 1), i0, i1, ty);
        RowI[s * K + t] = make_int2((i0 - a0) * A.SW, (i1 - a0) * A.SW);
        RowT[s * K + t] = make_double2(ty, 1.0 - ty);
      }
    }
    __syncthreads();

    // ---------------- phase 2: horizontal running sums + box terms in order
    if (NR > 0) {
      const int k = t / segs;
      const int sg = t % segs;
      if (k < nrow) {
        const int cbase = A.R + sg * L;
        double Hsum[NRA], HW[NRA];
#pragma unroll
        for (int i = 0; i < NR; i++) {
          const float* vrow = Vs + (i * K + k) * LDC + cbase;
          const float* wrow = Ws + (i * K + k) * LDC + cbase;
          double h = 0.0, hw = 0.0;
          const int r = A.rad[i];
#pragma unroll 8
          for (int d = -r; d <= r; d++) {
            h += (double)vrow[d];
            if (NANAW) hw += (double)wrow[d];
          }
          Hsum[i] = h;
          HW[i] = hw;
        }
        for (int m = 0; m < L; m++) {
          const int lc = cbase + m;
          const float z = Zs[k * LDZ + lc - A.R];
          float acc = 0.0f;
#pragma unroll
          for (int i = 0; i < NR; i++) {
            const int r = A.rad[i];
            if (m > 0) {
              const float* vrow = Vs + (i * K + k) * LDC;
              Hsum[i] += (double)vrow[lc + r] - (double)vrow[lc - r - 1];
              if (NANAW) {
                const float* wrow = Ws + (i * K + k) * LDC;
                HW[i] += (double)wrow[lc + r] - (double)wrow[lc - r - 1];
              }
            }
            const float fv = (float)(Hsum[i] * A.inv[i]);
            float bm;
            if (NANAW) {
              const float fw = (float)(HW[i] * A.inv[i]);
              bm = fw > 0.0f ? fv / fw : 0.0f;
            } else {
              bm = fv;
            }
            const float diff = A.wt[i] * (z - bm);
            acc = (i == 0) ? diff : acc + diff;
          }
          // each output column belongs to exactly one thread: in-place is safe
          Zs[k * LDZ + lc - A.R] = acc;
        }
      }
      __syncthreads();
    }

    // ---------------- phase 3: decimated terms, field, post, store
    if (t < cols) {
      const PostK p1 = load_post(A.o1);
      const bool two = A.o2.data != nullptr;
      float zq[K];
      if (rows_in) {
        const float* rz = Wp + (int64_t)wyc * ld + wx0 + t;
#pragma unroll
        for (int k = 0; k < K; k++) {
          zq[k] = *rz;
          rz += ld;
        }
      } else {
#pragma unroll
        for (int k = 0; k < K; k++) zq[k] = urow(Wp, min(wyc + k, H - 1), ld)[wx0 + t];
      }
#pragma unroll
      for (int k = 0; k < K; k++) {
        if (k < nrow) {
          const float z = zq[k];
          float acc = (NR > 0) ? Zs[k * LDZ + t] : 0.0f;
          for (int j = 0; j < nz; j++) {
            const int2 ri = RowI[j * K + k];
            const double2 ty = RowT[j * K + k];
            const double* h = Hs_t + A.src[j].lds_off + t;
            const float mean = (float)(h[ri.x] * ty.y + h[ri.y] * ty.x);
            const float diff = A.wt[NR + j] * (z - mean);
            acc = (NR == 0 && j == 0) ? diff : acc + diff;
          }
          if (isnan(z)) acc = __builtin_nanf("");
          if (A.field_src >= 0) {
            const int s = A.field_src;
            const int2 ri = RowI[s * K + k];
            const double2 ty = RowT[s * K + k];
            const double* h = Hs_t + A.src[s].lds_off + t;
            const float up = (float)(h[ri.x] * ty.y + h[ri.y] * ty.x);
            const float large = A.wl * z - up;
            acc = (A.n_terms > 0) ? acc + large : large;
          }
          const float v = apply_post(p1, acc);
          const int yr = y0 + ch * K + k, xr = x0 + t;
          store_out(A.o1, yr, xr, v);
          if (two) store_out(A.o2, yr, xr, v);
        }
      }
    }
    __syncthreads();
  }
}

inline size_t fast_lds_bytes(const UsmArgs& A, bool nanaw) {
  const int LDC = A.SW + 2 * A.R + 1;
  size_t lds = sizeof(float) * (((size_t)A.nr * FAST_K * LDC * (nanaw ? 2 : 1) +
                                 (size_t)FAST_K * (A.SW + 1) + 3) & ~(size_t)3);
  lds += sizeof(double) * (size_t)A.src_floats;  // staged rows (float64)
  return lds + (sizeof(int2) + sizeof(double2)) * (size_t)A.n_src * FAST_K;
}

// Validation, coarse-grid preparation (decimation, void fill, coarse box
// means, enqueued on s) and the main-pass arguments of one fsg_topousm call.
// *fast tells whether the fast path applies; *n_blocks is the main pass's
// SW x SH block count.  Nothing of the main pass is launched; with
// launch_prep = false nothing at all is (the arguments alone, for planning).
int usm_prepare(const fsg_window* win, const fsg_region* reg, const fsg_topousm_params* p,
                const fsg_output* out, const fsg_output* out2, void* workspace,
                size_t workspace_bytes, hipStream_t s, UsmArgs* A, bool* fast, int* n_blocks,
                bool launch_prep = true);

// Coarse preparation of many NaN-free windows as a few batched launches: the
// cell-mean decimations, then per coarse-radius level the two box passes.
struct DecJob {
  WinDev W;
  int f;
  float* out;
  int oh, ow;
  int nbx;   // blocks per coarse row (of 1024 source columns for f >= 4, else 256)
};
struct BoxJob {
  const float* a;
  const float* b;
  int64_t lda;
  int h, w, r;
  double inv;
  float* oa;
  float* ob;
  int64_t ldo;
  int nbx;   // blocks per band
  int nsub;  // 64-row sub-bands per band (0 or 1: one; the NaN-free passes only)
};
inline int dec_blocks(const DecJob& J) { return J.nbx * J.oh; }
inline int box_band(const BoxJob& J) { return 64 * (J.nsub > 1 ? J.nsub : 1); }
inline int box_blocks(const BoxJob& J) { return J.nbx * ((J.h + box_band(J) - 1) / box_band(J)); }

// The jobs the coarse preparation of this NaN-free window launches one by one
// (decimations in dec; level L's box passes in box0[L] / box1[L]).  Returns
// FSG_E_UNSUPPORTED when a coarse radius takes the Gaussian branch (r <= 1),
// which stays on the per-window path.
int usm_prep_jobs(const fsg_window* win, const fsg_topousm_params* p, void* workspace,
                  std::vector<DecJob>& dec, std::vector<std::vector<BoxJob>>& box0,
                  std::vector<std::vector<BoxJob>>& box1);
// Batched wavefront main pass over many windows (fsg_batch): one launch of
// k_usm_wave's blocks for the fast-path, NaN-free windows of the most common
// instance among those that take the default radii (2, 8, 32).
struct WaveTileIn {
  const fsg_window* win;
  const fsg_topousm_params* p;
  void* workspace;            // the window's fsg_topousm workspace (its table area)
  const UsmArgs* args;        // usm_prepare's arguments (fast path, NaN-free)
};
struct WaveBatch {
  int inst = -1;              // kernel instance (number of bilinear sources), -1: none
  int n_tiles = 0, n_blocks = 0;
  void* d_args = nullptr;     // the windows' wavefront-kernel arguments
  int* d_pre = nullptr;       // block prefix (n_tiles + 1), each count a multiple of 8
};
// Builds the coefficient tables (once: they depend on the geometry only,
// synchronously on the null stream) and the device descriptors; taken[i] = 1
// for the windows it takes.  wb->n_tiles = 0 when it takes none.
int wave_batch_create(const std::vector<WaveTileIn>& in, std::vector<char>& taken, WaveBatch* wb);
int wave_batch_launch(const WaveBatch& wb, hipStream_t s);
void wave_batch_destroy(WaveBatch* wb);

int launch_dec_batch(const DecJob* jobs, const int* pre, int n_jobs, int n_blocks,
                     hipStream_t s);
int launch_box_batch(int pass, const BoxJob* jobs, const int* pre, int n_jobs, int n_blocks,
                     hipStream_t s);

}  // namespace fsg
