// Sparse frames: one 2D level of an inner Wi x Hi image inside an outer
// Wo x Ho frame (the reference's DWT_ARR_SPARSE layout) in one launch.
//
// The reference transforms every outer row y < Ho over the inner length Wi,
// writing L to [0, Wl) and H from the outer half Wod on, and then every outer
// column x < Wo over the inner length Hi likewise (src/libdwt.c:6331-6400;
// inverse 6683-6747), garbage rows and columns included.  A level is split
// into parts that each read only the level's input and write only the
// level's output buffer Q (a separate buffer), so they run side by side as
// workgroups of one launch:
//   forward (Wl = ceil(Wi/2), Wh = floor(Wi/2), likewise Hl, Hh; Wod, Hod the
//   outer halves)
//     A  inner tiles: rows [0, Hi) with their columns, the packed level
//        kernel's tile body, the four subbands written at the outer frame's
//        Mallat positions (SparseDet);
//     B  garbage columns [Wl, Wod) and [Wod + Wh, Wo): their column transform
//        (the column-signal kernel's body), L rows [0, Hl), H from Hod;
//     C  rows [Hl, Hod) and [Hod + Hh, Ho), which the column pass leaves: their
//        row transform alone (the row-signal kernel's body);
//     D  garbage columns x those rows: the input, unchanged;
//   inverse (output level: outer Wo x Ho, inner Wi x Hi)
//     A  inner tiles: output [0, Hi) x [0, Wi), LL from the previous level's
//        output, details from the frame;
//     B  columns [Wi, Wo): their column inverse;
//     C  rows [Hi, Ho): their row inverse;
//     D  [Hi, Ho) x [Wi, Wo): the input, unchanged.
// With zero padding, B, C and D write zeros (the reference zeroes exactly
// those positions after each level).  The decomposition is replayed against
// the oracle with its 1D lines in tests/test_sparse_model.py.
//
// Input of the inverse level: the previous (coarser) level's output covers
// [0, Hod) x [0, Wod); everything else comes from the frame.  The parts read
// per column or per row from one of the two (selected on the host), and D per
// element.
#include "line_body.h"

namespace dwt {

// Detail subbands of a sparse level: the tile bodies address them at the
// packed positions (HL at column Wl + k, LH at row Hl + m); the outer frame
// puts them at Wod + k and Hod + m.
template <typename T>
struct SparseDet {
  T* __restrict__ p;
  long long pitch;
  int dr, dc;  // Hod - Hl, Wod - Wl
  __device__ __forceinline__ long long at(int band, int r, int c) const {
    return (long long)(r + (band != kHL ? dr : 0)) * pitch + c + (band != kLH ? dc : 0);
  }
  __device__ __forceinline__ void store(int band, int r, int c, T v) const { st_global(p + at(band, r, c), v); }
  __device__ __forceinline__ T load(int band, int r, int c) const {
    return gload_ptr(static_cast<const T*>(p))[at(band, r, c)];
  }
};

// Parts B (column ranges) and C (row ranges) of one launch, and the
// rectangles of part D, with their workgroup counts.
template <typename T>
struct SparseLine {
  const T* lsrc;  // B forward: input at the range's first column; C forward: input row
  long long lp;   // inverse: source of the L half (previous level or frame)
  const T* hsrc;  // inverse: source of the H half (frame)
  long long hp;
  T* dst;         // output at the range's start
  long long dp;
  int count, blocks;  // columns / rows of the range; workgroups
};
struct SparseRect {
  int y0, h, x0, w, blocks;
};
constexpr int kSparseRects = 6;

template <typename T>
struct SparseArgs {
  const T* in;   // forward: level input; inverse: the frame (details)
  long long ip;
  const T* prev;  // inverse: previous level's output ([0, Hod) x [0, Wod))
  long long pp;
  T* q;  // level output
  long long qp;
  int Wo, Ho, Wi, Hi;
  int nstrips, ngroups, nseg, a_blocks;  // part A tiles (8-row segments)
  SparseLine<T> col[2], row[2];          // parts B, C
  int col_hbase, row_hoff, nstrips_row;
  SparseRect rect[kSparseRects];  // part D
  int nrect, rect_zero;
  int sel_y, sel_x;  // part D reads prev where y < sel_y and x < sel_x, else in
};

template <typename T>
__device__ __forceinline__ void sparse_rect(const SparseArgs<T>& a, int b) {
  for (int i = 0; i < a.nrect; ++i) {
    const SparseRect& r = a.rect[i];
    if (b >= r.blocks) {
      b -= r.blocks;
      continue;
    }
    for (int y = r.y0 + b; y < r.y0 + r.h; y += r.blocks) {
      const bool top = y < a.sel_y;
      const Gbl<T>* s0 = gload_ptr(a.prev) + (long long)y * a.pp;
      const Gbl<T>* s1 = gload_ptr(a.in) + (long long)y * a.ip;
      Gbl<T>* d = gstore_ptr(a.q) + (long long)y * a.qp;
      for (int x = r.x0 + (int)threadIdx.x; x < r.x0 + r.w; x += (int)blockDim.x)
        d[x] = a.rect_zero ? T(0) : (top && x < a.sel_x ? s0[x] : s1[x]);
    }
    return;
  }
}

// Workgroup b of the parts after A, in order: B ranges, C ranges, D
// rectangles, then riders.
template <bool FWD, typename T, int NS>
__device__ __forceinline__ void sparse_lines(const SparseArgs<T>& a, int b, const RiderCopy<T>& rc,
                                             const Coeffs<T>& cf) {
  for (int i = 0; i < 2; ++i) {
    const SparseLine<T>& l = a.col[i];
    if (b < l.blocks) {
      if (FWD)
        col_fwd_body<T, NS, 8, true>(b, l.lsrc, l.lp, l.dst, l.dp, l.dst, l.dp, a.Hi, l.count, 0, a.col_hbase, cf);
      else
        col_inv_body<T, NS, 8, true>(b, l.lsrc, l.lp, l.hsrc, l.hp, l.dst, l.dp, a.Hi, l.count, 0, a.col_hbase,
                                     cf);
      return;
    }
    b -= l.blocks;
  }
  for (int i = 0; i < 2; ++i) {
    const SparseLine<T>& l = a.row[i];
    if (b < l.blocks) {
      const int w = __builtin_amdgcn_readfirstlane(b * kWavesPerBlock + (int)(threadIdx.x >> 6));
      if (FWD)
        row_fwd_body<T, NS, true>(w, l.lsrc, l.lp, l.dst, l.dp, l.dst + a.row_hoff, l.dp, a.Wi, l.count,
                                  a.nstrips_row, cf);
      else
        row_inv_body<T, NS, true>(w, l.lsrc, l.lp, l.hsrc, l.hp, l.dst, l.dp, a.Wi, l.count, a.nstrips_row, cf);
      return;
    }
    b -= l.blocks;
  }
  int nd = 0;
  for (int i = 0; i < a.nrect; ++i) nd += a.rect[i].blocks;
  if (b < nd) {
    sparse_rect(a, b);
    return;
  }
  b -= nd;
  rider_copy(rc, b, rc.blocks);
}

template <typename T, int NS>
__global__ void __launch_bounds__(256)
sparse_fwd_kernel(SparseArgs<T> a, Coeffs<T> cf, RiderCopy<T> rc) {
  const int b = blockIdx.x;
  if (b >= a.a_blocks) {
    sparse_lines<true, T, NS>(a, b - a.a_blocks, rc, cf);
    return;
  }
  const int t = xcd_remap(b, a.a_blocks);
  const int s = t / a.ngroups, g = t - s * a.ngroups;
  const int strip = g * kWavesPerBlock + (threadIdx.x >> 6);
  if (strip >= a.nstrips) return;
  const int Wl = (a.Wi + 1) >> 1, Hl = (a.Hi + 1) >> 1;
  fwd_tile<T, NS, -8>(strip, s, 8, a.Wi, a.Hi, GlobalRows<T, true>{a.in, a.ip, a.Hi}, GlobalStore<T>{a.q, a.qp},
                      SparseDet<T>{a.q, a.qp, ((a.Ho + 1) >> 1) - Hl, ((a.Wo + 1) >> 1) - Wl}, cf);
}

template <typename T, int NS>
__global__ void __launch_bounds__(256)
sparse_inv_kernel(SparseArgs<T> a, Coeffs<T> cf, RiderCopy<T> rc) {
  const int b = blockIdx.x;
  if (b >= a.a_blocks) {
    sparse_lines<false, T, NS>(a, b - a.a_blocks, rc, cf);
    return;
  }
  const int t = xcd_remap(b, a.a_blocks);
  const int s = t / a.ngroups, g = t - s * a.ngroups;
  const int strip = g * kWavesPerBlock + (threadIdx.x >> 6);
  if (strip >= a.nstrips) return;
  const int Wl = (a.Wi + 1) >> 1, Hl = (a.Hi + 1) >> 1;
  inv_tile<T, NS, -8>(strip, s, 8, a.Wi, a.Hi, GlobalLoad<T>{a.prev, a.pp},
                      SparseDet<T>{const_cast<T*>(a.in), a.ip, ((a.Ho + 1) >> 1) - Hl, ((a.Wo + 1) >> 1) - Wl},
                      GlobalOut<T, true>{a.q, a.qp}, cf);
}

// ---- host ---------------------------------------------------------------------------
namespace {

int col_blocks(int count, int n) {
  if (count <= 0) return 0;
  return ((count + 128 * kWavesPerBlock - 1) / (128 * kWavesPerBlock)) * ((n + 7) / 8);
}
int row_blocks(int count, int nstrips) {
  if (count <= 0) return 0;
  const long long waves = (long long)nstrips * ((count + kRowsPerWave - 1) / kRowsPerWave);
  return (int)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
}
template <typename T>
void add_rect(SparseArgs<T>& a, int y0, int y1, int x0, int x1) {
  if (y1 <= y0 || x1 <= x0) return;
  SparseRect& r = a.rect[a.nrect++];
  r.y0 = y0;
  r.h = y1 - y0;
  r.x0 = x0;
  r.w = x1 - x0;
  r.blocks = std::min(r.h, 256);
}
template <typename T>
int total_blocks(const SparseArgs<T>& a, const RiderCopy<T>& rc) {
  long long n = a.a_blocks;
  for (int i = 0; i < 2; ++i) n += a.col[i].blocks + a.row[i].blocks;
  for (int i = 0; i < a.nrect; ++i) n += a.rect[i].blocks;
  return (int)(n + rc.blocks);
}
template <typename T>
void tiles(SparseArgs<T>& a, int nsteps) {
  const int core = 2 * kWave - 2 * nsteps;
  a.nstrips = (a.Wi + core - 1) / core;
  a.ngroups = (a.nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
  a.nseg = (a.Hi + 7) / 8;
  a.a_blocks = a.ngroups * a.nseg;
  a.nstrips_row = (a.Wi + core - 1) / core;
}

}  // namespace

template <typename T>
void launch_sparse_fwd(const T* in, long long ip, T* q, long long qp, int Wo, int Ho, int Wi, int Hi, int zp,
                       const Coeffs<T>& cf, const RiderCopy<T>* rider, hipStream_t st) {
  SparseArgs<T> a{};
  a.in = in;
  a.ip = ip;
  a.prev = in;
  a.pp = ip;
  a.q = q;
  a.qp = qp;
  a.Wo = Wo;
  a.Ho = Ho;
  a.Wi = Wi;
  a.Hi = Hi;
  tiles(a, cf.nsteps);
  const int Wl = (Wi + 1) / 2, Wh = Wi / 2, Hl = (Hi + 1) / 2, Hh = Hi / 2;
  const int Wod = (Wo + 1) / 2, Hod = (Ho + 1) / 2;
  const int gx[2][2] = {{Wl, Wod}, {Wod + Wh, Wo}};      // garbage columns
  const int ry[2][2] = {{Hl, Hod}, {Hod + Hh, Ho}};      // rows the column pass leaves
  a.col_hbase = Hod;
  a.row_hoff = Wod;
  a.sel_y = a.sel_x = 0;  // part D reads the input
  if (!zp) {
    for (int i = 0; i < 2; ++i) {
      SparseLine<T>& c = a.col[i];
      c.count = gx[i][1] - gx[i][0];
      c.lsrc = in + gx[i][0];
      c.lp = ip;
      c.dst = q + gx[i][0];
      c.dp = qp;
      c.blocks = col_blocks(c.count, Hi);
      SparseLine<T>& r = a.row[i];
      r.count = ry[i][1] - ry[i][0];
      r.lsrc = in + (long long)ry[i][0] * ip;
      r.lp = ip;
      r.dst = q + (long long)ry[i][0] * qp;
      r.dp = qp;
      r.blocks = row_blocks(r.count, a.nstrips_row);
    }
    for (int i = 0; i < 2; ++i)
      for (int k = 0; k < 2; ++k) add_rect(a, ry[i][0], ry[i][1], gx[k][0], gx[k][1]);
  } else {
    // zeros everywhere but the inner subbands
    a.rect_zero = 1;
    for (int i = 0; i < 2; ++i) add_rect(a, ry[i][0], ry[i][1], 0, Wo);
    const int cy[2][2] = {{0, Hl}, {Hod, Hod + Hh}};
    for (int i = 0; i < 2; ++i)
      for (int k = 0; k < 2; ++k) add_rect(a, cy[i][0], cy[i][1], gx[k][0], gx[k][1]);
  }
  RiderCopy<T> rc{};
  if (rider && rider->r1 > rider->r0 && rider->blocks > 0) rc = *rider;
  const dim3 grid(total_blocks(a, rc)), block(256);
  if (cf.nsteps == 4)
    hipLaunchKernelGGL((sparse_fwd_kernel<T, 4>), grid, block, 0, st, a, cf, rc);
  else
    hipLaunchKernelGGL((sparse_fwd_kernel<T, 2>), grid, block, 0, st, a, cf, rc);
}

template <typename T>
void launch_sparse_inv(const T* prev, long long pp, const T* frame, long long fp, T* q, long long qp, int Wo,
                       int Ho, int Wi, int Hi, int zp, const Coeffs<T>& cf, const RiderCopy<T>* rider,
                       hipStream_t st) {
  SparseArgs<T> a{};
  a.in = frame;
  a.ip = fp;
  a.prev = prev;
  a.pp = pp;
  a.q = q;
  a.qp = qp;
  a.Wo = Wo;
  a.Ho = Ho;
  a.Wi = Wi;
  a.Hi = Hi;
  tiles(a, cf.nsteps);
  const int Wod = (Wo + 1) / 2, Hod = (Ho + 1) / 2;
  a.col_hbase = Hod;
  a.row_hoff = 0;
  a.sel_y = Hod;
  a.sel_x = Wod;
  if (!zp) {
    // columns [Wi, Wo) and rows [Hi, Ho), each split where the L half stops
    // coming from the previous level's output
    const int cx[2][2] = {{Wi, std::max(Wi, std::min(Wo, Wod))}, {std::max(Wi, Wod), Wo}};
    const int ry[2][2] = {{Hi, std::max(Hi, std::min(Ho, Hod))}, {std::max(Hi, Hod), Ho}};
    for (int i = 0; i < 2; ++i) {
      SparseLine<T>& c = a.col[i];
      c.count = cx[i][1] - cx[i][0];
      c.lsrc = (i == 0 ? prev : frame) + cx[i][0];
      c.lp = i == 0 ? pp : fp;
      c.hsrc = frame + cx[i][0];
      c.hp = fp;
      c.dst = q + cx[i][0];
      c.dp = qp;
      c.blocks = col_blocks(c.count, Hi);
      SparseLine<T>& r = a.row[i];
      r.count = ry[i][1] - ry[i][0];
      r.lsrc = (i == 0 ? prev + (long long)ry[i][0] * pp : frame + (long long)ry[i][0] * fp);
      r.lp = i == 0 ? pp : fp;
      r.hsrc = frame + (long long)ry[i][0] * fp + Wod;
      r.hp = fp;
      r.dst = q + (long long)ry[i][0] * qp;
      r.dp = qp;
      r.blocks = row_blocks(r.count, a.nstrips_row);
    }
    add_rect(a, Hi, Ho, Wi, Wo);
  } else {
    a.rect_zero = 1;
    add_rect(a, 0, Hi, Wi, Wo);
    add_rect(a, Hi, Ho, 0, Wo);
  }
  RiderCopy<T> rc{};
  if (rider && rider->r1 > rider->r0 && rider->blocks > 0) rc = *rider;
  const dim3 grid(total_blocks(a, rc)), block(256);
  if (cf.nsteps == 4)
    hipLaunchKernelGGL((sparse_inv_kernel<T, 4>), grid, block, 0, st, a, cf, rc);
  else
    hipLaunchKernelGGL((sparse_inv_kernel<T, 2>), grid, block, 0, st, a, cf, rc);
}

template void launch_sparse_fwd<float>(const float*, long long, float*, long long, int, int, int, int, int,
                                       const Coeffs<float>&, const RiderCopy<float>*, hipStream_t);
template void launch_sparse_fwd<double>(const double*, long long, double*, long long, int, int, int, int, int,
                                        const Coeffs<double>&, const RiderCopy<double>*, hipStream_t);
template void launch_sparse_inv<float>(const float*, long long, const float*, long long, float*, long long, int,
                                       int, int, int, int, const Coeffs<float>&, const RiderCopy<float>*,
                                       hipStream_t);
template void launch_sparse_inv<double>(const double*, long long, const double*, long long, double*, long long,
                                        int, int, int, int, int, const Coeffs<double>&, const RiderCopy<double>*,
                                        hipStream_t);

}  // namespace dwt
