// Full-depth 1D transforms of many strided signals in one call: the
// decomposition a caller gets from dwt_cdf97_f_ex_stride_s applied level by
// level, each level writing L to [0, ceil(n/2)) and H to [ceil(n/2), n) of
// the signal, until n < 2 (the inverse undoes it level by level).
//
// Two layouts have their own kernels, one launch per level over all signals:
//   * contiguous samples (each signal a row): the 2D tile body's row pass; a
//     wavefront owns a 128-sample window of four rows, lifting in registers
//     with the neighbour sample by DPP;
//   * adjacent signals (the signals are the columns of a matrix, sample
//     stride = row pitch): a column-only version of the 2D tile body's
//     vertical sliding window; a wavefront owns 128 adjacent signals (two per
//     lane, pair-wise loads) and a segment of rows.
// The arithmetic is the reference's line lifting (src/libdwt.c:5512-5603,
// 5779-5870): updates x + c * (l + r) as separate operations, (2c) * n at a
// line end, scaling after (forward) or before (inverse) the lifting.
#include <algorithm>
#include <vector>

#include "line_body.h"

namespace dwt {

// ---- row kernels (contiguous signals, one level per launch; line_body.h) -----------
template <typename T, int NS, bool VEC>
__global__ void __launch_bounds__(256)
row_fwd_kernel(const T* __restrict__ src, long long sp, T* __restrict__ ll, long long lp,
               T* __restrict__ hh, long long hp, int n, int count, int nstrips, Coeffs<T> cf) {
  const int w = __builtin_amdgcn_readfirstlane((int)blockIdx.x * kWavesPerBlock + (int)(threadIdx.x >> 6));
  row_fwd_body<T, NS, VEC>(w, src, sp, ll, lp, hh, hp, n, count, nstrips, cf);
}

template <typename T, int NS, bool VEC>
__global__ void __launch_bounds__(256)
row_inv_kernel(const T* __restrict__ ll, long long lp, const T* __restrict__ hh, long long hp,
               T* __restrict__ dst, long long dp, int n, int count, int nstrips, Coeffs<T> cf) {
  const int w = __builtin_amdgcn_readfirstlane((int)blockIdx.x * kWavesPerBlock + (int)(threadIdx.x >> 6));
  row_inv_body<T, NS, VEC>(w, ll, lp, hh, hp, dst, dp, n, count, nstrips, cf);
}

// ---- resident row kernels (level 0 in place) -----------------------------------------
// A workgroup owns R whole rows: it reads them into LDS, waits, and only then
// writes the level's output, so the output may overlap the rows it read (the
// H half of level 0 goes straight to its final place, no scratch copy).  The
// windows (R rows x nstrips) are shared among the workgroup's wavefronts and
// computed exactly as in row_fwd_kernel / row_inv_kernel.  Row u sits at
// lds[u * ln], ln = n rounded up to a multiple of four.
constexpr int kResBytes = 32768;     // LDS per workgroup (rows packed up to this)
constexpr int kResMaxBytes = 65536;  // longest row held whole

template <typename T, int VW> struct VecN { typedef T type __attribute__((ext_vector_type(VW))); };
template <typename T> struct VecN<T, 1> { typedef T type; };

// Read two segments of each of `rows` rows into LDS: row u's m1 elements at
// g1 + u * p1 to l + u * ln, its m2 elements at g2 + u * p2 to l + u * ln + m1.
// VW elements per access (pointers, pitches, m1, m2 and ln multiples of VW).
template <typename T, int VW>
__device__ __forceinline__ void res_fill(T* l, int ln, int rows, const T* g1, long long p1, int m1,
                                         const T* g2, long long p2, int m2) {
  // kFill loads are issued before their LDS stores, so that many of a wave's
  // loads are in flight at once
  constexpr int kFill = 8;
  constexpr int kThreads = kWave * kWavesPerBlock;  // the kernels' block size
  using V = typename VecN<T, VW>::type;
  const int v1 = m1 / VW, v = v1 + m2 / VW, total = rows * v;
  for (int base = threadIdx.x; base < total; base += kFill * kThreads) {
    V tmp[kFill];
    int at[kFill];
#pragma unroll
    for (int f = 0; f < kFill; ++f) {
      const int idx = base + f * kThreads;
      at[f] = -1;
      if (idx < total) {
        const int u = idx / v, i = idx - u * v;
        const T* g = i < v1 ? g1 + u * p1 + (long long)i * VW : g2 + u * p2 + (long long)(i - v1) * VW;
        tmp[f] = *reinterpret_cast<const Gbl<V>*>(gload_ptr(g));
        at[f] = u * ln + i * VW;
      }
    }
#pragma unroll
    for (int f = 0; f < kFill; ++f)
      if (at[f] >= 0) *reinterpret_cast<V*>(l + at[f]) = tmp[f];
  }
}

template <typename T, int NS, int VW>
__global__ void __launch_bounds__(256)
row_res_fwd_kernel(const T* src, long long sp, T* ll, long long lp, T* hh, long long hp, int n,
                   int count, int R, int nstrips, Coeffs<T> cf) {
  extern __shared__ __align__(16) unsigned char res_smem[];
  T* lds = reinterpret_cast<T*>(res_smem);
  constexpr int CORE = 2 * kWave - 2 * NS;
  const int ln = (n + 3) & ~3;
  const int r0 = blockIdx.x * R;
  const int rows = min(R, count - r0);
  res_fill<T, VW>(lds, ln, rows, src + (long long)r0 * sp, sp, n, src, 0, 0);
  __syncthreads();
  // the wave index is uniform: window bookkeeping and row addresses stay scalar
  const int lane = threadIdx.x & 63, nw = kWavesPerBlock;  // launched with 256 threads
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int half = (n + 1) >> 1, nh = n >> 1;
  const bool core = lane >= NS / 2 && lane < kWave - NS / 2;
  for (int it = wave; it < rows * nstrips; it += nw) {
    const int u = it / nstrips, strip = it - u * nstrips;
    const int Xe = strip * CORE - NS + 2 * lane;
    const bool interior = strip * CORE - NS >= 0 && strip * CORE - NS + 2 * kWave <= n;
    const T* row = lds + u * ln;
    T e, o;
    if (interior) {
      using V2 = typename Vec2<T>::type;
      const V2 v = *reinterpret_cast<const V2*>(row + Xe);
      e = v.x;
      o = v.y;
    } else {
      e = row[mirror_near(Xe, n)];
      o = row[mirror_near(Xe + 1, n)];
    }
    row_fwd<T, NS>(e, o, cf);
    const int k = Xe >> 1;
    T* lrow = ll + (long long)(r0 + u) * lp;
    T* hrow = hh + (long long)(r0 + u) * hp;
    if (core) {
      if (k < half) st_global(lrow + k, e);
      if (k < nh) st_global(hrow + k, o);
    }
  }
}

// Inverse level on n output samples: L (ceil(n/2)) from ll, H (n/2) from hh;
// output to dst.  In LDS, H follows L at offset ceil(n/2).
template <typename T, int NS, int VW, bool VEC>
__global__ void __launch_bounds__(256)
row_res_inv_kernel(const T* ll, long long lp, const T* hh, long long hp, T* dst, long long dp, int n,
                   int count, int R, int nstrips, Coeffs<T> cf) {
  extern __shared__ __align__(16) unsigned char res_smem[];
  T* lds = reinterpret_cast<T*>(res_smem);
  constexpr int CORE = 2 * kWave - 2 * NS;
  const int ln = (n + 3) & ~3;
  const int half = (n + 1) >> 1, nh = n >> 1;
  const int r0 = blockIdx.x * R;
  const int rows = min(R, count - r0);
  res_fill<T, VW>(lds, ln, rows, ll + (long long)r0 * lp, lp, half, hh + (long long)r0 * hp, hp, nh);
  __syncthreads();
  // the wave index is uniform: window bookkeeping and row addresses stay scalar
  const int lane = threadIdx.x & 63, nw = kWavesPerBlock;  // launched with 256 threads
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const bool core = lane >= NS / 2 && lane < kWave - NS / 2;
  for (int it = wave; it < rows * nstrips; it += nw) {
    const int u = it / nstrips, strip = it - u * nstrips;
    const int Xe = strip * CORE - NS + 2 * lane;
    const int ke = mirror_near(Xe, n) >> 1, ko = mirror_near(Xe + 1, n) >> 1;
    const T* row = lds + u * ln;
    T e = row[ke], o = row[half + ko];
    row_inv<T, NS>(e, o, cf);
    if (core) {
      T* orow = dst + (long long)(r0 + u) * dp;
      T* out = orow + Xe;
      if (VEC && Xe + 1 < n) {
        using V2 = typename Vec2<T>::type;
        V2 v;
        v.x = e;
        v.y = o;
        *reinterpret_cast<Gbl<V2>*>(gstore_ptr(out)) = v;
      } else {
        if (Xe < n) st_global(out, e);
        if (Xe + 1 < n) st_global(out + 1, o);
      }
    }
  }
}

// ---- column kernels (adjacent signals; line_body.h) -----------------------------------
// plain dispatch order (consecutive workgroups on different XCDs): measured
// faster here than keeping neighbouring workgroups on one XCD
template <typename T, int NS, int SS, bool VEC>
__global__ void __launch_bounds__(256)
col_fwd_kernel(const T* __restrict__ src, long long sp, T* __restrict__ ll, long long llp,
               T* __restrict__ det, long long dp, int n, int count, int nseg, int seg0, int hbase,
               Coeffs<T> cf) {
  col_fwd_body<T, NS, SS, VEC>(blockIdx.x, src, sp, ll, llp, det, dp, n, count, seg0, hbase, cf);
}

template <typename T, int NS, int SS, bool VEC>
__global__ void __launch_bounds__(256)
col_inv_kernel(const T* __restrict__ ll, long long llp, const T* __restrict__ det, long long dp,
               T* __restrict__ dst, long long dstp, int n, int count, int nseg, int seg0, int hbase,
               Coeffs<T> cf) {
  col_inv_body<T, NS, SS, VEC>(blockIdx.x, ll, llp, det, dp, dst, dstp, n, count, seg0, hbase, cf);
}

// ---- host launchers ------------------------------------------------------------------
#ifndef DWT_COL_ROWS
#define DWT_COL_ROWS 8
#endif
constexpr int kColRows = DWT_COL_ROWS;  // rows per segment of the column kernels

template <typename T, int NS, bool VEC>
static void col_go(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c, long long cp,
                   int n, int count, int seg0, int nseg, int hbase, const Coeffs<T>& cf,
                   hipStream_t st) {
  if (nseg <= 0) return;
  const int ngroups = (count + 128 * kWavesPerBlock - 1) / (128 * kWavesPerBlock);
  const dim3 grid(ngroups * nseg), block(256);
  if (fwd)  // a = src, b = ll (written), c = det
    hipLaunchKernelGGL((col_fwd_kernel<T, NS, kColRows, VEC>), grid, block, 0, st, a, ap,
                       const_cast<T*>(b), bp, c, cp, n, count, nseg, seg0, hbase, cf);
  else  // a = ll, b = det, c = dst
    hipLaunchKernelGGL((col_inv_kernel<T, NS, kColRows, VEC>), grid, block, 0, st, a, ap, b, bp, c,
                       cp, n, count, nseg, seg0, hbase, cf);
}

template <typename T, int NS, bool VEC>
static void row_go(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c, long long cp,
                   int n, int count, const Coeffs<T>& cf, hipStream_t st) {
  constexpr int CORE = 2 * kWave - 2 * NS;
  const int nstrips = (n + CORE - 1) / CORE;
  const long long waves = (long long)nstrips * ((count + kRowsPerWave - 1) / kRowsPerWave);
  const dim3 grid((unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock)), block(kWave * kWavesPerBlock);
  if (fwd)  // a = src, b = L out, c = H out
    hipLaunchKernelGGL((row_fwd_kernel<T, NS, VEC>), grid, block, 0, st, a, ap, const_cast<T*>(b), bp,
                       c, cp, n, count, nstrips, cf);
  else  // a = L in, b = H in, c = output
    hipLaunchKernelGGL((row_inv_kernel<T, NS, VEC>), grid, block, 0, st, a, ap, b, bp, c, cp, n, count,
                       nstrips, cf);
}

template <typename T, int NS, int VW>
static void res_go(bool fwd, bool vec, const T* a, long long ap, const T* b, long long bp, T* c,
                   long long cp, int n, int count, const Coeffs<T>& cf, hipStream_t st) {
  constexpr int CORE = 2 * kWave - 2 * NS;
  const int nstrips = (n + CORE - 1) / CORE;
  const long long row_bytes = (long long)((n + 3) & ~3) * (long long)sizeof(T);
  const int R = (int)std::max(1LL, std::min((long long)count, kResBytes / row_bytes));
  const dim3 grid((unsigned)((count + R - 1) / R)), block(kWave * kWavesPerBlock);
  const size_t lds = (size_t)(R * row_bytes);
  if (fwd)  // a = src, b = L out, c = H out
    hipLaunchKernelGGL((row_res_fwd_kernel<T, NS, VW>), grid, block, lds, st, a, ap, const_cast<T*>(b),
                       bp, c, cp, n, count, R, nstrips, cf);
  else if (vec)  // a = L in, b = H in, c = output
    hipLaunchKernelGGL((row_res_inv_kernel<T, NS, VW, true>), grid, block, lds, st, a, ap, b, bp, c, cp,
                       n, count, R, nstrips, cf);
  else
    hipLaunchKernelGGL((row_res_inv_kernel<T, NS, VW, false>), grid, block, lds, st, a, ap, b, bp, c, cp,
                       n, count, R, nstrips, cf);
}

template <typename T, int NS>
static void res_width(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c, long long cp,
                      int n, int count, const Coeffs<T>& cf, hipStream_t st) {
  const long long es = sizeof(T);
  auto fits = [&](int vw) {
    auto ok = [&](const void* p, long long pitch, long long m) {
      return reinterpret_cast<uintptr_t>(p) % (uintptr_t)(vw * es) == 0 && pitch % vw == 0 && m % vw == 0;
    };
    return fwd ? ok(a, ap, n) : ok(a, ap, (n + 1) / 2) && ok(b, bp, n / 2);
  };
  const bool vec = !fwd && reinterpret_cast<uintptr_t>(c) % (uintptr_t)(2 * es) == 0 && cp % 2 == 0;
  if (es == 4 && fits(4)) res_go<T, NS, 4>(fwd, vec, a, ap, b, bp, c, cp, n, count, cf, st);
  else if (fits(2)) res_go<T, NS, 2>(fwd, vec, a, ap, b, bp, c, cp, n, count, cf, st);
  else res_go<T, NS, 1>(fwd, vec, a, ap, b, bp, c, cp, n, count, cf, st);
}

// Rows shorter than 64 samples: one window each, mostly idle lanes; the
// windowed kernel through scratch measured faster there (16 and 32 samples).
bool row_resident_fits(int n, int elem) {
  return n >= 64 && (long long)((n + 3) & ~3) * elem <= kResMaxBytes;
}

template <typename T>
bool launch_row_resident(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c,
                         long long cp, int n, int count, const Coeffs<T>& cf, hipStream_t st) {
  if (!row_resident_fits(n, (int)sizeof(T))) return false;
  if (count > 0) {
    if (cf.nsteps == 4) res_width<T, 4>(fwd, a, ap, b, bp, c, cp, n, count, cf, st);
    else res_width<T, 2>(fwd, a, ap, b, bp, c, cp, n, count, cf, st);
  }
  return true;
}

template bool launch_row_resident<float>(bool, const float*, long long, const float*, long long, float*,
                                         long long, int, int, const Coeffs<float>&, hipStream_t);
template bool launch_row_resident<double>(bool, const double*, long long, const double*, long long,
                                          double*, long long, int, int, const Coeffs<double>&,
                                          hipStream_t);

template <typename T>
void launch_row_level(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c,
                      long long cp, int n, int count, const Coeffs<T>& cf, bool vec,
                      hipStream_t st) {
  if (cf.nsteps == 4) {
    if (vec) row_go<T, 4, true>(fwd, a, ap, b, bp, c, cp, n, count, cf, st);
    else row_go<T, 4, false>(fwd, a, ap, b, bp, c, cp, n, count, cf, st);
  } else {
    if (vec) row_go<T, 2, true>(fwd, a, ap, b, bp, c, cp, n, count, cf, st);
    else row_go<T, 2, false>(fwd, a, ap, b, bp, c, cp, n, count, cf, st);
  }
}

template void launch_row_level<float>(bool, const float*, long long, const float*, long long, float*,
                                      long long, int, int, const Coeffs<float>&, bool, hipStream_t);
template void launch_row_level<double>(bool, const double*, long long, const double*, long long,
                                       double*, long long, int, int, const Coeffs<double>&, bool,
                                       hipStream_t);

template <typename T>
void launch_col_segments(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c,
                         long long cp, int n, int count, int seg0, int nseg, int hbase,
                         const Coeffs<T>& cf, bool vec, hipStream_t st) {
  if (cf.nsteps == 4) {
    if (vec) col_go<T, 4, true>(fwd, a, ap, b, bp, c, cp, n, count, seg0, nseg, hbase, cf, st);
    else col_go<T, 4, false>(fwd, a, ap, b, bp, c, cp, n, count, seg0, nseg, hbase, cf, st);
  } else {
    if (vec) col_go<T, 2, true>(fwd, a, ap, b, bp, c, cp, n, count, seg0, nseg, hbase, cf, st);
    else col_go<T, 2, false>(fwd, a, ap, b, bp, c, cp, n, count, seg0, nseg, hbase, cf, st);
  }
}

template <typename T>
void launch_col_level(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c,
                      long long cp, int n, int count, const Coeffs<T>& cf, bool vec,
                      hipStream_t st) {
  launch_col_segments<T>(fwd, a, ap, b, bp, c, cp, n, count, 0, col_segments(n), (n + 1) / 2, cf, vec,
                         st);
}

int col_segments(int n) { return (n + kColRows - 1) / kColRows; }
int col_segment_rows() { return kColRows; }

// Rows one segment reads and writes (the column kernels' index arithmetic).
// Forward: reads input rows, writes H rows m (at hbase + m).  Inverse: reads
// H rows m (at hbase + m; its L rows come from scratch), writes output rows.
// Each list sorted, without duplicates; returns the counts.
void col_segment_access(bool fwd, int n, int nsteps, int seg, std::vector<int>& reads,
                        std::vector<int>& writes) {
  const int SS = kColRows, NS = nsteps, NP = SS / 2 + NS;
  reads.clear();
  writes.clear();
  for (int u = 0; u < 2 * NP; ++u) {
    const int r = col_load_row(seg, u, n, SS, NS);
    if (fwd)
      reads.push_back(r);
    else if (r & 1)
      reads.push_back(r >> 1);
  }
  if (fwd) {
    const int half = (n + 1) >> 1, nh = n >> 1;
    const int m0 = (seg * SS) >> 1, m1 = std::min(m0 + SS / 2, half);
    for (int m = m0; m < m1; ++m)
      if (m < nh) writes.push_back(m);
  } else {
    for (int r = seg * SS; r < std::min(seg * SS + SS, n); ++r) writes.push_back(r);
  }
  std::sort(reads.begin(), reads.end());
  reads.erase(std::unique(reads.begin(), reads.end()), reads.end());
}

template void launch_col_segments<float>(bool, const float*, long long, const float*, long long, float*,
                                         long long, int, int, int, int, int, const Coeffs<float>&, bool,
                                         hipStream_t);
template void launch_col_segments<double>(bool, const double*, long long, const double*, long long,
                                          double*, long long, int, int, int, int, int,
                                          const Coeffs<double>&, bool, hipStream_t);

template void launch_col_level<float>(bool, const float*, long long, const float*, long long, float*,
                                      long long, int, int, const Coeffs<float>&, bool, hipStream_t);
template void launch_col_level<double>(bool, const double*, long long, const double*, long long,
                                       double*, long long, int, int, const Coeffs<double>&, bool,
                                       hipStream_t);

}  // namespace dwt
