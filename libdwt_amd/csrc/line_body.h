// Device bodies of the 1D line kernels (one level over many signals): rows
// of contiguous samples (a wavefront's 128-sample window, lifting in
// registers) and adjacent signals as the columns of a matrix (the vertical
// sliding window).  Used by the 1D batch kernels (lines.hip), the generic 2D
// path's column passes and the sparse-frame level kernels (sparse.hip).
#pragma once

#include "tile_body.h"
#include "kernels.h"

namespace dwt {

// ---- row kernels (contiguous signals, one level per launch) ---------------------------
// A signal is a row of a count x n matrix.  A wavefront owns a 128-sample
// window of kRowsPerWave rows: lane l holds samples (2l, 2l+1) of the window
// and the lifting runs in registers with the neighbour sample from the
// adjacent lane by DPP (the 2D tile body's row pass, row_fwd / row_inv); the
// core 128 - 2 NS samples of the window are stored.  Forward level on n
// samples: L to ll (pitch lp) [0, ceil(n/2)), H to hh (pitch hp) [0, n/2).
constexpr int kRowsPerWave = 4;

// Wavefront w of a launch over count rows (windows numbered strip-fastest).
template <typename T, int NS, bool VEC>
__device__ __forceinline__ void row_fwd_body(int w, const T* __restrict__ src, long long sp, T* __restrict__ ll,
                                             long long lp, T* __restrict__ hh, long long hp, int n, int count,
                                             int nstrips, const Coeffs<T>& cf) {
  constexpr int CORE = 2 * kWave - 2 * NS;
  const int strip = w % nstrips, r0 = (w / nstrips) * kRowsPerWave;
  if (r0 >= count) return;
  const int lane = threadIdx.x & 63;
  const int Xe = strip * CORE - NS + 2 * lane;
  const bool interior = __builtin_amdgcn_readfirstlane(strip * CORE - NS >= 0 && strip * CORE - NS + 2 * kWave <= n);
  const int xe = interior ? Xe : mirror(Xe, n), xo = interior ? Xe + 1 : mirror(Xe + 1, n);
  const int half = (n + 1) >> 1, nh = n >> 1, k = Xe >> 1;
  const bool core = lane >= NS / 2 && lane < kWave - NS / 2;
  T e[kRowsPerWave], o[kRowsPerWave];
  const Gbl<T>* s = gload_ptr(src);
#pragma unroll
  for (int u = 0; u < kRowsPerWave; ++u) {
    const int r = min(r0 + u, count - 1);  // rows past the end repeat the last row (not stored)
    const Gbl<T>* row = s + (long long)r * sp;
    if (VEC && interior) {
      using V2 = typename GPair<T>::type;
      const V2 v = *(const Gbl<V2>*)(row + Xe);
      e[u] = v.x;
      o[u] = v.y;
    } else {
      e[u] = row[xe];
      o[u] = row[xo];
    }
  }
#pragma unroll
  for (int u = 0; u < kRowsPerWave; ++u) {
    row_fwd<T, NS>(e[u], o[u], cf);
    const int r = r0 + u;
    if (core && r < count) {
      if (k < half) st_global(ll + (long long)r * lp + k, e[u]);
      if (k < nh) st_global(hh + (long long)r * hp + k, o[u]);
    }
  }
}

// Inverse level on n output samples: L from ll (pitch lp), H from hh (pitch
// hp); output to dst (pitch dp).
template <typename T, int NS, bool VEC>
__device__ __forceinline__ void row_inv_body(int w, const T* __restrict__ ll, long long lp, const T* __restrict__ hh,
                                             long long hp, T* __restrict__ dst, long long dp, int n, int count,
                                             int nstrips, const Coeffs<T>& cf) {
  constexpr int CORE = 2 * kWave - 2 * NS;
  const int strip = w % nstrips, r0 = (w / nstrips) * kRowsPerWave;
  if (r0 >= count) return;
  const int lane = threadIdx.x & 63;
  const int Xe = strip * CORE - NS + 2 * lane;
  const int ke = mirror(Xe, n) >> 1, ko = mirror(Xe + 1, n) >> 1;  // parity is kept by mirroring
  const bool core = lane >= NS / 2 && lane < kWave - NS / 2;
  T e[kRowsPerWave], o[kRowsPerWave];
  const Gbl<T>* lg = gload_ptr(ll);
  const Gbl<T>* hg = gload_ptr(hh);
#pragma unroll
  for (int u = 0; u < kRowsPerWave; ++u) {
    const int r = min(r0 + u, count - 1);
    e[u] = lg[(long long)r * lp + ke];
    o[u] = hg[(long long)r * hp + ko];
  }
#pragma unroll
  for (int u = 0; u < kRowsPerWave; ++u) {
    row_inv<T, NS>(e[u], o[u], cf);
    const int r = r0 + u;
    if (core && r < count) {
      T* out = dst + (long long)r * dp + Xe;
      if (VEC && Xe + 1 < n) {
        using V2 = typename GPair<T>::type;
        V2 v;
        v.x = e[u];
        v.y = o[u];
        *reinterpret_cast<Gbl<V2>*>(gstore_ptr(out)) = v;
      } else {
        if (Xe < n) st_global(out, e[u]);
        if (Xe + 1 < n) st_global(out + 1, o[u]);
      }
    }
  }
}

// ---- column kernels (adjacent signals) ---------------------------------------------
// Stores of a lane's two adjacent columns: one pair-wise store when the row
// allows it (vec: an element-aligned base pointer; uniform).
template <typename T>
__device__ __forceinline__ bool pair_ok(const T* p, long long) {
  // a pair moves as one element-aligned access (GPair): any pitch will do
  return __builtin_amdgcn_readfirstlane((int)((reinterpret_cast<uintptr_t>(p) % sizeof(T)) == 0));
}
template <typename T>
__device__ __forceinline__ void st_pair(T* p, T a, T b, bool va, bool vb, bool vec) {
  if (vec && vb) {
    using V2 = typename GPair<T>::type;
    V2 v;
    v.x = a;
    v.y = b;
    *reinterpret_cast<Gbl<V2>*>(gstore_ptr(p)) = v;
  } else {
    if (va) st_global(p, a);
    if (vb) st_global(p + 1, b);
  }
}

// The u-th row (u < 2 (SS / 2 + NS)) a column segment `seg` of a level on n
// rows loads: forward, an input row; inverse, an interleaved row (odd: H row
// r >> 1, even: L row r >> 1).  Mirrored at the ends.  Shared with the host
// (dwt_gpu_col_segment_rows), so the staged in-place level 0 is checked
// against exactly these rows.
__host__ __device__ inline int col_load_row(int seg, int u, int n, int SS, int NS) {
  return mirror_near(2 * (((seg * SS) >> 1) - NS / 2) + u, n);
}

// A signal is a column: sample i of signal x at row i.  Forward level on n
// rows: src (row pitch sp) -> L rows to ll (pitch llp), H row m to det row
// hbase + m (pitch dp; hbase = ceil(n/2) puts H in place).  count signals;
// wave = 128 columns.  A launch computes segments [seg0, seg0 + nseg) of SS
// rows (SS / 2 output pairs each); segment s reads input rows
// [SS s - NS, SS (s + 1) + NS) (mirrored at the ends).
// Workgroup b of a launch over count columns.
template <typename T, int NS, int SS, bool VEC>
__device__ __forceinline__ void col_fwd_body(int b, const T* __restrict__ src, long long sp, T* __restrict__ ll,
                                             long long llp, T* __restrict__ det, long long dp, int n, int count,
                                             int seg0, int hbase, const Coeffs<T>& cf) {
  constexpr int LAG = NS / 2, NP = SS / 2 + NS;
  // row-major: consecutive workgroups take adjacent column groups of one row
  // segment, so the workgroups in flight read whole rows
  const int ngr = (count + 128 * kWavesPerBlock - 1) / (128 * kWavesPerBlock);
  const int sl = b / ngr, g = b - sl * ngr, seg = seg0 + sl;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int x = (g * kWavesPerBlock + wave) * 128 + 2 * lane;
  if (__builtin_amdgcn_readfirstlane((g * kWavesPerBlock + wave) * 128) >= count) return;
  const bool va = x < count, vb = x + 1 < count;
  const bool vl = pair_ok(ll, llp), vd = pair_ok(det, dp);
  const int half = (n + 1) >> 1, nh = n >> 1;
  const int m0 = (seg * SS) >> 1, m1 = min(m0 + SS / 2, half);
  const int pbeg = m0 - LAG;
  const Gbl<T>* s = gload_ptr(src);
  T ra[2 * NP], rb[2 * NP];
#pragma unroll
  for (int u = 0; u < 2 * NP; ++u) {
    const Gbl<T>* row = s + (long long)col_load_row(seg, u, n, SS, NS) * sp + x;
    if (VEC && vb) {
      using V2 = typename GPair<T>::type;
      const V2 v = *(const Gbl<V2>*)row;
      ra[u] = v.x;
      rb[u] = v.y;
    } else {
      ra[u] = va ? row[0] : T(0);
      rb[u] = vb ? row[1] : T(0);
    }
  }
  ColFwd<T, NS> ca, cb;
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    T La, Ha, Lb, Hb;
    ca.push(ra[2 * u], ra[2 * u + 1], cf, La, Ha);
    cb.push(rb[2 * u], rb[2 * u + 1], cf, Lb, Hb);
    const int m = pbeg + u - LAG;
    if (m >= m0 && m < m1) {
      st_pair(ll + (long long)m * llp + x, La, Lb, va, vb, vl);
      if (m < nh) {
        st_pair(det + (long long)(hbase + m) * dp + x, Ha, Hb, va, vb, vd);
      }
    }
  }
}

// Inverse level on n output rows: L rows from ll (pitch llp), H row m from
// det row hbase + m (pitch dp); output rows to dst (pitch dstp).  Segment s
// writes output rows [SS s, SS (s + 1)) and reads H rows
// [SS s / 2 - NS / 2, SS (s + 1) / 2 + NS / 2).
template <typename T, int NS, int SS, bool VEC>
__device__ __forceinline__ void col_inv_body(int b, const T* __restrict__ ll, long long llp, const T* __restrict__ det,
                                             long long dp, T* __restrict__ dst, long long dstp, int n, int count,
                                             int seg0, int hbase, const Coeffs<T>& cf) {
  constexpr int LAG = NS / 2, NP = SS / 2 + NS;
  // row-major: consecutive workgroups take adjacent column groups of one row
  // segment, so the workgroups in flight read whole rows
  const int ngr = (count + 128 * kWavesPerBlock - 1) / (128 * kWavesPerBlock);
  const int sl = b / ngr, g = b - sl * ngr, seg = seg0 + sl;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int x = (g * kWavesPerBlock + wave) * 128 + 2 * lane;
  if (__builtin_amdgcn_readfirstlane((g * kWavesPerBlock + wave) * 128) >= count) return;
  const bool va = x < count, vb = x + 1 < count;
  const bool vo = pair_ok(dst, dstp);
  const int y0 = seg * SS, y1 = min(y0 + SS, n);
  const int pbeg = (y0 >> 1) - LAG;
  const Gbl<T>* lp = gload_ptr(ll);
  const Gbl<T>* hp = gload_ptr(det);
  T ra[2 * NP], rb[2 * NP];
#pragma unroll
  for (int u = 0; u < 2 * NP; ++u) {
    const int Yr = col_load_row(seg, u, n, SS, NS);
    const int m = Yr >> 1;
    const Gbl<T>* row = (Yr & 1) ? hp + (long long)(hbase + m) * dp + x : lp + (long long)m * llp + x;
    if (VEC && vb) {
      using V2 = typename GPair<T>::type;
      const V2 v = *(const Gbl<V2>*)row;
      ra[u] = v.x;
      rb[u] = v.y;
    } else {
      ra[u] = va ? row[0] : T(0);
      rb[u] = vb ? row[1] : T(0);
    }
  }
  ColInv<T, NS> ca, cb;
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    T fa, sa, fb, sb;
    ca.push(ra[2 * u], ra[2 * u + 1], cf, fa, sa);
    cb.push(rb[2 * u], rb[2 * u + 1], cf, fb, sb);
    const int p = pbeg + u;
    const int r0 = NS == 4 ? 2 * p - 3 : 2 * p - 2;
    if (r0 >= y0 && r0 < y1) {
      st_pair(dst + (long long)r0 * dstp + x, fa, fb, va, vb, vo);
    }
    if (r0 + 1 >= y0 && r0 + 1 < y1) {
      st_pair(dst + (long long)(r0 + 1) * dstp + x, sa, sb, va, vb, vo);
    }
  }
}

}  // namespace dwt
