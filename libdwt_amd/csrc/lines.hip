// Full-depth 1D transforms of many strided signals in one call: the
// decomposition a caller gets from dwt_cdf97_f_ex_stride_s applied level by
// level, each level writing L to [0, ceil(n/2)) and H to [ceil(n/2), n) of
// the signal, until n < 2 (the inverse undoes it level by level).
//
// Two layouts have their own kernels:
//   * contiguous samples (each signal a row): one workgroup per signal holds
//     the whole signal in LDS and runs every level there, so a signal is read
//     and written once;
//   * adjacent signals (the signals are the columns of a matrix, sample
//     stride = row pitch): per level, one launch of a column-only version of
//     the 2D tile body's vertical sliding window; a wavefront owns 128
//     adjacent signals (two per lane, pair-wise loads) and a segment of rows.
// The arithmetic is the reference's line lifting (src/libdwt.c:5512-5603,
// 5779-5870): updates x + c * (l + r) as separate operations, (2c) * n at a
// line end, scaling after (forward) or before (inverse) the lifting.
#include "tile_body.h"
#include "kernels.h"

namespace dwt {

constexpr int kLineThreads = 256;
constexpr int kLineUnroll = 4;

// Lifting steps over x[0, n) (n >= 2) by the whole workgroup.  A step's
// updates are independent (their neighbours have the other parity), so each
// thread loads the operands of kLineUnroll positions before storing any.  A
// line end reads its one neighbour twice: c * (n + n) == (2c) * n, the
// reference's end-of-line form.
template <typename T>
__device__ __forceinline__ void lds_lift_line(Lds<T>* x, int n, const Coeffs<T>& cf) {
  constexpr int STEP = 2 * kLineThreads;
  for (int st = 0; st < cf.nsteps; ++st) {
    const int par = (cf.first_parity + st) & 1;
    const T c = cf.c[st];
    for (int i0 = par + 2 * (int)threadIdx.x; i0 < n; i0 += kLineUnroll * STEP) {
      T v[kLineUnroll], l[kLineUnroll], r[kLineUnroll];
#pragma unroll
      for (int u = 0; u < kLineUnroll; ++u) {
        const int i = i0 + u * STEP;
        if (i < n) {
          v[u] = x[i];
          l[u] = x[i > 0 ? i - 1 : i + 1];
          r[u] = x[i + 1 < n ? i + 1 : i - 1];
        }
      }
#pragma unroll
      for (int u = 0; u < kLineUnroll; ++u) {
        const int i = i0 + u * STEP;
        if (i < n) {
          T t = l[u] + r[u];
          t = c * t;
          x[i] = v[u] + t;
        }
      }
    }
    __syncthreads();
  }
}

constexpr int kLineBatch = 8;  // global loads in flight per thread

// x[0, n) = g[0, n): every thread issues kLineBatch loads before it stores
// any, so that the copy is not a chain of dependent memory round trips.
template <typename T>
__device__ __forceinline__ void copy_in(Lds<T>* x, const Gbl<T>* g, int n) {
  for (int i0 = threadIdx.x; i0 < n; i0 += kLineBatch * kLineThreads) {
    T v[kLineBatch];
#pragma unroll
    for (int u = 0; u < kLineBatch; ++u) {
      const int i = i0 + u * kLineThreads;
      if (i < n) v[u] = g[i];
    }
#pragma unroll
    for (int u = 0; u < kLineBatch; ++u) {
      const int i = i0 + u * kLineThreads;
      if (i < n) x[i] = v[u];
    }
  }
}

// One workgroup per signal: signal k at base + k * sig (elements), contiguous
// samples; `levels` levels while n >= 2.  LDS: buffers A (N elements) and B
// (ceil(N/2)); the level whose input (forward) or output (inverse) has n
// samples uses A when its index is even, B when odd.  Forward: lift in place,
// then H (odd samples x so) goes straight to its final place in global memory
// and L (even samples x se) to the other buffer for the next level.  Inverse:
// L (x se) from the buffer and H (x so) from global memory interleaved into
// the other buffer, then lifted.  Every sample is read once and written once.
template <typename T, bool FWD>
__global__ void __launch_bounds__(kLineThreads)
lines_lds_kernel(T* __restrict__ base, long long sig, int N, int levels, Coeffs<T> cf) {
  extern __shared__ __align__(16) unsigned char line_smem[];
  Lds<T>* const A = (Lds<T>*)line_smem;  // generic -> LDS address-space cast
  Lds<T>* const B = A + N;
  Gbl<T>* const g = gstore_ptr(base + (long long)blockIdx.x * sig);
  const int t = threadIdx.x;
  if (FWD) {
    copy_in(A, g, N);
    __syncthreads();
    int n = N, l = 0;
    for (; l < levels && n >= 2; ++l) {
      Lds<T>* X = (l & 1) ? B : A;
      Lds<T>* Y = (l & 1) ? A : B;
      lds_lift_line(X, n, cf);
      const int half = (n + 1) >> 1;
      for (int i = t; i < n; i += kLineThreads) {
        const T v = X[i];
        if (i & 1) g[half + (i >> 1)] = v * cf.so;
        else Y[i >> 1] = v * cf.se;
      }
      __syncthreads();
      n = half;
    }
    Lds<T>* X = (l & 1) ? B : A;
    for (int i = t; i < n; i += kLineThreads) g[i] = X[i];
  } else {
    int sizes[32];
    int L = 0;
    for (int n = N; n >= 2 && L < levels; n = (n + 1) >> 1) sizes[L++] = n;
    const int n0 = L ? (sizes[L - 1] + 1) >> 1 : N;  // coarsest L
    Lds<T>* X = (L & 1) ? B : A;
    copy_in(X, g, n0);
    __syncthreads();
    for (int l = L - 1; l >= 0; --l) {
      const int n = sizes[l], half = (n + 1) >> 1;
      Lds<T>* Xi = (l & 1) ? A : B;  // input: level l + 1's buffer
      Lds<T>* Y = (l & 1) ? B : A;
      // H from global memory, kLineBatch loads in flight per thread
      for (int k0 = t; k0 < n - half; k0 += kLineBatch * kLineThreads) {
        T h[kLineBatch];
#pragma unroll
        for (int u = 0; u < kLineBatch; ++u) {
          const int k = k0 + u * kLineThreads;
          if (k < n - half) h[u] = g[half + k];
        }
#pragma unroll
        for (int u = 0; u < kLineBatch; ++u) {
          const int k = k0 + u * kLineThreads;
          if (k < n - half) Y[2 * k + 1] = h[u] * cf.so;
        }
      }
      for (int k = t; k < half; k += kLineThreads) Y[2 * k] = Xi[k] * cf.se;
      __syncthreads();
      lds_lift_line(Y, n, cf);
    }
    for (int i = t; i < N; i += kLineThreads) g[i] = A[i];
  }
}

// ---- column kernels (adjacent signals) ---------------------------------------------
// A signal is a column: sample i of signal x at row i.  Forward level on n
// rows: src (row pitch sp) -> L rows to ll (pitch llp), H rows to det (pitch
// dp) starting at row ceil(n/2).  count signals; wave = 128 columns.
template <typename T, int NS, int SS, bool VEC>
__global__ void __launch_bounds__(256)
col_fwd_kernel(const T* __restrict__ src, long long sp, T* __restrict__ ll, long long llp,
               T* __restrict__ det, long long dp, int n, int count, int nseg, Coeffs<T> cf) {
  constexpr int LAG = NS / 2, NP = SS / 2 + NS;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int g = b / nseg, seg = b - g * nseg;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int x = (g * kWavesPerBlock + wave) * 128 + 2 * lane;
  if (__builtin_amdgcn_readfirstlane((g * kWavesPerBlock + wave) * 128) >= count) return;
  const bool va = x < count, vb = x + 1 < count;
  const int half = (n + 1) >> 1, nh = n >> 1;
  const int m0 = (seg * SS) >> 1, m1 = min(m0 + SS / 2, half);
  const int pbeg = m0 - LAG;
  const Gbl<T>* s = gload_ptr(src);
  T ra[2 * NP], rb[2 * NP];
#pragma unroll
  for (int u = 0; u < 2 * NP; ++u) {
    const Gbl<T>* row = s + (long long)mirror_near(2 * pbeg + u, n) * sp + x;
    if (VEC && vb) {
      using V2 = typename Vec2<T>::type;
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
      if (va) st_global(ll + (long long)m * llp + x, La);
      if (vb) st_global(ll + (long long)m * llp + x + 1, Lb);
      if (m < nh) {
        if (va) st_global(det + (long long)(half + m) * dp + x, Ha);
        if (vb) st_global(det + (long long)(half + m) * dp + x + 1, Hb);
      }
    }
  }
}

// Inverse level on n output rows: L rows from ll (pitch llp), H rows from det
// (pitch dp) starting at row ceil(n/2); output rows to dst (pitch dstp).
template <typename T, int NS, int SS, bool VEC>
__global__ void __launch_bounds__(256)
col_inv_kernel(const T* __restrict__ ll, long long llp, const T* __restrict__ det, long long dp,
               T* __restrict__ dst, long long dstp, int n, int count, int nseg, Coeffs<T> cf) {
  constexpr int LAG = NS / 2, NP = SS / 2 + NS;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int g = b / nseg, seg = b - g * nseg;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int x = (g * kWavesPerBlock + wave) * 128 + 2 * lane;
  if (__builtin_amdgcn_readfirstlane((g * kWavesPerBlock + wave) * 128) >= count) return;
  const bool va = x < count, vb = x + 1 < count;
  const int half = (n + 1) >> 1;
  const int y0 = seg * SS, y1 = min(y0 + SS, n);
  const int pbeg = (y0 >> 1) - LAG;
  const Gbl<T>* lp = gload_ptr(ll);
  const Gbl<T>* hp = gload_ptr(det);
  T ra[2 * NP], rb[2 * NP];
#pragma unroll
  for (int u = 0; u < 2 * NP; ++u) {
    const int Yr = mirror_near(2 * pbeg + u, n);
    const int m = Yr >> 1;
    const Gbl<T>* row = (Yr & 1) ? hp + (long long)(half + m) * dp + x : lp + (long long)m * llp + x;
    if (VEC && vb) {
      using V2 = typename Vec2<T>::type;
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
      if (va) st_global(dst + (long long)r0 * dstp + x, fa);
      if (vb) st_global(dst + (long long)r0 * dstp + x + 1, fb);
    }
    if (r0 + 1 >= y0 && r0 + 1 < y1) {
      if (va) st_global(dst + (long long)(r0 + 1) * dstp + x, sa);
      if (vb) st_global(dst + (long long)(r0 + 1) * dstp + x + 1, sb);
    }
  }
}

// ---- host launchers ------------------------------------------------------------------
// Longest signal of the LDS kernel: N + ceil(N/2) elements in 96 KB.
int lines_lds_max(int elem) { return (2 * 96 * 1024 / 3) / elem; }

template <typename T>
void launch_lines_lds(bool fwd, T* base, long long sig, int N, int count, int levels,
                      const Coeffs<T>& cf, hipStream_t st) {
  const size_t lds = sizeof(T) * ((size_t)N + (N + 1) / 2);
  if (fwd)
    hipLaunchKernelGGL((lines_lds_kernel<T, true>), dim3(count), dim3(kLineThreads), lds, st, base,
                       sig, N, levels, cf);
  else
    hipLaunchKernelGGL((lines_lds_kernel<T, false>), dim3(count), dim3(kLineThreads), lds, st,
                       base, sig, N, levels, cf);
}

constexpr int kColRows = 8;

template <typename T, int NS, bool VEC>
static void col_go(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c, long long cp,
                   int n, int count, const Coeffs<T>& cf, hipStream_t st) {
  const int nseg = (n + kColRows - 1) / kColRows;
  const int ngroups = (count + 128 * kWavesPerBlock - 1) / (128 * kWavesPerBlock);
  const dim3 grid(ngroups * nseg), block(256);
  if (fwd)  // a = src, b = ll (written), c = det
    hipLaunchKernelGGL((col_fwd_kernel<T, NS, kColRows, VEC>), grid, block, 0, st, a, ap,
                       const_cast<T*>(b), bp, c, cp, n, count, nseg, cf);
  else  // a = ll, b = det, c = dst
    hipLaunchKernelGGL((col_inv_kernel<T, NS, kColRows, VEC>), grid, block, 0, st, a, ap, b, bp, c,
                       cp, n, count, nseg, cf);
}

template <typename T>
void launch_col_level(bool fwd, const T* a, long long ap, const T* b, long long bp, T* c,
                      long long cp, int n, int count, const Coeffs<T>& cf, bool vec,
                      hipStream_t st) {
  if (cf.nsteps == 4) {
    if (vec) col_go<T, 4, true>(fwd, a, ap, b, bp, c, cp, n, count, cf, st);
    else col_go<T, 4, false>(fwd, a, ap, b, bp, c, cp, n, count, cf, st);
  } else {
    if (vec) col_go<T, 2, true>(fwd, a, ap, b, bp, c, cp, n, count, cf, st);
    else col_go<T, 2, false>(fwd, a, ap, b, bp, c, cp, n, count, cf, st);
  }
}

hipError_t configure_lines_kernels() {
  const void* fns[] = {reinterpret_cast<const void*>(&lines_lds_kernel<float, true>),
                       reinterpret_cast<const void*>(&lines_lds_kernel<float, false>),
                       reinterpret_cast<const void*>(&lines_lds_kernel<double, true>),
                       reinterpret_cast<const void*>(&lines_lds_kernel<double, false>)};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return e;
    }
  }
  return hipSuccess;
}

template void launch_lines_lds<float>(bool, float*, long long, int, int, int, const Coeffs<float>&,
                                      hipStream_t);
template void launch_lines_lds<double>(bool, double*, long long, int, int, int,
                                       const Coeffs<double>&, hipStream_t);
template void launch_col_level<float>(bool, const float*, long long, const float*, long long, float*,
                                      long long, int, int, const Coeffs<float>&, bool, hipStream_t);
template void launch_col_level<double>(bool, const double*, long long, const double*, long long,
                                       double*, long long, int, int, const Coeffs<double>&, bool,
                                       hipStream_t);

}  // namespace dwt
