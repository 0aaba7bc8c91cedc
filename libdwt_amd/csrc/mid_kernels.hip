// Fused middle levels: up to four consecutive levels of the packed 2D
// transform in one launch.
//
// Below a few MB a level is too small to keep the GPU busy, and every launch
// costs several microseconds of fixed latency (launch, the kernel-boundary L2
// write-back, the first loads from memory).  Here each workgroup owns a
// rectangle of the coarsest level's output (its "core", mapped to every finer
// level by doubling) and computes all fused levels for it, keeping the
// intermediate LL coefficients in LDS windows.  A window holds the core plus
// the halo the next level needs (the lifting's reach: NS samples per side at
// the level's input, three pairs for the inverse), so neighbouring workgroups
// recompute the overlap instead of exchanging it: no grid-wide
// synchronisation, one launch.
//
// Forward (levels j .. j+nl-1): the first fused level reads LL_j from global
// memory; each level writes its detail coefficients for the core only, in
// place in the Mallat layout, and its LL to the next level's window (the last
// level: the core of LL_{j+nl} to global memory).
// Inverse (levels j+nl-1 .. j): the coarsest fused level reads LL_{j+nl}
// from global memory; details are read from the Mallat layout wherever the
// halo needs them; each level writes its output window to LDS, the finest
// level its core of LL_j to global memory.
//
// The tile bodies are the same as the per-level kernels' (tile_body.h) with
// the strips and segments offset to the window, so the arithmetic, and with
// it every coefficient, is identical.  Reads from a window are clamped to it:
// a clamped value only reaches outputs outside the window, which are dropped.
#include "tile_body.h"
#include "kernels.h"

namespace dwt {

constexpr int kMidThreads = 1024;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

struct Span {
  int lo, hi;  // [lo, hi)
};

template <typename T>
struct WinRows {  // level input from an LDS window (rows y0.., columns x0..)
  const Lds<T>* p;
  int x0, y0, w, h;
  __device__ __forceinline__ void operator()(int Y, int, int xe, int xo, bool, T& e, T& o) const {
    const Lds<T>* rp = p + clampi(Y - y0, 0, h - 1) * w;
    e = rp[clampi(xe - x0, 0, w - 1)];
    o = rp[clampi(xo - x0, 0, w - 1)];
  }
};
template <typename T>
struct WinStore {  // LL coefficient into an LDS window, kept only inside it
  Lds<T>* p;
  int x0, y0, w, h;
  __device__ __forceinline__ void operator()(int m, int k, T v) const {
    const int r = m - y0, c = k - x0;
    if ((unsigned)r < (unsigned)h && (unsigned)c < (unsigned)w) p[r * w + c] = v;
  }
};
template <typename T>
struct WinLoad {  // LL coefficient from an LDS window
  const Lds<T>* p;
  int x0, y0, w, h;
  __device__ __forceinline__ T operator()(int m, int k) const {
    return p[clampi(m - y0, 0, h - 1) * w + clampi(k - x0, 0, w - 1)];
  }
};
template <typename T>
struct WinOut {  // inverse output samples into an LDS window, kept only inside it
  Lds<T>* p;
  int x0, y0, w, h;
  __device__ __forceinline__ void operator()(int r, int Xe, T e, T o, bool st_e, bool st_o) const {
    const int rr = r - y0, c = Xe - x0;
    if ((unsigned)rr >= (unsigned)h) return;
    if (st_e && (unsigned)c < (unsigned)w) p[rr * w + c] = e;
    if (st_o && (unsigned)(c + 1) < (unsigned)w) p[rr * w + c + 1] = o;
  }
};
template <typename T>
struct CoreStore {  // LL coefficient to global memory, kept only inside the core
  T* __restrict__ p;
  long long pitch;
  Span x, y;
  __device__ __forceinline__ void operator()(int m, int k, T v) const {
    if (m >= y.lo && m < y.hi && k >= x.lo && k < x.hi) st_global(p + (long long)m * pitch + k, v);
  }
};
template <typename T>
struct CoreDet {  // detail coefficient to the Mallat layout, kept only inside the core
  T* __restrict__ p;
  long long pitch;
  int Wd, Hd;
  Span x, y;  // core pairs
  __device__ __forceinline__ void store(int band, int r, int c, T v) const {
    const int m = band == kHL ? r : r - Hd, k = band == kLH ? c : c - Wd;
    if (m >= y.lo && m < y.hi && k >= x.lo && k < x.hi) st_global(p + (long long)r * pitch + c, v);
  }
};
template <typename T>
struct CoreOut {  // inverse output samples to global memory, kept only inside the core
  T* __restrict__ p;
  long long pitch;
  Span x, y;
  __device__ __forceinline__ void operator()(int r, int Xe, T e, T o, bool st_e, bool st_o) const {
    if (r < y.lo || r >= y.hi) return;
    T* rp = p + (long long)r * pitch;
    if (st_e && Xe >= x.lo && Xe < x.hi) st_global(rp + Xe, e);
    if (st_o && Xe + 1 >= x.lo && Xe + 1 < x.hi) st_global(rp + Xe + 1, o);
  }
};

__device__ __forceinline__ Span clip(int lo, int hi, int n) { return Span{max(lo, 0), min(hi, n)}; }

// ---- forward ----------------------------------------------------------------------
// Pairs a workgroup computes at fused level jj (0 = finest), columns or rows:
// the last level computes its core; a finer level the samples the next one
// reads, [2 lo - NS, 2 hi + NS).  n[jj]: pairs of level jj (ceil of half its
// width or height).
template <int NS>
__device__ inline void fwd_spans(int c, int cn, int nl, const int* n, Span* P, Span* core) {
  for (int jj = nl - 1; jj >= 0; --jj) {
    const int sh = nl - 1 - jj;
    core[jj] = clip((c * cn) << sh, ((c + 1) * cn) << sh, n[jj]);
    P[jj] = jj == nl - 1 ? core[jj] : clip(2 * P[jj + 1].lo - NS, 2 * P[jj + 1].hi + NS, n[jj]);
  }
}

template <typename T, int NS>
__global__ void __launch_bounds__(kMidThreads)
mid_fwd_kernel(const T* __restrict__ src, long long sp, T* __restrict__ det, long long dp,
               T* __restrict__ ll, long long llp, int W0, int H0, int nl, int cn, int ncx, int win1,
               Coeffs<T> cf) {
  extern __shared__ __align__(16) unsigned char mid_smem[];
  Lds<T>* const win[2] = {(Lds<T>*)mid_smem, (Lds<T>*)mid_smem + win1};  // generic -> LDS cast
  constexpr int CORE = 2 * kWave - 2 * NS;
  const int cx = blockIdx.x % ncx, cy = blockIdx.x / ncx;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nw = kMidThreads / kWave;
  int nx[kMidMaxLevels], ny[kMidMaxLevels];
  for (int jj = 0; jj < nl; ++jj) {
    nx[jj] = ceil_div_pow2(W0, jj + 1);
    ny[jj] = ceil_div_pow2(H0, jj + 1);
  }
  Span Px[kMidMaxLevels], Py[kMidMaxLevels], Cx[kMidMaxLevels], Cy[kMidMaxLevels];
  fwd_spans<NS>(cx, cn, nl, nx, Px, Cx);
  fwd_spans<NS>(cy, cn, nl, ny, Py, Cy);
  for (int jj = 0; jj < nl; ++jj) {
    const int W = ceil_div_pow2(W0, jj), H = ceil_div_pow2(H0, jj);
    const Span px = Px[jj], py = Py[jj];
    const int nstrips = (px.hi - px.lo + CORE / 2 - 1) / (CORE / 2);
    const CoreDet<T> cd{det, dp, nx[jj], ny[jj], Cx[jj], Cy[jj]};
    const bool last = jj == nl - 1;
    const int xo = 2 * px.lo, yo = 2 * py.lo;
    for_small_tiles<NS>(nstrips, 2 * (py.hi - py.lo), wave, nw, [&](auto sc, int strip, int seg) {
      constexpr int SR = decltype(sc)::value;
      if (jj == 0) {
        const GlobalRows<T, false> rows{src, sp, H};
        if (last)
          fwd_tile<T, NS, -SR>(strip, seg, SR, W, H, rows, CoreStore<T>{ll, llp, Cx[jj], Cy[jj]}, cd,
                               cf, xo, yo);
        else
          fwd_tile<T, NS, -SR>(strip, seg, SR, W, H, rows,
                               WinStore<T>{win[0], px.lo, py.lo, px.hi - px.lo, py.hi - py.lo}, cd,
                               cf, xo, yo);
      } else {
        const Span qx = Px[jj - 1], qy = Py[jj - 1];  // this level's input window
        const WinRows<T> rows{win[(jj - 1) & 1], qx.lo, qy.lo, qx.hi - qx.lo, qy.hi - qy.lo};
        if (last)
          fwd_tile<T, NS, -SR>(strip, seg, SR, W, H, rows, CoreStore<T>{ll, llp, Cx[jj], Cy[jj]}, cd,
                               cf, xo, yo);
        else
          fwd_tile<T, NS, -SR>(strip, seg, SR, W, H, rows,
                               WinStore<T>{win[jj & 1], px.lo, py.lo, px.hi - px.lo, py.hi - py.lo},
                               cd, cf, xo, yo);
      }
    });
    __syncthreads();
  }
}

// ---- inverse ----------------------------------------------------------------------
// Output samples a workgroup computes at fused level jj (0 = finest): the
// finest level its core; a coarser level the coefficient pairs the next finer
// one reads, [lo/2 - 3, (hi-1)/2 + 4).  n[jj]: samples of LL_{j+jj}.
__device__ inline void inv_spans(int c, int c0, int nl, const int* n, Span* O) {
  O[0] = clip(c * c0, (c + 1) * c0, n[0]);
  for (int jj = 1; jj < nl; ++jj) O[jj] = clip((O[jj - 1].lo >> 1) - 3, ((O[jj - 1].hi - 1) >> 1) + 4, n[jj]);
}

template <typename T, int NS>
__global__ void __launch_bounds__(kMidThreads)
mid_inv_kernel(const T* __restrict__ llsrc, long long lp, const T* __restrict__ det, long long dp,
               T* __restrict__ out, long long op, int W0, int H0, int nl, int c0, int ncx, int win1,
               Coeffs<T> cf) {
  extern __shared__ __align__(16) unsigned char mid_smem[];
  Lds<T>* const win[2] = {(Lds<T>*)mid_smem, (Lds<T>*)mid_smem + win1};  // generic -> LDS cast
  constexpr int CORE = 2 * kWave - 2 * NS;
  const int cx = blockIdx.x % ncx, cy = blockIdx.x / ncx;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nw = kMidThreads / kWave;
  int nx[kMidMaxLevels + 1], ny[kMidMaxLevels + 1];
  for (int jj = 0; jj <= nl; ++jj) {
    nx[jj] = ceil_div_pow2(W0, jj);
    ny[jj] = ceil_div_pow2(H0, jj);
  }
  Span Ox[kMidMaxLevels], Oy[kMidMaxLevels];
  inv_spans(cx, c0, nl, nx, Ox);
  inv_spans(cy, c0, nl, ny, Oy);
  const GlobalDet<T> gd{const_cast<T*>(det), dp};
  for (int jj = nl - 1; jj >= 0; --jj) {
    const int W = nx[jj], H = ny[jj];
    const Span ox = Ox[jj], oy = Oy[jj];
    const int xo = ox.lo & ~1, yo = oy.lo & ~1;
    const int nstrips = (ox.hi - xo + CORE - 1) / CORE;
    for_small_tiles<NS>(nstrips, oy.hi - yo, wave, nw, [&](auto sc, int strip, int seg) {
      constexpr int SR = decltype(sc)::value;
      if (jj == nl - 1) {
        const GlobalLoad<T> ld{llsrc, lp};
        if (jj == 0)
          inv_tile<T, NS, -SR>(strip, seg, SR, W, H, ld, gd, CoreOut<T>{out, op, ox, oy}, cf, xo, yo);
        else
          inv_tile<T, NS, -SR>(strip, seg, SR, W, H, ld, gd,
                               WinOut<T>{win[jj & 1], ox.lo, oy.lo, ox.hi - ox.lo, oy.hi - oy.lo}, cf,
                               xo, yo);
      } else {
        const Span ix = Ox[jj + 1], iy = Oy[jj + 1];  // this level's LL window
        const WinLoad<T> ld{win[(jj + 1) & 1], ix.lo, iy.lo, ix.hi - ix.lo, iy.hi - iy.lo};
        if (jj == 0)
          inv_tile<T, NS, -SR>(strip, seg, SR, W, H, ld, gd, CoreOut<T>{out, op, ox, oy}, cf, xo, yo);
        else
          inv_tile<T, NS, -SR>(strip, seg, SR, W, H, ld, gd,
                               WinOut<T>{win[jj & 1], ox.lo, oy.lo, ox.hi - ox.lo, oy.hi - oy.lo}, cf,
                               xo, yo);
      }
    });
    __syncthreads();
  }
}

// ---- host side ----------------------------------------------------------------------
// Window capacities (elements per side), from the unclipped span recurrences.
static void fwd_caps(int nl, int cn, int ns, int* cap) {
  cap[nl - 1] = cn;
  for (int jj = nl - 2; jj >= 0; --jj) cap[jj] = 2 * cap[jj + 1] + 2 * ns;
}
static void inv_caps(int nl, int c0, int* cap) {
  cap[0] = c0;
  for (int jj = 1; jj < nl; ++jj) cap[jj] = (cap[jj - 1] + 1) / 2 + 7;
}

// Core of the coarsest fused level (forward, LL pairs per side) and of the
// finest (inverse, samples per side), by element size.
static int mid_cn(int elem) { return elem == 4 ? 8 : 4; }
static int mid_c0(int elem) { return elem == 4 ? 64 : 32; }

size_t mid_lds_bytes(bool fwd, int nl, int nsteps, int elem) {
  int cap[kMidMaxLevels] = {0, 0, 0, 0};
  size_t n = 0;
  if (fwd) {
    fwd_caps(nl, mid_cn(elem), nsteps, cap);
    for (int jj = 0; jj + 1 < nl; ++jj) n = std::max(n, (size_t)cap[jj] * cap[jj]);
  } else {
    inv_caps(nl, mid_c0(elem), cap);
    for (int jj = 1; jj < nl; ++jj) n = std::max(n, (size_t)cap[jj] * cap[jj]);
  }
  // two windows (levels alternate between them), each of the largest size
  return 2 * n * elem;
}

constexpr int kMidLdsMax = 128 * 1024;

hipError_t configure_mid_kernels() {
  const void* fns[] = {reinterpret_cast<const void*>(&mid_fwd_kernel<float, 4>),
                       reinterpret_cast<const void*>(&mid_fwd_kernel<float, 2>),
                       reinterpret_cast<const void*>(&mid_fwd_kernel<double, 4>),
                       reinterpret_cast<const void*>(&mid_fwd_kernel<double, 2>),
                       reinterpret_cast<const void*>(&mid_inv_kernel<float, 4>),
                       reinterpret_cast<const void*>(&mid_inv_kernel<float, 2>),
                       reinterpret_cast<const void*>(&mid_inv_kernel<double, 4>),
                       reinterpret_cast<const void*>(&mid_inv_kernel<double, 2>)};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kMidLdsMax);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename T>
void launch_mid_fwd(const T* src, long long sp, T* det, long long dp, T* ll, long long llp, int W0,
                    int H0, int nl, const Coeffs<T>& cf, hipStream_t st) {
  const int elem = (int)sizeof(T), cn = mid_cn(elem);
  int cap[kMidMaxLevels] = {0, 0, 0, 0};
  fwd_caps(nl, cn, cf.nsteps, cap);
  const int win1 = nl >= 2 ? cap[0] * cap[0] : 0;
  const size_t lds = mid_lds_bytes(true, nl, cf.nsteps, elem);
  const int ncx = (ceil_div_pow2(W0, nl) + cn - 1) / cn, ncy = (ceil_div_pow2(H0, nl) + cn - 1) / cn;
  if (cf.nsteps == 4)
    hipLaunchKernelGGL((mid_fwd_kernel<T, 4>), dim3(ncx * ncy), dim3(kMidThreads), lds, st, src, sp,
                       det, dp, ll, llp, W0, H0, nl, cn, ncx, win1, cf);
  else
    hipLaunchKernelGGL((mid_fwd_kernel<T, 2>), dim3(ncx * ncy), dim3(kMidThreads), lds, st, src, sp,
                       det, dp, ll, llp, W0, H0, nl, cn, ncx, win1, cf);
}

template <typename T>
void launch_mid_inv(const T* llsrc, long long lp, const T* det, long long dp, T* out, long long op,
                    int W0, int H0, int nl, const Coeffs<T>& cf, hipStream_t st) {
  const int elem = (int)sizeof(T), c0 = mid_c0(elem);
  int cap[kMidMaxLevels] = {0, 0, 0, 0};
  inv_caps(nl, c0, cap);
  const int win1 = nl >= 2 ? cap[1] * cap[1] : 0;
  const size_t lds = mid_lds_bytes(false, nl, cf.nsteps, elem);
  const int ncx = (W0 + c0 - 1) / c0, ncy = (H0 + c0 - 1) / c0;
  if (cf.nsteps == 4)
    hipLaunchKernelGGL((mid_inv_kernel<T, 4>), dim3(ncx * ncy), dim3(kMidThreads), lds, st, llsrc, lp,
                       det, dp, out, op, W0, H0, nl, c0, ncx, win1, cf);
  else
    hipLaunchKernelGGL((mid_inv_kernel<T, 2>), dim3(ncx * ncy), dim3(kMidThreads), lds, st, llsrc, lp,
                       det, dp, out, op, W0, H0, nl, c0, ncx, win1, cf);
}

template void launch_mid_fwd<float>(const float*, long long, float*, long long, float*, long long,
                                    int, int, int, const Coeffs<float>&, hipStream_t);
template void launch_mid_fwd<double>(const double*, long long, double*, long long, double*, long long,
                                     int, int, int, const Coeffs<double>&, hipStream_t);
template void launch_mid_inv<float>(const float*, long long, const float*, long long, float*, long long,
                                    int, int, int, const Coeffs<float>&, hipStream_t);
template void launch_mid_inv<double>(const double*, long long, const double*, long long, double*,
                                     long long, int, int, int, const Coeffs<double>&, hipStream_t);

}  // namespace dwt
