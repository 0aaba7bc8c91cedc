// Packed fast path: one 2D lifting level per launch, one HBM read of the
// level input and one write of its four subbands; the single-workgroup
// block kernels for small pyramids; host launchers and the tile-shape table.
// The tile bodies (work decomposition, lifting in registers) are in
// tile_body.h.
#include "tile_body.h"
#include "kernels.h"

namespace dwt {

// ---- per-level kernels (many workgroups) -------------------------------------
template <typename T, int NS, bool VEC, int U>
__global__ void __launch_bounds__(256)
fwd_level_kernel(const T* __restrict__ src, long long sp, int W, int H, T* __restrict__ ll,
                 long long llp, T* __restrict__ det, long long dp, int S, int nstrips, int nseg,
                 Coeffs<T> cf) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int g = b / nseg, s = b - g * nseg;
  const int strip = g * kWavesPerBlock + (threadIdx.x >> 6);
  if (strip >= nstrips) return;
  fwd_tile<T, NS, U>(strip, s, S, W, H, GlobalRows<T, VEC>{src, sp, H}, GlobalStore<T>{ll, llp},
                     GlobalDet<T>{det, dp}, cf);
}

template <typename T, int NS, bool VEC, int U>
__global__ void __launch_bounds__(256)
inv_level_kernel(const T* __restrict__ ll, long long llp, const T* __restrict__ det, long long dp,
                 T* __restrict__ dst, long long dstp, int W, int H, int S, int nstrips, int nseg,
                 Coeffs<T> cf) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int g = b / nseg, s = b - g * nseg;
  const int strip = g * kWavesPerBlock + (threadIdx.x >> 6);
  if (strip >= nstrips) return;
  inv_tile<T, NS, U>(strip, s, S, W, H, GlobalLoad<T>{ll, llp},
                     GlobalDet<T>{const_cast<T*>(det), dp}, GlobalOut<T, VEC>{dst, dstp}, cf);
}

// ---- block kernels (one workgroup, a whole small pyramid in LDS) --------------
// The small levels of a transform are latency-bound: every kernel boundary on
// this multi-XCD GPU writes back and invalidates L2, and a level of a few KB
// cannot fill the machine.  A block kernel instead moves the whole small Mallat
// region (the LL input of the first small level, or all coefficients of the
// small levels for the inverse) into LDS with one cooperative sweep, runs every
// remaining level there, and writes the region back with one sweep: two global
// round trips for all small levels.  Because the whole region is read before
// anything is written, it may also work in place (an image small enough to fit
// is transformed completely by one launch).
//
// Within LDS a level is computed line-parallel by the whole workgroup, each
// lifting step a barrier-separated sweep (x + c * (l + r), a missing neighbour
// at a line end replaced by the existing one); the parity scaling is folded
// into the move between Mallat and interleaved order through the second LDS
// buffer.  Rows first, then columns, in both directions.
constexpr int kBlockThreads = 1024;

// Lifting steps on the H rows (length W >= 2) of buf: wave w takes rows
// w, w + nwaves, ...; lane l takes the samples of the step's parity at
// positions par + 2l, par + 2l + 128, ...
template <typename T>
__device__ __forceinline__ void lds_lift_rows(Lds<T>* buf, int pitch, int W, int H,
                                              const Coeffs<T>& cf) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int st = 0; st < cf.nsteps; ++st) {
    const int par = (cf.first_parity + st) & 1;
    const T c = cf.c[st];
    const T c2 = T(2) * c;
    for (int y = wave; y < H; y += nw) {
      Lds<T>* row = buf + y * pitch;
      for (int i = par + 2 * lane; i < W; i += 2 * kWave) {
        T t;
        if (i > 0 && i + 1 < W) {
          t = row[i - 1] + row[i + 1];
          t = c * t;
        } else {
          t = c2 * (i > 0 ? row[i - 1] : row[i + 1]);
        }
        row[i] = row[i] + t;
      }
    }
    __syncthreads();
  }
}

// Lifting steps on the W columns (length H >= 2) of buf: lane l takes columns
// l, l + 64, ... (consecutive lanes, consecutive LDS banks); wave w takes the
// step's positions par + 2w, par + 2w + 2 nwaves, ...
template <typename T>
__device__ __forceinline__ void lds_lift_cols(Lds<T>* buf, int pitch, int W, int H,
                                              const Coeffs<T>& cf) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int st = 0; st < cf.nsteps; ++st) {
    const int par = (cf.first_parity + st) & 1;
    const T c = cf.c[st];
    const T c2 = T(2) * c;
    for (int x = lane; x < W; x += kWave) {
      for (int i = par + 2 * wave; i < H; i += 2 * nw) {
        Lds<T>* e = buf + i * pitch + x;
        T t;
        if (i > 0 && i + 1 < H) {
          t = e[-pitch] + e[pitch];
          t = c * t;
        } else {
          t = c2 * (i > 0 ? e[-pitch] : e[pitch]);
        }
        e[0] = e[0] + t;
      }
    }
    __syncthreads();
  }
}

// Permutation with the parity scaling folded in, over the W x H region.
// Forward (deinterleave): out[.., j] = in[.., i] * (i odd ? so : se) with even
// i to j = i/2 and odd i to j = ceil(N/2) + i/2.  Inverse (interleave):
// out[.., j] = in[.., i] * (j odd ? so : se) with i < ceil(N/2) to j = 2i, the
// rest to j = 2(i - ceil(N/2)) + 1.  Along rows (ROWS) or columns.
template <typename T, bool ROWS, bool DEINTERLEAVE>
__device__ __forceinline__ void lds_permute_scale(const Lds<T>* in, Lds<T>* out, int pitch, int W,
                                                  int H, T se, T so) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int N = ROWS ? W : H, half = (N + 1) >> 1;
  for (int y = wave; y < H; y += nw) {
    for (int x = lane; x < W; x += kWave) {
      const int i = ROWS ? x : y;
      int j;
      bool odd;
      if (DEINTERLEAVE) {
        j = (i & 1) ? half + (i >> 1) : (i >> 1);
        odd = i & 1;
      } else {
        j = i < half ? 2 * i : 2 * (i - half) + 1;
        odd = j & 1;
      }
      const int yo = ROWS ? y : j, xo = ROWS ? j : x;
      out[yo * pitch + xo] = in[y * pitch + x] * (odd ? so : se);
    }
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void block_load(const T* __restrict__ src, long long sp, Lds<T>* b,
                                           int W, int H) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int y = wave; y < H; y += nw)
    for (int x = lane; x < W; x += kWave) b[y * W + x] = src[(long long)y * sp + x];
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void block_store(const Lds<T>* b, T* __restrict__ dst, long long dp,
                                            int W, int H) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int y = wave; y < H; y += nw)
    for (int x = lane; x < W; x += kWave) st_global(dst + (long long)y * dp + x, (T)b[y * W + x]);
}

// Forward: levels 0..nlev-1 of the W0 x H0 region read from src (pitch sp),
// result (Mallat layout) to dst (pitch dp); src may equal dst.
template <typename T>
__global__ void __launch_bounds__(kBlockThreads)
block_fwd_kernel(const T* __restrict__ src, long long sp, T* __restrict__ dst, long long dp, int W0,
                 int H0, int nlev, Coeffs<T> cf) {
  extern __shared__ __align__(16) unsigned char blk_smem[];
  Lds<T>* const b0 = (Lds<T>*)blk_smem;  // generic -> LDS address-space cast
  Lds<T>* const b1 = b0 + W0 * H0;
  block_load(src, sp, b0, W0, H0);
  for (int lev = 0; lev < nlev; ++lev) {
    const int W = ceil_div_pow2(W0, lev), H = ceil_div_pow2(H0, lev);
    lds_lift_rows(b0, W0, W, H, cf);  // rows: lift, then scale into Mallat order
    lds_permute_scale<T, true, true>(b0, b1, W0, W, H, cf.se, cf.so);
    lds_lift_cols(b1, W0, W, H, cf);  // columns
    lds_permute_scale<T, false, true>(b1, b0, W0, W, H, cf.se, cf.so);
  }
  block_store(b0, dst, dp, W0, H0);
}

// Inverse: the nlev levels whose finest output is W0 x H0; coefficients of the
// W0 x H0 region from src (pitch sp), output to dst (pitch dp); src may equal dst.
template <typename T>
__global__ void __launch_bounds__(kBlockThreads)
block_inv_kernel(const T* __restrict__ src, long long sp, T* __restrict__ dst, long long dp, int W0,
                 int H0, int nlev, Coeffs<T> cf) {
  extern __shared__ __align__(16) unsigned char blk_smem[];
  Lds<T>* const b0 = (Lds<T>*)blk_smem;  // generic -> LDS address-space cast
  Lds<T>* const b1 = b0 + W0 * H0;
  block_load(src, sp, b0, W0, H0);
  for (int lev = nlev - 1; lev >= 0; --lev) {
    const int W = ceil_div_pow2(W0, lev), H = ceil_div_pow2(H0, lev);
    lds_permute_scale<T, true, false>(b0, b1, W0, W, H, cf.se, cf.so);  // rows: order, scale,
    lds_lift_rows(b1, W0, W, H, cf);                                    // lift
    lds_permute_scale<T, false, false>(b1, b0, W0, W, H, cf.se, cf.so);  // columns
    lds_lift_cols(b0, W0, W, H, cf);
  }
  block_store(b0, dst, dp, W0, H0);
}

template <typename T>
__global__ void copy_rect_kernel(const T* __restrict__ src, long long sp, T* __restrict__ dst,
                                 long long dp, int w, int h) {
  const long long n = (long long)w * h;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int y = (int)(t / w), x = (int)(t - (long long)y * w);
    st_global(dst + y * dp + x, src[y * sp + x]);
  }
}

// ---- host launchers ---------------------------------------------------------

// Tile shape of one launch: S rows per segment (tall segments amortise the
// 2*HALO halo rows, short ones give a level more wavefronts) and U row pairs
// prefetched per wavefront (U = 8 keeps more bytes in flight and lets a short
// segment issue all its loads at once).  The table is indexed by the level's
// height class ceil(log2(H)); entries are set by dwt_gpu_tile_set (the
// autotuner) and default to the heuristic below.
struct TileShape {
  int S, U;
};
static TileShape g_tiles[2][32];  // [direction][height class]; S == 0: heuristic
static int height_class(int H) {
  int c = 0;
  while ((1 << c) < H && c < 31) ++c;
  return c;
}

static TileShape pick_tile(int H, int nstrips, bool fwd) {
  const TileShape t = g_tiles[fwd ? 0 : 1][height_class(H)];
  if (t.S > 0) return t;
  // default from autotuning sweeps on MI355X (4096^2 and 3001x2047, both
  // wavelets, both directions): 8-row segments loaded in a single shot were
  // fastest or within noise on every level size
  (void)nstrips;
  (void)H;
  return TileShape{8, 0};
}

int tile_set(int fwd, int hclass, int S, int U) {
  if (hclass < 0 || hclass >= 32) return -1;
  if (S != 0 && (S < 2 || (S & 1) || S > 4096)) return -1;
  // U = 4 or 8: rolling prefetch of U row pairs; U = 0: single shot, the
  // whole segment requested at once (S = 8, 16 or 32)
  if (S != 0 && U != 4 && U != 8 && !(U == 0 && (S == 8 || S == 16 || S == 32))) return -1;
  g_tiles[fwd ? 0 : 1][hclass] = TileShape{S, U};
  return 0;
}

int tile_get(int fwd, int H, int W, int nsteps, int* S, int* U) {
  const int core = 2 * kWave - 2 * nsteps;
  const TileShape t = pick_tile(H, (W + core - 1) / core, fwd != 0);
  *S = t.S;
  *U = t.U;
  return 0;
}

template <typename T>
static bool aligned_pair(const void* p, long long pitch) {
  return (reinterpret_cast<uintptr_t>(p) % (2 * sizeof(T))) == 0 && (pitch % 2) == 0;
}

// Instantiation for a tile shape: rolling prefetch (U = 4 / 8, any S) or
// single shot (U = 0, S in {8, 16, 32}, template U = -S).
template <typename T, int NS, bool VEC, class Launch>
static void with_shape(const TileShape& t, const Launch& launch) {
  if (t.U == 0) {
    if (t.S == 8) launch.template go<T, NS, VEC, -8>();
    else if (t.S == 16) launch.template go<T, NS, VEC, -16>();
    else launch.template go<T, NS, VEC, -32>();
  } else if (t.U == 8) {
    launch.template go<T, NS, VEC, 8>();
  } else {
    launch.template go<T, NS, VEC, 4>();
  }
}

template <typename T, class Launch>
static void with_variant(int nsteps, bool vec, const TileShape& t, const Launch& launch) {
  if (nsteps == 4) {
    if (vec) with_shape<T, 4, true>(t, launch);
    else with_shape<T, 4, false>(t, launch);
  } else {
    if (vec) with_shape<T, 2, true>(t, launch);
    else with_shape<T, 2, false>(t, launch);
  }
}

template <typename T>
struct FwdLaunch {
  const T* src; long long sp; int W, H; T* ll; long long llp; T* det; long long dp;
  int S, nstrips, nseg; dim3 grid, block; Coeffs<T> cf; hipStream_t st;
  template <typename TT, int NS, bool VEC, int U>
  void go() const {
    hipLaunchKernelGGL((fwd_level_kernel<TT, NS, VEC, U>), grid, block, 0, st, src, sp, W, H, ll,
                       llp, det, dp, S, nstrips, nseg, cf);
  }
};

template <typename T>
struct InvLaunch {
  const T* ll; long long llp; const T* det; long long dp; T* dst; long long dstp; int W, H;
  int S, nstrips, nseg; dim3 grid, block; Coeffs<T> cf; hipStream_t st;
  template <typename TT, int NS, bool VEC, int U>
  void go() const {
    hipLaunchKernelGGL((inv_level_kernel<TT, NS, VEC, U>), grid, block, 0, st, ll, llp, det, dp,
                       dst, dstp, W, H, S, nstrips, nseg, cf);
  }
};

template <typename T>
void launch_fwd_level(const T* src, long long sp, int W, int H, T* ll, long long llp, T* det,
                      long long dp, const Coeffs<T>& cf, hipStream_t st) {
  const int core = 2 * kWave - 2 * cf.nsteps;
  const int nstrips = (W + core - 1) / core;
  const TileShape t = pick_tile(H, nstrips, true);
  const int nseg = (H + t.S - 1) / t.S;
  const int ngroups = (nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
  const FwdLaunch<T> L{src, sp, W, H, ll, llp, det, dp, t.S, nstrips, nseg,
                       dim3(ngroups * nseg), dim3(kWave * kWavesPerBlock), cf, st};
  with_variant<T>(cf.nsteps, aligned_pair<T>(src, sp), t, L);
}

template <typename T>
void launch_inv_level(const T* ll, long long llp, const T* det, long long dp, T* dst,
                      long long dstp, int W, int H, const Coeffs<T>& cf, hipStream_t st) {
  const int core = 2 * kWave - 2 * cf.nsteps;
  const int nstrips = (W + core - 1) / core;
  const TileShape t = pick_tile(H, nstrips, false);
  const int nseg = (H + t.S - 1) / t.S;
  const int ngroups = (nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
  const InvLaunch<T> L{ll, llp, det, dp, dst, dstp, W, H, t.S, nstrips, nseg,
                       dim3(ngroups * nseg), dim3(kWave * kWavesPerBlock), cf, st};
  with_variant<T>(cf.nsteps, aligned_pair<T>(dst, dstp), t, L);
}

// ---- block (tail) launchers --------------------------------------------------------
// The block kernel takes over from the first level whose region fits the LDS
// budget (two buffers of W x H elements).  Level 0 may be included: the block
// kernel reads its whole region before writing, so it works in place.
static long long g_block_bytes = 16 * 1024;  // one LDS buffer
constexpr int kBlockLdsMax = 150 * 1024;

void tail_config(long long max_region_bytes, int) {
  g_block_bytes = max_region_bytes > kBlockLdsMax / 2 ? kBlockLdsMax / 2 : max_region_bytes;
}

int tail_start(int W, int H, int L, int elem) {
  if (g_block_bytes <= 0) return L;
  for (int j = 0; j < L; ++j) {
    const long long wj = ceil_div_pow2(W, j), hj = ceil_div_pow2(H, j);
    if (wj * hj * elem <= g_block_bytes) return j;
  }
  return L;
}

hipError_t configure_tail_kernels() {
  const void* fns[] = {reinterpret_cast<const void*>(&block_fwd_kernel<float>),
                       reinterpret_cast<const void*>(&block_fwd_kernel<double>),
                       reinterpret_cast<const void*>(&block_inv_kernel<float>),
                       reinterpret_cast<const void*>(&block_inv_kernel<double>)};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kBlockLdsMax);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return e;
    }
  }
  return hipSuccess;
}

template <typename T>
void launch_block_fwd(const T* src, long long sp, T* dst, long long dp, int W0, int H0, int nlev,
                      const Coeffs<T>& cf, hipStream_t st) {
  const size_t lds = 2 * (size_t)W0 * H0 * sizeof(T);
  hipLaunchKernelGGL((block_fwd_kernel<T>), dim3(1), dim3(kBlockThreads), lds, st, src, sp, dst, dp,
                     W0, H0, nlev, cf);
}

template <typename T>
void launch_block_inv(const T* src, long long sp, T* dst, long long dp, int W0, int H0, int nlev,
                      const Coeffs<T>& cf, hipStream_t st) {
  const size_t lds = 2 * (size_t)W0 * H0 * sizeof(T);
  hipLaunchKernelGGL((block_inv_kernel<T>), dim3(1), dim3(kBlockThreads), lds, st, src, sp, dst, dp,
                     W0, H0, nlev, cf);
}

template void launch_block_fwd<float>(const float*, long long, float*, long long, int, int, int,
                                      const Coeffs<float>&, hipStream_t);
template void launch_block_fwd<double>(const double*, long long, double*, long long, int, int, int,
                                       const Coeffs<double>&, hipStream_t);
template void launch_block_inv<float>(const float*, long long, float*, long long, int, int, int,
                                      const Coeffs<float>&, hipStream_t);
template void launch_block_inv<double>(const double*, long long, double*, long long, int, int, int,
                                       const Coeffs<double>&, hipStream_t);

template <typename T>
void launch_copy_rect(const T* src, long long sp, T* dst, long long dp, int w, int h,
                      hipStream_t st) {
  const long long n = (long long)w * h;
  if (n <= 0) return;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL((copy_rect_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, src, sp,
                     dst, dp, w, h);
}

template void launch_fwd_level<float>(const float*, long long, int, int, float*, long long,
                                      float*, long long, const Coeffs<float>&, hipStream_t);
template void launch_fwd_level<double>(const double*, long long, int, int, double*, long long,
                                       double*, long long, const Coeffs<double>&, hipStream_t);
template void launch_inv_level<float>(const float*, long long, const float*, long long, float*,
                                      long long, int, int, const Coeffs<float>&, hipStream_t);
template void launch_inv_level<double>(const double*, long long, const double*, long long,
                                       double*, long long, int, int, const Coeffs<double>&,
                                       hipStream_t);
template void launch_copy_rect<float>(const float*, long long, float*, long long, int, int,
                                      hipStream_t);
template void launch_copy_rect<double>(const double*, long long, double*, long long, int, int,
                                       hipStream_t);

}  // namespace dwt
