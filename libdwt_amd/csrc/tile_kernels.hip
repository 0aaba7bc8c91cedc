// Packed fast path: one 2D lifting level per launch, one HBM read of the
// level input and one write of its four subbands; copy kernels; host
// launchers and the tile-shape table.  (The single-workgroup block kernels
// for small pyramids are in block_kernels.hip.)
// The tile bodies (work decomposition, lifting in registers) are in
// tile_body.h.
#include "tile_body.h"
#include "kernels.h"

namespace dwt {

// ---- per-level kernels (many workgroups) -------------------------------------
// Tiles are numbered row-major (ngroups column groups per row segment).  For
// the forward level the XCDs take bands of kXcdBandSegs row segments in turn,
// so neighbouring row segments (which share halo rows) mostly run on one
// XCD's L2 while all XCDs work on nearby rows; grids that are not a whole
// number of rounds keep one contiguous run per XCD (xcd_remap), as the
// inverse level always does (bands measured slower there).
#ifndef DWT_XCD_BAND_SEGS
#define DWT_XCD_BAND_SEGS 16
#endif
constexpr int kXcdBandSegs = DWT_XCD_BAND_SEGS;
__device__ inline int xcd_band_remap(int b, int nb, int ngroups) {
  const int C = kXcdBandSegs * ngroups;
  if (nb % (8 * C) != 0) return xcd_remap(b, nb);
  const int x = b & 7, idx = b >> 3;
  return (x + 8 * (idx / C)) * C + idx % C;
}

// Workgroups past the level's tiles (blockIdx.x >= tile_blocks) move rows of
// a level-0 detail copy riding on this launch (RiderCopy, kernels.h).
template <typename T, int NS, bool VEC, int U>
__global__ void __launch_bounds__(256)
fwd_level_kernel(const T* __restrict__ src, long long sp, int W, int H, T* __restrict__ ll,
                 long long llp, T* __restrict__ det, long long dp, int S, int nstrips, int nseg,
                 Coeffs<T> cf, int tile_blocks, RiderCopy<T> rc) {
  if ((int)blockIdx.x >= tile_blocks) {
    rider_copy(rc, blockIdx.x - tile_blocks, gridDim.x - tile_blocks);
    return;
  }
  const int ngroups = (nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
  const int b = xcd_band_remap(blockIdx.x, tile_blocks, ngroups);
  const int s = b / ngroups, g = b - s * ngroups;
  const int strip = g * kWavesPerBlock + (threadIdx.x >> 6);
  if (strip >= nstrips) return;
  fwd_tile<T, NS, U>(strip, s, S, W, H, GlobalRows<T, VEC>{src, sp, H}, GlobalStore<T>{ll, llp},
                     GlobalDet<T>{det, dp}, cf);
}

template <typename T, int NS, bool VEC, int U>
__global__ void __launch_bounds__(256)
inv_level_kernel(const T* __restrict__ ll, long long llp, const T* __restrict__ det, long long dp,
                 T* __restrict__ dst, long long dstp, int W, int H, int S, int nstrips, int nseg,
                 Coeffs<T> cf, int tile_blocks, RiderCopy<T> rc) {
  if ((int)blockIdx.x >= tile_blocks) {
    rider_copy(rc, blockIdx.x - tile_blocks, gridDim.x - tile_blocks);
    return;
  }
  const int ngroups = (nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
  const int b = xcd_remap(blockIdx.x, tile_blocks);  // bands measured slower for the inverse
  const int s = b / ngroups, g = b - s * ngroups;
  const int strip = g * kWavesPerBlock + (threadIdx.x >> 6);
  if (strip >= nstrips) return;
  inv_tile<T, NS, U>(strip, s, S, W, H, GlobalLoad<T>{ll, llp},
                     GlobalDet<T>{const_cast<T*>(det), dp}, GlobalOut<T, VEC>{dst, dstp}, cf);
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

// The C channels of a W x H image (element strides: rows sr, pixels sx,
// channels sc) to / from C packed planes (plane c at planes + c * plane_step,
// row pitch pp).  One thread per pixel moves all of its channels, so an
// interleaved image is read (or written) once, in order.
template <typename T, bool GATHER>
__global__ void channels_kernel(T* __restrict__ img, long long sr, long long sx, long long sc, int C, int W,
                                int H, T* __restrict__ planes, long long pp, long long plane_step) {
  const long long n = (long long)W * H;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long y = t / W, x = t - y * W;
    T* px = img + y * sr + x * sx;
    T* pl = planes + y * pp + x;
    for (int c = 0; c < C; ++c) {
      if (GATHER) st_global(pl + c * plane_step, px[c * sc]);
      else st_global(px + c * sc, pl[c * plane_step]);
    }
  }
}

// Copy of a W x H frame between two element-strided layouts (row stride r,
// column stride c, in elements, either sign).  The index order follows the
// source's shorter stride when it gathers (xfast: x varies fastest), so one
// side of the copy is read or written in order.
template <typename T>
__global__ void copy_strided_kernel(const T* __restrict__ src, long long sr, long long sc,
                                    T* __restrict__ dst, long long dr, long long dc, int W, int H,
                                    bool xfast) {
  const long long n = (long long)W * H;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    long long x, y;
    if (xfast) {
      y = t / W;
      x = t - y * W;
    } else {
      x = t / H;
      y = t - x * H;
    }
    st_global(dst + y * dr + x * dc, src[y * sr + x * sc]);
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
  // wavelets, both directions, with the level-0 detail copy riding on the
  // small levels): 8-row segments loaded in a single shot, except 4-row ones
  // for the forward levels of at most 1024 rows and the inverse levels of at
  // most 512 rows, whose shorter tile chains finish a latency-bound level
  // sooner (0.5-1.1 us per level)
  (void)nstrips;
  if (H <= (fwd ? 1024 : 512)) return TileShape{4, 0};
  return TileShape{8, 0};
}

int tile_set(int fwd, int hclass, int S, int U) {
  if (hclass < 0 || hclass >= 32) return -1;
  if (S != 0 && (S < 2 || (S & 1) || S > 4096)) return -1;
  // U = 4 or 8: rolling prefetch of U row pairs; U = 0: single shot, the
  // whole segment requested at once (S = 2, 4, 8, 12, 16, 20, 24 or 32)
  const bool single = U == 0 && (S == 2 || S == 4 || S == 8 || S == 12 || S == 16 || S == 20 ||
                                 S == 24 || S == 32);
  if (S != 0 && U != 4 && U != 8 && !single) return -1;
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
static bool aligned_pair(const void* p, long long) {
  // a sample pair moves as one element-aligned access (GPair, tile_body.h),
  // so rows of an odd pitch, half of which start at an odd element, take the
  // vector path too; only a pointer that is not element-aligned does not
  return (reinterpret_cast<uintptr_t>(p) % sizeof(T)) == 0;
}

// Instantiation for a tile shape: rolling prefetch (U = 4 / 8, any S) or
// single shot (U = 0, S in {2, 4, 8, 12, 16, 20, 24, 32}, template U = -S).
template <typename T, int NS, bool VEC, class Launch>
static void with_shape(const TileShape& t, const Launch& launch) {
  if (t.U == 0) {
    switch (t.S) {
      case 2: launch.template go<T, NS, VEC, -2>(); break;
      case 4: launch.template go<T, NS, VEC, -4>(); break;
      case 8: launch.template go<T, NS, VEC, -8>(); break;
      case 12: launch.template go<T, NS, VEC, -12>(); break;
      case 16: launch.template go<T, NS, VEC, -16>(); break;
      case 20: launch.template go<T, NS, VEC, -20>(); break;
      case 24: launch.template go<T, NS, VEC, -24>(); break;
      default: launch.template go<T, NS, VEC, -32>(); break;
    }
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
  int S, nstrips, nseg; dim3 grid, block; Coeffs<T> cf; hipStream_t st; int tiles; RiderCopy<T> rc;
  template <typename TT, int NS, bool VEC, int U>
  void go() const {
    hipLaunchKernelGGL((fwd_level_kernel<TT, NS, VEC, U>), grid, block, 0, st, src, sp, W, H, ll,
                       llp, det, dp, S, nstrips, nseg, cf, tiles, rc);
  }
};

template <typename T>
struct InvLaunch {
  const T* ll; long long llp; const T* det; long long dp; T* dst; long long dstp; int W, H;
  int S, nstrips, nseg; dim3 grid, block; Coeffs<T> cf; hipStream_t st; int tiles; RiderCopy<T> rc;
  template <typename TT, int NS, bool VEC, int U>
  void go() const {
    hipLaunchKernelGGL((inv_level_kernel<TT, NS, VEC, U>), grid, block, 0, st, ll, llp, det, dp,
                       dst, dstp, W, H, S, nstrips, nseg, cf, tiles, rc);
  }
};

template <typename T>
static RiderCopy<T> rider_or_none(const RiderCopy<T>* r) {
  if (r && r->r1 > r->r0 && r->blocks > 0) return *r;
  RiderCopy<T> none{};
  return none;
}

template <typename T>
void launch_fwd_level(const T* src, long long sp, int W, int H, T* ll, long long llp, T* det,
                      long long dp, const Coeffs<T>& cf, hipStream_t st, const RiderCopy<T>* rider) {
  const int core = 2 * kWave - 2 * cf.nsteps;
  const int nstrips = (W + core - 1) / core;
  const TileShape t = pick_tile(H, nstrips, true);
  const int nseg = (H + t.S - 1) / t.S;
  const int ngroups = (nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
  const RiderCopy<T> rc = rider_or_none(rider);
  const int tiles = ngroups * nseg;
  const FwdLaunch<T> L{src, sp, W, H, ll, llp, det, dp, t.S, nstrips, nseg,
                       dim3(tiles + rc.blocks), dim3(kWave * kWavesPerBlock), cf, st, tiles, rc};
  with_variant<T>(cf.nsteps, aligned_pair<T>(src, sp), t, L);
}

template <typename T>
void launch_inv_level(const T* ll, long long llp, const T* det, long long dp, T* dst,
                      long long dstp, int W, int H, const Coeffs<T>& cf, hipStream_t st,
                      const RiderCopy<T>* rider) {
  const int core = 2 * kWave - 2 * cf.nsteps;
  const int nstrips = (W + core - 1) / core;
  const TileShape t = pick_tile(H, nstrips, false);
  const int nseg = (H + t.S - 1) / t.S;
  const int ngroups = (nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
  const RiderCopy<T> rc = rider_or_none(rider);
  const int tiles = ngroups * nseg;
  const InvLaunch<T> L{ll, llp, det, dp, dst, dstp, W, H, t.S, nstrips, nseg,
                       dim3(tiles + rc.blocks), dim3(kWave * kWavesPerBlock), cf, st, tiles, rc};
  with_variant<T>(cf.nsteps, aligned_pair<T>(dst, dstp), t, L);
}

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

template <typename T>
void launch_copy_strided(const T* src, long long sr, long long sc, T* dst, long long dr,
                         long long dc, int W, int H, hipStream_t st) {
  const long long n = (long long)W * H;
  if (n <= 0) return;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  // order by the strided side's shorter stride
  const long long r = (sc == 1 && sr > 0) ? dr : sr, c = (sc == 1 && sr > 0) ? dc : sc;
  const bool xfast = (c < 0 ? -c : c) <= (r < 0 ? -r : r);
  hipLaunchKernelGGL((copy_strided_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, src, sr, sc,
                     dst, dr, dc, W, H, xfast);
}

template void launch_copy_strided<float>(const float*, long long, long long, float*, long long,
                                         long long, int, int, hipStream_t);
template void launch_copy_strided<double>(const double*, long long, long long, double*, long long,
                                          long long, int, int, hipStream_t);

template <typename T>
void launch_channels(bool gather, T* img, long long sr, long long sx, long long sc, int C, int W, int H,
                     T* planes, long long pp, long long plane_step, hipStream_t st) {
  const long long n = (long long)W * H;
  if (n <= 0 || C <= 0) return;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (gather)
    hipLaunchKernelGGL((channels_kernel<T, true>), dim3((unsigned)blocks), dim3(256), 0, st, img, sr, sx, sc,
                       C, W, H, planes, pp, plane_step);
  else
    hipLaunchKernelGGL((channels_kernel<T, false>), dim3((unsigned)blocks), dim3(256), 0, st, img, sr, sx,
                       sc, C, W, H, planes, pp, plane_step);
}

template void launch_channels<float>(bool, float*, long long, long long, long long, int, int, int, float*,
                                     long long, long long, hipStream_t);
template void launch_channels<double>(bool, double*, long long, long long, long long, int, int, int,
                                      double*, long long, long long, hipStream_t);
template void launch_fwd_level<float>(const float*, long long, int, int, float*, long long,
                                      float*, long long, const Coeffs<float>&, hipStream_t,
                                      const RiderCopy<float>*);
template void launch_fwd_level<double>(const double*, long long, int, int, double*, long long,
                                       double*, long long, const Coeffs<double>&, hipStream_t,
                                       const RiderCopy<double>*);
template void launch_inv_level<float>(const float*, long long, const float*, long long, float*,
                                      long long, int, int, const Coeffs<float>&, hipStream_t,
                                      const RiderCopy<float>*);
template void launch_inv_level<double>(const double*, long long, const double*, long long,
                                       double*, long long, int, int, const Coeffs<double>&,
                                       hipStream_t, const RiderCopy<double>*);
template void launch_copy_rect<float>(const float*, long long, float*, long long, int, int,
                                      hipStream_t);
template void launch_copy_rect<double>(const double*, long long, double*, long long, int, int,
                                       hipStream_t);

}  // namespace dwt
