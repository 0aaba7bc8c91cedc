// Host-side launch interface of the device kernels.
#pragma once

#include "dwt_common.h"

namespace dwt {

// ---- generic line kernels (line_kernels.hip) --------------------------------
template <typename T>
void launch_line_pass(bool fwd, const char* a, const char* b, char* c, long long line_step,
                      long long es, int nlines, int N, const Coeffs<T>& cf, T* gtmp,
                      int gtmp_blocks, hipStream_t st);
template <typename T>
void launch_zero_pad(char* ptr, long long sx, long long sy, int fw, int fh, const int zx[4],
                     const int zy[4], hipStream_t st);
int line_lds_limit_bytes();
hipError_t configure_line_kernels();

// ---- packed fast path (tile_kernels.hip) ------------------------------------
// Forward level on a packed W x H region `src` (row pitch sp elements):
// LL (ceil(W/2) x ceil(H/2)) goes to `ll` (pitch llp); HL, LH, HH go to `det`
// (pitch dp) at their Mallat positions.  src must not overlap ll or det.
template <typename T>
void launch_fwd_level(const T* src, long long sp, int W, int H, T* ll, long long llp, T* det,
                      long long dp, const Coeffs<T>& cf, hipStream_t st);
// Inverse level: LL from `ll` (pitch llp), HL/LH/HH from `det` (pitch dp) at
// their Mallat positions of the W x H fine region; output to `dst` (pitch
// dstp).  dst must not overlap ll or the details read from det.
template <typename T>
void launch_inv_level(const T* ll, long long llp, const T* det, long long dp, T* dst,
                      long long dstp, int W, int H, const Coeffs<T>& cf, hipStream_t st);
// Tile-shape table of the level kernels, by direction (1 forward, 0 inverse)
// and height class ceil(log2(H)): S rows per segment, U prefetched row pairs
// (S = 0 restores the built-in heuristic).  tile_get reports the shape a
// launch of that size would use.
int tile_set(int fwd, int hclass, int S, int U);
int tile_get(int fwd, int H, int W, int nsteps, int* S, int* U);
// Block kernels: a whole small pyramid in one workgroup's LDS.  Forward:
// levels 0..nlev-1 of the W0 x H0 region read from src, written in Mallat
// layout to dst.  Inverse: the nlev levels whose finest output is W0 x H0,
// coefficients from src, output to dst.  src may equal dst (the region is read
// completely before it is written).  tail_start picks the first level of an
// L-level transform that the block kernel takes (L when none).
template <typename T>
void launch_block_fwd(const T* src, long long sp, T* dst, long long dp, int W0, int H0, int nlev,
                      const Coeffs<T>& cf, hipStream_t st);
template <typename T>
void launch_block_inv(const T* src, long long sp, T* dst, long long dp, int W0, int H0, int nlev,
                      const Coeffs<T>& cf, hipStream_t st);
int tail_start(int W, int H, int L, int elem);
void tail_config(long long max_region_bytes, int unused);
hipError_t configure_tail_kernels();
// Copy a w x h rectangle between packed buffers.
template <typename T>
void launch_copy_rect(const T* src, long long sp, T* dst, long long dp, int w, int h,
                      hipStream_t st);

}  // namespace dwt
