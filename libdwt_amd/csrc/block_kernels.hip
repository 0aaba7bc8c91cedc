// Block kernels: the small levels of a transform in one workgroup, the whole
// region held in LDS (the lifting and copies are in block_body.h).
#include "block_body.h"

#include <algorithm>

namespace dwt {

// Workgroup 0 transforms the region; any further workgroups move rows of a
// level-0 detail copy riding on the launch (RiderCopy, kernels.h), on the
// compute units the single transforming workgroup leaves idle.
template <typename T, int NS>
__global__ void __launch_bounds__(kBlockThreads)
block_fwd_kernel(const T* __restrict__ src, long long sp, T* __restrict__ dst, long long dp, int W0,
                 int H0, int nlev, Coeffs<T> cf, RiderCopy<T> rc) {
  if (blockIdx.x > 0) {
    rider_copy(rc, blockIdx.x - 1, gridDim.x - 1);
    return;
  }
  extern __shared__ __align__(16) unsigned char blk_smem[];
  Lds<T>* C = (Lds<T>*)blk_smem;  // generic -> LDS address-space cast
  const int pc = block_pitch(W0);
  Lds<T>* const dummy = C + pc * H0;
  block_copy<true>(C, pc, const_cast<T*>(src), sp, W0, H0);
  __syncthreads();
  for (int lev = 0; lev < nlev; ++lev) {
    const int W = ceil_div_pow2(W0, lev), H = ceil_div_pow2(H0, lev);
    block_pass_any<true, T, NS>(C, pc, 1, H, W, cf, dummy);  // rows
    __syncthreads();
    block_pass_any<true, T, NS>(C, 1, pc, W, H, cf, dummy);  // columns
    __syncthreads();
  }
  block_copy<false>(C, pc, dst, dp, W0, H0);
}

template <typename T, int NS>
__global__ void __launch_bounds__(kBlockThreads)
block_inv_kernel(const T* __restrict__ src, long long sp, T* __restrict__ dst, long long dp, int W0,
                 int H0, int nlev, Coeffs<T> cf, RiderCopy<T> rc) {
  if (blockIdx.x > 0) {
    rider_copy(rc, blockIdx.x - 1, gridDim.x - 1);
    return;
  }
  extern __shared__ __align__(16) unsigned char blk_smem[];
  Lds<T>* C = (Lds<T>*)blk_smem;  // generic -> LDS address-space cast
  const int pc = block_pitch(W0);
  Lds<T>* const dummy = C + pc * H0;
  block_copy<true>(C, pc, const_cast<T*>(src), sp, W0, H0);
  __syncthreads();
  for (int lev = nlev - 1; lev >= 0; --lev) {
    const int W = ceil_div_pow2(W0, lev), H = ceil_div_pow2(H0, lev);
    block_pass_any<false, T, NS>(C, pc, 1, H, W, cf, dummy);  // rows
    __syncthreads();
    block_pass_any<false, T, NS>(C, 1, pc, W, H, cf, dummy);  // columns
    __syncthreads();
  }
  block_copy<false>(C, pc, dst, dp, W0, H0);
}

// Sparse frames (outer W0 x H0, inner Wi0 x Hi0 at the region's first level;
// src/libdwt.c:6331-6400, 6683-6747): every level lifts the outer frame's rows
// over the inner length with H at the outer half, then its columns likewise,
// garbage rows and columns included, exactly as the reference's line loops
// do; optional zero padding after each level.  In place on one region.
template <typename T, int NS>
__global__ void __launch_bounds__(kBlockThreads)
block_sparse_fwd_kernel(T* __restrict__ a, long long ap, int W0, int H0, int Wi0, int Hi0, int nlev, int zp,
                        Coeffs<T> cf) {
  extern __shared__ __align__(16) unsigned char blk_smem[];
  Lds<T>* C = (Lds<T>*)blk_smem;  // generic -> LDS address-space cast
  const int pc = block_pitch(W0);
  Lds<T>* const dummy = C + pc * H0;
  block_copy<true>(C, pc, a, ap, W0, H0);
  __syncthreads();
  for (int lev = 0; lev < nlev; ++lev) {
    const int W = ceil_div_pow2(W0, lev), H = ceil_div_pow2(H0, lev);
    const int Wi = ceil_div_pow2(Wi0, lev), Hi = ceil_div_pow2(Hi0, lev);
    block_pass_any<true, T, NS>(C, pc, 1, H, Wi, cf, dummy, (W + 1) >> 1);  // rows y < H, length Wi
    __syncthreads();
    block_pass_any<true, T, NS>(C, 1, pc, W, Hi, cf, dummy, (H + 1) >> 1);  // columns x < W, length Hi
    __syncthreads();
    if (zp) {
      block_zero_pad<true, T>(C, pc, W, H, Wi, Hi);
      __syncthreads();
    }
  }
  block_copy<false>(C, pc, a, ap, W0, H0);
}

template <typename T, int NS>
__global__ void __launch_bounds__(kBlockThreads)
block_sparse_inv_kernel(T* __restrict__ a, long long ap, int W0, int H0, int Wi0, int Hi0, int nlev, int zp,
                        Coeffs<T> cf) {
  extern __shared__ __align__(16) unsigned char blk_smem[];
  Lds<T>* C = (Lds<T>*)blk_smem;  // generic -> LDS address-space cast
  const int pc = block_pitch(W0);
  Lds<T>* const dummy = C + pc * H0;
  block_copy<true>(C, pc, a, ap, W0, H0);
  __syncthreads();
  for (int lev = nlev - 1; lev >= 0; --lev) {
    // output level lev: outer W x H, inner Wi x Hi; H parts at the coarser
    // level's outer size
    const int W = ceil_div_pow2(W0, lev), H = ceil_div_pow2(H0, lev);
    const int Wi = ceil_div_pow2(Wi0, lev), Hi = ceil_div_pow2(Hi0, lev);
    block_pass_any<false, T, NS>(C, pc, 1, H, Wi, cf, dummy, (W + 1) >> 1);  // rows y < H
    __syncthreads();
    block_pass_any<false, T, NS>(C, 1, pc, W, Hi, cf, dummy, (H + 1) >> 1);  // columns x < W
    __syncthreads();
    if (zp) {
      block_zero_pad<false, T>(C, pc, W, H, Wi, Hi);
      __syncthreads();
    }
  }
  block_copy<false>(C, pc, a, ap, W0, H0);
}

// ---- host side ----------------------------------------------------------------
// The block kernel takes over from the first level j >= 1 whose W x H region
// is at most g_block_bytes, or from level 0 when the whole image is at most
// twice that (one launch then replaces every launch of the transform), as long
// as the region's rows and columns fit a wavefront's registers (<= 512
// samples) and the region fits one workgroup's LDS.
static long long g_block_bytes = 64 * 1024;
constexpr long long kBlockRegionMax = 150 * 1024;  // a region plus its LDS padding stays under 160 KiB
// the whole-image limit: twice the threshold, at most the largest region
static long long whole_image_limit() { return std::min(2 * g_block_bytes, kBlockRegionMax); }


void tail_config(long long max_region_bytes, int) {
  g_block_bytes = max_region_bytes > kBlockRegionMax ? kBlockRegionMax
                  : max_region_bytes < 0             ? 0
                                                     : max_region_bytes;
}

static long long block_lds_bytes(long long W0, long long H0, int elem) {
  return elem * ((long long)block_pitch((int)W0) * H0 + kBlockDummy);
}

int tail_start(int W, int H, int L, int elem) {
  if (g_block_bytes <= 0) return L;
  for (int j = 0; j < L; ++j) {
    const long long wj = ceil_div_pow2(W, j), hj = ceil_div_pow2(H, j);
    const long long limit = j == 0 ? whole_image_limit() : g_block_bytes;
    if (wj * hj * elem <= limit && wj <= kBlockMaxLine && hj <= kBlockMaxLine &&
        block_lds_bytes(wj, hj, elem) <= kBlockLdsMax)
      return j;
  }
  return L;
}

hipError_t configure_tail_kernels() {
  const void* fns[] = {reinterpret_cast<const void*>(&block_sparse_fwd_kernel<float, 4>),
                       reinterpret_cast<const void*>(&block_sparse_fwd_kernel<float, 2>),
                       reinterpret_cast<const void*>(&block_sparse_fwd_kernel<double, 4>),
                       reinterpret_cast<const void*>(&block_sparse_fwd_kernel<double, 2>),
                       reinterpret_cast<const void*>(&block_sparse_inv_kernel<float, 4>),
                       reinterpret_cast<const void*>(&block_sparse_inv_kernel<float, 2>),
                       reinterpret_cast<const void*>(&block_sparse_inv_kernel<double, 4>),
                       reinterpret_cast<const void*>(&block_sparse_inv_kernel<double, 2>),
                       reinterpret_cast<const void*>(&block_fwd_kernel<float, 4>),
                       reinterpret_cast<const void*>(&block_fwd_kernel<float, 2>),
                       reinterpret_cast<const void*>(&block_fwd_kernel<double, 4>),
                       reinterpret_cast<const void*>(&block_fwd_kernel<double, 2>),
                       reinterpret_cast<const void*>(&block_inv_kernel<float, 4>),
                       reinterpret_cast<const void*>(&block_inv_kernel<float, 2>),
                       reinterpret_cast<const void*>(&block_inv_kernel<double, 4>),
                       reinterpret_cast<const void*>(&block_inv_kernel<double, 2>)};
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
static RiderCopy<T> block_rider(const RiderCopy<T>* r) {
  if (r && r->r1 > r->r0 && r->blocks > 0) return *r;
  RiderCopy<T> none{};
  return none;
}

template <typename T>
void launch_block_fwd(const T* src, long long sp, T* dst, long long dp, int W0, int H0, int nlev,
                      const Coeffs<T>& cf, hipStream_t st, const RiderCopy<T>* rider) {
  const size_t lds = (size_t)block_lds_bytes(W0, H0, (int)sizeof(T));
  const RiderCopy<T> rc = block_rider(rider);
  const dim3 grid(1 + rc.blocks);
  if (cf.nsteps == 4)
    hipLaunchKernelGGL((block_fwd_kernel<T, 4>), grid, dim3(kBlockThreads), lds, st, src, sp, dst, dp, W0,
                       H0, nlev, cf, rc);
  else
    hipLaunchKernelGGL((block_fwd_kernel<T, 2>), grid, dim3(kBlockThreads), lds, st, src, sp, dst, dp, W0,
                       H0, nlev, cf, rc);
}

template <typename T>
void launch_block_inv(const T* src, long long sp, T* dst, long long dp, int W0, int H0, int nlev,
                      const Coeffs<T>& cf, hipStream_t st, const RiderCopy<T>* rider) {
  const size_t lds = (size_t)block_lds_bytes(W0, H0, (int)sizeof(T));
  const RiderCopy<T> rc = block_rider(rider);
  const dim3 grid(1 + rc.blocks);
  if (cf.nsteps == 4)
    hipLaunchKernelGGL((block_inv_kernel<T, 4>), grid, dim3(kBlockThreads), lds, st, src, sp, dst, dp, W0,
                       H0, nlev, cf, rc);
  else
    hipLaunchKernelGGL((block_inv_kernel<T, 2>), grid, dim3(kBlockThreads), lds, st, src, sp, dst, dp, W0,
                       H0, nlev, cf, rc);
}

// Sparse tail: the first level j >= 1 (or 0 for a whole small frame) whose
// outer region fits the block kernel (the packed thresholds), and from which
// every remaining level has inner sides of at least 2 (lines of one sample
// take the line kernel's N == 1 path); L when there is none.
int sparse_tail_start(int Wo, int Ho, int Wi, int Hi, int L, int elem) {
  if (g_block_bytes <= 0) return L;
  int jt = L;
  for (int j = L - 1; j >= 0; --j) {
    if (ceil_div_pow2(Wi, j) < 2 || ceil_div_pow2(Hi, j) < 2) break;
    const long long wj = ceil_div_pow2(Wo, j), hj = ceil_div_pow2(Ho, j);
    const long long limit = j == 0 ? whole_image_limit() : g_block_bytes;
    if (wj * hj * elem <= limit && wj <= kBlockMaxLine && hj <= kBlockMaxLine &&
        block_lds_bytes(wj, hj, elem) <= kBlockLdsMax)
      jt = j;
    else
      break;
  }
  return jt;
}

template <typename T>
void launch_block_sparse(bool fwd, T* a, long long ap, int W0, int H0, int Wi0, int Hi0, int nlev, int zp,
                         const Coeffs<T>& cf, hipStream_t st) {
  const size_t lds = (size_t)block_lds_bytes(W0, H0, (int)sizeof(T));
  const dim3 grid(1), block(kBlockThreads);
  if (fwd) {
    if (cf.nsteps == 4)
      hipLaunchKernelGGL((block_sparse_fwd_kernel<T, 4>), grid, block, lds, st, a, ap, W0, H0, Wi0, Hi0, nlev, zp, cf);
    else
      hipLaunchKernelGGL((block_sparse_fwd_kernel<T, 2>), grid, block, lds, st, a, ap, W0, H0, Wi0, Hi0, nlev, zp, cf);
  } else {
    if (cf.nsteps == 4)
      hipLaunchKernelGGL((block_sparse_inv_kernel<T, 4>), grid, block, lds, st, a, ap, W0, H0, Wi0, Hi0, nlev, zp, cf);
    else
      hipLaunchKernelGGL((block_sparse_inv_kernel<T, 2>), grid, block, lds, st, a, ap, W0, H0, Wi0, Hi0, nlev, zp, cf);
  }
}
template void launch_block_sparse<float>(bool, float*, long long, int, int, int, int, int, int, const Coeffs<float>&,
                                         hipStream_t);
template void launch_block_sparse<double>(bool, double*, long long, int, int, int, int, int, int,
                                          const Coeffs<double>&, hipStream_t);

template void launch_block_fwd<float>(const float*, long long, float*, long long, int, int, int,
                                      const Coeffs<float>&, hipStream_t, const RiderCopy<float>*);
template void launch_block_fwd<double>(const double*, long long, double*, long long, int, int, int,
                                       const Coeffs<double>&, hipStream_t, const RiderCopy<double>*);
template void launch_block_inv<float>(const float*, long long, float*, long long, int, int, int,
                                      const Coeffs<float>&, hipStream_t, const RiderCopy<float>*);
template void launch_block_inv<double>(const double*, long long, double*, long long, int, int, int,
                                       const Coeffs<double>&, hipStream_t, const RiderCopy<double>*);

}  // namespace dwt
