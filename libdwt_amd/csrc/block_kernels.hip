// Block kernels: the small levels of a transform in one workgroup, the whole
// region held in LDS (the lifting and copies are in block_body.h).
#include "block_body.h"

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
  block_copy<true>(C, pc, const_cast<T*>(src), sp, W0, H0);
  __syncthreads();
  for (int lev = 0; lev < nlev; ++lev) {
    const int W = ceil_div_pow2(W0, lev), H = ceil_div_pow2(H0, lev);
    block_pass_any<true, T, NS>(C, pc, 1, H, W, cf);  // rows
    __syncthreads();
    block_pass_any<true, T, NS>(C, 1, pc, W, H, cf);  // columns
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
  block_copy<true>(C, pc, const_cast<T*>(src), sp, W0, H0);
  __syncthreads();
  for (int lev = nlev - 1; lev >= 0; --lev) {
    const int W = ceil_div_pow2(W0, lev), H = ceil_div_pow2(H0, lev);
    block_pass_any<false, T, NS>(C, pc, 1, H, W, cf);  // rows
    __syncthreads();
    block_pass_any<false, T, NS>(C, 1, pc, W, H, cf);  // columns
    __syncthreads();
  }
  block_copy<false>(C, pc, dst, dp, W0, H0);
}

// ---- host side ----------------------------------------------------------------
// The block kernel takes over from the first level j >= 1 whose W x H region
// is at most g_block_bytes, or from level 0 when the whole image is at most
// twice that (one launch then replaces every launch of the transform), as long
// as the region's rows and columns fit a wavefront's registers (<= 512
// samples) and the region fits one workgroup's LDS.
static long long g_block_bytes = 64 * 1024;
constexpr long long kBlockRegionMax = 76 * 1024;  // twice this (whole images) stays under the LDS limit


void tail_config(long long max_region_bytes, int) {
  g_block_bytes = max_region_bytes > kBlockRegionMax ? kBlockRegionMax
                  : max_region_bytes < 0             ? 0
                                                     : max_region_bytes;
}

static long long block_lds_bytes(long long W0, long long H0, int elem) {
  return elem * (long long)block_pitch((int)W0) * H0;
}

int tail_start(int W, int H, int L, int elem) {
  if (g_block_bytes <= 0) return L;
  for (int j = 0; j < L; ++j) {
    const long long wj = ceil_div_pow2(W, j), hj = ceil_div_pow2(H, j);
    const long long limit = j == 0 ? 2 * g_block_bytes : g_block_bytes;
    if (wj * hj * elem <= limit && wj <= kBlockMaxLine && hj <= kBlockMaxLine &&
        block_lds_bytes(wj, hj, elem) <= kBlockLdsMax)
      return j;
  }
  return L;
}

hipError_t configure_tail_kernels() {
  const void* fns[] = {reinterpret_cast<const void*>(&block_fwd_kernel<float, 4>),
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

template void launch_block_fwd<float>(const float*, long long, float*, long long, int, int, int,
                                      const Coeffs<float>&, hipStream_t, const RiderCopy<float>*);
template void launch_block_fwd<double>(const double*, long long, double*, long long, int, int, int,
                                       const Coeffs<double>&, hipStream_t, const RiderCopy<double>*);
template void launch_block_inv<float>(const float*, long long, float*, long long, int, int, int,
                                      const Coeffs<float>&, hipStream_t, const RiderCopy<float>*);
template void launch_block_inv<double>(const double*, long long, double*, long long, int, int, int,
                                       const Coeffs<double>&, hipStream_t, const RiderCopy<double>*);

}  // namespace dwt
