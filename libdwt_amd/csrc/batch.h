// Batched transforms: device tables and launchers (batch.hip), host plans
// (capi.hip).
#pragma once

#include "dwt_common.h"

namespace dwt {

// Coefficient storage of a batch (forward output, inverse input).
constexpr int kStoreF32 = 0, kStoreF16 = 1, kStoreBF16 = 2;

// Item flag: forward, this is the image's last level and its LL goes to the
// coefficients (storage type); inverse, this is the image's coarsest level and
// its LL comes from the coefficients.
constexpr int kItemStorageLL = 1;
// Item flag: the FP32 rows it reads (forward) or writes (inverse) allow
// two-element vector access (element-aligned pointer; any pitch).
constexpr int kItemVec = 2;

// One image at one level of a batched launch.  Forward: src = level input
// (FP32, W x H), out = LL output (FP32 scratch, or the coefficients at the
// last level), det = the coefficients (storage type, Mallat positions).
// Inverse: src = LL input (FP32 scratch, or the coefficients at the coarsest
// level), det = the coefficients, out = the W x H output region (FP32).
// Pitches in elements of the buffer's type.  scale: the image's per-subband
// scales (16-bit storage), 3 per level then the final LL's.
struct BatchItem {
  const void* src;
  long long sp;
  void* out;
  long long op;
  void* det;
  long long dp;
  float* scale;
  int W, H, level, levels;
  int nstrips, ngroups, nseg, block0;
  int flags, reserved;
};

// One image of a batch for the max |x| pass and the scale computation.
struct BatchImage {
  const void* src;
  long long sp;
  float* scale;
  int W, H, levels, block0;
};
// The max |x| pass: a block takes maxabs_rows(W) whole rows of one image
// (about 16 pixels per thread per pass; one row, in passes of kMaxAbsPixels,
// when a row is longer), so no index needs a division.
constexpr int kMaxAbsPixels = 256 * 16;
__host__ __device__ inline int maxabs_rows(int W) {
  return W >= kMaxAbsPixels ? 1 : kMaxAbsPixels / (W > 0 ? W : 1);
}

void launch_batch_level(bool fwd, int storage, int S, int nblocks, const BatchItem* items,
                        const int* block_item, const Coeffs<float>& cf, hipStream_t st);
void launch_batch_maxabs(const BatchImage* imgs, const int* block_img, int nblocks, unsigned* amax,
                         hipStream_t st);
// gain: 3 per level (HL, LH, HH bound factors), then at 96 + J the bound of
// the LL after J levels.
void launch_batch_scales(const BatchImage* imgs, int n, const unsigned* amax, const float* gain,
                         hipStream_t st);
void analysis_gains(const Coeffs<double>& cf, double* gl, double* gh);

}  // namespace dwt
