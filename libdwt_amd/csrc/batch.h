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
// scales (16-bit storage), 3 per level then the final LL's.  band: 0, or the
// item is large: its nbp workgroups fill whole rounds of the launch order from
// block0 and are dealt to the XCDs in bands of `band` (batch_item_remap).
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
  int flags, band, nbp, reserved;
};

// The small levels of one image (levels `level` .. levels - 1, a W x H region
// of at most kBatchTailMax bytes), all in one workgroup's LDS.  Forward: src =
// the level's input (FP32: the image, or LL scratch), dst = the coefficients
// (storage type); inverse: src = the coefficients, dst = the FP32 output (the
// image, or LL scratch for the next finer level).  Pitches in elements.
struct BatchTail {
  const void* src;
  long long sp;
  void* dst;
  long long dp;
  float* scale;
  int W, H, level, nlev, levels, reserved;
};
// Largest tail region (bytes of FP32 samples); its LDS buffer (odd pitch)
// plus the scales stays within one workgroup's 160 KiB.
constexpr long long kBatchTailMax = 128 * 1024;

// Tiles of an item to its workgroups of four waves.  DWT_BATCH_PACK 1: the
// item's nstrips x nseg tiles, numbered row-major, four to a workgroup across
// segment boundaries, so a narrow image (one or two strips) fills its
// workgroups' waves; 0: each workgroup takes four strips of one segment.
#ifndef DWT_BATCH_PACK
#define DWT_BATCH_PACK 1
#endif
__host__ __device__ inline int batch_item_blocks(int nstrips, int nseg) {
#if DWT_BATCH_PACK
  return (nstrips * nseg + 3) / 4;
#else
  return ((nstrips + 3) / 4) * nseg;
#endif
}

// Workgroup order of the batched level kernels.  The hardware deals
// consecutive workgroups to the 8 XCDs round-robin.  Over the launch, each XCD
// takes kBatchBand consecutive logical workgroups in turn (the tail of a grid
// that is not a whole number of rounds keeps the plain order).  Measured on
// 256 uneven images against plain order (1), 4, 16, 32, 64 and 128 and one
// contiguous run per XCD: 8 was fastest (FP32 step 0.645 vs 0.680 ms).
constexpr int kBatchBand = 8;
constexpr int kBatchRound = 8 * kBatchBand;  // workgroups the order permutes among
__host__ __device__ inline int batch_band_remap(int b, int nb) {
  const int full = (nb / kBatchRound) * kBatchRound;
  if (b >= full) return b;
  const int x = b & 7, idx = b >> 3;
  return (x + 8 * (idx / kBatchBand)) * kBatchBand + idx % kBatchBand;
}
// A large item (at least kBatchSegBandMin workgroups) instead starts on a
// whole round and fills whole rounds (padding workgroups before it map to -1,
// those after it past its tiles), so the launch order above never mixes it
// with other items, and workgroup block0 + l runs on XCD l % 8.  Its own
// logical workgroups (row-major tiles) are cut into bands of about
// kBatchBandSegs segment rows (`band` workgroups), dealt to the XCDs in turn:
// vertically neighbouring segments, which share halo rows, mostly run on one
// XCD's L2 while all XCDs read nearby rows.  The workgroups past the last
// whole round of bands take contiguous runs per XCD.  Measured: the batch
// kernels on one 4096^2 image 0.140 -> 0.127 ms per fwd+inv (level 0 as fast
// as the single-image kernel), four 2048^2 images 0.126 -> 0.120 ms; at 1024^2
// and below (288 workgroups and fewer) the launch order above was as fast or
// faster, and per-item bands of 8 workgroups were slower.
#ifndef DWT_BATCH_BAND_SEGS
#define DWT_BATCH_BAND_SEGS 16
#endif
#ifndef DWT_BATCH_SEGBAND_MIN
#define DWT_BATCH_SEGBAND_MIN 1024
#endif
constexpr int kBatchBandSegs = DWT_BATCH_BAND_SEGS;      // segment rows per band
constexpr int kBatchSegBandMin = DWT_BATCH_SEGBAND_MIN;  // workgroups of a large item
__host__ __device__ inline int batch_item_remap(int l, int nbp, int band) {
  const int round = 8 * band, full = nbp / round * round;
  const int x = l & 7, r = l >> 3;
  if (l < full) return (x + 8 * (r / band)) * band + r % band;
  const int per = (nbp - full) >> 3;  // nbp and full are multiples of 8
  return full + x * per + (r - full / 8);
}

// Tile (segment s, strip) of wave `wave` of logical workgroup `local` of an
// item (DWT_BATCH_PACK dealing); false when the wave has none.
__host__ __device__ inline bool batch_tile_of(int local, int wave, const BatchItem& e, int& s, int& strip) {
#if DWT_BATCH_PACK
  const int t = local * 4 + wave;
  s = t / e.nstrips;
  strip = t - s * e.nstrips;
  return s < e.nseg;
#else
  s = local / e.ngroups;
  strip = (local - s * e.ngroups) * 4 + wave;
  return s < e.nseg && strip < e.nstrips;  // padding workgroups of a large item: s >= nseg
#endif
}

// Workgroup b of a batched level launch of nb workgroups -> (item `it`,
// logical workgroup `local` within it): a large item's whole rounds by its own
// bands, everything else by the launch order; false for padding workgroups
// (block_item -1).  The launch order permutes workgroups within whole rounds
// only, so its image of b lies in b's round: in a large item's round it
// finds that item, and the one table read serves both cases.
__host__ __device__ inline bool batch_block_of(int b, int nb, const BatchItem* items, const int* block_item,
                                               int& it, int& local) {
  const int bl = batch_band_remap(b, nb);
  it = block_item[bl];
  if (it < 0) return false;
  const BatchItem& e = items[it];
  local = e.band > 0 ? batch_item_remap(b - e.block0, e.nbp, e.band) : bl - e.block0;
  return true;
}

// One image of a batch for the max |x| pass and the scale computation.
struct BatchImage {
  const void* src;
  long long sp;
  float* scale;
  int W, H, levels, block0;
};
// The max |x| pass.  A contiguous image (pitch == width) is one flat range,
// a block taking kMaxAbsFlat consecutive pixels (eight 16-byte loads per
// thread); otherwise a block takes maxabs_rows(W) whole rows (about 16 pixels
// per thread per pass; one row, in passes of kMaxAbsPixels, when a row is
// longer), so no index needs a division.
constexpr int kMaxAbsPixels = 256 * 16;
constexpr int kMaxAbsFlat = 256 * 32;
__host__ __device__ inline int maxabs_rows(int W) {
  return W >= kMaxAbsPixels ? 1 : kMaxAbsPixels / (W > 0 ? W : 1);
}
__host__ __device__ inline long long maxabs_blocks(int W, int H, long long sp) {
  if (sp == W) return ((long long)W * H + kMaxAbsFlat - 1) / kMaxAbsFlat;
  const int rows = maxabs_rows(W);
  return (H + rows - 1) / rows;
}

void launch_batch_level(bool fwd, int storage, int S, int nblocks, const BatchItem* items,
                        const int* block_item, const Coeffs<float>& cf, hipStream_t st);
void launch_batch_maxabs(const BatchImage* imgs, const int* block_img, int nblocks, unsigned* amax,
                         hipStream_t st);
// gain: 3 per level (HL, LH, HH bound factors), then at 96 + J the bound of
// the LL after J levels.
// unit: every scale 1 (no partials read).
void launch_batch_scales(const BatchImage* imgs, int n, const unsigned* amax, const float* gain, int unit,
                         hipStream_t st);
void analysis_gains(const Coeffs<double>& cf, double* gl, double* gh);
// The tails of n images, one workgroup each (lds: dynamic LDS bytes, enough
// for the largest region and its scales).
void launch_batch_tail(bool fwd, int storage, int n, size_t lds, const BatchTail* tails,
                       const Coeffs<float>& cf, hipStream_t st);
// LDS bytes of a tail region of W x H samples: odd-pitch buffer, then the
// scales (or their inverses), then the line passes' dummy store slots.
size_t batch_tail_lds(int W, int H);

}  // namespace dwt
