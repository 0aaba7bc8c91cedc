// Host-side tables of a batched transform, shared by plans and the serving
// loop (batch_host.hip, server.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/libdwt.h"
#include "batch.h"

namespace dwt {

void batch_check_kind(int wavelet, int direction, int storage);
// Validate image i of a batch and clamp its levels (forward).
void batch_check_image(int i, dwt_gpu_image& im, bool fwd, int storage);
// Scratch of one image: its ping-pong LL buffers (B1 then B2).
size_t image_scratch_bytes(const dwt_gpu_image& im);
// First level of image (W x H, J levels) that the batched tail launch takes:
// the first whose region fits the tail threshold (dwt_gpu_batch_tail_set) and
// the block kernels' line and LDS limits; J when none does.
int batch_tail_start(int W, int H, int J);

struct BatchLevel {
  int level, items_off, nblocks, blocks_off, S;
};

// Per level (in execution order) the items of every image at that level and
// the block -> item map; for 16-bit storage also the images of the max |x|
// pass and the gain bounds, and the tails (every image's small levels, one
// launch).  Device layout: items | block map | images | image map | gains |
// amax | tails.
struct BatchTables {
  bool fwd = true;
  int wavelet = 0, storage = 0, n = 0, Lmax = 0;
  std::vector<BatchItem> items;
  std::vector<int> bmap;
  std::vector<BatchLevel> levels;  // in execution order
  std::vector<BatchTail> tails;    // the small levels, one workgroup per image
  size_t tail_lds = 0;
  std::vector<BatchImage> imgs;    // max |x| pass (forward, 16-bit storage)
  std::vector<int> imap;
  std::vector<float> gain;
  // device layout: items | block map | images | image map | gains | amax | tails
  size_t off_items = 0, off_bmap = 0, off_imgs = 0, off_imap = 0, off_gain = 0, off_amax = 0,
         off_tails = 0, bytes = 0;

  // images: validated; scratch[i]: image i's scratch (image_scratch_bytes)
  void build(int wavelet, bool fwd, int storage, const dwt_gpu_image* images, int n,
             char* const* scratch);
  void pack(char* host) const;  // into a buffer of `bytes` bytes
  void launch(char* dev, hipStream_t st) const;  // tables packed at dev
};

}  // namespace dwt
