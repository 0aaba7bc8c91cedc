// Host side of the batched transforms: plans (tables built once, executed
// any number of times) and their C entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "batch.h"
#include "block_body.h"
#include "batch_tables.h"
#include "host_common.h"
#include "kernels.h"

namespace dwt {

// ---- batched transforms ------------------------------------------------------------
// A batch is described by its tables: per level (in execution order) the
// items of every image at that level and the block -> item map, plus, for
// 16-bit storage, the images of the max |x| pass and the gain bounds.  Each
// image also needs its ping-pong scratch (LL of odd levels in B1, of even
// levels >= 2 in B2).  A plan builds the tables once and keeps them on the
// device; executing it only launches kernels, so it can be captured in a HIP
// graph and replayed.  The serving loop builds tables per admitted batch.
static int g_batch_rows[32] = {};  // segment rows per level (0: 8)
static long long g_batch_tail_bytes = 64 * 1024;  // tail regions up to this size (0: no tail)
// Bumped by every change of the two settings above: a cached plan built under
// other settings is not reused.
static std::atomic<int> g_batch_config_gen{0};

int batch_tail_start(int W, int H, int J) {
  if (g_batch_tail_bytes <= 0) return J;
  for (int j = 0; j < J; ++j) {
    const int w = ceil_div_pow2(W, j), h = ceil_div_pow2(H, j);
    if (4LL * w * h <= g_batch_tail_bytes && w <= kBlockMaxLine && h <= kBlockMaxLine &&
        batch_tail_lds(w, h) <= (size_t)kBlockLdsMax)
      return j;
  }
  return J;
}

static size_t storage_bytes(int storage) { return storage == kStoreF32 ? 4 : 2; }

// Validate one image of a batch and clamp its levels (forward).
void batch_check_image(int i, dwt_gpu_image& im, bool fwd, int storage) {
  const int W = im.width, H = im.height;
  const size_t cs = storage_bytes(storage);
  if (W < 2 || H < 2) fail(DWT_GPU_EINVAL, "image %d: width and height must be at least 2", i);
  if (!im.src || !im.dst) fail(DWT_GPU_EINVAL, "image %d: null buffer", i);
  if (im.src == im.dst) fail(DWT_GPU_EINVAL, "image %d: batches are out of place (src == dst)", i);
  const int jl = level_limit(W, H, 0);
  if (fwd) {
    if (im.levels < 0 || im.levels > jl) im.levels = jl;
  } else if (im.levels < 1 || im.levels > jl) {
    fail(DWT_GPU_EINVAL, "image %d: %d levels outside 1..%d", i, im.levels, jl);
  }
  if (im.levels < 1) fail(DWT_GPU_EINVAL, "image %d: no level to transform", i);
  if (storage != kStoreF32 && !im.scales) fail(DWT_GPU_EINVAL, "image %d: 16-bit storage needs scales", i);
  const long long ip = fwd ? im.src_pitch : im.dst_pitch;  // FP32 image side
  const long long cp = fwd ? im.dst_pitch : im.src_pitch;  // coefficient side
  const void* ipt = fwd ? im.src : (const void*)im.dst;
  if (ip < 4LL * W || ip % 4 != 0 || (reinterpret_cast<uintptr_t>(ipt) & 3) != 0)
    fail(DWT_GPU_EINVAL, "image %d: the FP32 image needs 4-byte alignment and pitch", i);
  if (cp < (long long)cs * W || cp % (long long)cs != 0)
    fail(DWT_GPU_EINVAL, "image %d: bad coefficient pitch", i);
}

void batch_check_kind(int wavelet, int direction, int storage) {
  if (wavelet != DWT_GPU_CDF97 && wavelet != DWT_GPU_CDF53) fail(DWT_GPU_EINVAL, "bad wavelet");
  if (direction != DWT_GPU_FORWARD && direction != DWT_GPU_INVERSE) fail(DWT_GPU_EINVAL, "bad direction");
  if (storage != kStoreF32 && storage != kStoreF16 && storage != kStoreBF16)
    fail(DWT_GPU_EINVAL, "bad coefficient storage");
}

// Scratch bytes of one image (B1 then B2, 256-byte aligned).
size_t image_scratch_bytes(const dwt_gpu_image& im) {
  const size_t b1 = (4 * pad_pitch(ceil_div_pow2(im.width, 1)) * ceil_div_pow2(im.height, 1) + 255) & ~size_t(255);
  const size_t b2 = im.levels >= 3 ? (4 * pad_pitch(ceil_div_pow2(im.width, 2)) * ceil_div_pow2(im.height, 2) + 255) & ~size_t(255) : 0;
  return b1 + b2;
}

void BatchTables::build(int wavelet_, bool fwd_, int storage_, const dwt_gpu_image* images, int n_,
                      char* const* scratch) {
  fwd = fwd_;
  wavelet = wavelet_;
  storage = storage_;
  n = n_;
  const size_t cs = storage_bytes(storage);
  const Coeffs<float> cf = make_coeffs<float>(wavelet == DWT_GPU_CDF97 ? CDF97 : CDF53, fwd);
  const int core = 2 * 64 - 2 * cf.nsteps;
  Lmax = 0;
  for (int i = 0; i < n; ++i) Lmax = std::max(Lmax, images[i].levels);
  items.clear();
  bmap.clear();
  levels.clear();
  tails.clear();
  tail_lds = 0;
  std::vector<int> jt(n);
  int Lp = 0;  // levels run by per-level launches
  for (int i = 0; i < n; ++i) {
    jt[i] = batch_tail_start(images[i].width, images[i].height, images[i].levels);
    Lp = std::max(Lp, jt[i]);
  }
  auto llbuf = [&](int i, int j) -> std::pair<float*, long long> {  // LL_j, j >= 1
    const dwt_gpu_image& im = images[i];
    const long long p1 = (long long)pad_pitch(ceil_div_pow2(im.width, 1));
    const long long p2 = (long long)pad_pitch(ceil_div_pow2(im.width, 2));
    const size_t b1 = (4 * (size_t)p1 * ceil_div_pow2(im.height, 1) + 255) & ~size_t(255);
    return (j & 1) ? std::make_pair((float*)scratch[i], p1)
                   : std::make_pair((float*)(scratch[i] + b1), p2);
  };
  for (int step = 0; step < Lp; ++step) {
    const int j = fwd ? step : Lp - 1 - step;
    BatchLevel lv{j, (int)items.size(), 0, (int)bmap.size(), g_batch_rows[j] ? g_batch_rows[j] : 8};
    int blocks = 0;
    for (int i = 0; i < n; ++i) {
      const dwt_gpu_image& im = images[i];
      const int J = im.levels;
      if (j >= jt[i]) continue;  // this level and the ones after it: the tail
      BatchItem e{};
      const long long ip = (fwd ? im.src_pitch : im.dst_pitch) / 4;
      const long long cp = (fwd ? im.dst_pitch : im.src_pitch) / (long long)cs;
      void* coef = fwd ? im.dst : const_cast<void*>(im.src);
      e.W = ceil_div_pow2(im.width, j);
      e.H = ceil_div_pow2(im.height, j);
      e.level = j;
      e.levels = J;
      e.scale = im.scales;
      e.det = coef;
      e.dp = cp;
      if (fwd) {
        if (j == 0) {
          e.src = im.src;
          e.sp = ip;
        } else {
          e.src = llbuf(i, j).first;
          e.sp = llbuf(i, j).second;
        }
        if (j == J - 1) {
          e.out = coef;
          e.op = cp;
          e.flags = kItemStorageLL;
        } else {
          e.out = llbuf(i, j + 1).first;
          e.op = llbuf(i, j + 1).second;
        }
      } else {
        if (j == J - 1) {
          e.src = coef;
          e.sp = cp;
          e.flags = kItemStorageLL;
        } else {
          e.src = llbuf(i, j + 1).first;
          e.sp = llbuf(i, j + 1).second;
        }
        if (j == 0) {
          e.out = im.dst;
          e.op = ip;
        } else {
          e.out = llbuf(i, j).first;
          e.op = llbuf(i, j).second;
        }
      }
      // vector access on the FP32 rows: a pair moves as one element-aligned
      // access (GPair), so every row of an element-aligned image allows it,
      // odd widths included
      const bool img_vec = (reinterpret_cast<uintptr_t>(fwd ? im.src : (const void*)im.dst) & 3) == 0;
      if (j > 0 || img_vec) e.flags |= kItemVec;
      e.nstrips = (e.W + core - 1) / core;
      e.ngroups = (e.nstrips + 3) / 4;
      e.nseg = (e.H + lv.S - 1) / lv.S;
      int nb = batch_item_blocks(e.nstrips, e.nseg);
      if (nb >= kBatchSegBandMin) {
        // a large item: whole rounds of the launch order (padding workgroups
        // before it map to -1; those after its tiles find no tile), dealt to
        // the XCDs in bands of segment rows
        const int pad = (kBatchRound - blocks % kBatchRound) % kBatchRound;
        bmap.insert(bmap.end(), pad, -1);
        blocks += pad;
        e.band = batch_item_blocks(e.nstrips, kBatchBandSegs);
        nb = (nb + kBatchRound - 1) / kBatchRound * kBatchRound;
      }
      e.nbp = nb;
      e.block0 = blocks;
      const int idx = (int)(items.size() - lv.items_off);
      bmap.insert(bmap.end(), nb, idx);
      blocks += nb;
      items.push_back(e);
    }
    lv.nblocks = blocks;
    levels.push_back(lv);
  }
  for (int i = 0; i < n; ++i) {
    const dwt_gpu_image& im = images[i];
    const int j = jt[i], J = im.levels;
    if (j >= J) continue;
    const long long ip = (fwd ? im.src_pitch : im.dst_pitch) / 4;
    const long long cp = (fwd ? im.dst_pitch : im.src_pitch) / (long long)cs;
    BatchTail t{};
    t.W = ceil_div_pow2(im.width, j);
    t.H = ceil_div_pow2(im.height, j);
    t.level = j;
    t.nlev = J - j;
    t.levels = J;
    t.scale = im.scales;
    // the FP32 side: the image at level 0, else LL_j in the ping-pong scratch
    void* fp = j == 0 ? (fwd ? const_cast<void*>(im.src) : im.dst) : (void*)llbuf(i, j).first;
    const long long fpp = j == 0 ? ip : llbuf(i, j).second;
    void* coef = fwd ? im.dst : const_cast<void*>(im.src);
    t.src = fwd ? fp : coef;
    t.sp = fwd ? fpp : cp;
    t.dst = fwd ? coef : fp;
    t.dp = fwd ? cp : fpp;
    tails.push_back(t);
    tail_lds = std::max(tail_lds, batch_tail_lds(t.W, t.H));
  }
  imgs.clear();
  imap.clear();
  gain.assign(96 + 33, 1.0f);
  if (fwd && storage != kStoreF32) {
    // the max |x| pass reads the images last to first, so that the ones it
    // read last -- still in the 256 MB Infinity Cache -- are the ones the
    // level-0 launch (image-major, first image first) reads first
    imgs.resize(n);
    for (int i = n - 1; i >= 0; --i) {
      const dwt_gpu_image& im = images[i];
      imgs[i] = BatchImage{im.src, im.src_pitch / 4, im.scales, im.width, im.height, im.levels, (int)imap.size()};
      // BF16 needs no max |x| (its scales are 1, batch.hip)
      const long long nb = storage == kStoreBF16 ? 0 : maxabs_blocks(im.width, im.height, im.src_pitch / 4);
      imap.insert(imap.end(), (size_t)nb, i);
    }
    double gl = 0, gh = 0;
    analysis_gains(make_coeffs<double>(wavelet == DWT_GPU_CDF97 ? CDF97 : CDF53, true), &gl, &gh);
    for (int j = 0; j < 32; ++j) {
      const double in = std::pow(gl * gl, j);  // bound of LL_j relative to max |x|
      gain[3 * j + 0] = (float)(in * gh * gl);
      gain[3 * j + 1] = (float)(in * gl * gh);
      gain[3 * j + 2] = (float)(in * gh * gh);
    }
    for (int j = 0; j <= 32; ++j) gain[96 + j] = (float)std::pow(gl * gl, j);
  }
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  off_items = 0;
  off_bmap = al(sizeof(BatchItem) * items.size());
  off_imgs = off_bmap + al(sizeof(int) * bmap.size());
  off_imap = off_imgs + al(sizeof(BatchImage) * imgs.size());
  off_gain = off_imap + al(sizeof(int) * imap.size());
  off_amax = off_gain + al(sizeof(float) * gain.size());
  off_tails = off_amax + al(sizeof(unsigned) * (imap.empty() ? (size_t)1 : imap.size()));
  bytes = off_tails + al(sizeof(BatchTail) * (tails.empty() ? (size_t)1 : tails.size()));
}

void BatchTables::pack(char* hp) const {
  std::memcpy(hp + off_items, items.data(), sizeof(BatchItem) * items.size());
  std::memcpy(hp + off_bmap, bmap.data(), sizeof(int) * bmap.size());
  if (!imgs.empty()) std::memcpy(hp + off_imgs, imgs.data(), sizeof(BatchImage) * imgs.size());
  if (!imap.empty()) std::memcpy(hp + off_imap, imap.data(), sizeof(int) * imap.size());
  std::memcpy(hp + off_gain, gain.data(), sizeof(float) * gain.size());
  if (!tails.empty()) std::memcpy(hp + off_tails, tails.data(), sizeof(BatchTail) * tails.size());
}

void BatchTables::launch(char* dev, hipStream_t st) const {
  const Coeffs<float> cf = make_coeffs<float>(wavelet == DWT_GPU_CDF97 ? CDF97 : CDF53, fwd);
  if (fwd && storage != kStoreF32) {
    unsigned* amax = reinterpret_cast<unsigned*>(dev + off_amax);
    const BatchImage* im = reinterpret_cast<const BatchImage*>(dev + off_imgs);
    LAUNCH(launch_batch_maxabs(im, reinterpret_cast<const int*>(dev + off_imap), (int)imap.size(), amax, st);
           launch_batch_scales(im, n, amax, reinterpret_cast<const float*>(dev + off_gain),
                               storage == kStoreBF16 ? 1 : 0, st));
  }
  auto tail = [&] {
    if (tails.empty()) return;
    ProfScope ps(fwd ? 12 : 13, -1, st);
    LAUNCH(launch_batch_tail(fwd, storage, (int)tails.size(), tail_lds,
                             reinterpret_cast<const BatchTail*>(dev + off_tails), cf, st));
  };
  if (!fwd) tail();  // the inverse starts at the coarsest levels
  for (const BatchLevel& lv : levels) {
    ProfScope ps(fwd ? 8 : 9, lv.level, st);
    LAUNCH(launch_batch_level(fwd, storage, lv.S, lv.nblocks,
                              reinterpret_cast<const BatchItem*>(dev + off_items) + lv.items_off,
                              reinterpret_cast<const int*>(dev + off_bmap) + lv.blocks_off, cf, st));
  }
  if (fwd) tail();
}

struct BatchPlan {
  std::vector<dwt_gpu_image> desc;  // as given, levels clamped
  int config_gen = 0;               // g_batch_config_gen when built
  BatchTables t;
  void* dev = nullptr;
  void* pinned = nullptr;
  void* scratch = nullptr;
  ~BatchPlan() {
    if (dev) hipFree(dev);
    if (pinned) hipHostFree(pinned);
    if (scratch) hipFree(scratch);
  }
};

static BatchPlan* batch_plan_create(int wavelet, int direction, int storage,
                                    dwt_gpu_image* images, int n, hipStream_t st) {
  init_once();
  batch_check_kind(wavelet, direction, storage);
  if (n <= 0 || !images) fail(DWT_GPU_EINVAL, "empty batch");
  const bool fwd = direction == DWT_GPU_FORWARD;
  std::unique_ptr<BatchPlan> pl(new BatchPlan());
  std::vector<size_t> off(n);
  size_t total = 0;
  for (int i = 0; i < n; ++i) {
    batch_check_image(i, images[i], fwd, storage);
    off[i] = total;
    total += image_scratch_bytes(images[i]);
  }
  pl->desc.assign(images, images + n);
  pl->config_gen = g_batch_config_gen.load();
  if (hipMalloc(&pl->scratch, std::max<size_t>(total, 256)) != hipSuccess) {
    (void)hipGetLastError();
    pl->scratch = nullptr;
    fail(DWT_GPU_EMEMORY, "batch scratch of %zu bytes", total);
  }
  std::vector<char*> base(n);
  for (int i = 0; i < n; ++i) base[i] = static_cast<char*>(pl->scratch) + off[i];
  pl->t.build(wavelet, fwd, storage, images, n, base.data());
  if (hipHostMalloc(&pl->pinned, pl->t.bytes, hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&pl->dev, pl->t.bytes) != hipSuccess) {
    (void)hipGetLastError();
    fail(DWT_GPU_EMEMORY, "batch tables of %zu bytes", pl->t.bytes);
  }
  std::memset(pl->pinned, 0, pl->t.bytes);
  pl->t.pack(static_cast<char*>(pl->pinned));
  HIPCHK(hipMemcpyAsync(pl->dev, pl->pinned, pl->t.bytes, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // a plan may then run on any stream
  return pl.release();
}

static void batch_plan_execute(const BatchPlan& pl, hipStream_t st) {
  pl.t.launch(static_cast<char*>(pl.dev), st);
}

// dwt_gpu_batch_2d keeps the plans of its most recent batches.  A plan is
// held by shared ownership: a caller keeps its reference until its launches
// are enqueued, so another thread that evicts the plan meanwhile only drops
// the cache's reference.  The last owner frees it, after a device
// synchronisation (its tables may still be read by launches on any stream).
using PlanRef = std::shared_ptr<BatchPlan>;
static void release_plan(BatchPlan* p) {
  hipDeviceSynchronize();
  delete p;
}
static std::mutex g_batch_mu;
static std::vector<PlanRef>* g_batch_cache = new std::vector<PlanRef>();
static constexpr size_t kBatchCache = 4;

static PlanRef batch_plan_cached(int wavelet, int direction, int storage, dwt_gpu_image* images, int n,
                                 hipStream_t st) {
  PlanRef evicted;  // released after the lock, by whichever owner is last
  std::lock_guard<std::mutex> lk(g_batch_mu);
  for (size_t i = 0; i < g_batch_cache->size(); ++i) {
    PlanRef p = (*g_batch_cache)[i];
    if (p->t.wavelet != wavelet || p->t.fwd != (direction == DWT_GPU_FORWARD) ||
        p->t.storage != storage || p->t.n != n || p->config_gen != g_batch_config_gen.load())
      continue;
    bool same = true;
    for (int k = 0; k < n && same; ++k) {
      dwt_gpu_image a = images[k];
      const dwt_gpu_image& b = p->desc[k];
      if (p->t.fwd && (a.levels < 0 || a.levels > b.levels)) a.levels = b.levels;  // clamped alike
      same = std::memcmp(&a, &b, sizeof a) == 0;
    }
    if (!same) continue;
    for (int k = 0; k < n; ++k) images[k].levels = p->desc[k].levels;
    g_batch_cache->erase(g_batch_cache->begin() + (long)i);
    g_batch_cache->insert(g_batch_cache->begin(), p);
    return p;
  }
  PlanRef p(batch_plan_create(wavelet, direction, storage, images, n, st), release_plan);
  g_batch_cache->insert(g_batch_cache->begin(), p);
  if (g_batch_cache->size() > kBatchCache) {
    evicted = std::move(g_batch_cache->back());
    g_batch_cache->pop_back();
  }
  return p;
}

}  // namespace dwt

using namespace dwt;

extern "C" {

int dwt_gpu_batch_plan_create(int wavelet, int direction, int storage, dwt_gpu_image* images,
                              int n, void* stream, void** plan) {
  try {
    g_err.clear();
    if (!plan) fail(DWT_GPU_EINVAL, "null plan pointer");
    *plan = batch_plan_create(wavelet, direction, storage, images, n, (hipStream_t)stream);
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_batch_plan_execute(void* plan, void* stream) {
  try {
    g_err.clear();
    if (!plan) fail(DWT_GPU_EINVAL, "null plan");
    batch_plan_execute(*static_cast<BatchPlan*>(plan), (hipStream_t)stream);
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_batch_plan_destroy(void* plan) {
  if (plan) {
    hipDeviceSynchronize();
    delete static_cast<BatchPlan*>(plan);
  }
  return DWT_GPU_OK;
}

int dwt_gpu_batch_2d(int wavelet, int direction, int storage, dwt_gpu_image* images, int n,
                     void* stream) {
  try {
    g_err.clear();
    const PlanRef p = batch_plan_cached(wavelet, direction, storage, images, n, (hipStream_t)stream);
    batch_plan_execute(*p, (hipStream_t)stream);
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_batch_tail_set(long long max_region_bytes) {
  if (max_region_bytes < 0 || max_region_bytes > kBatchTailMax) {
    g_err = "dwt_gpu_batch_tail_set: 0 (off) .. 131072 bytes";
    return DWT_GPU_EINVAL;
  }
  g_batch_tail_bytes = max_region_bytes;
  ++g_batch_config_gen;
  return DWT_GPU_OK;
}

int dwt_gpu_batch_order_check(int wavelet, int direction, const dwt_gpu_image* images, int n) {
  try {
    g_err.clear();
    batch_check_kind(wavelet, direction, kStoreF32);
    if (n <= 0 || !images) fail(DWT_GPU_EINVAL, "empty batch");
    const bool fwd = direction == DWT_GPU_FORWARD;
    std::vector<dwt_gpu_image> desc(images, images + n);
    std::vector<char*> base(n);
    for (int i = 0; i < n; ++i) {
      batch_check_image(i, desc[i], fwd, kStoreF32);
      base[i] = reinterpret_cast<char*>(uintptr_t(1) << 40);  // never dereferenced
    }
    BatchTables t;
    t.build(wavelet, fwd, kStoreF32, desc.data(), n, base.data());
    // every level launch: each tile (segment, strip) of each of its items is
    // computed by exactly one wave, through the kernels' own mapping
    for (size_t li = 0; li < t.levels.size(); ++li) {
      const BatchLevel& lv = t.levels[li];
      const BatchItem* items = t.items.data() + lv.items_off;
      const int* bmap = t.bmap.data() + lv.blocks_off;
      const size_t end = li + 1 < t.levels.size() ? (size_t)t.levels[li + 1].items_off : t.items.size();
      const int ni = (int)(end - lv.items_off);
      std::vector<std::vector<char>> seen(ni);
      for (int k = 0; k < ni; ++k) seen[k].assign((size_t)items[k].nstrips * items[k].nseg, 0);
      for (int b = 0; b < lv.nblocks; ++b) {
        int it, local;
        if (!batch_block_of(b, lv.nblocks, items, bmap, it, local)) continue;
        if (it < 0 || it >= ni || local < 0 || local >= items[it].nbp)
          fail(DWT_GPU_EINVAL, "level %d: workgroup %d maps outside its item", lv.level, b);
        for (int w = 0; w < 4; ++w) {
          int sg, strip;
          if (!batch_tile_of(local, w, items[it], sg, strip)) continue;
          if (sg < 0 || strip < 0 || strip >= items[it].nstrips)
            fail(DWT_GPU_EINVAL, "level %d: workgroup %d wave %d has a tile outside its item", lv.level, b, w);
          if (seen[it][(size_t)sg * items[it].nstrips + strip]++)
            fail(DWT_GPU_EINVAL, "level %d: tile (%d, %d) of item %d computed twice", lv.level, sg, strip, it);
        }
      }
      for (int k = 0; k < ni; ++k)
        for (size_t q = 0; q < seen[k].size(); ++q)
          if (!seen[k][q]) fail(DWT_GPU_EINVAL, "level %d: tile %zu of item %d not computed", lv.level, q, k);
    }
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_batch_rows_set(int level, int seg_rows) {
  if (level < 0 || level >= 32 || !(seg_rows == 0 || seg_rows == 8 || seg_rows == 16 || seg_rows == 32)) {
    g_err = "dwt_gpu_batch_rows_set: level 0..31, rows 0, 8, 16 or 32";
    return DWT_GPU_EINVAL;
  }
  g_batch_rows[level] = seg_rows;
  ++g_batch_config_gen;
  return DWT_GPU_OK;
}

}  // extern "C"
