// Batched transforms of many images of uneven sizes.
//
// Schedule: level-synchronous.  Level j of every image of the batch is one
// launch: the work items (four strips x one segment of S rows, as in the
// per-image level kernels) of all images at that level are numbered
// image-major and a workgroup finds its image through a per-block table, so a
// launch starts every workgroup at once with no atomics (a single ticket
// counter shared by tens of thousands of items serialises their start; the
// dataflow pyramid measured that).  A batch is then L launches per direction
// for L = the deepest image's level count, whatever the number of images.
//
// Layout: out of place.  The forward reads each image (FP32) and writes its
// coefficients in Mallat layout to dst; intermediate LL bands of levels
// 1 .. L-1 live in per-image ping-pong scratch.  The inverse reads
// coefficients from src and writes the image to dst.
//
// Storage: the coefficients (forward output, inverse input) are FP32, FP16 or
// BF16; lifting arithmetic is FP32 throughout.  A 16-bit coefficient is
// stored as round-to-nearest-even(c * s) for a per-subband power of two s
// (exact to apply and to undo).  FP16: s is chosen before the forward
// transform from the image's max |x| and the filters' l1 gains so that
// |c * s| <= 2^14: no FP16 overflow, and the subband's largest coefficient
// uses the top of the FP16 exponent range.  BF16 has FP32's exponent range,
// so a power-of-two scale changes only the exponent of what is stored, never
// the rounded mantissa: round(c * s) / s == round(c) for every c that is not
// an FP32 subnormal.  Its scales are 1, and the forward skips the pass over
// the images that finding max |x| costs.  Scales: 3 per level (HL, LH, HH)
// then one for the final LL, in a float array per image that the inverse
// reads back.
#include "block_body.h"
#include "kernels.h"
#include "batch.h"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace dwt {

// ---- 16-bit storage ------------------------------------------------------------
template <int SK>
__device__ __forceinline__ unsigned short to_storage(float v) {
  if constexpr (SK == kStoreF16) {
    const _Float16 h = (_Float16)v;  // round to nearest even
    return __builtin_bit_cast(unsigned short, h);
  } else {
    unsigned int u = __float_as_uint(v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
  }
}
template <int SK>
__device__ __forceinline__ float from_storage(unsigned short b) {
  if constexpr (SK == kStoreF16) return (float)__builtin_bit_cast(_Float16, b);
  else return __uint_as_float((unsigned int)b << 16);
}

// Detail subbands in 16-bit storage with a scale per band.
template <int SK>
struct Det16 {
  unsigned short* __restrict__ p;
  long long pitch;
  float s[3], inv[3];
  __device__ __forceinline__ void store(int band, int r, int c, float v) const {
    Gbl<unsigned short>* q = gstore_ptr(p);
    q[(long long)r * pitch + c] = to_storage<SK>(v * s[band]);
  }
  __device__ __forceinline__ float load(int band, int r, int c) const {
    return from_storage<SK>(gload_ptr(static_cast<const unsigned short*>(p))[(long long)r * pitch + c]) *
           inv[band];
  }
};
template <int SK>
struct Store16 {  // LL coefficient in 16-bit storage
  unsigned short* __restrict__ p;
  long long pitch;
  float s;
  __device__ __forceinline__ void operator()(int m, int k, float v) const {
    gstore_ptr(p)[(long long)m * pitch + k] = to_storage<SK>(v * s);
  }
};
template <int SK>
struct Load16 {
  const unsigned short* __restrict__ p;
  long long pitch;
  float inv;
  __device__ __forceinline__ float operator()(int m, int k) const {
    return from_storage<SK>(gload_ptr(p)[(long long)m * pitch + k]) * inv;
  }
};

// Image rows / output pairs with the two-element vector access chosen per
// item at run time (images of one batch may or may not be pair-aligned).
struct RowsRt {
  GlobalRows<float, true> v;
  bool vec;
  __device__ __forceinline__ void operator()(int Y, int Xe, int xe, int xo, bool interior, float& e,
                                             float& o) const {
    v(Y, Xe, xe, xo, interior && vec, e, o);
  }
};
struct OutRt {
  GlobalOut<float, true> v;
  GlobalOut<float, false> s;
  bool vec;
  __device__ __forceinline__ void operator()(int r, int Xe, float e, float o, bool st_e, bool st_o) const {
    if (vec) v(r, Xe, e, o, st_e, st_o);
    else s(r, Xe, e, o, st_e, st_o);
  }
};

// ---- level kernels ------------------------------------------------------------------
// This wave's tile (segment, strip) of an item (batch_item_blocks, batch.h).
static_assert(kWavesPerBlock == 4, "batch_item_blocks and batch_tile_of deal four tiles per workgroup");
__device__ __forceinline__ bool batch_tile(int local, const BatchItem& e, int& s, int& strip) {
  return batch_tile_of(local, __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), e, s, strip);
}

__device__ __forceinline__ bool batch_block(const BatchItem* __restrict__ items,
                                            const int* __restrict__ block_item, const BatchItem*& e,
                                            int& local) {
  int it;
  if (!batch_block_of(blockIdx.x, gridDim.x, items, block_item, it, local)) return false;
  e = &items[it];
  return true;
}

template <int NS, int SK, int SS>
__global__ void __launch_bounds__(256)
batch_fwd_kernel(const BatchItem* __restrict__ items, const int* __restrict__ block_item,
                 Coeffs<float> cf) {
  const BatchItem* ep;
  int local;
  if (!batch_block(items, block_item, ep, local)) return;
  const BatchItem& e = *ep;
  int s, strip;
  if (!batch_tile(local, e, s, strip)) return;
  const RowsRt rows{{static_cast<const float*>(e.src), e.sp, e.H},
                    __builtin_amdgcn_readfirstlane(e.flags & kItemVec) != 0};
  if constexpr (SK == kStoreF32) {
    fwd_tile<float, NS, -SS>(strip, s, SS, e.W, e.H, rows,
                             GlobalStore<float>{static_cast<float*>(e.out), e.op},
                             GlobalDet<float>{static_cast<float*>(e.det), e.dp}, cf);
  } else {
    const Gbl<float>* sc = gload_ptr(static_cast<const float*>(e.scale)) + 3 * e.level;
    const Det16<SK> det{static_cast<unsigned short*>(e.det), e.dp, {sc[0], sc[1], sc[2]},
                        {1.0f / sc[0], 1.0f / sc[1], 1.0f / sc[2]}};
    if (e.flags & kItemStorageLL)  // last level: LL into the coefficients
      fwd_tile<float, NS, -SS>(strip, s, SS, e.W, e.H, rows,
                               Store16<SK>{static_cast<unsigned short*>(e.out), e.op,
                                           gload_ptr(static_cast<const float*>(e.scale))[3 * e.levels]},
                               det, cf);
    else
      fwd_tile<float, NS, -SS>(strip, s, SS, e.W, e.H, rows,
                               GlobalStore<float>{static_cast<float*>(e.out), e.op}, det, cf);
  }
}

template <int NS, int SK, int SS>
__global__ void __launch_bounds__(256)
batch_inv_kernel(const BatchItem* __restrict__ items, const int* __restrict__ block_item,
                 Coeffs<float> cf) {
  const BatchItem* ep;
  int local;
  if (!batch_block(items, block_item, ep, local)) return;
  const BatchItem& e = *ep;
  int s, strip;
  if (!batch_tile(local, e, s, strip)) return;
  const OutRt out{{static_cast<float*>(e.out), e.op},
                  {static_cast<float*>(e.out), e.op},
                  __builtin_amdgcn_readfirstlane(e.flags & kItemVec) != 0};
  if constexpr (SK == kStoreF32) {
    inv_tile<float, NS, -SS>(strip, s, SS, e.W, e.H,
                             GlobalLoad<float>{static_cast<const float*>(e.src), e.sp},
                             GlobalDet<float>{static_cast<float*>(e.det), e.dp}, out, cf);
  } else {
    const Gbl<float>* sc = gload_ptr(static_cast<const float*>(e.scale)) + 3 * e.level;
    const Det16<SK> det{static_cast<unsigned short*>(e.det), e.dp, {sc[0], sc[1], sc[2]},
                        {1.0f / sc[0], 1.0f / sc[1], 1.0f / sc[2]}};
    if (e.flags & kItemStorageLL)  // coarsest level: LL from the coefficients
      inv_tile<float, NS, -SS>(strip, s, SS, e.W, e.H,
                               Load16<SK>{static_cast<const unsigned short*>(e.src), e.sp,
                                          1.0f / gload_ptr(static_cast<const float*>(e.scale))[3 * e.levels]},
                               det, out, cf);
    else
      inv_tile<float, NS, -SS>(strip, s, SS, e.W, e.H,
                               GlobalLoad<float>{static_cast<const float*>(e.src), e.sp}, det, out,
                               cf);
  }
}

// Max |x| of every block's rows, as the bit pattern (which orders
// non-negative floats): one partial per block, reduced per image by the
// scales kernel.  (One atomic max per block on a word per image serialised
// ~100 device-scope atomics per image: 200 us for 256 images.)
__global__ void __launch_bounds__(256)
batch_maxabs_kernel(const BatchImage* __restrict__ imgs, const int* __restrict__ block_img,
                    unsigned* __restrict__ amax) {
  __shared__ unsigned red[4];
  const int i = block_img[blockIdx.x];
  const BatchImage& im = imgs[i];
  if (im.sp == im.W) {
    // contiguous image: pixels [c0, c0 + kMaxAbsFlat) of the flat range,
    // element-aligned 16-byte loads (any base alignment), all issued first
    using F4 = float __attribute__((ext_vector_type(4), aligned(4)));
    constexpr int K = kMaxAbsFlat / 256 / 4;
    const long long n = (long long)im.W * im.H;
    const long long c0 = (long long)(blockIdx.x - im.block0) * kMaxAbsFlat;
    const int cnt = (int)min((long long)kMaxAbsFlat, n - c0);
    const Gbl<float>* src = gload_ptr(static_cast<const float*>(im.src)) + c0;
    float m = 0.0f;
    F4 v[K];
#pragma unroll
    for (int u = 0; u < K; ++u) {
      const int e = 4 * (threadIdx.x + 256 * u);
      v[u] = e + 3 < cnt ? *reinterpret_cast<const Gbl<F4>*>(src + e) : F4{0.0f, 0.0f, 0.0f, 0.0f};
    }
#pragma unroll
    for (int u = 0; u < K; ++u)
      m = fmaxf(m, fmaxf(fmaxf(fabsf(v[u].x), fabsf(v[u].y)), fmaxf(fabsf(v[u].z), fabsf(v[u].w))));
    // the last block's pixels past its whole groups of four
    for (int e = (cnt & ~3) + threadIdx.x; e < cnt; e += 256) m = fmaxf(m, fabsf(src[e]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = __float_as_uint(m);
    __syncthreads();
    if (threadIdx.x == 0) amax[blockIdx.x] = max(max(red[0], red[1]), max(red[2], red[3]));
    return;
  }
  const int W = im.W, R = maxabs_rows(W);
  const int y0 = (blockIdx.x - im.block0) * R;
  const int n = min(R, im.H - y0) * W;  // pixels of this block's rows
  const Gbl<float>* src = gload_ptr(static_cast<const float*>(im.src)) + (long long)y0 * im.sp;
  float m = 0.0f;
  const bool v4 = (W & 3) == 0 && (im.sp & 3) == 0 && (reinterpret_cast<uintptr_t>(im.src) & 15) == 0;
  // kMaxAbsPixels per pass; a block has more pixels only when it holds one
  // row longer than that
  for (int c0 = 0; c0 < n; c0 += kMaxAbsPixels) {
    const int lim = min(n, c0 + kMaxAbsPixels);
    if (v4) {
      // rows of whole 16-byte groups: four pixels per load, 32-bit offsets
      // from the block's first row
      using F4 = float __attribute__((ext_vector_type(4)));
      constexpr int K4 = kMaxAbsPixels / 256 / 4;
      const int W4 = W >> 2, lim4 = lim >> 2;
      const int sp4 = (int)(im.sp >> 2);
      const Gbl<F4>* src4 = reinterpret_cast<const Gbl<F4>*>(src);
      int e = (c0 >> 2) + threadIdx.x, y = e / W4, x = e - y * W4;
      int off[K4];
#pragma unroll
      for (int u = 0; u < K4; ++u) {
        off[u] = e < lim4 ? y * sp4 + x : -1;
        e += 256;
        x += 256;
        while (x >= W4) {
          x -= W4;
          ++y;
        }
      }
      F4 v[K4];
#pragma unroll
      for (int u = 0; u < K4; ++u) v[u] = off[u] >= 0 ? src4[off[u]] : F4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int u = 0; u < K4; ++u)
        m = fmaxf(m, fmaxf(fmaxf(fabsf(v[u].x), fabsf(v[u].y)), fmaxf(fabsf(v[u].z), fabsf(v[u].w))));
    } else {
      // (y, x) of pixel e advance with e by one subtraction loop instead of
      // a division; all of a thread's loads are issued before any is reduced
      // (one memory round trip per pass)
      constexpr int K = kMaxAbsPixels / 256;
      int e = c0 + threadIdx.x, y = e / W, x = e - y * W;
      long long off[K];
#pragma unroll
      for (int u = 0; u < K; ++u) {
        off[u] = e < lim ? (long long)y * im.sp + x : -1;
        e += 256;
        x += 256;
        while (x >= W) {
          x -= W;
          ++y;
        }
      }
      float v[K];
#pragma unroll
      for (int u = 0; u < K; ++u) v[u] = off[u] >= 0 ? fabsf(src[off[u]]) : 0.0f;
#pragma unroll
      for (int u = 0; u < K; ++u) m = fmaxf(m, v[u]);
    }
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = __float_as_uint(m);
  __syncthreads();
  if (threadIdx.x == 0) amax[blockIdx.x] = max(max(red[0], red[1]), max(red[2], red[3]));
}

// Per-subband scales from max |x|: s = 2^floor(log2(2^14 / (gain * max|x|))).
// One block per image: its threads reduce the image's per-block partials of
// batch_maxabs_kernel, then one thread per subband computes its scale.
// unit: every scale 1 (BF16 storage, below), no partials read.
__global__ void __launch_bounds__(256)
batch_scales_kernel(const BatchImage* __restrict__ imgs, int n, const unsigned* __restrict__ amax,
                    const float* __restrict__ gain, int unit) {
  __shared__ unsigned red[4];
  const int i = blockIdx.x;
  if (i >= n) return;
  const BatchImage& im = imgs[i];
  if (unit) {
    for (int k = threadIdx.x; k <= 3 * im.levels; k += 256) gstore_ptr(im.scale)[k] = 1.0f;
    return;
  }
  const int nb = (int)maxabs_blocks(im.W, im.H, im.sp);
  unsigned mb = 0;
  for (int b = threadIdx.x; b < nb; b += 256) mb = max(mb, amax[im.block0 + b]);
  for (int off = 32; off > 0; off >>= 1) mb = max(mb, (unsigned)__shfl_xor((int)mb, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mb;
  __syncthreads();
  const float mx = __uint_as_float(max(max(red[0], red[1]), max(red[2], red[3])));
  for (int k = threadIdx.x; k <= 3 * im.levels; k += 256) {
    // gain table: 3 per level then 32 LL entries (LL after j levels at 96 + j)
    const float gk = k < 3 * im.levels ? gain[k] : gain[96 + im.levels];
    const float bound = gk * mx;
    float s = 1.0f;
    if (bound > 0.0f && isfinite(bound)) {
      int e = 14 - (int)ceilf(log2f(bound));
      e = max(-60, min(60, e));
      s = ldexpf(1.0f, e);
    }
    gstore_ptr(im.scale)[k] = s;
  }
}

// ---- the small levels of every image: one workgroup per image -----------------------
// A level of a small image is a few tiles, and a launch per level for the
// whole batch is latency-bound (about 8 us each, whatever the batch).  Here
// the levels from the first whose region fits kBatchTailMax run as one
// launch, image b in workgroup b, with the block kernels' line-parallel
// lifting in LDS (block_body.h), whose arithmetic is the level kernels' bit
// for bit.  16-bit coefficients are converted on the way out (forward) or in
// (inverse) with the scale of the subband each position belongs to, as the
// level kernels' Det16 / Store16 / Load16 do.

// Level (0 .. nlev - 1, relative to the region) whose detail half holds
// position x of a region side of n samples after nlev levels; nlev for the
// final LL.
__device__ __forceinline__ int tail_level(int x, int n, int nlev) {
  for (int l = 0; l < nlev; ++l)
    if (x >= ceil_div_pow2(n, l + 1)) return l;
  return nlev;
}
// Index into the image's scales of position (x, y) of the region: 3 per level
// (HL, LH, HH) from level t.level on, then the final LL's at 3 * levels.
__device__ __forceinline__ int tail_scale_index(int lx, int ly, const BatchTail& t) {
  const int l = min(lx, ly);
  if (l == t.nlev) return 3 * t.levels;
  return 3 * (t.level + l) + (lx == l ? (ly == l ? kHH : kHL) : kLH);
}

// The region between the coefficients (16-bit, global) and the FP32 LDS
// buffer C (pitch pc); S holds the scales (forward: multiply) or their
// inverses (inverse: multiply), indexed as tail_scale_index.  As block_copy:
// wavefront w takes rows w, w + 16, ..., lanes consecutive samples in NCH
// chunks of 64, and a round issues all of its loads before the first is
// used (one memory round trip per round, not per row).
template <int SK, bool TO_LDS, int NCH>
__device__ __forceinline__ void tail_copy16_rows(Lds<float>* C, int pc, const BatchTail& t, const Lds<float>* S) {
  constexpr int RPR = kBlockRound / NCH;  // rows per round
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long gp = TO_LDS ? t.sp : t.dp;
  int lx[NCH];
  bool in[NCH];
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    in[c] = lane + c * kWave < t.W;
    lx[c] = tail_level(lane + c * kWave, t.W, t.nlev);
  }
  for (int y0 = wave; y0 < t.H; y0 += RPR * kBlockWaves) {
    float v[RPR][NCH];
#pragma unroll
    for (int r = 0; r < RPR; ++r) {
      const int y = y0 + r * kBlockWaves;
      if (y < t.H) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          const int x = lane + c * kWave;
          if constexpr (TO_LDS)
            v[r][c] = in[c] ? from_storage<SK>(gload_ptr(static_cast<const unsigned short*>(t.src))[y * gp + x]) : 0.0f;
          else
            v[r][c] = in[c] ? C[y * pc + x] : 0.0f;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RPR; ++r) {
      const int y = y0 + r * kBlockWaves;
      if (y < t.H) {
        const int ly = tail_level(y, t.H, t.nlev);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          const int x = lane + c * kWave;
          if (!in[c]) continue;
          const float s = S[tail_scale_index(lx[c], ly, t)];
          if constexpr (TO_LDS)
            C[y * pc + x] = v[r][c] * s;
          else
            gstore_ptr(static_cast<unsigned short*>(t.dst))[y * gp + x] = to_storage<SK>(v[r][c] * s);
        }
      }
    }
  }
}
template <int SK, bool TO_LDS>
__device__ __forceinline__ void tail_copy16(Lds<float>* C, int pc, const BatchTail& t, const Lds<float>* S) {
  if (t.W <= kWave)
    tail_copy16_rows<SK, TO_LDS, 1>(C, pc, t, S);
  else if (t.W <= 2 * kWave)
    tail_copy16_rows<SK, TO_LDS, 2>(C, pc, t, S);
  else if (t.W <= 4 * kWave)
    tail_copy16_rows<SK, TO_LDS, 4>(C, pc, t, S);
  else
    tail_copy16_rows<SK, TO_LDS, 8>(C, pc, t, S);
}

template <int NS, int SK>
__global__ void __launch_bounds__(kBlockThreads)
batch_tail_fwd_kernel(const BatchTail* __restrict__ tails, Coeffs<float> cf) {
  extern __shared__ __align__(16) unsigned char blk_smem[];
  const BatchTail t = tails[blockIdx.x];
  Lds<float>* C = (Lds<float>*)blk_smem;  // generic -> LDS address-space cast
  const int pc = block_pitch(t.W);
  Lds<float>* S = C + pc * t.H;
  Lds<float>* const dummy = S + 3 * 32 + 1;
  if constexpr (SK != kStoreF32)
    for (int k = threadIdx.x; k <= 3 * t.levels; k += kBlockThreads) S[k] = gload_ptr(t.scale)[k];
  block_copy<true>(C, pc, const_cast<float*>(static_cast<const float*>(t.src)), t.sp, t.W, t.H);
  __syncthreads();
  for (int lev = 0; lev < t.nlev; ++lev) {
    const int W = ceil_div_pow2(t.W, lev), H = ceil_div_pow2(t.H, lev);
    block_pass_any<true, float, NS>(C, pc, 1, H, W, cf, dummy);  // rows
    __syncthreads();
    block_pass_any<true, float, NS>(C, 1, pc, W, H, cf, dummy);  // columns
    __syncthreads();
  }
  if constexpr (SK == kStoreF32)
    block_copy<false>(C, pc, static_cast<float*>(t.dst), t.dp, t.W, t.H);
  else
    tail_copy16<SK, false>(C, pc, t, S);
}

template <int NS, int SK>
__global__ void __launch_bounds__(kBlockThreads)
batch_tail_inv_kernel(const BatchTail* __restrict__ tails, Coeffs<float> cf) {
  extern __shared__ __align__(16) unsigned char blk_smem[];
  const BatchTail t = tails[blockIdx.x];
  Lds<float>* C = (Lds<float>*)blk_smem;  // generic -> LDS address-space cast
  const int pc = block_pitch(t.W);
  Lds<float>* S = C + pc * t.H;
  Lds<float>* const dummy = S + 3 * 32 + 1;
  if constexpr (SK == kStoreF32) {
    block_copy<true>(C, pc, const_cast<float*>(static_cast<const float*>(t.src)), t.sp, t.W, t.H);
  } else {
    for (int k = threadIdx.x; k <= 3 * t.levels; k += kBlockThreads) S[k] = 1.0f / gload_ptr(t.scale)[k];
    __syncthreads();
    tail_copy16<SK, true>(C, pc, t, S);
  }
  __syncthreads();
  for (int lev = t.nlev - 1; lev >= 0; --lev) {
    const int W = ceil_div_pow2(t.W, lev), H = ceil_div_pow2(t.H, lev);
    block_pass_any<false, float, NS>(C, pc, 1, H, W, cf, dummy);  // rows
    __syncthreads();
    block_pass_any<false, float, NS>(C, 1, pc, W, H, cf, dummy);  // columns
    __syncthreads();
  }
  block_copy<false>(C, pc, static_cast<float*>(t.dst), t.dp, t.W, t.H);
}

// ---- host launchers --------------------------------------------------------------------
size_t batch_tail_lds(int W, int H) {
  return sizeof(float) * ((size_t)block_pitch(W) * H + 3 * 32 + 1 + kBlockDummy);
}


template <int NS, int SK>
static void tail_go(bool fwd, int n, size_t lds, const BatchTail* tails, const Coeffs<float>& cf,
                    hipStream_t st) {
  if (fwd)
    hipLaunchKernelGGL((batch_tail_fwd_kernel<NS, SK>), dim3(n), dim3(kBlockThreads), lds, st, tails, cf);
  else
    hipLaunchKernelGGL((batch_tail_inv_kernel<NS, SK>), dim3(n), dim3(kBlockThreads), lds, st, tails, cf);
}

void launch_batch_tail(bool fwd, int storage, int n, size_t lds, const BatchTail* tails,
                       const Coeffs<float>& cf, hipStream_t st) {
  if (n <= 0) return;
  if (cf.nsteps == 4) {
    if (storage == kStoreF16) tail_go<4, kStoreF16>(fwd, n, lds, tails, cf, st);
    else if (storage == kStoreBF16) tail_go<4, kStoreBF16>(fwd, n, lds, tails, cf, st);
    else tail_go<4, kStoreF32>(fwd, n, lds, tails, cf, st);
  } else {
    if (storage == kStoreF16) tail_go<2, kStoreF16>(fwd, n, lds, tails, cf, st);
    else if (storage == kStoreBF16) tail_go<2, kStoreBF16>(fwd, n, lds, tails, cf, st);
    else tail_go<2, kStoreF32>(fwd, n, lds, tails, cf, st);
  }
}

hipError_t configure_batch_tail() {
  const void* fns[] = {
      reinterpret_cast<const void*>(&batch_tail_fwd_kernel<4, kStoreF32>),
      reinterpret_cast<const void*>(&batch_tail_fwd_kernel<4, kStoreF16>),
      reinterpret_cast<const void*>(&batch_tail_fwd_kernel<4, kStoreBF16>),
      reinterpret_cast<const void*>(&batch_tail_fwd_kernel<2, kStoreF32>),
      reinterpret_cast<const void*>(&batch_tail_fwd_kernel<2, kStoreF16>),
      reinterpret_cast<const void*>(&batch_tail_fwd_kernel<2, kStoreBF16>),
      reinterpret_cast<const void*>(&batch_tail_inv_kernel<4, kStoreF32>),
      reinterpret_cast<const void*>(&batch_tail_inv_kernel<4, kStoreF16>),
      reinterpret_cast<const void*>(&batch_tail_inv_kernel<4, kStoreBF16>),
      reinterpret_cast<const void*>(&batch_tail_inv_kernel<2, kStoreF32>),
      reinterpret_cast<const void*>(&batch_tail_inv_kernel<2, kStoreF16>),
      reinterpret_cast<const void*>(&batch_tail_inv_kernel<2, kStoreBF16>)};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kBlockLdsMax);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return e;
    }
  }
  return hipSuccess;
}

template <int NS, int SK>
static void batch_go(bool fwd, int S, int nblocks, const BatchItem* items, const int* block_item,
                     const Coeffs<float>& cf, hipStream_t st) {
  const dim3 grid(nblocks), block(256);
  if (fwd) {
    if (S == 16) hipLaunchKernelGGL((batch_fwd_kernel<NS, SK, 16>), grid, block, 0, st, items, block_item, cf);
    else if (S == 32) hipLaunchKernelGGL((batch_fwd_kernel<NS, SK, 32>), grid, block, 0, st, items, block_item, cf);
    else hipLaunchKernelGGL((batch_fwd_kernel<NS, SK, 8>), grid, block, 0, st, items, block_item, cf);
  } else {
    if (S == 16) hipLaunchKernelGGL((batch_inv_kernel<NS, SK, 16>), grid, block, 0, st, items, block_item, cf);
    else if (S == 32) hipLaunchKernelGGL((batch_inv_kernel<NS, SK, 32>), grid, block, 0, st, items, block_item, cf);
    else hipLaunchKernelGGL((batch_inv_kernel<NS, SK, 8>), grid, block, 0, st, items, block_item, cf);
  }
}

void launch_batch_level(bool fwd, int storage, int S, int nblocks, const BatchItem* items,
                        const int* block_item, const Coeffs<float>& cf, hipStream_t st) {
  if (nblocks <= 0) return;
  if (cf.nsteps == 4) {
    if (storage == kStoreF16) batch_go<4, kStoreF16>(fwd, S, nblocks, items, block_item, cf, st);
    else if (storage == kStoreBF16) batch_go<4, kStoreBF16>(fwd, S, nblocks, items, block_item, cf, st);
    else batch_go<4, kStoreF32>(fwd, S, nblocks, items, block_item, cf, st);
  } else {
    if (storage == kStoreF16) batch_go<2, kStoreF16>(fwd, S, nblocks, items, block_item, cf, st);
    else if (storage == kStoreBF16) batch_go<2, kStoreBF16>(fwd, S, nblocks, items, block_item, cf, st);
    else batch_go<2, kStoreF32>(fwd, S, nblocks, items, block_item, cf, st);
  }
}

void launch_batch_maxabs(const BatchImage* imgs, const int* block_img, int nblocks, unsigned* amax,
                         hipStream_t st) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(batch_maxabs_kernel, dim3(nblocks), dim3(256), 0, st, imgs, block_img, amax);
}

void launch_batch_scales(const BatchImage* imgs, int n, const unsigned* amax, const float* gain, int unit,
                         hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(batch_scales_kernel, dim3(n), dim3(256), 0, st, imgs, n, amax, gain, unit);
}

// l1 norms of the one-level analysis lowpass and highpass filters (libdwt's
// normalisation), from the lifting steps applied to unit impulses of both
// parities far from the line ends.
void analysis_gains(const Coeffs<double>& cf, double* gl, double* gh) {
  const int N = 64;
  double sl = 0, sh = 0;
  for (int p = N / 2; p < N / 2 + 2; ++p) {
    double x[N] = {};
    x[p] = 1.0;
    for (int st = 0; st < cf.nsteps; ++st) {
      const int par = (cf.first_parity + st) & 1;
      for (int i = par; i < N; i += 2) {
        const double l = i > 0 ? x[i - 1] : x[i + 1];
        const double r = i + 1 < N ? x[i + 1] : x[i - 1];
        x[i] += cf.c[st] * (l + r);
      }
    }
    for (int i = 0; i < N; i += 2) sl += std::fabs(x[i] * cf.se);
    for (int i = 1; i < N; i += 2) sh += std::fabs(x[i] * cf.so);
  }
  *gl = sl;
  *gh = sh;
}

}  // namespace dwt
