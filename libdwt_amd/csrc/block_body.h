// Device code of the block kernels: every level of a region lifted
// line-parallel in one workgroup's LDS (rows, then columns), and the copies
// between global memory and the LDS buffer.  Shared by the single-image block
// kernels (block_kernels.hip) and the batched tail (batch.hip).
//
// Block kernels: the small levels of a transform in one workgroup, the whole
// region held in LDS.
//
// A level of a few thousand samples is bound by latency, not bandwidth: a
// separate launch per level costs about 3 us of launch and ramp on this GPU,
// and the streaming tile bodies of the level kernels (a wave walking down 16
// rows to finish 8) are long serial chains.  Here every level of the region is
// lifted line-parallel instead:
//   * row pass: each wavefront takes one row at a time; lane l holds PPL
//     consecutive sample pairs (2j, 2j+1), j = PPL*l + q, so a row of up to
//     128 * PPL samples is one wavefront's registers.  Each lifting step over
//     the whole row is a few VALU instructions: the neighbour pair across a
//     lane boundary comes by DPP, and a line end takes whole-sample symmetric
//     extension by using the lane's own (or its neighbour's) value;
//   * column pass: the same on columns, after a workgroup barrier;
//   * the level's L and H halves go back to the line in place, so the LL of
//     the next level is the top-left part of the same buffer, and after the
//     last level the buffer holds the region's coefficients in Mallat layout.
// Rows and then columns, in both directions, as the reference's drivers do
// (src/libdwt.c:6354-6375, 6706-6727); each update is x + c * (l + r) with the
// operations of the level kernels, so the result is bit-identical to them.
//
// Forward: the region is read from src, every level lifted in LDS, and the
// buffer written to dst once.  Inverse: the region's coefficients are read
// from src, the levels inverted coarsest first, and the finest output written
// to dst.  The region is read completely before any of it is written, so src
// may equal dst.
#pragma once

#include "tile_body.h"
#include "kernels.h"

namespace dwt {

constexpr int kBlockThreads = 1024;
constexpr int kBlockWaves = kBlockThreads / kWave;
constexpr int kBlockMaxPPL = 4;                       // lines of up to 512 samples
constexpr int kBlockMaxLine = 2 * kWave * kBlockMaxPPL;
constexpr int kBlockLdsMax = 160 * 1024;              // LDS of one workgroup on gfx950

// Row pitch of the LDS buffer: odd, so that a column pass (lane l reading
// rows 2j and 2j + 1 of one column) spreads over the LDS banks.
__host__ __device__ inline int block_pitch(int W0) { return W0 | 1; }

// Forward lifting of one line held as PPL pairs per lane: odd samples first,
// then even, NS steps, then scaling.  j0 = index of the lane's first pair,
// n = line length (>= 2).
template <typename T, int NS, int PPL>
__device__ __forceinline__ void lift_fwd(T (&e)[PPL], T (&o)[PPL], int j0, int n, const Coeffs<T>& cf) {
#pragma unroll
  for (int s = 0; s < NS; s += 2) {
    {
      // o_j += c (e_j + e_{j+1}); for the last odd sample of an even-length
      // line, e_{j+1} = x[n] mirrors to x[n-2] = e_j
      T en[PPL];
#pragma unroll
      for (int q = 0; q + 1 < PPL; ++q) en[q] = e[q + 1];
      en[PPL - 1] = from_next_lane(e[0]);
#pragma unroll
      for (int q = 0; q < PPL; ++q) {
        const T nb = 2 * (j0 + q) + 2 == n ? e[q] : en[q];
        T t = e[q] + nb;
        t = cf.c[s] * t;
        o[q] = o[q] + t;
      }
    }
    {
      // e_j += c (o_{j-1} + o_j); o_{-1} = x[-1] mirrors to x[1] = o_0, and
      // for the last even sample of an odd-length line o_j = x[n] mirrors to
      // x[n-2] = o_{j-1}
      T op[PPL];
      op[0] = from_prev_lane(o[PPL - 1]);
#pragma unroll
      for (int q = 1; q < PPL; ++q) op[q] = o[q - 1];
#pragma unroll
      for (int q = 0; q < PPL; ++q) {
        const int j = j0 + q;
        const T l = j == 0 ? o[q] : op[q];
        const T r = 2 * j + 1 == n ? l : o[q];
        T t = l + r;
        t = cf.c[s + 1] * t;
        e[q] = e[q] + t;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    e[q] = e[q] * cf.se;
    o[q] = o[q] * cf.so;
  }
}

// Inverse lifting of one interleaved line: scaling first, then even samples,
// then odd, NS steps.
template <typename T, int NS, int PPL>
__device__ __forceinline__ void lift_inv(T (&e)[PPL], T (&o)[PPL], int j0, int n, const Coeffs<T>& cf) {
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    e[q] = e[q] * cf.se;
    o[q] = o[q] * cf.so;
  }
#pragma unroll
  for (int s = 0; s < NS; s += 2) {
    {
      T op[PPL];
      op[0] = from_prev_lane(o[PPL - 1]);
#pragma unroll
      for (int q = 1; q < PPL; ++q) op[q] = o[q - 1];
#pragma unroll
      for (int q = 0; q < PPL; ++q) {
        const int j = j0 + q;
        const T l = j == 0 ? o[q] : op[q];
        const T r = 2 * j + 1 == n ? l : o[q];
        T t = l + r;
        t = cf.c[s] * t;
        e[q] = e[q] + t;
      }
    }
    {
      T en[PPL];
#pragma unroll
      for (int q = 0; q + 1 < PPL; ++q) en[q] = e[q + 1];
      en[PPL - 1] = from_next_lane(e[0]);
#pragma unroll
      for (int q = 0; q < PPL; ++q) {
        const T nb = 2 * (j0 + q) + 2 == n ? e[q] : en[q];
        T t = e[q] + nb;
        t = cf.c[s + 1] * t;
        o[q] = o[q] + t;
      }
    }
  }
}

// One pass of a level over `nlines` lines of length n: line i starts at
// base + i * lstep, its samples `es` apart.  Forward: the line's samples are
// lifted and its L coefficients written to positions [0, ceil(n/2)), its H
// coefficients from position hoff on (ceil(n/2) in a packed frame; the outer
// frame's half in a sparse one, src/libdwt.c:6354-6362).  Inverse: L from
// [0, ceil(n/2)) and H from hoff are lifted back to samples [0, n).  A line of at most 64 pairs takes LPL
// lanes (the power of two >= its pair count), so a wavefront lifts 64 / LPL
// lines at once; the DPP neighbour across two such lines is never used, as a
// line's end lanes take their mirror values.  A line is read completely
// before the wavefront writes it (every store depends on the lifted values,
// which depend on every lane's loads through the DPP exchange).
template <bool FWD, typename T, int NS, int PPL>
struct LinePass {
  int ld_e[PPL], ld_o[PPL], st_e[PPL], st_o[PPL];
  int j0, n, nl, nh;
  // A store the lane must not make goes to its own slot of the dummy area
  // instead: an address select, where a conditional store would be a branch
  // around it with exec-mask bookkeeping in every lifted line.
  __device__ __forceinline__ void lift(Lds<T>* line, bool store, const Coeffs<T>& cf, Lds<T>* dummy) const {
    T e[PPL], o[PPL];
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      e[q] = line[ld_e[q]];
      o[q] = line[ld_o[q]];
    }
    if (FWD)
      lift_fwd<T, NS, PPL>(e, o, j0, n, cf);
    else
      lift_inv<T, NS, PPL>(e, o, j0, n, cf);
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      Lds<T>* pe = store && j0 + q < nl ? line + st_e[q] : dummy;
      Lds<T>* po = store && j0 + q < nh ? line + st_o[q] : dummy;
      *pe = e[q];
      *po = o[q];
    }
  }
};

template <bool FWD, typename T, int NS, int PPL>
__device__ __forceinline__ void block_pass(Lds<T>* base, int lstep, int es, int nlines, int n, int hoff,
                                           int lshift, const Coeffs<T>& cf, Lds<T>* dummy) {
  const int lane = threadIdx.x & (kWave - 1);
  Lds<T>* const dl = dummy + lane;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  LinePass<FWD, T, NS, PPL> lp;
  lp.n = n;
  lp.nl = (n + 1) >> 1;
  lp.nh = n >> 1;
  const int lpw = kWave >> lshift;  // lines per wavefront (LPL = 1 << lshift lanes each)
  const int sub = lane >> lshift;
  lp.j0 = (lane & ((1 << lshift) - 1)) * PPL;
  // sample offsets of this lane's pairs within a line (clamped into the line:
  // lanes past its end read a valid sample and store nothing)
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    const int je = min(lp.j0 + q, lp.nl - 1), jo = min(lp.j0 + q, lp.nh - 1);
    lp.ld_e[q] = (FWD ? 2 * je : je) * es;
    lp.ld_o[q] = (FWD ? 2 * jo + 1 : hoff + jo) * es;
    lp.st_e[q] = (FWD ? je : 2 * je) * es;
    lp.st_o[q] = (FWD ? hoff + jo : 2 * jo + 1) * es;
  }
  const int sub_off = sub * lstep;
  // two lines per iteration (independent chains, interleaved by the compiler);
  // a lane of a line past the last one lifts the wavefront's first line again
  // and stores nothing
  const int step = kBlockWaves * lpw;
  for (int i0 = wave * lpw; i0 < nlines; i0 += 2 * step) {
    const int i1 = i0 + step;
    const bool v0 = i0 + sub < nlines, v1 = i1 + sub < nlines;
    Lds<T>* l0 = base + i0 * lstep + (v0 ? sub_off : 0);
    Lds<T>* l1 = base + (i1 < nlines ? i1 : i0) * lstep + (v1 ? sub_off : 0);
    lp.lift(l0, v0, cf, dl);
    lp.lift(l1, v1, cf, dl);
  }
}

// Pairs per lane for a line of n samples (n <= kBlockMaxLine).
__device__ __forceinline__ int block_ppl(int n) {
  const int pairs = (n + 1) >> 1;
  return pairs <= kWave ? 1 : pairs <= 2 * kWave ? 2 : 4;
}

// dummy: kBlockDummy elements of LDS outside the region (one slot per lane)
// that take the stores lanes must not make.  hoff < 0: the packed frame's
// ceil(n/2).
constexpr int kBlockDummy = kWave;
template <bool FWD, typename T, int NS>
__device__ __forceinline__ void block_pass_any(Lds<T>* base, int lstep, int es, int nlines, int n,
                                               const Coeffs<T>& cf, Lds<T>* dummy, int hoff = -1) {
  if (hoff < 0) hoff = (n + 1) >> 1;
  const int ppl = block_ppl(n);
  if (ppl == 1) {
    const int pairs = (n + 1) >> 1;
    int lshift = 0;
    while ((1 << lshift) < pairs) ++lshift;
    block_pass<FWD, T, NS, 1>(base, lstep, es, nlines, n, hoff, lshift, cf, dummy);
  } else if (ppl == 2) {
    block_pass<FWD, T, NS, 2>(base, lstep, es, nlines, n, hoff, 6, cf, dummy);
  } else {
    block_pass<FWD, T, NS, 4>(base, lstep, es, nlines, n, hoff, 6, cf, dummy);
  }
}

// Zero padding of one level of a sparse frame held in LDS (C, pitch pc), outer
// W x H, inner Wi x Hi (src/libdwt.c:6377-6397 forward, 6729-6745 inverse).
// Forward: every row's samples between its L part and the outer half and past
// its H part, then every column's likewise.  Inverse: everything past the
// inner size.  The positions form a few rectangles (full-height column strips
// and full-width row strips), which the workgroup clears one after another,
// visiting nothing else; a position in two of them is cleared twice.  Followed
// by a barrier in the caller.
template <typename T>
__device__ __forceinline__ void block_zero_rect(Lds<T>* C, int pc, int x0, int x1, int y0, int y1) {
  const int w = x1 - x0, h = y1 - y0;
  if (w <= 0 || h <= 0) return;
  for (int i = threadIdx.x; i < w * h; i += blockDim.x) {
    const int y = i / w, x = i - y * w;
    C[(y0 + y) * pc + x0 + x] = T(0);
  }
}
template <bool FWD, typename T>
__device__ __forceinline__ void block_zero_pad(Lds<T>* C, int pc, int W, int H, int Wi, int Hi) {
  if (FWD) {
    const int Wd = (W + 1) >> 1, Hd = (H + 1) >> 1;
    block_zero_rect(C, pc, (Wi + 1) >> 1, Wd, 0, H);  // columns between L and the outer half
    block_zero_rect(C, pc, Wd + (Wi >> 1), W, 0, H);  // columns past H
    block_zero_rect(C, pc, 0, W, (Hi + 1) >> 1, Hd);  // rows between L and the outer half
    block_zero_rect(C, pc, 0, W, Hd + (Hi >> 1), H);  // rows past H
  } else {
    block_zero_rect(C, pc, Wi, W, 0, H);
    block_zero_rect(C, pc, 0, W, Hi, H);
  }
}

// Copy the W0 x H0 region between global memory (pitch gp) and the LDS buffer
// (pitch pc), in either direction.  Wavefront w takes rows w, w + 16, ...,
// lanes consecutive samples of a row (NCH chunks of 64 per row), so a row's
// address is wave-uniform and the chunks are immediate offsets; a round moves
// 16 / NCH rows, every load of the round issued before the first is waited for.
constexpr int kBlockRound = 16;  // samples a lane moves per round
template <bool TO_LDS, typename T, int NCH>
__device__ __forceinline__ void block_copy_rows(Lds<T>* C, int pc, T* g_, long long gp, int W0, int H0) {
  constexpr int RPR = kBlockRound / NCH;  // rows per round
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  bool in[NCH];
#pragma unroll
  for (int c = 0; c < NCH; ++c) in[c] = lane + c * kWave < W0;
  for (int y0 = wave; y0 < H0; y0 += RPR * kBlockWaves) {
    T v[RPR][NCH];
#pragma unroll
    for (int r = 0; r < RPR; ++r) {
      const int y = y0 + r * kBlockWaves;
      if (y < H0) {
        const Gbl<T>* grow = gload_ptr(g_) + (long long)y * gp + lane;
        const Lds<T>* lrow = C + y * pc + lane;
#pragma unroll
        for (int c = 0; c < NCH; ++c) v[r][c] = in[c] ? (TO_LDS ? grow[c * kWave] : lrow[c * kWave]) : T(0);
      }
    }
#pragma unroll
    for (int r = 0; r < RPR; ++r) {
      const int y = y0 + r * kBlockWaves;
      if (y < H0) {
        Gbl<T>* grow = gstore_ptr(g_) + (long long)y * gp + lane;
        Lds<T>* lrow = C + y * pc + lane;
#pragma unroll
        for (int c = 0; c < NCH; ++c)
          if (in[c]) {
            if (TO_LDS)
              lrow[c * kWave] = v[r][c];
            else
              grow[c * kWave] = v[r][c];
          }
      }
    }
  }
}
template <bool TO_LDS, typename T>
__device__ __forceinline__ void block_copy(Lds<T>* C, int pc, T* g, long long gp, int W0, int H0) {
  if (W0 <= kWave)
    block_copy_rows<TO_LDS, T, 1>(C, pc, g, gp, W0, H0);
  else if (W0 <= 2 * kWave)
    block_copy_rows<TO_LDS, T, 2>(C, pc, g, gp, W0, H0);
  else if (W0 <= 4 * kWave)
    block_copy_rows<TO_LDS, T, 4>(C, pc, g, gp, W0, H0);
  else
    block_copy_rows<TO_LDS, T, 8>(C, pc, g, gp, W0, H0);
}

}  // namespace dwt
