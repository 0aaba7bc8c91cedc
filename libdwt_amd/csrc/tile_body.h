// Device code shared by the per-level kernels and the pyramid (dataflow)
// kernels: lifting arithmetic, DPP neighbour exchange, the tile bodies and
// their row source / destination policies.
//
// Work decomposition.  A wavefront owns a vertical strip of 128 window
// columns: lane l holds the sample pair (2l, 2l+1) of the window, which
// starts HALO samples left of the strip's core, so the core is the
// 128 - 2*HALO columns whose lifting is complete after NS steps (HALO = NS:
// 4 for CDF 9/7, 2 for CDF 5/3).  The wave walks down a segment of S rows
// (plus HALO rows above and below), two rows at a time:
//   * horizontal lifting of each row runs in registers; the neighbour sample
//     across a lane boundary comes from the adjacent lane through a DPP
//     wave shift, so no LDS is used;
//   * vertical lifting is a per-lane sliding window: each new row pair
//     advances every lifting step by one position and completes one output
//     pair (L, H) of every column, so the column pass costs no extra memory
//     traffic either.
// Line ends use whole-sample symmetric extension: rows and columns outside
// the image are read from their mirror images, which reproduces the
// reference's end-of-line update  x + (2c) * n  bit for bit.
//
// Four waves of a workgroup take four adjacent strips of the same segment.
// Tiles are numbered row-major and the level kernels remap the block index
// so that each XCD works on horizontal bands of the level, read row after
// row, whose halo rows and columns hit that XCD's L2 (tile_kernels.hip).
//
// A tile's strips and segments can start at an offset (xoff, yoff: even) so
// that a fused multi-level kernel computes a window of a level rather than
// the whole of it; every index stays in level coordinates.
#pragma once

#include <type_traits>

#include "dwt_common.h"
#include "kernels.h"

namespace dwt {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;

// Hardware assigns block b to XCD (b mod 8); give every XCD a contiguous run
// of logical tiles.
__device__ inline int xcd_remap(int b, int nb) {
  const int xcd = b & 7, idx = b >> 3;
  const int per = nb >> 3, rem = nb & 7;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// Value held by lane+1 / lane-1 of the wavefront, as a DPP modifier
// (wave_shl:1 / wave_shr:1) rather than an LDS-crossbar permute.  With the
// out-of-range lane reading zero (bound_ctrl) and full row / bank masks, the
// compiler folds the move into the add that consumes it (v_add_f32_dpp), so a
// lifting step's neighbour costs no instruction of its own.  The edge lanes
// (0 and 63) receive 0; they are halo lanes whose results are never stored.
__device__ inline int dpp_from_next(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, true);  // wave_shl:1
}
__device__ inline int dpp_from_prev(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, true);  // wave_shr:1
}
__device__ inline float from_next_lane(float v) {
  return __int_as_float(dpp_from_next(__float_as_int(v)));
}
__device__ inline float from_prev_lane(float v) {
  return __int_as_float(dpp_from_prev(__float_as_int(v)));
}
__device__ inline double from_next_lane(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_from_next((int)(b & 0xffffffffLL)), hi = dpp_from_next((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double from_prev_lane(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_from_prev((int)(b & 0xffffffffLL)), hi = dpp_from_prev((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Pointers typed in the global address space.  Kernel arguments are known to
// point to global memory, but a pointer read from a table (the pyramid
// kernels' task list) is generic, and a generic access compiles to a flat
// instruction; the policies below cast to these types so that every access
// is a global_* instruction whatever the pointer's origin.
template <typename T>
using Gbl = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ const Gbl<T>* gload_ptr(const T* p) {
  return (const Gbl<T>*)p;
}
template <typename T>
__device__ __forceinline__ Gbl<T>* gstore_ptr(T* p) {
  return (Gbl<T>*)p;
}

// Output stores.  DWT_STORE_POLICY selects the cache policy of every store
// of a level's results: 0 plain (write-back L2), 1 non-temporal, 2 agent
// scope (written through to memory, coherent across XCDs, so the end of the
// kernel has no dirty L2 lines to write back before the next level starts).
#ifndef DWT_STORE_POLICY
#define DWT_STORE_POLICY 0
#endif
template <typename T>
__device__ __forceinline__ void st_global(T* p, T v) {
#if DWT_STORE_POLICY == 1
  __builtin_nontemporal_store(v, p);
#elif DWT_STORE_POLICY == 2
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
  *gstore_ptr(p) = v;
#endif
}

// Rows of a RiderCopy (kernels.h) moved by one of its `nwg` extra
// workgroups, whose index among them is k: rows r0 + k, r0 + k + nwg, ...
typedef float Vec16f __attribute__((ext_vector_type(4)));
typedef double Vec16d __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ void rider_copy(const RiderCopy<T>& rc, int k, int nwg) {
  using V = typename std::conditional<sizeof(T) == 4, Vec16f, Vec16d>::type;
  constexpr int PER = 16 / sizeof(T);  // elements of one 16-byte vector
  for (int i = rc.r0 + k; i < rc.r1; i += nwg) {
    const int y = rc.rows ? rc.rows[i] : i;
    const int x0 = y < rc.ha ? rc.xa : 0, w = y < rc.ha ? rc.wa : rc.wb;
    const Gbl<T>* s = gload_ptr(rc.src) + (long long)y * rc.sp + x0;
    Gbl<T>* d = gstore_ptr(rc.dst) + (long long)y * rc.dp + x0;
    if (rc.vec) {
      const int nv = w / PER;
      for (int i = threadIdx.x; i < nv; i += blockDim.x)
        reinterpret_cast<Gbl<V>*>(d)[i] = reinterpret_cast<const Gbl<V>*>(s)[i];
      for (int i = nv * PER + threadIdx.x; i < w; i += blockDim.x) d[i] = s[i];
    } else {
      for (int i = threadIdx.x; i < w; i += blockDim.x) d[i] = s[i];
    }
  }
}

// Two-element vectors as clang ext vectors (HIP's vector classes have
// assignment operators only for generic-address-space objects).
template <typename T> struct Vec2;
template <> struct Vec2<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct Vec2<double> { typedef double type __attribute__((ext_vector_type(2))); };
// The same pair for global memory, aligned to one element only.  gfx950 does
// multi-dword global loads and stores at dword alignment (the compiler emits
// global_load_dwordx2 for a pointer known to be only 4-byte aligned), so a
// pair that starts at an odd element -- every other row of an image with an
// odd pitch -- still moves as one access instead of two.
template <typename T> struct GPair;
template <> struct GPair<float> { typedef float type __attribute__((ext_vector_type(2), aligned(4))); };
template <> struct GPair<double> { typedef double type __attribute__((ext_vector_type(2), aligned(8))); };

// Forward lifting of one row: lane holds (even, odd) samples; odd samples are
// updated first.  Result: (L, H) coefficient of the lane's pair.
template <typename T, int NS>
__device__ inline void row_fwd(T& e, T& o, const Coeffs<T>& cf) {
#pragma unroll
  for (int s = 0; s < NS; s += 2) {
    {
      T t = e + from_next_lane(e);
      t = cf.c[s] * t;
      o = o + t;
    }
    {
      T t = from_prev_lane(o) + o;
      t = cf.c[s + 1] * t;
      e = e + t;
    }
  }
  e = e * cf.se;
  o = o * cf.so;
}

// Inverse lifting of one interleaved row: scale, then even samples first.
template <typename T, int NS>
__device__ inline void row_inv(T& e, T& o, const Coeffs<T>& cf) {
  e = e * cf.se;
  o = o * cf.so;
#pragma unroll
  for (int s = 0; s < NS; s += 2) {
    {
      T t = from_prev_lane(o) + o;
      t = cf.c[s] * t;
      e = e + t;
    }
    {
      T t = e + from_next_lane(e);
      t = cf.c[s + 1] * t;
      o = o + t;
    }
  }
}

// Vertical forward lifting of one column as a sliding window.  push() takes
// the samples of rows 2p and 2p+1 and returns the finished pair p - NS/2.
template <typename T, int NS>
struct ColFwd {
  T e0 = 0, o0 = 0, o1 = 0, e2 = 0, o3 = 0;
  __device__ inline void push(T ze, T zo, const Coeffs<T>& cf, T& L, T& H) {
    T t1 = e0 + ze;
    t1 = cf.c[0] * t1;
    t1 = o0 + t1;          // step 1 at row 2p-1
    T t2 = o1 + t1;
    t2 = cf.c[1] * t2;
    t2 = e0 + t2;          // step 2 at row 2p-2
    if (NS == 4) {
      T t3 = e2 + t2;
      t3 = cf.c[2] * t3;
      t3 = o1 + t3;        // step 3 at row 2p-3
      T t4 = o3 + t3;
      t4 = cf.c[3] * t4;
      t4 = e2 + t4;        // step 4 at row 2p-4
      L = t4 * cf.se;
      H = t3 * cf.so;
      e2 = t2;
      o3 = t3;
    } else {
      L = t2 * cf.se;
      H = t1 * cf.so;
    }
    e0 = ze;
    o0 = zo;
    o1 = t1;
  }
};

// Vertical inverse lifting as a sliding window.  push() takes the (unscaled)
// coefficients of interleaved rows 2p (L) and 2p+1 (H) and returns two
// finished output rows: (2p-3, 2p-2) for NS = 4, (2p-2, 2p-1) for NS = 2,
// as (first, second).
template <typename T, int NS>
struct ColInv {
  T zp = 0, e1 = 0, o2 = 0, e3 = 0;
  __device__ inline void push(T ze, T zo, const Coeffs<T>& cf, T& first, T& second) {
    ze = ze * cf.se;
    zo = zo * cf.so;
    T a1 = zp + zo;
    a1 = cf.c[0] * a1;
    a1 = ze + a1;          // even step at row 2p
    T a2 = e1 + a1;
    a2 = cf.c[1] * a2;
    a2 = zp + a2;          // odd step at row 2p-1
    if (NS == 4) {
      T a3 = o2 + a2;
      a3 = cf.c[2] * a3;
      a3 = e1 + a3;        // even step at row 2p-2
      T a4 = e3 + a3;
      a4 = cf.c[3] * a4;
      a4 = o2 + a4;        // odd step at row 2p-3
      first = a4;
      second = a3;
      o2 = a2;
      e3 = a3;
    } else {
      first = e1;
      second = a2;
    }
    zp = zo;
    e1 = a1;
  }
};

// Detail subbands, by their Mallat position: HL top right (horizontal
// highpass), LH bottom left, HH bottom right.
constexpr int kHL = 0, kLH = 1, kHH = 2;

// ---- level bodies -----------------------------------------------------------
// One wavefront computes one tile: the core columns of strip `strip` and the
// rows of segment `seg` (S rows).  Where rows come from and where the LL
// coefficients go are policies (global or LDS), so that the per-level kernels
// and the single-workgroup tail kernel share the arithmetic.  Each policy
// addresses memory through a pointer of one known address space.
//
// LoadRow(Y, Xe, xe, xo, interior, e, o): samples (Xe, Xe + 1) of row Y of the
//   level input, Y in [0, H) (the tile body mirrors rows beyond the ends); xe /
//   xo are the mirrored column indices, `interior` says the wave's whole
//   window lies inside the row.
// StoreLL(m, k, v): LL coefficient (row m, column k) of the level output.
// Det.store(band, r, c, v) / Det.load(band, r, c): a detail coefficient at row
//   r, column c of the level's Mallat layout (band: kHL, kLH or kHH).
template <typename T, int NS, int U, bool INTERIOR, class LoadRow, class StoreLL, class Det>
__device__ __forceinline__ void fwd_tile_impl(int strip, int seg, int S, int W, int H,
                                              const LoadRow& load_row_in, const StoreLL& store_ll,
                                              const Det& det, const Coeffs<T>& cf, int xoff, int yoff) {
  constexpr int HALO = NS;
  constexpr int CORE = 2 * kWave - 2 * HALO;
  constexpr int LAG = NS / 2;
  const int lane = threadIdx.x & (kWave - 1);
  const int x0 = xoff + strip * CORE;
  const int Xe = x0 - HALO + 2 * lane;
  const int xe = INTERIOR ? Xe : mirror(Xe, W), xo = INTERIOR ? Xe + 1 : mirror(Xe + 1, W);
  const int Wd = (W + 1) >> 1, Wh = W >> 1, Hd = (H + 1) >> 1, Hh = H >> 1;
  const int k = Xe >> 1;
  const bool core = lane >= HALO / 2 && lane < kWave - HALO / 2;
  const bool st_a = core && k < Wd;
  const bool st_b = core && k < Wh;
  constexpr bool interior = INTERIOR;

  const int m0 = (yoff + seg * S) >> 1;
  const int m1 = min((yoff + seg * S + S) >> 1, Hd);
  const int pbeg = m0 - LAG, pend = m1 + LAG - 1;

  // INTERIOR: every row and column the tile reads lies inside the level
  auto load_row = [&](int Y, int Xe_, int xe_, int xo_, bool in, T& e, T& o) {
    load_row_in(INTERIOR ? Y : mirror_near(Y, H), Xe_, xe_, xo_, in, e, o);
  };
  ColFwd<T, NS> ca, cb;
  auto pair = [&](int p, T e0, T o0, T e1, T o1) {
    row_fwd<T, NS>(e0, o0, cf);
    row_fwd<T, NS>(e1, o1, cf);
    T La, Ha, Lb, Hb;
    ca.push(e0, e1, cf, La, Ha);
    cb.push(o0, o1, cf, Lb, Hb);
    const int m = p - LAG;
    if (m >= m0 && m < m1) {
      if (st_a) {
        store_ll(m, k, La);
        if (m < Hh) det.store(kLH, Hd + m, k, Ha);
      }
      if (st_b) {
        det.store(kHL, m, Wd + k, Lb);
        if (m < Hh) det.store(kHH, Hd + m, Wd + k, Hb);
      }
    }
  };

  if constexpr (U < 0) {
    // single shot: every row of the segment (S = -U rows plus the halo) is
    // requested before any is used, one memory round trip per wavefront
    constexpr int NP = -U / 2 + NS;
    T re[2 * NP], ro[2 * NP];
#pragma unroll
    for (int u = 0; u < 2 * NP; ++u) load_row(2 * pbeg + u, Xe, xe, xo, interior, re[u], ro[u]);
#pragma unroll
    for (int u = 0; u < NP; ++u) pair(pbeg + u, re[2 * u], ro[2 * u], re[2 * u + 1], ro[2 * u + 1]);
  } else {
    // rolling: U row pairs in flight while the previous U are computed
    T ce[2 * U], co[2 * U];
#pragma unroll
    for (int u = 0; u < 2 * U; ++u) load_row(2 * pbeg + u, Xe, xe, xo, interior, ce[u], co[u]);
    for (int pb = pbeg; pb <= pend; pb += U) {
      T ne[2 * U], no[2 * U];
#pragma unroll
      for (int u = 0; u < 2 * U; ++u)
        load_row(2 * (pb + U) + u, Xe, xe, xo, interior, ne[u], no[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) pair(pb + u, ce[2 * u], co[2 * u], ce[2 * u + 1], co[2 * u + 1]);
#pragma unroll
      for (int u = 0; u < 2 * U; ++u) {
        ce[u] = ne[u];
        co[u] = no[u];
      }
    }
  }
}

// Rows [first, last] that a tile whose segment starts at row ys reads
// (before mirroring):
// the segment's row pairs plus LAG pairs of halo on each side, and for the
// rolling mode the prefetch of up to 2U more pairs past the end.
template <int NS, int U>
__device__ __forceinline__ bool rows_inside(int ys, int S, int H, bool fwd) {
  constexpr int LAG = NS / 2;
  int first, last;
  if (fwd) {
    const int m0 = ys >> 1, m1 = min((ys + S) >> 1, (H + 1) >> 1);
    const int pbeg = m0 - LAG, pend = m1 + LAG - 1;
    first = 2 * pbeg;
    last = U < 0 ? 2 * (pbeg + (-U / 2 + NS)) - 1 : 2 * (pend + 2 * U) - 1;
  } else {
    const int y0 = ys, y1 = min(y0 + S, H);
    const int pbeg = (y0 >> 1) - LAG, pend = (y1 >> 1) + LAG + 1;
    first = 2 * pbeg;
    last = U < 0 ? 2 * (pbeg + (-U / 2 + NS)) - 1 : 2 * (pend + 2 * U) - 1;
  }
  return first >= 0 && last < H;
}

// The whole 128-column window of a strip either lies inside the rows (plain,
// vectorisable loads) or touches an end (mirrored columns); likewise the
// tile's rows.  The test is the same for every lane of the wavefront;
// readfirstlane makes it a scalar branch between two separately compiled
// bodies, so interior tiles (almost all of a large level) spend no
// instructions on mirroring.
template <typename T, int NS, int U, class LoadRow, class StoreLL, class Det>
__device__ __forceinline__ void fwd_tile(int strip, int seg, int S, int W, int H,
                                         const LoadRow& load_row, const StoreLL& store_ll,
                                         const Det& det, const Coeffs<T>& cf, int xoff = 0, int yoff = 0) {
  constexpr int HALO = NS;
  const int x0 = xoff + strip * (2 * kWave - 2 * HALO);
  const int interior = __builtin_amdgcn_readfirstlane((x0 - HALO >= 0) && (x0 - HALO + 2 * kWave <= W) &&
                                                      rows_inside<NS, U>(yoff + seg * S, S, H, true));
  if (interior)
    fwd_tile_impl<T, NS, U, true>(strip, seg, S, W, H, load_row, store_ll, det, cf, xoff, yoff);
  else
    fwd_tile_impl<T, NS, U, false>(strip, seg, S, W, H, load_row, store_ll, det, cf, xoff, yoff);
}

// Inverse tile: output columns of strip `strip`, rows of segment `seg` of the
// W x H fine region.  LoadLL(m, k): LL coefficient; details come from det at
// their Mallat positions.  StoreOut(r, Xe, e, o, st_e, st_o): output samples
// (Xe, Xe + 1) of row r.
template <typename T, int NS, int U, bool ROWS_IN, class LoadLL, class Det, class StoreOut>
__device__ __forceinline__ void inv_tile_impl(int strip, int seg, int S, int W, int H,
                                              const LoadLL& load_ll, const Det& det,
                                              const StoreOut& store_out, const Coeffs<T>& cf, int xoff, int yoff) {
  constexpr int HALO = NS;
  constexpr int CORE = 2 * kWave - 2 * HALO;
  constexpr int LAG = NS / 2;
  const int lane = threadIdx.x & (kWave - 1);
  const int x0 = xoff + strip * CORE;
  const int Xe = x0 - HALO + 2 * lane;
  const int ke = mirror(Xe, W) >> 1, ko = mirror(Xe + 1, W) >> 1;
  const int Wc = (W + 1) >> 1, Hc = (H + 1) >> 1;
  const bool core = lane >= HALO / 2 && lane < kWave - HALO / 2;
  const bool st_e = core && Xe < W;
  const bool st_o = core && Xe + 1 < W;

  const int y0 = yoff + seg * S;
  const int y1 = min(y0 + S, H);
  const int pbeg = (y0 >> 1) - LAG, pend = (y1 >> 1) + LAG + 1;

  auto load_row = [&](int Y, T& e, T& o) {
    const int Yr = ROWS_IN ? Y : mirror_near(Y, H);
    const int m = Yr >> 1;
    if (Yr & 1) {
      e = det.load(kLH, Hc + m, ke);
      o = det.load(kHH, Hc + m, Wc + ko);
    } else {
      e = load_ll(m, ke);
      o = det.load(kHL, m, Wc + ko);
    }
  };

  ColInv<T, NS> cle, clo;
  auto pair = [&](int p, T e0, T o0, T e1, T o1) {
    row_inv<T, NS>(e0, o0, cf);
    row_inv<T, NS>(e1, o1, cf);
    T fe, se_, fo, so_;
    cle.push(e0, e1, cf, fe, se_);
    clo.push(o0, o1, cf, fo, so_);
    const int r0 = NS == 4 ? 2 * p - 3 : 2 * p - 2;
    if (core) {
      if (r0 >= y0 && r0 < y1) store_out(r0, Xe, fe, fo, st_e, st_o);
      if (r0 + 1 >= y0 && r0 + 1 < y1) store_out(r0 + 1, Xe, se_, so_, st_e, st_o);
    }
  };

  if constexpr (U < 0) {
    constexpr int NP = -U / 2 + NS;  // pairs y0/2 - LAG .. y0/2 + S/2 + LAG - 1
    T re[2 * NP], ro[2 * NP];
#pragma unroll
    for (int u = 0; u < 2 * NP; ++u) load_row(2 * pbeg + u, re[u], ro[u]);
#pragma unroll
    for (int u = 0; u < NP; ++u) pair(pbeg + u, re[2 * u], ro[2 * u], re[2 * u + 1], ro[2 * u + 1]);
  } else {
    T ce[2 * U], co[2 * U];
#pragma unroll
    for (int u = 0; u < 2 * U; ++u) load_row(2 * pbeg + u, ce[u], co[u]);
    for (int pb = pbeg; pb <= pend; pb += U) {
      T ne[2 * U], no[2 * U];
#pragma unroll
      for (int u = 0; u < 2 * U; ++u) load_row(2 * (pb + U) + u, ne[u], no[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) pair(pb + u, ce[2 * u], co[2 * u], ce[2 * u + 1], co[2 * u + 1]);
#pragma unroll
      for (int u = 0; u < 2 * U; ++u) {
        ce[u] = ne[u];
        co[u] = no[u];
      }
    }
  }
}

// Rows of the tile inside the level (almost every tile): no mirroring.
template <typename T, int NS, int U, class LoadLL, class Det, class StoreOut>
__device__ __forceinline__ void inv_tile(int strip, int seg, int S, int W, int H,
                                         const LoadLL& load_ll, const Det& det,
                                         const StoreOut& store_out, const Coeffs<T>& cf, int xoff = 0,
                                         int yoff = 0) {
  if (__builtin_amdgcn_readfirstlane(rows_inside<NS, U>(yoff + seg * S, S, H, false)))
    inv_tile_impl<T, NS, U, true>(strip, seg, S, W, H, load_ll, det, store_out, cf, xoff, yoff);
  else
    inv_tile_impl<T, NS, U, false>(strip, seg, S, W, H, load_ll, det, store_out, cf, xoff, yoff);
}

// ---- small levels: tile height by latency ------------------------------------
// A level of a few tiles is latency-bound: its time is the number of tile
// rounds a wavefront runs times the row pairs of one tile (S/2 + NS, the
// tile's serial chain).  Short segments give more, shorter tiles; pick the
// height in {8, 4, 2} with the shortest chain for nw wavefronts, then run
// f(std::integral_constant<int, S>, strip, seg) for this wavefront's tiles.
template <int NS>
__device__ __forceinline__ int small_rows(int nstrips, int rows, int nw) {
  int best = 8, cost = 1 << 30;
  for (int S = 8; S >= 2; S >>= 1) {
    const int nt = nstrips * ((rows + S - 1) / S);
    const int c = ((nt + nw - 1) / nw) * (S / 2 + NS);
    if (c < cost) {
      cost = c;
      best = S;
    }
  }
  return best;
}
template <int NS, class F>
__device__ __forceinline__ void for_small_tiles(int nstrips, int rows, int wave, int nw, const F& f) {
  const int S = __builtin_amdgcn_readfirstlane(small_rows<NS>(nstrips, rows, nw));
  auto run = [&](auto sc) {
    constexpr int SR = decltype(sc)::value;
    const int ntiles = nstrips * ((rows + SR - 1) / SR);
    for (int t = wave; t < ntiles; t += nw) {
      const int seg = t / nstrips;
      f(sc, t - seg * nstrips, seg);
    }
  };
  if (S == 8)
    run(std::integral_constant<int, 8>{});
  else if (S == 4)
    run(std::integral_constant<int, 4>{});
  else
    run(std::integral_constant<int, 2>{});
}

// ---- row sources and destinations ------------------------------------------------
template <typename T, bool VEC>
struct GlobalRows {  // packed rows in global memory
  const T* __restrict__ p;
  long long pitch;
  int H;
  __device__ __forceinline__ void operator()(int Y, int Xe, int xe, int xo, bool interior, T& e,
                                             T& o) const {
    using V2 = typename GPair<T>::type;
    const Gbl<T>* rp = gload_ptr(p) + (long long)Y * pitch;
    if (VEC && interior) {
      const V2 v = *(const Gbl<V2>*)(rp + Xe);
      e = v.x;
      o = v.y;
    } else {
      e = rp[xe];
      o = rp[xo];
    }
  }
};
// LDS policies hold pointers typed in the LDS address space, so every access
// through them is a ds_* instruction: a generic pointer would let the compiler
// fold an LDS load and a global load of the two arms of a branch into one flat
// load, and flat access to LDS is what the codegen guard forbids.
template <typename T>
using Lds = __attribute__((address_space(3))) T;

template <typename T>
struct LdsRows {  // packed rows in LDS (pitch = W)
  const Lds<T>* p;
  int pitch, H;
  __device__ __forceinline__ void operator()(int Y, int, int xe, int xo, bool, T& e, T& o) const {
    const Lds<T>* rp = p + Y * pitch;
    e = rp[xe];
    o = rp[xo];
  }
};
template <typename T>
struct GlobalStore {  // element store, global
  T* __restrict__ p;
  long long pitch;
  __device__ __forceinline__ void operator()(int m, int k, T v) const {
    st_global(p + (long long)m * pitch + k, v);
  }
};
template <typename T>
struct LdsStore {
  Lds<T>* p;
  int pitch;
  __device__ __forceinline__ void operator()(int m, int k, T v) const { p[m * pitch + k] = v; }
};
template <typename T>
struct GlobalLoad {
  const T* __restrict__ p;
  long long pitch;
  __device__ __forceinline__ T operator()(int m, int k) const {
    return gload_ptr(p)[(long long)m * pitch + k];
  }
};
template <typename T>
struct LdsLoad {
  const Lds<T>* p;
  int pitch;
  __device__ __forceinline__ T operator()(int m, int k) const { return p[m * pitch + k]; }
};
template <typename T>
struct GlobalDet {  // detail subbands in the storage type, global
  T* __restrict__ p;
  long long pitch;
  __device__ __forceinline__ void store(int, int r, int c, T v) const {
    st_global(p + (long long)r * pitch + c, v);
  }
  __device__ __forceinline__ T load(int, int r, int c) const {
    return gload_ptr(static_cast<const T*>(p))[(long long)r * pitch + c];
  }
};
template <typename T, bool VEC>
struct GlobalOut {  // two adjacent output samples, global
  T* __restrict__ p;
  long long pitch;
  __device__ __forceinline__ void operator()(int r, int Xe, T e, T o, bool st_e, bool st_o) const {
    using V2 = typename GPair<T>::type;
    T* rp = p + (long long)r * pitch + Xe;
#if DWT_STORE_POLICY == 0
    if (VEC && st_o) {
      V2 v;
      v.x = e;
      v.y = o;
      *reinterpret_cast<Gbl<V2>*>(gstore_ptr(rp)) = v;
      return;
    }
#endif
    if (st_e) st_global(rp, e);
    if (st_o) st_global(rp + 1, o);
  }
};
// Write-through (agent-scope) stores for results that other workgroups read
// in the same launch: a store of this kind is in memory, visible on every XCD,
// once it is complete, so the producer needs no L2 write-back before it
// signals the consumers.
template <typename T>
__device__ __forceinline__ void st_coherent(T* p, T v) {
  __hip_atomic_store(gstore_ptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
struct CoherentStore {  // element store, written through
  T* __restrict__ p;
  long long pitch;
  __device__ __forceinline__ void operator()(int m, int k, T v) const {
    st_coherent(p + (long long)m * pitch + k, v);
  }
};
template <typename T, bool VEC>
struct CoherentOut {  // two adjacent output samples, written through
  T* __restrict__ p;
  long long pitch;
  __device__ __forceinline__ void operator()(int r, int Xe, T e, T o, bool st_e, bool st_o) const {
    T* rp = p + (long long)r * pitch + Xe;
    if constexpr (VEC && sizeof(T) == 4) {
      if (st_o) {  // an aligned pair as one 64-bit store
        const unsigned long long bits = (unsigned long long)__float_as_uint(e) |
                                        ((unsigned long long)__float_as_uint(o) << 32);
        st_coherent(reinterpret_cast<unsigned long long*>(rp), bits);
        return;
      }
    }
    if (st_e) st_coherent(rp, e);
    if (st_o) st_coherent(rp + 1, o);
  }
};
template <typename T>
struct LdsDet {  // detail subbands in LDS
  Lds<T>* p;
  int pitch;
  __device__ __forceinline__ void store(int, int r, int c, T v) const { p[r * pitch + c] = v; }
  __device__ __forceinline__ T load(int, int r, int c) const { return p[r * pitch + c]; }
};
template <typename T>
struct LdsOut {
  Lds<T>* p;
  int pitch;
  __device__ __forceinline__ void operator()(int r, int Xe, T e, T o, bool st_e, bool st_o) const {
    Lds<T>* rp = p + r * pitch + Xe;
    if (st_e) rp[0] = e;
    if (st_o) rp[1] = o;
  }
};

}  // namespace dwt
