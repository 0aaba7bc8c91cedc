// Packed fast path: one 2D lifting level per launch, one HBM read of the
// level input and one write of its four subbands.
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
// Four waves of a workgroup take four adjacent strips of the same segment,
// and the block index is remapped so that blocks resident on one XCD walk
// down neighbouring segments (their halo rows then hit that XCD's L2).
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

// Value held by lane+1 / lane-1 of the wavefront, as a DPP modifier on a
// VALU move (wave_shl:1 / wave_shr:1) rather than an LDS-crossbar permute, so
// a lifting step's neighbour costs a few cycles instead of an LDS round trip.
// The edge lanes (0 and 63) receive their own value; they are halo lanes whose
// results are never stored.
__device__ inline int dpp_from_next(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xF, 0xF, false);  // wave_shl:1
}
__device__ inline int dpp_from_prev(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xF, 0xF, false);  // wave_shr:1
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
  *p = v;
#endif
}

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

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

// ---- level bodies -----------------------------------------------------------
// One wavefront computes one tile: the core columns of strip `strip` and the
// rows of segment `seg` (S rows).  Where rows come from and where the LL
// coefficients go are policies (global or LDS), so that the per-level kernels
// and the single-workgroup tail kernel share the arithmetic.  Each policy
// addresses memory through a pointer of one known address space.
//
// LoadRow(Y, Xe, xe, xo, interior, e, o): samples (Xe, Xe + 1) of row Y of the
//   level input, Y possibly outside [0, H) (the policy mirrors it); xe / xo are
//   the mirrored column indices, `interior` says the wave's whole window lies
//   inside the row.
// StoreLL(m, k, v): LL coefficient (row m, column k) of the level output.
template <typename T, int NS, int U, bool INTERIOR, class LoadRow, class StoreLL>
__device__ __forceinline__ void fwd_tile_impl(int strip, int seg, int S, int W, int H,
                                              const LoadRow& load_row, const StoreLL& store_ll,
                                              T* __restrict__ det, long long dp,
                                              const Coeffs<T>& cf) {
  constexpr int HALO = NS;
  constexpr int CORE = 2 * kWave - 2 * HALO;
  constexpr int LAG = NS / 2;
  const int lane = threadIdx.x & (kWave - 1);
  const int x0 = strip * CORE;
  const int Xe = x0 - HALO + 2 * lane;
  const int xe = INTERIOR ? Xe : mirror(Xe, W), xo = INTERIOR ? Xe + 1 : mirror(Xe + 1, W);
  const int Wd = (W + 1) >> 1, Wh = W >> 1, Hd = (H + 1) >> 1, Hh = H >> 1;
  const int k = Xe >> 1;
  const bool core = lane >= HALO / 2 && lane < kWave - HALO / 2;
  const bool st_a = core && k < Wd;
  const bool st_b = core && k < Wh;
  constexpr bool interior = INTERIOR;

  const int m0 = (seg * S) >> 1;
  const int m1 = min((seg * S + S) >> 1, Hd);
  const int pbeg = m0 - LAG, pend = m1 + LAG - 1;

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
        if (m < Hh) st_global(det + (long long)(Hd + m) * dp + k, Ha);
      }
      if (st_b) {
        st_global(det + (long long)m * dp + Wd + k, Lb);
        if (m < Hh) st_global(det + (long long)(Hd + m) * dp + Wd + k, Hb);
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

// The whole 128-column window of a strip either lies inside the rows (plain,
// vectorisable loads) or touches an end (mirrored columns).  The test is the
// same for every lane of the wavefront; readfirstlane makes it a scalar branch
// between two separately compiled bodies.
template <typename T, int NS, int U, class LoadRow, class StoreLL>
__device__ __forceinline__ void fwd_tile(int strip, int seg, int S, int W, int H,
                                         const LoadRow& load_row, const StoreLL& store_ll,
                                         T* __restrict__ det, long long dp, const Coeffs<T>& cf) {
  constexpr int HALO = NS;
  const int x0 = strip * (2 * kWave - 2 * HALO);
  const int interior = __builtin_amdgcn_readfirstlane((x0 - HALO >= 0) && (x0 - HALO + 2 * kWave <= W));
  if (interior)
    fwd_tile_impl<T, NS, U, true>(strip, seg, S, W, H, load_row, store_ll, det, dp, cf);
  else
    fwd_tile_impl<T, NS, U, false>(strip, seg, S, W, H, load_row, store_ll, det, dp, cf);
}

// Inverse tile: output columns of strip `strip`, rows of segment `seg` of the
// W x H fine region.  LoadLL(m, k): LL coefficient; details come from det at
// their Mallat positions.  StoreOut(r, Xe, e, o, st_e, st_o): output samples
// (Xe, Xe + 1) of row r.
template <typename T, int NS, int U, class LoadLL, class StoreOut>
__device__ __forceinline__ void inv_tile(int strip, int seg, int S, int W, int H,
                                         const LoadLL& load_ll, const T* __restrict__ det,
                                         long long dp, const StoreOut& store_out,
                                         const Coeffs<T>& cf) {
  constexpr int HALO = NS;
  constexpr int CORE = 2 * kWave - 2 * HALO;
  constexpr int LAG = NS / 2;
  const int lane = threadIdx.x & (kWave - 1);
  const int x0 = strip * CORE;
  const int Xe = x0 - HALO + 2 * lane;
  const int ke = mirror(Xe, W) >> 1, ko = mirror(Xe + 1, W) >> 1;
  const int Wc = (W + 1) >> 1, Hc = (H + 1) >> 1;
  const bool core = lane >= HALO / 2 && lane < kWave - HALO / 2;
  const bool st_e = core && Xe < W;
  const bool st_o = core && Xe + 1 < W;

  const int y0 = seg * S;
  const int y1 = min(y0 + S, H);
  const int pbeg = (y0 >> 1) - LAG, pend = (y1 >> 1) + LAG + 1;

  auto load_row = [&](int Y, T& e, T& o) {
    const int Yr = mirror_near(Y, H);
    const int m = Yr >> 1;
    if (Yr & 1) {
      const T* r = det + (long long)(Hc + m) * dp;
      e = r[ke];
      o = r[Wc + ko];
    } else {
      e = load_ll(m, ke);
      o = det[(long long)m * dp + Wc + ko];
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

// ---- row sources and destinations ------------------------------------------------
template <typename T, bool VEC>
struct GlobalRows {  // packed rows in global memory
  const T* __restrict__ p;
  long long pitch;
  int H;
  __device__ __forceinline__ void operator()(int Y, int Xe, int xe, int xo, bool interior, T& e,
                                             T& o) const {
    using V2 = typename Vec2<T>::type;
    const T* rp = p + (long long)mirror_near(Y, H) * pitch;
    if (VEC && interior) {
      const V2 v = *reinterpret_cast<const V2*>(rp + Xe);
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
    const Lds<T>* rp = p + mirror_near(Y, H) * pitch;
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
  __device__ __forceinline__ T operator()(int m, int k) const { return p[(long long)m * pitch + k]; }
};
template <typename T>
struct LdsLoad {
  const Lds<T>* p;
  int pitch;
  __device__ __forceinline__ T operator()(int m, int k) const { return p[m * pitch + k]; }
};
template <typename T, bool VEC>
struct GlobalOut {  // two adjacent output samples, global
  T* __restrict__ p;
  long long pitch;
  __device__ __forceinline__ void operator()(int r, int Xe, T e, T o, bool st_e, bool st_o) const {
    using V2 = typename Vec2<T>::type;
    T* rp = p + (long long)r * pitch + Xe;
#if DWT_STORE_POLICY == 0
    if (VEC && st_o) {
      V2 v;
      v.x = e;
      v.y = o;
      *reinterpret_cast<V2*>(rp) = v;
      return;
    }
#endif
    if (st_e) st_global(rp, e);
    if (st_o) st_global(rp + 1, o);
  }
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

// ---- per-level kernels (many workgroups) -------------------------------------
template <typename T, int NS, bool VEC, int U>
__global__ void __launch_bounds__(256)
fwd_level_kernel(const T* __restrict__ src, long long sp, int W, int H, T* __restrict__ ll,
                 long long llp, T* __restrict__ det, long long dp, int S, int nstrips, int nseg,
                 Coeffs<T> cf) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int g = b / nseg, s = b - g * nseg;
  const int strip = g * kWavesPerBlock + (threadIdx.x >> 6);
  if (strip >= nstrips) return;
  fwd_tile<T, NS, U>(strip, s, S, W, H, GlobalRows<T, VEC>{src, sp, H}, GlobalStore<T>{ll, llp},
                     det, dp, cf);
}

template <typename T, int NS, bool VEC, int U>
__global__ void __launch_bounds__(256)
inv_level_kernel(const T* __restrict__ ll, long long llp, const T* __restrict__ det, long long dp,
                 T* __restrict__ dst, long long dstp, int W, int H, int S, int nstrips, int nseg,
                 Coeffs<T> cf) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int g = b / nseg, s = b - g * nseg;
  const int strip = g * kWavesPerBlock + (threadIdx.x >> 6);
  if (strip >= nstrips) return;
  inv_tile<T, NS, U>(strip, s, S, W, H, GlobalLoad<T>{ll, llp}, det, dp,
                     GlobalOut<T, VEC>{dst, dstp}, cf);
}

// ---- block kernels (one workgroup, a whole small pyramid in LDS) --------------
// The small levels of a transform are latency-bound: every kernel boundary on
// this multi-XCD GPU writes back and invalidates L2, and a level of a few KB
// cannot fill the machine.  A block kernel instead moves the whole small Mallat
// region (the LL input of the first small level, or all coefficients of the
// small levels for the inverse) into LDS with one cooperative sweep, runs every
// remaining level there, and writes the region back with one sweep: two global
// round trips for all small levels.  Because the whole region is read before
// anything is written, it may also work in place (an image small enough to fit
// is transformed completely by one launch).
//
// Within LDS a level is computed line-parallel by the whole workgroup, each
// lifting step a barrier-separated sweep (x + c * (l + r), a missing neighbour
// at a line end replaced by the existing one); the parity scaling is folded
// into the move between Mallat and interleaved order through the second LDS
// buffer.  Rows first, then columns, in both directions.
constexpr int kBlockThreads = 1024;

// Lifting steps on the H rows (length W >= 2) of buf: wave w takes rows
// w, w + nwaves, ...; lane l takes the samples of the step's parity at
// positions par + 2l, par + 2l + 128, ...
template <typename T>
__device__ __forceinline__ void lds_lift_rows(Lds<T>* buf, int pitch, int W, int H,
                                              const Coeffs<T>& cf) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int st = 0; st < cf.nsteps; ++st) {
    const int par = (cf.first_parity + st) & 1;
    const T c = cf.c[st];
    const T c2 = T(2) * c;
    for (int y = wave; y < H; y += nw) {
      Lds<T>* row = buf + y * pitch;
      for (int i = par + 2 * lane; i < W; i += 2 * kWave) {
        T t;
        if (i > 0 && i + 1 < W) {
          t = row[i - 1] + row[i + 1];
          t = c * t;
        } else {
          t = c2 * (i > 0 ? row[i - 1] : row[i + 1]);
        }
        row[i] = row[i] + t;
      }
    }
    __syncthreads();
  }
}

// Lifting steps on the W columns (length H >= 2) of buf: lane l takes columns
// l, l + 64, ... (consecutive lanes, consecutive LDS banks); wave w takes the
// step's positions par + 2w, par + 2w + 2 nwaves, ...
template <typename T>
__device__ __forceinline__ void lds_lift_cols(Lds<T>* buf, int pitch, int W, int H,
                                              const Coeffs<T>& cf) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int st = 0; st < cf.nsteps; ++st) {
    const int par = (cf.first_parity + st) & 1;
    const T c = cf.c[st];
    const T c2 = T(2) * c;
    for (int x = lane; x < W; x += kWave) {
      for (int i = par + 2 * wave; i < H; i += 2 * nw) {
        Lds<T>* e = buf + i * pitch + x;
        T t;
        if (i > 0 && i + 1 < H) {
          t = e[-pitch] + e[pitch];
          t = c * t;
        } else {
          t = c2 * (i > 0 ? e[-pitch] : e[pitch]);
        }
        e[0] = e[0] + t;
      }
    }
    __syncthreads();
  }
}

// Permutation with the parity scaling folded in, over the W x H region.
// Forward (deinterleave): out[.., j] = in[.., i] * (i odd ? so : se) with even
// i to j = i/2 and odd i to j = ceil(N/2) + i/2.  Inverse (interleave):
// out[.., j] = in[.., i] * (j odd ? so : se) with i < ceil(N/2) to j = 2i, the
// rest to j = 2(i - ceil(N/2)) + 1.  Along rows (ROWS) or columns.
template <typename T, bool ROWS, bool DEINTERLEAVE>
__device__ __forceinline__ void lds_permute_scale(const Lds<T>* in, Lds<T>* out, int pitch, int W,
                                                  int H, T se, T so) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int N = ROWS ? W : H, half = (N + 1) >> 1;
  for (int y = wave; y < H; y += nw) {
    for (int x = lane; x < W; x += kWave) {
      const int i = ROWS ? x : y;
      int j;
      bool odd;
      if (DEINTERLEAVE) {
        j = (i & 1) ? half + (i >> 1) : (i >> 1);
        odd = i & 1;
      } else {
        j = i < half ? 2 * i : 2 * (i - half) + 1;
        odd = j & 1;
      }
      const int yo = ROWS ? y : j, xo = ROWS ? j : x;
      out[yo * pitch + xo] = in[y * pitch + x] * (odd ? so : se);
    }
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void block_load(const T* __restrict__ src, long long sp, Lds<T>* b,
                                           int W, int H) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int y = wave; y < H; y += nw)
    for (int x = lane; x < W; x += kWave) b[y * W + x] = src[(long long)y * sp + x];
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void block_store(const Lds<T>* b, T* __restrict__ dst, long long dp,
                                            int W, int H) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int y = wave; y < H; y += nw)
    for (int x = lane; x < W; x += kWave) st_global(dst + (long long)y * dp + x, (T)b[y * W + x]);
}

// Forward: levels 0..nlev-1 of the W0 x H0 region read from src (pitch sp),
// result (Mallat layout) to dst (pitch dp); src may equal dst.
template <typename T>
__global__ void __launch_bounds__(kBlockThreads)
block_fwd_kernel(const T* __restrict__ src, long long sp, T* __restrict__ dst, long long dp, int W0,
                 int H0, int nlev, Coeffs<T> cf) {
  extern __shared__ __align__(16) unsigned char blk_smem[];
  Lds<T>* const b0 = (Lds<T>*)blk_smem;  // generic -> LDS address-space cast
  Lds<T>* const b1 = b0 + W0 * H0;
  block_load(src, sp, b0, W0, H0);
  for (int lev = 0; lev < nlev; ++lev) {
    const int W = ceil_div_pow2(W0, lev), H = ceil_div_pow2(H0, lev);
    lds_lift_rows(b0, W0, W, H, cf);  // rows: lift, then scale into Mallat order
    lds_permute_scale<T, true, true>(b0, b1, W0, W, H, cf.se, cf.so);
    lds_lift_cols(b1, W0, W, H, cf);  // columns
    lds_permute_scale<T, false, true>(b1, b0, W0, W, H, cf.se, cf.so);
  }
  block_store(b0, dst, dp, W0, H0);
}

// Inverse: the nlev levels whose finest output is W0 x H0; coefficients of the
// W0 x H0 region from src (pitch sp), output to dst (pitch dp); src may equal dst.
template <typename T>
__global__ void __launch_bounds__(kBlockThreads)
block_inv_kernel(const T* __restrict__ src, long long sp, T* __restrict__ dst, long long dp, int W0,
                 int H0, int nlev, Coeffs<T> cf) {
  extern __shared__ __align__(16) unsigned char blk_smem[];
  Lds<T>* const b0 = (Lds<T>*)blk_smem;  // generic -> LDS address-space cast
  Lds<T>* const b1 = b0 + W0 * H0;
  block_load(src, sp, b0, W0, H0);
  for (int lev = nlev - 1; lev >= 0; --lev) {
    const int W = ceil_div_pow2(W0, lev), H = ceil_div_pow2(H0, lev);
    lds_permute_scale<T, true, false>(b0, b1, W0, W, H, cf.se, cf.so);  // rows: order, scale,
    lds_lift_rows(b1, W0, W, H, cf);                                    // lift
    lds_permute_scale<T, false, false>(b1, b0, W0, W, H, cf.se, cf.so);  // columns
    lds_lift_cols(b0, W0, W, H, cf);
  }
  block_store(b0, dst, dp, W0, H0);
}

template <typename T>
__global__ void copy_rect_kernel(const T* __restrict__ src, long long sp, T* __restrict__ dst,
                                 long long dp, int w, int h) {
  const long long n = (long long)w * h;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int y = (int)(t / w), x = (int)(t - (long long)y * w);
    st_global(dst + y * dp + x, src[y * sp + x]);
  }
}

// ---- host launchers ---------------------------------------------------------

// Tile shape of one launch: S rows per segment (tall segments amortise the
// 2*HALO halo rows, short ones give a level more wavefronts) and U row pairs
// prefetched per wavefront (U = 8 keeps more bytes in flight and lets a short
// segment issue all its loads at once).  The table is indexed by the level's
// height class ceil(log2(H)); entries are set by dwt_gpu_tile_set (the
// autotuner) and default to the heuristic below.
struct TileShape {
  int S, U;
};
static TileShape g_tiles[2][32];  // [direction][height class]; S == 0: heuristic
static int height_class(int H) {
  int c = 0;
  while ((1 << c) < H && c < 31) ++c;
  return c;
}

static TileShape pick_tile(int H, int nstrips, bool fwd) {
  const TileShape t = g_tiles[fwd ? 0 : 1][height_class(H)];
  if (t.S > 0) return t;
  // default from autotuning sweeps on MI355X (4096^2 and 3001x2047, both
  // wavelets, both directions): 8-row segments loaded in a single shot were
  // fastest or within noise on every level size
  (void)nstrips;
  (void)H;
  return TileShape{8, 0};
}

int tile_set(int fwd, int hclass, int S, int U) {
  if (hclass < 0 || hclass >= 32) return -1;
  if (S != 0 && (S < 2 || (S & 1) || S > 4096)) return -1;
  // U = 4 or 8: rolling prefetch of U row pairs; U = 0: single shot, the
  // whole segment requested at once (S = 8, 16 or 32)
  if (S != 0 && U != 4 && U != 8 && !(U == 0 && (S == 8 || S == 16 || S == 32))) return -1;
  g_tiles[fwd ? 0 : 1][hclass] = TileShape{S, U};
  return 0;
}

int tile_get(int fwd, int H, int W, int nsteps, int* S, int* U) {
  const int core = 2 * kWave - 2 * nsteps;
  const TileShape t = pick_tile(H, (W + core - 1) / core, fwd != 0);
  *S = t.S;
  *U = t.U;
  return 0;
}

template <typename T>
static bool aligned_pair(const void* p, long long pitch) {
  return (reinterpret_cast<uintptr_t>(p) % (2 * sizeof(T))) == 0 && (pitch % 2) == 0;
}

// Instantiation for a tile shape: rolling prefetch (U = 4 / 8, any S) or
// single shot (U = 0, S in {8, 16, 32}, template U = -S).
template <typename T, int NS, bool VEC, class Launch>
static void with_shape(const TileShape& t, const Launch& launch) {
  if (t.U == 0) {
    if (t.S == 8) launch.template go<T, NS, VEC, -8>();
    else if (t.S == 16) launch.template go<T, NS, VEC, -16>();
    else launch.template go<T, NS, VEC, -32>();
  } else if (t.U == 8) {
    launch.template go<T, NS, VEC, 8>();
  } else {
    launch.template go<T, NS, VEC, 4>();
  }
}

template <typename T, class Launch>
static void with_variant(int nsteps, bool vec, const TileShape& t, const Launch& launch) {
  if (nsteps == 4) {
    if (vec) with_shape<T, 4, true>(t, launch);
    else with_shape<T, 4, false>(t, launch);
  } else {
    if (vec) with_shape<T, 2, true>(t, launch);
    else with_shape<T, 2, false>(t, launch);
  }
}

template <typename T>
struct FwdLaunch {
  const T* src; long long sp; int W, H; T* ll; long long llp; T* det; long long dp;
  int S, nstrips, nseg; dim3 grid, block; Coeffs<T> cf; hipStream_t st;
  template <typename TT, int NS, bool VEC, int U>
  void go() const {
    hipLaunchKernelGGL((fwd_level_kernel<TT, NS, VEC, U>), grid, block, 0, st, src, sp, W, H, ll,
                       llp, det, dp, S, nstrips, nseg, cf);
  }
};

template <typename T>
struct InvLaunch {
  const T* ll; long long llp; const T* det; long long dp; T* dst; long long dstp; int W, H;
  int S, nstrips, nseg; dim3 grid, block; Coeffs<T> cf; hipStream_t st;
  template <typename TT, int NS, bool VEC, int U>
  void go() const {
    hipLaunchKernelGGL((inv_level_kernel<TT, NS, VEC, U>), grid, block, 0, st, ll, llp, det, dp,
                       dst, dstp, W, H, S, nstrips, nseg, cf);
  }
};

template <typename T>
void launch_fwd_level(const T* src, long long sp, int W, int H, T* ll, long long llp, T* det,
                      long long dp, const Coeffs<T>& cf, hipStream_t st) {
  const int core = 2 * kWave - 2 * cf.nsteps;
  const int nstrips = (W + core - 1) / core;
  const TileShape t = pick_tile(H, nstrips, true);
  const int nseg = (H + t.S - 1) / t.S;
  const int ngroups = (nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
  const FwdLaunch<T> L{src, sp, W, H, ll, llp, det, dp, t.S, nstrips, nseg,
                       dim3(ngroups * nseg), dim3(kWave * kWavesPerBlock), cf, st};
  with_variant<T>(cf.nsteps, aligned_pair<T>(src, sp), t, L);
}

template <typename T>
void launch_inv_level(const T* ll, long long llp, const T* det, long long dp, T* dst,
                      long long dstp, int W, int H, const Coeffs<T>& cf, hipStream_t st) {
  const int core = 2 * kWave - 2 * cf.nsteps;
  const int nstrips = (W + core - 1) / core;
  const TileShape t = pick_tile(H, nstrips, false);
  const int nseg = (H + t.S - 1) / t.S;
  const int ngroups = (nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
  const InvLaunch<T> L{ll, llp, det, dp, dst, dstp, W, H, t.S, nstrips, nseg,
                       dim3(ngroups * nseg), dim3(kWave * kWavesPerBlock), cf, st};
  with_variant<T>(cf.nsteps, aligned_pair<T>(dst, dstp), t, L);
}

// ---- block (tail) launchers --------------------------------------------------------
// The block kernel takes over from the first level whose region fits the LDS
// budget (two buffers of W x H elements).  Level 0 may be included: the block
// kernel reads its whole region before writing, so it works in place.
static long long g_block_bytes = 16 * 1024;  // one LDS buffer
constexpr int kBlockLdsMax = 150 * 1024;

void tail_config(long long max_region_bytes, int) {
  g_block_bytes = max_region_bytes > kBlockLdsMax / 2 ? kBlockLdsMax / 2 : max_region_bytes;
}

int tail_start(int W, int H, int L, int elem) {
  if (g_block_bytes <= 0) return L;
  for (int j = 0; j < L; ++j) {
    const long long wj = ceil_div_pow2(W, j), hj = ceil_div_pow2(H, j);
    if (wj * hj * elem <= g_block_bytes) return j;
  }
  return L;
}

hipError_t configure_tail_kernels() {
  const void* fns[] = {reinterpret_cast<const void*>(&block_fwd_kernel<float>),
                       reinterpret_cast<const void*>(&block_fwd_kernel<double>),
                       reinterpret_cast<const void*>(&block_inv_kernel<float>),
                       reinterpret_cast<const void*>(&block_inv_kernel<double>)};
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
void launch_block_fwd(const T* src, long long sp, T* dst, long long dp, int W0, int H0, int nlev,
                      const Coeffs<T>& cf, hipStream_t st) {
  const size_t lds = 2 * (size_t)W0 * H0 * sizeof(T);
  hipLaunchKernelGGL((block_fwd_kernel<T>), dim3(1), dim3(kBlockThreads), lds, st, src, sp, dst, dp,
                     W0, H0, nlev, cf);
}

template <typename T>
void launch_block_inv(const T* src, long long sp, T* dst, long long dp, int W0, int H0, int nlev,
                      const Coeffs<T>& cf, hipStream_t st) {
  const size_t lds = 2 * (size_t)W0 * H0 * sizeof(T);
  hipLaunchKernelGGL((block_inv_kernel<T>), dim3(1), dim3(kBlockThreads), lds, st, src, sp, dst, dp,
                     W0, H0, nlev, cf);
}

template void launch_block_fwd<float>(const float*, long long, float*, long long, int, int, int,
                                      const Coeffs<float>&, hipStream_t);
template void launch_block_fwd<double>(const double*, long long, double*, long long, int, int, int,
                                       const Coeffs<double>&, hipStream_t);
template void launch_block_inv<float>(const float*, long long, float*, long long, int, int, int,
                                      const Coeffs<float>&, hipStream_t);
template void launch_block_inv<double>(const double*, long long, double*, long long, int, int, int,
                                       const Coeffs<double>&, hipStream_t);

template <typename T>
void launch_copy_rect(const T* src, long long sp, T* dst, long long dp, int w, int h,
                      hipStream_t st) {
  const long long n = (long long)w * h;
  if (n <= 0) return;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL((copy_rect_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, src, sp,
                     dst, dp, w, h);
}

template void launch_fwd_level<float>(const float*, long long, int, int, float*, long long,
                                      float*, long long, const Coeffs<float>&, hipStream_t);
template void launch_fwd_level<double>(const double*, long long, int, int, double*, long long,
                                       double*, long long, const Coeffs<double>&, hipStream_t);
template void launch_inv_level<float>(const float*, long long, const float*, long long, float*,
                                      long long, int, int, const Coeffs<float>&, hipStream_t);
template void launch_inv_level<double>(const double*, long long, const double*, long long,
                                       double*, long long, int, int, const Coeffs<double>&,
                                       hipStream_t);
template void launch_copy_rect<float>(const float*, long long, float*, long long, int, int,
                                      hipStream_t);
template void launch_copy_rect<double>(const double*, long long, double*, long long, int, int,
                                       hipStream_t);

}  // namespace dwt
