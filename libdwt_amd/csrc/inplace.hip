// In-place level 0 of the packed 2D path.
//
// Level 0 reads the image and writes three of its four subbands back into the
// image at their Mallat positions (forward), or reads them there and writes
// the reconstructed image over them (inverse).  A tile's writes land on rows
// that other tiles read: the forward tile of row segment s writes HL rows
// [m0, m1) (read by segments ~s/2) and LH / HH rows Hd + [m0, m1) (read by
// segments ~Hd/S + s/2); the inverse tile writes rows [y0, y1), which hold the
// details that segments ~2s or ~2(s - Hc/S) read.  The per-level launch
// therefore sends level-0 details through scratch (D), copied into or out of
// the image by workgroups riding on the small levels' launches.
//
// Here most of that copy goes.  Every workgroup (four strips of one row
// segment, as in the per-level kernel) loads its whole footprint into
// registers, computes, tells the others that its reads are done (a flag per
// tile holding the launch's epoch), waits until every tile that reads what it
// is about to overwrite has said the same, and only then stores.  The host
// orders the row segments so that a segment's readers come first
// (dependencies of the segment graph), greedily; where the graph has a cycle
// the part of a segment that closes it falls back to scratch:
//   * forward: the segment's HL rows, or its LH + HH rows, go to D and are
//     copied into the image after level 0 (riders on the later launches);
//   * inverse: the segment's output band is "preserved": copied from the
//     image to D before level 0 (riders on the earlier launches), and every
//     tile reads that band's details from D, so nobody waits for it.
// On a 4096 x 4096 image 14% (forward) and 21% (inverse) of the detail
// takes the scratch route, against all of it in the per-level launch.
//
// Placement and progress.  Workgroup b runs on XCD b mod 8 (the round-robin
// dispatch of an 8-XCD MI355X), so the tiles are dealt to XCDs by bands of
// eight row segments, XCD x taking the bands with (s / 8) mod 8 == x: row
// halos are then shared inside one XCD's L2 as in the per-level kernel
// (measured: without that locality level 0 takes 32 us instead of 25.7).
// XCD x's workgroups take its tiles in the global segment order, b = 8 k + x
// running the k-th.  A tile waits only on tiles of earlier segments.  With
// in-order dispatch on each XCD, the earliest tile (in the global order) whose
// reads are not done is resident: every workgroup ahead of it on its XCD
// holds an earlier tile, which has finished its reads and waits only on still
// earlier ones, all done, so those workgroups finish and free their slots.
// Every wait is still bounded by the constant-rate clock; past the limit the
// workgroup records an error (the results are then wrong and the host
// reports it) and goes on, so no launch can hang.
//
// Epochs.  The host gives every launch a new epoch (per scratch set, i.e. per
// stream and thread); a flag equal to it was set by this launch.  No counter
// is reset on the device, so a launch has no same-address atomics at all (one
// atomic on a single address costs ~11 ns and 4608 of them would serialise
// the launch).  A launch captured into a HIP graph would replay one epoch, so
// the drivers use the per-level path under stream capture.
//
// Measured on MI355X (4096 x 4096 CDF 9/7 J = 8): 37-39 us per direction for
// this level 0 against 25.7 us (forward) and 24.7 us (inverse) for the
// per-level kernel plus the scratch copy, whose rows ride on latency-bound
// small-level launches for ~4 us of the whole step.  Waiting for dependencies
// costs 7-10 us (a tile's readers are the tiles of the neighbouring dependency
// front, issued moments before it), holding and ordering 2-3 us.  So the
// drivers use it only when switched on (dwt_gpu_level0_set): it cuts the
// level-0 copy from ~200 MB to ~35 MB per forward + inverse of 4096 x 4096,
// at the cost of time.
//
// Coherence.  A flag is stored and loaded at agent scope (memory-side, seen
// by every XCD); data needs no cache maintenance: a waiting tile reads
// nothing the tile it waits on writes, and the stores it makes after the
// wait reach memory at the end of the kernel like any other level's output.
#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <vector>

#include "tile_body.h"
#include "kernels.h"

namespace dwt {

// ---- host: the plan -----------------------------------------------------------

namespace {

struct Iv {  // inclusive column interval, empty when a > b
  int a, b;
};
inline bool overlap(Iv x, Iv y) { return x.a <= x.b && y.a <= y.b && x.a <= y.b && y.a <= x.b; }

struct Geo {
  bool fwd;
  int W, H, NS, S, core, nstrips, ngroups, nseg;
  int Wd, Wh, Hd, Hh;  // ceil / floor halves (the inverse's Wc, Hc are Wd, Hd)
  Geo(bool f, int w, int h, int ns, int s) : fwd(f), W(w), H(h), NS(ns), S(s) {
    core = 2 * kWave - 2 * NS;
    nstrips = (W + core - 1) / core;
    ngroups = (nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
    nseg = (H + S - 1) / S;
    Wd = (W + 1) / 2;
    Wh = W / 2;
    Hd = (H + 1) / 2;
    Hh = H / 2;
  }
  int strips_in(int g) const { return std::min(kWavesPerBlock, nstrips - kWavesPerBlock * g); }
  // columns of the 128-sample windows of group g's strips, mirrored
  template <class F>
  void for_window_cols(int g, F f) const {
    for (int q = kWavesPerBlock * g; q < kWavesPerBlock * g + strips_in(g); ++q)
      for (int c = 0; c < 2 * kWave; ++c) f(q * core - NS + c);
  }
  // forward: input columns group g reads
  Iv fwd_read_cols(int g) const {
    Iv r{W, -1};
    for_window_cols(g, [&](int x) {
      const int m = mirror(x, W);
      r.a = std::min(r.a, m);
      r.b = std::max(r.b, m);
    });
    return r;
  }
  // forward: input rows segment s reads (single shot: S/2 + NS row pairs
  // from pair m0 - NS/2)
  std::vector<int> fwd_read_rows(int s) const {
    const int pbeg = ((s * S) >> 1) - NS / 2;
    std::vector<int> r;
    for (int u = 0; u < 2 * (S / 2 + NS); ++u) r.push_back(mirror(2 * pbeg + u, H));
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    return r;
  }
  // forward: detail output rows m in [m0, m1) of segment s
  void fwd_m(int s, int* m0, int* m1) const {
    *m0 = (s * S) >> 1;
    *m1 = std::min((s * S + S) >> 1, Hd);
  }
  // forward: half-resolution columns [k0, k1) of group g's cores
  void fwd_k(int g, int* k0, int* k1) const {
    *k0 = kWavesPerBlock * g * core / 2;
    *k1 = (kWavesPerBlock * g + strips_in(g)) * core / 2;
  }
  Iv fwd_hl_cols(int g) const {  // HL and HH columns
    int k0, k1;
    fwd_k(g, &k0, &k1);
    return Iv{Wd + k0, Wd + std::min(k1, Wh) - 1};
  }
  Iv fwd_lh_cols(int g) const {
    int k0, k1;
    fwd_k(g, &k0, &k1);
    return Iv{k0, std::min(k1, Wd) - 1};
  }
  // inverse: detail rows (image rows) segment s reads
  std::vector<int> inv_read_rows(int s) const {
    const int pbeg = ((s * S) >> 1) - NS / 2;
    std::vector<int> r;
    for (int u = 0; u < 2 * (S / 2 + NS); ++u) {
      const int Yr = mirror(2 * pbeg + u, H), m = Yr >> 1;
      r.push_back((Yr & 1) ? Hd + m : m);
    }
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    return r;
  }
  // inverse: columns group g reads in a row of the top half (HL) and of the
  // bottom half (LH and HH)
  void inv_read_cols(int g, Iv* top, Iv* lh, Iv* hh) const {
    Iv ke{W, -1}, ko{W, -1};
    for (int q = kWavesPerBlock * g; q < kWavesPerBlock * g + strips_in(g); ++q)
      for (int lane = 0; lane < kWave; ++lane) {
        const int Xe = q * core - NS + 2 * lane;
        const int e = mirror(Xe, W) >> 1, o = mirror(Xe + 1, W) >> 1;
        ke.a = std::min(ke.a, e);
        ke.b = std::max(ke.b, e);
        ko.a = std::min(ko.a, o);
        ko.b = std::max(ko.b, o);
      }
    *top = Iv{Wd + ko.a, Wd + ko.b};
    *lh = ke;
    *hh = Iv{Wd + ko.a, Wd + ko.b};
  }
  // inverse: output columns of group g
  Iv inv_write_cols(int g) const {
    const int x0 = kWavesPerBlock * g * core;
    const int x1 = std::min((kWavesPerBlock * g + strips_in(g)) * core, W);
    return Iv{x0, x1 - 1};
  }
};

// Greedy order of row segments: repeatedly take the segment with the
// cheapest unmet dependencies (cost buckets), smallest index first.
// need[s][p]: segments that must come before part p of s; weight[p][s]: cost
// of leaving part p of s unmet; forced[s][p]: part p always falls back.
// Returns the order; fallback[s] gets a bit per unmet part.
std::vector<int> greedy_order(int n, int nparts, const std::vector<std::vector<std::vector<int>>>& need,
                              const std::vector<std::vector<int>>& weight,
                              const std::vector<std::vector<char>>& forced, std::vector<int>& fallback) {
  std::vector<std::vector<int>> cnt(n, std::vector<int>(nparts, 0));
  std::vector<std::vector<std::pair<int, int>>> rev(n);  // s' -> (s, part) waiting on it
  for (int s = 0; s < n; ++s)
    for (int p = 0; p < nparts; ++p) {
      if (forced[s][p]) continue;
      for (int d : need[s][p]) {
        ++cnt[s][p];
        rev[d].push_back({s, p});
      }
    }
  auto cost = [&](int s) {
    int c = 0;
    for (int p = 0; p < nparts; ++p)
      if (cnt[s][p] > 0) c += weight[p][s];
    return c;
  };
  std::map<int, std::set<int>> bucket;
  std::vector<int> cur(n);
  for (int s = 0; s < n; ++s) {
    cur[s] = cost(s);
    bucket[cur[s]].insert(s);
  }
  std::vector<char> done(n, 0);
  std::vector<int> order;
  fallback.assign(n, 0);
  order.reserve(n);
  for (int it = 0; it < n; ++it) {
    auto b = bucket.begin();
    while (b->second.empty()) b = bucket.erase(b);
    const int s = *b->second.begin();
    b->second.erase(b->second.begin());
    for (int p = 0; p < nparts; ++p)
      if (forced[s][p] || cnt[s][p] > 0) fallback[s] |= 1 << p;
    done[s] = 1;
    order.push_back(s);
    for (auto [t, p] : rev[s]) {
      if (done[t]) continue;
      --cnt[t][p];
      const int c = cost(t);
      if (c != cur[t]) {
        bucket[cur[t]].erase(t);
        cur[t] = c;
        bucket[c].insert(t);
      }
    }
  }
  return order;
}

}  // namespace

int g_l0_band = 8;  // row segments per XCD band

L0Plan build_l0_plan(bool fwd, int W, int H, int nsteps, int S) {
  const Geo G(fwd, W, H, nsteps, S);
  L0Plan P;
  P.fwd = fwd;
  P.W = W;
  P.H = H;
  P.nsteps = nsteps;
  P.S = S;
  P.band = g_l0_band;
  P.nstrips = G.nstrips;
  P.ngroups = G.ngroups;
  P.nseg = G.nseg;
  const int n = G.nseg, ng = G.ngroups;
  // readers of every image row (segment level) and each group's columns
  std::vector<std::vector<int>> readers(H);
  std::vector<Iv> rc(ng), itop(ng), ilh(ng), ihh(ng);
  for (int s = 0; s < n; ++s)
    for (int y : fwd ? G.fwd_read_rows(s) : G.inv_read_rows(s)) readers[y].push_back(s);
  for (int g = 0; g < ng; ++g) {
    if (fwd)
      rc[g] = G.fwd_read_cols(g);
    else
      G.inv_read_cols(g, &itop[g], &ilh[g], &ihh[g]);
  }
  // the rows and columns each part of a tile overwrites in the image
  // forward parts: 0 = HL rows (right half), 1 = LH + HH rows; inverse: one
  // part, the output band
  const int nparts = fwd ? 2 : 1;
  auto part_rows = [&](int s, int p) {
    std::vector<int> r;
    if (fwd) {
      int m0, m1;
      G.fwd_m(s, &m0, &m1);
      for (int m = m0; m < m1; ++m)
        if (p == 0)
          r.push_back(m);
        else if (m < G.Hh)
          r.push_back(G.Hd + m);
    } else {
      for (int y = s * S; y < std::min(s * S + S, H); ++y) r.push_back(y);
    }
    return r;
  };
  // tile-level readers of a part of tile (s, g): every (s2, g2) != (s, g)
  // whose reads overlap what the part writes
  auto tile_readers = [&](int s, int g, int p, std::set<std::pair<int, int>>& out) {
    for (int y : part_rows(s, p)) {
      for (int s2 : readers[y]) {
        for (int g2 = 0; g2 < ng; ++g2) {
          if (s2 == s && g2 == g) continue;
          bool hit;
          if (fwd) {
            hit = p == 0 ? overlap(rc[g2], G.fwd_hl_cols(g))
                         : overlap(rc[g2], G.fwd_lh_cols(g)) || overlap(rc[g2], G.fwd_hl_cols(g));
          } else {
            const Iv w = G.inv_write_cols(g);
            hit = y < G.Hd ? overlap(itop[g2], w) : overlap(ilh[g2], w) || overlap(ihh[g2], w);
          }
          if (hit) out.insert({s2, g2});
        }
      }
    }
  };
  std::vector<std::vector<std::vector<int>>> need(n, std::vector<std::vector<int>>(nparts));
  std::vector<std::vector<char>> forced(n, std::vector<char>(nparts, 0));
  std::vector<std::vector<int>> weight(nparts, std::vector<int>(n, 1));
  for (int s = 0; s < n; ++s) {
    for (int p = 0; p < nparts; ++p) {
      std::set<int> segs;
      for (int y : part_rows(s, p))
        for (int s2 : readers[y]) segs.insert(s2);
      // another tile of the same segment reading what this part writes
      // cannot be ordered before it: that part always falls back
      for (int g = 0; g < ng && !forced[s][p]; ++g) {
        std::set<std::pair<int, int>> tr;
        tile_readers(s, g, p, tr);
        for (auto& t : tr)
          if (t.first == s) forced[s][p] = 1;
      }
      segs.erase(s);
      need[s][p].assign(segs.begin(), segs.end());
      if (fwd)
        weight[p][s] = p == 0 ? 1 : 2;  // HL is half a row, LH + HH a whole one
      else
        weight[p][s] = s * S < G.Hd ? 1 : 2;
    }
  }
  std::vector<int> fb;
  P.order = greedy_order(n, nparts, need, weight, forced, fb);
  P.segflags = fb;
  if (!fwd) {
    // bits 8 + u: row u of the segment's 2 (S/2 + NS) single-shot rows reads
    // its details from a preserved band (from D)
    for (int s = 0; s < n; ++s) {
      const int pbeg = ((s * S) >> 1) - nsteps / 2;
      for (int u = 0; u < 2 * (S / 2 + nsteps); ++u) {
        const int Yr = mirror(2 * pbeg + u, H), m = Yr >> 1;
        const int y = (Yr & 1) ? G.Hd + m : m;
        if (fb[y / S] & 1) P.segflags[s] |= 1 << (8 + u);
      }
    }
  }
  // tiles of each XCD, in the global segment order
  P.list_off.assign(9, 0);
  P.list.clear();
  for (int x = 0; x < 8; ++x) {
    P.list_off[x] = (int)P.list.size();
    for (int s : P.order)
      if ((s / P.band) % 8 == x)
        for (int g = 0; g < ng; ++g) P.list.push_back(s * ng + g);
  }
  P.list_off[8] = (int)P.list.size();
  int maxlen = 0;
  for (int x = 0; x < 8; ++x) maxlen = std::max(maxlen, P.list_off[x + 1] - P.list_off[x]);
  P.grid = 8 * maxlen;
  // dependencies of every tile (only of its in-place parts)
  const int ntiles = n * ng;
  P.dep_off.assign(ntiles + 1, 0);
  P.deps.clear();
  for (int t = 0; t < ntiles; ++t) {
    const int s = t / ng, g = t % ng;
    std::set<std::pair<int, int>> tr;
    for (int p = 0; p < nparts; ++p)
      if (!(fb[s] & (1 << p))) tile_readers(s, g, p, tr);
    P.dep_off[t] = (int)P.deps.size();
    for (auto& d : tr) P.deps.push_back(d.first * ng + d.second);
  }
  P.dep_off[ntiles] = (int)P.deps.size();
  // rows the riders copy: forward, the rows written to D; inverse, the
  // preserved bands
  std::set<int> rows;
  for (int s = 0; s < n; ++s)
    for (int p = 0; p < nparts; ++p)
      if (fb[s] & (1 << p))
        for (int y : part_rows(s, p)) rows.insert(y);
  P.rows.assign(rows.begin(), rows.end());
  return P;
}

// ---- device: the kernels --------------------------------------------------------

__device__ __forceinline__ int wave_index_l0() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Wait (wave 0, all lanes) until every dependency of `tile` has stored this
// launch's epoch in its flag; bounded by the constant-rate clock.
__device__ __forceinline__ void l0_wait(const L0Dev& ip, int tile, unsigned epoch) {
  const int lane = threadIdx.x & (kWave - 1);
  const int d0 = __builtin_amdgcn_readfirstlane(ip.dep_off[tile]);
  const int d1 = __builtin_amdgcn_readfirstlane(ip.dep_off[tile + 1]);
  const unsigned long long t0 = wall_clock64();
  bool timed_out = false;
  for (int b = d0; b < d1 && !timed_out; b += kWave) {
    const int i = b + lane;
    const int dep = i < d1 ? ip.deps[i] : -1;
    for (;;) {
      const unsigned v =
          dep >= 0 ? __hip_atomic_load(ip.flags + dep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : epoch;
      if (__ballot(v != epoch) == 0) break;
      __builtin_amdgcn_s_sleep(1);
      if (wall_clock64() - t0 > ip.spin_limit) {
        timed_out = true;
        break;
      }
    }
  }
  if (timed_out) __hip_atomic_store(ip.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every wavefront's loads have returned (s_waitcnt) before the barrier, so
// once wave 0 stores the flag no read of this tile is outstanding; after the
// wait and a second barrier the tile may overwrite what its readers read.
__device__ __forceinline__ void l0_signal_wait(const L0Dev& ip, int tile, unsigned epoch) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (wave_index_l0() == 0) {
    __hip_atomic_store(ip.flags + tile, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    l0_wait(ip, tile, epoch);
  }
  __syncthreads();
}

// Forward tile body of the single-shot shape with the three detail subbands
// held in registers: the same arithmetic, loads and LL stores as fwd_tile_impl
// (tile_body.h), the detail stores left to fwd_put.
template <typename T, int SS>
struct FwdHeld {
  T lh[SS / 2], hl[SS / 2], hh[SS / 2];
  int k, m0, m1;
  bool st_a, st_b;
};

template <typename T, int NS, int SS, bool INTERIOR, class LoadRow>
__device__ __forceinline__ void fwd_hold_impl(int strip, int seg, int W, int H, const LoadRow& load_row_in,
                                              T* __restrict__ ll, long long llp, const Coeffs<T>& cf,
                                              FwdHeld<T, SS>& h) {
  constexpr int HALO = NS;
  constexpr int CORE = 2 * kWave - 2 * HALO;
  constexpr int LAG = NS / 2;
  constexpr int NP = SS / 2 + NS;
  const int lane = threadIdx.x & (kWave - 1);
  const int x0 = strip * CORE;
  const int Xe = x0 - HALO + 2 * lane;
  const int xe = INTERIOR ? Xe : mirror(Xe, W), xo = INTERIOR ? Xe + 1 : mirror(Xe + 1, W);
  const int Wd = (W + 1) >> 1, Wh = W >> 1, Hd = (H + 1) >> 1;
  const int k = Xe >> 1;
  const bool core = lane >= HALO / 2 && lane < kWave - HALO / 2;
  h.k = k;
  h.st_a = core && k < Wd;
  h.st_b = core && k < Wh;
  h.m0 = (seg * SS) >> 1;
  h.m1 = min((seg * SS + SS) >> 1, Hd);
  const int pbeg = h.m0 - LAG;
  T re[2 * NP], ro[2 * NP];
#pragma unroll
  for (int u = 0; u < 2 * NP; ++u) {
    const int Y = 2 * pbeg + u;
    load_row_in(INTERIOR ? Y : mirror_near(Y, H), Xe, xe, xo, INTERIOR, re[u], ro[u]);
  }
  ColFwd<T, NS> ca, cb;
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    T e0 = re[2 * u], o0 = ro[2 * u], e1 = re[2 * u + 1], o1 = ro[2 * u + 1];
    row_fwd<T, NS>(e0, o0, cf);
    row_fwd<T, NS>(e1, o1, cf);
    T La, Ha, Lb, Hb;
    ca.push(e0, e1, cf, La, Ha);
    cb.push(o0, o1, cf, Lb, Hb);
    if (u >= NS) {  // pair pbeg + u completes output row m = pbeg + u - LAG = m0 + (u - NS)
      const int q = u - NS;
      const int m = h.m0 + q;
      if (m < h.m1 && h.st_a) st_global(ll + (long long)m * llp + k, La);
      h.lh[q] = Ha;
      h.hl[q] = Lb;
      h.hh[q] = Hb;
    }
  }
}

template <typename T, int NS, int SS, class LoadRow>
__device__ __forceinline__ void fwd_hold(int strip, int seg, int W, int H, const LoadRow& load_row, T* ll,
                                         long long llp, const Coeffs<T>& cf, FwdHeld<T, SS>& h) {
  constexpr int HALO = NS;
  const int x0 = strip * (2 * kWave - 2 * HALO);
  const int interior = __builtin_amdgcn_readfirstlane((x0 - HALO >= 0) && (x0 - HALO + 2 * kWave <= W) &&
                                                      rows_inside<NS, -SS>(seg * SS, SS, H, true));
  if (interior)
    fwd_hold_impl<T, NS, SS, true>(strip, seg, W, H, load_row, ll, llp, cf, h);
  else
    fwd_hold_impl<T, NS, SS, false>(strip, seg, W, H, load_row, ll, llp, cf, h);
}

// The held details to their Mallat positions: HL rows to `top` (pitch tp),
// LH and HH rows to `bot` (pitch bp) -- the image, or D for a part that
// falls back.
template <typename T, int SS>
__device__ __forceinline__ void fwd_put(const FwdHeld<T, SS>& h, int W, int H, T* top, long long tp, T* bot,
                                        long long bp) {
  const int Wd = (W + 1) >> 1, Hd = (H + 1) >> 1, Hh = H >> 1;
#pragma unroll
  for (int q = 0; q < SS / 2; ++q) {
    const int m = h.m0 + q;
    if (m < h.m1) {
      if (h.st_a && m < Hh) st_global(bot + (long long)(Hd + m) * bp + h.k, h.lh[q]);
      if (h.st_b) {
        st_global(top + (long long)m * tp + Wd + h.k, h.hl[q]);
        if (m < Hh) st_global(bot + (long long)(Hd + m) * bp + Wd + h.k, h.hh[q]);
      }
    }
  }
}

// Inverse tile body of the single-shot shape with the output rows held in
// registers: the arithmetic of inv_tile_impl (tile_body.h); details of a row
// in a preserved band come from D.
template <typename T, int SS>
struct InvHeld {
  T e[SS], o[SS];
  int Xe, y0, y1;
  bool core, st_e, st_o;
};

template <typename T>
struct L0Det {  // detail rows of the image, or of D for a preserved band
  const T* A;
  long long ap;
  const T* D;
  long long dp;
  int mask;  // bit 8 + u: read row u of this tile comes from D
  __device__ __forceinline__ T load(int u, int r, int c) const {
    return ((mask >> (8 + u)) & 1) ? gload_ptr(D)[(long long)r * dp + c] : gload_ptr(A)[(long long)r * ap + c];
  }
};

template <typename T, int NS, int SS, bool ROWS_IN>
__device__ __forceinline__ void inv_hold_impl(int strip, int seg, int W, int H, const T* __restrict__ ll,
                                              long long llp, const L0Det<T>& det, const Coeffs<T>& cf,
                                              InvHeld<T, SS>& h) {
  constexpr int HALO = NS;
  constexpr int CORE = 2 * kWave - 2 * HALO;
  constexpr int LAG = NS / 2;
  constexpr int NP = SS / 2 + NS;
  const int lane = threadIdx.x & (kWave - 1);
  const int x0 = strip * CORE;
  const int Xe = x0 - HALO + 2 * lane;
  const int ke = mirror(Xe, W) >> 1, ko = mirror(Xe + 1, W) >> 1;
  const int Wc = (W + 1) >> 1, Hc = (H + 1) >> 1;
  h.Xe = Xe;
  h.core = lane >= HALO / 2 && lane < kWave - HALO / 2;
  h.st_e = h.core && Xe < W;
  h.st_o = h.core && Xe + 1 < W;
  h.y0 = seg * SS;
  h.y1 = min(h.y0 + SS, H);
  const int pbeg = (h.y0 >> 1) - LAG;
  T re[2 * NP], ro[2 * NP];
#pragma unroll
  for (int u = 0; u < 2 * NP; ++u) {
    const int Y = 2 * pbeg + u;
    const int Yr = ROWS_IN ? Y : mirror_near(Y, H);
    const int m = Yr >> 1;
    if (Yr & 1) {
      re[u] = det.load(u, Hc + m, ke);
      ro[u] = det.load(u, Hc + m, Wc + ko);
    } else {
      re[u] = gload_ptr(ll)[(long long)m * llp + ke];
      ro[u] = det.load(u, m, Wc + ko);
    }
  }
  ColInv<T, NS> cle, clo;
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    T e0 = re[2 * u], o0 = ro[2 * u], e1 = re[2 * u + 1], o1 = ro[2 * u + 1];
    row_inv<T, NS>(e0, o0, cf);
    row_inv<T, NS>(e1, o1, cf);
    T fe, se_, fo, so_;
    cle.push(e0, e1, cf, fe, se_);
    clo.push(o0, o1, cf, fo, so_);
    // pair pbeg + u completes rows r0, r0 + 1 with r0 - y0 = 2u - NS - 3
    // (NS = 4) or 2u - NS - 2 (NS = 2)
    const int off = NS == 4 ? 2 * u - 7 : 2 * u - 4;
    if (off >= 0 && off < SS) {
      h.e[off] = fe;
      h.o[off] = fo;
    }
    if (off + 1 >= 0 && off + 1 < SS) {
      h.e[off + 1] = se_;
      h.o[off + 1] = so_;
    }
  }
}

template <typename T, int NS, int SS>
__device__ __forceinline__ void inv_hold(int strip, int seg, int W, int H, const T* ll, long long llp,
                                         const L0Det<T>& det, const Coeffs<T>& cf, InvHeld<T, SS>& h) {
  if (__builtin_amdgcn_readfirstlane(rows_inside<NS, -SS>(seg * SS, SS, H, false)))
    inv_hold_impl<T, NS, SS, true>(strip, seg, W, H, ll, llp, det, cf, h);
  else
    inv_hold_impl<T, NS, SS, false>(strip, seg, W, H, ll, llp, det, cf, h);
}

template <typename T, bool VEC, int SS>
__device__ __forceinline__ void inv_put(const InvHeld<T, SS>& h, T* A, long long ap) {
  if (!h.core) return;
  const GlobalOut<T, VEC> out{A, ap};
#pragma unroll
  for (int i = 0; i < SS; ++i)
    if (h.y0 + i < h.y1) out(h.y0 + i, h.Xe, h.e[i], h.o[i], h.st_e, h.st_o);
}

// Workgroup b: XCD list b mod 8, k-th tile of it (none when past its end).
__device__ __forceinline__ int l0_tile(const L0Dev& ip) {
  const int x = blockIdx.x & 7, k = blockIdx.x >> 3;
  const int n = ip.off[x + 1] - ip.off[x];
  return k < n ? __builtin_amdgcn_readfirstlane(ip.list[ip.off[x] + k]) : -1;
}

template <typename T, int NS, bool VEC, int SS>
__global__ void __launch_bounds__(256)
fwd_l0_kernel(T* A, long long ap, int W, int H, T* __restrict__ ll, long long llp, T* D, long long dp,
              Coeffs<T> cf, L0Dev ip) {
  const int tile = l0_tile(ip);
  if (tile < 0) return;
  const int s = tile / ip.ngroups, g = tile - s * ip.ngroups;
  const int strip = g * kWavesPerBlock + wave_index_l0();
  const bool active = strip < ip.nstrips;
  FwdHeld<T, SS> h;
  if (active)
    fwd_hold<T, NS, SS>(strip, s, W, H, GlobalRows<T, VEC>{A, ap, H}, ll, llp, cf, h);
  l0_signal_wait(ip, tile, ip.epoch);
  if (active) {
    const int f = __builtin_amdgcn_readfirstlane(ip.segflags[s]);
    fwd_put<T, SS>(h, W, H, (f & 1) ? D : A, (f & 1) ? dp : ap, (f & 2) ? D : A, (f & 2) ? dp : ap);
  }
}

template <typename T, int NS, bool VEC, int SS>
__global__ void __launch_bounds__(256)
inv_l0_kernel(const T* __restrict__ ll, long long llp, T* A, long long ap, const T* D, long long dp, int W,
              int H, Coeffs<T> cf, L0Dev ip) {
  const int tile = l0_tile(ip);
  if (tile < 0) return;
  const int s = tile / ip.ngroups, g = tile - s * ip.ngroups;
  const int strip = g * kWavesPerBlock + wave_index_l0();
  const bool active = strip < ip.nstrips;
  InvHeld<T, SS> h;
  const int mask = __builtin_amdgcn_readfirstlane(ip.segflags[s]);
  if (active) inv_hold<T, NS, SS>(strip, s, W, H, ll, llp, L0Det<T>{A, ap, D, dp, mask}, cf, h);
  l0_signal_wait(ip, tile, ip.epoch);
  if (active) inv_put<T, VEC, SS>(h, A, ap);
}

// ---- host: launchers --------------------------------------------------------------

template <typename T>
static bool pair_ok(const void* p, long long) {
  // pairs move as one element-aligned access (GPair): any row pitch will do
  return (reinterpret_cast<uintptr_t>(p) % sizeof(T)) == 0;
}

template <typename T, int NS, bool VEC>
static void fwd_l0_go(int S, dim3 grid, T* A, long long ap, int W, int H, T* ll, long long llp, T* D,
                      long long dp, const Coeffs<T>& cf, const L0Dev& ip, hipStream_t st) {
  if (S == 4)
    hipLaunchKernelGGL((fwd_l0_kernel<T, NS, VEC, 4>), grid, dim3(256), 0, st, A, ap, W, H, ll, llp, D, dp, cf, ip);
  else
    hipLaunchKernelGGL((fwd_l0_kernel<T, NS, VEC, 8>), grid, dim3(256), 0, st, A, ap, W, H, ll, llp, D, dp, cf, ip);
}

template <typename T>
void launch_fwd_l0(T* A, long long ap, int W, int H, T* ll, long long llp, T* D, long long dp, const Coeffs<T>& cf,
                   const L0Dev& ip, int S, int grid, hipStream_t st) {
  const bool vec = pair_ok<T>(A, ap);
  if (cf.nsteps == 4) {
    if (vec) fwd_l0_go<T, 4, true>(S, dim3(grid), A, ap, W, H, ll, llp, D, dp, cf, ip, st);
    else fwd_l0_go<T, 4, false>(S, dim3(grid), A, ap, W, H, ll, llp, D, dp, cf, ip, st);
  } else {
    if (vec) fwd_l0_go<T, 2, true>(S, dim3(grid), A, ap, W, H, ll, llp, D, dp, cf, ip, st);
    else fwd_l0_go<T, 2, false>(S, dim3(grid), A, ap, W, H, ll, llp, D, dp, cf, ip, st);
  }
}

template <typename T, int NS, bool VEC>
static void inv_l0_go(int S, dim3 grid, const T* ll, long long llp, T* A, long long ap, const T* D, long long dp,
                      int W, int H, const Coeffs<T>& cf, const L0Dev& ip, hipStream_t st) {
  if (S == 4)
    hipLaunchKernelGGL((inv_l0_kernel<T, NS, VEC, 4>), grid, dim3(256), 0, st, ll, llp, A, ap, D, dp, W, H, cf, ip);
  else
    hipLaunchKernelGGL((inv_l0_kernel<T, NS, VEC, 8>), grid, dim3(256), 0, st, ll, llp, A, ap, D, dp, W, H, cf, ip);
}

template <typename T>
void launch_inv_l0(const T* ll, long long llp, T* A, long long ap, const T* D, long long dp, int W, int H,
                   const Coeffs<T>& cf, const L0Dev& ip, int S, int grid, hipStream_t st) {
  const bool vec = pair_ok<T>(A, ap);
  if (cf.nsteps == 4) {
    if (vec) inv_l0_go<T, 4, true>(S, dim3(grid), ll, llp, A, ap, D, dp, W, H, cf, ip, st);
    else inv_l0_go<T, 4, false>(S, dim3(grid), ll, llp, A, ap, D, dp, W, H, cf, ip, st);
  } else {
    if (vec) inv_l0_go<T, 2, true>(S, dim3(grid), ll, llp, A, ap, D, dp, W, H, cf, ip, st);
    else inv_l0_go<T, 2, false>(S, dim3(grid), ll, llp, A, ap, D, dp, W, H, cf, ip, st);
  }
}

template <typename T>
__global__ void rider_only_kernel(RiderCopy<T> rc) {
  rider_copy(rc, blockIdx.x, gridDim.x);
}
template <typename T>
void launch_rider_alone(const RiderCopy<T>& rc, hipStream_t st) {
  if (rc.blocks <= 0 || rc.r1 <= rc.r0) return;
  hipLaunchKernelGGL((rider_only_kernel<T>), dim3(rc.blocks), dim3(256), 0, st, rc);
}

template void launch_fwd_l0<float>(float*, long long, int, int, float*, long long, float*, long long,
                                   const Coeffs<float>&, const L0Dev&, int, int, hipStream_t);
template void launch_fwd_l0<double>(double*, long long, int, int, double*, long long, double*, long long,
                                    const Coeffs<double>&, const L0Dev&, int, int, hipStream_t);
template void launch_inv_l0<float>(const float*, long long, float*, long long, const float*, long long, int, int,
                                   const Coeffs<float>&, const L0Dev&, int, int, hipStream_t);
template void launch_inv_l0<double>(const double*, long long, double*, long long, const double*, long long, int,
                                    int, const Coeffs<double>&, const L0Dev&, int, int, hipStream_t);
template void launch_rider_alone<float>(const RiderCopy<float>&, hipStream_t);
template void launch_rider_alone<double>(const RiderCopy<double>&, hipStream_t);

}  // namespace dwt
