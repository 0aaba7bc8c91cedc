// Pyramid kernels: many lifting levels (of one image or of many images) in one
// persistent launch.
//
// A launch gets a table of tasks, one per (image, level), in an order in
// which every task comes after the task that produces its input.  A task is
// cut into work items: one workgroup (four wavefronts, four adjacent strips)
// over one segment of S rows, items numbered segment-major.  Workgroup b takes
// items b, b + grid, b + 2 grid, ... in increasing order; the grid never
// exceeds the number of workgroups the GPU holds at once.
//
// Dataflow instead of a grid barrier between levels: when a workgroup has
// finished its item it increments the counter of its (task, segment); an item
// of a consumer task first waits until every producer segment whose rows it
// reads (the range is computed on the host from the same row arithmetic as
// the tile body, halo and mirrored rows included) has all its workgroups
// counted.  Levels therefore overlap: the first segments of level j + 1 run
// while level j finishes its last ones.
//
// Progress: an item only waits on items with smaller numbers.  The smallest
// unfinished item belongs to a workgroup that is processing it (its own
// earlier items are smaller, hence finished) and whose dependencies are all
// smaller, hence finished; all workgroups of the grid fit on the GPU together,
// so that workgroup is running (other kernels on the GPU end on their own).
// Every wait is still bounded by the GPU's constant-rate clock: past the limit
// the workgroup records an error flag and goes on (its results are then wrong
// and the host reports the flag), so no launch can hang.  A single atomic
// ticket counter was measured to serialise the start of a level's items (one
// ticket per ~11 ns across XCDs), hence the static assignment.
//
// Coherence across the eight XCDs without cache maintenance.  A release /
// acquire pair at agent scope would write back and invalidate a whole XCD's
// L2 per item, which costs more than the levels themselves.  Instead:
//   * results that later items read (the LL of each level, the inverse's
//     output regions) are stored write-through at agent scope, and every
//     wavefront waits for its stores to complete before the workgroup
//     increments its counter, so the data is in memory before the signal;
//   * no cache can hold a stale copy of such data: every buffer is written
//     once per launch, rows never share a cache line (pitches are multiples
//     of 64 elements), an item reads only rows its dependencies have
//     completed, and caches are invalidated at kernel start; so a consumer's
//     plain loads miss down to memory and see the final values.
// The detail subbands are not read inside a launch and use plain stores.
//
// The task table is copied into LDS by every workgroup once, so finding an
// item's task costs a few LDS reads.  The last workgroup to finish resets the
// segment counters to zero, so the next launch on the same buffers (or a
// graph replay) starts clean without a memset.
#include "tile_body.h"
#include "kernels.h"

#include <algorithm>
#include <cstddef>
#include <climits>
#include <mutex>

namespace dwt {

// Control flow of the persistent loop is kept wave-uniform.  A branch on
// threadIdx.x == 0 around atomics at the end and start of an iteration lets
// the compiler merge them into one lane-divergent region, and then lane 0
// leaves an inner loop while the wavefront's other lanes go on through the
// barrier with stale state, for ever.  So every atomic below is executed by
// all 64 lanes of wave 0 (only lane 0 adds a non-zero amount) and every value
// a branch or loop depends on is made uniform with readfirstlane.
__device__ __forceinline__ int wave_index() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

__device__ __forceinline__ unsigned wave_fetch_add(unsigned* p, unsigned v) {
  const unsigned inc = (threadIdx.x & (kWave - 1)) == 0 ? v : 0u;
  return __builtin_amdgcn_readfirstlane(
      __hip_atomic_fetch_add(p, inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Wait (all lanes of one wavefront) until producer segments r.x .. r.y of the
// task whose counters start at `base` each count `target` workgroups, then
// acquire.
__device__ __forceinline__ void pyr_wait(const PyrSync& sy, int base, int2 r, unsigned target) {
  const unsigned long long t0 = wall_clock64();
  bool timed_out = false;
  for (int q = r.x; q <= r.y && !timed_out; ++q) {
    for (;;) {
      const unsigned c = __builtin_amdgcn_readfirstlane(
          __hip_atomic_load(sy.cnt + base + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (c >= target) break;
      __builtin_amdgcn_s_sleep(1);
      if (wall_clock64() - t0 > sy.spin_limit) {
        timed_out = true;
        break;
      }
    }
  }
  if (timed_out) __hip_atomic_store(sy.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Optional per-item timeline (wave 0 of the workgroup writes it).
__device__ __forceinline__ void pyr_trace(const PyrSync& sy, int item, int slot) {
  if (sy.trace != nullptr && wave_index() == 0 && (threadIdx.x & (kWave - 1)) == 0)
    sy.trace[4 * (long long)item + slot] = slot == 3 ? (unsigned long long)blockIdx.x : wall_clock64();
}

__device__ __forceinline__ void pyr_done(const PyrSync& sy, int counter) {
  __builtin_amdgcn_s_waitcnt(0);  // this wavefront's write-through stores are complete
  __syncthreads();                // ... and every other wavefront's
  if (wave_index() == 0) {
    const unsigned inc = (threadIdx.x & (kWave - 1)) == 0 ? 1u : 0u;
    __hip_atomic_fetch_add(sy.cnt + counter, inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The last workgroup out clears the counters for the next launch.
__device__ __forceinline__ void pyr_finish(const PyrSync& sy, int* s_flag) {
  if (wave_index() == 0) {
    const unsigned inc = (threadIdx.x & (kWave - 1)) == 0 ? 1u : 0u;
    __builtin_amdgcn_s_waitcnt(0);  // this workgroup's counter updates are complete
    const unsigned d = __builtin_amdgcn_readfirstlane(
        __hip_atomic_fetch_add(sy.ctrl + 1, inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    *s_flag = d == gridDim.x - 1;
  }
  __syncthreads();
  if (!__builtin_amdgcn_readfirstlane(*s_flag)) return;
  for (int i = threadIdx.x; i < sy.ncnt; i += blockDim.x)
    __hip_atomic_store(sy.cnt + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x < 2) __hip_atomic_store(sy.ctrl + threadIdx.x, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int kTaskWords = sizeof(PyrTask) / 4;
static_assert(sizeof(PyrTask) % 4 == 0, "task words");

// A task's fields from the LDS copy, as wave-uniform values.
__device__ __forceinline__ PyrTask uniform_task(const Lds<int>* w) {
  PyrTask r;
  int* o = reinterpret_cast<int*>(&r);
#pragma unroll
  for (int i = 0; i < kTaskWords; ++i) o[i] = __builtin_amdgcn_readfirstlane(w[i]);
  return r;
}

// The persistent loop of both directions.  Forward: task = one level; src =
// level input (W x H), out = LL of the next level, det = detail subbands at
// their Mallat positions.  Inverse: src = LL of the coarser level, det =
// details at their Mallat positions of the W x H output region, out = that
// region.
template <bool FWD, typename T, int NS, bool VEC, int SS>
__device__ __forceinline__ void pyr_body(const PyrTask* __restrict__ tasks, int ntasks, int nitems,
                                         const PyrSync& sy, const Coeffs<T>& cf) {
  __shared__ int s_tasks[kPyrMaxTasks * kTaskWords];
  __shared__ int s_flag;
  Lds<int>* const lt = (Lds<int>*)s_tasks;  // generic -> LDS address-space cast
  {
    const Gbl<int>* g = gload_ptr(reinterpret_cast<const int*>(tasks));
    for (int i = threadIdx.x; i < ntasks * kTaskWords; i += blockDim.x) lt[i] = g[i];
  }
  __syncthreads();
  const int wave = wave_index();
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    pyr_trace(sy, item, 0);
    pyr_trace(sy, item, 3);
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (__builtin_amdgcn_readfirstlane(lt[mid * kTaskWords + (int)(offsetof(PyrTask, item0) / 4)]) <= item)
        lo = mid;
      else
        hi = mid - 1;
    }
    const PyrTask tk = uniform_task(lt + lo * kTaskWords);
    const int local = item - tk.item0;
    const int s = local / tk.ngroups, g = local - s * tk.ngroups;
    if (tk.dep >= 0) {
      if (wave == 0) {
        const PyrTask pt = uniform_task(lt + tk.dep * kTaskWords);
        pyr_wait(sy, pt.cnt0, sy.deps[tk.cnt0 + s], (unsigned)pt.ngroups);
      }
      __syncthreads();
    }
    pyr_trace(sy, item, 1);
    const int strip = g * kWavesPerBlock + wave;
    if (strip < tk.nstrips) {
      if constexpr (FWD)
        fwd_tile<T, NS, -SS>(strip, s, SS, tk.W, tk.H,
                             GlobalRows<T, VEC>{static_cast<const T*>(tk.src), tk.sp, tk.H},
                             CoherentStore<T>{static_cast<T*>(tk.out), tk.op},
                             GlobalDet<T>{static_cast<T*>(tk.det), tk.dp}, cf);
      else
        inv_tile<T, NS, -SS>(strip, s, SS, tk.W, tk.H,
                             GlobalLoad<T>{static_cast<const T*>(tk.src), tk.sp},
                             GlobalDet<T>{static_cast<T*>(tk.det), tk.dp},
                             CoherentOut<T, VEC>{static_cast<T*>(tk.out), tk.op}, cf);
    }
    pyr_done(sy, tk.cnt0 + s);
    pyr_trace(sy, item, 2);
  }
  pyr_finish(sy, &s_flag);
}

template <typename T, int NS, bool VEC, int SS>
__global__ void __launch_bounds__(256)
pyr_fwd_kernel(const PyrTask* __restrict__ tasks, int ntasks, int nitems, PyrSync sy, Coeffs<T> cf) {
  pyr_body<true, T, NS, VEC, SS>(tasks, ntasks, nitems, sy, cf);
}

template <typename T, int NS, bool VEC, int SS>
__global__ void __launch_bounds__(256)
pyr_inv_kernel(const PyrTask* __restrict__ tasks, int ntasks, int nitems, PyrSync sy, Coeffs<T> cf) {
  pyr_body<false, T, NS, VEC, SS>(tasks, ntasks, nitems, sy, cf);
}

__global__ void fill_u32_kernel(unsigned* p, unsigned v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---- host side ----------------------------------------------------------------

int pyr_layout(PyrTask* tasks, int ntasks, int nsteps) {
  const int core = 2 * kWave - 2 * nsteps;
  int item = 0, cnt = 0;
  for (int t = 0; t < ntasks; ++t) {
    PyrTask& k = tasks[t];
    k.nstrips = (k.W + core - 1) / core;
    k.ngroups = (k.nstrips + kWavesPerBlock - 1) / kWavesPerBlock;
    k.nseg = (k.H + k.S - 1) / k.S;
    k.item0 = item;
    k.cnt0 = cnt;
    item += k.ngroups * k.nseg;
    cnt += k.nseg;
  }
  return item;
}

// Rows of the level input that segment `seg` of a task loads, as the range
// [lo, hi] of producer segments that write them.  The row loop is the one of
// the single-shot tile bodies: 2 * NP rows from 2 * pbeg, mirrored.
void pyr_deps(const PyrTask& t, const PyrTask& producer, bool fwd, int nsteps, int2* out) {
  const int LAG = nsteps / 2, NP = t.S / 2 + nsteps;
  for (int seg = 0; seg < t.nseg; ++seg) {
    int lo = INT_MAX, hi = -1;
    const int pbeg = ((seg * t.S) >> 1) - LAG;
    for (int u = 0; u < 2 * NP; ++u) {
      const int y = mirror_near(2 * pbeg + u, t.H);
      int row;  // row of the producer's output
      if (fwd) {
        row = y;
      } else {
        if (y & 1) continue;  // detail row, not produced in this launch
        row = y >> 1;
      }
      lo = std::min(lo, row);
      hi = std::max(hi, row);
    }
    // forward producers write LL rows [s * S/2, s * S/2 + S/2) of segment s;
    // inverse producers write output rows [s * S, s * S + S)
    const int rows_per_seg = fwd ? producer.S / 2 : producer.S;
    int a = lo / rows_per_seg, b = hi / rows_per_seg;
    if (hi < 0) a = 0, b = -1;  // reads nothing the producer writes
    b = std::min(b, producer.nseg - 1);
    out[seg] = make_int2(a, b);
  }
}

template <typename T, int NS, bool VEC, int SS>
static void pyr_go(bool fwd, const PyrTask* d_tasks, int ntasks, int nitems, const PyrSync& sy,
                   const Coeffs<T>& cf, int grid_cap, hipStream_t st) {
  const void* fn = fwd ? reinterpret_cast<const void*>(&pyr_fwd_kernel<T, NS, VEC, SS>)
                       : reinterpret_cast<const void*>(&pyr_inv_kernel<T, NS, VEC, SS>);
  // resident workgroups of this instantiation (cached per kernel)
  static std::mutex mu;
  static int resident[2] = {0, 0};
  int& res = resident[fwd ? 1 : 0];
  {
    std::lock_guard<std::mutex> lk(mu);
    if (res == 0) {
      int per_cu = 0, dev = 0, cus = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess ||
          hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        per_cu = 4;
        cus = 256;
      }
      res = std::max(1, per_cu) * std::max(1, cus);
    }
  }
  int grid = std::min(nitems, res);
  if (grid_cap > 0) grid = std::min(grid, grid_cap);
  if (grid <= 0) return;
  if (fwd)
    hipLaunchKernelGGL((pyr_fwd_kernel<T, NS, VEC, SS>), dim3(grid), dim3(256), 0, st, d_tasks,
                       ntasks, nitems, sy, cf);
  else
    hipLaunchKernelGGL((pyr_inv_kernel<T, NS, VEC, SS>), dim3(grid), dim3(256), 0, st, d_tasks,
                       ntasks, nitems, sy, cf);
}

template <typename T, int NS, bool VEC>
static void pyr_s(int S, bool fwd, const PyrTask* d, int nt, int ni, const PyrSync& sy,
                  const Coeffs<T>& cf, int cap, hipStream_t st) {
  if (S == 16) pyr_go<T, NS, VEC, 16>(fwd, d, nt, ni, sy, cf, cap, st);
  else if (S == 32) pyr_go<T, NS, VEC, 32>(fwd, d, nt, ni, sy, cf, cap, st);
  else pyr_go<T, NS, VEC, 8>(fwd, d, nt, ni, sy, cf, cap, st);
}

template <typename T>
void launch_pyr(bool fwd, bool vec, int S, const PyrTask* d_tasks, int ntasks, int nitems,
                const PyrSync& sy, const Coeffs<T>& cf, int grid_cap, hipStream_t st) {
  if (cf.nsteps == 4) {
    if (vec) pyr_s<T, 4, true>(S, fwd, d_tasks, ntasks, nitems, sy, cf, grid_cap, st);
    else pyr_s<T, 4, false>(S, fwd, d_tasks, ntasks, nitems, sy, cf, grid_cap, st);
  } else {
    if (vec) pyr_s<T, 2, true>(S, fwd, d_tasks, ntasks, nitems, sy, cf, grid_cap, st);
    else pyr_s<T, 2, false>(S, fwd, d_tasks, ntasks, nitems, sy, cf, grid_cap, st);
  }
}

void launch_fill_u32(unsigned* p, unsigned v, long long n, hipStream_t st) {
  if (n <= 0) return;
  long long blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
}

template void launch_pyr<float>(bool, bool, int, const PyrTask*, int, int, const PyrSync&,
                                const Coeffs<float>&, int, hipStream_t);
template void launch_pyr<double>(bool, bool, int, const PyrTask*, int, int, const PyrSync&,
                                 const Coeffs<double>&, int, hipStream_t);

}  // namespace dwt
