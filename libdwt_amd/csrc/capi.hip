// C-ABI layer: libdwt's drop-in entry points and the dwt_gpu_* extensions.
//
// Each entry point classifies its pointer(s), runs the transform with the
// device kernels on the calling thread's stream, and synchronises before it
// returns (the reference's calls are synchronous).  Host (pageable) memory is
// staged: the frame or line is gathered into packed device memory, transformed
// there, and scattered back.  There is no CPU compute path: if the device work
// fails the call aborts (drop-in) or returns an error (extensions).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "dwt_common.h"
#include "kernels.h"
#include "host_common.h"

namespace dwt {

// ---- constants (src/libdwt.c:541-573) -----------------------------------------
namespace k {
static const float P1_s = 1.58613434342059;
static const float U1_s = -0.0529801185729;
static const float P2_s = -0.8829110755309;
static const float U2_s = 0.4435068520439;
static const float K_s = 1.1496043988602;
static const float S2_s = 1 / 1.1496043988602;
static const double P1_d = 1.58613434342059;
static const double U1_d = -0.0529801185729;
static const double P2_d = -0.8829110755309;
static const double U2_d = 0.4435068520439;
static const double K_d = 1.1496043988602;
static const double S2_d = 1 / 1.1496043988602;
static const float P53_s = 0.5, U53_s = 0.25;
static const float S153_s = 1.41421356237309504880, S253_s = 0.70710678118654752440;
static const double P53_d = 0.5, U53_d = 0.25;
static const double S153_d = 1.41421356237309504880, S253_d = 0.70710678118654752440;
}  // namespace k

template <>
Coeffs<float> make_coeffs<float>(Wavelet w, bool fwd) {
  Coeffs<float> c{};
  // volatile keeps the float quotient 1/K a run-time float division, the
  // value the reference's lifting path multiplies by (src/libdwt.c:1604)
  volatile float kk = k::K_s;
  const float invK = 1.0f / kk;
  if (w == CDF97) {
    c.nsteps = 4;
    if (fwd) {  // src/libdwt.c:5542
      c.c[0] = -k::P1_s; c.c[1] = k::U1_s; c.c[2] = -k::P2_s; c.c[3] = k::U2_s;
      c.se = k::K_s; c.so = invK; c.one = k::K_s; c.first_parity = 1;
    } else {    // src/libdwt.c:5811, 5793
      c.c[0] = -k::U2_s; c.c[1] = k::P2_s; c.c[2] = -k::U1_s; c.c[3] = k::P1_s;
      c.se = invK; c.so = k::K_s; c.one = k::S2_s; c.first_parity = 0;
    }
  } else {
    c.nsteps = 2;
    if (fwd) {  // src/libdwt.c:5576-5598
      c.c[0] = -k::P53_s; c.c[1] = k::U53_s;
      c.se = k::S153_s; c.so = k::S253_s; c.one = k::S153_s; c.first_parity = 1;
    } else {    // src/libdwt.c:5843-5866
      c.c[0] = -k::U53_s; c.c[1] = k::P53_s;
      c.se = k::S253_s; c.so = k::S153_s; c.one = k::S253_s; c.first_parity = 0;
    }
  }
  return c;
}

template <>
Coeffs<double> make_coeffs<double>(Wavelet w, bool fwd) {
  Coeffs<double> c{};
  if (w == CDF97) {
    c.nsteps = 4;
    if (fwd) {  // src/libdwt.c:1467-1525
      c.c[0] = -k::P1_d; c.c[1] = k::U1_d; c.c[2] = -k::P2_d; c.c[3] = k::U2_d;
      c.se = k::K_d; c.so = k::S2_d; c.one = k::K_d; c.first_parity = 1;
    } else {    // src/libdwt.c:5673-5731
      c.c[0] = -k::U2_d; c.c[1] = k::P2_d; c.c[2] = -k::U1_d; c.c[3] = k::P1_d;
      c.se = k::S2_d; c.so = k::K_d; c.one = k::S2_d; c.first_parity = 0;
    }
  } else {
    c.nsteps = 2;
    if (fwd) {  // src/libdwt.c:1527-1571
      c.c[0] = -k::P53_d; c.c[1] = k::U53_d;
      c.se = k::S153_d; c.so = k::S253_d; c.one = k::S153_d; c.first_parity = 1;
    } else {    // src/libdwt.c:5733-5777
      c.c[0] = -k::U53_d; c.c[1] = k::P53_d;
      c.se = k::S253_d; c.so = k::S153_d; c.one = k::S253_d; c.first_parity = 0;
    }
  }
  return c;
}

// ---- errors -------------------------------------------------------------------
thread_local std::string g_err;

void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  throw Failure{code};
}

void check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) fail(DWT_GPU_EHIP, "kernel launch: %s", hipGetErrorString(e));
}

// ---- device scratch, one set per (host thread, stream, device) ----------------
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
};

// A pyramid launch's task table and dependency ranges, uploaded once and
// reused by every later call with the same tasks (same sizes and pointers).
// The pinned host copy the upload reads is kept unchanged for the plan's life,
// so the upload is stream-ordered and safe to replay.
struct PyrPlan {
  std::vector<PyrTask> tasks;
  bool fwd = true, vec = true;
  int S = 8, nsteps = 4, nitems = 0, ncnt = 0;
  size_t deps_off = 0;
  void* dev = nullptr;
  void* pinned = nullptr;
};

// A cached in-place level-0 plan of one scratch set: the host plan, its
// tables in device memory (uploaded from a pinned copy kept for the plan's
// life, so the upload is stream-ordered) and their offsets.
struct L0Cached {
  L0Plan plan;
  void* dev = nullptr;
  void* pinned = nullptr;
  size_t bytes = 0;
  size_t o_list = 0, o_dep_off = 0, o_deps = 0, o_segflags = 0, o_rows = 0;
};

struct Scratch {
  // 0: LL of even levels, 1: LL of odd levels, 2: level-0 details,
  // 3: line temp, 4: host staging, 5: LL of the pyramid levels
  // 6: compare partials, 7: FP64 reference image of a verification,
  // 8: packed copy of a restaged frame
  DevBuf buf[9];
  void* cmp_host = nullptr;  // pinned copy of a compare result
  // staging of pageable host frames: two pinned chunks, each with an event
  // marking the end of the last copy engine transfer that used it
  void* pin[2] = {nullptr, nullptr};
  size_t pin_bytes = 0;
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  void ensure_pin(size_t bytes) {
    if (pin_bytes >= bytes) return;
    for (int i = 0; i < 2; ++i) {
      if (pin[i]) {
        HIPCHK(hipEventSynchronize(pin_ev[i]));
        hipHostFree(pin[i]);
        pin[i] = nullptr;
      }
      if (hipHostMalloc(&pin[i], bytes, hipHostMallocDefault) != hipSuccess) {
        hipGetLastError();
        pin[i] = nullptr;
        pin_bytes = 0;
        fail(DWT_GPU_EMEMORY, "pinned staging of %zu bytes failed", bytes);
      }
      if (!pin_ev[i]) HIPCHK(hipEventCreateWithFlags(&pin_ev[i], hipEventDisableTiming));
    }
    pin_bytes = bytes;
  }
  // pyramid state: plans (most recent first), and the sync words
  // ctrl[0..1], err[2], then per-segment counters
  std::vector<PyrPlan*> plans;
  unsigned* sync = nullptr;
  long long sync_cnt = 0;
  unsigned* err_host = nullptr;
  bool pyr_used = false;
  unsigned long long* trace = nullptr;  // per-item timeline of the last pyramid launch
  long long trace_n = 0, trace_items = 0;
  // in-place level 0: plans (most recent first); l0_words = error word,
  // padding to 256 bytes, then one flag per tile; the last launch's epoch
  std::vector<L0Cached*> l0_plans;
  unsigned* l0_words = nullptr;
  long long l0_cap = 0;
  unsigned l0_epoch = 0;
  hipStream_t stream = nullptr;
  // side stream for the level-0 detail copies, which touch memory disjoint
  // from what the other levels read and write, so they overlap those levels
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  void ensure_side() {
    if (side) return;
    HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
  }
  void* get(int i, size_t bytes) {
    DevBuf& b = buf[i];
    if (b.n < bytes) {
      if (b.p) {
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipFree(b.p));
        b.p = nullptr;
        b.n = 0;
      }
      size_t want = bytes + (bytes >> 3);
      want = (want + 4095) & ~size_t(4095);
      if (hipMalloc(&b.p, want) != hipSuccess) {
        hipGetLastError();
        b.p = nullptr;
        fail(DWT_GPU_EMEMORY, "hipMalloc of %zu bytes of scratch failed", want);
      }
      b.n = want;
    }
    return b.p;
  }
  long long bytes() const {
    long long s = 0;
    for (const auto& b : buf) s += (long long)b.n;
    for (const PyrPlan* pl : plans)
      s += (long long)(pl->deps_off + sizeof(int2) * (size_t)pl->ncnt);
    for (const L0Cached* c : l0_plans) s += (long long)c->bytes;
    return s + 4 * (4 + sync_cnt) + 4 * (l0_words ? 64 + l0_cap : 0);
  }
  void release() {
    for (auto& b : buf) {
      if (b.p) hipFree(b.p);
      b = DevBuf{};
    }
    for (PyrPlan* pl : plans) free_plan(pl);
    plans.clear();
    for (L0Cached* c : l0_plans) free_l0(c);
    l0_plans.clear();
    if (l0_words) hipFree(l0_words);
    l0_words = nullptr;
    l0_cap = 0;
    l0_epoch = 0;
    if (sync) hipFree(sync);
    sync = nullptr;
    sync_cnt = 0;
    if (err_host) hipHostFree(err_host);
    err_host = nullptr;
    if (cmp_host) hipHostFree(cmp_host);
    cmp_host = nullptr;
    for (int i = 0; i < 2; ++i) {
      if (pin[i]) hipHostFree(pin[i]);
      if (pin_ev[i]) hipEventDestroy(pin_ev[i]);
      pin[i] = nullptr;
      pin_ev[i] = nullptr;
    }
    pin_bytes = 0;
    if (trace) hipFree(trace);
    trace = nullptr;
    trace_n = trace_items = 0;
    if (side) hipStreamDestroy(side);
    if (fork) hipEventDestroy(fork);
    if (join) hipEventDestroy(join);
    side = nullptr;
    fork = join = nullptr;
  }
  static void free_plan(PyrPlan* pl) {
    if (pl->dev) hipFree(pl->dev);
    if (pl->pinned) hipHostFree(pl->pinned);
    delete pl;
  }
  static void free_l0(L0Cached* c) {
    if (c->dev) hipFree(c->dev);
    if (c->pinned) hipHostFree(c->pinned);
    delete c;
  }
};

static std::mutex g_scratch_mu;
using ScratchKey = std::tuple<std::thread::id, hipStream_t, int>;
static std::map<ScratchKey, Scratch*>* g_scratch = new std::map<ScratchKey, Scratch*>();
static const std::thread::id g_main_thread = std::this_thread::get_id();

// Frees every scratch set of the calling thread: at the exit of a thread
// other than the one that loaded the library (the process's main thread
// keeps its sets until the process ends, when the HIP runtime may already be
// gone), so short-lived threads do not leave device memory and streams behind.
static void release_thread_scratch() {
  std::vector<Scratch*> mine;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    const std::thread::id me = std::this_thread::get_id();
    for (auto it = g_scratch->begin(); it != g_scratch->end();) {
      if (std::get<0>(it->first) == me) {
        mine.push_back(it->second);
        it = g_scratch->erase(it);
      } else {
        ++it;
      }
    }
  }
  for (Scratch* s : mine) {
    // the stream may already be destroyed: wait for the whole device instead
    (void)hipDeviceSynchronize();
    s->release();
    delete s;
  }
  (void)hipGetLastError();
}
struct ThreadScratchGuard {
  bool armed = false;
  ~ThreadScratchGuard() {
    if (armed && std::this_thread::get_id() != g_main_thread) release_thread_scratch();
  }
};
static thread_local ThreadScratchGuard t_scratch_guard;

static Scratch& scratch_for(hipStream_t st) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  t_scratch_guard.armed = true;  // this thread's sets are freed at its exit
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto key = std::make_tuple(std::this_thread::get_id(), st, dev);
  auto it = g_scratch->find(key);
  if (it != g_scratch->end()) return *it->second;
  Scratch* s = new Scratch();
  s->stream = st;
  (*g_scratch)[key] = s;
  return *s;
}

// ---- optional per-launch timing (HIP events on the launch stream) ---------------
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static int g_prof_repeat = 1;
static std::vector<ProfRec>* g_prof = new std::vector<ProfRec>();

int prof_repeat() {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  return g_prof_on ? g_prof_repeat : 1;
}

ProfScope::ProfScope(int kind, int level, hipStream_t s, int reps) : st(s) {
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    on = g_prof_on;
  }
  if (!on) return;
  r.kind = kind;
  r.level = level;
  r.reps = reps;
  HIPCHK(hipEventCreate(&r.a));
  HIPCHK(hipEventCreate(&r.b));
  HIPCHK(hipEventRecord(r.a, st));
}
ProfScope::~ProfScope() {
  if (!on) return;
  hipEventRecord(r.b, st);
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof->push_back(r);
}

static std::once_flag g_init_once;
static hipError_t g_init_status = hipSuccess;
void init_once() {
  std::call_once(g_init_once, [] {
    g_init_status = configure_line_kernels();
    if (g_init_status == hipSuccess) g_init_status = configure_tail_kernels();
    if (g_init_status == hipSuccess) g_init_status = configure_mid_kernels();
    if (g_init_status == hipSuccess) g_init_status = configure_batch_tail();
  });
  if (g_init_status != hipSuccess)
    fail(DWT_GPU_EHIP, "raising the kernels' LDS limits: %s", hipGetErrorString(g_init_status));
}


// Two-element vector access is possible: pointer aligned to a pair, even pitch.
template <typename T>
static bool aligned_pair_host(const void* p, long long pitch) {
  return (reinterpret_cast<uintptr_t>(p) % (2 * sizeof(T))) == 0 && (pitch % 2) == 0;
}

// ---- pyramid launches ---------------------------------------------------------------
// Schedule of the levels after level 0: 0 = one launch per level plus the
// block kernel for the small ones (default), 1 = one pyramid launch.
static int g_schedule = 0;
// Sparse frames with contiguous rows: 1 = the large levels as fused sparse
// level launches (default), 0 = the generic per-line passes.
static int g_sparse_fused = 1;
static int g_trace_items = 0;  // > 0: pyramid launches record a per-item timeline
static int g_pyr_grid_cap = 0;  // > 0: at most this many workgroups per pyramid launch
static constexpr size_t kMaxPlans = 8;

static unsigned long long spin_limit_ticks() {
  static std::once_flag once;
  static unsigned long long ticks = 200000000ULL;  // 2 s at 100 MHz
  std::call_once(once, [] {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
      ticks = 2000ULL * (unsigned long long)khz;  // 2 s
    else
      (void)hipGetLastError();
  });
  return ticks;
}

// Run `tasks` (in producer-before-consumer order) as one pyramid launch on st.
template <typename T>
static void run_pyr(bool fwd, std::vector<PyrTask>& tasks, int S, bool vec, const Coeffs<T>& cf,
                    Scratch& sc, hipStream_t st) {
  for (auto& t : tasks) t.S = S;
  const int ntasks = (int)tasks.size();
  if (ntasks > kPyrMaxTasks) fail(DWT_GPU_EINVAL, "a pyramid holds at most %d levels", kPyrMaxTasks);
  const int nitems = pyr_layout(tasks.data(), ntasks, cf.nsteps);
  if (nitems == 0) return;
  const int ncnt = tasks.back().cnt0 + tasks.back().nseg;
  PyrPlan* plan = nullptr;
  for (size_t i = 0; i < sc.plans.size(); ++i) {
    PyrPlan* pl = sc.plans[i];
    if (pl->fwd == fwd && pl->vec == vec && pl->S == S && pl->nsteps == cf.nsteps &&
        pl->tasks.size() == tasks.size() &&
        std::memcmp(pl->tasks.data(), tasks.data(), sizeof(PyrTask) * tasks.size()) == 0) {
      plan = pl;
      sc.plans.erase(sc.plans.begin() + (long)i);
      sc.plans.insert(sc.plans.begin(), plan);
      break;
    }
  }
  if (!plan) {
    std::vector<int2> deps((size_t)ncnt, make_int2(0, -1));
    for (const PyrTask& t : tasks)
      if (t.dep >= 0) pyr_deps(t, tasks[t.dep], fwd, cf.nsteps, deps.data() + t.cnt0);
    plan = new PyrPlan();
    plan->tasks = tasks;
    plan->fwd = fwd;
    plan->vec = vec;
    plan->S = S;
    plan->nsteps = cf.nsteps;
    plan->nitems = nitems;
    plan->ncnt = ncnt;
    plan->deps_off = (sizeof(PyrTask) * tasks.size() + 255) & ~size_t(255);
    const size_t bytes = plan->deps_off + sizeof(int2) * deps.size();
    if (hipHostMalloc(&plan->pinned, bytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&plan->dev, bytes) != hipSuccess) {
      (void)hipGetLastError();
      Scratch::free_plan(plan);
      fail(DWT_GPU_EMEMORY, "allocating a pyramid task table of %zu bytes failed", bytes);
    }
    std::memcpy(plan->pinned, tasks.data(), sizeof(PyrTask) * tasks.size());
    std::memcpy((char*)plan->pinned + plan->deps_off, deps.data(), sizeof(int2) * deps.size());
    HIPCHK(hipMemcpyAsync(plan->dev, plan->pinned, bytes, hipMemcpyHostToDevice, st));
    sc.plans.insert(sc.plans.begin(), plan);
    if (sc.plans.size() > kMaxPlans) {
      HIPCHK(hipStreamSynchronize(st));  // the evicted plan may still be in use
      Scratch::free_plan(sc.plans.back());
      sc.plans.pop_back();
    }
  }
  if (sc.sync_cnt < ncnt) {
    if (sc.sync) {
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipFree(sc.sync));
      sc.sync = nullptr;
      sc.sync_cnt = 0;
    }
    const long long cap = std::max<long long>(ncnt, 4096);
    if (hipMalloc(&sc.sync, sizeof(unsigned) * (size_t)(4 + cap)) != hipSuccess) {
      (void)hipGetLastError();
      sc.sync = nullptr;
      fail(DWT_GPU_EMEMORY, "allocating pyramid counters failed");
    }
    sc.sync_cnt = cap;
    LAUNCH(launch_fill_u32(sc.sync, 0u, 4 + cap, st));  // a kernel, not a memset: capturable
  }
  PyrSync sy{};
  sy.ctrl = sc.sync;
  sy.err = sc.sync + 2;
  sy.cnt = sc.sync + 4;
  sy.deps = reinterpret_cast<const int2*>(static_cast<char*>(plan->dev) + plan->deps_off);
  sy.ncnt = ncnt;
  sy.spin_limit = spin_limit_ticks();
  sy.trace = nullptr;
  if (g_trace_items > 0) {
    if (!sc.trace || sc.trace_n < (long long)nitems) {
      if (sc.trace) {
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipFree(sc.trace));
        sc.trace = nullptr;
      }
      const long long n = std::max<long long>(nitems, g_trace_items);
      if (hipMalloc(&sc.trace, sizeof(unsigned long long) * 4 * (size_t)n) != hipSuccess) {
        (void)hipGetLastError();
        sc.trace = nullptr;
        fail(DWT_GPU_EMEMORY, "allocating the pyramid trace failed");
      }
      sc.trace_n = n;
    }
    sy.trace = sc.trace;
    sc.trace_items = nitems;
  }
  sc.pyr_used = true;
  LAUNCH(launch_pyr<T>(fwd, vec, S, static_cast<const PyrTask*>(plan->dev), ntasks, nitems, sy, cf,
                       g_pyr_grid_cap, st));
}

// Reads and clears the in-place level-0 error word of sc (synchronises st);
// true when a dependency wait of an earlier launch timed out.
static bool l0_failed(Scratch& sc, hipStream_t st) {
  if (!sc.l0_words) return false;
  if (!sc.err_host) HIPCHK(hipHostMalloc((void**)&sc.err_host, sizeof(unsigned), hipHostMallocDefault));
  HIPCHK(hipMemcpyAsync(sc.err_host, sc.l0_words, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (*sc.err_host == 0) return false;
  LAUNCH(launch_fill_u32(sc.l0_words, 0u, 1, st));
  HIPCHK(hipStreamSynchronize(st));
  return true;
}

// Reads and clears the pyramid error flag of sc (synchronises st); true when a
// dependency wait of an earlier launch timed out.
static bool pyr_failed(Scratch& sc, hipStream_t st) {
  if (!sc.pyr_used || !sc.sync) return false;
  if (!sc.err_host) HIPCHK(hipHostMalloc((void**)&sc.err_host, sizeof(unsigned), hipHostMallocDefault));
  HIPCHK(hipMemcpyAsync(sc.err_host, sc.sync + 2, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (*sc.err_host == 0) return false;
  LAUNCH(launch_fill_u32(sc.sync + 2, 0u, 1, st));
  HIPCHK(hipStreamSynchronize(st));
  return true;
}

// Segment rows of a pyramid launch whose largest level is W x H: the
// single-shot shape of the tile table, else 8.
static int pyr_rows(bool fwd, int W, int H, int nsteps) {
  int S = 8, U = 0;
  tile_get(fwd ? 1 : 0, H, W, nsteps, &S, &U);
  return (U == 0 && (S == 8 || S == 16 || S == 32)) ? S : 8;
}

// ---- generic path: per-level line passes in place on any byte strides --------
// Columns of a layout with contiguous rows (unit column stride, an element
// pitch) can run as adjacent signals through the coalesced column kernels;
// lines of one sample keep the line kernel (its N == 1 case).
template <typename T>
static bool columns_adjacent(const char* ptr, long long sx, long long sy, int n) {
  return sy == (long long)sizeof(T) && sx > 0 && sx % (long long)sizeof(T) == 0 &&
         reinterpret_cast<uintptr_t>(ptr) % sizeof(T) == 0 && n >= 2;
}

template <typename T>
static void generic_fwd_level(char* ptr, long long sx, long long sy, int Wo, int Ho, int Wi, int Hi,
                              int j, int zp, const Coeffs<T>& cf, Scratch& sc, hipStream_t st) {
  const int Wos = ceil_div_pow2(Wo, j), Hos = ceil_div_pow2(Ho, j);
  const int Wod = ceil_div_pow2(Wo, j + 1), Hod = ceil_div_pow2(Ho, j + 1);
  const int Wis = ceil_div_pow2(Wi, j), His = ceil_div_pow2(Hi, j);
  const int lim = line_lds_limit_bytes() / (int)sizeof(T);
  const int gblocks = 1024;
  // rows (src/libdwt.c:6354-6362), then columns (6367-6375)
  T* gt = Wis > lim ? (T*)sc.get(3, sizeof(T) * (size_t)Wis * gblocks) : nullptr;
  LAUNCH(launch_line_pass<T>(true, ptr, ptr, ptr + (long long)Wod * sy, sx, sy, Hos, Wis, cf, gt, gblocks, st));
  if (columns_adjacent<T>(ptr, sx, sy, His)) {
    // the Wos columns as adjacent signals, one coalesced launch; L and H rows
    // go to scratch rows [0, ceil(His/2)) and [Hod, Hod + His/2) (the
    // columns are read while other rows are written), then back into place
    T* A = reinterpret_cast<T*>(ptr);
    const long long p = sx / (long long)sizeof(T), Dp = (long long)pad_pitch(Wos);
    T* D = (T*)sc.get(2, sizeof(T) * (size_t)Dp * Hos);
    const int half = (His + 1) / 2;
    LAUNCH(launch_col_level<T>(true, A, p, D, Dp, D + (long long)(Hod - half) * Dp, Dp, His, Wos, cf,
                               aligned_pair_host<T>(A, p), st));
    LAUNCH(launch_copy_rect<T>(D, Dp, A, p, Wos, half, st);
           launch_copy_rect<T>(D + (long long)Hod * Dp, Dp, A + (long long)Hod * p, p, Wos, His / 2, st));
  } else {
    gt = His > lim ? (T*)sc.get(3, sizeof(T) * (size_t)His * gblocks) : nullptr;
    LAUNCH(launch_line_pass<T>(true, ptr, ptr, ptr + (long long)Hod * sx, sy, sx, Wos, His, cf, gt, gblocks, st));
  }
  if (zp) {  // src/libdwt.c:6377-6397
    const int zx[4] = {(Wis + 1) / 2, Wod, Wod + Wis / 2, Wos};
    const int zy[4] = {(His + 1) / 2, Hod, Hod + His / 2, Hos};
    LAUNCH(launch_zero_pad<T>(ptr, sx, sy, Wos, Hos, zx, zy, st));
  }
}

template <typename T>
static void generic_inv_level(char* ptr, long long sx, long long sy, int Wo, int Ho, int Wi, int Hi,
                              int j, int zp, const Coeffs<T>& cf, Scratch& sc, hipStream_t st) {
  const int Wos = ceil_div_pow2(Wo, j), Hos = ceil_div_pow2(Ho, j);
  const int Wod = ceil_div_pow2(Wo, j - 1), Hod = ceil_div_pow2(Ho, j - 1);
  const int Wid = ceil_div_pow2(Wi, j - 1), Hid = ceil_div_pow2(Hi, j - 1);
  const int lim = line_lds_limit_bytes() / (int)sizeof(T);
  const int gblocks = 1024;
  // rows (src/libdwt.c:6706-6714), then columns (6719-6727)
  T* gt = Wid > lim ? (T*)sc.get(3, sizeof(T) * (size_t)Wid * gblocks) : nullptr;
  LAUNCH(launch_line_pass<T>(false, ptr, ptr + (long long)Wos * sy, ptr, sx, sy, Hod, Wid, cf, gt, gblocks, st));
  if (columns_adjacent<T>(ptr, sx, sy, Hid)) {
    // the Wod columns as adjacent signals: L rows from [0, ceil(Hid/2)), H
    // rows from [Hos, ..); the Hid output rows to scratch, then into place
    T* A = reinterpret_cast<T*>(ptr);
    const long long p = sx / (long long)sizeof(T), Dp = (long long)pad_pitch(Wod);
    T* D = (T*)sc.get(2, sizeof(T) * (size_t)Dp * Hid);
    const int half = (Hid + 1) / 2;
    LAUNCH(launch_col_level<T>(false, A, p, A + (long long)(Hos - half) * p, p, D, Dp, Hid, Wod, cf,
                               aligned_pair_host<T>(A, p), st));
    LAUNCH(launch_copy_rect<T>(D, Dp, A, p, Wod, Hid, st));
  } else {
    gt = Hid > lim ? (T*)sc.get(3, sizeof(T) * (size_t)Hid * gblocks) : nullptr;
    LAUNCH(launch_line_pass<T>(false, ptr, ptr + (long long)Hos * sx, ptr, sy, sx, Wod, Hid, cf, gt, gblocks, st));
  }
  if (zp) {  // src/libdwt.c:6729-6745
    const int zx[4] = {Wid, Wod, 0, 0};
    const int zy[4] = {Hid, Hod, 0, 0};
    LAUNCH(launch_zero_pad<T>(ptr, sx, sy, Wod, Hod, zx, zy, st));
  }
}

// ---- packed fast path -----------------------------------------------------------
// A: W x H, row pitch p elements, levels [0, L) forward.
//  * Small levels: from level jt on (the first level whose W_jt x H_jt region
//    fits one LDS buffer) a single block-kernel launch does all remaining
//    levels inside one workgroup's LDS.  If jt == 0 the whole transform is
//    that one launch (the block kernel reads its region before writing).
//  * Level 0 (when jt > 0) reads A and writes its detail subbands to scratch D
//    (they overlap A's unread input) and LL_1 to scratch P[1]; its details are
//    copied into A on a side stream, overlapping the later levels, which touch
//    disjoint memory.
//  * Level 1 <= j < jt reads LL_j from ping-pong scratch P[j % 2] and writes
//    its details straight into A (that part of A was consumed by level 0) and
//    LL_{j+1} to P[(j + 1) % 2] (or to A when it is the last level).
// Pyramid schedule.  Level 0 is its own launch as above (A -> LL_1 in P[1],
// details in D, copied into A on the side stream); levels 1 .. L-1 are one
// pyramid launch: level j reads LL_j (P[1] for j = 1, else Q_j), writes its
// details into A and LL_{j+1} to Q_{j+1}, the last level's LL into A.  The
// Q_j (j >= 2) are distinct buffers, so no memory is reused inside the launch.
template <typename T>
struct PyrBuffers {
  T* q[32] = {};
  long long qp[32] = {};
};
template <typename T>
static PyrBuffers<T> pyr_buffers(int W, int H, int L, Scratch& sc) {
  PyrBuffers<T> b;
  size_t off[32] = {}, total = 0;
  for (int j = 2; j < L; ++j) {
    b.qp[j] = (long long)pad_pitch(ceil_div_pow2(W, j));
    off[j] = total;
    total += sizeof(T) * (size_t)b.qp[j] * ceil_div_pow2(H, j);
    total = (total + 255) & ~size_t(255);
  }
  if (total) {
    char* base = (char*)sc.get(5, total);
    for (int j = 2; j < L; ++j) b.q[j] = (T*)(base + off[j]);
  }
  return b;
}

template <typename T>
static void fast_fwd_pyr(T* A, long long p, int W, int H, int L, const Coeffs<T>& cf, Scratch& sc,
                         hipStream_t st) {
  const int W1 = ceil_div_pow2(W, 1), H1 = ceil_div_pow2(H, 1);
  const long long Dp = (long long)pad_pitch(W);
  T* D = (T*)sc.get(2, sizeof(T) * (size_t)Dp * H);
  const long long P1p = (long long)pad_pitch(W1);
  T* P1 = L >= 2 ? (T*)sc.get(1, sizeof(T) * (size_t)P1p * H1) : nullptr;
  const PyrBuffers<T> q = pyr_buffers<T>(W, H, L, sc);
  sc.ensure_side();
  {
    ProfScope ps(0, 0, st);
    LAUNCH(launch_fwd_level<T>(A, p, W, H, L >= 2 ? P1 : D, L >= 2 ? P1p : Dp, D, Dp, cf, st));
  }
  HIPCHK(hipEventRecord(sc.fork, st));
  HIPCHK(hipStreamWaitEvent(sc.side, sc.fork, 0));
  {
    ProfScope ps(2, 0, sc.side);
    LAUNCH(launch_copy_rect<T>(D + W1, Dp, A + W1, p, W - W1, H1, sc.side);
           launch_copy_rect<T>(D + (long long)H1 * Dp, Dp, A + (long long)H1 * p, p, W, H - H1, sc.side);
           if (L == 1) launch_copy_rect<T>(D, Dp, A, p, W1, H1, sc.side));
  }
  if (L >= 2) {
    std::vector<PyrTask> tasks;
    for (int j = 1; j < L; ++j) {
      PyrTask t{};
      t.src = j == 1 ? (const void*)P1 : (const void*)q.q[j];
      t.sp = j == 1 ? P1p : q.qp[j];
      t.out = j == L - 1 ? (void*)A : (void*)q.q[j + 1];
      t.op = j == L - 1 ? p : q.qp[j + 1];
      t.det = A;
      t.dp = p;
      t.W = ceil_div_pow2(W, j);
      t.H = ceil_div_pow2(H, j);
      t.dep = j == 1 ? -1 : (int)tasks.size() - 1;
      tasks.push_back(t);
    }
    ProfScope ps(6, 1, st);
    run_pyr<T>(true, tasks, pyr_rows(true, W1, H1, cf.nsteps), aligned_pair_host<T>(A, p), cf, sc, st);
  }
  HIPCHK(hipEventRecord(sc.join, sc.side));
  HIPCHK(hipStreamWaitEvent(st, sc.join, 0));
}

// Inverse of fast_fwd_pyr: A's level-0 details go to D on the side stream;
// one pyramid launch does levels L-1 .. 1 (LL_L from A, outputs in Q_j, LL_1
// in P[1]); then the level-0 inverse writes the image into A.
template <typename T>
static void fast_inv_pyr(T* A, long long p, int W, int H, int L, const Coeffs<T>& cf, Scratch& sc,
                         hipStream_t st) {
  const int W1 = ceil_div_pow2(W, 1), H1 = ceil_div_pow2(H, 1);
  const long long Dp = (long long)pad_pitch(W);
  T* D = (T*)sc.get(2, sizeof(T) * (size_t)Dp * H);
  const long long P1p = (long long)pad_pitch(W1);
  T* P1 = L >= 2 ? (T*)sc.get(1, sizeof(T) * (size_t)P1p * H1) : nullptr;
  const PyrBuffers<T> q = pyr_buffers<T>(W, H, L, sc);
  sc.ensure_side();
  HIPCHK(hipEventRecord(sc.fork, st));
  HIPCHK(hipStreamWaitEvent(sc.side, sc.fork, 0));
  {
    ProfScope ps(2, 0, sc.side);
    LAUNCH(launch_copy_rect<T>(A + W1, p, D + W1, Dp, W - W1, H1, sc.side);
           launch_copy_rect<T>(A + (long long)H1 * p, p, D + (long long)H1 * Dp, Dp, W, H - H1, sc.side);
           if (L == 1) launch_copy_rect<T>(A, p, D, Dp, W1, H1, sc.side));
  }
  HIPCHK(hipEventRecord(sc.join, sc.side));
  if (L >= 2) {
    std::vector<PyrTask> tasks;
    for (int j = L - 1; j >= 1; --j) {
      PyrTask t{};
      t.src = j == L - 1 ? (const void*)A : (const void*)q.q[j + 1];
      t.sp = j == L - 1 ? p : q.qp[j + 1];
      t.out = j == 1 ? (void*)P1 : (void*)q.q[j];
      t.op = j == 1 ? P1p : q.qp[j];
      t.det = A;
      t.dp = p;
      t.W = ceil_div_pow2(W, j);
      t.H = ceil_div_pow2(H, j);
      t.dep = j == L - 1 ? -1 : (int)tasks.size() - 1;
      tasks.push_back(t);
    }
    ProfScope ps(7, 1, st);
    run_pyr<T>(false, tasks, pyr_rows(false, W1, H1, cf.nsteps), aligned_pair_host<T>(A, p), cf, sc, st);
  }
  HIPCHK(hipStreamWaitEvent(st, sc.join, 0));
  {
    ProfScope ps(1, 0, st);
    LAUNCH(launch_inv_level<T>(L >= 2 ? P1 : D, L >= 2 ? P1p : Dp, D, Dp, A, p, W, H, cf, st));
  }
}

// Fused middle levels: kMidLevels levels in one launch from the first level
// j >= 1 whose region is at most g_mid_bytes, when they all come before the
// block kernel's first level jt.  Returns that level, or -1.  Off by default:
// on MI355X the three fused levels of a 4096 x 4096 image take 25 us (forward)
// and 21.5 us (inverse) in one launch, against 15.2 us and 17.3 us as three
// launches (kernel durations), because each workgroup runs its three levels' tiles as one serial
// chain on one CU where the per-level launches spread them over all CUs.
constexpr int kMidLevels = 3;
static long long g_mid_bytes = 0;
static int mid_start(int W, int H, int jt, int elem) {
  if (g_mid_bytes <= 0) return -1;
  for (int j = 1; j + kMidLevels <= jt; ++j)
    if ((long long)ceil_div_pow2(W, j) * ceil_div_pow2(H, j) * elem <= g_mid_bytes) return j;
  return -1;
}

// ---- in-place level 0 -----------------------------------------------------------
// The plan of a level-0 launch (cached per scratch set, most recent first) and
// its device view with a fresh epoch.
static int g_l0_inplace = 0;  // off: measured slower than the scratch route on MI355X
constexpr size_t kL0MaxPlans = 8;

static L0Cached& l0_cached(Scratch& sc, bool fwd, int W, int H, int nsteps, int S, hipStream_t st) {
  for (size_t i = 0; i < sc.l0_plans.size(); ++i) {
    L0Cached* c = sc.l0_plans[i];
    const L0Plan& q = c->plan;
    if (q.fwd == fwd && q.W == W && q.H == H && q.nsteps == nsteps && q.S == S && q.band == g_l0_band) {
      sc.l0_plans.erase(sc.l0_plans.begin() + (long)i);
      sc.l0_plans.insert(sc.l0_plans.begin(), c);
      return *c;
    }
  }
  L0Cached* c = new L0Cached();
  c->plan = build_l0_plan(fwd, W, H, nsteps, S);
  const L0Plan& q = c->plan;
  size_t off = 0;
  auto place = [&](size_t n) {
    const size_t at = off;
    off += (4 * n + 255) & ~size_t(255);
    return at;
  };
  c->o_list = place(q.list.size());
  c->o_dep_off = place(q.dep_off.size());
  c->o_deps = place(q.deps.size() + 1);
  c->o_segflags = place(q.segflags.size());
  c->o_rows = place(q.rows.size() + 1);
  c->bytes = off;
  if (hipHostMalloc(&c->pinned, off, hipHostMallocDefault) != hipSuccess || hipMalloc(&c->dev, off) != hipSuccess) {
    (void)hipGetLastError();
    Scratch::free_l0(c);
    fail(DWT_GPU_EMEMORY, "allocating a level-0 plan of %zu bytes failed", off);
  }
  char* h = static_cast<char*>(c->pinned);
  auto put = [&](size_t at, const std::vector<int>& v) {
    if (!v.empty()) std::memcpy(h + at, v.data(), 4 * v.size());
  };
  put(c->o_list, q.list);
  put(c->o_dep_off, q.dep_off);
  put(c->o_deps, q.deps);
  put(c->o_segflags, q.segflags);
  put(c->o_rows, q.rows);
  HIPCHK(hipMemcpyAsync(c->dev, c->pinned, off, hipMemcpyHostToDevice, st));
  sc.l0_plans.insert(sc.l0_plans.begin(), c);
  if (sc.l0_plans.size() > kL0MaxPlans) {
    HIPCHK(hipStreamSynchronize(st));  // the evicted plan may still be in use
    Scratch::free_l0(sc.l0_plans.back());
    sc.l0_plans.pop_back();
  }
  return *c;
}

static L0Dev l0_dev(Scratch& sc, const L0Cached& c, hipStream_t st) {
  const long long ntiles = (long long)c.plan.nseg * c.plan.ngroups;
  if (sc.l0_cap < ntiles) {
    if (sc.l0_words) {
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipFree(sc.l0_words));
      sc.l0_words = nullptr;
      sc.l0_cap = 0;
    }
    const long long cap = std::max<long long>(ntiles, 1 << 14);
    if (hipMalloc(&sc.l0_words, sizeof(unsigned) * (size_t)(64 + cap)) != hipSuccess) {
      (void)hipGetLastError();
      sc.l0_words = nullptr;
      fail(DWT_GPU_EMEMORY, "allocating level-0 flags failed");
    }
    sc.l0_cap = cap;
    sc.l0_epoch = 0;
    LAUNCH(launch_fill_u32(sc.l0_words, 0u, 64 + cap, st));  // a kernel, not a memset: capturable
  }
  if (++sc.l0_epoch == 0) {  // the epoch wrapped: no flag may hold the new one
    LAUNCH(launch_fill_u32(sc.l0_words + 64, 0u, sc.l0_cap, st));
    sc.l0_epoch = 1;
  }
  L0Dev d{};
  const char* base = static_cast<const char*>(c.dev);
  d.list = reinterpret_cast<const int*>(base + c.o_list);
  d.dep_off = reinterpret_cast<const int*>(base + c.o_dep_off);
  d.deps = reinterpret_cast<const int*>(base + c.o_deps);
  d.segflags = reinterpret_cast<const int*>(base + c.o_segflags);
  d.flags = sc.l0_words + 64;
  d.err = sc.l0_words;
  d.spin_limit = spin_limit_ticks();
  d.epoch = sc.l0_epoch;
  for (int x = 0; x < 9; ++x) d.off[x] = c.plan.list_off[x];
  d.ngroups = c.plan.ngroups;
  d.nstrips = c.plan.nstrips;
  return d;
}

// Segment rows of the in-place level 0 of a W x H image, or 0 when the
// per-level path runs it: the in-place kernels take the single-shot shapes of
// 4 and 8 rows; L >= 2 (LL_1 goes to scratch, not into the image); a launch
// under stream capture would replay one epoch; the allocation-free entry
// point (sc == null) has no room for plans.
static int l0_inplace_rows(const Scratch* sc, bool fwd, int W, int H, int L, int nsteps, hipStream_t st) {
  if (!g_l0_inplace || !sc || L < 2 || W < 2 || H < 2) return 0;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  if (cs != hipStreamCaptureStatusNone) return 0;
  int S = 0, U = 0;
  tile_get(fwd ? 1 : 0, H, W, nsteps, &S, &U);
  return U == 0 && (S == 4 || S == 8) ? S : 0;
}

// ---- the level-0 detail copy, carried by the small levels' launches -----------
// Level 0 of the in-place forward writes its detail subbands to scratch D
// (they overlap input that other workgroups have not read yet), and the last
// level of the inverse reads them from D (its output overwrites them).  The
// copy between D and the image rides on the launches of the later (forward) or
// earlier (inverse) levels as extra workgroups: those levels touch only the
// image's top-left LL_1 quadrant and the P buffers, disjoint from the copy,
// and leave most of the GPU idle (the block kernel runs on one compute unit).
// Weights: a level launch whose region is at most 8 MiB takes one share, the
// block launch three; larger levels carry none.  With no carrier, the copy is a
// launch of its own on the stream.
template <typename T>
static RiderCopy<T> detail_copy(const T* src, long long sp, T* dst, long long dp, int W, int H, int L) {
  RiderCopy<T> rc{};
  rc.src = src;
  rc.sp = sp;
  rc.dst = dst;
  rc.dp = dp;
  const int W1 = ceil_div_pow2(W, 1), H1 = ceil_div_pow2(H, 1);
  rc.xa = L == 1 ? 0 : W1;  // with one level, LL_1 moves too
  rc.wa = W - rc.xa;
  rc.ha = H1;
  rc.wb = W;
  rc.hb = H - H1;
  constexpr int per = 16 / (int)sizeof(T);
  rc.vec = (reinterpret_cast<uintptr_t>(src) % 16) == 0 && (reinterpret_cast<uintptr_t>(dst) % 16) == 0 &&
           sp % per == 0 && dp % per == 0 && rc.xa % per == 0;
  return rc;
}

#ifndef DWT_CARRIER_MIB
#define DWT_CARRIER_MIB 8
#endif
#ifndef DWT_BLOCK_CARRIER_WEIGHT
#define DWT_BLOCK_CARRIER_WEIGHT 3
#endif
constexpr long long kCarrierBytes = (long long)DWT_CARRIER_MIB << 20;
constexpr int kBlockCarrierWeight = DWT_BLOCK_CARRIER_WEIGHT;

// Assign rows [0, rows) of the copy to carriers by weight, in order; returns
// false when no launch can carry any (all weights 0).
static bool split_rows(int rows, const std::vector<int>& w, std::vector<int2>& out) {
  long long tot = 0;
  for (int x : w) tot += x;
  out.assign(w.size(), make_int2(0, 0));
  if (tot == 0) return false;
  long long acc = 0;
  int cur = 0;
  for (size_t i = 0; i < w.size(); ++i) {
    acc += w[i];
    const int end = (int)(rows * acc / tot);
    out[i] = make_int2(cur, end);
    cur = end;
  }
  return true;
}

// The rows [r.x, r.y) of rc as carried by a launch whose workgroups have
// `threads` threads (at most `cap` extra workgroups).
template <typename T>
static RiderCopy<T> rider_part(const RiderCopy<T>& rc, int2 r, int cap) {
  RiderCopy<T> x = rc;
  x.r0 = r.x;
  x.r1 = r.y;
  x.blocks = std::min(r.y - r.x, cap);
  return x;
}
constexpr int kLevelRiderCap = 2048;  // 256-thread workgroups
constexpr int kBlockRiderCap = 480;   // 1024-thread workgroups, two per compute unit

template <typename T>
static void copy_alone(const RiderCopy<T>& rc, hipStream_t st) {
  if (rc.rows) {
    RiderCopy<T> x = rc;
    x.blocks = std::min(rc.r1 - rc.r0, kLevelRiderCap);
    LAUNCH(launch_rider_alone<T>(x, st));
    return;
  }
  LAUNCH(launch_copy_rect<T>(rc.src + rc.xa, rc.sp, rc.dst + rc.xa, rc.dp, rc.wa, rc.ha, st);
         launch_copy_rect<T>(rc.src + (long long)rc.ha * rc.sp, rc.sp, rc.dst + (long long)rc.ha * rc.dp,
                             rc.dp, rc.wb, rc.hb, st));
}

// Scratch of the per-level schedule: D (level-0 details, W x H) and the
// ping-pong LL buffers P[0] (LL of even levels >= 2, W_2 x H_2) and P[1] (LL of
// odd levels, W_1 x H_1), pitches padded to 64 elements.  With base == null
// only the byte count is computed (each buffer 256-byte aligned).
template <typename T>
struct FastBufs {
  T* D = nullptr;
  T* P[2] = {nullptr, nullptr};
  long long Dp = 0, Pp[2] = {0, 0};
};
template <typename T>
static size_t fast_bufs_layout(int W, int H, int L, char* base, FastBufs<T>* b) {
  const int W1 = ceil_div_pow2(W, 1), H1 = ceil_div_pow2(H, 1);
  const int W2 = ceil_div_pow2(W, 2), H2 = ceil_div_pow2(H, 2);
  b->Dp = (long long)pad_pitch(W);
  b->Pp[0] = (long long)pad_pitch(W2);
  b->Pp[1] = (long long)pad_pitch(W1);
  const size_t sz[3] = {sizeof(T) * (size_t)b->Dp * H, L >= 2 ? sizeof(T) * (size_t)b->Pp[1] * H1 : 0,
                        L >= 3 ? sizeof(T) * (size_t)b->Pp[0] * H2 : 0};
  size_t off = 0, at[3];
  for (int i = 0; i < 3; ++i) {
    at[i] = off;
    off += (sz[i] + 255) & ~size_t(255);
  }
  if (base) {
    b->D = reinterpret_cast<T*>(base + at[0]);
    b->P[1] = sz[1] ? reinterpret_cast<T*>(base + at[1]) : nullptr;
    b->P[0] = sz[2] ? reinterpret_cast<T*>(base + at[2]) : nullptr;
  }
  return off;
}
template <typename T>
static FastBufs<T> fast_bufs(Scratch& sc, int W, int H, int L) {
  FastBufs<T> b;
  fast_bufs_layout<T>(W, H, L, nullptr, &b);
  b.D = (T*)sc.get(2, sizeof(T) * (size_t)b.Dp * H);
  if (L >= 2) b.P[1] = (T*)sc.get(1, sizeof(T) * (size_t)b.Pp[1] * ceil_div_pow2(H, 1));
  if (L >= 3) b.P[0] = (T*)sc.get(0, sizeof(T) * (size_t)b.Pp[0] * ceil_div_pow2(H, 2));
  return b;
}

// Forward levels [0, L) of a packed image on the given scratch buffers (unused
// when the block kernel takes the whole image).  Launches kernels only.  With
// a scratch set (sc), level 0 may run in place (l0_inplace_rows); the
// allocation-free entry point passes none.
template <typename T>
static void fast_fwd_on(T* A, long long p, int W, int H, int L, const Coeffs<T>& cf, const FastBufs<T>& fb,
                        hipStream_t st, Scratch* sc = nullptr) {
  const int jt = tail_start(W, H, L, (int)sizeof(T));
  if (jt == 0) {
    ProfScope ps(4, 0, st);
    LAUNCH(launch_block_fwd<T>(A, p, A, p, W, H, L, cf, st));
    return;
  }
  T* const D = fb.D;
  T* const P[2] = {fb.P[0], fb.P[1]};
  const long long Dp = fb.Dp, Pp[2] = {fb.Pp[0], fb.Pp[1]};
  const int jm = mid_start(W, H, jt, (int)sizeof(T));
  // carriers of the detail copy: the level launches 1 .. jt-1 (not the fused
  // middle levels) and the block launch
  std::vector<int> weight;
  for (int j = 1; j < jt; ++j) {
    const bool mid = jm > 0 && j >= jm && j < jm + kMidLevels;
    const long long bytes = (long long)ceil_div_pow2(W, j) * ceil_div_pow2(H, j) * (long long)sizeof(T);
    weight.push_back(!mid && bytes <= kCarrierBytes ? 1 : 0);
  }
  if (jt < L) weight.push_back(kBlockCarrierWeight);
  RiderCopy<T> rc = detail_copy<T>(D, Dp, A, p, W, H, L);
  // in place: only the rows that fell back to D are copied
  const int S0 = l0_inplace_rows(sc, true, W, H, L, cf.nsteps, st);
  const L0Cached* l0 = S0 ? &l0_cached(*sc, true, W, H, cf.nsteps, S0, st) : nullptr;
  int ncopy = H;
  if (l0) {
    rc.rows = reinterpret_cast<const int*>(static_cast<const char*>(l0->dev) + l0->o_rows);
    ncopy = (int)l0->plan.rows.size();
    rc.r0 = 0;
    rc.r1 = ncopy;
  }
  std::vector<int2> rows;
  const bool carried = split_rows(ncopy, weight, rows);
  for (int j = 0; j < jt; ++j) {
    const int Wj = ceil_div_pow2(W, j), Hj = ceil_div_pow2(H, j);
    const T* src = j == 0 ? A : P[j % 2];
    const long long sp = j == 0 ? p : Pp[j % 2];
    if (j == jm) {
      // levels jm .. jm+2 in one launch: LL_jm from P, details into A, LL_{jm+3}
      // to P (A when it is the last level)
      const int je = j + kMidLevels;
      T* ll = je == L ? A : P[je % 2];
      const long long llp = je == L ? p : Pp[je % 2];
      ProfScope ps(10, j, st);
      LAUNCH(launch_mid_fwd<T>(src, sp, A, p, ll, llp, Wj, Hj, kMidLevels, cf, st));
      j = je - 1;
      continue;
    }
    T* ll;
    long long llp;
    if (j == L - 1) {
      ll = j == 0 ? D : A;
      llp = j == 0 ? Dp : p;
    } else {
      ll = P[(j + 1) % 2];
      llp = Pp[(j + 1) % 2];
    }
    T* det = j == 0 ? D : A;
    const long long dp = j == 0 ? Dp : p;
    // level 0 reads A and writes only scratch, so a profiling pass may run
    // it several times back to back inside one event pair (per-launch
    // duration without the events' dispatch overhead)
    const int reps = j == 0 ? prof_repeat() : 1;
    ProfScope ps(0, j, st, reps);
    if (j == 0 && l0) {
      // in place (a profiling repeat transforms the image again: timing only)
      for (int r = 0; r < reps; ++r) {
        const L0Dev ip = l0_dev(*sc, *l0, st);
        LAUNCH(launch_fwd_l0<T>(A, p, W, H, ll, llp, D, Dp, cf, ip, S0, l0->plan.grid, st));
      }
    } else if (j == 0) {
      for (int r = 0; r < reps; ++r) LAUNCH(launch_fwd_level<T>(src, sp, Wj, Hj, ll, llp, det, dp, cf, st));
    } else {
      const RiderCopy<T> part = carried ? rider_part(rc, rows[j - 1], kLevelRiderCap) : RiderCopy<T>{};
      LAUNCH(launch_fwd_level<T>(src, sp, Wj, Hj, ll, llp, det, dp, cf, st, &part));
    }
  }
  if (jt < L) {
    // levels jt .. L-1 in one block launch: LL_jt from P[jt % 2], the region's
    // coefficients into A
    const RiderCopy<T> part = carried ? rider_part(rc, rows.back(), kBlockRiderCap) : RiderCopy<T>{};
    ProfScope ps(4, jt, st);
    LAUNCH(launch_block_fwd<T>(P[jt % 2], Pp[jt % 2], A, p, ceil_div_pow2(W, jt), ceil_div_pow2(H, jt), L - jt,
                               cf, st, &part));
  }
  if (!carried && ncopy > 0) {
    ProfScope ps(2, 0, st);
    copy_alone(rc, st);
  }
}

template <typename T>
static void fast_fwd(T* A, long long p, int W, int H, int L, const Coeffs<T>& cf, Scratch& sc,
                     hipStream_t st) {
  if (tail_start(W, H, L, (int)sizeof(T)) > 0 && g_schedule == 1) {
    fast_fwd_pyr<T>(A, p, W, H, L, cf, sc, st);
    return;
  }
  const FastBufs<T> fb = tail_start(W, H, L, (int)sizeof(T)) > 0 ? fast_bufs<T>(sc, W, H, L) : FastBufs<T>{};
  fast_fwd_on<T>(A, p, W, H, L, cf, fb, st, &sc);
}

// Inverse of fast_fwd: the block kernel does levels L .. jt+1 (writing LL_jt
// to P[jt % 2]); level j >= 2 writes LL_{j-1} to P; level 1 writes the image
// into A, so the level-0 details (and LL_1 when L == 1) are first copied out of
// A into D, carried by the launches before that last one.
template <typename T>
static void fast_inv_on(T* A, long long p, int W, int H, int L, const Coeffs<T>& cf, const FastBufs<T>& fb,
                        hipStream_t st, Scratch* sc = nullptr) {
  const int jt = tail_start(W, H, L, (int)sizeof(T));
  if (jt == 0) {
    ProfScope ps(5, 0, st);
    LAUNCH(launch_block_inv<T>(A, p, A, p, W, H, L, cf, st));
    return;
  }
  T* const D = fb.D;
  T* const P[2] = {fb.P[0], fb.P[1]};
  const long long Dp = fb.Dp, Pp[2] = {fb.Pp[0], fb.Pp[1]};
  const int jm = mid_start(W, H, jt, (int)sizeof(T));
  // carriers: the block launch, then the level launches jt .. 2 (each
  // producing LL_{j-1}; not the fused middle levels)
  std::vector<int> weight;
  if (jt < L) weight.push_back(kBlockCarrierWeight);
  for (int j = jt; j >= 2; --j) {
    const bool mid = jm > 0 && j - 1 >= jm && j - 1 < jm + kMidLevels;
    const long long bytes = (long long)ceil_div_pow2(W, j - 1) * ceil_div_pow2(H, j - 1) * (long long)sizeof(T);
    weight.push_back(!mid && bytes <= kCarrierBytes ? 1 : 0);
  }
  RiderCopy<T> rc = detail_copy<T>(A, p, D, Dp, W, H, L);
  // in place: only the preserved bands are copied
  const int S0 = l0_inplace_rows(sc, false, W, H, L, cf.nsteps, st);
  const L0Cached* l0 = S0 ? &l0_cached(*sc, false, W, H, cf.nsteps, S0, st) : nullptr;
  int ncopy = H;
  if (l0) {
    rc.rows = reinterpret_cast<const int*>(static_cast<const char*>(l0->dev) + l0->o_rows);
    ncopy = (int)l0->plan.rows.size();
    rc.r0 = 0;
    rc.r1 = ncopy;
  }
  std::vector<int2> rows;
  const bool carried = split_rows(ncopy, weight, rows);
  size_t carrier = 0;
  if (!carried && ncopy > 0) {
    ProfScope ps(2, 0, st);
    copy_alone(rc, st);
  }
  if (jt < L) {
    const RiderCopy<T> part = carried ? rider_part(rc, rows[carrier], kBlockRiderCap) : RiderCopy<T>{};
    ++carrier;
    ProfScope ps(5, jt, st);
    LAUNCH(launch_block_inv<T>(A, p, P[jt % 2], Pp[jt % 2], ceil_div_pow2(W, jt), ceil_div_pow2(H, jt), L - jt,
                               cf, st, &part));
  }
  for (int j = jt; j >= 1; --j) {
    const int Wf = ceil_div_pow2(W, j - 1), Hf = ceil_div_pow2(H, j - 1);
    if (jm > 0 && j == jm + kMidLevels) {
      // levels jm+2 .. jm in one launch: LL_{jm+3} from P (A when it is the
      // coarsest), details from A, LL_jm to P
      const T* llsrc = j == L ? A : P[j % 2];
      const long long lp = j == L ? p : Pp[j % 2];
      ProfScope ps(11, jm, st);
      LAUNCH(launch_mid_inv<T>(llsrc, lp, A, p, P[jm % 2], Pp[jm % 2], ceil_div_pow2(W, jm),
                               ceil_div_pow2(H, jm), kMidLevels, cf, st));
      carrier += kMidLevels;
      j = jm + 1;
      continue;
    }
    const T* ll;
    long long llp;
    if (j == L) {
      ll = L == 1 ? D : A;
      llp = L == 1 ? Dp : p;
    } else {
      ll = P[j % 2];
      llp = Pp[j % 2];
    }
    const T* det = j == 1 ? D : A;
    const long long dp = j == 1 ? Dp : p;
    T* dst = j == 1 ? A : P[(j - 1) % 2];
    const long long dstp = j == 1 ? p : Pp[(j - 1) % 2];
    ProfScope ps(1, j - 1, st);
    if (j >= 2) {
      const RiderCopy<T> part = carried ? rider_part(rc, rows[carrier], kLevelRiderCap) : RiderCopy<T>{};
      ++carrier;
      LAUNCH(launch_inv_level<T>(ll, llp, det, dp, dst, dstp, Wf, Hf, cf, st, &part));
    } else if (l0) {
      const L0Dev ip = l0_dev(*sc, *l0, st);
      LAUNCH(launch_inv_l0<T>(ll, llp, A, p, D, Dp, W, H, cf, ip, S0, l0->plan.grid, st));
    } else {
      LAUNCH(launch_inv_level<T>(ll, llp, det, dp, dst, dstp, Wf, Hf, cf, st));
    }
  }
}

template <typename T>
static void fast_inv(T* A, long long p, int W, int H, int L, const Coeffs<T>& cf, Scratch& sc,
                     hipStream_t st) {
  if (tail_start(W, H, L, (int)sizeof(T)) > 0 && g_schedule == 1) {
    fast_inv_pyr<T>(A, p, W, H, L, cf, sc, st);
    return;
  }
  const FastBufs<T> fb = tail_start(W, H, L, (int)sizeof(T)) > 0 ? fast_bufs<T>(sc, W, H, L) : FastBufs<T>{};
  fast_inv_on<T>(A, p, W, H, L, cf, fb, st, &sc);
}

// ---- 2D driver on device-accessible memory -------------------------------------
template <typename T>
static bool packed_layout(const char* ptr, long long sx, long long sy, int Wo, int Ho, int Wi,
                          int Hi) {
  return sy == (long long)sizeof(T) && sx > 0 && sx % (long long)sizeof(T) == 0 &&
         sx >= (long long)Wo * (long long)sizeof(T) &&
         (reinterpret_cast<uintptr_t>(ptr) % sizeof(T)) == 0 && Wo == Wi && Ho == Hi;
}

// A sparse frame (outer != inner) with contiguous rows at an element pitch:
// its small levels can run in LDS (launch_block_sparse).
template <typename T>
static bool sparse_rows_layout(const char* ptr, long long sx, long long sy, int Wo, int Ho, int Wi, int Hi) {
  return sy == (long long)sizeof(T) && sx > 0 && sx % (long long)sizeof(T) == 0 &&
         sx >= (long long)Wo * (long long)sizeof(T) && (reinterpret_cast<uintptr_t>(ptr) % sizeof(T)) == 0 &&
         (Wo != Wi || Ho != Hi);
}

// Levels [0, n) of a sparse frame whose inner sides stay >= 2: the fused
// sparse level launches take them.
static int sparse_fused_levels(int Wi, int Hi, int J) {
  int j = 0;
  while (j < J && ceil_div_pow2(Wi, j) >= 2 && ceil_div_pow2(Hi, j) >= 2) ++j;
  return j;
}

// Copy of a level's W x H output region from scratch into the frame, except
// the next level's Wn x Hn region (0 x 0: all of it), as a RiderCopy.
template <typename T>
static RiderCopy<T> region_copy(const T* src, long long sp, T* dst, long long dp, int W, int H, int Wn, int Hn) {
  RiderCopy<T> rc{};
  rc.src = src;
  rc.sp = sp;
  rc.dst = dst;
  rc.dp = dp;
  rc.xa = Hn > 0 ? Wn : 0;
  rc.wa = W - rc.xa;
  rc.ha = Hn > 0 ? Hn : H;
  rc.wb = W;
  rc.hb = H - rc.ha;
  rc.r0 = 0;
  rc.r1 = H;
  rc.blocks = std::min(H, 2048);
  constexpr int per = 16 / (int)sizeof(T);
  rc.vec = (reinterpret_cast<uintptr_t>(src) % 16) == 0 && (reinterpret_cast<uintptr_t>(dst) % 16) == 0 &&
           sp % per == 0 && dp % per == 0 && rc.xa % per == 0;
  return rc;
}

// Forward levels [0, jn) of a sparse frame with contiguous rows (A, element
// pitch p), each one launch reading the previous level's output and writing
// its own to scratch (Q0: level 0's outer size, Q1: level 1's, alternating).
// The part of a level's output the next level does not rewrite is copied into
// the frame by riders on the next launch; the last level's whole region by a
// copy of its own, so that what follows (the block tail or the generic levels)
// finds the frame complete.
template <typename T>
static void sparse_fwd_levels(T* A, long long p, int Wo, int Ho, int Wi, int Hi, int jn, int zp, const Coeffs<T>& cf,
                              Scratch& sc, hipStream_t st) {
  T* Q[2];
  long long Qp[2];
  for (int i = 0; i < 2 && i < jn; ++i) {
    Qp[i] = (long long)pad_pitch(ceil_div_pow2(Wo, i));
    Q[i] = (T*)sc.get(i == 0 ? 2 : 1, sizeof(T) * (size_t)Qp[i] * ceil_div_pow2(Ho, i));
  }
  const T* in = A;
  long long ip = p;
  RiderCopy<T> pending{};
  for (int j = 0; j < jn; ++j) {
    const int W = ceil_div_pow2(Wo, j), H = ceil_div_pow2(Ho, j);
    T* q = Q[j & 1];
    {
      ProfScope ps(0, j, st);
      LAUNCH(launch_sparse_fwd<T>(in, ip, q, Qp[j & 1], W, H, ceil_div_pow2(Wi, j), ceil_div_pow2(Hi, j), zp, cf,
                                  &pending, st));
    }
    pending = region_copy<T>(q, Qp[j & 1], A, p, W, H, (W + 1) / 2, (H + 1) / 2);
    in = q;
    ip = Qp[j & 1];
  }
  if (jn > 0) {
    const int W = ceil_div_pow2(Wo, jn - 1), H = ceil_div_pow2(Ho, jn - 1);
    const RiderCopy<T> all = region_copy<T>(Q[(jn - 1) & 1], Qp[(jn - 1) & 1], A, p, W, H, 0, 0);
    LAUNCH(launch_rider_alone<T>(all, st));
  }
}

// Inverse output levels jn - 1 .. 0 of a sparse frame (the coarser levels are
// in the frame already), each one launch reading the previous level's output
// and the frame's details and writing scratch.  Level 0 writes the frame
// itself: what it reads of the frame (everything outside the next level's
// region) is first preserved in scratch D by riders on the coarser levels'
// launches, which neither read nor write that part.  With one level only its
// output goes through scratch and is copied back.
template <typename T>
static void sparse_inv_levels(T* A, long long p, int Wo, int Ho, int Wi, int Hi, int jn, int zp, const Coeffs<T>& cf,
                              Scratch& sc, hipStream_t st) {
  if (jn <= 0) return;
  T* Q[2];
  long long Qp[2];
  for (int i = 0; i < 2 && i < jn; ++i) {
    Qp[i] = (long long)pad_pitch(ceil_div_pow2(Wo, i));
    Q[i] = (T*)sc.get(i == 0 ? 2 : 1, sizeof(T) * (size_t)Qp[i] * ceil_div_pow2(Ho, i));
  }
  const int Wod = (Wo + 1) / 2, Hod = (Ho + 1) / 2;
  T* D = nullptr;
  long long Dp = 0;
  RiderCopy<T> keep{};
  std::vector<int2> parts;
  if (jn >= 2) {
    Dp = (long long)pad_pitch(Wo);
    D = (T*)sc.get(0, sizeof(T) * (size_t)Dp * Ho);
    keep = region_copy<T>(A, p, D, Dp, Wo, Ho, Wod, Hod);
    split_rows(Ho, std::vector<int>(jn - 1, 1), parts);  // carried by levels jn - 1 .. 1
  }
  const T* prev = A;
  long long pp = p;
  for (int k = jn - 1; k >= 0; --k) {
    const bool direct = k == 0 && jn >= 2;
    T* q = direct ? A : Q[k & 1];
    const long long qp = direct ? p : Qp[k & 1];
    const RiderCopy<T> part = k > 0 && jn >= 2 ? rider_part(keep, parts[jn - 1 - k], kLevelRiderCap) : RiderCopy<T>{};
    {
      ProfScope ps(1, k, st);
      LAUNCH(launch_sparse_inv<T>(prev, pp, direct ? D : A, direct ? Dp : p, q, qp, ceil_div_pow2(Wo, k),
                                  ceil_div_pow2(Ho, k), ceil_div_pow2(Wi, k), ceil_div_pow2(Hi, k), zp, cf, &part,
                                  st));
    }
    prev = q;
    pp = qp;
  }
  if (jn == 1) {
    const RiderCopy<T> all = region_copy<T>(Q[0], Qp[0], A, p, Wo, Ho, 0, 0);
    LAUNCH(launch_rider_alone<T>(all, st));
  }
}

template <typename T>
static int fast_levels(int W, int H, int J) {
  int jf = 0;
  while (jf < J && ceil_div_pow2(W, jf) >= 2 && ceil_div_pow2(H, jf) >= 2) ++jf;
  return jf;
}

int level_limit(int Wo, int Ho, int d1) {
  const int mn = Wo < Ho ? Wo : Ho, mx = Wo < Ho ? Ho : Wo;
  return ceil_log2(d1 ? mx : mn);  // src/libdwt.c:6324
}

// Frames whose rows are not contiguous (interleaved channels, as the
// reference's OpenCV wrapper passes them per channel, src/cvdwt.cpp:67-147;
// transposed images) but whose strides are whole elements are restaged: the
// Wo x Ho frame is copied into packed scratch, transformed there on the
// packed paths, and copied back.  The transform depends only on the frame's
// values and sizes, so the result is the same as in place.
template <typename T>
static bool restage_layout(const char* ptr, long long sx, long long sy) {
  const long long es = sizeof(T);
  return sy != es && sx != 0 && sy != 0 && sx % es == 0 && sy % es == 0 &&
         reinterpret_cast<uintptr_t>(ptr) % es == 0;
}

template <typename T>
static void fwd2_device(char* ptr, long long sx, long long sy, int Wo, int Ho, int Wi, int Hi,
                        int* jp, int d1, int zp, Wavelet w, hipStream_t st);
template <typename T>
static void inv2_device(char* ptr, long long sx, long long sy, int Wo, int Ho, int Wi, int Hi,
                        int jmax, int d1, int zp, Wavelet w, hipStream_t st);

template <typename T>
static T* restage_in(char* ptr, long long sx, long long sy, int Wo, int Ho, long long* ps,
                     hipStream_t st) {
  const long long es = sizeof(T);
  *ps = (long long)pad_pitch(Wo);
  T* S = (T*)scratch_for(st).get(8, sizeof(T) * (size_t)(*ps) * Ho);
  LAUNCH(launch_copy_strided<T>(reinterpret_cast<const T*>(ptr), sx / es, sy / es, S, *ps, 1, Wo, Ho, st));
  return S;
}

template <typename T>
static void restage_out(const T* S, long long ps, char* ptr, long long sx, long long sy, int Wo, int Ho,
                        hipStream_t st) {
  const long long es = sizeof(T);
  LAUNCH(launch_copy_strided<T>(S, ps, 1, reinterpret_cast<T*>(ptr), sx / es, sy / es, Wo, Ho, st));
}

template <typename T>
static void fwd2_device(char* ptr, long long sx, long long sy, int Wo, int Ho, int Wi, int Hi,
                        int* jp, int d1, int zp, Wavelet w, hipStream_t st) {
  init_once();
  const int jl = level_limit(Wo, Ho, d1);
  if (*jp < 0 || *jp > jl) *jp = jl;  // src/libdwt.c:6326-6329
  const int J = *jp;
  if (J == 0 || Wo <= 0 || Ho <= 0) return;
  if (restage_layout<T>(ptr, sx, sy)) {
    long long ps;
    T* S = restage_in<T>(ptr, sx, sy, Wo, Ho, &ps, st);
    int j = J;
    fwd2_device<T>(reinterpret_cast<char*>(S), ps * (long long)sizeof(T), sizeof(T), Wo, Ho, Wi, Hi,
                   &j, d1, zp, w, st);
    restage_out<T>(S, ps, ptr, sx, sy, Wo, Ho, st);
    return;
  }
  Scratch& sc = scratch_for(st);
  const Coeffs<T> cf = make_coeffs<T>(w, true);
  int jf = 0;
  if (packed_layout<T>(ptr, sx, sy, Wo, Ho, Wi, Hi)) jf = fast_levels<T>(Wo, Ho, J);
  if (jf > 0) fast_fwd<T>(reinterpret_cast<T*>(ptr), sx / (long long)sizeof(T), Wo, Ho, jf, cf, sc, st);
  // sparse frames: the large levels as fused level launches, the small ones
  // in one block launch
  const bool sparse = jf == 0 && sparse_rows_layout<T>(ptr, sx, sy, Wo, Ho, Wi, Hi);
  const int jt = sparse ? sparse_tail_start(Wo, Ho, Wi, Hi, J, (int)sizeof(T)) : J;
  if (sparse && g_sparse_fused) {
    jf = std::min(jt, sparse_fused_levels(Wi, Hi, J));
    sparse_fwd_levels<T>(reinterpret_cast<T*>(ptr), sx / (long long)sizeof(T), Wo, Ho, Wi, Hi, jf, zp, cf, sc, st);
  }
  for (int j = jf; j < jt; ++j) generic_fwd_level<T>(ptr, sx, sy, Wo, Ho, Wi, Hi, j, zp, cf, sc, st);
  if (jt < J) {
    ProfScope ps(4, jt, st);
    LAUNCH(launch_block_sparse<T>(true, reinterpret_cast<T*>(ptr), sx / (long long)sizeof(T), ceil_div_pow2(Wo, jt),
                                  ceil_div_pow2(Ho, jt), ceil_div_pow2(Wi, jt), ceil_div_pow2(Hi, jt), J - jt, zp, cf,
                                  st));
  }
}

template <typename T>
static void inv2_device(char* ptr, long long sx, long long sy, int Wo, int Ho, int Wi, int Hi,
                        int jmax, int d1, int zp, Wavelet w, hipStream_t st) {
  init_once();
  int J = level_limit(Wo, Ho, d1);
  if (jmax >= 0 && jmax < J) J = jmax;  // src/libdwt.c:6678-6681
  if (J == 0 || Wo <= 0 || Ho <= 0) return;
  if (restage_layout<T>(ptr, sx, sy)) {
    long long ps;
    T* S = restage_in<T>(ptr, sx, sy, Wo, Ho, &ps, st);
    inv2_device<T>(reinterpret_cast<char*>(S), ps * (long long)sizeof(T), sizeof(T), Wo, Ho, Wi, Hi, J,
                   d1, zp, w, st);
    restage_out<T>(S, ps, ptr, sx, sy, Wo, Ho, st);
    return;
  }
  Scratch& sc = scratch_for(st);
  const Coeffs<T> cf = make_coeffs<T>(w, false);
  int jf = 0;
  if (packed_layout<T>(ptr, sx, sy, Wo, Ho, Wi, Hi)) jf = fast_levels<T>(Wo, Ho, J);
  // sparse frames: the coarsest levels in one block launch, the large ones as
  // fused level launches
  const bool sparse = jf == 0 && sparse_rows_layout<T>(ptr, sx, sy, Wo, Ho, Wi, Hi);
  const int jt = sparse ? sparse_tail_start(Wo, Ho, Wi, Hi, J, (int)sizeof(T)) : J;
  const int jn = sparse && g_sparse_fused ? std::min(jt, sparse_fused_levels(Wi, Hi, J)) : 0;
  if (jt < J) {
    ProfScope ps(5, jt, st);
    LAUNCH(launch_block_sparse<T>(false, reinterpret_cast<T*>(ptr), sx / (long long)sizeof(T), ceil_div_pow2(Wo, jt),
                                  ceil_div_pow2(Ho, jt), ceil_div_pow2(Wi, jt), ceil_div_pow2(Hi, jt), J - jt, zp, cf,
                                  st));
  }
  for (int j = jt; j > std::max(jf, jn); --j) generic_inv_level<T>(ptr, sx, sy, Wo, Ho, Wi, Hi, j, zp, cf, sc, st);
  if (jn > 0)
    sparse_inv_levels<T>(reinterpret_cast<T*>(ptr), sx / (long long)sizeof(T), Wo, Ho, Wi, Hi, jn, zp, cf, sc, st);
  if (jf > 0) fast_inv<T>(reinterpret_cast<T*>(ptr), sx / (long long)sizeof(T), Wo, Ho, jf, cf, sc, st);
}

// Every channel of a multi-channel image (channel c: the W x H frame at
// ptr + c * sc bytes, strides sx / sy as the drivers take them): the channels
// are gathered into packed planes in one pass, each plane is transformed on
// the packed path as the drivers would transform that channel alone, and the
// planes are scattered back in one pass.  (Per-channel driver calls, as the
// reference's OpenCV wrapper makes them, src/cvdwt.cpp:67-147, restage each
// channel separately and so read every pixel's whole group of channels once
// per channel.)
template <typename T>
static void run_channels(bool fwd, char* ptr, long long sx, long long sy, long long sc, int C, int W, int H,
                         int* jp, int d1, int zp, Wavelet w, hipStream_t st) {
  init_once();
  const long long es = sizeof(T);
  if (C < 1 || W < 0 || H < 0) fail(DWT_GPU_EINVAL, "channels %d, size %d x %d", C, W, H);
  if (sx % es || sy % es || sc % es || reinterpret_cast<uintptr_t>(ptr) % es)
    fail(DWT_GPU_EINVAL, "pointer and strides must be whole elements");
  const int jl = level_limit(W, H, d1);
  int J;
  if (fwd) {
    J = (*jp < 0 || *jp > jl) ? jl : *jp;  // src/libdwt.c:6326-6329
    *jp = J;
  } else {
    J = (*jp >= 0 && *jp < jl) ? *jp : jl;  // src/libdwt.c:6678-6681
  }
  if (J == 0 || W == 0 || H == 0) return;
  const long long pp = (long long)pad_pitch(W), step = pp * H;
  T* planes = (T*)scratch_for(st).get(8, sizeof(T) * (size_t)step * C);
  T* img = reinterpret_cast<T*>(ptr);
  LAUNCH(launch_channels<T>(true, img, sx / es, sy / es, sc / es, C, W, H, planes, pp, step, st));
  for (int c = 0; c < C; ++c) {
    char* plane = reinterpret_cast<char*>(planes + c * step);
    if (fwd) {
      int j = J;
      fwd2_device<T>(plane, pp * es, es, W, H, W, H, &j, d1, zp, w, st);
    } else {
      inv2_device<T>(plane, pp * es, es, W, H, W, H, J, d1, zp, w, st);
    }
  }
  LAUNCH(launch_channels<T>(false, img, sx / es, sy / es, sc / es, C, W, H, planes, pp, step, st));
}

// ---- pointer classification and staging --------------------------------------------
// Returns the device-usable address of p, or nullptr for pageable host memory.
static void* device_address(const void* p) {
  if (!p) return nullptr;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    hipGetLastError();
    return nullptr;
  }
  switch (a.type) {
    case hipMemoryTypeDevice:
    case hipMemoryTypeManaged:
    case hipMemoryTypeUnified:
      return const_cast<void*>(p);
    case hipMemoryTypeHost:
      if (a.devicePointer) {
        return static_cast<char*>(a.devicePointer) +
               (static_cast<const char*>(p) - static_cast<const char*>(a.hostPointer));
      }
      return nullptr;
    default:
      return nullptr;
  }
}

// Pageable host frames move through two pinned chunks of about kStageChunk
// bytes, in whole rows: upload, the host copies rows of chunk c into one
// pinned chunk while the copy engine uploads chunk c - 1 from the other;
// download, the copy engine fills one chunk while the host copies the other
// out.  (A pageable hipMemcpy goes through the runtime's own small staging
// and moved a 4096^2 FP32 frame at under 2 GB/s.)  Rows are contiguous when
// stride_y is the element size; any other element stride is gathered and
// scattered sample by sample on the host.
constexpr size_t kStageChunk = 8u << 20;

template <typename T>
static void host_rows_in(char* dst, const char* src, long long sx, long long sy, int Wo, int y0, int n) {
  const size_t pitch = (size_t)Wo * sizeof(T);
  for (int r = 0; r < n; ++r) {
    const char* s = src + (long long)(y0 + r) * sx;
    char* d = dst + (size_t)r * pitch;
    if (sy == (long long)sizeof(T))
      std::memcpy(d, s, pitch);
    else
      for (int x = 0; x < Wo; ++x) std::memcpy(d + (size_t)x * sizeof(T), s + (long long)x * sy, sizeof(T));
  }
}
template <typename T>
static void host_rows_out(char* dst, const char* src, long long sx, long long sy, int Wo, int y0, int n) {
  const size_t pitch = (size_t)Wo * sizeof(T);
  for (int r = 0; r < n; ++r) {
    char* d = dst + (long long)(y0 + r) * sx;
    const char* s = src + (size_t)r * pitch;
    if (sy == (long long)sizeof(T))
      std::memcpy(d, s, pitch);
    else
      for (int x = 0; x < Wo; ++x) std::memcpy(d + (long long)x * sy, s + (size_t)x * sizeof(T), sizeof(T));
  }
}

template <typename T>
static void stage_in(Scratch& sc, const char* host, long long sx, long long sy, int Wo, int Ho, char* dev,
                     hipStream_t st) {
  const size_t pitch = (size_t)Wo * sizeof(T);
  const int R = (int)std::max<size_t>(1, kStageChunk / pitch);  // rows per chunk
  sc.ensure_pin((size_t)R * pitch);
  for (int c = 0, y = 0; y < Ho; ++c, y += R) {
    const int k = c & 1, n = std::min(R, Ho - y);
    if (c >= 2) HIPCHK(hipEventSynchronize(sc.pin_ev[k]));  // its last upload is done
    host_rows_in<T>((char*)sc.pin[k], host, sx, sy, Wo, y, n);
    HIPCHK(hipMemcpyAsync(dev + (size_t)y * pitch, sc.pin[k], (size_t)n * pitch, hipMemcpyHostToDevice, st));
    HIPCHK(hipEventRecord(sc.pin_ev[k], st));
  }
}

template <typename T>
static void stage_out(Scratch& sc, const char* dev, char* host, long long sx, long long sy, int Wo, int Ho,
                      hipStream_t st) {
  const size_t pitch = (size_t)Wo * sizeof(T);
  const int R = (int)std::max<size_t>(1, kStageChunk / pitch);
  sc.ensure_pin((size_t)R * pitch);
  const int nc = (Ho + R - 1) / R;
  auto issue = [&](int c) {
    const int k = c & 1, y = c * R, n = std::min(R, Ho - y);
    HIPCHK(hipMemcpyAsync(sc.pin[k], dev + (size_t)y * pitch, (size_t)n * pitch, hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(sc.pin_ev[k], st));
  };
  for (int c = 0; c < nc && c < 2; ++c) issue(c);
  for (int c = 0; c < nc; ++c) {
    const int k = c & 1, y = c * R, n = std::min(R, Ho - y);
    HIPCHK(hipEventSynchronize(sc.pin_ev[k]));
    host_rows_out<T>(host, (const char*)sc.pin[k], sx, sy, Wo, y, n);
    if (c + 2 < nc) issue(c + 2);
  }
}

template <typename T>
static void run2d(bool fwd, void* ptr, int sx, int sy, int Wo, int Ho, int Wi, int Hi, int* jp,
                  int d1, int zp, Wavelet w, hipStream_t st, bool sync) {
  if (!ptr) fail(DWT_GPU_EINVAL, "null image pointer");
  if (Wo < 0 || Ho < 0 || Wi < 0 || Hi < 0) fail(DWT_GPU_EINVAL, "negative frame size");
  void* dptr = device_address(ptr);
  if (dptr) {
    if (fwd)
      fwd2_device<T>((char*)dptr, sx, sy, Wo, Ho, Wi, Hi, jp, d1, zp, w, st);
    else
      inv2_device<T>((char*)dptr, sx, sy, Wo, Ho, Wi, Hi, *jp, d1, zp, w, st);
    if (sync) {
      HIPCHK(hipStreamSynchronize(st));
      if (pyr_failed(scratch_for(st), st)) fail(DWT_GPU_EHIP, "pyramid dependency wait timed out");
    if (l0_failed(scratch_for(st), st)) fail(DWT_GPU_EHIP, "in-place level-0 dependency wait timed out");
    }
    return;
  }
  // pageable host memory: stage the outer frame as a packed device image
  if (Wo == 0 || Ho == 0) {
    if (fwd) {
      const int jl = level_limit(Wo, Ho, d1);
      if (*jp < 0 || *jp > jl) *jp = jl;
    }
    return;
  }
  Scratch& sc = scratch_for(st);
  const size_t pitch = (size_t)Wo * sizeof(T);
  char* dev = (char*)sc.get(4, pitch * (size_t)Ho);
  stage_in<T>(sc, (const char*)ptr, sx, sy, Wo, Ho, dev, st);
  if (fwd)
    fwd2_device<T>(dev, (long long)pitch, sizeof(T), Wo, Ho, Wi, Hi, jp, d1, zp, w, st);
  else
    inv2_device<T>(dev, (long long)pitch, sizeof(T), Wo, Ho, Wi, Hi, *jp, d1, zp, w, st);
  stage_out<T>(sc, dev, (char*)ptr, sx, sy, Wo, Ho, st);
  if (pyr_failed(sc, st)) fail(DWT_GPU_EHIP, "pyramid dependency wait timed out");
}

// One level of one strided line (src/libdwt.c:5512-5603, 5779-5870).
// Forward: a = src, b = dst_l, c = dst_h.  Inverse: a = src_l, b = src_h, c = dst.
template <typename T>
static void run_line(bool fwd, const void* a, const void* b, void* c, int N, int stride, Wavelet w) {
  if (N < 0) fail(DWT_GPU_EINVAL, "negative line length");
  if (stride == 0) fail(DWT_GPU_EINVAL, "zero stride");
  if (N == 0) return;
  if (!a || !b || !c) fail(DWT_GPU_EINVAL, "null line pointer");
  init_once();
  hipStream_t st = hipStreamPerThread;
  const Coeffs<T> cf = make_coeffs<T>(w, fwd);
  Scratch& sc = scratch_for(st);
  const int lim = line_lds_limit_bytes() / (int)sizeof(T);
  void* da = device_address(a);
  void* db = device_address(b);
  void* dc = device_address(c);
  if (da && db && dc) {
    T* gt = N > lim ? (T*)sc.get(3, sizeof(T) * (size_t)N) : nullptr;
    LAUNCH(launch_line_pass<T>(fwd, (const char*)da, (const char*)db, (char*)dc, 0, stride, 1, N, cf, gt, 1, st));
    HIPCHK(hipStreamSynchronize(st));
    return;
  }
  if (da || db || dc) fail(DWT_GPU_EMEMORY, "line pointers mix host and device memory");
  // host line: gather every input sample first (the reference gathers the
  // whole line into tmp before any store, so any aliasing is well defined),
  // transform packed on the device, then scatter the outputs
  const int nl = (N + 1) / 2, nh = N / 2;
  std::vector<T> in((size_t)N), out((size_t)N);
  if (fwd) {
    for (int i = 0; i < N; ++i) std::memcpy(&in[i], (const char*)a + (long long)i * stride, sizeof(T));
  } else {
    for (int i = 0; i < nl; ++i) std::memcpy(&in[i], (const char*)a + (long long)i * stride, sizeof(T));
    for (int i = 0; i < nh; ++i) std::memcpy(&in[nl + i], (const char*)b + (long long)i * stride, sizeof(T));
  }
  char* dev = (char*)sc.get(3, sizeof(T) * (size_t)N * 2 + sizeof(T) * (size_t)(N > lim ? N : 0));
  T* din = (T*)dev;
  T* dout = din + N;
  T* gt = N > lim ? dout + N : nullptr;
  HIPCHK(hipMemcpyAsync(din, in.data(), sizeof(T) * N, hipMemcpyHostToDevice, st));
  const long long es = sizeof(T);
  if (fwd)
    LAUNCH(launch_line_pass<T>(true, (const char*)din, (const char*)dout, (char*)(dout + nl), 0, es, 1, N, cf, gt, 1, st));
  else
    LAUNCH(launch_line_pass<T>(false, (const char*)din, (const char*)(din + nl), (char*)dout, 0, es, 1, N, cf, gt, 1, st));
  HIPCHK(hipMemcpyAsync(out.data(), dout, sizeof(T) * N, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (fwd) {
    for (int k = 0; k < nl; ++k) std::memcpy((char*)b + (long long)k * stride, &out[k], sizeof(T));
    for (int k = 0; k < nh; ++k) std::memcpy((char*)c + (long long)k * stride, &out[nl + k], sizeof(T));
  } else {
    for (int i = 0; i < N; ++i) std::memcpy((char*)c + (long long)i * stride, &out[i], sizeof(T));
  }
}

// Zero [k0, k1) of a strided line (src/libdwt.c:5925-5942, 5988-6004).
template <typename T>
static void run_zero(void* p, int k0, int k1, int stride) {
  if (k1 <= k0) return;
  if (!p) fail(DWT_GPU_EINVAL, "null pointer");
  if (stride == 0) fail(DWT_GPU_EINVAL, "zero stride");
  void* dp = device_address(p);
  if (!dp) {
    const T z = 0;
    for (int k = k0; k < k1; ++k) std::memcpy((char*)p + (long long)k * stride, &z, sizeof(T));
    return;
  }
  hipStream_t st = hipStreamPerThread;
  const int zx[4] = {k0, k1, 0, 0};
  const int zy[4] = {0, 0, 0, 0};
  LAUNCH(launch_zero_pad<T>((char*)dp, 0, stride, k1, 1, zx, zy, st));
  HIPCHK(hipStreamSynchronize(st));
}

// ---- full-depth 1D transforms of many signals ---------------------------------------
// Contiguous signals: later levels of at least this many samples also run on
// whole rows in LDS (5.2 vs 4.5 TB/s for the windowed kernel at 4096 samples;
// from 512 samples on no faster: 0.863 vs 0.886 ms at 16384 x 8192).
constexpr int kResLevelMin = 2048;
// Level 0 of signals stored as matrix columns, in place, as kColStages
// launches over segment ranges [s[k], s[k+1]) plus one over the top segments
// [s[kColStages], nseg) whose H goes through scratch.  With R rows per
// segment, h0 = ceil(n/2) and NS lifting steps:
//   forward, launched from the top down: segments [a, b) read input rows
//   [R a - NS, R b + NS) and write H rows h0 + [R a / 2, R b / 2), which
//   only segments >= b read (already done) when R b + NS <= h0 + R a / 2;
//   inverse, launched from the bottom up: segments [a, b) write output rows
//   [R a, R b) and read H rows h0 + [R a / 2 - NS / 2, ...), which lie above
//   everything written so far when R b <= h0 + R a / 2 - NS / 2.
// Each boundary is the largest the inequality allows.  Returns false when
// the signal is too short for strictly increasing boundaries.
constexpr int kColStages = 3;
static bool col_stages(bool fwd, int n, int nsteps, int* s) {
  const int R = col_segment_rows(), h0 = (n + 1) / 2, nseg = col_segments(n);
  const int slack = fwd ? nsteps : nsteps / 2;
  s[0] = 0;
  for (int k = 1; k <= kColStages; ++k) {
    s[k] = (h0 + R / 2 * s[k - 1] - slack) / R;
    if (s[k] <= s[k - 1]) return false;
  }
  return s[kColStages] < nseg && R / 2 * s[kColStages] - nsteps / 2 >= 0;
}

// Signal k, sample i at ptr + k * sig + i * smp (bytes).  Levels as the caller
// loop over dwt_cdf97_f_ex_stride_s: each level on the current n (L to
// [0, ceil(n/2)), H to [ceil(n/2), n)) while n >= 2.
template <typename T>
static void run_lines(bool fwd, char* ptr, long long sig, long long smp, int N, int count,
                      int* levels, Wavelet w, hipStream_t st) {
  init_once();
  if (N < 0 || count < 0) fail(DWT_GPU_EINVAL, "negative length or count");
  const int jl = ceil_log2(N);
  int L = *levels;
  if (fwd) {
    if (L < 0 || L > jl) L = jl;
    *levels = L;
  } else if (L < 0 || L > jl) {
    fail(DWT_GPU_EINVAL, "%d levels outside 0..%d", L, jl);
  }
  if (L == 0 || count == 0) return;
  const long long es = sizeof(T);
  const Coeffs<T> cf = make_coeffs<T>(w, fwd);
  Scratch& sc = scratch_for(st);
  const bool aligned = (reinterpret_cast<uintptr_t>(ptr) % es) == 0;
  std::vector<int> n(L + 1);
  n[0] = N;
  for (int j = 1; j <= L; ++j) n[j] = (n[j - 1] + 1) / 2;
  if (aligned && smp == es && sig % es == 0 && sig > 0) {
    // contiguous samples: one launch per level over all signals (rows), the
    // tile body's row lifting.  Level 0's H overlaps input that other
    // workgroups have not read yet: a row that fits one workgroup's LDS is
    // read whole before anything is written (level 0 in place); a longer one
    // goes through scratch, copied into place on the side stream.  (Per-level
    // launches measured faster at every length from 16 to 16384 samples than
    // one workgroup per signal running every level in LDS: 1.20 vs 1.57 ms at
    // 16384 x 8192, 6.9 vs 63.8 ms at 16 samples per signal.)
    T* A = reinterpret_cast<T*>(ptr);
    const long long p = sig / es;
    const bool vec = (reinterpret_cast<uintptr_t>(ptr) % (2 * es)) == 0 && p % 2 == 0;
    const int h0 = n[1];
    const bool res = row_resident_fits(N, (int)es);
    const long long Dp = (long long)pad_pitch(L == 1 ? N : N - h0);
    T* D = res ? nullptr : (T*)sc.get(2, sizeof(T) * (size_t)Dp * count);
    T* P[2] = {nullptr, nullptr};
    long long Pp[2] = {(long long)pad_pitch(n[L >= 3 ? 2 : 1]), (long long)pad_pitch(n[1])};
    if (L >= 2) P[1] = (T*)sc.get(1, sizeof(T) * (size_t)Pp[1] * count);
    if (L >= 3) P[0] = (T*)sc.get(0, sizeof(T) * (size_t)Pp[0] * count);
    if (!res) sc.ensure_side();
    if (fwd) {
      if (res) {
        LAUNCH(launch_row_resident<T>(true, A, p, L == 1 ? A : P[1], L == 1 ? p : Pp[1], A + h0, p, N, count,
                                      cf, st));
      } else {
        // L == 1: the whole level-0 output (L | H) goes to D and back
        LAUNCH(launch_row_level<T>(true, A, p, L == 1 ? D : P[1], L == 1 ? Dp : Pp[1],
                                   L == 1 ? D + h0 : D, Dp, N, count, cf, vec, st));
        HIPCHK(hipEventRecord(sc.fork, st));
        HIPCHK(hipStreamWaitEvent(sc.side, sc.fork, 0));
        LAUNCH(launch_copy_rect<T>(D, Dp, L == 1 ? A : A + h0, p, L == 1 ? N : N - h0, count, sc.side));
      }
      for (int j = 1; j < L; ++j) {
        T* out = j == L - 1 ? A : P[(j + 1) % 2];
        const long long op = j == L - 1 ? p : Pp[(j + 1) % 2];
        if (n[j] >= kResLevelMin && row_resident_fits(n[j], (int)es))
          LAUNCH(launch_row_resident<T>(true, P[j % 2], Pp[j % 2], out, op, A + n[j + 1], p, n[j], count, cf, st));
        else
          LAUNCH(launch_row_level<T>(true, P[j % 2], Pp[j % 2], out, op, A + n[j + 1], p, n[j], count, cf,
                                     true, st));
      }
      if (!res) {
        HIPCHK(hipEventRecord(sc.join, sc.side));
        HIPCHK(hipStreamWaitEvent(st, sc.join, 0));
      }
    } else {
      if (!res) {
        // level 0's H (all of level 0's input when L == 1) out of A first
        HIPCHK(hipEventRecord(sc.fork, st));
        HIPCHK(hipStreamWaitEvent(sc.side, sc.fork, 0));
        LAUNCH(launch_copy_rect<T>(L == 1 ? A : A + h0, p, D, Dp, L == 1 ? N : N - h0, count, sc.side));
        HIPCHK(hipEventRecord(sc.join, sc.side));
      }
      for (int j = L - 1; j >= 1; --j) {
        const T* lin = j == L - 1 ? A : P[(j + 1) % 2];
        const long long lp = j == L - 1 ? p : Pp[(j + 1) % 2];
        if (n[j] >= kResLevelMin && row_resident_fits(n[j], (int)es))
          LAUNCH(launch_row_resident<T>(false, lin, lp, A + n[j + 1], p, P[j % 2], Pp[j % 2], n[j], count, cf, st));
        else
          LAUNCH(launch_row_level<T>(false, lin, lp, A + n[j + 1], p, P[j % 2], Pp[j % 2], n[j], count, cf,
                                     true, st));
      }
      if (res) {
        LAUNCH(launch_row_resident<T>(false, L == 1 ? A : P[1], L == 1 ? p : Pp[1], A + h0, p, A, p, N, count,
                                      cf, st));
      } else {
        HIPCHK(hipStreamWaitEvent(st, sc.join, 0));
        LAUNCH(launch_row_level<T>(false, L == 1 ? D : P[1], L == 1 ? Dp : Pp[1], L == 1 ? D + h0 : D, Dp, A,
                                   p, N, count, cf, vec, st));
      }
    }
    return;
  }
  if (aligned && sig == es && smp % es == 0 && smp / es >= count) {
    // adjacent signals = the columns of a matrix of pitch p
    T* A = reinterpret_cast<T*>(ptr);
    const long long p = smp / es;
    const bool vec = (reinterpret_cast<uintptr_t>(ptr) % (2 * es)) == 0 && p % 2 == 0;
    const long long cp = (long long)pad_pitch(count);
    const int h0 = n[1], nh0 = N - h0;
    const int nseg0 = col_segments(N);
    int stg[kColStages + 1];
    const bool staged = L >= 2 && col_stages(fwd, N, cf.nsteps, stg);
    // staged: the top H rows [hs, nh0) pass through scratch (fwd: H rows from
    // the top segments' first pair; inv: also the rows the lowest of them reads)
    const int hs = staged ? (fwd ? stg[kColStages] * col_segment_rows() / 2
                                 : stg[kColStages] * col_segment_rows() / 2 - cf.nsteps / 2)
                          : 0;
    T* D = (T*)sc.get(2, sizeof(T) * (size_t)cp * (staged ? nh0 - hs : N));
    T* P[2] = {nullptr, nullptr};
    if (L >= 2) P[1] = (T*)sc.get(1, sizeof(T) * (size_t)cp * n[1]);
    if (L >= 3) P[0] = (T*)sc.get(0, sizeof(T) * (size_t)cp * n[2]);
    sc.ensure_side();
    if (fwd) {
      if (staged) {
        // level 0 in place, top segments first: segments [stg[k], stg[k+1])
        // write H rows that only segments >= stg[k+1] read (done before);
        // the top segments' H, over rows they read themselves, to scratch
        LAUNCH(launch_col_segments<T>(true, A, p, P[1], cp, D, cp, N, count, stg[kColStages],
                                      nseg0 - stg[kColStages], -hs, cf, vec, st));
        for (int k = kColStages - 1; k >= 0; --k)
          LAUNCH(launch_col_segments<T>(true, A, p, P[1], cp, A, p, N, count, stg[k], stg[k + 1] - stg[k], h0,
                                        cf, vec, st));
        HIPCHK(hipEventRecord(sc.fork, st));
        HIPCHK(hipStreamWaitEvent(sc.side, sc.fork, 0));
        LAUNCH(launch_copy_rect<T>(D, cp, A + (long long)(h0 + hs) * p, p, count, nh0 - hs, sc.side));
      } else {
        // level 0 into scratch (its H rows overlap unread input), copied into
        // A on the side stream; later levels read scratch and write A directly
        LAUNCH(launch_col_level<T>(true, A, p, L == 1 ? D : P[1], cp, D, cp, N, count, cf, vec, st));
        HIPCHK(hipEventRecord(sc.fork, st));
        HIPCHK(hipStreamWaitEvent(sc.side, sc.fork, 0));
        LAUNCH(launch_copy_rect<T>(D + (L == 1 ? 0 : (long long)h0 * cp), cp, A + (L == 1 ? 0 : (long long)h0 * p),
                                   p, count, L == 1 ? N : N - h0, sc.side));
      }
      for (int j = 1; j < L; ++j) {
        T* out = j == L - 1 ? A : P[(j + 1) % 2];
        const long long op = j == L - 1 ? p : cp;
        LAUNCH(launch_col_level<T>(true, P[j % 2], cp, out, op, A, p, n[j], count, cf, true, st));
      }
      HIPCHK(hipEventRecord(sc.join, sc.side));
      HIPCHK(hipStreamWaitEvent(st, sc.join, 0));
    } else {
      HIPCHK(hipEventRecord(sc.fork, st));
      HIPCHK(hipStreamWaitEvent(sc.side, sc.fork, 0));
      if (staged)
        LAUNCH(launch_copy_rect<T>(A + (long long)(h0 + hs) * p, p, D, cp, count, nh0 - hs, sc.side));
      else
        LAUNCH(launch_copy_rect<T>(A + (L == 1 ? 0 : (long long)h0 * p), p, D + (L == 1 ? 0 : (long long)h0 * cp),
                                   cp, count, L == 1 ? N : N - h0, sc.side));
      HIPCHK(hipEventRecord(sc.join, sc.side));
      for (int j = L - 1; j >= 1; --j) {
        const T* llp = j == L - 1 ? A : P[(j + 1) % 2];
        const long long lp = j == L - 1 ? p : cp;
        // H rows always come from A: pair-wise loads only if A allows them
        LAUNCH(launch_col_level<T>(false, llp, lp, A, p, P[j % 2], cp, n[j], count, cf, vec, st));
      }
      if (staged) {
        // level 0 in place, bottom segments first: segments [stg[k], stg[k+1])
        // overwrite only H rows below those that segments >= stg[k] read;
        // the top segments read their H from the scratch copy
        for (int k = 0; k < kColStages; ++k)
          LAUNCH(launch_col_segments<T>(false, P[1], cp, A, p, A, p, N, count, stg[k], stg[k + 1] - stg[k], h0,
                                        cf, vec, st));
        HIPCHK(hipStreamWaitEvent(st, sc.join, 0));
        LAUNCH(launch_col_segments<T>(false, P[1], cp, D, cp, A, p, N, count, stg[kColStages],
                                      nseg0 - stg[kColStages], -hs, cf, vec, st));
      } else {
        HIPCHK(hipStreamWaitEvent(st, sc.join, 0));
        LAUNCH(launch_col_level<T>(false, L == 1 ? D : P[1], cp, D, cp, A, p, N, count, cf, true, st));
      }
    }
    return;
  }
  // any other layout: one line-parallel launch per level, in place
  const int lim = line_lds_limit_bytes() / (int)sizeof(T);
  const int gblocks = 1024;
  for (int k = 0; k < L; ++k) {
    const int j = fwd ? k : L - 1 - k;
    const int m = n[j];
    T* gt = m > lim ? (T*)sc.get(3, sizeof(T) * (size_t)m * gblocks) : nullptr;
    if (fwd)
      LAUNCH(launch_line_pass<T>(true, ptr, ptr, ptr + (long long)n[j + 1] * smp, sig, smp, count, m, cf, gt, gblocks, st));
    else
      LAUNCH(launch_line_pass<T>(false, ptr, ptr + (long long)n[j + 1] * smp, ptr, sig, smp, count, m, cf, gt, gblocks, st));
  }
}

[[noreturn]] static void abort_with(const char* fn) {
  std::fprintf(stderr, "libdwt (MI355X): %s: %s\n", fn, g_err.c_str());
  std::fflush(stderr);
  std::abort();
}

// ---- allocation-free entry point: scratch in a static device array ------------
// dwt_gpu_2d_noalloc runs the packed fast path on scratch carved out of this
// file-scope array instead of allocated memory.  Every GPU the library runs on
// holds its own copy of it until the process ends; on the host the array's
// name is only a stand-in, so its device address is taken once per device with
// hipGetSymbolAddress and only that address is used.
constexpr size_t kNoAllocScratchBytes = (size_t)DWT_GPU_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_noalloc_mu;
static std::map<int, char*>* g_noalloc_ptr = new std::map<int, char*>();

static char* noalloc_scratch() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  auto it = g_noalloc_ptr->find(dev);
  if (it != g_noalloc_ptr->end()) return it->second;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    fail(DWT_GPU_EHIP, "hipGetSymbolAddress of the static scratch: %s", hipGetErrorString(e));
  }
  (*g_noalloc_ptr)[dev] = static_cast<char*>(ptr);
  return static_cast<char*>(ptr);
}

// The packed FP32 transform of dwt_gpu_2d_async on the static scratch; levels
// as the forward driver clamps them (decompose_one = 0, so every level has both
// sides >= 2 and takes the packed path).
static void run2d_noalloc(bool fwd, Wavelet w, void* ptr, int sx, int W, int H, int* jp, hipStream_t st) {
  init_once();
  if (!ptr) fail(DWT_GPU_EINVAL, "null image pointer");
  if (!jp) fail(DWT_GPU_EINVAL, "null j");
  if (W < 0 || H < 0) fail(DWT_GPU_EINVAL, "negative frame size");
  if (sx <= 0 || sx % (int)sizeof(float) != 0 || (long long)sx < (long long)W * (long long)sizeof(float))
    fail(DWT_GPU_EINVAL, "stride_x must be a positive multiple of 4 bytes and hold a row");
  if (reinterpret_cast<uintptr_t>(ptr) % sizeof(float) != 0) fail(DWT_GPU_EINVAL, "image pointer not 4-byte aligned");
  void* dptr = device_address(ptr);
  if (!dptr || dptr != ptr) fail(DWT_GPU_EMEMORY, "image pointer is not device memory");
  const int jl = level_limit(W, H, 0);
  int J = *jp;
  if (fwd) {
    if (J < 0 || J > jl) J = jl;  // src/libdwt.c:6326-6329
    *jp = J;
  } else if (J < 0 || J > jl) {
    J = jl;  // src/libdwt.c:6678-6681
  }
  if (J == 0 || W == 0 || H == 0) return;
  FastBufs<float> fb;
  if (tail_start(W, H, J, (int)sizeof(float)) > 0) {
    const size_t need = fast_bufs_layout<float>(W, H, J, nullptr, &fb);
    if (need > kNoAllocScratchBytes)
      fail(DWT_GPU_EINVAL, "%d x %d needs %zu bytes of scratch; the allocation-free entry point holds %zu", W, H,
           need, kNoAllocScratchBytes);
    fast_bufs_layout<float>(W, H, J, noalloc_scratch(), &fb);
  }
  const Coeffs<float> cf = make_coeffs<float>(w, fwd);
  float* A = static_cast<float*>(ptr);
  const long long p = sx / (long long)sizeof(float);
  if (fwd)
    fast_fwd_on<float>(A, p, W, H, J, cf, fb, st);
  else
    fast_inv_on<float>(A, p, W, H, J, cf, fb, st);
}

}  // namespace dwt

using namespace dwt;

#define DROPIN(fn, body)                 \
  do {                                   \
    try {                                \
      body;                              \
    } catch (const Failure&) {           \
      abort_with(fn);                    \
    }                                    \
  } while (0)

extern "C" {

// ---- 1D ---------------------------------------------------------------------------
#define LINE_FNS(W, WENUM, T, SUF)                                                              \
  void dwt_##W##_f_ex_stride_##SUF(const T* src, T* dst_l, T* dst_h, T* tmp, int N, int stride) { \
    (void)tmp;                                                                                  \
    DROPIN(__func__, run_line<T>(true, src, dst_l, dst_h, N, stride, WENUM));                   \
  }                                                                                             \
  void dwt_##W##_i_ex_stride_##SUF(const T* src_l, const T* src_h, T* dst, T* tmp, int N,         \
                                   int stride) {                                                \
    (void)tmp;                                                                                  \
    DROPIN(__func__, run_line<T>(false, src_l, src_h, dst, N, stride, WENUM));                  \
  }                                                                                             \
  void dwt_##W##_f_ex_##SUF(const T* src, T* dst_l, T* dst_h, T* tmp, int N) {                    \
    dwt_##W##_f_ex_stride_##SUF(src, dst_l, dst_h, tmp, N, (int)sizeof(T));                     \
  }                                                                                             \
  void dwt_##W##_i_ex_##SUF(const T* src_l, const T* src_h, T* dst, T* tmp, int N) {              \
    dwt_##W##_i_ex_stride_##SUF(src_l, src_h, dst, tmp, N, (int)sizeof(T));                     \
  }                                                                                             \
  void dwt_##W##_f_##SUF(const T* src, T* dst, T* tmp, int N) {                                 \
    dwt_##W##_f_ex_##SUF(src, dst, dst + (N + 1) / 2, tmp, N);                                  \
  }                                                                                             \
  void dwt_##W##_i_##SUF(const T* src, T* dst, T* tmp, int N) {                                 \
    dwt_##W##_i_ex_##SUF(src, src + (N + 1) / 2, dst, tmp, N);                                  \
  }

LINE_FNS(cdf97, CDF97, float, s)
LINE_FNS(cdf53, CDF53, float, s)
LINE_FNS(cdf97, CDF97, double, d)
LINE_FNS(cdf53, CDF53, double, d)

// ---- zero padding ------------------------------------------------------------------
#define ZERO_FNS(T, SUF)                                                                          \
  void dwt_zero_padding_f_stride_##SUF(T* dst_l, T* dst_h, int N, int N_dst_L, int N_dst_H,       \
                                       int stride) {                                              \
    DROPIN(__func__, {                                                                            \
      if (N < 0 || N_dst_L < 0 || N_dst_H < 0) fail(DWT_GPU_EINVAL, "negative size");             \
      run_zero<T>(dst_l, (N + 1) / 2, N_dst_L, stride);                                           \
      run_zero<T>(dst_h, N / 2, N_dst_H, stride);                                                 \
    });                                                                                           \
  }                                                                                               \
  void dwt_zero_padding_f_##SUF(T* dst_l, T* dst_h, int N, int N_dst_L, int N_dst_H) {            \
    dwt_zero_padding_f_stride_##SUF(dst_l, dst_h, N, N_dst_L, N_dst_H, (int)sizeof(T));           \
  }                                                                                               \
  void dwt_zero_padding_i_stride_##SUF(T* dst_l, int N, int N_dst, int stride) {                  \
    DROPIN(__func__, {                                                                            \
      if (N < 0 || N_dst < 0) fail(DWT_GPU_EINVAL, "negative size");                              \
      run_zero<T>(dst_l, N, N_dst, stride);                                                       \
    });                                                                                           \
  }                                                                                               \
  void dwt_zero_padding_i_##SUF(T* dst_l, int N, int N_dst) {                                     \
    dwt_zero_padding_i_stride_##SUF(dst_l, N, N_dst, (int)sizeof(T));                             \
  }

ZERO_FNS(float, s)
ZERO_FNS(double, d)

// ---- 2D ------------------------------------------------------------------------------
#define TWO_D_FNS(W, WENUM, T, SUF)                                                             \
  void dwt_##W##_2f_##SUF(void* ptr, int stride_x, int stride_y, int size_o_big_x,              \
                          int size_o_big_y, int size_i_big_x, int size_i_big_y, int* j_max_ptr, \
                          int decompose_one, int zero_padding) {                                \
    DROPIN(__func__, {                                                                          \
      if (!j_max_ptr) fail(DWT_GPU_EINVAL, "null j_max_ptr");                                   \
      run2d<T>(true, ptr, stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x,         \
               size_i_big_y, j_max_ptr, decompose_one, zero_padding, WENUM, hipStreamPerThread, \
               true);                                                                           \
    });                                                                                         \
  }                                                                                             \
  void dwt_##W##_2i_##SUF(void* ptr, int stride_x, int stride_y, int size_o_big_x,              \
                          int size_o_big_y, int size_i_big_x, int size_i_big_y, int j_max,      \
                          int decompose_one, int zero_padding) {                                \
    DROPIN(__func__, {                                                                          \
      int j = j_max;                                                                            \
      run2d<T>(false, ptr, stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x,        \
               size_i_big_y, &j, decompose_one, zero_padding, WENUM, hipStreamPerThread, true); \
    });                                                                                         \
  }

TWO_D_FNS(cdf97, CDF97, float, s)
TWO_D_FNS(cdf53, CDF53, float, s)
TWO_D_FNS(cdf97, CDF97, double, d)
TWO_D_FNS(cdf53, CDF53, double, d)

// ---- extensions ------------------------------------------------------------------------
int dwt_gpu_init(void) {
  try {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (n <= 0) fail(DWT_GPU_EHIP, "no HIP device");
    init_once();
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

const char* dwt_gpu_last_error(void) { return g_err.c_str(); }

const char* dwt_gpu_version(void) { return "libdwt-mi355x 0.1 (gfx950)"; }

int dwt_gpu_2d_async(int wavelet, int direction, int precision, void* ptr, int stride_x,
                     int stride_y, int size_o_big_x, int size_o_big_y, int size_i_big_x,
                     int size_i_big_y, int* j, int decompose_one, int zero_padding, void* stream) {
  try {
    g_err.clear();
    if (wavelet != DWT_GPU_CDF97 && wavelet != DWT_GPU_CDF53) fail(DWT_GPU_EINVAL, "bad wavelet");
    if (direction != DWT_GPU_FORWARD && direction != DWT_GPU_INVERSE) fail(DWT_GPU_EINVAL, "bad direction");
    if (precision != DWT_GPU_F32 && precision != DWT_GPU_F64) fail(DWT_GPU_EINVAL, "bad precision");
    if (!j) fail(DWT_GPU_EINVAL, "null j");
    if (!ptr) fail(DWT_GPU_EINVAL, "null image pointer");
    void* dptr = device_address(ptr);
    if (!dptr) fail(DWT_GPU_EMEMORY, "image pointer is not device-accessible memory");
    const Wavelet w = wavelet == DWT_GPU_CDF97 ? CDF97 : CDF53;
    const bool fwd = direction == DWT_GPU_FORWARD;
    hipStream_t st = (hipStream_t)stream;
    if (precision == DWT_GPU_F32)
      run2d<float>(fwd, dptr, stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x,
                   size_i_big_y, j, decompose_one, zero_padding, w, st, false);
    else
      run2d<double>(fwd, dptr, stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x,
                    size_i_big_y, j, decompose_one, zero_padding, w, st, false);
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_2d_noalloc(int wavelet, int direction, void* ptr, int stride_x, int size_x, int size_y, int* j,
                       void* stream) {
  try {
    g_err.clear();
    if (wavelet != DWT_GPU_CDF97 && wavelet != DWT_GPU_CDF53) fail(DWT_GPU_EINVAL, "bad wavelet");
    if (direction != DWT_GPU_FORWARD && direction != DWT_GPU_INVERSE) fail(DWT_GPU_EINVAL, "bad direction");
    run2d_noalloc(direction == DWT_GPU_FORWARD, wavelet == DWT_GPU_CDF97 ? CDF97 : CDF53, ptr, stride_x, size_x,
                  size_y, j, (hipStream_t)stream);
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_2d_channels(int wavelet, int direction, int precision, void* ptr, int stride_x, int stride_y,
                        int channel_stride, int channels, int size_x, int size_y, int* j, int decompose_one,
                        int zero_padding, void* stream) {
  try {
    g_err.clear();
    if (wavelet != DWT_GPU_CDF97 && wavelet != DWT_GPU_CDF53) fail(DWT_GPU_EINVAL, "bad wavelet");
    if (direction != DWT_GPU_FORWARD && direction != DWT_GPU_INVERSE) fail(DWT_GPU_EINVAL, "bad direction");
    if (precision != DWT_GPU_F32 && precision != DWT_GPU_F64) fail(DWT_GPU_EINVAL, "bad precision");
    if (!j) fail(DWT_GPU_EINVAL, "null j");
    if (!ptr) fail(DWT_GPU_EINVAL, "null image pointer");
    void* dptr = device_address(ptr);
    if (!dptr) fail(DWT_GPU_EMEMORY, "image pointer is not device-accessible memory");
    const Wavelet w = wavelet == DWT_GPU_CDF97 ? CDF97 : CDF53;
    const bool fwd = direction == DWT_GPU_FORWARD;
    hipStream_t st = (hipStream_t)stream;
    if (precision == DWT_GPU_F32)
      run_channels<float>(fwd, (char*)dptr, stride_x, stride_y, channel_stride, channels, size_x, size_y, j,
                          decompose_one, zero_padding, w, st);
    else
      run_channels<double>(fwd, (char*)dptr, stride_x, stride_y, channel_stride, channels, size_x, size_y, j,
                           decompose_one, zero_padding, w, st);
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

// ---- device-side compare and FP64 verification ------------------------------------
static void check_image(const void* p, int sx, int sy, int elem, const char* what) {
  if (!p) fail(DWT_GPU_EINVAL, "%s: null pointer", what);
  if (sx % elem || sy % elem || reinterpret_cast<uintptr_t>(p) % elem)
    fail(DWT_GPU_EINVAL, "%s: pointer and strides must be multiples of the element size", what);
}

static void compare_device(bool a64, const void* a, int sxa, int sya, bool b64, const void* b, int sxb,
                           int syb, int W, int H, double eps, double peak, dwt_gpu_compare_stats* out,
                           hipStream_t st) {
  std::memset(out, 0, sizeof(*out));
  out->psnr = INFINITY;
  if (W == 0 || H == 0) return;
  const long long n = (long long)W * H;
  Scratch& sc = scratch_for(st);
  char* s = (char*)sc.get(6, compare_scratch_bytes(n));
  if (!sc.cmp_host) HIPCHK(hipHostMalloc(&sc.cmp_host, compare_result_bytes(), hipHostMallocDefault));
  LAUNCH(launch_compare(a64, a, sxa, sya, b64, b, sxb, syb, W, H, eps, s, st));
  HIPCHK(hipMemcpyAsync(sc.cmp_host, s + compare_result_offset(n), compare_result_bytes(),
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  CompareResult r;
  compare_result(sc.cmp_host, &r);
  out->count = r.count;
  out->nonfinite = r.nonfinite;
  out->over_eps = r.over;
  out->max_abs = r.max_abs;
  out->sum_sq = r.sum_sq;
  out->peak_b = r.peak_b;
  const long long fin = r.count - r.nonfinite;
  out->mse = fin > 0 ? r.sum_sq / (double)fin : 0.0;
  const double pk = peak > 0 ? peak : r.peak_b;
  out->psnr = out->mse > 0 ? 10.0 * std::log10(pk * pk / out->mse) : INFINITY;
  for (int k = 0; k < DWT_GPU_ULP_BINS; ++k) out->ulp[k] = r.ulp[k];
  out->differ = (r.nonfinite > 0 || r.over > 0) ? 1 : 0;
}

int dwt_gpu_compare(int precision_a, const void* a, int stride_x_a, int stride_y_a, int precision_b,
                    const void* b, int stride_x_b, int stride_y_b, int size_x, int size_y, double eps,
                    double peak, dwt_gpu_compare_stats* stats, void* stream) {
  try {
    g_err.clear();
    init_once();
    for (int pr : {precision_a, precision_b})
      if (pr != DWT_GPU_F32 && pr != DWT_GPU_F64) fail(DWT_GPU_EINVAL, "bad precision");
    if (!stats) fail(DWT_GPU_EINVAL, "null stats");
    if (size_x < 0 || size_y < 0) fail(DWT_GPU_EINVAL, "negative size");
    const bool a64 = precision_a == DWT_GPU_F64, b64 = precision_b == DWT_GPU_F64;
    check_image(a, stride_x_a, stride_y_a, a64 ? 8 : 4, "image a");
    check_image(b, stride_x_b, stride_y_b, b64 ? 8 : 4, "image b");
    const void* da = device_address(a);
    const void* db = device_address(b);
    if (!da || !db) fail(DWT_GPU_EMEMORY, "compare needs device-accessible images");
    if (eps < 0) eps = (a64 && b64) ? 1e-6 : 1e-3;  // src/libdwt.c:1234, 1261
    compare_device(a64, da, stride_x_a, stride_y_a, b64, db, stride_x_b, stride_y_b, size_x, size_y,
                   eps, peak, stats, (hipStream_t)stream);
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_verify_2d(int wavelet, int direction, const void* input, const void* result, int stride_x,
                      int stride_y, int size_x, int size_y, int j, dwt_gpu_compare_stats* stats,
                      void* stream) {
  try {
    g_err.clear();
    init_once();
    if (wavelet != DWT_GPU_CDF97 && wavelet != DWT_GPU_CDF53) fail(DWT_GPU_EINVAL, "bad wavelet");
    if (direction != DWT_GPU_FORWARD && direction != DWT_GPU_INVERSE) fail(DWT_GPU_EINVAL, "bad direction");
    if (!stats) fail(DWT_GPU_EINVAL, "null stats");
    if (size_x < 0 || size_y < 0) fail(DWT_GPU_EINVAL, "negative size");
    check_image(input, stride_x, stride_y, 4, "input");
    check_image(result, stride_x, stride_y, 4, "result");
    const void* din = device_address(input);
    const void* dres = device_address(result);
    if (!din || !dres) fail(DWT_GPU_EMEMORY, "verification needs device-accessible images");
    hipStream_t st = (hipStream_t)stream;
    Scratch& sc = scratch_for(st);
    double* ref = (double*)sc.get(7, sizeof(double) * ((size_t)size_x * size_y + 1));
    LAUNCH(launch_to_f64(din, stride_x, stride_y, ref, size_x, size_y, st));
    int jj = j;
    run2d<double>(direction == DWT_GPU_FORWARD, ref, 8 * size_x, 8, size_x, size_y, size_x, size_y, &jj,
                  0, 0, wavelet == DWT_GPU_CDF97 ? CDF97 : CDF53, st, false);
    compare_device(false, dres, stride_x, stride_y, true, ref, 8 * size_x, 8, size_x, size_y, 1e-3, 0.0,
                   stats, st);
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_on = on != 0;
  g_prof_repeat = on > 1 ? on : 1;
  return DWT_GPU_OK;
}

int dwt_gpu_profile_read(int kind, int level, long long* launches, double* total_ms) {
  try {
    std::vector<ProfRec> recs;
    {
      std::lock_guard<std::mutex> lk(g_prof_mu);
      recs = *g_prof;
    }
    long long n = 0;
    double ms = 0;
    for (const auto& r : recs) {
      if (r.kind != kind || (level >= 0 && r.level != level)) continue;
      HIPCHK(hipEventSynchronize(r.b));
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
      ms += t;
      n += r.reps;
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_profile_reset(void) {
  std::vector<ProfRec> recs;
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    recs.swap(*g_prof);
  }
  for (auto& r : recs) {
    hipEventSynchronize(r.b);
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  return DWT_GPU_OK;
}

int dwt_gpu_tile_set(int forward, int height_class, int seg_rows, int prefetch_pairs) {
  if (tile_set(forward, height_class, seg_rows, prefetch_pairs) != 0) {
    g_err = "dwt_gpu_tile_set: bad class or shape";
    return DWT_GPU_EINVAL;
  }
  return DWT_GPU_OK;
}

int dwt_gpu_tile_get(int forward, int width, int height, int wavelet, int* seg_rows,
                     int* prefetch_pairs) {
  if (!seg_rows || !prefetch_pairs || width < 1 || height < 1) return DWT_GPU_EINVAL;
  return tile_get(forward, height, width, wavelet == DWT_GPU_CDF53 ? 2 : 4, seg_rows, prefetch_pairs);
}

int dwt_gpu_schedule_set(int mode) {
  if (mode != 0 && mode != 1) {
    g_err = "dwt_gpu_schedule_set: mode must be 0 or 1";
    return DWT_GPU_EINVAL;
  }
  g_schedule = mode;
  return DWT_GPU_OK;
}

int dwt_gpu_sparse_set(int fused) {
  if (fused != 0 && fused != 1) {
    g_err = "dwt_gpu_sparse_set: 0 or 1";
    return DWT_GPU_EINVAL;
  }
  g_sparse_fused = fused;
  return DWT_GPU_OK;
}

int dwt_gpu_pyr_grid_cap_set(int max_workgroups) {
  g_pyr_grid_cap = max_workgroups > 0 ? max_workgroups : 0;
  return DWT_GPU_OK;
}

int dwt_gpu_clock_rate_khz(void) {
  try {
    return (int)(spin_limit_ticks() / 2000ULL);
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_pyr_trace_enable(int max_items) {
  g_trace_items = max_items > 0 ? max_items : 0;
  return DWT_GPU_OK;
}

long long dwt_gpu_pyr_trace_read(void* stream, unsigned long long* out, long long max_items) {
  try {
    hipStream_t st = (hipStream_t)stream;
    Scratch& sc = scratch_for(st);
    HIPCHK(hipStreamSynchronize(st));
    if (!sc.trace || !out) return 0;
    const long long n = std::min(max_items, sc.trace_items);
    if (n > 0)
      HIPCHK(hipMemcpy(out, sc.trace, sizeof(unsigned long long) * 4 * (size_t)n, hipMemcpyDeviceToHost));
    return n;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_device_check(void* stream) {
  try {
    g_err.clear();
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(hipStreamSynchronize(st));
    if (pyr_failed(scratch_for(st), st)) fail(DWT_GPU_EHIP, "pyramid dependency wait timed out");
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_1d_batch(int wavelet, int direction, int precision, void* ptr, long long signal_stride,
                     long long sample_stride, int length, int count, int* levels, void* stream) {
  try {
    g_err.clear();
    if (wavelet != DWT_GPU_CDF97 && wavelet != DWT_GPU_CDF53) fail(DWT_GPU_EINVAL, "bad wavelet");
    if (direction != DWT_GPU_FORWARD && direction != DWT_GPU_INVERSE) fail(DWT_GPU_EINVAL, "bad direction");
    if (precision != DWT_GPU_F32 && precision != DWT_GPU_F64) fail(DWT_GPU_EINVAL, "bad precision");
    if (!levels) fail(DWT_GPU_EINVAL, "null levels");
    if (!ptr) fail(DWT_GPU_EINVAL, "null pointer");
    if (sample_stride == 0) fail(DWT_GPU_EINVAL, "zero sample stride");
    void* dptr = device_address(ptr);
    if (!dptr) fail(DWT_GPU_EMEMORY, "pointer is not device-accessible memory");
    const Wavelet w = wavelet == DWT_GPU_CDF97 ? CDF97 : CDF53;
    const bool fwd = direction == DWT_GPU_FORWARD;
    hipStream_t st = (hipStream_t)stream;
    if (precision == DWT_GPU_F32)
      run_lines<float>(fwd, (char*)dptr, signal_stride, sample_stride, length, count, levels, w, st);
    else
      run_lines<double>(fwd, (char*)dptr, signal_stride, sample_stride, length, count, levels, w, st);
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_col_plan(int forward, int n, int nsteps, int* stages, int* top_h_row, int* segments,
                     int* rows_per_segment) {
  if (!stages || !top_h_row || !segments || !rows_per_segment || n < 0 || (nsteps != 2 && nsteps != 4)) {
    g_err = "dwt_gpu_col_plan: bad argument";
    return DWT_GPU_EINVAL;
  }
  *segments = col_segments(n);
  *rows_per_segment = col_segment_rows();
  int s[kColStages + 1] = {};
  const bool staged = col_stages(forward != 0, n, nsteps, s);
  for (int k = 0; k <= kColStages; ++k) stages[k] = staged ? s[k] : 0;
  *top_h_row = staged ? (forward ? s[kColStages] * col_segment_rows() / 2
                                 : s[kColStages] * col_segment_rows() / 2 - nsteps / 2)
                      : 0;
  return staged ? 1 : 0;
}

int dwt_gpu_col_segment_rows(int forward, int n, int nsteps, int seg, int* reads, int max_reads, int* nreads,
                             int* writes, int max_writes, int* nwrites) {
  if (!reads || !writes || !nreads || !nwrites || n < 2 || seg < 0 || seg >= col_segments(n) ||
      (nsteps != 2 && nsteps != 4)) {
    g_err = "dwt_gpu_col_segment_rows: bad argument";
    return DWT_GPU_EINVAL;
  }
  std::vector<int> r, w;
  col_segment_access(forward != 0, n, nsteps, seg, r, w);
  if ((int)r.size() > max_reads || (int)w.size() > max_writes) {
    g_err = "dwt_gpu_col_segment_rows: output arrays too small";
    return DWT_GPU_EINVAL;
  }
  std::copy(r.begin(), r.end(), reads);
  std::copy(w.begin(), w.end(), writes);
  *nreads = (int)r.size();
  *nwrites = (int)w.size();
  return DWT_GPU_OK;
}

int dwt_gpu_level0_set(int inplace, int band) {
  if (inplace < 0 || inplace > 1 || band < 0 || band > 1 << 20) {
    g_err = "dwt_gpu_level0_set: bad argument";
    return DWT_GPU_EINVAL;
  }
  g_l0_inplace = inplace;
  if (band > 0) g_l0_band = band;
  return DWT_GPU_OK;
}

int dwt_gpu_level0_plan(int forward, int width, int height, int wavelet, int seg_rows, int* info, int* order,
                        int* segflags, int* list_off, int* list, int* dep_off, int* deps, int* rows) {
  if (!info || width < 2 || height < 2 || (wavelet != DWT_GPU_CDF97 && wavelet != DWT_GPU_CDF53) ||
      (seg_rows != 4 && seg_rows != 8)) {
    g_err = "dwt_gpu_level0_plan: bad argument";
    return DWT_GPU_EINVAL;
  }
  const L0Plan P = build_l0_plan(forward != 0, width, height, wavelet == DWT_GPU_CDF97 ? 4 : 2, seg_rows);
  info[0] = P.nseg;
  info[1] = P.ngroups;
  info[2] = P.nstrips;
  info[3] = P.nseg * P.ngroups;
  info[4] = (int)P.deps.size();
  info[5] = (int)P.rows.size();
  info[6] = P.grid;
  info[7] = P.band;
  auto out = [](const std::vector<int>& v, int* o) {
    if (o) std::copy(v.begin(), v.end(), o);
  };
  out(P.order, order);
  out(P.segflags, segflags);
  out(P.list_off, list_off);
  out(P.list, list);
  out(P.dep_off, dep_off);
  out(P.deps, deps);
  out(P.rows, rows);
  return DWT_GPU_OK;
}

int dwt_gpu_mid_set(long long max_level_bytes) {
  if (max_level_bytes < 0) {
    g_err = "dwt_gpu_mid_set: negative size";
    return DWT_GPU_EINVAL;
  }
  g_mid_bytes = max_level_bytes;
  return DWT_GPU_OK;
}

int dwt_gpu_tail_set(long long max_region_bytes) {
  tail_config(max_region_bytes, 0);
  return DWT_GPU_OK;
}

long long dwt_gpu_scratch_bytes(void* stream) {
  try {
    return scratch_for((hipStream_t)stream).bytes();
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_release_scratch(void* stream) {
  try {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    Scratch* s = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_scratch_mu);
      auto it = g_scratch->find(std::make_tuple(std::this_thread::get_id(), (hipStream_t)stream, dev));
      if (it == g_scratch->end()) return DWT_GPU_OK;
      s = it->second;
      g_scratch->erase(it);
    }
    const hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    s->release();
    delete s;
    if (e != hipSuccess) fail(DWT_GPU_EHIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

}  // extern "C"
