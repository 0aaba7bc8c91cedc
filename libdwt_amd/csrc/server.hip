// Serving loop: continuous batching of transform requests.
//
// Clients submit requests (one image each, any size, wavelet, direction,
// levels; buffers already resident in device memory) at any time.  A
// scheduler thread owns S streams ("slots").  Whenever a slot is free and
// requests are waiting, it admits waiting requests in arrival order as long as
// their scratch fits the server's HBM budget (an arena allocated once) and
// the batch is below its size limit, and launches them as one batch on that
// slot: the level-synchronous batched kernels, one group per (wavelet,
// direction, storage).  Requests that arrive meanwhile wait for the next free
// slot, so batches on different slots overlap on the device and a request
// never waits for a whole queue to drain.  A slot's batch completes when its
// event has fired; the scheduler then frees its scratch and records each
// request's latency (submission to observed completion, steady clock).
//
// A batch computes each request exactly as alone: items never read one
// another's memory and the arithmetic does not depend on the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "batch_tables.h"
#include "host_common.h"

namespace dwt {

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, size) of a device arena (host bookkeeping).
class Arena {
 public:
  explicit Arena(size_t size) { free_[0] = size; }
  // offset, or (size_t)-1 when no free range fits
  size_t alloc(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        return off;
      }
    }
    return (size_t)-1;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {  // merge right
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {  // merge left
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<size_t, size_t> free_;
};

class Server {
 public:
  Server(long long budget, int max_batch, int slots) : budget_((size_t)budget), max_batch_(max_batch), arena_((size_t)budget) {
    HIPCHK(hipGetDevice(&device_));
    if (hipMalloc(&base_, budget_) != hipSuccess) {
      (void)hipGetLastError();
      base_ = nullptr;
      fail(DWT_GPU_EMEMORY, "server arena of %lld bytes", budget);
    }
    slots_.resize(slots);
    for (Slot& s : slots_) {
      HIPCHK(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
    }
    thread_ = std::thread([this] { loop(); });
  }
  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    thread_.join();
    for (Slot& s : slots_) {
      hipStreamSynchronize(s.st);
      hipStreamDestroy(s.st);
      hipEventDestroy(s.ev);
      if (s.pinned) hipHostFree(s.pinned);
      if (s.dev) hipFree(s.dev);
    }
    if (base_) hipFree(base_);
  }

  long long submit(const dwt_gpu_request& r) {
    batch_check_kind(r.wavelet, r.direction, r.storage);
    Req q;
    q.r = r;
    batch_check_image(0, q.r.image, r.direction == DWT_GPU_FORWARD, r.storage);
    q.scratch_bytes = image_scratch_bytes(q.r.image);
    if (q.scratch_bytes > budget_) fail(DWT_GPU_EMEMORY, "request needs %zu bytes of scratch, budget %zu", q.scratch_bytes, budget_);
    std::lock_guard<std::mutex> lk(mu_);
    q.id = next_id_++;
    q.t_submit = now_s();
    pending_.push_back(q);
    ++outstanding_;
    cv_.notify_all();
    return q.id;
  }

  int poll(long long* ids, double* latency, int* status, int max) {
    std::lock_guard<std::mutex> lk(mu_);
    int k = 0;
    while (k < max && !done_.empty()) {
      const Done& d = done_.front();
      if (ids) ids[k] = d.id;
      if (latency) latency[k] = d.latency;
      if (status) status[k] = d.status;
      done_.pop_front();
      ++k;
    }
    return k;
  }

  void drain() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [this] { return outstanding_ == 0; });
  }

  void stats(long long* batches, long long* requests) {
    std::lock_guard<std::mutex> lk(mu_);
    if (batches) *batches = batches_;
    if (requests) *requests = completed_;
  }

 private:
  struct Req {
    long long id = 0;
    dwt_gpu_request r{};
    double t_submit = 0;
    size_t scratch_off = 0, scratch_bytes = 0;
  };
  struct Done {
    long long id;
    double latency;
    int status;
  };
  struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<Req> reqs;
    void* pinned = nullptr;
    void* dev = nullptr;
    size_t cap = 0;
  };

  // Complete the slot's batch (caller holds mu_).
  void finish(Slot& s, int status) {
    const double t = now_s();
    for (const Req& q : s.reqs) {
      arena_.release(q.scratch_off, q.scratch_bytes);
      done_.push_back(Done{q.id, t - q.t_submit, status});
      --outstanding_;
      ++completed_;
    }
    s.reqs.clear();
    s.busy = false;
    cv_.notify_all();
  }

  // Admit waiting requests into slot s (caller holds mu_); false when none fits.
  bool admit(Slot& s) {
    while (!pending_.empty() && (int)s.reqs.size() < max_batch_) {
      Req& q = pending_.front();
      const size_t off = arena_.alloc(q.scratch_bytes);
      if (off == (size_t)-1) break;  // admission: the budget is full
      q.scratch_off = off;
      s.reqs.push_back(q);
      pending_.pop_front();
    }
    s.busy = !s.reqs.empty();
    return s.busy;
  }

  // Build the tables of slot s's admitted requests and launch them (called
  // without mu_: only this thread touches a busy slot); returns a status.
  int launch(Slot& s) try {
    // one table set per (wavelet, direction, storage) group
    std::vector<BatchTables> groups;
    std::vector<size_t> offs;
    size_t total = 0;
    for (int w = 0; w < 2; ++w)
      for (int d = 0; d < 2; ++d)
        for (int sk = 0; sk < 3; ++sk) {
          std::vector<dwt_gpu_image> imgs;
          std::vector<char*> scr;
          for (const Req& q : s.reqs)
            if (q.r.wavelet == w && q.r.direction == d && q.r.storage == sk) {
              imgs.push_back(q.r.image);
              scr.push_back(static_cast<char*>(base_) + q.scratch_off);
            }
          if (imgs.empty()) continue;
          groups.emplace_back();
          groups.back().build(w, d == DWT_GPU_FORWARD, sk, imgs.data(), (int)imgs.size(), scr.data());
          offs.push_back(total);
          total += groups.back().bytes;
        }
    {
      if (s.cap < total) {
        if (s.pinned) HIPCHK(hipHostFree(s.pinned));
        if (s.dev) HIPCHK(hipFree(s.dev));
        s.pinned = s.dev = nullptr;
        s.cap = 0;
        const size_t cap = std::max(total + total / 2, size_t(1) << 20);
        HIPCHK(hipHostMalloc(&s.pinned, cap, hipHostMallocDefault));
        HIPCHK(hipMalloc(&s.dev, cap));
        s.cap = cap;
      }
      for (size_t g = 0; g < groups.size(); ++g) groups[g].pack(static_cast<char*>(s.pinned) + offs[g]);
      HIPCHK(hipMemcpyAsync(s.dev, s.pinned, total, hipMemcpyHostToDevice, s.st));
      for (size_t g = 0; g < groups.size(); ++g) groups[g].launch(static_cast<char*>(s.dev) + offs[g], s.st);
      HIPCHK(hipEventRecord(s.ev, s.st));
    }
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  } catch (const std::exception&) {
    return DWT_GPU_EMEMORY;
  }

  void loop() {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(mu_);
    while (true) {
      bool inflight = false;
      for (Slot& s : slots_) {
        if (!s.busy) continue;
        const hipError_t e = hipEventQuery(s.ev);
        if (e == hipSuccess) finish(s, DWT_GPU_OK);
        else if (e != hipErrorNotReady) finish(s, DWT_GPU_EHIP);
        else inflight = true;
      }
      for (Slot& s : slots_)
        if (!s.busy && !pending_.empty() && admit(s)) {
          lk.unlock();  // submitters are not held up while the tables are built
          const int rc = launch(s);
          lk.lock();
          if (rc != DWT_GPU_OK) {
            finish(s, rc);
          } else {
            ++batches_;
            inflight = true;
          }
        }
      if (stop_ && !inflight && pending_.empty()) break;
      if (inflight) {
        // poll the events again shortly; new submissions wake us earlier
        cv_.wait_for(lk, std::chrono::microseconds(20));
      } else {
        cv_.wait(lk, [this] { return stop_ || !pending_.empty(); });
      }
    }
  }

  size_t budget_;
  int max_batch_;
  int device_ = 0;
  void* base_ = nullptr;
  Arena arena_;
  std::vector<Slot> slots_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Req> pending_;
  std::deque<Done> done_;
  long long next_id_ = 1, outstanding_ = 0, batches_ = 0, completed_ = 0;
  bool stop_ = false;
  std::thread thread_;
};

}  // namespace dwt

using namespace dwt;

extern "C" {

int dwt_gpu_server_create(long long hbm_budget_bytes, int max_batch, int streams, void** server) {
  try {
    g_err.clear();
    if (!server) fail(DWT_GPU_EINVAL, "null server pointer");
    if (hbm_budget_bytes <= 0 || max_batch < 1 || streams < 1 || streams > 16)
      fail(DWT_GPU_EINVAL, "budget > 0, max_batch >= 1, 1 <= streams <= 16");
    init_once();
    *server = new Server(hbm_budget_bytes, max_batch, streams);
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_server_submit(void* server, const dwt_gpu_request* req, long long* id) {
  try {
    g_err.clear();
    if (!server || !req) fail(DWT_GPU_EINVAL, "null server or request");
    const long long k = static_cast<Server*>(server)->submit(*req);
    if (id) *id = k;
    return DWT_GPU_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_server_replay(void* server, const dwt_gpu_request* reqs, const double* at_s, int n,
                          long long* ids) {
  try {
    g_err.clear();
    if (!server || (n > 0 && (!reqs || !at_s))) fail(DWT_GPU_EINVAL, "null server, requests or times");
    Server* sv = static_cast<Server*>(server);
    using clk = std::chrono::steady_clock;
    const clk::time_point t0 = clk::now();
    for (int i = 0; i < n; ++i) {
      const clk::time_point due =
          t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(at_s[i]));
      // sleep through long gaps, spin through the last 50 us
      const clk::time_point wake = due - std::chrono::microseconds(50);
      if (clk::now() < wake) std::this_thread::sleep_until(wake);
      while (clk::now() < due) {
      }
      const long long k = sv->submit(reqs[i]);
      if (ids) ids[i] = k;
    }
    return n;
  } catch (const Failure& f) {
    return f.code;
  }
}

int dwt_gpu_server_poll(void* server, long long* ids, double* latency_s, int* status, int max) {
  if (!server || max < 0) return DWT_GPU_EINVAL;
  return static_cast<Server*>(server)->poll(ids, latency_s, status, max);
}

int dwt_gpu_server_drain(void* server) {
  if (!server) return DWT_GPU_EINVAL;
  static_cast<Server*>(server)->drain();
  return DWT_GPU_OK;
}

int dwt_gpu_server_stats(void* server, long long* batches, long long* requests) {
  if (!server) return DWT_GPU_EINVAL;
  static_cast<Server*>(server)->stats(batches, requests);
  return DWT_GPU_OK;
}

int dwt_gpu_server_destroy(void* server) {
  delete static_cast<Server*>(server);
  return DWT_GPU_OK;
}

}  // extern "C"
