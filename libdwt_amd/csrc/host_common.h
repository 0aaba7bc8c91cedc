// Host-side helpers shared by the C-ABI translation units: error reporting,
// launch checking, one-time initialisation, per-launch profiling.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/libdwt.h"

namespace dwt {

// Text of the calling thread's last failure (dwt_gpu_last_error).
extern thread_local std::string g_err;

// Thrown by fail(); the entry points turn it into an error code (extensions)
// or a message and abort (drop-in functions).
struct Failure {
  int code;
};
[[noreturn]] void fail(int code, const char* fmt, ...);

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) fail(DWT_GPU_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// hipGetLastError() returns the last error of *any* earlier HIP call of this
// thread and keeps it until read, so a failure left behind by other code would
// be reported as ours.  LAUNCH() discards such a stale error first and then
// checks only what the enclosed launches returned.
void check_launch();
#define LAUNCH(...)            \
  do {                         \
    (void)hipGetLastError();   \
    __VA_ARGS__;               \
    check_launch();            \
  } while (0)

// Kernel attributes (LDS limits), once per process.
void init_once();

// Row pitch (elements) of internal buffers: a multiple of 64 elements, so that
// rows never share a cache line.
inline size_t pad_pitch(size_t w) { return (w + 63) & ~size_t(63); }

// Levels the 2D drivers can do on a Wo x Ho frame (src/libdwt.c:6324).
int level_limit(int Wo, int Ho, int d1);

// Optional per-launch timing with HIP events on the launch stream.  Kinds:
// 0 forward level kernel, 1 inverse level kernel, 2 detail copy, 4 forward
// block kernel, 5 inverse block kernel, 6 forward pyramid, 7 inverse pyramid,
// 8 / 9 forward / inverse batched level launch, 10 / 11 forward / inverse
// fused middle levels.  Level: the level index the launch belongs to (the
// first level it does for a block kernel, pyramid or fused launch).
struct ProfRec {
  int kind, level;
  hipEvent_t a, b;
  int reps;  // launches between the two events
};
struct ProfScope {
  bool on = false;
  ProfRec r{};
  hipStream_t st;
  ProfScope(int kind, int level, hipStream_t s, int reps = 1);
  ~ProfScope();
};
// Launches of the level-0 forward kernel per profiled scope (1 unless
// profiling was enabled with a repeat count).
int prof_repeat();

}  // namespace dwt
