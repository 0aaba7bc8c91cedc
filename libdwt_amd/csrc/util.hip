// Host utilities of libdwt's interface (dwt_util_*), so that programs written
// against the reference (its examples: start, simple, simple-perf, subbands)
// compile and link unchanged against this library.  Each function restates
// the documented behaviour of the reference function it replaces (cited);
// images handled here are host memory, as in the reference, except in the
// performance test, which keeps its images in device memory.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <sys/resource.h>
#include <sys/time.h>
#include <sys/times.h>
#include <unistd.h>
#include <vector>

#include "../../include/libdwt.h"

namespace {

template <typename T>
T* at(void* ptr, int y, int x, int sx, int sy) {
  return reinterpret_cast<T*>(static_cast<char*>(ptr) + (long long)y * sx + (long long)x * sy);
}
template <typename T>
const T* at(const void* ptr, int y, int x, int sx, int sy) {
  return reinterpret_cast<const T*>(static_cast<const char*>(ptr) + (long long)y * sx + (long long)x * sy);
}

// Type-0 test pattern 2xy / (x^2 + y^2 + 1), x shifted right by `rand`; the
// numerator and denominator are integers converted for one division in T
// (src/libdwt.c:1076-1116).
template <typename T>
void fill(void* ptr, int sx, int sy, int w, int h, int rand) {
  for (int y = 0; y < h; ++y)
    for (int x0 = 0; x0 < w; ++x0) {
      const int x = x0 >> rand;
      *at<T>(ptr, y, x0, sx, sy) = (T)(2 * x * y) / (T)(x * x + y * y + 1);
    }
}

// 0 when every pixel pair is finite and within eps (src/libdwt.c:1223-1277).
template <typename T>
int compare(void* a, void* b, int sx, int sy, int w, int h, T eps) {
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const T p = *at<T>(a, y, x, sx, sy), q = *at<T>(b, y, x, sx, sy);
      if (!std::isfinite(p) || !std::isfinite(q) || std::fabs(p - q) > eps) return 1;
    }
  return 0;
}

// ASCII PGM (P2), 255 levels, value = (int)(255 px / max) clamped to
// [0, 255]; one warning per file for pixels outside [0, max] by more than
// 1e-3 (src/libdwt.c:7532-7599).
template <typename T>
int save_pgm(const char* name, T maxv, const void* ptr, int sx, int sy, int w, int h) {
  FILE* f = std::fopen(name, "w");
  if (!f) return 1;
  std::fprintf(f, "P2\n%i %i\n%i\n", w, h, 255);
  int err = 0;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const T px = *at<T>(ptr, y, x, sx, sy);
      int v = (int)(255 * px / maxv);
      if (px - (T)1e-3 > maxv && !err++)
        dwt_util_log(LOG_WARN, "%s: pixel intensity %f above %f (reported once)\n", __func__,
                     (double)px, (double)maxv);
      if (px > maxv) v = 255;
      if (px + (T)1e-3 < 0 && !err++)
        dwt_util_log(LOG_WARN, "%s: pixel intensity %f below 0 (reported once)\n", __func__, (double)px);
      if (px < 0) v = 0;
      if (std::fprintf(f, "%i\n", v) < 0) {
        std::fclose(f);
        return 1;
      }
    }
  std::fclose(f);
  if (err) dwt_util_log(LOG_WARN, "%s: %i pixels out of range\n", __func__, err);
  return 0;
}

// Viewable transform: log(1 + 100 |c|) / 10 (src/libdwt.c:8326-8384).
template <typename T>
void conv_show(const void* src, void* dst, int sx, int sy, int w, int h) {
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const T c = *at<T>(src, y, x, sx, sy);
      *at<T>(dst, y, x, sx, sy) = std::log((T)1 + std::fabs(c) * (T)100) / (T)10;
    }
}

// Fermat probable prime to base 2: 2^(N-1) == 1 (mod N) for odd N > 2, the
// test the reference's stride selection uses (src/libdwt.c:8069-8141).
bool probable_prime(int N) {
  if (N == 2) return true;
  if (N < 2 || !(N & 1)) return false;
  unsigned long long r = 1, b = 2 % (unsigned)N, e = (unsigned)(N - 1);
  while (e) {
    if (e & 1) r = r * b % (unsigned)N;
    b = b * b % (unsigned)N;
    e >>= 1;
  }
  return r == 1;
}
int next_prime(int N) {
  if (N <= 2) return 2;
  N |= 1;
  while (!probable_prime(N)) N += 2;
  return N;
}

template <typename T, typename F>
void perf(int stride_x, int stride_y, int size_o_big_x, int size_o_big_y, int size_i_big_x,
          int size_i_big_y, int j_max, int decompose_one, int zero_padding, int M, int N,
          int clock_type, F* fwd_secs, F* inv_secs) {
  if (M <= 0 || N <= 0 || !fwd_secs || !inv_secs) {
    dwt_util_log(LOG_ERR, "dwt_util_perf_cdf97_2: bad arguments\n");
    std::abort();
  }
  const size_t bytes = (size_t)stride_x * (size_t)size_o_big_y;
  std::vector<char> host(bytes, 0);
  fill<T>(host.data(), stride_x, stride_y, size_i_big_x, size_i_big_y, 0);
  std::vector<void*> img(M, nullptr);
  std::vector<int> j(M, j_max);
  for (int m = 0; m < M; ++m) {
    if (hipMalloc(&img[m], bytes) != hipSuccess ||
        hipMemcpy(img[m], host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
      dwt_util_log(LOG_ERR, "dwt_util_perf_cdf97_2: device memory\n");
      std::abort();
    }
  }
  *fwd_secs = *inv_secs = INFINITY;
  const double freq = (double)dwt_util_get_frequency(clock_type);
  for (int n = 0; n < N; ++n) {
    dwt_clock_t t0 = dwt_util_get_clock(clock_type);
    for (int m = 0; m < M; ++m) {
      if (sizeof(T) == 4)
        dwt_cdf97_2f_s(img[m], stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x,
                       size_i_big_y, &j[m], decompose_one, zero_padding);
      else
        dwt_cdf97_2f_d(img[m], stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x,
                       size_i_big_y, &j[m], decompose_one, zero_padding);
    }
    dwt_clock_t t1 = dwt_util_get_clock(clock_type);
    const F f = (F)((t1 - t0) / freq / M);
    if (f < *fwd_secs) *fwd_secs = f;
    t0 = dwt_util_get_clock(clock_type);
    for (int m = 0; m < M; ++m) {
      if (sizeof(T) == 4)
        dwt_cdf97_2i_s(img[m], stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x,
                       size_i_big_y, j[m], decompose_one, zero_padding);
      else
        dwt_cdf97_2i_d(img[m], stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x,
                       size_i_big_y, j[m], decompose_one, zero_padding);
    }
    t1 = dwt_util_get_clock(clock_type);
    const F iv = (F)((t1 - t0) / freq / M);
    if (iv < *inv_secs) *inv_secs = iv;
  }
  for (void* p : img) (void)hipFree(p);
}

// Frame of a test image for the performance sweeps (src/libdwt.c:8620-8700):
// the row pitch is taken from the power of two >= width (a prime >= it with
// opt_stride); DWT_ARR_SPARSE and DWT_ARR_SIMPLE put the image into a
// power-of-two outer frame, DWT_ARR_PACKED does not enlarge it.
template <typename T>
void get_sizes(enum dwt_array array_type, int size_x, int size_y, int opt_stride,
                      int* stride_x, int* stride_y, int* size_o_big_x, int* size_o_big_y,
                      int* size_i_big_x, int* size_i_big_y) {
  *stride_x = dwt_util_get_stride(dwt_util_pow2_ceil_log2(size_x) * (int)sizeof(T), opt_stride);
  *stride_y = sizeof(T);
  *size_o_big_x = *size_i_big_x = size_x;
  *size_o_big_y = *size_i_big_y = size_y;
  if (array_type == DWT_ARR_SPARSE || array_type == DWT_ARR_SIMPLE) {
    *size_o_big_x = dwt_util_pow2_ceil_log2(size_x);
    *size_o_big_y = dwt_util_pow2_ceil_log2(size_y);
  }
}
// Size sweeps (src/libdwt.c:8706-9036): x from min_x while x <= max_x, the
// next x = (int)(x * 1.28 + 1 - FLT_EPSILON) in single precision; an x-by-y
// image (y = x for the 2D sweep, 1 for the 1D sweep) per step; one line
// "<pixels>\t<seconds>" per size into each plot file (gnuplot data).
template <typename T, typename F>
void measure(bool two_d, enum dwt_array array_type, int min_x, int max_x, int opt_stride,
                    int j_max, int decompose_one, int zero_padding, int M, int N, int clock_type,
                    FILE* fwd_plot_data, FILE* inv_plot_data) {
  if (!(min_x > 0 && min_x < max_x) || M <= 0 || N <= 0 || !fwd_plot_data || !inv_plot_data) {
    dwt_util_log(LOG_ERR, "dwt_util_measure_perf_cdf97: bad arguments\n");
    std::abort();
  }
  const float growth = 1.28f;
  for (int x = min_x; x <= max_x; x = (int)((float)x * growth + 1.0f - FLT_EPSILON)) {
    const int y = two_d ? x : 1;
    int sx, sy, ox, oy, ix, iy;
    get_sizes<T>(array_type, x, y, opt_stride, &sx, &sy, &ox, &oy, &ix, &iy);
    dwt_util_log(LOG_DBG, "performance test for [%ix%i] in [%ix%i] with strides (%i, %i)...\n",
                 ix, iy, ox, oy, sx, sy);
    F fwd, inv;
    perf<T, F>(sx, sy, ox, oy, ix, iy, j_max, decompose_one, zero_padding, M, N, clock_type, &fwd,
               &inv);
    std::fprintf(fwd_plot_data, "%i\t%.10f\n", x * y, (double)fwd);
    std::fprintf(inv_plot_data, "%i\t%.10f\n", x * y, (double)inv);
  }
}
}  // namespace

extern "C" {

// src/libdwt.c:507-520
const char* dwt_util_version() { return "libdwt-mi355x 0.1"; }
const char* dwt_util_arch() { return "gfx950"; }

// src/libdwt.c:1159-1220
void dwt_util_test_image_fill_s(void* ptr, int stride_x, int stride_y, int size_i_big_x,
                                int size_i_big_y, int rand) {
  fill<float>(ptr, stride_x, stride_y, size_i_big_x, size_i_big_y, rand);
}
void dwt_util_test_image_fill_d(void* ptr, int stride_x, int stride_y, int size_i_big_x,
                                int size_i_big_y, int rand) {
  fill<double>(ptr, stride_x, stride_y, size_i_big_x, size_i_big_y, rand);
}
void dwt_util_test_image_zero_s(void* ptr, int stride_x, int stride_y, int size_i_big_x,
                                int size_i_big_y) {
  for (int y = 0; y < size_i_big_y; ++y)
    for (int x = 0; x < size_i_big_x; ++x) *at<float>(ptr, y, x, stride_x, stride_y) = 0.0f;
}

// Host memory of stride_x * size_o_big_y bytes (src/libdwt.c:1194-1220).
void dwt_util_alloc_image(void** pptr, int stride_x, int stride_y, int size_o_big_x,
                          int size_o_big_y) {
  (void)stride_y;
  (void)size_o_big_x;
  *pptr = std::malloc((size_t)stride_x * (size_t)size_o_big_y);
  if (!*pptr) {
    dwt_util_log(LOG_ERR, "Unable to allocate memory.\n");
    std::abort();
  }
}
void dwt_util_free_image(void** pptr) {
  std::free(*pptr);
  *pptr = nullptr;
}

int dwt_util_compare_s(void* ptr1, void* ptr2, int stride_x, int stride_y, int size_i_big_x,
                       int size_i_big_y) {
  return compare<float>(ptr1, ptr2, stride_x, stride_y, size_i_big_x, size_i_big_y, 1e-3f);
}
int dwt_util_compare_d(void* ptr1, void* ptr2, int stride_x, int stride_y, int size_i_big_x,
                       int size_i_big_y) {
  return compare<double>(ptr1, ptr2, stride_x, stride_y, size_i_big_x, size_i_big_y, 1e-6);
}

// Timers (src/libdwt.c:6834-7200): ticks of the selected source and their
// frequency; the automatic choice is clock_gettime.
int dwt_util_clock_autoselect() { return DWT_TIME_CLOCK_GETTIME; }

static clockid_t clock_of(int type) {
  switch (type) {
    case DWT_TIME_CLOCK_GETTIME_MONOTONIC: return CLOCK_MONOTONIC;
    case DWT_TIME_CLOCK_GETTIME_MONOTONIC_RAW: return CLOCK_MONOTONIC_RAW;
    case DWT_TIME_CLOCK_GETTIME_PROCESS_CPUTIME_ID: return CLOCK_PROCESS_CPUTIME_ID;
    case DWT_TIME_CLOCK_GETTIME_THREAD_CPUTIME_ID: return CLOCK_THREAD_CPUTIME_ID;
    default: return CLOCK_REALTIME;
  }
}

dwt_clock_t dwt_util_get_frequency(int type) {
  if (type == DWT_TIME_AUTOSELECT) type = dwt_util_clock_autoselect();
  switch (type) {
    case DWT_TIME_CLOCK: return (dwt_clock_t)CLOCKS_PER_SEC;
    case DWT_TIME_TIMES: return (dwt_clock_t)sysconf(_SC_CLK_TCK);
    case DWT_TIME_GETRUSAGE:
    case DWT_TIME_GETRUSAGE_SELF:
    case DWT_TIME_GETRUSAGE_CHILDREN:
    case DWT_TIME_GETRUSAGE_THREAD:
    case DWT_TIME_GETTIMEOFDAY: return 1000000;
    case DWT_TIME_IOCTL_RTC: return 1;
    default: return 1000000000;
  }
}

dwt_clock_t dwt_util_get_clock(int type) {
  if (type == DWT_TIME_AUTOSELECT) type = dwt_util_clock_autoselect();
  switch (type) {
    case DWT_TIME_CLOCK: return (dwt_clock_t)clock();
    case DWT_TIME_TIMES: {
      struct tms t;
      times(&t);
      return (dwt_clock_t)t.tms_utime;
    }
    case DWT_TIME_GETRUSAGE:
    case DWT_TIME_GETRUSAGE_SELF:
    case DWT_TIME_GETRUSAGE_CHILDREN:
    case DWT_TIME_GETRUSAGE_THREAD: {
      struct rusage u;
      getrusage(type == DWT_TIME_GETRUSAGE_CHILDREN ? RUSAGE_CHILDREN
                : type == DWT_TIME_GETRUSAGE_THREAD ? RUSAGE_THREAD : RUSAGE_SELF, &u);
      return (dwt_clock_t)u.ru_utime.tv_sec * 1000000 + u.ru_utime.tv_usec;
    }
    case DWT_TIME_GETTIMEOFDAY: {
      struct timeval tv;
      gettimeofday(&tv, nullptr);
      return (dwt_clock_t)tv.tv_sec * 1000000 + tv.tv_usec;
    }
    case DWT_TIME_IOCTL_RTC: return (dwt_clock_t)time(nullptr);
    default: {
      struct timespec ts;
      clock_gettime(clock_of(type), &ts);
      return (dwt_clock_t)ts.tv_sec * 1000000000 + ts.tv_nsec;
    }
  }
}

// One host thread drives the device; there are no ASVP workers
// (src/libdwt.c:7440-7465).
int dwt_util_get_num_threads() { return 1; }
int dwt_util_get_num_workers() { return 1; }
int dwt_util_get_max_threads() { return 1; }
int dwt_util_get_thread_num() { return 0; }

// Platform set-up: here the GPU (src/libdwt.c:7466-7530).
void dwt_util_init() {
  if (dwt_gpu_init() != DWT_GPU_OK)
    dwt_util_log(LOG_ERR, "dwt_util_init: %s\n", dwt_gpu_last_error());
}
void dwt_util_finish() { (void)hipDeviceSynchronize(); }

// The CPU lifting variants do not exist here (src/libdwt.c:7668).
void dwt_util_set_accel(int accel_type) { (void)accel_type; }

int dwt_util_save_to_pgm_s(const char* filename, float max_value, const void* ptr, int stride_x,
                           int stride_y, int size_i_big_x, int size_i_big_y) {
  return save_pgm<float>(filename, max_value, ptr, stride_x, stride_y, size_i_big_x, size_i_big_y);
}
int dwt_util_save_to_pgm_d(const char* filename, double max_value, const void* ptr, int stride_x,
                           int stride_y, int size_i_big_x, int size_i_big_y) {
  return save_pgm<double>(filename, max_value, ptr, stride_x, stride_y, size_i_big_x, size_i_big_y);
}

// printf to stderr with a level prefix (src/libdwt.c:7953-7986).
int dwt_util_log(enum dwt_util_loglevel level, const char* format, ...) {
  static const char* const prefix[] = {"", "DEBUG: ", "INFO: ", "WARNING: ", "ERROR: ", "TEST: "};
  const int l = (int)level >= 0 && (int)level <= (int)LOG_TEST ? (int)level : 0;
  flockfile(stderr);
  int n = std::fprintf(stderr, "%s", prefix[l]);
  va_list ap;
  va_start(ap, format);
  n += std::vfprintf(stderr, format, ap);
  va_end(ap);
  std::fflush(stderr);
  funlockfile(stderr);
  return n;
}

int dwt_util_is_aligned_8(const void* ptr) { return (reinterpret_cast<uintptr_t>(ptr) & 7) == 0; }

// src/libdwt.c:8014-8046
const char* dwt_util_node() {
  static char buf[256];
  if (gethostname(buf, sizeof buf - 1) != 0) return "unknown";
  buf[sizeof buf - 1] = 0;
  return buf;
}
const char* dwt_util_appname() {
  static char buf[4096];
  FILE* f = std::fopen("/proc/self/cmdline", "r");
  if (!f) return "unknown";
  const size_t n = std::fread(buf, 1, sizeof buf - 1, f);
  std::fclose(f);
  if (n == 0) return "unknown";
  buf[n] = 0;
  const char* slash = std::strrchr(buf, '/');
  return slash ? slash + 1 : buf;
}

// Strides: the smallest probable prime >= min_stride bytes avoids cache-set
// aliasing between rows (src/libdwt.c:8143-8151).
int dwt_util_is_prime(int N) { return probable_prime(N) ? 1 : 0; }
int dwt_util_next_prime(int N) { return next_prime(N); }
int dwt_util_get_opt_stride(int min_stride) { return next_prime(min_stride); }
int dwt_util_get_stride(int min_stride, int opt) { return opt ? next_prime(min_stride) : min_stride; }

// Address and size of subband `band` of level j_max in a transformed frame
// (src/libdwt.c:8153-8275).
void dwt_util_subband(void* ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                      int size_i_big_x, int size_i_big_y, int j_max, enum dwt_subbands band,
                      void** dst_ptr, int* dst_size_x, int* dst_size_y) {
  int hx = 0, hy = 0, lx = size_i_big_x, ly = size_i_big_y, ox = size_o_big_x, oy = size_o_big_y;
  for (int j = 1; j <= j_max; ++j) {
    hx = lx / 2;
    hy = ly / 2;
    lx = (lx + 1) / 2;
    ly = (ly + 1) / 2;
    ox = (ox + 1) / 2;
    oy = (oy + 1) / 2;
  }
  switch (band) {
    case DWT_LL: *dst_ptr = at<char>(ptr, 0, 0, stride_x, stride_y); *dst_size_x = lx; *dst_size_y = ly; break;
    case DWT_HL: *dst_ptr = at<char>(ptr, 0, ox, stride_x, stride_y); *dst_size_x = hx; *dst_size_y = ly; break;
    case DWT_LH: *dst_ptr = at<char>(ptr, oy, 0, stride_x, stride_y); *dst_size_x = lx; *dst_size_y = hy; break;
    case DWT_HH: *dst_ptr = at<char>(ptr, oy, ox, stride_x, stride_y); *dst_size_x = hx; *dst_size_y = hy; break;
  }
}
void dwt_util_subband_s(void* ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                        int size_i_big_x, int size_i_big_y, int j_max, enum dwt_subbands band,
                        void** dst_ptr, int* dst_size_x, int* dst_size_y) {
  dwt_util_subband(ptr, stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x, size_i_big_y,
                   j_max, band, dst_ptr, dst_size_x, dst_size_y);
}
void dwt_util_subband_d(void* ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                        int size_i_big_x, int size_i_big_y, int j_max, enum dwt_subbands band,
                        void** dst_ptr, int* dst_size_x, int* dst_size_y) {
  dwt_util_subband(ptr, stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x, size_i_big_y,
                   j_max, band, dst_ptr, dst_size_x, dst_size_y);
}

float* dwt_util_addr_coeff_s(void* ptr, int y, int x, int stride_x, int stride_y) {
  return at<float>(ptr, y, x, stride_x, stride_y);
}
double* dwt_util_addr_coeff_d(void* ptr, int y, int x, int stride_x, int stride_y) {
  return at<double>(ptr, y, x, stride_x, stride_y);
}

void dwt_util_conv_show_s(const void* src, void* dst, int stride_x, int stride_y, int size_i_big_x,
                          int size_i_big_y) {
  conv_show<float>(src, dst, stride_x, stride_y, size_i_big_x, size_i_big_y);
}
void dwt_util_conv_show_d(const void* src, void* dst, int stride_x, int stride_y, int size_i_big_x,
                          int size_i_big_y) {
  conv_show<double>(src, dst, stride_x, stride_y, size_i_big_x, size_i_big_y);
}

// Performance test (src/libdwt.c:8386-8483): M type-0 test images, N loops
// of M forward then M inverse transforms; the minimum over the loops of the
// mean time per image.  The images live in device memory (the transform
// under test is the device's; host staging would time PCIe).
void dwt_util_perf_cdf97_2_s(int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                             int size_i_big_x, int size_i_big_y, int j_max, int decompose_one,
                             int zero_padding, int M, int N, int clock_type, float* fwd_secs,
                             float* inv_secs) {
  perf<float, float>(stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x, size_i_big_y,
                     j_max, decompose_one, zero_padding, M, N, clock_type, fwd_secs, inv_secs);
}
void dwt_util_perf_cdf97_2_d(int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                             int size_i_big_x, int size_i_big_y, int j_max, int decompose_one,
                             int zero_padding, int M, int N, int clock_type, double* fwd_secs,
                             double* inv_secs) {
  perf<double, double>(stride_x, stride_y, size_o_big_x, size_o_big_y, size_i_big_x, size_i_big_y,
                       j_max, decompose_one, zero_padding, M, N, clock_type, fwd_secs, inv_secs);
}

// Smallest power of two >= x (src/libdwt.c:575-600).
int dwt_util_pow2_ceil_log2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

void dwt_util_get_sizes_s(enum dwt_array array_type, int size_x, int size_y, int opt_stride,
                          int* stride_x, int* stride_y, int* size_o_big_x, int* size_o_big_y,
                          int* size_i_big_x, int* size_i_big_y) {
  get_sizes<float>(array_type, size_x, size_y, opt_stride, stride_x, stride_y, size_o_big_x,
                   size_o_big_y, size_i_big_x, size_i_big_y);
}
void dwt_util_get_sizes_d(enum dwt_array array_type, int size_x, int size_y, int opt_stride,
                          int* stride_x, int* stride_y, int* size_o_big_x, int* size_o_big_y,
                          int* size_i_big_x, int* size_i_big_y) {
  get_sizes<double>(array_type, size_x, size_y, opt_stride, stride_x, stride_y, size_o_big_x,
                    size_o_big_y, size_i_big_x, size_i_big_y);
}

void dwt_util_measure_perf_cdf97_1_s(enum dwt_array array_type, int min_x, int max_x,
                                     int opt_stride, int j_max, int decompose_one,
                                     int zero_padding, int M, int N, int clock_type,
                                     FILE* fwd_plot_data, FILE* inv_plot_data) {
  measure<float, float>(false, array_type, min_x, max_x, opt_stride, j_max, decompose_one,
                        zero_padding, M, N, clock_type, fwd_plot_data, inv_plot_data);
}
void dwt_util_measure_perf_cdf97_1_d(enum dwt_array array_type, int min_x, int max_x,
                                     int opt_stride, int j_max, int decompose_one,
                                     int zero_padding, int M, int N, int clock_type,
                                     FILE* fwd_plot_data, FILE* inv_plot_data) {
  measure<double, double>(false, array_type, min_x, max_x, opt_stride, j_max, decompose_one,
                          zero_padding, M, N, clock_type, fwd_plot_data, inv_plot_data);
}
void dwt_util_measure_perf_cdf97_2_s(enum dwt_array array_type, int min_x, int max_x,
                                     int opt_stride, int j_max, int decompose_one,
                                     int zero_padding, int M, int N, int clock_type,
                                     FILE* fwd_plot_data, FILE* inv_plot_data) {
  measure<float, float>(true, array_type, min_x, max_x, opt_stride, j_max, decompose_one,
                        zero_padding, M, N, clock_type, fwd_plot_data, inv_plot_data);
}
void dwt_util_measure_perf_cdf97_2_d(enum dwt_array array_type, int min_x, int max_x,
                                     int opt_stride, int j_max, int decompose_one,
                                     int zero_padding, int M, int N, int clock_type,
                                     FILE* fwd_plot_data, FILE* inv_plot_data) {
  measure<double, double>(true, array_type, min_x, max_x, opt_stride, j_max, decompose_one,
                          zero_padding, M, N, clock_type, fwd_plot_data, inv_plot_data);
}

// OpenMP / ASVP worker counts do not apply (src/libdwt.c:7416-7438).
void dwt_util_set_num_threads(int num_threads) { (void)num_threads; }
void dwt_util_set_num_workers(int num_workers) { (void)num_workers; }

void dwt_util_abort() {
  std::fflush(stderr);
  std::abort();
}

// ceil(log2 x): the bit length of pow2ceil(x) - 1 (src/libdwt.c:618-627);
// x must be positive (the reference asserts it, src/libdwt.c:583).
int dwt_util_ceil_log2(int x) {
  if (x <= 0) {
    dwt_util_log(LOG_ERR, "dwt_util_ceil_log2: x must be positive\n");
    dwt_util_abort();
  }
  int r = 0;
  while ((1ll << r) < (long long)x) ++r;
  return r;
}

// Every timer type this library implements is available (src/libdwt.c:7259):
// the clock_gettime family, clock, times, getrusage, gettimeofday, and the
// RTC stand-in above.
int dwt_util_clock_available(int type) {
  switch (type) {
    case DWT_TIME_CLOCK_GETTIME:
    case DWT_TIME_CLOCK_GETTIME_REALTIME:
    case DWT_TIME_CLOCK_GETTIME_MONOTONIC:
    case DWT_TIME_CLOCK_GETTIME_MONOTONIC_RAW:
    case DWT_TIME_CLOCK_GETTIME_PROCESS_CPUTIME_ID:
    case DWT_TIME_CLOCK_GETTIME_THREAD_CPUTIME_ID:
    case DWT_TIME_CLOCK:
    case DWT_TIME_TIMES:
    case DWT_TIME_GETRUSAGE:
    case DWT_TIME_GETRUSAGE_SELF:
    case DWT_TIME_GETRUSAGE_CHILDREN:
    case DWT_TIME_GETRUSAGE_THREAD:
    case DWT_TIME_GETTIMEOFDAY:
    case DWT_TIME_IOCTL_RTC:
      return 0;
    default:
      return -1;
  }
}

// No ASVP workers here (src/libdwt.c:7407-7414).
int dwt_util_get_max_workers() { return 1; }

// Busy-wait on the automatically selected clock (src/libdwt.c:7375-7386).
void dwt_util_wait(int ms) {
  if (ms <= 0) {
    dwt_util_log(LOG_ERR, "dwt_util_wait: ms must be positive\n");
    dwt_util_abort();
  }
  const int type = dwt_util_clock_autoselect();
  const double freq = (double)dwt_util_get_frequency(type);
  const dwt_clock_t start = dwt_util_get_clock(type);
  while (1000.0 * (double)(dwt_util_get_clock(type) - start) / freq < (double)ms) {
  }
}

// Float vectors (src/libdwt.c:7689-7830).  A pair of values is "different"
// when either is not normal-or-zero or they differ by more than 1e-4; the
// difference is logged at LOG_ERR as the reference does.
static int vec_differ(float a, float b) {
  const auto ok = [](float v) { return v == 0.0f || std::isnormal(v); };
  if (!ok(a) || !ok(b)) {
    dwt_util_log(LOG_ERR, "%f or %f is not normal nor zero!\n", a, b);
    return 1;
  }
  if (std::fabs(a - b) > 1e-4f) {
    dwt_util_log(LOG_ERR, "%f should be %f!\n", a, b);
    return 1;
  }
  return 0;
}

float* dwt_util_allocate_vec_s(int size) {
  if (size < 0 || (size & 1)) {
    dwt_util_log(LOG_ERR, "dwt_util_allocate_vec_s: size must be even\n");
    dwt_util_abort();
  }
  void* p = nullptr;
  if (posix_memalign(&p, 8, sizeof(float) * (size_t)(size > 0 ? size : 2)) != 0) return nullptr;
  return static_cast<float*>(p);
}

int dwt_util_generate_vec_s(float* addr, int size) {
  for (int i = 0; i < size; ++i) addr[i] = (float)i;
  for (int i = 0; i < size; ++i)
    if (vec_differ(addr[i], (float)i)) return 1;
  return 0;
}

int dwt_util_zero_vec_s(float* addr, int size) {
  for (int i = 0; i < size; ++i) addr[i] = 0.0f;
  for (int i = 0; i < size; ++i)
    if (vec_differ(addr[i], 0.0f)) return 1;
  return 0;
}

int dwt_util_copy_vec_s(const float* src, float* dst, int size) {
  if (size > 0) std::memmove(dst, src, sizeof(float) * (size_t)size);
  for (int i = 0; i < size; ++i)
    if (vec_differ(dst[i], src[i])) return 1;
  return 0;
}

int dwt_util_cmp_vec_s(const float* a, const float* b, int size) {
  for (int i = size - 1; i >= 0; --i)
    if (vec_differ(a[i], b[i])) return 1;
  return 0;
}

// printf replacements (src/libdwt.c:7921-7951).
int dwt_util_vfprintf(FILE* stream, const char* format, va_list ap) { return std::vfprintf(stream, format, ap); }
int dwt_util_vprintf(const char* format, va_list ap) { return dwt_util_vfprintf(stdout, format, ap); }
int dwt_util_fprintf(FILE* stream, const char* format, ...) {
  va_list ap;
  va_start(ap, format);
  const int r = dwt_util_vfprintf(stream, format, ap);
  va_end(ap);
  return r;
}
int dwt_util_printf(const char* format, ...) {
  va_list ap;
  va_start(ap, format);
  const int r = dwt_util_vprintf(format, ap);
  va_end(ap);
  return r;
}

}  // extern "C"
