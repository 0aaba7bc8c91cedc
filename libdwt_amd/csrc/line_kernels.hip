// Generic line kernels: one workgroup per line, any byte strides (odd,
// unaligned, negative), any frame.  These reproduce libdwt's per-line
// gather / lift / scatter (src/libdwt.c:5512-5603, 5779-5870, 1467-1571,
// 5673-5777) for whole passes of lines at once, and serve the layouts the
// packed fast path does not take: sparse frames (outer > inner), zero padding,
// strided element pitches, lines of length 1, and the 1D line entry points.
#include "dwt_common.h"
#include "kernels.h"

namespace dwt {

template <typename T>
__device__ inline T load_any(const char* p) {
  if ((reinterpret_cast<uintptr_t>(p) & (sizeof(T) - 1)) == 0) return *reinterpret_cast<const T*>(p);
  T v;
  __builtin_memcpy(&v, p, sizeof(T));
  return v;
}

template <typename T>
__device__ inline void store_any(char* p, T v) {
  if ((reinterpret_cast<uintptr_t>(p) & (sizeof(T) - 1)) == 0) {
    *reinterpret_cast<T*>(p) = v;
  } else {
    __builtin_memcpy(p, &v, sizeof(T));
  }
}

// Lifting of one line held in x[0..N) (LDS or global), N >= 2, by the whole
// workgroup.  A missing neighbour at an end is the existing one, used as
// x + (2c) * n.
template <typename T>
__device__ void lift_line(T* x, int N, const Coeffs<T>& cf) {
  for (int s = 0; s < cf.nsteps; ++s) {
    const int par = (cf.first_parity + s) & 1;
    const T c = cf.c[s];
    const T c2 = T(2) * c;
    for (int i = par + 2 * (int)threadIdx.x; i < N; i += 2 * (int)blockDim.x) {
      if (i > 0 && i + 1 < N) {
        T t = x[i - 1] + x[i + 1];
        t = c * t;
        x[i] = x[i] + t;
      } else {
        T t = c2 * (i > 0 ? x[i - 1] : x[i + 1]);
        x[i] = x[i] + t;
      }
    }
    __syncthreads();
  }
}

template <typename T>
__device__ void scale_line(T* x, int N, T se, T so) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) x[i] = x[i] * ((i & 1) ? so : se);
  __syncthreads();
}

// Forward: read src (N samples), write L (ceil(N/2)) to dl and H (floor(N/2))
// to dh.  Inverse: read L from a, H from b, write N samples to c.
// Line `l` of the pass starts at base + l * line_step for each of a, b, c;
// samples are es bytes apart.
//
// The line buffer is either the dynamic LDS (IN_LDS) or, for lines longer
// than the LDS limit, a per-block slice of the global buffer gtmp.  The two
// are separate instantiations on purpose: a pointer that may point to either
// is a generic (flat) pointer, and flat access to LDS breaks when the
// compiler folds x[i-1] into base (x - 4) plus an instruction offset of 4 --
// for i = 0 that base lies below the start of the LDS aperture, which the
// hardware rejects as an aperture violation.  With IN_LDS the buffer is only
// ever addressed by ds_* instructions.
template <typename T, bool FWD, bool IN_LDS>
__global__ void __launch_bounds__(256) line_pass_kernel(const char* a, const char* b, char* c,
                                                        long long line_step, long long es,
                                                        int nlines, int N, Coeffs<T> cf,
                                                        T* gtmp) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* x;
  if constexpr (IN_LDS) {
    x = reinterpret_cast<T*>(smem_raw);
  } else {
    x = gtmp + (size_t)blockIdx.x * (size_t)N;
  }
  for (int line = blockIdx.x; line < nlines; line += gridDim.x) {
    const long long lo = (long long)line * line_step;
    const char* pa = a + lo;
    const char* pb = b + lo;
    char* pc = c + lo;
    if (N == 1) {
      // src/libdwt.c:5523-5528, 5790-5795: a single sample is only scaled
      if (threadIdx.x == 0) {
        const T v = load_any<T>(pa) * cf.one;
        store_any<T>(FWD ? const_cast<char*>(pb) : pc, v);
      }
      continue;
    }
    const int nl = (N + 1) / 2;
    if (FWD) {
      for (int i = threadIdx.x; i < N; i += blockDim.x) x[i] = load_any<T>(pa + (long long)i * es);
    } else {
      for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const int k = i >> 1;
        x[i] = (i & 1) ? load_any<T>(pb + (long long)k * es) : load_any<T>(pa + (long long)k * es);
      }
    }
    __syncthreads();
    if (FWD) {
      lift_line(x, N, cf);
      scale_line(x, N, cf.se, cf.so);
      char* dl = const_cast<char*>(pb);
      for (int k = threadIdx.x; k < nl; k += blockDim.x) store_any<T>(dl + (long long)k * es, x[2 * k]);
      for (int k = threadIdx.x; k < N / 2; k += blockDim.x) store_any<T>(pc + (long long)k * es, x[2 * k + 1]);
    } else {
      scale_line(x, N, cf.se, cf.so);
      lift_line(x, N, cf);
      for (int i = threadIdx.x; i < N; i += blockDim.x) store_any<T>(pc + (long long)i * es, x[i]);
    }
    __syncthreads();
  }
}

// Zero the padding of one level (src/libdwt.c:6377-6397, 6729-6745) over the
// frame [0, fw) x [0, fh): an element is zeroed when its column lies in
// [zx0, zx1) or [zx2, zx3) or its row lies in [zy0, zy1) or [zy2, zy3).
template <typename T>
__global__ void zero_pad_kernel(char* ptr, long long sx, long long sy, int fw, int fh,
                                int zx0, int zx1, int zx2, int zx3, int zy0, int zy1, int zy2,
                                int zy3) {
  const long long n = (long long)fw * fh;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(t % fw), y = (int)(t / fw);
    const bool zx = (x >= zx0 && x < zx1) || (x >= zx2 && x < zx3);
    const bool zy = (y >= zy0 && y < zy1) || (y >= zy2 && y < zy3);
    if (zx || zy) store_any<T>(ptr + y * sx + x * sy, T(0));
  }
}

template <typename T>
void launch_line_pass(bool fwd, const char* a, const char* b, char* c, long long line_step,
                      long long es, int nlines, int N, const Coeffs<T>& cf, T* gtmp,
                      int gtmp_blocks, hipStream_t st) {
  if (nlines <= 0 || N <= 0) return;
  if (gtmp) {
    const int grid = nlines < gtmp_blocks ? nlines : gtmp_blocks;
    if (fwd)
      hipLaunchKernelGGL((line_pass_kernel<T, true, false>), dim3(grid), dim3(256), 0, st, a, b, c,
                         line_step, es, nlines, N, cf, gtmp);
    else
      hipLaunchKernelGGL((line_pass_kernel<T, false, false>), dim3(grid), dim3(256), 0, st, a, b,
                         c, line_step, es, nlines, N, cf, gtmp);
    return;
  }
  const size_t lds = (size_t)N * sizeof(T);
  const int grid = nlines < 65535 * 8 ? nlines : 65535 * 8;
  if (fwd)
    hipLaunchKernelGGL((line_pass_kernel<T, true, true>), dim3(grid), dim3(256), lds, st, a, b, c,
                       line_step, es, nlines, N, cf, gtmp);
  else
    hipLaunchKernelGGL((line_pass_kernel<T, false, true>), dim3(grid), dim3(256), lds, st, a, b, c,
                       line_step, es, nlines, N, cf, gtmp);
}

template <typename T>
void launch_zero_pad(char* ptr, long long sx, long long sy, int fw, int fh, const int zx[4],
                     const int zy[4], hipStream_t st) {
  const long long n = (long long)fw * fh;
  if (n <= 0) return;
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((zero_pad_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, ptr, sx, sy,
                     fw, fh, zx[0], zx[1], zx[2], zx[3], zy[0], zy[1], zy[2], zy[3]);
}

int line_lds_limit_bytes() { return 64 * 1024; }

hipError_t configure_line_kernels() {
  const int bytes = line_lds_limit_bytes();
  const void* fns[4] = {reinterpret_cast<const void*>(&line_pass_kernel<float, true, true>),
                        reinterpret_cast<const void*>(&line_pass_kernel<float, false, true>),
                        reinterpret_cast<const void*>(&line_pass_kernel<double, true, true>),
                        reinterpret_cast<const void*>(&line_pass_kernel<double, false, true>)};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return e;
    }
  }
  return hipSuccess;
}

template void launch_line_pass<float>(bool, const char*, const char*, char*, long long, long long,
                                      int, int, const Coeffs<float>&, float*, int, hipStream_t);
template void launch_line_pass<double>(bool, const char*, const char*, char*, long long, long long,
                                       int, int, const Coeffs<double>&, double*, int, hipStream_t);
template void launch_zero_pad<float>(char*, long long, long long, int, int, const int*, const int*,
                                     hipStream_t);
template void launch_zero_pad<double>(char*, long long, long long, int, int, const int*, const int*,
                                      hipStream_t);

}  // namespace dwt
