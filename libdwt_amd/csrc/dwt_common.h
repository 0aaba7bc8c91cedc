// Shared device/host definitions for the MI355X (gfx950) lifting DWT.
//
// Arithmetic contract (bit-exact to libdwt's _s/_d output):
//   * every lifting update is  x = x + c * (l + r)  in the storage type,
//     evaluated as three separate IEEE operations (the library is built with
//     -ffp-contract=off, so no FMA is formed);
//   * a line end uses whole-sample symmetric extension, which reproduces the
//     reference's  x + (2c) * n  boundary form exactly (the two neighbours are
//     the same value, and 2c*n == c*(n+n) in IEEE arithmetic);
//   * forward scaling after lifting (even by se, odd by so), inverse scaling
//     before lifting; a line of length 1 is only multiplied by `one`.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dwt {

enum Wavelet { CDF97 = 0, CDF53 = 1 };

// Lifting constants of one direction of one wavelet in one precision.
// c[s] is applied at step s to samples of parity (first_parity + s) & 1.
template <typename T>
struct Coeffs {
  T c[4];
  T se, so;   // scale of even / odd samples
  T one;      // factor for a line of length 1
  int nsteps; // 4 (CDF 9/7) or 2 (CDF 5/3)
  int first_parity;  // 1 forward (odd samples first), 0 inverse
};

// Host-side constructors.  Values follow libdwt's constant table
// (src/libdwt.c:541-573) and the argument order of its lifting calls
// (src/libdwt.c:5542, 5811; 5576-5598, 5843-5866; 1467-1571, 5673-5777).
template <typename T>
Coeffs<T> make_coeffs(Wavelet w, bool forward);

// Whole-sample symmetric index: maps any integer onto [0, n) by reflecting
// about 0 and n-1 (period 2(n-1)).  Requires n >= 2.
__host__ __device__ inline int mirror(int i, int n) {
  const int p = 2 * (n - 1);
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - i;
}

// mirror() for an index that is usually inside [0, n): the in-range test is
// one compare, the modulo only runs near (or beyond) an end.
__host__ __device__ inline int mirror_near(int i, int n) {
  if (i >= 0 && i < n) return i;
  if (i < 0 && -i < n) return -i;
  if (i >= n && i <= 2 * n - 2) return 2 * n - 2 - i;
  return mirror(i, n);
}

__host__ __device__ inline int ceil_div_pow2(int i, int j) { return (i + (1 << j) - 1) >> j; }

inline int ceil_log2(int x) {
  int r = 0;
  while ((1 << r) < x) ++r;
  return r;
}

}  // namespace dwt
