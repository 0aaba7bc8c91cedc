/*
 * ORACLE -- test infrastructure only.
 *
 * Clean-room CPU restatement of libdwt's lifting DWT (CDF 9/7 and CDF 5/3,
 * FP32 "_s" and FP64 "_d", 1D strided lines and 2D multilevel in-place
 * drivers).  It exists to check the HIP product path; it is never linked
 * into, called by, or shipped with the product library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every function cites the reference lines it restates (paths relative to the
 * reference tree).  The FP32 path is written to be bit-exact to the
 * reference's "_s" output: each lifting update is evaluated as
 *     t = l + r;  t = c * t;  x = x + t
 * with no contraction (build with -ffp-contract=off), a missing neighbour at a
 * line end is replaced by the existing one and evaluated as x + (2c) * n, and
 * the scaling constants are the FP32 roundings the reference uses.
 *
 * Pinning (see tests/test_oracle.py):
 *   - golden fixtures produced by the reference itself (tests/golden/),
 *   - direct comparison with oracle/_ref (the reference compiled from its
 *     own sources) when that library is present,
 *   - perfect reconstruction, linearity, polynomial annihilation, DC gain and
 *     a brute-force FP64 convolution with the CDF filters on tiny inputs.
 */
#include <stddef.h>
#include <string.h>
#include <stdlib.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------------
 * Constants.  src/libdwt.c:541-573 (decimal literals rounded to the storage
 * type, exactly as the reference declares them).
 * ------------------------------------------------------------------------ */
static const float  P1_s = 1.58613434342059;
static const float  U1_s = -0.0529801185729;
static const float  P2_s = -0.8829110755309;
static const float  U2_s = 0.4435068520439;
static const float  K_s  = 1.1496043988602;
static const float  S2_s = 1 / 1.1496043988602;  /* N==1 inverse only (5793) */

static const double P1_d = 1.58613434342059;
static const double U1_d = -0.0529801185729;
static const double P2_d = -0.8829110755309;
static const double U2_d = 0.4435068520439;
static const double K_d  = 1.1496043988602;
static const double S2_d = 1 / 1.1496043988602;

static const float  P53_s  = 0.5;
static const float  U53_s  = 0.25;
static const float  S153_s = 1.41421356237309504880;
static const float  S253_s = 0.70710678118654752440;
static const double P53_d  = 0.5;
static const double U53_d  = 0.25;
static const double S153_d = 1.41421356237309504880;
static const double S253_d = 0.70710678118654752440;

/* A lifting scheme: nsteps updates, alternating parity, then (forward) or
 * preceded by (inverse) a per-parity scaling. */
typedef struct { int nsteps; float c[4]; int first_parity; float se, so, one; } scheme_s;
typedef struct { int nsteps; double c[4]; int first_parity; double se, so, one; } scheme_d;

/* Forward 9/7 _s: src/libdwt.c:5542 passes (-P1, U1, -P2, U2, K, +1) with
 * off=1 (first update on odd samples); scaling after lifting multiplies even
 * samples by K and odd samples by 1/K evaluated in float (5094-5104, 1630-1640,
 * 5157-5163, 5185-5192).  N==1 multiplies by K (5526). */
static scheme_s s97f_s(void) {
	scheme_s s = {4, {-P1_s, U1_s, -P2_s, U2_s}, 1, K_s, 1 / K_s, K_s};
	return s;
}
/* Inverse 9/7 _s: src/libdwt.c:5811 passes (-U2, P2, -U1, P1, K, -1) with
 * off=0; scaling before lifting: even by 1/K (float), odd by K (1601-1611,
 * 5073-5079, 5175-5178).  N==1 multiplies by S2 = (float)(1/K) (5793), which
 * differs by one ulp from the float quotient 1/K. */
static scheme_s s97i_s(void) {
	scheme_s s = {4, {-U2_s, P2_s, -U1_s, P1_s}, 0, 1 / K_s, K_s, S2_s};
	return s;
}
/* 5/3 _s forward: src/libdwt.c:5576-5598. */
static scheme_s s53f_s(void) {
	scheme_s s = {2, {-P53_s, U53_s, 0, 0}, 1, S153_s, S253_s, S153_s};
	return s;
}
/* 5/3 _s inverse: src/libdwt.c:5843-5866. */
static scheme_s s53i_s(void) {
	scheme_s s = {2, {-U53_s, P53_s, 0, 0}, 0, S253_s, S153_s, S253_s};
	return s;
}
/* 9/7 _d forward/inverse: src/libdwt.c:1467-1525, 5673-5731. */
static scheme_d s97f_d(void) {
	scheme_d s = {4, {-P1_d, U1_d, -P2_d, U2_d}, 1, K_d, S2_d, K_d};
	return s;
}
static scheme_d s97i_d(void) {
	scheme_d s = {4, {-U2_d, P2_d, -U1_d, P1_d}, 0, S2_d, K_d, S2_d};
	return s;
}
/* 5/3 _d: src/libdwt.c:1527-1571, 5733-5777. */
static scheme_d s53f_d(void) {
	scheme_d s = {2, {-P53_d, U53_d, 0, 0}, 1, S153_d, S253_d, S153_d};
	return s;
}
static scheme_d s53i_d(void) {
	scheme_d s = {2, {-U53_d, P53_d, 0, 0}, 0, S253_d, S153_d, S253_d};
	return s;
}

/* ---------------------------------------------------------------------------
 * Lifting on a contiguous line of N >= 2 samples.
 * Each step updates every sample of one parity from its two neighbours of the
 * other parity; at a line end the single neighbour is used twice, written as
 * x + (2c)*n like src/libdwt.c:5063,5071,5137,5156 and 1490-1498.
 * ------------------------------------------------------------------------ */
#define DEFINE_LIFT(T, S, SUF)                                                 \
static void lift_steps_##SUF(T *x, int N, const S *sc)                         \
{                                                                              \
	for (int s = 0; s < sc->nsteps; ++s) {                                     \
		const int par = (sc->first_parity + s) & 1;                            \
		const T c = sc->c[s];                                                  \
		const T c2 = 2 * c;                                                    \
		for (int i = par; i < N; i += 2) {                                     \
			const int hl = i > 0, hr = i + 1 < N;                              \
			if (hl && hr) {                                                    \
				T t = x[i - 1] + x[i + 1];                                     \
				t = c * t;                                                     \
				x[i] = x[i] + t;                                               \
			} else {                                                           \
				T t = c2 * (hl ? x[i - 1] : x[i + 1]);                         \
				x[i] = x[i] + t;                                               \
			}                                                                  \
		}                                                                      \
	}                                                                          \
}                                                                              \
static void scale_##SUF(T *x, int N, T se, T so)                               \
{                                                                              \
	for (int i = 0; i < N; i += 2) x[i] = x[i] * se;                           \
	for (int i = 1; i < N; i += 2) x[i] = x[i] * so;                           \
}

DEFINE_LIFT(float, scheme_s, s)
DEFINE_LIFT(double, scheme_d, d)

/* Element access through a byte stride (src/libdwt.c:888-934). */
#define AT(T, p, i, stride) (*(T *)((char *)(p) + (ptrdiff_t)(i) * (ptrdiff_t)(stride)))

/* ---------------------------------------------------------------------------
 * One forward level on a strided line: gather, lift, scatter L (ceil(N/2))
 * and H (floor(N/2)).  src/libdwt.c:5512-5557 (9/7 _s), 5559-5603 (5/3 _s),
 * 1467-1571 (_d).  The full gather into tmp before any store makes every
 * aliasing of src / dst_l / dst_h safe.
 * ------------------------------------------------------------------------ */
#define DEFINE_LINE(T, S, SUF)                                                 \
static void line_f_##SUF(const T *src, T *dl, T *dh, T *tmp, int N,            \
                         int stride, const S *sc)                              \
{                                                                              \
	if (N < 2) {                                                               \
		if (N == 1) AT(T, dl, 0, stride) = AT(T, src, 0, stride) * sc->one;    \
		return;                                                                \
	}                                                                          \
	for (int i = 0; i < N; ++i) tmp[i] = AT(T, src, i, stride);                \
	lift_steps_##SUF(tmp, N, sc);                                              \
	scale_##SUF(tmp, N, sc->se, sc->so);                                       \
	for (int k = 0; k < (N + 1) / 2; ++k) AT(T, dl, k, stride) = tmp[2 * k];   \
	for (int k = 0; k < N / 2; ++k) AT(T, dh, k, stride) = tmp[2 * k + 1];     \
}                                                                              \
static void line_i_##SUF(const T *sl, const T *sh, T *dst, T *tmp, int N,      \
                         int stride, const S *sc)                              \
{                                                                              \
	if (N < 2) {                                                               \
		if (N == 1) AT(T, dst, 0, stride) = AT(T, sl, 0, stride) * sc->one;    \
		return;                                                                \
	}                                                                          \
	for (int k = 0; k < (N + 1) / 2; ++k) tmp[2 * k] = AT(T, sl, k, stride);   \
	for (int k = 0; k < N / 2; ++k) tmp[2 * k + 1] = AT(T, sh, k, stride);     \
	scale_##SUF(tmp, N, sc->se, sc->so);                                       \
	lift_steps_##SUF(tmp, N, sc);                                              \
	for (int i = 0; i < N; ++i) AT(T, dst, i, stride) = tmp[i];                \
}

DEFINE_LINE(float, scheme_s, s)
DEFINE_LINE(double, scheme_d, d)

/* Public 1D entry points, same signatures as src/libdwt.h:219-348. */
void orc_cdf97_f_ex_stride_s(const float *src, float *dl, float *dh, float *tmp, int N, int stride)
{ scheme_s s = s97f_s(); line_f_s(src, dl, dh, tmp, N, stride, &s); }
void orc_cdf53_f_ex_stride_s(const float *src, float *dl, float *dh, float *tmp, int N, int stride)
{ scheme_s s = s53f_s(); line_f_s(src, dl, dh, tmp, N, stride, &s); }
void orc_cdf97_i_ex_stride_s(const float *sl, const float *sh, float *dst, float *tmp, int N, int stride)
{ scheme_s s = s97i_s(); line_i_s(sl, sh, dst, tmp, N, stride, &s); }
void orc_cdf53_i_ex_stride_s(const float *sl, const float *sh, float *dst, float *tmp, int N, int stride)
{ scheme_s s = s53i_s(); line_i_s(sl, sh, dst, tmp, N, stride, &s); }
void orc_cdf97_f_ex_stride_d(const double *src, double *dl, double *dh, double *tmp, int N, int stride)
{ scheme_d s = s97f_d(); line_f_d(src, dl, dh, tmp, N, stride, &s); }
void orc_cdf53_f_ex_stride_d(const double *src, double *dl, double *dh, double *tmp, int N, int stride)
{ scheme_d s = s53f_d(); line_f_d(src, dl, dh, tmp, N, stride, &s); }
void orc_cdf97_i_ex_stride_d(const double *sl, const double *sh, double *dst, double *tmp, int N, int stride)
{ scheme_d s = s97i_d(); line_i_d(sl, sh, dst, tmp, N, stride, &s); }
void orc_cdf53_i_ex_stride_d(const double *sl, const double *sh, double *dst, double *tmp, int N, int stride)
{ scheme_d s = s53i_d(); line_i_d(sl, sh, dst, tmp, N, stride, &s); }

/* ---------------------------------------------------------------------------
 * Zero padding.  src/libdwt.c:5925-5942 (forward: L tail from ceil(N/2) to
 * N_dst_L, H tail from floor(N/2) to N_dst_H) and 5988-6004 (inverse: from N
 * to N_dst).
 * ------------------------------------------------------------------------ */
#define DEFINE_ZP(T, SUF)                                                      \
void orc_zero_padding_f_stride_##SUF(T *dl, T *dh, int N, int NL, int NH,       \
                                      int stride)                              \
{                                                                              \
	for (int k = (N + 1) / 2; k < NL; ++k) AT(T, dl, k, stride) = 0;           \
	for (int k = N / 2; k < NH; ++k) AT(T, dh, k, stride) = 0;                 \
}                                                                              \
void orc_zero_padding_i_stride_##SUF(T *dl, int N, int Nd, int stride)          \
{                                                                              \
	for (int k = N; k < Nd; ++k) AT(T, dl, k, stride) = 0;                     \
}
DEFINE_ZP(float, s)
DEFINE_ZP(double, d)

/* ---------------------------------------------------------------------------
 * Index helpers.  src/libdwt.c:580-672: ceil_log2(x) = number of set bits of
 * pow2ceil(x)-1; ceil_div_pow2(i, j) = (i + 2^j - 1) >> j.
 * ------------------------------------------------------------------------ */
int orc_ceil_log2(int x)
{
	int p = 1, r = 0;
	while (p < x) { p <<= 1; ++r; }
	return r;
}
static int cdp2(int i, int j) { return (i + (1 << j) - 1) >> j; }

/* ---------------------------------------------------------------------------
 * 2D multilevel drivers, in place.  Forward: src/libdwt.c:6286-6408 (9/7 _s),
 * 6410-6491 (5/3 _s), 6120-6284 (_d).  Inverse: 6647-6755, 6757-6832,
 * 6493-6645.  Per level: every row y < Hos (length Wis), then every column
 * x < Wos (length His), then optional zero padding.  The inverse runs the
 * levels from J down to 1 and also does rows before columns.
 * ------------------------------------------------------------------------ */
#define DEFINE_2D(T, S, SUF)                                                   \
static void fwd2_##SUF(void *ptr, int sx, int sy, int Wo, int Ho, int Wi,      \
                       int Hi, int *jp, int d1, int zp, const S *sc)           \
{                                                                              \
	const int mn = Wo < Ho ? Wo : Ho, mx = Wo < Ho ? Ho : Wo;                  \
	const int jl = orc_ceil_log2(d1 ? mx : mn);                                \
	if (*jp < 0 || *jp > jl) *jp = jl;                                         \
	for (int j = 0; j < *jp; ++j) {                                            \
		const int Wos = cdp2(Wo, j), Hos = cdp2(Ho, j);                        \
		const int Wod = cdp2(Wo, j + 1), Hod = cdp2(Ho, j + 1);                \
		const int Wis = cdp2(Wi, j), His = cdp2(Hi, j);                        \
		_Pragma("omp parallel")                                                \
		{                                                                      \
			T *tmp = (T *)malloc(sizeof(T) * (size_t)(mx + 8));                \
			_Pragma("omp for schedule(static)")                                \
			for (int y = 0; y < Hos; ++y) {                                    \
				char *row = (char *)ptr + (ptrdiff_t)y * sx;                   \
				line_f_##SUF((T *)row, (T *)row,                               \
				             (T *)(row + (ptrdiff_t)Wod * sy), tmp, Wis, sy,   \
				             sc);                                              \
			}                                                                  \
			_Pragma("omp for schedule(static)")                                \
			for (int x = 0; x < Wos; ++x) {                                    \
				char *col = (char *)ptr + (ptrdiff_t)x * sy;                   \
				line_f_##SUF((T *)col, (T *)col,                               \
				             (T *)(col + (ptrdiff_t)Hod * sx), tmp, His, sx,   \
				             sc);                                              \
			}                                                                  \
			if (zp) {                                                          \
				_Pragma("omp for schedule(static)")                            \
				for (int y = 0; y < Hos; ++y) {                                \
					char *row = (char *)ptr + (ptrdiff_t)y * sx;               \
					orc_zero_padding_f_stride_##SUF((T *)row,                  \
					    (T *)(row + (ptrdiff_t)Wod * sy), Wis, Wod, Wos - Wod, \
					    sy);                                                   \
				}                                                              \
				_Pragma("omp for schedule(static)")                            \
				for (int x = 0; x < Wos; ++x) {                                \
					char *col = (char *)ptr + (ptrdiff_t)x * sy;               \
					orc_zero_padding_f_stride_##SUF((T *)col,                  \
					    (T *)(col + (ptrdiff_t)Hod * sx), His, Hod, Hos - Hod, \
					    sx);                                                   \
				}                                                              \
			}                                                                  \
			free(tmp);                                                         \
		}                                                                      \
	}                                                                          \
}                                                                              \
static void inv2_##SUF(void *ptr, int sx, int sy, int Wo, int Ho, int Wi,      \
                       int Hi, int jmax, int d1, int zp, const S *sc)          \
{                                                                              \
	const int mn = Wo < Ho ? Wo : Ho, mx = Wo < Ho ? Ho : Wo;                  \
	int J = orc_ceil_log2(d1 ? mx : mn);                                       \
	if (jmax >= 0 && jmax < J) J = jmax;                                       \
	for (int j = J; j > 0; --j) {                                              \
		const int Wos = cdp2(Wo, j), Hos = cdp2(Ho, j);                        \
		const int Wod = cdp2(Wo, j - 1), Hod = cdp2(Ho, j - 1);                \
		const int Wid = cdp2(Wi, j - 1), Hid = cdp2(Hi, j - 1);                \
		_Pragma("omp parallel")                                                \
		{                                                                      \
			T *tmp = (T *)malloc(sizeof(T) * (size_t)(mx + 8));                \
			_Pragma("omp for schedule(static)")                                \
			for (int y = 0; y < Hod; ++y) {                                    \
				char *row = (char *)ptr + (ptrdiff_t)y * sx;                   \
				line_i_##SUF((T *)row, (T *)(row + (ptrdiff_t)Wos * sy),       \
				             (T *)row, tmp, Wid, sy, sc);                      \
			}                                                                  \
			_Pragma("omp for schedule(static)")                                \
			for (int x = 0; x < Wod; ++x) {                                    \
				char *col = (char *)ptr + (ptrdiff_t)x * sy;                   \
				line_i_##SUF((T *)col, (T *)(col + (ptrdiff_t)Hos * sx),       \
				             (T *)col, tmp, Hid, sx, sc);                      \
			}                                                                  \
			if (zp) {                                                          \
				_Pragma("omp for schedule(static)")                            \
				for (int y = 0; y < Hod; ++y)                                  \
					orc_zero_padding_i_stride_##SUF(                           \
					    (T *)((char *)ptr + (ptrdiff_t)y * sx), Wid, Wod, sy); \
				_Pragma("omp for schedule(static)")                            \
				for (int x = 0; x < Wod; ++x)                                  \
					orc_zero_padding_i_stride_##SUF(                           \
					    (T *)((char *)ptr + (ptrdiff_t)x * sy), Hid, Hod, sx); \
			}                                                                  \
			free(tmp);                                                         \
		}                                                                      \
	}                                                                          \
}

DEFINE_2D(float, scheme_s, s)
DEFINE_2D(double, scheme_d, d)

#define PUB2(NAME, SUF, SCH)                                                   \
void orc_##NAME##_2f_##SUF(void *p, int sx, int sy, int Wo, int Ho, int Wi,    \
                           int Hi, int *j, int d1, int zp)                     \
{ scheme_##SUF s = SCH##f_##SUF(); fwd2_##SUF(p, sx, sy, Wo, Ho, Wi, Hi, j, d1, zp, &s); } \
void orc_##NAME##_2i_##SUF(void *p, int sx, int sy, int Wo, int Ho, int Wi,    \
                           int Hi, int j, int d1, int zp)                      \
{ scheme_##SUF s = SCH##i_##SUF(); inv2_##SUF(p, sx, sy, Wo, Ho, Wi, Hi, j, d1, zp, &s); }

PUB2(cdf97, s, s97)
PUB2(cdf53, s, s53)
PUB2(cdf97, d, s97)
PUB2(cdf53, d, s53)

/* Thread-count control for timing runs. */
void orc_set_threads(int n)
{
#ifdef _OPENMP
	if (n > 0) omp_set_num_threads(n);
#else
	(void)n;
#endif
}
int orc_get_threads(void)
{
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

/* Bit patterns of the FP32 constants, for the constant-table test. */
void orc_constants_s(float out[10])
{
	out[0] = P1_s; out[1] = U1_s; out[2] = P2_s; out[3] = U2_s; out[4] = K_s;
	out[5] = 1 / K_s; out[6] = S2_s; out[7] = S153_s; out[8] = S253_s; out[9] = 1 / S153_s;
}
