/*
 * libdwt drop-in C interface, served by hand-written HIP kernels on an
 * AMD Instinct MI355X (gfx950).
 *
 * Every function of the first part keeps the signature and semantics of the
 * reference libdwt function it replaces (the citation after each group is the
 * reference declaration in src/libdwt.h):
 *   - strides are in bytes and may be odd, unaligned or negative;
 *     stride_x is the distance between rows, stride_y between columns;
 *   - the 2D drivers work in place on the outer frame size_o x size_o, the
 *     nested image size_i x size_i is what gets transformed; *j_max_ptr is
 *     clamped to the achievable depth and written back (forward);
 *   - calls are synchronous: the result is in memory when the call returns;
 *   - results are bit-identical to the reference (FP32 "_s" and FP64 "_d").
 * Pointers may be device memory (hipMalloc), managed memory, pinned host
 * memory (hipHostMalloc) or ordinary host memory; the last is staged through
 * device memory.  The `tmp` argument of the 1D functions is accepted and not
 * used (device scratch is managed internally).
 * On a HIP error or invalid argument a message is printed to stderr and the
 * process aborts, as the reference does on its assertion and
 * dwt_util_abort() paths; the dwt_gpu_* extensions return an error code
 * instead.
 */
#ifndef LIBDWT_H
#define LIBDWT_H

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- one level, contiguous (deprecated forms: L and H concatenated) ------
 * reference: src/libdwt.h:29-120 (dwt_cdf97_f_d .. dwt_cdf53_i_s) */
void dwt_cdf97_f_d(const double *src, double *dst, double *tmp, int N);
void dwt_cdf53_f_d(const double *src, double *dst, double *tmp, int N);
void dwt_cdf97_f_s(const float *src, float *dst, float *tmp, int N);
void dwt_cdf53_f_s(const float *src, float *dst, float *tmp, int N);
void dwt_cdf97_i_d(const double *src, double *dst, double *tmp, int N);
void dwt_cdf53_i_d(const double *src, double *dst, double *tmp, int N);
void dwt_cdf97_i_s(const float *src, float *dst, float *tmp, int N);
void dwt_cdf53_i_s(const float *src, float *dst, float *tmp, int N);

/* ---- one level, separate L/H channels -------------------------------------
 * reference: src/libdwt.h:134-198 (forward _ex), 250-312 (inverse _ex) */
void dwt_cdf97_f_ex_d(const double *src, double *dst_l, double *dst_h, double *tmp, int N);
void dwt_cdf53_f_ex_d(const double *src, double *dst_l, double *dst_h, double *tmp, int N);
void dwt_cdf97_f_ex_s(const float *src, float *dst_l, float *dst_h, float *tmp, int N);
void dwt_cdf53_f_ex_s(const float *src, float *dst_l, float *dst_h, float *tmp, int N);
void dwt_cdf97_i_ex_d(const double *src_l, const double *src_h, double *dst, double *tmp, int N);
void dwt_cdf53_i_ex_d(const double *src_l, const double *src_h, double *dst, double *tmp, int N);
void dwt_cdf97_i_ex_s(const float *src_l, const float *src_h, float *dst, float *tmp, int N);
void dwt_cdf53_i_ex_s(const float *src_l, const float *src_h, float *dst, float *tmp, int N);

/* ---- one level on a strided line ------------------------------------------
 * reference: src/libdwt.h:205-240 (forward), 319-348 (inverse) */
void dwt_cdf97_f_ex_stride_d(const double *src, double *dst_l, double *dst_h, double *tmp, int N, int stride);
void dwt_cdf53_f_ex_stride_d(const double *src, double *dst_l, double *dst_h, double *tmp, int N, int stride);
void dwt_cdf97_f_ex_stride_s(const float *src, float *dst_l, float *dst_h, float *tmp, int N, int stride);
void dwt_cdf53_f_ex_stride_s(const float *src, float *dst_l, float *dst_h, float *tmp, int N, int stride);
void dwt_cdf97_i_ex_stride_d(const double *src_l, const double *src_h, double *dst, double *tmp, int N, int stride);
void dwt_cdf53_i_ex_stride_d(const double *src_l, const double *src_h, double *dst, double *tmp, int N, int stride);
void dwt_cdf97_i_ex_stride_s(const float *src_l, const float *src_h, float *dst, float *tmp, int N, int stride);
void dwt_cdf53_i_ex_stride_s(const float *src_l, const float *src_h, float *dst, float *tmp, int N, int stride);

/* ---- zero padding ----------------------------------------------------------
 * reference: src/libdwt.h:355-448 */
void dwt_zero_padding_f_d(double *dst_l, double *dst_h, int N, int N_dst_L, int N_dst_H);
void dwt_zero_padding_f_s(float *dst_l, float *dst_h, int N, int N_dst_L, int N_dst_H);
void dwt_zero_padding_f_stride_d(double *dst_l, double *dst_h, int N, int N_dst_L, int N_dst_H, int stride);
void dwt_zero_padding_f_stride_s(float *dst_l, float *dst_h, int N, int N_dst_L, int N_dst_H, int stride);
void dwt_zero_padding_i_d(double *dst_l, int N, int N_dst);
void dwt_zero_padding_i_s(float *dst_l, int N, int N_dst);
void dwt_zero_padding_i_stride_d(double *dst_l, int N, int N_dst, int stride);
void dwt_zero_padding_i_stride_s(float *dst_l, int N, int N_dst, int stride);

/* ---- 2D multilevel, in place ----------------------------------------------
 * reference: src/libdwt.h:455-520 (forward), 527-592 (inverse) */
void dwt_cdf97_2f_d(void *ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                    int size_i_big_x, int size_i_big_y, int *j_max_ptr, int decompose_one,
                    int zero_padding);
void dwt_cdf53_2f_d(void *ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                    int size_i_big_x, int size_i_big_y, int *j_max_ptr, int decompose_one,
                    int zero_padding);
void dwt_cdf97_2f_s(void *ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                    int size_i_big_x, int size_i_big_y, int *j_max_ptr, int decompose_one,
                    int zero_padding);
void dwt_cdf53_2f_s(void *ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                    int size_i_big_x, int size_i_big_y, int *j_max_ptr, int decompose_one,
                    int zero_padding);
void dwt_cdf97_2i_d(void *ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                    int size_i_big_x, int size_i_big_y, int j_max, int decompose_one,
                    int zero_padding);
void dwt_cdf53_2i_d(void *ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                    int size_i_big_x, int size_i_big_y, int j_max, int decompose_one,
                    int zero_padding);
void dwt_cdf97_2i_s(void *ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                    int size_i_big_x, int size_i_big_y, int j_max, int decompose_one,
                    int zero_padding);
void dwt_cdf53_2i_s(void *ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                    int size_i_big_x, int size_i_big_y, int j_max, int decompose_one,
                    int zero_padding);

/* ---- utilities used by programs written against libdwt --------------------
 * Host-side helpers with the reference's names and behaviour (image
 * allocation in host memory, test pattern, comparison, PGM output, timers,
 * logging, strides, subband addressing), so that its example programs build
 * unchanged against this library.  reference: src/libdwt.h:600-1270 */
enum dwt_timer_types {
	DWT_TIME_CLOCK_GETTIME,
	DWT_TIME_CLOCK_GETTIME_REALTIME,
	DWT_TIME_CLOCK_GETTIME_MONOTONIC,
	DWT_TIME_CLOCK_GETTIME_MONOTONIC_RAW,
	DWT_TIME_CLOCK_GETTIME_PROCESS_CPUTIME_ID,
	DWT_TIME_CLOCK_GETTIME_THREAD_CPUTIME_ID,
	DWT_TIME_CLOCK,
	DWT_TIME_TIMES,
	DWT_TIME_GETRUSAGE,
	DWT_TIME_GETRUSAGE_SELF,
	DWT_TIME_GETRUSAGE_CHILDREN,
	DWT_TIME_GETRUSAGE_THREAD,
	DWT_TIME_GETTIMEOFDAY,
	DWT_TIME_IOCTL_RTC,
	DWT_TIME_AUTOSELECT
};
typedef int64_t dwt_clock_t;
enum dwt_util_loglevel { LOG_NONE = 0, LOG_DBG, LOG_INFO, LOG_WARN, LOG_ERR, LOG_TEST };
enum dwt_subbands { DWT_LL, DWT_HL, DWT_LH, DWT_HH };

void dwt_util_test_image_fill_s(void *ptr, int stride_x, int stride_y, int size_i_big_x,
                                int size_i_big_y, int rand);
void dwt_util_test_image_fill_d(void *ptr, int stride_x, int stride_y, int size_i_big_x,
                                int size_i_big_y, int rand);
void dwt_util_test_image_zero_s(void *ptr, int stride_x, int stride_y, int size_i_big_x,
                                int size_i_big_y);
void dwt_util_alloc_image(void **pptr, int stride_x, int stride_y, int size_o_big_x,
                          int size_o_big_y);
void dwt_util_free_image(void **pptr);
int dwt_util_compare_s(void *ptr1, void *ptr2, int stride_x, int stride_y, int size_i_big_x,
                       int size_i_big_y);
int dwt_util_compare_d(void *ptr1, void *ptr2, int stride_x, int stride_y, int size_i_big_x,
                       int size_i_big_y);
const char *dwt_util_version(void);
const char *dwt_util_arch(void);
int dwt_util_clock_autoselect(void);
dwt_clock_t dwt_util_get_frequency(int type);
dwt_clock_t dwt_util_get_clock(int type);
int dwt_util_get_num_threads(void);
int dwt_util_get_num_workers(void);
int dwt_util_get_max_threads(void);
int dwt_util_get_thread_num(void);
void dwt_util_init(void);
void dwt_util_finish(void);
void dwt_util_set_accel(int accel_type);
int dwt_util_save_to_pgm_s(const char *filename, float max_value, const void *ptr, int stride_x,
                           int stride_y, int size_i_big_x, int size_i_big_y);
int dwt_util_save_to_pgm_d(const char *filename, double max_value, const void *ptr, int stride_x,
                           int stride_y, int size_i_big_x, int size_i_big_y);
int dwt_util_log(enum dwt_util_loglevel level, const char *format, ...);
int dwt_util_is_aligned_8(const void *ptr);
const char *dwt_util_node(void);
const char *dwt_util_appname(void);
int dwt_util_is_prime(int N);
int dwt_util_next_prime(int N);
int dwt_util_get_opt_stride(int min_stride);
int dwt_util_get_stride(int min_stride, int opt);
void dwt_util_subband(void *ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                      int size_i_big_x, int size_i_big_y, int j_max, enum dwt_subbands band,
                      void **dst_ptr, int *dst_size_x, int *dst_size_y);
void dwt_util_subband_s(void *ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                        int size_i_big_x, int size_i_big_y, int j_max, enum dwt_subbands band,
                        void **dst_ptr, int *dst_size_x, int *dst_size_y);
void dwt_util_subband_d(void *ptr, int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                        int size_i_big_x, int size_i_big_y, int j_max, enum dwt_subbands band,
                        void **dst_ptr, int *dst_size_x, int *dst_size_y);
float *dwt_util_addr_coeff_s(void *ptr, int y, int x, int stride_x, int stride_y);
double *dwt_util_addr_coeff_d(void *ptr, int y, int x, int stride_x, int stride_y);
void dwt_util_conv_show_s(const void *src, void *dst, int stride_x, int stride_y,
                          int size_i_big_x, int size_i_big_y);
void dwt_util_conv_show_d(const void *src, void *dst, int stride_x, int stride_y,
                          int size_i_big_x, int size_i_big_y);
void dwt_util_set_num_threads(int num_threads);
void dwt_util_set_num_workers(int num_workers);
void dwt_util_abort(void);
/* Performance harness (reference: src/libdwt.h:1250-1412); the M images of a
 * performance test live in device memory here. */
enum dwt_array { DWT_ARR_SIMPLE, DWT_ARR_SPARSE, DWT_ARR_PACKED };
void dwt_util_perf_cdf97_2_s(int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                             int size_i_big_x, int size_i_big_y, int j_max, int decompose_one,
                             int zero_padding, int M, int N, int clock_type, float *fwd_secs,
                             float *inv_secs);
void dwt_util_perf_cdf97_2_d(int stride_x, int stride_y, int size_o_big_x, int size_o_big_y,
                             int size_i_big_x, int size_i_big_y, int j_max, int decompose_one,
                             int zero_padding, int M, int N, int clock_type, double *fwd_secs,
                             double *inv_secs);
int dwt_util_pow2_ceil_log2(int x);
/* More host utilities (reference: src/libdwt.h:606, 744, 803, 950, 961-1001,
 * 1010-1036): ceil(log2 x); whether a timer type is available (0) or not
 * (-1); the worker count (1: no ASVP workers); an active wait of ms
 * milliseconds; float vectors (allocate 8-byte aligned, even size; fill with
 * 0, 1, 2, ...; zero; copy; compare -- the last four return non-zero when a
 * value is not normal-or-zero or differs by more than 1e-4); printf
 * replacements. */
int dwt_util_ceil_log2(int x);
int dwt_util_clock_available(int type);
int dwt_util_get_max_workers(void);
void dwt_util_wait(int ms);
float *dwt_util_allocate_vec_s(int size);
int dwt_util_generate_vec_s(float *addr, int size);
int dwt_util_zero_vec_s(float *addr, int size);
int dwt_util_copy_vec_s(const float *src, float *dst, int size);
int dwt_util_cmp_vec_s(const float *a, const float *b, int size);
int dwt_util_vfprintf(FILE *stream, const char *format, va_list ap);
int dwt_util_vprintf(const char *format, va_list ap);
int dwt_util_fprintf(FILE *stream, const char *format, ...);
int dwt_util_printf(const char *format, ...);
void dwt_util_get_sizes_s(enum dwt_array array_type, int size_x, int size_y, int opt_stride,
                          int *stride_x, int *stride_y, int *size_o_big_x, int *size_o_big_y,
                          int *size_i_big_x, int *size_i_big_y);
void dwt_util_get_sizes_d(enum dwt_array array_type, int size_x, int size_y, int opt_stride,
                          int *stride_x, int *stride_y, int *size_o_big_x, int *size_o_big_y,
                          int *size_i_big_x, int *size_i_big_y);
void dwt_util_measure_perf_cdf97_1_s(enum dwt_array array_type, int min_x, int max_x,
                                     int opt_stride, int j_max, int decompose_one,
                                     int zero_padding, int M, int N, int clock_type,
                                     FILE *fwd_plot_data, FILE *inv_plot_data);
void dwt_util_measure_perf_cdf97_1_d(enum dwt_array array_type, int min_x, int max_x,
                                     int opt_stride, int j_max, int decompose_one,
                                     int zero_padding, int M, int N, int clock_type,
                                     FILE *fwd_plot_data, FILE *inv_plot_data);
void dwt_util_measure_perf_cdf97_2_s(enum dwt_array array_type, int min_x, int max_x,
                                     int opt_stride, int j_max, int decompose_one,
                                     int zero_padding, int M, int N, int clock_type,
                                     FILE *fwd_plot_data, FILE *inv_plot_data);
void dwt_util_measure_perf_cdf97_2_d(enum dwt_array array_type, int min_x, int max_x,
                                     int opt_stride, int j_max, int decompose_one,
                                     int zero_padding, int M, int N, int clock_type,
                                     FILE *fwd_plot_data, FILE *inv_plot_data);

/* ===========================================================================
 * Extensions (no counterpart in the reference).  They return 0 on success or
 * a negative DWT_GPU_E* code; dwt_gpu_last_error() describes the last failure
 * of the calling thread.
 * ======================================================================== */
enum dwt_gpu_wavelet { DWT_GPU_CDF97 = 0, DWT_GPU_CDF53 = 1 };
enum dwt_gpu_direction { DWT_GPU_FORWARD = 0, DWT_GPU_INVERSE = 1 };
enum dwt_gpu_precision { DWT_GPU_F32 = 0, DWT_GPU_F64 = 1 };
enum dwt_gpu_status {
	DWT_GPU_OK = 0,
	DWT_GPU_EINVAL = -1,   /* bad argument */
	DWT_GPU_EMEMORY = -2,  /* pointer not device-accessible / allocation failed */
	DWT_GPU_EHIP = -3      /* HIP runtime error */
};

/* Initialise the library on the current device (optional; done on first use). */
int dwt_gpu_init(void);
/* Text of the calling thread's last error ("" when none). */
const char *dwt_gpu_last_error(void);
/* Version string of this library. */
const char *dwt_gpu_version(void);

/* 2D multilevel transform of device-accessible memory, enqueued on `stream`
 * (a hipStream_t; NULL = the null stream) without synchronising.  Arguments
 * as dwt_cdf97_2f_s / dwt_cdf97_2i_s; *j is read (inverse) or clamped and
 * written (forward).  Calls on one stream are ordered; scratch memory is kept
 * per (host thread, stream), so concurrent calls must use distinct streams or
 * distinct threads. */
int dwt_gpu_2d_async(int wavelet, int direction, int precision, void *ptr, int stride_x,
                     int stride_y, int size_o_big_x, int size_o_big_y, int size_i_big_x,
                     int size_i_big_y, int *j, int decompose_one, int zero_padding,
                     void *stream);

/* Allocation-free variant of dwt_gpu_2d_async.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * The call allocates nothing (no hipMalloc / hipFree, no streams or events, no
 * library routine that allocates); its scratch is a static device array of
 * DWT_GPU_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until
 * the process ends.  Frames: packed FP32 in device memory (ptr and stride_x
 * multiples of 4 bytes, element stride 4 bytes, outer frame = inner frame,
 * decompose_one = 0, zero_padding irrelevant).  *j as dwt_gpu_2d_async (read
 * for the inverse, clamped and written for the forward).  An image whose
 * scratch exceeds the array -- W x H x 4 x (1 + 1/4 + 1/16) bytes with padded
 * pitches, about 12.7 Mpixel; an image the single-workgroup kernel takes
 * whole (at most 128 KiB, sides at most 512) needs none -- is rejected with
 * DWT_GPU_EINVAL before anything is enqueued.  Asynchronous on
 * `stream`.  Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (other streams, other host threads) race; calls on one
 * stream are ordered. */
#define DWT_GPU_NOALLOC_SCRATCH_BYTES (64ll << 20)
int dwt_gpu_2d_noalloc(int wavelet, int direction, void *ptr, int stride_x, int size_x, int size_y,
                       int *j, void *stream);

/* Every channel of a multi-channel image in one call.  Channel c is the
 * size_x x size_y frame at ptr + c * channel_stride (bytes), rows stride_x
 * and pixels stride_y bytes apart, as the drivers take them (outer frame =
 * inner frame); the channels must not overlap.  Each channel is transformed
 * exactly as dwt_gpu_2d_async on that frame would (same *j semantics,
 * decompose_one, zero_padding).  No counterpart in the reference: it replaces
 * the per-channel driver calls of its OpenCV wrapper (src/cvdwt.cpp:67-147)
 * for interleaved images, gathering the channels into packed planes in one
 * pass and scattering them back in one pass.  Asynchronous on `stream`. */
int dwt_gpu_2d_channels(int wavelet, int direction, int precision, void *ptr, int stride_x,
                        int stride_y, int channel_stride, int channels, int size_x, int size_y,
                        int *j, int decompose_one, int zero_padding, void *stream);

/* Accuracy statistics of a device-side compare (dwt_gpu_compare). */
#define DWT_GPU_ULP_BINS 32
typedef struct dwt_gpu_compare_stats {
	long long count;      /* pixel pairs compared */
	long long nonfinite;  /* pairs with a NaN or an infinity on either side */
	long long over_eps;   /* finite pairs with |a - b| > eps */
	double max_abs;       /* largest |a - b| over the finite pairs */
	double sum_sq;        /* sum of (a - b)^2 over the finite pairs */
	double mse;           /* sum_sq / (count - nonfinite); 0 when none is finite */
	double peak_b;        /* largest |b| over the finite pairs */
	double psnr;          /* 10 log10(peak^2 / mse) dB (peak: the argument, else
	                         peak_b); +inf when mse == 0 */
	long long ulp[DWT_GPU_ULP_BINS]; /* distance in ulps of a's precision
	                         between a and b rounded to it: bin 0 equal, bin k
	                         [2^(k-1), 2^k), the last bin everything above */
	int differ;           /* dwt_util_compare_s / _d's verdict: 1 when a pair is
	                         non-finite or further apart than eps */
	int reserved;
} dwt_gpu_compare_stats;

/* Compare two images in device memory, element by element, without copying
 * them to the host: dwt_util_compare_s / _d (src/libdwt.c:1223-1277) with
 * statistics.  precision_a / _b: DWT_GPU_F32 or DWT_GPU_F64 (may differ);
 * byte strides as dwt_cdf97_2f_s (element-aligned).  eps < 0 selects the
 * reference's threshold: 1e-3 when either image is FP32, 1e-6 when both are
 * FP64; the |a - b| > eps test runs in FP32 when both images are FP32, as the
 * reference's.  Synchronises `stream`; returns DWT_GPU_OK or an error code. */
int dwt_gpu_compare(int precision_a, const void *a, int stride_x_a, int stride_y_a,
                    int precision_b, const void *b, int stride_x_b, int stride_y_b,
                    int size_x, int size_y, double eps, double peak,
                    dwt_gpu_compare_stats *stats, void *stream);

/* FP64 verification of an FP32 2D transform on the GPU: `input` (FP32, the
 * transform's input) is converted to FP64 and transformed in FP64 on the
 * device (j levels; forward: j as dwt_cdf97_2f_d clamps it), and `result`
 * (FP32, what the FP32 transform produced from `input`) is compared with it
 * (eps 1e-3, peak = the largest |coefficient| of the FP64 result).  Both
 * images are device memory with the same byte strides and a packed frame
 * (size_o == size_i).  Synchronises `stream`. */
int dwt_gpu_verify_2d(int wavelet, int direction, const void *input, const void *result,
                      int stride_x, int stride_y, int size_x, int size_y, int j,
                      dwt_gpu_compare_stats *stats, void *stream);

/* Per-launch timing with HIP events recorded on the launch stream around each
 * kernel of the packed path.  kind: 0 forward level kernel, 1 inverse level
 * kernel, 2 level-0 detail copy, 4 forward block kernel, 5 inverse block
 * kernel, 6 forward pyramid, 7 inverse pyramid (levels after level 0),
 * 8 / 9 forward / inverse batched level launch, 10 / 11 forward / inverse
 * fused middle levels (dwt_gpu_mid_set), 12 / 13 forward / inverse batched
 * tail (dwt_gpu_batch_tail_set; level -1);
 * level: the level index, the first level it does for a block kernel or a
 * pyramid (-1 = all).
 * on: 0 off, 1 on, k >= 2 on with every level-0 forward kernel of the packed
 * path launched k times back to back inside one event pair (it reads the
 * image and writes only scratch, so results are unchanged); profile_read
 * counts launches, so its average is then the back-to-back kernel duration.
 * profile_read synchronises on the recorded events. */
int dwt_gpu_profile_enable(int on);
int dwt_gpu_profile_read(int kind, int level, long long *launches, double *total_ms);
int dwt_gpu_profile_reset(void);

/* Tile shape of the packed level kernels: a wavefront walks seg_rows rows
 * (even, 2..4096) of a 128-column strip, prefetching prefetch_pairs (4 or 8)
 * row pairs, or (prefetch_pairs 0) requests all rows of its segment at once
 * (seg_rows 2, 4, 8, 12, 16, 20, 24 or 32).  Shapes are kept per direction (forward 1 / inverse 0) and level
 * height class ceil(log2(height)); seg_rows = 0 restores the built-in choice.
 * dwt_gpu_tile_get reports the shape a level of that size would use. */
int dwt_gpu_tile_set(int forward, int height_class, int seg_rows, int prefetch_pairs);
int dwt_gpu_tile_get(int forward, int width, int height, int wavelet, int *seg_rows,
                     int *prefetch_pairs);

/* The small levels of a transform run as one single-workgroup launch that
 * keeps the whole small pyramid in LDS, from the first level j >= 1 whose
 * region is at most max_region_bytes (default 64 KiB, at most 150 KiB; 0
 * disables it) and whose sides are at most 512 samples.  An image of at most
 * twice that (at most 150 KiB) is transformed completely by that one launch. */
int dwt_gpu_tail_set(long long max_region_bytes);

/* Level 0 of the packed path sends its detail subbands through scratch (one
 * copy of them per direction, carried by the smaller levels' launches).
 * inplace = 1 runs it in place instead: each workgroup loads its tile,
 * signals that its reads are done, waits for the tiles that read what it
 * overwrites, then stores, and only the rows that close a cycle of that
 * dependency order pass through scratch (about a sixth of the copy).  Same
 * results; measured slower on MI355X (the waits cost more than the copy), so
 * off by default.  band (> 0) sets the row segments per XCD band of the
 * in-place order (default 8; 0 keeps it).  Launches under stream capture and
 * the allocation-free entry point always use the scratch route. */
int dwt_gpu_level0_set(int inplace, int band);

/* Test support (no counterpart in the reference): the in-place level-0 plan of
 * a width x height image with seg_rows-row tiles (4 or 8).  info[8] receives
 * {segments, column groups, strips, tiles, dependencies, copied rows,
 * workgroups, band}; each non-null array receives: order[segments] (the
 * segment order), segflags[segments] (forward: bit 0 HL rows, bit 1 LH and HH
 * rows through scratch; inverse: bit 0 band preserved in scratch, bit 8 + u
 * single-shot read row u of the segment's tiles comes from scratch),
 * list_off[9] and list[tiles] (tiles of XCD x, tile = segment * groups +
 * group, for workgroups 8k + x), dep_off[tiles + 1] and deps[dependencies]
 * (tiles whose reads a tile waits for), rows[copied rows] (image rows the
 * copy carries). */
int dwt_gpu_level0_plan(int forward, int width, int height, int wavelet, int seg_rows, int *info, int *order,
                        int *segflags, int *list_off, int *list, int *dep_off, int *deps, int *rows);

/* Optionally, three middle levels of the packed path run as one launch (each
 * workgroup computes overlapping windows of the three levels in LDS) from the
 * first level j >= 1 whose region is at most max_level_bytes, when all three
 * come before the single-workgroup launch.  Default 0 (off): measured slower
 * than one launch per level on MI355X. */
int dwt_gpu_mid_set(long long max_level_bytes);

/* Schedule of the levels after level 0 on the packed path: 0 (default) runs
 * one launch per level plus a single-workgroup launch for the small levels;
 * 1 runs them as one persistent launch in which each level's work items start
 * as soon as the rows they read are finished (dataflow, no barrier between
 * levels). */
int dwt_gpu_schedule_set(int mode);
/* Sparse frames (size_o != size_i) whose rows are contiguous: 1 (default)
 * runs each large level as one launch (inner tiles, garbage rows and columns
 * as the reference transforms them, into scratch; copied back by riders) and
 * the small levels as one block launch; 0 runs the large levels as the
 * generic per-line passes.  The results are identical. */
int dwt_gpu_sparse_set(int fused);
/* Synchronise `stream` and report whether a persistent launch of this thread
 * on it hit its wait time limit since the last check (DWT_GPU_EHIP; the
 * results of that launch are then invalid).  The drop-in calls check this
 * themselves before returning. */
int dwt_gpu_device_check(void *stream);

/* Timeline of pyramid launches: with max_items > 0 every later pyramid launch
 * records, per work item, {start, dependencies met, done} in ticks of the
 * GPU's constant-rate clock (dwt_gpu_clock_rate_khz) and the index of the
 * workgroup that ran it.  dwt_gpu_pyr_trace_read copies the records of the
 * last launch on `stream` (synchronising it) and returns their count. */
int dwt_gpu_pyr_trace_enable(int max_items);
/* Rate of the GPU's constant-rate clock in kHz. */
int dwt_gpu_clock_rate_khz(void);
/* At most max_workgroups workgroups per pyramid launch (0: as many as fit). */
int dwt_gpu_pyr_grid_cap_set(int max_workgroups);
long long dwt_gpu_pyr_trace_read(void *stream, unsigned long long *out, long long max_items);

/* ---- full-depth 1D transforms of many signals ---------------------------------
 * count signals of `length` samples, in place: signal k, sample i at
 * ptr + k * signal_stride + i * sample_stride (bytes, device-accessible
 * memory).  The result is that of the caller loop over
 * dwt_cdf97_f_ex_stride_s (src/libdwt.h:219-226) with L and H written back
 * to [0, ceil(n/2)) and [ceil(n/2), n) of each signal for n = length,
 * ceil(length/2), ... while n >= 2; the inverse undoes it.  *levels: forward,
 * requested (< 0 or too large: full depth) and written back; inverse, the
 * levels to undo.  Asynchronous on `stream`.  Fast paths, one launch per
 * level: contiguous samples (level 0 in place for 64 <= length <= 16384
 * FP32 / 8192 FP64, through scratch otherwise), and adjacent signals
 * (signal_stride = element size: the columns of a matrix; level 0 in place
 * but for its top eighth).  Any other layout runs one generic launch per
 * level. */
int dwt_gpu_1d_batch(int wavelet, int direction, int precision, void *ptr, long long signal_stride,
                     long long sample_stride, int length, int count, int *levels, void *stream);

/* ---- batches of images of uneven sizes ------------------------------------
 * One call transforms many packed single-channel images, out of place, with
 * one launch per level for the whole batch (and one for every image's small
 * levels, dwt_gpu_batch_tail_set).  Each image: src -> dst;
 * forward: src is the FP32 image, dst receives its coefficients in Mallat
 * layout (the layout of dwt_cdf97_2f_s with size_o == size_i); inverse: src
 * holds the coefficients, dst receives the FP32 image.  The coefficients are
 * stored as FP32 (bit-identical to the single-image transform), FP16 or BF16;
 * the lifting arithmetic is FP32.  For 16-bit storage every subband is stored
 * multiplied by a power-of-two scale that the forward writes to `scales` (3
 * per level: HL, LH, HH, then the final LL: 3 * levels + 1 floats in device
 * memory); the inverse reads them from there.  FP16: chosen from the image's
 * max |x| so that |coefficient * scale| <= 2^14.  BF16: 1 (BF16 has FP32's
 * exponent range, so a power-of-two scale would not change the stored
 * mantissa).  Pitches are in bytes, multiples of the element size.
 * Images need width and height >= 2; levels as j_max of dwt_cdf97_2f_s with
 * decompose_one = 0 (forward: < 0 or too large means the maximum, written
 * back; inverse: the levels to undo). */
typedef struct dwt_gpu_image {
	const void *src;
	void *dst;
	long long src_pitch;
	long long dst_pitch;
	int width;
	int height;
	int levels;
	int reserved;
	float *scales;
} dwt_gpu_image;
enum dwt_gpu_storage { DWT_GPU_STORE_F32 = 0, DWT_GPU_STORE_F16 = 1, DWT_GPU_STORE_BF16 = 2 };

/* One-shot batch on `stream` (asynchronous).  The plans of recent batches
 * (same descriptors) are kept and reused. */
int dwt_gpu_batch_2d(int wavelet, int direction, int storage, dwt_gpu_image *images, int n,
                     void *stream);
/* Plans: create once (tables uploaded, scratch allocated; synchronises
 * `stream`), execute any number of times on one stream at a time (only kernel
 * launches, so it can be captured in a HIP graph), destroy. */
int dwt_gpu_batch_plan_create(int wavelet, int direction, int storage, dwt_gpu_image *images,
                              int n, void *stream, void **plan);
int dwt_gpu_batch_plan_execute(void *plan, void *stream);
int dwt_gpu_batch_plan_destroy(void *plan);
/* Segment rows (8, 16 or 32; 0: default 8) of the batched launches of one
 * level, for plans created afterwards. */
int dwt_gpu_batch_rows_set(int level, int seg_rows);
/* Batched tail: from the first level whose region is at most max_region_bytes
 * of FP32 samples (and at most 512 samples a side), every remaining level of
 * every image runs in one launch, one workgroup per image, in LDS
 * (bit-identical to the per-level launches).  0 turns it off; at most 131072;
 * default 65536.  For plans created afterwards. */
int dwt_gpu_batch_tail_set(long long max_region_bytes);

/* ---- serving loop -----------------------------------------------------------
 * A server takes transform requests (one image each, buffers resident in
 * device memory, semantics of one image of dwt_gpu_batch_2d) from any thread
 * at any time and runs them with continuous batching: a scheduler thread
 * owns `streams` streams; whenever one is free it admits waiting requests in
 * arrival order while their scratch fits the HBM budget (allocated once, at
 * creation) and the batch holds at most max_batch requests, and launches
 * them as one batch on that stream.  Requests arriving meanwhile join the
 * next batch, so batches on different streams overlap.  Each result equals
 * the result of the request alone.  poll() returns completed requests (id,
 * latency from submission to observed completion in seconds, status). */
typedef struct dwt_gpu_request {
	int wavelet;    /* DWT_GPU_CDF97 / DWT_GPU_CDF53 */
	int direction;  /* DWT_GPU_FORWARD / DWT_GPU_INVERSE */
	int storage;    /* DWT_GPU_STORE_* of the coefficients */
	int reserved;
	dwt_gpu_image image;
} dwt_gpu_request;
int dwt_gpu_server_create(long long hbm_budget_bytes, int max_batch, int streams, void **server);
int dwt_gpu_server_submit(void *server, const dwt_gpu_request *request, long long *id);
/* Native load generator: submit reqs[i] at at_s[i] seconds (ascending) after
 * the call starts, from the calling thread (sleeping through long gaps,
 * spinning through the last 50 us); ids[i] receives each request's id.
 * Returns n once all are submitted, or an error code. */
int dwt_gpu_server_replay(void *server, const dwt_gpu_request *reqs, const double *at_s, int n,
                          long long *ids);
/* Up to max completed requests; returns their count. */
int dwt_gpu_server_poll(void *server, long long *ids, double *latency_s, int *status, int max);
/* Wait until every submitted request has completed. */
int dwt_gpu_server_drain(void *server);
int dwt_gpu_server_stats(void *server, long long *batches, long long *requests);
/* Waits for the requests in flight, then frees the server. */
int dwt_gpu_server_destroy(void *server);

/* Test support (no counterpart in the reference): the staged in-place level 0
 * of dwt_gpu_1d_batch for signals stored as matrix columns.  dwt_gpu_col_plan
 * returns 1 when level 0 of an n-row level runs in place as segment ranges
 * [stages[k], stages[k+1]) (k < 3; forward launched from the highest range
 * down, inverse from the lowest up) plus the top segments [stages[3],
 * segments) whose H rows from *top_h_row on pass through scratch, 0 when it
 * runs through scratch entirely.  dwt_gpu_col_segment_rows lists the rows one
 * segment reads and writes, computed by the column kernels' own index
 * arithmetic: forward, input rows read and H rows written (at ceil(n/2) + m);
 * inverse, H rows read and output rows written. */
int dwt_gpu_col_plan(int forward, int n, int nsteps, int *stages, int *top_h_row, int *segments,
                     int *rows_per_segment);
int dwt_gpu_col_segment_rows(int forward, int n, int nsteps, int seg, int *reads, int max_reads,
                             int *nreads, int *writes, int max_writes, int *nwrites);

/* Test support (no counterpart in the reference): builds the tables of an FP32
 * batch (dwt_gpu_batch_plan_create's descriptors; the buffers are never
 * touched, so any distinct non-null pointers do) on the host and replays every
 * level launch's workgroup -> (image, tile) mapping of the batched level
 * kernels; returns DWT_GPU_OK when each tile of every image and level is
 * computed by exactly one wave, an error code (message in
 * dwt_gpu_last_error) otherwise. */
int dwt_gpu_batch_order_check(int wavelet, int direction, const dwt_gpu_image *images, int n);

/* Bytes of device scratch the calling thread currently holds for `stream`.
 * Scratch is kept per (host thread, stream, device) and reused by later calls;
 * a thread's sets are freed when the thread exits (except the thread that
 * loaded the library, whose sets live until the process ends) or by
 * dwt_gpu_release_scratch.  A caller that destroys a stream should release
 * its scratch first. */
long long dwt_gpu_scratch_bytes(void *stream);
/* Release the calling thread's scratch for `stream` (synchronises it). */
int dwt_gpu_release_scratch(void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LIBDWT_H */
