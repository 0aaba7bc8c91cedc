/* A C program using the library exactly as a libdwt program does: host
 * images from dwt_util_alloc_image at a prime row pitch, the type-0 test
 * pattern, forward and inverse CDF 9/7 and 5/3 in place, dwt_util_compare_s,
 * subband addressing, and the device-resident performance harness.  Prints
 * "roundtrip ok" / "perf ok" lines; exits non-zero on any failure. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "libdwt.h"

static int roundtrip(int w, int h, int wavelet) {
	const int stride_x = dwt_util_get_opt_stride(w * (int)sizeof(float));
	const int stride_y = sizeof(float);
	void *a, *b;
	dwt_util_alloc_image(&a, stride_x, stride_y, w, h);
	dwt_util_alloc_image(&b, stride_x, stride_y, w, h);
	dwt_util_test_image_fill_s(a, stride_x, stride_y, w, h, 0);
	dwt_util_test_image_fill_s(b, stride_x, stride_y, w, h, 0);
	int j = -1;
	if (wavelet == 97)
		dwt_cdf97_2f_s(a, stride_x, stride_y, w, h, w, h, &j, 0, 0);
	else
		dwt_cdf53_2f_s(a, stride_x, stride_y, w, h, w, h, &j, 0, 0);
	/* the LL band of the last level holds the image's mean times 2^j */
	void *ll;
	int lx, ly;
	dwt_util_subband_s(a, stride_x, stride_y, w, h, w, h, j, DWT_LL, &ll, &lx, &ly);
	const int differs_after_fwd = dwt_util_compare_s(a, b, stride_x, stride_y, w, h);
	if (wavelet == 97)
		dwt_cdf97_2i_s(a, stride_x, stride_y, w, h, w, h, j, 0, 0);
	else
		dwt_cdf53_2i_s(a, stride_x, stride_y, w, h, w, h, j, 0, 0);
	const int differs = dwt_util_compare_s(a, b, stride_x, stride_y, w, h);
	dwt_util_log(LOG_INFO, "%dx%d cdf%d pitch %d: j=%d LL %dx%d %s\n", w, h, wavelet, stride_x, j,
	             lx, ly, differs ? "images differ" : "roundtrip ok");
	dwt_util_free_image(&a);
	dwt_util_free_image(&b);
	return differs || !differs_after_fwd || lx < 1 || ly < 1;
}

int main(void) {
	dwt_util_init();
	int bad = 0;
	bad |= roundtrip(512, 512, 97);
	bad |= roundtrip(640, 481, 53);
	bad |= roundtrip(1021, 67, 97);
	float fwd = 0, inv = 0;
	dwt_util_perf_cdf97_2_s(dwt_util_get_opt_stride(1024 * 4), 4, 1024, 1024, 1024, 1024, -1, 0, 0,
	                        2, 3, dwt_util_clock_autoselect(), &fwd, &inv);
	dwt_util_log(LOG_INFO, "perf 1024x1024: fwd %.6f s inv %.6f s\n", fwd, inv);
	if (!(fwd > 0 && inv > 0 && isfinite(fwd) && isfinite(inv)))
		bad = 1;
	FILE *ff = fopen("perf_fwd.txt", "w"), *fi = fopen("perf_inv.txt", "w");
	if (!ff || !fi)
		return 2;
	dwt_util_measure_perf_cdf97_2_s(DWT_ARR_PACKED, 16, 100, 1, -1, 0, 0, 1, 2,
	                                dwt_util_clock_autoselect(), ff, fi);
	fclose(ff);
	fclose(fi);
	dwt_util_finish();
	printf(bad ? "FAILED\n" : "perf ok\n");
	return bad;
}
