/* Uses the dwt_util_* helpers a libdwt program may call besides the
 * transforms (ceil_log2, clock availability, worker count, active wait,
 * float vectors, printf family).  Linked unchanged against this library and
 * against the reference built from its own sources; both must print the same. */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

#include "libdwt.h"

static int say(const char *fmt, ...) {
	va_list ap;
	va_start(ap, fmt);
	int r = dwt_util_vprintf(fmt, ap);
	va_end(ap);
	return r;
}

int main(void) {
	for (int x = 1; x <= 70; x += 3)
		dwt_util_printf("ceil_log2(%d)=%d\n", x, dwt_util_ceil_log2(x));
	dwt_util_printf("clock_gettime available: %d\n", dwt_util_clock_available(DWT_TIME_CLOCK_GETTIME));
	dwt_util_printf("max workers: %d\n", dwt_util_get_max_workers());
	dwt_util_wait(2);
	float *a = dwt_util_allocate_vec_s(64), *b = dwt_util_allocate_vec_s(64);
	dwt_util_printf("aligned: %d\n", dwt_util_is_aligned_8(a) && dwt_util_is_aligned_8(b));
	dwt_util_printf("generate: %d\n", dwt_util_generate_vec_s(a, 64));
	dwt_util_printf("zero: %d\n", dwt_util_zero_vec_s(b, 64));
	dwt_util_printf("cmp differ: %d\n", dwt_util_cmp_vec_s(a, b, 64) != 0);
	dwt_util_printf("copy: %d\n", dwt_util_copy_vec_s(a, b, 64));
	dwt_util_printf("cmp equal: %d\n", dwt_util_cmp_vec_s(a, b, 64));
	b[17] += 5e-5f;
	dwt_util_printf("cmp within 1e-4: %d\n", dwt_util_cmp_vec_s(a, b, 64));
	b[17] += 1e-3f;
	dwt_util_printf("cmp beyond 1e-4: %d\n", dwt_util_cmp_vec_s(a, b, 64) != 0);
	dwt_util_fprintf(stdout, "sum %g %s\n", a[63] + b[1], "ok");
	say("vprintf %d\n", 42);
	free(a);
	free(b);
	return 0;
}
