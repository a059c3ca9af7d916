/* Builds as plain C99 against include/hpxfft_hip.h and links
 * libhpxfft_hip.so: the ABI header is valid C, and its entry points resolve.
 * Without a GPU it calls only the functions that need none: the version, the
 * plan-flag strings and the argument checks (a NULL plan pointer). */
#include <stdio.h>
#include <string.h>

#include "hpxfft_hip.h"

int main(void)
{
    int effort = -1;
    if (hpxfft_plan_flag_from_string("measure", &effort) != HPXFFT_OK || effort != HPXFFT_MEASURE) return 1;
    if (hpxfft_plan_flag_from_string("bogus", &effort) != HPXFFT_INVALID_ARG) return 2;
    if (strcmp(hpxfft_last_error(), "Invalid FFTW plan flag string") != 0) return 3;
    if (hpxfft_plan_r2c_2d(NULL, 8, 14, 16, HPXFFT_F64, HPXFFT_ESTIMATE, NULL) != HPXFFT_INVALID_ARG) return 4;
    if (hpxfft_exec_r2c_2d(NULL, NULL, NULL) != HPXFFT_INVALID_ARG) return 5;
    if (hpxfft_plan_destroy(NULL) != HPXFFT_OK) return 6;
    printf("%s\n", hpxfft_version());
    return 0;
}
