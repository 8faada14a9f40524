/* Reads every frame of a job directory with the drop-in executables' frame
 * reader (activechi4_amd/host/frame_reader.h) and prints a checksum; built
 * with AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_frame_reader_sanitized.py to feed it malformed input.
 * usage: frame_reader_harness <dir> <N> <header_count 0|1> <bin 0|1>
 * exit: 0 read, 2 index file refused, 3 data file refused, 4 frame refused */
#include <stdio.h>
#include <stdlib.h>
#include "../../activechi4_amd/host/frame_reader.h"

int main(int argc, char **argv) {
    if (argc != 5) return 1;
    const int n = atoi(argv[2]);
    int nf = 0;
    int *fidx = fr_read_index(argv[1], atoi(argv[3]), &nf);
    if (!fidx) return 2;
    frame_reader r;
    if (fr_open(&r, argv[1], n, atoi(argv[4]))) { free(fidx); return 3; }
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    int *ty = malloc(sizeof(int) * n);
    double sum = 0.0;
    long long tsum = 0;
    int rc = 0;
    for (int f = 0; f < nf && !rc; f++) {
        if (fr_next(&r, x, y, z, ty)) { rc = 4; break; }
        for (int i = 0; i < n; i++) { sum += x[i] + 2.0 * y[i] + 3.0 * z[i]; tsum += ty[i]; }
    }
    if (!rc) printf("%d %.17g %lld\n", nf, sum, tsum);
    fr_close(&r);
    free(x); free(y); free(z); free(ty); free(fidx);
    return rc;
}
