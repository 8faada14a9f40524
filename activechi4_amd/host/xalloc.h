/*
 * xalloc.h -- allocation for the drop-in executables: an allocation that
 * fails ends the program with a message, as the reference's programs end
 * on errors.
 */
#pragma once
#include <stdio.h>
#include <stdlib.h>

static inline void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "out of memory (%zu bytes)\n", n);
        exit(1);
    }
    return p;
}

static inline void *xcalloc(size_t n, size_t size) {
    void *p = calloc(n ? n : 1, size ? size : 1);
    if (!p) {
        fprintf(stderr, "out of memory (%zu x %zu bytes)\n", n, size);
        exit(1);
    }
    return p;
}
