/* Oracle (test infrastructure only): the layout of a code, shared by its C files. */
#ifndef OR_INTERNAL_H
#define OR_INTERNAL_H
#include <stdint.h>

struct or_code {
    int32_t n, m, k, k_info, e, bg, z, ils, core;
    int32_t* row_start; /* m + 1 */
    int32_t* cols;      /* edges */
    int32_t* solve_start; /* core + 1: rows of the inverse of the core block */
    int32_t* solve;
    int32_t* map;       /* e output positions (NR codes) */
};

#endif
