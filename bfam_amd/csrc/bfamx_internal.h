/* Internal helpers shared by the host C files and the HIP runtime. */
#ifndef BFAMX_INTERNAL_H
#define BFAMX_INTERNAL_H

#include "bfamx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* record an error (thread-local message), abort if requested; returns code */
int bfamx_fail(int code, const char *fmt, ...)
    __attribute__((format(printf, 2, 3)));
void bfamx_clear_error(void);

/* bits of the face permutation code used by bfamx_face_connectivity */
#define BFAMX_PERM_FLIP_A 1
#define BFAMX_PERM_FLIP_B 2
#define BFAMX_PERM_SWAP 4

/* index within the neighbour face list of face node (a, b), n1 = N+1 */
static inline int bfamx_perm_apply(int code, int n1, int a, int b)
{
  int x = (code & BFAMX_PERM_SWAP) ? b : a;
  int y = (code & BFAMX_PERM_SWAP) ? a : b;
  if (code & BFAMX_PERM_FLIP_A)
    x = n1 - 1 - x;
  if (code & BFAMX_PERM_FLIP_B)
    y = n1 - 1 - y;
  return x + y * n1;
}

#ifdef __cplusplus
}
#endif
#endif
