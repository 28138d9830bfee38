/* Error reporting: bfam aborts on every error (bfam.h:36-57); bfamx returns a
 * status code, keeps a per-thread message, and aborts only on request. */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

#include "../bfamx_internal.h"

static __thread char g_msg[512];
static int g_abort_on_error = 0;

const char *bfamx_last_error(void) { return g_msg; }

void bfamx_set_abort_on_error(int on) { g_abort_on_error = on; }

void bfamx_clear_error(void) { g_msg[0] = '\0'; }

int bfamx_fail(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  if (g_abort_on_error)
  {
    fprintf(stderr, "bfamx: error %d: %s\n", code, g_msg);
    abort();
  }
  return code;
}
