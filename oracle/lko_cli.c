/* lko_cli: runs the CPU oracle on an .in file read from stdin, like
   `lakcluster_ver18 < file.in` (lakcluster_ver18.c:238-370).  Test infrastructure. */
#include "lko_oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
int main(int argc, char **argv)
{
  size_t cap = 1 << 16, n = 0; char *buf = malloc(cap); int ch;
  while ((ch = getchar()) != EOF) { if (n + 1 >= cap) { cap *= 2; buf = realloc(buf, cap); } buf[n++] = (char)ch; }
  buf[n] = 0;
  int max_iter = argc > 1 ? atoi(argv[1]) : -1;
  void *h = lko_run_text(buf, NULL, max_iter, 1);
  char err[256]; int st = lko_status(h, err, sizeof err);
  if (st) { fprintf(stderr, "lko_cli: %s\n", err); lko_free(h); return 2; }
  double t[3]; lko_times(h, t);
  printf(" %% oracle: %d iterations, load %.2fs init %.2fs loop %.2fs, ops %lld\n", lko_n_iter(h), t[0], t[1], t[2], lko_ops(h));
  lko_free(h); free(buf);
  return 0;
}
