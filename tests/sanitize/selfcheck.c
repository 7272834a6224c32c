/* selfcheck.c -- test infrastructure only: one heap read past the end of an
 * allocation, which the sanitizer build must report (the test expects a
 * non-zero exit with "heap-buffer-overflow"); it shows the instrumentation
 * of the other programs is live.  CPU only. */
#include <stdio.h>
#include <stdlib.h>

int main(int argc, char **argv)
{
   (void)argv;
   double *a = malloc(4 * sizeof(double));
   for (int i = 0; i < 4; ++i) a[i] = i;
   const double v = a[3 + argc];     /* argc >= 1: one past the end */
   printf("%g\n", v);
   free(a);
   return 0;
}
