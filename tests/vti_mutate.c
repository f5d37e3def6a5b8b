/*
 * Mutation test of the VTI reader (pdfs_vti_read), built by tests/test_vti.py
 * with gcc -fsanitize=address,undefined and run on the CPU.
 *
 * usage: vti_mutate <seed.vti> <iterations> <work file>
 *
 * Each iteration applies 1-4 random edits to the seed file (byte
 * replacement, truncation, deletion of up to 16 bytes, insertion of a token
 * chosen to stress the header and number parsing: huge, negative and
 * non-finite numbers, quotes, attribute names, closing tags) and reads the
 * result. The reader must return one of its documented codes; a file it
 * accepts must have in-range extents and a ghost layer of location -1 and
 * distance +inf. Memory errors, undefined behaviour and leaks end the run
 * through the sanitizers. Prints a histogram of the return codes.
 */
#include "pdfs_vti.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static uint64_t rnd(void) {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}

static const char *TOKENS[] = {"999999999", "-5", "2147483647", "1e400", "nan", "-inf", "0", " ",
                               "\n", "\"", "=", "Spacing=\"", "WholeExtent=\"0 ", "4294967295",
                               "1073741823", "-2147483648", "0x", "1e-320", "</DataArray>", ""};

static int check_ghosts(const pdfs_vti *v) {
  const long gx = v->x + 2L, gy = v->y + 2L, gz = v->z + 2L;
  for (long k = 0; k < gz; k++)
    for (long j = 0; j < gy; j++)
      for (long i = 0; i < gx; i++) {
        const int ghost = k == 0 || k == gz - 1 || j == 0 || j == gy - 1 || i == 0 || i == gx - 1;
        const long p = (k * gy + j) * gx + i;
        if (ghost && (v->location[p] != -1 || !(isinf(v->distance[p]) && v->distance[p] > 0)))
          return 0;
      }
  return 1;
}

int main(int argc, char **argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s seed.vti iterations workfile\n", argv[0]);
    return 2;
  }
  const long iters = atol(argv[2]);
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  fseek(f, 0, SEEK_END);
  const long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *seed = (char *)malloc((size_t)n + 1);
  char *buf = (char *)malloc((size_t)n + 4096);
  if (!seed || !buf || fread(seed, 1, (size_t)n, f) != (size_t)n) return 2;
  fclose(f);
  long hist[5] = {0, 0, 0, 0, 0};
  for (long it = 0; it < iters; it++) {
    long m = n;
    memcpy(buf, seed, (size_t)n);
    const int edits = 1 + (int)(rnd() % 4);
    for (int e = 0; e < edits; e++) {
      const int kind = (int)(rnd() % 5);
      const long at = m ? (long)(rnd() % (uint64_t)m) : 0;
      if (kind == 0 && m) {
        buf[at] = (char)(rnd() & 0xff);
      } else if (kind == 1) {
        m = at;
      } else if (kind == 2 && m > at + 1) {
        long len = 1 + (long)(rnd() % 16);
        if (at + len > m) len = m - at;
        memmove(buf + at, buf + at + len, (size_t)(m - at - len));
        m -= len;
      } else {
        const char *t = TOKENS[rnd() % (sizeof TOKENS / sizeof *TOKENS)];
        const long len = (long)strlen(t);
        if (m + len < n + 4000) {
          memmove(buf + at + len, buf + at, (size_t)(m - at));
          memcpy(buf + at, t, (size_t)len);
          m += len;
        }
      }
    }
    FILE *w = fopen(argv[3], "wb");
    if (!w || fwrite(buf, 1, (size_t)m, w) != (size_t)m) return 2;
    fclose(w);
    pdfs_vti v;
    const int rc = pdfs_vti_read(argv[3], &v);
    if (rc > 0 || rc < -4) {
      fprintf(stderr, "iteration %ld: undocumented return code %d\n", it, rc);
      return 1;
    }
    if (rc == 0) {
      const int ok = v.x >= 1 && v.y >= 1 && v.z >= 1 && v.x < (1 << 30) && v.y < (1 << 30) &&
                     v.z < (1 << 30) && check_ghosts(&v);
      pdfs_vti_free(&v);
      if (!ok) {
        fprintf(stderr, "iteration %ld: accepted file with bad extents or ghost layer\n", it);
        return 1;
      }
    }
    hist[-rc]++;
  }
  free(seed);
  free(buf);
  printf("ok %ld open %ld header %ld memory %ld data %ld\n", hist[0], hist[1], hist[2], hist[3],
         hist[4]);
  return 0;
}
