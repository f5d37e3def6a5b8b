/*
 * VTI input and output of the PDFS CLI.
 *
 * Reader: restates single_gpu_cuda/src/vti_parser.c:85-180 on an in-memory
 * copy of the file. The dimension line is the first line containing
 * "ImageData WholeExtent", parsed with the reference's sscanf format
 * (vti_parser.c:91-94). Data: skip 6 lines from the top (fgets of at most
 * 510 characters, vti_parser.c:62-71), read one int per interior point
 * ("%d " semantics: leading and trailing whitespace consumed), skip 2 more
 * lines, read one double per interior point. The padded (x+2)(y+2)(z+2)
 * arrays get location -1 and distance +inf on the ghost layer
 * (vti_parser.c:137-180). Unlike the reference, a short or malformed file
 * is an error instead of undefined values.
 *
 * Writer: restates single_gpu_cuda/src/file_writer.c:152-190 byte for byte
 * (header with Origin between WholeExtent and Spacing, "%f  " values, a
 * newline per row and per plane), formatted in parallel per z-plane.
 */
#define _POSIX_C_SOURCE 200809L
#include "pdfs_vti.h"

#include <errno.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char *slurp(const char *path, size_t *len)
{
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  if (fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    return NULL;
  }
  long n = ftell(f);
  rewind(f);
  char *buf = (char *)malloc((size_t)n + 1);
  if (!buf || fread(buf, 1, (size_t)n, f) != (size_t)n) {
    free(buf);
    fclose(f);
    return NULL;
  }
  buf[n] = 0;
  fclose(f);
  *len = (size_t)n;
  return buf;
}

/* fgets(s, 511, f): consume up to 510 characters or through a newline. */
static const char *fgets511(const char *p, const char *end)
{
  int k = 0;
  while (p < end && k < 510) {
    k++;
    if (*p++ == '\n') break;
  }
  return p;
}

static int is_ws(char c)
{
  return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f';
}

/* Serial form of the reference's fscanf("%d ") / fscanf("%lf ") loop: count
 * numbers from p, each followed by its whitespace; NULL on a bad number. */
static const char *parse_serial(const char *p, const char *end, size_t count, int as_int,
                                int *iv, double *dv)
{
  for (size_t t = 0; t < count; t++) {
    char *e;
    if (as_int) {
      errno = 0;
      long v = strtol(p, &e, 10);
      if (e == p || errno) return NULL;
      iv[t] = (int)v;
    } else {
      double v = strtod(p, &e);
      if (e == p) return NULL;
      dv[t] = v;
    }
    p = e;
    while (p < end && is_ws(*p)) p++;
  }
  return p;
}

/* The same parse in parallel. The text from p is cut into chunks at
 * whitespace; each chunk's tokens are counted, a prefix sum gives every
 * token its index, and the chunks holding tokens < count parse them into
 * place. A token that strtol / strtod does not consume whole (e.g. "1-2",
 * which the serial loop reads as two numbers) sends the whole block to the
 * serial parse, so the result is always the serial one. */
static const char *parse_block(const char *p, const char *end, size_t count, int as_int, int *iv,
                               double *dv)
{
  const size_t len = (size_t)(end - p);
  const int nthr = omp_get_max_threads();
  const int nch = len < ((size_t)1 << 20) || nthr < 2 ? 1 : 16 * nthr;
  if (nch == 1 || count == 0) return parse_serial(p, end, count, as_int, iv, dv);
  const char **cb = (const char **)malloc((size_t)(nch + 1) * sizeof(char *));
  size_t *cnt = (size_t *)calloc((size_t)nch + 1, sizeof(size_t));
  if (!cb || !cnt) {
    free(cb);
    free(cnt);
    return parse_serial(p, end, count, as_int, iv, dv);
  }
  /* chunk starts: at a token start (after whitespace), or the end */
  cb[0] = p;
  for (int c = 1; c <= nch; c++) {
    const char *q = p + len / nch * (size_t)c;
    if (c == nch) q = end;
    while (q < end && !is_ws(*q)) q++;
    while (q < end && is_ws(*q)) q++;
    cb[c] = q < cb[c - 1] ? cb[c - 1] : q;
  }
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nch; c++) {
    size_t n = 0;
    int in = 0;
    for (const char *q = cb[c]; q < cb[c + 1]; q++) {
      const int w = is_ws(*q);
      if (!w && !in) n++;
      in = !w;
    }
    cnt[c + 1] = n;
  }
  for (int c = 0; c < nch; c++) cnt[c + 1] += cnt[c];
  if (cnt[nch] < count) {  /* not enough tokens: let the serial parse report it */
    free(cb);
    free(cnt);
    return parse_serial(p, end, count, as_int, iv, dv);
  }
  int ok = 1;
  const char *after = NULL;
#pragma omp parallel for schedule(dynamic, 1) reduction(& : ok)
  for (int c = 0; c < nch; c++) {
    if (cnt[c] >= count) continue;
    size_t t = cnt[c];
    const char *q = cb[c];
    while (q < cb[c + 1] && is_ws(*q)) q++;
    while (q < cb[c + 1] && t < count && ok) {
      char *e;
      if (as_int) {
        errno = 0;
        long v = strtol(q, &e, 10);
        if (e == q || errno) ok = 0;
        else iv[t] = (int)v;
      } else {
        double v = strtod(q, &e);
        if (e == q) ok = 0;
        else dv[t] = v;
      }
      if (!ok) break;
      if (e < end && !is_ws(*e)) {  /* token not consumed whole */
        ok = 0;
        break;
      }
      q = e;
      while (q < end && is_ws(*q)) q++;
      if (++t == count) after = q;  /* one chunk holds the last token */
    }
  }
  free(cb);
  free(cnt);
  if (!ok || !after) return parse_serial(p, end, count, as_int, iv, dv);
  return after;
}

int pdfs_vti_read(const char *path, pdfs_vti *out)
{
  memset(out, 0, sizeof *out);
  size_t len = 0;
  char *buf = slurp(path, &len);
  if (!buf) return -1;
  const char *end = buf + len;
  /* vti_get_dimensions (vti_parser.c:85-97) */
  double d[6];
  int found = 0;
  for (const char *p = buf; p < end;) {
    const char *q = fgets511(p, end);
    size_t L = (size_t)(q - p);
    char line[512];
    memcpy(line, p, L);
    line[L] = 0;
    if (strstr(line, "ImageData WholeExtent")) {
      int got = sscanf(line,
                       "    <ImageData WholeExtent=\"0 %lf 0 %lf 0 %lf\" "
                       "Spacing=\"%lf %lf %lf\">",
                       &d[0], &d[1], &d[2], &d[3], &d[4], &d[5]);
      if (got != 6) {
        free(buf);
        return -2;
      }
      found = 1;
      break;
    }
    p = q;
  }
  if (!found) {
    free(buf);
    return -2;
  }
  /* phi3D.c:57-62; extents beyond 2^30 per axis or 2^36 padded points are
   * malformed (they could not be held, and their products would overflow) */
  for (int k = 0; k < 3; k++)
    if (!(d[k] >= 0.0 && d[k] < (double)(1 << 30))) {
      free(buf);
      return -2;
    }
  int X = (int)d[0] + 1, Y = (int)d[1] + 1, Z = (int)d[2] + 1;
  if (X < 1 || Y < 1 || Z < 1 ||
      (double)(X + 2) * (double)(Y + 2) * (double)(Z + 2) > (double)(1ull << 36)) {
    free(buf);
    return -2;
  }
  size_t np = (size_t)(X + 2) * (Y + 2) * (Z + 2);
  int *loc = (int *)malloc(np * sizeof(int));
  double *dist = (double *)malloc(np * sizeof(double));
  if (!loc || !dist) {
    free(loc);
    free(dist);
    free(buf);
    return -3;
  }
  /* vti_get_data (vti_parser.c:112-180): the interior values in order,
   * parsed in parallel, then placed into the padded arrays */
  const char *p = buf;
  for (int k = 0; k < 6; k++) p = fgets511(p, end);
  const int gx = X + 2, gy = Y + 2, gz = Z + 2;
  const size_t ni = (size_t)X * Y * Z;
  int *iv = (int *)malloc(ni * sizeof(int));
  double *dv = (double *)malloc(ni * sizeof(double));
  if (!iv || !dv) {
    free(iv);
    free(dv);
    free(loc);
    free(dist);
    free(buf);
    return -3;
  }
  p = parse_block(p, end, ni, 1, iv, NULL);
  if (p) {
    for (int k = 0; k < 2; k++) p = fgets511(p, end);
    p = parse_block(p, end, ni, 0, NULL, dv);
  }
  if (!p) {
    free(iv);
    free(dv);
    goto bad;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < gz; i++)
    for (int j = 0; j < gy; j++) {
      size_t t = ((size_t)i * gy + j) * gx;
      for (int k = 0; k < gx; k++, t++) {
        if (k == 0 || k == gx - 1 || j == 0 || j == gy - 1 || i == 0 || i == gz - 1) {
          loc[t] = -1;
          dist[t] = INFINITY;
        } else {
          const size_t q = ((size_t)(i - 1) * Y + (j - 1)) * X + (k - 1);
          loc[t] = iv[q];
          dist[t] = dv[q];
        }
      }
    }
  free(iv);
  free(dv);
  free(buf);
  out->x = X;
  out->y = Y;
  out->z = Z;
  out->dx = d[3];
  out->dy = d[4];
  out->dz = d[5];
  out->location = loc;
  out->distance = dist;
  return 0;
bad:
  free(loc);
  free(dist);
  free(buf);
  return -4;
}

void pdfs_vti_free(pdfs_vti *v)
{
  free(v->location);
  free(v->distance);
  v->location = NULL;
  v->distance = NULL;
}

/* file_writer.c:152-190 on the padded array (gx, gy, gz = padded dims). */
int pdfs_vti_write(const char *path, const double *data, int gx, int gy, int gz, double dx, double dy,
              double dz)
{
  FILE *fp = fopen(path, "w");
  if (!fp) return -1;
  fprintf(fp, "<?xml version=\"1.0\"?>\n");
  fprintf(fp, "<VTKFile type=\"ImageData\" version=\"0.1\" "
              "byte_order=\"LittleEndian\">\n");
  fprintf(fp, "\t<ImageData WholeExtent=\"0 %d 0 %d 0 %d\" Origin=\"0 0 0\" "
              "Spacing=\" %f %f %f  \">\n",
          gx - 1, gy - 1, gz - 1, dx, dy, dz);
  fprintf(fp, "\t\t<Piece Extent=\"0 %d 0 %d 0 %d\">\n", gx - 1, gy - 1, gz - 1);
  fprintf(fp, "\t\t\t<PointData Scalars=\"Distance Field\">\n");
  fprintf(fp, "\t\t\t\t<DataArray type=\"Float32\" Name=\"Distance Field\" "
              "format=\"ascii\">\n");
  const size_t plane = (size_t)gx * gy;
  int rc = 0;
  /* format planes in parallel chunks, write them in order */
  const int chunk = 16;
  char **bufs = (char **)calloc((size_t)chunk, sizeof(char *));
  size_t *lens = (size_t *)calloc((size_t)chunk, sizeof(size_t));
  if (!bufs || !lens) rc = -2;
  for (int i0 = 0; rc == 0 && i0 < gz; i0 += chunk) {
    int n = gz - i0 < chunk ? gz - i0 : chunk;
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int c = 0; c < n; c++) {
      /* "%f  " is at most ~330 chars for huge values; typical < 16 */
      size_t cap = plane * 24 + (size_t)gy + 2, used = 0;
      char *b = (char *)malloc(cap);
      if (!b) {
        err |= 1;
        continue;
      }
      const double *t = data + (size_t)(i0 + c) * plane;
      for (int j = 0; j < gy; j++) {
        for (int k = 0; k < gx; k++) {
          if (cap - used < 400) {
            cap *= 2;
            char *nb = (char *)realloc(b, cap);
            if (!nb) {
              err |= 1;
              break;
            }
            b = nb;
          }
          used += (size_t)snprintf(b + used, cap - used, "%f  ", *(t++));
        }
        b[used++] = '\n';
      }
      b[used++] = '\n';
      bufs[c] = b;
      lens[c] = used;
    }
    if (err) rc = -2;
    for (int c = 0; c < n; c++) {
      if (bufs[c] && rc == 0 && fwrite(bufs[c], 1, lens[c], fp) != lens[c]) rc = -3;
      free(bufs[c]);
      bufs[c] = NULL;
    }
  }
  free(bufs);
  free(lens);
  fprintf(fp, "\t\t\t\t</DataArray>\n");
  fprintf(fp, "\t\t\t</PointData>\n");
  fprintf(fp, "\t\t\t<CellData>\n");
  fprintf(fp, "\t\t\t</CellData>\n");
  fprintf(fp, "\t\t</Piece>\n");
  fprintf(fp, "\t</ImageData>\n");
  fprintf(fp, "</VTKFile>\n");
  if (fclose(fp) != 0 && rc == 0) rc = -3;
  return rc;
}
