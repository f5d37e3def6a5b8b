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
  /* phi3D.c:57-62 */
  int X = (int)d[0] + 1, Y = (int)d[1] + 1, Z = (int)d[2] + 1;
  if (X < 1 || Y < 1 || Z < 1) {
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
  /* vti_get_data (vti_parser.c:112-180) */
  const char *p = buf;
  for (int k = 0; k < 6; k++) p = fgets511(p, end);
  int gx = X + 2, gy = Y + 2, gz = Z + 2;
  size_t t = 0;
  for (int i = 0; i < gz; i++)
    for (int j = 0; j < gy; j++)
      for (int k = 0; k < gx; k++, t++) {
        if (k == 0 || k == gx - 1 || j == 0 || j == gy - 1 || i == 0 || i == gz - 1) {
          loc[t] = -1;
        } else {
          char *e;
          errno = 0;
          long v = strtol(p, &e, 10);
          if (e == p || errno) goto bad;
          loc[t] = (int)v;
          p = e;
          while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' ||
                             *p == '\f'))
            p++;
        }
      }
  for (int k = 0; k < 2; k++) p = fgets511(p, end);
  t = 0;
  for (int i = 0; i < gz; i++)
    for (int j = 0; j < gy; j++)
      for (int k = 0; k < gx; k++, t++) {
        if (k == 0 || k == gx - 1 || j == 0 || j == gy - 1 || i == 0 || i == gz - 1) {
          dist[t] = INFINITY;
        } else {
          char *e;
          double v = strtod(p, &e);
          if (e == p) goto bad;
          dist[t] = v;
          p = e;
          while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' ||
                             *p == '\f'))
            p++;
        }
      }
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
