/*
 * NetCDF-4 field files of the reference's MPI/OpenACC CLI, read and written
 * through the HDF5 C library (lib/libpdfs_nc.so; NetCDF-4 files are HDF5
 * files with its dimension-scale conventions, and libnetcdf is not needed).
 *
 * Layout (mpi_openacc/src/ncwrap.c:121-225, process_file.c:78-120):
 *   dimensions x, y, z, each with a coordinate variable of the same name
 *   (float, values delta * k) carrying a double attribute "delta";
 *   variables Flags (int) and Distance (double), dimensions (z, y, x), so
 *   the arrays are [z][y][x] with x fastest.
 */
#ifndef PDFS_NC_H
#define PDFS_NC_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int x, y, z;          /* lengths of dimensions x, y, z */
  double dx, dy, dz;    /* their "delta" attributes */
  int *flags;           /* [z][y][x] */
  double *distance;     /* [z][y][x] */
} pdfs_nc_field;

/* Read Flags, Distance and the x / y / z lengths and deltas, as ncGetDim
 * and ncGetInt / ncGetDouble do (ncwrap.c:121-141, 283-373). Returns 0, or
 * -1 cannot open, -2 a variable, dimension or attribute is missing or has
 * another shape (or an axis of 2^30 or more, or over 2^36 points), -3 out
 * of memory. Arrays are malloc'd; pdfs_nc_free. */
int pdfs_nc_read(const char *path, pdfs_nc_field *out);
void pdfs_nc_free(pdfs_nc_field *f);

/* Write a file of that layout (ncPutDim, ncPutInt / ncPutDouble:
 * ncwrap.c:156-225, 320-404): dimensions x, y, z of lengths x, y, z,
 * coordinate variables with delta, Flags and Distance [z][y][x]. Returns 0,
 * or -1 cannot create, -5 a write failed. */
int pdfs_nc_write(const char *path, const int *flags, const double *distance, int x, int y,
                  int z, double dx, double dy, double dz);

#ifdef __cplusplus
}
#endif
#endif
