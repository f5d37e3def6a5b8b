/*
 * NetCDF-4 field files through the HDF5 C library (include/pdfs_nc.h).
 *
 * A NetCDF-4 file is an HDF5 file in which a dimension with a coordinate
 * variable is a dataset of that name marked as an HDF5 dimension scale
 * (attributes CLASS and NAME, plus _Netcdf4Dimid with the dimension's id),
 * a variable is a dataset whose DIMENSION_LIST attribute references its
 * dimension scales, and attributes are HDF5 attributes. Links and
 * attributes are created with creation order tracked and indexed, as the
 * NetCDF library does. This file writes and reads exactly the objects the
 * reference's ncwrap.c creates and reads:
 *   ncPutDim (ncwrap.c:156-225): dims x, y, z (defined in that order, ids
 *     0, 1, 2); float coordinate variables x, y, z with values delta * k and
 *     a double attribute "delta"; Flags (NC_INT) and Distance (NC_DOUBLE)
 *     over (z, y, x);
 *   ncGetDim (ncwrap.c:121-141): the lengths of the three dimensions and the
 *     "delta" attribute of the variables with ids 0, 1, 2, which in files of
 *     this layout are the coordinate variables x, y, z (read here by name).
 */
#include "pdfs_nc.h"

#include <hdf5.h>
#include <hdf5_hl.h>
#include <stdlib.h>
#include <string.h>

static const char *AXES[3] = {"x", "y", "z"};

/* length of a 1-D dataset and its scalar double attribute "delta" */
static int read_axis(hid_t f, const char *name, int *len, double *delta)
{
  int rc = -2;
  hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
  if (d < 0) return -2;
  hid_t sp = H5Dget_space(d);
  hsize_t n = 0;
  if (sp >= 0 && H5Sget_simple_extent_ndims(sp) == 1 && H5Sget_simple_extent_dims(sp, &n, NULL) == 1 &&
      n > 0 && n < (hsize_t)1 << 30) {
    hid_t a = H5Aopen(d, "delta", H5P_DEFAULT);
    if (a >= 0) {
      hid_t as = H5Aget_space(a);
      if (as >= 0 && H5Sget_simple_extent_npoints(as) == 1 &&
          H5Aread(a, H5T_NATIVE_DOUBLE, delta) >= 0) {
        *len = (int)n;
        rc = 0;
      }
      if (as >= 0) H5Sclose(as);
      H5Aclose(a);
    }
  }
  if (sp >= 0) H5Sclose(sp);
  H5Dclose(d);
  return rc;
}

/* a (z, y, x) dataset of the given shape, read as mem_type into buf */
static int read_var(hid_t f, const char *name, const hsize_t want[3], hid_t mem_type, void *buf)
{
  int rc = -2;
  hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
  if (d < 0) return -2;
  hid_t sp = H5Dget_space(d);
  hsize_t n[3] = {0, 0, 0};
  if (sp >= 0 && H5Sget_simple_extent_ndims(sp) == 3 && H5Sget_simple_extent_dims(sp, n, NULL) == 3 &&
      n[0] == want[0] && n[1] == want[1] && n[2] == want[2] &&
      H5Dread(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) >= 0)
    rc = 0;
  if (sp >= 0) H5Sclose(sp);
  H5Dclose(d);
  return rc;
}

int pdfs_nc_read(const char *path, pdfs_nc_field *out)
{
  if (!path || !out) return -2;
  memset(out, 0, sizeof *out);
  H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  int len[3];
  double dlt[3];
  int rc = 0;
  for (int i = 0; i < 3 && !rc; i++) rc = read_axis(f, AXES[i], &len[i], &dlt[i]);
  /* each axis < 2^30 (read_axis); at most 2^36 points in all, so neither
   * this size nor the padded one overflows */
  if (!rc && (double)len[0] * (double)len[1] * (double)len[2] > (double)(1ull << 36)) rc = -2;
  if (!rc) {
    const size_t n = (size_t)len[0] * len[1] * len[2];
    const hsize_t shape[3] = {(hsize_t)len[2], (hsize_t)len[1], (hsize_t)len[0]};
    out->flags = (int *)malloc(n * sizeof(int));
    out->distance = (double *)malloc(n * sizeof(double));
    if (!out->flags || !out->distance) rc = -3;
    if (!rc) rc = read_var(f, "Flags", shape, H5T_NATIVE_INT, out->flags);
    if (!rc) rc = read_var(f, "Distance", shape, H5T_NATIVE_DOUBLE, out->distance);
  }
  H5Fclose(f);
  if (rc) {
    pdfs_nc_free(out);
    return rc;
  }
  out->x = len[0];
  out->y = len[1];
  out->z = len[2];
  out->dx = dlt[0];
  out->dy = dlt[1];
  out->dz = dlt[2];
  return 0;
}

void pdfs_nc_free(pdfs_nc_field *f)
{
  if (!f) return;
  free(f->flags);
  free(f->distance);
  f->flags = NULL;
  f->distance = NULL;
}

/* one-value attribute: a 1-element array, as the NetCDF library stores a
 * user attribute of length 1, or a scalar, as it stores _Netcdf4Dimid */
static int put_att(hid_t obj, const char *name, hid_t file_type, hid_t mem_type, const void *v,
                   int scalar)
{
  const hsize_t one = 1;
  hid_t sp = scalar ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, &one, NULL);
  if (sp < 0) return -5;
  hid_t a = H5Acreate2(obj, name, file_type, sp, H5P_DEFAULT, H5P_DEFAULT);
  int rc = a >= 0 && H5Awrite(a, mem_type, v) >= 0 ? 0 : -5;
  if (a >= 0) H5Aclose(a);
  H5Sclose(sp);
  return rc;
}

static hid_t create_ds(hid_t f, const char *name, hid_t type, int rank, const hsize_t *dims)
{
  hid_t sp = H5Screate_simple(rank, dims, NULL);
  hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
  H5Pset_attr_creation_order(dcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  hid_t d = sp >= 0 && dcpl >= 0 ? H5Dcreate2(f, name, type, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT) : -1;
  if (dcpl >= 0) H5Pclose(dcpl);
  if (sp >= 0) H5Sclose(sp);
  return d;
}

int pdfs_nc_write(const char *path, const int *flags, const double *distance, int x, int y,
                  int z, double dx, double dy, double dz)
{
  if (!path || !flags || !distance || x < 1 || y < 1 || z < 1) return -5;
  H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
  hid_t fcpl = H5Pcreate(H5P_FILE_CREATE);
  H5Pset_link_creation_order(fcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  H5Pset_attr_creation_order(fcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, fcpl, H5P_DEFAULT);
  H5Pclose(fcpl);
  if (f < 0) return -1;
  const int len[3] = {x, y, z};
  const double dlt[3] = {dx, dy, dz};
  hid_t axis[3] = {-1, -1, -1};
  int rc = 0;
  /* ncPutDim: dimensions x, y, z with coordinate variables, values delta * k */
  for (int i = 0; i < 3 && !rc; i++) {
    const hsize_t n = (hsize_t)len[i];
    axis[i] = create_ds(f, AXES[i], H5T_IEEE_F32LE, 1, &n);
    float *v = (float *)malloc((size_t)len[i] * sizeof(float));
    if (axis[i] < 0 || !v) rc = -5;
    for (int k = 0; k < len[i] && !rc; k++) v[k] = (float)(dlt[i] * k);
    if (!rc && H5Dwrite(axis[i], H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0) rc = -5;
    free(v);
    if (!rc && H5DSset_scale(axis[i], AXES[i]) < 0) rc = -5;
    if (!rc) rc = put_att(axis[i], "_Netcdf4Dimid", H5T_STD_I32LE, H5T_NATIVE_INT, &i, 1);
    if (!rc) rc = put_att(axis[i], "delta", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &dlt[i], 0);
  }
  /* Flags and Distance over (z, y, x) */
  const hsize_t shape[3] = {(hsize_t)z, (hsize_t)y, (hsize_t)x};
  const char *vname[2] = {"Flags", "Distance"};
  const hid_t ftype[2] = {H5T_STD_I32LE, H5T_IEEE_F64LE};
  const hid_t mtype[2] = {H5T_NATIVE_INT, H5T_NATIVE_DOUBLE};
  const void *data[2] = {flags, distance};
  for (int v = 0; v < 2 && !rc; v++) {
    hid_t d = create_ds(f, vname[v], ftype[v], 3, shape);
    if (d < 0 || H5Dwrite(d, mtype[v], H5S_ALL, H5S_ALL, H5P_DEFAULT, data[v]) < 0 ||
        H5DSattach_scale(d, axis[2], 0) < 0 || H5DSattach_scale(d, axis[1], 1) < 0 ||
        H5DSattach_scale(d, axis[0], 2) < 0)
      rc = -5;
    if (d >= 0) H5Dclose(d);
  }
  for (int i = 0; i < 3; i++)
    if (axis[i] >= 0) H5Dclose(axis[i]);
  if (H5Fclose(f) < 0 && !rc) rc = -5;
  return rc;
}
