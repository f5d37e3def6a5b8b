/*
 * pdfs_vti.h — VTI input/output of the PDFS CLI (C-ABI, host only).
 *
 * pdfs_vti_read  replaces single_gpu_cuda/src/vti_parser.c:85-180
 *                (vti_get_dimensions + vti_get_data, with the ghost fill)
 *                and the allocation of create_phi_function (phi3D.c:50-82).
 * pdfs_vti_write replaces single_gpu_cuda/src/file_writer.c:100-190
 *                (file_generate + file_gen_vti).
 */
#ifndef PDFS_VTI_H
#define PDFS_VTI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int x, y, z;        /* interior point counts */
  double dx, dy, dz;  /* spacing */
  int *location;      /* padded (x+2)(y+2)(z+2), ghost = -1 */
  double *distance;   /* padded, ghost = +inf */
} pdfs_vti;

/* 0 on success; -1 open, -2 header (missing, or extents of 2^30 or more per
 * axis or over 2^36 padded points), -3 memory, -4 data. */
int pdfs_vti_read(const char *path, pdfs_vti *out);
void pdfs_vti_free(pdfs_vti *v);
/* Writes the padded array of padded dims gx, gy, gz; 0 on success. */
int pdfs_vti_write(const char *path, const double *data, int gx, int gy, int gz, double dx,
                   double dy, double dz);

#ifdef __cplusplus
}
#endif
#endif
