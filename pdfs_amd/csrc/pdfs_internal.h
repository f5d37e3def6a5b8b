/* Internal helpers shared by the host C modules of libpdfs. */
#ifndef PDFS_INTERNAL_H
#define PDFS_INTERNAL_H

#include <stdint.h>

#include "pdfs.h"
#include "pdfs_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

/* set the thread's last-error message (pdfs_last_error) and return code */
int pdfsi_fail(int code, const char *fmt, ...);

/* nonzero unless every spacing is finite and positive */
int pdfsi_bad_h(double dx, double dy, double dz);

/* FP64 residual judge of n dense FP32 fields (dims[3i..3i+2] = x, y, z) in
 * one launch on the context's stream; flagged[i] = failing points of field
 * i; flags[i] (optional, may be NULL) receive one bit per point. slow (or
 * NULL for unit speed): dense FP32 slowness per field. */
int pdfsi_judge_batch(pdfs_ctx *c, int n, const float *const *v, const float *const *init,
                      const float *const *slow, const int *dims, uint32_t *const *flags,
                      double dx, double dy, double dz, long long *flagged);

/* FP64 polish of FP32 fields with an init (those without are left alone):
 * passes of u = min(init, G64(u)) until the last ordering of a pass changes no
 * value by more than o->polish_tol * hmin, the result rounded back to FP32; adds to
 * s->polish_* and s->ms_polish. slows (or NULL): per-field FP32 slowness of
 * |grad u| = s. */
int pdfsi_polish(pdfs_ctx *c, int nf, float *const *fields, const float *const *inits,
                 const float *const *slows, const int *dims, double dx, double dy, double dz,
                 const pdfs_opts *o, pdfs_stats *s);

/* pdfs_solve_batch_f32_dev with per-field outcomes: field_flagged[f] = the
 * final judge's failing points of field f (0 when certified during the
 * passes or without an init), -1 when the field did not converge. */
int pdfsi_batch_f32(pdfs_ctx *c, int nf, float *const *fields, const float *const *inits,
                    const int *dims, double dx, double dy, double dz, const pdfs_opts *o,
                    pdfs_stats *st, int *field_passes, long long *field_flagged);

/* FP32 spacing constants in the oracle's operation order. */
void pdfsi_make_sp32(double dx, double dy, double dz, pdfs_sp32 *s);
/* default column variant of an FP32 sweep of nf fields (x, y of the first) */
int pdfsi_default_tile(int nf, int x, int y);
/* Ticket order of an overlapped segment (orderings sn0..sn1 in one launch):
 * nf fields with tnx[f] x tny[f] column tiles and fz[f] planes; nc active
 * columns cols[3 i] = (field, a, b) in the virtual frame; TX x TY tiles;
 * grid concurrent columns. Writes nc (sn1 - sn0 + 1) tickets of 5 ints
 * (slot, field, a, b, physical column: the field's base plus pa * tny + pb)
 * in a topological order of their dependencies (upwind in the slot; itself
 * and its physical neighbours in the slot before; all of the slot two
 * before), earliest simulated ready time first. Returns the tickets written
 * (nc (sn1 - sn0 + 1) when complete), -1 when out of memory. Host only. */
int pdfsi_ov_order(int nf, const int *tnx, const int *tny, const int *fz, int nc,
                   const int *cols, int TX, int TY, int sn0, int sn1, int grid, int *out);

/* Overlapped passes (PDFS_OVERLAP_*) of the FP64 reference-arithmetic
 * solves that take no options (pdfs_solve_ref64_dev, run_fsm,
 * pdfs_calc_dist_field); PDFS_OVERLAP_AUTO when the context is made. */
int pdfsi_set_overlap64(pdfs_ctx *c, int overlap);

/* Workgroups of the column kernel (tile variant, sweep mode) resident on
 * the device at once: the persistent grid of a launch. Needs the GPU. */
int pdfsi_col_resident(int tile, int mode);

#ifdef __cplusplus
}
#endif
#endif
