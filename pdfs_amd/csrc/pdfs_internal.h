/* Internal helpers shared by the host C modules of libpdfs. */
#ifndef PDFS_INTERNAL_H
#define PDFS_INTERNAL_H

#include <stdint.h>

#include "pdfs.h"

/* set the thread's last-error message (pdfs_last_error) and return code */
int pdfsi_fail(int code, const char *fmt, ...);

/* nonzero unless every spacing is finite and positive */
int pdfsi_bad_h(double dx, double dy, double dz);

/* FP64 residual judge of n dense FP32 fields (dims[3i..3i+2] = x, y, z) in
 * one launch on the context's stream; flagged[i] = failing points of field
 * i; flags[i] (optional, may be NULL) receive one bit per point. */
int pdfsi_judge_batch(pdfs_ctx *c, int n, const float *const *v, const float *const *init,
                      const int *dims, uint32_t *const *flags, double dx, double dy, double dz,
                      long long *flagged);

#endif
