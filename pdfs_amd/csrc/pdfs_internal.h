/* Internal helpers shared by the host C modules of libpdfs. */
#ifndef PDFS_INTERNAL_H
#define PDFS_INTERNAL_H

/* set the thread's last-error message (pdfs_last_error) and return code */
int pdfsi_fail(int code, const char *fmt, ...);

#endif
