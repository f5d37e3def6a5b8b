// One ordering of the FP32 column sweep with no device allocation: the
// sweep's scratch (launch state, field and ticket tables, change counters,
// the epoch-tagged edge entries) lives in one file-scope device array per
// GPU instead of a context's allocations. For callers that cannot allocate:
// code under HIP graph capture, allocation-free serving loops.
//
// Every call on a GPU shares that GPU's array, which keeps the last call's
// contents; calls that overlap on one GPU race (issue them on one stream).
// Edge entries carry the epoch of the launch that wrote them, and the epoch
// in the array only grows, so entries of earlier calls are never taken for
// current ones and the edge part needs no clearing.
#include <hip/hip_runtime.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include "pdfs_internal.h"
#include "pdfs_kernels.h"

// Layout of the scratch array (bytes): launch state and counters, the field
// table entry, the ticket table, then the edge entries.
#define NA_BYTES (64u << 20)
#define NA_STATE 0u          // 16 x u32 launch state (err at [8]), changed, streak[2]
#define NA_FIELD 256u        // one pdfsk_field
#define NA_COLS 512u         // ticket table: 3 ints per column
#define NA_COLS_MAX 87381u   // columns (1 MiB of tickets)
#define NA_EDGE ((1u << 20) + 1024u)  // after the ticket table, 256-byte aligned
static_assert(NA_COLS + 3u * 4u * NA_COLS_MAX <= NA_EDGE && NA_EDGE % 256u == 0u,
              "scratch layout");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

#define NA_MAX_DEVICES 64
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static unsigned char *g_ptr[NA_MAX_DEVICES];        // device address of scratch
static int g_resident[NA_MAX_DEVICES][PDFSK_TILE_COUNT][2];

// Per device, once, with the device current: the array's device address
// (the host-side name is only a stand-in) and the resident workgroup counts.
static int na_setup(int dev, unsigned char **ptr) {
  pthread_mutex_lock(&g_lock);
  int rc = PDFS_OK;
  if (!g_ptr[dev]) {
    void *p = NULL;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      rc = pdfsi_fail(PDFS_EHIP, "pdfs_sweep_f32_noalloc: hipGetSymbolAddress: %s",
                      hipGetErrorString(e));
    } else {
      for (int v = 0; v < PDFSK_TILE_COUNT; v++)
        for (int iso = 0; iso < 2; iso++)
          g_resident[dev][v][iso] = pdfsk_col_max_resident(
              v, iso ? PDFSK_MODE_F32_ISO : PDFSK_MODE_F32_ANISO);
      g_ptr[dev] = (unsigned char *)p;
    }
  }
  *ptr = g_ptr[dev];
  pthread_mutex_unlock(&g_lock);
  return rc;
}

// Field table, tickets and counters of one field. Tickets run in the
// expected ready-time order of the columns (key a * wx + b * wy, ties by a):
// both upwind neighbours of a column have smaller keys, so they hold earlier
// tickets. A column's rank is the count of columns before it in that order,
// summed per column index a'.
__global__ void k_na_setup(unsigned char *base, float *f, int X, int Y, int Z, int nTx, int nTy,
                           uint32_t ey_off, int wx, int wy) {
  unsigned *state = (unsigned *)(base + NA_STATE);
  int *cols = (int *)(base + NA_COLS);
  const int nc = nTx * nTy;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nc; t += gridDim.x * blockDim.x) {
    const int a = t / nTy, b = t % nTy;
    const long key = (long)a * wx + (long)b * wy;
    long rank = 0;
    for (int a2 = 0; a2 < nTx; a2++) {
      const long r = key - (long)a2 * wx;  // b2 * wy < r, or == r with a2 < a
      long cnt = r <= 0 ? 0 : (r + wy - 1) / wy;
      if (r >= 0 && r % wy == 0 && a2 < a) cnt += 1;
      rank += cnt < nTy ? cnt : nTy;
    }
    cols[3 * rank] = 0;
    cols[3 * rank + 1] = a;
    cols[3 * rank + 2] = b;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    pdfsk_field *fd = (pdfsk_field *)(base + NA_FIELD);
    fd->f = f;
    fd->init = NULL;
    fd->slow = NULL;
    fd->X = X;
    fd->Y = Y;
    fd->Z = Z;
    fd->nTx = nTx;
    fd->nTy = nTy;
    fd->ex_off = 0;
    fd->ey_off = ey_off;
    fd->fxk = 0;
    state[4] = (unsigned)nc;                 // column tickets
    if (state[1] == 0u) state[1] = 1u;       // epochs start at 1 (zeroed entries: 0)
    state[8] = 0u;                           // edge-wait timeout flag
    state[16] = 0u;                          // changed
    state[17] = 0u;                          // streak: counter, unchanged run
    state[18] = 0u;
  }
}

__global__ void k_na_finish(const unsigned char *base, unsigned long long *result) {
  const unsigned *state = (const unsigned *)(base + NA_STATE);
  result[0] += state[16];
  result[1] |= state[8];
}

extern "C" int pdfs_sweep_f32_noalloc(float *field, int x, int y, int z, double dx, double dy,
                                      double dz, int sn, int tile, void *stream,
                                      unsigned long long *result) {
  if (!field || !result || x < 1 || y < 1 || z < 1 || sn < 1 || sn > 8 || pdfsi_bad_h(dx, dy, dz))
    return pdfsi_fail(PDFS_EINVAL, "pdfs_sweep_f32_noalloc: bad arguments");
  if ((long long)x * y * z > PDFSK_MAX_POINTS)
    return pdfsi_fail(PDFS_EINVAL, "pdfs_sweep_f32_noalloc: %lld points exceed the FP32 limit",
                      (long long)x * y * z);
  if (tile < 0 || tile >= PDFSK_TILE_COUNT) tile = pdfsi_default_tile(1, x, y);
  int TX, TY;
  pdfsk_tile_dims(tile, &TX, &TY);
  const long long nTx = (x + TX - 1) / TX, nTy = (y + TY - 1) / TY, nc = nTx * nTy;
  const long long ns = (long long)z + TX + TY - 1;
  const long long ex = nc * ns * TY, ey = nc * ns * TX;
  const long long nx = ex + 2 * PDFSK_EDGE_PAD, ny = ey + 2 * PDFSK_EDGE_PAD;
  if (nc > (long long)NA_COLS_MAX || (nx + ny) * 8 > (long long)(NA_BYTES - NA_EDGE))
    return pdfsi_fail(PDFS_EINVAL,
                      "pdfs_sweep_f32_noalloc: %dx%dx%d needs %lld bytes of edge entries and %lld "
                      "columns; the scratch holds %u bytes and %u columns",
                      x, y, z, (nx + ny) * 8, nc, NA_BYTES - NA_EDGE, NA_COLS_MAX);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES)
    return pdfsi_fail(PDFS_EHIP, "pdfs_sweep_f32_noalloc: no device");
  unsigned char *base = NULL;
  int rc = na_setup(dev, &base);
  if (rc) return rc;
  pdfs_sp32 sp;
  pdfsi_make_sp32(dx, dy, dz, &sp);
  int grid = g_resident[dev][tile][sp.iso ? 1 : 0];
  if (grid <= 0) return pdfsi_fail(PDFS_EHIP, "pdfs_sweep_f32_noalloc: kernel cannot be resident");
  if (grid > nc) grid = (int)nc;
  hipStream_t st = (hipStream_t)stream;
  const int sg = (int)((nc + 255) / 256 < 1024 ? (nc + 255) / 256 : 1024);
  // the field's entries start each face part (edge_y has its own base)
  hipLaunchKernelGGL(k_na_setup, dim3(sg), dim3(256), 0, st, base, field, x, y, z, (int)nTx,
                     (int)nTy, 0u, TX + 6, TY + 6);
  if (hipGetLastError() != hipSuccess)
    return pdfsi_fail(PDFS_EHIP, "pdfs_sweep_f32_noalloc: setup launch failed");
  pdfsk_batch b;
  memset(&b, 0, sizeof b);
  b.fields = (pdfsk_field *)(base + NA_FIELD);
  b.cols = (int *)(base + NA_COLS);
  b.state = (unsigned *)(base + NA_STATE);
  b.err = b.state + 8;
  b.changed = b.state + 16;
  b.streak = b.state + 17;
  b.nfields = 1;
  b.edge_x = (uint64_t *)(base + NA_EDGE);
  b.edge_y = b.edge_x + nx;
  b.edge_x_bytes = (size_t)(nx + ny) * 8;
  b.edge_y_bytes = (size_t)ny * 8;
  const int mode = sp.iso ? PDFSK_MODE_F32_ISO : PDFSK_MODE_F32_ANISO;
  if (pdfsk_sweep_batch(&b, sn, mode, &sp, NULL, tile, grid, st))
    return pdfsi_fail(PDFS_EHIP, "pdfs_sweep_f32_noalloc: sweep launch failed");
  hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(1), 0, st, base, result);
  if (hipGetLastError() != hipSuccess)
    return pdfsi_fail(PDFS_EHIP, "pdfs_sweep_f32_noalloc: finish launch failed");
  return PDFS_OK;
}
