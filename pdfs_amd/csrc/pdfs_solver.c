/*
 * Host C orchestration of the MI355X-native FSM: the public C-ABI of
 * include/pdfs.h. Calls the gfx950 kernels through the thin C layer of
 * pdfs_kernels.h; owns device scratch, epochs, tickets and convergence.
 */
#include "pdfs.h"
#include "pdfs_internal.h"
#include "pdfs_kernels.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct pdfs_ctx {
  int dev;
  hipStream_t st;
  int resident[PDFSK_TILE_COUNT][2];
  /* batched column sweep: persistent device state, edge allocations, plan */
  pdfsk_batch bt;
  size_t cap_ex, cap_ey;           /* edge entries excluding padding */
  pdfsk_field *h_fields;           /* host copy of the field table */
  int *h_cols;
  int nf, ncols, tile, cap_f, cap_cols;
  unsigned *h_changed, *prev_changed; /* pinned readback, last values */
  unsigned *h_streak;              /* per field: counter, run of unchanged orderings */
  unsigned *h_err;
  hipGraphExec_t gexec[3];         /* captured pass (orderings 1-8), halves 1-4 and 5-8 */
  long long tune_keys[32];         /* tile autotuner cache: size class -> tile */
  int tune_tiles[32], ntune;
  int g_tile[3], g_iso[3], g_grid[3];
  /* FP64 path / judge */
  unsigned long long *d_u64; /* [0] maxdiff bits */
  unsigned long long *h_pin64;
  /* batched judge: field table (host staging, device) and per-field counts */
  pdfsk_judge_field *jt_host, *jt_dev;
  unsigned long long *jc_dev, *jc_host;
  int cap_jt;
  double *d_params;
  hipEvent_t ev0, ev1;
  hipEvent_t kev[16];
};

static __thread char g_err[512];

const char *pdfs_last_error(void) { return g_err; }

static int fail(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

/* spacings of the status-returning entry points: finite and positive (the
 * comparisons are false for NaN) */
static int bad_h(double dx, double dy, double dz)
{
  return !(dx > 0.0 && dx < HUGE_VAL) || !(dy > 0.0 && dy < HUGE_VAL) ||
         !(dz > 0.0 && dz < HUGE_VAL);
}

int pdfsi_bad_h(double dx, double dy, double dz) { return bad_h(dx, dy, dz); }

int pdfsi_fail(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

void *pdfs_ctx_stream(pdfs_ctx *c) { return c ? (void *)c->st : NULL; }

#define HIPCHK(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(PDFS_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,              \
                  hipGetErrorString(e_));                                            \
  } while (0)

#define KCHK(call)                                                                   \
  do {                                                                               \
    int r_ = (call);                                                                 \
    if (r_ != 0)                                                                     \
      return fail(PDFS_EHIP, "%s:%d %s: launch failed (%d: %s)", __FILE__, __LINE__, \
                  #call, r_, hipGetErrorString((hipError_t)r_));                     \
  } while (0)

/* HIP call inside a function that owns host allocations: record the error
 * and leave through its `out` label */
#define HIPGO(call)                                                                  \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      rc = fail(PDFS_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,                \
                hipGetErrorString(e_));                                              \
      goto out;                                                                      \
    }                                                                                \
  } while (0)

void pdfs_default_opts(pdfs_opts *o)
{
  memset(o, 0, sizeof *o);
  o->precision = PDFS_F32;
  o->max_passes = 100;
  o->tol = 1e-12;
  o->judge = 1;
  o->tile = -1;
  o->graph = 0;
  o->certify = 0;
  o->certify_from = 1;
}

static int create_events(hipEvent_t *ev, int n)
{
  for (int i = 0; i < n; i++)
    if (hipEventCreate(&ev[i]) != hipSuccess) return -1;
  return 0;
}

pdfs_ctx *pdfs_ctx_create(void)
{
  pdfs_ctx *c = (pdfs_ctx *)calloc(1, sizeof *c);
  if (!c) return NULL;
  if (hipGetDevice(&c->dev) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess ||
      hipMalloc((void **)&c->bt.state, 16 * sizeof(unsigned)) != hipSuccess ||
      hipMalloc((void **)&c->d_u64, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc((void **)&c->h_err, 16 * sizeof(unsigned), 0) != hipSuccess ||
      hipHostMalloc((void **)&c->h_pin64, 4 * sizeof(unsigned long long), 0) != hipSuccess ||
      hipMalloc((void **)&c->d_params, 1024 * sizeof(double)) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      create_events(c->kev, 16) != 0) {
    fail(PDFS_EHIP, "pdfs_ctx_create: HIP initialisation failed");
    free(c);
    return NULL;
  }
  c->bt.err = c->bt.state + 8;
  /* epochs start at 1 (zeroed edge entries carry epoch 0) and only grow for
   * the lifetime of the context */
  unsigned init[16] = {0, 1, 0, 0, 0};
  if (hipMemcpy(c->bt.state, init, sizeof init, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c->d_u64, 0, 4 * sizeof(unsigned long long)) != hipSuccess) {
    free(c);
    return NULL;
  }
  for (int v = 0; v < PDFSK_TILE_COUNT; v++)
    for (int iso = 0; iso < 2; iso++) c->resident[v][iso] = pdfsk_col_max_resident(v, iso);
  c->tile = -1;
  return c;
}

void pdfs_ctx_destroy(pdfs_ctx *c)
{
  if (!c) return;
  hipStreamSynchronize(c->st);
  for (int g = 0; g < 3; g++)
    if (c->gexec[g]) hipGraphExecDestroy(c->gexec[g]);
  if (c->bt.edge_x) hipFree(c->bt.edge_x); /* edge_y lies in the same allocation */
  hipFree(c->bt.fields);
  hipFree(c->bt.cols);
  hipFree(c->bt.changed);
  hipFree(c->bt.streak);
  free(c->h_streak);
  hipFree(c->bt.state);
  hipFree(c->d_u64);
  hipFree(c->jt_dev);
  hipFree(c->jc_dev);
  free(c->jt_host);
  hipHostFree(c->jc_host);
  hipFree(c->d_params);
  hipHostFree(c->h_changed);
  hipHostFree(c->h_err);
  hipHostFree(c->h_pin64);
  free(c->prev_changed);
  free(c->h_fields);
  free(c->h_cols);
  hipEventDestroy(c->ev0);
  hipEventDestroy(c->ev1);
  for (int i = 0; i < 16; i++) hipEventDestroy(c->kev[i]);
  hipStreamDestroy(c->st);
  free(c);
}

/* FP32 spacing constants in the oracle's operation order. */
static void make_sp32(double dx, double dy, double dz, pdfs_sp32 *s)
{
  memset(s, 0, sizeof *s);
  s->iso = (dx == dy && dy == dz);
  float h = (float)dx;
  float h2 = h * h;
  s->h = h;
  s->two_h2 = 2.0f * h2;
  s->three_h2 = 3.0f * h2;
  s->third = 1.0f / 3.0f;
  s->d[0] = (float)dx;
  s->d[1] = (float)dy;
  s->d[2] = (float)dz;
  for (int i = 0; i < 3; i++) s->w[i] = 1.0f / (s->d[i] * s->d[i]);
  s->W = s->w[0] + s->w[1] + s->w[2];
}

/* Size class of a batch for the tile cache: field count and log2 buckets of
 * the mean extents. */
static long long tune_key(int nf, const int *dims)
{
  double sx = 0, sy = 0, sz = 0;
  for (int f = 0; f < nf; f++) {
    sx += dims[3 * f];
    sy += dims[3 * f + 1];
    sz += dims[3 * f + 2];
  }
  int bx = (int)log2(sx / nf + 1), by = (int)log2(sy / nf + 1), bz = (int)log2(sz / nf + 1);
  int bn = (int)log2((double)nf + 1);
  return ((long long)bn << 24) | (bx << 16) | (by << 8) | bz;
}

/* Measured on MI355X (spheres, FP32): 32x16 tiles (512 threads, one
 * workgroup per CU) sweep 512^3 in 1.67 ms per ordering against 2.14 ms for
 * 16x16 at three per CU, and a 512-field batch in 160 ms against 166 ms;
 * at 256^3 and below 16x16 is as fast or faster. */
static int default_tile(int nf, int x, int y)
{
  if (nf > 1 || (long long)x * y >= 512LL * 512) return PDFSK_TILE_32x16;
  return PDFSK_TILE_16x16;
}

/* the online tuner gives each candidate one of orderings 2..7 */
_Static_assert(PDFSK_TILE_COUNT <= 6, "tile candidates must fit orderings 2..7 of a pass");

static int tune_lookup(pdfs_ctx *c, long long key)
{
  for (int i = 0; i < c->ntune; i++)
    if (c->tune_keys[i] == key) return c->tune_tiles[i];
  return -1;
}

static void tune_store(pdfs_ctx *c, long long key, int tile)
{
  int i = c->ntune < 32 ? c->ntune++ : (int)(key & 31);
  c->tune_keys[i] = key;
  c->tune_tiles[i] = tile;
}

/* ---- batched column sweep: plans ---------------------------------------
 * A plan is a list of dense FP32 fields swept by the same launches. Column
 * tickets run in the order the columns become ready: column (a, b) of a
 * field can start about a * (TX + E) + b * (TY + E) steps after the field's
 * first column (the hyperplane skew across a tile plus E steps of edge
 * latency per boundary), so tickets are sorted by that key across all
 * active fields. Both upwind neighbours of a column have smaller keys, so
 * every column's producers hold earlier tickets and the persistent grid
 * never waits on a column that no running workgroup has taken. */
#define PDFS_TICKET_EDGE_STEPS 6
typedef struct {
  int key, f, a, b;
} col_key;
static int col_key_cmp(const void *x, const void *y)
{
  const col_key *p = (const col_key *)x, *q = (const col_key *)y;
  if (p->key != q->key) return p->key < q->key ? -1 : 1;
  if (p->f != q->f) return p->f < q->f ? -1 : 1;
  return p->a < q->a ? -1 : (p->a > q->a);
}

static int grow_dev(void **p, int *cap, int need, size_t elem)
{
  if (need <= *cap) return 0;
  hipFree(*p);
  *p = NULL;
  if (hipMalloc(p, (size_t)need * elem) != hipSuccess) return -1;
  *cap = need;
  return 0;
}

static int plan_upload_cols(pdfs_ctx *c, const int *active)
{
  int TX = 0, TY = 0;
  pdfsk_tile_dims(c->tile, &TX, &TY);
  const int wx = TX + PDFS_TICKET_EDGE_STEPS, wy = TY + PDFS_TICKET_EDGE_STEPS;
  int n_all = 0;
  for (int f = 0; f < c->nf; f++)
    if (!active || active[f]) n_all += c->h_fields[f].nTx * c->h_fields[f].nTy;
  col_key *keys = (col_key *)malloc((size_t)(n_all > 0 ? n_all : 1) * sizeof(col_key));
  if (!keys) return fail(PDFS_ENOMEM, "plan: host alloc");
  int t = 0;
  for (int f = 0; f < c->nf; f++) {
    if (active && !active[f]) continue;
    const pdfsk_field *fd = &c->h_fields[f];
    for (int a = 0; a < fd->nTx; a++)
      for (int b = 0; b < fd->nTy; b++) {
        keys[t].key = a * wx + b * wy;
        keys[t].f = f;
        keys[t].a = a;
        keys[t].b = b;
        t++;
      }
  }
  qsort(keys, (size_t)t, sizeof(col_key), col_key_cmp);
  for (int i = 0; i < t; i++) {
    c->h_cols[3 * i] = keys[i].f;
    c->h_cols[3 * i + 1] = keys[i].a;
    c->h_cols[3 * i + 2] = keys[i].b;
  }
  free(keys);
  unsigned n = (unsigned)t;
  HIPCHK(hipMemcpyAsync(c->bt.cols, c->h_cols, (size_t)(3 * (t > 0 ? t : 1)) * sizeof(int),
                        hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->bt.state + 4, &n, sizeof n, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->ncols = t;
  return PDFS_OK;
}

/* Face entries of a batch's edge allocation for a tile (one row per column
 * step, z + TX + TY - 1 steps); 0 if the batch does not fit the kernel's
 * 32-bit buffer offsets (2 GiB with padding). */
static int plan_fits(int nf, const int *dims, int tile)
{
  int TX, TY;
  pdfsk_tile_dims(tile, &TX, &TY);
  size_t ex = 0, ey = 0;
  for (int f = 0; f < nf; f++) {
    const int x = dims[3 * f], y = dims[3 * f + 1], z = dims[3 * f + 2];
    const size_t nc = (size_t)((x + TX - 1) / TX) * (size_t)((y + TY - 1) / TY);
    ex += nc * (size_t)(z + TX + TY - 1) * TY;
    ey += nc * (size_t)(z + TX + TY - 1) * TX;
  }
  return (ex + ey + 4 * PDFSK_EDGE_PAD) * 8 < (1ull << 31);
}

/* keep: re-tile the same batch mid-solve, keeping the per-field change
 * counters and unchanged-ordering runs (they do not depend on the tile) */
static int plan_build_ex(pdfs_ctx *c, int nf, float *const *fields, const int *dims, int tile,
                         int keep)
{
  if (nf < 1) return fail(PDFS_EINVAL, "empty batch");
  int TX, TY;
  pdfsk_tile_dims(tile, &TX, &TY);
  if (!plan_fits(nf, dims, tile))
    return fail(PDFS_EINVAL, "batch too large for one sweep (edge buffers exceed 2 GiB)");
  if (nf > c->nf || !c->h_fields) {
    free(c->h_fields);
    free(c->prev_changed);
    free(c->h_streak);
    c->h_fields = (pdfsk_field *)calloc((size_t)nf, sizeof(pdfsk_field));
    c->prev_changed = (unsigned *)calloc((size_t)nf, sizeof(unsigned));
    c->h_streak = (unsigned *)calloc((size_t)2 * nf, sizeof(unsigned));
    hipHostFree(c->h_changed);
    c->h_changed = NULL;
    if (!c->h_fields || !c->prev_changed || !c->h_streak ||
        hipHostMalloc((void **)&c->h_changed, (size_t)nf * sizeof(unsigned), 0) != hipSuccess)
      return fail(PDFS_ENOMEM, "plan: host alloc");
  }
  size_t ex = 0, ey = 0;
  int ncols = 0;
  for (int f = 0; f < nf; f++) {
    int x = dims[3 * f], y = dims[3 * f + 1], z = dims[3 * f + 2];
    if (!fields[f] || x < 1 || y < 1 || z < 1)
      return fail(PDFS_EINVAL, "field %d: bad pointer or dims", f);
    if ((long long)x * y * z > PDFSK_MAX_POINTS)
      return fail(PDFS_EINVAL, "field %d: %lld points exceed the FP32 sweep limit %lld", f,
                  (long long)x * y * z, (long long)PDFSK_MAX_POINTS);
    pdfsk_field *fd = &c->h_fields[f];
    fd->f = fields[f];
    fd->X = x;
    fd->Y = y;
    fd->Z = z;
    fd->nTx = (x + TX - 1) / TX;
    fd->nTy = (y + TY - 1) / TY;
    fd->ex_off = (uint32_t)ex;
    fd->ey_off = (uint32_t)ey;
    fd->pad = 0;
    size_t nc = (size_t)fd->nTx * fd->nTy;
    /* one row of face entries per column step (z + TX + TY - 1 steps) */
    ex += nc * (size_t)(z + TX + TY - 1) * TY;
    ey += nc * (size_t)(z + TX + TY - 1) * TX;
    ncols += (int)nc;
  }
  if (ex > c->cap_ex || ey > c->cap_ey) {
    /* one allocation: the x-face part, then the y-face part, each with a
     * padding of PDFSK_EDGE_PAD entries on both sides; the kernel's single
     * buffer descriptor covers both, so one instruction moves a column's
     * west and south entries */
    if (c->bt.edge_x) hipFree(c->bt.edge_x);
    c->bt.edge_x = c->bt.edge_y = NULL;
    const size_t cx = ex > c->cap_ex ? ex : c->cap_ex, cy = ey > c->cap_ey ? ey : c->cap_ey;
    const size_t nx = cx + 2 * PDFSK_EDGE_PAD, ny = cy + 2 * PDFSK_EDGE_PAD;
    HIPCHK(hipMalloc((void **)&c->bt.edge_x, (nx + ny) * sizeof(uint64_t)));
    HIPCHK(hipMemsetAsync(c->bt.edge_x, 0, (nx + ny) * sizeof(uint64_t), c->st));
    c->bt.edge_y = c->bt.edge_x + nx;
    c->bt.edge_x_bytes = (nx + ny) * sizeof(uint64_t);
    c->bt.edge_y_bytes = ny * sizeof(uint64_t);
    c->cap_ex = cx;
    c->cap_ey = cy;
  }
  int cap_f = c->cap_f, cap_s = 2 * c->cap_f;
  if (grow_dev((void **)&c->bt.fields, &c->cap_f, nf, sizeof(pdfsk_field)) ||
      grow_dev((void **)&c->bt.changed, &cap_f, nf, sizeof(unsigned)) ||
      grow_dev((void **)&c->bt.streak, &cap_s, 2 * nf, sizeof(unsigned)) ||
      grow_dev((void **)&c->bt.cols, &c->cap_cols, 3 * ncols, sizeof(int)))
    return fail(PDFS_ENOMEM, "plan: device alloc");
  free(c->h_cols);
  c->h_cols = (int *)malloc((size_t)3 * ncols * sizeof(int));
  if (!c->h_cols) return fail(PDFS_ENOMEM, "plan: host alloc");
  c->nf = nf;
  c->tile = tile;
  HIPCHK(hipMemcpyAsync(c->bt.fields, c->h_fields, (size_t)nf * sizeof(pdfsk_field),
                        hipMemcpyHostToDevice, c->st));
  if (!keep) {
    HIPCHK(hipMemsetAsync(c->bt.changed, 0, (size_t)nf * sizeof(unsigned), c->st));
    HIPCHK(hipMemsetAsync(c->bt.streak, 0, (size_t)2 * nf * sizeof(unsigned), c->st));
    memset(c->prev_changed, 0, (size_t)nf * sizeof(unsigned));
  }
  c->bt.nfields = nf;
  for (int g = 0; g < 3; g++)
    if (c->gexec[g]) {
      hipGraphExecDestroy(c->gexec[g]);
      c->gexec[g] = NULL;
    }
  return plan_upload_cols(c, NULL);
}

static int plan_build(pdfs_ctx *c, int nf, float *const *fields, const int *dims, int tile)
{
  return plan_build_ex(c, nf, fields, dims, tile, 0);
}

static int launch_grid(pdfs_ctx *c, int iso)
{
  int grid = c->resident[c->tile][iso];
  if (grid <= 0) return -1;
  return grid > c->ncols ? (c->ncols > 0 ? c->ncols : 1) : grid;
}

/* Orderings sn0..sn1 of the current plan, optionally profiled per launch
 * or replayed from a captured graph (a whole pass, or one of its halves). */
static int run_orderings(pdfs_ctx *c, const pdfs_sp32 *sp, int sn0, int sn1, int profile,
                         int graph, double *kernel_ms)
{
  int grid = launch_grid(c, sp->iso);
  if (grid <= 0) return fail(PDFS_EHIP, "column kernel cannot be resident");
  const int gi = sn0 == 1 && sn1 == 8 ? 0 : sn0 == 1 && sn1 == 4 ? 1 : sn0 == 5 && sn1 == 8 ? 2 : -1;
  if (graph && gi >= 0 && !profile) {
    if (!c->gexec[gi] || c->g_tile[gi] != c->tile || c->g_iso[gi] != sp->iso ||
        c->g_grid[gi] != grid) {
      if (c->gexec[gi]) {
        hipGraphExecDestroy(c->gexec[gi]);
        c->gexec[gi] = NULL;
      }
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
      for (int sn = sn0; sn <= sn1; sn++)
        if (pdfsk_sweep_batch_f32(&c->bt, sn, sp, c->tile, grid, c->st)) {
          hipStreamEndCapture(c->st, &g);
          return fail(PDFS_EHIP, "launch during capture failed");
        }
      HIPCHK(hipStreamEndCapture(c->st, &g));
      hipError_t e = hipGraphInstantiate(&c->gexec[gi], g, NULL, NULL, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) return fail(PDFS_EHIP, "graph instantiate: %s", hipGetErrorString(e));
      c->g_tile[gi] = c->tile;
      c->g_iso[gi] = sp->iso;
      c->g_grid[gi] = grid;
    }
    HIPCHK(hipGraphLaunch(c->gexec[gi], c->st));
    return PDFS_OK;
  }
  for (int sn = sn0; sn <= sn1; sn++) {
    if (profile) HIPCHK(hipEventRecord(c->kev[2 * (sn - 1)], c->st));
    KCHK(pdfsk_sweep_batch_f32(&c->bt, sn, sp, c->tile, grid, c->st));
    if (profile) HIPCHK(hipEventRecord(c->kev[2 * (sn - 1) + 1], c->st));
  }
  if (profile) {
    HIPCHK(hipEventSynchronize(c->kev[2 * (sn1 - 1) + 1]));
    for (int sn = sn0; sn <= sn1; sn++) {
      float tk = 0.f;
      hipEventElapsedTime(&tk, c->kev[2 * (sn - 1)], c->kev[2 * (sn - 1) + 1]);
      *kernel_ms += tk;
    }
  }
  return PDFS_OK;
}

/* Per-field points changed since the last read; returns the timeout flag. */
static int read_changed(pdfs_ctx *c, long long *delta, unsigned *err)
{
  HIPCHK(hipMemcpyAsync(c->h_changed, c->bt.changed, (size_t)c->nf * sizeof(unsigned),
                        hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(c->h_err, c->bt.err, sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(c->h_streak, c->bt.streak, (size_t)2 * c->nf * sizeof(unsigned),
                        hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int f = 0; f < c->nf; f++) {
    delta[f] = (long long)(unsigned)(c->h_changed[f] - c->prev_changed[f]);
    c->prev_changed[f] = c->h_changed[f];
  }
  *err = c->h_err[0];
  if (*err) {
    unsigned z = 0;
    hipMemcpyAsync(c->bt.err, &z, sizeof z, hipMemcpyHostToDevice, c->st);
    hipStreamSynchronize(c->st);
  }
  return PDFS_OK;
}

int pdfs_sweep_f32_dev(pdfs_ctx *c, float *field, int x, int y, int z, double dx, double dy,
                       double dz, int sn, int tile, long long *changed)
{
  if (!c || !field || x < 1 || y < 1 || z < 1 || sn < 1 || sn > 8 || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_sweep_f32_dev: bad arguments");
  if (tile < 0 || tile >= PDFSK_TILE_COUNT) tile = default_tile(1, x, y);
  pdfs_sp32 sp;
  make_sp32(dx, dy, dz, &sp);
  int dims[3] = {x, y, z}, rc;
  if ((rc = plan_build(c, 1, &field, dims, tile))) return rc;
  if ((rc = run_orderings(c, &sp, sn, sn, 0, 0, NULL))) return rc;
  long long d = 0;
  unsigned err = 0;
  if ((rc = read_changed(c, &d, &err))) return rc;
  if (err) return fail(PDFS_ETIMEOUT, "column pipeline: edge wait timed out");
  if (changed) *changed += d;
  return PDFS_OK;
}

/* FP64 residual judge of n fields in one launch; flagged[i] = failing points
 * of field i. flags[i] (optional) receive one bit per point. */
int pdfsi_judge_batch(pdfs_ctx *c, int n, const float *const *v, const float *const *init,
                      const int *dims, uint32_t *const *flags, double dx, double dy, double dz,
                      long long *flagged)
{
  if (!c || n < 1 || !v || !init || !dims) return fail(PDFS_EINVAL, "judge: bad arguments");
  if (n + 1 > c->cap_jt) {
    HIPCHK(hipStreamSynchronize(c->st));
    hipFree(c->jt_dev);
    hipFree(c->jc_dev);
    free(c->jt_host);
    hipHostFree(c->jc_host);
    c->jt_dev = NULL;
    c->jc_dev = NULL;
    c->jc_host = NULL;
    c->cap_jt = 0;
    c->jt_host = (pdfsk_judge_field *)malloc((size_t)(n + 1) * sizeof(pdfsk_judge_field));
    if (!c->jt_host) return fail(PDFS_ENOMEM, "judge: host alloc");
    HIPCHK(hipMalloc((void **)&c->jt_dev, (size_t)(n + 1) * sizeof(pdfsk_judge_field)));
    HIPCHK(hipMalloc((void **)&c->jc_dev, (size_t)(n + 1) * sizeof(unsigned long long)));
    HIPCHK(hipHostMalloc((void **)&c->jc_host, (size_t)(n + 1) * sizeof(unsigned long long), 0));
    c->cap_jt = n + 1;
  }
  long long tiles = 0;
  for (int i = 0; i < n; i++) {
    const int x = dims[3 * i], y = dims[3 * i + 1], z = dims[3 * i + 2];
    if (!v[i] || !init[i] || x < 1 || y < 1 || z < 1)
      return fail(PDFS_EINVAL, "judge: field %d: bad pointer or dims", i);
    pdfsk_judge_field *t = &c->jt_host[i];
    t->v = v[i];
    t->init = init[i];
    t->flags = flags ? flags[i] : NULL;
    t->X = x;
    t->Y = y;
    t->Z = z;
    t->tile0 = (int)tiles;
    tiles += pdfsk_judge_tiles(x, y, z);
    if (tiles >= (1LL << 31) - 8) return fail(PDFS_EINVAL, "judge: batch too large");
    if (t->flags)
      KCHK(pdfsk_fill_u32(t->flags, 0u, ((size_t)x * y * z + 31) / 32, c->st));
  }
  memset(&c->jt_host[n], 0, sizeof c->jt_host[n]);
  c->jt_host[n].tile0 = (int)tiles;
  HIPCHK(hipMemcpyAsync(c->jt_dev, c->jt_host, (size_t)(n + 1) * sizeof(pdfsk_judge_field),
                        hipMemcpyHostToDevice, c->st));
  KCHK(pdfsk_fill_u32((uint32_t *)c->jc_dev, 0u, (size_t)2 * n, c->st));
  KCHK(pdfsk_judge_batch_f32(c->jt_dev, n, (long)tiles, dx, dy, dz, c->jc_dev, c->st));
  HIPCHK(hipMemcpyAsync(c->jc_host, c->jc_dev, (size_t)n * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, c->st));
  /* also keeps the host table from being rewritten while its copy is pending */
  HIPCHK(hipStreamSynchronize(c->st));
  for (int i = 0; i < n; i++) flagged[i] = (long long)c->jc_host[i];
  return PDFS_OK;
}

int pdfs_selftest_sqrt_f32(pdfs_ctx *c, unsigned long long *mismatches)
{
  if (!c || !mismatches) return fail(PDFS_EINVAL, "pdfs_selftest_sqrt_f32: bad arguments");
  unsigned long long *d = NULL;
  HIPCHK(hipMalloc((void **)&d, sizeof *d));
  int rc = PDFS_OK;
  if (hipMemsetAsync(d, 0, sizeof *d, c->st) != hipSuccess ||
      pdfsk_check_sqrt_cr(0x0F800000u, 0x7F800000u, d, c->st) ||
      hipMemcpyAsync(mismatches, d, sizeof *d, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    rc = fail(PDFS_EHIP, "pdfs_selftest_sqrt_f32: launch failed");
  hipFree(d);
  return rc;
}

int pdfs_judge_f32_dev(pdfs_ctx *c, const float *field, const float *init, int x, int y, int z,
                       double dx, double dy, double dz, uint32_t *flags, long long *flagged)
{
  if (!c || !field || !init || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_judge_f32_dev: bad arguments");
  const int dims[3] = {x, y, z};
  long long fl = 0;
  int rc = pdfsi_judge_batch(c, 1, &field, &init, dims, flags ? &flags : NULL, dx, dy, dz, &fl);
  if (rc == PDFS_OK && flagged) *flagged = fl;
  return rc;
}

/* The FP32 batch solve without the spacing check (run_fsm and the CLI take
 * the caller's spacings as the reference does). */
static int batch_f32(pdfs_ctx *c, int nf, float *const *fields, const float *const *inits,
                     const int *dims, double dx, double dy, double dz, const pdfs_opts *o,
                     pdfs_stats *st, int *field_passes, double *field_ms);

int pdfs_solve_batch_f32_dev(pdfs_ctx *c, int nf, float *const *fields,
                             const float *const *inits, const int *dims, double dx, double dy,
                             double dz, const pdfs_opts *o, pdfs_stats *st, int *field_passes,
                             double *field_ms)
{
  if (!c || nf < 1 || !fields || !dims || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_solve_batch_f32_dev: bad arguments");
  return batch_f32(c, nf, fields, inits, dims, dx, dy, dz, o, st, field_passes, field_ms);
}

static int batch_f32(pdfs_ctx *c, int nf, float *const *fields, const float *const *inits,
                     const int *dims, double dx, double dy, double dz, const pdfs_opts *o,
                     pdfs_stats *st, int *field_passes, double *field_ms)
{
  if (!c || nf < 1 || !fields || !dims)
    return fail(PDFS_EINVAL, "pdfs_solve_batch_f32_dev: bad arguments");
  pdfs_opts dflt;
  if (!o) {
    pdfs_default_opts(&dflt);
    o = &dflt;
  }
  pdfs_stats s;
  memset(&s, 0, sizeof s);
  /* tile: explicit, default, or tuned online (PDFS_TILE_AUTO): inside the
   * first pass, after one untimed ordering, each candidate sweeps one
   * ordering (results do not depend on the tile, and every ordering of the
   * first pass sweeps the whole batch); the fastest sweeps the rest and is
   * cached for the batch's size class */
  long long key = tune_key(nf, dims);
  int tuning = 0;
  int tile;
  if (o->tile >= 0 && o->tile < PDFSK_TILE_COUNT) {
    tile = o->tile;
  } else if (o->tile == PDFS_TILE_AUTO) {
    tile = tune_lookup(c, key);
    if (tile < 0) {
      /* the first candidate whose edge buffers fit */
      tuning = 1;
      for (tile = 0; tile < PDFSK_TILE_COUNT - 1 && !plan_fits(nf, dims, tile); tile++) {
      }
    }
  } else {
    tile = default_tile(nf, dims[0], dims[1]);
  }
  pdfs_sp32 sp;
  make_sp32(dx, dy, dz, &sp);
  int rc;
  if ((rc = plan_build(c, nf, fields, dims, tile))) return rc;
  long long *delta = (long long *)calloc((size_t)nf, sizeof(long long));
  /* per-field change count of the current pass (summed over its segments),
   * and the points the field's last judge flagged */
  long long *pdel = (long long *)calloc((size_t)nf, sizeof(long long));
  long long *lastfl = (long long *)malloc((size_t)nf * sizeof(long long));
  int *active = (int *)malloc((size_t)nf * sizeof(int));
  /* fields whose passes ended on a passing judge: the final judge skips them */
  char *certified = (char *)calloc((size_t)nf, 1);
  /* fields judged together: indices, pointers, dims, counts, outcome */
  int *jidx = (int *)malloc((size_t)nf * sizeof(int));
  int *jdone = (int *)malloc((size_t)nf * sizeof(int));
  int *jd = (int *)malloc((size_t)3 * nf * sizeof(int));
  const float **jv = (const float **)malloc((size_t)nf * sizeof(float *));
  const float **ji = (const float **)malloc((size_t)nf * sizeof(float *));
  long long *jfl = (long long *)malloc((size_t)nf * sizeof(long long));
  if (!delta || !pdel || !lastfl || !active || !certified || !jidx || !jdone || !jd || !jv ||
      !ji || !jfl) {
    free(delta);
    free(pdel);
    free(lastfl);
    free(active);
    free(certified);
    free(jidx);
    free(jdone);
    free(jd);
    free(jv);
    free(ji);
    free(jfl);
    return fail(PDFS_ENOMEM, "batch: host alloc");
  }
  for (int f = 0; f < nf; f++) {
    active[f] = 1;
    lastfl[f] = -1;  /* not judged yet */
    if (field_passes) field_passes[f] = 0;
    if (field_ms) field_ms[f] = 0.0;
  }
  int maxp = o->fixed_passes > 0 ? o->fixed_passes : (o->max_passes > 0 ? o->max_passes : 100);
  int nactive = nf;
  double ms = 0.0, kms = 0.0;
  long long last_total = 0;
  int counted_passes = 0; /* passes counted through a fixed point, not run */
  rc = PDFS_OK;
  HIPGO(hipEventRecord(c->ev0, c->st));
  /* A certified solve judges each field at the end of every pass from
   * certify_from on, and also after the first half of a pass once its last
   * judge flagged at most 1/1024 of its points: such a field often passes a
   * few orderings before the pass ends. The plain stop stays a whole pass
   * without change. */
  const int certifying = o->certify && inits && o->fixed_passes <= 0;
  const int cfrom = o->certify_from > 0 ? o->certify_from : 1;
  for (int p = 0; p < maxp && nactive > 0; p++) {
    int halves = 0;
    for (int f = 0; f < nf && certifying && !tuning && p >= cfrom && !halves; f++)
      halves = active[f] && inits[f] && lastfl[f] >= 0 &&
               lastfl[f] * 1024 <= (long long)dims[3 * f] * dims[3 * f + 1] * dims[3 * f + 2];
    const int nseg = halves ? 2 : 1;
    for (int f = 0; f < nf; f++) pdel[f] = 0;
    for (int g = 0; g < nseg && nactive > 0 && rc == PDFS_OK; g++) {
      const int sn0 = halves ? 1 + 4 * g : 1, sn1 = halves ? 4 + 4 * g : 8;
      const int nord = sn1 - sn0 + 1, last_seg = g == nseg - 1;
      long long npts = 0;
      for (int f = 0; f < nf; f++)
        if (active[f]) npts += (long long)dims[3 * f] * dims[3 * f + 1] * dims[3 * f + 2];
      if (tuning) {
        /* ordering 1 untimed, candidate v on ordering v + 2, the best on 8 */
        double cand_ms[PDFSK_TILE_COUNT], tk = 0.0;
        if ((rc = run_orderings(c, &sp, 1, 1, o->profile, 0, &kms))) break;
        for (int v = 0; v < PDFSK_TILE_COUNT && !rc; v++) {
          cand_ms[v] = 1e30;
          if (!plan_fits(nf, dims, v)) {
            /* its ordering still runs, on the current tile */
            rc = run_orderings(c, &sp, v + 2, v + 2, o->profile, 0, &kms);
            continue;
          }
          if (c->tile != v && (rc = plan_build_ex(c, nf, fields, dims, v, 1))) break;
          tk = 0.0;
          rc = run_orderings(c, &sp, v + 2, v + 2, 1, 0, &tk);
          cand_ms[v] = tk;
          if (o->profile) kms += tk;
        }
        if (rc) break;
        int best = 0;
        for (int v = 1; v < PDFSK_TILE_COUNT; v++)
          if (cand_ms[v] < cand_ms[best]) best = v;
        tuning = 0;
        tune_store(c, key, best);
        tile = best;
        if (c->tile != best && (rc = plan_build_ex(c, nf, fields, dims, best, 1))) break;
        if ((rc = run_orderings(c, &sp, PDFSK_TILE_COUNT + 2, 8, o->profile, 0, &kms))) break;
      } else if ((rc = run_orderings(c, &sp, sn0, sn1, o->profile, o->graph, &kms))) {
        break;
      }
      HIPGO(hipEventRecord(c->ev1, c->st));
      unsigned err = 0;
      if ((rc = read_changed(c, delta, &err))) break;
      float t = 0.f;
      hipEventElapsedTime(&t, c->ev0, c->ev1);
      ms = t;
      if (g == 0) s.passes++;
      s.orderings += nord;
      s.point_updates += nord * npts;
      if (err) {
        rc = fail(PDFS_ETIMEOUT, "column pipeline: edge wait timed out");
        break;
      }
      last_total = 0;
      int changed_set = 0;
      /* certified stop: the FP64 residual judge passes on a field (needs its
       * init); plain stop: a pass changed nothing */
      const int judge_now = certifying && (last_seg ? p + 1 >= cfrom : 1);
      float tj0 = 0.f;
      if (judge_now) {
        HIPGO(hipEventRecord(c->ev1, c->st));
        HIPGO(hipEventSynchronize(c->ev1));
        hipEventElapsedTime(&tj0, c->ev0, c->ev1);
      }
      int nj = 0;
      for (int f = 0; f < nf; f++) {
        jdone[f] = 0;
        if (!active[f]) continue;
        pdel[f] += delta[f];
        last_total += pdel[f];
        if (field_passes) field_passes[f] = p + 1;
        jdone[f] = last_seg && o->fixed_passes <= 0 && pdel[f] == 0;
        {
          /* the device skipped the orderings of this segment that followed
           * the field's PDFSK_CONVERGED_RUN-th consecutive unchanged one */
          const unsigned run = c->h_streak[2 * f + 1];
          long long skipped =
              run > PDFSK_CONVERGED_RUN ? (long long)(run - PDFSK_CONVERGED_RUN) : 0;
          if (skipped > nord) skipped = nord;
          s.point_updates -= skipped * dims[3 * f] * dims[3 * f + 1] * dims[3 * f + 2];
          if (nf == 1) s.orderings -= (int)skipped;
          /* A pass that ends in an unchanged ordering left the field at its
           * fixed point, so the next pass would change nothing: the plain
           * stop's pass is counted (passes, and a last pass without change)
           * but not launched. */
          if (!jdone[f] && last_seg && o->fixed_passes <= 0 && run >= PDFSK_CONVERGED_RUN &&
              p + 2 <= maxp) {
            jdone[f] = 1;
            last_total -= pdel[f];
            if (field_passes) field_passes[f] = p + 2;
            if (p + 2 > counted_passes) counted_passes = p + 2;
          }
        }
        const int near = last_seg || (lastfl[f] >= 0 && lastfl[f] * 1024 <=
                                                   (long long)dims[3 * f] * dims[3 * f + 1] * dims[3 * f + 2]);
        if (!jdone[f] && judge_now && inits[f] && near) {
          jidx[nj] = f;
          jv[nj] = fields[f];
          ji[nj] = inits[f];
          memcpy(jd + 3 * nj, dims + 3 * f, 3 * sizeof(int));
          nj++;
        }
      }
      /* all fields to certify in one judge launch */
      if (nj > 0) {
        if ((rc = pdfsi_judge_batch(c, nj, jv, ji, jd, NULL, dx, dy, dz, jfl))) break;
        s.judge_runs += nj;
        for (int k = 0; k < nj; k++) {
          jdone[jidx[k]] = jfl[k] == 0;
          certified[jidx[k]] = (char)(jfl[k] == 0);
          lastfl[jidx[k]] = jfl[k];
        }
      }
      for (int f = 0; f < nf; f++) {
        if (!active[f]) continue;
        if (jdone[f]) {
          active[f] = 0;
          nactive--;
          changed_set = 1;
          if (field_ms) field_ms[f] = ms;
        } else if (field_ms) {
          field_ms[f] = ms;
        }
      }
      if (judge_now) {
        HIPGO(hipEventRecord(c->ev1, c->st));
        HIPGO(hipEventSynchronize(c->ev1));
        float tj1 = 0.f;
        hipEventElapsedTime(&tj1, c->ev0, c->ev1);
        s.ms_judge += tj1 - tj0;
        ms = tj1;
      }
      if (changed_set && nactive > 0 && (rc = plan_upload_cols(c, active))) break;
    }
    if (rc) break;
  }
  if (counted_passes > s.passes) s.passes = counted_passes;
  s.tile = c->tile;
  s.last_changed = last_total;
  s.ms_sweeps = ms - s.ms_judge;
  s.ms_sweep_kernels = kms;
  if (rc == PDFS_OK && o->judge && inits) {
    float tj_before = (float)s.ms_judge;
    int judged = 0;
    HIPGO(hipEventRecord(c->ev0, c->st));
    int nj = 0;
    for (int f = 0; f < nf; f++) {
      if (!inits[f] || certified[f]) continue;
      jv[nj] = fields[f];
      ji[nj] = inits[f];
      memcpy(jd + 3 * nj, dims + 3 * f, 3 * sizeof(int));
      nj++;
    }
    if (nj > 0 && (rc = pdfsi_judge_batch(c, nj, jv, ji, jd, NULL, dx, dy, dz, jfl)) == PDFS_OK) {
      for (int k = 0; k < nj; k++) s.judge_flagged += jfl[k];
      judged = nj;
    }
    HIPGO(hipEventRecord(c->ev1, c->st));
    HIPGO(hipEventSynchronize(c->ev1));
    float t = 0.f;
    hipEventElapsedTime(&t, c->ev0, c->ev1);
    s.ms_judge = tj_before + t;
    s.judge_runs += judged;
  }
  s.ms_total = s.ms_sweeps + s.ms_judge;
  if (st) *st = s;
  if (rc == PDFS_OK && o->fixed_passes <= 0 && nactive > 0)
    rc = fail(PDFS_ENOTCONV, "%d field(s) not converged after %d passes", nactive, s.passes);
out:
  free(delta);
  free(pdel);
  free(lastfl);
  free(active);
  free(certified);
  free(jidx);
  free(jdone);
  free(jd);
  free(jv);
  free(ji);
  free(jfl);
  return rc;
}

/* One FP32 field, without the spacing check. */
static int single_f32(pdfs_ctx *c, float *field, const float *init, int x, int y, int z,
                      double dx, double dy, double dz, const pdfs_opts *o, pdfs_stats *st)
{
  if (!c || !field || x < 1 || y < 1 || z < 1)
    return fail(PDFS_EINVAL, "pdfs_solve_f32_dev: bad arguments");
  int dims[3] = {x, y, z};
  return batch_f32(c, 1, &field, &init, dims, dx, dy, dz, o, st, NULL, NULL);
}

int pdfs_solve_f32_dev(pdfs_ctx *c, float *field, const float *init, int x, int y, int z,
                       double dx, double dy, double dz, const pdfs_opts *o, pdfs_stats *st)
{
  if (!c || !field || x < 1 || y < 1 || z < 1 || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_solve_f32_dev: bad arguments");
  return single_f32(c, field, init, x, y, z, dx, dy, dz, o, st);
}

static int solve_ref64(pdfs_ctx *c, double *padded, const double *slow, int x, int y, int z,
                       double dx, double dy, double dz, int passes, double tol, pdfs_stats *st)
{
  pdfs_stats s;
  memset(&s, 0, sizeof s);
  int maxp = passes > 0 ? passes : 1000;
  long long npts = (long long)x * y * z;
  float ms = 0.f;
  for (int p = 0; p < maxp; p++) {
    HIPCHK(hipMemsetAsync(c->d_u64, 0, sizeof(unsigned long long), c->st));
    HIPCHK(hipEventRecord(c->ev0, c->st));
    for (int sn = 1; sn <= 8; sn++)
      KCHK(pdfsk_sweep_ref64(padded, slow, x, y, z, sn, dx, dy, dz, c->d_u64, c->st));
    HIPCHK(hipEventRecord(c->ev1, c->st));
    HIPCHK(hipMemcpyAsync(c->h_pin64, c->d_u64, sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    float t = 0.f;
    hipEventElapsedTime(&t, c->ev0, c->ev1);
    ms += t;
    double md;
    unsigned long long b = c->h_pin64[0];
    memcpy(&md, &b, sizeof md);
    s.passes++;
    s.orderings += 8;
    s.point_updates += 8 * npts;
    s.last_maxdiff = md;
    if (passes <= 0 && md <= tol) break;
  }
  s.ms_sweeps = ms;
  s.ms_total = ms;
  if (st) *st = s;
  return PDFS_OK;
}

int pdfs_solve_ref64_dev(pdfs_ctx *c, double *padded, int x, int y, int z, double dx, double dy,
                         double dz, int passes, double tol, pdfs_stats *st)
{
  if (!c || !padded || x < 1 || y < 1 || z < 1 || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_solve_ref64_dev: bad arguments");
  return solve_ref64(c, padded, NULL, x, y, z, dx, dy, dz, passes, tol, st);
}

int pdfs_solve_speed64_dev(pdfs_ctx *c, double *padded, const double *slowness, int x, int y,
                           int z, double dx, double dy, double dz, int passes, double tol,
                           pdfs_stats *st)
{
  if (!c || !padded || !slowness || x < 1 || y < 1 || z < 1 || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_solve_speed64_dev: bad arguments");
  return solve_ref64(c, padded, slowness, x, y, z, dx, dy, dz, passes, tol, st);
}

/* ---- host padded FP64 entry points ---- */
static pdfs_ctx *g_ctx;

static pdfs_ctx *default_ctx(void)
{
  if (!g_ctx) g_ctx = pdfs_ctx_create();
  return g_ctx;
}

/* pdfs_solve without the spacing check: run_fsm takes the spacings the
 * caller's Phi holds, as the reference does */
static int solve_host(double *distance, int x, int y, int z, double dx, double dy, double dz,
                      const pdfs_opts *o, pdfs_stats *st);

int pdfs_solve(double *distance, int x, int y, int z, double dx, double dy, double dz,
               const pdfs_opts *o, pdfs_stats *st)
{
  if (!distance || x < 1 || y < 1 || z < 1 || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_solve: bad arguments");
  return solve_host(distance, x, y, z, dx, dy, dz, o, st);
}

static int solve_host(double *distance, int x, int y, int z, double dx, double dy, double dz,
                      const pdfs_opts *o, pdfs_stats *st)
{
  if (!distance || x < 1 || y < 1 || z < 1)
    return fail(PDFS_EINVAL, "pdfs_solve: bad arguments");
  pdfs_opts dflt;
  if (!o) {
    pdfs_default_opts(&dflt);
    o = &dflt;
  }
  pdfs_ctx *c = default_ctx();
  if (!c) return fail(PDFS_EHIP, "pdfs_solve: no HIP device");
  size_t np = (size_t)(x + 2) * (y + 2) * (z + 2), n = (size_t)x * y * z;
  double *d = NULL;
  float *f = NULL, *f0 = NULL;
  int rc = PDFS_OK;
  HIPCHK(hipMalloc((void **)&d, np * sizeof(double)));
  if (hipMemcpy(d, distance, np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail(PDFS_EHIP, "pdfs_solve: H2D copy");
    goto out;
  }
  if (o->precision == PDFS_F64_REF) {
    rc = solve_ref64(c, d, NULL, x, y, z, dx, dy, dz, o->fixed_passes > 0 ? o->fixed_passes : 0,
                     o->tol, st);
  } else {
    if (hipMalloc((void **)&f, n * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&f0, n * sizeof(float)) != hipSuccess) {
      rc = fail(PDFS_ENOMEM, "pdfs_solve: device alloc");
      goto out;
    }
    if ((rc = pdfs_pack_dev(c, d, f, x, y, z))) goto out;
    if (hipMemcpyAsync(f0, f, n * sizeof(float), hipMemcpyDeviceToDevice, c->st) != hipSuccess) {
      rc = fail(PDFS_EHIP, "pdfs_solve: D2D copy");
      goto out;
    }
    rc = single_f32(c, f, f0, x, y, z, dx, dy, dz, o, st);
    if (rc && rc != PDFS_ENOTCONV) goto out;
    int rc2 = pdfs_unpack_dev(c, f, d, x, y, z);
    if (rc2) {
      rc = rc2;
      goto out;
    }
  }
  if (hipStreamSynchronize(c->st) != hipSuccess ||
      hipMemcpy(distance, d, np * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(PDFS_EHIP, "pdfs_solve: D2H copy");
out:
  hipFree(d);
  hipFree(f);
  hipFree(f0);
  return rc;
}

/* single_gpu_cuda/src/phi3D_fsm.cu:64-106: device copy in, `iterations`
 * passes, copy out, print the elapsed time; errors exit like
 * cudaCheckError. */
void run_fsm(Phi *pf, int iterations)
{
  pdfs_opts o;
  pdfs_default_opts(&o);
  o.precision = PDFS_F64_REF;
  o.fixed_passes = iterations > 0 ? iterations : 0;
  o.tol = 1e-12;
  pdfs_stats s;
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  hipEventRecord(a, 0);
  int rc = solve_host(pf->distance, pf->x, pf->y, pf->z, pf->dx, pf->dy, pf->dz, &o, &s);
  hipEventRecord(b, 0);
  hipEventSynchronize(b);
  if (rc != PDFS_OK) {
    printf("\nHIP failure in run_fsm: '%s'\n", pdfs_last_error());
    exit(EXIT_FAILURE);
  }
  float ms = 0.f;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  printf("Parallel FSM time: %f s.\n", ms / 1000.0);
}

/* mpi_openacc/src/phi3D_fsm.c:67: one ordering sn, in place. */
void runFSM(PhiACC *pf, int sn)
{
  pdfs_ctx *c = default_ctx();
  if (!c || sn < 1 || sn > 8) {
    printf("runFSM: invalid call\n");
    exit(EXIT_FAILURE);
  }
  size_t np = (size_t)(pf->x + 2) * (pf->y + 2) * (pf->z + 2);
  double *d = NULL;
  if (hipMalloc((void **)&d, np * sizeof(double)) != hipSuccess ||
      hipMemcpy(d, pf->distance, np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      pdfsk_sweep_ref64(d, NULL, pf->x, pf->y, pf->z, sn, pf->dx, pf->dy, pf->dz, NULL, c->st) != 0 ||
      hipStreamSynchronize(c->st) != hipSuccess ||
      hipMemcpy(pf->distance, d, np * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
    printf("\nHIP failure in runFSM\n");
    exit(EXIT_FAILURE);
  }
  hipFree(d);
}

int pdfs_pack_dev(pdfs_ctx *c, const double *padded, float *dense, int x, int y, int z)
{
  KCHK(pdfsk_pack_f64(padded, dense, x, y, z, c->st));
  return PDFS_OK;
}

int pdfs_unpack_dev(pdfs_ctx *c, const float *dense, double *padded, int x, int y, int z)
{
  KCHK(pdfsk_unpack_f64(dense, padded, x, y, z, c->st));
  return PDFS_OK;
}

/* ---- synthetic fields: host parameter draws (splitmix64, Steele, Lea &
 * Flood 2014) in the documented draw order, per-cell evaluation on the GPU */
static uint64_t splitmix64(uint64_t *s)
{
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double u01(uint64_t *s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

/* Synthetic field with explicit object parameters: terrain (nobj hills,
 * amplitude U(p_lo, p_hi) h, width U(q_lo, q_hi) h), spheres (nobj spheres,
 * radius U(p_lo, p_hi) h) or a point source (parameters unused). */
int pdfs_generate_ex_dev(pdfs_ctx *c, int kind, uint64_t seed, int x, int y, int z, double h,
                         int nobj, double p_lo, double p_hi, double q_lo, double q_hi,
                         float *dense, double *padded)
{
  if (!c || x < 1 || y < 1 || z < 1 || !(h > 0)) return fail(PDFS_EINVAL, "generate: bad args");
  if (nobj < 0 || nobj > 256) return fail(PDFS_EINVAL, "generate: 0..256 objects");
  uint64_t s = seed;
  double prm[1024];
  if (kind == PDFS_GEN_POINT) {
    int k = 1 + (int)(u01(&s) * x);
    int j = 1 + (int)(u01(&s) * y);
    int i = 1 + (int)(u01(&s) * z);
    KCHK(pdfsk_gen_point(dense, padded, x, y, z, k, j, i, c->st));
  } else if (kind == PDFS_GEN_TERRAIN) {
    if (nobj > 64) return fail(PDFS_EINVAL, "generate: at most 64 hills");
    for (int q = 0; q < nobj; q++) {
      prm[4 * q] = (p_lo + (p_hi - p_lo) * u01(&s)) * h;
      prm[4 * q + 1] = (q_lo + (q_hi - q_lo) * u01(&s)) * h;
      prm[4 * q + 2] = u01(&s) * (double)(x - 1) * h;
      prm[4 * q + 3] = u01(&s) * (double)(y - 1) * h;
    }
    if (nobj) HIPCHK(hipMemcpyAsync(c->d_params, prm, 4 * nobj * sizeof(double),
                                    hipMemcpyHostToDevice, c->st));
    KCHK(pdfsk_gen_terrain(dense, padded, x, y, z, h, c->d_params, nobj, c->st));
  } else if (kind == PDFS_GEN_SPHERES) {
    for (int q = 0; q < nobj; q++) {
      prm[4 * q] = u01(&s) * (double)(x - 1) * h;
      prm[4 * q + 1] = u01(&s) * (double)(y - 1) * h;
      prm[4 * q + 2] = u01(&s) * (double)(z - 1) * h;
      prm[4 * q + 3] = (p_lo + (p_hi - p_lo) * u01(&s)) * h;
    }
    if (nobj) HIPCHK(hipMemcpyAsync(c->d_params, prm, 4 * nobj * sizeof(double),
                                    hipMemcpyHostToDevice, c->st));
    KCHK(pdfsk_gen_spheres(dense, padded, x, y, z, h, c->d_params, nobj, c->st));
  } else {
    return fail(PDFS_EINVAL, "generate: unknown kind %d", kind);
  }
  HIPCHK(hipStreamSynchronize(c->st));
  return PDFS_OK;
}

/* Configuration defaults: terrain 8 hills, amplitude U(5,40) h, width
 * U(10,60) h; spheres 64, radius U(4,32) h. */
int pdfs_generate_dev(pdfs_ctx *c, int kind, uint64_t seed, int x, int y, int z, double h,
                      float *dense, double *padded)
{
  if (kind == PDFS_GEN_TERRAIN)
    return pdfs_generate_ex_dev(c, kind, seed, x, y, z, h, 8, 5.0, 40.0, 10.0, 60.0, dense,
                                padded);
  if (kind == PDFS_GEN_SPHERES)
    return pdfs_generate_ex_dev(c, kind, seed, x, y, z, h, 64, 4.0, 32.0, 0.0, 0.0, dense,
                                padded);
  return pdfs_generate_ex_dev(c, kind, seed, x, y, z, h, 0, 0.0, 0.0, 0.0, 0.0, dense, padded);
}

/* ---- CLI pipeline: single_gpu_cuda/src/phi3D.c:107-126 (calc_dist_field) ----
 * update_distance -> `passes` passes of the sweeps -> set_distance_negative_
 * inside -> adjust_boundary, all on the device. */
int pdfs_calc_dist_field(Phi *pf, int passes, int precision)
{
  if (!pf || !pf->distance || !pf->location || pf->x < 1 || pf->y < 1 || pf->z < 1)
    return fail(PDFS_EINVAL, "pdfs_calc_dist_field: bad arguments");
  pdfs_ctx *c = default_ctx();
  if (!c) return fail(PDFS_EHIP, "pdfs_calc_dist_field: no HIP device");
  const int x = pf->x, y = pf->y, z = pf->z;
  size_t np = (size_t)(x + 2) * (y + 2) * (z + 2), n = (size_t)x * y * z;
  double *d = NULL;
  int *loc = NULL;
  float *f = NULL;
  int rc = PDFS_OK;
  if (hipMalloc((void **)&d, np * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&loc, np * sizeof(int)) != hipSuccess) {
    rc = fail(PDFS_ENOMEM, "pdfs_calc_dist_field: device alloc");
    goto out;
  }
  if (hipMemcpy(d, pf->distance, np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(loc, pf->location, np * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail(PDFS_EHIP, "pdfs_calc_dist_field: H2D copy");
    goto out;
  }
  if (pdfsk_update_distance(loc, d, (long)np, c->st)) {
    rc = fail(PDFS_EHIP, "update_distance launch");
    goto out;
  }
  if (precision == PDFS_F64_REF) {
    if ((rc = solve_ref64(c, d, NULL, x, y, z, pf->dx, pf->dy, pf->dz, passes, 1e-12, NULL)))
      goto out;
  } else {
    pdfs_opts o;
    pdfs_default_opts(&o);
    o.fixed_passes = passes > 0 ? passes : 0;
    o.judge = 0;
    if (hipMalloc((void **)&f, n * sizeof(float)) != hipSuccess) {
      rc = fail(PDFS_ENOMEM, "pdfs_calc_dist_field: device alloc");
      goto out;
    }
    if ((rc = pdfs_pack_dev(c, d, f, x, y, z))) goto out;
    rc = single_f32(c, f, NULL, x, y, z, pf->dx, pf->dy, pf->dz, &o, NULL);
    if (rc && rc != PDFS_ENOTCONV) goto out;
    if ((rc = pdfs_unpack_dev(c, f, d, x, y, z))) goto out;
  }
  if (pdfsk_set_negative_inside(loc, d, (long)np, c->st) ||
      pdfsk_adjust_boundary(d, x, y, z, c->st)) {
    rc = fail(PDFS_EHIP, "post-processing launch");
    goto out;
  }
  if (hipStreamSynchronize(c->st) != hipSuccess ||
      hipMemcpy(pf->distance, d, np * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(PDFS_EHIP, "pdfs_calc_dist_field: D2H copy");
out:
  hipFree(d);
  hipFree(loc);
  hipFree(f);
  return rc;
}

/* Single pre/post steps on host padded arrays (used by tests and callers
 * that keep their own pipeline). op: 0 update_distance, 1 set_negative_
 * inside, 2 adjust_boundary (loc unused). */
int pdfs_prepost(int op, const int *location, double *distance, int x, int y, int z)
{
  pdfs_ctx *c = default_ctx();
  if (!c || !distance || (op != 2 && !location) || op < 0 || op > 2)
    return fail(PDFS_EINVAL, "pdfs_prepost: bad arguments");
  size_t np = (size_t)(x + 2) * (y + 2) * (z + 2);
  double *d = NULL;
  int *loc = NULL;
  int rc = PDFS_OK;
  if (hipMalloc((void **)&d, np * sizeof(double)) != hipSuccess ||
      (op != 2 && hipMalloc((void **)&loc, np * sizeof(int)) != hipSuccess)) {
    rc = fail(PDFS_ENOMEM, "pdfs_prepost: device alloc");
    goto out;
  }
  if (hipMemcpy(d, distance, np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      (loc && hipMemcpy(loc, location, np * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)) {
    rc = fail(PDFS_EHIP, "pdfs_prepost: H2D copy");
    goto out;
  }
  int k = op == 0   ? pdfsk_update_distance(loc, d, (long)np, c->st)
          : op == 1 ? pdfsk_set_negative_inside(loc, d, (long)np, c->st)
                    : pdfsk_adjust_boundary(d, x, y, z, c->st);
  if (k || hipStreamSynchronize(c->st) != hipSuccess ||
      hipMemcpy(distance, d, np * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(PDFS_EHIP, "pdfs_prepost: kernel or copy");
out:
  hipFree(d);
  hipFree(loc);
  return rc;
}
