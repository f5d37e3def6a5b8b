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

/* A ticket table of the overlapped passes (device, nt tickets of 5 ints, cap
 * allocated), the inputs it was ordered from (key: tile, grid, orderings,
 * fields, columns) and when it was last used. */
typedef struct {
  int *tab;
  int nt, cap;
  int *key;
  int key_n, key_cap;
  long long used;
} ov_table;
/* tables kept per segment: the plans a solve alternates between (its sweep
 * and its polish, a batch as fields drop out) order their tickets once */
#define PDFS_OV_TABLES 4

struct pdfs_ctx {
  int dev;
  hipStream_t st;
  int resident[PDFSK_TILE_COUNT][PDFSK_MODE_COUNT];
  /* batched column sweep: persistent device state, edge allocations, plan */
  pdfsk_batch bt;
  size_t cap_ex, cap_ey;           /* edge entries excluding padding */
  pdfsk_field *h_fields;           /* host copy of the field table */
  int *h_cols;
  int *h_cols_all; /* every column of the plan's fields, in ticket order */
  int ncols_all;
  int nf, ncols, tile, cap_f, cap_cols;
  int nact; /* fields with columns in the plan */
  int overlap64; /* opts.overlap of the FP64 entry points without options (default AUTO) */
  int words;                       /* 32-bit words per value of the current plan (1 or 2) */
  unsigned *h_changed, *prev_changed; /* pinned readback, last values */
  unsigned *h_streak;              /* per field: counter, run of unchanged orderings */
  unsigned *h_err;
  hipGraphExec_t gexec[3];         /* captured pass (orderings 1-8), halves 1-4 and 5-8 */
  long long tune_keys[32];         /* tile autotuner cache: size class -> tile */
  int tune_tiles[32], ntune;
  int g_tile[3], g_mode[3], g_grid[3];
  /* FP64 column sweeps: per-field max change (device, pinned readback),
   * dense FP64 working fields, conversion table */
  unsigned long long *h_maxdiff;
  int cap_md;
  double *pbuf;
  size_t pcap;
  pdfsk_conv *conv_host, *conv_dev;
  int cap_conv;
  /* FP64 path / judge */
  unsigned long long *d_u64; /* [0] maxdiff bits */
  unsigned long long *h_pin64;
  /* batched judge: field table (host staging, device) and per-field counts */
  pdfsk_judge_field *jt_host, *jt_dev;
  unsigned long long *jc_dev, *jc_host;
  int cap_jt;
  double *d_params;
  hipEvent_t ev0, ev1;
  hipEvent_t kev[18];   /* [0..15] per-ordering profile, [16..17] polish */
  /* overlapped passes (pdfsk_sweep_batch_ov): whether the current plan may
   * use them (its fields tile exactly and its edge allocation holds two
   * regions), whether the solve in progress does, ticket tables of the
   * segments 1-8, 1-4 and 5-8 for the current active set, done flags */
  int ov_ok, ov_use, last_ov;
  long long launches;  /* sweep launches since the context was created */
  ov_table ov_cache[3][PDFS_OV_TABLES]; /* per segment (1-8, 1-4, 5-8) */
  int ov_cur[3], ov_valid[3];           /* the table of the current plan, if valid */
  long long ov_clock;
  unsigned *ov_done;
  unsigned *ov_chg;  /* [8][nf] per-slot change counters of an overlapped launch */
  size_t ov_chg_cap;
  size_t ov_done_cap;
  int ov_ncolp;
  size_t ov_rx, ov_ry;
  /* fixed-point solve in progress: per-field quantum exponents of the plan
   * (NULL otherwise), conversion table (host staging, device), order keys */
  const int *fxk;
  pdfsk_fx *fx_host, *fx_dev;
  uint32_t *fx_keys, *h_fx_keys;
  int cap_fx;
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

/* kernel launch inside a function that owns host allocations */
#define KGO(call)                                                                    \
  do {                                                                               \
    int r_ = (call);                                                                 \
    if (r_ != 0) {                                                                   \
      rc = fail(PDFS_EHIP, "%s:%d %s: launch failed (%d)", __FILE__, __LINE__, #call, r_); \
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
  o->polish = 1;
  o->polish_tol = 2e-5;
  o->tile64 = -1;
  o->storage = PDFS_STORE_FIXED;
  o->judge_tol = 2e-6;
  o->stop_quanta = 4;
  o->overlap = PDFS_OVERLAP_AUTO;
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
  c->overlap64 = PDFS_OVERLAP_AUTO;
  for (int i = 0; i < 3; i++) c->ov_cur[i] = -1;
  if (hipGetDevice(&c->dev) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess ||
      hipMalloc((void **)&c->bt.state, 32 * sizeof(unsigned)) != hipSuccess ||
      hipMalloc((void **)&c->d_u64, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc((void **)&c->h_err, 16 * sizeof(unsigned), 0) != hipSuccess ||
      hipHostMalloc((void **)&c->h_pin64, 4 * sizeof(unsigned long long), 0) != hipSuccess ||
      hipMalloc((void **)&c->d_params, 1024 * sizeof(double)) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      create_events(c->kev, 18) != 0) {
    fail(PDFS_EHIP, "pdfs_ctx_create: HIP initialisation failed");
    free(c);
    return NULL;
  }
  c->bt.err = c->bt.state + 8;
  /* epochs start at 1 (zeroed edge entries carry epoch 0) and only grow for
   * the lifetime of the context */
  unsigned init[32] = {0, 1, 0, 0, 0};
  if (hipMemcpy(c->bt.state, init, sizeof init, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c->d_u64, 0, 4 * sizeof(unsigned long long)) != hipSuccess) {
    free(c);
    return NULL;
  }
  for (int v = 0; v < PDFSK_TILE_COUNT; v++)
    for (int m = 0; m < PDFSK_MODE_COUNT; m++) c->resident[v][m] = pdfsk_col_max_resident(v, m);
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
  hipFree(c->bt.maxdiff);
  hipHostFree(c->h_maxdiff);
  hipFree(c->pbuf);
  hipFree(c->conv_dev);
  free(c->conv_host);
  for (int i = 0; i < 3; i++)
    for (int e = 0; e < PDFS_OV_TABLES; e++) {
      hipFree(c->ov_cache[i][e].tab);
      free(c->ov_cache[i][e].key);
    }
  hipFree(c->ov_done);
  hipFree(c->ov_chg);
  free(c->fx_host);
  hipFree(c->fx_dev);
  hipFree(c->fx_keys);
  hipHostFree(c->h_fx_keys);
  hipHostFree(c->h_streak);
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
  free(c->h_cols_all);
  hipEventDestroy(c->ev0);
  hipEventDestroy(c->ev1);
  for (int i = 0; i < 18; i++) hipEventDestroy(c->kev[i]);
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

void pdfsi_make_sp32(double dx, double dy, double dz, pdfs_sp32 *s) { make_sp32(dx, dy, dz, s); }

static int mode32(const pdfs_sp32 *sp)
{
  return sp->iso ? PDFSK_MODE_F32_ISO : PDFSK_MODE_F32_ANISO;
}

/* FP64 spacing constants in the oracle's operation order. */
static void make_sp64(double dx, double dy, double dz, pdfs_sp64 *s)
{
  memset(s, 0, sizeof *s);
  s->iso = (dx == dy && dy == dz);
  s->d[0] = dx;
  s->d[1] = dy;
  s->d[2] = dz;
  for (int i = 0; i < 3; i++) s->w[i] = 1.0 / (s->d[i] * s->d[i]);
  s->W = s->w[0] + s->w[1] + s->w[2];
  s->h = dx;
  const double h2 = dx * dx;
  s->two_h2 = 2.0 * h2;
  s->three_h2 = 3.0 * h2;
  s->third = 1.0 / 3.0;
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

int pdfsi_default_tile(int nf, int x, int y) { return default_tile(nf, x, y); }

/* the online tuner gives each candidate one of orderings 2..7 */
_Static_assert(PDFSK_TUNE_COUNT <= 6, "tile candidates must fit orderings 2..7 of a pass");

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

/* Ticket order of the plan's columns: ascending a (TX + e) + b (TY + e),
 * e = PDFS_TICKET_EDGE_STEPS, then field, then a (a total order). It is
 * sorted once per plan (active NULL); a subset of the fields (fields
 * dropping out of a batch) takes the sorted list filtered, which is the
 * order a sort of the subset would give, in O(columns). Re-sorting the
 * active columns when fields converged had left the GPU idle for up to
 * 1.2 ms between passes of a batch of 512 fields. */
static int plan_upload_cols(pdfs_ctx *c, const int *active)
{
  int nact = 0;
  for (int f = 0; f < c->nf; f++)
    if (!active || active[f]) nact += c->h_fields[f].nTx * c->h_fields[f].nTy > 0;
  int t = 0;
  if (!active) {
    int TX = 0, TY = 0;
    pdfsk_tile_dims(c->tile, &TX, &TY);
    const int wx = TX + PDFS_TICKET_EDGE_STEPS, wy = TY + PDFS_TICKET_EDGE_STEPS;
    int n_all = 0;
    for (int f = 0; f < c->nf; f++) n_all += c->h_fields[f].nTx * c->h_fields[f].nTy;
    col_key *keys = (col_key *)malloc((size_t)(n_all > 0 ? n_all : 1) * sizeof(col_key));
    if (!keys) return fail(PDFS_ENOMEM, "plan: host alloc");
    for (int f = 0; f < c->nf; f++) {
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
      c->h_cols_all[3 * i] = keys[i].f;
      c->h_cols_all[3 * i + 1] = keys[i].a;
      c->h_cols_all[3 * i + 2] = keys[i].b;
    }
    free(keys);
    c->ncols_all = t;
    memcpy(c->h_cols, c->h_cols_all, (size_t)3 * t * sizeof(int));
  } else {
    for (int i = 0; i < c->ncols_all; i++) {
      const int *k = c->h_cols_all + 3 * i;
      if (!active[k[0]]) continue;
      c->h_cols[3 * t] = k[0];
      c->h_cols[3 * t + 1] = k[1];
      c->h_cols[3 * t + 2] = k[2];
      t++;
    }
  }
  unsigned n = (unsigned)t;
  HIPCHK(hipMemcpyAsync(c->bt.cols, c->h_cols, (size_t)(3 * (t > 0 ? t : 1)) * sizeof(int),
                        hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->bt.state + 4, &n, sizeof n, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->ncols = t;
  c->nact = nact;
  for (int i = 0; i < 3; i++) c->ov_valid[i] = 0;
  return PDFS_OK;
}

/* ---- overlapped passes --------------------------------------------------
 * One launch sweeps the orderings sn0..sn1 of a pass (pdfsk_sweep_batch_ov).
 * Its tickets (slot, field, a, b, physical column) follow a simulated
 * schedule: columns take Z + TX + TY - 2 steps, start TX (TY) +
 * PDFS_TICKET_EDGE_STEPS steps after their upwind neighbour in the slot and
 * after the slot before has finished them and their physical neighbours,
 * on as many concurrent columns as the grid holds. Tickets are emitted in a
 * topological order of those dependencies (Kahn, earliest simulated ready
 * time first; the whole slot o - 2 before any ticket of slot o), so every
 * ticket's dependencies hold earlier tickets. */
/* Simulated times are whole column steps: integer keys, ties broken by id. */
typedef struct {
  long long key;
  int id;
} ov_item;

static inline int ov_less(ov_item a, ov_item b)
{
  return a.key < b.key || (a.key == b.key && a.id < b.id);
}

static void ov_push(ov_item *h, int *n, long long key, int id)
{
  const ov_item it = {key, id};
  int i = (*n)++;
  while (i > 0) {
    const int p = (i - 1) / 2;
    if (ov_less(h[p], it)) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = it;
}

static ov_item ov_pop(ov_item *h, int *n)
{
  const ov_item top = h[0], last = h[--(*n)];
  int i = 0;
  for (;;) {
    const int l = 2 * i + 1;
    if (l >= *n) break;
    const int m = l + 1 < *n && ov_less(h[l + 1], h[l]) ? l + 1 : l;
    if (!ov_less(h[m], last)) break;
    h[i] = h[m];
    i = m;
  }
  if (*n > 0) h[i] = last;
  return top;
}

static void ov_dirs(int sn, int *fx, int *fy)
{
  const int desc = (sn == 2 || sn == 5 || sn == 7 || sn == 8);
  *fx = desc ^ (sn == 4 || sn == 8);
  *fy = desc ^ (sn == 2 || sn == 6);
}

int pdfsi_ov_order(int nf, const int *tnx, const int *tny, const int *fz, int nc,
                   const int *cols, int TX, int TY, int sn0, int sn1, int grid, int *out)
{
  const int ns = sn1 - sn0 + 1, nt = ns * nc;
  const long long LAG = PDFS_TICKET_EDGE_STEPS;
  if (ns < 1 || ns > 8 || nc < 0 || grid < 1) return -1;
  /* x / y reflections of each slot's ordering */
  int dfx[8], dfy[8];
  for (int o = 0; o < ns; o++) ov_dirs(sn0 + o, &dfx[o], &dfy[o]);
  /* column index of (f, a, b) in the active list; physical column bases */
  int *pbase = (int *)malloc((size_t)(nf + 1) * sizeof(int));
  int *cidx = NULL, *left = (int *)calloc((size_t)(nt > 0 ? nt : 1), sizeof(int));
  long long *ready = (long long *)calloc((size_t)(nt > 0 ? nt : 1), sizeof(long long));
  long long slotend[8] = {0};
  int emitted[8] = {0};
  ov_item *heap = (ov_item *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(ov_item));
  ov_item *slots = (ov_item *)malloc((size_t)grid * sizeof(ov_item));
  int ne = -1;
  if (!pbase || !left || !ready || !heap || !slots) goto done;
  pbase[0] = 0;
  for (int f = 0; f < nf; f++) pbase[f + 1] = pbase[f] + tnx[f] * tny[f];
  cidx = (int *)malloc((size_t)(pbase[nf] > 0 ? pbase[nf] : 1) * sizeof(int));
  if (!cidx) goto done;
  for (int i = 0; i < pbase[nf]; i++) cidx[i] = -1;
  for (int i = 0; i < nc; i++) {
    const int f = cols[3 * i], a = cols[3 * i + 1], b = cols[3 * i + 2];
    cidx[pbase[f] + a * tny[f] + b] = i;
  }
  /* dependency counts: upwind in the slot, the slot before (self and
   * physical neighbours) */
  for (int o = 0; o < ns; o++)
    for (int i = 0; i < nc; i++) {
      const int f = cols[3 * i], a = cols[3 * i + 1], b = cols[3 * i + 2];
      const int nTx = tnx[f], nTy = tny[f];
      int d = (a > 0) + (b > 0);
      if (o > 0) {
        const int pa = dfx[o] ? nTx - 1 - a : a, pb = dfy[o] ? nTy - 1 - b : b;
        d += 1 + (pa > 0) + (pa + 1 < nTx) + (pb > 0) + (pb + 1 < nTy);
      }
      left[o * nc + i] = d;
    }
  int nh = 0, nslot = 0;
  for (int g = 0; g < grid; g++) ov_push(slots, &nslot, 0, g);
  for (int i = 0; i < nc; i++)
    if (left[i] == 0) ov_push(heap, &nh, 0, i);
  ne = 0;
  while (nh > 0) {
    const ov_item it = ov_pop(heap, &nh);
    const int t = it.id, o = t / nc, i = t - o * nc;
    const int f = cols[3 * i], a = cols[3 * i + 1], b = cols[3 * i + 2];
    const int nTx = tnx[f], nTy = tny[f], cb = pbase[f];
    const ov_item sl = ov_pop(slots, &nslot);
    const long long st0 = it.key > sl.key ? it.key : sl.key;
    const long long fin = st0 + (fz[f] + TX + TY - 2);
    ov_push(slots, &nslot, fin, sl.id);
    if (fin > slotend[o]) slotend[o] = fin;
    const int pa = dfx[o] ? nTx - 1 - a : a, pb = dfy[o] ? nTy - 1 - b : b;
    out[5 * ne] = o;
    out[5 * ne + 1] = f;
    out[5 * ne + 2] = a;
    out[5 * ne + 3] = b;
    out[5 * ne + 4] = cb + pa * nTy + pb;
    ne++;
    emitted[o]++;
    /* dependents: downwind in the slot */
    for (int k = 0; k < 2; k++) {
      const int a2 = a + (k == 0), b2 = b + (k == 1);
      if (a2 >= nTx || b2 >= nTy) continue;
      const int j = o * nc + cidx[cb + a2 * nTy + b2];
      const long long r = st0 + (k == 0 ? TX : TY) + LAG;
      if (r > ready[j]) ready[j] = r;
      if (--left[j] == 0 && (o < 2 || emitted[o - 2] == nc)) ov_push(heap, &nh, ready[j], j);
    }
    /* dependents in the next slot: the tiles at and next to this one */
    if (o + 1 < ns) {
      static const int dq[5][2] = {{0, 0}, {1, 0}, {-1, 0}, {0, 1}, {0, -1}};
      for (int k = 0; k < 5; k++) {
        const int qa = pa + dq[k][0], qb = pb + dq[k][1];
        if (qa < 0 || qa >= nTx || qb < 0 || qb >= nTy) continue;
        const int va = dfx[o + 1] ? nTx - 1 - qa : qa, vb = dfy[o + 1] ? nTy - 1 - qb : qb;
        const int j = (o + 1) * nc + cidx[cb + va * nTy + vb];
        if (fin > ready[j]) ready[j] = fin;
        if (--left[j] == 0 && (o + 1 < 2 || emitted[o - 1] == nc)) ov_push(heap, &nh, ready[j], j);
      }
    }
    /* slot o complete: the slot two later may start */
    if (emitted[o] == nc && o + 2 < ns)
      for (int j = (o + 2) * nc; j < (o + 3) * nc; j++) {
        if (slotend[o] > ready[j]) ready[j] = slotend[o];
        if (left[j] == 0) ov_push(heap, &nh, ready[j], j);
      }
  }
done:
  free(pbase);
  free(cidx);
  free(left);
  free(ready);
  free(heap);
  free(slots);
  return ne;
}

/* The inputs a segment's ticket table is made from, in `key` (NULL: only
 * the length): tile, grid, orderings, per-field tiles and depths, the
 * active columns in ticket order. */
static int ov_key_fill(const pdfs_ctx *c, int sn0, int sn1, int grid, int *key)
{
  const int n = 6 + 3 * c->nf + 3 * c->ncols;
  if (!key) return n;
  key[0] = c->tile;
  key[1] = grid;
  key[2] = sn0;
  key[3] = sn1;
  key[4] = c->nf;
  key[5] = c->ncols;
  for (int f = 0; f < c->nf; f++) {
    key[6 + 3 * f] = c->h_fields[f].nTx;
    key[7 + 3 * f] = c->h_fields[f].nTy;
    key[8 + 3 * f] = c->h_fields[f].Z;
  }
  memcpy(key + 6 + 3 * c->nf, c->h_cols, (size_t)3 * c->ncols * sizeof(int));
  return n;
}

static int ov_build(pdfs_ctx *c, int seg, int sn0, int sn1, int grid)
{
  int TX, TY;
  pdfsk_tile_dims(c->tile, &TX, &TY);
  const int ns = sn1 - sn0 + 1, nc = c->ncols, nt = ns * nc;
  const int kn = ov_key_fill(c, sn0, sn1, grid, NULL);
  int *key = (int *)malloc((size_t)kn * sizeof(int));
  if (!key) return fail(PDFS_ENOMEM, "overlap plan: host alloc");
  ov_key_fill(c, sn0, sn1, grid, key);
  ov_table *tabs = c->ov_cache[seg];
  int e = -1, built = 0, rc = PDFS_OK;
  int *tnx = NULL, *out = NULL;
  /* a table ordered from the same inputs: use it */
  for (int i = 0; i < PDFS_OV_TABLES && e < 0; i++)
    if (tabs[i].key_n == kn && tabs[i].nt == nt &&
        !memcmp(key, tabs[i].key, (size_t)kn * sizeof(int)))
      e = i;
  if (e < 0) {
    /* else order the tickets into the least recently used table (its
     * upload follows, on the stream, any launch still reading it) */
    e = 0;
    for (int i = 1; i < PDFS_OV_TABLES; i++)
      if (tabs[i].used < tabs[e].used) e = i;
    ov_table *t = &tabs[e];
    t->key_n = 0;
    built = 1;
    if (kn > t->key_cap) {
      free(t->key);
      t->key_cap = 0;
      if (!(t->key = (int *)malloc((size_t)kn * sizeof(int)))) {
        rc = fail(PDFS_ENOMEM, "overlap plan: host alloc");
        goto done;
      }
      t->key_cap = kn;
    }
    tnx = (int *)malloc((size_t)(c->nf > 0 ? c->nf : 1) * 3 * sizeof(int));
    out = (int *)malloc((size_t)5 * (nt > 0 ? nt : 1) * sizeof(int));
    if (!tnx || !out) {
      rc = fail(PDFS_ENOMEM, "overlap plan: host alloc");
      goto done;
    }
    for (int f = 0; f < c->nf; f++) {
      tnx[f] = c->h_fields[f].nTx;
      tnx[c->nf + f] = c->h_fields[f].nTy;
      tnx[2 * c->nf + f] = c->h_fields[f].Z;
    }
    const int ne = pdfsi_ov_order(c->nf, tnx, tnx + c->nf, tnx + 2 * c->nf, nc, c->h_cols, TX,
                                  TY, sn0, sn1, grid, out);
    if (ne != nt) {
      rc = fail(PDFS_EINVAL, "overlap plan: %d of %d tickets ordered", ne, nt);
      goto done;
    }
    if (nt > t->cap) {
      if (hipStreamSynchronize(c->st) != hipSuccess) {
        rc = fail(PDFS_EHIP, "overlap plan: synchronize");
        goto done;
      }
      hipFree(t->tab);
      t->tab = NULL;
      t->cap = 0;
      if (hipMalloc((void **)&t->tab, (size_t)5 * nt * sizeof(int)) != hipSuccess) {
        rc = fail(PDFS_ENOMEM, "overlap plan: device alloc");
        goto done;
      }
      t->cap = nt;
    }
    if (hipMemcpyAsync(t->tab, out, (size_t)5 * nt * sizeof(int), hipMemcpyHostToDevice,
                       c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess) {
      rc = fail(PDFS_EHIP, "overlap plan: upload");
      goto done;
    }
    t->nt = nt;
    memcpy(t->key, key, (size_t)kn * sizeof(int));
    t->key_n = kn;
  }
  tabs[e].used = ++c->ov_clock;
  /* a captured launch of this segment holds another table's pointer and size */
  if ((built || e != c->ov_cur[seg]) && c->gexec[seg]) {
    hipGraphExecDestroy(c->gexec[seg]);
    c->gexec[seg] = NULL;
  }
  c->ov_cur[seg] = e;
  c->ov_valid[seg] = 1;
done:
  free(key);
  free(tnx);
  free(out);
  return rc;
}

/* Whether the automatic choice takes overlapped passes for the active plan:
 * at most PDFS_OV_AUTO_FIELDS fields, and the longest field's critical path
 * through one ordering (its upwind column lags plus one column) at least
 * PDFS_OV_AUTO_DEPTH (in tenths) of the ordering's column steps per resident
 * workgroup. Overlap then fills the drain of one ordering with the start of
 * the next; with more work per workgroup, or many fields, it does not pay
 * (measured: one 512^3 field 1.38x kernel time, two 1.13x, eight 0.95x;
 * 32 fields of 128^3 0.65x). */
static int ov_auto(const pdfs_ctx *c, int grid)
{
  int TX, TY;
  pdfsk_tile_dims(c->tile, &TX, &TY);
  const double lx = TX + PDFS_TICKET_EDGE_STEPS, ly = TY + PDFS_TICKET_EDGE_STEPS;
  int nact, last = -1;
  double work = 0.0, cp = 0.0;
  for (int i = 0; i < c->ncols; i++) {
    const int f = c->h_cols[3 * i];
    const pdfsk_field *fd = &c->h_fields[f];
    const double col = fd->Z + TX + TY - 2;
    work += col;
    if (f != last) {
      /* the columns of a field need not be adjacent in ticket order */
      last = f;
      const double p = (fd->nTx - 1) * lx + (fd->nTy - 1) * ly + col;
      if (p > cp) cp = p;
    }
  }
  nact = c->nact;
  if (nact == 0 || nact > PDFS_OV_AUTO_FIELDS) return 0;
  const double slots = grid < c->ncols ? grid : c->ncols;
  return 10.0 * cp * slots >= (double)PDFS_OV_AUTO_DEPTH * work;
}

/* Face entries of a batch's edge allocation for a tile (one row per column
 * step, z + TX + TY - 1 steps, `words` entries per value); 0 if the batch
 * does not fit the kernel's 32-bit buffer offsets (2 GiB with padding). */
static int plan_fits_w(int nf, const int *dims, int tile, int words)
{
  int TX, TY;
  pdfsk_tile_dims(tile, &TX, &TY);
  size_t ex = 0, ey = 0;
  for (int f = 0; f < nf; f++) {
    const int x = dims[3 * f], y = dims[3 * f + 1], z = dims[3 * f + 2];
    const size_t nc = (size_t)((x + TX - 1) / TX) * (size_t)((y + TY - 1) / TY);
    ex += nc * (size_t)(z + TX + TY - 1) * TY * words;
    ey += nc * (size_t)(z + TX + TY - 1) * TX * words;
  }
  return (ex + ey + 4 * PDFSK_EDGE_PAD) * 8 < (1ull << 31);
}

static int plan_fits(int nf, const int *dims, int tile) { return plan_fits_w(nf, dims, tile, 1); }

/* keep: re-tile the same batch mid-solve, keeping the per-field change
 * counters and unchanged-ordering runs (they do not depend on the tile).
 * words: 1 for FP32 fields, 2 for FP64 fields; inits (polish) or NULL. */
static int plan_build_ex(pdfs_ctx *c, int nf, void *const *fields, const float *const *inits,
                         const float *const *slows, const int *dims, int tile, int keep,
                         int words)
{
  if (nf < 1) return fail(PDFS_EINVAL, "empty batch");
  int TX, TY;
  pdfsk_tile_dims(tile, &TX, &TY);
  if (!plan_fits_w(nf, dims, tile, words))
    return fail(PDFS_EINVAL, "batch too large for one sweep (edge buffers exceed 2 GiB)");
  if (nf > c->nf || !c->h_fields) {
    free(c->h_fields);
    free(c->prev_changed);
    hipHostFree(c->h_streak);
    c->h_fields = (pdfsk_field *)calloc((size_t)nf, sizeof(pdfsk_field));
    c->prev_changed = (unsigned *)calloc((size_t)nf, sizeof(unsigned));
    c->h_streak = NULL;
    hipHostFree(c->h_changed);
    c->h_changed = NULL;
    if (!c->h_fields || !c->prev_changed ||
        hipHostMalloc((void **)&c->h_changed, (size_t)nf * sizeof(unsigned), 0) != hipSuccess ||
        hipHostMalloc((void **)&c->h_streak, (size_t)2 * nf * sizeof(unsigned), 0) !=
            hipSuccess)
      return fail(PDFS_ENOMEM, "plan: host alloc");
    memset(c->h_streak, 0, (size_t)2 * nf * sizeof(unsigned));
  }
  const long long maxpts = words == 2 ? PDFSK_MAX_POINTS64 : PDFSK_MAX_POINTS;
  size_t ex = 0, ey = 0;
  int ncols = 0;
  for (int f = 0; f < nf; f++) {
    int x = dims[3 * f], y = dims[3 * f + 1], z = dims[3 * f + 2];
    if (!fields[f] || x < 1 || y < 1 || z < 1)
      return fail(PDFS_EINVAL, "field %d: bad pointer or dims", f);
    if ((long long)x * y * z > maxpts)
      return fail(PDFS_EINVAL, "field %d: %lld points exceed the %s sweep limit %lld", f,
                  (long long)x * y * z, words == 2 ? "FP64" : "FP32", maxpts);
    pdfsk_field *fd = &c->h_fields[f];
    fd->f = fields[f];
    fd->init = inits ? inits[f] : NULL;
    fd->slow = slows ? slows[f] : NULL;
    if (slows && !fd->slow) return fail(PDFS_EINVAL, "field %d: no slowness", f);
    fd->X = x;
    fd->Y = y;
    fd->Z = z;
    fd->nTx = (x + TX - 1) / TX;
    fd->nTy = (y + TY - 1) / TY;
    fd->ex_off = (uint32_t)ex;
    fd->ey_off = (uint32_t)ey;
    fd->fxk = c->fxk ? c->fxk[f] : 0;
    size_t nc = (size_t)fd->nTx * fd->nTy;
    /* one row of face entries per column step (z + TX + TY - 1 steps) */
    ex += nc * (size_t)(z + TX + TY - 1) * TY * words;
    ey += nc * (size_t)(z + TX + TY - 1) * TX * words;
    ncols += (int)nc;
  }
  /* overlapped passes need every field to tile exactly (each ordering then
   * tiles the same physical columns) and a second edge region */
  int ov = 1;
  for (int f = 0; f < nf && ov; f++)
    ov = dims[3 * f] % TX == 0 && dims[3 * f + 1] % TY == 0;
  if (ov && (2 * (ex + ey) + 4 * PDFSK_EDGE_PAD) * 8 >= (1ull << 31)) ov = 0;
  c->ov_ok = ov;
  c->ov_rx = ov ? ex * 8 : 0;
  c->ov_ry = ov ? ey * 8 : 0;
  for (int i = 0; i < 3; i++) c->ov_valid[i] = 0;
  if (ov) {
    ex *= 2;
    ey *= 2;
    if ((size_t)8 * ncols > c->ov_done_cap) {
      HIPCHK(hipStreamSynchronize(c->st));
      hipFree(c->ov_done);
      c->ov_done = NULL;
      c->ov_done_cap = 0;
      HIPCHK(hipMalloc((void **)&c->ov_done, (size_t)8 * ncols * sizeof(unsigned)));
      /* epochs start at 1: zeroed flags are never taken for a finished slot */
      HIPCHK(hipMemsetAsync(c->ov_done, 0, (size_t)8 * ncols * sizeof(unsigned), c->st));
      c->ov_done_cap = (size_t)8 * ncols;
    }
    c->ov_ncolp = ncols;
    if ((size_t)8 * nf > c->ov_chg_cap) {
      HIPCHK(hipStreamSynchronize(c->st));
      hipFree(c->ov_chg);
      c->ov_chg = NULL;
      c->ov_chg_cap = 0;
      HIPCHK(hipMalloc((void **)&c->ov_chg, (size_t)8 * nf * sizeof(unsigned)));
      c->ov_chg_cap = (size_t)8 * nf;
    }
    /* zero between launches (each launch leaves them zero) */
    KCHK(pdfsk_fill_u32(c->ov_chg, 0u, (size_t)8 * nf, c->st));
  }
  if (ex > c->cap_ex || ey > c->cap_ey) {
    /* one allocation: the x-face part, then the y-face part, each with a
     * padding of PDFSK_EDGE_PAD entries on both sides; the kernel's single
     * buffer descriptor covers both, so one instruction moves a column's
     * west and south entries. Entries carry the epoch they were written
     * in, which only grows, so entries of earlier plans are never taken
     * for current ones. */
    const size_t cx = ex > c->cap_ex ? ex : c->cap_ex, cy = ey > c->cap_ey ? ey : c->cap_ey;
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->bt.edge_x) hipFree(c->bt.edge_x);
    c->bt.edge_x = c->bt.edge_y = NULL;
    c->cap_ex = c->cap_ey = 0;
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
  if (nf > c->cap_md) {
    HIPCHK(hipStreamSynchronize(c->st));
    hipFree(c->bt.maxdiff);
    hipHostFree(c->h_maxdiff);
    c->bt.maxdiff = NULL;
    c->h_maxdiff = NULL;
    c->cap_md = 0;
    if (hipMalloc((void **)&c->bt.maxdiff, (size_t)2 * nf * sizeof(unsigned long long)) !=
            hipSuccess ||
        hipHostMalloc((void **)&c->h_maxdiff, (size_t)2 * nf * sizeof(unsigned long long), 0) !=
            hipSuccess)
      return fail(PDFS_ENOMEM, "plan: maxdiff alloc");
    c->cap_md = nf;
  }
  free(c->h_cols);
  free(c->h_cols_all);
  c->h_cols = (int *)malloc((size_t)3 * ncols * sizeof(int));
  c->h_cols_all = (int *)malloc((size_t)3 * ncols * sizeof(int));
  c->ncols_all = 0;
  if (!c->h_cols || !c->h_cols_all) return fail(PDFS_ENOMEM, "plan: host alloc");
  c->nf = nf;
  c->tile = tile;
  c->words = words;
  HIPCHK(hipMemcpyAsync(c->bt.fields, c->h_fields, (size_t)nf * sizeof(pdfsk_field),
                        hipMemcpyHostToDevice, c->st));
  if (!keep) {
    KCHK(pdfsk_fill_u32(c->bt.changed, 0u, (size_t)nf, c->st));
    KCHK(pdfsk_fill_u32(c->bt.streak, 0u, (size_t)2 * nf, c->st));
    memset(c->prev_changed, 0, (size_t)nf * sizeof(unsigned));
  }
  c->bt.maxdiff_last = c->bt.maxdiff + c->cap_md;
  /* a re-tile mid-pass keeps the pass's largest changes as it keeps the
   * change counters */
  if (!keep) KCHK(pdfsk_fill_u32((uint32_t *)c->bt.maxdiff, 0u, (size_t)4 * c->cap_md, c->st));
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
  return plan_build_ex(c, nf, (void *const *)fields, NULL, NULL, dims, tile, 0, 1);
}

int pdfsi_set_overlap64(pdfs_ctx *c, int overlap)
{
  if (!c || overlap < PDFS_OVERLAP_OFF || overlap > PDFS_OVERLAP_ALWAYS) return PDFS_EINVAL;
  c->overlap64 = overlap;
  return PDFS_OK;
}

int pdfsi_col_resident(int tile, int mode)
{
  if (tile < 0 || tile >= PDFSK_TILE_COUNT || mode < 0 || mode >= PDFSK_MODE_COUNT) return -1;
  return pdfsk_col_max_resident(tile, mode);
}

static int launch_grid(pdfs_ctx *c, int mode)
{
  int grid = c->resident[c->tile][mode];
  if (grid <= 0) return -1;
  return grid > c->ncols ? (c->ncols > 0 ? c->ncols : 1) : grid;
}

/* Orderings sn0..sn1 of the current plan in a sweep mode, optionally
 * profiled per launch or replayed from a captured graph (a whole pass, or
 * one of its halves). */
static int ov_build(pdfs_ctx *c, int seg, int sn0, int sn1, int grid);

/* Whether whole passes (and their halves) of the current plan run as
 * overlapped launches in `mode` (automatic: ov_auto). */
static int ov_takes(pdfs_ctx *c, int mode, int grid)
{
  return c->ov_use && c->ov_ok && grid > 0 && mode >= 0 && mode < PDFSK_MODE_COUNT &&
         (c->ov_use == PDFS_OVERLAP_ALWAYS || ov_auto(c, grid));
}

static int run_orderings(pdfs_ctx *c, int mode, const pdfs_sp32 *sp, const pdfs_sp64 *sp64,
                         int sn0, int sn1, int profile, int graph, double *kernel_ms)
{
  int grid = launch_grid(c, mode);
  if (grid <= 0) return fail(PDFS_EHIP, "column kernel cannot be resident");
  const int gi = sn0 == 1 && sn1 == 8 ? 0 : sn0 == 1 && sn1 == 4 ? 1 : sn0 == 5 && sn1 == 8 ? 2 : -1;
  /* overlapped: the segment's orderings in one launch */
  const int ov = gi >= 0 && ov_takes(c, mode, grid);
  c->last_ov = ov;
  if (ov) {
    int rc;
    if (!c->ov_valid[gi] && (rc = ov_build(c, gi, sn0, sn1, grid))) return rc;
    const ov_table *tb = &c->ov_cache[gi][c->ov_cur[gi]];
    if (graph && !profile) {
      const int key = 1000 + mode;
      if (!c->gexec[gi] || c->g_tile[gi] != c->tile || c->g_mode[gi] != key ||
          c->g_grid[gi] != grid) {
        if (c->gexec[gi]) {
          hipGraphExecDestroy(c->gexec[gi]);
          c->gexec[gi] = NULL;
        }
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        if (pdfsk_sweep_batch_ov(&c->bt, sn0, sn1, tb->tab, tb->nt, c->ov_ncolp,
                                 c->ov_done, c->bt.state + 16, c->ov_chg, c->ov_rx, c->ov_ry,
                                 mode, sp, sp64, c->tile, grid, c->st)) {
          hipStreamEndCapture(c->st, &g);
          return fail(PDFS_EHIP, "launch during capture failed");
        }
        HIPCHK(hipStreamEndCapture(c->st, &g));
        hipError_t e = hipGraphInstantiate(&c->gexec[gi], g, NULL, NULL, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return fail(PDFS_EHIP, "graph instantiate: %s", hipGetErrorString(e));
        c->g_tile[gi] = c->tile;
        c->g_mode[gi] = key;
        c->g_grid[gi] = grid;
      }
      HIPCHK(hipGraphLaunch(c->gexec[gi], c->st));
      c->launches++;
      return PDFS_OK;
    }
    if (profile) HIPCHK(hipEventRecord(c->kev[0], c->st));
    KCHK(pdfsk_sweep_batch_ov(&c->bt, sn0, sn1, tb->tab, tb->nt, c->ov_ncolp,
                              c->ov_done, c->bt.state + 16, c->ov_chg, c->ov_rx, c->ov_ry, mode,
                              sp, sp64, c->tile, grid, c->st));
    c->launches++;
    if (profile) {
      HIPCHK(hipEventRecord(c->kev[1], c->st));
      HIPCHK(hipEventSynchronize(c->kev[1]));
      float tk = 0.f;
      hipEventElapsedTime(&tk, c->kev[0], c->kev[1]);
      *kernel_ms += tk;
    }
    return PDFS_OK;
  }
  if (graph && gi >= 0 && !profile) {
    if (!c->gexec[gi] || c->g_tile[gi] != c->tile || c->g_mode[gi] != mode ||
        c->g_grid[gi] != grid) {
      if (c->gexec[gi]) {
        hipGraphExecDestroy(c->gexec[gi]);
        c->gexec[gi] = NULL;
      }
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
      for (int sn = sn0; sn <= sn1; sn++)
        if (pdfsk_sweep_batch(&c->bt, sn, mode, sp, sp64, c->tile, grid, c->st)) {
          hipStreamEndCapture(c->st, &g);
          return fail(PDFS_EHIP, "launch during capture failed");
        }
      HIPCHK(hipStreamEndCapture(c->st, &g));
      hipError_t e = hipGraphInstantiate(&c->gexec[gi], g, NULL, NULL, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) return fail(PDFS_EHIP, "graph instantiate: %s", hipGetErrorString(e));
      c->g_tile[gi] = c->tile;
      c->g_mode[gi] = mode;
      c->g_grid[gi] = grid;
    }
    HIPCHK(hipGraphLaunch(c->gexec[gi], c->st));
    c->launches += sn1 - sn0 + 1;
    return PDFS_OK;
  }
  for (int sn = sn0; sn <= sn1; sn++) {
    if (profile) HIPCHK(hipEventRecord(c->kev[2 * (sn - 1)], c->st));
    KCHK(pdfsk_sweep_batch(&c->bt, sn, mode, sp, sp64, c->tile, grid, c->st));
    c->launches++;
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

/* Per-field points changed since the last read; returns the timeout flag.
 * maxdiff (FP64 plans, or NULL): per-field largest change since the last
 * read, and last (or NULL) that of ordering 8 alone; both are then reset. */
static int read_changed_md3(pdfs_ctx *c, long long *delta, unsigned *err, double *maxdiff,
                            double *last, int as_int);
static int read_changed_md2(pdfs_ctx *c, long long *delta, unsigned *err, double *maxdiff,
                            double *last)
{
  return read_changed_md3(c, delta, err, maxdiff, last, 0);
}

/* as_int: the maxdiff slots hold integers (fixed point: quanta), not FP64
 * bits */
static int read_changed_md3(pdfs_ctx *c, long long *delta, unsigned *err, double *maxdiff,
                            double *last, int as_int)
{
  /* one launch writes the counters into pinned host memory (and zeroes the
   * largest changes, which maxdiff_last follows at cap_md) */
  KCHK(pdfsk_read_counters(c->bt.changed, c->bt.err, c->bt.streak,
                           maxdiff ? c->bt.maxdiff : NULL, c->nf, c->cap_md, c->h_changed,
                           c->h_err, c->h_streak, c->h_maxdiff, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int f = 0; f < c->nf; f++) {
    delta[f] = (long long)(unsigned)(c->h_changed[f] - c->prev_changed[f]);
    c->prev_changed[f] = c->h_changed[f];
    if (maxdiff) {
      double a, b;
      if (as_int) {
        a = (double)c->h_maxdiff[f];
        b = (double)c->h_maxdiff[c->cap_md + f];
      } else {
        memcpy(&a, &c->h_maxdiff[f], sizeof(double));
        memcpy(&b, &c->h_maxdiff[c->cap_md + f], sizeof(double));
      }
      maxdiff[f] = b > a ? b : a;
      if (last) last[f] = b;
    }
  }
  *err = c->h_err[0];
  if (*err) {
    unsigned z = 0;
    hipMemcpyAsync(c->bt.err, &z, sizeof z, hipMemcpyHostToDevice, c->st);
    hipStreamSynchronize(c->st);
  }
  return PDFS_OK;
}

static int read_changed_md(pdfs_ctx *c, long long *delta, unsigned *err, double *maxdiff)
{
  return read_changed_md2(c, delta, err, maxdiff, NULL);
}

static int read_changed(pdfs_ctx *c, long long *delta, unsigned *err)
{
  return read_changed_md2(c, delta, err, NULL, NULL);
}

/* FP64 column variant: explicit (TX + TY <= 32), else 8x8. Measured per
 * polish ordering: 512^3 1.57 ms (16x16 1.88, 16x8 1.83, 16x4 1.73);
 * 128x128x512 0.36 ms (16x16 0.56). */
static int tile64_of(const pdfs_opts *o)
{
  if (o && o->tile64 >= 0 && o->tile64 < PDFSK_TILE_COUNT &&
      pdfsk_tile_supports(o->tile64, PDFSK_MODE_F64_REF))
    return o->tile64;
  return PDFSK_TILE_8x8;
}

/* FP64 column sweeps of nf dense FP64 fields in mode PDFSK_MODE_F64_*.
 * fixed_passes > 0: exactly that many passes; else passes until a pass's
 * largest change (stop_on_last: that of its last ordering) is <= tol, at
 * most maxp (PDFS_ENOTCONV beyond). A field
 * whose last ordering changed nothing is at its fixed point: the device
 * skips its later orderings, and the pass after it, which would change
 * nothing, is counted but not launched. field_passes / field_maxdiff (or
 * NULL) receive each field's pass count and its last pass's largest change.
 * st: passes (the largest), orderings, point_updates, last_maxdiff (the
 * largest of the fields' last passes), ms_sweeps. */
static int sweep64_batch(pdfs_ctx *c, int mode, int nf, double *const *fields,
                         const float *const *inits, const float *const *slows, const int *dims,
                         const pdfs_sp64 *sp64,
                         int tile, int fixed_passes, double tol, int stop_on_last, int maxp,
                         int graph, int profile, int overlap, pdfs_stats *st, int *field_passes,
                         double *field_maxdiff)
{
  double kms = 0.0;
  pdfs_stats s;
  memset(&s, 0, sizeof s);
  int rc = plan_build_ex(c, nf, (void *const *)fields, inits, slows, dims, tile, 0, 2);
  if (rc) return rc;
  /* overlapped passes (whole passes end at ordering 8, so the polish's stop
   * on the last ordering's change reads that slot's own maximum) */
  c->ov_use = overlap;
  const long long launches0 = c->launches;
  long long *delta = (long long *)calloc((size_t)nf, sizeof(long long));
  double *md = (double *)calloc((size_t)2 * nf, sizeof(double));
  double *md8 = md ? md + nf : NULL;
  double *fmd = (double *)calloc((size_t)nf, sizeof(double));
  int *active = (int *)malloc((size_t)nf * sizeof(int));
  int *fp = (int *)calloc((size_t)nf, sizeof(int));
  if (!delta || !md || !fmd || !active || !fp) {
    rc = fail(PDFS_ENOMEM, "sweep64: host alloc");
    goto out;
  }
  for (int f = 0; f < nf; f++) active[f] = 1;
  if (fixed_passes > 0) maxp = fixed_passes;
  if (maxp < 1) maxp = 1;
  int nactive = nf;
  HIPGO(hipEventRecord(c->ev0, c->st));
  for (int p = 0; p < maxp && nactive > 0; p++) {
    long long npts = 0;
    for (int f = 0; f < nf; f++)
      if (active[f]) npts += (long long)dims[3 * f] * dims[3 * f + 1] * dims[3 * f + 2];
    if ((rc = run_orderings(c, mode, NULL, sp64, 1, 8, profile, graph, &kms))) break;
    unsigned err = 0;
    if ((rc = read_changed_md2(c, delta, &err, md, md8))) break;
    if (err) {
      rc = fail(PDFS_ETIMEOUT, "column pipeline: edge wait timed out");
      break;
    }
    s.orderings += 8;
    s.point_updates += 8 * npts;
    int changed_set = 0;
    for (int f = 0; f < nf; f++) {
      if (!active[f]) continue;
      const long long nfp = (long long)dims[3 * f] * dims[3 * f + 1] * dims[3 * f + 2];
      const unsigned run = c->h_streak[2 * f + 1];
      long long skipped = run > PDFSK_CONVERGED_RUN ? (long long)(run - PDFSK_CONVERGED_RUN) : 0;
      if (c->last_ov) skipped = run >= 2 ? (run < 8u ? run : 8) - 2 : 0;
      if (skipped > 8) skipped = 8;
      s.point_updates -= skipped * nfp;
      fp[f] = p + 1;
      fmd[f] = md[f];
      int done =
          fixed_passes > 0 ? p + 1 >= fixed_passes : (stop_on_last ? md8[f] : md[f]) <= tol;
      if (!done && run >= PDFSK_CONVERGED_RUN) {
        /* at its fixed point: every later pass changes nothing */
        if (fixed_passes > 0) {
          fp[f] = fixed_passes;
          done = 1;
        } else if (p + 2 <= maxp) {
          fp[f] = p + 2;
          fmd[f] = 0.0;
          done = 1;
        }
      }
      if (done) {
        active[f] = 0;
        nactive--;
        changed_set = 1;
      }
    }
    if (changed_set && nactive > 0 && (rc = plan_upload_cols(c, active))) break;
  }
  HIPGO(hipEventRecord(c->ev1, c->st));
  HIPGO(hipEventSynchronize(c->ev1));
  {
    float t = 0.f;
    hipEventElapsedTime(&t, c->ev0, c->ev1);
    s.ms_sweeps = t;
    s.ms_total = t;
  }
  for (int f = 0; f < nf; f++) {
    if (fp[f] > s.passes) s.passes = fp[f];
    if (fmd[f] > s.last_maxdiff || fmd[f] != fmd[f]) s.last_maxdiff = fmd[f];
    if (field_passes) field_passes[f] = fp[f];
    if (field_maxdiff) field_maxdiff[f] = fmd[f];
  }
  s.tile = tile;
  s.ms_sweep_kernels = kms;
  s.sweep_launches = (int)(c->launches - launches0);
  c->ov_use = 0;
  if (rc == PDFS_OK && nactive > 0)
    rc = fail(PDFS_ENOTCONV, "%d field(s) not converged after %d passes", nactive, maxp);
out:
  if (st) *st = s;
  free(delta);
  free(md);
  free(fmd);
  free(active);
  free(fp);
  return rc;
}

/* Dense FP64 working fields in the context's buffer: offsets (in doubles,
 * 256-byte aligned) of n fields; grows the buffer as needed. */
static int pbuf_layout(pdfs_ctx *c, int n, const long long *npts, size_t *off)
{
  size_t tot = 0;
  for (int i = 0; i < n; i++) {
    off[i] = tot;
    tot += ((size_t)npts[i] + 31) / 32 * 32;
  }
  if (tot > c->pcap) {
    HIPCHK(hipStreamSynchronize(c->st));
    hipFree(c->pbuf);
    c->pbuf = NULL;
    c->pcap = 0;
    if (hipMalloc((void **)&c->pbuf, tot * sizeof(double)) != hipSuccess)
      return fail(PDFS_ENOMEM, "FP64 working fields: %zu bytes", tot * sizeof(double));
    c->pcap = tot;
  }
  return PDFS_OK;
}

/* FP64 polish of one group of fields that fits one FP64 plan. */
static int polish_group(pdfs_ctx *c, int k, float *const *fields, const float *const *inits,
                        const float *const *slows, const int *dims, const pdfs_sp64 *sp64,
                        double tol, int maxp, int tile, int graph, int profile, int overlap,
                        pdfs_stats *ps)
{
  long long *np = (long long *)malloc((size_t)k * sizeof(long long));
  size_t *off = (size_t *)malloc((size_t)k * sizeof(size_t));
  double **f64 = (double **)malloc((size_t)k * sizeof(double *));
  int rc = PDFS_OK;
  long long nmax = 0;
  if (!np || !off || !f64) {
    rc = fail(PDFS_ENOMEM, "polish: host alloc");
    goto out;
  }
  for (int i = 0; i < k; i++) {
    np[i] = (long long)dims[3 * i] * dims[3 * i + 1] * dims[3 * i + 2];
    if (np[i] > nmax) nmax = np[i];
  }
  if ((rc = pbuf_layout(c, k, np, off))) goto out;
  if (k > c->cap_conv) {
    HIPGO(hipStreamSynchronize(c->st));
    hipFree(c->conv_dev);
    free(c->conv_host);
    c->conv_dev = NULL;
    c->cap_conv = 0;
    c->conv_host = (pdfsk_conv *)malloc((size_t)k * sizeof(pdfsk_conv));
    if (!c->conv_host) {
      rc = fail(PDFS_ENOMEM, "polish: host alloc");
      goto out;
    }
    HIPGO(hipMalloc((void **)&c->conv_dev, (size_t)k * sizeof(pdfsk_conv)));
    c->cap_conv = k;
  }
  for (int i = 0; i < k; i++) {
    f64[i] = c->pbuf + off[i];
    c->conv_host[i].f32 = fields[i];
    c->conv_host[i].f64 = f64[i];
    c->conv_host[i].n = np[i];
  }
  HIPGO(hipMemcpyAsync(c->conv_dev, c->conv_host, (size_t)k * sizeof(pdfsk_conv),
                       hipMemcpyHostToDevice, c->st));
  HIPGO(hipStreamSynchronize(c->st));
  KGO(pdfsk_convert_batch(c->conv_dev, k, nmax, 0, c->st));
  {
    const int mode = slows ? PDFSK_MODE_F64_POL_SPEED
                           : sp64->iso ? PDFSK_MODE_F64_POL_ISO : PDFSK_MODE_F64_POL_ANISO;
    rc = sweep64_batch(c, mode, k, f64, inits, slows, dims, sp64, tile, 0, tol, 1, maxp, graph,
                       profile, overlap, ps, NULL, NULL);
    if (rc && rc != PDFS_ENOTCONV) goto out;
    const int rc0 = rc;
    rc = PDFS_OK;
    KGO(pdfsk_convert_batch(c->conv_dev, k, nmax, 1, c->st));
    HIPGO(hipStreamSynchronize(c->st));
    rc = rc0;
  }
out:
  free(np);
  free(off);
  free(f64);
  return rc;
}

/* FP64 polish of converged FP32 fields (those with an init): widen to FP64,
 * passes of u = min(init, G64(u)) on the column pipeline until a pass
 * changes no value by more than polish_tol * hmin in its last ordering, then
 * round back to FP32.
 * It removes the storage rounding that FP32 sweeps accumulate along
 * characteristics. Fields go in groups whose FP64 edge entries fit one plan;
 * a field beyond the FP64 column limit (2^28 points) is left as it is and
 * counted in polish_skipped. */
int pdfsi_polish(pdfs_ctx *c, int nf, float *const *fields, const float *const *inits,
                 const float *const *slows, const int *dims, double dx, double dy, double dz,
                 const pdfs_opts *o, pdfs_stats *s)
{
  if (!inits) return PDFS_OK;
  int tile = tile64_of(o);
  if (slows && !pdfsk_tile_supports(tile, PDFSK_MODE_F64_POL_SPEED)) tile = PDFSK_TILE_8x8;
  pdfs_sp64 sp64;
  make_sp64(dx, dy, dz, &sp64);
  const double hmin = dx < dy ? (dx < dz ? dx : dz) : (dy < dz ? dy : dz);
  const double tol = (o->polish_tol > 0 ? o->polish_tol : 2e-5) * hmin;
  const int maxp = o->max_passes > 0 ? o->max_passes : 100;
  float **gf = (float **)malloc((size_t)nf * sizeof(float *));
  const float **gi = (const float **)malloc((size_t)nf * sizeof(float *));
  const float **gs = (const float **)malloc((size_t)nf * sizeof(float *));
  int *gd = (int *)malloc((size_t)3 * nf * sizeof(int));
  int rc = PDFS_OK, k = 0;
  float tms = 0.f;
  if (!gf || !gi || !gs || !gd) {
    rc = fail(PDFS_ENOMEM, "polish: host alloc");
    goto out;
  }
  HIPGO(hipEventRecord(c->kev[16], c->st));
  for (int f = 0; f <= nf; f++) {
    if (f < nf) {
      if (!inits[f]) continue;
      if ((long long)dims[3 * f] * dims[3 * f + 1] * dims[3 * f + 2] > PDFSK_MAX_POINTS64 ||
          !plan_fits_w(1, dims + 3 * f, tile, 2)) {
        s->polish_skipped++;
        continue;
      }
      memcpy(gd + 3 * k, dims + 3 * f, 3 * sizeof(int));
    }
    /* run the group when field f does not fit with it, or at the end */
    if (k > 0 && (f == nf || !plan_fits_w(k + 1, gd, tile, 2))) {
      pdfs_stats ps;
      memset(&ps, 0, sizeof ps);
      const int r = polish_group(c, k, gf, gi, slows ? gs : NULL, gd, &sp64, tol, maxp, tile,
                                 o->graph, o->profile, o->overlap, &ps);
      if (ps.passes > s->polish_passes) s->polish_passes = ps.passes;
      s->polish_orderings += ps.orderings;
      s->polish_point_updates += ps.point_updates;
      s->ms_polish_kernels += ps.ms_sweep_kernels;
      if (ps.last_maxdiff > s->polish_maxdiff) s->polish_maxdiff = ps.last_maxdiff;
      if (r != PDFS_OK && r != PDFS_ENOTCONV) {
        rc = r;
        goto out;
      }
      if (r == PDFS_ENOTCONV) rc = r;
      if (f < nf) memcpy(gd, dims + 3 * f, 3 * sizeof(int));
      k = 0;
    }
    if (f < nf) {
      gf[k] = fields[f];
      gi[k] = inits[f];
      gs[k] = slows ? slows[f] : NULL;
      k++;
    }
  }
  HIPGO(hipEventRecord(c->kev[17], c->st));
  HIPGO(hipEventSynchronize(c->kev[17]));
  hipEventElapsedTime(&tms, c->kev[16], c->kev[17]);
  s->ms_polish += tms;
out:
  free(gf);
  free(gi);
  free(gs);
  free(gd);
  return rc;
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
  if ((rc = run_orderings(c, mode32(&sp), &sp, NULL, sn, sn, 0, 0, NULL))) return rc;
  long long d = 0;
  unsigned err = 0;
  if ((rc = read_changed(c, &d, &err))) return rc;
  if (err) return fail(PDFS_ETIMEOUT, "column pipeline: edge wait timed out");
  if (changed) *changed += d;
  return PDFS_OK;
}

/* FP64 residual judge of n fields in one launch; flagged[i] = failing points
 * of field i. flags[i] (optional) receive one bit per point. */
/* The judge of a batch; fxb (or NULL): the fields are fixed point, field i
 * holding q of value = fxb[i] + q fxd[i], judged with tolerance fxtol[i]. */
static int judge_batch_ex(pdfs_ctx *c, int n, const float *const *v, const float *const *init,
                          const float *const *slow, const int *dims, uint32_t *const *flags,
                          double dx, double dy, double dz, long long *flagged,
                          const double *fxb, const double *fxd, const double *fxtol);

int pdfsi_judge_batch(pdfs_ctx *c, int n, const float *const *v, const float *const *init,
                      const float *const *slow, const int *dims, uint32_t *const *flags,
                      double dx, double dy, double dz, long long *flagged)
{
  return judge_batch_ex(c, n, v, init, slow, dims, flags, dx, dy, dz, flagged, NULL, NULL, NULL);
}

static int judge_batch_ex(pdfs_ctx *c, int n, const float *const *v, const float *const *init,
                          const float *const *slow, const int *dims, uint32_t *const *flags,
                          double dx, double dy, double dz, long long *flagged,
                          const double *fxb, const double *fxd, const double *fxtol)
{
  if (fxb && (slow || dx != dy || dy != dz))
    return fail(PDFS_EINVAL, "judge: fixed-point fields need equal spacings and unit speed");
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
    if (!v[i] || !init[i] || (slow && !slow[i]) || x < 1 || y < 1 || z < 1)
      return fail(PDFS_EINVAL, "judge: field %d: bad pointer or dims", i);
    /* 32-bit element offsets in the judge kernel */
    if ((long long)x * y * z > PDFSK_MAX_POINTS)
      return fail(PDFS_EINVAL, "judge: field %d: %lld points exceed the limit %lld", i,
                  (long long)x * y * z, (long long)PDFSK_MAX_POINTS);
    pdfsk_judge_field *t = &c->jt_host[i];
    t->v = v[i];
    t->init = init[i];
    t->slow = slow ? slow[i] : NULL;
    t->flags = flags ? flags[i] : NULL;
    t->X = x;
    t->Y = y;
    t->Z = z;
    t->tile0 = (int)tiles;
    t->fx_base = fxb ? fxb[i] : 0.0;
    t->fx_d = fxb ? fxd[i] : 0.0;
    t->fx_tol = fxb ? fxtol[i] : 0.0;
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
  if (fxb)
    KCHK(pdfsk_judge_batch_fx(c->jt_dev, n, (long)tiles, dx, c->jc_dev, c->st));
  else
    KCHK(pdfsk_judge_batch_f32(c->jt_dev, n, (long)tiles, dx, dy, dz, slow != NULL, c->jc_dev,
                               c->st));
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
  int rc =
      pdfsi_judge_batch(c, 1, &field, &init, NULL, dims, flags ? &flags : NULL, dx, dy, dz, &fl);
  if (rc == PDFS_OK && flagged) *flagged = fl;
  return rc;
}

int pdfs_judge_fixed_dev(pdfs_ctx *c, const uint32_t *q, const float *init, int x, int y, int z,
                         double h, double base, double d, double tol, uint32_t *flags,
                         long long *flagged)
{
  if (!c || !q || !init || bad_h(h, h, h) || !(d > 0.0) || !(tol >= 0.0) || base != base)
    return fail(PDFS_EINVAL, "pdfs_judge_fixed_dev: bad arguments");
  const int dims[3] = {x, y, z};
  const float *v = (const float *)q;
  long long fl = 0;
  int rc = judge_batch_ex(c, 1, &v, &init, NULL, dims, flags ? &flags : NULL, h, h, h, &fl, &base,
                          &d, &tol);
  if (rc == PDFS_OK && flagged) *flagged = fl;
  return rc;
}

/* The FP32 batch solve without the spacing check (run_fsm and the CLI take
 * the caller's spacings as the reference does). */
/* Fixed-point storage of a batch (pdfs_opts.storage). fx_encode finds each
 * field's smallest non-NaN value (one reduction launch and a readback),
 * picks the field's quantum d = h 2^-k, k = pdfsk_fx_bits, and converts the
 * fields in place; *ok = 0 (nothing converted) when a field holds -inf.
 * fx_decode converts them back (the device table of the encode is kept). */
static int fx_encode(pdfs_ctx *c, int nf, float *const *fields, const int *dims, double h,
                     int *k, double *base, double *d, int *ok)
{
  *ok = 0;
  if (nf > c->cap_fx) {
    HIPCHK(hipStreamSynchronize(c->st));
    free(c->fx_host);
    hipFree(c->fx_dev);
    hipFree(c->fx_keys);
    hipHostFree(c->h_fx_keys);
    c->fx_host = NULL;
    c->fx_dev = NULL;
    c->fx_keys = NULL;
    c->h_fx_keys = NULL;
    c->cap_fx = 0;
    c->fx_host = (pdfsk_fx *)malloc((size_t)nf * sizeof(pdfsk_fx));
    if (!c->fx_host) return fail(PDFS_ENOMEM, "fixed point: host alloc");
    HIPCHK(hipMalloc((void **)&c->fx_dev, (size_t)nf * sizeof(pdfsk_fx)));
    HIPCHK(hipMalloc((void **)&c->fx_keys, (size_t)nf * sizeof(uint32_t)));
    HIPCHK(hipHostMalloc((void **)&c->h_fx_keys, (size_t)nf * sizeof(uint32_t), 0));
    c->cap_fx = nf;
  }
  long long nmax = 0;
  for (int f = 0; f < nf; f++) {
    const long long n = (long long)dims[3 * f] * dims[3 * f + 1] * dims[3 * f + 2];
    c->fx_host[f].f = fields[f];
    c->fx_host[f].n = n;
    c->fx_host[f].base = 0.0;
    c->fx_host[f].d = 1.0;
    c->fx_host[f].key = c->fx_keys + f;
    if (n > nmax) nmax = n;
  }
  HIPCHK(hipMemcpyAsync(c->fx_dev, c->fx_host, (size_t)nf * sizeof(pdfsk_fx),
                        hipMemcpyHostToDevice, c->st));
  KCHK(pdfsk_fill_u32(c->fx_keys, 0xFFFFFFFFu, (size_t)nf, c->st));
  KCHK(pdfsk_fx_min_batch(c->fx_dev, nf, nmax, c->st));
  HIPCHK(hipMemcpyAsync(c->h_fx_keys, c->fx_keys, (size_t)nf * sizeof(uint32_t),
                        hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int f = 0; f < nf; f++) {
    const uint32_t key = c->h_fx_keys[f];
    /* no non-NaN value, or all +inf: every point encodes as +inf */
    double b = key == 0xFFFFFFFFu ? 0.0 : (double)pdfsk_fx_key_value(key);
    if (b == -HUGE_VAL) return PDFS_OK;
    if (b == HUGE_VAL) b = 0.0;
    k[f] = pdfsk_fx_bits(dims[3 * f], dims[3 * f + 1], dims[3 * f + 2]);
    base[f] = b;
    d[f] = ldexp(h, -k[f]);
    c->fx_host[f].base = b;
    c->fx_host[f].d = d[f];
  }
  HIPCHK(hipMemcpyAsync(c->fx_dev, c->fx_host, (size_t)nf * sizeof(pdfsk_fx),
                        hipMemcpyHostToDevice, c->st));
  KCHK(pdfsk_fx_encode_batch(c->fx_dev, nf, nmax, c->st));
  *ok = 1;
  return PDFS_OK;
}

static int fx_decode(pdfs_ctx *c, int nf)
{
  long long nmax = 0;
  for (int f = 0; f < nf; f++)
    if (c->fx_host[f].n > nmax) nmax = c->fx_host[f].n;
  KCHK(pdfsk_fx_decode_batch(c->fx_dev, nf, nmax, c->st));
  return PDFS_OK;
}

static int batch_f32(pdfs_ctx *c, int nf, float *const *fields, const float *const *inits,
                     const float *const *slows, const int *dims, double dx, double dy,
                     double dz, const pdfs_opts *o, pdfs_stats *st, int *field_passes,
                     double *field_ms, long long *field_flagged);

int pdfs_solve_batch_f32_dev(pdfs_ctx *c, int nf, float *const *fields,
                             const float *const *inits, const int *dims, double dx, double dy,
                             double dz, const pdfs_opts *o, pdfs_stats *st, int *field_passes,
                             double *field_ms)
{
  if (!c || nf < 1 || !fields || !dims || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_solve_batch_f32_dev: bad arguments");
  return batch_f32(c, nf, fields, inits, NULL, dims, dx, dy, dz, o, st, field_passes, field_ms,
                   NULL);
}

int pdfs_solve_batch_speed_f32_dev(pdfs_ctx *c, int nf, float *const *fields,
                                   const float *const *inits, const float *const *slowness,
                                   const int *dims, double dx, double dy, double dz,
                                   const pdfs_opts *o, pdfs_stats *st, int *field_passes,
                                   double *field_ms)
{
  if (!c || nf < 1 || !fields || !slowness || !dims || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_solve_batch_speed_f32_dev: bad arguments");
  return batch_f32(c, nf, fields, inits, slowness, dims, dx, dy, dz, o, st, field_passes,
                   field_ms, NULL);
}

int pdfsi_batch_f32(pdfs_ctx *c, int nf, float *const *fields, const float *const *inits,
                    const int *dims, double dx, double dy, double dz, const pdfs_opts *o,
                    pdfs_stats *st, int *field_passes, long long *field_flagged)
{
  return batch_f32(c, nf, fields, inits, NULL, dims, dx, dy, dz, o, st, field_passes, NULL,
                   field_flagged);
}

static int batch_f32(pdfs_ctx *c, int nf, float *const *fields, const float *const *inits,
                     const float *const *slows, const int *dims, double dx, double dy,
                     double dz, const pdfs_opts *o, pdfs_stats *st, int *field_passes,
                     double *field_ms, long long *field_flagged)
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
  /* a slowness field: the speed mode, whose column variants exclude 32x16
   * (its aux delay line does not fit the LDS); no online tuning */
  const int speed = slows != NULL;
  if (o->tile >= 0 && o->tile < PDFSK_TILE_COUNT &&
      (!speed || pdfsk_tile_supports(o->tile, PDFSK_MODE_F32_SPEED))) {
    tile = o->tile;
  } else if (speed) {
    tile = default_tile(nf, dims[0], dims[1]);
    if (!pdfsk_tile_supports(tile, PDFSK_MODE_F32_SPEED)) tile = PDFSK_TILE_32x8;
  } else if (o->tile == PDFS_TILE_AUTO) {
    tile = tune_lookup(c, key);
    if (tile < 0) {
      /* the first candidate whose edge buffers fit */
      tuning = 1;
      for (tile = 0; tile < PDFSK_TUNE_COUNT - 1 && !plan_fits(nf, dims, tile); tile++) {
      }
    }
  } else {
    tile = default_tile(nf, dims[0], dims[1]);
  }
  pdfs_sp32 sp;
  make_sp32(dx, dy, dz, &sp);
  int m32 = speed ? PDFSK_MODE_F32_SPEED : mode32(&sp);
  int rc;
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
  const float **js = (const float **)malloc((size_t)nf * sizeof(float *));
  long long *jfl = (long long *)malloc((size_t)nf * sizeof(long long));
  /* fixed-point storage: per-field quantum exponent, base, quantum and judge
   * tolerance; [3 nf ..]: the same three for the fields judged together */
  int *fxk = (int *)malloc((size_t)nf * sizeof(int));
  double *fxv = (double *)malloc((size_t)8 * nf * sizeof(double));
  if (!delta || !pdel || !lastfl || !active || !certified || !jidx || !jdone || !jd || !jv ||
      !ji || !js || !jfl || !fxk || !fxv) {
    free(fxk);
    free(fxv);
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
    free(js);
    free(jfl);
    return fail(PDFS_ENOMEM, "batch: host alloc");
  }
  /* fixed-point storage (pdfs_opts.storage): encode in place, sweep in the
   * fixed-point mode, judge the fixed-point values, decode at the end */
  int fixed = 0, decoded = 0;
  double *fxb = fxv, *fxd = fxv + nf, *fxt = fxv + 2 * nf;
  double *jfb = fxv + 3 * nf, *jfd = fxv + 4 * nf, *jft = fxv + 5 * nf;
  /* fixed point: the largest drop (quanta) of the current pass, per field,
   * and of the last segment read back */
  double *pmq = fxv + 6 * nf, *fmd = fxv + 7 * nf;
  if (o->storage == PDFS_STORE_FIXED && !speed && o->fixed_passes <= 0 && sp.iso) {
    HIPGO(hipEventRecord(c->kev[16], c->st));
    if ((rc = fx_encode(c, nf, fields, dims, dx, fxk, fxb, fxd, &fixed))) goto out;
    HIPGO(hipEventRecord(c->kev[17], c->st));
    if (fixed) {
      HIPGO(hipEventSynchronize(c->kev[17]));
      float tc = 0.f;
      hipEventElapsedTime(&tc, c->kev[16], c->kev[17]);
      s.ms_convert = tc;
      m32 = PDFSK_MODE_FX_ISO;
      c->fxk = fxk;
      for (int f = 0; f < nf; f++) fxt[f] = o->judge_tol * dx + 4.0 * fxd[f];
    }
  }
  s.storage = fixed ? PDFS_STORE_FIXED : PDFS_STORE_F32;
  if ((rc = plan_build_ex(c, nf, (void *const *)fields, NULL, slows, dims, tile, 0, 1))) goto out;
  /* whole and half passes as one launch each where the plan allows it */
  c->ov_use = o->overlap;
  const long long launches0 = c->launches;
  for (int f = 0; f < nf; f++) {
    active[f] = 1;
    lastfl[f] = -1;  /* not judged yet */
    if (field_flagged) field_flagged[f] = 0;
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
   * few orderings before the pass ends (c4 3% faster, eight 512^3 fields
   * 2.7%). Also where the pass runs as overlapped launches, although the
   * two halves lose the overlap between them (one 512^3 field, which
   * certifies at the pass end: 16.5 ms against 15.9 without halves), so
   * that a field's certified result is the same alone and in a batch. The
   * plain stop stays a whole pass without change. */
  const int certifying = o->certify && inits && o->fixed_passes <= 0;
  const int cfrom = o->certify_from > 0 ? o->certify_from : 1;
  for (int p = 0; p < maxp && nactive > 0; p++) {
    int halves = 0;
    for (int f = 0; f < nf && certifying && !tuning && p >= cfrom && !halves; f++)
      halves = active[f] && inits[f] && lastfl[f] >= 0 &&
               lastfl[f] * 1024 <= (long long)dims[3 * f] * dims[3 * f + 1] * dims[3 * f + 2];
    const int nseg = halves ? 2 : 1;
    for (int f = 0; f < nf; f++) pdel[f] = 0;
    for (int f = 0; f < nf; f++) pmq[f] = 0.0;
    for (int g = 0; g < nseg && nactive > 0 && rc == PDFS_OK; g++) {
      const int sn0 = halves ? 1 + 4 * g : 1, sn1 = halves ? 4 + 4 * g : 8;
      const int nord = sn1 - sn0 + 1, last_seg = g == nseg - 1;
      long long npts = 0;
      for (int f = 0; f < nf; f++)
        if (active[f]) npts += (long long)dims[3 * f] * dims[3 * f + 1] * dims[3 * f + 2];
      if (tuning) {
        /* ordering 1 untimed, candidate v on ordering v + 2, the best on 8 */
        double cand_ms[PDFSK_TUNE_COUNT], tk = 0.0;
        if ((rc = run_orderings(c, m32, &sp, NULL, 1, 1, o->profile, 0, &kms))) break;
        for (int v = 0; v < PDFSK_TUNE_COUNT && !rc; v++) {
          cand_ms[v] = 1e30;
          if (!plan_fits(nf, dims, v)) {
            /* its ordering still runs, on the current tile */
            rc = run_orderings(c, m32, &sp, NULL, v + 2, v + 2, o->profile, 0, &kms);
            continue;
          }
          if (c->tile != v &&
              (rc = plan_build_ex(c, nf, (void *const *)fields, NULL, NULL, dims, v, 1, 1)))
            break;
          tk = 0.0;
          rc = run_orderings(c, m32, &sp, NULL, v + 2, v + 2, 1, 0, &tk);
          cand_ms[v] = tk;
          if (o->profile) kms += tk;
        }
        if (rc) break;
        int best = 0;
        for (int v = 1; v < PDFSK_TUNE_COUNT; v++)
          if (cand_ms[v] < cand_ms[best]) best = v;
        tuning = 0;
        tune_store(c, key, best);
        tile = best;
        if (c->tile != best &&
            (rc = plan_build_ex(c, nf, (void *const *)fields, NULL, NULL, dims, best, 1, 1)))
          break;
        if ((rc = run_orderings(c, m32, &sp, NULL, PDFSK_TUNE_COUNT + 2, 8, o->profile, 0, &kms)))
          break;
      } else if ((rc = run_orderings(c, m32, &sp, NULL, sn0, sn1, o->profile, o->graph, &kms))) {
        break;
      }
      HIPGO(hipEventRecord(c->ev1, c->st));
      unsigned err = 0;
      if ((rc = fixed ? read_changed_md3(c, delta, &err, fmd, NULL, 1)
                      : read_changed(c, delta, &err)))
        break;
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
        if (fixed && fmd[f] > pmq[f]) pmq[f] = fmd[f];
        last_total += pdel[f];
        if (field_passes) field_passes[f] = p + 1;
        /* plain stop: a pass that changed nothing, or in fixed point none of
         * its values by more than stop_quanta quanta */
        jdone[f] = last_seg && o->fixed_passes <= 0 &&
                   (pdel[f] == 0 || (fixed && pmq[f] <= (double)o->stop_quanta));
        {
          /* the device skipped the orderings of this segment that followed
           * the field's PDFSK_CONVERGED_RUN-th consecutive unchanged one */
          const unsigned run = c->h_streak[2 * f + 1];
          long long skipped =
              run > PDFSK_CONVERGED_RUN ? (long long)(run - PDFSK_CONVERGED_RUN) : 0;
          /* an overlapped launch skips the slots from two after its first
           * unchanged one */
          if (c->last_ov) skipped = run >= 2 ? (run < (unsigned)nord ? run : nord) - 2 : 0;
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
          js[nj] = speed ? slows[f] : NULL;
          jfb[nj] = fxb[f];
          jfd[nj] = fxd[f];
          jft[nj] = fxt[f];
          memcpy(jd + 3 * nj, dims + 3 * f, 3 * sizeof(int));
          nj++;
        }
      }
      /* all fields to certify in one judge launch */
      if (nj > 0) {
        if ((rc = judge_batch_ex(c, nj, jv, ji, speed ? js : NULL, jd, NULL, dx, dy, dz, jfl,
                                 fixed ? jfb : NULL, jfd, jft)))
          break;
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
  s.sweep_launches = (int)(c->launches - launches0);
  c->ov_use = 0;
  s.tile = c->tile;
  s.last_changed = last_total;
  s.ms_sweeps = ms - s.ms_judge;
  s.ms_sweep_kernels = kms;
  /* FP64 polish of the converged FP32 fields (they keep their FP32 tile
   * choice; the plan is rebuilt for the FP64 sweeps) */
  if ((rc == PDFS_OK || rc == PDFS_ENOTCONV) && o->polish && inits && o->fixed_passes <= 0 &&
      !fixed) {
    const int rc0 = rc;
    rc = pdfsi_polish(c, nf, fields, inits, slows, dims, dx, dy, dz, o, &s);
    if (rc == PDFS_OK) rc = rc0;
  }
  if (rc == PDFS_OK && o->judge && inits) {
    float tj_before = (float)s.ms_judge;
    int judged = 0;
    HIPGO(hipEventRecord(c->ev0, c->st));
    int nj = 0;
    for (int f = 0; f < nf; f++) {
      if (!inits[f] || certified[f]) continue;
      jidx[nj] = f;
      jv[nj] = fields[f];
      ji[nj] = inits[f];
      js[nj] = speed ? slows[f] : NULL;
      jfb[nj] = fxb[f];
      jfd[nj] = fxd[f];
      jft[nj] = fxt[f];
      memcpy(jd + 3 * nj, dims + 3 * f, 3 * sizeof(int));
      nj++;
    }
    if (nj > 0 && (rc = judge_batch_ex(c, nj, jv, ji, speed ? js : NULL, jd, NULL, dx, dy, dz,
                                       jfl, fixed ? jfb : NULL, jfd, jft)) == PDFS_OK) {
      for (int k = 0; k < nj; k++) {
        s.judge_flagged += jfl[k];
        if (field_flagged) field_flagged[jidx[k]] = jfl[k];
      }
      judged = nj;
    }
    HIPGO(hipEventRecord(c->ev1, c->st));
    HIPGO(hipEventSynchronize(c->ev1));
    float t = 0.f;
    hipEventElapsedTime(&t, c->ev0, c->ev1);
    s.ms_judge = tj_before + t;
    s.judge_runs += judged;
  }
  if (fixed) {
    /* back to FP32 (also after an error, so the caller's fields hold FP32) */
    int rd;
    if (hipEventRecord(c->kev[16], c->st) != hipSuccess || (rd = fx_decode(c, nf)) ||
        hipEventRecord(c->kev[17], c->st) != hipSuccess ||
        hipEventSynchronize(c->kev[17]) != hipSuccess) {
      if (rc == PDFS_OK || rc == PDFS_ENOTCONV)
        rc = fail(PDFS_EHIP, "fixed point: decode failed");
    } else {
      float tc = 0.f;
      hipEventElapsedTime(&tc, c->kev[16], c->kev[17]);
      s.ms_convert += tc;
    }
    decoded = 1;
  }
  s.ms_total = s.ms_sweeps + s.ms_judge + s.ms_polish + s.ms_convert;
  if (st) *st = s;
  if (field_flagged)
    for (int f = 0; f < nf; f++)
      if (active[f] && o->fixed_passes <= 0) field_flagged[f] = -1;
  if (rc == PDFS_OK && o->fixed_passes <= 0 && nactive > 0)
    rc = fail(PDFS_ENOTCONV, "%d field(s) not converged after %d passes", nactive, s.passes);
out:
  if (fixed && !decoded) fx_decode(c, nf);
  c->fxk = NULL;
  free(fxk);
  free(fxv);
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
  free(js);
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
  return batch_f32(c, 1, &field, &init, NULL, dims, dx, dy, dz, o, st, NULL, NULL, NULL);
}

int pdfs_solve_f32_dev(pdfs_ctx *c, float *field, const float *init, int x, int y, int z,
                       double dx, double dy, double dz, const pdfs_opts *o, pdfs_stats *st)
{
  if (!c || !field || x < 1 || y < 1 || z < 1 || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_solve_f32_dev: bad arguments");
  return single_f32(c, field, init, x, y, z, dx, dy, dz, o, st);
}

int pdfs_solve_speed_f32_dev(pdfs_ctx *c, float *field, const float *init,
                             const float *slowness, int x, int y, int z, double dx, double dy,
                             double dz, const pdfs_opts *o, pdfs_stats *st)
{
  if (!c || !field || !slowness || x < 1 || y < 1 || z < 1 || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_solve_speed_f32_dev: bad arguments");
  const int dims[3] = {x, y, z};
  return batch_f32(c, 1, &field, &init, &slowness, dims, dx, dy, dz, o, st, NULL, NULL, NULL);
}

int pdfs_judge_speed_f32_dev(pdfs_ctx *c, const float *field, const float *init,
                             const float *slowness, int x, int y, int z, double dx, double dy,
                             double dz, uint32_t *flags, long long *flagged)
{
  if (!c || !field || !init || !slowness || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_judge_speed_f32_dev: bad arguments");
  const int dims[3] = {x, y, z};
  long long fl = 0;
  int rc = pdfsi_judge_batch(c, 1, &field, &init, &slowness, dims, flags ? &flags : NULL, dx, dy,
                             dz, &fl);
  if (rc == PDFS_OK && flagged) *flagged = fl;
  return rc;
}

/* Whether a padded FP64 field runs on the column pipeline: its ghost layer
 * is all +inf (the reference's solver input, single_gpu_cuda/src/
 * vti_parser.c:137-180; the column kernel's ghosts are implicit +inf) and it
 * fits the kernel's buffer offsets. Other fields run the per-level kernel,
 * which reads the ghost cells. */
static int col64_ok(pdfs_ctx *c, const double *padded, int x, int y, int z, int tile, int *ok)
{
  *ok = 0;
  const int dims[3] = {x, y, z};
  if ((long long)x * y * z > PDFSK_MAX_POINTS64 || !plan_fits_w(1, dims, tile, 2)) return PDFS_OK;
  HIPCHK(hipMemsetAsync(c->d_u64 + 1, 0, sizeof(unsigned long long), c->st));
  KCHK(pdfsk_count_ghost_not_inf(padded, x, y, z, c->d_u64 + 1, c->st));
  HIPCHK(hipMemcpyAsync(c->h_pin64 + 1, c->d_u64 + 1, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  *ok = c->h_pin64[1] == 0;
  return PDFS_OK;
}

/* Reference-arithmetic FP64 solve of a padded device field, in place:
 * `passes` passes (> 0), or passes until the largest change of a pass is
 * <= tol, at most maxp (PDFS_ENOTCONV beyond). On the column pipeline
 * (bit-identical to the reference's per-level schedule) when col64_ok;
 * with a slowness field, or otherwise, on the per-level kernel. */
static int solve_ref64(pdfs_ctx *c, double *padded, const double *slow, int x, int y, int z,
                       double dx, double dy, double dz, int passes, double tol, int maxp,
                       int tile, int overlap, pdfs_stats *st)
{
  pdfs_stats s;
  memset(&s, 0, sizeof s);
  if (maxp < 1) maxp = 1000;
  long long npts = (long long)x * y * z;
  int ok = 0, rc;
  if (!slow && (rc = col64_ok(c, padded, x, y, z, tile, &ok))) return rc;
  if (ok) {
    size_t off = 0;
    if ((rc = pbuf_layout(c, 1, &npts, &off))) return rc;
    double *dn = c->pbuf;
    const int dims[3] = {x, y, z};
    pdfs_sp64 sp64;
    make_sp64(dx, dy, dz, &sp64);
    KCHK(pdfsk_pack64(padded, dn, x, y, z, c->st));
    rc = sweep64_batch(c, PDFSK_MODE_F64_REF, 1, &dn, NULL, NULL, dims, &sp64, tile,
                       passes > 0 ? passes : 0, tol, 0, maxp, 0, 0, overlap, &s, NULL, NULL);
    if (rc && rc != PDFS_ENOTCONV) return rc;
    KCHK(pdfsk_unpack64(dn, padded, x, y, z, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (st) *st = s;
    return rc;
  }
  int fixed = passes > 0;
  if (fixed) maxp = passes;
  float ms = 0.f;
  double md = 0.0;
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
    unsigned long long b = c->h_pin64[0];
    memcpy(&md, &b, sizeof md);
    s.passes++;
    s.orderings += 8;
    s.point_updates += 8 * npts;
    s.last_maxdiff = md;
    if (!fixed && md <= tol) break;
  }
  s.ms_sweeps = ms;
  s.ms_total = ms;
  if (st) *st = s;
  if (!fixed && !(md <= tol))
    return fail(PDFS_ENOTCONV, "FP64 solve not converged after %d passes (max change %g)", maxp,
                md);
  return PDFS_OK;
}

int pdfs_solve_ref64_dev(pdfs_ctx *c, double *padded, int x, int y, int z, double dx, double dy,
                         double dz, int passes, double tol, pdfs_stats *st)
{
  if (!c || !padded || x < 1 || y < 1 || z < 1 || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_solve_ref64_dev: bad arguments");
  return solve_ref64(c, padded, NULL, x, y, z, dx, dy, dz, passes, tol, 1000, tile64_of(NULL),
                     c->overlap64, st);
}

int pdfs_solve_speed64_dev(pdfs_ctx *c, double *padded, const double *slowness, int x, int y,
                           int z, double dx, double dy, double dz, int passes, double tol,
                           pdfs_stats *st)
{
  if (!c || !padded || !slowness || x < 1 || y < 1 || z < 1 || bad_h(dx, dy, dz))
    return fail(PDFS_EINVAL, "pdfs_solve_speed64_dev: bad arguments");
  return solve_ref64(c, padded, slowness, x, y, z, dx, dy, dz, passes, tol, 1000,
                     tile64_of(NULL), c->overlap64, st);
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
                     o->tol, o->max_passes, tile64_of(o), o->overlap, st);
    if (rc && rc != PDFS_ENOTCONV) goto out;
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
  o.max_passes = 1000;
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
    printf("\n%s in run_fsm: '%s'\n", rc == PDFS_ENOTCONV ? "No convergence" : "HIP failure",
           pdfs_last_error());
    exit(EXIT_FAILURE);
  }
  float ms = 0.f;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  printf("Parallel FSM time: %f s.\n", ms / 1000.0);
}

/* One reference ordering sn of a padded device field, in place: on the
 * column pipeline when col64_ok, else on the per-level kernel. */
static int ordering_ref64(pdfs_ctx *c, double *padded, int x, int y, int z, double dx, double dy,
                          double dz, int sn)
{
  int ok = 0, rc;
  if ((rc = col64_ok(c, padded, x, y, z, tile64_of(NULL), &ok))) return rc;
  if (!ok) {
    KCHK(pdfsk_sweep_ref64(padded, NULL, x, y, z, sn, dx, dy, dz, NULL, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return PDFS_OK;
  }
  long long npts = (long long)x * y * z;
  size_t off = 0;
  if ((rc = pbuf_layout(c, 1, &npts, &off))) return rc;
  double *dn = c->pbuf;
  const int dims[3] = {x, y, z};
  pdfs_sp64 sp64;
  make_sp64(dx, dy, dz, &sp64);
  KCHK(pdfsk_pack64(padded, dn, x, y, z, c->st));
  if ((rc = plan_build_ex(c, 1, (void *const *)&dn, NULL, NULL, dims, tile64_of(NULL), 0, 2)))
    return rc;
  if ((rc = run_orderings(c, PDFSK_MODE_F64_REF, NULL, &sp64, sn, sn, 0, 0, NULL))) return rc;
  long long d = 0;
  unsigned err = 0;
  double md = 0.0;
  if ((rc = read_changed_md(c, &d, &err, &md))) return rc;
  if (err) return fail(PDFS_ETIMEOUT, "column pipeline: edge wait timed out");
  KCHK(pdfsk_unpack64(dn, padded, x, y, z, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return PDFS_OK;
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
      ordering_ref64(c, d, pf->x, pf->y, pf->z, pf->dx, pf->dy, pf->dz, sn) != PDFS_OK ||
      hipMemcpy(pf->distance, d, np * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
    printf("\nHIP failure in runFSM: '%s'\n", pdfs_last_error());
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
    if ((rc = solve_ref64(c, d, NULL, x, y, z, pf->dx, pf->dy, pf->dz, passes, 1e-12, 1000,
                          tile64_of(NULL), c->overlap64, NULL)))
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
