/*
 * Thin C-ABI layer over the gfx950 HIP kernels (pdfs_kernels.hip).
 * Called from the host C solver (pdfs_solver.c). Device pointers only.
 *
 * Internal FP32 layout ("dense"): interior only, [Z][Y][X], x fastest, the
 * ghost layer implicit +inf. The FP64 reference-schedule path keeps the
 * reference's padded (X+2)(Y+2)(Z+2) layout with explicit ghosts.
 */
#ifndef PDFS_KERNELS_H
#define PDFS_KERNELS_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FP32 spacing constants, computed on the host in IEEE single precision in
 * the same operation order as the oracle's difference-form solve. */
typedef struct {
  float h, two_h2, three_h2, third; /* isotropic */
  float d[3], w[3], W;               /* anisotropic: spacing, 1/d^2, sum */
  int iso;
} pdfs_sp32;

/* FP64 spacing constants of the difference form, computed on the host in
 * the oracle's operation order (h2 = h h, 2 h2, 3 h2, 1/3; w = 1/d^2,
 * W = w0 + w1 + w2). */
typedef struct {
  double d[3], w[3], W;
  double h, two_h2, three_h2, third;
  int iso;
} pdfs_sp64;

/* Sweep modes of the column pipeline. */
enum {
  PDFSK_MODE_F32_ISO = 0,      /* FP32 difference form, u = min(u, G(u)) */
  PDFSK_MODE_F32_ANISO = 1,
  PDFSK_MODE_F64_REF = 2,      /* FP64 reference arithmetic, u = min(u, G(u)) */
  PDFSK_MODE_F64_POL_ISO = 3,  /* FP64 difference form, u = min(init, G(u)) */
  PDFSK_MODE_F64_POL_ANISO = 4,
  PDFSK_MODE_F32_SPEED = 5,    /* FP32 difference form, spacings scaled by the point's slowness */
  PDFSK_MODE_F64_POL_SPEED = 6,/* FP64 polish, spacings scaled by the point's slowness */
  PDFSK_MODE_FX_ISO = 7,       /* 32-bit fixed-point values (quantum h 2^-fxk, UINT32_MAX = +inf),
                                  FP32 difference form on their differences, u = min(u, G(u)) */
  PDFSK_MODE_COUNT = 8
};
/* modes whose values are FP64 */
#define PDFSK_MODE_IS_F64(m) \
  ((m) == PDFSK_MODE_F64_REF || (m) == PDFSK_MODE_F64_POL_ISO || (m) == PDFSK_MODE_F64_POL_ANISO || \
   (m) == PDFSK_MODE_F64_POL_SPEED)

/* Entries of padding before and after each edge array: prefetch indices of
 * the column kernel stay inside the allocation. */
#define PDFSK_EDGE_PAD 4096

/* Largest dense FP32 field of the column kernel: byte offsets are 32-bit
 * and out-of-range planes must map past the buffer end (x*y*z*4 < 2^31). */
#define PDFSK_MAX_POINTS ((1ll << 31) / 4 - 1)

/* Field table entry of a batched sweep (one per field; a single solve is a
 * batch of one). Dense FP32 field [Z][Y][X]; its east/north edge blocks
 * start at entry ex_off / ey_off of the edge allocations. */
typedef struct {
  void *f;            /* float (FP32 modes), double (FP64 modes) or uint32 (fixed point) */
  const float *init;  /* polish modes: dense FP32 init */
  const float *slow;  /* speed modes: dense FP32 slowness (1 / speed) */
  int X, Y, Z;
  int nTx, nTy;
  uint32_t ex_off, ey_off;
  int fxk;            /* fixed-point mode: the field's quantum is h 2^-fxk (0 otherwise) */
} pdfsk_field;

/* Fixed-point storage of a field (PDFSK_MODE_FX_ISO): value = base + q d with
 * d = h 2^-k, q = UINT32_MAX for +inf. k = 32 - ceil(log2(X + Y + Z)): every
 * converged value lies in [base, base + (X + Y + Z - 3) h] when base is the
 * smallest value of the start field (an axis-parallel path of 1-D updates
 * bounds it), so it is below UINT32_MAX d. */
int pdfsk_fx_bits(int X, int Y, int Z);
/* Per-field work of the fixed-point conversions, one launch each over all
 * fields: f the field (FP32 in, uint32 while solving, FP32 out), n points,
 * base and d as above; key: per-field order key of the smallest non-NaN
 * value (pdfsk_fx_min: UINT32_MAX-initialised by the caller). */
typedef struct {
  float *f;
  long long n;
  double base, d;
  uint32_t *key;
} pdfsk_fx;
/* key = min over the field of an order-preserving key of its non-NaN FP32
 * values; pdfsk_fx_key_value inverts it on the host (UINT32_MAX: no value). */
int pdfsk_fx_min_batch(const pdfsk_fx *tab, int nf, long long nmax, hipStream_t st);
float pdfsk_fx_key_value(uint32_t key);
/* In place: q = rint((v - base) / d) when below UINT32_MAX, else (and for
 * NaN) UINT32_MAX; back: v = (float)(base + q d), UINT32_MAX -> +inf. */
int pdfsk_fx_encode_batch(const pdfsk_fx *tab, int nf, long long nmax, hipStream_t st);
int pdfsk_fx_decode_batch(const pdfsk_fx *tab, int nf, long long nmax, hipStream_t st);

/* Largest dense FP64 field of the column kernel (x*y*z*8 < 2^31). */
#define PDFSK_MAX_POINTS64 ((1ll << 31) / 8 - 1)

/* Device state of the batched column sweep. state: [0] ticket counter,
 * [1] epoch, [2] ticket base, [3] finished workgroups, [4] column tickets;
 * advanced on the device by the last workgroup of each launch, so launches
 * (and graph replays) take no per-launch host arguments. */
typedef struct {
  pdfsk_field *fields;   /* [nfields] */
  int *cols;             /* [ncols][3] ticket -> (field, a, b), upwind-first */
  unsigned *state;       /* [8] */
  unsigned *changed;     /* [nfields] changed-point counters (accumulate) */
  unsigned *streak;      /* [nfields][2]: counter value after the last ordering, and the
                            number of consecutive orderings that changed nothing */
  int nfields;
  unsigned long long *maxdiff; /* [nfields] FP64 modes: max |old - new| bits; fixed point: the
                                  largest drop in quanta, as an integer (accumulate) */
  unsigned long long *maxdiff_last; /* [nfields] or NULL: ordering 8 accumulates here instead */
  unsigned *err;         /* bounded-spin timeout flag */
  uint64_t *edge_x, *edge_y;        /* face parts (index -PDFSK_EDGE_PAD) of one allocation
                                       that starts at edge_x; edge_y follows the x part */
  size_t edge_x_bytes, edge_y_bytes;  /* whole allocation from edge_x; y part alone */
} pdfsk_batch;

/* A field whose last ordering changed nothing is at the fixed point: every
 * point met u <= G(neighbours) when it was visited, and as nothing changed
 * those neighbours are still the current values, so no ordering can change
 * it again. Later orderings skip its columns (streak[2f+1] >= this run). */
#define PDFSK_CONVERGED_RUN 1u

/* Variants of the column pipeline (autotuner chooses one): tile shape and
 * workgroups per CU (max = as many as fit). */
enum {
  PDFSK_TILE_16x16 = 0,     /* 256 threads, max per CU */
  PDFSK_TILE_32x8 = 1,      /* 256 threads, max per CU */
  PDFSK_TILE_16x8 = 2,      /* 128 threads, max per CU */
  PDFSK_TILE_32x16 = 3,     /* 512 threads, 1 per CU (LDS) */
  PDFSK_TILE_16x16_1CU = 4, /* 256 threads, 1 per CU */
  PDFSK_TILE_16x8_2CU = 5,  /* 128 threads, 2 per CU */
  PDFSK_TILE_8x8 = 6,       /* 64 threads, max per CU */
  PDFSK_TILE_16x4 = 7,      /* 64 threads, max per CU */
  PDFSK_TILE_COUNT = 8
};
/* variants the online FP32 tuner tries (one per ordering 2..7 of a pass) */
enum {
  PDFSK_TUNE_COUNT = 6
};
void pdfsk_tile_dims(int variant, int *tx, int *ty);

/* Whether a tile variant runs a sweep mode (FP64 modes: 16-wide tiles). */
int pdfsk_tile_supports(int variant, int mode);
/* Resident workgroups of the column kernel for a tile variant and mode (0 if
 * unsupported or on error). */
int pdfsk_col_max_resident(int variant, int mode);

/* One ordering sn (1..8) over every field of the batch, in place; grid =
 * persistent workgroups (at most the resident count). sp: FP32 constants
 * (FP32 modes), sp64: FP64 constants (FP64 modes). */
int pdfsk_sweep_batch(const pdfsk_batch *b, int sn, int mode, const pdfs_sp32 *sp,
                      const pdfs_sp64 *sp64, int variant, int grid, hipStream_t st);

/* Orderings sn0..sn1 of a pass in one persistent launch (any sweep mode;
 * every field's extents multiples of the tile, so each ordering tiles the
 * same physical columns). FP64 modes: maxdiff takes every slot's largest
 * change; a launch that ends at ordering 8 also puts its last slot's into
 * maxdiff_last. otab: nt
 * tickets of 5 ints (ordering slot sn - sn0, field, virtual tile a, b,
 * physical column index in [0, ncolp)) in an order where every ticket's
 * dependencies come first: its upwind columns in the slot, itself and its 4
 * physical neighbours in the slot before, and all columns of the slot two
 * before. done: [8][ncolp] epochs (zeroed once, never reset); cnt: [8]
 * counters, zero between launches (the launch leaves them zero); chg:
 * [8][nfields] per-slot change counters, likewise zero between launches,
 * from which the launch leaves each field's streak as one launch per
 * ordering would. The edge allocation holds two regions, the second rx_bytes
 * (x part) / ry_bytes (y part) after the first; slot o uses region o % 2.
 * Epochs advance by the number of slots. */
int pdfsk_sweep_batch_ov(const pdfsk_batch *b, int sn0, int sn1, const int *otab, int nt,
                         int ncolp, unsigned *done, unsigned *cnt, unsigned *chg, size_t rx_bytes,
                         size_t ry_bytes, int mode, const pdfs_sp32 *sp, const pdfs_sp64 *sp64,
                         int variant, int grid, hipStream_t st);

/* One ordering sn of the reference hyperplane schedule, FP64, on the padded
 * array, one launch per level (reference arithmetic, bitwise parity).
 * If maxdiff != NULL, atomically maxes |old - new| (as uint64 bits).
 * slow (padded slowness 1/F, or NULL for unit speed) scales each point's
 * spacings. */
int pdfsk_sweep_ref64(double *d, const double *slow, int X, int Y, int Z, int sn, double dx,
                      double dy, double dz, unsigned long long *maxdiff, hipStream_t st);

/* FP64 residual judge of dense FP32 fields against their inits, all fields
 * of a batch in one launch. Field f owns judge tiles [tile0, next tile0);
 * entry nf is a sentinel whose tile0 is the total; flags (one bit per point,
 * [ceil(N/32)] words zeroed by the caller) optional per field. */
typedef struct {
  const float *v, *init;
  const float *slow;  /* dense FP32 slowness, or NULL for unit speed (same for every field) */
  uint32_t *flags;
  int X, Y, Z, tile0;
  /* fixed-point fields (pdfsk_judge_batch_f32 with fx): v holds q of
   * value = fx_base + q fx_d, and a point passes when |value - e| <= fx_tol */
  double fx_base, fx_d, fx_tol;
} pdfsk_judge_field;
/* judge tiles of one field (the tile0 stride) */
unsigned pdfsk_judge_tiles(int X, int Y, int Z);
/* counts[f] += points of field f failing the FP64 residual (counts zeroed by
 * the caller); tab on the device, ntiles = tab[nf].tile0 */
int pdfsk_judge_batch_f32(const pdfsk_judge_field *tab, int nf, long ntiles, double dx, double dy,
                          double dz, int speed, unsigned long long *counts, hipStream_t st);
/* the same for fixed-point fields (isotropic unit speed) */
int pdfsk_judge_batch_fx(const pdfsk_judge_field *tab, int nf, long ntiles, double h,
                         unsigned long long *counts, hipStream_t st);

/* Padded FP64 <-> dense FP64 (interior; unpack leaves the ghosts alone). */
int pdfsk_pack64(const double *padded, double *dense, int X, int Y, int Z, hipStream_t st);
int pdfsk_unpack64(const double *dense, double *padded, int X, int Y, int Z, hipStream_t st);
/* *count += ghost cells of a padded field whose bits are not +inf (count
 * zeroed by the caller). */
int pdfsk_count_ghost_not_inf(const double *padded, int X, int Y, int Z,
                              unsigned long long *count, hipStream_t st);
/* Element-wise conversions over the fields of a table, one launch:
 * dir 0: dst64[i] = (double)src32[i]; dir 1: dst32[i] = (float)src64[i]
 * (round to nearest even). */
typedef struct {
  float *f32;
  double *f64;
  long long n;
} pdfsk_conv;
int pdfsk_convert_batch(const pdfsk_conv *tab, int nf, long long nmax, int dir, hipStream_t st);

/* After a sweep segment: changed[nf], err[0], streak[2 nf] and (md != NULL)
 * md[f], md[cap_md + f] copied to pinned host memory (visible once the
 * stream is synchronized), then md[f], md[cap_md + f] zeroed. One launch. */
int pdfsk_read_counters(const unsigned *changed, const unsigned *err, const unsigned *streak,
                        unsigned long long *md, int nf, int cap_md, unsigned *h_changed,
                        unsigned *h_err, unsigned *h_streak, unsigned long long *h_md,
                        hipStream_t st);

/* Layout conversions: padded FP64 <-> dense FP32 and dense copies. */
int pdfsk_pack_f64(const double *padded, float *dense, int X, int Y, int Z,
                   hipStream_t st);
int pdfsk_unpack_f64(const float *dense, double *padded, int X, int Y, int Z,
                     hipStream_t st);
int pdfsk_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st);
/* count += floats with bit patterns in [lo, hi) where the sweep's FP32
 * square root differs from the correctly rounded one (the FP64 root rounded
 * to FP32); count zeroed by the caller */
int pdfsk_check_sqrt_cr(uint32_t lo, uint32_t hi, unsigned long long *count, hipStream_t st);

/* Synthetic initial fields, dense FP32 and/or padded FP64 (either may be
 * NULL). params: see pdfs_gen.c for the packing. */
int pdfsk_gen_point(float *dense, double *padded, int X, int Y, int Z, int k,
                    int j, int i, hipStream_t st);
int pdfsk_gen_terrain(float *dense, double *padded, int X, int Y, int Z,
                      double h, const double *dev_params, int nhills,
                      hipStream_t st);
int pdfsk_gen_spheres(float *dense, double *padded, int X, int Y, int Z,
                      double h, const double *dev_params, int nspheres,
                      hipStream_t st);

/* CLI pre/post processing on device arrays (padded FP64 + int location). */
int pdfsk_update_distance(const int *loc, double *d, long n, hipStream_t st);
int pdfsk_set_negative_inside(const int *loc, double *d, long n, hipStream_t st);
int pdfsk_adjust_boundary(double *d, int X, int Y, int Z, hipStream_t st);

#ifdef __cplusplus
}
#endif
#endif
