/*
 * pdfs.h — C-ABI of the MI355X-native parallel Fast Sweeping Method.
 *
 * Drop-in for the solver surface of GEM3D/PDFS:
 *   run_fsm  replaces single_gpu_cuda/src/phi3D_fsm.h:69-71 (definition
 *            phi3D_fsm.cu:64-106), called from single_gpu_cuda/src/phi3D.c:119.
 *   runFSM   replaces mpi_openacc/src/phi3D_fsm.h:73 (definition
 *            phi3D_fsm.c:67), called from mpi_openacc/src/block_fsm.c:298,332.
 * The pdfs_* entry points are new (status codes, options, statistics,
 * device-resident fields, the FP64 judge, synthetic fields).
 *
 * Plain C types and pointers only. Link with -lpdfs (pdfs_amd/lib).
 */
#ifndef PDFS_H
#define PDFS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field struct of the single-GPU build, same layout (x86-64 SysV: x,y,z at
 * 0,4,8; dx,dy,dz at 16,24,32; location 40; distance 48; F 56; size 64).
 * single_gpu_cuda/src/phi3D.h:45-55. `distance` is the padded
 * (x+2)(y+2)(z+2) FP64 array, x fastest, ghost layer +inf on entry. */
#ifndef PDFS_NO_PHI_TYPEDEF
typedef struct {
  int x, y, z;
  double dx, dy, dz;
  int *location;
  double *distance;
  double F;
} Phi;
#endif

/* Field struct of the MPI/OpenACC build (no `location`; size 56).
 * mpi_openacc/src/phi3D.h:45-54. */
typedef struct {
  int x, y, z;
  double dx, dy, dz;
  double *distance;
  double F;
} PhiACC;

/* Reference semantics: exactly `iterations` passes of orderings sn = 1..8 in
 * the reference's order, FP64, reference arithmetic, in place; ghost cells
 * never written. Extension (a departure from the reference, which runs zero
 * passes for iterations <= 0, phi3D_fsm.cu:143): iterations <= 0 runs passes
 * until no value changes by more than 1e-12, at most 1000. Errors, and no
 * convergence within 1000 passes, print a message and exit(EXIT_FAILURE) as
 * cudaCheckError does (single_gpu_cuda/src/phi3D_fsm.h:39-46). Prints
 * "Parallel FSM time: %f s." (phi3D_fsm.cu:104). The sweeps run on the FP64
 * column pipeline when the ghost layer is all +inf (the reference's input)
 * and the field has fewer than 2^28 points, else on one launch per
 * hyperplane; both are bit-identical to the reference. */
void run_fsm(Phi *pf, int iterations);
/* run_fsm, runFSM, pdfs_solve and pdfs_calc_dist_field share one
 * process-wide context (created on first use, kept until exit): like the
 * reference, they are not to be called from several threads at once. */

/* One ordering sn (1..8) of the reference, FP64, in place on the padded
 * array of an OpenACC-layout field. */
void runFSM(PhiACC *pf, int sn);

/* ---- status-returning API ---- */
enum { PDFS_OK = 0, PDFS_EINVAL = -1, PDFS_EHIP = -2, PDFS_ENOMEM = -3, PDFS_ETIMEOUT = -4,
       PDFS_ENOTCONV = -5 };
enum { PDFS_TILE_AUTO = -2 };
enum { PDFS_F64_REF = 0, /* FP64, reference arithmetic and schedule */
       PDFS_F32 = 1 };   /* FP32 difference form, column pipeline */
/* Storage of the FP32 path's fields while they are swept (opts.storage). */
enum { PDFS_STORE_F32 = 0,   /* FP32 values (then the FP64 polish, opts.polish) */
       PDFS_STORE_FIXED = 1 }; /* 32-bit fixed point (default; see pdfs_opts.storage) */
/* Overlapped passes (opts.overlap). */
enum { PDFS_OVERLAP_OFF = 0,    /* one launch per ordering */
       PDFS_OVERLAP_AUTO = 1,   /* where it pays: few fields, long critical path (default) */
       PDFS_OVERLAP_ALWAYS = 2, /* wherever the fields tile exactly */
       PDFS_OV_AUTO_FIELDS = 2, /* AUTO: at most this many active fields, */
       PDFS_OV_AUTO_DEPTH = 6 };/* and critical path >= this/10 of the steps per workgroup */

typedef struct {
  int precision;   /* PDFS_F64_REF or PDFS_F32 */
  int max_passes;  /* cap on 8-ordering passes (default 100) */
  int fixed_passes;/* > 0: run exactly this many passes, no convergence test */
  double tol;      /* FP64: stop when a pass changes nothing by more than tol */
  int judge;       /* FP32: run the FP64 residual judge after convergence */
  int tile;        /* column variant: 0 16x16, 1 32x8, 2 16x8, 3 32x16, 4 16x16 one
                      workgroup per CU, 5 16x8 two per CU; -1 default; PDFS_TILE_AUTO tuned
                      online on orderings of the first pass and cached per size class */
  int quiet;       /* no stdout */
  int profile;     /* time every sweep launch with events on the solver stream */
  int graph;       /* replay each pass (8 persistent launches) from a captured hipGraph */
  int certify;     /* FP32: also stop a field once the FP64 residual judge passes on it */
  int certify_from;/* first pass after which the judge is run (default 1); from then on it
                      also runs after every half pass (orderings 4 and 8) */
  int polish;      /* FP32 (converged solves of fields with an init; default 1): after the
                      FP32 stop, FP64 passes of u = min(init, G64(u)) from the FP32 field
                      until the last ordering of a pass changes no value by more than
                      polish_tol * hmin, then the result is rounded to FP32. FP32 storage
                      rounding accumulates along characteristics (at 512^3 up to ~3e-4 h);
                      one polish pass leaves ~1e-6 h (the last ordering's largest change
                      is ~10x the error it leaves). 0: the FP32 fixed point itself. */
  double polish_tol; /* default 2e-5 */
  int tile64;      /* FP64 column variant (16-wide tiles: 0 16x16, 2 16x8, 4, 5); -1 default */
  int storage;     /* FP32 path, converged solves (fixed_passes <= 0) with equal spacings and unit
                      speed: PDFS_STORE_FIXED (default) stores each field while it is swept as
                      32-bit fixed point, value = base + q h 2^-k with base the field's smallest
                      value and k = 32 - ceil(log2(x + y + z)) (q = 2^32 - 1 is +inf; values
                      beyond the range, which no converged value reaches, are +inf while they
                      last). The local solve is the FP32 difference form on the integer
                      differences, rounded to the nearest quantum, so storage adds no error
                      that grows with the distance: at 512^3 the result is within ~3e-5 h of
                      the FP64 solution with no polish (FP32 storage: ~3e-4 h before the
                      polish). Fields are converted in place (FP32 in and out); NaN in the
                      start field is treated as +inf. The polish does not run; the judge (and
                      certification) tests |value - min(init, G64(value))| <= judge_tol h +
                      4 quanta on the fixed-point values. Falls back to PDFS_STORE_F32 for
                      other solves and for a start field holding -inf. */
  double judge_tol;/* fixed-point storage: judge tolerance in units of h (default 2e-6; the
                      certified stop at 512^3 leaves about 10x this much error) */
  int stop_quanta; /* fixed-point storage, plain stop: a field ends on a pass that changes no
                      value by more than this many quanta (default 4; 0: a pass that changes
                      nothing, the FP32 storage rule) */
  int overlap;     /* PDFS_OVERLAP_* (FP32 path, fixed point, speed fields, and the FP64
                      reference arithmetic of pdfs_solve; not the FP64 polish): sweep a pass
                      (or half pass) in one persistent launch whose orderings overlap (a
                      column starts ordering k + 1 once it and its four neighbours have
                      finished ordering k). Possible when every
                      field's extents are multiples of the column tile and the edge regions
                      fit. AUTO (the default) takes it for at most PDFS_OV_AUTO_FIELDS
                      active fields whose longest critical path through an ordering is at
                      least PDFS_OV_AUTO_DEPTH / 10 of the ordering's column steps per
                      resident workgroup (one 512^3 field: 1.7); ALWAYS wherever possible;
                      else one launch per ordering. Results are the same bits. Values
                      other than OFF and ALWAYS act as AUTO. */
} pdfs_opts;

typedef struct {
  /* FP32 path: a field converges on a pass that changes nothing. Once one
   * ordering changes nothing the field is at the fixed point: the device
   * skips its remaining orderings in that pass, and the next pass, which
   * would change nothing, is counted but not launched; passes,
   * last_changed and the result are those of sweeping every ordering.
   * orderings counts the orderings swept (for a single field without the
   * skipped ones; for a batch, the launches); point_updates excludes skipped
   * orderings. */
  int passes, orderings;
  long long point_updates;   /* interior points x orderings executed */
  long long last_changed;    /* points changed in the last pass */
  double last_maxdiff;       /* FP64: max |delta| of the last pass */
  long long judge_flagged;   /* FP64 judge: points failing the residual */
  double ms_sweeps, ms_judge, ms_total;
  double ms_sweep_kernels;   /* profile: sum of per-launch sweep kernel times */
  int tile;
  int judge_runs;            /* judge launches (certification checks + final) */
  int fallbacks;             /* warm start: regions reset to the init and re-solved */
  double ms_warm;            /* warm start: network + upsample time */
  int polish_passes;         /* FP64 polish passes (counted as the FP32 passes are) */
  int polish_orderings;      /* FP64 polish orderings launched */
  double polish_maxdiff;     /* largest change of the last polish pass */
  double ms_polish;          /* polish time, conversions included */
  int polish_skipped;        /* fields left unpolished: 2^28 points or more (FP64 column
                                limit) */
  long long polish_point_updates; /* interior points x polish orderings executed */
  double ms_polish_kernels;  /* profile: sum of per-launch polish kernel times */
  int storage;               /* PDFS_STORE_FIXED if the fields were swept in fixed point */
  int sweep_launches;        /* sweep kernel launches of the solve (the polish's excluded) */
  double ms_convert;         /* fixed point: FP32 <-> fixed conversions (with the base readback) */
} pdfs_stats;

void pdfs_default_opts(pdfs_opts *o);

/* Solve on a host padded FP64 array (reference layout), in place. */
int pdfs_solve(double *distance, int x, int y, int z, double dx, double dy, double dz,
               const pdfs_opts *o, pdfs_stats *st);

/* ---- device-resident API (pointers are HIP device memory) ---- */
typedef struct pdfs_ctx pdfs_ctx;
pdfs_ctx *pdfs_ctx_create(void);
void pdfs_ctx_destroy(pdfs_ctx *c);
const char *pdfs_last_error(void);
/* The context's HIP stream (hipStream_t): every *_dev call runs on it. */
void *pdfs_ctx_stream(pdfs_ctx *c);

/* Dense FP32 field [z][y][x] (interior only, ghosts implicit +inf), in place. */
int pdfs_solve_f32_dev(pdfs_ctx *c, float *field, const float *init_or_null, int x, int y,
                       int z, double dx, double dy, double dz, const pdfs_opts *o,
                       pdfs_stats *st);
/* Batch of independent dense FP32 fields (uneven sizes), one persistent
 * launch per ordering covering every field still changing; each field stops
 * when a pass changes none of its values. dims: 3 ints per field; inits
 * (or NULL) for the judge; field_passes / field_ms (or NULL) receive each
 * field's pass count and completion time since the batch start. */
int pdfs_solve_batch_f32_dev(pdfs_ctx *c, int nfields, float *const *fields,
                             const float *const *inits, const int *dims, double dx, double dy,
                             double dz, const pdfs_opts *o, pdfs_stats *st, int *field_passes,
                             double *field_ms);
/* One ordering over a dense FP32 field; *changed += points changed. */
int pdfs_sweep_f32_dev(pdfs_ctx *c, float *field, int x, int y, int z, double dx, double dy,
                       double dz, int sn, int tile, long long *changed);
/* Padded FP64 field, reference arithmetic, `passes` passes (<= 0: until
 * max |delta| <= tol, at most 1000 passes, else PDFS_ENOTCONV), in place. */
int pdfs_solve_ref64_dev(pdfs_ctx *c, double *padded, int x, int y, int z, double dx,
                         double dy, double dz, int passes, double tol, pdfs_stats *st);
/* Spatially varying speed (no counterpart in the reference, which declares
 * Phi.F and SPEED, phi3D.h:37, 54, but never uses them): the Eikonal
 * |grad u| = 1/F with the padded slowness array s = 1/F (device, same
 * layout as padded; interior values finite and > 0). Each point's update is
 * the reference arithmetic's local solve with spacings dx s, dy s, dz s, on
 * the reference hyperplane schedule; passes as pdfs_solve_ref64_dev. */
int pdfs_solve_speed64_dev(pdfs_ctx *c, double *padded, const double *slowness, int x, int y,
                           int z, double dx, double dy, double dz, int passes, double tol,
                           pdfs_stats *st);
/* Spatially varying speed on the FP32 column pipeline (no counterpart in
 * the reference, which declares Phi.F and SPEED, phi3D.h:37, 54, but never
 * uses them): pdfs_solve_f32_dev for the Eikonal |grad u| = s with a dense
 * FP32 slowness s = 1/F per point (device, [z][y][x] like the field; values
 * finite and > 0). Each update is the FP32 difference-form solve with the
 * point's spacings d_i s; the FP64 polish (fields with an init) and the
 * residual judge use the same scaling, the judge's spacing tolerance being
 * hmin s 2^-21. Options and statistics as pdfs_solve_f32_dev, except that
 * the column variant 32x16 (tile 3) and PDFS_TILE_AUTO fall back to the
 * default variant (the slowness delay line does not fit 32x16's LDS). */
int pdfs_solve_speed_f32_dev(pdfs_ctx *c, float *field, const float *init_or_null,
                             const float *slowness, int x, int y, int z, double dx, double dy,
                             double dz, const pdfs_opts *o, pdfs_stats *st);
/* Batch form: slowness[f] for fields[f]; otherwise pdfs_solve_batch_f32_dev. */
int pdfs_solve_batch_speed_f32_dev(pdfs_ctx *c, int nfields, float *const *fields,
                                   const float *const *inits, const float *const *slowness,
                                   const int *dims, double dx, double dy, double dz,
                                   const pdfs_opts *o, pdfs_stats *st, int *field_passes,
                                   double *field_ms);
/* pdfs_judge_f32_dev of |grad u| = s (the judge pdfs_solve_speed_f32_dev runs). */
int pdfs_judge_speed_f32_dev(pdfs_ctx *c, const float *field, const float *init,
                             const float *slowness, int x, int y, int z, double dx, double dy,
                             double dz, uint32_t *flags, long long *flagged);
/* One ordering over a dense FP32 field, in place, with no device
 * allocation. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, allocation-free
 * serving loops). Same sweep and result as pdfs_sweep_f32_dev; its scratch
 * (launch state, tickets, counters, edge entries) is one 64 MiB device array
 * per GPU, held until the process ends. Fields whose edge entries exceed
 * 63 MiB (or more than 87381 columns) are rejected with PDFS_EINVAL (for
 * 16x16 tiles, about 384^3; tile as in pdfs_sweep_f32_dev). Launches on
 * `stream` (a hipStream_t, NULL: the null stream) without synchronising:
 * result (device, 2 words) gets result[0] += points changed and result[1] |=
 * 1 if an edge wait timed out. Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU race, so issue them on one stream. */
int pdfs_sweep_f32_noalloc(float *field, int x, int y, int z, double dx, double dy, double dz,
                           int sn, int tile, void *stream, unsigned long long *result);

/* FP64 residual judge: a point fails when |v - min(init, G64(v))| >
 * |e| 2^-22 + hmin 2^-21 (e = min(init, G64(v)); equal infinities pass, an
 * infinity against a finite value fails, and so does a point whose value or
 * a neighbour's is NaN). flags: bit p of
 * [ceil(x*y*z/32)] zeroed words is set for each failing point p, or NULL.
 * *flagged = number of failing points. Fields of 2^29 points or more are
 * rejected with PDFS_EINVAL (the FP32 field limit). */
int pdfs_judge_f32_dev(pdfs_ctx *c, const float *field, const float *init, int x, int y, int z,
                       double dx, double dy, double dz, uint32_t *flags,
                       long long *flagged);
/* The judge of a field held in fixed point (no counterpart in the reference;
 * the judge that fixed-point solves run, pdfs_opts.storage): q (device,
 * dense [z][y][x]) holds value = base + q d, q = 2^32 - 1 for +inf; init is
 * the dense FP32 init; spacing h. A point fails when
 * |value - min(init, G64(value))| > tol (equal infinities pass). */
int pdfs_judge_fixed_dev(pdfs_ctx *c, const uint32_t *q, const float *init, int x, int y, int z,
                         double h, double base, double d, double tol, uint32_t *flags,
                         long long *flagged);

/* Self-test (no counterpart in the reference): the FP32 sweep's square root
 * (hardware estimate plus a one-ulp correction) against the correctly
 * rounded root (FP64 root rounded to FP32) on every float in [2^-96, +inf),
 * the radicands the sweep can take; *mismatches = 0 when they agree
 * everywhere. */
int pdfs_selftest_sqrt_f32(pdfs_ctx *c, unsigned long long *mismatches);

/* Synthetic initial fields; dense FP32 and/or padded FP64 outputs (NULL to
 * skip). kind: 0 point source, 1 terrain, 2 spheres. */
enum { PDFS_GEN_POINT = 0, PDFS_GEN_TERRAIN = 1, PDFS_GEN_SPHERES = 2 };
int pdfs_generate_dev(pdfs_ctx *c, int kind, uint64_t seed, int x, int y, int z, double h,
                      float *dense, double *padded);

/* Synthetic field with explicit object parameters: terrain (nobj hills,
 * amplitude U(p_lo,p_hi) h, width U(q_lo,q_hi) h), spheres (nobj, radius
 * U(p_lo,p_hi) h), point source (parameters unused). */
int pdfs_generate_ex_dev(pdfs_ctx *c, int kind, uint64_t seed, int x, int y, int z, double h,
                         int nobj, double p_lo, double p_hi, double q_lo, double q_hi,
                         float *dense, double *padded);

/* Layout conversion between the padded FP64 and dense FP32 device arrays. */
int pdfs_pack_dev(pdfs_ctx *c, const double *padded, float *dense, int x, int y, int z);
int pdfs_unpack_dev(pdfs_ctx *c, const float *dense, double *padded, int x, int y, int z);

/* CLI pipeline of single_gpu_cuda/src/phi3D.c:107-126 (calc_dist_field) on
 * the device: update_distance (0/NaN -> 90000), `passes` sweep passes
 * (reference: 1; <= 0: until converged), location 1 -> -1, ghosts <- nearest
 * interior. precision: PDFS_F64_REF (reference arithmetic) or PDFS_F32. */
int pdfs_calc_dist_field(Phi *pf, int passes, int precision);
/* One pre/post step on host padded arrays: op 0 update_distance
 * (phi3D.c:165-180), 1 set_distance_negative_inside (phi3D.c:190-205),
 * 2 adjust_boundary (phi3D.c:221-242). */
int pdfs_prepost(int op, const int *location, double *distance, int x, int y, int z);

/* ---- learned warm start (no counterpart in the reference) ----
 * A 3-D conv encoder-decoder (4 levels, 16/32/64/64 channels, 3x3x3, BF16
 * weights and activations, FP32 accumulation on the matrix cores) maps the
 * source mask at 1/4 resolution to the distance at the coarse points
 * (4z, 4y, 4x). Coarse tensors are channels-last [nb][Dz][Dy][Dx][16], each
 * coarse side ceil(side/4) rounded up to a multiple of 8 (pdfs_net_grid).
 * A network belongs to the context it was created with and runs on its
 * stream. */
typedef struct pdfs_net pdfs_net;
pdfs_net *pdfs_net_create(pdfs_ctx *c, unsigned long long seed); /* He-initialised */
/* May be called before or after pdfs_ctx_destroy of its context; every other
 * pdfs_net_* call needs the context alive. */
void pdfs_net_destroy(pdfs_net *n);
int pdfs_net_load(pdfs_net *n, const char *path);
int pdfs_net_save(pdfs_net *n, const char *path);
long long pdfs_net_num_params(const pdfs_net *n);
/* which: 0 FP32 master parameters, 1 gradients of the last train step */
int pdfs_net_get(pdfs_net *n, int which, float *host);
int pdfs_net_set(pdfs_net *n, const float *host);
/* target scale (prediction = scale * distance in cells) and the upper-bound
 * inflation of the warm start w = min(init, a * d + b * hmax) */
void pdfs_net_get_calibration(const pdfs_net *n, float *scale, float *a, float *b);
void pdfs_net_set_calibration(pdfs_net *n, float scale, float a, float b);
/* cutoff (in h): where the raw prediction d reaches cut * hmax, beyond the
 * distances the network resolves, the warm start keeps the init (0: none) */
float pdfs_net_get_cut(const pdfs_net *n);
void pdfs_net_set_cut(pdfs_net *n, float cut);
void pdfs_net_grid(int x, int y, int z, int *Dz, int *Dy, int *Dx);
/* mask (BF16 [Dz][Dy][Dx][16]) of a dense init; target (FP32 [Dz][Dy][Dx]) of
 * a converged field */
int pdfs_net_features_dev(pdfs_net *n, const float *init, int x, int y, int z, uint16_t *mask);
int pdfs_net_target_dev(pdfs_net *n, const float *u, int x, int y, int z, float *target);
/* pred (optional): FP32 [nb][Dz][Dy][Dx][16], channel 0 the prediction */
int pdfs_net_forward_dev(pdfs_net *n, const uint16_t *mask, int nb, int Dz, int Dy, int Dx,
                         float *pred);
/* forward, MSE loss, backward, Adam (lr > 0; lr = 0 leaves the gradients) */
int pdfs_net_train_step_dev(pdfs_net *n, const uint16_t *mask, const float *target, int nb,
                            int Dz, int Dy, int Dx, float lr, float *loss);
/* w = min(init, a * d + b * hmax), d the upsampled prediction */
int pdfs_warm_start_f32_dev(pdfs_net *n, const float *init, float *w, int x, int y, int z,
                            double hmax);
/* Warm starts of nf uneven fields (dims[3f..3f+2] = x, y, z) in one launch
 * per network layer; per field bit-identical to pdfs_warm_start_f32_dev.
 * No counterpart in the reference (the learned warm start is new). */
int pdfs_warm_start_batch_f32_dev(pdfs_net *n, int nf, const float *const *inits,
                                  float *const *ws, const int *dims, double hmax);
/* warm start + certified FP32 solve; on a judge failure the field is reset to
 * its init and solved plainly (stats.fallbacks = 1) */
int pdfs_solve_warm_f32_dev(pdfs_ctx *c, pdfs_net *n, float *field, const float *init, int x,
                            int y, int z, double dx, double dy, double dz, const pdfs_opts *o,
                            pdfs_stats *st);
/* batch form: warm starts, one certified batch solve, per-field judge, and a
 * plain batch re-solve of flagged fields (stats.fallbacks = their count) */
int pdfs_solve_batch_warm_f32_dev(pdfs_ctx *c, pdfs_net *n, int nf, float *const *fields,
                                  const float *const *inits, const int *dims, double dx,
                                  double dy, double dz, const pdfs_opts *o, pdfs_stats *st,
                                  int *field_passes, int *field_fallback);

#ifdef __cplusplus
}
#endif
#endif
