/*
 * CPU oracle for the parallel Fast Sweeping Method (FSM) of GEM3D/PDFS.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (pdfs_amd/) never links, loads or calls it.
 *
 * Layout everywhere: padded (X+2)(Y+2)(Z+2) arrays, x fastest, then y, then z,
 * ghost layer = +inf during the solve (single_gpu_cuda/src/vti_parser.c:137-180,
 * single_gpu_cuda/src/phi3D_fsm.cu:209-225, mpi_openacc/src/phi3D_fsm.c:112).
 *
 * Pinning of each function (what checks it besides its own restatement):
 *   orc_godunov_ref      golden vectors from the compiled reference runFSM
 *                        (tests/golden), closed forms, invariants.
 *   orc_sweep_ref64      compiled reference runFSM (oracle/_ref), bitwise;
 *                        golden vectors; lexicographic Gauss-Seidel (K4).
 *   orc_sweep_lex        bitwise equal to orc_sweep_ref64 (schedule
 *                        independence), brute force on tiny grids.
 *   orc_godunov_f32/_d64 closed forms (1-D m+h, 2-D m+h/sqrt2, 3-D m+h/sqrt3),
 *                        general and 2-D/3-D-boundary roots against an
 *                        independent root finder, shift invariance,
 *                        monotonicity.
 *   orc_sweep_ref64_speed / orc_solve_ref64_speed
 *                        slowness 1 = orc_sweep_ref64 and slowness 2 = the
 *                        doubled-spacing sweep, bitwise; a layered medium's
 *                        exact travel times; monotonicity in the slowness.
 *   orc_solve_*          point-source axis exactness (n*h), fixed-point
 *                        residual 0, agreement with the compiled reference.
 *   orc_update_distance, orc_set_negative_inside, orc_adjust_boundary
 *                        known-answer tests written from the reference's
 *                        branch structure (0->90000, NaN->90000, ghosts).
 *   generators           closed-form checks (sphere surface distances,
 *                        terrain column heights) and determinism.
 *   orc_judge_f32        planted defects (values below truth) are flagged;
 *                        a converged field passes.
 *   orc_sweep_polish64 / orc_polish64
 *                        its limit equals the monotone FP64 solve's
 *                        (orc_solve_d64) to 1e-12; the exact fixed point is
 *                        left unchanged; a region planted 0.5 h below the
 *                        solution is raised back (a planted min(u, G)
 *                        update fails this); point-source axis values n h.
 *   orc_godunov_fx / orc_sweep_fx / orc_solve_fx (fixed-point storage)
 *                        1-D updates exact (axis values n 2^k quanta),
 *                        2-D / 3-D equal-minima roots h/sqrt2, h/sqrt3 to
 *                        half a quantum plus the FP32 rounding of the root
 *                        offset, general roots within 1.5 quanta (plus that
 *                        rounding) of an independent root finder, +inf
 *                        saturation; shift of the start field;
 *                        converged fields within 5e-5 h of the FP64 solve
 *                        (orc_solve_d64); encode / decode against numpy.
 *   orc_judge_fx         a converged fixed-point field passes, planted
 *                        defects of a few quanta beyond the tolerance fail.
 */
#ifndef PDFS_FSM_ORACLE_H
#define PDFS_FSM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Godunov local solves. */
double orc_godunov_ref(double cur, double mx, double my, double mz,
                       double dx, double dy, double dz);
double orc_godunov_d64(double cur, double mx, double my, double mz,
                       double dx, double dy, double dz);
float orc_godunov_f32(float cur, float mx, float my, float mz,
                      double dx, double dy, double dz);

/* One ordering sn (1..8) over the padded field, in place. */
void orc_sweep_ref64(double *d, int X, int Y, int Z, double dx, double dy,
                     double dz, int sn);
void orc_sweep_lex64(double *d, int X, int Y, int Z, double dx, double dy,
                     double dz, int sn);
void orc_sweep_f32(float *d, int X, int Y, int Z, double dx, double dy,
                   double dz, int sn);
void orc_sweep_d64(double *d, int X, int Y, int Z, double dx, double dy,
                   double dz, int sn);
/* |grad u| = s, s = slowness (1 / speed) per point: the FP32 / FP64
 * difference-form solves with the point's spacings d_i s (anisotropic
 * weights 1/d_i^2 and their sum scaled by 1/s^2). s padded, interior read. */
float orc_godunov_f32_speed(float cur, float mx, float my, float mz, double dx, double dy,
                            double dz, float sl);
double orc_godunov_d64_speed(double cur, double mx, double my, double mz, double dx,
                             double dy, double dz, double sl);
void orc_sweep_f32_speed(float *d, const float *s, int X, int Y, int Z, double dx, double dy,
                         double dz, int sn);
void orc_sweep_d64_speed(double *d, const double *s, int X, int Y, int Z, double dx, double dy,
                         double dz, int sn);
int orc_solve_f32_speed(float *d, const float *s, int X, int Y, int Z, double dx, double dy,
                        double dz, int max_passes, long *last_changed);

/* FP64 polish: one ordering of u = min(init, G64(u)) (difference form; may
 * raise values), returning the largest change of any update; and passes of
 * it until the largest change of a pass's last ordering is <= tol (at most
 * max_passes), returning the passes run and (last_maxdiff) the last pass's
 * largest change. init: padded FP64 (ghosts unused). */
double orc_sweep_polish64(double *d, const double *init, int X, int Y, int Z, double dx,
                          double dy, double dz, int sn);
int orc_polish64(double *d, const double *init, int X, int Y, int Z, double dx, double dy,
                 double dz, double tol, int max_passes, double *last_maxdiff);
/* the polish of |grad u| = s (s padded FP32, or NULL for 1) */
double orc_sweep_polish64_speed(double *d, const double *init, const float *s, int X, int Y,
                                int Z, double dx, double dy, double dz, int sn);
int orc_polish64_speed(double *d, const double *init, const float *s, int X, int Y, int Z,
                       double dx, double dy, double dz, double tol, int max_passes,
                       double *last_maxdiff);

/* Passes of sn = 1..8 until converged (or max_passes). Return passes run. */
int orc_solve_ref64(double *d, int X, int Y, int Z, double dx, double dy,
                    double dz, double tol, int max_passes, double *last_maxdiff);
/* Spatially varying speed: slowness s (padded, = 1/F) scales the point's
 * spacings; the same schedule and local solve as orc_sweep_ref64. */
void orc_sweep_ref64_speed(double *d, const double *s, int X, int Y, int Z, double dx,
                           double dy, double dz, int sn);
int orc_solve_ref64_speed(double *d, const double *s, int X, int Y, int Z, double dx,
                          double dy, double dz, double tol, int max_passes, double *last_maxdiff);
int orc_solve_f32(float *d, int X, int Y, int Z, double dx, double dy,
                  double dz, int max_passes, long *last_changed);
int orc_solve_d64(double *d, int X, int Y, int Z, double dx, double dy,
                  double dz, double tol, int max_passes);

/* Pre/post processing of the CLI path. */
void orc_update_distance(const int *loc, double *d, size_t n);
void orc_set_negative_inside(const int *loc, double *d, size_t n);
void orc_adjust_boundary(double *d, int X, int Y, int Z);
void orc_fill_ghost(double *d, int X, int Y, int Z, double v);

/* Synthetic initial fields (padded, ghost = +inf). */
void orc_gen_point(double *d, int X, int Y, int Z, uint64_t seed);
void orc_gen_terrain(double *d, int X, int Y, int Z, double h, uint64_t seed,
                     int nhills, double amp_lo, double amp_hi, double sig_lo,
                     double sig_hi);
void orc_gen_spheres(double *d, int X, int Y, int Z, double h, uint64_t seed,
                     int nspheres, double r_lo, double r_hi);

/* FP64 residual judge of an FP32 field: flags[p] = 1 where
 * |v - min(init, G64(v))| > tol. Returns the number of flagged points. */
long orc_judge_f32(const float *v, const float *init, int X, int Y, int Z,
                   double dx, double dy, double dz, uint8_t *flags);
/* with slowness s (padded FP32, or NULL): spacing tolerance hmin s 2^-21 */
long orc_judge_f32_speed(const float *v, const float *init, const float *s, int X, int Y, int Z,
                         double dx, double dy, double dz, uint8_t *flags);
/* max |v - min(init, G(v))| over the interior, FP64 difference form. */
double orc_residual_f64(const double *v, const double *init, int X, int Y,
                        int Z, double dx, double dy, double dz);

uint64_t orc_splitmix64(uint64_t *state);

/* Fixed-point storage (new; no reference counterpart): value = base + q d,
 * d = h 2^-k, q = UINT32_MAX for +inf; k = 32 - ceil(log2(X + Y + Z)).
 * Arrays of q are padded like the FP arrays, ghosts UINT32_MAX. */
int orc_fx_bits(int X, int Y, int Z);
/* smallest non-NaN value of n FP32 values (+inf if none) */
float orc_fx_min(const float *v, size_t n);
void orc_fx_encode(const float *v, uint32_t *q, size_t n, double base, double d);
void orc_fx_decode(const uint32_t *q, float *v, size_t n, double base, double d);
/* local solve on quanta: Hq = 2^k */
uint32_t orc_godunov_fx(uint32_t cur, uint32_t mx, uint32_t my, uint32_t mz, int k);
/* one ordering in place; returns the points changed and (maxdrop, or NULL)
 * the largest drop old - new of an update, in quanta */
long orc_sweep_fx(uint32_t *q, int X, int Y, int Z, int k, int sn, uint32_t *maxdrop);
/* passes until a pass changes nothing or none of its drops exceeds
 * stop_quanta (or max_passes); returns passes run */
int orc_solve_fx(uint32_t *q, int X, int Y, int Z, int k, int stop_quanta, int max_passes,
                 long *last_changed);
/* FP64 residual judge of a fixed-point field (spacing h): a point fails when
 * |value - min(init, G64(value))| > tol; returns the failing points. */
long orc_judge_fx(const uint32_t *q, const float *init, int X, int Y, int Z, double h,
                  double base, double d, double tol, uint8_t *flags);

#ifdef __cplusplus
}
#endif
#endif
