/*
 * CPU oracle for the parallel Fast Sweeping Method of GEM3D/PDFS.
 *
 * TEST INFRASTRUCTURE ONLY (the checker). Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product. Plain C99, host only. Build: oracle/Makefile.
 *
 * Every function cites the reference file:line it restates. Function-level
 * pinning is listed in fsm_oracle.h.
 */
#include "fsm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MIN(a, b) ((a) < (b) ? (a) : (b))
#define ORC_MAX(a, b) ((a) > (b) ? (a) : (b))

static inline size_t idx3(int X, int Y, int i, int j, int k)
{
  return ((size_t)i * (size_t)(Y + 2) + (size_t)j) * (size_t)(X + 2) + (size_t)k;
}

/* ------------------------------------------------------------------------ */
/* Godunov local solve, reference arithmetic.                               */
/* Restates single_gpu_cuda/src/phi3D_fsm.cu:249-295 and                   */
/* mpi_openacc/src/phi3D_fsm.c:132-182: bubble sort of the (min, spacing)   */
/* pairs by min with strict '>' swaps, then the 1-D / 2-D / 3-D roots in    */
/* the expanded radicand form, evaluated in the same operation order, then  */
/* min(cur, new) with the reference's '<' macro. The expression order must  */
/* stay as written for bitwise parity with the reference.                   */
/* ------------------------------------------------------------------------ */
double orc_godunov_ref(double cur, double mx, double my, double mz,
                       double dx, double dy, double dz)
{
  double m[3] = {mx, my, mz};
  double d[3] = {dx, dy, dz};
  for (int i = 1; i < 3; i++) {
    for (int j = 0; j < 3 - i; j++) {
      if (m[j] > m[j + 1]) {
        double tm = m[j], td = d[j];
        m[j] = m[j + 1];
        d[j] = d[j + 1];
        m[j + 1] = tm;
        d[j + 1] = td;
      }
    }
  }
  double m_0 = m[0], m_1 = m[1], m_2 = m[2];
  double d_0 = d[0], d_1 = d[1], d_2 = d[2];
  double m2_0 = m_0 * m_0, m2_1 = m_1 * m_1, m2_2 = m_2 * m_2;
  double d2_0 = d_0 * d_0, d2_1 = d_1 * d_1, d2_2 = d_2 * d_2;

  double u = m_0 + d_0;
  if (u > m_1) {
    double s = sqrt(-m2_0 + 2 * m_0 * m_1 - m2_1 + d2_0 + d2_1);
    u = (m_1 * d2_0 + m_0 * d2_1 + d_0 * d_1 * s) / (d2_0 + d2_1);
    if (u > m_2) {
      double a = sqrt(-m2_0 * d2_1 - m2_0 * d2_2 + 2 * m_0 * m_1 * d2_2 -
                      m2_1 * d2_0 - m2_1 * d2_2 + 2 * m_0 * m_2 * d2_1 -
                      m2_2 * d2_0 - m2_2 * d2_1 + 2 * m_1 * m_2 * d2_0 +
                      d2_0 * d2_1 + d2_0 * d2_2 + d2_1 * d2_2);
      u = (m_2 * d2_0 * d2_1 + m_1 * d2_0 * d2_2 + m_0 * d2_1 * d2_2 +
           d_0 * d_1 * d_2 * a) /
          (d2_0 * d2_1 + d2_0 * d2_2 + d2_1 * d2_2);
    }
  }
  return ORC_MIN(cur, u);
}

/* ------------------------------------------------------------------------ */
/* Godunov local solve, shift-invariant difference form (new; no reference  */
/* counterpart). Same sort and branch structure as the reference, but the   */
/* roots are written as u = m0 + t(m1 - m0, m2 - m0, spacing), so FP32 keeps */
/* the O(h) radicand at any distance magnitude. Isotropic spacing uses      */
/* constant multiplies and one square root: with a = m1 - m0, b = m2 - m0,  */
/* the 2-D root m0 + (a + sqrt(2h^2 - a^2)) / 2 exceeds m2 exactly when     */
/* 2h^2 - a^2 > (2b - a)^2 (both sides of r > 2b - a are >= 0), so that    */
/* test picks the 2-D or 3-D radicand before the root is taken, and        */
/* u = m0 + (P + sqrt(R)) K with (P, K) = (a, 1/2) or (a + b, 1/3).        */
/* Anisotropic spacing uses the general weighted form.                      */
/* FP32: a 3-D radicand at or below 2^-96 is flushed to 0 (its root is below */
/* 2^-48 and cannot move any value >= 2^-24), so the GPU square root never  */
/* sees a denormal or tiny input. FP64 clamps at 0.                         */
/* The HIP kernels evaluate exactly these operations in the same order.     */
/* Slowness sl (the _speed variants; 1 otherwise, where every scaling below */
/* is exact): the spacings become d_i sl, the isotropic h sl, and the       */
/* anisotropic weights 1/d_i^2 and their sum are scaled by 1/sl^2.          */
/* ------------------------------------------------------------------------ */
#define DIFF_FORM_BODY(T, SQRT, FLUSH)                                          \
  T m[3] = {mx, my, mz};                                                        \
  T d[3] = {(T)dx, (T)dy, (T)dz};                                               \
  if (dx == dy && dy == dz) {                                                   \
    T h = (T)dx * sl;                                                           \
    T h2 = h * h;                                                               \
    T two_h2 = (T)2 * h2;                                                       \
    T three_h2 = (T)3 * h2;                                                     \
    T lo = ORC_MIN(m[0], m[1]);                                                 \
    T hi = ORC_MAX(m[0], m[1]);                                                 \
    T m0 = ORC_MIN(lo, m[2]);                                                   \
    T mid = ORC_MAX(lo, m[2]);                                                  \
    T m1 = ORC_MIN(hi, mid);                                                    \
    T m2 = ORC_MAX(hi, mid);                                                    \
    T u = m0 + h;                                                               \
    if (u > m1) {                                                               \
      T a = m1 - m0;                                                            \
      T b = m2 - m0;                                                            \
      T r2 = two_h2 - a * a;                                                    \
      T e = b + (b - a);                                                        \
      T R = r2, P = a, K = (T)0.5;                                              \
      if (r2 > e * e) {                                                         \
        T D = three_h2 - (T)2 * ((a * a + b * b) - a * b);                      \
        R = D > (T)FLUSH ? D : (T)0;                                            \
        P = a + b;                                                              \
        K = (T)1 / (T)3;                                                        \
      }                                                                         \
      u = m0 + (P + SQRT(R)) * K;                                               \
    }                                                                           \
    return cur < u ? cur : u;                                                   \
  }                                                                             \
  {                                                                             \
    T w[3] = {(T)1 / (d[0] * d[0]), (T)1 / (d[1] * d[1]), (T)1 / (d[2] * d[2])}; \
    T W = w[0] + w[1] + w[2];                                                   \
    const T r = (T)1 / (sl * sl);                                               \
    for (int i = 0; i < 3; i++) {                                               \
      d[i] = d[i] * sl;                                                         \
      w[i] = w[i] * r;                                                          \
    }                                                                           \
    W = W * r;                                                                  \
    for (int i = 1; i < 3; i++) {                                               \
      for (int j = 0; j < 3 - i; j++) {                                         \
        if (m[j] > m[j + 1]) {                                                  \
          T tm = m[j], td = d[j], tw = w[j];                                    \
          m[j] = m[j + 1];                                                      \
          d[j] = d[j + 1];                                                      \
          w[j] = w[j + 1];                                                      \
          m[j + 1] = tm;                                                        \
          d[j + 1] = td;                                                        \
          w[j + 1] = tw;                                                        \
        }                                                                       \
      }                                                                         \
    }                                                                           \
    T u = m[0] + d[0];                                                          \
    if (u > m[1]) {                                                             \
      T a = m[1] - m[0];                                                        \
      T d20 = d[0] * d[0], d21 = d[1] * d[1];                                   \
      T r = SQRT(d20 + d21 - a * a);                                            \
      u = m[0] + (a * d20 + d[0] * d[1] * r) / (d20 + d21);                     \
      if (u > m[2]) {                                                           \
        T b = m[2] - m[0];                                                      \
        T P = a * w[1] + b * w[2];                                              \
        T Q = a * a * w[1] + b * b * w[2] - (T)1;                               \
        T D = P * P - W * Q;                                                    \
        D = D > (T)FLUSH ? D : (T)0;                                            \
        u = m[0] + (P + SQRT(D)) / W;                                           \
      }                                                                         \
    }                                                                           \
    return cur < u ? cur : u;                                                   \
  }

float orc_godunov_f32_speed(float cur, float mx, float my, float mz, double dx, double dy,
                            double dz, float sl)
{
  DIFF_FORM_BODY(float, sqrtf, 0x1p-96f)
}

double orc_godunov_d64_speed(double cur, double mx, double my, double mz, double dx,
                             double dy, double dz, double sl)
{
  DIFF_FORM_BODY(double, sqrt, 0.0)
}

float orc_godunov_f32(float cur, float mx, float my, float mz, double dx,
                      double dy, double dz)
{
  return orc_godunov_f32_speed(cur, mx, my, mz, dx, dy, dz, 1.0f);
}

double orc_godunov_d64(double cur, double mx, double my, double mz, double dx,
                       double dy, double dz)
{
  return orc_godunov_d64_speed(cur, mx, my, mz, dx, dy, dz, 1.0);
}

/* ------------------------------------------------------------------------ */
/* Ordering tables. Restates mpi_openacc/src/phi3D_fsm.c:195-221 and        */
/* single_gpu_cuda/src/phi3D_fsm.cu:148-159: reflected axes                 */
/* x for sn in {4,8}, y for {2,6}, z for {3,7}; descending levels for       */
/* sn in {2,5,7,8}.                                                         */
/* ------------------------------------------------------------------------ */
typedef struct {
  int xo, yo, zo; /* reflection offsets (0 = none) */
  int incr;       /* 1 = ascending levels */
} orc_order;

static orc_order make_order(int X, int Y, int Z, int sn)
{
  orc_order o;
  o.xo = (sn == 4 || sn == 8) ? X + 1 : 0;
  o.yo = (sn == 2 || sn == 6) ? Y + 1 : 0;
  o.zo = (sn == 3 || sn == 7) ? Z + 1 : 0;
  o.incr = !(sn == 2 || sn == 5 || sn == 7 || sn == 8);
  return o;
}

/* Reference neighbour gathering: mpi_openacc/src/phi3D_fsm.c:136-140. */
static inline double ref_point(const double *d, size_t idx, size_t extX,
                               size_t extXY, double dx, double dy, double dz)
{
  double minX = ORC_MIN(d[idx - 1], d[idx + 1]);
  double minY = ORC_MIN(d[idx - extX], d[idx + extX]);
  double minZ = ORC_MIN(d[idx - extXY], d[idx + extXY]);
  return orc_godunov_ref(d[idx], minX, minY, minZ, dx, dy, dz);
}

/* Hyperplane sweep of one ordering: restates mpi_openacc/src/phi3D_fsm.c:
 * 83-119 (level loop, x/y bounds, z = level - x - y, reflection, update). */
void orc_sweep_ref64(double *d, int X, int Y, int Z, double dx, double dy,
                     double dz, int sn)
{
  orc_order o = make_order(X, Y, Z, sn);
  size_t extX = (size_t)X + 2, extXY = extX * (size_t)(Y + 2);
  int first = 3, last = X + Y + Z;
  int start = o.incr ? first : last;
  int end = o.incr ? last + 1 : first - 1;
  for (int level = start; level != end; level += o.incr ? 1 : -1) {
    int xs = ORC_MAX(1, level - (Y + Z));
    int ys = ORC_MAX(1, level - (X + Z));
    int xe = ORC_MIN(X, level - (first - 1));
    int ye = ORC_MIN(Y, level - (first - 1));
    for (int x = xs; x <= xe; x++) {
      int k = abs(x - o.xo);
      for (int y = ys; y <= ye; y++) {
        int z = level - (x + y);
        if (z > 0 && z <= Z) {
          int j = abs(y - o.yo);
          int i = abs(z - o.zo);
          size_t id = (size_t)i * extXY + (size_t)j * extX + (size_t)k;
          d[id] = ref_point(d, id, extX, extXY, dx, dy, dz);
        }
      }
    }
  }
}

/* Spatially varying speed (new; the reference declares Phi.F / SPEED,
 * phi3D.h:37, 54, and never uses them): |grad u| = s with slowness s = 1/F
 * per point. The point's update is the reference's local solve with its
 * spacings scaled by its slowness (dx s, dy s, dz s), since
 * sum ((u - m_i) / d_i)^2 = s^2 is sum ((u - m_i) / (d_i s))^2 = 1; the
 * hyperplane schedule is orc_sweep_ref64's. s is the padded array of the
 * field (only interior points are read). */
void orc_sweep_ref64_speed(double *d, const double *s, int X, int Y, int Z, double dx,
                           double dy, double dz, int sn)
{
  orc_order o = make_order(X, Y, Z, sn);
  size_t extX = (size_t)X + 2, extXY = extX * (size_t)(Y + 2);
  int first = 3, last = X + Y + Z;
  int start = o.incr ? first : last;
  int end = o.incr ? last + 1 : first - 1;
  for (int level = start; level != end; level += o.incr ? 1 : -1) {
    int xs = ORC_MAX(1, level - (Y + Z));
    int ys = ORC_MAX(1, level - (X + Z));
    int xe = ORC_MIN(X, level - (first - 1));
    int ye = ORC_MIN(Y, level - (first - 1));
    for (int x = xs; x <= xe; x++) {
      int k = abs(x - o.xo);
      for (int y = ys; y <= ye; y++) {
        int z = level - (x + y);
        if (z > 0 && z <= Z) {
          int j = abs(y - o.yo);
          int i = abs(z - o.zo);
          size_t id = (size_t)i * extXY + (size_t)j * extX + (size_t)k;
          const double sp = s[id];
          d[id] = ref_point(d, id, extX, extXY, dx * sp, dy * sp, dz * sp);
        }
      }
    }
  }
}

/* Physical traversal direction of ordering sn per axis: +1 ascending. */
static void order_dirs(int sn, int *sx, int *sy, int *sz)
{
  orc_order o = make_order(1, 1, 1, sn);
  int lv = o.incr ? 1 : -1;
  *sx = o.xo ? -lv : lv;
  *sy = o.yo ? -lv : lv;
  *sz = o.zo ? -lv : lv;
}

/* Lexicographic Gauss-Seidel in ordering sn's octant direction, reference
 * arithmetic. Equal to orc_sweep_ref64 bit for bit: on a level x+y+z=L the
 * 7-point stencil only reads levels L-1 (already new) and L+1 (still old),
 * in both schedules. */
void orc_sweep_lex64(double *d, int X, int Y, int Z, double dx, double dy,
                     double dz, int sn)
{
  int sx, sy, sz;
  order_dirs(sn, &sx, &sy, &sz);
  size_t extX = (size_t)X + 2, extXY = extX * (size_t)(Y + 2);
  for (int a = 1; a <= Z; a++) {
    int i = sz > 0 ? a : Z + 1 - a;
    for (int b = 1; b <= Y; b++) {
      int j = sy > 0 ? b : Y + 1 - b;
      for (int c = 1; c <= X; c++) {
        int k = sx > 0 ? c : X + 1 - c;
        size_t id = idx3(X, Y, i, j, k);
        d[id] = ref_point(d, id, extX, extXY, dx, dy, dz);
      }
    }
  }
}

#define LEX_SWEEP_BODY(T, SOLVE, SL)                                         \
  int sx, sy, sz;                                                            \
  order_dirs(sn, &sx, &sy, &sz);                                             \
  size_t extX = (size_t)X + 2, extXY = extX * (size_t)(Y + 2);               \
  for (int a = 1; a <= Z; a++) {                                             \
    int i = sz > 0 ? a : Z + 1 - a;                                          \
    for (int b = 1; b <= Y; b++) {                                           \
      int j = sy > 0 ? b : Y + 1 - b;                                        \
      for (int c = 1; c <= X; c++) {                                         \
        int k = sx > 0 ? c : X + 1 - c;                                      \
        size_t id = idx3(X, Y, i, j, k);                                     \
        T mx = ORC_MIN(d[id - 1], d[id + 1]);                                \
        T my = ORC_MIN(d[id - extX], d[id + extX]);                          \
        T mz = ORC_MIN(d[id - extXY], d[id + extXY]);                        \
        d[id] = SOLVE(d[id], mx, my, mz, dx, dy, dz, SL);                    \
      }                                                                      \
    }                                                                        \
  }

void orc_sweep_f32(float *d, int X, int Y, int Z, double dx, double dy,
                   double dz, int sn)
{
  LEX_SWEEP_BODY(float, orc_godunov_f32_speed, 1.0f)
}

/* FP32 sweep of |grad u| = s (new; the reference declares Phi.F / SPEED,  */
/* phi3D.h:37, 54, and never uses them): orc_sweep_f32 with each point's    */
/* spacings scaled by its slowness s (padded FP32, interior read).          */
void orc_sweep_f32_speed(float *d, const float *s, int X, int Y, int Z, double dx, double dy,
                         double dz, int sn)
{
  LEX_SWEEP_BODY(float, orc_godunov_f32_speed, s[id])
}

void orc_sweep_d64(double *d, int X, int Y, int Z, double dx, double dy,
                   double dz, int sn)
{
  LEX_SWEEP_BODY(double, orc_godunov_d64_speed, 1.0)
}

/* FP64 difference-form sweep of |grad u| = s (s padded FP64). */
void orc_sweep_d64_speed(double *d, const double *s, int X, int Y, int Z, double dx, double dy,
                         double dz, int sn)
{
  LEX_SWEEP_BODY(double, orc_godunov_d64_speed, s[id])
}

/* FP64 polish (new; no reference counterpart). A Gauss-Seidel step of the */
/* fixed-point map u = min(init, G64(u)) in ordering sn's octant order, the  */
/* local solve being the difference form (orc_godunov_d64) applied to the   */
/* point's init instead of its current value: unlike the reference's        */
/* min(u, G(u)) (mpi_openacc/src/phi3D_fsm.c:180) it can raise a value, so  */
/* it converges to the FP64 solution from an FP32 solution that lies below  */
/* it in places. Neighbour minima in the reference's order (lower index     */
/* first, mpi_openacc/src/phi3D_fsm.c:138-140). Returns the largest         */
/* |old - new| of any update (a NaN change counts as +inf).                 */
/* ------------------------------------------------------------------------ */
double orc_sweep_polish64(double *d, const double *init, int X, int Y, int Z, double dx,
                          double dy, double dz, int sn)
{
  return orc_sweep_polish64_speed(d, init, NULL, X, Y, Z, dx, dy, dz, sn);
}

/* orc_sweep_polish64 of |grad u| = s: s (padded FP32, or NULL for 1) scales
 * each point's spacings as in orc_sweep_f32_speed. */
double orc_sweep_polish64_speed(double *d, const double *init, const float *s, int X, int Y,
                                int Z, double dx, double dy, double dz, int sn)
{
  int sx, sy, sz;
  order_dirs(sn, &sx, &sy, &sz);
  size_t extX = (size_t)X + 2, extXY = extX * (size_t)(Y + 2);
  double md = 0.0;
  for (int a = 1; a <= Z; a++) {
    int i = sz > 0 ? a : Z + 1 - a;
    for (int b = 1; b <= Y; b++) {
      int j = sy > 0 ? b : Y + 1 - b;
      for (int c = 1; c <= X; c++) {
        int k = sx > 0 ? c : X + 1 - c;
        size_t id = idx3(X, Y, i, j, k);
        double mx = ORC_MIN(d[id - 1], d[id + 1]);
        double my = ORC_MIN(d[id - extX], d[id + extX]);
        double mz = ORC_MIN(d[id - extXY], d[id + extXY]);
        double u = orc_godunov_d64_speed(init[id], mx, my, mz, dx, dy, dz,
                                         s ? (double)s[id] : 1.0);
        if (u != d[id]) {
          double df = fabs(d[id] - u);
          if (df != df) df = INFINITY;
          if (df > md) md = df;
        }
        d[id] = u;
      }
    }
  }
  return md;
}

int orc_polish64(double *d, const double *init, int X, int Y, int Z, double dx, double dy,
                 double dz, double tol, int max_passes, double *last_maxdiff)
{
  return orc_polish64_speed(d, init, NULL, X, Y, Z, dx, dy, dz, tol, max_passes, last_maxdiff);
}

int orc_polish64_speed(double *d, const double *init, const float *s, int X, int Y, int Z,
                       double dx, double dy, double dz, double tol, int max_passes,
                       double *last_maxdiff)
{
  int passes = 0;
  double md = 0.0, m8 = 0.0;
  while (passes < max_passes) {
    md = 0.0;
    for (int sn = 1; sn <= 8; sn++) {
      double m = orc_sweep_polish64_speed(d, init, s, X, Y, Z, dx, dy, dz, sn);
      if (m > md) md = m;
      m8 = m;
    }
    passes++;
    if (m8 <= tol) break;
  }
  if (last_maxdiff) *last_maxdiff = md;
  return passes;
}

/* ------------------------------------------------------------------------ */
/* Convergence loops (new: the reference runs exactly one pass,             */
/* single_gpu_cuda/src/phi3D.c:118, mpi_openacc/src/block_fsm.c:296).       */
/* The pass body is the 1x1x1 reduction of block_fsm.c:296-298: runFSM for  */
/* sn = 1..8.                                                               */
/* ------------------------------------------------------------------------ */
static double max_abs_diff64(const double *a, const double *b, size_t n)
{
  double m = 0.0;
  for (size_t p = 0; p < n; p++) {
    if (a[p] == b[p]) continue;
    double df = fabs(a[p] - b[p]);
    if (!(df <= m)) m = df; /* NaN (inf - inf) counts as a change */
  }
  return m;
}

int orc_solve_ref64(double *d, int X, int Y, int Z, double dx, double dy,
                    double dz, double tol, int max_passes, double *last_maxdiff)
{
  size_t n = (size_t)(X + 2) * (Y + 2) * (Z + 2);
  double *prev = (double *)malloc(n * sizeof(double));
  int passes = 0;
  double md = 0.0;
  while (passes < max_passes) {
    memcpy(prev, d, n * sizeof(double));
    for (int sn = 1; sn <= 8; sn++) orc_sweep_ref64(d, X, Y, Z, dx, dy, dz, sn);
    passes++;
    md = max_abs_diff64(prev, d, n);
    if (md <= tol) break;
  }
  free(prev);
  if (last_maxdiff) *last_maxdiff = md;
  return passes;
}

/* Passes of orc_sweep_ref64_speed until max |delta| <= tol. */
int orc_solve_ref64_speed(double *d, const double *s, int X, int Y, int Z, double dx,
                          double dy, double dz, double tol, int max_passes, double *last_maxdiff)
{
  size_t n = (size_t)(X + 2) * (Y + 2) * (Z + 2);
  double *prev = (double *)malloc(n * sizeof(double));
  int passes = 0;
  double md = 0.0;
  while (passes < max_passes) {
    memcpy(prev, d, n * sizeof(double));
    for (int sn = 1; sn <= 8; sn++) orc_sweep_ref64_speed(d, s, X, Y, Z, dx, dy, dz, sn);
    passes++;
    md = max_abs_diff64(prev, d, n);
    if (md <= tol) break;
  }
  free(prev);
  if (last_maxdiff) *last_maxdiff = md;
  return passes;
}

int orc_solve_d64(double *d, int X, int Y, int Z, double dx, double dy,
                  double dz, double tol, int max_passes)
{
  size_t n = (size_t)(X + 2) * (Y + 2) * (Z + 2);
  double *prev = (double *)malloc(n * sizeof(double));
  int passes = 0;
  while (passes < max_passes) {
    memcpy(prev, d, n * sizeof(double));
    for (int sn = 1; sn <= 8; sn++) orc_sweep_d64(d, X, Y, Z, dx, dy, dz, sn);
    passes++;
    if (max_abs_diff64(prev, d, n) <= tol) break;
  }
  free(prev);
  return passes;
}

int orc_solve_f32(float *d, int X, int Y, int Z, double dx, double dy,
                  double dz, int max_passes, long *last_changed)
{
  return orc_solve_f32_speed(d, NULL, X, Y, Z, dx, dy, dz, max_passes, last_changed);
}

/* Passes of orc_sweep_f32_speed (s NULL: orc_sweep_f32) until a pass
 * changes no value. */
int orc_solve_f32_speed(float *d, const float *s, int X, int Y, int Z, double dx, double dy,
                        double dz, int max_passes, long *last_changed)
{
  size_t n = (size_t)(X + 2) * (Y + 2) * (Z + 2);
  float *prev = (float *)malloc(n * sizeof(float));
  int passes = 0;
  long changed = 0;
  while (passes < max_passes) {
    memcpy(prev, d, n * sizeof(float));
    for (int sn = 1; sn <= 8; sn++) {
      if (s)
        orc_sweep_f32_speed(d, s, X, Y, Z, dx, dy, dz, sn);
      else
        orc_sweep_f32(d, X, Y, Z, dx, dy, dz, sn);
    }
    passes++;
    changed = 0;
    for (size_t p = 0; p < n; p++) changed += memcmp(&prev[p], &d[p], sizeof(float)) != 0;
    if (changed == 0) break;
  }
  free(prev);
  if (last_changed) *last_changed = changed;
  return passes;
}

/* ------------------------------------------------------------------------ */
/* Pre/post processing of the single-GPU CLI path.                          */
/* ------------------------------------------------------------------------ */

/* single_gpu_cuda/src/phi3D.c:165-180: every non-ghost value that is not
 * +inf is kept if nonzero and replaced by 90000 if exactly zero or NaN. The
 * location==1 && d==+inf branch is unreachable under the outer test. */
void orc_update_distance(const int *loc, double *d, size_t n)
{
  for (size_t p = 0; p < n; p++) {
    if (loc[p] != -1 && d[p] != INFINITY) {
      d[p] = (loc[p] == 1 && d[p] == INFINITY) ? -1
             : (d[p] > 0.0 || d[p] < 0.0)     ? d[p]
                                              : 90000;
    }
  }
}

/* single_gpu_cuda/src/phi3D.c:190-205. */
void orc_set_negative_inside(const int *loc, double *d, size_t n)
{
  for (size_t p = 0; p < n; p++) {
    if (loc[p] != -1 && d[p] != INFINITY && loc[p] == 1) d[p] = -1;
  }
}

/* single_gpu_cuda/src/phi3D.c:221-242: every ghost cell takes the value of
 * the interior cell reached by clamping each index into [1, dim]. */
void orc_adjust_boundary(double *d, int X, int Y, int Z)
{
  int x = X + 2, y = Y + 2, z = Z + 2;
  for (int i = 0; i < z; i++)
    for (int j = 0; j < y; j++)
      for (int k = 0; k < x; k++) {
        int I = (i == z - 1) ? i - 1 : (i == 0) ? i + 1 : i;
        int J = (j == y - 1) ? j - 1 : (j == 0) ? j + 1 : j;
        int K = (k == x - 1) ? k - 1 : (k == 0) ? k + 1 : k;
        if (i != I || j != J || k != K)
          d[idx3(X, Y, i, j, k)] = d[idx3(X, Y, I, J, K)];
      }
}

/* single_gpu_cuda/src/vti_parser.c:137-180: the padded border holds +inf. */
void orc_fill_ghost(double *d, int X, int Y, int Z, double v)
{
  for (int i = 0; i < Z + 2; i++)
    for (int j = 0; j < Y + 2; j++)
      for (int k = 0; k < X + 2; k++)
        if (i == 0 || i == Z + 1 || j == 0 || j == Y + 1 || k == 0 || k == X + 1)
          d[idx3(X, Y, i, j, k)] = v;
}

/* ------------------------------------------------------------------------ */
/* Synthetic generators (new). splitmix64 (Steele, Lea & Flood 2014) drives */
/* all parameters; grid point (k, j, i) (1-based interior index) sits at    */
/* ((k-1)h, (j-1)h, (i-1)h). Unknown cells hold 1e6, sources 0.             */
/* ------------------------------------------------------------------------ */
uint64_t orc_splitmix64(uint64_t *s)
{
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double u01(uint64_t *s) { return (double)(orc_splitmix64(s) >> 11) * 0x1.0p-53; }

#define FAR_VALUE 1e6

void orc_gen_point(double *d, int X, int Y, int Z, uint64_t seed)
{
  uint64_t s = seed;
  int k = 1 + (int)(u01(&s) * X);
  int j = 1 + (int)(u01(&s) * Y);
  int i = 1 + (int)(u01(&s) * Z);
  size_t n = (size_t)(X + 2) * (Y + 2) * (Z + 2);
  for (size_t p = 0; p < n; p++) d[p] = FAR_VALUE;
  orc_fill_ghost(d, X, Y, Z, INFINITY);
  d[idx3(X, Y, i, j, k)] = 0.0;
}

/* Terrain: H(x,y) = sum_k A_k exp(-((x-cx_k)^2 + (y-cy_k)^2) / (2 s_k^2)).
 * Draw order per hill: A, sigma, cx, cy. Cells with z <= H are 0, cells with
 * 0 < z - H <= 1.5h hold z - H, all others 1e6. */
void orc_gen_terrain(double *d, int X, int Y, int Z, double h, uint64_t seed,
                     int nhills, double amp_lo, double amp_hi, double sig_lo,
                     double sig_hi)
{
  uint64_t s = seed;
  double A[64], S[64], CX[64], CY[64];
  if (nhills > 64) nhills = 64;
  for (int q = 0; q < nhills; q++) {
    A[q] = (amp_lo + (amp_hi - amp_lo) * u01(&s)) * h;
    S[q] = (sig_lo + (sig_hi - sig_lo) * u01(&s)) * h;
    CX[q] = u01(&s) * (double)(X - 1) * h;
    CY[q] = u01(&s) * (double)(Y - 1) * h;
  }
  for (int j = 1; j <= Y; j++)
    for (int k = 1; k <= X; k++) {
      double x = (double)(k - 1) * h, y = (double)(j - 1) * h;
      double H = 0.0;
      for (int q = 0; q < nhills; q++) {
        double ex = x - CX[q], ey = y - CY[q];
        H += A[q] * exp(-(ex * ex + ey * ey) / (2.0 * S[q] * S[q]));
      }
      for (int i = 1; i <= Z; i++) {
        double z = (double)(i - 1) * h;
        double off = z - H;
        d[idx3(X, Y, i, j, k)] = off <= 0.0 ? 0.0 : (off <= 1.5 * h ? off : FAR_VALUE);
      }
    }
  orc_fill_ghost(d, X, Y, Z, INFINITY);
}

/* Spheres: draw order per sphere: cx, cy, cz, r. s = min_k(|p - c_k| - r_k);
 * s <= 0 -> 0, s <= 1.5h -> s, else 1e6. */
void orc_gen_spheres(double *d, int X, int Y, int Z, double h, uint64_t seed,
                     int nspheres, double r_lo, double r_hi)
{
  uint64_t s = seed;
  double C[256][4];
  if (nspheres > 256) nspheres = 256;
  for (int q = 0; q < nspheres; q++) {
    C[q][0] = u01(&s) * (double)(X - 1) * h;
    C[q][1] = u01(&s) * (double)(Y - 1) * h;
    C[q][2] = u01(&s) * (double)(Z - 1) * h;
    C[q][3] = (r_lo + (r_hi - r_lo) * u01(&s)) * h;
  }
  for (int i = 1; i <= Z; i++)
    for (int j = 1; j <= Y; j++)
      for (int k = 1; k <= X; k++) {
        double x = (double)(k - 1) * h, y = (double)(j - 1) * h, z = (double)(i - 1) * h;
        double best = INFINITY;
        for (int q = 0; q < nspheres; q++) {
          double ex = x - C[q][0], ey = y - C[q][1], ez = z - C[q][2];
          double sd = sqrt(ex * ex + ey * ey + ez * ez) - C[q][3];
          if (sd < best) best = sd;
        }
        d[idx3(X, Y, i, j, k)] = best <= 0.0 ? 0.0 : (best <= 1.5 * h ? best : FAR_VALUE);
      }
  orc_fill_ghost(d, X, Y, Z, INFINITY);
}

/* ------------------------------------------------------------------------ */
/* FP64 residual judge (new). At every interior point the Godunov root of   */
/* the stored FP32 neighbours is re-solved in FP64 and the stored value must */
/* equal min(init, G) within tol = |e| 2^-22 + hmin 2^-21 (a few FP32 ulps   */
/* of the value plus a few ulps of the spacing).                            */
/* ------------------------------------------------------------------------ */
long orc_judge_f32(const float *v, const float *init, int X, int Y, int Z,
                   double dx, double dy, double dz, uint8_t *flags)
{
  return orc_judge_f32_speed(v, init, NULL, X, Y, Z, dx, dy, dz, flags);
}

/* orc_judge_f32 of |grad u| = s (s padded FP32, or NULL for 1): the FP64
 * solve scales the point's spacings by s, and the spacing part of the
 * tolerance becomes hmin s 2^-21. */
long orc_judge_f32_speed(const float *v, const float *init, const float *s, int X, int Y, int Z,
                         double dx, double dy, double dz, uint8_t *flags)
{
  size_t extX = (size_t)X + 2, extXY = extX * (size_t)(Y + 2);
  double hmin = ORC_MIN(dx, ORC_MIN(dy, dz));
  long bad = 0;
  for (int i = 1; i <= Z; i++)
    for (int j = 1; j <= Y; j++)
      for (int k = 1; k <= X; k++) {
        size_t id = idx3(X, Y, i, j, k);
        double mx = ORC_MIN((double)v[id - 1], (double)v[id + 1]);
        double my = ORC_MIN((double)v[id - extX], (double)v[id + extX]);
        double mz = ORC_MIN((double)v[id - extXY], (double)v[id + extXY]);
        const double sl = s ? (double)s[id] : 1.0;
        double e = orc_godunov_d64_speed((double)init[id], mx, my, mz, dx, dy, dz, sl);
        double val = (double)v[id];
        int ok;
        if (isinf(e) || isinf(val)) ok = (e == val);
        else ok = fabs(val - e) <= fabs(e) * 0x1.0p-22 + hmin * sl * 0x1.0p-21;
        if (flags) flags[id] = (uint8_t)!ok;
        bad += !ok;
      }
  return bad;
}

double orc_residual_f64(const double *v, const double *init, int X, int Y,
                        int Z, double dx, double dy, double dz)
{
  size_t extX = (size_t)X + 2, extXY = extX * (size_t)(Y + 2);
  double worst = 0.0;
  for (int i = 1; i <= Z; i++)
    for (int j = 1; j <= Y; j++)
      for (int k = 1; k <= X; k++) {
        size_t id = idx3(X, Y, i, j, k);
        double mx = ORC_MIN(v[id - 1], v[id + 1]);
        double my = ORC_MIN(v[id - extX], v[id + extX]);
        double mz = ORC_MIN(v[id - extXY], v[id + extXY]);
        double e = orc_godunov_d64(init[id], mx, my, mz, dx, dy, dz);
        if (e == v[id]) continue;
        double r = fabs(v[id] - e);
        if (!(r <= worst)) worst = r;
      }
  return worst;
}

/* ------------------------------------------------------------------------ */
/* Fixed-point storage (new; no reference counterpart). Values are unsigned */
/* multiples q of a quantum d = h 2^-k above a base (the start field's      */
/* smallest value), UINT32_MAX standing for +inf. The local solve is        */
/* orc_godunov_f32's isotropic difference form evaluated in FP32 on the     */
/* integer differences a = m1 - m0, b = m2 - m0 in units of the quantum     */
/* (Hq = 2^k), its root offset rounded to the nearest quantum; additions    */
/* saturate at +inf.                                                        */
/* ------------------------------------------------------------------------ */
#define ORC_FX_INF 0xFFFFFFFFu

int orc_fx_bits(int X, int Y, int Z)
{
  long long s = (long long)X + Y + Z;
  int c = 0;
  while ((1ll << c) < s) c++;
  return 32 - c;
}

float orc_fx_min(const float *v, size_t n)
{
  float m = INFINITY;
  for (size_t i = 0; i < n; i++)
    if (v[i] == v[i] && v[i] < m) m = v[i];
  return m;
}

void orc_fx_encode(const float *v, uint32_t *q, size_t n, double base, double d)
{
  for (size_t i = 0; i < n; i++) {
    double r = rint(((double)v[i] - base) / d);
    q[i] = r < 4294967295.0 ? (uint32_t)r : ORC_FX_INF;
  }
}

void orc_fx_decode(const uint32_t *q, float *v, size_t n, double base, double d)
{
  for (size_t i = 0; i < n; i++)
    v[i] = q[i] == ORC_FX_INF ? INFINITY : (float)(base + (double)q[i] * d);
}

static inline uint32_t fx_add_sat(uint32_t a, uint32_t b)
{
  uint64_t s = (uint64_t)a + b;
  return s > ORC_FX_INF ? ORC_FX_INF : (uint32_t)s;
}

uint32_t orc_godunov_fx(uint32_t cur, uint32_t mx, uint32_t my, uint32_t mz, int k)
{
  const uint32_t Hq = 1u << k;
  const float H = (float)Hq;
  const float two_h2 = 2.0f * (H * H), three_h2 = 3.0f * (H * H);
  uint32_t lo = ORC_MIN(mx, my), hi = ORC_MAX(mx, my);
  uint32_t m0 = ORC_MIN(lo, mz), mid = ORC_MAX(lo, mz);
  uint32_t m1 = ORC_MIN(hi, mid), m2 = ORC_MAX(hi, mid);
  uint32_t u = fx_add_sat(m0, Hq);
  if (u > m1) {
    float a = (float)(m1 - m0);
    float b = (float)(m2 - m0);
    float r2 = two_h2 - a * a;
    float e = b + (b - a);
    float R = r2, P = a, K = 0.5f;
    if (r2 > e * e) {
      float D = three_h2 - 2.0f * ((a * a + b * b) - a * b);
      R = D > 0x1p-96f ? D : 0.0f;
      P = a + b;
      K = 1.0f / 3.0f;
    }
    float t = (P + sqrtf(R)) * K;
    uint32_t tq = t >= 0.0f ? (uint32_t)rintf(t) : 0u;
    u = fx_add_sat(m0, tq);
  }
  return cur < u ? cur : u;
}

long orc_sweep_fx(uint32_t *q, int X, int Y, int Z, int k, int sn, uint32_t *maxdrop)
{
  uint32_t md = 0;
  int sx, sy, sz;
  order_dirs(sn, &sx, &sy, &sz);
  size_t extX = (size_t)X + 2, extXY = extX * (size_t)(Y + 2);
  long changed = 0;
  for (int a = 1; a <= Z; a++) {
    int i = sz > 0 ? a : Z + 1 - a;
    for (int b = 1; b <= Y; b++) {
      int j = sy > 0 ? b : Y + 1 - b;
      for (int c = 1; c <= X; c++) {
        int x = sx > 0 ? c : X + 1 - c;
        size_t id = idx3(X, Y, i, j, x);
        uint32_t mx = ORC_MIN(q[id - 1], q[id + 1]);
        uint32_t my = ORC_MIN(q[id - extX], q[id + extX]);
        uint32_t mz = ORC_MIN(q[id - extXY], q[id + extXY]);
        uint32_t u = orc_godunov_fx(q[id], mx, my, mz, k);
        changed += u != q[id];
        if (q[id] - u > md) md = q[id] - u;
        q[id] = u;
      }
    }
  }
  if (maxdrop) *maxdrop = md;
  return changed;
}

int orc_solve_fx(uint32_t *q, int X, int Y, int Z, int k, int stop_quanta, int max_passes,
                 long *last_changed)
{
  int passes = 0;
  long changed = 0;
  while (passes < max_passes) {
    changed = 0;
    uint32_t pmd = 0;
    for (int sn = 1; sn <= 8; sn++) {
      uint32_t md;
      changed += orc_sweep_fx(q, X, Y, Z, k, sn, &md);
      if (md > pmd) pmd = md;
    }
    passes++;
    if (changed == 0 || pmd <= (uint32_t)stop_quanta) break;
  }
  if (last_changed) *last_changed = changed;
  return passes;
}

long orc_judge_fx(const uint32_t *q, const float *init, int X, int Y, int Z, double h,
                  double base, double d, double tol, uint8_t *flags)
{
  size_t extX = (size_t)X + 2, extXY = extX * (size_t)(Y + 2);
  long bad = 0;
#define FXV(w) ((w) == ORC_FX_INF ? INFINITY : base + (double)(w) * d)
  for (int i = 1; i <= Z; i++)
    for (int j = 1; j <= Y; j++)
      for (int k = 1; k <= X; k++) {
        size_t id = idx3(X, Y, i, j, k);
        double mx = ORC_MIN(FXV(q[id - 1]), FXV(q[id + 1]));
        double my = ORC_MIN(FXV(q[id - extX]), FXV(q[id + extX]));
        double mz = ORC_MIN(FXV(q[id - extXY]), FXV(q[id + extXY]));
        double e = orc_godunov_d64((double)init[id], mx, my, mz, h, h, h);
        double val = FXV(q[id]);
        int ok = (val == e) || fabs(val - e) <= tol;
        if (flags) flags[id] = (uint8_t)!ok;
        bad += !ok;
      }
#undef FXV
  return bad;
}
