// gfx950 kernels of the PDFS parallel Fast Sweeping Method.
//
// Hot path: the column-pipeline sweep (k_sweep_col). The (x, y) plane of a
// field is cut into TX x TY column tiles; one workgroup owns one column and
// streams it through z, sweeping the i+j+k = const hyperplanes of the column
// one per step out of an LDS ring of TX+TY+2 z-planes (coalesced plane loads
// and stores, one Godunov update per thread per step). Columns exchange
// their downstream edge values through epoch-tagged 64-bit entries written
// and read with relaxed agent-scope atomics (single-copy atomic, no cache
// maintenance), so a column starts a level as soon as its upwind neighbours
// have published it. Columns are handed out by a ticket counter in
// upwind-first order, which keeps the persistent grid deadlock free.
//
// Every schedule that respects the sweep's dependency order gives the same
// result as the reference's per-level kernel
// (single_gpu_cuda/src/phi3D_fsm.cu:199-233): on hyperplane L the 7-point
// stencil reads L-1 (new) and L+1 (old) only.
//
// Compiled with -ffp-contract=off semantics via pragmas on every arithmetic
// routine, so the FP32 and FP64 results match the CPU oracle bit for bit.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>
#include <string.h>

#include "pdfs_kernels.h"

#define PDFS_INF __builtin_inff()

// ---------------------------------------------------------------------------
// Godunov local solves
// ---------------------------------------------------------------------------

// Reference arithmetic (single_gpu_cuda/src/phi3D_fsm.cu:249-295): bubble
// sort of (min, spacing) pairs, expanded-radicand 1-D/2-D/3-D roots in the
// reference's operation order, min(cur, new) with '<'.
__device__ __forceinline__ double god_ref64(double cur, double mx, double my, double mz,
                                            double dx, double dy, double dz) {
#pragma clang fp contract(off)
  double m[3] = {mx, my, mz};
  double d[3] = {dx, dy, dz};
#pragma unroll
  for (int i = 1; i < 3; i++) {
#pragma unroll
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
    double s = __builtin_sqrt(-m2_0 + 2 * m_0 * m_1 - m2_1 + d2_0 + d2_1);
    u = (m_1 * d2_0 + m_0 * d2_1 + d_0 * d_1 * s) / (d2_0 + d2_1);
    if (u > m_2) {
      double a = __builtin_sqrt(-m2_0 * d2_1 - m2_0 * d2_2 + 2 * m_0 * m_1 * d2_2 -
                                m2_1 * d2_0 - m2_1 * d2_2 + 2 * m_0 * m_2 * d2_1 -
                                m2_2 * d2_0 - m2_2 * d2_1 + 2 * m_1 * m_2 * d2_0 +
                                d2_0 * d2_1 + d2_0 * d2_2 + d2_1 * d2_2);
      u = (m_2 * d2_0 * d2_1 + m_1 * d2_0 * d2_2 + m_0 * d2_1 * d2_2 + d_0 * d_1 * d_2 * a) /
          (d2_0 * d2_1 + d2_0 * d2_2 + d2_1 * d2_2);
    }
  }
  return cur < u ? cur : u;
}

// Square root in the operand's own precision (correctly rounded).
__device__ __forceinline__ float sqrt_t(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double sqrt_t(double x) { return __builtin_sqrt(x); }
// Lower limit of a kept 3-D radicand: FP32 flushes radicands <= 2^-96 to 0
// (their root cannot move a value >= 2^-24), FP64 clamps at 0.
template <typename T> __device__ __forceinline__ T radicand_floor();
template <> __device__ __forceinline__ float radicand_floor<float>() { return 0x1p-96f; }
template <> __device__ __forceinline__ double radicand_floor<double>() { return 0.0; }

// Shift-invariant difference form u = m0 + t(m1 - m0, m2 - m0, spacing).
// Isotropic spacing: sorting network + constant multiplies, no division, one
// square root. With a = m1 - m0 and b = m2 - m0 (b >= a >= 0), the 2-D root
// m0 + (a + sqrt(2h^2 - a^2)) / 2 exceeds m2 exactly when
// 2h^2 - a^2 > (2b - a)^2, so the 2-D / 3-D choice is made on the radicands
// and u = m0 + (P + sqrt(R)) K with (R, P, K) = (2h^2 - a^2, a, 1/2) or
// (3-D radicand, a + b, 1/3).
template <typename T>
__device__ __forceinline__ T god_diff_iso(T cur, T mx, T my, T mz, T h, T two_h2, T three_h2,
                                          T third) {
#pragma clang fp contract(off)
  T lo = mx < my ? mx : my;
  T hi = mx > my ? mx : my;
  T m0 = lo < mz ? lo : mz;
  T mid = lo > mz ? lo : mz;
  T m1 = hi < mid ? hi : mid;
  T m2 = hi > mid ? hi : mid;
  T u = m0 + h;
  if (u > m1) {
    T a = m1 - m0;
    T b = m2 - m0;
    T r2 = two_h2 - a * a;
    T e = b + (b - a);
    T R = r2, P = a, K = (T)0.5;
    if (r2 > e * e) {  // the 2-D root would exceed m2: 3-D
      T D = three_h2 - (T)2 * ((a * a + b * b) - a * b);
      R = D > radicand_floor<T>() ? D : (T)0;
      P = a + b;
      K = third;
    }
    u = m0 + (P + sqrt_t(R)) * K;
  }
  return cur < u ? cur : u;
}

// Correctly rounded FP32 square root for x == 0 or x >= 2^-96: v_sqrt_f32
// (within one ulp) followed by a one-ulp correction from the two FMA
// residuals x - (s -/+ ulp) s. Inputs here never reach the denormal range
// (see radicand_floor), so no rescaling is needed. x = 0 gives 0 (the s - 1
// candidate is a NaN pattern whose residual test fails, the s + 1 residual is
// 0); x < 0 gives NaN, which the callers discard.
__device__ __forceinline__ float sqrt_cr_nonneg(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u);
  const float sp = __uint_as_float(__float_as_uint(s) + 1u);
  const float rm = __builtin_fmaf(-sm, s, x);
  const float rp = __builtin_fmaf(-sp, s, x);
  float r = rm <= 0.0f ? sm : s;
  return rp > 0.0f ? sp : r;
}

// Branch-free isotropic FP32 solve: the same operations as the branching
// difference form, with the 1-D / 2-D / 3-D choice made by selects.
// Inputs are non-negative or +inf (no NaN, no -0): the per-axis minima and
// the 3-value sort (min3 / med3 / max3) equal the '<' / '>' selects of the
// oracle. A negative 2-D radicand (a^2 > 2h^2) implies a > h, so u1 <= m1
// and the root candidate (NaN then) is not selected; with m2 = +inf the 3-D
// test is false and its (NaN) radicand is not selected either.
__device__ __forceinline__ float god_iso32(float c, float xm, float xp, float ym, float yp,
                                           float zm, float zp, float h, float two_h2,
                                           float three_h2, float third) {
#pragma clang fp contract(off)
  const float mx = __builtin_elementwise_minimum(xm, xp);
  const float my = __builtin_elementwise_minimum(ym, yp);
  const float mz = __builtin_elementwise_minimum(zm, zp);
  const float m0 = __builtin_elementwise_minimum(__builtin_elementwise_minimum(mx, my), mz);
  const float m2 = __builtin_elementwise_maximum(__builtin_elementwise_maximum(mx, my), mz);
  const float m1 = __builtin_amdgcn_fmed3f(mx, my, mz);
  const float u1 = m0 + h;
  const float a = m1 - m0;
  const float b = m2 - m0;
  const float aa = a * a;
  const float r2 = two_h2 - aa;
  const float e = b + (b - a);
  float D = three_h2 - 2.0f * ((aa + b * b) - a * b);
  D = D > 0x1p-96f ? D : 0.0f;
  const bool three = r2 > e * e;
  const float R = three ? D : r2;
  const float P = three ? a + b : a;
  const float K = three ? third : 0.5f;
  const float t = m0 + (P + sqrt_cr_nonneg(R)) * K;
  const float u = u1 > m1 ? t : u1;
  return c < u ? c : u;
}

// Fixed-point isotropic solve. Values are unsigned multiples of a quantum
// (the field's h 2^-k); UINT32_MAX is +inf. The per-axis minima and their
// sort are exact integer min / max (min3 / med3 / max3); the differences
// a = m1 - m0 and b = m2 - m0 go to FP32 (exact below 2^24 quanta), and the
// root offset t of god_iso32's difference form is evaluated in FP32 in units
// of the quantum (Hq = 2^k, 2 Hq^2 and 3 Hq^2 exact) and rounded to the
// nearest quantum; u = m0 + t saturates at +inf. The 1-D candidate m0 + Hq
// is exact, so the stored values carry no rounding from their magnitude:
// the quantum bounds what one update adds, wherever the value lies.
// The radicand is clamped at 0: the 3-D one like god_iso32's (a nonzero
// radicand of integers in quanta is at least 1, so god_iso32's flush below
// 2^-96 changes nothing here), the 2-D one only where it is negative, which
// means a > h, so u1 <= m1 and the root is not selected. t is then finite,
// non-negative and below 2^32 (P < 2^33, K <= 1/2), and converts exactly.
__device__ __forceinline__ uint32_t god_fx_iso(uint32_t c, uint32_t xm, uint32_t xp,
                                               uint32_t ym, uint32_t yp, uint32_t zm,
                                               uint32_t zp, uint32_t Hq, float two_h2,
                                               float three_h2, float third) {
#pragma clang fp contract(off)
  const uint32_t mx = __builtin_elementwise_min(xm, xp);
  const uint32_t my = __builtin_elementwise_min(ym, yp);
  const uint32_t mz = __builtin_elementwise_min(zm, zp);
  const uint32_t lo = __builtin_elementwise_min(mx, my), hi = __builtin_elementwise_max(mx, my);
  const uint32_t m0 = __builtin_elementwise_min(lo, mz);
  const uint32_t m2 = __builtin_elementwise_max(hi, mz);
  const uint32_t m1 = __builtin_elementwise_max(lo, __builtin_elementwise_min(hi, mz));
  const uint32_t u1 = __builtin_elementwise_add_sat(m0, Hq);
  const float a = (float)(m1 - m0);
  const float b = (float)(m2 - m0);
  const float aa = a * a;
  const float r2 = two_h2 - aa;
  const float e = b + (b - a);
  const float D = three_h2 - 2.0f * ((aa + b * b) - a * b);
  const bool three = r2 > e * e;
  const float R = __builtin_elementwise_maximum(three ? D : r2, 0.0f);
  const float P = three ? a + b : a;
  const float K = three ? third : 0.5f;
  const float t = (P + sqrt_cr_nonneg(R)) * K;
  const uint32_t ut = __builtin_elementwise_add_sat(m0, (uint32_t)__builtin_rintf(t));
  const uint32_t u = u1 > m1 ? ut : u1;
  return c < u ? c : u;
}

// Anisotropic difference form: bubble sort of (m, d, 1/d^2) triples, 2-D
// root t = (a d0^2 + d0 d1 sqrt(d0^2 + d1^2 - a^2)) / (d0^2 + d1^2), 3-D root
// t = (P + sqrt(P^2 - W Q)) / W with P = a w1 + b w2, Q = a^2 w1 + b^2 w2 - 1.
template <typename T>
__device__ __forceinline__ T god_diff_aniso(T cur, T mx, T my, T mz, const T dd[3],
                                            const T ww[3], T W) {
#pragma clang fp contract(off)
  T m[3] = {mx, my, mz};
  T d[3] = {dd[0], dd[1], dd[2]};
  T w[3] = {ww[0], ww[1], ww[2]};
#pragma unroll
  for (int i = 1; i < 3; i++) {
#pragma unroll
    for (int j = 0; j < 3 - i; j++) {
      if (m[j] > m[j + 1]) {
        T tm = m[j], td = d[j], tw = w[j];
        m[j] = m[j + 1];
        d[j] = d[j + 1];
        w[j] = w[j + 1];
        m[j + 1] = tm;
        d[j + 1] = td;
        w[j + 1] = tw;
      }
    }
  }
  T u = m[0] + d[0];
  if (u > m[1]) {
    T a = m[1] - m[0];
    T d20 = d[0] * d[0], d21 = d[1] * d[1];
    T r = sqrt_t(d20 + d21 - a * a);
    u = m[0] + (a * d20 + d[0] * d[1] * r) / (d20 + d21);
    if (u > m[2]) {
      T b = m[2] - m[0];
      T P = a * w[1] + b * w[2];
      T Q = a * a * w[1] + b * b * w[2] - (T)1;
      T D = P * P - W * Q;
      D = D > radicand_floor<T>() ? D : (T)0;
      u = m[0] + (P + sqrt_t(D)) / W;
    }
  }
  return cur < u ? cur : u;
}

template <bool ISO>
__device__ __forceinline__ float god32(float c, float xm, float xp, float ym, float yp, float zm,
                                       float zp, const pdfs_sp32 &s) {
  if (ISO) return god_iso32(c, xm, xp, ym, yp, zm, zp, s.h, s.two_h2, s.three_h2, s.third);
  return god_diff_aniso<float>(c, xm < xp ? xm : xp, ym < yp ? ym : yp, zm < zp ? zm : zp, s.d,
                               s.w, s.W);
}

// FP64 difference form (judge). Spacing constants in double, same op order
// as the oracle's double instantiation.
// Branch-free isotropic FP64 solve for the residual judge: the operations of
// god_diff_iso<double> after its sort, with the 1-D / 2-D / 3-D choice made
// by selects, so the value is the same; lanes of a wave do not diverge over
// the branches. m0 <= m1 <= m2 are the sorted per-axis minima, non-negative
// or +inf: a negative 2-D radicand (NaN root) implies u1 <= m1, so the NaN
// candidate is never selected; with m2 = +inf the 3-D test is false.
__device__ __forceinline__ double god_iso64_sorted(double cur, double m0, double m1, double m2,
                                                   double h, double two_h2, double three_h2,
                                                   double third) {
#pragma clang fp contract(off)
  const double u1 = m0 + h;
  const double a = m1 - m0;
  const double b = m2 - m0;
  const double r2 = two_h2 - a * a;
  const double e = b + (b - a);
  double D = three_h2 - 2.0 * ((a * a + b * b) - a * b);
  D = D > 0.0 ? D : 0.0;
  const bool three = r2 > e * e;
  const double R = three ? D : r2;
  const double P = three ? a + b : a;
  const double K = three ? third : 0.5;
  const double t = m0 + (P + __builtin_sqrt(R)) * K;
  const double u = u1 > m1 ? t : u1;
  return cur < u ? cur : u;
}

__device__ __forceinline__ double god_d64(double cur, double mx, double my, double mz,
                                          double dx, double dy, double dz) {
#pragma clang fp contract(off)
  if (dx == dy && dy == dz) {
    double h = dx, h2 = h * h;
    return god_diff_iso<double>(cur, mx, my, mz, h, 2.0 * h2, 3.0 * h2, 1.0 / 3.0);
  }
  double dd[3] = {dx, dy, dz};
  double ww[3] = {1.0 / (dx * dx), 1.0 / (dy * dy), 1.0 / (dz * dz)};
  double W = ww[0] + ww[1] + ww[2];
  return god_diff_aniso<double>(cur, mx, my, mz, dd, ww, W);
}

// Anisotropic FP64 solve with the spacings scaled by the slowness s: the
// weights 1/d_i^2 and their sum are scaled by r = 1/s^2, in the oracle's
// operation order.
__device__ __forceinline__ double god_d64_speed(double cur, double mx, double my, double mz,
                                                double dx, double dy, double dz, double s) {
#pragma clang fp contract(off)
  const double w0 = 1.0 / (dx * dx), w1 = 1.0 / (dy * dy), w2 = 1.0 / (dz * dz);
  const double W = w0 + w1 + w2;
  const double r = 1.0 / (s * s);
  double dd[3] = {dx * s, dy * s, dz * s};
  double ww[3] = {w0 * r, w1 * r, w2 * r};
  return god_diff_aniso<double>(cur, mx, my, mz, dd, ww, W * r);
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// Buffer-op cache policy bit sc1 (gfx940+): the access is coherent at agent
// scope (all XCDs); 64-bit aligned accesses are single-copy atomic.
#define PDFSK_SC1 16

// ---------------------------------------------------------------------------
// Relaxed agent-scope 64-bit atomics: single-copy atomic, coherent across the
// XCDs, no L2 writeback or invalidate.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ld_edge(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait for an edge entry of the current epoch (2 s of the 100 MHz
// constant clock). SLEEP is the pause between polls in units of 64 clocks;
// the shared error flag is read on every 16th poll only, so a waiting lane
// puts little traffic on the memory side. On timeout the error flag is
// raised and the stale word used, so every wave still exits. Returns the
// entry's 32-bit payload (an FP32 value, or one half of an FP64 value).
template <int SLEEP>
__device__ __noinline__ uint32_t wait_edge(const uint64_t *p, uint32_t epoch, unsigned *err) {
  const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
  uint64_t e = ld_edge(p);
  unsigned n = 0;
  while ((uint32_t)(e >> 32) != epoch) {
    __builtin_amdgcn_s_sleep(SLEEP);
    e = ld_edge(p);
    if ((++n & 15u) == 0u) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (__builtin_amdgcn_s_memrealtime() - t0 > 200000000ull) {
        atomicOr(err, 1u);
        break;
      }
    }
  }
  return (uint32_t)e;
}

// ---------------------------------------------------------------------------
// Edge queue depth and start-gate slack, in steps. Measured at 512^3 with the
// 32x16 tile: depth 2 with slack 0..2 gives the shortest ordering (1.08 ms;
// depth 4: 1.11-1.15 ms, depth 8: 1.25 ms); depth 1 stalls every step on the
// in-order memory counter (the edge load then waits for plane loads issued a
// step before it).
#ifndef PDFSK_UE
#define PDFSK_UE 2
#endif
// Overlapped passes: one ordering's critical path matters less when the next
// ordering fills its drain, so a deeper queue pays (512^3: 17.56 against
// 18.01 ms of sweeps with depth 2; smaller single fields within 0.5%). With
// one launch per ordering, depth 3 costs c4's short columns 32%. Depth 4,
// and start-gate slack 1 or 3 at depth 3, measured no better.
#ifndef PDFSK_UE_OV
#define PDFSK_UE_OV 3
#endif
#ifndef PDFSK_GATE_EXTRA
#define PDFSK_GATE_EXTRA 2
#endif
#ifndef PDFSK_GATE_EXTRA_OV
#define PDFSK_GATE_EXTRA_OV PDFSK_GATE_EXTRA
#endif

// Isotropic FP64 difference form with its own sort: the operations of
// god_diff_iso<double> (the oracle's orc_godunov_d64 for equal spacings),
// with the 1-D / 2-D / 3-D choice made by selects.
__device__ __forceinline__ double god_iso64(double cur, double mx, double my, double mz, double h,
                                            double two_h2, double three_h2, double third) {
#pragma clang fp contract(off)
  const double lo = mx < my ? mx : my;
  const double hi = mx > my ? mx : my;
  const double m0 = lo < mz ? lo : mz;
  const double mid = lo > mz ? lo : mz;
  const double m1 = hi < mid ? hi : mid;
  const double m2 = hi > mid ? hi : mid;
  return god_iso64_sorted(cur, m0, m1, m2, h, two_h2, three_h2, third);
}

// The same solve with the per-axis minima and the sort on v_min_f64 /
// v_max_f64 (one instruction each instead of a compare and two selects).
// Equal to god_iso64 (and the oracle's '<' / '>' selects) for operands that
// are not NaN; a +0 / -0 tie may give the other zero, which no sum with a
// spacing can tell apart.
__device__ __forceinline__ double god_iso64_fminmax(double cur, double xm, double xp, double ym,
                                                    double yp, double zm, double zp, double h,
                                                    double two_h2, double three_h2,
                                                    double third) {
#pragma clang fp contract(off)
  const double mx = __builtin_fmin(xm, xp), my = __builtin_fmin(ym, yp),
               mz = __builtin_fmin(zm, zp);
  const double lo = __builtin_fmin(mx, my);
  const double hi = __builtin_fmax(mx, my);
  const double m0 = __builtin_fmin(lo, mz);
  const double mid = __builtin_fmax(lo, mz);
  const double m1 = __builtin_fmin(hi, mid);
  const double m2 = __builtin_fmax(hi, mid);
  return god_iso64_sorted(cur, m0, m1, m2, h, two_h2, three_h2, third);
}

// Sweep modes of the column pipeline (PDFSK_MODE_*):
//   F32_ISO / F32_ANISO   FP32 difference form, u = min(u, G(u))
//   F64_REF               FP64, the reference's arithmetic
//                         (single_gpu_cuda/src/phi3D_fsm.cu:249-295),
//                         u = min(u, G(u)): bit-identical to the reference
//   F64_POL_ISO / _ANISO  FP64 difference form, u = min(init, G(u)): a
//                         Gauss-Seidel step of the fixed-point map itself,
//                         which may raise a value as well as lower it
//   F32_SPEED, F64_POL_SPEED
//                         the FP32 and polish solves of |grad u| = s with a
//                         slowness s per point: the point's spacings are
//                         d_i s (isotropic: h s; anisotropic: d_i s and
//                         weights w_i / s^2, W / s^2)
//   FX_ISO                32-bit fixed-point values (god_fx_iso), isotropic
//                         spacing, u = min(u, G(u)); the field's quantum
//                         exponent is in its table entry
// FP64 values cross column boundaries as two 32-bit halves, each in its own
// epoch-tagged entry (two edge-wave lanes per value). The per-point inputs
// (init, slowness) are "aux" streams: loaded with the field, plane by
// plane, and kept in a per-thread LDS delay line until the thread's
// hyperplane reaches them.
template <bool F64, bool FX> struct ValT { typedef float T; };
template <> struct ValT<true, false> { typedef double T; };
template <> struct ValT<false, true> { typedef uint32_t T; };
// +inf of a value type (fixed point: UINT32_MAX)
template <typename T> __device__ __forceinline__ T inf_of() { return (T)__builtin_inff(); }
template <> __device__ __forceinline__ uint32_t inf_of<uint32_t>() { return 0xFFFFFFFFu; }
template <int M> struct SMode {
  static constexpr bool F64 = PDFSK_MODE_IS_F64(M);
  static constexpr bool FX = M == PDFSK_MODE_FX_ISO;
  static constexpr bool POL = M == PDFSK_MODE_F64_POL_ISO || M == PDFSK_MODE_F64_POL_ANISO ||
                              M == PDFSK_MODE_F64_POL_SPEED;
  static constexpr bool SPEED = M == PDFSK_MODE_F32_SPEED || M == PDFSK_MODE_F64_POL_SPEED;
  static constexpr bool ISO = M == PDFSK_MODE_F32_ISO || M == PDFSK_MODE_F64_POL_ISO;
  static constexpr int W = F64 ? 2 : 1;  // 32-bit words per value
  static constexpr int NAUX = (POL ? 1 : 0) + (SPEED ? 1 : 0);  // init, then slowness
  static constexpr int AUX_SLOW = POL ? 1 : 0;
  typedef typename ValT<F64, FX>::T T;
};

__device__ __forceinline__ float buf_ld(float, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ double buf_ld(double, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ uint32_t buf_ld(uint32_t, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
__device__ __forceinline__ void buf_st(uint32_t v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, 0);
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off, 0, 0);
}
__device__ __forceinline__ void buf_st(double v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, 0);
}

// Column-pipeline sweep, one ordering.
//
// Virtual coordinates make every ordering (+,+,+): virtual v in [1, dim]
// maps to physical flip ? dim - v : v - 1. Workgroup = one TX x TY column
// tile plus one edge wave; compute thread (tx, ty) updates virtual z =
// l + 1 - tx - ty at step l, so a step sweeps one i+j+k hyperplane of the
// tile. The LDS ring holds R = TX+TY+2 z-planes of the tile with a halo on
// all four sides: east column / north row hold old values (loaded with the
// plane by compute lanes), west column / south row hold the upwind columns'
// results of the same hyperplane, which the edge wave brings in. Plane l+3
// enters the ring at step l and plane l-(TX+TY-2) leaves it (coalesced
// store) at step l.
//
// The edge wave moves every value that crosses a column boundary: during
// step l it loads the epoch-tagged entries the tile's west / south lanes
// need at step l+1 (waiting for an entry its upwind column has not
// published yet) and writes them into the halo cells, and it publishes the
// east / north results of step l-1 from the ring. Compute waves therefore
// issue only plane loads and stores, and the device-coherent edge traffic
// is one load and one store instruction per step with its own in-order
// memory counter. All field and edge accesses are buffer loads/stores with
// a range-checked descriptor: lanes outside the field, outside [1, Z] or
// without an entry use an out-of-range offset (the load returns 0 and puts
// no request on the memory side, the store is dropped), so steps have no
// lane branches.
// ---------------------------------------------------------------------------
// Field table entry (one per field of a batch; a single solve is a batch
// of one). Dense field [Z][Y][X] of the mode's value type; polish modes read
// the dense FP32 init; edge blocks start at ex_off / ey_off (entries from the
// face parts' index 0, padding included).
struct FieldDesc {
  void *f;
  const float *init;
  const float *slow;
  int X, Y, Z;
  int nTx, nTy;
  uint32_t ex_off, ey_off;
  int fxk;  // fixed-point mode: quantum h 2^-fxk
};

// Launch state in device memory, advanced by the last workgroup of each
// launch so that consecutive launches (and graph replays) need no host
// arguments: [0] ticket counter, [1] epoch, [2] ticket base, [3] finished
// workgroups, [4] number of column tickets of the batch.
enum { ST_TICKET = 0, ST_EPOCH = 1, ST_BASE = 2, ST_DONE = 3, ST_NCOLS = 4 };

struct ColArgs {
  const FieldDesc *fields;
  const int *cols;  // ticket -> (field, a, b)
  unsigned *state;
  int fx, fy, fz;   // the ordering's reflections (same for every field)
  uint64_t *edge_x, *edge_y;  // face parts (entry 0 = index -PDFSK_EDGE_PAD) of one allocation
  uint32_t edge_x_bytes, edge_y_bytes;  // whole allocation from edge_x; y part alone
  unsigned *changed;  // per field
  unsigned *streak;   // per field: [2f] counter after the last ordering, [2f+1] unchanged run
  unsigned long long *maxdiff;  // per field, FP64 modes: max |old - new| (bits) of any update
  int nfields;
  unsigned *err;
  pdfs_sp32 sp;
  pdfs_sp64 sp64;
  // overlapped pass (k_sweep_col<..., true>): tickets of orderings 0..ov_n-1
  // of the launch, 5 ints each (ordering slot, field, a, b, physical column)
  int ov_n, ov_nt;          // orderings in the launch, tickets
  int ov_dirs;              // bits 3 o + {0, 1, 2}: reflections x, y, z of ordering slot o
  int ov_ncolp;             // physical columns of the plan (stride of ov_done per slot)
  uint32_t ov_rx, ov_ry;    // bytes between the two edge regions (x part, y part)
  unsigned *ov_done;        // [8][ov_ncolp]: epoch of the last ordering each column finished
  unsigned *ov_cnt;         // [8]: columns finished per ordering slot of this launch
  unsigned *ov_chg;         // [8][nfields]: columns that changed a value, per slot and field
  unsigned long long *ov_md_last;  // FP64 modes, a launch that ends at ordering 8: its largest
                                   // change per field (also counted in maxdiff), else NULL
};


template <int TX, int TY, int W>
struct ColGeom {
  static constexpr int NT = TX * TY;         // compute threads
  static constexpr int NB = NT + 64;         // plus the edge wave
  static constexpr int R = TX + TY + 2;      // ring depth (z-planes)
  static constexpr int PW = TX + 2;          // row pitch: x = -1 .. TX
  static constexpr int PA = PW * (TY + 2);   // rows y = -1 .. TY
  // plane stride in values: 2 mod 32 LDS banks (in 32-bit words)
  static constexpr int PSW = PA * W + ((2 - (PA * W) % 32) + 32) % 32;
  static constexpr int PS = PSW / W;
  static constexpr int DUMMY = R * PS;       // one dummy cell per thread
  static constexpr int LDS_VALUES = DUMMY + NB;
  static constexpr int cell(int x, int y) { return (y + 1) * PW + x + 1; }
};

// Bounded wait until a 32-bit word (written with relaxed agent-scope
// atomics) holds at least `target` (wrapping compare), or the shared error
// flag is raised (2 s timeout, like wait_edge).
__device__ __noinline__ void wait_at_least(const unsigned *p, unsigned target, unsigned *err) {
  const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
  unsigned n = 0;
  while ((int)(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
    __builtin_amdgcn_s_sleep(8);
    if ((++n & 15u) == 0u) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (__builtin_amdgcn_s_memrealtime() - t0 > 200000000ull) {
        atomicOr(err, 1u);
        break;
      }
    }
  }
}

// OV: one launch sweeps several consecutive orderings of a pass ("slots"),
// their columns interleaved in one ticket order. A column starts ordering
// slot o once (1) its upwind columns of o publish their edge entries (the
// start gate and edge waits, as in a single ordering), (2) it and its four
// neighbours have finished slot o - 1 (done flags: release by the producer,
// acquire here, so their plane stores are visible across the XCDs), and (3)
// every column has finished slot o - 2, whose edge region slot o reuses.
// Tickets are ordered so that every dependency holds an earlier ticket,
// which keeps the persistent grid deadlock free.
template <int TX, int TY, int U, int M, bool OV = false>
__global__ __launch_bounds__(TX *TY + 64) void k_sweep_col(ColArgs a) {
  using S = SMode<M>;
  using T = typename S::T;
  constexpr int W = S::W;
  // queue depths: plane loads U steps ahead (HBM latency), edge entries UE
  // steps ahead; GATE = start lag beyond the inherent skew
  constexpr int UE = OV ? PDFSK_UE_OV : PDFSK_UE;
  constexpr int GATE = UE + (OV ? PDFSK_GATE_EXTRA_OV : PDFSK_GATE_EXTRA);
  using G = ColGeom<TX, TY, W>;
  constexpr int NT = G::NT, R = G::R, PW = G::PW, PS = G::PS;
  static_assert(NT % 64 == 0 && W * (TX + TY) <= 64, "compute waves and edge-wave lanes");
  static_assert(G::PSW % W == 0, "plane stride in whole values");
  __shared__ T lds[G::LDS_VALUES];
  // aux streams (init, slowness): per-thread delay lines (plane vz in slot
  // vz mod R)
  constexpr int NA = S::NAUX;
  __shared__ float iring[NA ? NA * R * NT : 1];
  __shared__ int s_col;
  __shared__ unsigned s_changed;
  __shared__ unsigned s_skip;
  T *const ring = lds;
  uint32_t *const lds32 = (uint32_t *)lds;

  const int tid = threadIdx.x;
  const bool edge_wave = tid >= NT;  // wave-uniform: NT is a multiple of 64
  const T INF = inf_of<T>();
  const unsigned ncols = OV ? (unsigned)a.ov_nt
                           : __hip_atomic_load(a.state + ST_NCOLS, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t epoch = __hip_atomic_load(a.state + ST_EPOCH, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
  const unsigned tbase = __hip_atomic_load(a.state + ST_BASE, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
  // one descriptor over both face parts (edge_y follows edge_x)
  const __amdgpu_buffer_rsrc_t re =
      __builtin_amdgcn_make_buffer_rsrc(a.edge_x, (short)0, (int)a.edge_x_bytes, 0x00020000);
  const uint32_t ybase = (uint32_t)((const char *)a.edge_y - (const char *)a.edge_x);
  // stores that do not happen go to an out-of-range offset; the empty asm
  // keeps the compiler from turning the select into a branch around the
  // store, so every step issues the same memory operations
  auto opq = [](uint32_t v) -> uint32_t {
    asm volatile("" : "+v"(v));
    return v;
  };

  for (;;) {
    if (tid == 0) {
      s_col = (int)(atomicAdd(a.state + ST_TICKET, 1u) - tbase);
      s_changed = 0;
      s_skip = 0;
    }
    __syncthreads();
    // the ticket and everything derived from it are wave-uniform: keep them
    // in scalar registers (a divergent buffer descriptor would make every
    // buffer operation a waterfall loop)
    const int col = __builtin_amdgcn_readfirstlane(s_col);
    if ((unsigned)col >= ncols) break;
    constexpr int CW = OV ? 5 : 3;  // ints per ticket
    const int oslot = OV ? __builtin_amdgcn_readfirstlane(a.cols[CW * col]) : 0;
    const int fi = __builtin_amdgcn_readfirstlane(a.cols[CW * col + (OV ? 1 : 0)]);
    if (!OV) {
      // skip the columns of a converged field (its whole field skips, so no
      // column waits on one of them); the barrier keeps thread 0 from
      // taking the next ticket before every wave has read this one
      const unsigned run = __builtin_amdgcn_readfirstlane(__hip_atomic_load(
          a.streak + 2 * fi + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (run >= PDFSK_CONVERGED_RUN) {
        __syncthreads();
        continue;
      }
    }
    const int ca = __builtin_amdgcn_readfirstlane(a.cols[CW * col + (OV ? 2 : 1)]);
    const int cb = __builtin_amdgcn_readfirstlane(a.cols[CW * col + (OV ? 3 : 2)]);
    // the ordering this ticket sweeps: reflections, epoch, edge region
    const int fx = OV ? (a.ov_dirs >> (3 * oslot)) & 1 : a.fx;
    const int fy = OV ? (a.ov_dirs >> (3 * oslot + 1)) & 1 : a.fy;
    const int fz = OV ? (a.ov_dirs >> (3 * oslot + 2)) & 1 : a.fz;
    const uint32_t ep = epoch + (uint32_t)oslot;
    const uint32_t rgx = OV && (oslot & 1) ? a.ov_rx : 0u, rgy = OV && (oslot & 1) ? a.ov_ry : 0u;
    FieldDesc fd = a.fields[fi];
    {
      const uint64_t fp = (uint64_t)fd.f;
      const uint32_t flo = __builtin_amdgcn_readfirstlane((uint32_t)fp);
      const uint32_t fhi = __builtin_amdgcn_readfirstlane((uint32_t)(fp >> 32));
      fd.f = (void *)(((uint64_t)fhi << 32) | flo);
      const uint64_t ip = (uint64_t)fd.init;
      const uint32_t ilo = __builtin_amdgcn_readfirstlane((uint32_t)ip);
      const uint32_t ihi = __builtin_amdgcn_readfirstlane((uint32_t)(ip >> 32));
      fd.init = (const float *)(((uint64_t)ihi << 32) | ilo);
      const uint64_t sp = (uint64_t)fd.slow;
      const uint32_t slo = __builtin_amdgcn_readfirstlane((uint32_t)sp);
      const uint32_t shi = __builtin_amdgcn_readfirstlane((uint32_t)(sp >> 32));
      fd.slow = (const float *)(((uint64_t)shi << 32) | slo);
      fd.X = __builtin_amdgcn_readfirstlane(fd.X);
      fd.Y = __builtin_amdgcn_readfirstlane(fd.Y);
      fd.Z = __builtin_amdgcn_readfirstlane(fd.Z);
      fd.nTy = __builtin_amdgcn_readfirstlane(fd.nTy);
      fd.ex_off = __builtin_amdgcn_readfirstlane(fd.ex_off);
      fd.ey_off = __builtin_amdgcn_readfirstlane(fd.ey_off);
      fd.fxk = __builtin_amdgcn_readfirstlane(fd.fxk);
    }
    const int X = fd.X, Y = fd.Y, Z = fd.Z, nTy = fd.nTy;
    const int NL = Z + TX + TY - 2;  // steps 0..NL; the last one only stores
    const int NS = Z + TX + TY - 1;  // steps (entry rows) per column
    // Edge entries, byte offsets from edge_x. A column's east-face block
    // holds one row of TY values per column step: the lane (TX-1, r) result
    // of step l' is value l' * TY + r, stored as W entries (its 32-bit
    // words), so entry (l' * TY + r) * W + w (north face alike, TX values per
    // step). Lane (0, r) of the downstream column needs at its step s the
    // value of plane s + 1 - r, which its upwind column computed at step
    // s + TX - 1 for every r: a step's entries are written together and read
    // together.
    constexpr uint32_t PADB = PDFSK_EDGE_PAD * 8u;
    const uint32_t exb = PADB + fd.ex_off * 8u + rgx, eyb = ybase + PADB + fd.ey_off * 8u + rgy;
    const int self = ca * nTy + cb;
    // OV: this column's physical tile and flag index; wait for the slot's
    // dependencies (edge-wave lanes 0..5), then make the producers' plane
    // stores visible to every wave of this workgroup
    const int colp = OV ? __builtin_amdgcn_readfirstlane(a.cols[CW * col + 4]) : 0;
    if (OV && oslot > 0) {
      const int nTx = __builtin_amdgcn_readfirstlane(fd.nTx);
      const int pa = fx ? nTx - 1 - ca : ca, pb = fy ? nTy - 1 - cb : cb;
      if (edge_wave) {
        const int e = tid - NT;
        const unsigned *prev = a.ov_done + (size_t)(oslot - 1) * (unsigned)a.ov_ncolp;
        int q = -1;
        if (e == 0) q = colp;
        else if (e == 1 && pa > 0) q = colp - nTy;
        else if (e == 2 && pa + 1 < nTx) q = colp + nTy;
        else if (e == 3 && pb > 0) q = colp - 1;
        else if (e == 4 && pb + 1 < nTy) q = colp + 1;
        if (q >= 0) wait_at_least(prev + q, ep - 1u, a.err);
        // wave 0 of every column of slot o - 2 counted 64; that slot's
        // change counts are then final (each column adds to them before its
        // release). A field that slot changed nothing in is at its fixed
        // point: its columns skip this slot (every one of them reads the
        // same count, so no column waits on a skipped one's edges).
        if (e == 5 && oslot >= 2) {
          wait_at_least(a.ov_cnt + (oslot - 2), 64u * (ncols / (unsigned)a.ov_n), a.err);
          s_skip = __hip_atomic_load(a.ov_chg + (oslot - 2) * a.nfields + fi, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT) == 0u;
        }
      }
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    const bool skip_col = OV && oslot >= 2 && __builtin_amdgcn_readfirstlane(s_skip) != 0u;

    if (skip_col) {
      // at the fixed point: nothing to sweep; the tail still reports the
      // column done
    } else if (edge_wave) {
      // ---- edge wave: value lane v < TX serves column c = v of the south /
      // north faces, value lane TX + r row r of the west / east faces; lane
      // e = v * W + w moves word w of its value; other lanes idle
      const int e = tid - NT;
      const int ev = e / W, ew = e % W;
      const bool isS = ev < TX, isW = !isS && ev < TX + TY;
      const int k = isS ? ev : (isW ? ev - TX : 0);  // the served lane's skew (c or r)
      const bool up = isS ? cb > 0 : (isW && ca > 0);  // an upwind column exists
      // halo word of a boundary lane (+inf in the value's type)
      const uint32_t infw =
          S::FX ? 0xFFFFFFFFu : (W == 1 ? 0x7F800000u : (ew ? 0x7FF00000u : 0u));
      // entry of step s: coff + s * cinc
      uint32_t coff = 0xFFFFFFF0u, cinc = 0u;
      if (isS && cb > 0) {
        coff = eyb + (uint32_t)((((ca * nTy + cb - 1) * NS + TY - 1) * TX + k) * W + ew) * 8u;
        cinc = TX * W * 8u;
      } else if (isW && ca > 0) {
        coff = exb + (uint32_t)(((((ca - 1) * nTy + cb) * NS + TX - 1) * TY + k) * W + ew) * 8u;
        cinc = TY * W * 8u;
      }
      // halo cell written (plane of step s: s + 1 - k) and result cell
      // published (plane of producer step l': l' + 1 - pk)
      const int hcl = isS ? G::cell(k, -1) : (isW ? G::cell(-1, k) : G::DUMMY + tid);
      const int pcl = isS ? G::cell(k, TY - 1) : (isW ? G::cell(TX - 1, k) : G::DUMMY + tid);
      const int pk = isS ? k + TY - 1 : TX - 1 + k;
      const uint32_t pslot = (isS || isW) ? PS : 0u;  // slot stride (idle lanes: fixed dummy)
      uint32_t poff = isS ? eyb + (uint32_t)((self * NS * TX + k) * W + ew) * 8u
                          : (isW ? exb + (uint32_t)((self * NS * TY + k) * W + ew) * 8u
                                 : 0xFFFFFFF0u);
      const uint32_t pinc = isS ? TX * W * 8u : (isW ? TY * W * 8u : 0u);
      // slot (in values) of plane p: (p mod R) * PS
      auto slot_of = [&](int p) -> uint32_t { return (uint32_t)(((p % R) + R) % R) * pslot; };
      uint32_t hs = slot_of(1 - k);   // halo slot of step 0
      uint32_t ps = slot_of(1 - pk);  // result slot of producer step 0
      auto next = [&](uint32_t s) -> uint32_t {
        const uint32_t n = s + pslot;
        return __builtin_elementwise_min(n, n - (uint32_t)R * PS);
      };
      // Start gate: wait until the upwind columns have published the entries
      // of step GATE. A column that starts at least GATE > UE steps behind
      // its producers finds the entries published in the steady state
      // instead of catching up and paying a memory round trip on every step.
      // The pause between polls keeps waiting columns off the memory side.
      // (Upwind columns never wait on this one, so the wait ends; it is
      // bounded like every edge wait.)
      {
        const int gl = Z - 1 < GATE ? Z - 1 : GATE;
        if (up && (unsigned)(gl - k) < (unsigned)Z) {
          const uint64_t *p = a.edge_x + (coff + (uint32_t)gl * cinc) / 8u;
          if ((uint32_t)(ld_edge(p) >> 32) != ep) wait_edge<8>(p, ep, a.err);
        }
      }
      // consume the entry of step s (loaded UE steps earlier) into the halo
      // cell; a stale entry (its upwind column is behind) is waited for
      auto consume = [&](const int s, const u32x2 q) __attribute__((always_inline)) {
        uint32_t v = q.x;
        if (__builtin_expect(up && (unsigned)(s - k) < (unsigned)Z && q.y != ep, 0))
          v = wait_edge<1>(a.edge_x + (coff + (uint32_t)s * cinc) / 8u, ep, a.err);
        lds32[(hcl + hs) * W + ew] = up ? v : infw;
        hs = next(hs);
      };
      // publish the result of producer step lp (complete since the barrier
      // that ended step lp)
      auto publish = [&](const int lp) __attribute__((always_inline)) {
        const uint32_t v = lds32[(pcl + ps) * W + ew];
        ps = next(ps);
        const u32x2 q = {v, ep};
        const bool live = (isS || isW) && (unsigned)(lp - pk) < (unsigned)Z;
        __builtin_amdgcn_raw_buffer_store_b64(q, re, opq(live ? poff : 0xFFFFFFF0u), 0, PDFSK_SC1);
        poff += pinc;
      };
      u32x2 q[UE];
#pragma unroll
      for (int j = 0; j < UE; j++)
        q[j] = __builtin_amdgcn_raw_buffer_load_b64(re, coff + (uint32_t)j * cinc, 0, PDFSK_SC1);
      consume(0, q[0]);
      q[0] = __builtin_amdgcn_raw_buffer_load_b64(re, coff + (uint32_t)UE * cinc, 0, PDFSK_SC1);
      __syncthreads();
      // step l: halo cells of step l + 1 (queue entry (l + 1) % UE, then
      // refilled with step l + 1 + UE) and the results of step l - 1
      auto estep = [&](const int l, const int j) __attribute__((always_inline)) {
        consume(l + 1, q[j]);
        q[j] = __builtin_amdgcn_raw_buffer_load_b64(re, coff + (uint32_t)(l + 1 + UE) * cinc, 0,
                                                   PDFSK_SC1);
        if (l >= 1) publish(l - 1);
        __syncthreads();
      };
      // blocks of UE steps, so the queue index is a constant
      int l = 0;
      for (; l + UE - 1 <= NL; l += UE) {
#pragma unroll
        for (int j = 0; j < UE; j++) estep(l + j, (j + 1) % UE);
      }
#pragma unroll
      for (int j = 0; j < UE; j++)
        if (l + j <= NL) estep(l + j, (j + 1) % UE);
      publish(NL);
    } else {
      // ---- compute waves
      const int tx = tid % TX, ty = tid / TX;
      const uint32_t XYb = (uint32_t)X * (uint32_t)Y * (uint32_t)sizeof(T);
      const uint32_t dZ = fz ? (uint32_t)(0u - XYb) : XYb;  // byte step of virtual z + 1
      const __amdgpu_buffer_rsrc_t fr = __builtin_amdgcn_make_buffer_rsrc(
          fd.f, (short)0, (int)(XYb * (uint32_t)Z), 0x00020000);
      // byte offset of virtual plane vz (negative / >= N*4 wrap to out of range)
      auto zb = [&](int vz) -> uint32_t { return (uint32_t)(fz ? Z - vz : vz - 1) * XYb; };
      auto xyi = [&](int vx, int vy) -> uint32_t {
        const int px = fx ? X - vx : vx - 1, py = fy ? Y - vy : vy - 1;
        return (uint32_t)(py * X + px);
      };
      const int vx = ca * TX + tx + 1, vy = cb * TY + ty + 1;
      const bool dom = vx <= X && vy <= Y;
      // fixed-point mode: the spacing in quanta (2^fxk) and its exact squares
      const uint32_t fxq = S::FX ? 1u << fd.fxk : 0u;
      const float fxh = (float)fxq;
      const float fx_two_h2 = 2.0f * (fxh * fxh), fx_three_h2 = 3.0f * (fxh * fxh);
      const int skew = tx + ty;
      const int lo = G::cell(tx, ty);  // ring cell of this thread

      // interior plane loads / stores
      const uint32_t lxy = dom ? xyi(vx, vy) * (uint32_t)sizeof(T) : 0x80000000u;
      uint32_t ld = zb(3) + lxy;  // next plane to load: 3
      uint32_t st = dom ? zb(1) + lxy : 0xFFFFFFF0u;
      const uint32_t st_inc = dom ? dZ : 0u;
      // halo loaders: lanes [0, TY) east column (row tid), [TY, TY+TX) north
      // row; the LDS cell of slot s is lds[hcell + s * hmul] (other lanes: a
      // fixed dummy cell, hmul = 0)
      int hcell = G::DUMMY + tid;
      int hmul = 0;
      bool hval = false;
      uint32_t hxy = 0x80000000u;
      if (tid < TY) {
        const int hx = ca * TX + TX + 1, hy = cb * TY + tid + 1;
        hcell = G::cell(TX, tid);
        hmul = PS;
        hval = hx <= X && hy <= Y;
        if (hval) hxy = xyi(hx, hy) * (uint32_t)sizeof(T);
      } else if (tid < TY + TX) {
        const int q = tid - TY;
        const int hx = ca * TX + q + 1, hy = cb * TY + TY + 1;
        hcell = G::cell(q, TY);
        hmul = PS;
        hval = hx <= X && hy <= Y;
        if (hval) hxy = xyi(hx, hy) * (uint32_t)sizeof(T);
      }
      uint32_t hl = zb(3) + hxy;
      const int hmask = hmul ? -1 : 0;  // slot offset applies to halo lanes only

      // aux streams (polish: the init; speed: the slowness) of this thread's
      // column, loaded plane by plane with the field (coalesced) and kept in
      // its LDS delay line until the thread's hyperplane reaches them
      const uint32_t XY4 = (uint32_t)X * (uint32_t)Y * 4u;
      __amdgpu_buffer_rsrc_t ir[NA ? NA : 1];
#pragma unroll
      for (int k = 0; k < NA; k++)
        ir[k] = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(S::POL && k == 0 ? fd.init : fd.slow), (short)0, (int)(XY4 * (uint32_t)Z),
            0x00020000);
      const uint32_t dZI = fz ? (uint32_t)(0u - XY4) : XY4;
      auto zbi = [&](int vz) -> uint32_t { return (uint32_t)(fz ? Z - vz : vz - 1) * XY4; };
      const uint32_t lxyi = dom ? xyi(vx, vy) * 4u : 0x80000000u;
      uint32_t ldi = zbi(3) + lxyi;

      // prologue: planes 0 (ghost), 1 and 2 into the ring, then the plane
      // queue (planes 3 .. U+2)
      ring[lo] = INF;
      lds[hcell] = INF;
      {
        const T v1 = buf_ld(T(), fr, zb(1) + lxy);
        const T h1 = buf_ld(T(), fr, zb(1) + hxy);
        const T v2 = buf_ld(T(), fr, zb(2) + lxy);
        const T h2 = buf_ld(T(), fr, zb(2) + hxy);
        const bool live2 = Z >= 2;
        ring[PS + lo] = dom ? v1 : INF;
        lds[hcell + hmul] = hval ? h1 : INF;
        ring[2 * PS + lo] = (dom && live2) ? v2 : INF;
        lds[hcell + 2 * hmul] = (hval && live2) ? h2 : INF;
#pragma unroll
        for (int k = 0; k < NA; k++) {
          iring[(k * R + 1) * NT + tid] = buf_ld(float(), ir[k], zbi(1) + lxyi);
          iring[(k * R + 2) * NT + tid] = buf_ld(float(), ir[k], zbi(2) + lxyi);
        }
      }
      T qP[U], qH[U];
      float qI[NA ? NA : 1][U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        qP[j] = buf_ld(T(), fr, ld);
        qH[j] = buf_ld(T(), fr, hl);
        ld += dZ;
        hl += dZ;
#pragma unroll
        for (int k = 0; k < NA; k++) qI[k][j] = buf_ld(float(), ir[k], ldi);
        ldi += dZI;
      }
      __syncthreads();
      // Ring addressing: plane vz of a cell lives in slot vz mod R. so_c is
      // the offset of slot (vz mod R) for this thread's current plane
      // vz = l + 1 - skew; the step wraps it with an unsigned min (si_c
      // alike in the init delay line).
      constexpr uint32_t RPS = (uint32_t)R * PS, RNT = (uint32_t)R * NT;
      auto wrap_next = [](uint32_t s) -> uint32_t {
        const uint32_t n = s + PS;
        return __builtin_elementwise_min(n, n - RPS);
      };
      const uint32_t slot0 = (uint32_t)(((1 - skew) % R + R) % R);
      uint32_t so_c = slot0 * PS;
      uint32_t si_c = slot0 * NT;
      // registers of step 0: own plane vz (c) and vz + 1 (zp), the east and
      // north cells' plane vz (xp, yp); zm is the own result of the step
      // before
      T c, zp, xp, yp;
      {
        const T *pc = ring + so_c + lo;
        c = pc[0];
        xp = pc[1];
        yp = pc[PW];
        zp = ring[wrap_next(so_c) + lo];
      }
      T zm = INF;
      unsigned nchg = 0;
      T md = (T)0;  // FP64 modes: largest |old - new| of this thread's updates
      uint32_t mdq = 0u;  // fixed point: largest drop (quanta) of this thread's updates
      // uniform slots: plane l+3 is written at step l, plane l-(TX+TY-2) is
      // stored at step l
      int slot_w = 3 % R, slot_s = (R - (TX + TY - 2) % R) % R;

      // One step l. Between the barrier that ended step l-1 and the one
      // that ends step l: (a) read the west / south neighbours' results of
      // step l-1 (the edge wave's halo values at the tile's west / south
      // border) and the finished plane l-(TX+TY-2); (b) write plane l+3 into
      // the ring and refill the plane queue; (d) read the old values step
      // l+1 needs (none of them is written in this step by another thread);
      // (c) update this thread's hyperplane point, publish it to the ring,
      // store the finished plane. `steady`: every lane of the field is
      // active, plane l+3 is inside the field and the stores have begun.
      auto step = [&](const int l, const int j, const bool steady)
                      __attribute__((always_inline)) {
        // (a)
        T *const pc = ring + so_c + lo;
        const T xm = pc[-1], ym = pc[-PW];
        const T sval = ring[slot_s * PS + lo];
        // (b)
        {
          const bool live = steady || l + 3 <= Z;
          ring[slot_w * PS + lo] = (dom && live) ? qP[j] : INF;
          lds[hcell + ((slot_w * PS) & hmask)] = (hval && live) ? qH[j] : INF;
        }
        qP[j] = buf_ld(T(), fr, ld);
        qH[j] = buf_ld(T(), fr, hl);
        ld += dZ;
        hl += dZ;
        float aux[NA ? NA : 1];
#pragma unroll
        for (int k = 0; k < NA; k++) {
          iring[k * R * NT + slot_w * NT + tid] = qI[k][j];
          qI[k][j] = buf_ld(float(), ir[k], ldi);
          aux[k] = iring[k * R * NT + si_c + tid];
        }
        ldi += NA ? dZI : 0u;
        const float ini = S::POL ? aux[0] : 0.0f;
        // (d)
        const uint32_t so_n = wrap_next(so_c);
        const T *pn1 = ring + so_n + lo;
        const T xp1 = pn1[1], yp1 = pn1[PW];
        const T zp1 = ring[wrap_next(so_n) + lo];
        // (c)
        const int vz = l + 1 - skew;
        const bool act = steady ? dom : (dom && (unsigned)(vz - 1) < (unsigned)Z);
        T u;
        if constexpr (M == PDFSK_MODE_F32_SPEED) {
          // spacings scaled by the point's slowness s
          const float sl = aux[S::AUX_SLOW];
          if (a.sp.iso) {
            const float hs = a.sp.h * sl, h2 = hs * hs;
            u = god_iso32(c, xm, xp, ym, yp, zm, zp, hs, 2.0f * h2, 3.0f * h2, a.sp.third);
          } else {
            const float r = 1.0f / (sl * sl);
            const float dd[3] = {a.sp.d[0] * sl, a.sp.d[1] * sl, a.sp.d[2] * sl};
            const float ww[3] = {a.sp.w[0] * r, a.sp.w[1] * r, a.sp.w[2] * r};
            u = god_diff_aniso<float>(c, xm < xp ? xm : xp, ym < yp ? ym : yp, zm < zp ? zm : zp,
                                      dd, ww, a.sp.W * r);
          }
        } else if constexpr (S::FX) {
          u = god_fx_iso(c, xm, xp, ym, yp, zm, zp, fxq, fx_two_h2, fx_three_h2, a.sp.third);
        } else if constexpr (!S::F64) {
          u = god32<S::ISO>(c, xm, xp, ym, yp, zm, zp, a.sp);
        } else if constexpr (M == PDFSK_MODE_F64_POL_SPEED) {
          const double sl = (double)aux[S::AUX_SLOW];
          if (a.sp64.iso) {
            const double hs = a.sp64.h * sl, h2 = hs * hs;
            u = god_iso64_fminmax((double)ini, xm, xp, ym, yp, zm, zp, hs, 2.0 * h2, 3.0 * h2,
                                  a.sp64.third);
          } else {
            const double mx = fx ? (xp < xm ? xp : xm) : (xm < xp ? xm : xp);
            const double my = fy ? (yp < ym ? yp : ym) : (ym < yp ? ym : yp);
            const double mz = fz ? (zp < zm ? zp : zm) : (zm < zp ? zm : zp);
            const double r = 1.0 / (sl * sl);
            const double dd[3] = {a.sp64.d[0] * sl, a.sp64.d[1] * sl, a.sp64.d[2] * sl};
            const double ww[3] = {a.sp64.w[0] * r, a.sp64.w[1] * r, a.sp64.w[2] * r};
            u = god_diff_aniso<double>((double)ini, mx, my, mz, dd, ww, a.sp64.W * r);
          }
        } else if constexpr (M == PDFSK_MODE_F64_POL_ISO) {
          u = god_iso64_fminmax((double)ini, xm, xp, ym, yp, zm, zp, a.sp64.h, a.sp64.two_h2,
                                a.sp64.three_h2, a.sp64.third);
        } else {
          // per-axis minima in the physical order of the reference's
          // min(lower index, higher index) (mpi_openacc/src/utilities.h:35)
          const T mx = fx ? (xp < xm ? xp : xm) : (xm < xp ? xm : xp);
          const T my = fy ? (yp < ym ? yp : ym) : (ym < yp ? ym : yp);
          const T mz = fz ? (zp < zm ? zp : zm) : (zm < zp ? zm : zp);
          if constexpr (M == PDFSK_MODE_F64_REF)
            u = god_ref64(c, mx, my, mz, a.sp64.d[0], a.sp64.d[1], a.sp64.d[2]);
          else
            u = god_diff_aniso<double>((double)ini, mx, my, mz, a.sp64.d, a.sp64.w, a.sp64.W);
        }
        if constexpr (S::F64) {
          const T df = __builtin_fabs(c - u);
          const T dd = df == df ? df : (T)__builtin_inff();
          md = (act && u != c && dd > md) ? dd : md;
        }
        nchg += (act && u != c) ? 1u : 0u;
        if constexpr (S::FX) {
          // u <= c (a min update); lanes off the field's z range are masked
          const uint32_t dq = act ? (uint32_t)(c - u) : 0u;
          mdq = dq > mdq ? dq : mdq;
        }
        // lanes outside the field keep their cells at +inf (the ghost
        // values their neighbours read); inactive lanes inside it write a
        // value no active lane reads before the slot is refilled
        pc[0] = dom ? u : INF;
        {
          const bool sto = steady || l >= TX + TY - 1;
          buf_st(sval, fr, steady ? st : opq(sto ? st : 0xFFFFFFF0u));
          st += sto ? st_inc : 0u;
        }
        zm = (steady || act) ? u : INF;
        c = zp;
        zp = zp1;
        xp = xp1;
        yp = yp1;
        so_c = so_n;
        if (NA) {
          const uint32_t n = si_c + NT;
          si_c = __builtin_elementwise_min(n, n - RNT);
        }
        slot_w = slot_w == R - 1 ? 0 : slot_w + 1;
        slot_s = slot_s == R - 1 ? 0 : slot_s + 1;
        __syncthreads();
      };
      // U steps per block: queue entry j serves step l0 + j. The first block
      // is peeled so that in the loop every queued load is U steps old on
      // entry as on the back edge (the compiler's wait counts are then exact
      // instead of draining each step).
      auto block = [&](const int l0, const bool guard, const bool steady)
                       __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < U; j++)
          if (!guard || l0 + j <= NL) step(l0 + j, j, steady);
      };
      static_assert(TX + TY - 2 >= U - 1, "first block must be full (NL >= U - 1)");
      block(0, false, false);
      int l0 = U;
      for (; l0 + U - 1 <= NL; l0 += U) {
        if (l0 >= TX + TY - 1 && l0 + U + 2 <= Z)
          block(l0, false, true);
        else
          block(l0, false, false);
      }
      if (l0 <= NL) block(l0, true, false);
      if (nchg) atomicAdd(&s_changed, nchg);
      if constexpr (S::F64) {
        // wave max of the change (non-negative doubles order as their bits)
        unsigned long long b = (unsigned long long)__double_as_longlong(md);
        for (int o = 32; o > 0; o >>= 1) {
          const unsigned long long t = __shfl_xor(b, o);
          b = t > b ? t : b;
        }
        if ((tid & 63) == 0 && b) {
          atomicMax(a.maxdiff + fi, b);
          if (OV && a.ov_md_last && oslot == a.ov_n - 1) atomicMax(a.ov_md_last + fi, b);
        }
      }
      if constexpr (S::FX) {
        // wave max of the drop in quanta (an integer in the 64-bit slot)
        uint32_t b = mdq;
        for (int o = 32; o > 0; o >>= 1) {
          const uint32_t t = (uint32_t)__shfl_xor((int)b, o);
          b = t > b ? t : b;
        }
        if ((tid & 63) == 0 && b) atomicMax(a.maxdiff + fi, (unsigned long long)b);
      }
    }
    // OV: every wave's plane stores have completed before the barrier; one
    // agent-scope release then makes them visible across the XCDs ahead of
    // the done flag that the next ordering's columns wait for
    if (OV) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0 && s_changed) {
      atomicAdd(a.changed + fi, s_changed);
      if (OV) atomicAdd(a.ov_chg + oslot * a.nfields + fi, 1u);
    }
    // The release is wave-uniform (wave 0, every lane: the same flag value,
    // 64 counts per column). A lane-0-only tail here let the compiler split
    // wave 0 around the loop's barriers: lanes 1..63 re-entered the loop
    // top's barrier before lane 0 took the next ticket, so every wave read
    // the stale ticket again, without end.
    if (OV && __builtin_amdgcn_readfirstlane(tid) == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_store(a.ov_done + (size_t)oslot * (unsigned)a.ov_ncolp + colp, ep,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(a.ov_cnt + oslot, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  // The last workgroup to finish advances the launch state for the next
  // launch and each field's run of orderings without change. (Release: this
  // workgroup's counter updates before its done count; the last workgroup
  // then sees every workgroup's.)
  __shared__ int s_last;
  if (tid == 0) {
    __threadfence();
    s_last = atomicAdd(a.state + ST_DONE, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (s_last) {
    __threadfence();
    for (int f = tid; f < a.nfields; f += blockDim.x) {
      const unsigned cur =
          __hip_atomic_load(a.changed + f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned last = a.streak[2 * f], run = a.streak[2 * f + 1];
      a.streak[2 * f] = cur;
      if (!OV) {
        a.streak[2 * f + 1] = cur == last ? (run < 0xFFFFu ? run + 1u : run) : 0u;
      } else {
        // the run one launch per ordering would leave: the slots without a
        // change after the last slot with one, or the whole launch added
        unsigned trail = 0u;
        bool any = false;
        for (int o = a.ov_n - 1; o >= 0; o--) {
          unsigned *p = a.ov_chg + o * a.nfields + f;
          const unsigned c = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (c) any = true;
          if (!any) trail++;
          *p = 0u;
        }
        const unsigned add = run + (unsigned)a.ov_n;
        a.streak[2 * f + 1] = any ? trail : (add < 0xFFFFu ? add : 0xFFFFu);
      }
    }
    if (OV && tid < 8) a.ov_cnt[tid] = 0u;
    if (tid == 0) {
      a.state[ST_EPOCH] = epoch + (OV ? (uint32_t)a.ov_n : 1u);
      a.state[ST_BASE] = tbase + ncols + gridDim.x;
      a.state[ST_DONE] = 0u;
    }
  }
}

// ---------------------------------------------------------------------------
// Reference hyperplane schedule, FP64, one launch per level
// (single_gpu_cuda/src/phi3D_fsm.cu:161-185 geometry, :199-233 body).
// ---------------------------------------------------------------------------
// With `slow` (padded, slowness s = 1/F per point) the point's spacings are
// scaled by its slowness: the Eikonal |grad u| = s (a speed field; the
// reference declares Phi.F and never uses it).
__global__ __launch_bounds__(256) void k_level_ref64(double *d, const double *__restrict__ slow,
                                                     int X, int Y, int Z,
                                                     int level, int xs, int ys, int xe,
                                                     int ye, int xo, int yo, int zo,
                                                     double dx, double dy, double dz,
                                                     unsigned long long *maxdiff) {
  const int x = xs + blockIdx.x * blockDim.x + threadIdx.x;
  const int y = ys + blockIdx.y * blockDim.y + threadIdx.y;
  double diff = 0.0;
  if (x <= xe && y <= ye) {
    const int z = level - (x + y);
    if (z > 0 && z <= Z) {
      const int k = abs(x - xo), j = abs(y - yo), i = abs(z - zo);
      const long ex = X + 2, exy = ex * (Y + 2);
      const long id = (long)i * exy + (long)j * ex + k;
      const double c = d[id];
      const double l = d[id - 1], r = d[id + 1];
      const double dn = d[id - ex], up = d[id + ex];
      const double fr = d[id - exy], bk = d[id + exy];
      const double mx = l < r ? l : r;
      const double my = dn < up ? dn : up;
      const double mz = fr < bk ? fr : bk;
      const double sp = slow ? slow[id] : 1.0;
      const double u = slow ? god_ref64(c, mx, my, mz, dx * sp, dy * sp, dz * sp)
                            : god_ref64(c, mx, my, mz, dx, dy, dz);
      d[id] = u;
      if (u != c) {
        double df = __builtin_fabs(c - u);
        diff = (df == df) ? df : __builtin_inf();
      }
    }
  }
  if (maxdiff) {
    // wave max then one atomic per wave
    unsigned long long b = __double_as_longlong(diff);
    for (int o = 32; o > 0; o >>= 1) {
      unsigned long long t = __shfl_xor(b, o);
      b = t > b ? t : b;
    }
    if ((threadIdx.x + threadIdx.y * blockDim.x) % 64 == 0 && b) atomicMax(maxdiff, b);
  }
}

// ---------------------------------------------------------------------------
// FP64 residual judge of a dense FP32 field (new). tol = |e| 2^-22 + hmin 2^-21.
// A workgroup owns a 64 x 4 (x, y) tile of columns over JUDGE_Z planes; each
// thread walks its column in z. A point's eight operands (its value, six
// neighbours and init) are loaded while the plane before it is judged; of
// them only the plane above and the init are new to the caches, and no
// index division is needed.
// Tiles are dealt XCD-contiguously: workgroups go to the 8 XCDs round robin
// (workgroup b runs on XCD b % 8), so workgroup b takes tile
// (b % 8) * per + b / 8 and each XCD sweeps one contiguous part of the field,
// whose neighbour rows then sit in its own L2. A point that fails sets its
// flag bit (atomic OR, flags zeroed by the caller); the count is summed per
// wave.
// ---------------------------------------------------------------------------
#define JUDGE_Z 32

// Element i (< 2^29) of a device array, loaded as global memory with a
// 32-bit byte offset from a wave-uniform base. Pointers read from a table in
// memory are generic to the compiler, which would otherwise emit flat loads
// (counted against both the vector-memory and the LDS counters) with 64-bit
// per-lane address arithmetic.
__device__ __forceinline__ float ld_global(const float *base, uint32_t i) {
  typedef const __attribute__((address_space(1))) char gchar;
  typedef const __attribute__((address_space(1))) float gfloat;
  return *(gfloat *)((gchar *)base + (i << 2));
}
// tiles of a field (64 x 4 columns x JUDGE_Z planes)
__host__ __device__ inline uint32_t judge_tiles(int X, int Y, int Z) {
  return (uint32_t)((X + 63) / 64) * (uint32_t)((Y + 3) / 4) * (uint32_t)((Z + JUDGE_Z - 1) / JUDGE_Z);
}

// Per-lane constants of a judge column.
struct JudgeLane {
  uint32_t XY, col;                // plane size; (x, y) offset of the column
  uint32_t oxm, oxp, oym, oyp;     // in-plane neighbour offsets (0 outside)
  int Z;
  bool in, hxm, hxp, hym, hyp;     // judged lane; neighbour inside the field
  double htol;                     // hmin 2^-21
  double fb, fd, ftol;             // fixed-point fields: base, quantum, tolerance
};
// The operands of one point: its value, its six neighbours' and its init.
struct JudgeOps {
  float zm, cur, zp, xm, xp, ym, yp, ini, sl;
};

// Operands of plane zl (clamped to the field: the planes past its end are
// loaded but never judged). Out-of-field neighbours load the point itself
// and are replaced by +inf in judge_plane.
template <bool SPEED>
__device__ __forceinline__ JudgeOps judge_load(const float *v, const float *init,
                                               const float *slow, const JudgeLane &L, int zl) {
  const int zc = zl < L.Z ? zl : L.Z - 1;
  const uint32_t p = (uint32_t)zc * L.XY + L.col;
  JudgeOps o;
  o.zm = ld_global(v, zc > 0 ? p - L.XY : p);
  o.cur = ld_global(v, p);
  o.zp = ld_global(v, zc + 1 < L.Z ? p + L.XY : p);
  o.xm = ld_global(v, p - L.oxm);
  o.xp = ld_global(v, p + L.oxp);
  o.ym = ld_global(v, p - L.oym);
  o.yp = ld_global(v, p + L.oyp);
  o.ini = ld_global(init, p);
  o.sl = SPEED ? ld_global(slow, p) : 1.0f;
  return o;
}

// Judge the point of plane z (< Z) whose operands are o: sets its flag bit
// when it fails and returns the number of failing points of the wave.
template <bool ISO, bool SPEED>
__device__ __forceinline__ unsigned judge_plane(const JudgeOps &o, const JudgeLane &L, int z,
                                                double dx, double dy, double dz,
                                                uint32_t *flags) {
#pragma clang fp contract(off)
  const float FINF = __builtin_inff();
  const float zm = z > 0 ? o.zm : FINF, zp = z + 1 < L.Z ? o.zp : FINF;
  const float xm = L.hxm ? o.xm : FINF, xp = L.hxp ? o.xp : FINF;
  const float ym = L.hym ? o.ym : FINF, yp = L.hyp ? o.yp : FINF;
  // Per axis the smaller neighbour, and for the isotropic solve the sort,
  // in FP32: FP32 values are exact in FP64, so both commute with the
  // conversion (non-NaN fields; a NaN operand makes the point fail).
  const float mx = __builtin_elementwise_minimum(xm, xp);
  const float my = __builtin_elementwise_minimum(ym, yp);
  const float mz = __builtin_elementwise_minimum(zm, zp);
  double e;
  // speed: the point's spacings scaled by its slowness
  const double sl = SPEED ? (double)o.sl : 1.0;
  if (ISO) {
    const float m0 = __builtin_elementwise_minimum(__builtin_elementwise_minimum(mx, my), mz);
    const float m2 = __builtin_elementwise_maximum(__builtin_elementwise_maximum(mx, my), mz);
    const float m1 = __builtin_amdgcn_fmed3f(mx, my, mz);
    const double h = SPEED ? dx * sl : dx;
    e = god_iso64_sorted((double)o.ini, (double)m0, (double)m1, (double)m2, h, 2.0 * (h * h),
                         3.0 * (h * h), 1.0 / 3.0);
  } else if (SPEED) {
    e = god_d64_speed((double)o.ini, (double)mx, (double)my, (double)mz, dx, dy, dz, sl);
  } else {
    e = god_d64((double)o.ini, (double)mx, (double)my, (double)mz, dx, dy, dz);
  }
  // |v - e| <= |e| 2^-22 + hmin 2^-21, and equal infinities pass: with |e|
  // capped at DBL_MAX an infinite e against a finite v (or the reverse)
  // fails, as does a NaN.
  const double dv = (double)o.cur;
  const double ae = __builtin_fabs(e) < 0x1.fffffffffffffp+1023 ? __builtin_fabs(e)
                                                                : 0x1.fffffffffffffp+1023;
  const double htol = SPEED ? L.htol * sl : L.htol;
  const bool bad = L.in && !(dv == e) && !(__builtin_fabs(dv - e) <= ae * 0x1.0p-22 + htol);
  if (bad && flags) {
    const uint32_t p = (uint32_t)z * L.XY + L.col;
    // a global (not flat) atomic: a flat one in the loop would make the
    // compiler wait for every load in flight, prefetches included
    typedef __attribute__((address_space(1))) uint32_t gu32;
    __hip_atomic_fetch_or((gu32 *)(flags + (p >> 5)), 1u << (p & 31u), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
  }
  return (unsigned)__popcll(__ballot(bad));
}

// A point of a fixed-point field (value = base + q d, q = UINT32_MAX for
// +inf; the operand words hold q): the per-axis minima and the sort are
// taken on q (the decoding is increasing), the FP64 isotropic solve on the
// decoded values; the point fails when |value - e| > ftol (equal
// infinities pass, NaN fails). Decoding is base + (double)q * d without
// contraction, as the fixed-to-FP32 conversion computes it.
__device__ __forceinline__ unsigned judge_plane_fx(const JudgeOps &o, const JudgeLane &L, int z,
                                                   double h, uint32_t *flags) {
#pragma clang fp contract(off)
  const uint32_t QI = 0xFFFFFFFFu;
  const uint32_t zm = z > 0 ? __float_as_uint(o.zm) : QI;
  const uint32_t zp = z + 1 < L.Z ? __float_as_uint(o.zp) : QI;
  const uint32_t xm = L.hxm ? __float_as_uint(o.xm) : QI, xp = L.hxp ? __float_as_uint(o.xp) : QI;
  const uint32_t ym = L.hym ? __float_as_uint(o.ym) : QI, yp = L.hyp ? __float_as_uint(o.yp) : QI;
  const uint32_t mx = xm < xp ? xm : xp, my = ym < yp ? ym : yp, mz = zm < zp ? zm : zp;
  const uint32_t lo = mx < my ? mx : my, hi = mx > my ? mx : my;
  const uint32_t m0 = lo < mz ? lo : mz, mid = lo > mz ? lo : mz;
  const uint32_t m1 = hi < mid ? hi : mid, m2 = hi > mid ? hi : mid;
  const double inf = __builtin_inf();
  auto dec = [&](uint32_t q) -> double { return q == QI ? inf : L.fb + (double)q * L.fd; };
  const double e = god_iso64_sorted((double)o.ini, dec(m0), dec(m1), dec(m2), h, 2.0 * (h * h),
                                    3.0 * (h * h), 1.0 / 3.0);
  const double dv = dec(__float_as_uint(o.cur));
  const bool bad = L.in && !(dv == e) && !(__builtin_fabs(dv - e) <= L.ftol);
  if (bad && flags) {
    const uint32_t p = (uint32_t)z * L.XY + L.col;
    typedef __attribute__((address_space(1))) uint32_t gu32;
    __hip_atomic_fetch_or((gu32 *)(flags + (p >> 5)), 1u << (p & 31u), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
  }
  return (unsigned)__popcll(__ballot(bad));
}

// Fields of a batch: field f owns tiles [tab[f].tile0, tab[f+1].tile0)
// (tab[nf].tile0 = total); flags optional per field. FX: fixed-point fields
// (isotropic, unit speed; dx is the spacing).
template <bool ISO, bool SPEED, bool FX = false>
__global__ __launch_bounds__(256) void k_judge_f32(const pdfsk_judge_field *__restrict__ tab,
                                                   int nf, double dx, double dy, double dz,
                                                   unsigned long long *counts) {
#pragma clang fp contract(off)
  const uint32_t per = gridDim.x / 8u;
  uint32_t t = (blockIdx.x % 8u) * per + blockIdx.x / 8u;
  if (t >= (uint32_t)tab[nf].tile0) return;  // the grid is a multiple of 8 workgroups
  int f = 0;
  {
    int hi = nf;
    while (hi - f > 1) {
      const int mid = (f + hi) >> 1;
      if ((uint32_t)tab[mid].tile0 <= t) f = mid; else hi = mid;
    }
  }
  const pdfsk_judge_field fd = tab[f];
  const float *__restrict__ v = fd.v;
  const float *__restrict__ init = fd.init;
  const float *__restrict__ slow = fd.slow;
  const int X = fd.X, Y = fd.Y, Z = fd.Z;
  t -= (uint32_t)fd.tile0;
  const int ntx = (X + 63) >> 6, nty = (Y + 3) >> 2;
  const int bx = (int)(t % (uint32_t)ntx);
  t /= (uint32_t)ntx;
  const int by = (int)(t % (uint32_t)nty);
  const int bz = (int)(t / (uint32_t)nty);
  const int x = bx * 64 + (int)(threadIdx.x & 63), y = by * 4 + (int)(threadIdx.x >> 6);
  const int z0 = bz * JUDGE_Z, z1 = min(z0 + JUDGE_Z, Z);
  // Lanes past the field's x / y end load the clamped column and judge
  // nothing, so every load is in bounds and none needs a branch.
  const int xc = min(x, X - 1), yc = min(y, Y - 1);
  JudgeLane L;
  L.in = x < X && y < Y;
  // FP32 fields hold fewer than 2^29 points: 32-bit index arithmetic
  L.XY = (uint32_t)X * (uint32_t)Y;
  L.Z = Z;
  L.col = (uint32_t)yc * (uint32_t)X + (uint32_t)xc;
  // in-plane neighbours outside the field are +inf: their offset is 0 (the
  // point itself is loaded) and the value is replaced
  L.hxm = xc > 0;
  L.hxp = xc < X - 1;
  L.hym = yc > 0;
  L.hyp = yc < Y - 1;
  L.oxm = L.hxm ? 1u : 0u;
  L.oxp = L.hxp ? 1u : 0u;
  L.oym = L.hym ? (uint32_t)X : 0u;
  L.oyp = L.hyp ? (uint32_t)X : 0u;
  const double hmin = dx < dy ? (dx < dz ? dx : dz) : (dy < dz ? dy : dz);
  L.htol = hmin * 0x1.0p-21;
  L.fb = fd.fx_base;
  L.fd = fd.fx_d;
  L.ftol = fd.fx_tol;
  unsigned long long nbad = 0;
  auto judge = [&](const JudgeOps &o, int zz) -> unsigned {
    if constexpr (FX)
      return judge_plane_fx(o, L, zz, dx, fd.flags);
    else
      return judge_plane<ISO, SPEED>(o, L, zz, dx, dy, dz, fd.flags);
  };
  // Two operand sets, alternately loaded one plane ahead of the plane being
  // judged: the loads of plane z + 1 are in flight while plane z is solved.
  JudgeOps A = judge_load<SPEED>(v, init, slow, L, z0), B;
  int z = z0;
  for (; z + 1 < z1; z += 2) {
    B = judge_load<SPEED>(v, init, slow, L, z + 1);
    nbad += judge(A, z);
    A = judge_load<SPEED>(v, init, slow, L, z + 2);
    nbad += judge(B, z + 1);
  }
  if (z < z1) nbad += judge(A, z);
  if ((threadIdx.x & 63) == 0 && nbad) atomicAdd(counts + f, nbad);
}

// ---------------------------------------------------------------------------
// Layout conversion and fills.
// ---------------------------------------------------------------------------
__global__ void k_pack_f64(const double *__restrict__ pd, float *__restrict__ dn, int X,
                           int Y, int Z) {
  const long N = (long)X * Y * Z;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long)gridDim.x * blockDim.x) {
    const int x = (int)(p % X);
    const long q = p / X;
    const int y = (int)(q % Y), z = (int)(q / Y);
    dn[p] = (float)pd[((long)(z + 1) * (Y + 2) + (y + 1)) * (X + 2) + (x + 1)];
  }
}

__global__ void k_unpack_f64(const float *__restrict__ dn, double *__restrict__ pd, int X,
                             int Y, int Z) {
  const long N = (long)X * Y * Z;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long)gridDim.x * blockDim.x) {
    const int x = (int)(p % X);
    const long q = p / X;
    const int y = (int)(q % Y), z = (int)(q / Y);
    pd[((long)(z + 1) * (Y + 2) + (y + 1)) * (X + 2) + (x + 1)] = (double)dn[p];
  }
}

__global__ void k_pack64(const double *__restrict__ pd, double *__restrict__ dn, int X, int Y,
                         int Z) {
  const long N = (long)X * Y * Z;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long)gridDim.x * blockDim.x) {
    const int x = (int)(p % X);
    const long q = p / X;
    const int y = (int)(q % Y), z = (int)(q / Y);
    dn[p] = pd[((long)(z + 1) * (Y + 2) + (y + 1)) * (X + 2) + (x + 1)];
  }
}

__global__ void k_unpack64(const double *__restrict__ dn, double *__restrict__ pd, int X, int Y,
                           int Z) {
  const long N = (long)X * Y * Z;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long)gridDim.x * blockDim.x) {
    const int x = (int)(p % X);
    const long q = p / X;
    const int y = (int)(q % Y), z = (int)(q / Y);
    pd[((long)(z + 1) * (Y + 2) + (y + 1)) * (X + 2) + (x + 1)] = dn[p];
  }
}

// Ghost cells (any index 0 or dim + 1) whose bits are not those of +inf.
__global__ void k_count_ghost_not_inf(const double *__restrict__ pd, int X, int Y, int Z,
                                      unsigned long long *count) {
  const long n = (long)(X + 2) * (Y + 2) * (Z + 2);
  unsigned long long bad = 0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long)gridDim.x * blockDim.x) {
    const int k = (int)(p % (X + 2));
    const long q = p / (X + 2);
    const int j = (int)(q % (Y + 2)), i = (int)(q / (Y + 2));
    const bool ghost = i == 0 || i == Z + 1 || j == 0 || j == Y + 1 || k == 0 || k == X + 1;
    bad += ghost && __double_as_longlong(pd[p]) != 0x7FF0000000000000ll;
  }
  if (bad) atomicAdd(count, bad);
}

// dir 0: widen FP32 -> FP64 (exact); dir 1: narrow FP64 -> FP32 (nearest
// even). Field f is handled by the workgroups of row blockIdx.y.
__global__ void k_convert_batch(const pdfsk_conv *__restrict__ tab, int dir) {
  const pdfsk_conv t = tab[blockIdx.y];
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < t.n;
       p += (long long)gridDim.x * blockDim.x) {
    if (dir == 0)
      t.f64[p] = (double)t.f32[p];
    else
      t.f32[p] = (float)t.f64[p];
  }
}

// Fixed-point storage. Order key of an FP32 value: non-negative values map
// to their bits with the sign bit set, negative values to their inverted
// bits, so unsigned key order is value order (-inf < ... < -0 < +0 < ... <
// +inf); NaN is skipped. Field f is handled by the workgroups of row
// blockIdx.y.
__device__ __forceinline__ uint32_t fx_key(float v) {
  const uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
// The conversion kernels move 16 bytes per lane where the field is 16-byte
// aligned (four values; elementwise the same operations), then the rest
// one value per lane.
__device__ __forceinline__ uint32_t fx_key_valid(float v) {
  return v == v ? fx_key(v) : 0xFFFFFFFFu;
}
__device__ __forceinline__ uint32_t fx_key_min4(float4 a) {
  return min(min(fx_key_valid(a.x), fx_key_valid(a.y)), min(fx_key_valid(a.z), fx_key_valid(a.w)));
}
// Four 16-byte loads in flight per lane; the wave minima meet in LDS, so one
// atomic per workgroup reaches the field's key.
__global__ __launch_bounds__(256) void k_fx_min_batch(const pdfsk_fx *__restrict__ tab) {
  __shared__ uint32_t s_min[4];
  const pdfsk_fx t = tab[blockIdx.y];
  // the grid is sized for the largest field of the batch: workgroups past
  // this field's points leave at once (uniformly, before the barrier)
  if ((long long)blockIdx.x * blockDim.x >= t.n) return;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  uint32_t k0 = 0xFFFFFFFFu, k1 = 0xFFFFFFFFu, k2 = 0xFFFFFFFFu, k3 = 0xFFFFFFFFu;
  long long head = 0;
  if (((uintptr_t)t.f & 15u) == 0) {
    const float4 *f4 = (const float4 *)t.f;
    const long long n4 = t.n >> 2;
    long long p = tid;
    for (; p + 3 * stride < n4; p += 4 * stride) {
      const float4 a = f4[p], b = f4[p + stride], c = f4[p + 2 * stride],
                   d = f4[p + 3 * stride];
      k0 = min(k0, fx_key_min4(a));
      k1 = min(k1, fx_key_min4(b));
      k2 = min(k2, fx_key_min4(c));
      k3 = min(k3, fx_key_min4(d));
    }
    for (; p < n4; p += stride) k0 = min(k0, fx_key_min4(f4[p]));
    head = n4 << 2;
  }
  for (long long p = head + tid; p < t.n; p += stride) k0 = min(k0, fx_key_valid(t.f[p]));
  uint32_t k = min(min(k0, k1), min(k2, k3));
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t o2 = (uint32_t)__shfl_xor((int)k, o);
    k = o2 < k ? o2 : k;
  }
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    k = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
    if (k != 0xFFFFFFFFu) atomicMin(t.key, k);
  }
}
// q = rint((v - base) / d) below UINT32_MAX, else UINT32_MAX (and for NaN,
// +inf); contraction off so that the oracle's C gives the same bits
__device__ __forceinline__ uint32_t fx_enc1(float v, double base, double d) {
#pragma clang fp contract(off)
  const double r = __builtin_rint(((double)v - base) / d);
  return r < 4294967295.0 ? (uint32_t)r : 0xFFFFFFFFu;
}
__device__ __forceinline__ float fx_dec1(uint32_t v, double base, double d) {
#pragma clang fp contract(off)
  return v == 0xFFFFFFFFu ? __builtin_inff() : (float)(base + (double)v * d);
}
__global__ void k_fx_encode_batch(const pdfsk_fx *__restrict__ tab) {
  const pdfsk_fx t = tab[blockIdx.y];
  uint32_t *q = (uint32_t *)t.f;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long head = 0;
  if (((uintptr_t)t.f & 15u) == 0) {
    const long long n4 = t.n >> 2;
    for (long long p = tid; p < n4; p += stride) {
      const float4 a = ((const float4 *)t.f)[p];
      uint4 o;
      o.x = fx_enc1(a.x, t.base, t.d);
      o.y = fx_enc1(a.y, t.base, t.d);
      o.z = fx_enc1(a.z, t.base, t.d);
      o.w = fx_enc1(a.w, t.base, t.d);
      ((uint4 *)q)[p] = o;
    }
    head = n4 << 2;
  }
  for (long long p = head + tid; p < t.n; p += stride) q[p] = fx_enc1(t.f[p], t.base, t.d);
}
__global__ void k_fx_decode_batch(const pdfsk_fx *__restrict__ tab) {
  const pdfsk_fx t = tab[blockIdx.y];
  const uint32_t *q = (const uint32_t *)t.f;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long head = 0;
  if (((uintptr_t)t.f & 15u) == 0) {
    const long long n4 = t.n >> 2;
    for (long long p = tid; p < n4; p += stride) {
      const uint4 a = ((const uint4 *)q)[p];
      float4 o;
      o.x = fx_dec1(a.x, t.base, t.d);
      o.y = fx_dec1(a.y, t.base, t.d);
      o.z = fx_dec1(a.z, t.base, t.d);
      o.w = fx_dec1(a.w, t.base, t.d);
      ((float4 *)t.f)[p] = o;
    }
    head = n4 << 2;
  }
  for (long long p = head + tid; p < t.n; p += stride) t.f[p] = fx_dec1(q[p], t.base, t.d);
}

// The counters the host reads after a sweep segment, in one launch: each
// field's changed count, streak pair and (with md) largest changes, written
// to pinned host memory; the largest changes are then zeroed for the next
// segment.
__global__ void k_read_counters(const unsigned *__restrict__ changed, const unsigned *err,
                                const unsigned *__restrict__ streak, unsigned long long *md,
                                int nf, int cap_md, unsigned *h_changed, unsigned *h_err,
                                unsigned *h_streak, unsigned long long *h_md) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
    h_changed[f] = changed[f];
    h_streak[2 * f] = streak[2 * f];
    h_streak[2 * f + 1] = streak[2 * f + 1];
    if (md) {
      h_md[f] = md[f];
      h_md[cap_md + f] = md[cap_md + f];
      md[f] = 0ull;
      md[cap_md + f] = 0ull;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) h_err[0] = err[0];
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// Synthetic initial fields (same formulas and draw order as the host
// parameter generator; grid point (k, j, i) at ((k-1)h, (j-1)h, (i-1)h)).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void put_val(float *dn, double *pd, int X, int Y, int x0, int y0,
                                        int z0, double val) {
  // x0, y0, z0 are 0-based interior coordinates
  if (dn) dn[((long)z0 * Y + y0) * X + x0] = (float)val;
  if (pd) pd[((long)(z0 + 1) * (Y + 2) + (y0 + 1)) * (X + 2) + (x0 + 1)] = val;
}

__global__ void k_gen_ghost64(double *pd, int X, int Y, int Z) {
  const long n = (long)(X + 2) * (Y + 2) * (Z + 2);
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long)gridDim.x * blockDim.x) {
    const int k = (int)(p % (X + 2));
    const long q = p / (X + 2);
    const int j = (int)(q % (Y + 2)), i = (int)(q / (Y + 2));
    if (i == 0 || i == Z + 1 || j == 0 || j == Y + 1 || k == 0 || k == X + 1)
      pd[p] = __builtin_inf();
  }
}

__global__ void k_gen_point(float *dn, double *pd, int X, int Y, int Z, int k, int j, int i) {
  const long N = (long)X * Y * Z;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long)gridDim.x * blockDim.x) {
    const int x = (int)(p % X);
    const long q = p / X;
    const int y = (int)(q % Y), z = (int)(q / Y);
    const double val = (x == k - 1 && y == j - 1 && z == i - 1) ? 0.0 : 1e6;
    put_val(dn, pd, X, Y, x, y, z, val);
  }
}

// params: nhills x {A, S, CX, CY}
__global__ void k_gen_terrain(float *dn, double *pd, int X, int Y, int Z, double h,
                              const double *__restrict__ prm, int nh) {
#pragma clang fp contract(off)
  const long NXY = (long)X * Y;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < NXY;
       p += (long)gridDim.x * blockDim.x) {
    const int x0 = (int)(p % X), y0 = (int)(p / X);
    const double x = (double)x0 * h, y = (double)y0 * h;
    double H = 0.0;
    for (int q = 0; q < nh; q++) {
      const double A = prm[4 * q], S = prm[4 * q + 1], cx = prm[4 * q + 2], cy = prm[4 * q + 3];
      const double ex = x - cx, ey = y - cy;
      H += A * exp(-(ex * ex + ey * ey) / (2.0 * S * S));
    }
    for (int z0 = 0; z0 < Z; z0++) {
      const double off = (double)z0 * h - H;
      const double val = off <= 0.0 ? 0.0 : (off <= 1.5 * h ? off : 1e6);
      put_val(dn, pd, X, Y, x0, y0, z0, val);
    }
  }
}

// params: ns x {cx, cy, cz, r}
__global__ void k_gen_spheres(float *dn, double *pd, int X, int Y, int Z, double h,
                              const double *__restrict__ prm, int ns) {
#pragma clang fp contract(off)
  const long N = (long)X * Y * Z;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (long)gridDim.x * blockDim.x) {
    const int x0 = (int)(p % X);
    const long q = p / X;
    const int y0 = (int)(q % Y), z0 = (int)(q / Y);
    const double x = (double)x0 * h, y = (double)y0 * h, z = (double)z0 * h;
    double best = __builtin_inf();
    for (int s = 0; s < ns; s++) {
      const double ex = x - prm[4 * s], ey = y - prm[4 * s + 1], ez = z - prm[4 * s + 2];
      const double sd = __builtin_sqrt(ex * ex + ey * ey + ez * ez) - prm[4 * s + 3];
      if (sd < best) best = sd;
    }
    const double val = best <= 0.0 ? 0.0 : (best <= 1.5 * h ? best : 1e6);
    put_val(dn, pd, X, Y, x0, y0, z0, val);
  }
}

// ---------------------------------------------------------------------------
// CLI pre/post processing on the padded FP64 field and int location array.
// ---------------------------------------------------------------------------
// single_gpu_cuda/src/phi3D.c:165-180: non-ghost values that are not +inf are
// kept when nonzero, and 0 or NaN becomes 90000 (the location==1 && +inf
// branch of the reference is unreachable under its outer test).
__global__ void k_update_distance(const int *__restrict__ loc, double *__restrict__ d, long n) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long)gridDim.x * blockDim.x) {
    const double v = d[p];
    if (loc[p] != -1 && v != __builtin_inf()) d[p] = (v > 0.0 || v < 0.0) ? v : 90000.0;
  }
}

// single_gpu_cuda/src/phi3D.c:190-205: location 1 (inside) becomes -1.
__global__ void k_set_negative_inside(const int *__restrict__ loc, double *__restrict__ d,
                                      long n) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long)gridDim.x * blockDim.x) {
    if (loc[p] != -1 && d[p] != __builtin_inf() && loc[p] == 1) d[p] = -1.0;
  }
}

// single_gpu_cuda/src/phi3D.c:221-242: every ghost cell takes the value of
// the interior cell reached by clamping each index into [1, dim]. Reads
// only interior cells and writes only ghost cells, so one pass is exact.
__global__ void k_adjust_boundary(double *d, int X, int Y, int Z) {
  const int x = X + 2, y = Y + 2, z = Z + 2;
  const long n = (long)x * y * z;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long)gridDim.x * blockDim.x) {
    const int k = (int)(p % x);
    const long q = p / x;
    const int j = (int)(q % y), i = (int)(q / y);
    const int I = i == z - 1 ? i - 1 : (i == 0 ? i + 1 : i);
    const int J = j == y - 1 ? j - 1 : (j == 0 ? j + 1 : j);
    const int K = k == x - 1 ? k - 1 : (k == 0 ? k + 1 : k);
    if (i != I || j != J || k != K) d[p] = d[((long)I * y + J) * x + K];
  }
}

// ---------------------------------------------------------------------------
// C-ABI launchers
// ---------------------------------------------------------------------------
static inline int hip_rc(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }
static inline int grid_for(long n, int bs) {
  long g = (n + bs - 1) / bs;
  if (g > 65536L * 16) g = 65536L * 16;
  return (int)(g < 1 ? 1 : g);
}

// Column pipeline variants: tile shape and a cap on workgroups per CU (0 =
// as many as fit). A step of one column is latency-bound (barrier, LDS
// round trip, dependent Godunov chain); several columns per CU mostly
// compete for the same issue slots, so capped variants can be faster.
struct ColVariant {
  int tx, ty, per_cu;
};
static const ColVariant kVariants[PDFSK_TILE_COUNT] = {
    {16, 16, 0}, {32, 8, 0}, {16, 8, 0}, {32, 16, 0}, {16, 16, 1}, {16, 8, 2}, {8, 8, 0},
    {16, 4, 0}};

extern "C" void pdfsk_tile_dims(int variant, int *tx, int *ty) {
  const ColVariant &v = kVariants[(unsigned)variant < PDFSK_TILE_COUNT ? variant : 0];
  *tx = v.tx;
  *ty = v.ty;
}

#ifndef PDFSK_U
#define PDFSK_U 8
#endif
#ifndef PDFSK_U64
#define PDFSK_U64 8
#endif
// plane-queue depth of a mode. Measured at 512^3 on 8x8 tiles: depth 4 for
// the FP64 modes with one aux stream, with fewer registers, is no faster
// than 8 (1.63 against 1.58 ms per polish ordering: both at two waves per
// SIMD), and depth 2 (three waves per SIMD) is slower (1.80 ms). The speed
// polish (two aux streams) needs 256 VGPRs at depth 8, one wave per SIMD;
// at depth 4 it takes 201, two waves per SIMD: 2.15 -> 1.83 ms.
template <int M> constexpr int queue_depth() {
  return SMode<M>::F64 ? (SMode<M>::NAUX > 1 && PDFSK_U64 > 4 ? 4 : PDFSK_U64) : PDFSK_U;
}

// Whether a tile shape can run a mode: an FP64 value takes two edge-wave
// lanes (W (TX + TY) <= 64) and the LDS (ring, aux delay lines) must fit the
// CU's 160 KB. FP32 modes take every shape but 32 x 16 with an aux stream;
// FP64 modes the 16-wide shapes, and not 16 x 16 with two aux streams.
template <int TX, int TY, int M>
constexpr bool col_fits() {
  using S = SMode<M>;
  using G = ColGeom<TX, TY, S::W>;
  const long lds = (long)G::LDS_VALUES * (long)sizeof(typename S::T) +
                   4L * (S::NAUX ? S::NAUX * G::R * G::NT : 1) + 8;
  return S::W * (TX + TY) <= 64 && lds <= 160L * 1024;
}

// The tile shapes of kVariants, as template arguments: calls f with an
// object whose TX / TY members are the shape's, or returns dflt for a shape
// that cannot run mode M.
template <int X_, int Y_> struct TileShape {
  static constexpr int TX = X_, TY = Y_;
};
template <int M, typename R, typename F>
static R with_tile(const ColVariant &v, R dflt, F f) {
  auto go = [&](auto shape) -> R {
    using Sh = decltype(shape);
    if constexpr (col_fits<Sh::TX, Sh::TY, M>())
      return f(shape);
    else
      return dflt;
  };
  if (v.tx == 8 && v.ty == 8) return go(TileShape<8, 8>());
  if (v.tx == 16 && v.ty == 4) return go(TileShape<16, 4>());
  if (v.tx == 16 && v.ty == 8) return go(TileShape<16, 8>());
  if (v.tx == 16 && v.ty == 16) return go(TileShape<16, 16>());
  if (v.tx == 32 && v.ty == 8) return go(TileShape<32, 8>());
  if (v.tx == 32 && v.ty == 16) return go(TileShape<32, 16>());
  return dflt;
}

// f<M>() for a runtime mode
template <typename F>
static auto with_mode(int mode, F f) {
  switch (mode) {
    case PDFSK_MODE_F32_ISO: return f(std::integral_constant<int, PDFSK_MODE_F32_ISO>());
    case PDFSK_MODE_F32_ANISO: return f(std::integral_constant<int, PDFSK_MODE_F32_ANISO>());
    case PDFSK_MODE_F64_REF: return f(std::integral_constant<int, PDFSK_MODE_F64_REF>());
    case PDFSK_MODE_F64_POL_ISO: return f(std::integral_constant<int, PDFSK_MODE_F64_POL_ISO>());
    case PDFSK_MODE_F64_POL_ANISO:
      return f(std::integral_constant<int, PDFSK_MODE_F64_POL_ANISO>());
    case PDFSK_MODE_F32_SPEED: return f(std::integral_constant<int, PDFSK_MODE_F32_SPEED>());
    case PDFSK_MODE_FX_ISO: return f(std::integral_constant<int, PDFSK_MODE_FX_ISO>());
    default: return f(std::integral_constant<int, PDFSK_MODE_F64_POL_SPEED>());
  }
}

extern "C" int pdfsk_tile_supports(int variant, int mode) {
  if ((unsigned)variant >= PDFSK_TILE_COUNT || (unsigned)mode >= PDFSK_MODE_COUNT) return 0;
  return with_mode(mode, [&](auto m) {
    return with_tile<decltype(m)::value, int>(kVariants[variant], 0, [](auto) { return 1; });
  });
}

extern "C" int pdfsk_col_max_resident(int variant, int mode) {
  if (!pdfsk_tile_supports(variant, mode)) return 0;
  const ColVariant &v = kVariants[variant];
  return with_mode(mode, [&](auto m) {
    constexpr int M = decltype(m)::value;
    return with_tile<M, int>(v, 0, [&](auto shape) {
      using Sh = decltype(shape);
      int nb = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &nb, k_sweep_col<Sh::TX, Sh::TY, queue_depth<M>(), M>, Sh::TX * Sh::TY + 64,
              0) != hipSuccess)
        return 0;
      int dev = 0, ncu = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
      if (v.per_cu > 0 && v.per_cu < nb) nb = v.per_cu;
      return nb * ncu;
    });
  });
}

extern "C" int pdfsk_sweep_batch(const pdfsk_batch *b, int sn, int mode, const pdfs_sp32 *sp,
                                 const pdfs_sp64 *sp64, int variant, int grid, hipStream_t st) {
  if (sn < 1 || sn > 8 || grid < 1 || !pdfsk_tile_supports(variant, mode)) return -1;
  if ((PDFSK_MODE_IS_F64(mode) || mode == PDFSK_MODE_FX_ISO) && !b->maxdiff) return -1;
  static_assert(sizeof(FieldDesc) == sizeof(pdfsk_field), "field table layout");
  // Ordering sn as (+,+,+) in virtual coordinates: physical direction per
  // axis (mpi_openacc/src/phi3D_fsm.c:206-218 reflections x{4,8} y{2,6}
  // z{3,7}, descending levels for {2,5,7,8}).
  const int desc = (sn == 2 || sn == 5 || sn == 7 || sn == 8);
  const int rx = (sn == 4 || sn == 8), ry = (sn == 2 || sn == 6), rz = (sn == 3 || sn == 7);
  ColArgs a;
  memset(&a, 0, sizeof a);
  a.fields = (const FieldDesc *)b->fields;
  a.cols = b->cols;
  a.state = b->state;
  a.fx = desc ^ rx;
  a.fy = desc ^ ry;
  a.fz = desc ^ rz;
  a.edge_x = b->edge_x;
  a.edge_y = b->edge_y;
  a.edge_x_bytes = (uint32_t)b->edge_x_bytes;
  a.edge_y_bytes = (uint32_t)b->edge_y_bytes;
  a.changed = b->changed;
  a.streak = b->streak;
  a.maxdiff = (sn == 8 && b->maxdiff_last) ? b->maxdiff_last : b->maxdiff;
  a.nfields = b->nfields;
  a.err = b->err;
  if (sp) a.sp = *sp;
  if (sp64) a.sp64 = *sp64;
  with_mode(mode, [&](auto m) {
    constexpr int M = decltype(m)::value;
    return with_tile<M, int>(kVariants[variant], 0, [&](auto shape) {
      using Sh = decltype(shape);
      hipLaunchKernelGGL((k_sweep_col<Sh::TX, Sh::TY, queue_depth<M>(), M>), dim3(grid),
                         dim3(Sh::TX * Sh::TY + 64), 0, st, a);
      return 0;
    });
  });
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_sweep_batch_ov(const pdfsk_batch *b, int sn0, int sn1, const int *otab,
                                    int nt, int ncolp, unsigned *done, unsigned *cnt, unsigned *chg,
                                    size_t rx_bytes, size_t ry_bytes, int mode,
                                    const pdfs_sp32 *sp, const pdfs_sp64 *sp64, int variant,
                                    int grid, hipStream_t st) {
  if (sn0 < 1 || sn1 > 8 || sn1 < sn0 || grid < 1 || nt < 1 || !otab || !done || !cnt || !chg ||
      !pdfsk_tile_supports(variant, mode))
    return -1;
  if ((PDFSK_MODE_IS_F64(mode) || mode == PDFSK_MODE_FX_ISO) && !b->maxdiff) return -1;
  ColArgs a;
  memset(&a, 0, sizeof a);
  a.fields = (const FieldDesc *)b->fields;
  a.cols = otab;
  a.state = b->state;
  a.edge_x = b->edge_x;
  a.edge_y = b->edge_y;
  a.edge_x_bytes = (uint32_t)b->edge_x_bytes;
  a.edge_y_bytes = (uint32_t)b->edge_y_bytes;
  a.changed = b->changed;
  a.streak = b->streak;
  a.maxdiff = b->maxdiff;
  a.nfields = b->nfields;
  a.err = b->err;
  if (sp) a.sp = *sp;
  if (sp64) a.sp64 = *sp64;
  a.ov_n = sn1 - sn0 + 1;
  a.ov_nt = nt;
  a.ov_ncolp = ncolp;
  a.ov_rx = (uint32_t)rx_bytes;
  a.ov_ry = (uint32_t)ry_bytes;
  a.ov_done = done;
  a.ov_cnt = cnt;
  a.ov_chg = chg;
  a.ov_md_last = sn1 == 8 ? b->maxdiff_last : nullptr;
  for (int sn = sn0; sn <= sn1; sn++) {
    const int desc = (sn == 2 || sn == 5 || sn == 7 || sn == 8);
    const int rx = (sn == 4 || sn == 8), ry = (sn == 2 || sn == 6), rz = (sn == 3 || sn == 7);
    const int o = sn - sn0;
    a.ov_dirs |= ((desc ^ rx) | ((desc ^ ry) << 1) | ((desc ^ rz) << 2)) << (3 * o);
  }
  auto go = [&](auto m) {
    constexpr int M = decltype(m)::value;
    return with_tile<M, int>(kVariants[variant], 0, [&](auto shape) {
      using Sh = decltype(shape);
      hipLaunchKernelGGL((k_sweep_col<Sh::TX, Sh::TY, queue_depth<M>(), M, true>), dim3(grid),
                         dim3(Sh::TX * Sh::TY + 64), 0, st, a);
      return 0;
    });
  };
  with_mode(mode, [&](auto m) { return go(m); });
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_sweep_ref64(double *d, const double *slow, int X, int Y, int Z, int sn,
                                 double dx, double dy, double dz, unsigned long long *maxdiff,
                                 hipStream_t st) {
  if (sn < 1 || sn > 8) return -1;
  // mpi_openacc/src/phi3D_fsm.c:195-221 ordering table, :92-98 level bounds.
  const int xo = (sn == 4 || sn == 8) ? X + 1 : 0;
  const int yo = (sn == 2 || sn == 6) ? Y + 1 : 0;
  const int zo = (sn == 3 || sn == 7) ? Z + 1 : 0;
  const int incr = !(sn == 2 || sn == 5 || sn == 7 || sn == 8);
  const int first = 3, last = X + Y + Z;
  const int start = incr ? first : last, end = incr ? last + 1 : first - 1;
  for (int level = start; level != end; level += incr ? 1 : -1) {
    int xs = level - (Y + Z);
    xs = xs > 1 ? xs : 1;
    int ys = level - (X + Z);
    ys = ys > 1 ? ys : 1;
    int xe = level - (first - 1);
    xe = xe < X ? xe : X;
    int ye = level - (first - 1);
    ye = ye < Y ? ye : Y;
    const int xr = xe - xs + 1, yr = ye - ys + 1;
    if (xr <= 0 || yr <= 0) continue;
    dim3 bs(64, 4);
    dim3 gs((xr + 63) / 64, (yr + 3) / 4);
    hipLaunchKernelGGL(k_level_ref64, gs, bs, 0, st, d, slow, X, Y, Z, level, xs, ys, xe, ye,
                       xo, yo, zo, dx, dy, dz, maxdiff);
  }
  return hip_rc(hipGetLastError());
}

extern "C" unsigned pdfsk_judge_tiles(int X, int Y, int Z) { return judge_tiles(X, Y, Z); }

extern "C" int pdfsk_judge_batch_f32(const pdfsk_judge_field *tab, int nf, long ntiles, double dx,
                                     double dy, double dz, int speed, unsigned long long *counts,
                                     hipStream_t st) {
  if (nf < 1 || ntiles < 1 || ntiles >= (1L << 31) - 8) return -1;
  const long g = (ntiles + 7) / 8 * 8;  // a multiple of the 8 XCDs
  const bool iso = dx == dy && dy == dz;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)g), dim3(256), 0, st, tab, nf, dx, dy, dz, counts);
  };
  if (speed)
    iso ? go(k_judge_f32<true, true>) : go(k_judge_f32<false, true>);
  else
    iso ? go(k_judge_f32<true, false>) : go(k_judge_f32<false, false>);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_judge_batch_fx(const pdfsk_judge_field *tab, int nf, long ntiles, double h,
                                    unsigned long long *counts, hipStream_t st) {
  if (nf < 1 || ntiles < 1 || ntiles >= (1L << 31) - 8) return -1;
  const long g = (ntiles + 7) / 8 * 8;  // a multiple of the 8 XCDs
  hipLaunchKernelGGL((k_judge_f32<true, false, true>), dim3((unsigned)g), dim3(256), 0, st, tab,
                     nf, h, h, h, counts);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_pack_f64(const double *padded, float *dense, int X, int Y, int Z,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_pack_f64, dim3(grid_for((long)X * Y * Z, 256)), dim3(256), 0, st, padded,
                     dense, X, Y, Z);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_unpack_f64(const float *dense, double *padded, int X, int Y, int Z,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_unpack_f64, dim3(grid_for((long)X * Y * Z, 256)), dim3(256), 0, st,
                     dense, padded, X, Y, Z);
  return hip_rc(hipGetLastError());
}

// Every float bit pattern in [lo, hi): count where the sweep's square root
// (sqrt_cr_nonneg) differs from the FP64 square root rounded to FP32. That
// is the correctly rounded FP32 root (double carries more than 2 * 24 + 2
// bits, so rounding the double root again is innocuous for sqrt), computed
// by a different method (the FP64 square root sequence), so the check does
// not share the FP32 expansion's arithmetic.
__global__ __launch_bounds__(256) void k_check_sqrt_cr(uint32_t lo, uint32_t n,
                                                       unsigned long long *count) {
  unsigned long long bad = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float x = __uint_as_float(lo + i);
    bad += __float_as_uint(sqrt_cr_nonneg(x)) != __float_as_uint((float)__builtin_sqrt((double)x));
  }
  if (bad) atomicAdd(count, bad);
}

extern "C" int pdfsk_check_sqrt_cr(uint32_t lo, uint32_t hi, unsigned long long *count,
                                   hipStream_t st) {
  if (hi <= lo) return -1;
  hipLaunchKernelGGL(k_check_sqrt_cr, dim3(256 * 64), dim3(256), 0, st, lo, hi - lo, count);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_pack64(const double *padded, double *dense, int X, int Y, int Z,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_pack64, dim3(grid_for((long)X * Y * Z, 256)), dim3(256), 0, st, padded,
                     dense, X, Y, Z);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_unpack64(const double *dense, double *padded, int X, int Y, int Z,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_unpack64, dim3(grid_for((long)X * Y * Z, 256)), dim3(256), 0, st, dense,
                     padded, X, Y, Z);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_count_ghost_not_inf(const double *padded, int X, int Y, int Z,
                                         unsigned long long *count, hipStream_t st) {
  const long n = (long)(X + 2) * (Y + 2) * (Z + 2);
  hipLaunchKernelGGL(k_count_ghost_not_inf, dim3(grid_for(n, 256) < 4096 ? grid_for(n, 256) : 4096),
                     dim3(256), 0, st, padded, X, Y, Z, count);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_convert_batch(const pdfsk_conv *tab, int nf, long long nmax, int dir,
                                   hipStream_t st) {
  if (nf < 1 || nf > 65535 || nmax < 1 || (dir != 0 && dir != 1)) return -1;
  long long gx = (nmax + 255) / 256;
  if (gx > 2048) gx = 2048;
  hipLaunchKernelGGL(k_convert_batch, dim3((unsigned)gx, (unsigned)nf), dim3(256), 0, st, tab,
                     dir);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_fx_bits(int X, int Y, int Z) {
  const long long s = (long long)X + Y + Z;
  int c = 0;  // ceil(log2(s))
  while ((1ll << c) < s) c++;
  return 32 - c;
}

extern "C" float pdfsk_fx_key_value(uint32_t key) {
  const uint32_t b = (key & 0x80000000u) ? (key & 0x7FFFFFFFu) : ~key;
  float v;
  memcpy(&v, &b, sizeof v);
  return v;
}

// per_lane: points per lane the grid is sized for (the largest field's);
// capped: at most 2^16 workgroups over the batch. 2048 workgroups per field
// at most.
static int fx_launch(void (*kern)(const pdfsk_fx *), const pdfsk_fx *tab, int nf,
                     long long nmax, hipStream_t st, int per_lane = 1, bool capped = false) {
  if (nf < 1 || nf > 65535 || nmax < 1) return -1;
  long long gx = (nmax + 256LL * per_lane - 1) / (256LL * per_lane);
  if (capped && gx > 65536 / nf) gx = 65536 / nf;
  if (gx > 2048) gx = 2048;
  hipLaunchKernelGGL(kern, dim3((unsigned)gx, (unsigned)nf), dim3(256), 0, st, tab);
  return hip_rc(hipGetLastError());
}
extern "C" int pdfsk_fx_min_batch(const pdfsk_fx *tab, int nf, long long nmax, hipStream_t st) {
  // At least 16 points per lane and at most 2^16 workgroups: each workgroup
  // waits for its field's table entry and ends in a reduction and an atomic,
  // which on a batch of many small fields (512 fields of 32^3 .. 192^3) cost
  // more than the loads when a workgroup takes 256 points (1.8 ms against
  // 0.56 ms for 2.9 GB). The conversions, with no reduction, keep a
  // workgroup per 256 points: with the cap they ran 15% slower.
  return fx_launch(k_fx_min_batch, tab, nf, nmax, st, 16, true);
}
extern "C" int pdfsk_fx_encode_batch(const pdfsk_fx *tab, int nf, long long nmax,
                                     hipStream_t st) {
  return fx_launch(k_fx_encode_batch, tab, nf, nmax, st);
}
extern "C" int pdfsk_fx_decode_batch(const pdfsk_fx *tab, int nf, long long nmax,
                                     hipStream_t st) {
  return fx_launch(k_fx_decode_batch, tab, nf, nmax, st);
}

extern "C" int pdfsk_read_counters(const unsigned *changed, const unsigned *err,
                                   const unsigned *streak, unsigned long long *md, int nf,
                                   int cap_md, unsigned *h_changed, unsigned *h_err,
                                   unsigned *h_streak, unsigned long long *h_md, hipStream_t st) {
  if (nf < 1 || (md && cap_md < nf)) return -1;
  hipLaunchKernelGGL(k_read_counters, dim3(grid_for(nf, 256)), dim3(256), 0, st, changed, err,
                     streak, md, nf, cap_md, h_changed, h_err, h_streak, h_md);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for((long)n, 256)), dim3(256), 0, st, p, v, n);
  return hip_rc(hipGetLastError());
}

static int gen_ghost(double *padded, int X, int Y, int Z, hipStream_t st) {
  if (!padded) return 0;
  hipLaunchKernelGGL(k_gen_ghost64, dim3(grid_for((long)(X + 2) * (Y + 2) * (Z + 2), 256)),
                     dim3(256), 0, st, padded, X, Y, Z);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_gen_point(float *dense, double *padded, int X, int Y, int Z, int k, int j,
                               int i, hipStream_t st) {
  if (gen_ghost(padded, X, Y, Z, st)) return -1;
  hipLaunchKernelGGL(k_gen_point, dim3(grid_for((long)X * Y * Z, 256)), dim3(256), 0, st, dense,
                     padded, X, Y, Z, k, j, i);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_gen_terrain(float *dense, double *padded, int X, int Y, int Z, double h,
                                 const double *prm, int nh, hipStream_t st) {
  if (gen_ghost(padded, X, Y, Z, st)) return -1;
  hipLaunchKernelGGL(k_gen_terrain, dim3(grid_for((long)X * Y, 256)), dim3(256), 0, st, dense,
                     padded, X, Y, Z, h, prm, nh);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_gen_spheres(float *dense, double *padded, int X, int Y, int Z, double h,
                                 const double *prm, int ns, hipStream_t st) {
  if (gen_ghost(padded, X, Y, Z, st)) return -1;
  hipLaunchKernelGGL(k_gen_spheres, dim3(grid_for((long)X * Y * Z, 256)), dim3(256), 0, st,
                     dense, padded, X, Y, Z, h, prm, ns);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_update_distance(const int *loc, double *d, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_update_distance, dim3(grid_for(n, 256)), dim3(256), 0, st, loc, d, n);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_set_negative_inside(const int *loc, double *d, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_set_negative_inside, dim3(grid_for(n, 256)), dim3(256), 0, st, loc, d, n);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_adjust_boundary(double *d, int X, int Y, int Z, hipStream_t st) {
  hipLaunchKernelGGL(k_adjust_boundary, dim3(grid_for((long)(X + 2) * (Y + 2) * (Z + 2), 256)),
                     dim3(256), 0, st, d, X, Y, Z);
  return hip_rc(hipGetLastError());
}
