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
#include <stdint.h>

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

// Shift-invariant difference form u = m0 + t(m1 - m0, m2 - m0, spacing).
// Isotropic spacing: sorting network + constant multiplies, no division.
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
    T r = sqrt_t(two_h2 - a * a);
    u = m0 + (T)0.5 * (a + r);
    if (u > m2) {
      T b = m2 - m0;
      T D = three_h2 - (T)2 * ((a * a + b * b) - a * b);
      D = D > (T)0 ? D : (T)0;
      u = m0 + ((a + b) + sqrt_t(D)) * third;
    }
  }
  return cur < u ? cur : u;
}

__device__ __forceinline__ float god_diff_iso32(float cur, float mx, float my, float mz,
                                                float h, float two_h2, float three_h2,
                                                float third) {
#pragma clang fp contract(off)
  float lo = mx < my ? mx : my;
  float hi = mx > my ? mx : my;
  float m0 = lo < mz ? lo : mz;
  float mid = lo > mz ? lo : mz;
  float m1 = hi < mid ? hi : mid;
  float m2 = hi > mid ? hi : mid;
  float u = m0 + h;
  if (u > m1) {
    float a = m1 - m0;
    float r = __builtin_sqrtf(two_h2 - a * a);
    u = m0 + 0.5f * (a + r);
    if (u > m2) {
      float b = m2 - m0;
      float D = three_h2 - 2.0f * ((a * a + b * b) - a * b);
      D = D > 0.0f ? D : 0.0f;
      u = m0 + ((a + b) + __builtin_sqrtf(D)) * third;
    }
  }
  return cur < u ? cur : u;
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
      D = D > (T)0 ? D : (T)0;
      u = m[0] + (P + sqrt_t(D)) / W;
    }
  }
  return cur < u ? cur : u;
}

template <bool ISO>
__device__ __forceinline__ float god32(float cur, float mx, float my, float mz,
                                       const pdfs_sp32 &s) {
  if (ISO) return god_diff_iso32(cur, mx, my, mz, s.h, s.two_h2, s.three_h2, s.third);
  return god_diff_aniso<float>(cur, mx, my, mz, s.d, s.w, s.W);
}

// FP64 difference form (judge). Spacing constants in double, same op order
// as the oracle's double instantiation.
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

// ---------------------------------------------------------------------------
// Relaxed agent-scope 64-bit atomics: single-copy atomic, coherent across the
// XCDs, no L2 writeback or invalidate.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ld_edge(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_edge(uint64_t *p, float v, uint32_t epoch) {
  uint64_t e = ((uint64_t)epoch << 32) | (uint64_t)__float_as_uint(v);
  __hip_atomic_store(p, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait for an edge entry of the current epoch (2 s of the 100 MHz
// constant clock). On timeout the error flag is raised and the stale value
// used, so every wave still exits.
__device__ __noinline__ float wait_edge(const uint64_t *p, uint64_t e, uint32_t epoch,
                                        unsigned *err) {
  uint64_t t0 = __builtin_amdgcn_s_memrealtime();
  while ((uint32_t)(e >> 32) != epoch) {
    __builtin_amdgcn_s_sleep(1);
    e = ld_edge(p);
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    if (__builtin_amdgcn_s_memrealtime() - t0 > 200000000ull) {
      atomicOr(err, 1u);
      break;
    }
  }
  return __uint_as_float((uint32_t)e);
}

// ---------------------------------------------------------------------------
// Column-pipeline sweep, FP32, one ordering.
// ---------------------------------------------------------------------------
struct ColArgs {
  float *f;
  int X, Y, Z;
  int fx, fy, fz;  // physical = flip ? dim - v : v - 1 (v = virtual 1-based)
  int nTx, nTy;
  const int *order;
  unsigned *ticket;
  unsigned ticket_base;
  uint64_t *edge_x, *edge_y;
  uint32_t epoch;
  unsigned *changed;
  unsigned *err;
  pdfs_sp32 sp;
};

template <int TX, int TY, int U, bool ISO>
__global__ __launch_bounds__(TX *TY) void k_sweep_col(ColArgs a) {
  constexpr int NT = TX * TY;
  constexpr int R = TX + TY + 2;  // ring depth (z-planes)
  __shared__ float ring[R * NT];
  __shared__ int s_col;
  __shared__ unsigned s_changed;

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int X = a.X, Y = a.Y, Z = a.Z;
  const long XY = (long)X * Y;
  const int ncols = a.nTx * a.nTy;

  for (;;) {
    if (tid == 0) {
      s_col = (int)(atomicAdd(a.ticket, 1u) - a.ticket_base);
      s_changed = 0;
    }
    __syncthreads();
    const int col = s_col;
    if ((unsigned)col >= (unsigned)ncols) break;
    const int ca = a.order[2 * col], cb = a.order[2 * col + 1];
    const int vx = ca * TX + tx + 1, vy = cb * TY + ty + 1;
    const bool inx = vx <= X, iny = vy <= Y, active = inx && iny;
    const int px = a.fx ? X - vx : vx - 1;
    const int py = a.fy ? Y - vy : vy - 1;
    const long poff = (long)py * X + px;
    const int sxp = a.fx ? -1 : 1;             // physical step of virtual x+1
    const long syp = a.fy ? -(long)X : (long)X; // physical step of virtual y+1
    const bool east_ok = (tx == TX - 1) && iny && (vx + 1 <= X);
    const bool north_ok = (ty == TY - 1) && inx && (vy + 1 <= Y);
    const bool west_edge = (tx == 0) && active && ca > 0;
    const bool south_edge = (ty == 0) && active && cb > 0;
    const uint64_t *ew = a.edge_x + ((long)((ca - 1) * a.nTy + cb) * Z) * TY + ty;
    const uint64_t *es = a.edge_y + ((long)(ca * a.nTy + cb - 1) * Z) * TX + tx;
    uint64_t *mx_out = a.edge_x + ((long)(ca * a.nTy + cb) * Z) * TY + ty;
    uint64_t *my_out = a.edge_y + ((long)(ca * a.nTy + cb) * Z) * TX + tx;
    const bool pub_x = (tx == TX - 1) && active && ca + 1 < a.nTx;
    const bool pub_y = (ty == TY - 1) && active && cb + 1 < a.nTy;
    auto pz = [&](int vz) -> long { return (long)(a.fz ? Z - vz : vz - 1); };

    // Queues of in-flight loads, U steps deep (fully unrolled indices).
    float qP[U], qE[U], qN[U];
    uint64_t qW[U], qS[U];
    const int skew = tx + ty;

    auto load_plane = [&](int vz) -> float {
      return (active && vz >= 1 && vz <= Z) ? a.f[pz(vz) * XY + poff] : PDFS_INF;
    };
    auto issue = [&](int j, int l) {
      // plane needed at step l: z = l + 2; halos needed at step l: z = l + 1 - skew
      qP[j] = load_plane(l + 2);
      const int vz = l + 1 - skew;
      const bool zok = vz >= 1 && vz <= Z;
      qE[j] = (east_ok && zok) ? a.f[pz(vz) * XY + poff + sxp] : PDFS_INF;
      qN[j] = (north_ok && zok) ? a.f[pz(vz) * XY + poff + syp] : PDFS_INF;
      qW[j] = (west_edge && zok) ? ld_edge(ew + (long)(vz - 1) * TY) : 0ull;
      qS[j] = (south_edge && zok) ? ld_edge(es + (long)(vz - 1) * TX) : 0ull;
    };

    // Prologue: ghost plane 0 and plane 1 in the ring; queue steps 0..U-1.
    ring[0 * NT + tid] = PDFS_INF;
    ring[1 * NT + tid] = load_plane(1);
#pragma unroll
    for (int j = 0; j < U; j++) issue(j, j);

    unsigned my_changed = 0;
    const int NL = Z + TX + TY - 2;  // steps 0..NL (last step only stores)
    int slot_l2 = 2;                 // (l + 2) % R
    for (int l0 = 0; l0 <= NL; l0 += U) {
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int l = l0 + j;
        if (l <= NL) {
          // (1) plane l+2 into the ring, then refill this queue entry
          ring[slot_l2 * NT + tid] = qP[j];
          const float hE = qE[j], hN = qN[j];
          const uint64_t eW = qW[j], eS = qS[j];
          issue(j, l + U);
          __syncthreads();
          // (2) store the finished plane zs
          const int zs = l - (TX + TY - 2);
          if (zs >= 1 && active) {
            int s = zs % R;
            a.f[pz(zs) * XY + poff] = ring[s * NT + tid];
          }
          // (3) the hyperplane point of this thread
          const int vz = l + 1 - skew;
          if (active && vz >= 1 && vz <= Z) {
            const int sc = vz % R;
            const int sm = sc == 0 ? R - 1 : sc - 1;
            const int sp = sc == R - 1 ? 0 : sc + 1;
            const float c = ring[sc * NT + tid];
            const float zm = ring[sm * NT + tid];
            const float zp = ring[sp * NT + tid];
            float xm, xp, ym, yp;
            if (tx > 0) xm = ring[sc * NT + tid - 1];
            else xm = (ca > 0) ? wait_edge(ew + (long)(vz - 1) * TY, eW, a.epoch, a.err)
                               : PDFS_INF;
            xp = (tx < TX - 1) ? ring[sc * NT + tid + 1] : hE;
            if (ty > 0) ym = ring[sc * NT + tid - TX];
            else ym = (cb > 0) ? wait_edge(es + (long)(vz - 1) * TX, eS, a.epoch, a.err)
                               : PDFS_INF;
            yp = (ty < TY - 1) ? ring[sc * NT + tid + TX] : hN;
            const float mxv = xm < xp ? xm : xp;
            const float myv = ym < yp ? ym : yp;
            const float mzv = zm < zp ? zm : zp;
            const float u = god32<ISO>(c, mxv, myv, mzv, a.sp);
            my_changed += (u != c);
            ring[sc * NT + tid] = u;
            if (pub_x) st_edge(mx_out + (long)(vz - 1) * TY, u, a.epoch);
            if (pub_y) st_edge(my_out + (long)(vz - 1) * TX, u, a.epoch);
          }
          slot_l2 = slot_l2 == R - 1 ? 0 : slot_l2 + 1;
        }
      }
    }
    if (my_changed) atomicAdd(&s_changed, my_changed);
    __syncthreads();
    if (tid == 0 && s_changed) atomicAdd(a.changed, s_changed);
  }
}

// ---------------------------------------------------------------------------
// Reference hyperplane schedule, FP64, one launch per level
// (single_gpu_cuda/src/phi3D_fsm.cu:161-185 geometry, :199-233 body).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_level_ref64(double *d, int X, int Y, int Z,
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
      const double u = god_ref64(c, mx, my, mz, dx, dy, dz);
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
// One thread per point, 64 consecutive points per wave -> two flag words.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_judge_f32(const float *__restrict__ v,
                                                   const float *__restrict__ init, int X,
                                                   int Y, int Z, double dx, double dy,
                                                   double dz, uint32_t *flags,
                                                   unsigned long long *count) {
#pragma clang fp contract(off)
  const long N = (long)X * Y * Z;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (p < N) {
    const long XY = (long)X * Y;
    const int x = (int)(p % X);
    const long q = p / X;
    const int y = (int)(q % Y);
    const int z = (int)(q / Y);
    const float c = v[p];
    const double xm = x > 0 ? (double)v[p - 1] : __builtin_inf();
    const double xp = x < X - 1 ? (double)v[p + 1] : __builtin_inf();
    const double ym = y > 0 ? (double)v[p - X] : __builtin_inf();
    const double yp = y < Y - 1 ? (double)v[p + X] : __builtin_inf();
    const double zm = z > 0 ? (double)v[p - XY] : __builtin_inf();
    const double zp = z < Z - 1 ? (double)v[p + XY] : __builtin_inf();
    const double e = god_d64((double)init[p], xm < xp ? xm : xp, ym < yp ? ym : yp,
                             zm < zp ? zm : zp, dx, dy, dz);
    const double val = (double)c;
    const double hmin = dx < dy ? (dx < dz ? dx : dz) : (dy < dz ? dy : dz);
    if (__builtin_isinf(e) || __builtin_isinf(val)) bad = !(e == val);
    else bad = !(__builtin_fabs(val - e) <= __builtin_fabs(e) * 0x1.0p-22 + hmin * 0x1.0p-21);
  }
  const unsigned long long m = __ballot(bad);
  const int lane = threadIdx.x & 63;
  const long w0 = (((long)blockIdx.x * blockDim.x + threadIdx.x) - lane) >> 5;
  if (lane == 0) {
    if ((w0 << 5) < N) flags[w0] |= (uint32_t)m;
    if (((w0 + 1) << 5) < N) flags[w0 + 1] |= (uint32_t)(m >> 32);
    if (m) atomicAdd(count, (unsigned long long)__popcll(m));
  }
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
// C-ABI launchers
// ---------------------------------------------------------------------------
static inline int hip_rc(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }
static inline int grid_for(long n, int bs) {
  long g = (n + bs - 1) / bs;
  if (g > 65536L * 16) g = 65536L * 16;
  return (int)(g < 1 ? 1 : g);
}

extern "C" void pdfsk_tile_dims(int variant, int *tx, int *ty) {
  switch (variant) {
    case PDFSK_TILE_32x8: *tx = 32; *ty = 8; break;
    case PDFSK_TILE_16x8: *tx = 16; *ty = 8; break;
    default: *tx = 16; *ty = 16; break;
  }
}

#define PDFSK_U 4

template <int TX, int TY, bool ISO>
static int max_resident() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep_col<TX, TY, PDFSK_U, ISO>,
                                                   TX * TY, 0) != hipSuccess)
    return 0;
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  return nb * ncu;
}

extern "C" int pdfsk_col_max_resident(int variant, int iso) {
  switch (variant) {
    case PDFSK_TILE_32x8: return iso ? max_resident<32, 8, true>() : max_resident<32, 8, false>();
    case PDFSK_TILE_16x8: return iso ? max_resident<16, 8, true>() : max_resident<16, 8, false>();
    default: return iso ? max_resident<16, 16, true>() : max_resident<16, 16, false>();
  }
}

template <int TX, int TY>
static void launch_col(const ColArgs &a, int grid, hipStream_t st, int iso) {
  if (iso)
    hipLaunchKernelGGL((k_sweep_col<TX, TY, PDFSK_U, true>), dim3(grid), dim3(TX * TY), 0, st, a);
  else
    hipLaunchKernelGGL((k_sweep_col<TX, TY, PDFSK_U, false>), dim3(grid), dim3(TX * TY), 0, st, a);
}

extern "C" int pdfsk_sweep_col_f32(float *field, int X, int Y, int Z, int sn,
                                   const pdfs_sp32 *sp, const pdfsk_col_scratch *s,
                                   int variant, uint32_t epoch, unsigned ticket_base,
                                   int grid, hipStream_t st) {
  if (sn < 1 || sn > 8 || X < 1 || Y < 1 || Z < 1) return -1;
  int TX, TY;
  pdfsk_tile_dims(variant, &TX, &TY);
  // Ordering sn as (+,+,+) in virtual coordinates: physical direction per
  // axis (mpi_openacc/src/phi3D_fsm.c:206-218 reflections x{4,8} y{2,6}
  // z{3,7}, descending levels for {2,5,7,8}).
  const int desc = (sn == 2 || sn == 5 || sn == 7 || sn == 8);
  const int rx = (sn == 4 || sn == 8), ry = (sn == 2 || sn == 6), rz = (sn == 3 || sn == 7);
  ColArgs a;
  a.f = field;
  a.X = X;
  a.Y = Y;
  a.Z = Z;
  a.fx = desc ^ rx;
  a.fy = desc ^ ry;
  a.fz = desc ^ rz;
  a.nTx = (X + TX - 1) / TX;
  a.nTy = (Y + TY - 1) / TY;
  a.order = s->order;
  a.ticket = s->ticket;
  a.ticket_base = ticket_base;
  a.edge_x = s->edge_x;
  a.edge_y = s->edge_y;
  a.epoch = epoch;
  a.changed = s->changed;
  a.err = s->err;
  a.sp = *sp;
  switch (variant) {
    case PDFSK_TILE_32x8: launch_col<32, 8>(a, grid, st, sp->iso); break;
    case PDFSK_TILE_16x8: launch_col<16, 8>(a, grid, st, sp->iso); break;
    default: launch_col<16, 16>(a, grid, st, sp->iso); break;
  }
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_sweep_ref64(double *d, int X, int Y, int Z, int sn, double dx,
                                 double dy, double dz, unsigned long long *maxdiff,
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
    hipLaunchKernelGGL(k_level_ref64, gs, bs, 0, st, d, X, Y, Z, level, xs, ys, xe, ye, xo,
                       yo, zo, dx, dy, dz, maxdiff);
  }
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsk_judge_f32(const float *v, const float *init, int X, int Y, int Z,
                               double dx, double dy, double dz, uint32_t *flags,
                               unsigned long long *count, hipStream_t st) {
  const long N = (long)X * Y * Z;
  const int bs = 256;
  const long g = (N + bs - 1) / bs;
  hipLaunchKernelGGL(k_judge_f32, dim3((unsigned)g), dim3(bs), 0, st, v, init, X, Y, Z, dx, dy,
                     dz, flags, count);
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
