// gfx950 kernels of the learned warm start: a 3-D conv encoder-decoder in
// BF16 with FP32 accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Activations: channels-last [n][z][y][x][C] BF16, C a multiple of 16.
// Weights: [27][Cout][Cin] BF16 (tap = (dz*3 + dy)*3 + dx, cin fastest),
// an FP32 master copy lives with the optimizer.
//
// MFMA 16x16x32 operand layout (wave64): lane l = c + 16 g (c = l % 16,
// g = l / 16). A (16 x 32, M x K): lane holds A[c][8g .. 8g+7]. B (K x N):
// lane holds B[8g .. 8g+7][c]. D (M x N, FP32): lane holds D[4g .. 4g+3][c].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "pdfs_nn.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));  // the same 16 bytes, as raw bits
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                  __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

static inline int hip_rc(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }
static inline unsigned grid1(long n, int bs) {
  long g = (n + bs - 1) / bs;
  if (g > 1048576) g = 1048576;
  return (unsigned)(g < 1 ? 1 : g);
}

// ---------------------------------------------------------------------------
// 3x3x3 convolution, stride 1, zero padding 1, as an implicit GEMM:
// M = positions, N = Cout, K = 27 taps x Cin, on LDS-staged input tiles.
// Block (4 waves): an 8 x 8 x TZ tile of output positions of one sample, as
// 4 TZ groups of 16 positions (8 along x by 2 along y); wave w owns groups
// w TZ .. w TZ + TZ - 1, each against all Cout (NT tiles of 16), so every
// weight fragment feeds TZ MFMAs. TZ = 4, or 1 on the coarse levels (more
// blocks). The input tile with a
// one-point halo (10 x 10 x (TZ + 2) positions) is staged in LDS one channel
// chunk at a time:
// * K32 (Cin a multiple of 32): chunks of 32 channels; one MFMA step is one
//   tap x 32 channels, lane group g taking channels 8g .. 8g+7;
// * otherwise chunks of 16 channels and taps in pairs: lane groups 0, 1 take
//   tap 2p (channels 8g ..), groups 2, 3 tap 2p + 1 (tap 27 is zero), so the
//   32-deep K of one MFMA spans two taps.
// The K steps are unrolled with two register sets: the A fragments (LDS)
// and weight fragments (L2) of step s + 1 load while step s's MFMAs run.
// out = act(conv + bias); act: 0 linear, 1 ReLU. out (BF16) and out_f32
// each optional. With flip = 1 the taps are mirrored (dz,dy,dx) ->
// (2-dz, 2-dy, 2-dx): with transposed weights this is the input gradient.
// ---------------------------------------------------------------------------
constexpr int CT_HX = 10, CT_HY = 10;
// LDS tile of a (TZ, K32) kernel, in 16-byte vectors: one plane of positions
// per 8-channel vector, plane pitch HP + 2 vectors ((HP + 2) * 16 B is 32 B
// mod 128 B for HP = 100 (TZ + 2), so consecutive planes start 8 banks
// apart); an A-fragment read takes consecutive vectors of one plane.
template <int TZ> struct ConvTile {
  static constexpr int HZ = TZ + 2, HP = CT_HX * CT_HY * HZ, HPS = HP + 2;
};
template <int TZ, bool K32> constexpr int conv_tile_vecs() { return ConvTile<TZ>::HPS * (K32 ? 4 : 2); }

__device__ __forceinline__ int tap_shift(int tap, int flip) {
  const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
  return flip ? ((2 - dz) * CT_HY + (2 - dy)) * CT_HX + (2 - dx)
              : (dz * CT_HY + dy) * CT_HX + dx;
}

// One output tile (bx, by, bz) of sample b; tile = the block's LDS staging
// area.
template <int NT, bool K32, int TZ>
__device__ __forceinline__ void conv3_tile_body(const uint16_t *__restrict__ in, int Cin,
                                                const uint16_t *__restrict__ W,
                                                const float *__restrict__ bias, int Cout,
                                                uint16_t *__restrict__ out,
                                                float *__restrict__ out_f32, long b, int Dz,
                                                int Dy, int Dx, int bx, int by, int bz, int relu,
                                                int flip, s16x8 *tile) {
  constexpr int HX = CT_HX, HY = CT_HY, HP = ConvTile<TZ>::HP, HPS = ConvTile<TZ>::HPS;
  constexpr int CQ = K32 ? 4 : 2;       // 8-channel vectors per position and chunk
  constexpr int CK = 8 * CQ;            // channels per chunk
  constexpr int NSTEP = K32 ? 27 : 14;  // MFMA K steps per chunk
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = lane & 15, g = lane >> 4;
  const int x0 = bx * 8, y0 = by * 8, z0 = bz * TZ;
  f32x4 acc[TZ][NT];
#pragma unroll
  for (int i = 0; i < TZ; i++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[i][t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  // this lane's A row in each of its groups q: x = c & 7, y = 2 (q & 3) + (c >> 3), z = q >> 2
  int arow[TZ];
#pragma unroll
  for (int i = 0; i < TZ; i++) {
    const int q = TZ * wave + i;
    arow[i] = ((q >> 2) * HY + 2 * (q & 3) + (c >> 3)) * HX + (c & 7);
  }
  const s16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
  // K step s of chunk k for this lane: its tap, and its 8-channel vector of the chunk
  auto step_tap = [&](int s) -> int { return K32 ? s : 2 * s + (g >> 1); };
  const int vq = K32 ? g : (g & 1);
  // weight fragments of step s (tap 27, past the last pair, is zero: its
  // load reads tap 26 and is discarded)
  const uint16_t *const wl = W + (long)c * Cin + 8 * vq;
  auto load_w = [&](int k, int s, s16x8 *bw) __attribute__((always_inline)) {
    const int tap = step_tap(s);
    const int tc = tap < 27 ? tap : 26;
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const s16x8 v = *(const s16x8 *)(wl + ((long)tc * Cout + 16 * t) * Cin + CK * k);
      bw[t] = tap < 27 ? v : zero8;
    }
  };
  auto load_a = [&](int s, s16x8 *a) __attribute__((always_inline)) {
    const int tap = step_tap(s);
    const int sh = tap_shift(tap < 27 ? tap : 26, flip);
#pragma unroll
    for (int i = 0; i < TZ; i++) a[i] = tile[vq * HPS + arow[i] + sh];
  };
  const int nchunk = Cin / CK;
  for (int k = 0; k < nchunk; k++) {
    s16x8 a[2][TZ], w[2][NT];
    load_w(k, 0, w[0]);
    // stage the chunk in batches of up to SB vectors per thread: a batch's
    // loads all issue before its LDS writes (the first batch's before the
    // barrier that ends the previous chunk's reads)
    constexpr int NV = (HP * CQ + 255) / 256, SB = 6;
    auto stage_load = [&](int j) -> s16x8 {
      const int i = (int)threadIdx.x + 256 * j;
      const int v = i % CQ, p = i / CQ;
      const int hx = p % HX, hy = (p / HX) % HY, hz = p / (HX * HY);
      const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
      s16x8 r = zero8;
      if (i < HP * CQ && (unsigned)gx < (unsigned)Dx && (unsigned)gy < (unsigned)Dy &&
          (unsigned)gz < (unsigned)Dz)
        r = *(const s16x8 *)(in + (((b * Dz + gz) * Dy + gy) * (long)Dx + gx) * Cin + CK * k +
                             8 * v);
      return r;
    };
    auto stage_store = [&](int j, s16x8 v) __attribute__((always_inline)) {
      const int i = (int)threadIdx.x + 256 * j;
      if (i < HP * CQ) tile[(i % CQ) * HPS + i / CQ] = v;
    };
#pragma unroll
    for (int j0 = 0; j0 < NV; j0 += SB) {
      s16x8 val[SB];
#pragma unroll
      for (int j = 0; j < SB; j++)
        if (j0 + j < NV) val[j] = stage_load(j0 + j);
      if (j0 == 0) __syncthreads();
#pragma unroll
      for (int j = 0; j < SB; j++)
        if (j0 + j < NV) stage_store(j0 + j, val[j]);
    }
    __syncthreads();
    load_a(0, a[0]);
#pragma unroll
    for (int s = 0; s < NSTEP; s++) {
      const int cur = s & 1;
      if (s + 1 < NSTEP) {
        load_w(k, s + 1, w[cur ^ 1]);
        load_a(s + 1, a[cur ^ 1]);
      }
#pragma unroll
      for (int i = 0; i < TZ; i++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[i][t] = mfma32(a[cur][i], w[cur][t], acc[i][t]);
    }
  }
  // D: lane (c, g) holds rows 4g .. 4g+3 (positions of the group) of column c
#pragma unroll
  for (int i = 0; i < TZ; i++) {
    const int q = TZ * wave + i;
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int co = 16 * t + c;
      const float bv = bias ? bias[co] : 0.f;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int r = 4 * g + e;
        const int x = x0 + (r & 7), y = y0 + 2 * (q & 3) + (r >> 3), z = z0 + (q >> 2);
        if (x >= Dx || y >= Dy || z >= Dz) continue;
        const long pos = ((b * Dz + z) * Dy + y) * (long)Dx + x;
        float v = acc[i][t][e] + bv;
        if (relu) v = v > 0.f ? v : 0.f;
        if (out) out[pos * Cout + co] = f2bf(v);
        if (out_f32) out_f32[pos * Cout + co] = v;
      }
    }
  }
}


template <int NT, bool K32, int TZ>
__global__ __launch_bounds__(256) void k_conv3_tile(const uint16_t *__restrict__ in, int Cin,
                                                    const uint16_t *__restrict__ W,
                                                    const float *__restrict__ bias, int Cout,
                                                    uint16_t *__restrict__ out,
                                                    float *__restrict__ out_f32, int n, int Dz,
                                                    int Dy, int Dx, int relu, int flip) {
  __shared__ s16x8 tile[conv_tile_vecs<TZ, K32>()];
  const int ntx = (Dx + 7) >> 3, nty = (Dy + 7) >> 3, ntz = (Dz + TZ - 1) / TZ;
  int bid = blockIdx.x;
  const int bx = bid % ntx;
  bid /= ntx;
  const int by = bid % nty;
  bid /= nty;
  const int bz = bid % ntz;
  const long b = bid / ntz;
  conv3_tile_body<NT, K32, TZ>(in, Cin, W, bias, Cout, out, out_f32, b, Dz, Dy, Dx, bx, by, bz,
                               relu, flip, tile);
}

// Uneven batch: segment f owns blocks [seg[f].blk, seg[f+1].blk) and the
// packed positions from seg[f].pos (seg[nf] is the sentinel).
__device__ __forceinline__ int seg_by_blk(const pdfsnn_seg *seg, int nf, int v) {
  int lo = 0, hi = nf;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (seg[mid].blk <= v) lo = mid; else hi = mid;
  }
  return lo;
}

// The batch's blocks are 8 x 8 x 4 tiles (its block tables count them so).
template <int NT, bool K32>
__global__ __launch_bounds__(256) void k_conv3_tile_rag(const uint16_t *__restrict__ in, int Cin,
                                                        const uint16_t *__restrict__ W,
                                                        const float *__restrict__ bias, int Cout,
                                                        uint16_t *__restrict__ out,
                                                        float *__restrict__ out_f32,
                                                        const pdfsnn_seg *__restrict__ seg,
                                                        int nf, int relu) {
  __shared__ s16x8 tile[conv_tile_vecs<4, K32>()];
  const int f = seg_by_blk(seg, nf, (int)blockIdx.x);
  const pdfsnn_seg sg = seg[f];
  const int ntx = (sg.Dx + 7) >> 3, nty = (sg.Dy + 7) >> 3;
  int bid = (int)blockIdx.x - sg.blk;
  const int bx = bid % ntx;
  bid /= ntx;
  const int by = bid % nty;
  const int bz = bid / nty;
  const long p0 = sg.pos;
  conv3_tile_body<NT, K32, 4>(in + p0 * Cin, Cin, W, bias, Cout, out ? out + p0 * Cout : nullptr,
                              out_f32 ? out_f32 + p0 * Cout : nullptr, 0, sg.Dz, sg.Dy, sg.Dx,
                              bx, by, bz, relu, 0, tile);
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[tap][co][ci] += sum_p dY[p][co] X[p + s(tap)][ci], and
// (db non-null) the bias gradient db[co] += sum_p dY[p][co], summed by the
// centre tap's blocks from their staged dY.
// Block: one tap, a run of positions; 64-position chunks staged in LDS
// transposed (channel-major: [co][pos], [ci][pos]), so an operand fragment of
// 8 consecutive positions is one 16-byte read; every wave accumulates its
// (co, ci) tiles with MFMA (M = co, N = ci, K = positions, two 32-position
// steps per chunk), and adds them to the FP32 dW at the end.
// ---------------------------------------------------------------------------
constexpr int WG_PP = 64 + 8;  // position pitch of the transposed chunks (bf16), 16-byte aligned

__global__ __launch_bounds__(256) void k_conv3_wgrad(const uint16_t *__restrict__ X, int Cin,
                                                     const uint16_t *__restrict__ dY, int Cout,
                                                     float *__restrict__ dW,
                                                     float *__restrict__ db, int n, int Dz,
                                                     int Dy, int Dx, int chunks_per_block) {
  __shared__ __attribute__((aligned(16))) uint16_t sY[64 * WG_PP];   // [co][pos]
  __shared__ __attribute__((aligned(16))) uint16_t sX[128 * WG_PP];  // [ci][pos]
  const int tap = blockIdx.y;
  const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = lane & 15, g = lane >> 4;
  const long M = (long)n * Dz * Dy * Dx;
  const int nco = Cout / 16, nci = Cin / 16, ntile = nco * nci;
  // up to 32 tiles (Cout <= 64, Cin <= 128) over 4 waves: 8 per wave
  f32x4 acc[8];
#pragma unroll
  for (int i = 0; i < 8; i++) acc[i] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const bool bsum = db && tap == 13 && (int)threadIdx.x < Cout;
  float bacc = 0.f;
  for (int ch = 0; ch < chunks_per_block; ch++) {
    const long pbase = ((long)blockIdx.x * chunks_per_block + ch) * 64;
    if (pbase >= M) break;
    __syncthreads();
    // stage dY[pbase .. +64][0..Cout) and X[shifted][0..Cin), transposed:
    // a thread loads 8 channels (16 bytes) of one position and writes them
    // to 8 channel rows
    const int nvy = Cout / 8, nvx = Cin / 8;
    for (int i = threadIdx.x; i < 64 * nvy; i += 256) {
      const int r = i / nvy, c8 = 8 * (i % nvy);
      const long p = pbase + r;
      s16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
      if (p < M) v = *(const s16x8 *)(dY + p * Cout + c8);
#pragma unroll
      for (int j = 0; j < 8; j++) sY[(c8 + j) * WG_PP + r] = (uint16_t)v[j];
    }
    for (int i = threadIdx.x; i < 64 * nvx; i += 256) {
      const int r = i / nvx, c8 = 8 * (i % nvx);
      const long p = pbase + r;
      s16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
      if (p < M) {
        const int x = (int)(p % Dx);
        long q = p / Dx;
        const int y = (int)(q % Dy);
        q /= Dy;
        const int z = (int)(q % Dz);
        const long b = q / Dz;
        const int zz = z + dz - 1, yy = y + dy - 1, xx = x + dx - 1;
        if ((unsigned)zz < (unsigned)Dz && (unsigned)yy < (unsigned)Dy &&
            (unsigned)xx < (unsigned)Dx)
          v = *(const s16x8 *)(X + (((b * Dz + zz) * Dy + yy) * Dx + xx) * (long)Cin + c8);
      }
#pragma unroll
      for (int j = 0; j < 8; j++) sX[(c8 + j) * WG_PP + r] = (uint16_t)v[j];
    }
    __syncthreads();
    if (bsum)
      for (int r = 0; r < 64; r++) bacc += bf2f(sY[threadIdx.x * WG_PP + r]);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int tl = wave + 4 * i;
      if (tl >= ntile) break;
      const int tco = tl / nci, tci = tl % nci;
#pragma unroll
      for (int kk = 0; kk < 2; kk++) {
        const s16x8 a = *(const s16x8 *)(sY + (16 * tco + c) * WG_PP + 32 * kk + 8 * g);
        const s16x8 b = *(const s16x8 *)(sX + (16 * tci + c) * WG_PP + 32 * kk + 8 * g);
        acc[i] = mfma32(a, b, acc[i]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int tl = wave + 4 * i;
    if (tl >= ntile) break;
    const int tco = tl / nci, tci = tl % nci;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int co = 16 * tco + 4 * g + e, ci = 16 * tci + c;
      atomicAdd(&dW[((long)tap * Cout + co) * Cin + ci], acc[i][e]);
    }
  }
  if (bsum) atomicAdd(&db[threadIdx.x], bacc);
}

// ---------------------------------------------------------------------------
// Pooling, upsampling, concat, ReLU backward, loss, optimizer, conversions.
// ---------------------------------------------------------------------------
// 2x2x2 average pool: out[b][z][y][x][c] = mean of in[b][2z..][2y..][2x..][c]
__global__ void k_pool2(const uint16_t *in, uint16_t *out, int n, int Dz, int Dy, int Dx,
                        int C) {
  const int oz = Dz / 2, oy = Dy / 2, ox = Dx / 2;
  const long N = (long)n * oz * oy * ox * C;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long q = i / C;
    const int x = (int)(q % ox);
    q /= ox;
    const int y = (int)(q % oy);
    q /= oy;
    const int z = (int)(q % oz);
    const long b = q / oz;
    float s = 0.f;
    for (int k = 0; k < 8; k++) {
      const int zz = 2 * z + (k >> 2), yy = 2 * y + ((k >> 1) & 1), xx = 2 * x + (k & 1);
      s += bf2f(in[(((b * Dz + zz) * Dy + yy) * Dx + xx) * (long)C + c]);
    }
    out[i] = f2bf(0.125f * s);
  }
}

// gradient of the average pool: din = dout / 8 broadcast (din is fine-level)
__global__ void k_pool2_bwd(const float *dout, float *din, int n, int Dz, int Dy, int Dx,
                            int C, int accumulate) {
  const long N = (long)n * Dz * Dy * Dx * C;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long q = i / C;
    const int x = (int)(q % Dx);
    q /= Dx;
    const int y = (int)(q % Dy);
    q /= Dy;
    const int z = (int)(q % Dz);
    const long b = q / Dz;
    const float v = 0.125f * dout[((((b * (Dz / 2) + z / 2) * (Dy / 2) + y / 2) * (Dx / 2)) +
                                   x / 2) * (long)C + c];
    din[i] = accumulate ? din[i] + v : v;
  }
}

// nearest x2 upsample of src (coarse, Cs channels) into channels [0, Cs) of
// dst (fine, Cd channels), and copy skip (fine, Ck channels) into [Cs, Cs+Ck)
// One thread per 8 channels (16 bytes) of an output position; Cs and Ck are
// multiples of 16.
__global__ void k_up2_concat(const uint16_t *src, int Cs, const uint16_t *skip, int Ck,
                             uint16_t *dst, int n, int Dz, int Dy, int Dx) {
  const int Cv = (Cs + Ck) / 8;
  const long N = (long)n * Dz * Dy * Dx * Cv;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x) {
    const int c = 8 * (int)(i % Cv);
    const long p = i / Cv;
    s16x8 v;
    if (c < Cs) {
      const int x = (int)(p % Dx);
      long q = p / Dx;
      const int y = (int)(q % Dy);
      q /= Dy;
      const int z = (int)(q % Dz);
      const long b = q / Dz;
      v = *(const s16x8 *)(src + ((((b * (Dz / 2) + z / 2) * (Dy / 2) + y / 2) * (Dx / 2)) + x / 2) *
                                     (long)Cs + c);
    } else {
      v = *(const s16x8 *)(skip + p * Ck + (c - Cs));
    }
    *(s16x8 *)(dst + 8 * i) = v;
  }
}

// backward of k_up2_concat: dsrc (coarse) = sum of the 8 children's first Cs
// channels; dskip (fine) = channels [Cs, Cs+Ck)
__global__ void k_up2_concat_bwd(const float *ddst, int Cs, int Ck, float *dsrc,
                                 float *dskip, int n, int Dz, int Dy, int Dx) {
  const int Cd = Cs + Ck;
  const int oz = Dz / 2, oy = Dy / 2, ox = Dx / 2;
  const long Nc = (long)n * oz * oy * ox * Cs;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < Nc;
       i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % Cs);
    long q = i / Cs;
    const int x = (int)(q % ox);
    q /= ox;
    const int y = (int)(q % oy);
    q /= oy;
    const int z = (int)(q % oz);
    const long b = q / oz;
    float s = 0.f;
    for (int k = 0; k < 8; k++) {
      const int zz = 2 * z + (k >> 2), yy = 2 * y + ((k >> 1) & 1), xx = 2 * x + (k & 1);
      s += ddst[(((b * Dz + zz) * Dy + yy) * Dx + xx) * (long)Cd + c];
    }
    dsrc[i] = s;
  }
  const long Nf = (long)n * Dz * Dy * Dx * Ck;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < Nf;
       i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % Ck);
    const long p = i / Ck;
    dskip[i] = ddst[p * Cd + Cs + c];
  }
}

// dy_pre (BF16) = g (FP32) masked by the ReLU output y (y > 0), or unmasked
__global__ void k_relu_bwd(const float *g, const uint16_t *y, uint16_t *out, long N) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x)
    out[i] = (y == nullptr || bf2f(y[i]) > 0.f) ? f2bf(g[i]) : (uint16_t)0;
}

// head output channel 0 (FP32, C channels) vs target: loss += sum (o-t)^2 / N,
// grad (BF16, C channels, zero except channel 0) = 2 (o - t) / N
__global__ void k_mse(const float *o, int C, const float *t, long N, uint16_t *grad,
                      float *loss) {
  __shared__ float part[256];
  float s = 0.f;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x) {
    const float d = o[i * C] - t[i];
    s += d * d;
    for (int c = 0; c < C; c++) grad[i * C + c] = c == 0 ? f2bf(2.f * d / (float)N) : (uint16_t)0;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) part[threadIdx.x] += part[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, part[0] / (float)N);
}

// Adam (Kingma & Ba 2015) on FP32 masters; refresh the BF16 weights [27][Co][Ci]
// and the flipped-transposed copy [27][Ci][Co] used by the input gradient.
__global__ void k_adam(float *w, const float *gw, float *m, float *v, long N, float lr, float b1,
                       float b2, float eps, float c1, float c2, uint16_t *wb, uint16_t *wt,
                       int Co, int Ci) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x) {
    const float gi = gw[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) {
      const uint16_t h = f2bf(wi);
      wb[i] = h;
      const int ci = (int)(i % Ci);
      const long q = i / Ci;
      const int co = (int)(q % Co);
      const int tap = (int)(q / Co);
      wt[((long)tap * Ci + ci) * Co + co] = h;
    }
  }
}

__global__ void k_f32_to_bf16(const float *x, uint16_t *y, long N) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

__global__ void k_transpose_w(const uint16_t *wb, uint16_t *wt, int Co, int Ci) {
  const long N = 27L * Co * Ci;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x) {
    const int ci = (int)(i % Ci);
    const long q = i / Ci;
    const int co = (int)(q % Co);
    const int tap = (int)(q / Co);
    wt[((long)tap * Ci + ci) * Co + co] = wb[i];
  }
}

// ---------------------------------------------------------------------------
// Warm-start input and output.
// ---------------------------------------------------------------------------
// Source mask at 1/4 resolution: fraction of known cells (init < far) in
// each 4x4x4 block of the dense field [Z][Y][X]; channel 0 of a 16-channel
// BF16 tensor [Dz][Dy][Dx][16]. The coarse grid may extend past the field
// (rounded up for pooling): blocks outside it are empty.
__global__ void k_mask_down4(const float *f, int X, int Y, int Z, uint16_t *m, int Dz, int Dy,
                             int Dx, float far) {
  const long N = (long)Dz * Dy * Dx;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i % Dx);
    const long q = i / Dx;
    const int y = (int)(q % Dy), z = (int)(q / Dy);
    int known = 0, cnt = 0;
    for (int k = 0; k < 64; k++) {
      const int zz = 4 * z + (k >> 4), yy = 4 * y + ((k >> 2) & 3), xx = 4 * x + (k & 3);
      if (zz < Z && yy < Y && xx < X) {
        cnt++;
        known += f[((long)zz * Y + yy) * X + xx] < far;
      }
    }
    for (int c = 0; c < 16; c++) m[i * 16 + c] = c == 0 ? f2bf(cnt ? (float)known / 64.f : 0.f)
                                                         : (uint16_t)0;
  }
}

// Training target at 1/4 resolution: the converged field at block corners
// (4z, 4y, 4x), scaled.
__global__ void k_target_down4(const float *u, int X, int Y, int Z, float *t, int Dz, int Dy,
                               int Dx, float scale) {
  const long N = (long)Dz * Dy * Dx;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i % Dx);
    const long q = i / Dx;
    const int y = (int)(q % Dy), z = (int)(q / Dy);
    const int zz = min(4 * z, Z - 1), yy = min(4 * y, Y - 1), xx = min(4 * x, X - 1);
    t[i] = scale * u[((long)zz * Y + yy) * X + xx];
  }
}

// Warm start: trilinear interpolation of the coarse prediction (corner
// samples at 4i, FP32 channel 0 of a C-channel tensor) back to the dense
// field, inflated to an upper bound w = a * d + b, capped by the init.
__global__ void k_warm_up4(const float *pred, int C, int Dz, int Dy, int Dx, float inv_scale,
                           float a, float b, float cut, const float *init, float *w, int X, int Y,
                           int Z) {
#pragma clang fp contract(off)
  const long N = (long)X * Y * Z;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i % X);
    const long q = i / X;
    const int y = (int)(q % Y), z = (int)(q / Y);
    const float fx = 0.25f * x, fy = 0.25f * y, fz = 0.25f * z;
    const int x0 = min((int)fx, Dx - 1), y0 = min((int)fy, Dy - 1), z0 = min((int)fz, Dz - 1);
    const int x1 = min(x0 + 1, Dx - 1), y1 = min(y0 + 1, Dy - 1), z1 = min(z0 + 1, Dz - 1);
    const float tx = fx - x0, ty = fy - y0, tz = fz - z0;
    auto P = [&](int zz, int yy, int xx) { return pred[(((long)zz * Dy + yy) * Dx + xx) * C]; };
    const float c00 = P(z0, y0, x0) * (1 - tx) + P(z0, y0, x1) * tx;
    const float c01 = P(z0, y1, x0) * (1 - tx) + P(z0, y1, x1) * tx;
    const float c10 = P(z1, y0, x0) * (1 - tx) + P(z1, y0, x1) * tx;
    const float c11 = P(z1, y1, x0) * (1 - tx) + P(z1, y1, x1) * tx;
    const float c0 = c00 * (1 - ty) + c01 * ty, c1 = c10 * (1 - ty) + c11 * ty;
    const float d = (c0 * (1 - tz) + c1 * tz) * inv_scale;
    const float f0 = init[i];
    // beyond the cutoff the prediction has not seen a source: keep the init
    const float up = d >= cut ? f0 : a * (d > 0.f ? d : 0.f) + b;
    w[i] = f0 < up ? f0 : up;
  }
}

// ---------------------------------------------------------------------------
// Uneven batches: the same per-field arithmetic as the kernels above, with
// each element's field found from the segment tables.
// ---------------------------------------------------------------------------
// Element-wise kernels of an uneven batch: segment f owns the blocks
// [seg[f].eblk, seg[f+1].eblk), RAG_EP positions of the level each (the
// warm start's dense end: fld[f].blk, RAG_EF points each), so a block finds
// its field once and works on one field only.
__device__ __forceinline__ int seg_by_eblk(const pdfsnn_seg *seg, int nf, int v) {
  int lo = 0, hi = nf;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (seg[mid].eblk <= v) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_pool2_rag(const uint16_t *in, uint16_t *out,
                                                   const pdfsnn_seg *si, const pdfsnn_seg *so,
                                                   int nf, int C) {
  const int f = seg_by_eblk(so, nf, (int)blockIdx.x);
  const pdfsnn_seg a = si[f], o = so[f];
  const int p0 = ((int)blockIdx.x - o.eblk) * RAG_EP;
  const int np = min(RAG_EP, o.Dz * o.Dy * o.Dx - p0);
  for (int e = threadIdx.x; e < np * C; e += 256) {
    const int c = e % C;
    int q = p0 + e / C;
    const int x = q % o.Dx;
    q /= o.Dx;
    const int y = q % o.Dy, z = q / o.Dy;
    float s = 0.f;
    for (int k = 0; k < 8; k++) {
      const int zz = 2 * z + (k >> 2), yy = 2 * y + ((k >> 1) & 1), xx = 2 * x + (k & 1);
      s += bf2f(in[((long)a.pos + (zz * a.Dy + yy) * a.Dx + xx) * C + c]);
    }
    out[((long)o.pos + p0) * C + e] = f2bf(0.125f * s);
  }
}

__global__ __launch_bounds__(256) void k_up2_concat_rag(const uint16_t *src, int Cs,
                                                        const uint16_t *skip, int Ck,
                                                        uint16_t *dst, const pdfsnn_seg *sf,
                                                        const pdfsnn_seg *sc, int nf) {
  const int Cd = Cs + Ck, Cv = Cd / 8;
  const int f = seg_by_eblk(sf, nf, (int)blockIdx.x);
  const pdfsnn_seg a = sf[f];
  const long cpos = sc[f].pos;
  const int p0 = ((int)blockIdx.x - a.eblk) * RAG_EP;
  const int np = min(RAG_EP, a.Dz * a.Dy * a.Dx - p0);
  const long g0 = (long)a.pos + p0;  // packed position of the block's first point
  // one thread per 8 channels (16 bytes) of a position
  for (int e = threadIdx.x; e < np * Cv; e += 256) {
    const int c = 8 * (e % Cv);
    const int pl = e / Cv;
    s16x8 v;
    if (c < Cs) {
      int q = p0 + pl;
      const int x = q % a.Dx;
      q /= a.Dx;
      const int y = q % a.Dy, z = q / a.Dy;
      v = *(const s16x8 *)(src + (cpos + ((z / 2) * (a.Dy / 2) + y / 2) * (a.Dx / 2) + x / 2) * Cs +
                           c);
    } else {
      v = *(const s16x8 *)(skip + (g0 + pl) * Ck + (c - Cs));
    }
    *(s16x8 *)(dst + (g0 + pl) * Cd + c) = v;
  }
}

__global__ __launch_bounds__(256) void k_mask_down4_rag(const pdfsnn_fld *fld,
                                                        const pdfsnn_seg *s0, int nf, uint16_t *m,
                                                        float far) {
  const int f = seg_by_eblk(s0, nf, (int)blockIdx.x);
  const pdfsnn_seg a = s0[f];
  const pdfsnn_fld d = fld[f];
  const int p0 = ((int)blockIdx.x - a.eblk) * RAG_EP;
  const int np = min(RAG_EP, a.Dz * a.Dy * a.Dx - p0);
  // 4 threads per coarse position, 16 cells each
  for (int e = threadIdx.x; e < np * 4; e += 256) {
    const int part = e & 3;
    int q = p0 + (e >> 2);
    const int x = q % a.Dx;
    q /= a.Dx;
    const int y = q % a.Dy, z = q / a.Dy;
    int known = 0, cnt = 0;
    for (int k = 16 * part; k < 16 * part + 16; k++) {
      const int zz = 4 * z + (k >> 4), yy = 4 * y + ((k >> 2) & 3), xx = 4 * x + (k & 3);
      if (zz < d.Z && yy < d.Y && xx < d.X) {
        cnt++;
        known += d.init[((long)zz * d.Y + yy) * d.X + xx] < far;
      }
    }
    // the four parts of a position are consecutive lanes of one wave
    known += __shfl_xor(known, 1);
    known += __shfl_xor(known, 2);
    cnt += __shfl_xor(cnt, 1);
    cnt += __shfl_xor(cnt, 2);
    const long i = (long)a.pos + p0 + (e >> 2);
    uint16_t *mi = m + i * 16;
    for (int c = 4 * part; c < 4 * part + 4; c++)
      mi[c] = c == 0 ? f2bf(cnt ? (float)known / 64.f : 0.f) : (uint16_t)0;
  }
}

__device__ __forceinline__ int fld_by_blk(const pdfsnn_fld *fld, int nf, int v) {
  int lo = 0, hi = nf;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (fld[mid].blk <= v) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_warm_up4_rag(const float *pred, int C,
                                                      const pdfsnn_seg *s0, const pdfsnn_fld *fld,
                                                      int nf, float inv_scale, float a, float b,
                                                      float cut) {
#pragma clang fp contract(off)
  const int f = fld_by_blk(fld, nf, (int)blockIdx.x);
  const pdfsnn_fld d = fld[f];
  const pdfsnn_seg sg = s0[f];
  const int Dz = sg.Dz, Dy = sg.Dy, Dx = sg.Dx;
  const float *pr = pred + (long)sg.pos * C;
  const int i0 = ((int)blockIdx.x - d.blk) * RAG_EF;
  const int n = min(RAG_EF, d.X * d.Y * d.Z - i0);
  for (int e = threadIdx.x; e < n; e += 256) {
    const int li = i0 + e;
    const int x = li % d.X;
    const int q = li / d.X;
    const int y = q % d.Y, z = q / d.Y;
    const float fx = 0.25f * x, fy = 0.25f * y, fz = 0.25f * z;
    const int x0 = min((int)fx, Dx - 1), y0 = min((int)fy, Dy - 1), z0 = min((int)fz, Dz - 1);
    const int x1 = min(x0 + 1, Dx - 1), y1 = min(y0 + 1, Dy - 1), z1 = min(z0 + 1, Dz - 1);
    const float tx = fx - x0, ty = fy - y0, tz = fz - z0;
    auto P = [&](int zz, int yy, int xx) { return pr[(((long)zz * Dy + yy) * Dx + xx) * C]; };
    const float c00 = P(z0, y0, x0) * (1 - tx) + P(z0, y0, x1) * tx;
    const float c01 = P(z0, y1, x0) * (1 - tx) + P(z0, y1, x1) * tx;
    const float c10 = P(z1, y0, x0) * (1 - tx) + P(z1, y0, x1) * tx;
    const float c11 = P(z1, y1, x0) * (1 - tx) + P(z1, y1, x1) * tx;
    const float c0 = c00 * (1 - ty) + c01 * ty, c1 = c10 * (1 - ty) + c11 * ty;
    const float dd = (c0 * (1 - tz) + c1 * tz) * inv_scale;
    const float f0 = d.init[li];
    const float up = dd >= cut ? f0 : a * (dd > 0.f ? dd : 0.f) + b;
    d.w[li] = f0 < up ? f0 : up;
  }
}

// ---------------------------------------------------------------------------
// C-ABI launchers
// ---------------------------------------------------------------------------
// f(NT, K32) for a conv's channel counts: NT = Cout / 16 output tiles (1,
// 2, 3, 4, 6 or 8: the layers' counts), K32 when Cin is a multiple of 32.
// Returns -1 for another Cout.
template <typename F>
static int conv3_dispatch(int Cin, int Cout, F f) {
  const int nt = Cout / 16;
  auto with_nt = [&](auto k32) {
    switch (nt) {
      case 1: f(std::integral_constant<int, 1>(), k32); return 0;
      case 2: f(std::integral_constant<int, 2>(), k32); return 0;
      case 3: f(std::integral_constant<int, 3>(), k32); return 0;
      case 4: f(std::integral_constant<int, 4>(), k32); return 0;
      case 6: f(std::integral_constant<int, 6>(), k32); return 0;
      case 8: f(std::integral_constant<int, 8>(), k32); return 0;
      default: return -1;
    }
  };
  if (Cin % 32 == 0) return with_nt(std::integral_constant<bool, true>());
  return with_nt(std::integral_constant<bool, false>());
}

extern "C" int pdfsnn_conv3(const uint16_t *in, int Cin, const uint16_t *W, const float *bias,
                            int Cout, uint16_t *out, float *out_f32, int n, int Dz, int Dy,
                            int Dx, int relu, int flip, hipStream_t st) {
  if (Cin % 16 || Cout % 16 || Cout > 128 || Cin > 128 || Cin < 16 || Cout < 16) return -1;
  if (n < 1 || Dz < 1 || Dy < 1 || Dx < 1) return -1;
  // output tile depth 4, or 1 on the coarse levels (4x the blocks) when the
  // grid would otherwise give under 64 blocks or the conv has one channel
  // chunk and under 512 blocks. Measured at a 128^3 forward: depth 1 takes
  // the 16^3 layers from 45 to 30 us and the 32^3 first layer from 31 to 21
  // us, but its other 32^3 layers (2 and 4 chunks) from 46 / 77 to 55 / 103
  // us; depth 8 was slower everywhere (register-bound occupancy).
  auto nblocks = [&](int tz) {
    return (long)n * ((Dz + tz - 1) / tz) * ((Dy + 7) / 8) * ((Dx + 7) / 8);
  };
  const long nb4 = nblocks(4);
  const int tz = (nb4 < 64 || (nb4 < 512 && Cin <= 32)) ? 1 : 4;
  const long blocks = nblocks(tz);
  if (blocks >= (1L << 31)) return -1;
  const dim3 grid((unsigned)blocks), blk(256);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, blk, 0, st, in, Cin, W, bias, Cout, out, out_f32, n, Dz, Dy,
                       Dx, relu, flip);
  };
  if (conv3_dispatch(Cin, Cout, [&](auto nt, auto k32) {
        constexpr int NT = decltype(nt)::value;
        constexpr bool K32 = decltype(k32)::value;
        if (tz == 4)
          go(k_conv3_tile<NT, K32, 4>);
        else
          go(k_conv3_tile<NT, K32, 1>);
      }))
    return -1;
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_conv3_wgrad(const uint16_t *X, int Cin, const uint16_t *dY, int Cout,
                                  float *dW, float *db, int n, int Dz, int Dy, int Dx,
                                  hipStream_t st) {
  if (Cin % 16 || Cout % 16 || Cout > 64 || Cin > 128 || (Cout / 16) * (Cin / 16) > 32) return -1;
  const long M = (long)n * Dz * Dy * Dx;
  const long nchunks = (M + 63) / 64;
  const int per = 16;
  const unsigned bx = (unsigned)((nchunks + per - 1) / per);
  hipLaunchKernelGGL(k_conv3_wgrad, dim3(bx, 27), dim3(256), 0, st, X, Cin, dY, Cout, dW, db, n,
                     Dz, Dy, Dx, per);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_pool2(const uint16_t *in, uint16_t *out, int n, int Dz, int Dy, int Dx,
                            int C, hipStream_t st) {
  hipLaunchKernelGGL(k_pool2, dim3(grid1((long)n * Dz * Dy * Dx * C / 8, 256)), dim3(256), 0, st,
                     in, out, n, Dz, Dy, Dx, C);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_pool2_bwd(const float *dout, float *din, int n, int Dz, int Dy, int Dx,
                                int C, int accumulate, hipStream_t st) {
  hipLaunchKernelGGL(k_pool2_bwd, dim3(grid1((long)n * Dz * Dy * Dx * C, 256)), dim3(256), 0,
                     st, dout, din, n, Dz, Dy, Dx, C, accumulate);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_up2_concat(const uint16_t *src, int Cs, const uint16_t *skip, int Ck,
                                 uint16_t *dst, int n, int Dz, int Dy, int Dx, hipStream_t st) {
  if (Cs % 16 || Ck % 16) return -1;
  hipLaunchKernelGGL(k_up2_concat, dim3(grid1((long)n * Dz * Dy * Dx * (Cs + Ck) / 8, 256)),
                     dim3(256), 0, st, src, Cs, skip, Ck, dst, n, Dz, Dy, Dx);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_up2_concat_bwd(const float *ddst, int Cs, int Ck, float *dsrc,
                                     float *dskip, int n, int Dz, int Dy, int Dx,
                                     hipStream_t st) {
  hipLaunchKernelGGL(k_up2_concat_bwd, dim3(grid1((long)n * Dz * Dy * Dx * Ck, 256)), dim3(256),
                     0, st, ddst, Cs, Ck, dsrc, dskip, n, Dz, Dy, Dx);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_relu_bwd(const float *g, const uint16_t *y, uint16_t *out, long N,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid1(N, 256)), dim3(256), 0, st, g, y, out, N);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_mse(const float *o, int C, const float *t, long N, uint16_t *grad,
                          float *loss, hipStream_t st) {
  hipLaunchKernelGGL(k_mse, dim3(grid1(N, 256) < 256 ? grid1(N, 256) : 256), dim3(256), 0, st, o,
                     C, t, N, grad, loss);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_adam(float *w, const float *gw, float *m, float *v, long N, float lr,
                           float b1, float b2, float eps, int step, uint16_t *wb, uint16_t *wt,
                           int Co, int Ci, hipStream_t st) {
  float c1 = 1.f, c2 = 1.f;
  for (int i = 0; i < step; i++) {
    c1 *= b1;
    c2 *= b2;
  }
  hipLaunchKernelGGL(k_adam, dim3(grid1(N, 256)), dim3(256), 0, st, w, gw, m, v, N, lr, b1, b2,
                     eps, 1.f - c1, 1.f - c2, wb, wt, Co, Ci);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_f32_to_bf16(const float *x, uint16_t *y, long N, hipStream_t st) {
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid1(N, 256)), dim3(256), 0, st, x, y, N);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_transpose_w(const uint16_t *wb, uint16_t *wt, int Co, int Ci,
                                  hipStream_t st) {
  hipLaunchKernelGGL(k_transpose_w, dim3(grid1(27L * Co * Ci, 256)), dim3(256), 0, st, wb, wt,
                     Co, Ci);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_mask_down4(const float *f, int X, int Y, int Z, uint16_t *m, int Dz,
                                 int Dy, int Dx, float far, hipStream_t st) {
  hipLaunchKernelGGL(k_mask_down4, dim3(grid1((long)Dz * Dy * Dx, 256)), dim3(256), 0, st, f, X,
                     Y, Z, m, Dz, Dy, Dx, far);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_target_down4(const float *u, int X, int Y, int Z, float *t, int Dz,
                                   int Dy, int Dx, float scale, hipStream_t st) {
  hipLaunchKernelGGL(k_target_down4, dim3(grid1((long)Dz * Dy * Dx, 256)), dim3(256), 0, st, u,
                     X, Y, Z, t, Dz, Dy, Dx, scale);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_warm_up4(const float *pred, int C, int Dz, int Dy, int Dx, float inv_scale,
                               float a, float b, float cut, const float *init, float *w, int X,
                               int Y, int Z, hipStream_t st) {
  hipLaunchKernelGGL(k_warm_up4, dim3(grid1((long)X * Y * Z, 256)), dim3(256), 0, st, pred, C,
                     Dz, Dy, Dx, inv_scale, a, b, cut, init, w, X, Y, Z);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_conv3_rag(const uint16_t *in, int Cin, const uint16_t *W, const float *bias,
                                int Cout, uint16_t *out, float *out_f32, const pdfsnn_seg *seg,
                                int nf, int nblocks, int relu, hipStream_t st) {
  if (Cin % 16 || Cout % 16 || Cout > 128 || Cin > 128 || Cin < 16 || Cout < 16) return -1;
  if (nf < 1 || nblocks < 1) return -1;
  const dim3 grid((unsigned)nblocks), blk(256);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, blk, 0, st, in, Cin, W, bias, Cout, out, out_f32, seg, nf,
                       relu);
  };
  if (conv3_dispatch(Cin, Cout, [&](auto nt, auto k32) {
        go(k_conv3_tile_rag<decltype(nt)::value, decltype(k32)::value>);
      }))
    return -1;
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_pool2_rag(const uint16_t *in, uint16_t *out, const pdfsnn_seg *seg_in,
                                const pdfsnn_seg *seg_out, int nf, int eblocks_out, int C,
                                hipStream_t st) {
  if (nf < 1 || eblocks_out < 1) return -1;
  hipLaunchKernelGGL(k_pool2_rag, dim3((unsigned)eblocks_out), dim3(256), 0, st, in, out, seg_in,
                     seg_out, nf, C);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_up2_concat_rag(const uint16_t *src, int Cs, const uint16_t *skip, int Ck,
                                     uint16_t *dst, const pdfsnn_seg *seg_fine,
                                     const pdfsnn_seg *seg_coarse, int nf, int eblocks_fine,
                                     hipStream_t st) {
  if (nf < 1 || eblocks_fine < 1 || Cs % 16 || Ck % 16) return -1;
  hipLaunchKernelGGL(k_up2_concat_rag, dim3((unsigned)eblocks_fine), dim3(256), 0, st, src, Cs,
                     skip, Ck, dst, seg_fine, seg_coarse, nf);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_mask_down4_rag(const pdfsnn_fld *fld, const pdfsnn_seg *seg0, int nf,
                                     int eblocks0, uint16_t *m, float far, hipStream_t st) {
  if (nf < 1 || eblocks0 < 1) return -1;
  hipLaunchKernelGGL(k_mask_down4_rag, dim3((unsigned)eblocks0), dim3(256), 0, st, fld, seg0, nf,
                     m, far);
  return hip_rc(hipGetLastError());
}

extern "C" int pdfsnn_warm_up4_rag(const float *pred, int C, const pdfsnn_seg *seg0,
                                   const pdfsnn_fld *fld, int nf, int fblocks, float inv_scale,
                                   float a, float b, float cut, hipStream_t st) {
  if (nf < 1 || fblocks < 1) return -1;
  hipLaunchKernelGGL(k_warm_up4_rag, dim3((unsigned)fblocks), dim3(256), 0, st, pred, C, seg0,
                     fld, nf, inv_scale, a, b, cut);
  return hip_rc(hipGetLastError());
}
