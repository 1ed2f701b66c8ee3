// MLP surrogate kernels: bf16 matrix-core products with FP32 accumulation, Fourier features, SiLU,
// loss, bias gradients and Adam, plus the batch gather of training data and table prediction.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "surrogate.hpp"

namespace sfsim {

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

constexpr int kEw = 256;  // elementwise block size

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// SiLU and its derivative with the hardware reciprocal (1 ulp) instead of an IEEE division: the
// results are rounded to bf16 for the products anyway.
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }

__device__ __forceinline__ float dsilu(float z) {
  float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// One wavefront computes a 32 x 32 tile of C with v_mfma_f32_32x32x8bf16_1k.  Operand layout of that
// instruction: lane l supplies A(l % 32, 4 (l / 32) + i) and B(4 (l / 32) + i, l % 32), i = 0..3;
// accumulator element e of lane l is C(8 (e / 4) + 4 (l / 32) + e % 4, l % 32).
// Split K: workgroup z covers k in [z kc, min(K, (z + 1) kc)) and writes its partial product to
// C + z * split_stride (the caller sums the partials in a fixed order).
__global__ void __launch_bounds__(64) gemm_kernel(const uint16_t* __restrict__ A, long long sam, long long sak,
                                                  const uint16_t* __restrict__ B, long long sbk, long long sbn,
                                                  float* __restrict__ C, int ldc, const float* __restrict__ bias,
                                                  int act, uint16_t* __restrict__ outb, int M, int N, int K, int kc,
                                                  long long split_stride) {
  const int lane = threadIdx.x;
  const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  const int r = lane & 31, kq = (lane >> 5) * 4;
  const int am = m0 + r, bn = n0 + r;
  const bool arow = am < M, bcol = bn < N;
  const int kbeg = blockIdx.z * kc, kend = min(K, kbeg + kc);
  C += blockIdx.z * split_stride;
  float16_t acc;
#pragma unroll
  for (int e = 0; e < 16; ++e) acc[e] = 0.f;
  for (int k0 = kbeg; k0 < kend; k0 += 8) {
    short4_t va, vb;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int k = k0 + kq + i;
      bool kin = k < kend;
      va[i] = (short)((arow && kin) ? A[am * sam + k * sak] : 0);
      vb[i] = (short)((bcol && kin) ? B[k * sbk + bn * sbn] : 0);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(va, vb, acc, 0, 0, 0);
  }
  const int n = n0 + r;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    int m = m0 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
    if (m < M && n < N) {
      float v = acc[e] + (bias ? bias[n] : 0.f);
      long long o = (long long)m * ldc + n;
      if (C) C[o] = v;
      if (outb) outb[o] = f2bf(act == 1 ? silu(v) : v);
    }
  }
}

// The 17 Fourier features of one input x: x, then sin / cos (2^b pi x) for b = 0..7 (bf16).
__device__ inline void encode_value(float v, uint16_t* out) {
  out[0] = f2bf(v);
  float a = 3.14159265358979f * v;
  for (int b = 0; b < 8; ++b) {
    float sn, cs;
    sincosf(a, &sn, &cs);
    out[1 + 2 * b] = f2bf(sn);
    out[2 + 2 * b] = f2bf(cs);
    a *= 2.0f;  // exact
  }
}

// Features of one input row: the 17 features of each of the 10 inputs, zero padded to 192.
__device__ inline void encode_row(const float* x, uint16_t* out) {
  for (int i = 0; i < kMlpIn; ++i) encode_value(x[i], out + kMlpBands * i);
  for (int f = kMlpBands * kMlpIn; f < kMlpFeat; ++f) out[f] = 0;
}

__global__ void encode_rows_kernel(const float* __restrict__ in, uint16_t* __restrict__ X, int rows) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  encode_row(in + (long long)r * kMlpIn, X + (long long)r * kMlpFeat);
}

__device__ inline void entry_coords(const DevParams& P, long long entry, float* x) {
  const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  int ia = (int)(entry % NA);
  long long q = entry / NA;
  int is = (int)(q % NL);
  q /= NL;
  int ie = (int)(q % NE);
  int ih = (int)(q / NE);
  x[0] = (float)ih / (NH - 1);
  x[1] = (float)ie / (NE - 1);
  x[2] = (float)is / (NL - 1);
  x[3] = (float)ia / (NA - 1);
}

__global__ void gather_kernel(const DevParams P, const AtmosphereCode* __restrict__ codes, int natm,
                              const float* __restrict__ targets, uint64_t seed, int step, int rows,
                              uint16_t* __restrict__ X, float* __restrict__ T) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const long long entries = (long long)P.nS[0] * P.nS[1] * P.nS[2] * P.nS[3];
  uint64_t h = splitmix64(seed ^ splitmix64(((uint64_t)step << 32) | (uint64_t)r));
  int atm = (int)(h % (uint64_t)natm);
  long long entry = (long long)((h >> 16) % (uint64_t)entries);
  float x[kMlpIn];
  entry_coords(P, entry, x);
  for (int i = 0; i < 6; ++i) x[4 + i] = codes[atm].v[i];
  encode_row(x, X + (long long)r * kMlpFeat);
  const float* t = targets + ((long long)atm * entries + entry) * kMlpUsed;
  for (int o = 0; o < kMlpOut; ++o) T[(long long)r * kMlpOut + o] = o < kMlpUsed ? t[o] : 0.f;
}

struct SumTables {
  const float4* p[kMlpOrders];
};

__global__ void targets_kernel(const SumTables sums, long long entries, float* __restrict__ t) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= entries) return;
  for (int k = 0; k < kMlpOrders; ++k) {
    const float4 v = sums.p[k][e];
    t[e * kMlpUsed + 3 * k + 0] = logf(fmaxf(v.x, 0.f) + kMlpEps) + kMlpShift;
    t[e * kMlpUsed + 3 * k + 1] = logf(fmaxf(v.y, 0.f) + kMlpEps) + kMlpShift;
    t[e * kMlpUsed + 3 * k + 2] = logf(fmaxf(v.z, 0.f) + kMlpEps) + kMlpShift;
  }
}

// dY = mask (Y - T) / rows (FP32 and bf16), loss partials 0.5 sum d^2 / rows per block.
__global__ void __launch_bounds__(kEw) loss_kernel(const float* __restrict__ Y, const float* __restrict__ T, int rows,
                                                   float* __restrict__ dYf, uint16_t* __restrict__ dY,
                                                   float* __restrict__ partial) {
  __shared__ float red[kEw / 64];
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < (long long)rows * kMlpOut) {
    int o = (int)(i % kMlpOut);
    float d = o < kMlpUsed ? (Y[i] - T[i]) : 0.f;
    l = 0.5f * d * d / rows;
    float g = d / rows;
    dYf[i] = g;
    dY[i] = f2bf(g);
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int w = 0; w < kEw / 64; ++w) s += red[w];
    partial[blockIdx.x] = s;
  }
}

__global__ void dsilu_kernel(const float* __restrict__ dA, const float* __restrict__ Z, long long n,
                             float* __restrict__ dZf, uint16_t* __restrict__ dZ) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float g = dA[i] * dsilu(Z[i]);
  dZf[i] = g;
  dZ[i] = f2bf(g);
}

// out[c] = mean over rows of T[row][c] for the kMlpUsed target columns (row stride kMlpUsed), FP64
// partial sums in a fixed order: the output biases' initial values.
__global__ void __launch_bounds__(kEw) target_mean_kernel(const float* __restrict__ T, long long rows,
                                                          float* __restrict__ out) {
  __shared__ double red[kEw / 64];
  const int c = blockIdx.x;
  double s = 0.0;
  for (long long r = threadIdx.x; r < rows; r += blockDim.x) s += T[r * kMlpUsed + c];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kEw / 64; ++w) t += red[w];
    out[c] = (float)(t / (double)rows);
  }
}

// out[c] = sum over rows of G[row][c] (one workgroup per column, fixed summation order).
__global__ void __launch_bounds__(kEw) colsum_kernel(const float* __restrict__ G, int rows, int cols,
                                                     float* __restrict__ out) {
  __shared__ float red[kEw / 64];
  int c = blockIdx.x;
  float s = 0.f;
  for (int r = threadIdx.x; r < rows; r += blockDim.x) s += G[(long long)r * cols + c];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < kEw / 64; ++w) t += red[w];
    out[c] = t;
  }
}

__global__ void adam_kernel(float* __restrict__ theta, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1,
                            float c2, uint16_t* __restrict__ theta_bf) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float t = theta[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  theta[i] = t;
  theta_bf[i] = f2bf(t);
}

__global__ void to_bf16_kernel(const float* __restrict__ in, uint16_t* __restrict__ out, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f2bf(in[i]);
}

// ---- fused inference over a scatter grid ---------------------------------------------------------
// Layer 1 is linear in the features and the features of an entry split into those of its 4 grid
// coordinates and those of the 6 atmosphere inputs, which one prediction holds fixed.  So
//   Z1(entry) = b1 + W1_atm^T f(atm) + sum_c W1_c^T f(x_c(entry)),
// and the terms are tabulated once per prediction: row 0 = b1 + W1_atm^T f(atm), then one row of 128
// per value of each grid coordinate (heights, elevations, sun elevations, headings).
__global__ void layer1_tables_kernel(const DevParams P, AtmosphereCode code, const uint16_t* __restrict__ wbf,
                                     const float* __restrict__ theta, float* __restrict__ ztab) {
  const int j = threadIdx.x;  // hidden unit (blockDim.x == kMlpHidden)
  int row = blockIdx.x;
  uint16_t f[kMlpBands];
  const uint16_t* W1 = wbf + MlpLayout::W1;
  float acc;
  if (row == 0) {
    acc = theta[MlpLayout::b1 + j];
    for (int i = 0; i < 6; ++i) {
      encode_value(code.v[i], f);
      for (int q = 0; q < kMlpBands; ++q)
        acc += bf2f(W1[(size_t)(kMlpBands * (4 + i) + q) * kMlpHidden + j]) * bf2f(f[q]);
    }
  } else {
    int v = row - 1, c = 0;
    while (v >= P.nS[c]) v -= P.nS[c++];
    encode_value((float)v / (float)(P.nS[c] - 1), f);
    acc = 0.f;
    for (int q = 0; q < kMlpBands; ++q) acc += bf2f(W1[(size_t)(kMlpBands * c + q) * kMlpHidden + j]) * bf2f(f[q]);
  }
  ztab[(size_t)blockIdx.x * kMlpHidden + j] = acc;
}

constexpr int kFusedWaves = 4;
constexpr int kLdsRow = kMlpHidden + 8;  // padded row (bf16) of the transposed weights in LDS, 16-byte aligned
constexpr int kFusedLds = (2 * kMlpHidden + 32) * kLdsRow * 2;

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// Column order of the transposed weights in LDS: within each block of 16 input units the middle two
// groups of 4 swap places, so that the 8 units a lane half supplies to one K = 16 step (see below) are
// contiguous: one 16-byte LDS read per operand.
__device__ __forceinline__ int lds_column(int k) {
  const int g = (k >> 2) & 3;
  const int gs = g == 1 ? 2 : (g == 2 ? 1 : g);
  return (k & ~15) | (gs << 2) | (k & 3);
}

// bf16 (round to nearest even, v_cvt_pk_bf16_f32) of SiLU of 8 values.
__device__ __forceinline__ bf16x8_t silu_bf16x8(float4 lo, float4 hi) {
  bf16x8_t r;
  r[0] = (__bf16)silu(lo.x);
  r[1] = (__bf16)silu(lo.y);
  r[2] = (__bf16)silu(lo.z);
  r[3] = (__bf16)silu(lo.w);
  r[4] = (__bf16)silu(hi.x);
  r[5] = (__bf16)silu(hi.y);
  r[6] = (__bf16)silu(hi.z);
  r[7] = (__bf16)silu(hi.w);
  return r;
}

__device__ __forceinline__ float4 add5(float4 a, float4 b, float4 c, float4 d, float4 e) {
  return make_float4(a.x + b.x + c.x + d.x + e.x, a.y + b.y + c.y + d.y + e.y, a.z + b.z + c.z + d.z + e.z,
                     a.w + b.w + c.w + d.w + e.w);
}

// One wavefront evaluates 32 entries at a time with the hidden units as the row dimension of every
// product (Z^T = W^T A^T) on v_mfma_f32_32x32x16_bf16.  Its accumulator leaves lane l (half h = l / 32)
// with hidden units 8 g + 4 h + i (g = 0..3, i = 0..3) of each block of 32, for entry l % 32.  A K = 16
// step t over the previous layer's units 16 t .. 16 t + 15 takes from lane half h the 8 units
// 16 t + 4 h + i and 16 t + 8 + 4 h + i (the same register positions in A and B pair up, whatever unit
// they stand for), which that lane already holds: activations never leave the registers.  W2^T, W3^T and
// W4^T (padded to 32 rows, columns in lds_column order) are staged once per workgroup in LDS;
// workgroups loop over the tiles.
struct PredTables {
  float4* p[kMlpOrders];
};

__global__ void __launch_bounds__(64 * kFusedWaves, 2) fused_predict_kernel(
    const DevParams P, const uint16_t* __restrict__ wbf, const float* __restrict__ theta,
    const float* __restrict__ ztab, long long entries, const PredTables out) {
  __shared__ __align__(16) __bf16 lds[kFusedLds / 2];
  __bf16* w2t = lds;
  __bf16* w3t = w2t + kMlpHidden * kLdsRow;
  __bf16* w4t = w3t + kMlpHidden * kLdsRow;
  {
    // 16-byte loads of 8 consecutive weights W[k][n .. n + 7] (all issued before the LDS scatter),
    // transposed into W^T rows; W4 rows 16..31 of W4^T are zero
    static_assert(MlpLayout::W2 % 8 == 0 && MlpLayout::W3 % 8 == 0 && MlpLayout::W4 % 8 == 0, "16-byte rows");
    constexpr int kVec = kMlpHidden * kMlpHidden / 8 / (64 * kFusedWaves);  // loads per thread per matrix
    uint4 v2[kVec], v3[kVec];
    const uint4* s2 = (const uint4*)(wbf + MlpLayout::W2);
    const uint4* s3 = (const uint4*)(wbf + MlpLayout::W3);
    const uint4* s4 = (const uint4*)(wbf + MlpLayout::W4);
#pragma unroll
    for (int q = 0; q < kVec; ++q) {
      v2[q] = s2[threadIdx.x + q * 64 * kFusedWaves];
      v3[q] = s3[threadIdx.x + q * 64 * kFusedWaves];
    }
    const uint4 v4 = s4[threadIdx.x];  // 128 x 16 weights = 256 vectors of 8
#pragma unroll
    for (int q = 0; q < kVec; ++q) {
      const int i = 8 * (threadIdx.x + q * 64 * kFusedWaves);
      const int k = i / kMlpHidden, n = i % kMlpHidden, c = lds_column(k);
      const uint16_t* a2 = (const uint16_t*)&v2[q];
      const uint16_t* a3 = (const uint16_t*)&v3[q];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        w2t[(n + j) * kLdsRow + c] = __builtin_bit_cast(__bf16, a2[j]);
        w3t[(n + j) * kLdsRow + c] = __builtin_bit_cast(__bf16, a3[j]);
      }
    }
    {
      const int i = 8 * threadIdx.x;
      const int k = i / kMlpOut, n = i % kMlpOut, c = lds_column(k);
      const uint16_t* a4 = (const uint16_t*)&v4;
#pragma unroll
      for (int j = 0; j < 8; ++j) w4t[(n + j) * kLdsRow + c] = __builtin_bit_cast(__bf16, a4[j]);
    }
    for (int i = threadIdx.x; i < (32 - kMlpOut) * kMlpHidden; i += blockDim.x)
      w4t[(kMlpOut + i / kMlpHidden) * kLdsRow + i % kMlpHidden] = __builtin_bit_cast(__bf16, (uint16_t)0);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 31, h = lane >> 5;
  const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  const float* b2 = theta + MlpLayout::b2;
  const float* b3 = theta + MlpLayout::b3;
  const float* b4 = theta + MlpLayout::b4;
  const long long tiles = (entries + 31) / 32;
  for (long long t = (long long)blockIdx.x * kFusedWaves + wave; t < tiles; t += (long long)gridDim.x * kFusedWaves) {
    const long long e = t * 32 + r;
    const unsigned ec = (unsigned)(e < entries ? e : entries - 1);  // < 2^31: the host caps the grid
    const int ia = (int)(ec % (unsigned)NA);
    unsigned q = ec / (unsigned)NA;
    const int is = (int)(q % (unsigned)NL);
    q /= (unsigned)NL;
    const int ie = (int)(q % (unsigned)NE);
    const int ih = (int)(q / (unsigned)NE);
    const float4* z0 = (const float4*)ztab;
    const float4* zh = (const float4*)(ztab + (size_t)(1 + ih) * kMlpHidden);
    const float4* ze = (const float4*)(ztab + (size_t)(1 + NH + ie) * kMlpHidden);
    const float4* zs = (const float4*)(ztab + (size_t)(1 + NH + NE + is) * kMlpHidden);
    const float4* za = (const float4*)(ztab + (size_t)(1 + NH + NE + NL + ia) * kMlpHidden);
    bf16x8_t b[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int lo = 4 * s + h, hi = 4 * s + 2 + h;  // float4 index of units 16 s + 4 h and 16 s + 8 + 4 h
      b[s] = silu_bf16x8(add5(z0[lo], zh[lo], ze[lo], zs[lo], za[lo]), add5(z0[hi], zh[hi], ze[hi], zs[hi], za[hi]));
    }
#pragma unroll 1
    for (int layer = 0; layer < 2; ++layer) {
      const __bf16* wt = layer ? w3t : w2t;
      const float* bias = layer ? b3 : b2;
      float16_t acc[4];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int x = 0; x < 16; ++x) acc[mt][x] = 0.f;
#pragma unroll
      for (int s = 0; s < 8; ++s)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          const bf16x8_t a = *(const bf16x8_t*)(wt + (32 * mt + r) * kLdsRow + 16 * s + 8 * h);
          acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[s], acc[mt], 0, 0, 0);
        }
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int mt = s >> 1, g = 2 * (s & 1);
        const float4 bl = *(const float4*)(bias + 32 * mt + 8 * g + 4 * h);
        const float4 bh = *(const float4*)(bias + 32 * mt + 8 * (g + 1) + 4 * h);
        b[s] = silu_bf16x8(make_float4(acc[mt][4 * g] + bl.x, acc[mt][4 * g + 1] + bl.y, acc[mt][4 * g + 2] + bl.z,
                                       acc[mt][4 * g + 3] + bl.w),
                           make_float4(acc[mt][4 * g + 4] + bh.x, acc[mt][4 * g + 5] + bh.y,
                                       acc[mt][4 * g + 6] + bh.z, acc[mt][4 * g + 7] + bh.w));
      }
    }
    float16_t y;
#pragma unroll
    for (int x = 0; x < 16; ++x) y[x] = 0.f;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const bf16x8_t a = *(const bf16x8_t*)(w4t + r * kLdsRow + 16 * s + 8 * h);
      y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[s], y, 0, 0, 0);
    }
    // lane (r, h) holds outputs 4 h + i (y[i]) and 8 + 4 h + i (y[4 + i]) of entry r, i = 0..3
    float o[kMlpOut];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float u = __shfl(y[i], r + 32, 64), v = __shfl(y[4 + i], r + 32, 64);
      o[i] = y[i];
      o[4 + i] = u;
      o[8 + i] = y[4 + i];
      o[12 + i] = v;
    }
    if (h == 0 && e < entries) {
#pragma unroll
      for (int k = 0; k < kMlpOrders; ++k) {
        if (out.p[k]) {
          float v[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) v[c] = fmaxf(expf(o[3 * k + c] + b4[3 * k + c] - kMlpShift) - kMlpEps, 0.f);
          out.p[k][e] = make_float4(v[0], v[1], v[2], 0.f);
        }
      }
    }
  }
}

// C[i] = sum over z of partial[z][i], z in order.
__global__ void split_sum_kernel(const float* __restrict__ partial, int splits, long long n, float* __restrict__ C) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += partial[z * n + i];
  C[i] = s;
}

inline unsigned ew_blocks(long long n) { return (unsigned)((n + kEw - 1) / kEw); }

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc((void**)p, count * sizeof(T) < 256 ? 256 : count * sizeof(T));
}

}  // namespace

hipError_t mfma_gemm(const uint16_t* A, long long sam, long long sak, const uint16_t* B, long long sbk,
                     long long sbn, float* C, int ldc, const float* bias, int act, uint16_t* out_bf16, int M, int N,
                     int K, hipStream_t stream) {
  dim3 grid((unsigned)((N + 31) / 32), (unsigned)((M + 31) / 32));
  hipLaunchKernelGGL(gemm_kernel, grid, dim3(64), 0, stream, A, sam, sak, B, sbk, sbn, C, ldc, bias, act, out_bf16, M,
                     N, K, K, 0LL);
  return hipGetLastError();
}

// Weight gradients (M x N small, K = rows large): K split over kMaxSplits workgroups per tile, partial
// products in `partial` (kMaxSplits x M x N floats), then summed in split order: deterministic.
hipError_t mfma_gemm_split_k(const uint16_t* A, long long sam, long long sak, const uint16_t* B, long long sbk,
                             long long sbn, float* C, int M, int N, int K, float* partial, hipStream_t stream) {
  int kc = (K + kMaxSplits - 1) / kMaxSplits;
  kc = (kc + 7) / 8 * 8;
  if (kc < 64) kc = 64;
  const int splits = (K + kc - 1) / kc;
  dim3 grid((unsigned)((N + 31) / 32), (unsigned)((M + 31) / 32), (unsigned)splits);
  hipLaunchKernelGGL(gemm_kernel, grid, dim3(64), 0, stream, A, sam, sak, B, sbk, sbn, partial, N, nullptr, 0,
                     nullptr, M, N, K, kc, (long long)M * N);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long long n = (long long)M * N;
  hipLaunchKernelGGL(split_sum_kernel, dim3((unsigned)((n + kEw - 1) / kEw)), dim3(kEw), 0, stream, partial, splits, n,
                     C);
  return hipGetLastError();
}

#define SG_TRY(expr)                         \
  do {                                       \
    hipError_t e_ = (expr);                  \
    if (e_ != hipSuccess) return e_;         \
  } while (0)

hipError_t surrogate_alloc(Surrogate& s, int batch) {
  surrogate_free(s);
  const size_t n = MlpLayout::total;
  SG_TRY(dalloc(&s.theta, n));
  SG_TRY(dalloc(&s.theta_bf, n));
  SG_TRY(dalloc(&s.adam_m, n));
  SG_TRY(dalloc(&s.adam_v, n));
  SG_TRY(dalloc(&s.grad, n));
  SG_TRY(dalloc(&s.X, (size_t)batch * kMlpFeat));
  for (int l = 0; l < 3; ++l) {
    SG_TRY(dalloc(&s.Z[l], (size_t)batch * kMlpHidden));
    SG_TRY(dalloc(&s.A[l], (size_t)batch * kMlpHidden));
  }
  SG_TRY(dalloc(&s.Y, (size_t)batch * kMlpOut));
  SG_TRY(dalloc(&s.T, (size_t)batch * kMlpOut));
  SG_TRY(dalloc(&s.dY, (size_t)batch * kMlpOut));
  SG_TRY(dalloc(&s.dA, (size_t)batch * kMlpHidden));
  SG_TRY(dalloc(&s.dZf, (size_t)batch * kMlpHidden));
  SG_TRY(dalloc(&s.dZ, (size_t)batch * kMlpHidden));
  SG_TRY(dalloc(&s.loss, (size_t)(batch * kMlpOut) / kEw + 2));
  SG_TRY(dalloc(&s.partial, (size_t)kMaxSplits * kMlpFeat * kMlpHidden));
  s.batch_capacity = batch;
  return hipSuccess;
}

void surrogate_free(Surrogate& s) {
  void* ptrs[] = {s.theta, s.theta_bf, s.adam_m, s.adam_v, s.grad, s.X,  s.Z[0], s.Z[1], s.Z[2], s.A[0], s.A[1],
                  s.A[2],  s.Y,        s.T,      s.dY,     s.dA,   s.dZf, s.dZ,   s.loss, s.ztab, s.partial};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  s = Surrogate();
}

hipError_t surrogate_init(Surrogate& s, uint64_t seed, hipStream_t stream) {
  std::vector<float> h(MlpLayout::total, 0.f);
  uint64_t state = seed;
  auto uniform = [&]() {
    state = splitmix64(state);
    return (float)((state >> 11) * (1.0 / 9007199254740992.0));
  };
  auto fill = [&](size_t off, int fan_in, int fan_out, float gain) {
    float a = gain * std::sqrt(6.0f / (float)(fan_in + fan_out));
    for (size_t i = 0; i < (size_t)fan_in * fan_out; ++i) h[off + i] = (2.f * uniform() - 1.f) * a;
  };
  fill(MlpLayout::W1, kMlpFeat, kMlpHidden, 1.f);
  fill(MlpLayout::W2, kMlpHidden, kMlpHidden, 1.f);
  fill(MlpLayout::W3, kMlpHidden, kMlpHidden, 1.f);
  fill(MlpLayout::W4, kMlpHidden, kMlpOut, 0.1f);
  SG_TRY(hipMemcpyAsync(s.theta, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, stream));
  SG_TRY(hipMemsetAsync(s.adam_m, 0, h.size() * sizeof(float), stream));
  SG_TRY(hipMemsetAsync(s.adam_v, 0, h.size() * sizeof(float), stream));
  s.adam_step = 0;
  s.trained = false;
  SG_TRY(surrogate_sync_bf16(s, stream));
  return hipStreamSynchronize(stream);
}

hipError_t surrogate_sync_bf16(Surrogate& s, hipStream_t stream) {
  long long n = (long long)MlpLayout::total;
  hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_blocks(n)), dim3(kEw), 0, stream, s.theta, s.theta_bf, n);
  return hipGetLastError();
}

hipError_t surrogate_encode_rows(Surrogate& s, const float* inputs_dev, int rows, hipStream_t stream) {
  hipLaunchKernelGGL(encode_rows_kernel, dim3(ew_blocks(rows)), dim3(kEw), 0, stream, inputs_dev, s.X, rows);
  return hipGetLastError();
}

hipError_t surrogate_forward(Surrogate& s, int rows, hipStream_t stream) {
  const uint16_t* W[4] = {s.theta_bf + MlpLayout::W1, s.theta_bf + MlpLayout::W2, s.theta_bf + MlpLayout::W3,
                          s.theta_bf + MlpLayout::W4};
  const float* b[4] = {s.theta + MlpLayout::b1, s.theta + MlpLayout::b2, s.theta + MlpLayout::b3,
                       s.theta + MlpLayout::b4};
  // hidden layers: Z = A_prev W + b, A = SiLU(Z)
  SG_TRY(mfma_gemm(s.X, kMlpFeat, 1, W[0], kMlpHidden, 1, s.Z[0], kMlpHidden, b[0], 1, s.A[0], rows, kMlpHidden,
                   kMlpFeat, stream));
  SG_TRY(mfma_gemm(s.A[0], kMlpHidden, 1, W[1], kMlpHidden, 1, s.Z[1], kMlpHidden, b[1], 1, s.A[1], rows, kMlpHidden,
                   kMlpHidden, stream));
  SG_TRY(mfma_gemm(s.A[1], kMlpHidden, 1, W[2], kMlpHidden, 1, s.Z[2], kMlpHidden, b[2], 1, s.A[2], rows, kMlpHidden,
                   kMlpHidden, stream));
  SG_TRY(mfma_gemm(s.A[2], kMlpHidden, 1, W[3], kMlpOut, 1, s.Y, kMlpOut, b[3], 0, nullptr, rows, kMlpOut, kMlpHidden,
                   stream));
  return hipSuccess;
}

hipError_t surrogate_backward(Surrogate& s, int rows, hipStream_t stream) {
  const long long nout = (long long)rows * kMlpOut, nh = (long long)rows * kMlpHidden;
  // dY = mask (Y - T) / rows; dYf reuses dZf's storage (rows x 16 <= rows x 128)
  float* dYf = s.dZf;
  hipLaunchKernelGGL(loss_kernel, dim3(ew_blocks(nout)), dim3(kEw), 0, stream, s.Y, s.T, rows, dYf, s.dY, s.loss);
  SG_TRY(hipGetLastError());
  // layer 4: gW4 = A3^T dY, gb4 = colsum(dY), dA3 = dY W4^T
  SG_TRY(mfma_gemm_split_k(s.A[2], 1, kMlpHidden, s.dY, kMlpOut, 1, s.grad + MlpLayout::W4, kMlpHidden, kMlpOut,
                           rows, s.partial, stream));
  hipLaunchKernelGGL(colsum_kernel, dim3(kMlpOut), dim3(kEw), 0, stream, dYf, rows, kMlpOut, s.grad + MlpLayout::b4);
  SG_TRY(hipGetLastError());
  SG_TRY(mfma_gemm(s.dY, kMlpOut, 1, s.theta_bf + MlpLayout::W4, 1, kMlpOut, s.dA, kMlpHidden, nullptr, 0, nullptr,
                   rows, kMlpHidden, kMlpOut, stream));
  // hidden layers 3, 2, 1
  const size_t Woff[3] = {MlpLayout::W1, MlpLayout::W2, MlpLayout::W3};
  const size_t boff[3] = {MlpLayout::b1, MlpLayout::b2, MlpLayout::b3};
  for (int l = 2; l >= 0; --l) {
    hipLaunchKernelGGL(dsilu_kernel, dim3(ew_blocks(nh)), dim3(kEw), 0, stream, s.dA, s.Z[l], nh, s.dZf, s.dZ);
    SG_TRY(hipGetLastError());
    const uint16_t* prev = l == 0 ? s.X : s.A[l - 1];
    const int kin = l == 0 ? kMlpFeat : kMlpHidden;
    SG_TRY(mfma_gemm_split_k(prev, 1, kin, s.dZ, kMlpHidden, 1, s.grad + Woff[l], kin, kMlpHidden, rows, s.partial,
                             stream));
    hipLaunchKernelGGL(colsum_kernel, dim3(kMlpHidden), dim3(kEw), 0, stream, s.dZf, rows, kMlpHidden,
                       s.grad + boff[l]);
    SG_TRY(hipGetLastError());
    if (l > 0)
      SG_TRY(mfma_gemm(s.dZ, kMlpHidden, 1, s.theta_bf + Woff[l], 1, kMlpHidden, s.dA, kMlpHidden, nullptr, 0,
                       nullptr, rows, kMlpHidden, kMlpHidden, stream));
  }
  return hipSuccess;
}

hipError_t surrogate_adam(Surrogate& s, float lr, hipStream_t stream) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  s.adam_step += 1;
  float c1 = 1.f - std::pow(b1, (float)s.adam_step), c2 = 1.f - std::pow(b2, (float)s.adam_step);
  long long n = (long long)MlpLayout::total;
  hipLaunchKernelGGL(adam_kernel, dim3(ew_blocks(n)), dim3(kEw), 0, stream, s.theta, s.grad, s.adam_m, s.adam_v, n, lr,
                     b1, b2, eps, c1, c2, s.theta_bf);
  return hipGetLastError();
}

hipError_t surrogate_gather_batch(Surrogate& s, const DevParams& P, const AtmosphereCode* codes_dev, int natm,
                                  const float* targets_dev, uint64_t seed, int step, int rows, hipStream_t stream) {
  hipLaunchKernelGGL(gather_kernel, dim3(ew_blocks(rows)), dim3(kEw), 0, stream, P, codes_dev, natm, targets_dev, seed,
                     step, rows, s.X, s.T);
  return hipGetLastError();
}

hipError_t surrogate_init_output_bias(Surrogate& s, const float* targets_dev, long long rows, hipStream_t stream) {
  hipLaunchKernelGGL(target_mean_kernel, dim3(kMlpUsed), dim3(kEw), 0, stream, targets_dev, rows,
                     s.theta + MlpLayout::b4);
  return hipGetLastError();
}

hipError_t surrogate_make_targets(const float4* const* sums, long long entries, float* targets_dev, hipStream_t stream) {
  SumTables st;
  for (int k = 0; k < kMlpOrders; ++k) st.p[k] = sums[k];
  hipLaunchKernelGGL(targets_kernel, dim3(ew_blocks(entries)), dim3(kEw), 0, stream, st, entries, targets_dev);
  return hipGetLastError();
}

hipError_t surrogate_reserve_tables(Surrogate& s, const DevParams& P) {
  const int rows = 1 + P.nS[0] + P.nS[1] + P.nS[2] + P.nS[3];
  if ((size_t)rows * kMlpHidden > s.ztab_capacity) {
    if (s.ztab) SG_TRY(hipFree(s.ztab));
    s.ztab = nullptr;
    s.ztab_capacity = 0;
    SG_TRY(dalloc(&s.ztab, (size_t)rows * kMlpHidden));
    s.ztab_capacity = (size_t)rows * kMlpHidden;
  }
  return hipSuccess;
}

hipError_t surrogate_predict_tables(Surrogate& s, const DevParams& P, const AtmosphereCode& code, float4* const* out,
                                    hipStream_t stream) {
  const long long entries = (long long)P.nS[0] * P.nS[1] * P.nS[2] * P.nS[3];
  const int rows = 1 + P.nS[0] + P.nS[1] + P.nS[2] + P.nS[3];
  SG_TRY(surrogate_reserve_tables(s, P));
  hipLaunchKernelGGL(layer1_tables_kernel, dim3(rows), dim3(kMlpHidden), 0, stream, P, code, s.theta_bf, s.theta,
                     s.ztab);
  SG_TRY(hipGetLastError());
  const long long tiles = (entries + 31) / 32;
  const long long groups = (tiles + kFusedWaves - 1) / kFusedWaves;
  const unsigned grid = (unsigned)std::min<long long>(groups, 512);  // 2 resident workgroups per CU
  PredTables pt;
  for (int k = 0; k < kMlpOrders; ++k) pt.p[k] = out[k];
  hipLaunchKernelGGL(fused_predict_kernel, dim3(grid), dim3(64 * kFusedWaves), 0, stream, P, s.theta_bf,
                     s.theta, s.ztab, entries, pt);
  return hipGetLastError();
}

AtmosphereCode atmosphere_code(const double* rayleigh_base, double rayleigh_scale, double mie_base, double mie_scale,
                               double mie_g, double albedo) {
  auto unit = [](double v, double lo, double hi) { return (float)std::fmin(std::fmax((v - lo) / (hi - lo), 0.0), 1.0); };
  AtmosphereCode c;
  c.v[0] = unit(rayleigh_scale, 6000.0, 10000.0);
  c.v[1] = unit(rayleigh_base[0] / 5.8e-6, 0.5, 2.0);
  c.v[2] = unit(mie_scale, 800.0, 2000.0);
  c.v[3] = unit(mie_base, 5e-6, 5e-5);
  c.v[4] = unit(mie_g, 0.6, 0.85);
  c.v[5] = unit(albedo, 0.1, 0.5);
  return c;
}

}  // namespace sfsim
