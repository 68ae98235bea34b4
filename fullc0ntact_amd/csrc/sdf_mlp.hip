// Learned signed-distance surrogate (network in sdf_mlp.hpp): GPU training
// by this file's own forward/backward kernels (FP32, deterministic), the
// reference FP32 forward, and the BF16 MFMA inference kernel (FP32
// accumulation) used to cull far-field boundary queries.
#include "fc_internal.hpp"
#include "distmap.hpp"
#include "sdf_mlp.hpp"
#include "radix_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

struct fc_sdf {
  fc_context* ctx = nullptr;
  uint32_t B = 0;
  std::vector<float> hostFreq;   // B x 48
  std::vector<float> hostFrame;  // B x 4: center xyz, halfExtent
  float* params = nullptr;       // B x kSdfStride
  float* adamM = nullptr;
  float* adamV = nullptr;
  float* freq = nullptr;
  float* frame = nullptr;
  uint16_t* packed = nullptr;    // B x kSdfPackedBf16
  bool packedValid = false;
  uint32_t adamStep = 0;
  float* partial = nullptr;      // gradient partials
  size_t partialCap = 0;         // floats
  float* lossPartial = nullptr;
  size_t lossCap = 0;
  float* grad = nullptr;         // B x kSdfStride (reduced gradient)
  float* lossOut = nullptr;      // B
};

namespace fc {
// A rigid body of a boundary: world -> model rotation (row major), the
// world position of the model origin, and the body's distance map.
struct BodyDev {
  float w2m[9];
  float origin[3];
  int net;  // surrogate network of the body's shape
  DistMapView map;
};
}  // namespace fc

struct fc_body_set {
  fc_context* ctx = nullptr;
  uint32_t B = 0;
  int lattice[3] = {0, 0, 0};
  float latticeOrigin[3] = {0, 0, 0};
  float latticeCell = 0.0f;
  fc::BodyDev* bodies = nullptr;  // device, B
  int32_t maxNet = 0;              // largest network index of a body
  int32_t* cellBody = nullptr;    // device, lattice cells
  float* margins = nullptr;       // device, B
  uint32_t* counters = nullptr;   // device, 3
  uint16_t* netOf = nullptr;      // device, capacity: bucket of each sphere
  uint32_t* order = nullptr;      // device, capacity: spheres in bucket order
  uint32_t* band = nullptr;       // device, capacity: candidates of the cull
  uint32_t* bins = nullptr;       // device: spheres in a box, then per cull wave its candidates
  uint32_t* blockBins = nullptr;  // device: per bucketing block and bucket, count then start
  size_t blockBinsCap = 0;
  size_t capacity = 0;
};

namespace fc {

namespace {

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));

// Eight floats rounded to bfloat16 (round to nearest even, packed
// conversions) as one MFMA operand register group.
__device__ __forceinline__ bf16x8 pack8(float a0, float a1, float a2, float a3, float a4, float a5,
                                        float a6, float a7) {
  bf16v8 v;
  v[0] = (__bf16)a0;
  v[1] = (__bf16)a1;
  v[2] = (__bf16)a2;
  v[3] = (__bf16)a3;
  v[4] = (__bf16)a4;
  v[5] = (__bf16)a5;
  v[6] = (__bf16)a6;
  v[7] = (__bf16)a7;
  return __builtin_bit_cast(bf16x8, v);
}
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;  // samples per training tile
constexpr int kTB = 256;   // training / inference block
constexpr int kWPad = 10500;

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + __expf(-z)); }
// SiLU of the inference path: the hardware exp2 and reciprocal (about 1 ulp
// each) instead of the IEEE division of the training arithmetic.
__device__ __forceinline__ float silu_fast(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.4426950408889634f));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoidf_(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoidf_(z);
  return s * (1.0f + z * (1.0f - s));
}

// sin(2 pi t) and cos(2 pi t) with the hardware's revolution-based
// instructions after reducing t to [-1/2, 1/2].
__device__ __forceinline__ void sincos_rev(float t, float& s, float& c) {
  const float r = t - rintf(t);
  s = __builtin_amdgcn_sinf(r);
  c = __builtin_amdgcn_cosf(r);
}

// Normalised model-frame coordinates of a query for a body's frame.
__device__ __forceinline__ void sdf_norm(const float* fr, float qx, float qy, float qz, float& x,
                                         float& y, float& z) {
  x = (qx - fr[0]) / fr[3];
  y = (qy - fr[1]) / fr[3];
  z = (qz - fr[2]) / fr[3];
}

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// ---------------------------------------------------------------------------
// Reference FP32 forward, one lane per query.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kTB) void k_sdf_eval_f32(const float* __restrict__ params,
                                                      const float* __restrict__ freq,
                                                      const float* __restrict__ frame,
                                                      const uint32_t* __restrict__ bodies,
                                                      uint32_t numBodies,
                                                      const float* __restrict__ pts, uint32_t n,
                                                      float* __restrict__ out) {
  const uint32_t q = blockIdx.x * kTB + threadIdx.x;
  if (q >= n) return;
  const uint32_t b = bodies ? bodies[q] : 0u;
  if (b >= numBodies) {
    out[q] = __int_as_float(0x7fc00000);
    return;
  }
  const float* P = params + (size_t)b * kSdfStride;
  const float* F = freq + (size_t)b * kSdfFreqs * 3;
  const float* fr = frame + (size_t)b * 4;
  float x, y, z;
  sdf_norm(fr, pts[3 * q], pts[3 * q + 1], pts[3 * q + 2], x, y, z);
  float f[kSdfIn];
  for (int k = 0; k < kSdfFreqs; ++k) {
    const float t = F[3 * k] * x + F[3 * k + 1] * y + F[3 * k + 2] * z;
    sincos_rev(t, f[k], f[kSdfFreqs + k]);
  }
  float h[kSdfHidden], g[kSdfHidden];
  for (int o = 0; o < kSdfHidden; ++o) {
    float a = P[kSdfOffB1 + o];
    for (int i = 0; i < kSdfIn; ++i) a += f[i] * P[kSdfOffW1 + i * kSdfHidden + o];
    h[o] = silu(a);
  }
  for (int o = 0; o < kSdfHidden; ++o) {
    float a = P[kSdfOffB2 + o];
    for (int i = 0; i < kSdfHidden; ++i) a += h[i] * P[kSdfOffW2 + i * kSdfHidden + o];
    g[o] = silu(a);
  }
  float yv = P[kSdfOffB4];
  for (int o = 0; o < kSdfHidden; ++o) {
    float a = P[kSdfOffB3 + o];
    for (int i = 0; i < kSdfHidden; ++i) a += g[i] * P[kSdfOffW3 + i * kSdfHidden + o];
    yv += silu(a) * P[kSdfOffW4 + o];
  }
  out[q] = yv * fr[3];
}

// ---------------------------------------------------------------------------
// Training: forward + backward of 64-sample tiles of one body's minibatch.
// Block (blk, body) handles tiles blk*tilesPerBlock .. +tilesPerBlock-1 and
// writes the sum of their weight gradients (dL/dp of L = 1/M sum (y - t)^2,
// M = minibatch size) and of their squared errors.  Thread t owns the 4x4
// output tile (t / 16, t % 16) of every 64x64 product and the matching
// gradient entries, accumulated in registers over the block's tiles.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kTB) void k_sdf_grad_tiles(
    const float* __restrict__ params, const float* __restrict__ freq,
    const float* __restrict__ frame, const float* __restrict__ pts,
    const float* __restrict__ labels, uint32_t samplesPerBody, uint32_t batchStart,
    uint32_t batchSize, uint32_t tilesPerBlock, float invM, float* __restrict__ partial,
    float* __restrict__ lossPartial) {
  __shared__ float W[kWPad];
  __shared__ float X0[kTile * kSdfIn];
  __shared__ float Z1[kTile * kSdfHidden], A1[kTile * kSdfHidden];
  __shared__ float Z2[kTile * kSdfHidden], A2[kTile * kSdfHidden];
  __shared__ float Z3[kTile * kSdfHidden], A3[kTile * kSdfHidden];
  __shared__ float Y[kTile];
  __shared__ float FR[kSdfFreqs * 3 + 4];
  __shared__ float red[kTB / 64];

  const uint32_t body = blockIdx.y, blk = blockIdx.x, nblk = gridDim.x;
  const int t = threadIdx.x, rb = t >> 4, cb = t & 15;
  const float* P = params + (size_t)body * kSdfStride;
  for (int i = t; i < kSdfParams; i += kTB) W[i] = P[i];
  if (t < kSdfFreqs * 3) FR[t] = freq[(size_t)body * kSdfFreqs * 3 + t];
  if (t < 4) FR[kSdfFreqs * 3 + t] = frame[(size_t)body * 4 + t];

  float gW1[2][4] = {}, gW2[4][4] = {}, gW3[4][4] = {};
  float gW4 = 0.0f, gb1 = 0.0f, gb2 = 0.0f, gb3 = 0.0f, gb4 = 0.0f, lossAcc = 0.0f;
  const uint32_t batchEnd = batchStart + batchSize;

  for (uint32_t u = 0; u < tilesPerBlock; ++u) {
    const uint32_t tile = blk * tilesPerBlock + u;
    const uint32_t s0 = batchStart + tile * kTile;
    __syncthreads();
    // features
    for (int idx = t; idx < kTile * kSdfFreqs; idx += kTB) {
      const int s = idx / kSdfFreqs, k = idx % kSdfFreqs;
      const uint32_t gs = s0 + s;
      float sv = 0.0f, cv = 0.0f;
      if (gs < batchEnd && gs < samplesPerBody) {
        const float* q = pts + ((size_t)body * samplesPerBody + gs) * 3;
        float x, y, z;
        sdf_norm(FR + kSdfFreqs * 3, q[0], q[1], q[2], x, y, z);
        sincos_rev(FR[3 * k] * x + FR[3 * k + 1] * y + FR[3 * k + 2] * z, sv, cv);
      }
      X0[s * kSdfIn + k] = sv;
      X0[s * kSdfIn + kSdfFreqs + k] = cv;
    }
    __syncthreads();
    // ---- forward: Z = A W + b, A' = SiLU(Z), 4x4 tile per thread
    {
      float acc[4][4];
      for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) acc[a][c] = W[kSdfOffB1 + 4 * cb + c];
      for (int i = 0; i < kSdfIn; ++i) {
        float w[4], x[4];
        for (int c = 0; c < 4; ++c) w[c] = W[kSdfOffW1 + i * kSdfHidden + 4 * cb + c];
        for (int a = 0; a < 4; ++a) x[a] = X0[(4 * rb + a) * kSdfIn + i];
        for (int a = 0; a < 4; ++a)
          for (int c = 0; c < 4; ++c) acc[a][c] += x[a] * w[c];
      }
      for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) {
          const int e = (4 * rb + a) * kSdfHidden + 4 * cb + c;
          Z1[e] = acc[a][c];
          A1[e] = silu(acc[a][c]);
        }
    }
    __syncthreads();
    const float* Ain[2] = {A1, A2};
    float* Zout[2] = {Z2, Z3};
    float* Aout[2] = {A2, A3};
#pragma unroll
    for (int L = 0; L < 2; ++L) {
      const int offW = L == 0 ? kSdfOffW2 : kSdfOffW3, offB = L == 0 ? kSdfOffB2 : kSdfOffB3;
      float acc[4][4];
      for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) acc[a][c] = W[offB + 4 * cb + c];
      for (int i = 0; i < kSdfHidden; ++i) {
        float w[4], x[4];
        for (int c = 0; c < 4; ++c) w[c] = W[offW + i * kSdfHidden + 4 * cb + c];
        for (int a = 0; a < 4; ++a) x[a] = Ain[L][(4 * rb + a) * kSdfHidden + i];
        for (int a = 0; a < 4; ++a)
          for (int c = 0; c < 4; ++c) acc[a][c] += x[a] * w[c];
      }
      for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) {
          const int e = (4 * rb + a) * kSdfHidden + 4 * cb + c;
          Zout[L][e] = acc[a][c];
          Aout[L][e] = silu(acc[a][c]);
        }
      __syncthreads();
    }
    // ---- output and loss: sample s = t / 4, quarter t % 4 of the 64 terms
    {
      const int s = t >> 2, part = t & 3;
      float yv = 0.0f;
      for (int i = 16 * part; i < 16 * part + 16; ++i) yv += A3[s * kSdfHidden + i] * W[kSdfOffW4 + i];
      yv += __shfl_xor(yv, 1);
      yv += __shfl_xor(yv, 2);
      if (part == 0) {
        const uint32_t gs = s0 + s;
        float dy = 0.0f;
        if (gs < batchEnd && gs < samplesPerBody) {
          const float target = labels[(size_t)body * samplesPerBody + gs] / FR[kSdfFreqs * 3 + 3];
          const float err = (yv + W[kSdfOffB4]) - target;
          lossAcc += err * err;
          dy = 2.0f * err * invM;
        }
        Y[s] = dy;
      }
    }
    __syncthreads();
    // ---- backward
    // dW4, db4 (from A3, Y) before A3 is overwritten
    if (t < kSdfHidden) {
      float a = 0.0f;
      for (int s = 0; s < kTile; ++s) a += A3[s * kSdfHidden + t] * Y[s];
      gW4 += a;
    }
    if (t == 0) {
      float a = 0.0f;
      for (int s = 0; s < kTile; ++s) a += Y[s];
      gb4 += a;
    }
    // dZ3 = dy w4 SiLU'(Z3), in place in Z3
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 4; ++c) {
        const int s = 4 * rb + a, o = 4 * cb + c;
        const int e = s * kSdfHidden + o;
        Z3[e] = Y[s] * W[kSdfOffW4 + o] * silu_grad(Z3[e]);
      }
    __syncthreads();
    // layers 3 and 2: dW += A_in^T dZ, db += sum dZ, dZ_in = (dZ W^T) SiLU'(Z_in)
    float* dZ = Z3;
#pragma unroll
    for (int L = 1; L >= 0; --L) {
      const float* Aprev = L == 1 ? A2 : A1;
      const float* Zprev = L == 1 ? Z2 : Z1;
      const int offW = L == 1 ? kSdfOffW3 : kSdfOffW2;
      float(*gW)[4] = L == 1 ? gW3 : gW2;
      for (int s = 0; s < kTile; ++s) {
        float x[4], d[4];
        for (int a = 0; a < 4; ++a) x[a] = Aprev[s * kSdfHidden + 4 * rb + a];
        for (int c = 0; c < 4; ++c) d[c] = dZ[s * kSdfHidden + 4 * cb + c];
        for (int a = 0; a < 4; ++a)
          for (int c = 0; c < 4; ++c) gW[a][c] += x[a] * d[c];
      }
      if (t < kSdfHidden) {
        float a = 0.0f;
        for (int s = 0; s < kTile; ++s) a += dZ[s * kSdfHidden + t];
        if (L == 1) gb3 += a;
        else gb2 += a;
      }
      // dA_prev[s][i] = sum_o dZ[s][o] W[i][o]; tile (s-block rb, i-block cb)
      float acc[4][4] = {};
      for (int o = 0; o < kSdfHidden; ++o) {
        float d[4], w[4];
        for (int a = 0; a < 4; ++a) d[a] = dZ[(4 * rb + a) * kSdfHidden + o];
        for (int c = 0; c < 4; ++c) w[c] = W[offW + (4 * cb + c) * kSdfHidden + o];
        for (int a = 0; a < 4; ++a)
          for (int c = 0; c < 4; ++c) acc[a][c] += d[a] * w[c];
      }
      float* dst = L == 1 ? A3 : Z3;  // A3 is dead after dW4; Z3 (dZ3) after this layer
      __syncthreads();
      for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) {
          const int e = (4 * rb + a) * kSdfHidden + 4 * cb + c;
          dst[e] = acc[a][c] * silu_grad(Zprev[e]);
        }
      __syncthreads();
      dZ = dst;
    }
    // layer 1: dW1 += X0^T dZ1 (inputs 2rb, 2rb+1), db1
    for (int s = 0; s < kTile; ++s) {
      float x[2], d[4];
      for (int a = 0; a < 2; ++a) x[a] = X0[s * kSdfIn + 2 * rb + a];
      for (int c = 0; c < 4; ++c) d[c] = dZ[s * kSdfHidden + 4 * cb + c];
      for (int a = 0; a < 2; ++a)
        for (int c = 0; c < 4; ++c) gW1[a][c] += x[a] * d[c];
    }
    if (t < kSdfHidden) {
      float a = 0.0f;
      for (int s = 0; s < kTile; ++s) a += dZ[s * kSdfHidden + t];
      gb1 += a;
    }
  }
  // write this block's sums
  float* G = partial + ((size_t)body * nblk + blk) * kSdfStride;
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 4; ++c) G[kSdfOffW1 + (2 * rb + a) * kSdfHidden + 4 * cb + c] = gW1[a][c];
  for (int a = 0; a < 4; ++a)
    for (int c = 0; c < 4; ++c) {
      G[kSdfOffW2 + (4 * rb + a) * kSdfHidden + 4 * cb + c] = gW2[a][c];
      G[kSdfOffW3 + (4 * rb + a) * kSdfHidden + 4 * cb + c] = gW3[a][c];
    }
  if (t < kSdfHidden) {
    G[kSdfOffB1 + t] = gb1;
    G[kSdfOffB2 + t] = gb2;
    G[kSdfOffB3 + t] = gb3;
    G[kSdfOffW4 + t] = gW4;
  }
  if (t == 0) G[kSdfOffB4] = gb4;
  // block sum of squared errors (lanes 4s hold the sample terms)
  float l = lossAcc;
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if ((t & 63) == 0) red[t >> 6] = l;
  __syncthreads();
  if (t == 0) {
    float s = 0.0f;
    for (int w = 0; w < kTB / 64; ++w) s += red[w];
    lossPartial[(size_t)body * nblk + blk] = s;
  }
}

// Sum of the block partials per parameter (fixed order), then either the
// plain gradient (adam == 0) or an Adam step on the parameters.
__global__ __launch_bounds__(kTB) void k_sdf_reduce_step(
    const float* __restrict__ partial, const float* __restrict__ lossPartial, uint32_t nblk,
    uint32_t B, float* __restrict__ grad, float* __restrict__ loss, int adam,
    float* __restrict__ params, float* __restrict__ m, float* __restrict__ v, float lr,
    float corr1, float corr2, float invM) {
  const size_t idx = (size_t)blockIdx.x * kTB + threadIdx.x;
  if (idx >= (size_t)B * kSdfStride) return;
  const uint32_t b = (uint32_t)(idx / kSdfStride);
  const uint32_t p = (uint32_t)(idx % kSdfStride);
  if (p >= (uint32_t)kSdfParams) return;  // padding
  float g = 0.0f;
  for (uint32_t k = 0; k < nblk; ++k) g += partial[((size_t)b * nblk + k) * kSdfStride + p];
  if (p == 0) {
    float l = 0.0f;
    for (uint32_t k = 0; k < nblk; ++k) l += lossPartial[(size_t)b * nblk + k];
    loss[b] += l * invM;
  }
  if (!adam) {
    grad[idx] = g;
    return;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float mi = b1 * m[idx] + (1.0f - b1) * g;
  const float vi = b2 * v[idx] + (1.0f - b2) * g * g;
  m[idx] = mi;
  v[idx] = vi;
  params[idx] -= lr * (mi * corr1) / (sqrtf(vi * corr2) + eps);
}

// BF16 A-operand fragments of the inference kernel (see sdf_krow).
__global__ __launch_bounds__(kTB) void k_sdf_pack(const float* __restrict__ params, uint32_t B,
                                                  uint16_t* __restrict__ packed) {
  const size_t idx = (size_t)blockIdx.x * kTB + threadIdx.x;
  if (idx >= (size_t)B * kSdfPackedBf16) return;
  const uint32_t b = (uint32_t)(idx / kSdfPackedBf16);
  const uint32_t r = (uint32_t)(idx % kSdfPackedBf16);
  const int f = (int)(r / kSdfFragElems), lane = (int)(r % kSdfFragElems) / 8, j = (int)(r % 8);
  const int i = lane & 15, g = lane >> 4;
  int ob, kb, offW;
  if (f < 4) {
    ob = f;
    kb = 0;
    offW = kSdfOffW1;
  } else if (f < 12) {
    ob = (f - 4) >> 1;
    kb = (f - 4) & 1;
    offW = kSdfOffW2;
  } else {
    ob = (f - 12) >> 1;
    kb = (f - 12) & 1;
    offW = kSdfOffW3;
  }
  const int k = kb * 32 + sdf_krow(g, j), o = ob * 16 + i;
  packed[idx] = bf16_rne(params[(size_t)b * kSdfStride + offW + k * kSdfHidden + o]);
}

__device__ __forceinline__ bf16x8 load_frag(const uint16_t* __restrict__ packed, uint32_t body,
                                            int f, int lane) {
  const uint4 u = *reinterpret_cast<const uint4*>(packed + (size_t)body * kSdfPackedBf16 +
                                                  (size_t)f * kSdfFragElems + lane * 8);
  bf16x8 r;
  r[0] = (short)(u.x & 0xffff);
  r[1] = (short)(u.x >> 16);
  r[2] = (short)(u.y & 0xffff);
  r[3] = (short)(u.y >> 16);
  r[4] = (short)(u.z & 0xffff);
  r[5] = (short)(u.z >> 16);
  r[6] = (short)(u.w & 0xffff);
  r[7] = (short)(u.w >> 16);
  return r;
}

// Next layer's B operand from two 16-row accumulator blocks (after bias
// and SiLU): elements 0..3 from the first, 4..7 from the second.
__device__ __forceinline__ bf16x8 act_operand(const floatx4& lo, const floatx4& hi,
                                              const float4& blo, const float4& bhi) {
  return pack8(silu_fast(lo[0] + blo.x), silu_fast(lo[1] + blo.y), silu_fast(lo[2] + blo.z),
               silu_fast(lo[3] + blo.w), silu_fast(hi[0] + bhi.x), silu_fast(hi[1] + bhi.y),
               silu_fast(hi[2] + bhi.z), silu_fast(hi[3] + bhi.w));
}

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }

// One 16-query tile through the network with 16x16x32 BF16 MFMAs: the
// queries are the columns, the layer outputs the rows, so every lane group
// g = lane/16 holds rows 4g..4g+3 of each 16-row block for its query column
// lane%16.  Returns the network output y for the lane's column.
template <class Frags>
__device__ __forceinline__ float sdf_tile_mfma_f(const Frags& frag, const float* __restrict__ P,
                                                 const float* __restrict__ F, float x, float y,
                                                 float z, int g) {
  float sv[4], cv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = 4 * g + j;
    sincos_rev(F[3 * k] * x + F[3 * k + 1] * y + F[3 * k + 2] * z, sv[j], cv[j]);
  }
  const bf16x8 bop = pack8(sv[0], sv[1], sv[2], sv[3], cv[0], cv[1], cv[2], cv[3]);
  floatx4 d[4];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob)
    d[ob] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(frag(ob), bop, floatx4{0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
  for (int L = 0; L < 2; ++L) {
    const int offB = L == 0 ? kSdfOffB1 : kSdfOffB2;
    const int f0 = L == 0 ? 4 : 12;
    bf16x8 in0 = act_operand(d[0], d[1], ld4(P + offB + 4 * g), ld4(P + offB + 16 + 4 * g));
    bf16x8 in1 = act_operand(d[2], d[3], ld4(P + offB + 32 + 4 * g), ld4(P + offB + 48 + 4 * g));
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      floatx4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(frag(f0 + 2 * ob), in0,
                                                            floatx4{0, 0, 0, 0}, 0, 0, 0);
      d[ob] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(frag(f0 + 2 * ob + 1), in1, acc, 0, 0, 0);
    }
  }
  float yv = 0.0f;
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    const float4 bb = ld4(P + kSdfOffB3 + ob * 16 + 4 * g);
    const float4 ww = ld4(P + kSdfOffW4 + ob * 16 + 4 * g);
    yv += silu_fast(d[ob][0] + bb.x) * ww.x;
    yv += silu_fast(d[ob][1] + bb.y) * ww.y;
    yv += silu_fast(d[ob][2] + bb.z) * ww.z;
    yv += silu_fast(d[ob][3] + bb.w) * ww.w;
  }
  yv += __shfl_xor(yv, 16);
  yv += __shfl_xor(yv, 32);
  return yv + P[kSdfOffB4];
}

// The tile with the network's fragments held in registers.
__device__ __forceinline__ float sdf_tile_mfma(const bf16x8 (&frag)[kSdfFrags],
                                               const float* __restrict__ P,
                                               const float* __restrict__ F, float x, float y,
                                               float z, int g) {
  return sdf_tile_mfma_f([&](int f) { return frag[f]; }, P, F, x, y, z, g);
}

// The tile with each fragment loaded where it is used (the network's 20 KB
// of packed weights stay in the L1/L2 caches across the tiles of a wave),
// which frees the registers that hold them.
__device__ __forceinline__ float sdf_tile_mfma_loaded(const uint16_t* __restrict__ packed,
                                                      uint32_t net, int lane,
                                                      const float* __restrict__ P,
                                                      const float* __restrict__ F, float x,
                                                      float y, float z, int g) {
  return sdf_tile_mfma_f([&](int f) { return load_frag(packed, net, f, lane); }, P, F, x, y, z,
                         g);
}

// Wave-level inference over 16-query tiles (grid-stride).  Queries of one
// tile that belong to different bodies are run once per distinct body with
// that body's fragments; out = y * halfExtent (NaN for a body id out of
// range).
__global__ __launch_bounds__(kTB) void k_sdf_infer(const uint16_t* __restrict__ packed,
                                                   const float* __restrict__ params,
                                                   const float* __restrict__ freq,
                                                   const float* __restrict__ frame,
                                                   const uint32_t* __restrict__ bodies,
                                                   uint32_t numBodies,
                                                   const float* __restrict__ pts, uint32_t n,
                                                   float* __restrict__ out) {
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const uint32_t wave = (blockIdx.x * kTB + threadIdx.x) >> 6;
  const uint32_t nWaves = (gridDim.x * kTB) >> 6;
  const uint32_t tiles = (n + 15) / 16;
  // contiguous runs of tiles per wave: consecutive tiles share a network,
  // so the weight fragments are reloaded only where the network changes
  const uint32_t per = (tiles + nWaves - 1) / nWaves;
  const uint32_t t0 = wave * per, t1 = t0 + per < tiles ? t0 + per : tiles;
  bf16x8 frag[kSdfFrags];
  int cur = -1;
  for (uint32_t tile = t0; tile < t1; ++tile) {
    const uint32_t q = tile * 16 + col;
    const bool valid = q < n;
    uint32_t body = valid ? (bodies ? bodies[q] : 0u) : 0u;
    float qx = 0.0f, qy = 0.0f, qz = 0.0f;
    if (valid) {
      qx = pts[3 * q];
      qy = pts[3 * q + 1];
      qz = pts[3 * q + 2];
      if (body >= numBodies) {
        if (g == 0) out[q] = __int_as_float(0x7fc00000);
      }
    }
    bool remaining = valid && body < numBodies;
    while (true) {
      const uint64_t mask = __ballot(remaining);
      if (mask == 0) break;
      const int leadLane = __ffsll((unsigned long long)mask) - 1;
      const uint32_t lead = (uint32_t)__shfl((int)body, leadLane);
      if ((int)lead != cur) {
#pragma unroll
        for (int f = 0; f < kSdfFrags; ++f) frag[f] = load_frag(packed, lead, f, lane);
        cur = (int)lead;
      }
      const float* P = params + (size_t)lead * kSdfStride;
      const float* F = freq + (size_t)lead * kSdfFreqs * 3;
      const float* fr = frame + (size_t)lead * 4;
      float x, y, z;
      sdf_norm(fr, qx, qy, qz, x, y, z);
      const float yv = sdf_tile_mfma(frag, P, F, x, y, z, g);
      const bool mine = remaining && body == lead;
      if (mine && g == 0) out[q] = yv * fr[3];
      remaining = remaining && !mine;
    }
  }
}

// Lane-layout probe of the 16x16x32 BF16 MFMA with the K permutation of
// sdf_krow: D = A B for A 16x32 and B 32x16 row-major bf16, D 16x16 f32.
__global__ void k_mfma_probe(const uint16_t* __restrict__ A, const uint16_t* __restrict__ Bm,
                             float* __restrict__ D) {
  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15;
  bf16x8 a, b;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = sdf_krow(g, j);
    a[j] = (short)A[i * 32 + k];
    b[j] = (short)Bm[k * 16 + i];
  }
  const floatx4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, floatx4{0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
  for (int t = 0; t < 4; ++t) D[(4 * g + t) * 16 + i] = d[t];
}

// ---------------------------------------------------------------------------
// Boundary contacts of spheres with a set of rigid bodies.
// ---------------------------------------------------------------------------

// Body of the lattice cell holding p (-1: none), and p in that body's model
// frame, w2m (p - origin) with each row accumulated from zero
// (Matrix3x3 * Vector3, math/matrix3x3.h:313-321).
__device__ __forceinline__ int body_of(const BodyDev* __restrict__ bodies,
                                       const int32_t* __restrict__ cellBody, int lx, int ly,
                                       int lz, float ox, float oy, float oz, float cell,
                                       uint32_t B, v3f p, v3f& q) {
  const int cx = (int)floorf((p.x - ox) / cell), cy = (int)floorf((p.y - oy) / cell),
            cz = (int)floorf((p.z - oz) / cell);
  if (!(cx >= 0 && cx < lx && cy >= 0 && cy < ly && cz >= 0 && cz < lz)) return -1;
  const int b = cellBody[(cz * ly + cy) * lx + cx];
  if (b < 0 || (uint32_t)b >= B) return -1;
  const BodyDev& D = bodies[b];
  const float dx = p.x - D.origin[0], dy = p.y - D.origin[1], dz = p.z - D.origin[2];
  float r[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float s = 0.0f;
    s += D.w2m[3 * i + 0] * dx;
    s += D.w2m[3 * i + 1] * dy;
    s += D.w2m[3 * i + 2] * dz;
    r[i] = s;
  }
  q = make_v3(r[0], r[1], r[2]);
  return b;
}

__device__ __forceinline__ bool in_box(const DistMapView& m, v3f q) {
  return m.bmin[0] <= q.x && q.x <= m.bmax[0] && m.bmin[1] <= q.y && q.y <= m.bmax[1] &&
         m.bmin[2] <= q.z && q.z <= m.bmax[2];
}

struct LatticeArg {
  int lx, ly, lz;
  float ox, oy, oz, cell;
};

// Spheres are first grouped by the network of their body (one bucket per
// network plus one for spheres outside every box) so that each 16-sphere
// MFMA tile of the cull sees one network: count, scan, scatter, with plain
// LDS atomics.  The order inside a bucket is arbitrary (it follows the
// atomics); every sphere's result depends only on its own column of the
// tile, so the results do not.
// Bucketing blocks: 256 threads, 16 spheres per thread (few blocks, so the
// per-block bucket counts the scan walks stay few).
constexpr uint32_t kBT = 256;
constexpr uint32_t kBucketChunk = 16 * kBT;

// Also clears candidate[i] and contact[i] of every sphere (in index order;
// the band query sets them for the cull's candidates).  The spheres are
// taken kBucketBatch at a time with their loads issued together.
constexpr uint32_t kBucketBatch = 16;

__global__ __launch_bounds__(kBT) void k_bucket_count(
    const BodyDev* __restrict__ bodies, const int32_t* __restrict__ cellBody, LatticeArg L,
    uint32_t B, const float4* __restrict__ pos, uint32_t n, uint32_t nets,
    uint16_t* __restrict__ netOf, uint32_t* __restrict__ blockCounts,
    uint8_t* __restrict__ candidate, uint8_t* __restrict__ contact) {
  extern __shared__ uint32_t s_cnt[];
  for (uint32_t k = threadIdx.x; k <= nets; k += kBT) s_cnt[k] = 0u;
  __syncthreads();
  for (uint32_t u0 = 0; u0 < kBucketChunk / kBT; u0 += kBucketBatch) {
    float4 p[kBucketBatch];
#pragma unroll
    for (uint32_t u = 0; u < kBucketBatch; ++u) {
      const uint32_t i = blockIdx.x * kBucketChunk + (u0 + u) * kBT + threadIdx.x;
      p[u] = i < n ? pos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    uint32_t net[kBucketBatch];
#pragma unroll
    for (uint32_t u = 0; u < kBucketBatch; ++u) {
      const uint32_t i = blockIdx.x * kBucketChunk + (u0 + u) * kBT + threadIdx.x;
      net[u] = nets;  // nets = outside every box
      if (i < n) {
        v3f q;
        const int b = body_of(bodies, cellBody, L.lx, L.ly, L.lz, L.ox, L.oy, L.oz, L.cell, B,
                              xyz(p[u]), q);
        if (b >= 0 && in_box(bodies[b].map, q)) net[u] = (uint32_t)bodies[b].net;
        netOf[i] = (uint16_t)net[u];
        candidate[i] = 0;
        contact[i] = 0;
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < kBucketBatch; ++u) {
      const uint32_t i = blockIdx.x * kBucketChunk + (u0 + u) * kBT + threadIdx.x;
      if (i < n) atomicAdd(&s_cnt[net[u]], 1u);
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k <= nets; k += kBT)
    blockCounts[(size_t)k * gridDim.x + blockIdx.x] = s_cnt[k];
}

// Exclusive scan of the per-block bucket counts, laid out bucket by bucket
// (blockCounts[k * blocks + b]): one block per bucket k scans its row in
// chunks of kScanThreads coalesced entries, so that afterwards entry (k, b)
// is where block b's spheres start inside bucket k, and totals[k] holds the
// bucket's size.  No global atomics: a few hundred blocks adding to the same
// few dozen bucket counters serialise on them.
constexpr uint32_t kScanThreads = 1024;

__global__ __launch_bounds__(kScanThreads) void k_bucket_scan(uint32_t* __restrict__ blockCounts,
                                                              uint32_t blocks,
                                                              uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_scratch[kScanThreads / 64];
  __shared__ uint32_t s_chunk;
  uint32_t* row = blockCounts + (size_t)blockIdx.x * blocks;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < blocks; c0 += kScanThreads) {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t v = i < blocks ? row[i] : 0u;
    const uint32_t ex = scan_exclusive<kScanThreads>(v, s_scratch);
    if (i < blocks) row[i] = carry + ex;
    if (threadIdx.x == kScanThreads - 1) s_chunk = ex + v;
    __syncthreads();
    carry += s_chunk;
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Scatter of the spheres into bucket order: the bucket bases are the
// exclusive scan of the bucket totals (done by every block, in chunks of
// kBT buckets), plus the block's start inside each bucket; inBox (written by
// block 0) is the start of the last bucket, the spheres outside every box.
__global__ __launch_bounds__(kBT) void k_bucket_scatter(const uint16_t* __restrict__ netOf,
                                                        uint32_t n, uint32_t nets,
                                                        const uint32_t* __restrict__ blockStart,
                                                        const uint32_t* __restrict__ totals,
                                                        uint32_t* __restrict__ order,
                                                        uint32_t* __restrict__ inBox) {
  extern __shared__ uint32_t s_cnt[];
  uint32_t* s_base = s_cnt + nets + 1;
  __shared__ uint32_t s_scratch[kBT / 64];
  __shared__ uint32_t s_chunk;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 <= nets; c0 += kBT) {
    const uint32_t k = c0 + threadIdx.x;
    const uint32_t v = k <= nets ? totals[k] : 0u;
    const uint32_t ex = scan_exclusive<kBT>(v, s_scratch);
    if (k <= nets) s_base[k] = carry + ex + blockStart[(size_t)k * gridDim.x + blockIdx.x];
    if (k <= nets) s_cnt[k] = 0u;
    if (k == nets && blockIdx.x == 0) *inBox = carry + ex;
    if (threadIdx.x == kBT - 1) s_chunk = ex + v;
    __syncthreads();
    carry += s_chunk;
    __syncthreads();
  }
  const uint32_t base = blockIdx.x * kBucketChunk;
#pragma unroll
  for (uint32_t u = 0; u < kBucketChunk / kBT; ++u) {
    const uint32_t i = base + u * kBT + threadIdx.x;
    if (i < n) {
      const uint32_t net = (uint32_t)netOf[i];
      order[s_base[net] + atomicAdd(&s_cnt[net], 1u)] = i;
    }
  }
}

// Whether the cull loads each weight fragment where it is used (fewer
// registers, more waves per SIMD) instead of holding all 20 per wave.
#ifndef FC_CULL_LOADED_FRAGS
#define FC_CULL_LOADED_FRAGS 1
#endif
constexpr bool kCullLoadedFrags = FC_CULL_LOADED_FRAGS != 0;

// Surrogate cull over the bucketed spheres: the BF16 network predicts each
// in-box sphere's distance, and those within radius + margin[body] are
// listed for the exact query: wave w lists its candidates from band[16 t0]
// on (t0 = its first tile) and their number in waveCount[w].
__global__ __launch_bounds__(kTB) void k_boundary_cull(
    const uint16_t* __restrict__ packed, const float* __restrict__ params,
    const float* __restrict__ freq, const float* __restrict__ frame,
    const BodyDev* __restrict__ bodies, const int32_t* __restrict__ cellBody, LatticeArg L,
    uint32_t B, const float* __restrict__ margins, const float4* __restrict__ pos,
    const uint32_t* __restrict__ order, const uint32_t* __restrict__ inBoxCount, float radius,
    uint32_t* __restrict__ band, uint32_t* __restrict__ waveCount) {
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const uint32_t wave = (blockIdx.x * kTB + threadIdx.x) >> 6;
  const uint32_t nWaves = (gridDim.x * kTB) >> 6;
  // only the in-box buckets (the spheres outside every box come last and
  // were cleared by the bucketing), shared evenly among the waves
  const uint32_t n = *inBoxCount;
  const uint32_t tiles = (n + 15) / 16;
  // contiguous runs of tiles per wave: consecutive tiles share a network,
  // so the weight fragments are reloaded only where the network changes
  const uint32_t per = (tiles + nWaves - 1) / nWaves;
  const uint32_t t0 = wave * per, t1 = t0 + per < tiles ? t0 + per : tiles;
  bf16x8 frag[kSdfFrags];
  int cur = -1;
  // the wave's candidates are listed from band[16 t0] on (its slots' range)
  uint32_t listed = 0;
  for (uint32_t tile = t0; tile < t1; ++tile) {
    const uint32_t slot = tile * 16 + col;
    uint32_t i = 0;
    int body = -1, net = -1;
    v3f q = make_v3(0.0f, 0.0f, 0.0f);
    if (slot < n) {
      i = order[slot];
      body = body_of(bodies, cellBody, L.lx, L.ly, L.lz, L.ox, L.oy, L.oz, L.cell, B,
                     xyz(pos[i]), q);
      if (body >= 0 && !in_box(bodies[body].map, q)) body = -1;
      if (body >= 0) net = bodies[body].net;
    }
    bool cand = false;
    bool remaining = body >= 0;
    while (true) {
      const uint64_t mask = __ballot(remaining);
      if (mask == 0) break;
      const int leadLane = __ffsll((unsigned long long)mask) - 1;
      const int lead = __shfl(net, leadLane);
      if (!kCullLoadedFrags && lead != cur) {
#pragma unroll
        for (int f = 0; f < kSdfFrags; ++f) frag[f] = load_frag(packed, (uint32_t)lead, f, lane);
        cur = lead;
      }
      const float* P = params + (size_t)lead * kSdfStride;
      const float* F = freq + (size_t)lead * kSdfFreqs * 3;
      const float* fr = frame + (size_t)lead * 4;
      float x, y, z;
      sdf_norm(fr, q.x, q.y, q.z, x, y, z);
      const float yv = kCullLoadedFrags
                           ? sdf_tile_mfma_loaded(packed, (uint32_t)lead, lane, P, F, x, y, z, g)
                           : sdf_tile_mfma(frag, P, F, x, y, z, g);
      const bool mine = remaining && net == lead;
      if (mine) cand = yv * fr[3] <= radius + margins[body];
      remaining = remaining && !mine;
    }
    // candidates go to the wave's band list
    const bool listMe = slot < n && g == 0 && cand;
    const uint64_t cmask = __ballot(listMe);
    if (listMe) band[t0 * 16u + listed + (uint32_t)__popcll(cmask & ((1ull << lane) - 1ull))] = i;
    listed += (uint32_t)__popcll(cmask);
  }
  if (lane == 0) waveCount[wave] = listed;
}

// The exact query of the cull's candidates: block w walks the list of cull
// wave w (waveCount[w] entries from band[16 t0(w)] on, t0 as in the cull);
// contact[i] = 1 when the map distance is < radius.  counters (optional):
// [0] = spheres in a box, [1] += candidates, [2] += contacts.
constexpr uint32_t kBandThreads = 128;

__global__ __launch_bounds__(kBandThreads) void k_boundary_band(
    const BodyDev* __restrict__ bodies, const int32_t* __restrict__ cellBody, LatticeArg L,
    uint32_t B, const float4* __restrict__ pos, const uint32_t* __restrict__ band,
    const uint32_t* __restrict__ waveCount, uint32_t cullWaves,
    const uint32_t* __restrict__ inBoxCount, float radius, uint8_t* __restrict__ candidate,
    uint8_t* __restrict__ contact, uint32_t* __restrict__ counters) {
  const uint32_t w = blockIdx.x;
  const uint32_t tiles = (*inBoxCount + 15) / 16;
  const uint32_t per = (tiles + cullWaves - 1) / cullWaves;
  const uint32_t first = w * per * 16u;
  const uint32_t m = (w * per < tiles) ? waveCount[w] : 0u;
  if (counters && threadIdx.x == 0) {
    if (w == 0) counters[0] = *inBoxCount;
    if (m) atomicAdd(&counters[1], m);
  }
  for (uint32_t e = threadIdx.x; e < m; e += kBandThreads) {
    const uint32_t i = band[first + e];
    v3f q;
    const int b = body_of(bodies, cellBody, L.lx, L.ly, L.lz, L.ox, L.oy, L.oz, L.cell, B,
                          xyz(pos[i]), q);
    uint8_t c = 0;
    if (b >= 0) {  // always: the cull found the same body
      float d = 0.0f;
      v3f cp;
      distmap_query(bodies[b].map, q, d, cp, nullptr);
      c = d < radius ? 1 : 0;
    }
    candidate[i] = 1;
    contact[i] = c;
    if (counters && c) atomicAdd(&counters[2], 1u);
  }
}

// Exact contact test, one lane per sphere: contact[i] = 1 when the sphere's
// body's map distance is < radius (checkDistMap's test, boundaries.cuh:51).
// counters (optional): spheres in a box, exact queries, contacts.
__global__ __launch_bounds__(kTB) void k_boundary_exact(
    const BodyDev* __restrict__ bodies, const int32_t* __restrict__ cellBody, LatticeArg L,
    uint32_t B, const float4* __restrict__ pos, uint32_t n, float radius,
    uint8_t* __restrict__ contact, uint32_t* __restrict__ counters) {
  const uint32_t i = blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  v3f q;
  const int b = body_of(bodies, cellBody, L.lx, L.ly, L.lz, L.ox, L.oy, L.oz, L.cell, B,
                        xyz(pos[i]), q);
  const bool inside = b >= 0 && in_box(bodies[b].map, q);
  uint8_t c = 0;
  const bool query = inside;
  if (query) {
    float d = 0.0f;
    v3f cp;
    distmap_query(bodies[b].map, q, d, cp, nullptr);
    c = d < radius ? 1 : 0;
  }
  contact[i] = c;
  if (counters) {
    if (inside) atomicAdd(&counters[0], 1u);
    if (query) atomicAdd(&counters[1], 1u);
    if (c) atomicAdd(&counters[2], 1u);
  }
}

// Margin statistics of one network against one distance map, over the map's
// lattice vertices (the points where the map holds exact distances).  The
// lattice is walked in blocks of 4 x 2 x 2 vertices, one 16-vertex MFMA
// tile each (lane column c holds vertex (c & 3, (c >> 2) & 1, c >> 3) of the
// block); blocks with no vertex within radius + band of the surface are
// skipped.  The BF16 network is evaluated, as the cull evaluates it, and for
// the vertices v with d(v) < radius + band the kernel takes
//   stats[0] = max (y(v) - radius), stats[3] = max |y(v) - d(v)|,
//   stats[2] = the number of such vertices,
//   stats[1] = max |y(v') - y(v)| / |v' - v| over the x, y and z neighbours
//              v' of v in the same block (the network's largest secant
//              slope along the lattice edges at the band),
// as order-preserving integer keys of the floats (stats[2] a plain count).
// stats must start at {key(-inf), 0, 0, 0}.
__device__ __forceinline__ uint32_t float_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ __launch_bounds__(kTB) void k_sdf_map_margin(
    const uint16_t* __restrict__ packed, const float* __restrict__ params,
    const float* __restrict__ freq, const float* __restrict__ frame, uint32_t net,
    DistMapView m, float radius, float band, uint32_t* __restrict__ stats) {
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const uint32_t wave = (blockIdx.x * kTB + threadIdx.x) >> 6;
  const uint32_t nWaves = (gridDim.x * kTB) >> 6;
  const uint32_t nx = (uint32_t)m.cx + 1u, ny = (uint32_t)m.cy + 1u, nz = (uint32_t)m.cz + 1u;
  const uint32_t bx = (nx + 3u) / 4u, by = (ny + 1u) / 2u, bz = (nz + 1u) / 2u;
  const uint64_t blocks = (uint64_t)bx * by * bz;
  const uint32_t dx = (uint32_t)col & 3u, dy = ((uint32_t)col >> 2) & 1u, dz = (uint32_t)col >> 3;
  const float* P = params + (size_t)net * kSdfStride;
  const float* F = freq + (size_t)net * kSdfFreqs * 3;
  const float* fr = frame + (size_t)net * 4;
  bf16x8 frag[kSdfFrags];
#pragma unroll
  for (int f = 0; f < kSdfFrags; ++f) frag[f] = load_frag(packed, net, f, lane);
  float over = -INFINITY, slope = 0.0f, err = 0.0f;
  uint32_t count = 0;
  for (uint64_t blk = wave; blk < blocks; blk += nWaves) {
    const uint32_t kx = (uint32_t)(blk % bx), ky = (uint32_t)((blk / bx) % by),
                   kz = (uint32_t)(blk / ((uint64_t)bx * by));
    const uint32_t ix = 4u * kx + dx, iy = 2u * ky + dy, iz = 2u * kz + dz;
    const bool valid = ix < nx && iy < ny && iz < nz;
    float d = INFINITY, qx = 0.0f, qy = 0.0f, qz = 0.0f;
    if (valid) {
      d = m.cpd[((uint64_t)iz * ny + iy) * nx + ix].w;
      qx = m.xs[ix];
      qy = m.ys[iy];
      qz = m.zs[iz];
    }
    const bool inBand = valid && d < radius + band;
    if (__ballot(inBand) == 0ull) continue;
    float x, y, z;
    sdf_norm(fr, qx, qy, qz, x, y, z);
    const float yv = sdf_tile_mfma(frag, P, F, x, y, z, g) * fr[3];
    // neighbours along x (column + 1), y (+ 4) and z (+ 8) inside the block
    const int base = lane & 48;
    const int cX = (col + 1) & 15, cY = (col + 4) & 15, cZ = (col + 8) & 15;
    const float yX = __shfl(yv, base | cX), yY = __shfl(yv, base | cY), yZ = __shfl(yv, base | cZ);
    const float pX = __shfl(qx, base | cX), pY = __shfl(qy, base | cY), pZ = __shfl(qz, base | cZ);
    const bool vX = __shfl((int)valid, base | cX) != 0, vY = __shfl((int)valid, base | cY) != 0,
               vZ = __shfl((int)valid, base | cZ) != 0;
    const bool bX = __shfl((int)inBand, base | cX) != 0, bY = __shfl((int)inBand, base | cY) != 0,
               bZ = __shfl((int)inBand, base | cZ) != 0;
    if (g == 0 && valid) {
      if (inBand) {
        over = fmaxf(over, yv - radius);
        err = fmaxf(err, fabsf(yv - d));
        ++count;
      }
      if (dx < 3u && vX && (inBand || bX)) slope = fmaxf(slope, fabsf(yX - yv) / (pX - qx));
      if (dy == 0u && vY && (inBand || bY)) slope = fmaxf(slope, fabsf(yY - yv) / (pY - qy));
      if (dz == 0u && vZ && (inBand || bZ)) slope = fmaxf(slope, fabsf(yZ - yv) / (pZ - qz));
    }
  }
  // wave maxima, then one atomic per wave
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    over = fmaxf(over, __shfl_xor(over, off));
    slope = fmaxf(slope, __shfl_xor(slope, off));
    err = fmaxf(err, __shfl_xor(err, off));
    count += __shfl_xor(count, off);
  }
  if (lane == 0) {
    atomicMax(&stats[0], float_key(over));
    atomicMax(&stats[1], float_key(slope));
    atomicAdd(&stats[2], count);
    atomicMax(&stats[3], float_key(err));
  }
}

// SplitMix64 stream for the initial weights and frequencies.
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

fc_status ensure_partials(fc_sdf* m, size_t floats, size_t losses) {
  if (m->partialCap < floats) {
    FC_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    hipFree(m->partial);
    m->partial = nullptr;
    m->partialCap = 0;
    FC_HIP_TRY(hipMalloc(&m->partial, floats * sizeof(float)));
    m->partialCap = floats;
  }
  if (m->lossCap < losses) {
    FC_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    hipFree(m->lossPartial);
    m->lossPartial = nullptr;
    m->lossCap = 0;
    FC_HIP_TRY(hipMalloc(&m->lossPartial, losses * sizeof(float)));
    m->lossCap = losses;
  }
  return FC_SUCCESS;
}

void free_sdf(fc_sdf* m) {
  hipFree(m->params);
  hipFree(m->adamM);
  hipFree(m->adamV);
  hipFree(m->freq);
  hipFree(m->frame);
  hipFree(m->packed);
  hipFree(m->partial);
  hipFree(m->lossPartial);
  hipFree(m->grad);
  hipFree(m->lossOut);
}

constexpr uint32_t kTilesPerBlock = 8;
// Waves of the inference kernel (the 20 weight fragments in registers,
// 170 VGPRs: two waves per SIMD of the 256 CUs) and of the cull (fragments
// loaded where used, under 100 VGPRs: four waves per SIMD).
constexpr uint32_t kMaxInferWaves = 2048;
#ifndef FC_CULL_WAVES
#define FC_CULL_WAVES 4096
#endif
constexpr uint32_t kMaxCullWaves = FC_CULL_WAVES;

// Blocks (of 4 waves) of the MFMA kernels for n queries: at most the waves
// that are resident at once, each wave taking at least 16 tiles, so no wave
// waits for a second round.
uint32_t mfma_blocks(uint32_t n, uint32_t maxWaves = kMaxInferWaves) {
  const uint32_t tiles = (n + 15) / 16;
  uint32_t waves = (tiles + 15) / 16;
  if (waves > maxWaves) waves = maxWaves;
  const uint32_t blocks = (waves + 3) / 4;
  return blocks == 0 ? 1u : blocks;
}

// One gradient evaluation (+ optional Adam step) over samples
// [batchStart, batchStart + batchSize) of every body.
fc_status sdf_step(fc_sdf* m, const float* pts, const float* labels, uint32_t samplesPerBody,
                   uint32_t batchStart, uint32_t batchSize, int adam, float lr) {
  hipStream_t st = m->ctx->stream;
  const uint32_t tiles = (batchSize + kTile - 1) / kTile;
  const uint32_t nblk = (tiles + kTilesPerBlock - 1) / kTilesPerBlock;
  FC_TRY(ensure_partials(m, (size_t)m->B * nblk * kSdfStride, (size_t)m->B * nblk));
  const float invM = 1.0f / (float)batchSize;
  hipLaunchKernelGGL(k_sdf_grad_tiles, dim3(nblk, m->B), dim3(kTB), 0, st, m->params, m->freq,
                     m->frame, pts, labels, samplesPerBody, batchStart, batchSize,
                     kTilesPerBlock, invM, m->partial, m->lossPartial);
  FC_HIP_TRY(hipGetLastError());
  float c1 = 1.0f, c2 = 1.0f;
  if (adam) {
    ++m->adamStep;
    c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)m->adamStep)));
    c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)m->adamStep)));
    m->packedValid = false;
  }
  const size_t total = (size_t)m->B * kSdfStride;
  hipLaunchKernelGGL(k_sdf_reduce_step, dim3((unsigned)((total + kTB - 1) / kTB)), dim3(kTB), 0,
                     st, m->partial, m->lossPartial, nblk, m->B, m->grad, m->lossOut, adam,
                     m->params, m->adamM, m->adamV, lr, c1, c2, invM);
  FC_HIP_TRY(hipGetLastError());
  return FC_SUCCESS;
}

fc_status sdf_pack(fc_sdf* m) {
  if (m->packedValid) return FC_SUCCESS;
  const size_t total = (size_t)m->B * kSdfPackedBf16;
  hipLaunchKernelGGL(k_sdf_pack, dim3((unsigned)((total + kTB - 1) / kTB)), dim3(kTB), 0,
                     m->ctx->stream, m->params, m->B, m->packed);
  FC_HIP_TRY(hipGetLastError());
  m->packedValid = true;
  return FC_SUCCESS;
}

}  // namespace

}  // namespace fc

using namespace fc;

extern "C" {

uint32_t fc_sdf_num_params(void) { return (uint32_t)kSdfParams; }
uint32_t fc_sdf_param_stride(void) { return (uint32_t)kSdfStride; }

fc_status fc_sdf_create(fc_context* ctx, uint32_t numBodies, const float* centers,
                        const float* halfExtents, float freqSigma, uint64_t seed,
                        fc_sdf** out) {
  if (!ctx || !centers || !halfExtents || !out)
    return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  *out = nullptr;
  if (numBodies == 0) return set_error(FC_ERROR_INVALID_VALUE, "numBodies must be >= 1");
  for (uint32_t b = 0; b < numBodies; ++b)
    if (!(halfExtents[b] > 0.0f)) return set_error(FC_ERROR_INVALID_VALUE, "halfExtent must be > 0");
  fc_sdf* m = new fc_sdf();
  m->ctx = ctx;
  m->B = numBodies;
  m->hostFreq.resize((size_t)numBodies * kSdfFreqs * 3);
  m->hostFrame.resize((size_t)numBodies * 4);
  std::vector<float> P((size_t)numBodies * kSdfStride, 0.0f);
  Rng rng{seed};
  for (uint32_t b = 0; b < numBodies; ++b) {
    for (int k = 0; k < kSdfFreqs * 3; k += 2) {  // Box-Muller
      const double u1 = 1.0 - rng.u01(), u2 = rng.u01();
      const double rad = std::sqrt(-2.0 * std::log(u1));
      m->hostFreq[(size_t)b * kSdfFreqs * 3 + k] = (float)(freqSigma * rad * std::cos(6.283185307179586 * u2));
      if (k + 1 < kSdfFreqs * 3)
        m->hostFreq[(size_t)b * kSdfFreqs * 3 + k + 1] =
            (float)(freqSigma * rad * std::sin(6.283185307179586 * u2));
    }
    for (int a = 0; a < 3; ++a) m->hostFrame[(size_t)b * 4 + a] = centers[3 * b + a];
    m->hostFrame[(size_t)b * 4 + 3] = halfExtents[b];
    float* p = P.data() + (size_t)b * kSdfStride;
    auto glorot = [&](int off, int fin, int fout) {
      const double lim = std::sqrt(6.0 / (double)(fin + fout));
      for (int i = 0; i < fin * fout; ++i) p[off + i] = (float)((2.0 * rng.u01() - 1.0) * lim);
    };
    glorot(kSdfOffW1, kSdfIn, kSdfHidden);
    glorot(kSdfOffW2, kSdfHidden, kSdfHidden);
    glorot(kSdfOffW3, kSdfHidden, kSdfHidden);
    glorot(kSdfOffW4, kSdfHidden, 1);
  }
  const size_t np = (size_t)numBodies * kSdfStride;
  hipStream_t st = ctx->stream;
  hipError_t e = hipMalloc(&m->params, np * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->adamM, np * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->adamV, np * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->grad, np * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->freq, m->hostFreq.size() * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->frame, m->hostFrame.size() * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->packed, (size_t)numBodies * kSdfPackedBf16 * 2);
  if (e == hipSuccess) e = hipMalloc(&m->lossOut, numBodies * sizeof(float));
  if (e == hipSuccess) e = hipMemcpyAsync(m->params, P.data(), np * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(m->adamM, 0, np * 4, st);
  if (e == hipSuccess) e = hipMemsetAsync(m->adamV, 0, np * 4, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(m->freq, m->hostFreq.data(), m->hostFreq.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(m->frame, m->hostFrame.data(), m->hostFrame.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    free_sdf(m);
    delete m;
    return hip_error(e, "sdf allocation");
  }
  *out = m;
  return FC_SUCCESS;
}

fc_status fc_sdf_destroy(fc_sdf* m) {
  if (!m) return FC_SUCCESS;
  // No use of m->ctx: the context may already be gone; hipFree waits for
  // the device.
  free_sdf(m);
  delete m;
  return FC_SUCCESS;
}

fc_status fc_sdf_get_params(fc_sdf* m, float* params, float* freqs, float* frames) {
  if (!m) return set_error(FC_ERROR_INVALID_VALUE, "sdf is NULL");
  if (params) {
    FC_HIP_TRY(hipMemcpyAsync(params, m->params, (size_t)m->B * kSdfStride * 4,
                              hipMemcpyDeviceToHost, m->ctx->stream));
    FC_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
  }
  if (freqs) std::memcpy(freqs, m->hostFreq.data(), m->hostFreq.size() * 4);
  if (frames) std::memcpy(frames, m->hostFrame.data(), m->hostFrame.size() * 4);
  return FC_SUCCESS;
}

fc_status fc_sdf_set_params(fc_sdf* m, const float* params) {
  if (!m || !params) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  FC_HIP_TRY(hipMemcpyAsync(m->params, params, (size_t)m->B * kSdfStride * 4,
                            hipMemcpyHostToDevice, m->ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
  m->packedValid = false;
  return FC_SUCCESS;
}

fc_status fc_sdf_loss_grad(fc_sdf* m, const float* points, const float* labels,
                           uint32_t samplesPerBody, float* grad, float* loss) {
  if (!m || !points || !labels) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (samplesPerBody == 0) return set_error(FC_ERROR_INVALID_VALUE, "no samples");
  FC_HIP_TRY(hipMemsetAsync(m->lossOut, 0, m->B * sizeof(float), m->ctx->stream));
  FC_TRY(sdf_step(m, points, labels, samplesPerBody, 0, samplesPerBody, 0, 0.0f));
  if (grad)
    FC_HIP_TRY(hipMemcpyAsync(grad, m->grad, (size_t)m->B * kSdfStride * 4,
                              hipMemcpyDeviceToHost, m->ctx->stream));
  if (loss)
    FC_HIP_TRY(hipMemcpyAsync(loss, m->lossOut, m->B * sizeof(float), hipMemcpyDeviceToHost,
                              m->ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
  return FC_SUCCESS;
}

fc_status fc_sdf_train(fc_sdf* m, const float* points, const float* labels,
                       uint32_t samplesPerBody, uint32_t batchPerBody, uint32_t epochs, float lr,
                       float lrDecay, float* lossHistory) {
  if (!m || !points || !labels) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (samplesPerBody == 0 || batchPerBody == 0)
    return set_error(FC_ERROR_INVALID_VALUE, "empty training set or batch");
  const uint32_t batches = (samplesPerBody + batchPerBody - 1) / batchPerBody;
  std::vector<float> l(m->B);
  float rate = lr;
  for (uint32_t e = 0; e < epochs; ++e, rate *= lrDecay) {
    FC_HIP_TRY(hipMemsetAsync(m->lossOut, 0, m->B * sizeof(float), m->ctx->stream));
    for (uint32_t k = 0; k < batches; ++k) {
      const uint32_t start = k * batchPerBody;
      const uint32_t size = std::min(batchPerBody, samplesPerBody - start);
      FC_TRY(sdf_step(m, points, labels, samplesPerBody, start, size, 1, rate));
    }
    if (lossHistory) {
      FC_HIP_TRY(hipMemcpyAsync(l.data(), m->lossOut, m->B * sizeof(float),
                                hipMemcpyDeviceToHost, m->ctx->stream));
      FC_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
      for (uint32_t b = 0; b < m->B; ++b) lossHistory[(size_t)e * m->B + b] = l[b] / (float)batches;
    }
  }
  FC_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
  return FC_SUCCESS;
}

fc_status fc_sdf_eval(fc_sdf* m, const uint32_t* bodies, const float* points, uint32_t n,
                      float* out, int precision) {
  if (!m) return set_error(FC_ERROR_INVALID_VALUE, "sdf is NULL");
  if (n == 0) return FC_SUCCESS;
  if (!points || !out) return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  hipStream_t st = m->ctx->stream;
  if (precision == FC_SDF_FP32) {
    hipLaunchKernelGGL(k_sdf_eval_f32, dim3((n + kTB - 1) / kTB), dim3(kTB), 0, st, m->params,
                       m->freq, m->frame, bodies, m->B, points, n, out);
  } else if (precision == FC_SDF_BF16) {
    FC_TRY(sdf_pack(m));
    const uint32_t blocks = mfma_blocks(n);
    hipLaunchKernelGGL(k_sdf_infer, dim3(blocks), dim3(kTB), 0, st, m->packed, m->params,
                       m->freq, m->frame, bodies, m->B, points, n, out);
  } else {
    return set_error(FC_ERROR_INVALID_VALUE, "precision must be FC_SDF_FP32 or FC_SDF_BF16");
  }
  FC_HIP_TRY(hipGetLastError());
  return FC_SUCCESS;
}

fc_status fc_sdf_map_margin(fc_sdf* m, uint32_t net, const fc_distmap* map, float radius,
                            float band, float out[4]) {
  if (!m || !map || !out) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (net >= m->B) return set_error(FC_ERROR_INVALID_VALUE, "network index out of range");
  if (!(band >= 0.0f)) return set_error(FC_ERROR_INVALID_VALUE, "band must be >= 0");
  FC_TRY(sdf_pack(m));
  hipStream_t st = m->ctx->stream;
  uint32_t* d = nullptr;
  FC_HIP_TRY(hipMalloc(&d, 4 * sizeof(uint32_t)));
  const uint32_t init[4] = {0x007fffffu /* key of -inf */, 0u, 0u, 0u};
  hipError_t e = hipMemcpyAsync(d, init, sizeof(init), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sdf_map_margin, dim3(2048), dim3(kTB), 0, st, m->packed, m->params,
                       m->freq, m->frame, net, distmap_view(map), radius, band, d);
    e = hipGetLastError();
  }
  uint32_t h[4] = {0, 0, 0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  hipFree(d);
  if (e != hipSuccess) return hip_error(e, "sdf map margin");
  auto unkey = [](uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  };
  out[0] = unkey(h[0]);
  out[1] = unkey(h[1]);
  out[2] = (float)h[2];
  out[3] = unkey(h[3]);
  return FC_SUCCESS;
}

fc_status fc_body_set_create(fc_context* ctx, uint32_t numBodies, const fc_distmap* const* maps,
                             const float* model2world, const float* origins,
                             const int32_t* bodyNet, const int32_t lattice[3],
                             const float latticeOrigin[3], float latticeCell,
                             const int32_t* cellBody, fc_body_set** out) {
  if (!ctx || !maps || !model2world || !origins || !lattice || !latticeOrigin || !cellBody || !out)
    return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  *out = nullptr;
  if (numBodies == 0) return set_error(FC_ERROR_INVALID_VALUE, "numBodies must be >= 1");
  if (lattice[0] < 1 || lattice[1] < 1 || lattice[2] < 1 || !(latticeCell > 0.0f))
    return set_error(FC_ERROR_INVALID_VALUE, "invalid body lattice");
  const size_t cells = (size_t)lattice[0] * lattice[1] * lattice[2];
  std::vector<BodyDev> host(numBodies);
  int32_t maxNet = 0;
  for (uint32_t b = 0; b < numBodies; ++b) {
    if (!maps[b]) return set_error(FC_ERROR_INVALID_VALUE, "a body has no distance map");
    host[b].net = bodyNet ? bodyNet[b] : (int)b;
    if (host[b].net < 0) return set_error(FC_ERROR_INVALID_VALUE, "negative network index");
    maxNet = std::max(maxNet, host[b].net);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) host[b].w2m[3 * i + j] = model2world[9 * b + 3 * j + i];
    for (int a = 0; a < 3; ++a) host[b].origin[a] = origins[3 * b + a];
    host[b].map = distmap_view(maps[b]);
  }
  for (size_t c = 0; c < cells; ++c)
    if (cellBody[c] >= (int32_t)numBodies)
      return set_error(FC_ERROR_INVALID_VALUE, "cellBody names a body that does not exist");
  fc_body_set* s = new fc_body_set();
  s->ctx = ctx;
  s->B = numBodies;
  for (int a = 0; a < 3; ++a) {
    s->lattice[a] = lattice[a];
    s->latticeOrigin[a] = latticeOrigin[a];
  }
  s->latticeCell = latticeCell;
  s->maxNet = maxNet;
  hipStream_t st = ctx->stream;
  hipError_t e = hipMalloc(&s->bodies, numBodies * sizeof(BodyDev));
  if (e == hipSuccess) e = hipMalloc(&s->cellBody, cells * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&s->margins, numBodies * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&s->counters, 3 * sizeof(uint32_t));
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->bodies, host.data(), numBodies * sizeof(BodyDev), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->cellBody, cellBody, cells * sizeof(int32_t), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    hipFree(s->bodies);
    hipFree(s->cellBody);
    hipFree(s->margins);
    hipFree(s->counters);
    delete s;
    return hip_error(e, "body set allocation");
  }
  *out = s;
  return FC_SUCCESS;
}

fc_status fc_body_set_destroy(fc_body_set* s) {
  if (!s) return FC_SUCCESS;
  hipFree(s->netOf);
  hipFree(s->order);
  hipFree(s->band);
  hipFree(s->bins);
  hipFree(s->blockBins);
  hipFree(s->bodies);
  hipFree(s->cellBody);
  hipFree(s->margins);
  hipFree(s->counters);
  delete s;
  return FC_SUCCESS;
}

fc_status fc_boundary_contacts(fc_context* ctx, fc_body_set* set, fc_sdf* surrogate,
                               const float* margins, const float* pos, uint32_t n, float radius,
                               uint8_t* candidate, uint8_t* contact, uint32_t counts[3]) {
  if (!ctx || !set) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (n == 0) {
    if (counts) counts[0] = counts[1] = counts[2] = 0;
    return FC_SUCCESS;
  }
  if (!pos || !contact) return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  hipStream_t st = ctx->stream;
  const LatticeArg L{set->lattice[0], set->lattice[1], set->lattice[2], set->latticeOrigin[0],
                     set->latticeOrigin[1], set->latticeOrigin[2], set->latticeCell};
  StageTimer* timer = ctx->timer.enabled ? &ctx->timer : nullptr;
  if (surrogate) {
    if ((uint32_t)set->maxNet >= surrogate->B)
      return set_error(FC_ERROR_INVALID_VALUE, "a body names a network the surrogate lacks");
    if (!candidate) return set_error(FC_ERROR_INVALID_VALUE, "the surrogate path needs candidate");
    std::vector<float> m(set->B, 0.0f);
    if (margins) std::memcpy(m.data(), margins, set->B * sizeof(float));
    FC_HIP_TRY(hipMemcpyAsync(set->margins, m.data(), set->B * sizeof(float),
                              hipMemcpyHostToDevice, st));
    FC_TRY(sdf_pack(surrogate));
    const uint32_t nets = (uint32_t)set->maxNet + 1;
    if (nets > 4095u) return set_error(FC_ERROR_CAPACITY, "at most 4095 networks");
    if (set->capacity < n) {
      FC_HIP_TRY(hipStreamSynchronize(st));
      hipFree(set->netOf);
      hipFree(set->order);
      hipFree(set->band);
      set->netOf = nullptr;
      set->order = nullptr;
      set->band = nullptr;
      set->capacity = 0;
      FC_HIP_TRY(hipMalloc(&set->netOf, (size_t)n * sizeof(uint16_t)));
      FC_HIP_TRY(hipMalloc(&set->order, (size_t)n * sizeof(uint32_t)));
      FC_HIP_TRY(hipMalloc(&set->band, (size_t)n * sizeof(uint32_t)));
      set->capacity = n;
    }
    if (!set->bins) FC_HIP_TRY(hipMalloc(&set->bins, (1 + kMaxCullWaves) * sizeof(uint32_t)));
    uint32_t* inBox = set->bins;
    uint32_t* waveCount = set->bins + 1;
    const unsigned blocks = (n + kBucketChunk - 1) / kBucketChunk;
    const size_t entries = (size_t)(nets + 1) * blocks + (nets + 1);  // + bucket totals
    if (set->blockBinsCap < entries) {
      FC_HIP_TRY(hipStreamSynchronize(st));
      hipFree(set->blockBins);
      set->blockBins = nullptr;
      set->blockBinsCap = 0;
      FC_HIP_TRY(hipMalloc(&set->blockBins, entries * sizeof(uint32_t)));
      set->blockBinsCap = entries;
    }
    if (counts) FC_HIP_TRY(hipMemsetAsync(set->counters, 0, 3 * sizeof(uint32_t), st));
    if (timer) timer->begin(st, "boundary_bucket");
    hipLaunchKernelGGL(k_bucket_count, dim3(blocks), dim3(kBT), (nets + 1) * sizeof(uint32_t), st,
                       set->bodies, set->cellBody, L, set->B, (const float4*)pos, n, nets,
                       set->netOf, set->blockBins, candidate, contact);
    uint32_t* totals = set->blockBins + (size_t)(nets + 1) * blocks;
    hipLaunchKernelGGL(k_bucket_scan, dim3(nets + 1), dim3(kScanThreads), 0, st, set->blockBins,
                       (uint32_t)blocks, totals);
    hipLaunchKernelGGL(k_bucket_scatter, dim3(blocks), dim3(kBT), 2 * (nets + 1) * sizeof(uint32_t),
                       st, set->netOf, n, nets, set->blockBins, totals, set->order, inBox);
    if (timer) timer->end(st);
    FC_HIP_TRY(hipGetLastError());
    const uint32_t cblocks = mfma_blocks(n, kMaxCullWaves);
    if (timer) timer->begin(st, "boundary_cull");
    hipLaunchKernelGGL(k_boundary_cull, dim3(cblocks), dim3(kTB), 0, st, surrogate->packed,
                       surrogate->params, surrogate->freq, surrogate->frame, set->bodies,
                       set->cellBody, L, set->B, set->margins, (const float4*)pos, set->order,
                       inBox, radius, set->band, waveCount);
    if (timer) timer->end(st);
    FC_HIP_TRY(hipGetLastError());
    if (timer) timer->begin(st, "boundary_exact_band");
    hipLaunchKernelGGL(k_boundary_band, dim3(cblocks * 4), dim3(kBandThreads), 0, st,
                       set->bodies, set->cellBody, L, set->B, (const float4*)pos, set->band,
                       waveCount, cblocks * 4, inBox, radius, candidate, contact,
                       counts ? set->counters : nullptr);
    if (timer) timer->end(st);
    FC_HIP_TRY(hipGetLastError());
  } else {
    if (counts) FC_HIP_TRY(hipMemsetAsync(set->counters, 0, 3 * sizeof(uint32_t), st));
    if (timer) timer->begin(st, "boundary_exact");
    hipLaunchKernelGGL(k_boundary_exact, dim3((n + kTB - 1) / kTB), dim3(kTB), 0, st,
                       set->bodies, set->cellBody, L, set->B, (const float4*)pos, n, radius,
                       contact, counts ? set->counters : nullptr);
    if (timer) timer->end(st);
    FC_HIP_TRY(hipGetLastError());
  }
  if (counts) {
    FC_HIP_TRY(hipMemcpyAsync(counts, set->counters, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    FC_HIP_TRY(hipStreamSynchronize(st));
  }
  return FC_SUCCESS;
}

fc_status fc_mfma_probe(fc_context* ctx, const uint16_t* a, const uint16_t* b, float* d) {
  if (!ctx || !a || !b || !d) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, ctx->stream, a, b, d);
  FC_HIP_TRY(hipGetLastError());
  FC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return FC_SUCCESS;
}

}  // extern "C"
