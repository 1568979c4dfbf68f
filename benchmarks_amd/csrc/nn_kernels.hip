// nn_kernels.hip -- kernels of the learned preconditioner's 3D U-Net: the
// convolutions of the forward pass on BF16 MFMA with FP32 accumulation, the
// backward pass (data and weight gradients) in FP32, the loss and Adam.
//
// Tensors are channels-last [z][y][x][c] over a level's nodal grid.  A
// convolution is a correlation with zero padding 1 over the 27 taps
// t = tx + 3 ty + 9 tz (offset (tx-1, ty-1, tz-1)):
//   out[p][o] = sum_{t,c} W[o][t][c] in[s p + off_t][c]      (stride s = 1, 2)
// and, with up-sampling, in = up(v), up(v)[i] = v[i >> 1] per direction.
//
// Forward (implicit GEMM on v_mfma_f32_16x16x32_bf16): a wavefront owns 16
// consecutive output nodes and all output channels; the K dimension
// (t, c) runs in chunks of 32: lane l gathers 8 consecutive channels of one
// tap of node l % 16 (one 16-byte load, channels-last) as the A operand and
// 8 consecutive k of output channel l % 16 of the weights [o][k] as the B
// operand.  D register r of lane l is node 4 (l / 16) + r, channel l % 16.
#include <algorithm>

#include "mfx_internal.hpp"
#include "nn.hpp"

namespace mfx {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kBlock = 256;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct Dims {
  uint32_t x, y, z;
};

__device__ __forceinline__ void decode(uint64_t i, const Dims &d, int &x, int &y, int &z) {
  x = (int)(i % d.x);
  const uint64_t q = i / d.x;
  y = (int)(q % d.y);
  z = (int)(q / d.y);
}

__device__ __forceinline__ bool inside(int x, int y, int z, const Dims &d) {
  return x >= 0 && y >= 0 && z >= 0 && x < (int)d.x && y < (int)d.y && z < (int)d.z;
}

__device__ __forceinline__ uint64_t lin(int x, int y, int z, const Dims &d) {
  return (uint64_t)x + (uint64_t)d.x * ((uint64_t)y + (uint64_t)d.y * (uint64_t)z);
}

// Index of the input node of output node (x,y,z) at tap t, or -1 outside.
// od: output grid, id: input tensor grid (the coarse grid when UP).
template <int STRIDE, bool UP>
__device__ __forceinline__ int64_t tap_source(int x, int y, int z, int t, const Dims &od,
                                              const Dims &id) {
  const int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = t / 9 - 1;
  int qx = STRIDE * x + ox, qy = STRIDE * y + oy, qz = STRIDE * z + oz;
  if (UP) {
    if (!inside(qx, qy, qz, od)) return -1;
    return (int64_t)lin(qx >> 1, qy >> 1, qz >> 1, id);
  }
  if (!inside(qx, qy, qz, id)) return -1;
  return (int64_t)lin(qx, qy, qz, id);
}

template <int CIN, int COUT, int STRIDE, bool UP, bool SKIP>
__global__ void __launch_bounds__(kBlock)
k_conv_mfma(const uint16_t *__restrict__ in, const uint16_t *__restrict__ W,
            const uint16_t *__restrict__ skip, uint16_t *__restrict__ out, Dims od, Dims id) {
  static_assert(CIN % 16 == 0 && COUT % 16 == 0, "channel counts must be multiples of 16");
  constexpr int K = 27 * CIN, KP = (K + 31) / 32 * 32, NT = COUT / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t P = (uint64_t)od.x * od.y * od.z;
  const uint64_t base = ((uint64_t)blockIdx.x * (kBlock / 64) + wave) * 16;
  if (base >= P) return;
  const uint64_t node = base + (lane & 15);
  int x = 0, y = 0, z = 0;
  const bool valid = node < P;
  if (valid) decode(node, od, x, y, z);
  const int kl = 8 * (lane >> 4);
  v4f acc[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) acc[n] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int kc = 0; kc < KP; kc += 32) {
    const int k = kc + kl;
    uint4 araw = make_uint4(0, 0, 0, 0);
    if (valid && k < K) {
      const int t = k / CIN, c = k - t * CIN;
      const int64_t src = tap_source<STRIDE, UP>(x, y, z, t, od, id);
      if (src >= 0) araw = *reinterpret_cast<const uint4 *>(in + (uint64_t)src * CIN + c);
    }
    const v8bf a = __builtin_bit_cast(v8bf, araw);
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const uint4 braw =
          *reinterpret_cast<const uint4 *>(W + (uint64_t)(n * 16 + (lane & 15)) * KP + k);
      acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(v8bf, braw), acc[n], 0,
                                                       0, 0);
    }
  }
  // epilogue: node base + 4 (lane / 16) + r, channel n * 16 + lane % 16
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint64_t o = base + 4 * (lane >> 4) + r;
      if (o >= P) continue;
      const int ch = n * 16 + (lane & 15);
      float v = acc[n][r];
      if (SKIP) v += bf2f(skip[o * COUT + ch]);
      out[o * COUT + ch] = f2bf(fmaxf(v, 0.f));
    }
}

// first layer: one FP32 input channel -> 16, ReLU; thread per node
__global__ void __launch_bounds__(kBlock)
k_conv_first(const float *__restrict__ in, const float *__restrict__ W, uint16_t *__restrict__ out,
             Dims d) {
  const uint64_t P = (uint64_t)d.x * d.y * d.z;
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= P) return;
  int x, y, z;
  decode(i, d, x, y, z);
  float v[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) {
    const int64_t s = tap_source<1, false>(x, y, z, t, d, d);
    v[t] = s >= 0 ? in[s] : 0.f;
  }
  uint16_t o[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    float a = 0.f;
#pragma unroll
    for (int t = 0; t < 27; ++t) a += W[c * 27 + t] * v[t];
    o[c] = f2bf(fmaxf(a, 0.f));
  }
  uint32_t pk[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) pk[c] = (uint32_t)o[2 * c] | ((uint32_t)o[2 * c + 1] << 16);
  uint4 *dst = reinterpret_cast<uint4 *>(out + i * 16);
  dst[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
  dst[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
}

// last layer: 16 -> 1 FP32 output, no activation; thread per node
__global__ void __launch_bounds__(kBlock)
k_conv_last(const uint16_t *__restrict__ in, const float *__restrict__ W, float *__restrict__ out,
            Dims d) {
  const uint64_t P = (uint64_t)d.x * d.y * d.z;
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= P) return;
  int x, y, z;
  decode(i, d, x, y, z);
  float a = 0.f;
  for (int t = 0; t < 27; ++t) {
    const int64_t s = tap_source<1, false>(x, y, z, t, d, d);
    if (s < 0) continue;
    const uint4 *src = reinterpret_cast<const uint4 *>(in + (uint64_t)s * 16);
    const uint4 lo = src[0], hi = src[1];
    const uint32_t w32[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const uint16_t h = (uint16_t)(w32[c >> 1] >> (16 * (c & 1)));
      a += W[t * 16 + c] * bf2f(h);
    }
  }
  out[i] = a;
}

// ---------------------------------------------------------------- backward
// dX[q][c] = sum_{t,o} W[o][t][c] G[(q - off_t) / s][o] over the taps whose
// source lands on q; thread per (q, c).  dX is on the convolution's input
// grid (the fine grid of an up-sampled input).
template <int STRIDE>
__global__ void __launch_bounds__(kBlock)
k_conv_adjoint(const float *__restrict__ G, const float *__restrict__ W, float *__restrict__ dX,
               Dims od, Dims xd, int cin, int cout) {
  const uint64_t n = (uint64_t)xd.x * xd.y * xd.z * cin;
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % cin);
  int x, y, z;
  decode(i / cin, xd, x, y, z);
  float a = 0.f;
  for (int t = 0; t < 27; ++t) {
    const int rx = x - (t % 3 - 1), ry = y - ((t / 3) % 3 - 1), rz = z - (t / 9 - 1);
    if (STRIDE == 2 && ((rx | ry | rz) & 1)) continue;
    const int px = rx / STRIDE, py = ry / STRIDE, pz = rz / STRIDE;
    if (rx < 0 || ry < 0 || rz < 0 || !inside(px, py, pz, od)) continue;
    const float *g = G + lin(px, py, pz, od) * cout;
    const float *w = W + (uint64_t)t * cin + c;
    for (int o = 0; o < cout; ++o) a += w[(uint64_t)o * 27 * cin] * g[o];
  }
  dX[i] = a;
}

// coarse[j][c] = (sum over the fine children i of j of fine[i][c]) * mask
// (mask: coarse activation > 0), optionally plus add[j][c] before the mask
__global__ void __launch_bounds__(kBlock)
k_up_adjoint_mask(const float *__restrict__ fine, Dims fd, const float *__restrict__ add,
                  const uint16_t *__restrict__ act, float *__restrict__ coarse, Dims cd, int C) {
  const uint64_t n = (uint64_t)cd.x * cd.y * cd.z * C;
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % C);
  int x, y, z;
  decode(i / C, cd, x, y, z);
  float a = add ? add[i] : 0.f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int fx = 2 * x + dx, fy = 2 * y + dy, fz = 2 * z + dz;
        if (inside(fx, fy, fz, fd)) a += fine[lin(fx, fy, fz, fd) * C + c];
      }
  coarse[i] = bf2f(act[i]) > 0.f ? a : 0.f;
}

// out = (a + b) * (act > 0), elementwise (b optional)
__global__ void k_add_mask(const float *__restrict__ a, const float *__restrict__ b,
                           const uint16_t *__restrict__ act, float *__restrict__ out, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float v = a[i] + (b ? b[i] : 0.f);
  out[i] = bf2f(act[i]) > 0.f ? v : 0.f;
}

// partial weight gradients: block (t, part) sums over its node range
//   dW[o][t][c] += G[p][o] * in(s p + off_t)[c]
// in: bf16 channels-last (or FP32 when cin == 1 and in32 is set)
template <int STRIDE, bool UP>
__global__ void __launch_bounds__(kBlock)
k_conv_wgrad(const uint16_t *__restrict__ in, const float *__restrict__ in32,
             const float *__restrict__ G, float *__restrict__ part, Dims od, Dims id, int cin,
             int cout, uint32_t n_parts) {
  constexpr int PC = 16;  // nodes staged per round
  __shared__ float sX[PC][64];
  __shared__ float sG[PC][64];
  const int t = blockIdx.x;
  const uint32_t pi = blockIdx.y;
  const uint64_t P = (uint64_t)od.x * od.y * od.z;
  const uint64_t p0 = P * pi / n_parts, p1 = P * (pi + 1) / n_parts;
  const int nout = cin * cout;  // <= 4096: up to 16 per thread
  float acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = 0.f;
  for (uint64_t pb = p0; pb < p1; pb += PC) {
    __syncthreads();
    for (int e = threadIdx.x; e < PC * 64; e += kBlock) {
      const int j = e / 64, c = e % 64;
      const uint64_t p = pb + j;
      float xv = 0.f, gv = 0.f;
      if (p < p1) {
        int x, y, z;
        decode(p, od, x, y, z);
        if (c < cin) {
          const int64_t s = tap_source<STRIDE, UP>(x, y, z, t, od, id);
          if (s >= 0) xv = in32 ? in32[s] : bf2f(in[(uint64_t)s * cin + c]);
        }
        if (c < cout) gv = G[p * cout + c];
      }
      sX[j][c] = xv;
      sG[j][c] = gv;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int idx = threadIdx.x + k * kBlock;
      if (idx >= nout) break;
      const int o = idx / cin, c = idx - o * cin;
      float s = acc[k];
#pragma unroll
      for (int j = 0; j < PC; ++j) s += sG[j][o] * sX[j][c];
      acc[k] = s;
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int idx = threadIdx.x + k * kBlock;
    if (idx >= nout) break;
    const int o = idx / cin, c = idx - o * cin;
    part[((uint64_t)pi * cout + o) * 27 * cin + (uint64_t)t * cin + c] = acc[k];
  }
}

// The same weight gradient on MFMA (bf16 inputs, FP32 accumulation), for
// channel counts that are multiples of 16: per tap t an implicit GEMM
// dW_t[o][c] = sum_p G[p][o] X[src(p, t)][c] with M = output channels,
// N = input channels and K = grid points, 32 points per
// v_mfma_f32_16x16x32_bf16.  Each wave owns every (o, c) tile and a
// strided share of the part's 32-point chunks; the output gradients G are
// rounded to bf16 (the activations X are stored in bf16 already).  The four
// waves are summed in LDS into the same per-part partials as k_conv_wgrad.
template <int CIN, int COUT, int STRIDE, bool UP>
__global__ void __launch_bounds__(kBlock)
k_conv_wgrad_mfma(const uint16_t *__restrict__ in, const float *__restrict__ G,
                  float *__restrict__ part, Dims od, Dims id, uint32_t n_parts) {
  static_assert(CIN % 16 == 0 && COUT % 16 == 0, "channel counts must be multiples of 16");
  constexpr int OT = COUT / 16, CT = CIN / 16, NW = kBlock / 64;
  __shared__ float red[NW][OT * CT * 4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = blockIdx.x;
  const uint32_t pi = blockIdx.y;
  const uint64_t P = (uint64_t)od.x * od.y * od.z;
  const uint64_t p0 = P * pi / n_parts, p1 = P * (pi + 1) / n_parts;
  const int kq = 8 * (lane >> 4), li = lane & 15;
  v4f acc[OT][CT];
#pragma unroll
  for (int a = 0; a < OT; ++a)
#pragma unroll
    for (int b = 0; b < CT; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  for (uint64_t cb = p0 + 32 * (uint64_t)wave; cb < p1; cb += 32 * NW) {
    // this lane's 8 points of the chunk: k = kq .. kq + 7
    int64_t src[8];
    uint64_t pt[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      pt[q] = cb + kq + q;
      src[q] = -1;
      if (pt[q] < p1) {
        int x, y, z;
        decode(pt[q], od, x, y, z);
        src[q] = tap_source<STRIDE, UP>(x, y, z, t, od, id);
      }
    }
    v8bf bx[CT], ag[OT];
#pragma unroll
    for (int b = 0; b < CT; ++b) {
      uint16_t h[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        h[q] = src[q] >= 0 ? in[(uint64_t)src[q] * CIN + b * 16 + li] : (uint16_t)0;
      const uint4 raw = make_uint4(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16,
                                   h[4] | (uint32_t)h[5] << 16, h[6] | (uint32_t)h[7] << 16);
      bx[b] = __builtin_bit_cast(v8bf, raw);
    }
#pragma unroll
    for (int a = 0; a < OT; ++a) {
      uint16_t h[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) h[q] = pt[q] < p1 ? f2bf(G[pt[q] * COUT + a * 16 + li]) : 0;
      const uint4 raw = make_uint4(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16,
                                   h[4] | (uint32_t)h[5] << 16, h[6] | (uint32_t)h[7] << 16);
      ag[a] = __builtin_bit_cast(v8bf, raw);
    }
#pragma unroll
    for (int a = 0; a < OT; ++a)
#pragma unroll
      for (int b = 0; b < CT; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ag[a], bx[b], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < OT; ++a)
#pragma unroll
    for (int b = 0; b < CT; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[wave][(a * CT + b) * 4 + r][lane] = acc[a][b][r];
  __syncthreads();
  // D layout: register r of lane l holds (o, c) = (16 a + 4 (l / 16) + r, 16 b + l % 16)
  for (int e = threadIdx.x; e < OT * CT * 4 * 64; e += kBlock) {
    const int tr = e / 64, l = e % 64;
    float v = 0.f;
#pragma unroll
    for (int w = 0; w < NW; ++w) v += red[w][tr][l];
    const int tile = tr / 4, r = tr % 4, a = tile / CT, b = tile % CT;
    const int o = a * 16 + 4 * (l >> 4) + r, c = b * 16 + (l & 15);
    part[((uint64_t)pi * COUT + o) * 27 * CIN + (uint64_t)t * CIN + c] = v;
  }
}

// Weight gradient of the two 1-channel layers (stride 1, no up-sampling):
// the input layer (1 -> 16, FP32 input in32) and the head (16 -> 1, bf16
// input in, output gradient with one channel).  Per tap only 16 weights, so
// every thread accumulates all 16 over its share of the part's points; the
// sums go through wave shuffles and LDS into the same per-part partials.
template <bool HEAD>
__global__ void __launch_bounds__(kBlock)
k_conv_wgrad_thin(const uint16_t *__restrict__ in, const float *__restrict__ in32,
                  const float *__restrict__ G, float *__restrict__ part, Dims od,
                  uint32_t n_parts) {
  __shared__ float red[kBlock / 64][16];
  const int t = blockIdx.x;
  const uint32_t pi = blockIdx.y;
  const uint64_t P = (uint64_t)od.x * od.y * od.z;
  const uint64_t p0 = P * pi / n_parts, p1 = P * (pi + 1) / n_parts;
  float acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = 0.f;
  for (uint64_t p = p0 + threadIdx.x; p < p1; p += kBlock) {
    int x, y, z;
    decode(p, od, x, y, z);
    const int64_t src = tap_source<1, false>(x, y, z, t, od, od);
    if (src < 0) continue;
    if constexpr (HEAD) {  // dW[0][t][c] = sum_p dy[p] act[p + off_t][c]
      const float g = G[p];
      const uint16_t *a = in + (uint64_t)src * 16;
#pragma unroll
      for (int k = 0; k < 16; ++k) acc[k] += g * bf2f(a[k]);
    } else {  // dW[o][t][0] = sum_p G[p][o] x[p + off_t]
      const float xv = in32[src];
      const float *g = G + p * 16;
#pragma unroll
      for (int k = 0; k < 16; ++k) acc[k] += g[k] * xv;
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    float v = acc[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    acc[k] = v;
  }
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < 16; ++k) red[threadIdx.x >> 6][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < 16) {
    const int k = threadIdx.x;
    float v = 0.f;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) v += red[w][k];
    // HEAD: cout 1, cin 16 (c = k); else cout 16 (o = k), cin 1
    const uint64_t idx = HEAD ? ((uint64_t)pi * 27 + t) * 16 + k
                              : ((uint64_t)pi * 16 + k) * 27 + t;
    part[idx] = v;
  }
}

// grad[w] += sum over parts of part[part][w] (fixed order)
__global__ void k_wgrad_reduce(const float *__restrict__ part, float *__restrict__ grad, uint32_t nw,
                               uint32_t n_parts) {
  const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= nw) return;
  float s = 0.f;
  for (uint32_t k = 0; k < n_parts; ++k) s += part[(uint64_t)k * nw + w];
  grad[w] += s;
}

// head backward: g[p][c] = (sum_t W[t][c] dy[p - off_t]) * (act[p][c] > 0)
__global__ void __launch_bounds__(kBlock)
k_head_backward(const float *__restrict__ dy, const float *__restrict__ W,
                const uint16_t *__restrict__ act, float *__restrict__ g, Dims d) {
  const uint64_t n = (uint64_t)d.x * d.y * d.z * 16;
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % 16);
  int x, y, z;
  decode(i / 16, d, x, y, z);
  float a = 0.f;
  for (int t = 0; t < 27; ++t) {
    const int px = x - (t % 3 - 1), py = y - ((t / 3) % 3 - 1), pz = z - (t / 9 - 1);
    if (inside(px, py, pz, d)) a += W[t * 16 + c] * dy[lin(px, py, pz, d)];
  }
  g[i] = bf2f(act[i]) > 0.f ? a : 0.f;
}

// loss pieces: dy = scale (y - target) on free nodes (0 on masked nodes),
// partial sums of (y - target)^2 per block
__global__ void __launch_bounds__(kBlock)
k_mse(const float *__restrict__ y, const float *__restrict__ target,
      const uint8_t *__restrict__ mask, float scale, float *__restrict__ dy,
      double *__restrict__ part, uint64_t n) {
  __shared__ double red[kBlock / 64];
  double s = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    const float d = (mask && mask[i]) ? 0.f : y[i] - target[i];
    dy[i] = scale * d;
    s += (double)d * d;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) a += red[w];
    part[blockIdx.x] = a;
  }
}

__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, uint32_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// layer weights [o][t][c] FP32 -> bf16 [o][KP] (k = t cin + c, zero padded)
// and the FP32 value of each bf16 weight [o][t][c]
__global__ void k_pack_weights(const float *__restrict__ w, uint16_t *__restrict__ wb,
                               float *__restrict__ wr, int cin, int cout, int kp) {
  const int n = cout * kp;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int o = i / kp, k = i - o * kp;
  const int K = 27 * cin;
  uint16_t h = 0;
  if (k < K) {
    h = f2bf(w[o * K + k]);
    wr[o * K + k] = bf2f(h);
  }
  wb[i] = h;
}

__global__ void k_scale_to_f32(const double *__restrict__ r, float *__restrict__ out, double s,
                               uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = (float)(r[i] * s);
}

__global__ void k_scale_to_f64(const float *__restrict__ y, double *__restrict__ z, double s,
                               uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) z[i] = (double)y[i] * s;
}

__global__ void k_axpy_f32(float *__restrict__ x, float a, const float *__restrict__ y, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] += a * y[i];
}

inline unsigned blocks(uint64_t n) { return (unsigned)std::max<uint64_t>(1, (n + kBlock - 1) / kBlock); }
inline Dims D3(const uint32_t *d) { return Dims{d[0], d[1], d[2]}; }

template <int CIN, int COUT, int STRIDE, bool UP, bool SKIP>
void mfma_launch(const uint16_t *in, const uint16_t *W, const uint16_t *skip, uint16_t *out,
                 const uint32_t *od, const uint32_t *id, hipStream_t s) {
  const uint64_t P = (uint64_t)od[0] * od[1] * od[2];
  const uint64_t waves = (P + 15) / 16;
  hipLaunchKernelGGL((k_conv_mfma<CIN, COUT, STRIDE, UP, SKIP>), dim3(blocks(waves * 64)),
                     dim3(kBlock), 0, s, in, W, skip, out, D3(od), D3(id));
}

}  // namespace

void nn_conv_forward(int layer, const uint16_t *in, const float *in32, const uint16_t *wb,
                     const float *wr, const uint16_t *skip, uint16_t *out, float *out32,
                     const uint32_t *od, const uint32_t *id, hipStream_t s) {
  const uint64_t P = (uint64_t)od[0] * od[1] * od[2];
  switch (layer) {
    case 0: hipLaunchKernelGGL(k_conv_first, dim3(blocks(P)), dim3(kBlock), 0, s, in32, wr, out, D3(od)); break;
    case 1: mfma_launch<16, 32, 2, false, false>(in, wb, nullptr, out, od, id, s); break;
    case 2: mfma_launch<32, 64, 2, false, false>(in, wb, nullptr, out, od, id, s); break;
    case 3: mfma_launch<64, 64, 2, false, false>(in, wb, nullptr, out, od, id, s); break;
    case 4: mfma_launch<64, 64, 1, true, true>(in, wb, skip, out, od, id, s); break;
    case 5: mfma_launch<64, 32, 1, true, true>(in, wb, skip, out, od, id, s); break;
    case 6: mfma_launch<32, 16, 1, true, true>(in, wb, skip, out, od, id, s); break;
    case 7: hipLaunchKernelGGL(k_conv_last, dim3(blocks(P)), dim3(kBlock), 0, s, in, wr, out32, D3(od)); break;
    default: throw Error{MFX_ERR_INVALID_ARG, "layer out of range"};
  }
  MFX_HIP(hipGetLastError());
}

void nn_conv_adjoint(int stride, const float *G, const float *W, float *dX, const uint32_t *od,
                     const uint32_t *xd, int cin, int cout, hipStream_t s) {
  const uint64_t n = (uint64_t)xd[0] * xd[1] * xd[2] * cin;
  if (stride == 1)
    hipLaunchKernelGGL(k_conv_adjoint<1>, dim3(blocks(n)), dim3(kBlock), 0, s, G, W, dX, D3(od),
                       D3(xd), cin, cout);
  else
    hipLaunchKernelGGL(k_conv_adjoint<2>, dim3(blocks(n)), dim3(kBlock), 0, s, G, W, dX, D3(od),
                       D3(xd), cin, cout);
  MFX_HIP(hipGetLastError());
}

void nn_up_adjoint_mask(const float *fine, const uint32_t *fd, const float *add,
                        const uint16_t *act, float *coarse, const uint32_t *cd, int C,
                        hipStream_t s) {
  const uint64_t n = (uint64_t)cd[0] * cd[1] * cd[2] * C;
  hipLaunchKernelGGL(k_up_adjoint_mask, dim3(blocks(n)), dim3(kBlock), 0, s, fine, D3(fd), add, act,
                     coarse, D3(cd), C);
  MFX_HIP(hipGetLastError());
}

void nn_add_mask(const float *a, const float *b, const uint16_t *act, float *out, uint64_t n,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_add_mask, dim3(blocks(n)), dim3(kBlock), 0, s, a, b, act, out, n);
  MFX_HIP(hipGetLastError());
}

// the MFMA weight gradient for the layer shapes of the U-Net; false if the
// shape has no MFMA instance (then the VALU kernel runs)
template <int CI, int CO, int ST, bool UPF>
bool wgrad_mfma_if(int stride, bool up, const uint16_t *in, const float *G, float *part,
                   const uint32_t *od, const uint32_t *id, int cin, int cout, uint32_t n_parts,
                   hipStream_t s) {
  if (!(in && cin == CI && cout == CO && stride == ST && up == UPF)) return false;
  hipLaunchKernelGGL((k_conv_wgrad_mfma<CI, CO, ST, UPF>), dim3(27, n_parts), dim3(kBlock), 0, s,
                     in, G, part, D3(od), D3(id), n_parts);
  return true;
}
bool wgrad_mfma(int stride, bool up, const uint16_t *in, const float *G, float *part,
                const uint32_t *od, const uint32_t *id, int cin, int cout, uint32_t n_parts,
                hipStream_t s) {
  return wgrad_mfma_if<32, 16, 1, true>(stride, up, in, G, part, od, id, cin, cout, n_parts, s) ||
         wgrad_mfma_if<64, 32, 1, true>(stride, up, in, G, part, od, id, cin, cout, n_parts, s) ||
         wgrad_mfma_if<64, 64, 1, true>(stride, up, in, G, part, od, id, cin, cout, n_parts, s) ||
         wgrad_mfma_if<64, 64, 2, false>(stride, up, in, G, part, od, id, cin, cout, n_parts, s) ||
         wgrad_mfma_if<32, 64, 2, false>(stride, up, in, G, part, od, id, cin, cout, n_parts, s) ||
         wgrad_mfma_if<16, 32, 2, false>(stride, up, in, G, part, od, id, cin, cout, n_parts, s);
}

void nn_conv_wgrad(int stride, bool up, const uint16_t *in, const float *in32, const float *G,
                   float *part, float *grad, const uint32_t *od, const uint32_t *id, int cin,
                   int cout, uint32_t n_parts, hipStream_t s) {
  MFX_REQUIRE(cin <= 64 && cout <= 64 && cin * cout <= 16 * kBlock, "weight tile too large");
  const dim3 grid(27, n_parts);
  // the U-Net's layers with 16-multiple channel counts run on MFMA, the two
  // 1-channel layers on the thin kernel
  if (stride == 1 && !up && cin == 16 && cout == 1 && in) {
    hipLaunchKernelGGL(k_conv_wgrad_thin<true>, grid, dim3(kBlock), 0, s, in, nullptr, G, part,
                       D3(od), n_parts);
  } else if (stride == 1 && !up && cin == 1 && cout == 16 && in32) {
    hipLaunchKernelGGL(k_conv_wgrad_thin<false>, grid, dim3(kBlock), 0, s, nullptr, in32, G,
                       part, D3(od), n_parts);
  } else if (!wgrad_mfma(stride, up, in, G, part, od, id, cin, cout, n_parts, s)) {
    if (stride == 2)
      hipLaunchKernelGGL((k_conv_wgrad<2, false>), grid, dim3(kBlock), 0, s, in, in32, G, part,
                         D3(od), D3(id), cin, cout, n_parts);
    else if (up)
      hipLaunchKernelGGL((k_conv_wgrad<1, true>), grid, dim3(kBlock), 0, s, in, in32, G, part,
                         D3(od), D3(id), cin, cout, n_parts);
    else
      hipLaunchKernelGGL((k_conv_wgrad<1, false>), grid, dim3(kBlock), 0, s, in, in32, G, part,
                         D3(od), D3(id), cin, cout, n_parts);
  }
  const uint32_t nw = 27u * cin * cout;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks(nw)), dim3(kBlock), 0, s, part, grad, nw, n_parts);
  MFX_HIP(hipGetLastError());
}

void nn_head_backward(const float *dy, const float *W, const uint16_t *act, float *g,
                      const uint32_t *d, hipStream_t s) {
  const uint64_t n = (uint64_t)d[0] * d[1] * d[2] * 16;
  hipLaunchKernelGGL(k_head_backward, dim3(blocks(n)), dim3(kBlock), 0, s, dy, W, act, g, D3(d));
  MFX_HIP(hipGetLastError());
}

uint32_t nn_mse_blocks(uint64_t n) {
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(1, (n + kBlock - 1) / kBlock), 1024);
}

void nn_mse(const float *y, const float *target, const uint8_t *mask, float scale, float *dy,
            double *part, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(nn_mse_blocks(n)), dim3(kBlock), 0, s, y, target, mask, scale, dy,
                     part, n);
  MFX_HIP(hipGetLastError());
}

void nn_adam(float *w, float *m, float *v, const float *g, uint32_t n, float lr, float b1, float b2,
             float eps, int step, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(kBlock), 0, s, w, m, v, g, n, lr, b1, b2, eps,
                     c1, c2);
  MFX_HIP(hipGetLastError());
}

void nn_pack_weights(const float *w, uint16_t *wb, float *wr, int cin, int cout, int kp,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_pack_weights, dim3(blocks((uint64_t)cout * kp)), dim3(kBlock), 0, s, w, wb,
                     wr, cin, cout, kp);
  MFX_HIP(hipGetLastError());
}

void nn_scale_to_f32(const double *r, float *out, double s_, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_to_f32, dim3(blocks(n)), dim3(kBlock), 0, s, r, out, s_, n);
  MFX_HIP(hipGetLastError());
}

void nn_axpy_f32(float *x, float a, const float *y, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy_f32, dim3(blocks(n)), dim3(kBlock), 0, s, x, a, y, n);
  MFX_HIP(hipGetLastError());
}

void nn_scale_to_f64(const float *y, double *z, double s_, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_to_f64, dim3(blocks(n)), dim3(kBlock), 0, s, y, z, s_, n);
  MFX_HIP(hipGetLastError());
}

}  // namespace mfx
