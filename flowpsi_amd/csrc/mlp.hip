// Learned initial guess for the block-SGS / line-SGS solve: a two-hop message-passing
// MLP of width 64 with bf16 GEMM operands on MFMA (v_mfma_f32_16x16x32_bf16,
// fp32 accumulation), fp32 activations and fp32 master weights.
//
//   h0  = relu(W0 x + b0)                    x: 32 cell features
//   a1  = mean of h0 over face neighbours
//   h1  = relu(W1 [h0; a1] + b1)
//   a2  = mean of h1 over face neighbours
//   h2  = relu(W2 [h1; a2] + b2)
//   out = W3 h2 + b3                         16 outputs, 5 used
//   X_guess = X_jacobi + sx * out[0:5]
//
// Every reduction (weight gradients, loss, feature scales) is a fixed-order
// two-pass reduction, so training is bitwise reproducible.
#include "mlp.h"
#include "blocks.h"

#include <algorithm>

namespace fpsi {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even float -> bf16 bits (finite inputs)
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }

__device__ __forceinline__ bf16x8 pack8(const float *v) {
  bf16x8 r;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    unsigned short b = f2bf(v[i]);
    r[i] = __builtin_bit_cast(__bf16, b);
  }
  return r;
}
__device__ __forceinline__ bf16x8 pack8b(const unsigned short *v) {
  bf16x8 r;
#pragma unroll
  for (int i = 0; i < 8; ++i) r[i] = __builtin_bit_cast(__bf16, v[i]);
  return r;
}

static constexpr int TPB = 256;
static inline int nblk(long n, int t = TPB) { return (int)((n + t - 1) / t); }

// ---------------------------------------------------------------------------
// dense layer forward: out[N][64] = act(in[N][K] W^T + b), W bf16 [64][K].
// `in` is the concatenation of in0[N][K0] and in1[N][K-K0] (in1 may be null).
// One wave computes 16 cells x 64 outputs; a block holds 4 waves.
// Lane layout (16x16x32): A row = lane%16, k = 8*(lane/16)+[0,8);
// B col = lane%16, same k; D row = 4*(lane/16)+i, col = lane%16.
// ---------------------------------------------------------------------------
template <int K, int NOUT, bool RELU>
__global__ void __launch_bounds__(256) k_dense_fwd(int N, const float *__restrict__ in0, int K0,
                                                   const float *__restrict__ in1,
                                                   const unsigned short *__restrict__ W,
                                                   const float *__restrict__ b, float *__restrict__ out) {
  __shared__ unsigned short sW[NOUT * K];
  for (int i = threadIdx.x; i < NOUT * K; i += blockDim.x) sW[i] = W[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = (blockIdx.x * 4 + wave) * 16;
  if (row0 >= N) return;
  const int r = lane & 15, kq = (lane >> 4) * 8;
  const int cell = min(row0 + r, N - 1);
  f32x4 acc[NOUT / 16];
#pragma unroll
  for (int t = 0; t < NOUT / 16; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < K; kk += 32) {
    const int k = kk + kq;
    float v[8];
    const float *src = (k < K0) ? in0 + (size_t)cell * K0 + k : in1 + (size_t)cell * (K - K0) + (k - K0);
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = src[i];
    const bf16x8 a = pack8(v);
#pragma unroll
    for (int t = 0; t < NOUT / 16; ++t) {
      const bf16x8 bb = pack8b(&sW[(t * 16 + r) * K + k]);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NOUT / 16; ++t) {
    const int o = t * 16 + r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = row0 + (lane >> 4) * 4 + i;
      if (row < N) {
        float y = acc[t][i] + b[o];
        if (RELU) y = y > 0.f ? y : 0.f;
        out[(size_t)row * NOUT + o] = y;
      }
    }
  }
}

// One message-passing hop of the forward pass for the guess (inference):
// h_out = relu(W [h_in; mean of h_in over face neighbours] + b), with the
// neighbour mean formed in registers in k_agg's order and arithmetic, so
// the result is bitwise that of k_agg + k_dense_fwd<128, 64, true> without
// the mean's round trip through memory.  FINAL keeps the hop's output on
// chip: it goes through LDS (one 16 x 64 tile per wave) into the output
// layer out = W3 h + b3 (bitwise k_dense_fwd<64, 16, false>), and the
// guess X = X0 + sd * out[0:5] is written.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <bool FINAL>
__global__ void __launch_bounds__(256) k_hop_fwd(int N, const int *__restrict__ off, const int *__restrict__ nbr,
                                                 const float *__restrict__ hin, const unsigned short *__restrict__ W,
                                                 const float *__restrict__ b, float *__restrict__ hout,
                                                 const unsigned short *__restrict__ W3, const float *__restrict__ b3,
                                                 const float *__restrict__ raw, const float *__restrict__ norm,
                                                 double *__restrict__ X) {
  __shared__ unsigned short sW[64 * 128];
  __shared__ unsigned short sW3[FINAL ? 16 * 64 : 1];
  __shared__ float sh[FINAL ? 4 * 16 * 64 : 1];
  for (int i = threadIdx.x; i < 64 * 128; i += blockDim.x) sW[i] = W[i];
  if constexpr (FINAL)
    for (int i = threadIdx.x; i < 16 * 64; i += blockDim.x) sW3[i] = W3[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = (blockIdx.x * 4 + wave) * 16;
  if (row0 >= N) return;
  const int r = lane & 15, kq = (lane >> 4) * 8;
  const int cell = min(row0 + r, N - 1);
  const int j0 = off[cell], j1 = off[cell + 1];
  const float inv = j1 > j0 ? 1.0f / (float)(j1 - j0) : 0.f;
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 128; kk += 32) {
    const int k = kk + kq;
    float v[8];
    if (k < 64) {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = hin[(size_t)cell * 64 + k + i];
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = 0.f;
      for (int j = j0; j < j1; ++j) {
        const float *hn = hin + (size_t)nbr[j] * 64 + (k - 64);
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] += hn[i];
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] *= inv;
    }
    const bf16x8 a = pack8(v);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bf16x8 bb = pack8b(&sW[(t * 16 + r) * 128 + k]);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[t], 0, 0, 0);
    }
  }
  if constexpr (!FINAL) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int o = t * 16 + r;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = row0 + (lane >> 4) * 4 + i;
        if (row < N) {
          float y = acc[t][i] + b[o];
          y = y > 0.f ? y : 0.f;
          hout[(size_t)row * 64 + o] = y;
        }
      }
    }
  } else {
    float *tile = sh + wave * 16 * 64;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int o = t * 16 + r;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float y = acc[t][i] + b[o];
        tile[((lane >> 4) * 4 + i) * 64 + o] = y > 0.f ? y : 0.f;
      }
    }
    wave_lds_sync();
    f32x4 o3 = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 64; kk += 32) {
      float v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = tile[r * 64 + kk + kq + i];
      const bf16x8 a = pack8(v);
      const bf16x8 bb = pack8b(&sW3[r * 64 + kk + kq]);
      o3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, o3, 0, 0, 0);
    }
    if (r < 5) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = row0 + (lane >> 4) * 4 + i;
        if (row < N) {
          const float out = o3[i] + b3[r];
          X[(size_t)row * 5 + r] = (double)raw[(size_t)row * 32 + r] + (double)norm[21 + r] * (double)out;
        }
      }
    }
  }
}

// input gradient: din[N][K] = g[N][64] W[64][K] (W given transposed, WT [K][64])
template <int K>
__global__ void __launch_bounds__(256) k_dense_bwd_in(int N, const float *__restrict__ g,
                                                      const unsigned short *__restrict__ WT,
                                                      float *__restrict__ din) {
  __shared__ unsigned short sW[K * 64];
  for (int i = threadIdx.x; i < K * 64; i += blockDim.x) sW[i] = WT[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = (blockIdx.x * 4 + wave) * 16;
  if (row0 >= N) return;
  const int r = lane & 15, kq = (lane >> 4) * 8;
  const int cell = min(row0 + r, N - 1);
  f32x4 acc[K / 16];
#pragma unroll
  for (int t = 0; t < K / 16; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 64; kk += 32) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = g[(size_t)cell * 64 + kk + kq + i];
    const bf16x8 a = pack8(v);
#pragma unroll
    for (int t = 0; t < K / 16; ++t) {
      const bf16x8 bb = pack8b(&sW[(t * 16 + r) * 64 + kk + kq]);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < K / 16; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = row0 + (lane >> 4) * 4 + i;
      if (row < N) din[(size_t)row * K + t * 16 + r] = acc[t][i];
    }
}

// weight gradient partials: for block p over cells [p*CPB, (p+1)*CPB):
// part[p][o][k] = sum_c g[c][o] in[c][k], part_b[p][o] = sum_c g[c][o].
// Wave w owns output rows [16w, 16w+16) (NOUT/16 waves used).
static constexpr int CPB = 2048;
template <int K, int NOUT>
__global__ void __launch_bounds__(256) k_wgrad(int N, const float *__restrict__ g, int gs, const float *__restrict__ in0,
                                               int K0, const float *__restrict__ in1,
                                               float *__restrict__ part, float *__restrict__ partb) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = blockIdx.x * CPB, c1 = min(N, c0 + CPB);
  float *pw = part + (size_t)blockIdx.x * NOUT * K;
  float *pb = partb + (size_t)blockIdx.x * NOUT;
  if (wave * 16 >= NOUT) return;
  const int r = lane & 15, kq = (lane >> 4) * 8;
  const int o = wave * 16 + r;
  f32x4 acc[K / 16];
#pragma unroll
  for (int t = 0; t < K / 16; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (int cb = c0; cb < c1; cb += 32) {
    float va[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int c = cb + kq + i;
      va[i] = c < c1 ? g[(size_t)c * gs + o] : 0.f;
    }
    // bias: lanes 0..15 of each k-quarter hold disjoint cells of output o
#pragma unroll
    for (int i = 0; i < 8; ++i) bsum += va[i];
    const bf16x8 a = pack8(va);
#pragma unroll
    for (int t = 0; t < K / 16; ++t) {
      const int k = t * 16 + r;
      float vb[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int c = cb + kq + i;
        float x = 0.f;
        if (c < c1) x = (k < K0) ? in0[(size_t)c * K0 + k] : in1[(size_t)c * (K - K0) + (k - K0)];
        vb[i] = x;
      }
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, pack8(vb), acc[t], 0, 0, 0);
    }
  }
  // D row = output 4*(lane/16)+i of this wave's tile, col = input t*16 + lane%16
#pragma unroll
  for (int t = 0; t < K / 16; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) pw[(size_t)(wave * 16 + (lane >> 4) * 4 + i) * K + t * 16 + r] = acc[t][i];
  // bias partial: combine the four k-quarters of output o in fixed order
  const float q1 = __shfl(bsum, (lane & 15) + 16, 64);
  const float q2 = __shfl(bsum, (lane & 15) + 32, 64);
  const float q3 = __shfl(bsum, (lane & 15) + 48, 64);
  if (lane < 16) pb[o] = ((bsum + q1) + q2) + q3;
}

// out[i] = sum_p part[p][i] for a few components with many partials, one
// workgroup per component: thread t sums parts t, t + TPB, ... in order,
// then a fixed-order tree over the workgroup
__global__ void __launch_bounds__(256) k_sum_parts_wg(int nparts, int len, const float *__restrict__ part,
                                                      float *__restrict__ out) {
  __shared__ float sh[256];
  const int i = blockIdx.x;
  float s = 0.f;
  for (int p = threadIdx.x; p < nparts; p += 256) s += part[(size_t)p * len + i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = sh[0];
}

// sum partials in block order: out[i] = sum_p part[p][i] (one thread per
// component, for many components with few partials each)
__global__ void k_sum_parts(int nparts, int len, const float *__restrict__ part, float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.f;
  for (int p = 0; p < nparts; ++p) s += part[(size_t)p * len + i];
  out[i] = s;
}

// neighbour mean of a 64-wide field: a[c] = (1/deg) sum_{n in N(c)} h[n]
__global__ void k_agg(int N, const int *__restrict__ off, const int *__restrict__ nbr,
                      const float *__restrict__ h, float *__restrict__ a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = t >> 4, q = (t & 15) * 4;
  if (c >= N) return;
  const int j0 = off[c], j1 = off[c + 1];
  float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int j = j0; j < j1; ++j) {
    const float *hn = h + (size_t)nbr[j] * 64 + q;
    s0 += hn[0]; s1 += hn[1]; s2 += hn[2]; s3 += hn[3];
  }
  const float inv = j1 > j0 ? 1.0f / (float)(j1 - j0) : 0.f;
  float *o = a + (size_t)c * 64 + q;
  o[0] = s0 * inv; o[1] = s1 * inv; o[2] = s2 * inv; o[3] = s3 * inv;
}

// transpose of k_agg (the neighbour relation is symmetric):
// dh[n] += sum_{c in N(n)} da[c] / deg(c)
__global__ void k_agg_bwd(int N, const int *__restrict__ off, const int *__restrict__ nbr,
                          const float *__restrict__ da, float *__restrict__ dh) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = t >> 4, q = (t & 15) * 4;
  if (n >= N) return;
  const int j0 = off[n], j1 = off[n + 1];
  float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int j = j0; j < j1; ++j) {
    const int c = nbr[j];
    const float inv = 1.0f / (float)(off[c + 1] - off[c]);
    const float *d = da + (size_t)c * 64 + q;
    s0 += d[0] * inv; s1 += d[1] * inv; s2 += d[2] * inv; s3 += d[3] * inv;
  }
  float *o = dh + (size_t)n * 64 + q;
  o[0] += s0; o[1] += s1; o[2] += s2; o[3] += s3;
}

// split a [N][128] input gradient into its two 64-wide halves: the first is
// the gradient of h (masked by its relu), the second of the neighbour mean.
__global__ void k_split(int N, const float *__restrict__ dcat, float *__restrict__ dh, float *__restrict__ da) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * 64) return;
  const int c = t >> 6, k = t & 63;
  dh[t] = dcat[(size_t)c * 128 + k];
  da[t] = dcat[(size_t)c * 128 + 64 + k];
}

// g = dh * relu'(h)  (h is the post-activation, > 0 where the unit is on)
__global__ void k_relu_mask(int n, const float *__restrict__ h, float *__restrict__ g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && !(h[t] > 0.f)) g[t] = 0.f;
}

// loss = mean over cells and 5 components of (out - y)^2 ; dout = d loss / d out
__global__ void k_loss(int N, const float *__restrict__ out, const float *__restrict__ y,
                       float *__restrict__ dout, float *__restrict__ part) {
  __shared__ float sh[TPB];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.f;
  if (c < N) {
    const float scale = 2.0f / (5.0f * (float)N);
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      float d = 0.f;
      if (k < 5) {
        const float e = out[(size_t)c * 16 + k] - y[(size_t)c * 8 + k];
        s += e * e;
        d = scale * e;
      }
      dout[(size_t)c * 32 + k] = d;
    }
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Adam on the fp32 master weights, refreshes the bf16 copy; mask zeroes
// parameters that are structurally absent (unused output rows)
__global__ void k_adam(int n, float lr, float b1, float b2, float eps, float c1, float c2,
                       const float *__restrict__ grad, const unsigned char *__restrict__ mask,
                       float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       unsigned short *__restrict__ wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = mask[i] ? grad[i] : 0.f;
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  const float wi = mask[i] ? w[i] - lr * mh / (sqrtf(vh) + eps) : 0.f;
  w[i] = wi;
  wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(int n, const float *__restrict__ w, unsigned short *__restrict__ wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = f2bf(w[i]);
}

// W3T[64][32] from W3[16][64], zero columns 16..31 (input gradient of layer 3)
__global__ void k_w3t(const unsigned short *__restrict__ W3, unsigned short *__restrict__ wt3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 64 * 32) return;
  const int h = i / 32, k = i % 32;
  wt3[i] = k < 16 ? W3[k * 64 + h] : (unsigned short)0;
}

// transposed bf16 copy WT[K][64] of W[64][K]
__global__ void k_transpose_bf16(int K, const unsigned short *__restrict__ W, unsigned short *__restrict__ WT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 64 * K) return;
  const int o = i / K, k = i % K;
  WT[k * 64 + o] = W[i];
}

// ---------------------------------------------------------------------------
// features and guess
// ---------------------------------------------------------------------------
// raw per-cell quantities [N][32]: X0 (the plain start), the previous
// Newton update, log|U_kk| of the diagonal factors (0..14); then
// (k_feat_resid) the residual of the plain start r0 = B - A X0 and the
// Jacobi correction D^-1 r0 (16..25)
template <typename FL>
__global__ void k_feat_raw(int N, const double *__restrict__ X, const double *__restrict__ Xprev,
                           const FL *__restrict__ lu, float *__restrict__ raw) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float *o = raw + (size_t)c * 32;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    o[k] = (float)X[(size_t)c * 5 + k];
    o[5 + k] = (float)Xprev[(size_t)c * 5 + k];
    o[10 + k] = logf(fabsf((float)lu[(size_t)c * 25 + k * 6]) + 1e-30f);
  }
  o[15] = 0.f;
}

// r0 and D^-1 r0 from the diagonal factors (pivoted under block SGS, piv
// null: unpivoted FP64 factors of line SGS)
template <typename FL>
__global__ void k_feat_resid(int N, const double *__restrict__ r, const FL *__restrict__ lu,
                             const int *__restrict__ piv, float *__restrict__ raw) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double b[5], x[5];
  FL a[25];
#pragma unroll
  for (int k = 0; k < 5; ++k) b[k] = r[(size_t)c * 5 + k];
#pragma unroll
  for (int i = 0; i < 25; ++i) a[i] = lu[(size_t)c * 25 + i];
  lusolve5<FL>(a, piv ? piv[c] : PIV_ID, b, x);
  float *o = raw + (size_t)c * 32;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    o[16 + k] = (float)b[k];
    o[21 + k] = (float)x[k];
  }
#pragma unroll
  for (int k = 26; k < 32; ++k) o[k] = 0.f;
}

// feature scales: part[p][0..31] = sums of squares (X0, previous update, r0, D^-1 r0)
// and sums (log|U_kk|)
__device__ __forceinline__ bool feat_squared(int k) { return k < 10 || (k >= 16 && k < 26); }
__global__ void k_feat_stats(int N, const float *__restrict__ raw, float *__restrict__ part) {
  __shared__ float sh[TPB][32];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    float v = 0.f;
    if (c < N && k != 15 && k < 26) {
      const float x = raw[(size_t)c * 32 + k];
      v = feat_squared(k) ? x * x : x;
    }
    sh[threadIdx.x][k] = v;
  }
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 32; ++k) sh[threadIdx.x][k] += sh[threadIdx.x + w][k];
    __syncthreads();
  }
  if (threadIdx.x < 32) part[(size_t)blockIdx.x * 32 + threadIdx.x] = sh[0][threadIdx.x];
}

// x[c] = normalised features: X0, the previous update and D^-1 r0 in units
// of the RMS of D^-1 r0 per equation (the unit of the target, so that the
// solution's dependence on them is the same map for every system), r0/sr,
// log|U_kk| - mean, geometry (4)
__global__ void k_feat(int N, const float *__restrict__ raw, const float *__restrict__ geo,
                       const float *__restrict__ norm, float *__restrict__ x) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const float *r = raw + (size_t)c * 32;
  float *o = x + (size_t)c * 32;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const float sd = norm[21 + k];
    o[k] = r[k] / sd;
    o[5 + k] = r[5 + k] / sd;
    o[10 + k] = r[10 + k] - norm[10 + k];
    o[19 + k] = r[16 + k] / norm[16 + k];
    o[24 + k] = r[21 + k] / sd;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) o[15 + k] = geo[(size_t)c * 4 + k];
#pragma unroll
  for (int k = 29; k < 32; ++k) o[k] = 0.f;
}

// training target y = (dq* - X0) / sd, in units of the Jacobi correction's
// RMS per equation
__global__ void k_target(int N, const double *__restrict__ Xs, const float *__restrict__ raw,
                         const float *__restrict__ norm, float *__restrict__ y) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    y[(size_t)c * 8 + k] =
        k < 5 ? (float)((Xs[(size_t)c * 5 + k] - (double)raw[(size_t)c * 32 + k]) / norm[21 + k]) : 0.f;
}

// ---------------------------------------------------------------------------
// host-side orchestration
// ---------------------------------------------------------------------------
MlpLayout mlp_layout() {
  MlpLayout L;
  L.w0 = 0;
  L.b0 = L.w0 + 64 * 32;
  L.w1 = L.b0 + 64;
  L.b1 = L.w1 + 64 * 128;
  L.w2 = L.b1 + 64;
  L.b2 = L.w2 + 64 * 128;
  L.w3 = L.b2 + 64;
  L.b3 = L.w3 + 16 * 64;
  L.n = L.b3 + 16;
  return L;
}

int mlp_parts(int N) { return (N + CPB - 1) / CPB; }

void mlp_forward(const MlpBufs &m, int N, hipStream_t st) {
  const MlpLayout L = mlp_layout();
  const int gb = nblk(N, 64);
  hipLaunchKernelGGL((k_dense_fwd<32, 64, true>), dim3(gb), dim3(256), 0, st, N, m.x, 32, (const float *)nullptr,
                     m.wb + L.w0, m.w + L.b0, m.h0);
  hipLaunchKernelGGL(k_agg, dim3(nblk((long)N * 16)), dim3(TPB), 0, st, N, m.nb_off, m.nb_idx, m.h0, m.a1);
  hipLaunchKernelGGL((k_dense_fwd<128, 64, true>), dim3(gb), dim3(256), 0, st, N, m.h0, 64, m.a1, m.wb + L.w1,
                     m.w + L.b1, m.h1);
  hipLaunchKernelGGL(k_agg, dim3(nblk((long)N * 16)), dim3(TPB), 0, st, N, m.nb_off, m.nb_idx, m.h1, m.a2);
  hipLaunchKernelGGL((k_dense_fwd<128, 64, true>), dim3(gb), dim3(256), 0, st, N, m.h1, 64, m.a2, m.wb + L.w2,
                     m.w + L.b2, m.h2);
  hipLaunchKernelGGL((k_dense_fwd<64, 16, false>), dim3(gb), dim3(256), 0, st, N, m.h2, 64, (const float *)nullptr,
                     m.wb + L.w3, m.w + L.b3, m.out);
}

// the guess X = X0 + sd * out[0:5] from the normalised features m.x: the
// forward pass of mlp_forward, bitwise, with the neighbour means in
// registers and the last hop, the output layer and the guess fused
void mlp_infer_guess(const MlpBufs &m, int N, double *X, hipStream_t st) {
  const MlpLayout L = mlp_layout();
  const int gb = nblk(N, 64);
  hipLaunchKernelGGL((k_dense_fwd<32, 64, true>), dim3(gb), dim3(256), 0, st, N, m.x, 32, (const float *)nullptr,
                     m.wb + L.w0, m.w + L.b0, m.h0);
  hipLaunchKernelGGL(k_hop_fwd<false>, dim3(gb), dim3(256), 0, st, N, m.nb_off, m.nb_idx, (const float *)m.h0,
                     m.wb + L.w1, m.w + L.b1, m.h1, (const unsigned short *)nullptr, (const float *)nullptr,
                     (const float *)nullptr, (const float *)nullptr, (double *)nullptr);
  hipLaunchKernelGGL(k_hop_fwd<true>, dim3(gb), dim3(256), 0, st, N, m.nb_off, m.nb_idx, (const float *)m.h1,
                     m.wb + L.w2, m.w + L.b2, (float *)nullptr, m.wb + L.w3, m.w + L.b3, (const float *)m.raw,
                     (const float *)m.norm, X);
}

// loss + gradient of all parameters into m.grad (fixed-order reductions)
void mlp_backward(const MlpBufs &m, int N, hipStream_t st) {
  const MlpLayout L = mlp_layout();
  const int gb = nblk(N, 64);
  const int np = mlp_parts(N);
  hipLaunchKernelGGL(k_loss, dim3(nblk(N)), dim3(TPB), 0, st, N, m.out, m.y, m.dout, m.lpart);
  hipLaunchKernelGGL(k_sum_parts_wg, dim3(1), dim3(256), 0, st, nblk(N), 1, m.lpart, m.loss);
  // transposed bf16 weights for the input gradients
  hipLaunchKernelGGL(k_transpose_bf16, dim3(nblk(64 * 128)), dim3(TPB), 0, st, 128, m.wb + L.w1, m.wt1);
  hipLaunchKernelGGL(k_transpose_bf16, dim3(nblk(64 * 128)), dim3(TPB), 0, st, 128, m.wb + L.w2, m.wt2);
  // layer 3: g = dout [N][16]
  hipLaunchKernelGGL((k_wgrad<64, 16>), dim3(np), dim3(256), 0, st, N, m.dout, 32, m.h2, 64,
                     (const float *)nullptr, m.wpart, m.bpart);
  hipLaunchKernelGGL(k_sum_parts, dim3(nblk(16 * 64)), dim3(TPB), 0, st, np, 16 * 64, m.wpart, m.grad + L.w3);
  hipLaunchKernelGGL(k_sum_parts_wg, dim3(16), dim3(256), 0, st, np, 16, m.bpart, m.grad + L.b3);
  // dh2 = W3^T dout: a 64-wide output from the 16-wide gradient, done as a
  // dense layer over the transposed (zero-padded) W3
  hipLaunchKernelGGL(k_w3t, dim3(nblk(64 * 32)), dim3(TPB), 0, st, m.wb + L.w3, m.wt3);
  hipLaunchKernelGGL((k_dense_fwd<32, 64, false>), dim3(gb), dim3(256), 0, st, N, m.dout, 32,
                     (const float *)nullptr, m.wt3, m.zero64, m.g2);
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk((long)N * 64)), dim3(TPB), 0, st, N * 64, m.h2, m.g2);
  // layer 2
  hipLaunchKernelGGL((k_wgrad<128, 64>), dim3(np), dim3(256), 0, st, N, m.g2, 64, m.h1, 64, m.a2, m.wpart, m.bpart);
  hipLaunchKernelGGL(k_sum_parts, dim3(nblk(64 * 128)), dim3(TPB), 0, st, np, 64 * 128, m.wpart, m.grad + L.w2);
  hipLaunchKernelGGL(k_sum_parts_wg, dim3(64), dim3(256), 0, st, np, 64, m.bpart, m.grad + L.b2);
  hipLaunchKernelGGL(k_dense_bwd_in<128>, dim3(gb), dim3(256), 0, st, N, m.g2, m.wt2, m.dcat);
  hipLaunchKernelGGL(k_split, dim3(nblk((long)N * 64)), dim3(TPB), 0, st, N, m.dcat, m.g1, m.da);
  hipLaunchKernelGGL(k_agg_bwd, dim3(nblk((long)N * 16)), dim3(TPB), 0, st, N, m.nb_off, m.nb_idx, m.da, m.g1);
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk((long)N * 64)), dim3(TPB), 0, st, N * 64, m.h1, m.g1);
  // layer 1
  hipLaunchKernelGGL((k_wgrad<128, 64>), dim3(np), dim3(256), 0, st, N, m.g1, 64, m.h0, 64, m.a1, m.wpart, m.bpart);
  hipLaunchKernelGGL(k_sum_parts, dim3(nblk(64 * 128)), dim3(TPB), 0, st, np, 64 * 128, m.wpart, m.grad + L.w1);
  hipLaunchKernelGGL(k_sum_parts_wg, dim3(64), dim3(256), 0, st, np, 64, m.bpart, m.grad + L.b1);
  hipLaunchKernelGGL(k_dense_bwd_in<128>, dim3(gb), dim3(256), 0, st, N, m.g1, m.wt1, m.dcat);
  hipLaunchKernelGGL(k_split, dim3(nblk((long)N * 64)), dim3(TPB), 0, st, N, m.dcat, m.g0, m.da);
  hipLaunchKernelGGL(k_agg_bwd, dim3(nblk((long)N * 16)), dim3(TPB), 0, st, N, m.nb_off, m.nb_idx, m.da, m.g0);
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk((long)N * 64)), dim3(TPB), 0, st, N * 64, m.h0, m.g0);
  // layer 0
  hipLaunchKernelGGL((k_wgrad<32, 64>), dim3(np), dim3(256), 0, st, N, m.g0, 64, m.x, 32,
                     (const float *)nullptr, m.wpart, m.bpart);
  hipLaunchKernelGGL(k_sum_parts, dim3(nblk(64 * 32)), dim3(TPB), 0, st, np, 64 * 32, m.wpart, m.grad + L.w0);
  hipLaunchKernelGGL(k_sum_parts_wg, dim3(64), dim3(256), 0, st, np, 64, m.bpart, m.grad + L.b0);
}

}  // namespace fpsi

namespace fpsi {
void mlp_features(const MlpBufs &m, int N, int fp32, const double *X, const double *Xprev, const void *lu,
                  int lu_f64, const int *piv, const double *r0, hipStream_t st) {
  const dim3 gb(nblk(N)), tb(TPB);
  if (fp32 && !lu_f64) {
    hipLaunchKernelGGL(k_feat_raw<float>, gb, tb, 0, st, N, X, Xprev, (const float *)lu, m.raw);
    hipLaunchKernelGGL(k_feat_resid<float>, gb, tb, 0, st, N, r0, (const float *)lu, piv, m.raw);
  } else {
    hipLaunchKernelGGL(k_feat_raw<double>, gb, tb, 0, st, N, X, Xprev, (const double *)lu, m.raw);
    hipLaunchKernelGGL(k_feat_resid<double>, gb, tb, 0, st, N, r0, (const double *)lu, piv, m.raw);
  }
}
void mlp_feature_stats(const MlpBufs &m, int N, float *out32, hipStream_t st) {
  hipLaunchKernelGGL(k_feat_stats, dim3(nblk(N)), dim3(TPB), 0, st, N, m.raw, m.lpart);
  hipLaunchKernelGGL(k_sum_parts_wg, dim3(32), dim3(256), 0, st, nblk(N), 32, m.lpart, out32);
}
void mlp_normalise(const MlpBufs &m, int N, float *x, hipStream_t st) {
  hipLaunchKernelGGL(k_feat, dim3(nblk(N)), dim3(TPB), 0, st, N, m.raw, m.geo, m.norm, x);
}
void mlp_target(const MlpBufs &m, int N, const double *Xs, float *y, hipStream_t st) {
  hipLaunchKernelGGL(k_target, dim3(nblk(N)), dim3(TPB), 0, st, N, Xs, m.raw, m.norm, y);
}
void mlp_adam(const MlpBufs &m, float lr, int t, hipStream_t st) {
  const MlpLayout L = mlp_layout();
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3(nblk(L.n)), dim3(TPB), 0, st, L.n, lr, b1, b2, 1e-8f, c1, c2, m.grad, m.mask,
                     m.w, m.am, m.av, m.wb);
}
void mlp_refresh_bf16(const MlpBufs &m, hipStream_t st) {
  const MlpLayout L = mlp_layout();
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(L.n)), dim3(TPB), 0, st, L.n, m.w, m.wb);
}
}  // namespace fpsi
