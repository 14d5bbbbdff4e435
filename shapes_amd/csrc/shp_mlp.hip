// shp_mlp.hip - the learned warm start: a 14-128-128-128-2 SiLU MLP that
// predicts the initial (normal, friction) impulses of contacts that were not
// in last frame's cache, where the reference starts at zero
// (shapes/src/Physics/Solvers/Contact.hs:91-99).  Weights are BF16 on the
// device, every product accumulates in FP32 on the matrix cores
// (v_mfma_f32_16x16x16_bf16), biases and activations are FP32 and are
// rounded to BF16 only where they feed the next product.
//
// Kernels:
//   k_mlp_forward   fused forward of tiles of 64 rows (one workgroup of 4
//                   waves per tile, activations in LDS); optionally stores
//                   pre-activations and BF16 activations for training
//   k_mlp_backward  fused backward over tiles; weight gradients accumulate
//                   in MFMA accumulators across a workgroup's tiles and are
//                   written as one partial per workgroup (no atomics)
//   k_mlp_reduce    sums the partials in a fixed order (deterministic)
//   k_mlp_adam      Adam update of the FP32 master weights + BF16 copy
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/shapes_gpu.h"
#include "shp_engine.h"
#include "shp_mlp.h"
#include "shp_records.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Operand fragments of v_mfma_f32_16x16x16_bf16: lane l holds
//   A[row = l%16][k = 4*(l/16) + 0..3]   B[k = 4*(l/16) + 0..3][col = l%16]
// and receives D[row = 4*(l/16) + i][col = l%16], i = 0..3.
__device__ __forceinline__ bf16x4 frag_a(const uint16_t* A, int lda, int r0, int k0) {
  int l = threadIdx.x & 63;
  const uint16_t* p = A + (r0 + (l & 15)) * lda + k0 + 4 * (l >> 4);
  bf16x4 v;
  v[0] = (short)p[0];
  v[1] = (short)p[1];
  v[2] = (short)p[2];
  v[3] = (short)p[3];
  return v;
}
// A stored transposed: element (r, k) at A[k * lda + r]
__device__ __forceinline__ bf16x4 frag_at(const uint16_t* A, int lda, int r0, int k0) {
  int l = threadIdx.x & 63;
  int r = r0 + (l & 15), k = k0 + 4 * (l >> 4);
  bf16x4 v;
  v[0] = (short)A[(k + 0) * lda + r];
  v[1] = (short)A[(k + 1) * lda + r];
  v[2] = (short)A[(k + 2) * lda + r];
  v[3] = (short)A[(k + 3) * lda + r];
  return v;
}
// B row-major K x N: element (k, n) at B[k * ldb + n]
__device__ __forceinline__ bf16x4 frag_b(const uint16_t* B, int ldb, int k0, int c0) {
  int l = threadIdx.x & 63;
  int n = c0 + (l & 15), k = k0 + 4 * (l >> 4);
  bf16x4 v;
  v[0] = (short)B[(k + 0) * ldb + n];
  v[1] = (short)B[(k + 1) * ldb + n];
  v[2] = (short)B[(k + 2) * ldb + n];
  v[3] = (short)B[(k + 3) * ldb + n];
  return v;
}
// B given as its transpose: element (k, n) at B[n * ldb + k]
__device__ __forceinline__ bf16x4 frag_bt(const uint16_t* B, int ldb, int k0, int c0) {
  int l = threadIdx.x & 63;
  const uint16_t* p = B + (c0 + (l & 15)) * ldb + k0 + 4 * (l >> 4);
  bf16x4 v;
  v[0] = (short)p[0];
  v[1] = (short)p[1];
  v[2] = (short)p[2];
  v[3] = (short)p[3];
  return v;
}

struct MlpDev {
  // BF16 weights, row-major [in][out] with padded shapes
  const uint16_t *w1, *w2, *w3, *w4;  // 16x128, 128x128, 128x128, 128x16
  const float *b1, *b2, *b3, *b4;     // 128, 128, 128, 16
};

struct TrainBufs {
  float *z1, *z2, *z3;        // pre-activations, rows x 128
  uint16_t *h1, *h2, *h3;     // BF16 activations, rows x 128
  float* y;                   // outputs, rows x 16 (cols 0,1 used)
};

#define TILE 64

// One 64-row tile: X (64 x 16 BF16) in LDS -> Y (64 x 16 FP32).
// Wave w owns rows 16w..16w+15.
// The inference kernel's LDS copies of the weights: transposed ([out][in])
// with rows of WT1 (first layer) or WT2 entries, padded so that its
// transposing stores spread over the LDS banks while a fragment stays one
// aligned 8-byte read.
constexpr int WT1 = 20, WT2 = 132;

// WT: the weights of m transposed and padded as above instead of
// row-major [in][out].
template <bool WT>
__device__ void forward_tile(const MlpDev& m, uint16_t* sx, uint16_t* s1, uint16_t* s2,
                             float* yout /* 64 x 16 */, const TrainBufs* tb, int row0) {
  int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  int r0 = 16 * wv;
  // layer 1: K = 16
  for (int ct = 0; ct < 8; ++ct) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma16(frag_a(sx, 16, r0, 0),
                 WT ? frag_bt(m.w1, WT1, 0, 16 * ct) : frag_b(m.w1, 128, 0, 16 * ct), acc);
    int col = 16 * ct + (l & 15);
    for (int i = 0; i < 4; ++i) {
      int row = r0 + 4 * (l >> 4) + i;
      float z = acc[i] + m.b1[col];
      uint16_t h = f2bf(silu(z));
      s1[row * 128 + col] = h;
      if (tb) {
        tb->z1[(size_t)(row0 + row) * 128 + col] = z;
        tb->h1[(size_t)(row0 + row) * 128 + col] = h;
      }
    }
  }
  __syncthreads();
  // layers 2 and 3: K = 128
  for (int layer = 0; layer < 2; ++layer) {
    const uint16_t* in = layer == 0 ? s1 : s2;
    uint16_t* out = layer == 0 ? s2 : s1;
    const uint16_t* W = layer == 0 ? m.w2 : m.w3;
    const float* bias = layer == 0 ? m.b2 : m.b3;
    float* zs = tb ? (layer == 0 ? tb->z2 : tb->z3) : nullptr;
    uint16_t* hs = tb ? (layer == 0 ? tb->h2 : tb->h3) : nullptr;
    f32x4 acc[8];
    for (int ct = 0; ct < 8; ++ct) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < 128; k0 += 16) {
      bf16x4 a = frag_a(in, 128, r0, k0);
      for (int ct = 0; ct < 8; ++ct)
        acc[ct] = mfma16(a, WT ? frag_bt(W, WT2, k0, 16 * ct) : frag_b(W, 128, k0, 16 * ct), acc[ct]);
    }
    __syncthreads();  // every wave has read `in` before `out` (== in of layer 1) is overwritten
    for (int ct = 0; ct < 8; ++ct) {
      int col = 16 * ct + (l & 15);
      for (int i = 0; i < 4; ++i) {
        int row = r0 + 4 * (l >> 4) + i;
        float z = acc[ct][i] + bias[col];
        uint16_t h = f2bf(silu(z));
        out[row * 128 + col] = h;
        if (tb) {
          zs[(size_t)(row0 + row) * 128 + col] = z;
          hs[(size_t)(row0 + row) * 128 + col] = h;
        }
      }
    }
    __syncthreads();
  }
  // layer 4: 128 -> 16 (2 used), input in s1
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < 128; k0 += 16)
    acc = mfma16(frag_a(s1, 128, r0, k0), WT ? frag_bt(m.w4, WT2, k0, 0) : frag_b(m.w4, 16, k0, 0), acc);
  int col = l & 15;
  for (int i = 0; i < 4; ++i) {
    int row = r0 + 4 * (l >> 4) + i;
    yout[row * 16 + col] = acc[i] + m.b4[col];
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_mlp_forward(MlpDev m, const uint16_t* X /* rows x 16 */,
                                                     int rows, float* Y /* rows x 16 */,
                                                     TrainBufs tb, int store) {
  __shared__ uint16_t sx[TILE * 16];
  __shared__ uint16_t s1[TILE * 128];
  __shared__ uint16_t s2[TILE * 128];
  __shared__ float sy[TILE * 16];
  for (int t = blockIdx.x; t * TILE < rows; t += gridDim.x) {
    int row0 = t * TILE;
    for (int i = threadIdx.x; i < TILE * 16; i += 256) sx[i] = X[(size_t)row0 * 16 + i];
    __syncthreads();
    forward_tile<false>(m, sx, s1, s2, sy, store ? &tb : nullptr, row0);
    for (int i = threadIdx.x; i < TILE * 16; i += 256) Y[(size_t)row0 * 16 + i] = sy[i];
    __syncthreads();
  }
}

// Backward of the mean loss over `rows` rows (rows % 64 == 0) given the
// forward buffers: squared error (huber = 0), dY = 2 (Y - T) / n_real, or
// twice the Huber loss with threshold `huber`, dY = 2 clamp(Y - T, +-huber)
// / n_real (the same gradient where |Y - T| <= huber).
// Workgroup g accumulates the weight gradients of its tiles in MFMA
// accumulators and writes partial g (MLP_NPARAM_PAD floats).
__global__ __launch_bounds__(256) void k_mlp_backward(MlpDev m, const uint16_t* X, const float* T,
                                                      int rows, int n_real, TrainBufs tb,
                                                      float* partial, float huber) {
  __shared__ uint16_t s_dy[TILE * 16];
  __shared__ uint16_t s_da[TILE * 128];
  __shared__ uint16_t s_db[TILE * 128];
  __shared__ float s_bias[3 * 128 + 16];
  int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  // this wave's output tiles of each weight gradient
  //   dW4 128x16: 8 row tiles -> wave w owns row tiles {2w, 2w+1}
  //   dW3, dW2 128x128: 64 tiles -> wave w owns row tiles {2w, 2w+1} x 8 cols
  //   dW1 16x128: 8 col tiles -> wave w owns col tiles {2w, 2w+1}
  f32x4 g4[2], g3[16], g2[16], g1[2];
  for (int i = 0; i < 2; ++i) g4[i] = g1[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int i = 0; i < 16; ++i) g3[i] = g2[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int i = threadIdx.x; i < 3 * 128 + 16; i += 256) s_bias[i] = 0.f;
  float inv_n = 2.0f / (float)n_real;
  for (int t = blockIdx.x; t * TILE < rows; t += gridDim.x) {
    int row0 = t * TILE;
    __syncthreads();
    for (int i = threadIdx.x; i < TILE * 16; i += 256) {
      int r = i >> 4, cc = i & 15;
      float d = 0.f;
      if (cc < 2 && row0 + r < n_real) {
        float e = tb.y[(size_t)(row0 + r) * 16 + cc] - T[(size_t)(row0 + r) * 2 + cc];
        if (huber > 0.f) e = fminf(fmaxf(e, -huber), huber);
        d = e * inv_n;
      }
      s_dy[i] = f2bf(d);
      if (cc < 2) atomicAdd(&s_bias[3 * 128 + cc], d);
    }
    __syncthreads();
    // dW4 += H3^T dY   (H3 rows x 128 global, transposed read)
    for (int i = 0; i < 2; ++i) {
      int rt = 2 * wv + i;
      for (int k0 = 0; k0 < TILE; k0 += 16)
        g4[i] = mfma16(frag_at(tb.h3 + (size_t)row0 * 128, 128, 16 * rt, k0), frag_b(s_dy, 16, k0, 0), g4[i]);
    }
    // dZ3 = (dY W4^T) * silu'(Z3)   (64 x 128): wave w rows 16w.., 8 col tiles, K = 16
    const uint16_t* Ws[3] = {m.w3, m.w2, nullptr};
    uint16_t* dz_cur = s_da;
    {
      int r0 = 16 * wv;
      for (int ct = 0; ct < 8; ++ct) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma16(frag_a(s_dy, 16, r0, 0), frag_bt(m.w4, 16, 0, 16 * ct), acc);
        int col = 16 * ct + (l & 15);
        float bsum = 0.f;
        for (int i = 0; i < 4; ++i) {
          int row = r0 + 4 * (l >> 4) + i;
          float d = acc[i] * dsilu(tb.z3[(size_t)(row0 + row) * 128 + col]);
          dz_cur[row * 128 + col] = f2bf(d);
          bsum += d;
        }
        atomicAdd(&s_bias[2 * 128 + col], bsum);
      }
    }
    __syncthreads();
    // layers 3 and 2: dW_L += H_{L-1}^T dZ_L ; dZ_{L-1} = (dZ_L W_L^T) * silu'(Z_{L-1})
    for (int layer = 0; layer < 2; ++layer) {
      const uint16_t* hprev = layer == 0 ? tb.h2 : tb.h1;
      const float* zprev = layer == 0 ? tb.z2 : tb.z1;
      f32x4* g = layer == 0 ? g3 : g2;
      uint16_t* dz_next = dz_cur == s_da ? s_db : s_da;
      for (int i = 0; i < 2; ++i) {
        int rt = 2 * wv + i;
        for (int k0 = 0; k0 < TILE; k0 += 16) {
          bf16x4 a = frag_at(hprev + (size_t)row0 * 128, 128, 16 * rt, k0);
          for (int ct = 0; ct < 8; ++ct) g[8 * i + ct] = mfma16(a, frag_b(dz_cur, 128, k0, 16 * ct), g[8 * i + ct]);
        }
      }
      int r0 = 16 * wv;
      f32x4 acc[8];
      for (int ct = 0; ct < 8; ++ct) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
      for (int k0 = 0; k0 < 128; k0 += 16) {
        bf16x4 a = frag_a(dz_cur, 128, r0, k0);
        for (int ct = 0; ct < 8; ++ct) acc[ct] = mfma16(a, frag_bt(Ws[layer], 128, k0, 16 * ct), acc[ct]);
      }
      for (int ct = 0; ct < 8; ++ct) {
        int col = 16 * ct + (l & 15);
        float bsum = 0.f;
        for (int i = 0; i < 4; ++i) {
          int row = r0 + 4 * (l >> 4) + i;
          float d = acc[ct][i] * dsilu(zprev[(size_t)(row0 + row) * 128 + col]);
          dz_next[row * 128 + col] = f2bf(d);
          bsum += d;
        }
        atomicAdd(&s_bias[(1 - layer) * 128 + col], bsum);
      }
      __syncthreads();
      dz_cur = dz_next;
    }
    // dW1 += X^T dZ1   (16 x 128): wave w col tiles {2w, 2w+1}
    for (int i = 0; i < 2; ++i) {
      int ct = 2 * wv + i;
      for (int k0 = 0; k0 < TILE; k0 += 16)
        g1[i] = mfma16(frag_at(X + (size_t)row0 * 16, 16, 0, k0), frag_b(dz_cur, 128, k0, 16 * ct), g1[i]);
    }
  }
  __syncthreads();
  float* P = partial + (size_t)blockIdx.x * MLP_NPARAM_PAD;
  int col = l & 15;
  for (int i = 0; i < 2; ++i) {
    int rt = 2 * wv + i;
    for (int q = 0; q < 4; ++q) {
      int row = 4 * (l >> 4) + q;
      P[MLP_OFF_W4 + (16 * rt + row) * 16 + col] = g4[i][q];
      P[MLP_OFF_W1 + row * 128 + 16 * (2 * wv + i) + col] = g1[i][q];
      for (int ct = 0; ct < 8; ++ct) {
        P[MLP_OFF_W3 + (16 * rt + row) * 128 + 16 * ct + col] = g3[8 * i + ct][q];
        P[MLP_OFF_W2 + (16 * rt + row) * 128 + 16 * ct + col] = g2[8 * i + ct][q];
      }
    }
  }
  for (int i = threadIdx.x; i < 128; i += 256) {
    P[MLP_OFF_B1 + i] = s_bias[i];
    P[MLP_OFF_B2 + i] = s_bias[128 + i];
    P[MLP_OFF_B3 + i] = s_bias[256 + i];
  }
  if (threadIdx.x < 16) P[MLP_OFF_B4 + threadIdx.x] = threadIdx.x < 2 ? s_bias[384 + threadIdx.x] : 0.f;
}

__global__ void k_mlp_reduce(const float* partial, int parts, float* grad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NPARAM_PAD) return;
  float s = 0.f;
  for (int p = 0; p < parts; ++p) s += partial[(size_t)p * MLP_NPARAM_PAD + i];
  grad[i] = s;
}

// Adam on the padded FP32 master weights; padding entries stay zero.
__global__ void k_mlp_adam(float* w, float* mom, float* vel, const float* grad, const uint8_t* mask,
                           uint16_t* wbf, float lr, float b1c, float b2c) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NPARAM_PAD) return;
  if (!mask[i]) return;
  float g = grad[i];
  float m = 0.9f * mom[i] + 0.1f * g;
  float v = 0.999f * vel[i] + 0.001f * g * g;
  mom[i] = m;
  vel[i] = v;
  float x = w[i] - lr * (m / b1c) / (sqrtf(v / b2c) + 1e-8f);
  w[i] = x;
  wbf[i] = f2bf(x);
}

__global__ void k_f32_to_bf16(const float* in, uint16_t* out, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f2bf(in[i]);
}

// features (n x 14 FP32) -> padded BF16 rows (rows x 16), zero rows beyond n
__global__ void k_pack_features(const float* f, long n, long rows, uint16_t* X) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * 16) return;
  long r = i >> 4;
  int cc = (int)(i & 15);
  X[i] = (r < n && cc < MLP_IN) ? f2bf(f[r * MLP_IN + cc]) : (uint16_t)0;
}

}  // namespace

// ------------------------------------------------------------------ host

MlpDev mlp_view(const MlpParams& p) {
  MlpDev m;
  m.w1 = p.wbf + MLP_OFF_W1;
  m.w2 = p.wbf + MLP_OFF_W2;
  m.w3 = p.wbf + MLP_OFF_W3;
  m.w4 = p.wbf + MLP_OFF_W4;
  m.b1 = p.w + MLP_OFF_B1;
  m.b2 = p.w + MLP_OFF_B2;
  m.b3 = p.w + MLP_OFF_B3;
  m.b4 = p.w + MLP_OFF_B4;
  return m;
}

// padded index of logical parameter k of the published blob layout
int mlp_padded_index(int k) {
  if (k < MLP_IN * 128) return MLP_OFF_W1 + k;  // W1 rows 0..13 of 16
  k -= MLP_IN * 128;
  if (k < 128) return MLP_OFF_B1 + k;
  k -= 128;
  if (k < 128 * 128) return MLP_OFF_W2 + k;
  k -= 128 * 128;
  if (k < 128) return MLP_OFF_B2 + k;
  k -= 128;
  if (k < 128 * 128) return MLP_OFF_W3 + k;
  k -= 128 * 128;
  if (k < 128) return MLP_OFF_B3 + k;
  k -= 128;
  if (k < 128 * 2) return MLP_OFF_W4 + (k / 2) * 16 + (k % 2);
  k -= 128 * 2;
  return MLP_OFF_B4 + k;
}

shp_status mlp_params_alloc(MlpParams* p) {
  if (hipMalloc(&p->w, MLP_NPARAM_PAD * sizeof(float)) != hipSuccess) return SHP_E_OOM;
  if (hipMalloc(&p->wbf, MLP_NPARAM_PAD * sizeof(uint16_t)) != hipSuccess) return SHP_E_OOM;
  return SHP_OK;
}

void mlp_params_free(MlpParams* p) {
  if (p->w) (void)hipFree(p->w);
  if (p->wbf) (void)hipFree(p->wbf);
  p->w = nullptr;
  p->wbf = nullptr;
}

shp_status mlp_params_upload(MlpParams* p, const float* params, hipStream_t s) {
  std::vector<float> pad(MLP_NPARAM_PAD, 0.0f);
  for (int k = 0; k < MLP_NPARAM; ++k) pad[mlp_padded_index(k)] = params[k];
  if (hipMemcpyAsync(p->w, pad.data(), pad.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess) return SHP_E_HIP;
  k_f32_to_bf16<<<(MLP_NPARAM_PAD + 255) / 256, 256, 0, s>>>(p->w, p->wbf, MLP_NPARAM_PAD);
  if (hipStreamSynchronize(s) != hipSuccess) return SHP_E_HIP;
  return SHP_OK;
}

void mlp_launch_forward(const MlpParams& p, const uint16_t* X, int rows, float* Y, hipStream_t s) {
  TrainBufs tb{};
  int tiles = (rows + TILE - 1) / TILE;
  int grid = tiles < 1024 ? (tiles > 0 ? tiles : 1) : 1024;
  k_mlp_forward<<<grid, 256, 0, s>>>(mlp_view(p), X, rows, Y, tb, 0);
}

// ---- engine hooks ------------------------------------------------------------

// Learned warm start of the step's cache misses: tiles of 64 misses, the
// features rounded to BF16, the forward pass on the matrix cores, then the
// output head lambda_n = max(y0, 0) / mn, lambda_t = y1 / mt projected onto
// |lambda_t| <= mu * lambda_n.  The contact is marked warm (2) so the
// solver applies the predicted impulses like a cached solution (the
// coloured solver's manifold record, already written, gets them too); a
// non-finite prediction keeps the zero start.
__global__ __launch_bounds__(256) void k_mlp_infer(Ctx c, MlpDev m) {
  __shared__ uint16_t sx[TILE * 16];
  __shared__ uint16_t s1[TILE * 128];
  __shared__ uint16_t s2[TILE * 128];
  __shared__ float sy[TILE * 16];
  // the weights, transposed, and biases, once per workgroup for all its
  // tiles: a B fragment is then one 8-byte LDS read
  __shared__ __align__(16) uint16_t s_w1[128 * WT1], s_w2[128 * WT2], s_w3[128 * WT2], s_w4[16 * WT2];
  __shared__ float s_b[3 * 128 + 16];
  int n = c.counters[CNT_MISS];
  if (n > c.miss_cap) n = c.miss_cap;
  if ((int)blockIdx.x * TILE >= n) return;  // no tile: skip the staging (no barrier follows)
  for (int i = threadIdx.x; i < 128 * 128; i += 256) {
    const int k = i >> 7, o = i & 127;  // read row-major [k][o], coalesced
    s_w2[o * WT2 + k] = m.w2[i];
    s_w3[o * WT2 + k] = m.w3[i];
    if (k < 16) s_w1[o * WT1 + k] = m.w1[i];
    if (o < 16) s_w4[o * WT2 + k] = m.w4[k * 16 + o];
  }
  for (int i = threadIdx.x; i < 3 * 128 + 16; i += 256)
    s_b[i] = i < 128 ? m.b1[i] : i < 256 ? m.b2[i - 128] : i < 384 ? m.b3[i - 256] : m.b4[i - 384];
  __syncthreads();
  MlpDev ml;
  ml.w1 = s_w1;
  ml.w2 = s_w2;
  ml.w3 = s_w3;
  ml.w4 = s_w4;
  ml.b1 = s_b;
  ml.b2 = s_b + 128;
  ml.b3 = s_b + 256;
  ml.b4 = s_b + 384;
  for (int t = blockIdx.x; t * TILE < n; t += gridDim.x) {
    int row0 = t * TILE;
    // the output head's inputs of this lane's row (first wavefront), loaded
    // with the features so that the head waits on no load after the layers
    static_assert(TILE == 64, "one row per lane of the first wavefront");
    const bool row_ok = threadIdx.x < TILE && row0 + (int)threadIdx.x < n;
    int hk = 0, hmr = -1;
    double hmcn = 1.0, hmct = 1.0, hmu = 0.0;
    if (row_ok) {
      hk = c.miss_idx[row0 + threadIdx.x];
      hmr = c.solver == SHP_SOLVER_COLORED_F32 ? c.miss_rec[row0 + threadIdx.x] : -1;
      hmcn = c.mcn[hk];
      hmct = c.mct[hk];
      hmu = (c.mu[c.ga[hk]] + c.mu[c.gb[hk]]) / 2.0;
    }
    if (c.features_in_contacts) {
      for (int i = threadIdx.x; i < TILE * 16; i += 256) {
        int r = i >> 4, q = i & 15;
        float f = (row0 + r < n && q < MLP_IN) ? c.miss_feat[(size_t)(row0 + r) * MLP_IN + q] : 0.0f;
        sx[i] = f2bf(f);
      }
    } else if (threadIdx.x < TILE) {
      // one row per lane of the first wavefront: the features of the
      // stored contact, as the contact kernel would compute them (the
      // velocities are those it saw: nothing has changed them since)
      const int r = threadIdx.x;
      float f[MLP_IN];
      for (int q = 0; q < MLP_IN; ++q) f[q] = 0.0f;
      if (row_ok) {
        const int k = hk;
        const int ga = c.ga[k], gb = c.gb[k];
        const double v[6] = {c.vx[ga], c.vy[ga], c.w[ga], c.vx[gb], c.vy[gb], c.w[gb]};
        const double sgn = c.flip[k] ? -1.0 : 1.0;
        mlp_features(sgn * c.cnx[k], sgn * c.cny[k], c.depth[k], v, c.rax[k], c.ray[k], c.rbx[k],
                     c.rby[k], c.iml[ga], c.iml[gb], c.imr[ga], c.imr[gb],
                     (c.mu[ga] + c.mu[gb]) / 2.0, f);
      }
      for (int q = 0; q < 16; ++q) sx[r * 16 + q] = f2bf(q < MLP_IN ? f[q] : 0.0f);
    }
    __syncthreads();
    forward_tile<true>(ml, sx, s1, s2, sy, nullptr, row0);
    if (threadIdx.x < TILE) {  // the first wavefront: one row per lane
      bool set = false;
      if (row_ok) {
        int r = threadIdx.x;
        int k = hk;
        double ln = (double)sy[r * 16];
        ln = ln > 0.0 ? ln / hmcn : 0.0;
        double lt = (double)sy[r * 16 + 1] / hmct;
        double mu = hmu;
        double thr = ln * mu;
        lt = lt > thr ? thr : (lt < -thr ? -thr : lt);
        if (ln - ln == 0.0 && lt - lt == 0.0) {
          c.ln[k] = ln;
          c.lt[k] = lt;
          c.warm[k] = 2;
          set = true;
          if (hmr >= 0) shp::patch_record_warm(shp::packf(c), hmr >> 1, hmr & 1, ln, lt);
        }
      }
      const unsigned long long setm = __ballot(set);  // one atomic per tile
      if (threadIdx.x == 0 && setm) atomicAdd(&c.counters[CNT_LEARNED], (int)__popcll(setm));
    }
    __syncthreads();
  }
}

// After the solve: append (features, lambda_n * mn, lambda_t * mt) of the
// step's cache misses to the training log.
__global__ void k_mlp_log(Ctx c) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  int n = c.counters[CNT_MISS];
  if (n > c.miss_cap) n = c.miss_cap;
  if (s >= n) return;
  unsigned long long pos = atomicAdd(c.log_count, 1ull);
  if ((long)pos >= c.log_cap) return;
  int k = c.miss_idx[s];
  for (int q = 0; q < MLP_IN; ++q) c.log_feat[pos * MLP_IN + q] = c.miss_feat[(size_t)s * MLP_IN + q];
  c.log_tgt[pos * 2] = (float)(c.ln[k] * c.mcn[k]);
  c.log_tgt[pos * 2 + 1] = (float)(c.lt[k] * c.mct[k]);
}

void mlp_launch_infer(shp_batch* b) {
  MlpDev m;
  m.w1 = b->ctx.mlp_wbf + MLP_OFF_W1;
  m.w2 = b->ctx.mlp_wbf + MLP_OFF_W2;
  m.w3 = b->ctx.mlp_wbf + MLP_OFF_W3;
  m.w4 = b->ctx.mlp_wbf + MLP_OFF_W4;
  m.b1 = b->ctx.mlp_w + MLP_OFF_B1;
  m.b2 = b->ctx.mlp_w + MLP_OFF_B2;
  m.b3 = b->ctx.mlp_w + MLP_OFF_B3;
  m.b4 = b->ctx.mlp_w + MLP_OFF_B4;
  k_mlp_infer<<<b->plan.g_infer, 256, 0, b->stream>>>(b->ctx, m);
}

// sx, s1, s2 (BF16) and sy (FP32) of k_mlp_infer
size_t mlp_infer_static_lds() {
  return lds16(TILE * 16 * 2) + 2 * lds16(TILE * 128 * 2) + lds16(TILE * 16 * 4) +
         lds16(128 * WT1 * 2) + 2 * lds16(128 * WT2 * 2) + lds16(16 * WT2 * 2) +
         lds16((3 * 128 + 16) * 4);
}

void mlp_launch_log(shp_batch* b) {
  k_mlp_log<<<b->plan.g_log, 256, 0, b->stream>>>(b->ctx);
}

// ---- training handle -------------------------------------------------------

struct shp_mlp {
  int device = 0;  // the GPU the handle's weights live on
  hipStream_t stream = nullptr;
  MlpParams p{};
  float *mom = nullptr, *vel = nullptr, *grad = nullptr, *partial = nullptr;
  uint8_t* mask = nullptr;
  int parts = 64;
  int cap_rows = 0;
  uint16_t* X = nullptr;   // cap_rows x 16
  float* T = nullptr;      // cap_rows x 2
  float* Fin = nullptr;    // cap_rows x 14 staging
  TrainBufs tb{};
  long adam_t = 0;
  // training set on device
  uint16_t* dX = nullptr;
  float* dT = nullptr;
  long n_data = 0;
  float huber = 0.f;  // 0: squared error
};

static shp_status mlp_fail(shp_status s, const char* m) {
  shp_set_error(m);
  return s;
}

// ---- forward pass that allocates nothing (shp_mlp_forward_noalloc) --------
// Its staging - the padded BF16 input rows and the FP32 output rows of up
// to MLP_NOALLOC_MAX_ROWS rows (96 B a row, 6 MiB) - lives in this
// file-scope device array instead of the handle's buffers.  Each GPU holds
// its own copy until the process ends; every call on a GPU shares it.
#define MLP_NOALLOC_MAX_ROWS 65536
static __device__ __align__(256) unsigned char scratch[(size_t)MLP_NOALLOC_MAX_ROWS * 16 * (2 + 4)];
#define MLP_MAX_DEVICES 64
static void* g_scratch_ptr[MLP_MAX_DEVICES];  // device address of `scratch` per GPU
static std::mutex g_scratch_mu;

// The device address of `scratch` on the current GPU, taken once per GPU
// (its host-side name is not a device address).
static shp_status scratch_setup(int dev) {
  if (dev < 0 || dev >= MLP_MAX_DEVICES) return mlp_fail(SHP_E_UNSUPPORTED, "device index");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return SHP_OK;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
    return mlp_fail(SHP_E_HIP, "hipGetSymbolAddress(scratch) failed");
  g_scratch_ptr[dev] = ptr;
  return SHP_OK;
}

// out[2 r + k] = y[16 r + k], k < 2, r < n
__global__ void k_take_outputs(const float* y, int n, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2 * n) out[i] = y[(size_t)(i >> 1) * 16 + (i & 1)];
}

#define MCHK(x) \
  do { if ((x) != hipSuccess) return mlp_fail(SHP_E_HIP, #x); } while (0)

static shp_status mlp_reserve(shp_mlp* h, int rows) {
  if (rows <= h->cap_rows) return SHP_OK;
  rows = ((rows + TILE - 1) / TILE) * TILE;
  void* ptrs[] = {h->X, h->T, h->Fin, h->tb.z1, h->tb.z2, h->tb.z3, h->tb.h1, h->tb.h2, h->tb.h3, h->tb.y};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  size_t R = rows;
  MCHK(hipMalloc(&h->X, R * 16 * 2));
  MCHK(hipMalloc(&h->T, R * 2 * 4));
  MCHK(hipMalloc(&h->Fin, R * MLP_IN * 4));
  MCHK(hipMalloc(&h->tb.z1, R * 128 * 4));
  MCHK(hipMalloc(&h->tb.z2, R * 128 * 4));
  MCHK(hipMalloc(&h->tb.z3, R * 128 * 4));
  MCHK(hipMalloc(&h->tb.h1, R * 128 * 2));
  MCHK(hipMalloc(&h->tb.h2, R * 128 * 2));
  MCHK(hipMalloc(&h->tb.h3, R * 128 * 2));
  MCHK(hipMalloc(&h->tb.y, R * 16 * 4));
  h->cap_rows = rows;
  return SHP_OK;
}

// rows of X/T are taken from the device training set by index
__global__ void k_gather_rows(const uint16_t* dX, const float* dT, const int32_t* idx, int n, int rows,
                              uint16_t* X, float* T) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * 16) return;
  int r = i >> 4, cc = i & 15;
  if (r < n) {
    long src = idx[r];
    X[i] = dX[src * 16 + cc];
    if (cc < 2) T[r * 2 + cc] = dT[src * 2 + cc];
  } else {
    X[i] = 0;
    if (cc < 2) T[r * 2 + cc] = 0.f;
  }
}

static shp_status mlp_grad_rows(shp_mlp* h, int n) {
  int rows = ((n + TILE - 1) / TILE) * TILE;
  int tiles = rows / TILE;
  int fg = tiles < 1024 ? tiles : 1024;
  MlpDev m = mlp_view(h->p);
  k_mlp_forward<<<fg, 256, 0, h->stream>>>(m, h->X, rows, h->tb.y, h->tb, 1);
  int parts = tiles < h->parts ? tiles : h->parts;
  k_mlp_backward<<<parts, 256, 0, h->stream>>>(m, h->X, h->T, rows, n, h->tb, h->partial, h->huber);
  k_mlp_reduce<<<(MLP_NPARAM_PAD + 255) / 256, 256, 0, h->stream>>>(h->partial, parts, h->grad);
  MCHK(hipGetLastError());
  return SHP_OK;
}

extern "C" {

int32_t shp_mlp_param_count(void) { return MLP_NPARAM; }

shp_status shp_mlp_create(const float* params, shp_mlp** out) {
  if (!params || !out) return mlp_fail(SHP_E_INVAL, "null argument");
  shp_mlp* h = new shp_mlp();
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return mlp_fail(SHP_E_HIP, "no HIP device");
  }
  if (scratch_setup(h->device) != SHP_OK) {
    delete h;
    return SHP_E_HIP;
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return mlp_fail(SHP_E_HIP, "stream");
  }
  shp_status st = mlp_params_alloc(&h->p);
  if (st == SHP_OK) st = mlp_params_upload(&h->p, params, h->stream);
  if (st == SHP_OK) {
    if (hipMalloc(&h->mom, MLP_NPARAM_PAD * 4) != hipSuccess || hipMalloc(&h->vel, MLP_NPARAM_PAD * 4) != hipSuccess ||
        hipMalloc(&h->grad, MLP_NPARAM_PAD * 4) != hipSuccess ||
        hipMalloc(&h->partial, (size_t)h->parts * MLP_NPARAM_PAD * 4) != hipSuccess ||
        hipMalloc(&h->mask, MLP_NPARAM_PAD) != hipSuccess)
      st = SHP_E_OOM;
  }
  if (st == SHP_OK) {
    std::vector<uint8_t> mask(MLP_NPARAM_PAD, 0);
    for (int k = 0; k < MLP_NPARAM; ++k) mask[mlp_padded_index(k)] = 1;
    std::vector<float> zero(MLP_NPARAM_PAD, 0.f);
    if (hipMemcpy(h->mask, mask.data(), MLP_NPARAM_PAD, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->mom, zero.data(), MLP_NPARAM_PAD * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->vel, zero.data(), MLP_NPARAM_PAD * 4, hipMemcpyHostToDevice) != hipSuccess)
      st = SHP_E_HIP;
  }
  if (st != SHP_OK) {
    shp_mlp_destroy(h);
    return mlp_fail(st, "mlp allocation failed");
  }
  *out = h;
  return SHP_OK;
}

shp_status shp_mlp_destroy(shp_mlp* h) {
  if (!h) return SHP_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  mlp_params_free(&h->p);
  void* ptrs[] = {h->mom, h->vel, h->grad, h->partial, h->mask, h->X, h->T, h->Fin, h->tb.z1,
                  h->tb.z2, h->tb.z3, h->tb.h1, h->tb.h2, h->tb.h3, h->tb.y, h->dX, h->dT};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return SHP_OK;
}

shp_status shp_mlp_get_params(shp_mlp* h, float* params) {
  if (!h || !params) return mlp_fail(SHP_E_INVAL, "null argument");
  std::vector<float> pad(MLP_NPARAM_PAD);
  MCHK(hipStreamSynchronize(h->stream));
  MCHK(hipMemcpy(pad.data(), h->p.w, MLP_NPARAM_PAD * 4, hipMemcpyDeviceToHost));
  for (int k = 0; k < MLP_NPARAM; ++k) params[k] = pad[mlp_padded_index(k)];
  return SHP_OK;
}

shp_status shp_mlp_forward(shp_mlp* h, const float* feats, int32_t n, float* out) {
  if (!h || n < 0 || (n > 0 && (!feats || !out))) return mlp_fail(SHP_E_INVAL, "bad argument");
  if (n == 0) return SHP_OK;
  shp_status st = mlp_reserve(h, n);
  if (st != SHP_OK) return st;
  int rows = ((n + TILE - 1) / TILE) * TILE;
  MCHK(hipMemcpyAsync(h->Fin, feats, (size_t)n * MLP_IN * 4, hipMemcpyHostToDevice, h->stream));
  k_pack_features<<<(rows * 16 + 255) / 256, 256, 0, h->stream>>>(h->Fin, n, rows, h->X);
  mlp_launch_forward(h->p, h->X, rows, h->tb.y, h->stream);
  std::vector<float> y((size_t)rows * 16);
  MCHK(hipMemcpyAsync(y.data(), h->tb.y, y.size() * 4, hipMemcpyDeviceToHost, h->stream));
  MCHK(hipStreamSynchronize(h->stream));
  for (int r = 0; r < n; ++r) {
    out[2 * r] = y[(size_t)r * 16];
    out[2 * r + 1] = y[(size_t)r * 16 + 1];
  }
  return SHP_OK;
}

shp_status shp_mlp_forward_noalloc(shp_mlp* h, const float* d_feats, int32_t n, float* d_out,
                                   void* stream) {
  if (!h || n < 0 || (n > 0 && (!d_feats || !d_out))) return mlp_fail(SHP_E_INVAL, "bad argument");
  if (n > MLP_NOALLOC_MAX_ROWS)
    return mlp_fail(SHP_E_INVAL, "n exceeds the 65536 rows of the static scratch");
  if (n == 0) return SHP_OK;
  int dev = -1;
  MCHK(hipGetDevice(&dev));
  if (dev != h->device) return mlp_fail(SHP_E_INVAL, "the current device is not the handle's");
  void* base = g_scratch_ptr[dev];
  if (!base) return mlp_fail(SHP_E_HIP, "scratch not set up on this device");
  uint16_t* X = (uint16_t*)base;                                                    // rows x 16 BF16
  float* Y = (float*)((unsigned char*)base + (size_t)MLP_NOALLOC_MAX_ROWS * 16 * 2);  // rows x 16
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  const int rows = ((n + TILE - 1) / TILE) * TILE;
  // the rows n..rows-1 are packed as zeros, so nothing needs clearing
  k_pack_features<<<(rows * 16 + 255) / 256, 256, 0, s>>>(d_feats, n, rows, X);
  mlp_launch_forward(h->p, X, rows, Y, s);
  k_take_outputs<<<(2 * n + 255) / 256, 256, 0, s>>>(Y, n, d_out);
  MCHK(hipGetLastError());
  return SHP_OK;
}

shp_status shp_mlp_gradient(shp_mlp* h, const float* feats, const float* targets, int32_t n,
                            float* grad_out, float* loss_out) {
  if (!h || n <= 0 || !feats || !targets || !grad_out) return mlp_fail(SHP_E_INVAL, "bad argument");
  shp_status st = mlp_reserve(h, n);
  if (st != SHP_OK) return st;
  int rows = ((n + TILE - 1) / TILE) * TILE;
  MCHK(hipMemcpyAsync(h->Fin, feats, (size_t)n * MLP_IN * 4, hipMemcpyHostToDevice, h->stream));
  std::vector<float> t((size_t)rows * 2, 0.f);
  for (int i = 0; i < 2 * n; ++i) t[i] = targets[i];
  MCHK(hipMemcpyAsync(h->T, t.data(), t.size() * 4, hipMemcpyHostToDevice, h->stream));
  k_pack_features<<<(rows * 16 + 255) / 256, 256, 0, h->stream>>>(h->Fin, n, rows, h->X);
  st = mlp_grad_rows(h, n);
  if (st != SHP_OK) return st;
  std::vector<float> g(MLP_NPARAM_PAD), y((size_t)rows * 16);
  MCHK(hipMemcpyAsync(g.data(), h->grad, g.size() * 4, hipMemcpyDeviceToHost, h->stream));
  MCHK(hipMemcpyAsync(y.data(), h->tb.y, y.size() * 4, hipMemcpyDeviceToHost, h->stream));
  MCHK(hipStreamSynchronize(h->stream));
  for (int k = 0; k < MLP_NPARAM; ++k) grad_out[k] = g[mlp_padded_index(k)];
  if (loss_out) {
    double l = 0.0;
    for (int r = 0; r < n; ++r)
      for (int q = 0; q < 2; ++q) {
        double d = (double)y[(size_t)r * 16 + q] - targets[2 * r + q];
        l += d * d;
      }
    *loss_out = (float)(l / n);
  }
  return SHP_OK;
}

shp_status shp_mlp_set_loss(shp_mlp* h, float huber_delta) {
  if (!h || !(huber_delta >= 0.f))
    return mlp_fail(SHP_E_INVAL, "huber_delta must be >= 0");
  h->huber = huber_delta;
  return SHP_OK;
}

shp_status shp_mlp_set_data(shp_mlp* h, const float* feats, const float* targets, int64_t n) {
  if (!h || n <= 0 || !feats || !targets) return mlp_fail(SHP_E_INVAL, "bad argument");
  if (h->dX) (void)hipFree(h->dX);
  if (h->dT) (void)hipFree(h->dT);
  h->dX = nullptr;
  h->dT = nullptr;
  MCHK(hipMalloc(&h->dX, (size_t)n * 16 * 2));
  MCHK(hipMalloc(&h->dT, (size_t)n * 2 * 4));
  float* stage = nullptr;
  const long chunk = 1 << 20;
  MCHK(hipMalloc(&stage, (size_t)chunk * MLP_IN * 4));
  for (long o = 0; o < n; o += chunk) {
    long m = n - o < chunk ? n - o : chunk;
    MCHK(hipMemcpy(stage, feats + o * MLP_IN, (size_t)m * MLP_IN * 4, hipMemcpyHostToDevice));
    k_pack_features<<<(int)((m * 16 + 255) / 256), 256, 0, h->stream>>>(stage, m, m, h->dX + o * 16);
    MCHK(hipStreamSynchronize(h->stream));
  }
  (void)hipFree(stage);
  MCHK(hipMemcpy(h->dT, targets, (size_t)n * 2 * 4, hipMemcpyHostToDevice));
  h->n_data = n;
  return SHP_OK;
}

// `steps` Adam steps on mini-batches whose row indices (steps x batch,
// into the training set) are given by the caller; loss_out[s] receives the
// mean squared error of step s's batch before its update.
shp_status shp_mlp_train(shp_mlp* h, const int32_t* indices, int32_t steps, int32_t batch, float lr,
                         float* loss_out) {
  if (!h || !indices || steps <= 0 || batch <= 0 || batch % TILE) return mlp_fail(SHP_E_INVAL, "bad argument");
  if (!h->dX) return mlp_fail(SHP_E_INVAL, "no training data");
  for (long i = 0; i < (long)steps * batch; ++i)
    if (indices[i] < 0 || indices[i] >= h->n_data) return mlp_fail(SHP_E_INVAL, "training index out of range");
  shp_status st = mlp_reserve(h, batch);
  if (st != SHP_OK) return st;
  int32_t* didx = nullptr;
  MCHK(hipMalloc(&didx, (size_t)steps * batch * 4));
  MCHK(hipMemcpy(didx, indices, (size_t)steps * batch * 4, hipMemcpyHostToDevice));
  std::vector<float> y((size_t)batch * 16), t((size_t)batch * 2);
  for (int s = 0; s < steps; ++s) {
    k_gather_rows<<<(batch * 16 + 255) / 256, 256, 0, h->stream>>>(h->dX, h->dT, didx + (size_t)s * batch,
                                                                  batch, batch, h->X, h->T);
    st = mlp_grad_rows(h, batch);
    if (st != SHP_OK) break;
    ++h->adam_t;
    float b1c = 1.0f - powf(0.9f, (float)h->adam_t);
    float b2c = 1.0f - powf(0.999f, (float)h->adam_t);
    k_mlp_adam<<<(MLP_NPARAM_PAD + 255) / 256, 256, 0, h->stream>>>(h->p.w, h->mom, h->vel, h->grad, h->mask,
                                                                    h->p.wbf, lr, b1c, b2c);
    if (loss_out) {
      MCHK(hipMemcpyAsync(y.data(), h->tb.y, y.size() * 4, hipMemcpyDeviceToHost, h->stream));
      MCHK(hipMemcpyAsync(t.data(), h->T, t.size() * 4, hipMemcpyDeviceToHost, h->stream));
      MCHK(hipStreamSynchronize(h->stream));
      double l = 0.0;
      for (int r = 0; r < batch; ++r)
        for (int q = 0; q < 2; ++q) {
          double d = (double)y[(size_t)r * 16 + q] - t[(size_t)r * 2 + q];
          l += d * d;
        }
      loss_out[s] = (float)(l / batch);
    }
  }
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(didx);
  return st;
}

}  // extern "C"
