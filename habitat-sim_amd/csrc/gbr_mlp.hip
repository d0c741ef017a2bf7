// Chunk-visibility MLP (32 -> 256 -> 256 -> 1, ReLU) on gfx950 MFMA:
// BF16 operands, FP32 accumulation, FP32 master weights with Adam.
//
// Every matrix product is an "NT" GEMM, C[M][N] = sum_k A[M][K] * Bt[N][K],
// with both operands K-contiguous so that a wave's MFMA fragments are single
// 16-byte loads:
//   forward   H1 = relu(X  W1^T + b1)        A = X   [B][32],  Bt = W1  [256][32]
//             H2 = relu(H1 W2^T + b2)        A = H1  [B][256], Bt = W2  [256][256]
//   backward  dW2 = dH2^T H1                 A = dH2^T [256][B], Bt = H1^T [256][B]
//             dH1 = (dH2 W2) o [H1 > 0]      A = dH2 [B][256], Bt = W2^T [256][256]
//             dW1 = dH1^T X                  A = dH1^T [256][B], Bt = X^T  [32][B]
// The last layer (256 -> 1), the loss and the bias gradients are reductions
// (no atomics: each output has one writer).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>

#include "gbr_mlp.h"

namespace gbr {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// float -> bf16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ int rows_of(int M, const uint32_t* M_dev) {
  return M_dev ? min(M, (int)*M_dev) : M;
}

// --------------------------------------------------------------------------
// NT GEMM, one 32x32 output tile per wave (v_mfma_f32_32x32x16_bf16):
// lane l supplies A[m0 + l%32][k + 8*(l/32) .. +8] and Bt[n0 + l%32][same k];
// accumulator v holds C[m0 + 8*(v/4) + 4*(l/32) + v%4][n0 + l%32].
// Requirements (checked on the host): K % 16 == 0, lda % 8 == 0,
// ldb % 8 == 0, 16-byte aligned base pointers.
// --------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_gemm_nt(GemmArgs g) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int M = rows_of(g.M, g.M_dev);
  const int tiles_n = (g.N + 31) / 32;
  const int tile = blockIdx.x * 4 + wave;
  const int tiles_m = (M + 31) / 32;
  if (tile >= tiles_m * tiles_n) return;
  const int m0 = (tile / tiles_n) * 32, n0 = (tile % tiles_n) * 32;
  const int am = m0 + (lane & 31), bn = n0 + (lane & 31);
  const int kofs = 8 * (lane >> 5);
  const bool arow = am < M, brow = bn < g.N;
  const uint16_t* ap = g.A + (size_t)am * g.lda + kofs;
  const uint16_t* bp = g.Bt + (size_t)bn * g.ldb + kofs;
  floatx16 acc = {};
  for (int k = 0; k < g.K; k += 16) {
    bf16x8 a = {}, b = {};
    if (arow) a = *reinterpret_cast<const bf16x8*>(ap + k);
    if (brow) b = *reinterpret_cast<const bf16x8*>(bp + k);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  const int n = bn;
  if (n >= g.N) return;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int m = m0 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
    if (m >= M) continue;
    float x = acc[v] * g.alpha;
    if (g.C) g.C[(size_t)m * g.ldc + n] = x;
    if (g.bias) x = x + g.bias[n];
    if (g.mask) x = bf2f(g.mask[(size_t)m * g.ldmask + n]) > 0.0f ? x : 0.0f;
    if (g.relu) x = x > 0.0f ? x : 0.0f;
    if (g.D) g.D[(size_t)m * g.ldd + n] = f2bf(x);
    if (g.Dt) g.Dt[(size_t)n * g.lddt + m] = f2bf(x);
  }
}

// logit[r] = sum_k H2[r][k] * w3[k] + b3, one wave per row
__global__ __launch_bounds__(256) void k_last_layer(const uint16_t* __restrict__ H2,
                                                    const uint16_t* __restrict__ w3,
                                                    const float* __restrict__ b3,
                                                    float* __restrict__ logits, int rows,
                                                    const uint32_t* __restrict__ rows_dev) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= rows_of(rows, rows_dev)) return;
  float s = 0.0f;
  for (int k = lane; k < kMlpHidden; k += 64)
    s += bf2f(H2[(size_t)row * kMlpHidden + k]) * bf2f(w3[k]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) logits[row] = s + *b3;
}

// weighted binary cross-entropy with logits, per row:
//   loss_r = w_r [log(1 + e^-|z|) + max(z, 0) - z y],  dl_r = w_r (sigmoid(z) - y) / rows
// with w_r = pos_weight for visible rows, 1 otherwise
__global__ __launch_bounds__(256) void k_head(const float* __restrict__ logits,
                                              const uint8_t* __restrict__ labels, float pos_weight,
                                              int rows, float* __restrict__ dl,
                                              float* __restrict__ loss) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= rows) return;
  const float z = logits[r];
  const float y = labels[r] ? 1.0f : 0.0f;
  const float w = labels[r] ? pos_weight : 1.0f;
  const float sig = 1.0f / (1.0f + expf(-z));
  dl[r] = w * (sig - y) / (float)rows;
  loss[r] = w * (log1pf(expf(-fabsf(z))) + fmaxf(z, 0.0f) - z * y);
}

// dH2 = [H2 > 0] dl[r] w3[k], row-major [rows][256] and transposed [256][rows]
// (two passes so that each store is coalesced)
__global__ __launch_bounds__(256) void k_dh2(const uint16_t* __restrict__ H2,
                                             const uint16_t* __restrict__ H2t,
                                             const float* __restrict__ dl,
                                             const uint16_t* __restrict__ w3, int rows,
                                             uint16_t* __restrict__ dH2,
                                             uint16_t* __restrict__ dH2t) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t n = (size_t)rows * kMlpHidden;
  if (i < n) {
    const size_t r = i / kMlpHidden, k = i % kMlpHidden;
    dH2[i] = bf2f(H2[i]) > 0.0f ? f2bf(dl[r] * bf2f(w3[k])) : (uint16_t)0;
  } else if (i < 2 * n) {
    const size_t j = i - n, k = j / rows, r = j % rows;
    dH2t[j] = bf2f(H2t[j]) > 0.0f ? f2bf(dl[r] * bf2f(w3[k])) : (uint16_t)0;
  }
}

// out[j] = sum_r Mt[j][r] * (w ? w[r] : 1), one wave per j
__global__ __launch_bounds__(256) void k_row_sums(const uint16_t* __restrict__ Mt,
                                                  const float* __restrict__ w, int n, int rows,
                                                  float* __restrict__ out) {
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= n) return;
  float s = 0.0f;
  for (int r = lane; r < rows; r += 64) {
    const float x = bf2f(Mt[(size_t)j * rows + r]);
    s += w ? x * w[r] : x;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) out[j] = s;
}

// out[0] = sum of v[0..n) (one workgroup)
__global__ __launch_bounds__(256) void k_sum(const float* __restrict__ v, int n,
                                             float* __restrict__ out) {
  __shared__ float s_part[4];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += 256) s += v[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = (s_part[0] + s_part[1]) + (s_part[2] + s_part[3]);
}

// minibatch: row i of the batch = dataset row hash(seed, step, i) mod total,
// copied row-major and transposed, with its label
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_gather(const uint16_t* __restrict__ X,
                                                const uint8_t* __restrict__ Y, uint64_t total,
                                                uint64_t seed, uint64_t step, int batch,
                                                uint16_t* __restrict__ Xb,
                                                uint16_t* __restrict__ Xbt,
                                                uint8_t* __restrict__ Yb) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= batch) return;
  const uint64_t src = mix64(seed ^ mix64(step * 0x9e3779b97f4a7c15ull + (uint64_t)i)) % total;
#pragma unroll 4
  for (int k = 0; k < kMlpIn; ++k) {
    const uint16_t x = X[src * kMlpIn + k];
    Xb[(size_t)i * kMlpIn + k] = x;
    Xbt[(size_t)k * batch + i] = x;
  }
  Yb[i] = Y[src];
}

// float -> bf16 conversion (weights after the optimiser step), optionally
// also transposed ([rows][cols] -> [cols][rows])
__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ in, size_t n,
                                                 uint16_t* __restrict__ out,
                                                 uint16_t* __restrict__ out_t, int rows, int cols) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint16_t b = f2bf(in[i]);
  out[i] = b;
  if (out_t) out_t[(i % cols) * rows + i / cols] = b;
}

// Adam (Kingma & Ba 2015) on fp32 master weights; bc1/bc2 = bias corrections
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              size_t n, float lr, float b1, float b2, float eps,
                                              float bc1, float bc2) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// ---- launchers --------------------------------------------------------------

static unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

hipError_t mlp_gemm(const GemmArgs& g, hipStream_t s) {
  const int tiles = ((g.M + 31) / 32) * ((g.N + 31) / 32);
  if (tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gemm_nt, dim3((tiles + 3) / 4), dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t mlp_last_layer(const uint16_t* H2, const uint16_t* w3, const float* b3, float* logits,
                          int rows, const uint32_t* rows_dev, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  hipLaunchKernelGGL(k_last_layer, dim3((rows + 3) / 4), dim3(256), 0, s, H2, w3, b3, logits, rows,
                     rows_dev);
  return hipGetLastError();
}

hipError_t mlp_head(const float* logits, const uint8_t* labels, float pos_weight, int rows,
                    float* dl, float* loss, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  hipLaunchKernelGGL(k_head, dim3(blocks_for(rows)), dim3(256), 0, s, logits, labels, pos_weight,
                     rows, dl, loss);
  return hipGetLastError();
}

hipError_t mlp_dh2(const uint16_t* H2, const uint16_t* H2t, const float* dl, const uint16_t* w3,
                   int rows, uint16_t* dH2, uint16_t* dH2t, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dh2, dim3(blocks_for(2 * (size_t)rows * kMlpHidden)), dim3(256), 0, s, H2,
                     H2t, dl, w3, rows, dH2, dH2t);
  return hipGetLastError();
}

hipError_t mlp_row_sums(const uint16_t* Mt, const float* w, int n, int rows, float* out,
                        hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_row_sums, dim3((n + 3) / 4), dim3(256), 0, s, Mt, w, n, rows, out);
  return hipGetLastError();
}

hipError_t mlp_sum(const float* v, int n, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, v, n, out);
  return hipGetLastError();
}

hipError_t mlp_gather(const uint16_t* X, const uint8_t* Y, uint64_t total, uint64_t seed,
                      uint64_t step, int batch, uint16_t* Xb, uint16_t* Xbt, uint8_t* Yb,
                      hipStream_t s) {
  if (batch == 0 || total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3(blocks_for(batch)), dim3(256), 0, s, X, Y, total, seed, step,
                     batch, Xb, Xbt, Yb);
  return hipGetLastError();
}

hipError_t mlp_to_bf16(const float* in, size_t n, uint16_t* out, uint16_t* out_t, int rows, int cols,
                       hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, s, in, n, out, out_t, rows, cols);
  return hipGetLastError();
}

hipError_t mlp_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1,
                    float b2, float eps, float bc1, float bc2, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps,
                     bc1, bc2);
  return hipGetLastError();
}

}  // namespace gbr

namespace gbr {

// --------------------------------------------------------------------------
// Fused scorer for rendering: logits of up to min(rows, *rows_dev) feature
// rows in one pass, one wave per 32 rows.  Layer 1 (K = 32) and layer 2
// (K = 256) on v_mfma_f32_32x32x16_bf16 in two halves of 128 columns; H1
// (bf16, after bias and ReLU) is staged in LDS as the A operand of layer 2;
// H2 is rounded to bf16 like the training forward, multiplied by w3 and
// reduced across the 32 lanes that hold a row's columns.
// --------------------------------------------------------------------------

constexpr int kScoreRows = 32;              // rows per wave and round
constexpr int kH1Stride = kMlpHidden + 8;   // padded LDS row (bank spread)
constexpr int kW2Group = 128;               // W2 rows (output columns) staged at a time
static_assert(kMlpHidden % kW2Group == 0, "hidden width in groups of 4 tiles");

// Per round a workgroup scores 4 tiles of 32 rows (one per wave): layer 1
// from L2-resident W1 into the wave's H1 tile in LDS, then layer 2 in groups
// of 128 output columns, each group of W2 rows staged in LDS once for the 4
// waves; the output layer folds into the layer-2 epilogue.
__global__ __launch_bounds__(256) void k_mlp_score(const uint16_t* __restrict__ X,
                                                   const uint16_t* __restrict__ W1,
                                                   const float* __restrict__ b1,
                                                   const uint16_t* __restrict__ W2,
                                                   const float* __restrict__ b2,
                                                   const uint16_t* __restrict__ w3,
                                                   const float* __restrict__ b3, int rows,
                                                   const uint32_t* __restrict__ rows_dev,
                                                   float* __restrict__ logits) {
  __shared__ uint16_t s_h1[4][kScoreRows * kH1Stride];
  __shared__ uint16_t s_w2[kW2Group * kH1Stride];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n_rows = rows_of(rows, rows_dev);
  const int n_tiles = (n_rows + kScoreRows - 1) / kScoreRows;
  uint16_t* h1 = s_h1[wave];
  const int lr = lane & 31, kh = 8 * (lane >> 5);
  // block-uniform rounds: every wave reaches every barrier
  for (int tb = blockIdx.x * 4; tb < n_tiles; tb += gridDim.x * 4) {
    const int row0 = (tb + wave) * kScoreRows;
    const bool active = tb + wave < n_tiles;
    if (active) {
      // layer 1: A fragments of the wave's 32 rows (zero past the last row)
      bf16x8 xa[2];
      const int xr = row0 + lr;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        bf16x8 z = {};
        xa[s] = xr < n_rows ? *reinterpret_cast<const bf16x8*>(X + (size_t)xr * kMlpIn + 16 * s + kh)
                            : z;
      }
#pragma unroll
      for (int grp = 0; grp < kMlpHidden / kW2Group; ++grp) {
        floatx16 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          acc[t] = floatx16{};
          const int n = grp * kW2Group + t * 32 + lr;
#pragma unroll
          for (int s = 0; s < 2; ++s) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(W1 + (size_t)n * kMlpIn + 16 * s + kh);
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xa[s], b, acc[t], 0, 0, 0);
          }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int n = grp * kW2Group + t * 32 + lr;
          const float bias = b1[n];
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const int m = 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
            float x = acc[t][v] + bias;
            x = x > 0.0f ? x : 0.0f;
            h1[m * kH1Stride + n] = f2bf(x);
          }
        }
      }
    }
    // layer 2 and the output layer
    float part[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) part[v] = 0.0f;
#pragma unroll 1
    for (int grp = 0; grp < kMlpHidden / kW2Group; ++grp) {
      __syncthreads();  // H1 written; the previous group's readers are done with s_w2
      {
        // all loads of a thread in flight before its LDS stores
        constexpr int kStage = kW2Group * (kMlpHidden / 8) / 256;
        bf16x8 buf[kStage];
#pragma unroll
        for (int j = 0; j < kStage; ++j) {
          const int c = threadIdx.x + 256 * j;
          const int r = c / (kMlpHidden / 8), k = (c % (kMlpHidden / 8)) * 8;
          buf[j] = *reinterpret_cast<const bf16x8*>(W2 + (size_t)(grp * kW2Group + r) * kMlpHidden + k);
        }
#pragma unroll
        for (int j = 0; j < kStage; ++j) {
          const int c = threadIdx.x + 256 * j;
          const int r = c / (kMlpHidden / 8), k = (c % (kMlpHidden / 8)) * 8;
          *reinterpret_cast<bf16x8*>(s_w2 + r * kH1Stride + k) = buf[j];
        }
      }
      __syncthreads();
      if (active) {
        floatx16 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = floatx16{};
#pragma unroll 4
        for (int s = 0; s < kMlpHidden / 16; ++s) {
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(h1 + lr * kH1Stride + 16 * s + kh);
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const bf16x8 b =
                *reinterpret_cast<const bf16x8*>(s_w2 + (t * 32 + lr) * kH1Stride + 16 * s + kh);
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
          }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int n = grp * kW2Group + t * 32 + lr;
          const float bias = b2[n], wn = bf2f(w3[n]);
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            float x = acc[t][v] + bias;
            x = x > 0.0f ? x : 0.0f;
            part[v] += bf2f(f2bf(x)) * wn;
          }
        }
      }
    }
    if (active) {
      // sum each row over the 32 lanes of its half-wave
#pragma unroll
      for (int v = 0; v < 16; ++v)
        for (int o = 1; o < 32; o <<= 1) part[v] += __shfl_xor(part[v], o);
      if (lr == 0) {
        const float bb = *b3;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int m = row0 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
          if (m < n_rows) logits[m] = part[v] + bb;
        }
      }
    }
    __syncthreads();  // the next round rewrites the H1 tiles
  }
}

hipError_t mlp_score(const uint16_t* X, const uint16_t* W1, const float* b1, const uint16_t* W2,
                     const float* b2, const uint16_t* w3, const float* b3, int rows,
                     const uint32_t* rows_dev, float* logits, int max_blocks, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  const int tiles = (rows + kScoreRows - 1) / kScoreRows;
  const int blocks = max_blocks > 0 ? std::min((tiles + 3) / 4, max_blocks) : (tiles + 3) / 4;
  hipLaunchKernelGGL(k_mlp_score, dim3(blocks), dim3(256), 0, s, X, W1, b1, W2, b2, w3, b3, rows,
                     rows_dev, logits);
  return hipGetLastError();
}

}  // namespace gbr
