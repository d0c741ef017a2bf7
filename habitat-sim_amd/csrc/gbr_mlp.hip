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
// Fused scorer for rendering: logits of min(rows, *rows_dev) feature rows in
// one persistent pass with the weights in registers.  A workgroup of 8 waves
// holds the model: wave w owns hidden columns [32w, 32w + 32) of both layers,
// with its slices of W1 (2 MFMA B fragments) and W2 (16 fragments) in VGPRs
// for the whole kernel.  Per tile of 32 rows: layer 1 (K = 32, 2
// v_mfma_f32_32x32x16_bf16) -> bias, ReLU, bf16 -> the tile's H1 in LDS;
// layer 2 (K = 256, 16 MFMAs, A fragments from LDS) -> bias, ReLU, bf16,
// times w3 -> per-column products in LDS; then each row's logit is summed in
// k_last_layer's order (lane l: columns l, l+64, l+128, l+192, then a
// butterfly over 64 lanes).  The K loops run k = 0, 16, ... like k_gemm_nt,
// so H1, H2 and the logits equal the training forward's bit for bit.
// --------------------------------------------------------------------------

constexpr int kScoreRows = 32;                  // rows per tile
constexpr int kScoreWaves = kMlpHidden / 32;    // one wave per 32 hidden columns
constexpr int kScoreThreads = kScoreWaves * 64;
constexpr int kH1Stride = kMlpHidden + 8;       // padded LDS row of H1 (bf16)
constexpr int kPartStride = kMlpHidden + 4;     // padded LDS row of the products (fp32)
static_assert(kMlpIn == 32 && kMlpHidden % 32 == 0, "scorer tiling");
static_assert(kScoreRows % kScoreWaves == 0, "rows of the logit reduction per wave");

__global__ __launch_bounds__(kScoreThreads) void k_mlp_score(const uint16_t* __restrict__ X,
                                                             const uint16_t* __restrict__ W1,
                                                             const float* __restrict__ b1,
                                                             const uint16_t* __restrict__ W2,
                                                             const float* __restrict__ b2,
                                                             const uint16_t* __restrict__ w3,
                                                             const float* __restrict__ b3,
                                                             int rows,
                                                             const uint32_t* __restrict__ rows_dev,
                                                             float* __restrict__ logits) {
  __shared__ uint16_t s_h1[kScoreRows * kH1Stride];
  __shared__ float s_part[kScoreRows * kPartStride];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 31, kh = 8 * (lane >> 5);
  const int n = wave * 32 + lr;  // this lane's hidden column
  const int n_rows = rows_of(rows, rows_dev);
  const int n_tiles = (n_rows + kScoreRows - 1) / kScoreRows;
  if ((int)blockIdx.x >= n_tiles) return;  // (whole workgroup)
  // the wave's weight slices, loaded once
  bf16x8 w1f[2], w2f[kMlpHidden / 16];
#pragma unroll
  for (int s = 0; s < 2; ++s)
    w1f[s] = *reinterpret_cast<const bf16x8*>(W1 + (size_t)n * kMlpIn + 16 * s + kh);
#pragma unroll
  for (int s = 0; s < kMlpHidden / 16; ++s)
    w2f[s] = *reinterpret_cast<const bf16x8*>(W2 + (size_t)n * kMlpHidden + 16 * s + kh);
  const float bias1 = b1[n], bias2 = b2[n], wn = bf2f(w3[n]), bias3 = *b3;
  // the A fragments of a tile's feature rows, fetched one tile ahead (rows
  // past the last read the last row; they are zeroed when used)
  bf16x8 xnext[2];
  auto load_x = [&](int tile, bf16x8 (&xa)[2]) {
    const int rr = min(tile * kScoreRows + lr, n_rows - 1);
#pragma unroll
    for (int s = 0; s < 2; ++s)
      xa[s] = *reinterpret_cast<const bf16x8*>(X + (size_t)rr * kMlpIn + 16 * s + kh);
  };
  load_x(blockIdx.x, xnext);
  for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const int row0 = tile * kScoreRows;
    bf16x8 xa[2] = {xnext[0], xnext[1]};
    if (row0 + lr >= n_rows) xa[0] = xa[1] = bf16x8{};
    if (tile + (int)gridDim.x < n_tiles) load_x(tile + gridDim.x, xnext);
    // layer 1: the tile's 32 rows (zero past the last row) x the wave's columns
    floatx16 acc = {};
#pragma unroll
    for (int s = 0; s < 2; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xa[s], w1f[s], acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int m = 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
      float x = acc[v] + bias1;
      x = x > 0.0f ? x : 0.0f;
      s_h1[m * kH1Stride + n] = f2bf(x);
    }
    __syncthreads();  // H1 complete (and the previous tile's products are consumed)
    // layer 2: A fragments of every row from LDS, the wave's W2 slice from VGPRs
    acc = floatx16{};
#pragma unroll
    for (int s = 0; s < kMlpHidden / 16; ++s) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(s_h1 + lr * kH1Stride + 16 * s + kh);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, w2f[s], acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int m = 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
      float x = acc[v] + bias2;
      x = x > 0.0f ? x : 0.0f;
      s_part[m * kPartStride + n] = bf2f(f2bf(x)) * wn;
    }
    __syncthreads();  // products complete; H1 may be rewritten
    // logits: wave w sums rows w, w + 8, ... in k_last_layer's order
#pragma unroll
    for (int q = 0; q < kScoreRows / kScoreWaves; ++q) {
      const int m = wave + q * kScoreWaves;
      float sum = 0.0f;
#pragma unroll
      for (int k = 0; k < kMlpHidden; k += 64) sum += s_part[m * kPartStride + k + lane];
      for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
      if (lane == 0 && row0 + m < n_rows) logits[row0 + m] = sum + bias3;
    }
  }
}

hipError_t mlp_score(const uint16_t* X, const uint16_t* W1, const float* b1, const uint16_t* W2,
                     const float* b2, const uint16_t* w3, const float* b3, int rows,
                     const uint32_t* rows_dev, float* logits, int max_blocks, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  const int tiles = (rows + kScoreRows - 1) / kScoreRows;
  const int blocks = max_blocks > 0 ? std::min(tiles, max_blocks) : tiles;
  hipLaunchKernelGGL(k_mlp_score, dim3(blocks), dim3(kScoreThreads), 0, s, X, W1, b1, W2, b2, w3,
                     b3, rows, rows_dev, logits);
  return hipGetLastError();
}

}  // namespace gbr
