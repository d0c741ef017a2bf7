// Chunk-visibility MLP: shapes, GEMM arguments and launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gbr {

constexpr int kMlpIn = 32;
constexpr int kMlpHidden = 256;  // GBR_VIS_HIDDEN

// C[M][N] = alpha * sum_k A[m][k] * Bt[n][k]  (bf16 in, fp32 accumulate)
// rows M = min(M, *M_dev) when M_dev is set (a row count produced on the
// device); epilogue on x = alpha*acc (stored to C if C):
//   x += bias[n]; x = mask[m][n] > 0 ? x : 0; x = relu(x); D[m][n] = bf16(x);
//   Dt[n][m] = bf16(x)
struct GemmArgs {
  const uint16_t* A = nullptr;
  const uint16_t* Bt = nullptr;
  int M = 0, N = 0, K = 0;
  int lda = 0, ldb = 0;
  const uint32_t* M_dev = nullptr;
  float alpha = 1.0f;
  float* C = nullptr;
  int ldc = 0;
  const float* bias = nullptr;
  const uint16_t* mask = nullptr;
  int ldmask = 0;
  int relu = 0;
  uint16_t* D = nullptr;
  int ldd = 0;
  uint16_t* Dt = nullptr;
  int lddt = 0;
};

hipError_t mlp_gemm(const GemmArgs& g, hipStream_t s);
hipError_t mlp_last_layer(const uint16_t* H2, const uint16_t* w3, const float* b3, float* logits,
                          int rows, const uint32_t* rows_dev, hipStream_t s);
hipError_t mlp_head(const float* logits, const uint8_t* labels, float pos_weight, int rows,
                    float* dl, float* loss, hipStream_t s);
hipError_t mlp_dh2(const uint16_t* H2, const uint16_t* H2t, const float* dl, const uint16_t* w3,
                   int rows, uint16_t* dH2, uint16_t* dH2t, hipStream_t s);
hipError_t mlp_row_sums(const uint16_t* Mt, const float* w, int n, int rows, float* out,
                        hipStream_t s);
hipError_t mlp_sum(const float* v, int n, float* out, hipStream_t s);
hipError_t mlp_gather(const uint16_t* X, const uint8_t* Y, uint64_t total, uint64_t seed,
                      uint64_t step, int batch, uint16_t* Xb, uint16_t* Xbt, uint8_t* Yb,
                      hipStream_t s);
hipError_t mlp_to_bf16(const float* in, size_t n, uint16_t* out, uint16_t* out_t, int rows,
                       int cols, hipStream_t s);
// logits of rows [0, min(rows, *rows_dev)) of X for rendering, in one fused
// pass with no intermediates in memory; at most max_blocks workgroups (0: one
// per 128 rows), each looping over 32-row tiles.  H1 and H2 are the training
// forward's bit for bit; the logit sums in another order.
hipError_t mlp_score(const uint16_t* X, const uint16_t* W1, const float* b1, const uint16_t* W2,
                     const float* b2, const uint16_t* w3, const float* b3, int rows,
                     const uint32_t* rows_dev, float* logits, int max_blocks, hipStream_t s);
hipError_t mlp_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1,
                    float b2, float eps, float bc1, float bc2, hipStream_t s);

}  // namespace gbr
