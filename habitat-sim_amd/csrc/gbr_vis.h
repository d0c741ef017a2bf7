// Learned visibility proposal: the chunk-visibility MLP's parameters, its
// training set of (features, visible) rows, and the GPU training and
// inference passes.  Host code; the kernels are in gbr_mlp.hip.
#pragma once

#include <stdint.h>

#include "gbr.h"
#include "gbr_host.h"
#include "gbr_mlp.h"

namespace gbr {

// parameter layout of the fp32 master copy: W1 [256][32], b1 [256],
// W2 [256][256], b2 [256], w3 [256], b3
constexpr size_t kVisW1 = 0;
constexpr size_t kVisB1 = kVisW1 + (size_t)kMlpHidden * kMlpIn;
constexpr size_t kVisW2 = kVisB1 + kMlpHidden;
constexpr size_t kVisB2 = kVisW2 + (size_t)kMlpHidden * kMlpHidden;
constexpr size_t kVisW3 = kVisB2 + kMlpHidden;
constexpr size_t kVisB3 = kVisW3 + kMlpHidden;
constexpr size_t kVisParams = kVisB3 + 1;

struct VisModel {
  bool ready = false;  // weights set (initialised, trained or loaded)
  DevBuf<float> w, g, m, v;  // master weights, gradients, Adam moments
  DevBuf<uint16_t> W1b, W2b, W2tb, w3b;  // bf16 operands (W2 also transposed)
  uint64_t adam_t = 0;

  // training set
  DevBuf<uint16_t> Xd;
  DevBuf<uint8_t> Yd;
  uint64_t rows = 0, cap = 0;
  DevBuf<uint32_t> positives;  // device count of visible rows

  // minibatch scratch
  int bcap = 0;
  DevBuf<uint16_t> Xb, Xbt, H1, H1t, H2, H2t, dH2, dH2t, dH1, dH1t;
  DevBuf<uint8_t> Yb;
  DevBuf<float> logits, dl, lossr, red;

  // per-frame inference scratch (rows = candidate work items)
  DevBuf<uint16_t> Xi;
  DevBuf<float> logits_i;
  DevBuf<uint32_t> prop;

  gbr_status init(uint64_t seed, hipStream_t s);
  gbr_status set_weights(const float* host, hipStream_t s);
  gbr_status get_weights(float* host, hipStream_t s);
  gbr_status reserve_dataset(uint64_t rows_cap);
  gbr_status reserve_infer(size_t rows);
  // forward over rows [0, min(rows, *rows_dev)) of X into logits
  gbr_status forward(const uint16_t* X, int rows, const uint32_t* rows_dev, uint16_t* H1o,
                     uint16_t* H1to, uint16_t* H2o, uint16_t* H2to, float* out, hipStream_t s);
  // rendering: logits of rows [0, min(rows, *rows_dev)) of X, fused
  gbr_status score(const uint16_t* X, int rows, const uint32_t* rows_dev, float* out,
                   int max_blocks, hipStream_t s);
  // loss and gradients (into g) of the batch already in Xb/Xbt/Yb
  gbr_status backward(int batch, float pos_weight, hipStream_t s);
  gbr_status adam_step(const gbr_vis_fit_params& fp, hipStream_t s);
  gbr_status refresh_bf16(hipStream_t s);
  gbr_status fit(const gbr_vis_fit_params& fp, gbr_vis_fit_report* rep, hipStream_t s);
  gbr_status loss_grad(const float* X, const uint8_t* labels, int rows, float pos_weight,
                       float* loss, float* grad, hipStream_t s);
  gbr_status predict(const float* X, int rows, float* out, hipStream_t s);
  void release();

 private:
  gbr_status reserve_batch(int batch);
};

}  // namespace gbr
