// Learned visibility proposal: parameter handling, training loop and the
// forward/backward launch sequences of the chunk-visibility MLP.
#include "gbr_vis.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace gbr {

namespace {

// float -> bf16 bits, round to nearest even (as the device's conversion)
uint16_t host_f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

GemmArgs gemm(const uint16_t* A, int lda, const uint16_t* Bt, int ldb, int M, int N, int K) {
  GemmArgs g;
  g.A = A;
  g.lda = lda;
  g.Bt = Bt;
  g.ldb = ldb;
  g.M = M;
  g.N = N;
  g.K = K;
  return g;
}

}  // namespace

void VisModel::release() {
  for (DevBuf<float>* b : {&w, &g, &m, &v, &logits, &dl, &lossr, &red, &logits_i}) b->release();
  for (DevBuf<uint16_t>* b : {&W1b, &W2b, &W2tb, &w3b, &Xd, &Xb, &Xbt, &H1, &H1t, &H2, &H2t, &dH2,
                              &dH2t, &dH1, &dH1t, &Xi})
    b->release();
  Yd.release();
  Yb.release();
  positives.release();
  prop.release();
  ready = false;
  rows = cap = 0;
  bcap = 0;
}

gbr_status VisModel::refresh_bf16(hipStream_t s) {
  TRY(W1b.reserve((size_t)kMlpHidden * kMlpIn));
  TRY(W2b.reserve((size_t)kMlpHidden * kMlpHidden));
  TRY(W2tb.reserve((size_t)kMlpHidden * kMlpHidden));
  TRY(w3b.reserve(kMlpHidden));
  HIP_TRY(mlp_to_bf16(w.ptr + kVisW1, (size_t)kMlpHidden * kMlpIn, W1b.ptr, nullptr, kMlpHidden,
                      kMlpIn, s));
  HIP_TRY(mlp_to_bf16(w.ptr + kVisW2, (size_t)kMlpHidden * kMlpHidden, W2b.ptr, W2tb.ptr,
                      kMlpHidden, kMlpHidden, s));
  HIP_TRY(mlp_to_bf16(w.ptr + kVisW3, kMlpHidden, w3b.ptr, nullptr, 1, kMlpHidden, s));
  return GBR_OK;
}

gbr_status VisModel::set_weights(const float* host, hipStream_t s) {
  TRY(w.reserve(kVisParams));
  TRY(g.reserve(kVisParams));
  TRY(m.reserve(kVisParams));
  TRY(v.reserve(kVisParams));
  HIP_TRY(hipMemcpyAsync(w.ptr, host, kVisParams * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemsetAsync(m.ptr, 0, kVisParams * sizeof(float), s));
  HIP_TRY(hipMemsetAsync(v.ptr, 0, kVisParams * sizeof(float), s));
  adam_t = 0;
  TRY(refresh_bf16(s));
  HIP_TRY(hipStreamSynchronize(s));
  ready = true;
  return GBR_OK;
}

gbr_status VisModel::get_weights(float* host, hipStream_t s) {
  if (!ready) return fail(GBR_ERR_STATE, "visibility model: no weights (gbr_vis_init first)");
  HIP_TRY(hipMemcpyAsync(host, w.ptr, kVisParams * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return GBR_OK;
}

// He-uniform weights (fan-in), zero biases, the output bias at the prior
// log-odds of a chunk being visible (about one in four)
gbr_status VisModel::init(uint64_t seed, hipStream_t s) {
  std::vector<float> h(kVisParams, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](size_t off, size_t n, int fan_in) {
    const float lim = std::sqrt(6.0f / (float)fan_in);
    std::uniform_real_distribution<float> u(-lim, lim);
    for (size_t i = 0; i < n; ++i) h[off + i] = u(rng);
  };
  fill(kVisW1, (size_t)kMlpHidden * kMlpIn, kMlpIn);
  fill(kVisW2, (size_t)kMlpHidden * kMlpHidden, kMlpHidden);
  fill(kVisW3, kMlpHidden, kMlpHidden);
  for (size_t i = 0; i < kMlpHidden; ++i) h[kVisW3 + i] *= 0.1f;
  h[kVisB3] = -1.0f;
  return set_weights(h.data(), s);
}

gbr_status VisModel::reserve_dataset(uint64_t rows_cap) {
  if (rows_cap <= cap && Xd.ptr) return GBR_OK;
  TRY(Xd.reserve((size_t)rows_cap * kMlpIn, true));
  TRY(Yd.reserve(rows_cap, true));
  cap = rows_cap;
  if (!positives.ptr) {
    TRY(positives.reserve(1));
    HIP_TRY(hipMemset(positives.ptr, 0, sizeof(uint32_t)));
  }
  return GBR_OK;
}

gbr_status VisModel::reserve_infer(size_t n) {
  n = std::max<size_t>((n + 31) & ~size_t(31), 32);
  TRY(Xi.reserve(n * kMlpIn));
  TRY(logits_i.reserve(n));
  TRY(prop.reserve(n));
  return GBR_OK;
}

gbr_status VisModel::reserve_batch(int batch) {
  if (batch <= bcap) return GBR_OK;
  const size_t b = (size_t)batch;
  TRY(Xb.reserve(b * kMlpIn));
  TRY(Xbt.reserve(b * kMlpIn));
  for (DevBuf<uint16_t>* x : {&H1, &H1t, &H2, &H2t, &dH2, &dH2t, &dH1, &dH1t})
    TRY(x->reserve(b * kMlpHidden));
  TRY(Yb.reserve(b));
  TRY(logits.reserve(b));
  TRY(dl.reserve(b));
  TRY(lossr.reserve(b));
  TRY(red.reserve(4));
  bcap = batch;
  return GBR_OK;
}

gbr_status VisModel::forward(const uint16_t* X, int nrows, const uint32_t* rows_dev,
                             uint16_t* H1o, uint16_t* H1to, uint16_t* H2o, uint16_t* H2to,
                             float* out, hipStream_t s) {
  GemmArgs g1 = gemm(X, kMlpIn, W1b.ptr, kMlpIn, nrows, kMlpHidden, kMlpIn);
  g1.M_dev = rows_dev;
  g1.bias = w.ptr + kVisB1;
  g1.relu = 1;
  g1.D = H1o;
  g1.ldd = kMlpHidden;
  g1.Dt = H1to;
  g1.lddt = nrows;
  HIP_TRY(mlp_gemm(g1, s));
  GemmArgs g2 = gemm(H1o, kMlpHidden, W2b.ptr, kMlpHidden, nrows, kMlpHidden, kMlpHidden);
  g2.M_dev = rows_dev;
  g2.bias = w.ptr + kVisB2;
  g2.relu = 1;
  g2.D = H2o;
  g2.ldd = kMlpHidden;
  g2.Dt = H2to;
  g2.lddt = nrows;
  HIP_TRY(mlp_gemm(g2, s));
  HIP_TRY(mlp_last_layer(H2o, w3b.ptr, w.ptr + kVisB3, out, nrows, rows_dev, s));
  return GBR_OK;
}

gbr_status VisModel::score(const uint16_t* X, int nrows, const uint32_t* rows_dev, float* out,
                           int max_blocks, hipStream_t s) {
  HIP_TRY(mlp_score(X, W1b.ptr, w.ptr + kVisB1, W2b.ptr, w.ptr + kVisB2, w3b.ptr, w.ptr + kVisB3,
                    nrows, rows_dev, out, max_blocks, s));
  return GBR_OK;
}

// gradients of the mean weighted loss over the batch in Xb/Xbt/Yb; the loss
// lands in red[0]
gbr_status VisModel::backward(int b, float pos_weight, hipStream_t s) {
  TRY(forward(Xb.ptr, b, nullptr, H1.ptr, H1t.ptr, H2.ptr, H2t.ptr, logits.ptr, s));
  HIP_TRY(mlp_head(logits.ptr, Yb.ptr, pos_weight, b, dl.ptr, lossr.ptr, s));
  HIP_TRY(mlp_sum(lossr.ptr, b, red.ptr, s));
  HIP_TRY(mlp_sum(dl.ptr, b, g.ptr + kVisB3, s));
  HIP_TRY(mlp_row_sums(H2t.ptr, dl.ptr, kMlpHidden, b, g.ptr + kVisW3, s));
  HIP_TRY(mlp_dh2(H2.ptr, H2t.ptr, dl.ptr, w3b.ptr, b, dH2.ptr, dH2t.ptr, s));
  // dW2 = dH2^T H1
  GemmArgs gw2 = gemm(dH2t.ptr, b, H1t.ptr, b, kMlpHidden, kMlpHidden, b);
  gw2.C = g.ptr + kVisW2;
  gw2.ldc = kMlpHidden;
  HIP_TRY(mlp_gemm(gw2, s));
  HIP_TRY(mlp_row_sums(dH2t.ptr, nullptr, kMlpHidden, b, g.ptr + kVisB2, s));
  // dH1 = (dH2 W2) o [H1 > 0]
  GemmArgs gh1 = gemm(dH2.ptr, kMlpHidden, W2tb.ptr, kMlpHidden, b, kMlpHidden, kMlpHidden);
  gh1.mask = H1.ptr;
  gh1.ldmask = kMlpHidden;
  gh1.D = dH1.ptr;
  gh1.ldd = kMlpHidden;
  gh1.Dt = dH1t.ptr;
  gh1.lddt = b;
  HIP_TRY(mlp_gemm(gh1, s));
  // dW1 = dH1^T X
  GemmArgs gw1 = gemm(dH1t.ptr, b, Xbt.ptr, b, kMlpHidden, kMlpIn, b);
  gw1.C = g.ptr + kVisW1;
  gw1.ldc = kMlpIn;
  HIP_TRY(mlp_gemm(gw1, s));
  HIP_TRY(mlp_row_sums(dH1t.ptr, nullptr, kMlpHidden, b, g.ptr + kVisB1, s));
  return GBR_OK;
}

gbr_status VisModel::adam_step(const gbr_vis_fit_params& fp, hipStream_t s) {
  ++adam_t;
  const float bc1 = 1.0f - std::pow(fp.beta1, (float)adam_t);
  const float bc2 = 1.0f - std::pow(fp.beta2, (float)adam_t);
  HIP_TRY(mlp_adam(w.ptr, g.ptr, m.ptr, v.ptr, kVisParams, fp.learning_rate, fp.beta1, fp.beta2,
                   fp.eps, bc1, bc2, s));
  return refresh_bf16(s);
}

gbr_status VisModel::fit(const gbr_vis_fit_params& fp, gbr_vis_fit_report* rep, hipStream_t s) {
  if (!ready) return fail(GBR_ERR_STATE, "vis_fit: no weights (gbr_vis_init first)");
  if (rows == 0) return fail(GBR_ERR_STATE, "vis_fit: no training rows (gbr_vis_collect first)");
  if (fp.steps < 1 || fp.batch < 32 || fp.batch > 65536 || fp.batch % 32)
    return fail(GBR_ERR_ARGUMENT, "vis_fit: steps >= 1 and batch a multiple of 32 in [32, 65536]");
  TRY(reserve_batch(fp.batch));
  std::vector<float> losses((size_t)fp.steps, 0.0f);
  DevBuf<float> dloss;
  TRY(dloss.reserve((size_t)fp.steps));
  for (int k = 0; k < fp.steps; ++k) {
    HIP_TRY(mlp_gather(Xd.ptr, Yd.ptr, rows, fp.seed, (uint64_t)k, fp.batch, Xb.ptr, Xbt.ptr,
                       Yb.ptr, s));
    TRY(backward(fp.batch, fp.pos_weight, s));
    HIP_TRY(hipMemcpyAsync(dloss.ptr + k, red.ptr, sizeof(float), hipMemcpyDeviceToDevice, s));
    TRY(adam_step(fp, s));
  }
  HIP_TRY(hipMemcpyAsync(losses.data(), dloss.ptr, losses.size() * sizeof(float),
                         hipMemcpyDeviceToHost, s));
  uint32_t pos = 0;
  HIP_TRY(hipMemcpyAsync(&pos, positives.ptr, sizeof pos, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  dloss.release();
  if (rep) {
    rep->loss_first = losses[0] / (float)fp.batch;
    const int tail = std::max(1, fp.steps / 10);
    double t = 0.0;
    for (int k = fp.steps - tail; k < fp.steps; ++k) t += losses[(size_t)k];
    rep->loss_last = (float)(t / tail / fp.batch);
    rep->rows = rows;
    rep->positives = pos;
  }
  return GBR_OK;
}

gbr_status VisModel::loss_grad(const float* X, const uint8_t* labels, int n, float pos_weight,
                               float* loss, float* grad, hipStream_t s) {
  if (!ready) return fail(GBR_ERR_STATE, "vis_loss_grad: no weights (gbr_vis_init first)");
  if (n < 32 || n > 65536 || n % 32)
    return fail(GBR_ERR_ARGUMENT, "vis_loss_grad: rows must be a multiple of 32 in [32, 65536]");
  TRY(reserve_batch(n));
  std::vector<uint16_t> xb((size_t)n * kMlpIn), xbt((size_t)n * kMlpIn);
  for (int r = 0; r < n; ++r)
    for (int k = 0; k < kMlpIn; ++k) {
      const uint16_t b = host_f2bf(X[(size_t)r * kMlpIn + k]);
      xb[(size_t)r * kMlpIn + k] = b;
      xbt[(size_t)k * n + r] = b;
    }
  HIP_TRY(hipMemcpyAsync(Xb.ptr, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(Xbt.ptr, xbt.data(), xbt.size() * 2, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(Yb.ptr, labels, (size_t)n, hipMemcpyHostToDevice, s));
  TRY(backward(n, pos_weight, s));
  float l = 0.0f;
  HIP_TRY(hipMemcpyAsync(&l, red.ptr, sizeof l, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(grad, g.ptr, kVisParams * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *loss = l / (float)n;
  return GBR_OK;
}

gbr_status VisModel::predict(const float* X, int n, float* out, hipStream_t s) {
  if (!ready) return fail(GBR_ERR_STATE, "vis_predict: no weights (gbr_vis_init first)");
  if (n < 1 || n > 65536) return fail(GBR_ERR_ARGUMENT, "vis_predict: rows must be in [1, 65536]");
  const int np = (n + 31) & ~31;
  TRY(reserve_batch(np));
  std::vector<uint16_t> xb((size_t)np * kMlpIn, 0);
  for (size_t i = 0; i < (size_t)n * kMlpIn; ++i) xb[i] = host_f2bf(X[i]);
  HIP_TRY(hipMemcpyAsync(Xb.ptr, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, s));
  // the renderer's fused scorer, on few workgroups so that each loops over
  // several row tiles as it does on a frame's candidate rows
  TRY(score(Xb.ptr, n, nullptr, logits.ptr, 16, s));
  HIP_TRY(hipMemcpyAsync(out, logits.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return GBR_OK;
}

}  // namespace gbr
