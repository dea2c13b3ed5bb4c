// rank_model.hip -- the learned rank predictor: a 16-64-64-64-1 FP32 MLP
// (ReLU hidden layers, linear output) with hand-written inference, training
// (mean squared error, backpropagation) and Adam kernels.
//
// It predicts log2(1 + rank) of a box's ACA from the box's geometry, its
// block shape, the kernel and the tolerance.  Assembly uses the prediction
// only to size and bucket the ACA scratch; the ACA still decides every rank
// and reruns a box at full capacity when the predicted capacity is reached
// before its stopping test passes.  No reference counterpart (the
// reference's ACA, ACA.cpp:27-133, always reserves min(rows, columns)).
//
// Layout: activations row-major [sample][feature]; parameters one float
// vector, W[out][in] row-major per layer, W1 b1 W2 b2 W3 b3 W4 b4.  Every
// reduction runs in a fixed order (no atomics), so training is
// deterministic for a given seed.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "engine.hpp"

namespace afmm {

namespace {

// parameter offsets
constexpr int kOffW[4] = {0, 1088, 5248, 9408};
constexpr int kOffB[4] = {1024, 5184, 9344, 9472};
constexpr int kDin[4] = {kRmIn, kRmHid, kRmHid, kRmHid};
constexpr int kDout[4] = {kRmHid, kRmHid, kRmHid, 1};
constexpr int kRmChunk = 256;  // samples per partial sum of the weight gradients

// Y[s][o] = b[o] + sum_i W[o][i] X[s][i] (i ascending), ReLU'd for hidden
// layers.  W and b staged in LDS; thread (s, o) over a grid-stride loop.
template <bool RELU>
__global__ __launch_bounds__(256) void k_rm_dense(const float* __restrict__ X, int din,
                                                  const float* __restrict__ W,
                                                  const float* __restrict__ b,
                                                  float* __restrict__ Y, int dout, int B) {
  __shared__ float sw[kRmHid * kRmHid];
  __shared__ float sb[kRmHid];
  for (int u = threadIdx.x; u < din * dout; u += blockDim.x) sw[u] = W[u];
  for (int u = threadIdx.x; u < dout; u += blockDim.x) sb[u] = b[u];
  __syncthreads();
  const int64_t tot = (int64_t)B * dout;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(t / dout), o = (int)(t - (int64_t)s * dout);
    const float* x = X + (int64_t)s * din;
    float acc = sb[o];
    for (int i = 0; i < din; i++) acc = fmaf(sw[o * din + i], x[i], acc);
    Y[t] = RELU ? fmaxf(acc, 0.0f) : acc;
  }
}

// d[s] = 2 (out[s] - y[s]) / B; part[block] = sum of (out - y)^2 over the
// block's samples (ascending), for the reported loss.
__global__ __launch_bounds__(256) void k_rm_loss(const float* __restrict__ out,
                                                 const float* __restrict__ y, int B,
                                                 float* __restrict__ d,
                                                 double* __restrict__ part) {
  __shared__ double red[256];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0;
  if (s < B) {
    const float e = out[s] - y[s];
    d[s] = 2.0f * e / (float)B;
    e2 = (double)e * (double)e;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dX[s][i] = sum_o dZ[s][o] W[o][i] (o ascending), times [Xprev[s][i] > 0]
// (the ReLU of the layer below).
__global__ __launch_bounds__(256) void k_rm_back_data(const float* __restrict__ dZ, int dout,
                                                      const float* __restrict__ W, int din,
                                                      const float* __restrict__ Xprev,
                                                      float* __restrict__ dX, int B) {
  __shared__ float sw[kRmHid * kRmHid];
  for (int u = threadIdx.x; u < din * dout; u += blockDim.x) sw[u] = W[u];
  __syncthreads();
  const int64_t tot = (int64_t)B * din;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(t / din), i = (int)(t - (int64_t)s * din);
    const float* dz = dZ + (int64_t)s * dout;
    float acc = 0.0f;
    for (int o = 0; o < dout; o++) acc = fmaf(dz[o], sw[o * din + i], acc);
    dX[t] = Xprev[t] > 0.0f ? acc : 0.0f;
  }
}

// Partial weight gradients of chunk c = blockIdx.y (samples
// [c kRmChunk, (c + 1) kRmChunk), ascending): part[c][o (din + 1) + i] =
// sum_s dZ[s][o] X[s][i], column din holding sum_s dZ[s][o] (the bias).
__global__ __launch_bounds__(256) void k_rm_back_weight(const float* __restrict__ dZ, int dout,
                                                        const float* __restrict__ X, int din,
                                                        int B, float* __restrict__ part) {
  const int w = din + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= dout * w) return;
  const int o = t / w, i = t - o * w;
  const int s0 = blockIdx.y * kRmChunk, s1 = min(B, s0 + kRmChunk);
  float acc = 0.0f;
  for (int s = s0; s < s1; s++) {
    const float xv = i < din ? X[(int64_t)s * din + i] : 1.0f;
    acc = fmaf(dZ[(int64_t)s * dout + o], xv, acc);
  }
  part[(int64_t)blockIdx.y * dout * w + t] = acc;
}

// g[W(o, i)] and g[b(o)] = the chunks' partial sums in chunk order.
__global__ __launch_bounds__(256) void k_rm_sum_parts(const float* __restrict__ part, int nchunk,
                                                      int dout, int din, float* __restrict__ gW,
                                                      float* __restrict__ gb) {
  const int w = din + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= dout * w) return;
  float acc = 0.0f;
  for (int c = 0; c < nchunk; c++) acc += part[(int64_t)c * dout * w + t];
  const int o = t / w, i = t - o * w;
  if (i < din) gW[o * din + i] = acc;
  else gb[o] = acc;
}

// Adam (Kingma & Ba 2015, Algorithm 1); bc1 = 1 - beta1^t, bc2 = 1 - beta2^t.
__global__ __launch_bounds__(256) void k_rm_adam(float* __restrict__ p, const float* __restrict__ g,
                                                 float* __restrict__ m, float* __restrict__ v,
                                                 int n, float lr, float b1, float b2, float eps,
                                                 float bc1, float bc2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt;
  const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt / bc1) / (sqrtf(vt / bc2) + eps);
}

// Xb[u] = X[idx[u]], yb[u] = y[idx[u]] (a shuffled minibatch).
__global__ __launch_bounds__(256) void k_rm_gather(const float* __restrict__ X,
                                                   const float* __restrict__ y,
                                                   const int* __restrict__ idx, int B,
                                                   float* __restrict__ Xb, float* __restrict__ yb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * kRmIn) return;
  const int u = t / kRmIn, f = t - u * kRmIn;
  const int64_t s = idx[u];
  Xb[t] = X[s * kRmIn + f];
  if (f == 0) yb[u] = y[s];
}

unsigned grid_for(int64_t work) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(work, 256), 4096));
}

}  // namespace

RankModel::~RankModel() {
  if (s) {
    (void)hipSetDevice(device);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
}

RankModel* rank_model_create(int device, uint64_t seed) {
  auto m = std::make_unique<RankModel>();
  m->device = device;
  AFMM_HIP(hipSetDevice(device));
  AFMM_HIP(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
  // He-uniform weights (limit sqrt(6 / fan_in)), zero biases
  std::vector<float> p(kRmParams, 0.0f);
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  for (int l = 0; l < 4; l++) {
    const double lim = std::sqrt(6.0 / kDin[l]);
    for (int k = 0; k < kDin[l] * kDout[l]; k++) p[kOffW[l] + k] = (float)(lim * u(rng));
  }
  m->params.alloc(sizeof(float) * kRmParams);
  m->grads.alloc(sizeof(float) * kRmParams);
  m->adam_m.alloc(sizeof(float) * kRmParams);
  m->adam_v.alloc(sizeof(float) * kRmParams);
  AFMM_HIP(hipMemcpy(m->params.p, p.data(), sizeof(float) * kRmParams, hipMemcpyHostToDevice));
  AFMM_HIP(hipMemset(m->adam_m.p, 0, sizeof(float) * kRmParams));
  AFMM_HIP(hipMemset(m->adam_v.p, 0, sizeof(float) * kRmParams));
  return m.release();
}

void rank_model_set_params(RankModel& m, const float* p) {
  AFMM_HIP(hipSetDevice(m.device));
  AFMM_HIP(hipMemcpy(m.params.p, p, sizeof(float) * kRmParams, hipMemcpyHostToDevice));
  AFMM_HIP(hipMemset(m.adam_m.p, 0, sizeof(float) * kRmParams));
  AFMM_HIP(hipMemset(m.adam_v.p, 0, sizeof(float) * kRmParams));
  m.step = 0;
}

void rank_model_get_params(RankModel& m, float* p) {
  AFMM_HIP(hipSetDevice(m.device));
  AFMM_HIP(hipMemcpy(p, m.params.p, sizeof(float) * kRmParams, hipMemcpyDeviceToHost));
}

namespace {

// Activations for B samples: A0 = X (B x 16), A1..A3 (B x 64), out (B).
void ensure_acts(RankModel& m, int B) {
  if (B <= m.cap) return;
  for (int l = 0; l < 4; l++) m.act[l].alloc(sizeof(float) * (size_t)B * kDin[l]);
  m.out.alloc(sizeof(float) * (size_t)B);
  m.yb.alloc(sizeof(float) * (size_t)B);
  m.dout.alloc(sizeof(float) * (size_t)B);
  m.dact[0].alloc(sizeof(float) * (size_t)B * kRmHid);
  m.dact[1].alloc(sizeof(float) * (size_t)B * kRmHid);
  m.loss_part.alloc(sizeof(double) * (size_t)ceil_div(B, 256));
  m.wpart.alloc(sizeof(float) * (size_t)ceil_div(B, kRmChunk) * kRmHid * (kRmHid + 1));
  m.cap = B;
}

// Forward pass of B samples already in act[0]; outputs in out.
void forward(RankModel& m, int B, hipStream_t s) {
  const float* P = m.params.as<float>();
  for (int l = 0; l < 4; l++) {
    const float* X = m.act[l].as<float>();
    float* Y = l < 3 ? m.act[l + 1].as<float>() : m.out.as<float>();
    const unsigned g = grid_for((int64_t)B * kDout[l]);
    if (l < 3) k_rm_dense<true><<<g, 256, 0, s>>>(X, kDin[l], P + kOffW[l], P + kOffB[l], Y, kDout[l], B);
    else k_rm_dense<false><<<g, 256, 0, s>>>(X, kDin[l], P + kOffW[l], P + kOffB[l], Y, kDout[l], B);
    AFMM_LAUNCH_CHECK();
  }
}

// Loss gradient into grads for the B samples in act[0] / yb; returns the
// launch count's loss partials (summed on the host by the caller).
void backward(RankModel& m, int B, hipStream_t s) {
  const float* P = m.params.as<float>();
  float* G = m.grads.as<float>();
  k_rm_loss<<<ceil_div(B, 256), 256, 0, s>>>(m.out.as<float>(), m.yb.as<float>(), B,
                                              m.dout.as<float>(), m.loss_part.as<double>());
  AFMM_LAUNCH_CHECK();
  const float* dZ = m.dout.as<float>();
  int cur = 0;  // dact buffer the next dX goes to
  const int nchunk = ceil_div(B, kRmChunk);
  for (int l = 3; l >= 0; l--) {
    const int din = kDin[l], dout = kDout[l];
    const float* X = m.act[l].as<float>();
    dim3 gw(ceil_div(dout * (din + 1), 256), nchunk);
    k_rm_back_weight<<<gw, 256, 0, s>>>(dZ, dout, X, din, B, m.wpart.as<float>());
    AFMM_LAUNCH_CHECK();
    k_rm_sum_parts<<<ceil_div(dout * (din + 1), 256), 256, 0, s>>>(
        m.wpart.as<float>(), nchunk, dout, din, G + kOffW[l], G + kOffB[l]);
    AFMM_LAUNCH_CHECK();
    if (l > 0) {
      float* dX = m.dact[cur].as<float>();
      k_rm_back_data<<<grid_for((int64_t)B * din), 256, 0, s>>>(dZ, dout, P + kOffW[l], din, X,
                                                                 dX, B);
      AFMM_LAUNCH_CHECK();
      dZ = dX;
      cur ^= 1;
    }
  }
}

double loss_sum(RankModel& m, int B, hipStream_t s) {
  std::vector<double> part(ceil_div(B, 256));
  AFMM_HIP(hipMemcpyAsync(part.data(), m.loss_part.p, sizeof(double) * part.size(),
                          hipMemcpyDeviceToHost, s));
  AFMM_HIP(hipStreamSynchronize(s));
  double t = 0.0;
  for (double v : part) t += v;
  return t;
}

}  // namespace

void rank_model_predict_device(RankModel& m, const float* d_x, int64_t n, float* d_out,
                               hipStream_t s) {
  constexpr int64_t kBatch = 1 << 16;
  for (int64_t i0 = 0; i0 < n; i0 += kBatch) {
    const int B = (int)std::min(kBatch, n - i0);
    ensure_acts(m, B);
    AFMM_HIP(hipMemcpyAsync(m.act[0].p, d_x + i0 * kRmIn, sizeof(float) * B * kRmIn,
                            hipMemcpyDeviceToDevice, s));
    forward(m, B, s);
    AFMM_HIP(hipMemcpyAsync(d_out + i0, m.out.p, sizeof(float) * B, hipMemcpyDeviceToDevice, s));
  }
}

void rank_model_predict(RankModel& m, const float* x, int64_t n, float* out) {
  AFMM_HIP(hipSetDevice(m.device));
  if (n <= 0) return;
  DevBuf dx, dy;
  dx.alloc(sizeof(float) * n * kRmIn);
  dy.alloc(sizeof(float) * n);
  AFMM_HIP(hipMemcpyAsync(dx.p, x, sizeof(float) * n * kRmIn, hipMemcpyHostToDevice, m.s));
  rank_model_predict_device(m, dx.as<float>(), n, dy.as<float>(), m.s);
  AFMM_HIP(hipMemcpyAsync(out, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost, m.s));
  AFMM_HIP(hipStreamSynchronize(m.s));
}

double rank_model_gradient(RankModel& m, const float* x, const float* y, int n, float* grad) {
  AFMM_HIP(hipSetDevice(m.device));
  if (n <= 0) throw Error(1, "gradient of an empty batch");
  ensure_acts(m, n);
  hipStream_t s = m.s;
  AFMM_HIP(hipMemcpyAsync(m.act[0].p, x, sizeof(float) * (size_t)n * kRmIn,
                          hipMemcpyHostToDevice, s));
  AFMM_HIP(hipMemcpyAsync(m.yb.p, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, s));
  forward(m, n, s);
  backward(m, n, s);
  AFMM_HIP(hipMemcpyAsync(grad, m.grads.p, sizeof(float) * kRmParams, hipMemcpyDeviceToHost, s));
  return loss_sum(m, n, s) / n;
}

double rank_model_train(RankModel& m, const float* x, const float* y, int64_t n, int epochs,
                        int batch, float lr, uint64_t seed) {
  AFMM_HIP(hipSetDevice(m.device));
  if (n <= 0 || epochs <= 0 || batch <= 0) throw Error(1, "empty training set");
  batch = (int)std::min<int64_t>(batch, n);
  hipStream_t s = m.s;
  DevBuf dx, dy, didx;
  dx.alloc(sizeof(float) * (size_t)n * kRmIn);
  dy.alloc(sizeof(float) * (size_t)n);
  didx.alloc(sizeof(int) * (size_t)n);
  AFMM_HIP(hipMemcpyAsync(dx.p, x, sizeof(float) * (size_t)n * kRmIn, hipMemcpyHostToDevice, s));
  AFMM_HIP(hipMemcpyAsync(dy.p, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, s));
  ensure_acts(m, batch);
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double last = 0.0;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(idx.begin(), idx.end(), rng);
    AFMM_HIP(hipMemcpyAsync(didx.p, idx.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
    AFMM_HIP(hipStreamSynchronize(s));  // idx is reshuffled next epoch
    double tot = 0.0;
    for (int64_t i0 = 0; i0 + batch <= n; i0 += batch) {
      k_rm_gather<<<ceil_div(batch * kRmIn, 256), 256, 0, s>>>(
          dx.as<float>(), dy.as<float>(), didx.as<int>() + i0, batch, m.act[0].as<float>(),
          m.yb.as<float>());
      AFMM_LAUNCH_CHECK();
      forward(m, batch, s);
      backward(m, batch, s);
      m.step++;
      const float bc1 = 1.0f - std::pow(b1, (float)m.step);
      const float bc2 = 1.0f - std::pow(b2, (float)m.step);
      k_rm_adam<<<ceil_div(kRmParams, 256), 256, 0, s>>>(
          m.params.as<float>(), m.grads.as<float>(), m.adam_m.as<float>(), m.adam_v.as<float>(),
          kRmParams, lr, b1, b2, eps, bc1, bc2);
      AFMM_LAUNCH_CHECK();
      if (ep == epochs - 1) tot += loss_sum(m, batch, s);
    }
    if (ep == epochs - 1) last = tot / (double)(n / batch * (int64_t)batch);
  }
  AFMM_HIP(hipStreamSynchronize(s));
  return last;
}

}  // namespace afmm
