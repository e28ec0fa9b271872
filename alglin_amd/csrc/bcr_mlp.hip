// GPU forward, backward and Adam kernels of the selector perceptron
// (layout in bcr_mlp.hh).
//
//   mlp_sample_kernel  one thread per sample: standardize, forward through the
//                      three layers with the weights in LDS, softmax
//                      cross-entropy; in training mode also the per-sample
//                      back-propagated deltas dZ3 (n x 2), dZ2, dZ1 (n x 64)
//                      and the activations A1, A2 the weight gradients need
//   mlp_grad_partial   parameter gradients as sums over samples, split into
//                      fixed sample chunks (grid: parameters x chunks)
//   mlp_grad_final     chunk partials summed in chunk order (deterministic);
//                      the extra slot kMlpParams carries the mean loss
//   mlp_adam_kernel    Adam (beta1 0.9, beta2 0.999, eps 1e-8), bias-corrected
#include "bcr_mlp.hh"

#include <cmath>
#include <cstring>
#include <fstream>
#include <random>

namespace bcr {
namespace {

constexpr int I = kMlpIn, H = kMlpHidden, O = kMlpOut;
constexpr int oW1 = 0, ob1 = oW1 + I * H, oW2 = ob1 + H, ob2 = oW2 + H * H, oW3 = ob2 + H, ob3 = oW3 + H * O;
constexpr int kChunks = 64;

__global__ void __launch_bounds__(64) mlp_sample_kernel(const float* __restrict__ P, const float* __restrict__ norm,
                                                         const float* __restrict__ X, const int32_t* __restrict__ y,
                                                         int n, float* __restrict__ A1, float* __restrict__ A2,
                                                         float* __restrict__ dZ1, float* __restrict__ dZ2,
                                                         float* __restrict__ dZ3, float* __restrict__ loss,
                                                         float* __restrict__ p1) {
  __shared__ float w[kMlpParams];
  for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) w[i] = P[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[I];
#pragma unroll
  for (int i = 0; i < I; ++i) x[i] = (X[size_t(s) * I + i] - norm[i]) / norm[I + i];
  float a1[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float z = w[ob1 + j];
#pragma unroll
    for (int i = 0; i < I; ++i) z += x[i] * w[oW1 + i * H + j];
    a1[j] = z > 0.f ? z : 0.f;
  }
  float z2[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float z = w[ob2 + j];
    for (int k = 0; k < H; ++k) z += a1[k] * w[oW2 + k * H + j];
    z2[j] = z;
  }
  float z3[O];
#pragma unroll
  for (int c = 0; c < O; ++c) {
    float z = w[ob3 + c];
#pragma unroll
    for (int k = 0; k < H; ++k) z += (z2[k] > 0.f ? z2[k] : 0.f) * w[oW3 + k * O + c];
    z3[c] = z;
  }
  const float zm = fmaxf(z3[0], z3[1]);
  const float e0 = expf(z3[0] - zm), e1 = expf(z3[1] - zm);
  const float lse = zm + logf(e0 + e1);
  const float q1 = e1 / (e0 + e1);
  if (p1) p1[s] = q1;
  if (!y) return;
  const int lab = y[s];
  loss[s] = lse - z3[lab];
  const float inv_n = 1.f / float(n);
  const float d3[O] = {((1.f - q1) - (lab == 0 ? 1.f : 0.f)) * inv_n, (q1 - (lab == 1 ? 1.f : 0.f)) * inv_n};
  dZ3[size_t(s) * O + 0] = d3[0];
  dZ3[size_t(s) * O + 1] = d3[1];
  float d2[H];
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const float da = d3[0] * w[oW3 + k * O + 0] + d3[1] * w[oW3 + k * O + 1];
    d2[k] = z2[k] > 0.f ? da : 0.f;
    dZ2[size_t(s) * H + k] = d2[k];
    A2[size_t(s) * H + k] = z2[k] > 0.f ? z2[k] : 0.f;
    A1[size_t(s) * H + k] = a1[k];
  }
  for (int k = 0; k < H; ++k) {
    float da = 0.f;
    for (int j = 0; j < H; ++j) da += d2[j] * w[oW2 + k * H + j];
    dZ1[size_t(s) * H + k] = a1[k] > 0.f ? da : 0.f;
  }
}

// partial[c][q] = sum over samples of chunk c of the term of parameter q
__global__ void __launch_bounds__(256) mlp_grad_partial(const float* __restrict__ norm, const float* __restrict__ X,
                                                         int n, const float* __restrict__ A1,
                                                         const float* __restrict__ A2, const float* __restrict__ dZ1,
                                                         const float* __restrict__ dZ2,
                                                         const float* __restrict__ dZ3,
                                                         const float* __restrict__ loss, float* __restrict__ part) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (q > kMlpParams) return;
  const int per = (n + kChunks - 1) / kChunks;
  const int s0 = c * per, s1 = min(n, s0 + per);
  float acc = 0.f;
  if (q < ob1) {
    const int i = q / H, j = q % H;
    for (int s = s0; s < s1; ++s) acc += (X[size_t(s) * I + i] - norm[i]) / norm[I + i] * dZ1[size_t(s) * H + j];
  } else if (q < oW2) {
    const int j = q - ob1;
    for (int s = s0; s < s1; ++s) acc += dZ1[size_t(s) * H + j];
  } else if (q < ob2) {
    const int k = (q - oW2) / H, j = (q - oW2) % H;
    for (int s = s0; s < s1; ++s) acc += A1[size_t(s) * H + k] * dZ2[size_t(s) * H + j];
  } else if (q < oW3) {
    const int j = q - ob2;
    for (int s = s0; s < s1; ++s) acc += dZ2[size_t(s) * H + j];
  } else if (q < ob3) {
    const int k = (q - oW3) / O, cc = (q - oW3) % O;
    for (int s = s0; s < s1; ++s) acc += A2[size_t(s) * H + k] * dZ3[size_t(s) * O + cc];
  } else if (q < kMlpParams) {
    const int cc = q - ob3;
    for (int s = s0; s < s1; ++s) acc += dZ3[size_t(s) * O + cc];
  } else {
    for (int s = s0; s < s1; ++s) acc += loss[s];
    acc /= float(n);
  }
  part[size_t(c) * (kMlpParams + 1) + q] = acc;
}

__global__ void __launch_bounds__(256) mlp_grad_final(const float* __restrict__ part, float* __restrict__ G,
                                                       float* __restrict__ loss_out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > kMlpParams) return;
  float acc = 0.f;
  for (int c = 0; c < kChunks; ++c) acc += part[size_t(c) * (kMlpParams + 1) + q];
  if (q < kMlpParams) G[q] = acc;
  else *loss_out = acc;
}

__global__ void __launch_bounds__(256) mlp_adam_kernel(float* __restrict__ P, const float* __restrict__ G,
                                                        float* __restrict__ M, float* __restrict__ V, float lr,
                                                        float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kMlpParams) return;
  const float g = G[q];
  const float m = 0.9f * M[q] + 0.1f * g;
  const float v = 0.999f * V[q] + 0.001f * g * g;
  M[q] = m;
  V[q] = v;
  P[q] -= lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
}

}  // namespace

Mlp::Mlp(uint64_t seed) {
  hip_check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreate");
  std::vector<float> p(kMlpParams, 0.f);
  std::mt19937_64 g(seed);
  auto glorot = [&](int off, int fan_in, int fan_out) {
    const float a = std::sqrt(6.f / float(fan_in + fan_out));
    std::uniform_real_distribution<float> u(-a, a);
    for (int i = 0; i < fan_in * fan_out; ++i) p[off + i] = u(g);
  };
  glorot(oW1, I, H);
  glorot(oW2, H, H);
  glorot(oW3, H, O);
  P_.resize(kMlpParams);
  G_.resize(kMlpParams);
  M_.resize(kMlpParams);
  V_.resize(kMlpParams);
  norm_.resize(2 * I);
  loss_.resize(1);
  set_params(p.data());
  mu_.assign(I, 0.f);
  sig_.assign(I, 1.f);
  set_normalization(mu_.data(), sig_.data());
}

Mlp::~Mlp() {
  if (st_) (void)hipStreamDestroy(st_);
}

void Mlp::get_params(float* host) const {
  hip_check(hipMemcpyAsync(host, P_.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st_), "params D2H");
  hip_check(hipStreamSynchronize(st_), "params sync");
}

void Mlp::set_params(float const* host) {
  hip_check(hipMemcpyAsync(P_.p, host, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st_), "params H2D");
  hip_check(launch_convert<float, float>(nullptr, M_.p, kMlpParams, st_), "adam m");
  hip_check(launch_convert<float, float>(nullptr, V_.p, kMlpParams, st_), "adam v");
  hip_check(hipStreamSynchronize(st_), "params sync");
  step_ = 0;
}

void Mlp::set_normalization(float const* mu, float const* sig) {
  mu_.assign(mu, mu + I);
  sig_.assign(sig, sig + I);
  for (float& s : sig_)
    if (!(s > 0.f)) s = 1.f;
  std::vector<float> h(mu_);
  h.insert(h.end(), sig_.begin(), sig_.end());
  hip_check(hipMemcpyAsync(norm_.p, h.data(), sizeof(float) * 2 * I, hipMemcpyHostToDevice, st_), "norm H2D");
  hip_check(hipStreamSynchronize(st_), "norm sync");
}

void Mlp::get_normalization(float* mu, float* sig) const {
  std::memcpy(mu, mu_.data(), sizeof(float) * I);
  std::memcpy(sig, sig_.data(), sizeof(float) * I);
}

void Mlp::load(std::string const& path) {
  std::ifstream f(path, std::ios::binary);
  std::vector<float> v(2 * I + kMlpParams);
  if (!f.read(reinterpret_cast<char*>(v.data()), std::streamsize(v.size() * sizeof(float))))
    throw HipError("selector: cannot read " + path);
  set_normalization(v.data(), v.data() + I);
  set_params(v.data() + 2 * I);
}

void Mlp::save(std::string const& path) const {
  std::vector<float> v(2 * I + kMlpParams);
  std::memcpy(v.data(), mu_.data(), sizeof(float) * I);
  std::memcpy(v.data() + I, sig_.data(), sizeof(float) * I);
  get_params(v.data() + 2 * I);
  std::ofstream f(path, std::ios::binary);
  if (!f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(float))))
    throw HipError("selector: cannot write " + path);
}

void Mlp::ensure(int n) {
  if (n <= cap_) return;
  A1_.resize(size_t(n) * H);
  A2_.resize(size_t(n) * H);
  dZ1_.resize(size_t(n) * H);
  dZ2_.resize(size_t(n) * H);
  dZ3_.resize(size_t(n) * O);
  Ls_.resize(size_t(n));
  part_.resize(size_t(kChunks) * (kMlpParams + 1));   // gradient partials
  cap_ = n;
}

void Mlp::forward_backward(float const* X, int32_t const* y, int n) {
  ensure(n);
  hipLaunchKernelGGL(mlp_sample_kernel, dim3((n + 63) / 64), dim3(64), 0, st_, P_.p, norm_.p, X, y, n, A1_.p, A2_.p,
                     dZ1_.p, dZ2_.p, dZ3_.p, Ls_.p, static_cast<float*>(nullptr));
  hip_check(hipGetLastError(), "mlp forward/backward");
  hipLaunchKernelGGL(mlp_grad_partial, dim3((kMlpParams + 1 + 255) / 256, kChunks), dim3(256), 0, st_, norm_.p, X, n,
                     A1_.p, A2_.p, dZ1_.p, dZ2_.p, dZ3_.p, Ls_.p, part_.p);
  hip_check(hipGetLastError(), "mlp gradient");
  hipLaunchKernelGGL(mlp_grad_final, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, st_, part_.p, G_.p, loss_.p);
  hip_check(hipGetLastError(), "mlp gradient sum");
}

float Mlp::loss_grad(float const* X, int32_t const* y, int n, float* grad_host) {
  if (n <= 0) throw HipError("mlp: no samples");
  forward_backward(X, y, n);
  float l = 0;
  hip_check(hipMemcpyAsync(&l, loss_.p, sizeof(float), hipMemcpyDeviceToHost, st_), "loss D2H");
  if (grad_host)
    hip_check(hipMemcpyAsync(grad_host, G_.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st_), "grad D2H");
  hip_check(hipStreamSynchronize(st_), "mlp sync");
  return l;
}

std::vector<float> Mlp::train(float const* X, int32_t const* y, int n, int epochs, float lr) {
  std::vector<float> hist;
  for (int e = 0; e < epochs; ++e) {
    forward_backward(X, y, n);
    float l = 0;
    hip_check(hipMemcpyAsync(&l, loss_.p, sizeof(float), hipMemcpyDeviceToHost, st_), "loss D2H");
    hip_check(hipStreamSynchronize(st_), "mlp sync");
    hist.push_back(l);
    ++step_;
    const float c1 = 1.f - std::pow(0.9f, float(step_)), c2 = 1.f - std::pow(0.999f, float(step_));
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, st_, P_.p, G_.p, M_.p, V_.p, lr,
                       c1, c2);
    hip_check(hipGetLastError(), "adam");
  }
  hip_check(hipStreamSynchronize(st_), "train sync");
  return hist;
}

void Mlp::predict(float const* X, int n, float* p1) {
  if (n <= 0) return;
  hipLaunchKernelGGL(mlp_sample_kernel, dim3((n + 63) / 64), dim3(64), 0, st_, P_.p, norm_.p, X,
                     static_cast<const int32_t*>(nullptr), n, static_cast<float*>(nullptr),
                     static_cast<float*>(nullptr), static_cast<float*>(nullptr), static_cast<float*>(nullptr),
                     static_cast<float*>(nullptr), static_cast<float*>(nullptr), p1);
  hip_check(hipGetLastError(), "mlp predict");
  hip_check(hipStreamSynchronize(st_), "predict sync");
}

}  // namespace bcr
