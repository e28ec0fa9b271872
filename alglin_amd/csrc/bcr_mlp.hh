// The learned LU-vs-QRP selector: a 12 -> 64 -> 64 -> 2 perceptron (ReLU,
// softmax), FP32, trained and evaluated on the GPU by the kernels of
// bcr_mlp.hip.  Not part of the reference; it decides per system whether the
// pair eliminations use partial-pivot LU (class 0) or column-pivoted QR
// (class 1).
//
// Parameter vector (kMlpParams floats): W1 (12 x 64, row-major: W1[i*64+j]),
// b1 (64), W2 (64 x 64), b2 (64), W3 (64 x 2), b3 (2).  Inputs are
// standardized with a fixed per-feature mean and scale before W1.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "bcr_engine.hh"
#include "bcr_launch.hh"

namespace bcr {

constexpr int kMlpIn = 12, kMlpHidden = 64, kMlpOut = 2;
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpHidden * kMlpOut + kMlpOut;

class Mlp {
 public:
  explicit Mlp(uint64_t seed);
  ~Mlp();
  Mlp(Mlp const&) = delete;
  Mlp& operator=(Mlp const&) = delete;

  void get_params(float* host) const;
  void set_params(float const* host);
  void set_normalization(float const* mu, float const* sig);
  void get_normalization(float* mu, float* sig) const;
  // file: 12 means, 12 scales, then the parameter vector (little-endian float32)
  void load(std::string const& path);
  void save(std::string const& path) const;

  // mean cross-entropy over n samples and its gradient (full batch);
  // X: n x 12 row-major device floats, y: n device labels (0/1)
  float loss_grad(float const* X, int32_t const* y, int n, float* grad_host);
  // full-batch Adam; returns the loss of every epoch (before its update)
  std::vector<float> train(float const* X, int32_t const* y, int n, int epochs, float lr);
  // probabilities of class 1 (device output, n floats)
  void predict(float const* X, int n, float* p1);
  hipStream_t stream() const { return st_; }

 private:
  void ensure(int n);
  void forward_backward(float const* X, int32_t const* y, int n);
  hipStream_t st_ = nullptr;
  DevBuf<float> P_, G_, M_, V_, norm_, A1_, A2_, dZ1_, dZ2_, dZ3_, Ls_, part_, loss_;
  int cap_ = 0, step_ = 0;
  std::vector<float> mu_, sig_;
};

}  // namespace bcr
