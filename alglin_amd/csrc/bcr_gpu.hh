// Device state of alglin_amd::BorderedCR<T> (the Gpu member declared in
// include/alglin_amd/BorderedCR.hh).  The engines are created on first use,
// so the host surface (loaders, accessors, Mv, sparse_*, check_matrix, info)
// works without touching the GPU.
#pragma once

#include <memory>
#include <vector>

#include "alglin_amd/BorderedCR.hh"
#include "bcr_engine.hh"
#include "bcr_refine.hh"

namespace alglin_amd {

// the object's stream (created on first use, destroyed after the engines)
struct OwnedStream {
  hipStream_t st = nullptr;
  hipStream_t get() {
    if (!st) bcr::hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    return st;
  }
  ~OwnedStream() {
    if (st) (void)hipStreamDestroy(st);
  }
};

template <typename T>
struct BorderedCR<T>::Gpu {
  OwnedStream stream;   // one stream per object: both engines run on it
  bcr::Dims want;       // the object's dimensions
  int want_P = 1;       // the object's partition count
  std::unique_ptr<bcr::Engine<T, T>> eng;   // factors in T (the reference's arithmetic)
  std::unique_ptr<bcr::MixedEngine> mix;    // FP32 factors of the FP64 operator (T = double, MIXED)
  bcr::Dims eng_dims, mix_dims;
  int eng_P = 0, mix_P = 0;
  bcr::RefineReport report;
  std::vector<int> col_status;
  bcr::DevBuf<T> xbuf;          // host solves: the right-hand sides in HBM
  bcr::DevBuf<double> fbx;      // FP64 fallback: the columns it solves again
  bcr::DevBuf<double> fbb;      // and their right-hand sides
  bool timing = false;          // HIP-event timing of every launch group
  int pair_variant = 0;         // LU pair kernel choice (alglin_bcr_set_pair_kernel)

  // the engine in T, allocated for the current dimensions and partition
  bcr::Engine<T, T>& fp() {
    if (!eng) {
      eng.reset(new bcr::Engine<T, T>(stream.get()));
      eng->set_timing(timing);
      eng->set_pair_variant(pair_variant);
    }
    if (!(eng_dims == want)) {
      eng->allocate(want);
      eng_dims = want;
      eng_P = 0;
    }
    if (eng_P != want_P) {
      eng->set_partition(want_P);
      eng_P = want_P;
    }
    return *eng;
  }
  // the mixed engine (on the stream of the T engine: one stream per object)
  bcr::MixedEngine& mx() {
    if (!mix) {
      mix.reset(new bcr::MixedEngine(stream.get()));
      mix->engine().set_timing(timing);
      mix->engine().set_pair_variant(pair_variant);
    }
    if (!(mix_dims == want)) {
      mix->allocate(want);
      mix_dims = want;
      mix_P = 0;
    }
    if (mix_P != want_P) {
      mix->engine().set_partition(want_P);
      mix_P = want_P;
    }
    return *mix;
  }
  bool has_mix() const { return mix != nullptr; }
  void set_timing(bool on) {
    timing = on;
    if (eng) eng->set_timing(on);
    if (mix) mix->engine().set_timing(on);
  }
  void set_pair_variant(int v) {
    pair_variant = v;
    if (eng) eng->set_pair_variant(v);
    if (mix) mix->engine().set_pair_variant(v);
  }
};

}  // namespace alglin_amd
