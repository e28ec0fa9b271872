// GPU engine of one bordered BABD system: device buffers and the launch
// sequence of factorize / solve / residual.  No torch types; plain HIP.
//
//   TA  element type of the pristine operator (double, or float for the
//       BorderedCR<float> surface)
//   TW  element type of the cyclic-reduction factors (== TA for the plain
//       path, float for the mixed-precision path with TA = double)
//
// The pristine blocks are never modified: the first reduction level reads
// them and writes the factor arrays, later levels work in the factor arrays.
// That keeps A intact for the FP64 residual of iterative refinement without
// a copy (the reference overwrites D/E/B/C/F in place, cc:1099-1153).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "bcr_common.hh"

namespace bcr {

struct Dims {
  int N = 0, n = 0, qr = 0, qx = 0, nr = 0, nx = 0;
  int Nr() const { return 2 * n + nr + qr; }
  int Nc() const { return 2 * n + nx + qx; }
  int64_t nrows() const { return int64_t(n) * (N + 1) + qr + nr; }
  int64_t ncols() const { return int64_t(n) * (N + 1) + qx + nx; }
  bool operator==(Dims const& o) const {
    return N == o.N && n == o.n && qr == o.qr && qx == o.qx && nr == o.nr && nx == o.nx;
  }
};

class HipError : public std::runtime_error {
 public:
  explicit HipError(std::string const& s) : std::runtime_error(s) {}
};

void hip_check(hipError_t e, char const* what);

// A device buffer owned by the engine, or a slice of a region the engine
// does not own (borrow): then resize never allocates, it only checks.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t count = 0;
  bool borrowed = false;
  void resize(size_t c);
  void borrow(T* ptr, size_t c);
  void release();
  ~DevBuf() { release(); }
};

struct FactorStats {
  int levels = 0;
  int fast_path = 0;   // 1 when the register-resident LU kernel ran
};

template <typename TA, typename TW>
class Engine {
 public:
  Engine();
  // on the caller's stream (not owned, not created)
  explicit Engine(hipStream_t external);
  ~Engine();
  Engine(Engine const&) = delete;
  Engine& operator=(Engine const&) = delete;

  void allocate(Dims const& d);
  // as allocate() for pure cyclic reduction with LU eliminations and an LU
  // last block, solving up to nrhs columns, every working array a slice of
  // [base, base + bytes) (256-byte aligned): no device allocation, now or in
  // the factorize / solve that follow.  Throws when the slices do not fit.
  void allocate_in(Dims const& d, int nrhs, char* base, size_t bytes);
  // bytes allocate_in needs for these dimensions
  static size_t bytes_in(Dims const& d, int nrhs, int max_chunks);
  int max_chunks() const { return max_chunks_; }
  Dims const& dims() const { return d_; }
  // elimination order of BorderedCR with num_parallel_block = P
  // (BABD_BorderedCR.cc:139-158): the blocks are split into P ranges, each
  // range is reduced by cyclic reduction on its own (cc:1055-1171), then the
  // chain of the 2P-1 range-end rows (cc:1175-1266); P = 1 (the default) is
  // pure cyclic reduction over all blocks.  P is shrunk like the reference's
  // until 2P <= N.
  void set_partition(int P);
  int partition() const { return used_P_; }
  std::vector<int32_t> const& partition_blocks() const { return iblock_; }
  hipStream_t stream() const { return st_; }
  // run on another engine's stream (not owned) so that the two engines of one
  // solver object order their work on a single stream
  void share_stream(hipStream_t s);

  void set_method(int m) { method_ = m; }
  // last-block factorization: kLastLU, kLastLUPQ, kLastQR, kLastQRP, or
  // 4..7 (SVD, LSS, LSY, PINV: the SVD path)
  void set_last_method(int m) { last_method_ = m; }
  int last_method() const { return last_method_; }
  // second last-block solver (SVD path) used when the first fails
  void set_last2(bool on) { last2_ = on; }
  // LU solves with n <= 32 use the wave-synchronous sweep kernels unless off
  void set_sweep_fast(bool on) { sweep_fast_ = on; }
  // LU pair kernel choice: 0 default (see level_factor), 3 the
  // register-tiled kernel for every n, 4 column-split waves, 5 the
  // matrix-core pair (FP32 factors)
  void set_pair_variant(int v) { pair_variant_ = v; }
  int method() const { return method_; }

  // pristine operator: upload from host block arrays (reference layout)
  void upload(TA const* D, TA const* E, TA const* B, TA const* C, TA const* H, TA const* Cq, TA const* F);
  // or use caller-owned device arrays as the pristine operator (no copy)
  void bind_device(TA const* D, TA const* E, TA const* B, TA const* C, TA const* H, TA const* Cq, TA const* F);

  // enqueue the factorization on the engine stream (no host sync)
  void factorize_async();
  // status after factorize_async: 0 ok, 1 zero pivot in a pair, 2 singular last block
  int32_t status_sync();

  // in-place solve on device X (ldx x nrhs), X type TX = TA or double
  template <typename TX> void solve_async(TX* x, int nrhs, int64_t ldx);

  // r = b - A x (FP64, pristine operator); all device pointers
  void residual_async(double const* x, double const* b, double* r);
  // R_q = B_q - A X_q for nrhs columns (X_q = x + q ldx, B_q, R_q likewise):
  // one pass over the blocks for up to four columns
  void residual_multi_async(double const* x, int nrhs, int64_t ldx, double const* b, int64_t ldb, double* r,
                            int64_t ldr);
  // y = A x
  void mv_async(double const* x, double* y);
  // out = |A| |x| + |b| (FP64, pristine operator): the row scales of the
  // componentwise backward error
  void abs_bound_async(double const* x, double const* b, double* out);
  // *out = max_i |r_i| / den_i over i < len (device scalar)
  void max_ratio_async(double const* r, double const* den, int64_t len, double* out);
  // true once an operator is uploaded or bound
  bool has_operator() const { return pD_ != nullptr; }
  // ||v||_2 of a device vector into *out (device)
  void nrm2_async(double const* v, int64_t len, double* out);

  // internals (device pointers) for tests
  TW const* factor_T() const { return Tm_.p; }
  int32_t const* factor_perm() const { return perm_.p; }
  TW const* work_D() const { return wD_.p; }
  TW const* work_E() const { return wE_.p; }
  TW const* work_B() const { return wB_.p; }
  TW const* work_C() const { return wC_.p; }
  double const* last_block() const { return Hm_.p; }
  TW const* pending() const { return Cpend_.p; }
  size_t pending_count() const { return Cpend_.count; }
  FactorStats const& stats() const { return stats_; }

  // per-phase timing with HIP events on the engine stream (see alglin_amd.h)
  enum Phase { kPair = 0, kForward = 1, kBackward = 2, kLast = 3, kReduce = 4, kResidual = 5 };
  void set_timing(bool on) { timing_ = on; }
  void timing(double out[8], bool reset);

 private:
  template <typename F> void timed(int phase, F&& f);
  struct Rec { int phase; hipEvent_t a, b; };
  bool timing_ = false;
  std::vector<Rec> recs_;
  std::vector<hipEvent_t> event_pool_;
  double acc_[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  hipEvent_t get_event();

  // one elimination launch: stride L (pure cyclic reduction) or a pair
  // table slice [off, off + 3 npairs) of pairtab_ (off >= 0)
  struct Level {
    int L = 1, npairs = 0, slot = 0;
    bool first = false;
    int64_t off = -1;
  };
  void build_schedule();
  void level_factor(Level const& l);
  std::vector<Level> sched_;
  std::vector<int32_t> copy_rows_, copy_cols_;   // untouched by the first launch: copied from A
  std::vector<int32_t> iblock_;
  DevBuf<int32_t> pairtab_;
  int want_P_ = 1, used_P_ = 1;
  Dims d_;
  int method_ = kLU;
  int last_method_ = kLastLU;
  bool last2_ = false;
  bool sweep_fast_ = true;
  int pair_variant_ = 0;
  hipStream_t st_ = nullptr;
  bool own_stream_ = true;
  size_t lds_limit_ = 65536;
  int max_chunks_ = 1024;
  FactorStats stats_;
  // pristine (owned copy or bound)
  DevBuf<TA> aD_, aE_, aB_, aC_, aH_, aCq_, aF_;
  TA const *pD_ = nullptr, *pE_ = nullptr, *pB_ = nullptr, *pC_ = nullptr, *pH_ = nullptr, *pCq_ = nullptr,
           *pF_ = nullptr;
  // factors / working
  DevBuf<TW> wD_, wE_, wB_, wC_, Tm_, Cpend_, Fpart_, gscratch_, tmpw_;
  DevBuf<int32_t> perm_, ipiv_, jpiv_, status_;
  DevBuf<double> Hm2_, svdW_, svdVj_, svdU_, svdV_, svdS_;
  DevBuf<int32_t> svd_rank_;
  DevBuf<double> Feff_, Hm_, tau_, tmpd_, Cpart_, Csum_, Xpartd_;
  DevBuf<TW> Xpartw_;
  bool pend_clean_ = false;
};

}  // namespace bcr
