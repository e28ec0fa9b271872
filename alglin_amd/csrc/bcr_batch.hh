// Batched bordered cyclic reduction: many independent BABD systems of uneven
// block size n and block count N factored and solved together.
//
// Every system keeps the single-system algorithm (the reference's
// BorderedCR::factorize / solve at num_parallel_block = 1,
// BABD_BorderedCR.cc:803-839, 1489-1542).  What is batched is the launch:
// at reduction level L every system with N > L has pairs to eliminate, and one
// launch per kernel class covers the pairs of all of them (workgroup ->
// (system, local block) through a per-level table built once by plan()).
// Kernel classes follow the block size (register kernels for n <= 32, the
// LDS-panel kernel above) and the elimination method of the system (LU, QR,
// QRP), so systems of different shapes and methods share the same launches.
// The last blocks, the border reductions and the residual SpMV run one
// workgroup (or a few) per system in single launches.
//
// The whole factorize + solve sequence has a fixed launch list for a given
// plan, so it is captured once into a HIP graph and replayed as one launch.
//
// Packed storage: array kind k of system s starts at element offset(k, s)
// (kinds: D and E, B, C, H, Cq, F, x); blocks inside a slice follow the
// reference layout (hxx:972-998).
#pragma once

#include <map>
#include <vector>

#include "bcr_batch_kernels.hh"
#include "bcr_engine.hh"

namespace bcr {

// factorization kernel classes of a batch: 0..4 LU by block size (n <= 4, 8,
// 16, 32, above), 5 QR, 6 QRP
constexpr int kFacClasses = 7;

struct SysDims {
  int32_t N, n, qr, qx, nr, nx;
  int32_t method;   // kLU, kQR, kQRP (elimination of the pairs)
  int64_t nrows() const { return int64_t(n) * (N + 1) + qr + nr; }
  int64_t ncols() const { return int64_t(n) * (N + 1) + qx + nx; }
  int Nr() const { return 2 * n + qr + nr; }
  int Nc() const { return 2 * n + qx + nx; }
};

enum BatchKind { kbDE = 0, kbB = 1, kbC = 2, kbH = 3, kbCq = 4, kbF = 5, kbX = 6, kbKinds = 7 };

template <typename TA, typename TW>
class BatchEngine {
 public:
  BatchEngine();
  ~BatchEngine();
  BatchEngine(BatchEngine const&) = delete;
  BatchEngine& operator=(BatchEngine const&) = delete;

  // validates the shapes, computes the packed offsets, allocates the factor
  // storage; throws HipError on an unsupported shape
  void plan(std::vector<SysDims> const& dims);
  int nsys() const { return int(dims_.size()); }
  // restrict factorize/solve/residual/norms to the systems with mask[s] != 0
  // (the others' factors and vectors are left untouched)
  void set_active(std::vector<char> const& mask);
  std::vector<int> const& active() const { return active_; }
  // pair elimination method of system s (kLU, kQR, kQRP)
  void set_method(int s, int method);
  int method(int s) const { return dims_[size_t(s)].method; }
  // kernel of an LU pair class (3: 16 < n <= 32, 4: 32 < n <= 64): -1 the
  // default, kKernMfma the matrix-core pair (FP32 factors), kKernRank1 the
  // rank-1 kernels (two-wave split / column-split with `waves` waves per
  // pair, 0 = default), kKernTiled the register-tiled kernel (class 4)
  static constexpr int kKernMfma = 0, kKernRank1 = 1, kKernTiled = 2;
  void set_class_kernel(int cls, int kernel, int waves);
  size_t lds_limit() const { return lds_limit_; }
  // false when the pair panel of d would need LDS beyond lds_limit (the
  // shapes no register kernel of its class covers)
  static bool panel_fits(SysDims const& d, size_t lds_limit);
  SysDims const& dims(int s) const { return dims_[s]; }
  int64_t total(int kind) const { return tot_[kind]; }
  int64_t offset(int kind, int s) const { return off_[kind][s]; }

  // the pristine operator (packed device arrays, kinds as above)
  void bind_device(TA const* D, TA const* E, TA const* B, TA const* C, TA const* H, TA const* Cq, TA const* F);
  // fills the bound arrays with the synthetic Gaussian systems (needs TA =
  // double); sing_k (per system, may be empty): pair column made rank
  // deficient up to eps (k >= 1), or none (k < 1)
  void generate(uint64_t seed, double shift, std::vector<int32_t> const& sing_k = {}, double eps = 0.0);

  void factorize_async();
  // per-system status after factorize: bit 0 zero pivot in a pair, bit 1 singular last block
  void status(int32_t* host_out);
  // in-place solve of the packed vector X (offsets kbX, 1 right-hand side)
  void solve_async(double* X);
  // R = B - A X, or (absval) |A| |X| + |B|, per active system (FP64, the
  // pristine operator)
  void residual_async(double const* X, double const* B, double* R, bool absval = false);
  // out[s] = max_i |R_i| / (|A| |X| + |B|)_i per active system (the
  // componentwise backward error; den: x-layout scratch)
  void backward_errors_async(double const* X, double const* B, double const* R, double* den, double* out);
  // out[2s] = ||B_s||, out[2s+1] = ||R_s|| (device array of 2*nsys)
  void norms_async(double const* B, double const* R, double* out);

  // selector features of the active systems after factorize + solve +
  // norms (out: 12 floats per system, at out[12*s])
  void features_async(double const* norms, float* out);
  // dst_s = src_s (kind x slices) for the active systems
  void copy_active_async(double const* src, double* dst);
  // dst_s += src_s over the active systems' solution entries (ncols each)
  void add_active_async(double const* src, double* dst);

  // builds every launch table these buffers need (no launch); a graph
  // capture of factorize/solve/residual must not allocate
  void prepare(double* X, double const* B, double* R);

  hipStream_t stream() const { return st_; }
  size_t factor_bytes() const { return factor_bytes_; }
  // per-system flags (device, nsys): the residual, norms, solve and add
  // launches skip a system whose flag is nonzero; factorize_async clears them
  int32_t* skip_flags() { return skip_.p; }
  // The refinement's tail: narrow the active set to `mask` (launch tables
  // for those systems only), keeping the current tables -- and so every
  // graph captured over them -- aside; widen() brings them back unchanged.
  // Nothing may be in flight on the engine's streams at either call.
  void narrow(std::vector<char> const& mask);
  void widen();
  // kernels of the 16 < n <= 32 and 32 < n <= 64 LU classes: 0 default (the
  // matrix-core pair where it applies: n <= 48, nr <= 8, nx <= 16; else the
  // rank-1 kernels), 4 the rank-1 kernels (two-wave split, column-split),
  // 5 the matrix-core pair
  void set_pair_variant(int v) { pair_variant_ = v; }
  // factorization and sweeps on one stream (1) or with the n > 32 class on
  // a second stream beside the others (2, the default), the second stream
  // forked from reduction level `fork_level` on (the levels before it run
  // one launch after the other)
  void set_streams(int k, int fork_level = 0) {
    streams_ = k > 1 ? 2 : 1;
    fork_level_ = fork_level < 0 ? 0 : fork_level;
  }
  int streams() const { return streams_; }

  // one eager factorization with HIP events around every pair-elimination
  // launch: per launch its level, kernel class, workgroups, duration and
  // algorithmic bytes / flops (reads of the level-0 sources in TA, the rest
  // in TW)
  struct LaunchTime {
    int level, cls, nblocks;
    double ms, bytes, flops;
  };
  std::vector<LaunchTime> time_factorize();

 private:
  int pair_variant_ = 0;
  struct Launch {
    const void* args;
    const int32_t* entry;
    const int32_t* first;
    int nblocks, cls, method, nmax, nxmax, nrmax;
    double bytes = 0, flops = 0;   // algorithmic HBM bytes and flops of a factor launch
  };
  struct SolveTables {
    std::vector<std::vector<Launch>> fwd;   // per level
    std::vector<std::vector<Launch>> bwd;
    const ReduceArgs<double, double>* xb = nullptr;
    const LastSolveArgs<double>* last = nullptr;
    DevBuf<char> blob;
  };
  struct MvTables {
    Launch rows{};
    const ReduceArgs<double, double>* csum = nullptr;
    const MvArgs<TA>* tail = nullptr;
    const double* const* csum_ptr = nullptr;
    DevBuf<char> blob;
  };
  struct NormTables {
    const NormArgs* args = nullptr;
    DevBuf<char> blob;
  };
  struct AddTables {
    const CopyArgs* args = nullptr;
    DevBuf<char> blob;
  };

  void factorize_impl(std::vector<hipEvent_t>* ev);
  void build_factor_tables();
  const CopyArgs* add_table(double const* src, double* dst);
  SolveTables& solve_tables(double* X);
  MvTables& mv_tables(double const* X, double const* B, double* R);

  void invalidate();
  std::vector<SysDims> dims_;
  std::vector<int> active_;
  std::vector<int64_t> off_[kbKinds];
  int64_t tot_[kbKinds] = {0, 0, 0, 0, 0, 0, 0};
  // factor-side offsets
  std::vector<int64_t> offT_, offPerm_, offCpend_, offFpart_, offHm_, offPiv_, offXpart_, offCpart_, offCsum_, offPst_;
  int maxN_ = 0, levels_ = 0, Ncmax_ = 0, Nrmax_ = 0;
  bool square_ = true;   // every last block square (Nr == Nc)
  size_t lds_limit_ = 65536, factor_bytes_ = 0;
  int cls_kernel_[kFacClasses] = {-1, -1, -1, -1, -1, -1, -1};
  int cls_waves_[kFacClasses] = {0, 0, 0, 0, 0, 0, 0};
  hipStream_t st_ = nullptr;
  // A second stream for the n > 32 class: the classes touch disjoint
  // systems, so their level sequences run concurrently (forked from and
  // joined to st_ by events, inside a captured graph too).
  hipStream_t st2_ = nullptr;
  hipEvent_t fork_ = nullptr, join_ = nullptr;
  void fork();
  void join();
  int streams_ = 2, fork_level_ = 0;
  hipStream_t class_stream(int cls, bool concurrent) const { return concurrent && cls == 4 ? st2_ : st_; }

  TA const *pD_ = nullptr, *pE_ = nullptr, *pB_ = nullptr, *pC_ = nullptr, *pH_ = nullptr, *pCq_ = nullptr,
           *pF_ = nullptr;
  DevBuf<TW> wD_, wE_, wB_, wC_, Tm_, Cpend_, Fpart_;
  DevBuf<int32_t> perm_, ipiv_, status_, skip_;
  DevBuf<double> Feff_, Hm_, Xpart_, Cpart_, Csum_;

  // factor tables
  std::vector<std::vector<Launch>> fac_;   // per level
  const PrepArgs<TA, TW>* prep_ = nullptr;
  const ReduceArgs<TW, double>* fred_ = nullptr;
  const AssembleArgs<TW, TA>* asm_ = nullptr;
  const LastSolveArgs<double>* lu_ = nullptr;
  int32_t* const* stptr_ = nullptr;
  DevBuf<char> fblob_, scratch_;
  bool tables_ok_ = false;

  std::map<const void*, SolveTables> solve_cache_;
  std::map<std::vector<const void*>, MvTables> mv_cache_;
  std::map<std::vector<const void*>, NormTables> norm_cache_;
  std::map<std::vector<const void*>, AddTables> add_cache_;
  // the feature launches' records and tables (active set and norms pointer
  // of the call that built them), and the per-panel statistics
  DevBuf<char> feat_blob_;
  DevBuf<float> pstat_;   // 3 floats per panel j = 1 .. N-1 of every system (offPst_)
  // the last factorization left every active panel's statistics in pstat_
  // (FP32 LU classes on the kernels that write them)
  bool stats_ok_ = false;
  const FeatArgs<TW>* feat_args_ = nullptr;
  BatchTable<FeatArgs<TW>> feat_tab_{};
  int feat_wgs_ = 0, feat_nsys_ = 0;
  const double* feat_norms_ = nullptr;
  bool feat_ok_ = false;

  // the tables set aside by narrow()
  struct Stash {
    bool has = false;
    std::vector<int> active;
    std::vector<std::vector<Launch>> fac;
    const PrepArgs<TA, TW>* prep = nullptr;
    const ReduceArgs<TW, double>* fred = nullptr;
    const AssembleArgs<TW, TA>* asm_ = nullptr;
    const LastSolveArgs<double>* lu = nullptr;
    int32_t* const* stptr = nullptr;
    DevBuf<char> fblob;
    bool tables_ok = false;
    std::map<const void*, SolveTables> solve;
    std::map<std::vector<const void*>, MvTables> mv;
    std::map<std::vector<const void*>, NormTables> norm;
    std::map<std::vector<const void*>, AddTables> add;
  } stash_;
  void swap_stash();
};

// The batched solver a caller sees: FP64 cyclic reduction, or FP32 factors
// with FP64 iterative refinement per system (all systems swept together,
// until every system reaches tol or stalls).
struct BatchReport {
  int sweeps = 0;              // refinement sweeps after the first solve
  double max_rel_residual = 0; // over systems
  int failed = 0;              // systems that did not reach tol
  double ms_total = 0;         // wall time of the solve call (ms)
};

class BatchSolver {
 public:
  explicit BatchSolver(bool mixed);
  ~BatchSolver();
  void plan(std::vector<SysDims> const& dims);
  int64_t total(int kind) const;
  int64_t offset(int kind, int s) const;
  void bind_device(double const* D, double const* E, double const* B, double const* C, double const* H,
                   double const* Cq, double const* F);
  void generate(uint64_t seed, double shift, std::vector<int32_t> const& sing_k = {}, double eps = 0.0);
  // B, X: packed device vectors (kind x); X = A^{-1} B per system
  BatchReport solve(double const* B, double* X, double tol, int max_iter, bool use_graph);
  // mixed solver with per-system method choice: every system starts with FP32
  // LU; after the first solve the selector (if given) scores each system.
  // With `fallback`, the systems it scores above `threshold` go straight to
  // the FP64 path (LU, then QRP for what LU misses) and the others are refined
  // in FP64 and join that path if they still miss tol; without, the flagged
  // systems are refactored with FP32 QRP and all are refined in FP64
  BatchReport solve_selected(double const* B, double* X, double tol, int max_iter, class Mlp* selector,
                             float threshold, bool fallback);
  // features of the last solve_selected (12 per system, host) and the
  // systems sent to QRP / to the FP64 fallback
  std::vector<float> const& features() const { return feat_; }
  std::vector<char> const& chose_qrp() const { return chose_qrp_; }
  std::vector<char> const& fell_back() const { return fell_back_; }
  // per-system methods for solve(): kLU / kQR / kQRP
  void set_methods(std::vector<int32_t> const& m);
  // kernel of the 32 < n <= 64 LU class (BatchEngine::set_pair_variant)
  void set_pair_variant(int v);
  // 1: every launch on the batch's stream; 2 (default): the n > 32 LU class
  // on a second stream from reduction level fork_level on (drops the
  // captured graphs)
  void set_streams(int k, int fork_level);
  // tuned kernel of an LU pair class for the FP32-factor engine or the FP64
  // one (BatchEngine::set_class_kernel); kept for an engine created later
  void set_class_kernel(bool fp32_factors, int cls, int kernel, int waves);
  std::vector<int32_t> const& status() const { return status_; }
  std::vector<double> const& rel_residual() const { return rel_; }
  std::vector<int32_t> const& sweeps() const { return sweeps_; }
  hipStream_t stream() const;
  bool mixed() const { return mixed_; }
  // one eager factorization of the solver's engine (FP32 factors when
  // mixed) with every pair-elimination launch timed by HIP events
  std::vector<BatchEngine<double, float>::LaunchTime> time_factorize();

 private:
  template <typename E> BatchReport run(E& eng, double const* B, double* X, double tol, int max_iter, bool use_graph);
  template <typename E> void ensure_graphs(E& eng, double const* B, double* X);
  template <typename E> int mark_floor(E& eng, double const* B, double* X, double tol);
  std::vector<char> floor_;          // per system: stopped at the FP64 residual floor
  DevBuf<double> den_, berr_;        // |A||x| + |b| rows, per-system backward errors
  bool graphs_ = true;   // solve_selected replays captured launch lists
  int streams_ = 2, fork_level_ = 0;
  int tune32_[kFacClasses][2] = {{-1, 0}, {-1, 0}, {-1, 0}, {-1, 0}, {-1, 0}, {-1, 0}, {-1, 0}};
  int tune64_[kFacClasses][2] = {{-1, 0}, {-1, 0}, {-1, 0}, {-1, 0}, {-1, 0}, {-1, 0}, {-1, 0}};
  void apply_tuning(BatchEngine<double, double>& e);
  const void* g_eng_ = nullptr;   // the engine the captured lists launch on
  bool mixed_;
  BatchEngine<double, double>* e64_ = nullptr;
  BatchEngine<double, float>* e32_ = nullptr;
  DevBuf<double> R_, norms_;
  // device side of the refinement test (per system)
  DevBuf<double> drel_, dbest_;
  DevBuf<int32_t> dstat_, dsweeps_, dstag_, dopen_;
  std::vector<int32_t> status_, sweeps_;
  std::vector<double> rel_;
  std::vector<float> feat_;
  std::vector<char> chose_qrp_, fell_back_;
  DevBuf<float> dfeat_, dprob_;
  std::vector<SysDims> dims_;
  double const *oD_ = nullptr, *oE_ = nullptr, *oB_ = nullptr, *oC_ = nullptr, *oH_ = nullptr, *oCq_ = nullptr,
               *oF_ = nullptr;
  // refinement loop over the systems of `eng` not marked done (device
  // convergence test; done systems are skipped by every launch).
  // have_residual: R_ and norms_ already hold b - A x of the current X, so
  // the first sweep starts with the convergence test
  template <typename E>
  void refine(E& eng, double const* B, double* X, double tol, int max_iter, std::vector<char>& done, BatchReport& rep,
              int giveup = 0, bool have_residual = false);
  // captured graphs: factor+solve, residual+norms, correction
  struct Graphs {
    const double* B = nullptr;
    double* X = nullptr;
    hipGraphExec_t fs = nullptr, res = nullptr, corr = nullptr;
  } g_;
  void drop_graphs();
};

}  // namespace bcr
