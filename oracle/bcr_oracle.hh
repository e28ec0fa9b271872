// ORACLE — test infrastructure only. Nothing in the product links this file.
//
// CPU restatement of alglin::BorderedCR<real> (src/BABD_BorderedCR.cc,
// src/Alglin/BABD_BorderedCR.hxx).  Every method cites the reference lines it
// follows.  It is used three ways and only these: (1) the checker in tests/,
// (2) the checker in __graft_entry__.smoke(), (3) the timed CPU baseline of
// bench.py (reported beside the GPU figure, never as the GPU figure).
//
// Storage follows the reference's block layout (hxx:972-998): D_k, E_k are
// n x n column-major, B_k n x nx, C_k nr x n (k = 0..N), H = [H0 HN Hq Hp]
// (n+qr) x Nc, Cq nr x qx, F nr x nx.  The elimination order of the
// threaded path (num_parallel_block > 1) is the reference's range partition
// (cc:139-158) + reduced cyclic reduction (cc:1175-1266); with one block
// range it is pure cyclic reduction (cc:832-834).
#pragma once

#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>

#include "dense.hh"
#include "lastblock.hh"

namespace oracle {

// Minimal fork-join pool with the run/wait shape of Utils::ThreadPoolBase
// (used at cc:816-818, 1501-1503).
class ForkJoinPool {
 public:
  explicit ForkJoinPool(int nthreads) : n_(std::max(1, nthreads)) {
    for (int i = 0; i < n_; ++i) workers_.emplace_back([this] { loop(); });
  }
  ~ForkJoinPool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }
  int thread_count() const { return n_; }
  void run(std::function<void()> f) {
    {
      std::lock_guard<std::mutex> g(mu_);
      tasks_.push_back(std::move(f));
      ++pending_;
    }
    cv_.notify_one();
  }
  void wait() {
    std::unique_lock<std::mutex> g(mu_);
    done_cv_.wait(g, [this] { return pending_ == 0; });
  }

 private:
  void loop() {
    for (;;) {
      std::function<void()> f;
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [this] { return stop_ || !tasks_.empty(); });
        if (stop_ && tasks_.empty()) return;
        f = std::move(tasks_.back());
        tasks_.pop_back();
      }
      f();
      {
        std::lock_guard<std::mutex> g(mu_);
        if (--pending_ == 0) done_cv_.notify_all();
      }
    }
  }
  int n_;
  std::vector<std::thread> workers_;
  std::vector<std::function<void()>> tasks_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  int pending_{0};
  bool stop_{false};
};

template <class T>
class BorderedCR {
 public:
  enum Choice { LU = 0, QR = 1, QRP = 2 };                                     // hxx:111-115
  enum LastChoice { L_LU = 0, L_LUPQ, L_QR, L_QRP, L_SVD, L_LSS, L_LSY, L_PINV };  // hxx:118-127
  enum LastChoice2 { L2_NONE = 0, L2_SVD, L2_LSS, L2_LSY, L2_PINV };           // hxx:130-136

  // hxx:373-384: a pool of `nthreads` (<= 1 means the single-thread fake
  // pool), num_parallel_block = 0 means "as many as threads".
  explicit BorderedCR(int nthreads, integer num_parallel_block = 0) {
    if (nthreads > 1) {
      pool_.reset(new ForkJoinPool(nthreads));
      max_pb_ = num_parallel_block == 0 ? nthreads : std::min(nthreads, num_parallel_block);
    } else {
      max_pb_ = 1;
    }
  }

  int thread_count() const { return pool_ ? pool_->thread_count() : 1; }

  // hxx:407-423
  void set_num_parallel_block(integer npb) {
    max_pb_ = std::max(integer(1), std::min(integer(thread_count()), npb));
    if (N_ > 0) {
      // re-partition; the loaded blocks are kept (their storage does not
      // depend on the partition)
      auto D = D_, E = E_, B = B_, C = C_, Cq = Cq_, H = H_;
      std::vector<T> F(F_.begin(), F_.begin() + size_t(nr_) * nx_);
      integer N = N_;
      N_ = -1;
      allocate(N, n_, qr_, qx_, nr_, nx_);
      D_ = D; E_ = E; B_ = B; C_ = C; Cq_ = Cq; H_ = H;
      std::copy(F.begin(), F.end(), F_.begin());
    }
  }
  void set_factorize_use_thread(bool v) { fac_thread_ = v; }
  void set_solve_use_thread(bool v) { sol_thread_ = v; }
  void select(Choice c) { sel_ = c; }
  void select_last(LastChoice c) { last_ = c; }
  void select_last2(LastChoice2 c) { last2_ = c; }

  // cc:42-170 (storage sizes and the block-range partition)
  void allocate(integer N, integer n, integer qr, integer qx, integer nr, integer nx) {
    integer const NBLK = std::max(integer(1), max_pb_);
    if (N == N_ && n == n_ && qr == qr_ && qx == qx_ && nr == nr_ && nx == nx_ &&
        integer(kBlock_.size()) == NBLK)
      return;  // cc:55-61
    N_ = N; n_ = n; qr_ = qr; qx_ = qx; nr_ = nr; nx_ = nx;
    Nr_ = 2 * n + nr + qr;
    Nc_ = 2 * n + nx + qx;
    Tsize_ = 2 * n * n + n;
    D_.assign(size_t(N) * n * n, T(0));
    E_.assign(size_t(N) * n * n, T(0));
    B_.assign(size_t(N) * n * nx, T(0));
    C_.assign(size_t(N + 1) * nr * n, T(0));
    Cq_.assign(size_t(nr) * qx, T(0));
    F_.assign(size_t(NBLK) * nr * nx, T(0));
    H_.assign(size_t(n + qr) * Nc_, T(0));
    Tm_.assign(size_t(N) * Tsize_, T(0));
    Perm_.assign(size_t(N) * n, 0);
    Hmat_.assign(size_t(Nr_) * Nc_, T(0));
    xb_thread_.assign(size_t(NBLK) * nr, T(0));
    perm_thread_.assign(size_t(NBLK) * n, 0);
    iBlock_.assign(2 * NBLK, 0);
    kBlock_.assign(NBLK, 0);
    // cc:141-158: every range keeps at least 2 blocks
    used_pb_ = NBLK;
    reduced_nblk_ = 2 * NBLK;
    while (reduced_nblk_ > N) { reduced_nblk_ -= 2; --used_pb_; }
    if (used_pb_ <= 1) {
      used_pb_ = 1;
      iBlock_[0] = 0;
      iBlock_[1] = N;
    } else {
      iBlock_[0] = 0;
      iBlock_[1] = N / used_pb_;
      for (integer t = 1; t < used_pb_; ++t) {
        iBlock_[2 * t] = iBlock_[2 * t - 1] + 1;
        iBlock_[2 * t + 1] = integer((int64_t(t + 1) * N) / used_pb_);
      }
      --reduced_nblk_;
    }
  }

  // dims (hxx:617-650)
  integer nrows() const { return n_ * (N_ + 1) + qr_ + nr_; }
  integer ncols() const { return n_ * (N_ + 1) + qx_ + nx_; }
  integer Nr() const { return Nr_; }
  integer Nc() const { return Nc_; }
  integer used_parallel_block() const { return used_pb_; }

  // element accessors (hxx:972-998)
  T& D(integer k, integer i, integer j) { return D_[size_t(k) * n_ * n_ + i + j * n_]; }
  T& E(integer k, integer i, integer j) { return E_[size_t(k) * n_ * n_ + i + j * n_]; }
  T& B(integer k, integer i, integer j) { return B_[size_t(k) * n_ * nx_ + i + j * n_]; }
  T& C(integer k, integer i, integer j) { return C_[size_t(k) * nr_ * n_ + i + j * nr_]; }
  T& Cq(integer i, integer j) { return Cq_[i + j * nr_]; }
  T& F(integer i, integer j) { return F_[i + j * nr_]; }
  T& H(integer i, integer j) { return H_[i + j * (n_ + qr_)]; }

  T* Dptr() { return D_.data(); }
  T* Eptr() { return E_.data(); }
  T* Bptr() { return B_.data(); }
  T* Cptr() { return C_.data(); }
  T* Cqptr() { return Cq_.data(); }
  T* Fptr() { return F_.data(); }
  T* Hptr() { return H_.data(); }
  T* Tptr() { return Tm_.data(); }
  integer* Permptr() { return Perm_.data(); }
  T* Hmatptr() { return Hmat_.data(); }
  size_t Dsize() const { return D_.size(); }
  size_t Bsize() const { return B_.size(); }
  size_t Csize() const { return C_.size(); }
  size_t Cqsize() const { return Cq_.size(); }
  size_t Fsize() const { return size_t(nr_) * nx_; }
  size_t Hsize() const { return H_.size(); }
  size_t Tsize() const { return Tm_.size(); }
  size_t Permsize() const { return Perm_.size(); }
  size_t Hmatsize() const { return Hmat_.size(); }
  integer kblock0() const { return kBlock_.empty() ? 0 : kBlock_[0]; }
  std::string const& last_error() const { return last_error_; }

  // cc:2118-2166: res += alpha*A*x in the block layout
  void add_Mv(T alpha, T const* x, T* res) const {
    integer const n = n_, N = N_;
    T const* xe = x + size_t(N) * n;
    T const* xq = xe + n;
    T const* xb = xq + qx_;
    T* yy = res;
    for (integer k = 0; k < N; ++k) {
      T const* xk = x + size_t(k) * n;
      gemv_n(n, n, alpha, &D_[size_t(k) * n * n], n, xk, 1, T(1), yy, 1);
      gemv_n(n, n, alpha, &E_[size_t(k) * n * n], n, xk + n, 1, T(1), yy, 1);
      gemv_n(n, nx_, alpha, &B_[size_t(k) * n * nx_], n, xb, 1, T(1), yy, 1);
      yy += n;
    }
    integer const m = n + qr_;
    T const* Hp = H_.data();
    gemv_n(m, n, alpha, Hp, m, x, 1, T(1), yy, 1);              Hp += m * n;
    gemv_n(m, n, alpha, Hp, m, xe, 1, T(1), yy, 1);             Hp += m * n;
    gemv_n(m, qx_, alpha, Hp, m, xq, 1, T(1), yy, 1);           Hp += m * qx_;
    gemv_n(m, nx_, alpha, Hp, m, xb, 1, T(1), yy, 1);
    if (nr_ > 0) {
      yy += m;
      gemv_n(nr_, nx_, alpha, F_.data(), nr_, xb, 1, T(1), yy, 1);
      for (integer k = 0; k <= N; ++k)
        gemv_n(nr_, n, alpha, &C_[size_t(k) * nr_ * n], nr_, x + size_t(k) * n, 1, T(1), yy, 1);
      gemv_n(nr_, qx_, alpha, Cq_.data(), nr_, x + size_t(N + 1) * n, 1, T(1), yy, 1);
    }
  }
  void Mv(T const* x, T* res) const {  // cc:2111-2116
    for (integer i = 0; i < nrows(); ++i) res[i] = 0;
    add_Mv(T(1), x, res);
  }

  // cc:2338-2340 (hxx:1079-1088)
  integer sparse_nnz() const {
    return n_ * N_ * (2 * n_ + nx_ + nr_) + nr_ * (n_ + qx_ + nx_) + (n_ + qr_) * (2 * n_ + qx_ + nx_);
  }

  // cc:2342-2413: pattern and values in the reference's block order
  void sparse(integer* I, integer* J, T* V, integer offs) const {
    integer const n = n_, N = N_;
    integer kk = 0;
    auto put = [&](integer i, integer j, T v) { I[kk] = i + offs; J[kk] = j + offs; V[kk] = v; ++kk; };
    for (integer k = 0; k < N; ++k) {
      for (integer ij = 0; ij < n * n; ++ij) put(k * n + ij % n, k * n + ij / n, D_[size_t(k) * n * n + ij]);
      for (integer ij = 0; ij < n * n; ++ij) put(k * n + ij % n, (k + 1) * n + ij / n, E_[size_t(k) * n * n + ij]);
      for (integer ij = 0; ij < n * nx_; ++ij)
        put(k * n + ij % n, (N + 1) * n + qx_ + ij / n, B_[size_t(k) * n * nx_ + ij]);
    }
    integer const nqr = n + qr_;
    for (integer ij = 0; ij < nqr * Nc_; ++ij) {
      integer j = ij / nqr;
      if (j >= n) j += N * n - n;
      put(N * n + ij % nqr, j, H_[ij]);
    }
    integer const i0 = (N + 1) * n + qr_;
    for (integer k = 0; k <= N; ++k)
      for (integer ij = 0; ij < nr_ * n; ++ij) put(i0 + ij % nr_, k * n + ij / nr_, C_[size_t(k) * nr_ * n + ij]);
    for (integer ij = 0; ij < nr_ * qx_; ++ij) put(i0 + ij % nr_, (N + 1) * n + ij / nr_, Cq_[ij]);
    for (integer ij = 0; ij < nr_ * nx_; ++ij) put(i0 + ij % nr_, (N + 1) * n + qx_ + ij / nr_, F_[ij]);
  }

  // cc:803-839
  bool factorize() {
    factorized_ = false;
    ok_ = true;
    integer const npb = used_pb_;
    if (npb > 1) {
      std::fill(F_.begin() + size_t(nr_) * nx_, F_.begin() + size_t(npb) * nr_ * nx_, T(0));
      if (fac_thread_ && pool_) {
        for (integer t = 0; t < npb; ++t) pool_->run([this, t] { factorize_block(t); });
        pool_->wait();
      } else {
        for (integer t = 0; t < npb; ++t) factorize_block(t);
      }
      for (integer t = 1; t < npb; ++t)
        for (integer i = 0; i < nr_ * nx_; ++i) F_[i] += F_[size_t(t) * nr_ * nx_ + i];
      if (ok_) ok_ = factorize_reduced();
    } else {
      iBlock_[0] = 0;
      iBlock_[1] = N_;
      factorize_block(0);
    }
    if (ok_) ok_ = load_and_factorize_last();
    factorized_ = ok_;
    return ok_;
  }

  // cc:1489-1542 (in place: b in, x out; x holds max(nrows,ncols) entries)
  bool solve(T* x) {
    integer const npb = used_pb_;
    T* xb = x + size_t(N_ + 1) * n_ + qr_;
    ok_ = true;
    if (npb > 1) {
      if (sol_thread_ && pool_) {
        for (integer t = 0; t < npb; ++t) pool_->run([this, t, x] { forward(t, x); });
        pool_->wait();
        if (!ok_) return false;
      } else {
        for (integer t = 0; t < npb; ++t) if (!forward(t, x)) return false;
      }
      for (integer t = 0; t < npb; ++t)
        for (integer i = 0; i < nr_; ++i) xb[i] += xb_thread_[size_t(t) * nr_ + i];
      if (!forward_reduced(x, xb)) return false;
    } else {
      if (!forward(0, x)) return false;
      for (integer i = 0; i < nr_; ++i) xb[i] += xb_thread_[i];
    }
    if (!solve_last(1, x, 0)) return false;
    if (npb > 1) {
      backward_reduced(1, x, 0);
      if (sol_thread_ && pool_) {
        for (integer t = 0; t < npb; ++t) pool_->run([this, t, x] { backward(t, 1, x, 0); });
        pool_->wait();
      } else {
        for (integer t = 0; t < npb; ++t) backward(t, 1, x, 0);
      }
    } else {
      backward(0, 1, x, 0);
    }
    return true;
  }

  // cc:1546-1605 (multi-RHS, column-major with leading dimension ld)
  bool solve(integer nrhs, T* X, integer ld) {
    integer const npb = used_pb_;
    integer const nn = nr_ * nrhs;
    std::vector<T> work(size_t(npb) * nn, T(0));
    T* XB = X + size_t(N_ + 1) * n_ + qr_;
    ok_ = true;
    if (npb > 1) {
      if (sol_thread_ && pool_) {
        for (integer t = 0; t < npb; ++t)
          pool_->run([this, t, nrhs, X, ld, &work, nn] { forward_n(t, nrhs, X, ld, work.data() + size_t(t) * nn); });
        pool_->wait();
      } else {
        for (integer t = 0; t < npb; ++t) forward_n(t, nrhs, X, ld, work.data() + size_t(t) * nn);
      }
      for (integer t = 0; t < npb; ++t)
        for (integer c = 0; c < nrhs; ++c)
          for (integer i = 0; i < nr_; ++i) XB[i + size_t(c) * ld] += work[size_t(t) * nn + i + c * nr_];
      if (ok_) ok_ = forward_n_reduced(nrhs, X, ld);
    } else {
      forward_n(0, nrhs, X, ld, work.data());
      for (integer c = 0; c < nrhs; ++c)
        for (integer i = 0; i < nr_; ++i) XB[i + size_t(c) * ld] += work[i + c * nr_];
    }
    if (ok_) ok_ = solve_last(nrhs, X, ld);
    if (!ok_) return false;
    if (npb > 1) {
      backward_reduced(nrhs, X, ld);
      if (sol_thread_ && pool_) {
        for (integer t = 0; t < npb; ++t) pool_->run([this, t, nrhs, X, ld] { backward(t, nrhs, X, ld); });
        pool_->wait();
      } else {
        for (integer t = 0; t < npb; ++t) backward(t, nrhs, X, ld);
      }
    } else {
      backward(0, nrhs, X, ld);
    }
    return true;
  }

  // cc:208-235 (same fields; the tests match substrings of this text)
  std::string info() const {
    static char const* ch[] = {"CyclicReduction+LU", "CyclicReduction+QR", "CyclicReduction+QRP"};
    static char const* la[] = {"LastBlock LU", "LastBlock LUPQ", "LastBlock QR", "LastBlock QRP",
                               "LastBlock SVD", "LastBlock LSS", "LastBlock LSY", "LastBlock PINV"};
    static char const* l2[] = {"LastBlock NONE", "LastBlock SVD", "LastBlock LSS", "LastBlock LSY", "LastBlock PINV"};
    std::ostringstream s;
    s << "rows   = " << nrows() << "\ncols   = " << ncols() << "\nnblock = " << N_ << "\nn      = " << n_
      << "\nqr     = " << qr_ << "\nqx     = " << qx_ << "\nnr     = " << nr_ << "\nnx     = " << nx_
      << std::boolalpha << "\ntreads = " << thread_count() << "/nblk:" << used_pb_ << "/factorize:" << fac_thread_
      << "/solve:" << sol_thread_ << "\nintern = " << ch[sel_] << "\nlast   = " << la[last_]
      << "\nlast2  = " << l2[last2_] << "\n";
    return s.str();
  }

 private:
  // one pair elimination: rows (jp, j) eliminate x_j; cc:1090-1154 / 1194-1260
  bool eliminate_pair(integer j, integer jp, integer jpp, T* Fwork, integer th) {
    integer const n = n_;
    integer const n2 = 2 * n;
    T* Tp = &Tm_[size_t(j) * Tsize_];
    integer* P = &Perm_[size_t(j) * n];
    T* Dj = &D_[size_t(j) * n * n];
    T* Djp = &D_[size_t(jp) * n * n];
    T* Ej = &E_[size_t(j) * n * n];
    T* Ejp = &E_[size_t(jp) * n * n];
    gecopy(n, n, Ejp, n, Tp, n2);
    gecopy(n, n, Dj, n, Tp + n, n2);
    if (!buildT(Tp, P, th)) return false;
    gezero(n, n, Dj, n);
    gezero(n, n, Ejp, n);
    applyT(Tp, P, Djp, n, Dj, n, n);
    applyT(Tp, P, Ejp, n, Ej, n, n);
    if (nx_ > 0) applyT(Tp, P, &B_[size_t(jp) * n * nx_], n, &B_[size_t(j) * n * nx_], n, nx_);
    T* Cj = &C_[size_t(j) * nr_ * n];
    if (nr_ > 0) {
      T* Cjp = &C_[size_t(jp) * nr_ * n];
      if (sel_ == QRP) {  // cc:1109-1115
        integer i = n;
        do {
          --i;
          if (P[i] > i)
            for (integer r = 0; r < nr_; ++r) std::swap(Cj[r + i * nr_], Cj[r + P[i] * nr_]);
        } while (i > 0);
      }
      trsm_runn(nr_, n, Tp, n2, Cj, nr_);
      gemm_nn(nr_, n, n, T(-1), Cj, nr_, Dj, n, T(1), Cjp, nr_);
      T* Cpp = &C_[size_t(jpp) * nr_ * n];
      gemm_nn(nr_, n, n, T(-1), Cj, nr_, Ej, n, T(1), Cpp, nr_);
    }
    if (nr_ * nx_ > 0) gemm_nn(nr_, nx_, n, T(-1), Cj, nr_, &B_[size_t(j) * n * nx_], n, T(1), Fwork, nr_);
    return true;
  }

  // cc:845-898
  bool buildT(T* Tp, integer* iperm, integer th) {
    integer const n = n_, n2 = 2 * n;
    switch (sel_) {
      case LU:
        if (getrf(n2, n, Tp, n2, iperm) != 0) {
          set_error("BorderedCR::buildT [LU] getrf failed");
          return false;
        }
        break;
      case QR:
        geqrf(n2, n, Tp, n2, Tp + 2 * n * n);
        break;
      case QRP: {
        integer* Pt = &perm_thread_[size_t(th) * n];
        geqp3(n2, n, Tp, n2, Pt, Tp + 2 * n * n);
        // hxx:265-275: permutation -> exchange list (0-based)
        for (integer i = 0; i < n; ++i) {
          integer jj = i;
          while (jj < n) { if (Pt[jj] == i + 1) break; ++jj; }
          std::swap(Pt[jj], Pt[i]);
          iperm[i] = jj;
        }
        break;
      }
    }
    return true;
  }

  // cc:917-982: W = [TOP;BOTTOM] <- transform; TOP <- W(n:2n), BOTTOM <- W(0:n)
  void applyT(T const* Tp, integer const* iperm, T* TOP, integer ldTOP, T* BOTTOM, integer ldBOT,
              integer ncol) const {
    integer const n = n_, n2 = 2 * n;
    thread_local std::vector<T> W;
    if (W.size() < size_t(n2) * ncol) W.resize(size_t(n2) * ncol);
    gecopy(n, ncol, TOP, ldTOP, W.data(), n2);
    gecopy(n, ncol, BOTTOM, ldBOT, W.data() + n, n2);
    if (sel_ == LU) {
      laswp(ncol, W.data(), n2, 0, n - 1, iperm);
      trsm_llnu(n, ncol, Tp, n2, W.data(), n2);
      gemm_nn(n, ncol, n, T(-1), Tp + n, n2, W.data(), n2, T(1), W.data() + n, n2);
    } else {
      ormqr_lt(n2, ncol, n, Tp, n2, Tp + 2 * n * n, W.data(), n2);
    }
    gecopy(n, ncol, W.data() + n, n2, TOP, ldTOP);
    gecopy(n, ncol, W.data(), n2, BOTTOM, ldBOT);
  }

  // cc:1055-1171
  void factorize_block(integer t) {
    integer const ib = iBlock_[2 * t], eb = iBlock_[2 * t + 1], nblk = eb - ib;
    T* Fwork = &F_[size_t(t) * nr_ * nx_];
    integer k = 1;
    while (k < nblk) {
      for (integer j = ib + k; j < eb; j += 2 * k) {
        if (!eliminate_pair(j, j - k, std::min(j + k, eb), Fwork, t)) { ok_ = false; return; }
      }
      k *= 2;
    }
    kBlock_[t] = k;
  }

  // cc:1175-1266
  bool factorize_reduced() {
    integer k = 1;
    while (k < reduced_nblk_) {
      for (integer jj = k; jj < reduced_nblk_; jj += 2 * k) {
        integer const j = iBlock_[jj], jp = iBlock_[jj - k];
        integer const jpp = iBlock_[std::min(jj + k, reduced_nblk_)];
        if (!eliminate_pair(j, jp, jpp, F_.data(), 0)) return false;
      }
      k *= 2;
    }
    return true;
  }

  // cc:1281-1374
  bool load_and_factorize_last() {
    integer const n = n_, Nr = Nr_;
    T* W0 = Hmat_.data();
    T* WN = W0 + n * Nr;
    T* Wq = WN + n * Nr;
    T* Wp = Wq + qx_ * Nr;
    gecopy(n, n, D_.data(), n, W0, Nr);
    gecopy(n, n, E_.data(), n, WN, Nr);
    gezero(n, qx_, Wq, Nr);
    gecopy(n, nx_, B_.data(), n, Wp, Nr);
    gecopy(n + qr_, Nc_, H_.data(), n + qr_, Hmat_.data() + n, Nr);
    integer const offs = 2 * n + qr_;
    gecopy(nr_, n, C_.data(), nr_, W0 + offs, Nr);
    gecopy(nr_, n, &C_[size_t(N_) * nr_ * n], nr_, WN + offs, Nr);
    gecopy(nr_, qx_, Cq_.data(), nr_, Wq + offs, Nr);
    gecopy(nr_, nx_, F_.data(), nr_, Wp + offs, Nr);
    using K = typename DenseSolver<T>::Kind;
    // PINV with Nc > Nr factorizes the transpose in the reference
    // (cc:1343); the SVD-based solver here handles both shapes directly.
    bool ok = last_solver_.factorize(K(int(last_)), Nr_, Nc_, Hmat_.data(), Nr_);
    use_solver2_ = false;
    if (!ok) {
      use_solver2_ = true;
      static K const map2[] = {K::LU, K::SVD, K::LSS, K::LSY, K::PINV};
      if (last2_ != L2_NONE) ok = last_solver2_.factorize(map2[last2_], Nr_, Nc_, Hmat_.data(), Nr_);
    }
    if (!ok) set_error("BorderedCR<t_Value>::load_and_factorize_last failed");
    return ok;
  }

  // cc:1385-1479: the first block of the rhs is moved next to the H/C rows
  bool solve_last(integer nrhs, T* x, integer ld) {
    integer const n = n_;
    if (nrhs == 1) ld = std::max(nrows(), ncols());
    T* X = x + size_t(N_ - 1) * n;
    for (integer c = 0; c < nrhs; ++c)
      for (integer i = 0; i < n; ++i) std::swap(X[i + size_t(c) * ld], x[i + size_t(c) * ld]);
    DenseSolver<T> const& S = use_solver2_ ? last_solver2_ : last_solver_;
    if (use_solver2_ && last2_ == L2_NONE) {
      set_error("BorderedCR<t_Value>::solve_last failed");
      return false;
    }
    std::vector<T> tmp(std::max(Nr_, Nc_));
    for (integer c = 0; c < nrhs; ++c) {
      T* Xc = X + size_t(c) * ld;
      for (integer i = 0; i < Nr_; ++i) tmp[i] = Xc[i];
      S.solve(tmp.data());
      for (integer i = 0; i < Nc_; ++i) Xc[i] = tmp[i];
    }
    for (integer c = 0; c < nrhs; ++c)
      for (integer i = 0; i < n; ++i) std::swap(X[i + size_t(c) * ld], x[i + size_t(c) * ld]);
    return true;
  }

  // one pair of the forward sweep: cc:1640-1650 / 1696-1710
  void forward_pair(integer j, integer jp, integer nrhs, T* x, integer ld, T* work) const {
    integer const n = n_;
    T const* Tp = &Tm_[size_t(j) * Tsize_];
    integer const* P = &Perm_[size_t(j) * n];
    applyT(Tp, P, x + size_t(jp) * n, ld, x + size_t(j) * n, ld, nrhs);
    if (nr_ > 0)
      gemm_nn(nr_, nrhs, n, T(-1), &C_[size_t(j) * nr_ * n], nr_, x + size_t(j) * n, ld, T(1), work, nr_);
  }

  // cc:1615-1661 and cc:1665-1719 (nrhs = 1 uses ld = row count)
  bool forward(integer t, T* x) { forward_n(t, 1, x, std::max(nrows(), ncols()), &xb_thread_[size_t(t) * nr_]); return true; }
  void forward_n(integer t, integer nrhs, T* x, integer ld, T* work) {
    integer const ib = iBlock_[2 * t], eb = iBlock_[2 * t + 1], nblk = eb - ib;
    for (integer i = 0; i < nr_ * nrhs; ++i) work[i] = 0;
    integer k = 1;
    while (k < nblk) {
      for (integer jj = k; jj < nblk; jj += 2 * k) forward_pair(ib + jj, ib + jj - k, nrhs, x, ld, work);
      k *= 2;
    }
  }

  // cc:1723-1795
  bool forward_reduced(T* x, T* xb) { return forward_n_reduced_impl(1, x, std::max(nrows(), ncols()), xb); }
  bool forward_n_reduced(integer nrhs, T* x, integer ld) {
    return forward_n_reduced_impl(nrhs, x, ld, x + size_t(N_ + 1) * n_ + qr_);
  }
  bool forward_n_reduced_impl(integer nrhs, T* x, integer ld, T* xb) {
    integer const n = n_;
    integer k = 1;
    while (k < reduced_nblk_) {
      for (integer jj = k; jj < reduced_nblk_; jj += 2 * k) {
        integer const j = iBlock_[jj], jp = iBlock_[jj - k];
        T const* Tp = &Tm_[size_t(j) * Tsize_];
        applyT(Tp, &Perm_[size_t(j) * n], x + size_t(jp) * n, ld, x + size_t(j) * n, ld, nrhs);
        if (nr_ > 0)
          gemm_nn(nr_, nrhs, n, T(-1), &C_[size_t(j) * nr_ * n], nr_, x + size_t(j) * n, ld, T(1), xb, ld);
      }
      k *= 2;
    }
    return true;
  }

  // one pair of the backward sweep: cc:1830-1856 / 1900-1939
  void backward_pair(integer j, integer jp, integer jpp, integer nrhs, T* x, integer ld) const {
    integer const n = n_;
    T const* xn = x + size_t(N_ + 1) * n + qx_;
    T* xj = x + size_t(j) * n;
    gemm_nn(n, nrhs, n, T(-1), &D_[size_t(j) * n * n], n, x + size_t(jp) * n, ld, T(1), xj, ld);
    gemm_nn(n, nrhs, n, T(-1), &E_[size_t(j) * n * n], n, x + size_t(jpp) * n, ld, T(1), xj, ld);
    if (nx_ > 0) gemm_nn(n, nrhs, nx_, T(-1), &B_[size_t(j) * n * nx_], n, xn, ld, T(1), xj, ld);
    trsm_lunn(n, nrhs, &Tm_[size_t(j) * Tsize_], 2 * n, xj, ld);
    if (sel_ == QRP) {
      integer const* P = &Perm_[size_t(j) * n];
      for (integer i = 0; i < n; ++i)
        if (P[i] > i)
          for (integer c = 0; c < nrhs; ++c) std::swap(xj[i + size_t(c) * ld], xj[P[i] + size_t(c) * ld]);
    }
  }

  // cc:1805-1866 / cc:1870-1949
  void backward(integer t, integer nrhs, T* x, integer ld) {
    if (nrhs == 1) ld = std::max(nrows(), ncols());
    integer const ib = iBlock_[2 * t], eb = iBlock_[2 * t + 1];
    integer k = kBlock_[t];
    while ((k /= 2) > 0)
      for (integer j = ib + k; j < eb; j += 2 * k) backward_pair(j, j - k, std::min(j + k, eb), nrhs, x, ld);
  }

  // cc:1953-2075
  void backward_reduced(integer nrhs, T* x, integer ld) {
    if (nrhs == 1) ld = std::max(nrows(), ncols());
    integer k = 1;
    while (k < reduced_nblk_) k *= 2;
    while ((k /= 2) > 0)
      for (integer jj = k; jj < reduced_nblk_; jj += 2 * k)
        backward_pair(iBlock_[jj], iBlock_[jj - k], iBlock_[std::min(jj + k, reduced_nblk_)], nrhs, x, ld);
  }

  void set_error(char const* s) {
    std::lock_guard<std::mutex> g(err_mu_);
    last_error_ = s;
  }

  std::unique_ptr<ForkJoinPool> pool_;
  integer max_pb_{1}, used_pb_{1}, reduced_nblk_{0};
  integer N_{0}, n_{0}, qr_{0}, qx_{0}, nr_{0}, nx_{0}, Nr_{0}, Nc_{0}, Tsize_{0};
  bool fac_thread_{true}, sol_thread_{true}, factorized_{false}, use_solver2_{false};
  std::atomic<bool> ok_{true};
  Choice sel_{LU};
  LastChoice last_{L_LU};
  LastChoice2 last2_{L2_NONE};
  std::vector<T> D_, E_, B_, C_, Cq_, F_, H_, Tm_, Hmat_, xb_thread_;
  std::vector<integer> Perm_, perm_thread_, iBlock_, kBlock_;
  DenseSolver<T> last_solver_, last_solver2_;
  std::string last_error_{"no error"};
  std::mutex err_mu_;
};

}  // namespace oracle
