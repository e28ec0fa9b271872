// Launch sequences of the GPU engine (see bcr_engine.hh).
#include "bcr_engine.hh"

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "bcr_launch.hh"

namespace bcr {

namespace {
// the matrix-core pair (FP32 factors n <= 64, FP64 factors n <= 48)
template <typename TS, typename TW>
bool launch_mfma(FactorArgs<TS, TW> const& a, hipStream_t st, hipError_t* e) {
  return launch_pair_lu_mfma(a, st, e);
}
}  // namespace

void hip_check(hipError_t e, char const* what) {
  if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
void DevBuf<T>::resize(size_t c) {
  if (c <= count && p != nullptr) return;
  if (borrowed) throw HipError("working array of " + std::to_string(c) + " elements exceeds its fixed slice");
  release();
  if (c == 0) return;
  hip_check(hipMalloc(reinterpret_cast<void**>(&p), c * sizeof(T)), "hipMalloc");
  count = c;
}

template <typename T>
void DevBuf<T>::borrow(T* ptr, size_t c) {
  release();
  p = ptr;
  count = c;
  borrowed = true;
}

template <typename T>
void DevBuf<T>::release() {
  if (p && !borrowed) (void)hipFree(p);
  p = nullptr;
  count = 0;
  borrowed = false;
}

template struct DevBuf<double>;
template struct DevBuf<float>;
template struct DevBuf<int32_t>;
template struct DevBuf<char>;

template <typename TA, typename TW>
Engine<TA, TW>::Engine() {
  hip_check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreate");
  int dev = 0, lds = 0, cus = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
    lds_limit_ = size_t(lds);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
    max_chunks_ = 4 * cus;
}

template <typename TA, typename TW>
Engine<TA, TW>::Engine(hipStream_t external) : st_(external), own_stream_(false) {
  int dev = 0, lds = 0, cus = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
    lds_limit_ = size_t(lds);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
    max_chunks_ = 4 * cus;
}

template <typename TA, typename TW>
Engine<TA, TW>::~Engine() {
  for (auto& r : recs_) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : event_pool_) (void)hipEventDestroy(e);
  if (st_ && own_stream_) (void)hipStreamDestroy(st_);
}

template <typename TA, typename TW>
void Engine<TA, TW>::share_stream(hipStream_t s) {
  hip_check(hipStreamSynchronize(st_), "share_stream sync");
  if (own_stream_) hip_check(hipStreamDestroy(st_), "hipStreamDestroy");
  st_ = s;
  own_stream_ = false;
}

template <typename TA, typename TW>
hipEvent_t Engine<TA, TW>::get_event() {
  if (!event_pool_.empty()) {
    hipEvent_t e = event_pool_.back();
    event_pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  hip_check(hipEventCreate(&e), "hipEventCreate");
  return e;
}

template <typename TA, typename TW>
template <typename F>
void Engine<TA, TW>::timed(int phase, F&& f) {
  if (!timing_) { f(); return; }
  Rec r{phase, get_event(), get_event()};
  hip_check(hipEventRecord(r.a, st_), "hipEventRecord");
  f();
  hip_check(hipEventRecord(r.b, st_), "hipEventRecord");
  recs_.push_back(r);
  if (phase == kPair) acc_[6] += 1;
  if (phase == kForward || phase == kBackward) acc_[7] += 1;
}

template <typename TA, typename TW>
void Engine<TA, TW>::timing(double out[8], bool reset) {
  hip_check(hipStreamSynchronize(st_), "timing sync");
  for (auto& r : recs_) {
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
    acc_[r.phase] += ms;
    event_pool_.push_back(r.a);
    event_pool_.push_back(r.b);
  }
  recs_.clear();
  for (int i = 0; i < 8; ++i) out[i] = acc_[i];
  if (reset)
    for (double& v : acc_) v = 0;
}

template <typename TA, typename TW>
void Engine<TA, TW>::allocate(Dims const& d) {
  d_ = d;
  const size_t N = size_t(d.N), n = size_t(d.n), nr = size_t(d.nr), nx = size_t(d.nx);
  wD_.resize(std::max<size_t>(1, N * n * n));
  wE_.resize(std::max<size_t>(1, N * n * n));
  wB_.resize(std::max<size_t>(1, N * n * nx));
  wC_.resize(std::max<size_t>(1, (N + 1) * nr * n));
  Tm_.resize(std::max<size_t>(1, N * tsize(d.n)));
  perm_.resize(std::max<size_t>(1, N * n));
  Cpend_.resize(std::max<size_t>(1, 2 * (N + 1) * nr * n));
  Fpart_.resize(std::max<size_t>(1, N * nr * nx));
  Feff_.resize(std::max<size_t>(1, nr * nx));
  Hm_.resize(size_t(d.Nr()) * d.Nc());
  ipiv_.resize(std::max(d.Nr(), d.Nc()));
  jpiv_.resize(std::max(d.Nr(), d.Nc()));
  tau_.resize(std::max(d.Nr(), d.Nc()));
  status_.resize(4);
  tmpw_.resize(size_t(max_chunks_) * std::max<size_t>(1, nr * nx));
  tmpd_.resize(size_t(max_chunks_) * std::max<size_t>(1, nr * kRhsChunk));
  Cpart_.resize((N + 1) * std::max<size_t>(1, nr));
  Csum_.resize(std::max<size_t>(1, nr));
  hip_check(hipStreamSynchronize(st_), "allocate sync");
  build_schedule();
}

namespace {
// the slices of allocate_in, in order: element counts and element sizes
struct Carve {
  size_t off = 0;
  template <typename T>
  T* take(char* base, size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += std::max<size_t>(1, count) * sizeof(T);
    return p;
  }
};

template <typename TW, typename F>
void carve_all(Dims const& d, int nrhs, int max_chunks, F&& put) {
  const size_t N = size_t(d.N), n = size_t(d.n), nr = size_t(d.nr), nx = size_t(d.nx);
  const size_t m = nr * size_t(std::max(nrhs, 1));
  put(0, N * n * n);                 // wD
  put(1, N * n * n);                 // wE
  put(2, N * n * nx);                // wB
  put(3, (N + 1) * nr * n);          // wC
  put(4, N * tsize(d.n));            // Tm
  put(5, 2 * (N + 1) * nr * n);      // Cpend
  put(6, N * nr * nx);               // Fpart
  put(7, size_t(max_chunks) * std::max<size_t>(1, nr * nx));   // tmpw
  put(8, N * n);                     // perm (int32)
  put(9, size_t(std::max(d.Nr(), d.Nc())));   // ipiv (int32)
  put(10, 4);                        // status (int32)
  put(11, std::max<size_t>(1, nr * nx));   // Feff
  put(12, size_t(d.Nr()) * d.Nc());  // Hm
  put(13, size_t(max_chunks) * std::max<size_t>(1, std::max(nr * kRhsChunk, m)));   // tmpd
  put(14, (N + 1) * std::max<size_t>(1, nr));   // Cpart
  put(15, std::max<size_t>(1, nr));  // Csum
  put(16, std::max<size_t>(1, N * m));   // Xpartd
  (void)sizeof(TW);
}
}  // namespace

template <typename TA, typename TW>
size_t Engine<TA, TW>::bytes_in(Dims const& d, int nrhs, int max_chunks) {
  Carve c;
  carve_all<TW>(d, nrhs, max_chunks, [&](int k, size_t cnt) {
    if (k <= 7) c.take<TW>(nullptr, cnt);
    else if (k <= 10) c.take<int32_t>(nullptr, cnt);
    else c.take<double>(nullptr, cnt);
  });
  return (c.off + 255) & ~size_t(255);
}

template <typename TA, typename TW>
void Engine<TA, TW>::allocate_in(Dims const& d, int nrhs, char* base, size_t bytes) {
  const size_t need = bytes_in(d, nrhs, max_chunks_);
  if (need > bytes)
    throw HipError("working set of " + std::to_string(need) + " bytes exceeds the " + std::to_string(bytes) +
                   "-byte scratch array");
  d_ = d;
  Carve c;
  DevBuf<TW>* tw[] = {&wD_, &wE_, &wB_, &wC_, &Tm_, &Cpend_, &Fpart_, &tmpw_};
  DevBuf<int32_t>* ti[] = {&perm_, &ipiv_, &status_};
  DevBuf<double>* td[] = {&Feff_, &Hm_, &tmpd_, &Cpart_, &Csum_, &Xpartd_};
  carve_all<TW>(d, nrhs, max_chunks_, [&](int k, size_t cnt) {
    if (k <= 7) tw[k]->borrow(c.take<TW>(base, cnt), std::max<size_t>(1, cnt));
    else if (k <= 10) ti[k - 8]->borrow(c.take<int32_t>(base, cnt), std::max<size_t>(1, cnt));
    else td[k - 11]->borrow(c.take<double>(base, cnt), std::max<size_t>(1, cnt));
  });
  // every other array stays empty and refuses to grow: a path that would
  // need one (QR panels in global scratch, the SVD last block, a pair table)
  // fails instead of allocating
  for (DevBuf<TW>* b : {&gscratch_, &Xpartw_}) b->borrow(nullptr, 0);
  for (DevBuf<int32_t>* b : {&jpiv_, &svd_rank_, &pairtab_}) b->borrow(nullptr, 0);
  for (DevBuf<double>* b : {&tau_, &Hm2_, &svdW_, &svdVj_, &svdU_, &svdV_, &svdS_}) b->borrow(nullptr, 0);
  // pure cyclic reduction: the schedule needs no pair table on the device
  want_P_ = 1;
  build_schedule();
}

template <typename TA, typename TW>
void Engine<TA, TW>::set_partition(int P) {
  want_P_ = std::max(1, P);
  if (d_.N > 0) build_schedule();
}

// The launch list of one factorization (and, in the same order, of the
// forward sweep; reversed for the backward sweep).
template <typename TA, typename TW>
void Engine<TA, TW>::build_schedule() {
  const int N = d_.N;
  sched_.clear();
  copy_rows_.clear();
  copy_cols_.clear();
  // partition of cc:139-158: shrink P until 2P <= N
  int P = want_P_, red = 2 * want_P_;
  while (red > N) { red -= 2; --P; }
  if (P <= 1) {
    used_P_ = 1;
    iblock_ = {0, N};
    for (int L = 1; L < N; L *= 2) {
      Level l;
      int lg = 0;
      for (int t = L; t > 1; t >>= 1) ++lg;
      l.L = L; l.npairs = (N + L - 1) / (2 * L); l.slot = lg & 1; l.first = L == 1;
      sched_.push_back(l);
    }
    // blocks the first level does not touch: row N-1 when N is odd, C_N
    if (N > 1 && N % 2 == 1) { copy_rows_.push_back(N - 1); copy_cols_.push_back(N - 1); }
    copy_cols_.push_back(N);
    return;
  }
  used_P_ = P;
  iblock_.assign(size_t(2 * P), 0);
  iblock_[0] = 0;
  iblock_[1] = N / P;
  for (int t = 1; t < P; ++t) {
    iblock_[2 * t] = iblock_[2 * t - 1] + 1;
    iblock_[2 * t + 1] = int((int64_t(t + 1) * N) / P);
  }
  const int M = 2 * P - 1;   // reduced chain: rows iblock_[0..M-1], last column iblock_[M] = N
  std::vector<int32_t> tab;
  auto slot_of = [](int k) { int lg = 0; for (int t = k; t > 1; t >>= 1) ++lg; return lg & 1; };
  // ranges: every range's level k in one launch (cc:1072-1080)
  int maxlen = 0;
  for (int t = 0; t < P; ++t) maxlen = std::max(maxlen, iblock_[2 * t + 1] - iblock_[2 * t]);
  for (int k = 1; k < maxlen; k *= 2) {
    Level l;
    l.L = k; l.slot = slot_of(k); l.first = k == 1; l.off = int64_t(tab.size());
    for (int t = 0; t < P; ++t) {
      const int ib = iblock_[2 * t], eb = iblock_[2 * t + 1];
      for (int j = ib + k; j < eb; j += 2 * k) {
        tab.push_back(j - k); tab.push_back(j); tab.push_back(std::min(j + k, eb));
        ++l.npairs;
      }
    }
    if (l.npairs > 0) sched_.push_back(l);
  }
  // the reduced chain (cc:1181-1190): jp = iblock[jj-k], j = iblock[jj],
  // jpp = iblock[min(jj+k, M)]; the range-end blocks' pending border terms
  // are read as j at its first level
  for (int k = 1; k < M; k *= 2) {
    Level l;
    l.L = k; l.slot = slot_of(k); l.first = false; l.off = int64_t(tab.size());
    for (int jj = k; jj < M; jj += 2 * k) {
      tab.push_back(iblock_[jj - k]); tab.push_back(iblock_[jj]); tab.push_back(iblock_[std::min(jj + k, M)]);
      ++l.npairs;
    }
    sched_.push_back(l);
  }
  // rows and C blocks the first launch does not read from A are copied
  std::vector<char> row(size_t(N), 0), col(size_t(N) + 1, 0);
  if (!sched_.empty() && sched_[0].first)
    for (int p = 0; p < sched_[0].npairs; ++p) {
      const int jp = tab[size_t(sched_[0].off) + 3 * p], j = tab[size_t(sched_[0].off) + 3 * p + 1];
      row[jp] = row[j] = 1;
      col[jp] = col[j] = 1;
    }
  for (int r = 0; r < N; ++r) if (!row[r]) copy_rows_.push_back(r);
  for (int c = 0; c <= N; ++c) if (!col[c]) copy_cols_.push_back(c);
  pairtab_.resize(std::max<size_t>(1, tab.size()));
  hip_check(hipMemcpy(pairtab_.p, tab.data(), tab.size() * sizeof(int32_t), hipMemcpyHostToDevice), "pair table");
}

template <typename TA, typename TW>
void Engine<TA, TW>::upload(TA const* D, TA const* E, TA const* B, TA const* C, TA const* H, TA const* Cq,
                            TA const* F) {
  const size_t N = size_t(d_.N), n = size_t(d_.n), nr = size_t(d_.nr), nx = size_t(d_.nx);
  auto up = [&](DevBuf<TA>& b, TA const* src, size_t cnt) {
    b.resize(std::max<size_t>(1, cnt));
    if (cnt) hip_check(hipMemcpyAsync(b.p, src, cnt * sizeof(TA), hipMemcpyHostToDevice, st_), "H2D");
  };
  up(aD_, D, N * n * n);
  up(aE_, E, N * n * n);
  up(aB_, B, N * n * nx);
  up(aC_, C, (N + 1) * nr * n);
  up(aH_, H, (n + d_.qr) * size_t(d_.Nc()));
  up(aCq_, Cq, nr * d_.qx);
  up(aF_, F, nr * nx);
  pD_ = aD_.p; pE_ = aE_.p; pB_ = aB_.p; pC_ = aC_.p; pH_ = aH_.p; pCq_ = aCq_.p; pF_ = aF_.p;
}

template <typename TA, typename TW>
void Engine<TA, TW>::bind_device(TA const* D, TA const* E, TA const* B, TA const* C, TA const* H, TA const* Cq,
                                 TA const* F) {
  pD_ = D; pE_ = E; pB_ = B; pC_ = C; pH_ = H; pCq_ = Cq; pF_ = F;
}

template <typename TA, typename TW>
void Engine<TA, TW>::level_factor(Level const& l) {
  const int N = d_.N;
  const int npairs = l.npairs;
  const bool first = l.first;
  auto fill = [&](auto& a) {
    a.N = N; a.n = d_.n; a.nr = d_.nr; a.nx = d_.nx;
    a.L = l.L; a.npairs = npairs; a.slot_cur = l.slot; a.first = first ? 1 : 0;
    a.D = wD_.p; a.E = wE_.p; a.B = wB_.p; a.C = wC_.p;
    a.Tm = Tm_.p; a.perm = perm_.p; a.Cpend = Cpend_.p; a.Fpart = Fpart_.p; a.status = status_.p;
    a.pairs = l.off >= 0 ? pairtab_.p + l.off : nullptr;
  };
  auto run = [&](auto& a) {
    using A = std::remove_reference_t<decltype(a)>;
    (void)sizeof(A);
    hipError_t e = hipSuccess;
    // variants: 0 default (16 < n: the matrix-core pair where it applies,
    // FP32 factors n <= 64, FP64 factors n <= 48; else wave-per-row kernels
    // for n <= 32, column-split waves up to n = 64, register tiles beyond), 3
    // register tiles only, 4 column-split waves only, 5 the matrix-core pair
    // where it applies
    const int v = pair_variant_;
    const bool lu_fast =
        method_ == kLU &&
        ((v == 4 && launch_pair_lu_cols(a, st_, &e)) ||
         (v == 5 && launch_mfma(a, st_, &e)) ||
         (v != 3 && v != 4 && v != 5 &&
          ((d_.n > 16 && launch_mfma(a, st_, &e)) || launch_pair_lu_split(a, st_, &e) ||
           launch_pair_lu_fast(a, st_, &e) || launch_pair_lu_cols(a, st_, &e))) ||
         launch_pair_lu_tiled(a, st_, &e));
    if (lu_fast) {
      stats_.fast_path = 1;
      hip_check(e, "pair_lu_fast");
      return;
    }
    if (method_ != kLU && launch_pair_qr(method_, a, lds_limit_, st_, &e)) {
      hip_check(e, "pair_qr");
      return;
    }
    const size_t bytes = size_t(2 * d_.n + d_.nr) * (3 * d_.n + d_.nx) * sizeof(TW);
    if (bytes > lds_limit_) gscratch_.resize(size_t(npairs) * (bytes / sizeof(TW)));
    hip_check(launch_pair_generic(method_, a, gscratch_.p, lds_limit_, st_), "pair_generic");
  };
  if (first) {
    FactorArgs<TA, TW> a{};
    fill(a);
    a.sD = pD_; a.sE = pE_; a.sB = pB_; a.sC = pC_;
    run(a);
  } else {
    FactorArgs<TW, TW> a{};
    fill(a);
    a.sD = wD_.p; a.sE = wE_.p; a.sB = wB_.p; a.sC = wC_.p;
    run(a);
  }
}

template <typename TA, typename TW>
void Engine<TA, TW>::factorize_async() {
  const int N = d_.N, n = d_.n, nr = d_.nr, nx = d_.nx;
  const size_t nn = size_t(n) * n, rn = size_t(nr) * n, bn = size_t(n) * nx;
  if (!pD_) throw HipError("factorize: no operator loaded");
  if ((last_method_ == kLastLU || last_method_ == kLastLUPQ) && d_.Nr() != d_.Nc())
    throw HipError("factorize: the last-block LU needs Nr == Nc (qr+nr == qx+nx)");
  if ((last_method_ == kLastQR || last_method_ == kLastQRP) && d_.Nr() < d_.Nc())
    throw HipError("factorize: the last-block QR needs Nr >= Nc (qr+nr >= qx+nx)");
  // per-pair LDS tables of the generic and sweep kernels hold up to 64 pivots
  // and 16 border rows
  if (n > 64 || nr > 16) throw HipError("factorize: block size n <= 64 and nr <= 16 are supported");
  stats_ = FactorStats{};
  hip_check(launch_convert<int32_t, int32_t>(nullptr, status_.p, 4, st_), "status");
  // pending border contributions accumulate per factorization
  timed(kReduce, [&] { hip_check(launch_convert<TW, TW>(nullptr, Cpend_.p, int64_t(Cpend_.count), st_), "zero pending"); });
  // rows and border blocks the first launch does not read from A (row N-1
  // when N is odd and C_N in pure cyclic reduction; the range ends too when
  // partitioned) go to the working arrays as they are
  for (int r : copy_rows_) {
    const size_t k = size_t(r);
    hip_check(launch_convert<TA, TW>(pD_ + k * nn, wD_.p + k * nn, int64_t(nn), st_), "copy D");
    hip_check(launch_convert<TA, TW>(pE_ + k * nn, wE_.p + k * nn, int64_t(nn), st_), "copy E");
    hip_check(launch_convert<TA, TW>(pB_ + k * bn, wB_.p + k * bn, int64_t(bn), st_), "copy B");
  }
  for (int c : copy_cols_) {
    const size_t k = size_t(c);
    hip_check(launch_convert<TA, TW>(pC_ + k * rn, wC_.p + k * rn, int64_t(rn), st_), "copy C");
  }
  for (Level const& l : sched_) {
    timed(kPair, [&] { level_factor(l); });
    ++stats_.levels;
  }
  // F_eff = F + sum_j Fpart[j]  (cc:823-829 / 1146-1154)
  const int mF = nr * nx;
  if (mF > 0)
    timed(kReduce, [&] {
      hip_check(launch_convert<TA, double>(pF_, Feff_.p, mF, st_), "F");
      if (N > 1)
        hip_check(launch_reduce<TW, double>(Fpart_.p, 1, N - 1, mF, tmpw_.p, max_chunks_, Feff_.p, mF, mF, st_),
                  "reduce F");
    });
  AssembleArgs<TW, TA> as{};
  as.N = N; as.n = n; as.qr = d_.qr; as.qx = d_.qx; as.nr = nr; as.nx = nx; as.Nr = d_.Nr(); as.Nc = d_.Nc();
  as.D0 = wD_.p; as.E0 = wE_.p; as.B0 = wB_.p; as.C = wC_.p; as.Cpend = Cpend_.p;
  as.H = pH_; as.Cq = pCq_; as.Feff = Feff_.p; as.Hm = Hm_.p;
  if (N == 1) {  // no reduction level: row 0 and C_0 come straight from A
    hip_check(launch_convert<TA, TW>(pD_, wD_.p, int64_t(nn), st_), "copy D0");
    hip_check(launch_convert<TA, TW>(pE_, wE_.p, int64_t(nn), st_), "copy E0");
    hip_check(launch_convert<TA, TW>(pB_, wB_.p, int64_t(bn), st_), "copy B0");
    hip_check(launch_convert<TA, TW>(pC_, wC_.p, int64_t(rn), st_), "copy C0");
  }
  const bool svd_first = last_method_ >= kLastSVD;
  if (svd_first || last2_) {
    const int k = std::min(d_.Nr(), d_.Nc()), rows = std::max(d_.Nr(), d_.Nc());
    svdW_.resize(size_t(rows) * k);
    svdVj_.resize(size_t(k) * k);
    svdU_.resize(size_t(d_.Nr()) * k);
    svdV_.resize(size_t(d_.Nc()) * k);
    svdS_.resize(size_t(k));
    svd_rank_.resize(1);
    if (!svd_first) Hm2_.resize(size_t(d_.Nr()) * d_.Nc());
  }
  SvdArgs sv{};
  sv.Nr = d_.Nr(); sv.Nc = d_.Nc(); sv.W = svdW_.p; sv.Vj = svdVj_.p; sv.U = svdU_.p; sv.V = svdV_.p;
  sv.S = svdS_.p; sv.rank = svd_rank_.p; sv.status = status_.p;
  timed(kLast, [&] {
    hip_check(launch_assemble_last(as, st_), "assemble last");
    if (svd_first) {
      sv.A = Hm_.p;
      sv.conditional = 0;
      hip_check(launch_svd_last(sv, st_), "svd last");
      return;
    }
    if (last2_)
      hip_check(launch_convert<double, double>(Hm_.p, Hm2_.p, int64_t(d_.Nr()) * d_.Nc(), st_), "copy last block");
    switch (last_method_) {
      case kLastLU: hip_check(launch_lu_last(Hm_.p, d_.Nc(), ipiv_.p, status_.p, lds_limit_, st_), "lu last"); break;
      case kLastLUPQ:
        hip_check(launch_lupq_last(Hm_.p, d_.Nc(), ipiv_.p, jpiv_.p, status_.p, lds_limit_, st_), "lupq last");
        break;
      default:
        hip_check(launch_qr_last(last_method_ == kLastQRP, Hm_.p, d_.Nr(), d_.Nc(), tau_.p, ipiv_.p, status_.p,
                                 lds_limit_, st_),
                  "qr last");
    }
    if (last2_) {   // the second solver runs on the device only if the first failed
      sv.A = Hm2_.p;
      sv.conditional = 1;
      hip_check(launch_svd_last(sv, st_), "svd last (second solver)");
    }
  });
}

template <typename TA, typename TW>
int32_t Engine<TA, TW>::status_sync() {
  int32_t s = 0;
  hip_check(hipMemcpyAsync(&s, status_.p, sizeof(int32_t), hipMemcpyDeviceToHost, st_), "status D2H");
  hip_check(hipStreamSynchronize(st_), "sync");
  return s;
}

template <typename TA, typename TW>
template <typename TX>
void Engine<TA, TW>::solve_async(TX* x, int nrhs, int64_t ldx) {
  const int N = d_.N, n = d_.n, nr = d_.nr;
  SolveArgs<TW, TX> a{};
  a.N = N; a.n = n; a.nr = nr; a.nx = d_.nx; a.qr = d_.qr; a.qx = d_.qx;
  a.nrhs = nrhs; a.ldx = ldx; a.x = x;
  a.D = wD_.p; a.E = wE_.p; a.B = wB_.p; a.C = wC_.p; a.Tm = Tm_.p; a.perm = perm_.p;
  const size_t m = size_t(nr) * nrhs;
  TX* Xpart = nullptr;
  TX* tmp = nullptr;
  if (sizeof(TX) == sizeof(double)) {
    Xpartd_.resize(std::max<size_t>(1, size_t(N) * m));
    tmpd_.resize(std::max<size_t>(1, size_t(max_chunks_) * m));
    Xpart = reinterpret_cast<TX*>(Xpartd_.p);
    tmp = reinterpret_cast<TX*>(tmpd_.p);
  } else {
    Xpartw_.resize(std::max<size_t>(1, size_t(N) * m));
    tmpw_.resize(std::max<size_t>(size_t(max_chunks_) * std::max<size_t>(1, size_t(nr) * d_.nx), size_t(max_chunks_) * m));
    Xpart = reinterpret_cast<TX*>(Xpartw_.p);
    tmp = reinterpret_cast<TX*>(tmpw_.p);
  }
  a.Xpart = Xpart;
  for (Level const& l : sched_) {
    a.L = l.L;
    a.npairs = l.npairs;
    a.pairs = l.off >= 0 ? pairtab_.p + l.off : nullptr;
    timed(kForward, [&] {
      hipError_t e = hipSuccess;
      if (method_ == kLU && sweep_fast_ &&
          (launch_sweep_lu_fast(true, a, st_, &e) || launch_sweep_lu_stream(true, a, st_, &e)))
        hip_check(e, "forward_fast");
      else hip_check(launch_forward(method_, a, st_), "forward");
    });
  }
  if (N > 1 && m > 0)
    timed(kReduce, [&] {
      hip_check(launch_reduce<TX, TX>(Xpart, 1, N - 1, int(m), tmp, max_chunks_, x + size_t(N + 1) * n + d_.qr, nr,
                                      ldx, st_),
                "reduce xb");
    });
  timed(kLast, [&] {
    LastSolveArgs<TX> ls{};
    ls.N = N; ls.n = n; ls.qr = d_.qr; ls.qx = d_.qx; ls.nr = nr; ls.nx = d_.nx; ls.Nr = d_.Nr(); ls.Nc = d_.Nc();
    ls.method = std::min(last_method_, int(kLastSVD)); ls.nrhs = nrhs; ls.Hm = Hm_.p; ls.P = ipiv_.p;
    ls.Q = jpiv_.p; ls.tau = tau_.p; ls.x = x; ls.ldx = ldx;
    ls.status = last2_ ? status_.p : nullptr;
    ls.U = svdU_.p; ls.V = svdV_.p; ls.S = svdS_.p; ls.rank = svd_rank_.p;
    hip_check(launch_last_solve<TX>(ls, lds_limit_, st_), "last solve");
  });
  for (auto it = sched_.rbegin(); it != sched_.rend(); ++it) {
    a.L = it->L;
    a.npairs = it->npairs;
    a.pairs = it->off >= 0 ? pairtab_.p + it->off : nullptr;
    timed(kBackward, [&] {
      hipError_t e = hipSuccess;
      // the back substitution of a QR pair is the LU one (R where U sits);
      // QRP adds the pair's column exchanges (the streamed sweep's exch)
      a.exch = method_ == kQRP ? 1 : 0;
      if ((method_ == kLU || method_ == kQR) && sweep_fast_ &&
          (launch_sweep_lu_fast(false, a, st_, &e) || launch_sweep_lu_stream(false, a, st_, &e)))
        hip_check(e, "backward_fast");
      else if (method_ == kQRP && sweep_fast_ && launch_sweep_lu_stream(false, a, st_, &e))
        hip_check(e, "backward_stream");
      else hip_check(launch_backward(method_, a, st_), "backward");
      a.exch = 0;
    });
  }
}

template <typename TA, typename TW>
void Engine<TA, TW>::residual_async(double const* x, double const* b, double* r) {
  if (!pD_) throw HipError("residual: no operator loaded or bound");
  MvArgs<TA> m{};
  m.N = d_.N; m.n = d_.n; m.qr = d_.qr; m.qx = d_.qx; m.nr = d_.nr; m.nx = d_.nx;
  m.D = pD_; m.E = pE_; m.B = pB_; m.C = pC_; m.H = pH_; m.Cq = pCq_; m.F = pF_;
  m.x = x; m.b = b; m.out = r; m.Cpart = Cpart_.p;
  tmpd_.resize(std::max<size_t>(tmpd_.count, size_t(max_chunks_) * std::max(1, d_.nr)));
  timed(kResidual, [&] { hip_check(launch_mv(m, Csum_.p, tmpd_.p, max_chunks_, st_), "mv"); });
}

template <typename TA, typename TW>
void Engine<TA, TW>::residual_multi_async(double const* x, int nrhs, int64_t ldx, double const* b, int64_t ldb,
                                          double* r, int64_t ldr) {
  if (nrhs <= 1) {
    if (nrhs == 1) residual_async(x, b, r);
    return;
  }
  if (!pD_) throw HipError("residual: no operator loaded or bound");
  const size_t nr = size_t(std::max(1, d_.nr));
  if (Cpart_.count < size_t(nrhs) * (d_.N + 1) * nr) Cpart_.resize(size_t(nrhs) * (d_.N + 1) * nr);
  if (Csum_.count < size_t(nrhs) * nr) Csum_.resize(size_t(nrhs) * nr);
  MvArgs<TA> m{};
  m.N = d_.N; m.n = d_.n; m.qr = d_.qr; m.qx = d_.qx; m.nr = d_.nr; m.nx = d_.nx;
  m.D = pD_; m.E = pE_; m.B = pB_; m.C = pC_; m.H = pH_; m.Cq = pCq_; m.F = pF_;
  m.x = x; m.b = b; m.out = r; m.Cpart = Cpart_.p;
  m.nrhs = nrhs; m.ldx = ldx; m.ldb = ldb; m.ldo = ldr;
  tmpd_.resize(std::max<size_t>(tmpd_.count, size_t(max_chunks_) * std::max(1, d_.nr)));
  timed(kResidual, [&] { hip_check(launch_mv(m, Csum_.p, tmpd_.p, max_chunks_, st_), "mv"); });
}

template <typename TA, typename TW>
void Engine<TA, TW>::mv_async(double const* x, double* y) {
  residual_async(x, nullptr, y);
}

template <typename TA, typename TW>
void Engine<TA, TW>::abs_bound_async(double const* x, double const* b, double* out) {
  if (!pD_) throw HipError("abs_bound: no operator loaded or bound");
  MvArgs<TA> m{};
  m.N = d_.N; m.n = d_.n; m.qr = d_.qr; m.qx = d_.qx; m.nr = d_.nr; m.nx = d_.nx;
  m.D = pD_; m.E = pE_; m.B = pB_; m.C = pC_; m.H = pH_; m.Cq = pCq_; m.F = pF_;
  m.x = x; m.b = b; m.out = out; m.Cpart = Cpart_.p;
  tmpd_.resize(std::max<size_t>(tmpd_.count, size_t(max_chunks_) * std::max(1, d_.nr)));
  timed(kResidual, [&] { hip_check(launch_mv(m, Csum_.p, tmpd_.p, max_chunks_, st_, true), "abs mv"); });
}

template <typename TA, typename TW>
void Engine<TA, TW>::max_ratio_async(double const* r, double const* den, int64_t len, double* out) {
  tmpd_.resize(std::max<size_t>(tmpd_.count, size_t(max_chunks_)));
  hip_check(launch_max_ratio(r, den, len, tmpd_.p, max_chunks_, out, st_), "max ratio");
}

template <typename TA, typename TW>
void Engine<TA, TW>::nrm2_async(double const* v, int64_t len, double* out) {
  tmpd_.resize(std::max<size_t>(tmpd_.count, size_t(max_chunks_)));
  hip_check(launch_nrm2(v, len, tmpd_.p, max_chunks_, out, st_), "nrm2");
}

template class Engine<double, double>;
template class Engine<float, float>;
template class Engine<double, float>;
template void Engine<double, double>::solve_async<double>(double*, int, int64_t);
template void Engine<float, float>::solve_async<float>(float*, int, int64_t);
template void Engine<double, float>::solve_async<double>(double*, int, int64_t);

}  // namespace bcr
