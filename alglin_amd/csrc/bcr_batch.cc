// Batched cyclic reduction (see bcr_batch.hh).
#include "bcr_batch.hh"

#include "bcr_mlp.hh"
#include "bcr_refine.hh"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>

namespace bcr {
namespace {

// the matrix-core pair of the 16 < n <= 32 and 32 < n <= 48 LU classes
// (FP64 factors: n <= 48; FP32 factors also up to 64): the default (variant
// 0) and variant 5; variant 4 keeps the rank-1 kernels (two-wave split,
// column-split)
template <typename TS, typename TW, typename L>
bool try_mfma(int variant, BatchTable<FactorArgs<TS, TW>> const& t, L const& l, hipStream_t st, hipError_t* e) {
  return (variant == 0 || variant == 5) && launch_pair_lu_mfma_batched(t, l.nblocks, l.nmax, l.nxmax, l.nrmax, st, e);
}

// host staging of the launch tables; uploaded once as one device blob
struct Blob {
  std::vector<char> h;
  size_t add(const void* p, size_t bytes) {
    const size_t off = (h.size() + 255) & ~size_t(255);
    h.resize(off + std::max<size_t>(bytes, 1));
    if (bytes) std::memcpy(h.data() + off, p, bytes);
    return off;
  }
  template <typename T> size_t add(std::vector<T> const& v) { return add(v.data(), v.size() * sizeof(T)); }
  void upload(DevBuf<char>& d, hipStream_t st) {
    d.resize(std::max<size_t>(h.size(), 1));
    hip_check(hipMemcpyAsync(d.p, h.data(), h.size(), hipMemcpyHostToDevice, st), "table upload");
    hip_check(hipStreamSynchronize(st), "table upload sync");
  }
};

int levels_of(int N) {
  int k = 0;
  for (int L = 1; L < N; L *= 2) ++k;
  return k;
}

// factor kernel class of a system: 0..2 register kernels (n <= 4, 8, 16),
// 3 two-wave kernel (16 < n <= 32), 4 LDS-panel LU, 5 QR, 6 QRP
int factor_class(SysDims const& d) {
  if (d.method == kQR) return 5;
  if (d.method == kQRP) return 6;
  if (d.n <= 4 && d.nx <= 4 && d.nr <= 4) return 0;
  if (d.n <= 8 && d.nx <= 8 && d.nr <= 8) return 1;
  if (d.n <= 16 && d.nx <= 8 && d.nr <= 8) return 2;
  if (d.n <= 32 && d.nx <= 8 && d.nr <= 8) return 3;
  return 4;
}
constexpr int kFastNmax[3] = {4, 8, 16};

// sweep kernel class: 0..3 wave kernels for LU with n <= 4, 8, 16, 32;
// 4 generic LU, 5 generic QR, 6 generic QRP
int sweep_class(SysDims const& d) {
  if (d.method == kQR) return 5;
  if (d.method == kQRP) return 6;
  if (d.n <= 4) return 0;
  if (d.n <= 8) return 1;
  if (d.n <= 16) return 2;
  if (d.n <= 32) return 3;
  return 4;
}
constexpr int kSweepNmax[4] = {4, 8, 16, 32};

}  // namespace

template <typename TA, typename TW>
BatchEngine<TA, TW>::BatchEngine() {
  hip_check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipStreamCreateWithFlags(&st2_, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreateWithFlags(&join_, hipEventDisableTiming), "hipEventCreate");
  int dev = 0, lds = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
    lds_limit_ = size_t(lds);
}

template <typename TA, typename TW>
BatchEngine<TA, TW>::~BatchEngine() {
  if (fork_) (void)hipEventDestroy(fork_);
  if (join_) (void)hipEventDestroy(join_);
  if (st2_) (void)hipStreamDestroy(st2_);
  if (st_) (void)hipStreamDestroy(st_);
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::fork() {
  hip_check(hipEventRecord(fork_, st_), "fork");
  hip_check(hipStreamWaitEvent(st2_, fork_, 0), "fork");
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::join() {
  hip_check(hipEventRecord(join_, st2_), "join");
  hip_check(hipStreamWaitEvent(st_, join_, 0), "join");
}

// the launch of a system's class takes a whole panel in LDS only when no
// other kernel covers the shape: LU with n <= 32 (wave kernels), LU up to
// n = 64 through the column-split or register-tiled kernels; QR and QRP
// through the chunked QR kernel (T and a column chunk in LDS)
template <typename TA, typename TW>
bool BatchEngine<TA, TW>::panel_fits(SysDims const& d, size_t lds_limit) {
  const int c = factor_class(d);
  if (c < 4) return true;
  if (c == 4 && (pair_lu_cols_fits(d.n, d.nr, d.nx) || pair_lu_tiled_fits(d.n, d.nr, d.nx))) return true;
  int cw = 0;
  size_t bytes = 0;
  if (d.method != kLU && pair_qr_plan(d.n, d.nr, d.nx, sizeof(TW), lds_limit, &cw, &bytes)) return true;
  return size_t(2 * d.n + d.nr) * (3 * d.n + d.nx) * sizeof(TW) <= lds_limit;
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::plan(std::vector<SysDims> const& dims) {
  if (dims.empty()) throw HipError("batch plan: no systems");
  for (size_t s = 0; s < dims.size(); ++s) {
    SysDims const& d = dims[s];
    const std::string id = "batch plan: system " + std::to_string(s) + ": ";
    if (d.N < 1 || d.n < 1 || d.qr < 0 || d.qx < 0 || d.nr < 0 || d.nx < 0) throw HipError(id + "bad dimensions");
    if (d.n > 64 || d.nr > 16) throw HipError(id + "block size n <= 64 and nr <= 16 are supported");
    if (d.Nr() != d.Nc()) throw HipError(id + "the last-block LU needs qr + nr == qx + nx");
    if (d.method != kLU && d.method != kQR && d.method != kQRP) throw HipError(id + "method must be LU, QR or QRP");
    if (!panel_fits(d, lds_limit_)) throw HipError(id + "pair panel does not fit in LDS");
  }
  dims_ = dims;
  const int ns = nsys();
  for (auto& v : off_) v.assign(ns + 1, 0);
  for (auto* v : {&offT_, &offPerm_, &offCpend_, &offFpart_, &offHm_, &offPiv_, &offXpart_, &offCpart_, &offCsum_,
                  &offPst_})
    v->assign(ns + 1, 0);
  maxN_ = Ncmax_ = Nrmax_ = 0;
  square_ = true;
  for (int s = 0; s < ns; ++s) {
    SysDims const& d = dims_[s];
    const int64_t N = d.N, n = d.n, nr = d.nr, nx = d.nx;
    off_[kbDE][s + 1] = off_[kbDE][s] + N * n * n;
    off_[kbB][s + 1] = off_[kbB][s] + N * n * nx;
    off_[kbC][s + 1] = off_[kbC][s] + (N + 1) * nr * n;
    off_[kbH][s + 1] = off_[kbH][s] + (n + d.qr) * d.Nc();
    off_[kbCq][s + 1] = off_[kbCq][s] + nr * d.qx;
    off_[kbF][s + 1] = off_[kbF][s] + nr * nx;
    off_[kbX][s + 1] = off_[kbX][s] + std::max(d.nrows(), d.ncols());
    offT_[s + 1] = offT_[s] + N * int64_t(tsize(d.n));
    offPerm_[s + 1] = offPerm_[s] + N * n;
    offCpend_[s + 1] = offCpend_[s] + 2 * (N + 1) * nr * n;
    offFpart_[s + 1] = offFpart_[s] + N * nr * nx;
    offHm_[s + 1] = offHm_[s] + int64_t(d.Nr()) * d.Nc();
    offPiv_[s + 1] = offPiv_[s] + std::max(d.Nr(), d.Nc());
    offXpart_[s + 1] = offXpart_[s] + N * nr;
    offCpart_[s + 1] = offCpart_[s] + (N + 1) * nr;
    offCsum_[s + 1] = offCsum_[s] + nr;
    offPst_[s + 1] = offPst_[s] + 3 * std::max<int64_t>(N - 1, 0);
    maxN_ = std::max(maxN_, d.N);
    Ncmax_ = std::max(Ncmax_, d.Nc());
    Nrmax_ = std::max(Nrmax_, d.Nr());
    square_ = square_ && d.Nr() == d.Nc();
  }
  for (int k = 0; k < kbKinds; ++k) tot_[k] = off_[k][ns];
  levels_ = levels_of(maxN_);
  auto one = [](int64_t v) { return size_t(std::max<int64_t>(v, 1)); };
  wD_.resize(one(tot_[kbDE]));
  wE_.resize(one(tot_[kbDE]));
  wB_.resize(one(tot_[kbB]));
  wC_.resize(one(tot_[kbC]));
  Tm_.resize(one(offT_[ns]));
  perm_.resize(one(offPerm_[ns]));
  Cpend_.resize(one(offCpend_[ns]));
  Fpart_.resize(one(offFpart_[ns]));
  Feff_.resize(one(tot_[kbF]));
  Hm_.resize(one(offHm_[ns]));
  ipiv_.resize(one(offPiv_[ns]));
  status_.resize(size_t(ns));
  skip_.resize(size_t(ns));
  Xpart_.resize(one(offXpart_[ns]));
  Cpart_.resize(one(offCpart_[ns]));
  Csum_.resize(one(offCsum_[ns]));
  pstat_.resize(one(offPst_[ns]));
  stats_ok_ = false;
  factor_bytes_ = sizeof(TW) * (2 * tot_[kbDE] + tot_[kbB] + tot_[kbC] + offT_[ns] + offCpend_[ns] + offFpart_[ns]) +
                  4 * offPerm_[ns] + 8 * offHm_[ns];
  active_.resize(ns);
  for (int s = 0; s < ns; ++s) active_[s] = s;
  invalidate();
  pD_ = pE_ = pB_ = pC_ = pH_ = pCq_ = pF_ = nullptr;
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::invalidate() {
  tables_ok_ = false;
  solve_cache_.clear();
  mv_cache_.clear();
  norm_cache_.clear();
  add_cache_.clear();
  feat_ok_ = false;
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::set_active(std::vector<char> const& mask) {
  widen();
  active_.clear();
  for (int s = 0; s < nsys(); ++s)
    if (s < int(mask.size()) && mask[s]) active_.push_back(s);
  invalidate();
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::swap_stash() {
  stash_.active.swap(active_);
  stash_.fac.swap(fac_);
  std::swap(stash_.prep, prep_);
  std::swap(stash_.fred, fred_);
  std::swap(stash_.asm_, asm_);
  std::swap(stash_.lu, lu_);
  std::swap(stash_.stptr, stptr_);
  std::swap(stash_.fblob.p, fblob_.p);
  std::swap(stash_.fblob.count, fblob_.count);
  std::swap(stash_.tables_ok, tables_ok_);
  stash_.solve.swap(solve_cache_);
  stash_.mv.swap(mv_cache_);
  stash_.norm.swap(norm_cache_);
  stash_.add.swap(add_cache_);
  feat_ok_ = false;   // built for the active set of its call
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::narrow(std::vector<char> const& mask) {
  widen();
  swap_stash();   // the current tables aside; the live ones are now empty
  stash_.has = true;
  active_.clear();
  for (int s = 0; s < nsys(); ++s)
    if (s < int(mask.size()) && mask[s]) active_.push_back(s);
  tables_ok_ = false;
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::widen() {
  if (!stash_.has) return;
  swap_stash();   // the narrowed tables into the stash, freed below
  stash_.has = false;
  stash_.active.clear();
  stash_.fac.clear();
  stash_.prep = nullptr;
  stash_.fred = nullptr;
  stash_.asm_ = nullptr;
  stash_.lu = nullptr;
  stash_.stptr = nullptr;
  stash_.fblob.release();
  stash_.tables_ok = false;
  stash_.solve.clear();
  stash_.mv.clear();
  stash_.norm.clear();
  stash_.add.clear();
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::set_class_kernel(int cls, int kernel, int waves) {
  if (cls < 0 || cls >= kFacClasses) throw HipError("set_class_kernel: bad class");
  if (kernel < -1 || kernel > kKernTiled) throw HipError("set_class_kernel: bad kernel");
  cls_kernel_[cls] = kernel;
  cls_waves_[cls] = waves;
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::set_method(int s, int method) {
  if (s < 0 || s >= nsys()) throw HipError("set_method: bad system index");
  if (method != kLU && method != kQR && method != kQRP) throw HipError("set_method: method must be LU, QR or QRP");
  SysDims d = dims_[s];
  d.method = method;
  if (!panel_fits(d, lds_limit_)) throw HipError("set_method: pair panel does not fit in LDS");
  if (dims_[s].method != method) {
    dims_[s].method = method;
    invalidate();
  }
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::bind_device(TA const* D, TA const* E, TA const* B, TA const* C, TA const* H, TA const* Cq,
                                      TA const* F) {
  pD_ = D; pE_ = E; pB_ = B; pC_ = C; pH_ = H; pCq_ = Cq; pF_ = F;
  invalidate();
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::generate(uint64_t seed, double shift, std::vector<int32_t> const& sing_k, double eps) {
  if constexpr (std::is_same_v<TA, double>) {
    if (!pD_) throw HipError("batch generate: no operator bound");
    std::vector<GenArgs> g(nsys());
    for (int s = 0; s < nsys(); ++s) {
      SysDims const& d = dims_[s];
      GenArgs& a = g[s];
      a.sys = s; a.N = d.N; a.n = d.n; a.qr = d.qr; a.qx = d.qx; a.nr = d.nr; a.nx = d.nx; a.shift = shift;
      a.sing_k = s < int(sing_k.size()) && sing_k[s] >= 1 && sing_k[s] < d.N ? sing_k[s] : -1;
      a.eps = eps;
      a.D = const_cast<double*>(pD_) + off_[kbDE][s];
      a.E = const_cast<double*>(pE_) + off_[kbDE][s];
      a.B = const_cast<double*>(pB_) + off_[kbB][s];
      a.C = const_cast<double*>(pC_) + off_[kbC][s];
      a.H = const_cast<double*>(pH_) + off_[kbH][s];
      a.Cq = const_cast<double*>(pCq_) + off_[kbCq][s];
      a.F = const_cast<double*>(pF_) + off_[kbF][s];
    }
    Blob bb;
    const size_t o = bb.add(g);
    DevBuf<char> blob;
    bb.upload(blob, st_);
    hip_check(launch_gen_batched(reinterpret_cast<const GenArgs*>(blob.p + o), nsys(), seed, st_), "generate");
    hip_check(hipStreamSynchronize(st_), "generate sync");
  } else {
    throw HipError("batch generate: needs a double operator");
  }
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::build_factor_tables() {
  if (!pD_) throw HipError("batch factorize: no operator bound");
  const int ns = nsys();
  Blob bb;
  struct Pending { size_t args, entry, first; Launch l; };
  std::vector<std::vector<Pending>> pend(levels_);
  int lev = 0;
  for (int L = 1; L < maxN_; L *= 2, ++lev) {
    int lg = 0;
    for (int t = L; t > 1; t >>= 1) ++lg;
    for (int cls = 0; cls < kFacClasses; ++cls) {
      std::vector<FactorArgs<TA, TW>> a0;
      std::vector<FactorArgs<TW, TW>> a1;
      std::vector<int32_t> entry, first{0};
      int nmax = 0, nxmax = 0, nrmax = 0;
      std::vector<int> members;
      for (int s : active_) {
        SysDims const& d = dims_[s];
        if (d.N <= L || factor_class(d) != cls) continue;
        members.push_back(s);
        nmax = std::max(nmax, d.n);
        nxmax = std::max(nxmax, d.nx);
        nrmax = std::max(nrmax, d.nr);
      }
      if (members.empty()) continue;
      int ppb = 1;
      if (cls <= 2) ppb = pair_lu_fast_ppb(kFastNmax[cls], nxmax, nrmax);
      double lbytes = 0, lflops = 0;
      for (int s : members) {
        SysDims const& d = dims_[s];
        {
          // per pair: reads E_jp, D_jp, D_j, E_j, B_jp, B_j, C_j, C_jp; writes T, D, E (both rows),
          // B (both), C_j, C_jp, pending C_jpp, the F partial and ipiv
          const double n = d.n, nx = d.nx, nr = d.nr, np = (d.N + L - 1) / (2 * L);
          const double rd = 4 * n * n + 2 * n * nx + 2 * nr * n;
          const double wr = 6 * n * n + 2 * n * nx + 3 * nr * n + nr * nx;
          lbytes += np * (rd * (L == 1 ? sizeof(TA) : sizeof(TW)) + wr * sizeof(TW) + 4 * n);
          lflops += np * ((23.0 / 3.0) * n * n * n + 3 * n * n * nx + 5 * nr * n * n + 2 * nr * nx * n);
        }
        FactorArgs<TW, TW> a{};
        a.N = d.N; a.n = d.n; a.nr = d.nr; a.nx = d.nx; a.L = L;
        a.npairs = (d.N + L - 1) / (2 * L);
        a.slot_cur = lg & 1; a.first = L == 1 ? 1 : 0;
        a.D = wD_.p + off_[kbDE][s]; a.E = wE_.p + off_[kbDE][s];
        a.B = wB_.p + off_[kbB][s]; a.C = wC_.p + off_[kbC][s];
        a.Tm = Tm_.p + offT_[s]; a.perm = perm_.p + offPerm_[s];
        a.Cpend = Cpend_.p + offCpend_[s]; a.Fpart = Fpart_.p + offFpart_[s];
        a.status = status_.p + s;
        // FP32 LU panels: the selector's per-panel statistics, formed by the
        // pair kernels where they hold the panel anyway
        a.pstat = sizeof(TW) == 4 && cls <= 4 ? pstat_.p + offPst_[s] : nullptr;
        if (L == 1) {
          FactorArgs<TA, TW> b{};
          b.N = a.N; b.n = a.n; b.nr = a.nr; b.nx = a.nx; b.L = a.L; b.npairs = a.npairs;
          b.slot_cur = a.slot_cur; b.first = 1;
          b.sD = pD_ + off_[kbDE][s]; b.sE = pE_ + off_[kbDE][s]; b.sB = pB_ + off_[kbB][s]; b.sC = pC_ + off_[kbC][s];
          b.D = a.D; b.E = a.E; b.B = a.B; b.C = a.C; b.Tm = a.Tm; b.perm = a.perm; b.Cpend = a.Cpend;
          b.Fpart = a.Fpart; b.status = a.status; b.pstat = a.pstat;
          a0.push_back(b);
        } else {
          a.sD = a.D; a.sE = a.E; a.sB = a.B; a.sC = a.C;
          a1.push_back(a);
        }
        const int blocks = (a.npairs + ppb - 1) / ppb;
        const int e = int(first.size()) - 1;
        for (int b = 0; b < blocks; ++b) entry.push_back(e);
        first.push_back(first.back() + blocks);
      }
      Pending p{};
      p.args = L == 1 ? bb.add(a0) : bb.add(a1);
      p.entry = bb.add(entry);
      p.first = bb.add(first);
      p.l.nblocks = first.back();
      p.l.cls = cls;
      p.l.method = cls == 5 ? kQR : cls == 6 ? kQRP : kLU;
      p.l.nmax = cls <= 2 ? kFastNmax[cls] : nmax;
      p.l.nxmax = nxmax;
      p.l.nrmax = nrmax;
      p.l.bytes = lbytes;
      p.l.flops = lflops;
      pend[lev].push_back(p);
    }
  }
  // per-system records (active systems, in order)
  const int na = int(active_.size());
  std::vector<PrepArgs<TA, TW>> prep(na);
  std::vector<ReduceArgs<TW, double>> fred(na);
  std::vector<AssembleArgs<TW, TA>> as(na);
  std::vector<LastSolveArgs<double>> lu(na);
  std::vector<int32_t*> stp(na);
  for (int i = 0; i < na; ++i) {
    const int s = active_[i];
    SysDims const& d = dims_[s];
    PrepArgs<TA, TW>& p = prep[i];
    p.N = d.N; p.n = d.n; p.nr = d.nr; p.nx = d.nx;
    p.D = pD_ + off_[kbDE][s]; p.E = pE_ + off_[kbDE][s]; p.B = pB_ + off_[kbB][s]; p.C = pC_ + off_[kbC][s];
    p.wD = wD_.p + off_[kbDE][s]; p.wE = wE_.p + off_[kbDE][s]; p.wB = wB_.p + off_[kbB][s];
    p.wC = wC_.p + off_[kbC][s];
    ReduceArgs<TW, double>& f = fred[i];
    f.part = Fpart_.p + offFpart_[s]; f.first = 1; f.count = d.N - 1; f.m = d.nr * d.nx; f.mrow = std::max(1, f.m);
    f.ldd = f.mrow; f.dst = Feff_.p + off_[kbF][s];
    AssembleArgs<TW, TA>& a = as[i];
    a.N = d.N; a.n = d.n; a.qr = d.qr; a.qx = d.qx; a.nr = d.nr; a.nx = d.nx; a.Nr = d.Nr(); a.Nc = d.Nc();
    a.D0 = wD_.p + off_[kbDE][s]; a.E0 = wE_.p + off_[kbDE][s]; a.B0 = wB_.p + off_[kbB][s];
    a.C = wC_.p + off_[kbC][s]; a.Cpend = Cpend_.p + offCpend_[s];
    a.H = pH_ + off_[kbH][s]; a.Cq = pCq_ + off_[kbCq][s]; a.Feff = Feff_.p + off_[kbF][s];
    a.Hm = Hm_.p + offHm_[s];
    LastSolveArgs<double>& l = lu[i];
    l = LastSolveArgs<double>{};
    l.Nc = d.Nc(); l.Nr = d.Nr(); l.Hm = Hm_.p + offHm_[s]; l.P = ipiv_.p + offPiv_[s];
    stp[i] = status_.p + s;
  }
  const size_t oprep = bb.add(prep), ofred = bb.add(fred), oasm = bb.add(as), olu = bb.add(lu), ost = bb.add(stp);
  bb.upload(fblob_, st_);
  char* base = fblob_.p;
  fac_.assign(levels_, {});
  for (int k = 0; k < levels_; ++k)
    for (auto& p : pend[k]) {
      Launch l = p.l;
      l.args = base + p.args;
      l.entry = reinterpret_cast<const int32_t*>(base + p.entry);
      l.first = reinterpret_cast<const int32_t*>(base + p.first);
      fac_[k].push_back(l);
    }
  prep_ = reinterpret_cast<const PrepArgs<TA, TW>*>(base + oprep);
  fred_ = reinterpret_cast<const ReduceArgs<TW, double>*>(base + ofred);
  asm_ = reinterpret_cast<const AssembleArgs<TW, TA>*>(base + oasm);
  lu_ = reinterpret_cast<const LastSolveArgs<double>*>(base + olu);
  stptr_ = reinterpret_cast<int32_t* const*>(base + ost);
  tables_ok_ = true;
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::factorize_impl(std::vector<hipEvent_t>* ev) {
  if (!tables_ok_) build_factor_tables();
  const int ns = int(active_.size());
  if (ns == 0) return;
  hip_check(launch_convert<int32_t, int32_t>(nullptr, status_.p, nsys(), st_), "status");
  hip_check(launch_convert<int32_t, int32_t>(nullptr, skip_.p, nsys(), st_), "skip flags");
  hip_check(launch_convert<TW, TW>(nullptr, Cpend_.p, int64_t(Cpend_.count), st_), "zero pending");
  hip_check(launch_prep_batched<TA, TW>(prep_, ns, st_), "prep");
  // the n > 32 class on st2_ from level fork_level_ on, unless every launch
  // is timed (then one stream)
  const bool conc = ev == nullptr && streams_ > 1 && fork_level_ < levels_;
  bool forked = false;
  size_t li = 0;
  // the launches' kernels all write the panels' statistics (FP32 LU: the
  // register kernels n <= 16 and the matrix-core pair)
  bool stats = sizeof(TW) == 4;
  for (int k = 0; k < levels_; ++k)
    for (Launch const& l : fac_[k]) {
      if (conc && !forked && k >= fork_level_) {
        fork();
        forked = true;
      }
      hipStream_t s = class_stream(l.cls, forked);
      if (ev) hip_check(hipEventRecord((*ev)[2 * li], s), "event");
      hipError_t e = hipSuccess;
      auto run = [&](auto const* args) {
        using A = std::remove_cv_t<std::remove_pointer_t<decltype(args)>>;
        BatchTable<A> t{args, l.entry, l.first};
        bool ok = true;
        // the class's kernel: a tuned choice (set_class_kernel) or the
        // default of pair_variant_
        const int ck = cls_kernel_[l.cls];
        const int v = ck == kKernMfma ? 5 : ck >= 0 ? 4 : pair_variant_;
        bool wrote = false;   // the kernel writes the panels' statistics
        if (l.cls <= 2) {
          ok = launch_pair_lu_fast_batched(t, l.nblocks, l.nmax, l.nxmax, l.nrmax, s, &e);
          wrote = true;
        } else if ((l.cls == 3 || l.cls == 4) && try_mfma(v, t, l, s, &e)) {
          wrote = true;
        } else if (l.cls == 3) ok = launch_pair_lu_split_batched(t, l.nblocks, 32, l.nxmax, l.nrmax, s, &e);
        else if (l.cls == 4 && ck != kKernTiled && pair_lu_cols_fits(l.nmax, l.nrmax, l.nxmax))
          ok = launch_pair_lu_cols_batched(t, l.nblocks, l.nmax, l.nxmax, l.nrmax, s, &e, cls_waves_[l.cls]);
        else if (l.cls == 4 && pair_lu_tiled_fits(l.nmax, l.nrmax, l.nxmax))
          ok = launch_pair_lu_tiled_batched(t, l.nblocks, l.nmax, l.nxmax, l.nrmax, s, &e);
        else if (l.method != kLU &&
                 launch_pair_qr_batched(l.method, t, l.nblocks, l.nmax, l.nrmax, l.nxmax, lds_limit_, s, &e))
          ok = true;
        else {
          const size_t bytes = size_t(2 * l.nmax + l.nrmax) * (3 * l.nmax + l.nxmax) * sizeof(TW);
          e = launch_pair_generic_batched(l.method, t, l.nblocks, bytes, lds_limit_, s);
        }
        if (!ok) throw HipError("batch factorize: no kernel for a class");
        stats = stats && wrote;
      };
      if (k == 0) run(static_cast<const FactorArgs<TA, TW>*>(l.args));
      else run(static_cast<const FactorArgs<TW, TW>*>(l.args));
      hip_check(e, "batch pair elimination");
      if (ev) hip_check(hipEventRecord((*ev)[2 * li + 1], s), "event");
      ++li;
    }
  if (forked) join();
  stats_ok_ = stats;
  hip_check(launch_convert<TA, double>(pF_, Feff_.p, tot_[kbF], st_), "F");
  hip_check(launch_reduce_batched<TW, double>(fred_, ns, st_), "reduce F");
  hip_check(launch_assemble_last_batched<TW, TA>(asm_, ns, st_), "assemble last");
  hip_check(launch_lu_last_batched(lu_, stptr_, ns, Ncmax_, lds_limit_, st_), "lu last");
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::factorize_async() {
  factorize_impl(nullptr);
}

template <typename TA, typename TW>
std::vector<typename BatchEngine<TA, TW>::LaunchTime> BatchEngine<TA, TW>::time_factorize() {
  if (!tables_ok_) build_factor_tables();
  std::vector<LaunchTime> out;
  for (int k = 0; k < levels_; ++k)
    for (Launch const& l : fac_[k]) out.push_back(LaunchTime{k, l.cls, l.nblocks, 0.0, l.bytes, l.flops});
  std::vector<hipEvent_t> ev(2 * out.size(), nullptr);
  for (auto& e : ev) hip_check(hipEventCreate(&e), "event create");
  try {
    factorize_impl(&ev);
    hip_check(hipStreamSynchronize(st_), "timed factorize");
    for (size_t i = 0; i < out.size(); ++i) {
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]), "event time");
      out[i].ms = ms;
    }
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return out;
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::status(int32_t* host_out) {
  hip_check(hipMemcpyAsync(host_out, status_.p, sizeof(int32_t) * nsys(), hipMemcpyDeviceToHost, st_), "status");
  hip_check(hipStreamSynchronize(st_), "status sync");
}

template <typename TA, typename TW>
typename BatchEngine<TA, TW>::SolveTables& BatchEngine<TA, TW>::solve_tables(double* X) {
  auto it = solve_cache_.find(X);
  if (it != solve_cache_.end()) return it->second;
  if (!tables_ok_) build_factor_tables();
  SolveTables& T = solve_cache_[X];
  const int ns = nsys();
  Blob bb;
  struct Pending { size_t args, entry, first; Launch l; };
  std::vector<std::vector<Pending>> pend(levels_);
  int lev = 0;
  for (int L = 1; L < maxN_; L *= 2, ++lev) {
    for (int cls = 0; cls < 7; ++cls) {
      std::vector<SolveArgs<TW, double>> args;
      std::vector<int32_t> entry, first{0};
      int nmax = 0, nxmax = 0;
      for (int s : active_) {
        SysDims const& d = dims_[s];
        if (d.N <= L || sweep_class(d) != cls) continue;
        SolveArgs<TW, double> a{};
        a.N = d.N; a.n = d.n; a.nr = d.nr; a.nx = d.nx; a.qr = d.qr; a.qx = d.qx;
        a.L = L; a.npairs = (d.N + L - 1) / (2 * L); a.nrhs = 1;
        a.ldx = std::max(d.nrows(), d.ncols()); a.x = X + off_[kbX][s];
        a.D = wD_.p + off_[kbDE][s]; a.E = wE_.p + off_[kbDE][s]; a.B = wB_.p + off_[kbB][s];
        a.C = wC_.p + off_[kbC][s]; a.Tm = Tm_.p + offT_[s]; a.perm = perm_.p + offPerm_[s];
        a.Xpart = Xpart_.p + offXpart_[s];
        a.skip = skip_.p + s;
        args.push_back(a);
        nmax = std::max(nmax, d.n);
        nxmax = std::max(nxmax, d.nx);
        const int ppb = cls <= 3 ? sweep_lu_fast_pairs_per_block(kSweepNmax[cls]) : (cls == 4 && d.n <= 64 ? 4 : 1);
        const int blocks = (a.npairs + ppb - 1) / ppb;
        const int e = int(first.size()) - 1;
        for (int b = 0; b < blocks; ++b) entry.push_back(e);
        first.push_back(first.back() + blocks);
      }
      if (args.empty()) continue;
      Pending p{};
      p.args = bb.add(args);
      p.entry = bb.add(entry);
      p.first = bb.add(first);
      p.l.nblocks = first.back();
      p.l.cls = cls;
      p.l.method = cls == 5 ? kQR : cls == 6 ? kQRP : kLU;
      p.l.nmax = cls <= 3 ? kSweepNmax[cls] : nmax;
      p.l.nxmax = nxmax;
      pend[lev].push_back(p);
    }
  }
  const int na = int(active_.size());
  std::vector<ReduceArgs<double, double>> xb(na);
  std::vector<LastSolveArgs<double>> last(na);
  for (int i = 0; i < na; ++i) {
    const int s = active_[i];
    SysDims const& d = dims_[s];
    ReduceArgs<double, double>& r = xb[i];
    r.part = Xpart_.p + offXpart_[s]; r.first = 1; r.count = d.N - 1; r.m = d.nr; r.mrow = std::max(1, d.nr);
    r.ldd = std::max(d.nrows(), d.ncols());
    r.dst = X + off_[kbX][s] + int64_t(d.N + 1) * d.n + d.qr;
    r.skip = skip_.p + s;
    LastSolveArgs<double>& l = last[i];
    l.N = d.N; l.n = d.n; l.qr = d.qr; l.qx = d.qx; l.nr = d.nr; l.nx = d.nx; l.Nr = d.Nr(); l.Nc = d.Nc();
    l.method = kLastLU; l.nrhs = 1; l.Hm = Hm_.p + offHm_[s]; l.P = ipiv_.p + offPiv_[s]; l.Q = nullptr;
    l.tau = nullptr; l.x = X + off_[kbX][s]; l.ldx = std::max(d.nrows(), d.ncols());
    l.skip = skip_.p + s;
  }
  const size_t oxb = bb.add(xb), olast = bb.add(last);
  bb.upload(T.blob, st_);
  char* base = T.blob.p;
  T.fwd.assign(levels_, {});
  for (int k = 0; k < levels_; ++k)
    for (auto& p : pend[k]) {
      Launch l = p.l;
      l.args = base + p.args;
      l.entry = reinterpret_cast<const int32_t*>(base + p.entry);
      l.first = reinterpret_cast<const int32_t*>(base + p.first);
      T.fwd[k].push_back(l);
    }
  T.xb = reinterpret_cast<const ReduceArgs<double, double>*>(base + oxb);
  T.last = reinterpret_cast<const LastSolveArgs<double>*>(base + olast);
  return T;
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::solve_async(double* X) {
  SolveTables& T = solve_tables(X);
  const int ns = int(active_.size());
  if (ns == 0) return;
  // the n > 32 class's sweeps on st2_, concurrent with the other classes'
  const bool conc = streams_ > 1 && fork_level_ < levels_;
  bool forked = false;
  auto sweep = [&](Launch const& l, bool fwd) {
    BatchTable<SolveArgs<TW, double>> t{static_cast<const SolveArgs<TW, double>*>(l.args), l.entry, l.first};
    hipStream_t s = class_stream(l.cls, forked);
    hipError_t e = hipSuccess;
    if (l.cls <= 2 || (l.cls == 3 && fwd)) {
      if (!launch_sweep_lu_fast_batched(fwd, t, l.nblocks, l.nmax, 1, s, &e))
        throw HipError("batch solve: no sweep kernel");
    } else if (l.cls <= 4) {   // the 16 < n <= 32 backward too: one wave per pair, columns streamed
      if (!launch_sweep_lu_stream_batched(fwd, t, l.nblocks, l.nmax, s, &e))
        throw HipError("batch solve: no sweep kernel");
    } else {
      e = launch_sweep_generic_batched(fwd, l.method, t, l.nblocks, l.nmax, l.nxmax, s);
    }
    hip_check(e, fwd ? "batch forward" : "batch backward");
  };
  // forward: the levels before fork_level_ on one stream, then two; backward
  // (levels in reverse): two streams down to fork_level_, then one
  for (int k = 0; k < levels_; ++k) {
    if (conc && !forked && k >= fork_level_) {
      fork();
      forked = true;
    }
    for (Launch const& l : T.fwd[k]) sweep(l, true);
  }
  if (forked) join();
  forked = false;
  hip_check(launch_reduce_batched<double, double>(T.xb, ns, st_), "reduce xb");
  // the last blocks are LU-factored with one right-hand side here: one
  // wavefront per system when they are square and Nc <= 256
  hipError_t le = hipSuccess;
  if (!(square_ && launch_last_solve_lu_wave_batched<double>(T.last, ns, Ncmax_, st_, &le)))
    le = launch_last_solve_batched<double>(T.last, ns, Nrmax_, Ncmax_, lds_limit_, st_);
  hip_check(le, "last solve");
  if (conc && fork_level_ < levels_) {
    fork();
    forked = true;
  }
  for (int k = levels_ - 1; k >= 0; --k) {
    if (forked && k < fork_level_) {
      join();
      forked = false;
    }
    for (Launch const& l : T.fwd[k]) sweep(l, false);
  }
  if (forked) join();
}

template <typename TA, typename TW>
typename BatchEngine<TA, TW>::MvTables& BatchEngine<TA, TW>::mv_tables(double const* X, double const* B, double* R) {
  std::vector<const void*> key{X, B, R};
  auto it = mv_cache_.find(key);
  if (it != mv_cache_.end()) return it->second;
  if (!pD_) throw HipError("batch residual: no operator bound");
  MvTables& T = mv_cache_[key];
  const int na = int(active_.size());
  std::vector<MvArgs<TA>> mv(na);
  std::vector<int32_t> entry, first{0};
  std::vector<ReduceArgs<double, double>> cs(na);
  std::vector<const double*> csp(na);
  for (int i = 0; i < na; ++i) {
    const int s = active_[i];
    SysDims const& d = dims_[s];
    MvArgs<TA>& m = mv[i];
    m.N = d.N; m.n = d.n; m.qr = d.qr; m.qx = d.qx; m.nr = d.nr; m.nx = d.nx;
    m.D = pD_ + off_[kbDE][s]; m.E = pE_ + off_[kbDE][s]; m.B = pB_ + off_[kbB][s]; m.C = pC_ + off_[kbC][s];
    m.H = pH_ + off_[kbH][s]; m.Cq = pCq_ + off_[kbCq][s]; m.F = pF_ + off_[kbF][s];
    m.x = X + off_[kbX][s]; m.b = B ? B + off_[kbX][s] : nullptr; m.out = R + off_[kbX][s];
    m.Cpart = Cpart_.p + offCpart_[s];
    m.skip = skip_.p + s;
    const int per_block = 4 * mv_rows_per_wave(d.n, d.nr);   // block rows per 256-thread workgroup
    const int blocks = (d.N + 1 + per_block - 1) / per_block;
    for (int b = 0; b < blocks; ++b) entry.push_back(i);
    first.push_back(first.back() + blocks);
    ReduceArgs<double, double>& r = cs[i];
    r.part = Cpart_.p + offCpart_[s]; r.first = 0; r.count = d.N + 1; r.m = d.nr; r.mrow = std::max(1, d.nr);
    r.ldd = r.mrow; r.dst = Csum_.p + offCsum_[s];
    r.skip = skip_.p + s;
    csp[i] = Csum_.p + offCsum_[s];
  }
  Blob bb;
  const size_t om = bb.add(mv), oe = bb.add(entry), of = bb.add(first), oc = bb.add(cs), op = bb.add(csp);
  bb.upload(T.blob, st_);
  char* base = T.blob.p;
  T.rows.args = base + om;
  T.rows.entry = reinterpret_cast<const int32_t*>(base + oe);
  T.rows.first = reinterpret_cast<const int32_t*>(base + of);
  T.rows.nblocks = first.back();
  T.csum = reinterpret_cast<const ReduceArgs<double, double>*>(base + oc);
  T.tail = reinterpret_cast<const MvArgs<TA>*>(base + om);
  T.csum_ptr = reinterpret_cast<const double* const*>(base + op);
  return T;
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::residual_async(double const* X, double const* B, double* R, bool absval) {
  MvTables& T = mv_tables(X, B, R);
  const int ns = int(active_.size());
  if (ns == 0) return;
  BatchTable<MvArgs<TA>> t{static_cast<const MvArgs<TA>*>(T.rows.args), T.rows.entry, T.rows.first};
  hip_check(launch_mv_rows_batched<TA>(t, T.rows.nblocks, st_, absval), "batch mv rows");
  hip_check(launch_convert<double, double>(nullptr, Csum_.p, int64_t(Csum_.count), st_), "zero Csum");
  hip_check(launch_reduce_batched<double, double>(T.csum, ns, st_), "reduce Csum");
  hip_check(launch_mv_tail_batched<TA>(T.tail, T.csum_ptr, ns, st_, absval), "batch mv tail");
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::backward_errors_async(double const* X, double const* B, double const* R, double* den,
                                                double* out) {
  residual_async(X, B, den, true);   // den = |A| |X| + |B|
  norms_async(den, R, nullptr);      // the per-system (den, R) table
  auto it = norm_cache_.find(std::vector<const void*>{den, R});
  if (!active_.empty())
    hip_check(launch_berr_batched(it->second.args, int(active_.size()), out, st_), "batch backward errors");
}

template <typename TA, typename TW>
const CopyArgs* BatchEngine<TA, TW>::add_table(double const* src, double* dst) {
  std::vector<const void*> key{src, dst};
  auto it = add_cache_.find(key);
  if (it != add_cache_.end()) return it->second.args;
  AddTables& T = add_cache_[key];
  std::vector<CopyArgs> c(active_.size());
  for (size_t i = 0; i < active_.size(); ++i) {
    const int s = active_[i];
    c[i].src = src + off_[kbX][s];
    c[i].dst = dst + off_[kbX][s];
    c[i].len = dims_[s].ncols();
    c[i].skip = skip_.p + s;
  }
  Blob bb;
  const size_t o = bb.add(c);
  bb.upload(T.blob, st_);
  T.args = reinterpret_cast<const CopyArgs*>(T.blob.p + o);
  return T.args;
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::add_active_async(double const* src, double* dst) {
  const CopyArgs* t = add_table(src, dst);
  if (!active_.empty()) hip_check(launch_add_batched(t, int(active_.size()), st_), "x += d");
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::prepare(double* X, double const* B, double* R) {
  if (!tables_ok_) build_factor_tables();
  solve_tables(X);
  solve_tables(R);
  mv_tables(X, B, R);
  norms_async(B, R, nullptr);
  add_table(R, X);
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::norms_async(double const* B, double const* R, double* out) {
  std::vector<const void*> key{B, R};
  auto it = norm_cache_.find(key);
  if (it == norm_cache_.end()) {
    NormTables& T = norm_cache_[key];
    std::vector<NormArgs> a(active_.size());
    for (size_t i = 0; i < active_.size(); ++i) {
      const int s = active_[i];
      a[i].b = B + off_[kbX][s];
      a[i].r = R + off_[kbX][s];
      a[i].len = dims_[s].nrows();
      a[i].sys = s;
      a[i].skip = skip_.p + s;
    }
    Blob bb;
    const size_t o = bb.add(a);
    bb.upload(T.blob, st_);
    T.args = reinterpret_cast<const NormArgs*>(T.blob.p + o);
    it = norm_cache_.find(key);
  }
  if (out && !active_.empty())
    hip_check(launch_norms_batched(it->second.args, int(active_.size()), out, st_), "batch norms");
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::features_async(double const* norms, float* out) {
  if (!tables_ok_) build_factor_tables();
  if (active_.empty()) return;
  if (!feat_ok_ || feat_norms_ != norms) {
    // panels j = 1 .. N-1 of the active systems, in order, four per workgroup
    std::vector<FeatArgs<TW>> f(active_.size());
    std::vector<int32_t> entry, first(active_.size());
    for (size_t i = 0; i < active_.size(); ++i) {
      const int s = active_[i];
      SysDims const& d = dims_[s];
      FeatArgs<TW>& a = f[i];
      a.N = d.N; a.n = d.n; a.qr = d.qr; a.nr = d.nr; a.Nc = d.Nc(); a.sys = s;
      a.Tm = Tm_.p + offT_[s]; a.Hm = Hm_.p + offHm_[s]; a.norms = norms;
      a.pstat = pstat_.p + offPst_[s];
      first[i] = int32_t(entry.size());
      for (int k = 0; k < (d.N - 1 + 3) / 4; ++k) entry.push_back(int32_t(i));
    }
    Blob bb;
    const size_t oa = bb.add(f), oe = bb.add(entry), of = bb.add(first);
    bb.upload(feat_blob_, st_);
    feat_args_ = reinterpret_cast<const FeatArgs<TW>*>(feat_blob_.p + oa);
    feat_tab_ = BatchTable<FeatArgs<TW>>{feat_args_, reinterpret_cast<const int32_t*>(feat_blob_.p + oe),
                                         reinterpret_cast<const int32_t*>(feat_blob_.p + of)};
    feat_wgs_ = int(entry.size());
    feat_nsys_ = int(f.size());
    feat_norms_ = norms;
    feat_ok_ = true;
  }
  // the pair kernels of the last factorization formed the statistics where
  // they held the panels; else a pass over the T panels does
  if (!stats_ok_) hip_check(launch_panel_stats_batched<TW>(feat_tab_, feat_wgs_, st_), "panel statistics");
  hip_check(launch_features_batched<TW>(feat_args_, feat_nsys_, out, st_), "features");
}

template <typename TA, typename TW>
void BatchEngine<TA, TW>::copy_active_async(double const* src, double* dst) {
  std::vector<CopyArgs> c(active_.size());
  for (size_t i = 0; i < active_.size(); ++i) {
    const int s = active_[i];
    c[i].src = src + off_[kbX][s];
    c[i].dst = dst + off_[kbX][s];
    c[i].len = std::max(dims_[s].nrows(), dims_[s].ncols());
  }
  if (c.empty()) return;
  Blob bb;
  const size_t o = bb.add(c);
  bb.upload(scratch_, st_);
  hip_check(launch_copy_batched(reinterpret_cast<const CopyArgs*>(scratch_.p + o), int(c.size()), st_), "copy");
  hip_check(hipStreamSynchronize(st_), "copy sync");
}

template class BatchEngine<double, double>;
template class BatchEngine<double, float>;

// ---------------------------------------------------------------------------
BatchSolver::BatchSolver(bool mixed) : mixed_(mixed) {
  if (mixed_) e32_ = new BatchEngine<double, float>();
  else e64_ = new BatchEngine<double, double>();
}

BatchSolver::~BatchSolver() {
  drop_graphs();
  delete e64_;
  delete e32_;
}

void BatchSolver::drop_graphs() {
  for (hipGraphExec_t* g : {&g_.fs, &g_.res, &g_.corr})
    if (*g) { (void)hipGraphExecDestroy(*g); *g = nullptr; }
  g_.B = nullptr;
  g_.X = nullptr;
  g_eng_ = nullptr;
}

hipStream_t BatchSolver::stream() const { return mixed_ ? e32_->stream() : e64_->stream(); }

std::vector<BatchEngine<double, float>::LaunchTime> BatchSolver::time_factorize() {
  drop_graphs();
  if (mixed_) return e32_->time_factorize();
  std::vector<BatchEngine<double, float>::LaunchTime> out;
  for (auto const& t : e64_->time_factorize())
    out.push_back(BatchEngine<double, float>::LaunchTime{t.level, t.cls, t.nblocks, t.ms, t.bytes, t.flops});
  return out;
}

void BatchSolver::plan(std::vector<SysDims> const& dims) {
  drop_graphs();
  if (mixed_) {
    e32_->plan(dims);
    if (e64_) {
      e64_->plan(dims);
    } else {
      // the FP64 fallback (LU, then QRP) is planned when first needed; its
      // shapes are checked now so that a batch it cannot take fails here
      // rather than after the FP32 work
      const size_t lds = e32_->lds_limit();
      for (size_t k = 0; k < dims.size(); ++k) {
        SysDims d = dims[k];
        for (int m : {int(kLU), int(kQRP)}) {
          d.method = m;
          if (!BatchEngine<double, double>::panel_fits(d, lds))
            throw HipError("batch plan: system " + std::to_string(k) +
                           ": the FP64 fallback's pair panel does not fit in LDS");
        }
      }
    }
  } else {
    e64_->plan(dims);
  }
  dims_ = dims;
  R_.resize(size_t(std::max<int64_t>(total(kbX), 1)));
  norms_.resize(2 * dims.size());
  drel_.resize(dims.size());
  dbest_.resize(dims.size());
  dstat_.resize(dims.size());
  dsweeps_.resize(dims.size());
  dstag_.resize(dims.size());
  dopen_.resize(1);
}

int64_t BatchSolver::total(int kind) const { return mixed_ ? e32_->total(kind) : e64_->total(kind); }
int64_t BatchSolver::offset(int kind, int s) const { return mixed_ ? e32_->offset(kind, s) : e64_->offset(kind, s); }

void BatchSolver::bind_device(double const* D, double const* E, double const* B, double const* C, double const* H,
                              double const* Cq, double const* F) {
  drop_graphs();
  oD_ = D; oE_ = E; oB_ = B; oC_ = C; oH_ = H; oCq_ = Cq; oF_ = F;
  if (mixed_) e32_->bind_device(D, E, B, C, H, Cq, F);
  if (e64_) e64_->bind_device(D, E, B, C, H, Cq, F);
}

void BatchSolver::generate(uint64_t seed, double shift, std::vector<int32_t> const& sing_k, double eps) {
  if (mixed_) e32_->generate(seed, shift, sing_k, eps);
  else e64_->generate(seed, shift, sing_k, eps);
}

void BatchSolver::set_pair_variant(int v) {
  drop_graphs();
  if (e32_) e32_->set_pair_variant(v);
  if (e64_) e64_->set_pair_variant(v);
}

void BatchSolver::set_streams(int k, int fork_level) {
  drop_graphs();
  streams_ = k;
  fork_level_ = fork_level;
  if (e32_) e32_->set_streams(k, fork_level);
  if (e64_) e64_->set_streams(k, fork_level);
}

void BatchSolver::set_class_kernel(bool fp32_factors, int cls, int kernel, int waves) {
  if (cls < 0 || cls >= kFacClasses) throw HipError("set_class_kernel: bad class");
  drop_graphs();
  auto& t = fp32_factors ? tune32_ : tune64_;
  t[cls][0] = kernel;
  t[cls][1] = waves;
  if (fp32_factors && e32_) e32_->set_class_kernel(cls, kernel, waves);
  if (!fp32_factors && e64_) e64_->set_class_kernel(cls, kernel, waves);
}

void BatchSolver::apply_tuning(BatchEngine<double, double>& e) {
  for (int c = 0; c < kFacClasses; ++c) e.set_class_kernel(c, tune64_[c][0], tune64_[c][1]);
}

void BatchSolver::set_methods(std::vector<int32_t> const& m) {
  drop_graphs();
  for (size_t s = 0; s < m.size() && s < dims_.size(); ++s) {
    dims_[s].method = m[s];
    if (mixed_) e32_->set_method(int(s), m[s]);
    else e64_->set_method(int(s), m[s]);
  }
}

namespace {
template <typename F>
hipGraphExec_t capture(hipStream_t st, F&& body) {
  hipGraph_t g = nullptr;
  hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(st, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hip_check(hipStreamEndCapture(st, &g), "end capture");
  hipGraphExec_t ex = nullptr;
  hip_check(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0), "graph instantiate");
  (void)hipGraphDestroy(g);
  return ex;
}
}  // namespace

// Refinement sweeps over the systems not yet done: residual and norms, the
// convergence test on the device (converged / stalled: three sweeps without
// halving the best residual / factorization failed), correction solve.  A
// done system is skipped by every launch through the engine's skip flags, so
// the captured residual and correction graphs replay unchanged while the set
// of open systems shrinks; one int (the open count) comes back per sweep.
template <typename E>
void BatchSolver::refine(E& eng, double const* B, double* X, double tol, int max_iter, std::vector<char>& done,
                         BatchReport& rep, int giveup, bool have_residual) {
  const int ns = eng.nsys();
  hipStream_t st = eng.stream();
  // the captured lists belong to one engine: the FP64 fallback engine runs eager
  const bool graph = g_.res && g_.corr && g_eng_ == static_cast<const void*>(&eng);
  std::vector<int32_t> hd(ns);
  for (int s = 0; s < ns; ++s) hd[s] = done[s] ? 1 : 0;
  const size_t i4 = sizeof(int32_t) * size_t(ns), d8 = sizeof(double) * size_t(ns);
  hip_check(hipMemcpyAsync(eng.skip_flags(), hd.data(), i4, hipMemcpyHostToDevice, st), "done flags");
  hip_check(hipMemcpyAsync(dstat_.p, status_.data(), i4, hipMemcpyHostToDevice, st), "status");
  hip_check(hipMemcpyAsync(drel_.p, rel_.data(), d8, hipMemcpyHostToDevice, st), "residuals");
  hip_check(hipMemcpyAsync(dsweeps_.p, sweeps_.data(), i4, hipMemcpyHostToDevice, st), "sweeps");
  ConvergeArgs ca{norms_.p, dstat_.p, dbest_.p, dstag_.p, eng.skip_flags(), drel_.p, dsweeps_.p, dopen_.p, tol, 0,
                  ns, giveup};
  int32_t open = 0;
  bool narrowed = false;
  for (int it = 0;; ++it) {
    if (it == 0 && have_residual) {
      // the caller's residual of the first solve stands
    } else if (graph && !narrowed) {
      hip_check(hipGraphLaunch(g_.res, st), "graph residual");
    } else {
      eng.residual_async(X, B, R_.p);
      eng.norms_async(B, R_.p, norms_.p);
    }
    hip_check(launch_convert<int32_t, int32_t>(nullptr, dopen_.p, 1, st), "open count");
    ca.it = it;
    hip_check(launch_converge_batched(ca, st), "convergence test");
    hip_check(hipMemcpyAsync(&open, dopen_.p, sizeof(int32_t), hipMemcpyDeviceToHost, st), "open count");
    hip_check(hipStreamSynchronize(st), "open count sync");
    rep.sweeps = std::max(rep.sweeps, it);
    if (open == 0 || it >= max_iter) break;
    if (!narrowed && int64_t(open) * 16 < ns) {
      // few systems left: from here on launch only their pairs (every sweep
      // over the whole batch costs its launches and empty workgroups)
      hip_check(hipMemcpy(hd.data(), eng.skip_flags(), i4, hipMemcpyDeviceToHost), "done flags");
      std::vector<char> mask(ns);
      for (int s = 0; s < ns; ++s) mask[s] = hd[s] == 0;
      eng.narrow(mask);
      narrowed = true;
    }
    if (graph && !narrowed) {
      hip_check(hipGraphLaunch(g_.corr, st), "graph correction");
    } else {
      eng.solve_async(R_.p);
      eng.add_active_async(R_.p, X);
    }
  }
  hip_check(hipMemcpyAsync(rel_.data(), drel_.p, d8, hipMemcpyDeviceToHost, st), "residuals");
  hip_check(hipMemcpyAsync(sweeps_.data(), dsweeps_.p, i4, hipMemcpyDeviceToHost, st), "sweeps");
  hip_check(hipMemcpyAsync(hd.data(), eng.skip_flags(), i4, hipMemcpyDeviceToHost, st), "done flags");
  hip_check(hipStreamSynchronize(st), "refinement results");
  for (int s = 0; s < ns; ++s) done[s] = hd[s] != 0;
  if (narrowed) eng.widen();
}

// the factor + first solve, the residual + norms and the correction launch
// lists of the current tables, captured once per (B, X) pair
template <typename E>
void BatchSolver::ensure_graphs(E& eng, double const* B, double* X) {
  if (g_.B == B && g_.X == X && g_.fs && g_eng_ == static_cast<const void*>(&eng)) return;
  drop_graphs();
  hipStream_t st = eng.stream();
  const int64_t nx = eng.total(kbX);
  eng.prepare(X, B, R_.p);   // no allocation inside a capture
  g_.fs = capture(st, [&] {
    hip_check(hipMemcpyAsync(X, B, size_t(nx) * sizeof(double), hipMemcpyDeviceToDevice, st), "copy b");
    eng.factorize_async();
    eng.solve_async(X);
  });
  g_.res = capture(st, [&] {
    eng.residual_async(X, B, R_.p);
    eng.norms_async(B, R_.p, norms_.p);
  });
  g_.corr = capture(st, [&] {
    eng.solve_async(R_.p);
    eng.add_active_async(R_.p, X);
  });
  g_.B = B;
  g_.X = X;
  g_eng_ = &eng;
}

// Systems that ended above tol without a factorization failure: their
// componentwise backward error max_i |r_i| / (|A| |x| + |b|)_i, from the
// residual the refinement left in R_ (a system leaves the sweeps with its
// last residual in place).  omega <= 64 u marks the FP64 floor: x solves a
// system within a few roundings of (A, b), no FP64 path gets the residual
// lower, and the FP64 fallback is not tried.  Returns the floor count.
template <typename E>
int BatchSolver::mark_floor(E& eng, double const* B, double* X, double tol) {
  const int ns = eng.nsys();
  floor_.assign(ns, 0);
  std::vector<char> cand(ns, 0);
  int k = 0;
  for (int s = 0; s < ns; ++s)
    if (status_[s] == 0 && rel_[s] == rel_[s] && !(rel_[s] <= tol)) {
      cand[s] = 1;
      ++k;
    }
  if (k == 0) return 0;
  hipStream_t st = eng.stream();
  eng.narrow(cand);   // tables of the candidates only; widen() restores the others
  hip_check(launch_convert<int32_t, int32_t>(nullptr, eng.skip_flags(), ns, st), "clear skip flags");
  den_.resize(size_t(std::max<int64_t>(eng.total(kbX), 1)));
  berr_.resize(size_t(ns));
  eng.backward_errors_async(X, B, R_.p, den_.p, berr_.p);
  std::vector<double> om(ns, 0.0);
  hip_check(hipMemcpyAsync(om.data(), berr_.p, sizeof(double) * size_t(ns), hipMemcpyDeviceToHost, st), "omega");
  hip_check(hipStreamSynchronize(st), "omega sync");
  eng.widen();
  int f = 0;
  for (int s = 0; s < ns; ++s)
    if (cand[s] && om[s] <= MixedEngine::kFloorBerr) {
      floor_[s] = 1;
      ++f;
    }
  return f;
}

template <typename E>
BatchReport BatchSolver::run(E& eng, double const* B, double* X, double tol, int max_iter, bool use_graph) {
  const auto t0 = std::chrono::steady_clock::now();
  const int ns = eng.nsys();
  hipStream_t st = eng.stream();
  const int64_t nx = eng.total(kbX);
  if (int(eng.active().size()) != ns) {   // a previous sweep narrowed the set
    drop_graphs();
    eng.set_active(std::vector<char>(ns, 1));
  }
  if (use_graph) {
    ensure_graphs(eng, B, X);
    hip_check(hipGraphLaunch(g_.fs, st), "graph factor+solve");
  } else {
    drop_graphs();
    hip_check(hipMemcpyAsync(X, B, size_t(nx) * sizeof(double), hipMemcpyDeviceToDevice, st), "copy b");
    eng.factorize_async();
    eng.solve_async(X);
  }
  status_.assign(ns, 0);
  eng.status(status_.data());
  rel_.assign(ns, 0.0);
  sweeps_.assign(ns, 0);
  std::vector<char> done(ns, 0);
  BatchReport rep;
  refine(eng, B, X, tol, max_iter, done, rep);
  mark_floor(eng, B, X, tol);
  for (int s = 0; s < ns; ++s) {
    rep.max_rel_residual = std::max(rep.max_rel_residual, rel_[s]);
    if (status_[s] == 0 && !(rel_[s] <= tol)) status_[s] |= 4;
    if (floor_[s]) status_[s] |= 8;
    if (status_[s] != 0) ++rep.failed;
  }
  rep.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rep;
}

BatchReport BatchSolver::solve(double const* B, double* X, double tol, int max_iter, bool use_graph) {
  return mixed_ ? run(*e32_, B, X, tol, max_iter, use_graph) : run(*e64_, B, X, tol, max_iter, use_graph);
}

BatchReport BatchSolver::solve_selected(double const* B, double* X, double tol, int max_iter, Mlp* selector,
                                        float threshold, bool fallback) {
  if (!mixed_) throw HipError("solve_selected: needs the mixed batch solver");
  const auto t0 = std::chrono::steady_clock::now();
  auto& eng = *e32_;
  const int ns = eng.nsys();
  hipStream_t st = eng.stream();
  BatchReport rep;
  status_.assign(ns, 0);
  rel_.assign(ns, 0.0);
  sweeps_.assign(ns, 0);
  chose_qrp_.assign(ns, 0);
  fell_back_.assign(ns, 0);
  feat_.assign(size_t(ns) * 12, 0.f);
  // 1. FP32 LU everywhere, first solve and its residual.  With the FP64
  // fallback the launch lists stay those of every system with LU, so they are
  // captured once and replayed (like solve()); a previous call that moved
  // systems to QRP or narrowed the set rebuilds them
  bool all_lu = int(eng.active().size()) == ns;
  for (int s = 0; s < ns && all_lu; ++s) all_lu = eng.method(s) == kLU;
  if (!all_lu) {
    drop_graphs();
    for (int s = 0; s < ns; ++s) eng.set_method(s, kLU);
    eng.set_active(std::vector<char>(ns, 1));
  }
  const bool use_graph = fallback && graphs_;
  if (use_graph) {
    ensure_graphs(eng, B, X);
    hip_check(hipGraphLaunch(g_.fs, st), "graph factor+solve");
    hip_check(hipGraphLaunch(g_.res, st), "graph residual");
  } else {
    drop_graphs();
    hip_check(hipMemcpyAsync(X, B, size_t(eng.total(kbX)) * sizeof(double), hipMemcpyDeviceToDevice, st), "copy b");
    eng.factorize_async();
    eng.solve_async(X);
    eng.residual_async(X, B, R_.p);
    eng.norms_async(B, R_.p, norms_.p);
  }
  eng.status(status_.data());
  // 2. features (always: they are also the selector's training inputs), selector
  std::vector<char> direct(ns, 0);   // flagged systems sent straight to the FP64 path
  dfeat_.resize(size_t(ns) * 12);
  eng.features_async(norms_.p, dfeat_.p);
  hip_check(hipMemcpyAsync(feat_.data(), dfeat_.p, feat_.size() * sizeof(float), hipMemcpyDeviceToHost, st),
            "features");
  hip_check(hipStreamSynchronize(st), "features sync");
  if (selector) {
    dprob_.resize(size_t(ns));
    selector->predict(dfeat_.p, ns, dprob_.p);
    std::vector<float> p(ns);
    hip_check(hipMemcpy(p.data(), dprob_.p, p.size() * sizeof(float), hipMemcpyDeviceToHost), "probabilities");
    std::vector<char> mask(ns, 0);
    int nq = 0;
    for (int s = 0; s < ns; ++s)
      if (p[s] > threshold) {
        mask[s] = 1;
        chose_qrp_[s] = 1;
        ++nq;
      }
    if (fallback) {
      // with the FP64 path behind it, a flagged system goes there at once:
      // the systems FP32 LU cannot solve are nearly singular beyond FP32's
      // reach, FP32 QRP does not rescue them either, and the FP32 sweeps
      // they would stall in are the costly tail of the refinement
      direct = mask;
    } else if (nq > 0) {
      drop_graphs();   // the launch lists change with the methods
      for (int s = 0; s < ns; ++s)
        if (mask[s]) eng.set_method(s, kQRP);
      eng.set_active(mask);
      eng.copy_active_async(B, X);
      eng.factorize_async();
      eng.solve_async(X);
      std::vector<int32_t> st2(ns, 0);
      eng.status(st2.data());
      for (int s = 0; s < ns; ++s)
        if (mask[s]) status_[s] = st2[s];
      eng.set_active(std::vector<char>(ns, 1));
    }
  }
  // 3. FP64 refinement of every system with its factors; with the fallback
  // behind it, a system that would need more than six further sweeps at its
  // contraction rate is handed over at once instead of sweeping to max_iter
  std::vector<char> done(direct.begin(), direct.end());
  refine(eng, B, X, tol, max_iter, done, rep, fallback ? 6 : 0, /*have_residual=*/true);
  // systems at the FP64 floor stay where they are (no FP64 path does better;
  // the systems sent straight to the FP64 path never ran the sweeps and
  // carry no residual, so they are not candidates)
  mark_floor(eng, B, X, tol);
  // 4. the FP64 fallback for the systems that did not get there: first the
  // plain FP64 path (LU elimination, FP64 refinement), then FP64 QRP for what
  // FP64 LU cannot solve either (QRP's factors lose digits on nearly singular
  // pair columns; the refinement recovers them while the FP64 backward error
  // stays below one).  fell_back_ = 1 (FP64 LU) or 2 (FP64 QRP).
  auto failed = [&](std::vector<char> const& among, std::vector<char>& out) {
    int k = 0;
    for (int s = 0; s < ns; ++s) {
      out[s] = among[s] && !floor_[s] && (status_[s] != 0 || !(rel_[s] <= tol));
      k += out[s];
    }
    return k;
  };
  std::vector<char> all(ns, 1), bad(ns, 0);
  int nbad = failed(all, bad);
  for (int s = 0; s < ns; ++s)
    if (direct[s] && !bad[s]) {
      bad[s] = 1;
      ++nbad;
    }
  if (fallback && nbad > 0) {
    if (!e64_) {
      e64_ = new BatchEngine<double, double>();
      e64_->plan(dims_);
      e64_->bind_device(oD_, oE_, oB_, oC_, oH_, oCq_, oF_);
      apply_tuning(*e64_);
      e64_->set_streams(streams_, fork_level_);
    }
    auto& f = *e64_;
    hip_check(hipStreamSynchronize(st), "sync before fallback");
    for (int stage = 1; stage <= 2; ++stage) {
      for (int s = 0; s < ns; ++s) f.set_method(s, stage == 1 ? kLU : kQRP);
      f.set_active(bad);
      f.copy_active_async(B, X);
      f.factorize_async();
      f.solve_async(X);
      std::vector<int32_t> st3(ns, 0);
      f.status(st3.data());
      std::vector<char> done2(ns, 1);
      for (int s = 0; s < ns; ++s)
        if (bad[s]) {
          fell_back_[s] = char(stage);
          status_[s] = st3[s];
          done2[s] = 0;
        }
      BatchReport r2;
      refine(f, B, X, tol, max_iter, done2, r2);
      rep.sweeps = std::max(rep.sweeps, r2.sweeps);
      std::vector<char> still(ns, 0);
      if (stage == 1 && failed(bad, still) == 0) break;
      bad.swap(still);
    }
  }
  for (int s = 0; s < ns; ++s) {
    rep.max_rel_residual = std::max(rep.max_rel_residual, rel_[s]);
    if (status_[s] == 0 && !(rel_[s] <= tol)) status_[s] |= 4;
    if (floor_[s]) status_[s] |= 8;
    if (status_[s] != 0) ++rep.failed;
  }
  rep.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rep;
}

}  // namespace bcr
