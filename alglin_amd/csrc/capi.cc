// extern "C" boundary of alglin_amd (declarations in include/alglin_amd.h).
#include "../../include/alglin_amd.h"

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <type_traits>

#include "bcr_gpu.hh"
#include "bcr_batch.hh"
#include "bcr_mlp.hh"

using alglin_amd::BorderedCR;

struct alglin_bcr {
  int dtype = 0;
  std::unique_ptr<BorderedCR<double>> d;
  std::unique_ptr<BorderedCR<float>> f;
  std::string info, err;
  bool bound = false;
};

namespace {

template <typename Fn>
int guarded(alglin_bcr_t h, Fn&& fn) {
  if (!h) return -1;
  try {
    return fn();
  } catch (std::exception const& e) {
    h->err = e.what();
    return 1;
  }
}

template <typename Fn>
auto on(alglin_bcr_t h, Fn&& fn) {
  return h->dtype == 0 ? fn(*h->d, (double*)nullptr) : fn(*h->f, (float*)nullptr);
}

int on_(alglin_bcr_t h, int on) {
  if (h->dtype == 0) h->d->gpu().set_timing(on != 0);
  else h->f->gpu().set_timing(on != 0);
  return 0;
}

template <typename S>
int result(alglin_bcr_t h, S& s, bool ok) {
  if (!ok) h->err = s.last_error();
  return ok ? 0 : 1;
}

}  // namespace

extern "C" {

alglin_bcr_t alglin_bcr_create(int dtype) {
  try {
    auto* h = new alglin_bcr;
    h->dtype = dtype == 1 ? 1 : 0;
    if (h->dtype == 0) h->d.reset(new BorderedCR<double>(nullptr));
    else h->f.reset(new BorderedCR<float>(nullptr));
    return h;
  } catch (...) {
    return nullptr;
  }
}

void alglin_bcr_destroy(alglin_bcr_t h) { delete h; }

int alglin_bcr_allocate(alglin_bcr_t h, int32_t N, int32_t n, int32_t qr, int32_t qx, int32_t nr, int32_t nx) {
  return guarded(h, [&] {
    h->bound = false;
    return on(h, [&](auto& s, auto*) { s.allocate(N, n, qr, qx, nr, nx); return 0; });
  });
}

int alglin_bcr_select(alglin_bcr_t h, char const* internal, char const* last, char const* last2) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto*) {
      if (internal) s.select(internal);
      if (last) s.select_last(last);
      if (last2) s.select_last2(last2);
      return 0;
    });
  });
}

int alglin_bcr_dims(alglin_bcr_t h, int32_t out[10]) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto*) {
      int32_t v[10] = {s.number_of_blocks(), s.block_size(), s.dim_qr(), s.dim_qx(), s.dim_nr(),
                       s.dim_nx(), s.nrows(), s.ncols(), s.Nr(), s.Nc()};
      std::memcpy(out, v, sizeof(v));
      return 0;
    });
  });
}

int alglin_bcr_load(alglin_bcr_t h, char const* what, int32_t nbl, void const* A, int32_t ld) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      T const* p = static_cast<T const*>(A);
      std::string const w(what ? what : "");
      if (w == "B") s.load_B(nbl, p, ld);
      else if (w == "C") s.load_C(nbl, p, ld);
      else if (w == "D") s.load_D(nbl, p, ld);
      else if (w == "E") s.load_E(nbl, p, ld);
      else if (w == "DE") s.load_DE(nbl, p, ld);
      else if (w == "DEB") s.load_DEB(nbl, p, ld);
      else if (w == "F") s.load_F(p, ld);
      else if (w == "Cq") s.load_Cq(p, ld);
      else if (w == "CqF") s.load_CqF(p, ld);
      else if (w == "H") s.load_bottom(p, ld);
      else if (w == "add_B") s.add_to_B(nbl, p, ld);
      else if (w == "add_C") s.add_to_C(nbl, p, ld);
      else if (w == "add_C2") s.add_to_C2(nbl, p, ld);
      else if (w == "add_C2F") s.add_to_C2F(nbl, p, ld);
      else if (w == "add_F") s.add_to_F(p, ld);
      else throw std::invalid_argument("alglin_bcr_load: unknown block '" + w + "'");
      return 0;
    });
  });
}

int alglin_bcr_load_bottom(alglin_bcr_t h, void const* H0, int32_t ld0, void const* HN, int32_t ldN, void const* Hq,
                           int32_t ldQ, void const* Hp, int32_t ldP) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      s.load_bottom(static_cast<T const*>(H0), ld0, static_cast<T const*>(HN), ldN, static_cast<T const*>(Hq), ldQ,
                    static_cast<T const*>(Hp), ldP);
      return 0;
    });
  });
}

int alglin_bcr_load_bottom2(alglin_bcr_t h, void const* C0, int32_t ld0, void const* CN, int32_t ldN, void const* Cq,
                            int32_t ldCq, void const* F, int32_t ldF) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      s.load_bottom2(static_cast<T const*>(C0), ld0, static_cast<T const*>(CN), ldN, static_cast<T const*>(Cq), ldCq,
                     static_cast<T const*>(F), ldF);
      return 0;
    });
  });
}

int alglin_bcr_load_packed(alglin_bcr_t h, void const* D, void const* E, void const* B, void const* C, void const* H,
                           void const* Cq, void const* F) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      const size_t N = s.number_of_blocks(), n = s.block_size(), nr = s.dim_nr(), nx = s.dim_nx();
      auto cp = [](void const* src, T* dst, size_t cnt) { if (src && cnt) std::memcpy(dst, src, cnt * sizeof(T)); };
      cp(D, s.data_D(), N * n * n);
      cp(E, s.data_E(), N * n * n);
      cp(B, s.data_B(), N * n * nx);
      cp(C, s.data_C(), (N + 1) * nr * n);
      cp(H, s.data_H(), (n + s.dim_qr()) * size_t(s.Nc()));
      cp(Cq, s.data_Cq(), nr * s.dim_qx());
      cp(F, s.data_F(), nr * nx);
      return 0;
    });
  });
}

int alglin_bcr_factorize(alglin_bcr_t h) {
  return guarded(h, [&] { return on(h, [&](auto& s, auto*) { return result(h, s, s.factorize()); }); });
}

int alglin_bcr_solve(alglin_bcr_t h, void* x) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      return result(h, s, s.solve(static_cast<T*>(x)));
    });
  });
}

int alglin_bcr_solve_n(alglin_bcr_t h, int32_t nrhs, void* X, int32_t ld) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      return result(h, s, s.solve(nrhs, static_cast<T*>(X), ld));
    });
  });
}

int alglin_bcr_mv(alglin_bcr_t h, void const* x, void* y) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      s.Mv(static_cast<T const*>(x), static_cast<T*>(y));
      return 0;
    });
  });
}

int32_t alglin_bcr_sparse_nnz(alglin_bcr_t h) {
  if (!h) return -1;
  return on(h, [&](auto& s, auto*) { return s.sparse_nnz(); });
}

int alglin_bcr_sparse_pattern(alglin_bcr_t h, int32_t* I, int32_t* J, int32_t offs) {
  return guarded(h, [&] { return on(h, [&](auto& s, auto*) { s.sparse_pattern(I, J, offs); return 0; }); });
}

int alglin_bcr_sparse_values(alglin_bcr_t h, void* V) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      s.sparse_values(static_cast<T*>(V));
      return 0;
    });
  });
}

int alglin_bcr_sparse_load(alglin_bcr_t h, void const* V, int32_t const* I, int32_t r_offs, int32_t const* J,
                           int32_t c_offs, int32_t nnz, int do_error) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      s.sparse_load(static_cast<T const*>(V), I, r_offs, J, c_offs, nnz, do_error != 0);
      return 0;
    });
  });
}

int alglin_bcr_check_matrix(alglin_bcr_t h) {
  return guarded(h, [&] { return on(h, [&](auto& s, auto*) { s.check_matrix(); return 0; }); });
}

char const* alglin_bcr_last_error(alglin_bcr_t h) {
  if (!h) return "null handle";
  if (!h->err.empty()) return h->err.c_str();
  return on(h, [&](auto& s, auto*) { return s.last_error().data(); });
}

char const* alglin_bcr_info(alglin_bcr_t h) {
  if (!h) return "";
  h->info = on(h, [&](auto& s, auto*) { return s.info(); });
  return h->info.c_str();
}

int alglin_bcr_set_debug(alglin_bcr_t h, int flag) {
  return guarded(h, [&] { return on(h, [&](auto& s, auto*) { s.set_debug(flag != 0); return 0; }); });
}

int alglin_bcr_set_num_parallel_block(alglin_bcr_t h, int32_t npb) {
  return guarded(h, [&] { return on(h, [&](auto& s, auto*) { s.set_num_parallel_block(npb); return 0; }); });
}

int32_t alglin_bcr_used_parallel_block(alglin_bcr_t h) {
  if (!h) return -1;
  return on(h, [&](auto& s, auto*) { return s.used_parallel_block(); });
}

int alglin_bcr_dup(alglin_bcr_t h, alglin_bcr_t src) {
  return guarded(h, [&] {
    if (!src || src->dtype != h->dtype) throw std::invalid_argument("alglin_bcr_dup: source of another type");
    if (h->dtype == 0) h->d->dup(*src->d);
    else h->f->dup(*src->f);
    h->bound = false;
    return 0;
  });
}

int alglin_bcr_t_solve(alglin_bcr_t h, void* x) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      return s.t_solve(static_cast<T*>(x)) ? 0 : 1;
    });
  });
}

int64_t alglin_bcr_matlab_script(alglin_bcr_t h, char* buf, int64_t len) {
  if (!h) return -1;
  try {
    std::ostringstream o;
    on(h, [&](auto& s, auto*) { s.print_matlab_script(o); return 0; });
    std::string const t = o.str();
    if (buf && len > 0) {
      const size_t k = std::min<size_t>(t.size(), size_t(len - 1));
      std::memcpy(buf, t.data(), k);
      buf[k] = 0;
    }
    return int64_t(t.size()) + 1;
  } catch (std::exception const& e) {
    h->err = e.what();
    return -1;
  }
}

int alglin_bcr_get_packed(alglin_bcr_t h, void* D, void* E, void* B, void* C, void* H, void* Cq, void* F) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      const size_t N = s.number_of_blocks(), n = s.block_size(), nr = s.dim_nr(), nx = s.dim_nx();
      auto cp = [](void* dst, T const* src, size_t cnt) { if (dst && cnt) std::memcpy(dst, src, cnt * sizeof(T)); };
      cp(D, s.data_D(), N * n * n);
      cp(E, s.data_E(), N * n * n);
      cp(B, s.data_B(), N * n * nx);
      cp(C, s.data_C(), (N + 1) * nr * n);
      cp(H, s.data_H(), (n + s.dim_qr()) * size_t(s.Nc()));
      cp(Cq, s.data_Cq(), nr * s.dim_qx());
      cp(F, s.data_F(), nr * nx);
      return 0;
    });
  });
}

int alglin_bcr_bind_device(alglin_bcr_t h, void const* D, void const* E, void const* B, void const* C, void const* H,
                           void const* Cq, void const* F) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      s.bind_device(static_cast<T const*>(D), static_cast<T const*>(E), static_cast<T const*>(B),
                    static_cast<T const*>(C), static_cast<T const*>(H), static_cast<T const*>(Cq),
                    static_cast<T const*>(F));
      h->bound = true;
      return 0;
    });
  });
}

int alglin_bcr_factorize_device(alglin_bcr_t h, int sync) {
  return guarded(h, [&] {
    if (!h->bound) throw std::runtime_error("alglin_bcr_factorize_device: no device operator bound");
    return on(h, [&](auto& s, auto*) {
      const int32_t st = s.factorize_device(sync != 0);
      if (st) h->err = st & 1 ? "zero pivot in a cyclic-reduction elimination" : "singular last block";
      return st ? 1 : 0;
    });
  });
}

int alglin_bcr_solve_device(alglin_bcr_t h, void* X, int32_t nrhs, int64_t ld, int sync) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      s.solve_device(static_cast<T*>(X), nrhs, ld, sync != 0);
      const int rs = s.refine_report().status;
      if (s.mixed() && !bcr::refine_ok(rs)) {   // 0 converged, 4 FP64 QRP fallback, 5 FP64 floor
        h->err = "iterative refinement did not reach the requested residual";
        return 1;
      }
      return 0;
    });
  });
}

int alglin_bcr_set_precision(alglin_bcr_t h, char const* mode, double tol, int32_t max_iter) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto*) { s.select_precision(mode ? mode : "FP64", tol, max_iter); return 0; });
  });
}

int alglin_bcr_set_pair_kernel(alglin_bcr_t h, int variant) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto*) {
      s.gpu().set_pair_variant(variant);
      return 0;
    });
  });
}

int alglin_bcr_set_fallback(alglin_bcr_t h, int on_) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto*) { s.set_fallback(on_ != 0); return 0; });
  });
}

int alglin_bcr_refine_report(alglin_bcr_t h, double out[5]) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto*) {
      auto const r = s.refine_report();
      out[0] = r.iterations; out[1] = r.rel_residual; out[2] = r.first_residual; out[3] = r.status;
      out[4] = r.backward_error;
      return 0;
    });
  });
}

int alglin_bcr_residual_device(alglin_bcr_t h, double const* x, double const* b, double* r, int sync) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto*) {
      s.residual_device(x, b, r, sync != 0);
      return 0;
    });
  });
}

int alglin_bcr_sync(alglin_bcr_t h) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto*) {
      s.synchronize();
      return 0;
    });
  });
}

void* alglin_bcr_stream(alglin_bcr_t h) {
  if (!h) return nullptr;
  try {
    return on(h, [&](auto& s, auto*) { return s.stream(); });
  } catch (std::exception const& e) {
    h->err = e.what();
    return nullptr;
  }
}

int alglin_bcr_get_factors(alglin_bcr_t h, void* T_, int32_t* perm, void* D, void* E, void* B, void* C,
                           double* last_block, double* pending_sumsq) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      auto& e = s.gpu().fp();
      const size_t N = s.number_of_blocks(), n = s.block_size(), nr = s.dim_nr(), nx = s.dim_nx();
      auto cp = [&](void* dst, void const* src, size_t bytes) {
        if (dst && bytes) bcr::hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "D2H");
      };
      cp(T_, e.factor_T(), N * bcr::tsize(int(n)) * sizeof(T));
      cp(perm, e.factor_perm(), N * n * sizeof(int32_t));
      cp(D, e.work_D(), N * n * n * sizeof(T));
      cp(E, e.work_E(), N * n * n * sizeof(T));
      cp(B, e.work_B(), N * n * nx * sizeof(T));
      cp(C, e.work_C(), (N + 1) * nr * n * sizeof(T));
      cp(last_block, e.last_block(), size_t(s.Nr()) * s.Nc() * sizeof(double));
      if (pending_sumsq) {
        std::vector<T> pend(e.pending_count());
        cp(pend.data(), e.pending(), pend.size() * sizeof(T));
        double ss = 0;
        for (T v : pend) ss += double(v) * double(v);
        *pending_sumsq = ss;
      }
      return 0;
    });
  });
}

int alglin_bcr_fast_path(alglin_bcr_t h) {
  if (!h) return 0;
  return on(h, [&](auto& s, auto*) {
    auto& g = s.gpu();
    return g.eng ? g.eng->stats().fast_path : 0;
  });
}

int alglin_bcr_set_timing(alglin_bcr_t h, int on) {
  return guarded(h, [&] { return on_(h, on); });
}

int alglin_bcr_timing(alglin_bcr_t h, double out[8], int reset) {
  return guarded(h, [&] {
    return on(h, [&](auto& s, auto*) {
      auto& g = s.gpu();
      for (int i = 0; i < 8; ++i) out[i] = 0;
      if (g.eng) g.eng->timing(out, reset != 0);
      if (g.mix) {   // phases of the mixed engine add to those of the FP64 one
        double m[8];
        g.mix->engine().timing(m, reset != 0);
        for (int i = 0; i < 8; ++i) out[i] += m[i];
      }
      return 0;
    });
  });
}

void* alglin_dev_malloc(int64_t bytes) {
  void* p = nullptr;
  if (bytes <= 0) bytes = 1;
  if (hipMalloc(&p, size_t(bytes)) != hipSuccess) return nullptr;
  return p;
}

int alglin_dev_free(void* p) { return p && hipFree(p) == hipSuccess ? 0 : 1; }

// The solver streams are non-blocking (they do not order against the null
// stream), so every plain copy below completes on the device before it
// returns; work queued afterwards on any stream sees the data.
static int copy_sync(void* dst, void const* src, int64_t bytes, hipMemcpyKind kind) {
  if (bytes <= 0) return 0;
  if (hipMemcpy(dst, src, size_t(bytes), kind) != hipSuccess) return 1;
  return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}

int alglin_memcpy_htod(void* dst, void const* src, int64_t bytes) {
  return copy_sync(dst, src, bytes, hipMemcpyHostToDevice);
}

int alglin_memcpy_dtoh(void* dst, void const* src, int64_t bytes) {
  return copy_sync(dst, src, bytes, hipMemcpyDeviceToHost);
}

int alglin_memcpy_dtod(void* dst, void const* src, int64_t bytes) {
  return copy_sync(dst, src, bytes, hipMemcpyDeviceToDevice);
}

int alglin_memcpy_dtod_async(void* dst, void const* src, int64_t bytes, void* stream) {
  if (bytes <= 0) return 0;
  return hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)) ==
                 hipSuccess
             ? 0
             : 1;
}

int alglin_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : 1; }

int alglin_device_info(char* name, int32_t len, int32_t* cus, int64_t* hbm_bytes) {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 1;
  if (name && len > 0) {
    std::strncpy(name, prop.gcnArchName, size_t(len) - 1);
    name[len - 1] = 0;
  }
  if (cus) *cus = prop.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = int64_t(prop.totalGlobalMem);
  return 0;
}

// ---- reference C interface -----------------------------------------------------

namespace {
std::mutex g_mu;
std::map<BABD_intType, std::unique_ptr<BorderedCR<double>>> g_solvers;
std::string g_last_error = "no error";

BorderedCR<double>& solver_for(BABD_intType id) {
  auto& p = g_solvers[id];
  if (!p) p.reset(new BorderedCR<double>(nullptr));
  return *p;
}

void configure(BorderedCR<double>& s, BABD_intType mat_fact, BABD_intType last_block_fact) {
  switch (mat_fact) {
    case 0: case 2: s.select("LU"); break;
    case 3: s.select("QR"); break;
    case 4: s.select("QRP"); break;
    default: throw std::invalid_argument("BABD_factorize: mat_fact must be 0 (CR+LU), 2, 3 or 4");
  }
  static char const* names[] = {"LU", "LUPQ", "QR", "QRP", "SVD", "LSS", "LSY", "PINV"};
  if (last_block_fact < 0 || last_block_fact > 7) throw std::invalid_argument("BABD_factorize: bad last_block_fact");
  s.select_last(names[last_block_fact]);
}

int finish(BorderedCR<double>& s) {
  if (!s.factorize()) {
    g_last_error = s.last_error();
    return 1;
  }
  return 0;
}
}  // namespace

int BABD_factorize(BABD_intType id, BABD_intType mat_fact, BABD_intType last_block_fact, BABD_intType nblock,
                   BABD_intType n, BABD_intType qr, BABD_intType qx, BABD_realType const DE[], BABD_intType ldDE,
                   BABD_realType const H0[], BABD_intType ldH0, BABD_realType const HN[], BABD_intType ldHN,
                   BABD_realType const Hq[], BABD_intType ldHq) {
  std::lock_guard<std::mutex> g(g_mu);
  try {
    auto& s = solver_for(id);
    s.allocate(nblock, n, qr, qx, 0, 0);
    configure(s, mat_fact, last_block_fact);
    for (BABD_intType k = 0; k < nblock; ++k) s.load_DE(k, DE + size_t(2) * n * k * ldDE, ldDE);
    s.load_bottom(H0, ldH0, HN, ldHN, Hq, ldHq, nullptr, 1);
    return finish(s);
  } catch (std::exception const& e) {
    g_last_error = e.what();
    return 1;
  }
}

int BABD_factorize_bordered(BABD_intType id, BABD_intType mat_fact, BABD_intType last_block_fact, BABD_intType nblock,
                            BABD_intType n, BABD_intType qr, BABD_intType nr, BABD_intType qx, BABD_intType nx,
                            BABD_realType const DE[], BABD_intType ldDE, BABD_realType const H0[], BABD_intType ldH0,
                            BABD_realType const HN[], BABD_intType ldHN, BABD_realType const Hq[], BABD_intType ldHq,
                            BABD_realType const B[], BABD_intType ldB, BABD_realType const C[], BABD_intType ldC,
                            BABD_realType const D[], BABD_intType ldD) {
  std::lock_guard<std::mutex> g(g_mu);
  try {
    auto& s = solver_for(id);
    s.allocate(nblock, n, qr, qx, nr, nx);
    configure(s, mat_fact, last_block_fact);
    for (BABD_intType k = 0; k < nblock; ++k) {
      s.load_DE(k, DE + size_t(2) * n * k * ldDE, ldDE);
      // B is the full right border column block: rows k*n.. of B_k, then Hp
      s.load_B(k, B + size_t(k) * n, ldB);
    }
    s.load_bottom(H0, ldH0, HN, ldHN, Hq, ldHq, B + size_t(nblock) * n, ldB);
    // C is the full bottom border row block: C_0..C_N then Cq; D is the corner F
    for (BABD_intType k = 0; k <= nblock; ++k) s.load_C(k, C + size_t(k) * n * ldC, ldC);
    s.load_Cq(C + size_t(nblock + 1) * n * ldC, ldC);
    s.load_F(D, ldD);
    return finish(s);
  } catch (std::exception const& e) {
    g_last_error = e.what();
    return 1;
  }
}

int BABD_solve(BABD_intType id, BABD_realType rhs_sol[]) {
  std::lock_guard<std::mutex> g(g_mu);
  auto it = g_solvers.find(id);
  if (it == g_solvers.end()) {
    g_last_error = "BABD_solve: unknown mat_id";
    return 1;
  }
  if (!it->second->solve(rhs_sol)) {
    g_last_error = it->second->last_error();
    return 1;
  }
  return 0;
}

char const* BABD_get_last_error(void) { return g_last_error.c_str(); }

int BABD_free(BABD_intType id) {
  std::lock_guard<std::mutex> g(g_mu);
  return g_solvers.erase(id) ? 0 : 1;
}

}  // extern "C"

// ---- batched systems ---------------------------------------------------------
struct alglin_batch {
  std::unique_ptr<bcr::BatchSolver> s;
  std::string err{"no error"};
  bcr::BatchReport rep;
};

namespace {
template <typename Fn>
int bguard(alglin_batch_t h, Fn&& fn) {
  if (!h) return 1;
  try {
    return fn();
  } catch (std::exception const& e) {
    h->err = e.what();
    return 1;
  }
}
}  // namespace

extern "C" {

alglin_batch_t alglin_batch_create(int mixed) {
  try {
    auto* h = new alglin_batch;
    h->s.reset(new bcr::BatchSolver(mixed != 0));
    return h;
  } catch (...) {
    return nullptr;
  }
}

void alglin_batch_destroy(alglin_batch_t h) { delete h; }

int alglin_batch_plan(alglin_batch_t h, int32_t nsys, int32_t const* dims) {
  return bguard(h, [&] {
    if (nsys <= 0 || !dims) throw std::invalid_argument("alglin_batch_plan: no systems");
    std::vector<bcr::SysDims> d(nsys);
    for (int32_t s = 0; s < nsys; ++s) {
      int32_t const* v = dims + 7 * size_t(s);
      d[s] = bcr::SysDims{v[0], v[1], v[2], v[3], v[4], v[5], v[6]};
    }
    h->s->plan(d);
    return 0;
  });
}

int64_t alglin_batch_size(alglin_batch_t h, int32_t kind) {
  if (!h || kind < 0 || kind >= bcr::kbKinds) return -1;
  return h->s->total(kind);
}

int64_t alglin_batch_offset(alglin_batch_t h, int32_t kind, int32_t s) {
  if (!h || kind < 0 || kind >= bcr::kbKinds || s < 0) return -1;
  return h->s->offset(kind, s);
}

int alglin_batch_bind_device(alglin_batch_t h, double const* D, double const* E, double const* B, double const* C,
                             double const* H, double const* Cq, double const* F) {
  return bguard(h, [&] {
    h->s->bind_device(D, E, B, C, H, Cq, F);
    return 0;
  });
}

int alglin_batch_generate(alglin_batch_t h, uint64_t seed, double shift) {
  return bguard(h, [&] {
    h->s->generate(seed, shift);
    return 0;
  });
}

int alglin_batch_generate_singular(alglin_batch_t h, uint64_t seed, double shift, int32_t nsys,
                                   int32_t const* sing_k, double eps) {
  return bguard(h, [&] {
    h->s->generate(seed, shift, std::vector<int32_t>(sing_k, sing_k + nsys), eps);
    return 0;
  });
}

int alglin_batch_solve(alglin_batch_t h, double const* B, double* X, double tol, int32_t max_iter, int use_graph) {
  return bguard(h, [&] {
    h->rep = h->s->solve(B, X, tol, max_iter, use_graph != 0);
    return h->rep.failed == 0 ? 0 : 2;
  });
}

int alglin_batch_report(alglin_batch_t h, int32_t* status, double* rel_residual, int32_t* sweeps, double summary[4]) {
  return bguard(h, [&] {
    auto const& st = h->s->status();
    auto const& rel = h->s->rel_residual();
    auto const& sw = h->s->sweeps();
    for (size_t i = 0; i < st.size(); ++i) {
      if (status) status[i] = st[i];
      if (rel_residual) rel_residual[i] = rel[i];
      if (sweeps) sweeps[i] = sw[i];
    }
    if (summary) {
      summary[0] = h->rep.sweeps;
      summary[1] = h->rep.max_rel_residual;
      summary[2] = h->rep.failed;
      summary[3] = h->rep.ms_total;
    }
    return 0;
  });
}

char const* alglin_batch_last_error(alglin_batch_t h) { return h ? h->err.c_str() : "null handle"; }

int alglin_batch_time_factorize(alglin_batch_t h, double* rows, int32_t max_rows) {
  int n = -1;
  const int rc = bguard(h, [&] {
    auto t = h->s->time_factorize();
    n = int(t.size());
    for (int i = 0; i < n && i < max_rows; ++i) {
      double* r = rows + 6 * size_t(i);
      r[0] = t[i].level;
      r[1] = t[i].cls;
      r[2] = t[i].nblocks;
      r[3] = t[i].ms;
      r[4] = t[i].bytes;
      r[5] = t[i].flops;
    }
    return 0;
  });
  return rc == 0 ? n : -1;
}

void* alglin_batch_stream(alglin_batch_t h) { return h ? static_cast<void*>(h->s->stream()) : nullptr; }

int alglin_batch_set_pair_kernel(alglin_batch_t h, int variant) {
  return bguard(h, [&] {
    if (variant != 0 && variant != 4 && variant != 5) throw std::invalid_argument("pair kernel variant: 0, 4 or 5");
    h->s->set_pair_variant(variant);
    return 0;
  });
}

int alglin_batch_set_streams(alglin_batch_t h, int streams, int fork_level) {
  return bguard(h, [&] {
    if (streams != 1 && streams != 2) throw std::invalid_argument("streams: 1 or 2");
    if (fork_level < 0) throw std::invalid_argument("fork_level >= 0");
    h->s->set_streams(streams, fork_level);
    return 0;
  });
}

int alglin_batch_set_class_kernel(alglin_batch_t h, int fp32_factors, int cls, int kernel, int waves) {
  return bguard(h, [&] {
    h->s->set_class_kernel(fp32_factors != 0, cls, kernel, waves);
    return 0;
  });
}

int alglin_batch_set_methods(alglin_batch_t h, int32_t nsys, int32_t const* methods) {
  return bguard(h, [&] {
    h->s->set_methods(std::vector<int32_t>(methods, methods + nsys));
    return 0;
  });
}

// ---- learned selector ----------------------------------------------------------
struct alglin_mlp {
  std::unique_ptr<bcr::Mlp> m;
  std::string err{"no error"};
};

int32_t alglin_mlp_num_params(void) { return bcr::kMlpParams; }

alglin_mlp_t alglin_mlp_create(uint64_t seed) {
  try {
    auto* h = new alglin_mlp;
    h->m.reset(new bcr::Mlp(seed));
    return h;
  } catch (...) {
    return nullptr;
  }
}

void alglin_mlp_destroy(alglin_mlp_t h) { delete h; }

char const* alglin_mlp_last_error(alglin_mlp_t h) { return h ? h->err.c_str() : "null handle"; }

}  // extern "C"

namespace {
template <typename Fn>
int mguard(alglin_mlp_t h, Fn&& fn) {
  if (!h) return 1;
  try {
    return fn();
  } catch (std::exception const& e) {
    h->err = e.what();
    return 1;
  }
}
}  // namespace

extern "C" {

int alglin_mlp_get_params(alglin_mlp_t h, float* params, float* mu, float* sig) {
  return mguard(h, [&] {
    if (params) h->m->get_params(params);
    if (mu && sig) h->m->get_normalization(mu, sig);
    return 0;
  });
}

int alglin_mlp_set_params(alglin_mlp_t h, float const* params, float const* mu, float const* sig) {
  return mguard(h, [&] {
    if (params) h->m->set_params(params);
    if (mu && sig) h->m->set_normalization(mu, sig);
    return 0;
  });
}

int alglin_mlp_load(alglin_mlp_t h, char const* path) {
  return mguard(h, [&] {
    h->m->load(path);
    return 0;
  });
}

int alglin_mlp_save(alglin_mlp_t h, char const* path) {
  return mguard(h, [&] {
    h->m->save(path);
    return 0;
  });
}

int alglin_mlp_loss_grad(alglin_mlp_t h, float const* X, int32_t const* y, int32_t n, float* loss, float* grad) {
  return mguard(h, [&] {
    const float l = h->m->loss_grad(X, y, n, grad);
    if (loss) *loss = l;
    return 0;
  });
}

int alglin_mlp_train(alglin_mlp_t h, float const* X, int32_t const* y, int32_t n, int32_t epochs, float lr,
                     float* losses) {
  return mguard(h, [&] {
    auto hist = h->m->train(X, y, n, epochs, lr);
    if (losses) std::memcpy(losses, hist.data(), hist.size() * sizeof(float));
    return 0;
  });
}

int alglin_mlp_predict(alglin_mlp_t h, float const* X, int32_t n, float* p1) {
  return mguard(h, [&] {
    h->m->predict(X, n, p1);
    return 0;
  });
}

int alglin_batch_solve_selected(alglin_batch_t h, double const* B, double* X, double tol, int32_t max_iter,
                                alglin_mlp_t selector, float threshold, int fallback) {
  return bguard(h, [&] {
    h->rep = h->s->solve_selected(B, X, tol, max_iter, selector ? selector->m.get() : nullptr, threshold,
                                  fallback != 0);
    return h->rep.failed == 0 ? 0 : 2;
  });
}

int alglin_batch_selection(alglin_batch_t h, float* features, int32_t* chose_qrp, int32_t* fell_back) {
  return bguard(h, [&] {
    auto const& f = h->s->features();
    auto const& q = h->s->chose_qrp();
    auto const& b = h->s->fell_back();
    if (features) std::memcpy(features, f.data(), f.size() * sizeof(float));
    for (size_t i = 0; i < q.size(); ++i) {
      if (chose_qrp) chose_qrp[i] = q[i];
      if (fell_back) fell_back[i] = b[i];
    }
    return 0;
  });
}

}  // extern "C"
