// ORACLE — test infrastructure only. Nothing in the product links this file.
//
// C entry points over oracle::BorderedCR<double|float> so that tests/ (via
// ctypes), __graft_entry__.smoke() and bench.py's CPU-baseline leg can drive
// the CPU restatement.  A handle carries its element type; every data
// pointer is to that type.
#include <cstdio>
#include <cstring>
#include <memory>

#include "bcr_oracle.hh"

namespace {

struct Handle {
  int dtype;  // 0 = double, 1 = float
  std::unique_ptr<oracle::BorderedCR<double>> d;
  std::unique_ptr<oracle::BorderedCR<float>> f;
  std::string info;
};

template <class Fn>
auto dispatch(void* h, Fn&& fn) {
  Handle* H = static_cast<Handle*>(h);
  if (H->dtype == 0) return fn(*H->d, (double*)nullptr);
  return fn(*H->f, (float*)nullptr);
}

}  // namespace

extern "C" {

using oracle::integer;

void* orc_create(int dtype, int nthreads, int npb) {
  Handle* H = new Handle;
  H->dtype = dtype;
  if (dtype == 0) H->d.reset(new oracle::BorderedCR<double>(nthreads, npb));
  else            H->f.reset(new oracle::BorderedCR<float>(nthreads, npb));
  return H;
}

void orc_destroy(void* h) { delete static_cast<Handle*>(h); }

void orc_allocate(void* h, integer N, integer n, integer qr, integer qx, integer nr, integer nx) {
  dispatch(h, [&](auto& S, auto*) { S.allocate(N, n, qr, qx, nr, nx); return 0; });
}

void orc_set_num_parallel_block(void* h, integer npb) {
  dispatch(h, [&](auto& S, auto*) { S.set_num_parallel_block(npb); return 0; });
}

void orc_set_threads(void* h, int fac, int sol) {
  dispatch(h, [&](auto& S, auto*) { S.set_factorize_use_thread(fac != 0); S.set_solve_use_thread(sol != 0); return 0; });
}

void orc_select(void* h, int internal, int last, int last2) {
  dispatch(h, [&](auto& S, auto*) {
    using BC = std::remove_reference_t<decltype(S)>;
    S.select(typename BC::Choice(internal));
    S.select_last(typename BC::LastChoice(last));
    S.select_last2(typename BC::LastChoice2(last2));
    return 0;
  });
}

// Packed loads in the reference's internal block layout (hxx:972-998).
// Any pointer may be null (block left unchanged).
void orc_load(void* h, void const* D, void const* E, void const* B, void const* C, void const* H,
              void const* Cq, void const* F) {
  dispatch(h, [&](auto& S, auto* tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    auto cp = [](void const* src, T* dst, size_t cnt) { if (src && cnt) std::memcpy(dst, src, cnt * sizeof(T)); };
    integer N = 0, n = 0;
    (void)N; (void)n;
    size_t const nD = S.Dsize(), nB = S.Bsize(), nC = S.Csize(), nH = S.Hsize(), nCq = S.Cqsize(), nF = S.Fsize();
    cp(D, S.Dptr(), nD);
    cp(E, S.Eptr(), nD);
    cp(B, S.Bptr(), nB);
    cp(C, S.Cptr(), nC);
    cp(H, S.Hptr(), nH);
    cp(Cq, S.Cqptr(), nCq);
    cp(F, S.Fptr(), nF);
    return 0;
  });
}

// Read back the (possibly factorized) internal arrays.
void orc_get(void* h, void* D, void* E, void* B, void* C, void* F, void* Tm, integer* perm, void* Hmat) {
  dispatch(h, [&](auto& S, auto* tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    auto cp = [](T const* src, void* dst, size_t cnt) { if (dst && cnt) std::memcpy(dst, src, cnt * sizeof(T)); };
    cp(S.Dptr(), D, S.Dsize());
    cp(S.Eptr(), E, S.Dsize());
    cp(S.Bptr(), B, S.Bsize());
    cp(S.Cptr(), C, S.Csize());
    cp(S.Fptr(), F, S.Fsize());
    cp(S.Tptr(), Tm, S.Tsize());
    if (perm) std::memcpy(perm, S.Permptr(), S.Permsize() * sizeof(integer));
    cp(S.Hmatptr(), Hmat, S.Hmatsize());
    return 0;
  });
}

int orc_factorize(void* h) {
  return dispatch(h, [&](auto& S, auto*) { return S.factorize() ? 0 : 1; });
}

int orc_solve(void* h, void* x) {
  return dispatch(h, [&](auto& S, auto* tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    return S.solve(static_cast<T*>(x)) ? 0 : 1;
  });
}

int orc_solve_n(void* h, integer nrhs, void* X, integer ld) {
  return dispatch(h, [&](auto& S, auto* tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    return S.solve(nrhs, static_cast<T*>(X), ld) ? 0 : 1;
  });
}

void orc_mv(void* h, void const* x, void* y) {
  dispatch(h, [&](auto& S, auto* tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    S.Mv(static_cast<T const*>(x), static_cast<T*>(y));
    return 0;
  });
}

integer orc_sparse_nnz(void* h) {
  return dispatch(h, [&](auto& S, auto*) { return S.sparse_nnz(); });
}

void orc_sparse(void* h, integer* I, integer* J, void* V, integer offs) {
  dispatch(h, [&](auto& S, auto* tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    S.sparse(I, J, static_cast<T*>(V), offs);
    return 0;
  });
}

integer orc_dims(void* h, integer which) {
  return dispatch(h, [&](auto& S, auto*) {
    switch (which) {
      case 0: return S.nrows();
      case 1: return S.ncols();
      case 2: return S.Nr();
      case 3: return S.Nc();
      case 4: return S.used_parallel_block();
      case 5: return S.kblock0();
    }
    return integer(-1);
  });
}

char const* orc_info(void* h) {
  Handle* H = static_cast<Handle*>(h);
  H->info = dispatch(h, [&](auto& S, auto*) { return S.info(); });
  return H->info.c_str();
}

char const* orc_last_error(void* h) {
  return dispatch(h, [&](auto& S, auto*) { return S.last_error().c_str(); });
}

}  // extern "C"
