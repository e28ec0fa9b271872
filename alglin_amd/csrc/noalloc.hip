// Allocation-free factor + solve (alglin_bcr_factor_solve_noalloc).
//
// The working arrays of one FP64 cyclic reduction (LU eliminations, LU last
// block, pure cyclic-reduction order) are slices of `scratch`, a fixed
// device array of the library: no hipMalloc / hipFree, no allocating library
// call.  Each GPU the library runs on holds its own copy of the array (64 MiB)
// until the process ends.  The host name `scratch` is not a device address:
// its address on the current device comes from hipGetSymbolAddress, taken
// once per device and kept in g_scratch_ptr; only that pointer reaches the
// kernels.  Every call on one GPU uses the same array, so calls that overlap
// on one GPU race; the factorization clears what must start at zero (status
// word, pending border terms) with fill kernels at its start.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/alglin_amd.h"
#include "bcr_engine.hh"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// the array's address on the current device (set up once per device)
char* scratch_on_current_device(int* dev_out) {
  int dev = 0;
  bcr::hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw bcr::HipError("device index out of range");
  *dev_out = dev;
  std::lock_guard<std::mutex> g(g_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    bcr::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<char*>(p);
  }
  return g_scratch_ptr[dev];
}

void put_error(char* err, int32_t len, std::string const& msg) {
  if (err && len > 0) {
    const size_t k = std::min<size_t>(msg.size(), size_t(len) - 1);
    std::memcpy(err, msg.data(), k);
    err[k] = 0;
  }
}

}  // namespace

extern "C" int alglin_bcr_factor_solve_noalloc(int32_t const dims[6], double const* D, double const* E,
                                               double const* B, double const* C, double const* H, double const* Cq,
                                               double const* F, double* X, int32_t nrhs, int64_t ld, void* stream,
                                               int sync, char* err, int32_t errlen) {
  put_error(err, errlen, "");
  try {
    if (!dims || !X) throw std::invalid_argument("null dims or X");
    bcr::Dims d;
    d.N = dims[0]; d.n = dims[1]; d.qr = dims[2]; d.qx = dims[3]; d.nr = dims[4]; d.nx = dims[5];
    if (d.N < 1 || d.n < 1 || d.qr < 0 || d.qx < 0 || d.nr < 0 || d.nx < 0)
      throw std::invalid_argument("bad dimensions");
    if (d.Nr() != d.Nc()) throw std::invalid_argument("the LU last block needs qr + nr == qx + nx");
    const int64_t L = std::max(d.nrows(), d.ncols());
    if (nrhs < 1 || (nrhs > 1 && ld < L)) throw std::invalid_argument("nrhs >= 1 and ld >= max(nrows, ncols) needed");
    if (d.n > 64 || d.nr > 16) throw std::invalid_argument("block size n <= 64 and nr <= 16 are supported");
    int dev = 0;
    char* base = scratch_on_current_device(&dev);
    bcr::Engine<double, double> e(static_cast<hipStream_t>(stream));
    const size_t need = bcr::Engine<double, double>::bytes_in(d, nrhs, e.max_chunks());
    if (need > kScratchBytes) {
      put_error(err, errlen,
                "alglin_bcr_factor_solve_noalloc: the working set of this system (" + std::to_string(need) +
                    " bytes) exceeds the fixed scratch array (" + std::to_string(kScratchBytes) + " bytes)");
      return 3;
    }
    e.allocate_in(d, nrhs, base, kScratchBytes);
    e.set_method(bcr::kLU);
    e.set_last_method(bcr::kLastLU);
    e.bind_device(D, E, B, C, H, Cq, F);
    e.factorize_async();
    e.solve_async<double>(X, nrhs, nrhs > 1 ? ld : L);
    if (!sync) return 0;
    const int32_t st = e.status_sync() & 3;
    if (st & 1) put_error(err, errlen, "zero pivot in a cyclic-reduction elimination");
    else if (st & 2) put_error(err, errlen, "singular last block");
    return st & 1 ? 1 : (st & 2 ? 2 : 0);
  } catch (std::exception const& ex) {
    put_error(err, errlen, std::string("alglin_bcr_factor_solve_noalloc: ") + ex.what());
    return -1;
  }
}
