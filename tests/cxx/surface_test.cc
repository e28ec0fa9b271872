// A C++ caller of include/alglin_amd/BorderedCR.hh and BABD.hh, built with
// g++ against libalglin_amd.so.  It restates, on this library, the checks of
// the reference's surface tests:
//   host  - src_tests/test18-BorderedCR-coverage.cc:357-811 (every loader and
//           accessor against tagged values, raw and view overloads, zero_*),
//           test18:342-355 and test17:128-152 (partition and info strings,
//           bad select names), test17:55-124 (view loaders), test18:832-918
//           (sparse export/import, Mv, check_matrix, t_solve), dup, the
//           MATLAB script, and the BABD wrapper's selection (test22:320-360);
//   gpu   - test18:920-1059 (the closed-form stable system, every internal x
//           last-block choice, 1 and 3 right-hand sides, tol 5e-9), the
//           partitioned order (num_parallel_block 3) and the BABD wrapper's
//           cyclic-reduction solve (test22:320-370).
// Usage: surface_test host|gpu.  Exit status 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "alglin_amd/BABD.hh"
#include "alglin_amd/BorderedCR.hh"

using alglin_amd::integer;
using BCR = alglin_amd::BorderedCR<double>;
using MatW = BCR::MatW;

static int g_fail = 0, g_pass = 0;
// the reference takes its partition count only with a thread pool; the pool
// is never used here, any non-null pointer selects num_parallel_block
static char g_pool_tag;
static Utils::ThreadPoolBase* const kPool = reinterpret_cast<Utils::ThreadPoolBase*>(&g_pool_tag);

static void expect(bool ok, std::string const& what) {
  if (ok) ++g_pass;
  else { ++g_fail; std::printf("FAIL: %s\n", what.c_str()); }
}

static void expect_throw(std::function<void()> const& f, std::string const& what) {
  bool thrown = false;
  try { f(); } catch (std::exception const&) { thrown = true; }
  expect(thrown, what + " should throw");
}

// storage of rows x cols with `pad` extra rows (ld = rows + pad) filled with
// a sentinel, and a view of the block starting at row `off`
struct Storage {
  std::vector<double> a;
  MatW view;
  Storage(integer rows, integer cols, integer pad, integer off, double tag) : a(size_t(rows + pad) * cols, -999.0) {
    view.setup(a.data() + off, rows, cols, rows + pad);
    for (integer j = 0; j < cols; ++j)
      for (integer i = 0; i < rows; ++i) view(i, j) = tag + 10 * j + i;   // tagged: distinct per entry
  }
};

template <typename Get>
static void same(Get get, MatW const& v, std::string const& what, double scale = 1, MatW const* add1 = nullptr,
                 MatW const* add2 = nullptr) {
  bool ok = true;
  for (integer j = 0; j < v.ncols(); ++j)
    for (integer i = 0; i < v.nrows(); ++i) {
      double ref = scale * v(i, j);
      if (add1) ref += (*add1)(i, j);
      if (add2) ref += (*add2)(i, j);
      if (std::fabs(get(i, j) - ref) > 1e-12) ok = false;
    }
  expect(ok, what);
}

// test18:357-811
static void loaders() {
  const integer N = 3, n = 2, qr = 1, qx = 2, nr = 3, nx = 2, m = n + qr;
  BCR bcr;
  bcr.allocate(N, n, qr, qx, nr, nx);
  bcr.fill_zero();
  bcr.D(0, 0, 0) = 1; bcr.zero_D(); expect(bcr.D(0, 0, 0) == 0, "zero_D");
  bcr.E(0, 0, 0) = 2; bcr.zero_E(); expect(bcr.E(0, 0, 0) == 0, "zero_E");
  bcr.B(0, 0, 0) = 3; bcr.zero_B(); expect(bcr.B(0, 0, 0) == 0, "zero_B");
  bcr.F(0, 0) = 4; bcr.zero_F(); expect(bcr.F(0, 0) == 0, "zero_F");
  bcr.H(0, 0) = 5; bcr.zero_H(); expect(bcr.H(0, 0) == 0, "zero_H");
  bcr.C(0, 0, 0) = 6; bcr.zero_C(); expect(bcr.C(0, 0, 0) == 0, "zero_C");
  bcr.Cq(0, 0) = 7; bcr.zero_Cq(); expect(bcr.Cq(0, 0) == 0, "zero_Cq");
  bcr.D(0, 0, 0) = bcr.E(0, 0, 0) = bcr.B(0, 0, 0) = bcr.F(0, 0) = bcr.H(0, 0) = bcr.C(0, 0, 0) = bcr.Cq(0, 0) = 1;
  bcr.fill_zero();
  expect(bcr.D(0, 0, 0) == 0 && bcr.E(0, 0, 0) == 0 && bcr.B(0, 0, 0) == 0 && bcr.F(0, 0) == 0 &&
             bcr.H(0, 0) == 0 && bcr.C(0, 0, 0) == 0 && bcr.Cq(0, 0) == 0,
         "fill_zero");

  Storage b1(n, nx, 2, 1, 10), b2(n, nx, 1, 0, 20), b3(n, nx, 1, 0, 30);
  bcr.load_B(1, b1.view);
  same([&](integer i, integer j) { return bcr.B(1, i, j); }, b1.view, "load_B(MatW)");
  bcr.add_to_B(1, b2.view.data(), b2.view.ldim());
  same([&](integer i, integer j) { return bcr.B(1, i, j); }, b1.view, "add_to_B(raw)", 1, &b2.view);
  bcr.add_to_B(1, b3.view);
  same([&](integer i, integer j) { return bcr.B(1, i, j); }, b1.view, "add_to_B(MatW)", 1, &b2.view, &b3.view);

  Storage c1(nr, n, 1, 0, 40), c2(nr, n, 2, 1, 50), c3(nr, n, 1, 0, 60);
  bcr.load_C(2, c1.view.data(), c1.view.ldim());
  same([&](integer i, integer j) { return bcr.C(2, i, j); }, c1.view, "load_C(raw)");
  bcr.add_to_C(2, c2.view);
  same([&](integer i, integer j) { return bcr.C(2, i, j); }, c1.view, "add_to_C(MatW)", 1, &c2.view);
  bcr.add_to_C(2, c3.view.data(), c3.view.ldim());
  same([&](integer i, integer j) { return bcr.C(2, i, j); }, c1.view, "add_to_C(raw)", 1, &c2.view, &c3.view);
  Storage cl(nr, n, 0, 0, 45);
  bcr.load_C(1, cl.view);
  same([&](integer i, integer j) { return bcr.C(1, i, j); }, cl.view, "load_C(MatW)");

  Storage d1(n, n, 1, 0, 70), d2(n, n, 2, 1, 80), e1(n, n, 2, 1, 90), e2(n, n, 1, 0, 100);
  bcr.load_D(0, d1.view);
  same([&](integer i, integer j) { return bcr.D(0, i, j); }, d1.view, "load_D(MatW)");
  bcr.load_D(1, d2.view.data(), d2.view.ldim());
  same([&](integer i, integer j) { return bcr.D(1, i, j); }, d2.view, "load_D(raw)");
  bcr.load_E(0, e1.view.data(), e1.view.ldim());
  same([&](integer i, integer j) { return bcr.E(0, i, j); }, e1.view, "load_E(raw)");
  bcr.load_E(1, e2.view);
  same([&](integer i, integer j) { return bcr.E(1, i, j); }, e2.view, "load_E(MatW)");

  Storage de(n, 2 * n, 1, 0, 110), deb(n, 2 * n + nx, 2, 1, 120);
  bcr.load_DE(2, de.view.data(), de.view.ldim());
  same([&](integer i, integer j) { return j < n ? bcr.D(2, i, j) : bcr.E(2, i, j - n); }, de.view, "load_DE");
  bcr.load_DEB(1, deb.view.data(), deb.view.ldim());
  same([&](integer i, integer j) { return j < n ? bcr.D(1, i, j) : j < 2 * n ? bcr.E(1, i, j - n) : bcr.B(1, i, j - 2 * n); },
       deb.view, "load_DEB");

  Storage f1(nr, nx, 2, 1, 130), f2(nr, nx, 1, 0, 140), f3(nr, nx, 1, 0, 150);
  bcr.load_F(f1.view.data(), f1.view.ldim());
  same([&](integer i, integer j) { return bcr.F(i, j); }, f1.view, "load_F(raw)");
  bcr.add_to_F(f2.view);
  same([&](integer i, integer j) { return bcr.F(i, j); }, f1.view, "add_to_F(MatW)", 1, &f2.view);
  bcr.add_to_F(f3.view.data(), f3.view.ldim());
  same([&](integer i, integer j) { return bcr.F(i, j); }, f1.view, "add_to_F(raw)", 1, &f2.view, &f3.view);
  Storage fl(nr, nx, 0, 0, 135);
  bcr.load_F(fl.view);
  same([&](integer i, integer j) { return bcr.F(i, j); }, fl.view, "load_F(MatW)");

  Storage q1(nr, qx, 1, 0, 160), q2(nr, qx, 2, 1, 170), qf(nr, qx + nx, 2, 1, 180);
  bcr.load_Cq(q1.view);
  same([&](integer i, integer j) { return bcr.Cq(i, j); }, q1.view, "load_Cq(MatW)");
  bcr.load_Cq(q2.view.data(), q2.view.ldim());
  same([&](integer i, integer j) { return bcr.Cq(i, j); }, q2.view, "load_Cq(raw)");
  bcr.load_CqF(qf.view.data(), qf.view.ldim());
  same([&](integer i, integer j) { return j < qx ? bcr.Cq(i, j) : bcr.F(i, j - qx); }, qf.view, "load_CqF");

  Storage h0(m, n, 1, 0, 190), hn(m, n, 2, 1, 200), hq(m, qx, 1, 0, 210), hp(m, nx, 1, 0, 220);
  bcr.load_bottom(h0.view.data(), h0.view.ldim(), hn.view.data(), hn.view.ldim(), hq.view.data(), hq.view.ldim(),
                  hp.view.data(), hp.view.ldim());
  auto Hcol = [&](integer i, integer j, integer c0) { return bcr.H(i, c0 + j); };
  same([&](integer i, integer j) { return Hcol(i, j, 0); }, h0.view, "load_bottom/H0");
  same([&](integer i, integer j) { return Hcol(i, j, n); }, hn.view, "load_bottom/HN");
  same([&](integer i, integer j) { return Hcol(i, j, 2 * n); }, hq.view, "load_bottom/Hq");
  same([&](integer i, integer j) { return Hcol(i, j, 2 * n + qx); }, hp.view, "load_bottom/Hp");
  Storage h0b(m, n, 0, 0, 230), hnb(m, n, 0, 0, 240), hqb(m, qx, 0, 0, 250), hpb(m, nx, 0, 0, 260);
  bcr.load_bottom(h0b.view, hnb.view, hqb.view, hpb.view);
  same([&](integer i, integer j) { return Hcol(i, j, 2 * n + qx); }, hpb.view, "load_bottom(MatW)/Hp");
  Storage hall(m, bcr.Nc(), 1, 1, 270);
  bcr.load_bottom(hall.view);
  same([&](integer i, integer j) { return bcr.H(i, j); }, hall.view, "load_bottom(H MatW)");

  Storage c0(nr, n, 1, 0, 280), cN(nr, n, 0, 0, 290), cq(nr, qx, 2, 1, 300), ff(nr, nx, 1, 1, 310);
  bcr.load_bottom2(c0.view.data(), c0.view.ldim(), cN.view.data(), cN.view.ldim(), cq.view.data(), cq.view.ldim(),
                   ff.view.data(), ff.view.ldim());
  same([&](integer i, integer j) { return bcr.C(0, i, j); }, c0.view, "load_bottom2/C0");
  same([&](integer i, integer j) { return bcr.C(N, i, j); }, cN.view, "load_bottom2/CN");
  same([&](integer i, integer j) { return bcr.Cq(i, j); }, cq.view, "load_bottom2/Cq");
  same([&](integer i, integer j) { return bcr.F(i, j); }, ff.view, "load_bottom2/F");

  // block views write through to the storage (hxx:1002-1014)
  MatW w;
  bcr.D(2, w);
  w(1, 0) = 4242;
  expect(bcr.D(2, 1, 0) == 4242 && w.nrows() == n && w.ncols() == n, "D(k, MatW&) view");
  bcr.H(w);
  expect(w.nrows() == m && w.ncols() == bcr.Nc() && &w(0, 0) == &bcr.H(0, 0), "H(MatW&) view");

  // dimension and leading-dimension errors
  Storage bad(n + 1, nx, 0, 0, 1);
  expect_throw([&] { bcr.load_B(0, bad.view); }, "load_B with a (n+1) x nx view");
  expect_throw([&] { bcr.load_D(0, d1.view.data(), n - 1); }, "load_D with ld < n");
  expect_throw([&] { bcr.load_C(0, c1.view.data(), nr - 1); }, "load_C with ld < nr");
  expect_throw([&] { bcr.load_bottom(hn.view); }, "load_bottom with an m x n view");
}

// test17:55-124: bottom rows from one view, add_to_C2F from a view
static void views() {
  const integer N = 2, n = 2, qr = 1, qx = 1, nr = 2, nx = 1;
  BCR bcr;
  bcr.allocate(N, n, qr, qx, nr, nx);
  bcr.fill_zero();
  Storage bot(nr, bcr.Nc(), 2, 1, 100);
  bcr.load_bottom2(bot.view);
  same([&](integer i, integer j) {
         return j < n ? bcr.C(0, i, j) : j < 2 * n ? bcr.C(N, i, j - n) : j < 2 * n + qx ? bcr.Cq(i, j - 2 * n)
                                                                                           : bcr.F(i, j - 2 * n - qx);
       },
       bot.view, "load_bottom2(MatW)");
  bcr.fill_zero();
  Storage c2f(nr, 2 * n + nx, 1, 0, 200);
  bcr.add_to_C2F(1, c2f.view);
  same([&](integer i, integer j) { return j < n ? bcr.C(1, i, j) : j < 2 * n ? bcr.C(2, i, j - n) : bcr.F(i, j - 2 * n); },
       c2f.view, "add_to_C2F(MatW)");
  Storage c2(nr, 2 * n, 0, 0, 300);
  bcr.add_to_C2(0, c2.view.data(), c2.view.ldim());
  same([&](integer i, integer j) { return j < n ? bcr.C(0, i, j) : bcr.C(1, i, j - n) - c2f.view(i, j - n); }, c2.view,
       "add_to_C2(raw)");
}

// test17:128-152, test18:342-355
static void info_and_selection() {
  BCR a(kPool, 1);
  a.allocate(16, 2, 0, 0, 0, 0);
  expect(a.info().find("nblk:1") != std::string::npos, "info nblk:1");
  a.set_num_parallel_block(4);
  expect(a.info().find("nblk:4") != std::string::npos, "info nblk:4 after set_num_parallel_block(4)");
  a.set_num_parallel_block(2);
  expect(a.info().find("nblk:2") != std::string::npos, "info nblk:2 after set_num_parallel_block(2)");
  a.select_last_LU();
  a.select_last2_PINV();
  expect(a.info().find("last   = LastBlock LU") != std::string::npos, "info last");
  expect(a.info().find("last2  = LastBlock PINV") != std::string::npos, "info last2");

  BCR b;
  b.allocate(6, 2, 1, 2, 3, 2);
  b.select_QRP();
  b.select_last("PINV");
  b.set_num_parallel_block(4);   // 2P <= N shrinks 4 to 3
  std::string const s = b.info("  ");
  expect(s.find("nblk:3") != std::string::npos, "info nblk:3 (N = 6, 4 requested)");
  expect(s.find("CyclicReduction+QRP") != std::string::npos, "info internal solver");
  expect(s.find("LastBlock PINV") != std::string::npos, "info last solver");
  std::ostringstream o;
  b.info(o);
  expect(o.str().find("rows") != std::string::npos, "info(stream)");
  expect_throw([&] { b.select("BAD"); }, "select(BAD)");
  expect_throw([&] { b.select_last("BAD"); }, "select_last(BAD)");
  expect_throw([&] { b.select_last2("BAD"); }, "select_last2(BAD)");
  expect(b.info_algo() == "CyclicReduction+QRP and LastBlock PINV and LastBlock NONE", "info_algo");
}

// test18:163-230: the stable 6-block system
static void load_stable(BCR& s) {
  const integer N = 6, n = 2, qr = 1, qx = 2, nr = 3, nx = 2, m = n + qr;
  s.allocate(N, n, qr, qx, nr, nx);
  s.fill_zero();
  for (integer k = 0; k < N; ++k) {
    s.D(k, 0, 0) = 5.0 + 0.2 * k; s.D(k, 0, 1) = -0.08; s.D(k, 1, 0) = 0.15; s.D(k, 1, 1) = 4.7 + 0.2 * k;
    s.E(k, 0, 0) = -0.35; s.E(k, 0, 1) = 0.06; s.E(k, 1, 0) = 0.04; s.E(k, 1, 1) = -0.28;
    for (integer j = 0; j < nx; ++j)
      for (integer i = 0; i < n; ++i) s.B(k, i, j) = 0.04 * (1 + k + i + 2 * j);
    for (integer j = 0; j < n; ++j)
      for (integer i = 0; i < nr; ++i) s.C(k, i, j) = 0.03 * (1 + 2 * k + i + j);
  }
  for (integer j = 0; j < n; ++j)
    for (integer i = 0; i < nr; ++i) s.C(N, i, j) = 0.05 * (1 + i + j);
  for (integer j = 0; j < qx; ++j)
    for (integer i = 0; i < nr; ++i) s.Cq(i, j) = 0.07 * (1 + i + 2 * j);
  for (integer j = 0; j < nx; ++j)
    for (integer i = 0; i < nr; ++i) s.F(i, j) = 0.02 * (1 + i + j);
  s.F(0, 0) += 3.2;
  s.F(1, 1) += 2.8;
  for (integer j = 0; j < n; ++j)
    for (integer i = 0; i < m; ++i) {
      s.H(i, j) = 0.02 * (1 + i + j);
      s.H(i, n + j) = 0.03 * (1 + i + j);
    }
  s.H(0, n) += 5.1;
  s.H(1, n + 1) += 4.9;
  s.H(2, n) += 0.2;
  s.H(2, n + 1) -= 0.25;
  for (integer j = 0; j < qx; ++j)
    for (integer i = 0; i < m; ++i) s.H(i, 2 * n + j) = 0.04 * (1 + i + j);
  for (integer j = 0; j < nx; ++j)
    for (integer i = 0; i < m; ++i) s.H(i, 2 * n + qx + j) = 0.05 * (1 + i + j);
}

// test18:832-918 plus dup and the MATLAB script
static void sparse_and_errors() {
  BCR s;
  load_stable(s);
  const integer nnz = s.sparse_nnz(), nr = s.nrows(), nc = s.ncols();
  std::vector<integer> I(static_cast<size_t>(nnz)), J(static_cast<size_t>(nnz));
  std::vector<double> V(static_cast<size_t>(nnz));
  s.sparse_pattern(I.data(), J.data(), 0);
  s.sparse_values(V.data());
  // Mv equals the COO product
  std::vector<double> x(static_cast<size_t>(nc)), y(static_cast<size_t>(nr)), z(static_cast<size_t>(nr), 0.0);
  for (integer i = 0; i < nc; ++i) x[size_t(i)] = 0.25 + 0.5 * i;
  s.Mv(x.data(), y.data());
  for (integer k = 0; k < nnz; ++k) z[size_t(I[size_t(k)])] += V[size_t(k)] * x[size_t(J[size_t(k)])];
  double d = 0;
  for (integer i = 0; i < nr; ++i) d = std::max(d, std::fabs(y[size_t(i)] - z[size_t(i)]));
  expect(d < 1e-13, "Mv == COO product");
  // add_Mv accumulates
  std::vector<double> y2 = y;
  s.add_Mv(-1.0, x.data(), y2.data());
  double d2 = 0;
  for (double v : y2) d2 = std::max(d2, std::fabs(v));
  expect(d2 < 1e-13, "add_Mv(-1) cancels Mv");
  // round trip through sparse_load (1-based indices)
  BCR t;
  t.allocate(s.number_of_blocks(), s.block_size(), s.dim_qr(), s.dim_qx(), s.dim_nr(), s.dim_nx());
  std::vector<integer> I1(I), J1(J);
  for (auto& v : I1) ++v;
  for (auto& v : J1) ++v;
  t.sparse_load(V.data(), I1.data(), 1, J1.data(), 1, nnz);
  std::vector<double> V2(static_cast<size_t>(nnz));
  t.sparse_values(V2.data());
  expect(V2 == V, "sparse_load(sparse_values) round trip");
  // an entry outside the pattern: recorded, thrown on request
  integer bi = 0, bj = 3 * s.block_size();   // row 0 holds D_0, E_0 and B_0 only: column 3n is off-pattern
  double bv = 1;
  t.sparse_load(&bv, &bi, 0, &bj, 0, 1, false);
  expect(std::string(t.last_error()).find("out of pattern") != std::string::npos, "sparse_load records an off-pattern entry");
  expect_throw([&] { t.sparse_load(&bv, &bi, 0, &bj, 0, 1, true); }, "sparse_load(do_error) off-pattern");
  // dup copies the loaded blocks
  BCR u;
  u.dup(s);
  std::vector<double> V3(static_cast<size_t>(nnz));
  u.sparse_values(V3.data());
  expect(V3 == V && u.nrows() == nr, "dup");
  // check_matrix and t_solve
  s.check_matrix();
  ++g_pass;
  s.D(3, 1, 0) = std::nan("");
  expect_throw([&] { s.check_matrix(); }, "check_matrix on a NaN");
  expect(std::string(s.last_error()).find("matrix D") != std::string::npos, "check_matrix names the block");
  s.set_debug(true);
  expect(!s.factorize() && std::string(s.last_error()).find("check_matrix") != std::string::npos,
         "factorize with debug rejects a NaN");
  expect_throw([&] { s.t_solve(x.data()); }, "t_solve");
  expect_throw([&] { s.t_solve(1, x.data(), nc); }, "t_solve(nrhs)");
  expect(!s.solve(x.data()) && std::string(s.last_error()).find("not factorized") != std::string::npos,
         "solve before factorize fails");
  // MATLAB script (cc:2493-2534)
  std::ostringstream o;
  u.print_matlab_script(o);
  std::string const ms = o.str();
  expect(ms.rfind("I = [ 1", 0) == 0, "matlab script starts with I");
  expect(ms.find("MAT = sparse( I, J, V, nr, nc );") != std::string::npos, "matlab script sparse()");
  expect(ms.find("nnz = " + std::to_string(nnz) + ";") != std::string::npos, "matlab script nnz");
  expect(ms.find("nr  = " + std::to_string(nr) + ";") != std::string::npos, "matlab script nr");
}

// test22:320-360 (selection only: no GPU)
static void babd_wrapper_host() {
  alglin_amd::BABD<double> w;
  using C = alglin_amd::BABD<double>::BABD_Choice;
  expect(w.using_cyclic_reduction() && !w.using_diaz(), "BABD default backend is cyclic reduction");
  expect(w.selected_solver() == C::CYCLIC_REDUCTION_LU, "BABD default choice");
  expect_throw([&] { (void)w.diaz(); }, "BABD::diaz on the CR backend");
  w.selectSolver(C::CYCLIC_REDUCTION_QRP);
  expect(w.selected_solver() == C::CYCLIC_REDUCTION_QRP && w.cr().selected() == BCR::BORDERED_Choice::QRP,
         "BABD selectSolver(QRP)");
  expect(w.info().find("CyclicReduction+QRP") != std::string::npos, "BABD info mentions QRP");
  expect_throw([&] { w.selectSolver(C::DIAZ); }, "BABD selectSolver(DIAZ)");
  expect(w.selected_solver() == C::CYCLIC_REDUCTION_QRP, "a refused choice leaves the selection");
}

// ---- GPU ---------------------------------------------------------------------------

static double max_err(std::vector<double> const& a, std::vector<double> const& b, integer n) {
  double d = 0;
  for (integer i = 0; i < n; ++i) d = std::max(d, std::fabs(a[size_t(i)] - b[size_t(i)]));
  return d;
}

// test18:920-1059: x_i = 0.25 + 0.5 i, every internal x last-block choice
static void stable_closed_forms() {
  char const* internal[] = {"LU", "QR", "QRP"};
  char const* last[] = {"LU", "LUPQ", "QR", "QRP", "SVD", "LSS", "LSY", "PINV"};
  for (integer npb : {1, 3}) {
    for (char const* in : internal)
      for (char const* la : last) {
        BCR s(kPool, npb);
        load_stable(s);
        s.select(in);
        s.select_last(la);
        const integer nr = s.nrows(), nc = s.ncols(), L = std::max(nr, nc);
        std::vector<double> xref(static_cast<size_t>(nc)), b(static_cast<size_t>(L), 0.0);
        for (integer i = 0; i < nc; ++i) xref[size_t(i)] = 0.25 + 0.5 * i;
        s.Mv(xref.data(), b.data());
        std::string const tag = std::string(in) + "/" + la + " npb " + std::to_string(npb);
        const bool okf = s.factorize();
        expect(okf, "factorize " + tag + ": " + std::string(s.last_error()));
        if (!okf) continue;
        std::vector<double> x = b;
        expect(s.solve(x.data()) && max_err(x, xref, nc) < 5e-9, "solve " + tag);
        std::vector<double> r(static_cast<size_t>(nr));
        s.Mv(x.data(), r.data());
        expect(max_err(r, b, nr) < 5e-9, "residual " + tag);
        // three right-hand sides: X_c = (1 + c) xref - 0.1 c
        std::vector<double> X(size_t(L) * 3, 0.0), Xr(size_t(nc) * 3);
        for (integer c = 0; c < 3; ++c) {
          std::vector<double> xc(static_cast<size_t>(nc));
          for (integer i = 0; i < nc; ++i) xc[size_t(i)] = Xr[size_t(c) * nc + i] = (1 + c) * xref[size_t(i)] - 0.1 * c;
          s.Mv(xc.data(), X.data() + size_t(c) * L);
        }
        bool ok3 = s.solve(3, X.data(), L);
        double e3 = 0;
        for (integer c = 0; c < 3; ++c)
          for (integer i = 0; i < nc; ++i) e3 = std::max(e3, std::fabs(X[size_t(c) * L + i] - Xr[size_t(c) * nc + i]));
        expect(ok3 && e3 < 5e-9, "solve 3 rhs " + tag);
      }
  }
}

// test22:320-370: the wrapper's cyclic-reduction path with QRP
static void babd_wrapper_gpu() {
  alglin_amd::BABD<double> w;
  load_stable(w.cr());
  const integer N = w.ncols();
  std::vector<double> xref(static_cast<size_t>(N)), rhs(static_cast<size_t>(N));
  for (integer i = 0; i < N; ++i) xref[size_t(i)] = 1.0 + 0.2 * i;
  w.cr().Mv(xref.data(), rhs.data());
  w.selectSolver(alglin_amd::BABD<double>::BABD_Choice::CYCLIC_REDUCTION_QRP);
  expect(w.factorize_system(), "BABD factorize_system");
  std::vector<double> x = rhs;
  expect(w.solve_system(x.data()) && max_err(x, xref, N) < 5e-9, "BABD solve_system");
  std::vector<double> X(size_t(N) * 2);
  std::copy(rhs.begin(), rhs.end(), X.begin());
  std::copy(rhs.begin(), rhs.end(), X.begin() + N);
  expect(w.solve(2, X.data(), N) && max_err(std::vector<double>(X.begin() + N, X.end()), xref, N) < 5e-9,
         "BABD solve(nrhs)");
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "host";
  try {
    if (mode == "host") {
      loaders();
      views();
      info_and_selection();
      sparse_and_errors();
      babd_wrapper_host();
    } else if (mode == "gpu") {
      stable_closed_forms();
      babd_wrapper_gpu();
    } else {
      std::printf("usage: surface_test host|gpu\n");
      return 2;
    }
  } catch (std::exception const& e) {
    std::printf("FAIL: uncaught exception: %s\n", e.what());
    ++g_fail;
  }
  std::printf("%s: %d checks passed, %d failed\n", mode.c_str(), g_pass, g_fail);
  return g_fail == 0 ? 0 : 1;
}
