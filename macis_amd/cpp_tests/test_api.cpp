// Tests of the C++ mirror of the MACIS API (macis_amd/include), written the
// way the reference's Catch2 tests use that API (tests/double_loop.cxx,
// csr_hamiltonian.cxx, davidson.cxx, asci.cxx, sd_operations.cxx,
// src/sparsexx/tests/test_spmbv.cxx), with tighter tolerances.
//
// usage: test_api <golden dir> [--cpu-only]
#include <algorithm>
#include <cmath>
#include <iterator>
#include <cstdio>
#include <fstream>
#include <functional>
#include <string>
#include <vector>

#include <macis/asci/grow.hpp>
#include <macis/asci/learned_prescreen.hpp>
#include <macis/asci/refine.hpp>
#include <macis/csr_hamiltonian.hpp>
#include <macis/hip/hip_hamiltonian_generator.hpp>
#include <macis/sd_operations.hpp>
#include <macis/solvers/davidson.hpp>
#include <macis/solvers/selected_ci_diag.hpp>
#include <macis/util/detail/rdm_files.hpp>
#include <macis/util/fcidump.hpp>
#include <macis/util/transform.hpp>
#include <macis/wavefunction_io.hpp>
#include <sparsexx/spblas/spmbv.hpp>
#include <sparsexx/util/submatrix.hpp>

static int g_fail = 0, g_checks = 0;
#define REQUIRE(cond)                                                   \
  do {                                                                  \
    ++g_checks;                                                         \
    if(!(cond)) {                                                       \
      ++g_fail;                                                         \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
    }                                                                   \
  } while(0)
#define REQUIRE_NEAR(a, b, tol)                                                           \
  do {                                                                                    \
    ++g_checks;                                                                           \
    const double _a = (a), _b = (b);                                                      \
    if(!(std::abs(_a - _b) <= (tol))) {                                                   \
      ++g_fail;                                                                           \
      std::printf("  FAILED %s:%d: %s = %.15g vs %.15g\n", __FILE__, __LINE__, #a, _a, _b); \
    }                                                                                     \
  } while(0)

struct H2O {
  size_t norb;
  double E_core;
  std::vector<double> T, V;
};

static H2O load_h2o(const std::string& dir) {
  const std::string f = dir + "/h2o.ccpvdz.fci.dat";
  H2O h;
  h.norb = macis::read_fcidump_norb(f);
  h.E_core = macis::read_fcidump_core(f);
  h.T.assign(h.norb * h.norb, 0.);
  h.V.assign(h.norb * h.norb * h.norb * h.norb, 0.);
  macis::read_fcidump_1body(f, h.T.data(), h.norb);
  macis::read_fcidump_2body(f, h.V.data(), h.norb);
  return h;
}

// ---------------------------------------------------------------- CPU cases
static void test_sd_operations() {
  // sd_operations.cxx:217-237
  auto combs = macis::generate_combs<64>(4, 2);
  std::vector<unsigned long long> ref = {0x3, 0x5, 0x9, 0x6, 0xA, 0xC};
  REQUIRE(combs.size() == ref.size());
  for(size_t i = 0; i < ref.size(); ++i) REQUIRE(combs[i].to_ullong() == ref[i]);
  auto dets = macis::generate_hilbert_space<64>(4, 2, 2);
  REQUIRE(dets.size() == 36);
  REQUIRE(dets[0].to_ullong() == 0x300000003ull);
  REQUIRE(dets[1].to_ullong() == 0x500000003ull);
  REQUIRE(dets[35].to_ullong() == 0xC0000000Cull);
  auto hf = macis::canonical_hf_determinant<64>(4, 2);
  REQUIRE(hf.to_ullong() == 0x30000000Full);
  auto h128 = macis::canonical_hf_determinant<128>(3, 2);
  REQUIRE(macis::bitset_lo_word(h128).to_ullong() == 0x7ull);
  REQUIRE(macis::bitset_hi_word(h128).to_ullong() == 0x3ull);
  std::bitset<64> s(0xAFull);
  std::vector<uint32_t> occ, vir;
  macis::bitset_to_occ_vir(12, s, occ, vir);  // sd_operations.cxx:32-40
  REQUIRE((occ == std::vector<uint32_t>{0, 1, 2, 3, 5, 7}));
  REQUIRE((vir == std::vector<uint32_t>{4, 6, 8, 9, 10, 11}));
  REQUIRE(macis::bitset_less(std::bitset<64>(0x100000000ull), std::bitset<64>(0x200000000ull)));
  REQUIRE(!macis::bitset_less(std::bitset<64>(0x100000001ull), std::bitset<64>(0x000000002ull)));
  auto cisd = macis::generate_cisd_hilbert_space<64>(8, macis::canonical_hf_determinant<64>(4, 4));
  REQUIRE(cisd.size() == 1 + 2 * 16 + 2 * 36 + 256);
}

static void test_host_generator_elements(const H2O& h) {
  // The base-class matrix_element (double_loop.cxx:44-53), through a
  // generator that is never asked for a CSR.
  struct NoCsr : macis::HamiltonianGenerator<64> {
    using HamiltonianGenerator::HamiltonianGenerator;
    sparse_matrix_type<int32_t> make_csr_hamiltonian_block_32bit_(full_det_iterator, full_det_iterator,
                                                                  full_det_iterator, full_det_iterator,
                                                                  double) override {
      throw std::runtime_error("unused");
    }
    sparse_matrix_type<int64_t> make_csr_hamiltonian_block_64bit_(full_det_iterator, full_det_iterator,
                                                                  full_det_iterator, full_det_iterator,
                                                                  double) override {
      throw std::runtime_error("unused");
    }
    void form_rdms(full_det_iterator, full_det_iterator, full_det_iterator, full_det_iterator,
                   double*, matrix_span_t, rank4_span_t) override {}
  };
  auto T = h.T;
  auto V = h.V;
  NoCsr g(macis::matrix_span<double>(T.data(), h.norb, h.norb),
          macis::rank4_span<double>(V.data(), h.norb, h.norb, h.norb, h.norb));
  auto hf = macis::canonical_hf_determinant<64>(5, 5);
  REQUIRE_NEAR(g.matrix_element(hf, hf) + h.E_core, -76.0267803489191, 1e-10);
  auto s = hf;
  s.flip(0).flip(5);
  REQUIRE_NEAR(g.matrix_element(s, s), -6.488097259228e+01, 1e-10);
  bool threw = false;
  try {
    std::vector<macis::wfn_t<64>> d{hf};
    g.make_csr_hamiltonian_block<int16_t>(d.begin(), d.end(), d.begin(), d.end(), 0.);
  } catch(const std::runtime_error& e) {
    threw = std::string(e.what()) == "Unsupported index_t";
  }
  REQUIRE(threw);
}

static void test_file_formats(const std::string& dir) {
  // sd_operations.cxx:240-260: canonical strings
  auto state = macis::canonical_hf_determinant<64>(2, 2).flip(3).flip(4 + 32);
  REQUIRE(macis::to_canonical_string(state) == "220ud000000000000000000000000000");
  REQUIRE(macis::from_canonical_string<64>("220ud000000000000000000000000000") == state);
  REQUIRE(macis::from_canonical_string<64>("220ud") == state);
  // wavefunction files: write then read gives the same determinants and
  // coefficients bit for bit (17 significant digits)
  auto dets = macis::generate_hilbert_space<128>(6, 3, 2);
  std::vector<double> C(dets.size());
  for(size_t i = 0; i < C.size(); ++i) C[i] = std::sin(1.0 + i) / (1.0 + i);
  const std::string wf = "/tmp/macis_amd_wfn_test.dat";
  macis::write_wavefunction(wf, 6, dets, C);
  std::vector<std::bitset<128>> d2;
  std::vector<double> C2;
  macis::read_wavefunction(wf, d2, C2);
  REQUIRE(d2 == dets);
  REQUIRE(C2 == C);
  std::ifstream f(wf);
  std::string head, first;
  std::getline(f, head);
  std::getline(f, first);
  REQUIRE(head == std::to_string(dets.size()) + " 6 3 2");
  char ref[64];
  std::snprintf(ref, sizeof ref, "%30.16e ", C[0]);
  REQUIRE(first == std::string(ref) + macis::to_canonical_string(dets[0]) + " ");
  std::remove(wf.c_str());
  // binary RDMs: the reference's water CAS(8,8) file, read into padded
  // arrays and written back byte for byte
  const size_t n = 8, ld1 = 9, ld2 = 10;
  std::vector<double> D1(ld1 * n, -1.), D2(ld2 * ld2 * ld2 * n, -1.);
  macis::read_rdms_binary(dir + "/h2o.ccpvdz.cas88.rdms.bin", n, D1.data(), ld1, D2.data(), ld2);
  double tr = 0.;
  for(size_t p = 0; p < n; ++p) tr += D1[p * (ld1 + 1)];
  REQUIRE_NEAR(tr, 8.0, 1e-12);
  REQUIRE(D1[n] == -1.);  // padding beyond norb untouched (fcidump.cxx:212-222)
  const std::string rf = "/tmp/macis_amd_rdm_test.bin";
  macis::write_rdms_binary(rf, n, D1.data(), ld1, D2.data(), ld2);
  auto slurp = [](const std::string& fn) {
    std::ifstream in(fn, std::ios::binary);
    return std::string(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
  };
  REQUIRE(slurp(rf) == slurp(dir + "/h2o.ccpvdz.cas88.rdms.bin"));
  std::remove(rf.c_str());
  bool threw = false;
  try {
    macis::read_rdms_binary(dir + "/h2o.ccpvdz.cas88.rdms.bin", 7, D1.data(), ld1, D2.data(), ld2);
  } catch(const std::runtime_error& e) {
    threw = std::string(e.what()) == "NORB in RDM file doesn't match 7 8";
  }
  REQUIRE(threw);
}

static void test_transforms() {
  // tests/transform.cxx: two-index literals, four-index against the defining sum
  const size_t n = 4;
  const std::vector<double> A = {1, 5, 9, 13, 2, 6, 10, 14, 3, 7, 11, 15, 4, 8, 12, 16};
  const std::vector<double> I4 = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  const std::vector<double> P4 = {1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 1, 0, 0};
  const std::vector<double> S43 = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  const std::vector<double> SP43 = {1, 0, 0, 0, 0, 0, 0, 1, 0, 1, 0, 0};
  std::vector<double> B(n * n);
  macis::two_index_transform(n, n, A.data(), n, I4.data(), n, B.data(), n);
  REQUIRE(B == A);
  macis::two_index_transform(n, n, A.data(), n, P4.data(), n, B.data(), n);
  REQUIRE(B == std::vector<double>({1, 9, 13, 5, 3, 11, 15, 7, 4, 12, 16, 8, 2, 10, 14, 6}));
  std::vector<double> B3(9);
  macis::two_index_transform(n, 3, A.data(), n, S43.data(), n, B3.data(), 3);
  REQUIRE(B3 == std::vector<double>({1, 5, 9, 2, 6, 10, 3, 7, 11}));
  macis::two_index_transform(n, 3, A.data(), n, SP43.data(), n, B3.data(), 3);
  REQUIRE(B3 == std::vector<double>({1, 13, 5, 4, 16, 8, 2, 14, 6}));
  // four-index: Y(p,q,r,s) = sum X(i,j,k,l) C(i,p) C(j,q) C(k,r) C(l,s)
  std::vector<double> X(n * n * n * n);
  for(size_t i = 0; i < X.size(); ++i) X[i] = std::sin(0.37 * i + 1.0);
  auto check4 = [&](const std::vector<double>& Cm, size_t m) {
    std::vector<double> Y(m * m * m * m), ref(m * m * m * m, 0.);
    macis::four_index_transform(n, m, X.data(), n, Cm.data(), n, Y.data(), m);
    auto c = [&](size_t i, size_t p) { return Cm[i + n * p]; };
    for(size_t p = 0; p < m; ++p)
      for(size_t q = 0; q < m; ++q)
        for(size_t r = 0; r < m; ++r)
          for(size_t t = 0; t < m; ++t)
            for(size_t i = 0; i < n; ++i)
              for(size_t j = 0; j < n; ++j)
                for(size_t k = 0; k < n; ++k)
                  for(size_t l = 0; l < n; ++l)
                    ref[p + m * (q + m * (r + m * t))] += c(i, p) * c(j, q) * c(k, r) * c(l, t) *
                                                         X[i + n * (j + n * (k + n * l))];
    double e = 0.;
    for(size_t i = 0; i < Y.size(); ++i) e = std::max(e, std::abs(Y[i] - ref[i]));
    REQUIRE(e < 1e-13);
  };
  check4(I4, 4);
  check4(P4, 4);
  check4(S43, 3);
  check4(SP43, 3);
  // in place, as asci_grow calls it
  std::vector<double> Xi = X, Yo(X.size());
  macis::four_index_transform(n, n, X.data(), n, P4.data(), n, Yo.data(), n);
  macis::four_index_transform(n, n, Xi.data(), n, P4.data(), n, Xi.data(), n);
  REQUIRE(Xi == Yo);
}

// ---------------------------------------------------------------- GPU cases
static void test_rdms(const H2O& h, const std::string& dir) {
  // double_loop.cxx:162-177: the HF determinant's RDMs give E_HF
  const size_t n = h.norb, n2 = n * n, n4 = n2 * n2;
  auto T = h.T;
  auto V = h.V;
  macis::HipHamiltonianGenerator<64> gen(macis::matrix_span<double>(T.data(), n, n),
                                         macis::rank4_span<double>(V.data(), n, n, n, n));
  auto hf = macis::canonical_hf_determinant<64>(5, 5);
  std::vector<macis::wfn_t<64>> one{hf};
  std::vector<double> C1{1.}, ordm(n2, 0.), trdm(n4, 0.);
  gen.form_rdms(one.begin(), one.end(), one.begin(), one.end(), C1.data(),
                macis::matrix_span<double>(ordm.data(), n, n),
                macis::rank4_span<double>(trdm.data(), n, n, n, n));
  double E = 0.;
  for(size_t i = 0; i < n2; ++i) E += ordm[i] * T[i];
  for(size_t i = 0; i < n4; ++i) E += trdm[i] * V[i];
  REQUIRE_NEAR(E, gen.matrix_element(hf, hf), 1e-10);

  // CAS(8e,8o) with one inactive orbital against the reference's fixture
  // h2o.ccpvdz.cas88.rdms.bin (read as in fcidump.cxx:209-251)
  const size_t na = 8, ni = 1;
  std::vector<double> Ta(na * na), Va(na * na * na * na);
  auto v = [&](size_t p, size_t q, size_t r, size_t s) { return V[p + n * (q + n * (r + n * s))]; };
  for(size_t q = 0; q < na; ++q)
    for(size_t p = 0; p < na; ++p) {
      double t = T[(p + ni) + n * (q + ni)];
      for(size_t i = 0; i < ni; ++i) t += 2. * v(p + ni, q + ni, i, i) - v(p + ni, i, i, q + ni);
      Ta[p + na * q] = t;
      for(size_t s = 0; s < na; ++s)
        for(size_t r = 0; r < na; ++r)
          Va[p + na * (q + na * (r + na * s))] = v(p + ni, q + ni, r + ni, s + ni);
    }
  macis::HipHamiltonianGenerator<64> cas(macis::matrix_span<double>(Ta.data(), na, na),
                                         macis::rank4_span<double>(Va.data(), na, na, na, na));
  auto dets = macis::generate_hilbert_space<64>(na, 4, 4);
  std::vector<double> X;
  macis::selected_ci_diag<64, int32_t>(dets.begin(), dets.end(), cas, 1e-16, 20, 1e-8, X);
  std::vector<double> d1(na * na, 0.), d2(na * na * na * na, 0.);
  cas.form_rdms(dets.begin(), dets.end(), dets.begin(), dets.end(), X.data(),
                macis::matrix_span<double>(d1.data(), na, na),
                macis::rank4_span<double>(d2.data(), na, na, na, na));
  std::ifstream in(dir + "/h2o.ccpvdz.cas88.rdms.bin", std::ios::binary);
  int nread = 0;
  in.read(reinterpret_cast<char*>(&nread), sizeof(int));
  REQUIRE(nread == (int)na);
  std::vector<double> f1(d1.size()), f2(d2.size());
  in.read(reinterpret_cast<char*>(f1.data()), f1.size() * sizeof(double));
  in.read(reinterpret_cast<char*>(f2.data()), f2.size() * sizeof(double));
  REQUIRE(bool(in));
  double e1 = 0., e2 = 0.;
  for(size_t i = 0; i < d1.size(); ++i) e1 = std::max(e1, std::abs(d1[i] - f1[i]));
  for(size_t i = 0; i < d2.size(); ++i) e2 = std::max(e2, std::abs(d2[i] - f2[i]));
  std::printf("  CAS(8,8) RDMs vs fixture: max |dD1| %.2e, max |dD2| %.2e\n", e1, e2);
  REQUIRE(e1 < 1e-7);
  REQUIRE(e2 < 1e-7);
}

static void test_csr_hamiltonian(const H2O& h, const std::string& dir) {
  // csr_hamiltonian.cxx:18-71
  auto T = h.T;
  auto V = h.V;
  macis::HipHamiltonianGenerator<64> gen(macis::matrix_span<double>(T.data(), h.norb, h.norb),
                                         macis::rank4_span<double>(V.data(), h.norb, h.norb,
                                                                   h.norb, h.norb));
  auto hf = macis::canonical_hf_determinant<64>(5, 5);
  auto dets = macis::generate_cisd_hilbert_space(h.norb, hf);
  REQUIRE(dets.size() == 12636);
  auto H = macis::make_csr_hamiltonian_block<int32_t>(dets.begin(), dets.end(), dets.begin(),
                                                      dets.end(), gen, 1e-16);
  std::vector<int32_t> ref(H.rowptr().size());
  std::ifstream f(dir + "/h2o.ccpvdz.cisd.rowptr.bin", std::ios::binary);
  f.read((char*)ref.data(), ref.size() * sizeof(int32_t));
  REQUIRE(H.rowptr() == ref);
  REQUIRE(H.nnz() == 3517816);
  REQUIRE(H.indexing() == 0);
  auto H64 = macis::make_csr_hamiltonian<int64_t>(dets.begin(), dets.end(), gen, 1e-16);
  REQUIRE(H64.nnz() == H.nnz());
  bool same = true;
  for(int64_t i = 0; i < H.nnz(); ++i)
    same = same && H64.colind()[i] == H.colind()[i] && H64.nzval()[i] == H.nzval()[i];
  REQUIRE(same);
  // empty range -> empty matrix with indexing 0 (csr_hamiltonian.hpp:30-35)
  auto E = macis::make_csr_hamiltonian_block<int32_t>(dets.begin(), dets.begin(), dets.begin(),
                                                      dets.end(), gen, 1e-16);
  REQUIRE(E.m() == 0 && E.n() == 12636 && E.nnz() == 0);
}

static void test_davidson(const H2O& h) {
  // davidson.cxx:19-70
  auto T = h.T;
  auto V = h.V;
  macis::HipHamiltonianGenerator<64> gen(macis::matrix_span<double>(T.data(), h.norb, h.norb),
                                         macis::rank4_span<double>(V.data(), h.norb, h.norb,
                                                                   h.norb, h.norb));
  auto dets = macis::generate_cisd_hilbert_space(h.norb, macis::canonical_hf_determinant<64>(5, 5));
  auto H = macis::make_csr_hamiltonian<int32_t>(dets.begin(), dets.end(), gen, 1e-16);
  std::vector<double> X(H.n());
  macis::diagonal_guess(H.n(), H, X.data());
  auto D = sparsexx::extract_diagonal_elements(H);
  auto [niter, E0] = macis::davidson(H.n(), 15, macis::SparseMatrixOperator(H), D.data(), 1e-8,
                                     X.data());
  REQUIRE(niter > 0);
  REQUIRE_NEAR(E0 + h.E_core, -76.23197835987, 1e-9);
  double nrm = 0.;
  for(auto x : X) nrm += x * x;
  REQUIRE_NEAR(std::sqrt(nrm), 1.0, 1e-12);
  std::vector<double> AX(X.size());
  sparsexx::spblas::gespmbv(1, 1., H, X.data(), H.n(), 0., AX.data(), H.n());
  double xax = 0.;
  for(size_t i = 0; i < X.size(); ++i) xax += X[i] * AX[i];
  REQUIRE_NEAR(xax, E0, 1e-9);
  // selected_ci_diag on the device path gives the same energy
  std::vector<double> C;
  auto E1 = macis::selected_ci_diag<64, int32_t>(dets.begin(), dets.end(), gen, 1e-16, 15, 1e-8, C);
  REQUIRE_NEAR(E1, E0, 1e-10);
  // non-convergence is reported with the reference's message
  std::vector<double> Y(H.n(), 0.);
  macis::diagonal_guess(H.n(), H, Y.data());
  std::string msg;
  try {
    macis::davidson(H.n(), 3, macis::SparseMatrixOperator(H), D.data(), 1e-8, Y.data());
  } catch(const std::runtime_error& e) {
    msg = e.what();
  }
  REQUIRE(msg == "Davidson Did Not Converge!");
}

static void test_spmbv() {
  // src/sparsexx/tests/test_spmbv.cxx:14-53
  for(int indexing : {0, 1}) {
    sparsexx::csr_matrix<double, int32_t> A(6, 6, 16, indexing);
    const int i = indexing;
    A.rowptr() = {i, i + 2, i + 5, i + 8, i + 11, i + 14, i + 16};
    A.colind() = {i + 0, i + 1, i + 0, i + 1, i + 2, i + 1, i + 2, i + 3,
                  i + 2, i + 3, i + 4, i + 3, i + 4, i + 5, i + 4, i + 5};
    A.nzval() = {2, 1, 1, 2, 1, 1, 2, 1, 1, 2, 1, 1, 2, 1, 1, 2};
    std::vector<double> V = {1, 2, 3, 4, 5, 6}, AV(6, 1.);
    sparsexx::spblas::gespmbv(1, 1., A, V.data(), 6, 0., AV.data(), 6);
    std::vector<double> ref = {4, 8, 12, 16, 20, 17};
    REQUIRE(AV == ref);
  }
}

static void test_asci(const H2O& h) {
  // asci.cxx:209-276
  auto T = h.T;
  auto V = h.V;
  macis::HipHamiltonianGenerator<64> gen(macis::matrix_span<double>(T.data(), h.norb, h.norb),
                                         macis::rank4_span<double>(V.data(), h.norb, h.norb,
                                                                   h.norb, h.norb));
  macis::ASCISettings asci_settings;
  macis::MCSCFSettings mcscf_settings;
  std::vector<macis::wfn_t<64>> dets = {macis::canonical_hf_determinant<64>(5, 5)};
  std::vector<double> C = {1.0};
  double E0 = gen.matrix_element(dets[0], dets[0]);
  asci_settings.ntdets_max = 10000;
  std::tie(E0, dets, C) = macis::asci_grow(asci_settings, mcscf_settings, E0, std::move(dets),
                                           std::move(C), gen, h.norb);
  REQUIRE_NEAR(E0, -8.542926243842e+01, 1e-8);
  REQUIRE(dets.size() == 10000);
  REQUIRE(C.size() == 10000);
  std::printf("  grown to %zu determinants, E0 = %.12f\n", dets.size(), E0);
  double n2 = 0.;
  for(auto c : C) n2 += c * c;
  REQUIRE_NEAR(n2, 1.0, 1e-10);
  std::tie(E0, dets, C) = macis::asci_refine(asci_settings, mcscf_settings, E0, std::move(dets),
                                             std::move(C), gen, h.norb);
  REQUIRE_NEAR(E0, -8.542925964708e+01, 1e-8);
  REQUIRE(dets.size() == 10000);
  std::printf("  refined, E0 = %.12f\n", E0);
}

static void test_asci_rot(const H2O& h, double ref_E0) {
  // asci_grow with grow_with_rot (grow.hpp:67-160): RDMs on the device, natural
  // orbitals, rotated integrals re-uploaded, re-diagonalisation. ref_E0 (when
  // given) is the oracle's orc_asci_grow_rot energy for the same settings.
  auto T = h.T;
  auto V = h.V;
  macis::HipHamiltonianGenerator<64> gen(macis::matrix_span<double>(T.data(), h.norb, h.norb),
                                         macis::rank4_span<double>(V.data(), h.norb, h.norb,
                                                                   h.norb, h.norb));
  macis::ASCISettings asci_settings;
  macis::MCSCFSettings mcscf_settings;
  asci_settings.ntdets_max = 2000;
  asci_settings.grow_with_rot = true;
  asci_settings.rot_size_start = 100;
  std::vector<macis::wfn_t<64>> dets = {macis::canonical_hf_determinant<64>(5, 5)};
  std::vector<double> C = {1.0};
  double E0 = gen.matrix_element(dets[0], dets[0]);
  std::tie(E0, dets, C) = macis::asci_grow(asci_settings, mcscf_settings, E0, std::move(dets),
                                           std::move(C), gen, h.norb);
  std::printf("  grow_with_rot E0 = %.12f (%zu determinants)\n", E0, dets.size());
  REQUIRE(dets.size() == 2000);
  if(ref_E0 != 0.) REQUIRE_NEAR(E0, ref_E0, 1e-8);
  // the returned vector is the ground state of the same determinants in the
  // rotated orbitals (the last rotation follows the last selection, so this
  // energy need not be below E0): its Rayleigh quotient with the rotated
  // generator equals a fresh solve's energy
  std::vector<double> Cfresh;
  const double Efresh = macis::selected_ci_diag<64, int32_t>(dets.begin(), dets.end(), gen,
                                                             mcscf_settings.ci_matel_tol,
                                                             mcscf_settings.ci_max_subspace,
                                                             mcscf_settings.ci_res_tol, Cfresh);
  auto H = macis::make_csr_hamiltonian<int32_t>(dets.begin(), dets.end(), gen, 1e-16);
  std::vector<double> HC(C.size());
  sparsexx::spblas::gespmbv(1, 1., H, C.data(), H.n(), 0., HC.data(), H.n());
  double e = 0., n2 = 0.;
  for(size_t i = 0; i < C.size(); ++i) {
    e += C[i] * HC[i];
    n2 += C[i] * C[i];
  }
  std::printf("  re-diagonalised energy in natural orbitals %.12f\n", e / n2);
  REQUIRE_NEAR(n2, 1.0, 1e-10);
  REQUIRE_NEAR(e / n2, Efresh, 1e-8);
  // the caller's integrals were rotated in place and T stays symmetric
  double asym = 0.;
  for(size_t p = 0; p < h.norb; ++p)
    for(size_t q = 0; q < h.norb; ++q) asym = std::max(asym, std::abs(T[p + h.norb * q] - T[q + h.norb * p]));
  REQUIRE(asym < 1e-12);
  REQUIRE(std::abs(T[0] - h.T[0]) > 1e-8);
}

static void test_learned_prescreen(const H2O& h) {
  // LearnedPrescreen: plain steps HF -> 100 -> 800 -> 6400 are observed, the
  // network is trained on them, and the screened search of the next
  // (refinement) step is judged against the plain search of that step.
  auto T = h.T;
  auto V = h.V;
  macis::HipHamiltonianGenerator<64> gen(macis::matrix_span<double>(T.data(), h.norb, h.norb),
                                         macis::rank4_span<double>(V.data(), h.norb, h.norb,
                                                                   h.norb, h.norb));
  macis::ASCISettings s;
  macis::LearnedPrescreen<64> nn(gen, /*seed=*/3, /*keep_fraction=*/0.05, /*min_recall=*/0.5);
  std::vector<macis::wfn_t<64>> wfn = {macis::canonical_hf_determinant<64>(5, 5)};
  std::vector<double> C = {1.0};
  double E0 = gen.matrix_element(wfn[0], wfn[0]);
  auto step = [&](size_t nd, bool screened) {
    macis::reorder_ci_on_coeff(wfn, C);
    const size_t nc = std::min<size_t>(s.ncdets_max, wfn.size());
    auto out = screened ? nn.search(s, nd, wfn.begin(), wfn.begin() + nc, E0, C, h.norb)
                        : macis::asci_search(s, nd, wfn.begin(), wfn.begin() + nc, E0, C, h.norb,
                                             gen.T(), gen.G_red(), gen.V_red(), gen.G(), gen.V(),
                                             gen);
    return out;
  };
  // untrained: the plain search is the fallback
  {
    auto a = step(100, true);
    REQUIRE(nn.used_fallback());
    auto b = step(100, false);
    REQUIRE(a == b);
  }
  for(size_t nd : {100, 800, 6400}) {
    macis::reorder_ci_on_coeff(wfn, C);
    const size_t nc = std::min<size_t>(s.ncdets_max, wfn.size());
    nn.observe(s, wfn.begin(), wfn.begin() + nc, E0, C, h.norb, 4000000);
    wfn = step(nd, false);
    std::vector<double> X;
    E0 = macis::selected_ci_diag<64, int32_t>(wfn.begin(), wfn.end(), gen, 2.2e-16, 20, 1e-8, X);
    C = X;
  }
  nn.train(20, 4096, 1e-3f);
  std::printf("  validation recall %.4f\n", nn.recall());
  const size_t nd = wfn.size();
  auto plain = step(nd, false);
  auto screened = step(nd, true);
  REQUIRE(!nn.used_fallback());
  REQUIRE(screened.size() == plain.size());
  std::vector<double> Xp, Xs;
  const double Ep =
      macis::selected_ci_diag<64, int32_t>(plain.begin(), plain.end(), gen, 2.2e-16, 20, 1e-8, Xp);
  const double Es = macis::selected_ci_diag<64, int32_t>(screened.begin(), screened.end(), gen,
                                                         2.2e-16, 20, 1e-8, Xs);
  auto key = [](const macis::wfn_t<64>& d) { return d.to_ullong(); };
  std::vector<unsigned long long> kp, ks;
  for(auto& d : plain) kp.push_back(key(d));
  for(auto& d : screened) ks.push_back(key(d));
  std::sort(kp.begin(), kp.end());
  std::sort(ks.begin(), ks.end());
  std::vector<unsigned long long> both;
  std::set_intersection(kp.begin(), kp.end(), ks.begin(), ks.end(), std::back_inserter(both));
  std::printf("  recall %.4f, dE %+.3e Ha\n", (double)both.size() / kp.size(), Es - Ep);
  REQUIRE(std::abs(Es - Ep) <= 1e-4);
}

int main(int argc, char** argv) {
  std::setvbuf(stdout, nullptr, _IONBF, 0);  // every line reaches the log even on abort
  if(argc < 2) {
    std::printf("usage: %s <golden dir> [--cpu-only]\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  bool cpu_only = false;
  double rot_ref = 0.;  // --rot-ref=<E0 of the oracle's rotating grow>
  for(int a = 2; a < argc; ++a) {
    const std::string arg = argv[a];
    if(arg == "--cpu-only") cpu_only = true;
    if(arg.rfind("--rot-ref=", 0) == 0) rot_ref = std::stod(arg.substr(10));
  }
  auto h = load_h2o(dir);
  std::vector<std::pair<std::string, std::function<void()>>> cases = {
      {"sd_operations", [] { test_sd_operations(); }},
      {"host generator elements", [&] { test_host_generator_elements(h); }},
      {"file formats (wavefunction, binary RDMs)", [&] { test_file_formats(dir); }},
      {"orbital transforms", [] { test_transforms(); }},
  };
  if(!cpu_only) {
    cases.push_back({"CSR Hamiltonian (device)", [&] { test_csr_hamiltonian(h, dir); }});
    cases.push_back({"spmbv (device)", [] { test_spmbv(); }});
    cases.push_back({"Davidson (device)", [&] { test_davidson(h); }});
    cases.push_back({"RDMs (device)", [&] { test_rdms(h, dir); }});
    cases.push_back({"ASCI grow/refine (device)", [&] { test_asci(h); }});
    cases.push_back({"ASCI grow with natural-orbital rotation (device)",
                     [&] { test_asci_rot(h, rot_ref); }});
    cases.push_back({"learned ASCI pre-screen (device)", [&] { test_learned_prescreen(h); }});
  }
  for(auto& [name, fn] : cases) {
    std::printf("[case] %s\n", name.c_str());
    std::fflush(stdout);
    const int before = g_fail;
    try {
      fn();
    } catch(const std::exception& e) {
      ++g_fail;
      std::printf("  FAILED with exception: %s\n", e.what());
    }
    std::printf("  %s\n", g_fail == before ? "ok" : "FAILED");
  }
  std::printf("%d checks, %d failures\n", g_checks, g_fail);
  std::printf("[exit]\n");
  return g_fail ? 1 : 0;
}
