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
#include <random>
#include <unistd.h>
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

// A generator of the reference's base class with no CSR of its own: the
// per-pair members are all it is used for.
template <size_t N>
struct HostOnly : macis::HamiltonianGenerator<N> {
  using base = macis::HamiltonianGenerator<N>;
  using base::base;
  using typename base::full_det_iterator;
  using typename base::matrix_span_t;
  using typename base::rank4_span_t;
  typename base::template sparse_matrix_type<int32_t> make_csr_hamiltonian_block_32bit_(
      full_det_iterator, full_det_iterator, full_det_iterator, full_det_iterator,
      double) override {
    throw std::runtime_error("unused");
  }
  typename base::template sparse_matrix_type<int64_t> make_csr_hamiltonian_block_64bit_(
      full_det_iterator, full_det_iterator, full_det_iterator, full_det_iterator,
      double) override {
    throw std::runtime_error("unused");
  }
  void form_rdms(full_det_iterator, full_det_iterator, full_det_iterator, full_det_iterator,
                 double*, matrix_span_t, rank4_span_t) override {}
};

// Host RDMs of a wavefunction through the generator's rdm_contributions, the
// loop of double_loop.hpp:116-165 (pairs with popcount(bra^ket) <= 4 and
// |C_i C_j| > 1e-16).
template <size_t N, typename Gen>
static void host_rdms(Gen& g, const std::vector<macis::wfn_t<N>>& dets, const double* C,
                      macis::matrix_span<double> ordm, macis::rank4_span<double> trdm) {
  for(size_t i = 0; i < dets.size(); ++i) {
    const auto ba = g.alpha_string(dets[i]), bb = g.beta_string(dets[i]);
    const auto oa = macis::bits_to_indices(ba), ob = macis::bits_to_indices(bb);
    for(size_t j = 0; j < dets.size(); ++j) {
      if((dets[i] ^ dets[j]).count() > 4) continue;
      const double val = C[i] * C[j];
      if(std::abs(val) <= 1e-16) continue;
      const auto ka = g.alpha_string(dets[j]), kb = g.beta_string(dets[j]);
      g.rdm_contributions(ba, ka, ba ^ ka, bb, kb, bb ^ kb, oa, ob, val, ordm, trdm);
    }
  }
}

static void test_generator_members(const H2O& h) {
  // tests/double_loop.cxx:55-87: fast diagonals of excited determinants
  // from the HF diagonal, called as the reference's test calls them
  auto T = h.T;
  auto V = h.V;
  const size_t norb = h.norb, nocc = 5;
  HostOnly<64> ham_gen(macis::matrix_span<double>(T.data(), norb, norb),
                       macis::rank4_span<double>(V.data(), norb, norb, norb, norb));
  const auto hf_det = macis::canonical_hf_determinant<64>(nocc, nocc);
  const auto EHF = ham_gen.matrix_element(hf_det, hf_det);
  std::vector<uint32_t> occ = {0, 1, 2, 3, 4};
  {
    auto state = hf_det;
    state.flip(0).flip(nocc);
    const auto ES = ham_gen.matrix_element(state, state);
    REQUIRE_NEAR(ES, -6.488097259228e+01, 1e-10);
    REQUIRE_NEAR(ham_gen.fast_diag_single(occ, occ, 0, nocc, EHF), ES, 1e-10);
  }
  {
    auto state = hf_det;
    state.flip(0).flip(nocc).flip(1).flip(nocc + 1);
    const auto ED = ham_gen.matrix_element(state, state);
    REQUIRE_NEAR(ED, -6.314093508151e+01, 1e-10);
    REQUIRE_NEAR(ham_gen.fast_diag_ss_double(occ, occ, 0, 1, nocc, nocc + 1, EHF), ED, 1e-10);
  }
  {
    auto state = hf_det;
    state.flip(0).flip(nocc).flip(1 + 32).flip(nocc + 1 + 32);
    const auto ED = ham_gen.matrix_element(state, state);
    REQUIRE_NEAR(ED, -6.304547887231e+01, 1e-10);
    REQUIRE_NEAR(ham_gen.fast_diag_os_double(occ, occ, 0, 1, nocc, nocc + 1, EHF), ED, 1e-10);
  }
  // orbital energies: the vector form equals the scalar form; the
  // energy-argument fast diagonals equal the occupation-argument ones
  const auto eps = ham_gen.single_orbital_ens(norb, occ, occ);
  REQUIRE(eps.size() == norb);
  bool same = true;
  for(uint32_t p = 0; p < norb; ++p) same &= eps[p] == ham_gen.single_orbital_en(p, occ, occ);
  REQUIRE(same);
  REQUIRE(ham_gen.fast_diag_single(eps[2], eps[7], 2, 7, EHF) ==
          ham_gen.fast_diag_single(occ, occ, 2, 7, EHF));
  // every single and double of the HF determinant: fast diagonal == exact
  double worst = 0.;
  for(uint32_t i = 0; i < nocc; ++i)
    for(uint32_t a = nocc; a < norb; ++a) {
      auto s1 = hf_det;
      s1.flip(i).flip(a);
      worst = std::max(worst, std::abs(ham_gen.fast_diag_single(eps[i], eps[a], i, a, EHF) -
                                       ham_gen.matrix_element(s1, s1)));
      for(uint32_t j = 0; j < nocc; ++j)
        for(uint32_t b = nocc; b < norb; ++b) {
          auto s2 = s1;
          s2.flip(j + 32).flip(b + 32);
          worst = std::max(worst, std::abs(ham_gen.fast_diag_os_double(eps[i], eps[j], eps[a],
                                                                       eps[b], i, j, a, b, EHF) -
                                           ham_gen.matrix_element(s2, s2)));
          if(j > i && b > a) {
            auto s3 = s1;
            s3.flip(j).flip(b);
            worst = std::max(worst,
                             std::abs(ham_gen.fast_diag_ss_double(eps[i], eps[j], eps[a], eps[b],
                                                                  i, j, a, b, EHF) -
                                      ham_gen.matrix_element(s3, s3)));
          }
        }
    }
  REQUIRE(worst < 1e-10);
  // the per-class elements and the 8-argument dispatch equal the full-det
  // matrix_element bit for bit (matrix_elements.hpp:14-62)
  std::mt19937_64 rng(5);
  bool bits = true;
  int classes[5] = {0, 0, 0, 0, 0};
  auto dets = macis::generate_cisd_hilbert_space<64>(norb, hf_det);
  for(int t = 0; t < 200000; ++t) {
    // bras among the first 400 (HF, singles, the first doubles), kets anywhere
    const auto bra = dets[rng() % 400], ket = dets[rng() % dets.size()];
    const auto ba = ham_gen.alpha_string(bra), ka = ham_gen.alpha_string(ket);
    const auto bb = ham_gen.beta_string(bra), kb = ham_gen.beta_string(ket);
    const auto ea = ba ^ ka, eb = bb ^ kb;
    const auto oa = macis::bits_to_indices(ba), ob = macis::bits_to_indices(bb);
    const double ref = ham_gen.matrix_element(bra, ket);
    bits &= ham_gen.matrix_element(ba, ka, ea, bb, kb, eb, oa, ob) == ref;
    const size_t ca = ea.count(), cb = eb.count();
    if(ca + cb > 4) continue;
    double cls;
    if(ca == 4) cls = ham_gen.matrix_element_4(ba, ka, ea), ++classes[0];
    else if(cb == 4) cls = ham_gen.matrix_element_4(bb, kb, eb), ++classes[0];
    else if(ca == 2 && cb == 2) cls = ham_gen.matrix_element_22(ba, ka, ea, bb, kb, eb), ++classes[1];
    else if(ca == 2) cls = ham_gen.matrix_element_2(ba, ka, ea, oa, ob), ++classes[2];
    else if(cb == 2) cls = ham_gen.matrix_element_2(bb, kb, eb, ob, oa), ++classes[3];
    else cls = ham_gen.matrix_element_diag(oa, ob), ++classes[4];
    bits &= cls == ref;
  }
  REQUIRE(bits);
  std::printf("  pairs per class (4, 22, 2a, 2b, diag): %d %d %d %d %d\n", classes[0], classes[1],
              classes[2], classes[3], classes[4]);
  REQUIRE(classes[0] > 0 && classes[1] > 0 && classes[2] > 0 && classes[3] > 0 && classes[4] > 0);
  REQUIRE(ham_gen.alpha_string(hf_det).to_ullong() == 0x1Full);
  REQUIRE(ham_gen.beta_string(hf_det).to_ullong() == 0x1Full);
  // public intermediate spans over the generator's storage (impl.hpp:25-70)
  REQUIRE(&ham_gen.G_red_(0, 0, 0) == ham_gen.G_red());
  REQUIRE(ham_gen.G2_red_(3, 2) == 0.5 * ham_gen.G_pqrs_(3, 3, 2, 2));
  REQUIRE(ham_gen.V2_red_(3, 2) == ham_gen.V_pqrs_(3, 3, 2, 2));

  // double_loop.cxx:162-177 through the host rdm_contributions: the HF
  // determinant's RDMs reproduce E_HF
  {
    const size_t n2 = norb * norb, n4 = n2 * n2;
    std::vector<double> ordm(n2, 0.), trdm(n4, 0.);
    std::vector<macis::wfn_t<64>> one{hf_det};
    const double C1 = 1.;
    host_rdms<64>(ham_gen, one, &C1, macis::matrix_span<double>(ordm.data(), norb, norb),
                  macis::rank4_span<double>(trdm.data(), norb, norb, norb, norb));
    double E = 0.;
    for(size_t i = 0; i < n2; ++i) E += ordm[i] * T[i];
    for(size_t i = 0; i < n4; ++i) E += trdm[i] * V[i];
    REQUIRE_NEAR(E, EHF, 1e-10);
  }
  // double_loop.cxx:179-215 ("RDMS", HF section): wfn_t<128>, 34 orbitals,
  // D2 of a single determinant is the antisymmetrised product of D1
  {
    const size_t n = 34, n2 = n * n;
    std::vector<double> T0(n2, 0.), V0(n2 * n2, 0.), ordm(n2, 0.), trdm(n2 * n2, 0.);
    HostOnly<128> g(macis::matrix_span<double>(T0.data(), n, n),
                    macis::rank4_span<double>(V0.data(), n, n, n, n));
    macis::matrix_span<double> D1(ordm.data(), n, n);
    macis::rank4_span<double> D2(trdm.data(), n, n, n, n);
    std::vector<macis::wfn_t<128>> one{macis::canonical_hf_determinant<128>(nocc, nocc)};
    const double C1 = 1.;
    host_rdms<128>(g, one, &C1, D1, D2);
    for(size_t i = 0; i < nocc; ++i)
      for(size_t j = 0; j < nocc; ++j)
        for(size_t k = 0; k < nocc; ++k)
          for(size_t l = 0; l < nocc; ++l) {
            D2(i, j, l, k) -= 0.5 * D1(i, j) * D1(k, l);
            D2(i, j, l, k) += 0.25 * D1(i, l) * D1(k, j);
          }
    double sum = 0.;
    for(auto x : trdm) sum += std::abs(x);
    REQUIRE(sum < 1e-15);
    for(size_t i = 0; i < nocc; ++i) D1(i, i) -= 2.0;
    sum = 0.;
    for(auto x : ordm) sum += std::abs(x);
    REQUIRE(sum < 1e-15);
  }
  // rotate_hamiltonian_ordm: a diagonal 1-RDM with ascending occupations
  // rotates to the reversed orbital order (natural orbitals by descending
  // occupation), so T'(p,q) = T(n-1-p, n-1-q) and V likewise
  {
    const size_t n = 6;
    std::vector<double> Tr(n * n), Vr(n * n * n * n), D(n * n, 0.);
    for(size_t i = 0; i < Tr.size(); ++i) Tr[i] = std::sin(1.0 + i);
    for(size_t i = 0; i < Vr.size(); ++i) Vr[i] = std::cos(0.5 * i);
    const auto T0 = Tr, V0 = Vr;
    for(size_t p = 0; p < n; ++p) D[p + n * p] = 0.1 * (p + 1);
    HostOnly<64> g(macis::matrix_span<double>(Tr.data(), n, n),
                   macis::rank4_span<double>(Vr.data(), n, n, n, n));
    g.rotate_hamiltonian_ordm(D.data());
    double eT = 0., eV = 0.;
    const size_t m = n - 1;
    for(size_t q = 0; q < n; ++q)
      for(size_t p = 0; p < n; ++p) {
        eT = std::max(eT, std::abs(Tr[p + n * q] - T0[(m - p) + n * (m - q)]));
        for(size_t r = 0; r < n; ++r)
          for(size_t s = 0; s < n; ++s)
            eV = std::max(eV, std::abs(Vr[p + n * (q + n * (r + n * s))] -
                                       V0[(m - p) + n * ((m - q) + n * ((m - r) + n * (m - s)))]));
      }
    REQUIRE(eT < 1e-14);
    REQUIRE(eV < 1e-14);
    REQUIRE(g.G2_red_(0, 1) == 0.5 * g.G_pqrs_(0, 0, 1, 1));  // intermediates regenerated
  }
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

  // write_fcidump (fcidump.cxx:184-207) of the water integrals, read back:
  // 17 significant digits make the round trip exact
  {
    const std::string f = dir + "/h2o.ccpvdz.fci.dat";
    const size_t n = macis::read_fcidump_norb(f);
    const double Ec = macis::read_fcidump_core(f);
    std::vector<double> T(n * n, 0.), V(n * n * n * n, 0.);
    macis::read_fcidump_1body(f, T.data(), n);
    macis::read_fcidump_2body(f, V.data(), n);
    char tmpl[] = "/tmp/mhip_fcidump_XXXXXX";
    const int fd = mkstemp(tmpl);
    REQUIRE(fd >= 0);
    close(fd);
    std::remove(tmpl);  // write_fcidump appends, as the reference's sink does
    macis::write_fcidump(tmpl, n, T.data(), n, V.data(), n, Ec);
    std::ifstream in(tmpl);
    std::string first;
    std::getline(in, first);
    char ref[128];
    std::snprintf(ref, sizeof ref, "%8d %8d %8d %8d %25.16e", 1, 1, 1, 1, V[0]);
    REQUIRE(first == ref);
    size_t lines = 1;
    for(std::string l; std::getline(in, l);) ++lines;
    REQUIRE(lines == n * n * n * n + n * n + 1);
    REQUIRE(macis::read_fcidump_norb(tmpl) == n);
    REQUIRE(macis::read_fcidump_core(tmpl) == Ec);
    std::vector<double> T2(n * n, 0.), V2(n * n * n * n, 0.);
    macis::read_fcidump_1body(tmpl, T2.data(), n);
    macis::read_fcidump_2body(tmpl, V2.data(), n);
    REQUIRE(T2 == T);
    REQUIRE(V2 == V);
    // span readers (fcidump.hpp:54-71) into views wider than norb: the same
    // entries at the view's strides, the padding untouched
    const size_t m = n + 3;
    std::vector<double> Tp(m * m, -7.), Vp(m * m * m * m, -7.);
    macis::read_fcidump_1body(tmpl, macis::col_major_span<double, 2>(Tp.data(), m, m));
    macis::read_fcidump_2body(tmpl, macis::col_major_span<double, 4>(Vp.data(), m, m, m, m));
    bool same = true, pad = true;
    for(size_t q = 0; q < m; ++q)
      for(size_t p = 0; p < m; ++p) {
        if(p < n && q < n) same &= Tp[p + m * q] == T[p + n * q];
        else pad &= Tp[p + m * q] == -7.;
      }
    for(size_t l = 0; l < n; ++l)
      for(size_t k = 0; k < n; ++k)
        for(size_t j = 0; j < n; ++j)
          for(size_t i = 0; i < n; ++i)
            same &= Vp[i + m * (j + m * (k + m * l))] == V[i + n * (j + n * (k + n * l))];
    pad &= Vp[n + m * (n + m * (n + m * n))] == -7.;
    REQUIRE(same);
    REQUIRE(pad);
    std::remove(tmpl);
  }
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
  // spans wider than norb (the reference indexes any extents): the corner
  // receives the same RDMs, the padding is untouched
  {
    const size_t m = na + 2;
    std::vector<double> w1(m * m, -1.), w2(m * m * m * m, -1.);
    for(size_t q = 0; q < na; ++q)
      for(size_t p = 0; p < na; ++p) w1[p + m * q] = 0.;
    for(size_t l = 0; l < na; ++l)
      for(size_t k = 0; k < na; ++k)
        for(size_t j = 0; j < na; ++j)
          for(size_t i = 0; i < na; ++i) w2[i + m * (j + m * (k + m * l))] = 0.;
    cas.form_rdms(dets.begin(), dets.end(), dets.begin(), dets.end(), X.data(),
                  macis::matrix_span<double>(w1.data(), m, m),
                  macis::rank4_span<double>(w2.data(), m, m, m, m));
    double e = 0.;
    bool pad = true;
    for(size_t q = 0; q < m; ++q)
      for(size_t p = 0; p < m; ++p) {
        if(p < na && q < na) e = std::max(e, std::abs(w1[p + m * q] - d1[p + na * q]));
        else pad &= w1[p + m * q] == -1.;
      }
    for(size_t l = 0; l < na; ++l)
      for(size_t k = 0; k < na; ++k)
        for(size_t j = 0; j < na; ++j)
          for(size_t i = 0; i < na; ++i)
            e = std::max(e, std::abs(w2[i + m * (j + m * (k + m * l))] -
                                     d2[i + na * (j + na * (k + na * l))]));
    pad &= w2[na + m * (0 + m * (0 + m * 0))] == -1. && w2[m * m * m * m - 1] == -1.;
    REQUIRE(e < 1e-12);
    REQUIRE(pad);
  }
  // the generator's host rdm_contributions over the same pairs equal the
  // device form_rdms up to summation order
  std::vector<double> h1(na * na, 0.), h2(na * na * na * na, 0.);
  host_rdms<64>(cas, dets, X.data(), macis::matrix_span<double>(h1.data(), na, na),
                macis::rank4_span<double>(h2.data(), na, na, na, na));
  double g1 = 0., g2 = 0.;
  for(size_t i = 0; i < d1.size(); ++i) g1 = std::max(g1, std::abs(d1[i] - h1[i]));
  for(size_t i = 0; i < d2.size(); ++i) g2 = std::max(g2, std::abs(d2[i] - h2[i]));
  std::printf("  host rdm_contributions vs device form_rdms: %.2e %.2e\n", g1, g2);
  REQUIRE(g1 < 1e-12);
  REQUIRE(g2 < 1e-12);
  // rotate_hamiltonian_ordm to the natural orbitals keeps the CAS-CI energy
  // and makes the new 1-RDM diagonal with descending occupations
  double E0 = 0.;
  {
    std::vector<double> Xr;
    E0 = macis::selected_ci_diag<64, int32_t>(dets.begin(), dets.end(), cas, 1e-16, 20, 1e-8, Xr);
  }
  cas.rotate_hamiltonian_ordm(d1.data());
  std::vector<double> Xn;
  const double E1 =
      macis::selected_ci_diag<64, int32_t>(dets.begin(), dets.end(), cas, 1e-16, 20, 1e-8, Xn);
  REQUIRE_NEAR(E1, E0, 1e-9);
  std::vector<double> n1(na * na, 0.), nt(na * na * na * na, 0.);
  cas.form_rdms(dets.begin(), dets.end(), dets.begin(), dets.end(), Xn.data(),
                macis::matrix_span<double>(n1.data(), na, na),
                macis::rank4_span<double>(nt.data(), na, na, na, na));
  double off = 0.;
  for(size_t q = 0; q < na; ++q)
    for(size_t p = 0; p < na; ++p)
      if(p != q) off = std::max(off, std::abs(n1[p + na * q]));
  REQUIRE(off < 1e-6);
  bool desc = true;
  for(size_t p = 1; p < na; ++p) desc &= n1[p + na * p] <= n1[(p - 1) + na * (p - 1)] + 1e-9;
  REQUIRE(desc);
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
    // the device-resident overload: matrix uploaded once, device blocks
    auto& ctx = macis::hip::default_context();
    std::vector<int64_t> rp(A.rowptr().begin(), A.rowptr().end());
    mhip_csr* h = nullptr;
    macis::hip::check(mhip_csr_create_from_host(ctx.get(), 6, 6, 16, indexing, rp.data(),
                                                A.colind().data(), A.nzval().data(), &h));
    macis::hip::DeviceCSR dA(h);
    macis::hip::DeviceBuffer<double> dV(ctx, V.data(), 6), dAV(ctx, AV.data(), 6);
    sparsexx::spblas::gespmbv(1, 2., dA, dV.data(), 6, -1., dAV.data(), 6);
    std::vector<double> out(6);
    dAV.download(out.data(), 6);
    std::vector<double> ref2 = {4, 8, 12, 16, 20, 17};  // 2 A V - A V
    REQUIRE(out == ref2);
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
  // asci_search over a generator of the reference's base class (no device
  // copy of its own): the integral pointers of the signature go to the
  // device, and the result equals the device generator's
  auto T2 = h.T;
  auto V2 = h.V;
  HostOnly<64> host_gen(macis::matrix_span<double>(T2.data(), h.norb, h.norb),
                        macis::rank4_span<double>(V2.data(), h.norb, h.norb, h.norb, h.norb));
  macis::reorder_ci_on_coeff(dets, C);
  auto a = macis::asci_search(asci_settings, 3000, dets.begin(), dets.begin() + 100, E0, C,
                              h.norb, gen.T(), gen.G_red(), gen.V_red(), gen.G(), gen.V(), gen);
  auto b = macis::asci_search(asci_settings, 3000, dets.begin(), dets.begin() + 100, E0, C,
                              h.norb, host_gen.T(), host_gen.G_red(), host_gen.V_red(),
                              host_gen.G(), host_gen.V(), host_gen);
  REQUIRE(a.size() == 3000);
  REQUIRE(a == b);
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
  // LearnedPrescreen: the cores of the plain steps HF -> 100 -> 800 -> 6400
  // are observed, the network is trained on their block targets and judged
  // on the next (refinement) step against the plain search of that step.
  // The screen generates only the kept blocks; at keep_fraction 1 it must
  // reproduce the plain search, and a network whose recall is below
  // min_recall leaves the plain search in charge.
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
  auto plain_step = [&](size_t nd) {
    macis::reorder_ci_on_coeff(wfn, C);
    const size_t nc = std::min<size_t>(s.ncdets_max, wfn.size());
    return macis::asci_search(s, nd, wfn.begin(), wfn.begin() + nc, E0, C, h.norb, gen.T(),
                              gen.G_red(), gen.V_red(), gen.G(), gen.V(), gen);
  };
  auto screened_step = [&](size_t nd) {
    macis::reorder_ci_on_coeff(wfn, C);
    const size_t nc = std::min<size_t>(s.ncdets_max, wfn.size());
    return nn.search(s, nd, wfn.begin(), wfn.begin() + nc, E0, C, h.norb);
  };
  // untrained: the plain search is the fallback
  {
    auto a = screened_step(100);
    REQUIRE(nn.used_fallback());
    REQUIRE(a == plain_step(100));
  }
  for(size_t nd : {100, 800, 6400}) {
    macis::reorder_ci_on_coeff(wfn, C);
    const size_t nc = std::min<size_t>(s.ncdets_max, wfn.size());
    REQUIRE(nn.observe(s, wfn.begin(), wfn.begin() + nc, E0, h.norb) > 0);
    wfn = plain_step(nd);
    std::vector<double> X;
    E0 = macis::selected_ci_diag<64, int32_t>(wfn.begin(), wfn.end(), gen, 2.2e-16, 20, 1e-8, X);
    C = X;
  }
  nn.train(20, 4096, 1e-3f);
  const size_t nd = wfn.size();
  macis::reorder_ci_on_coeff(wfn, C);
  const size_t nc = std::min<size_t>(s.ncdets_max, wfn.size());
  // every block kept: the plain search's determinants
  nn.set_keep_fraction(1.0);
  REQUIRE(nn.validate(s, nd, wfn.begin(), wfn.begin() + nc, E0, C, h.norb) == 1.0);
  auto all = screened_step(nd);
  REQUIRE(!nn.used_fallback());
  REQUIRE(all == plain_step(nd));
  // 5 % of the blocks: judged against the plain search
  nn.set_keep_fraction(0.05);
  const double rec = nn.validate(s, nd, wfn.begin(), wfn.begin() + nc, E0, C, h.norb);
  auto plain = plain_step(nd);
  auto screened = screened_step(nd);
  REQUIRE(nn.used_fallback() == (rec < 0.5));
  REQUIRE(screened.size() == plain.size());
  std::vector<double> Xp, Xs;
  const double Ep =
      macis::selected_ci_diag<64, int32_t>(plain.begin(), plain.end(), gen, 2.2e-16, 20, 1e-8, Xp);
  const double Es = macis::selected_ci_diag<64, int32_t>(screened.begin(), screened.end(), gen,
                                                         2.2e-16, 20, 1e-8, Xs);
  std::printf("  keep 0.05: recall %.4f (%s), dE %+.3e Ha\n", rec,
              nn.used_fallback() ? "plain fallback" : "screened", Es - Ep);
  REQUIRE(rec > 0.3);
  if(nn.used_fallback()) REQUIRE(Es == Ep);
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
      {"generator members (fast diagonals, elements, RDM contributions, rotation)",
       [&] { test_generator_members(h); }},
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
