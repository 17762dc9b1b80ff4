// ASCI search (restating include/macis/asci/determinant_search.hpp:26-51,
// 394-663, world_size == 1). The candidate generation, accumulation and
// top-k run on the device (mhip_asci_search). With a HipHamiltonianGenerator
// the device copy of its integrals is used; with any other generator the
// reference signature's own integral pointers are uploaded: T_pq (norb x
// norb) and V_pqrs (norb^4), from which the device forms G, G_red, V_red and
// the diagonal intermediates by the same formulas (impl.hpp:25-70), so the
// scores equal those the generator's members would give.
#pragma once
#include <memory>
#include <vector>

#include <macis/hip/hip_hamiltonian_generator.hpp>

namespace macis {

struct ASCISettings {
  size_t ntdets_max = 1e5;
  size_t ntdets_min = 100;
  size_t ncdets_max = 100;
  double h_el_tol = 1e-8;
  double rv_prune_tol = 1e-8;
  size_t pair_size_max = 5e8;
  bool just_singles = false;
  size_t grow_factor = 8;
  size_t max_refine_iter = 6;
  double refine_energy_tol = 1e-6;
  bool grow_with_rot = false;
  size_t rot_size_start = 1000;
  int constraint_level = 2;
};

template <size_t N>
std::vector<wfn_t<N>> asci_search(ASCISettings asci_settings, size_t ndets_max,
                                  wavefunction_iterator_t<N> cdets_begin,
                                  wavefunction_iterator_t<N> cdets_end, const double E_ASCI,
                                  const std::vector<double>& C, size_t norb, const double* T_pq,
                                  const double* G_red, const double* V_red, const double* G_pqrs,
                                  const double* V_pqrs, HamiltonianGenerator<N>& ham_gen) {
  (void)G_red, (void)V_red, (void)G_pqrs;
  auto* hg = dynamic_cast<HipHamiltonianGenerator<N>*>(&ham_gen);
  std::unique_ptr<hip::Context> own;
  if(!hg) {
    if(!T_pq || !V_pqrs) throw std::runtime_error("asci_search: T_pq and V_pqrs are required");
    own = std::make_unique<hip::Context>(0);
    hip::check(mhip_set_integrals(own->get(), (int64_t)norb, T_pq, V_pqrs));
  }
  auto& ctx = hg ? hg->context() : *own;
  const size_t ncdets = std::distance(cdets_begin, cdets_end);
  mhip_asci_settings s;
  s.ncdets_max = asci_settings.ncdets_max;
  s.h_el_tol = asci_settings.h_el_tol;
  s.rv_prune_tol = asci_settings.rv_prune_tol;
  s.pair_size_max = asci_settings.pair_size_max;
  s.just_singles = asci_settings.just_singles;
  auto dc = hip::upload_dets<N>(ctx, cdets_begin, cdets_end);
  hip::DeviceBuffer<double> dC(ctx, C.data(), std::max<size_t>(ncdets, 1));
  hip::DeviceBuffer<uint64_t> out(ctx, std::max<size_t>(ndets_max, 1) * hip::words_per_det(N));
  int64_t nnew = 0;
  hip::check(mhip_asci_search(ctx.get(), hip::words_per_det(N), &s, ndets_max, dc.data(), ncdets,
                              E_ASCI, dC.data(), norb, out.data(), &nnew));
  return hip::download_dets<N>(out, nnew);
}

}  // namespace macis
