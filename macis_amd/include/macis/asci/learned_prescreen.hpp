// Learned ASCI block screen (no counterpart in the reference). The
// candidates of a core determinant fall into excitation blocks, one per hole
// set (mhip_asci_blocks); an MLP on the block's determinant (the core
// determinant with the holes cleared) predicts the log of the block's
// largest |h / (E - H_kk)|, and the screened search generates only the
// keep_fraction blocks with the best predicted log|C_core| + prediction.
// Their exact contributions are accumulated and ranked as in asci_search.
//
// Training pairs are the block targets of the cores of earlier steps
// (observe); training runs on the device (BF16 MFMA kernels). A network is
// trusted only after validate() has compared its screened search with the
// plain asci_search on a core and found a recall of at least min_recall;
// until then, and whenever that check fails, search() is the plain
// asci_search.
#pragma once
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include <macis/asci/determinant_search.hpp>

namespace macis {

template <size_t N>
class LearnedPrescreen {
 public:
  static constexpr int W = hip::words_per_det(N);

  LearnedPrescreen(HipHamiltonianGenerator<N>& gen, uint64_t seed = 0, double keep_fraction = 0.05,
                   double min_recall = 0.99)
      : gen_(gen), keep_(keep_fraction), min_recall_(min_recall) {
    hip::check(mhip_mlp_create(gen_.context().get(), seed, &mlp_));
  }
  ~LearnedPrescreen() {
    if(mlp_) mhip_mlp_destroy(mlp_);
  }
  LearnedPrescreen(const LearnedPrescreen&) = delete;
  LearnedPrescreen& operator=(const LearnedPrescreen&) = delete;

  // Keeps the (block determinant, target) pairs of the core as training
  // data (blocks without a screened-in candidate are skipped); returns the
  // number added.
  size_t observe(const ASCISettings& s, wavefunction_iterator_t<N> cbeg,
                 wavefunction_iterator_t<N> cend, double E0, size_t norb) {
    auto& ctx = gen_.context();
    const size_t nc = std::distance(cbeg, cend);
    if(!nc) return 0;
    auto dc = hip::upload_dets<N>(ctx, cbeg, cend);
    const mhip_asci_settings ms = to_c(s);
    int64_t nb = 0;
    hip::check(mhip_asci_blocks(ctx.get(), W, &ms, dc.data(), nc, E0, norb, nullptr, nullptr, 0,
                                &nb));
    hip::DeviceBuffer<uint64_t> dd(ctx, std::max<int64_t>(nb, 1) * W);
    hip::DeviceBuffer<float> dt(ctx, std::max<int64_t>(nb, 1));
    hip::check(mhip_asci_blocks(ctx.get(), W, &ms, dc.data(), nc, E0, norb, dd.data(), dt.data(),
                                nb, &nb));
    std::vector<uint64_t> w(nb * W);
    std::vector<float> t(nb);
    if(nb) {
      dd.download(w.data(), nb * W);
      dt.download(t.data(), nb);
    }
    size_t added = 0;
    for(int64_t i = 0; i < nb; ++i) {
      if(!std::isfinite(t[i])) continue;
      dets_.insert(dets_.end(), w.begin() + i * W, w.begin() + (i + 1) * W);
      target_.push_back(t[i]);
      ++added;
    }
    return added;
  }

  // Adam epochs over the observed pairs; returns the last mini-batch loss.
  // Targets below the clip_quantile quantile are raised to it, then
  // standardised; the screen undoes the standardisation (scale, shift).
  double train(int epochs, int64_t batch, float lr, uint64_t seed = 1,
               double clip_quantile = 0.0) {
    const size_t n = target_.size();
    if(!n) return 0.;
    float clip = -INFINITY;
    if(clip_quantile > 0.) {
      std::vector<float> sorted(target_);
      const size_t qi = std::min(n - 1, (size_t)(clip_quantile * (double)(n - 1)));
      std::nth_element(sorted.begin(), sorted.begin() + qi, sorted.end());
      clip = sorted[qi];
    }
    double mu = 0., var = 0.;
    for(float t : target_) mu += std::max(t, clip);
    mu /= n;
    for(float t : target_) var += (std::max(t, clip) - mu) * (std::max(t, clip) - mu);
    mu_ = mu;
    sd_ = std::sqrt(var / n) + 1e-12;
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    auto& ctx = gen_.context();
    batch = std::min<int64_t>(batch, n);
    // each epoch's permutation is uploaded once; the steps walk through it
    hip::DeviceBuffer<uint64_t> bd(ctx, n * W);
    hip::DeviceBuffer<float> bt(ctx, n);
    std::vector<uint64_t> hd(n * W);
    std::vector<float> ht(n);
    float loss = 0.f;
    const size_t nsteps = n / (size_t)batch;
    for(int e = 0; e < epochs; ++e) {
      std::shuffle(perm.begin(), perm.end(), rng);
      for(size_t i = 0; i < n; ++i) {
        const size_t p = perm[i];
        std::copy_n(dets_.begin() + p * W, W, hd.begin() + i * W);
        ht[i] = (float)((std::max(target_[p], clip) - mu_) / sd_);
      }
      bd.upload(hd.data(), n * W);
      bt.upload(ht.data(), n);
      // the loss (a wait for the device) is read on the last step only
      for(size_t b0 = 0, i = 0; b0 + batch <= n; b0 += batch, ++i)
        hip::check(mhip_mlp_train_step(ctx.get(), mlp_, W, bd.data() + b0 * W, bt.data() + b0,
                                       batch, lr,
                                       e + 1 == epochs && i + 1 == nsteps ? &loss : nullptr));
    }
    trained_ = true;
    recall_ = 0.;  // a retrained network is judged again
    return loss;
  }

  // Runs the screened and the plain search on one core and returns the
  // fraction of the plain search's new determinants that the screened one
  // also selects; the network is trusted when it is >= min_recall.
  double validate(const ASCISettings& s, size_t ndets_max, wavefunction_iterator_t<N> cbeg,
                  wavefunction_iterator_t<N> cend, double E0, const std::vector<double>& C,
                  size_t norb) {
    if(!trained_) return 0.;
    const size_t nc = std::distance(cbeg, cend);
    auto plain = asci_search(s, ndets_max, cbeg, cend, E0, C, norb, gen_.T(), gen_.G_red(),
                             gen_.V_red(), gen_.G(), gen_.V(), gen_);
    auto scr = screened(s, ndets_max, cbeg, cend, E0, C, norb);
    const size_t np = plain.size() - nc;
    std::vector<wfn_t<N>> a(plain.begin(), plain.begin() + np), b(scr.begin(), scr.end() - nc);
    auto less = bitset_less_comparator<N>{};
    std::sort(a.begin(), a.end(), less);
    std::sort(b.begin(), b.end(), less);
    std::vector<wfn_t<N>> both;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(both), less);
    recall_ = np ? (double)both.size() / np : 1.;
    return recall_;
  }

  // Screened search when the network is trusted, else the plain asci_search.
  std::vector<wfn_t<N>> search(const ASCISettings& s, size_t ndets_max,
                               wavefunction_iterator_t<N> cbeg, wavefunction_iterator_t<N> cend,
                               double E0, const std::vector<double>& C, size_t norb) {
    used_fallback_ = !trained_ || recall_ < min_recall_;
    if(used_fallback_)
      return asci_search(s, ndets_max, cbeg, cend, E0, C, norb, gen_.T(), gen_.G_red(),
                         gen_.V_red(), gen_.G(), gen_.V(), gen_);
    return screened(s, ndets_max, cbeg, cend, E0, C, norb);
  }

  bool used_fallback() const { return used_fallback_; }
  double recall() const { return recall_; }
  double keep_fraction() const { return keep_; }
  void set_keep_fraction(double k) {
    keep_ = k;
    recall_ = 0.;  // judged again at the new fraction
  }
  mhip_mlp* network() const { return mlp_; }
  size_t num_pairs() const { return target_.size(); }

 private:
  std::vector<wfn_t<N>> screened(const ASCISettings& s, size_t ndets_max,
                                 wavefunction_iterator_t<N> cbeg, wavefunction_iterator_t<N> cend,
                                 double E0, const std::vector<double>& C, size_t norb) {
    auto& ctx = gen_.context();
    const size_t nc = std::distance(cbeg, cend);
    auto dc = hip::upload_dets<N>(ctx, cbeg, cend);
    hip::DeviceBuffer<double> dC(ctx, C.data(), std::max<size_t>(nc, 1));
    hip::DeviceBuffer<uint64_t> out(ctx, std::max<size_t>(ndets_max, 1) * W);
    const mhip_asci_settings ms = to_c(s);
    int64_t nnew = 0;
    hip::check(mhip_asci_search_screened(ctx.get(), W, &ms, mlp_, (float)sd_, (float)mu_, keep_,
                                         ndets_max, dc.data(), nc, E0, dC.data(), norb,
                                         out.data(), &nnew));
    return hip::download_dets<N>(out, nnew);
  }
  static mhip_asci_settings to_c(const ASCISettings& s) {
    mhip_asci_settings m;
    m.ncdets_max = s.ncdets_max;
    m.h_el_tol = s.h_el_tol;
    m.rv_prune_tol = s.rv_prune_tol;
    m.pair_size_max = s.pair_size_max;
    m.just_singles = s.just_singles;
    return m;
  }
  HipHamiltonianGenerator<N>& gen_;
  mhip_mlp* mlp_ = nullptr;
  double keep_, min_recall_;
  std::vector<uint64_t> dets_;
  std::vector<float> target_;
  double mu_ = 0., sd_ = 1., recall_ = 0.;
  bool trained_ = false, used_fallback_ = true;
};

}  // namespace macis
