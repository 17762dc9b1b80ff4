// Learned ASCI pre-screen (no counterpart in the reference): an MLP on the
// occupation bits predicts log|rv| of asci_search candidates; the screened
// search keeps the best-predicted fraction and lets the exact scores pick
// the top-k. Training pairs come from plain searches (observe), training
// runs on the device (BF16 MFMA kernels), and the plain asci_search is the
// fallback whenever the network is untrained or its recall on the last
// observed search is below `min_recall`.
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

  // Scores all candidates of the core with the plain search and keeps them
  // (determinant, log|rv|) as training data; returns the number added.
  size_t observe(const ASCISettings& s, wavefunction_iterator_t<N> cbeg,
                 wavefunction_iterator_t<N> cend, double E0, const std::vector<double>& C,
                 size_t norb, size_t capacity) {
    auto& ctx = gen_.context();
    const size_t nc = std::distance(cbeg, cend);
    auto dc = hip::upload_dets<N>(ctx, cbeg, cend);
    hip::DeviceBuffer<double> dC(ctx, C.data(), std::max<size_t>(nc, 1));
    hip::DeviceBuffer<uint64_t> dd(ctx, std::max<size_t>(capacity, 1) * W);
    hip::DeviceBuffer<double> da(ctx, std::max<size_t>(capacity, 1));
    const mhip_asci_settings ms = to_c(s);
    int64_t n = 0;
    hip::check(mhip_asci_candidates(ctx.get(), W, &ms, dc.data(), nc, E0, dC.data(), norb,
                                    dd.data(), da.data(), capacity, &n));
    std::vector<uint64_t> w(n * W);
    std::vector<double> a(n);
    if(n) {
      dd.download(w.data(), n * W);
      da.download(a.data(), n);
    }
    dets_.insert(dets_.end(), w.begin(), w.end());
    last_dets_ = std::move(w);
    last_logrv_.resize(n);
    for(int64_t i = 0; i < n; ++i) {
      last_logrv_[i] = std::log(a[i]);
      target_.push_back((float)last_logrv_[i]);
    }
    return n;
  }

  // Adam epochs over the observed pairs; returns the last mini-batch loss.
  // Targets below the clip_quantile quantile are raised to it (the ranking
  // among candidates that cannot enter the top-k does not matter), then
  // standardised.
  double train(int epochs, int64_t batch, float lr, uint64_t seed = 1,
               double clip_quantile = 0.7) {
    const size_t n = target_.size();
    if(!n) return 0.;
    std::vector<float> sorted(target_);
    const size_t qi = std::min(n - 1, (size_t)(clip_quantile * (double)(n - 1)));
    std::nth_element(sorted.begin(), sorted.begin() + qi, sorted.end());
    clip_ = sorted[qi];
    double mu = 0., var = 0.;
    for(float t : target_) mu += std::max(t, clip_);
    mu /= n;
    for(float t : target_) var += (std::max(t, clip_) - mu) * (std::max(t, clip_) - mu);
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
        ht[i] = (float)((std::max(target_[p], clip_) - mu_) / sd_);
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
    recall_ = validate();
    return loss;
  }

  // Recall of the top-k (k = the last search's size minus the core) among
  // the best-predicted keep_fraction of the last observed candidates.
  double validate(size_t k = 0) {
    const size_t n = last_logrv_.size();
    if(!n) return 0.;
    if(!k) k = std::max<size_t>(1, (size_t)(0.01 * n));
    auto& ctx = gen_.context();
    hip::DeviceBuffer<uint64_t> dd(ctx, last_dets_.data(), n * W);
    hip::DeviceBuffer<float> dp(ctx, n);
    hip::check(mhip_mlp_predict(ctx.get(), mlp_, W, dd.data(), n, dp.data()));
    std::vector<float> pred(n);
    dp.download(pred.data(), n);
    std::vector<uint32_t> byrv(n), bypred(n);
    std::iota(byrv.begin(), byrv.end(), 0);
    std::iota(bypred.begin(), bypred.end(), 0);
    k = std::min(k, n);
    std::partial_sort(byrv.begin(), byrv.begin() + k, byrv.end(),
                      [&](uint32_t a, uint32_t b) { return last_logrv_[a] > last_logrv_[b]; });
    const size_t keep = std::min(n, std::max(k, (size_t)std::ceil(keep_ * n)));
    std::nth_element(bypred.begin(), bypred.begin() + keep - 1, bypred.end(),
                     [&](uint32_t a, uint32_t b) { return pred[a] > pred[b]; });
    std::vector<char> kept(n, 0);
    for(size_t i = 0; i < keep; ++i) kept[bypred[i]] = 1;
    size_t hit = 0;
    for(size_t i = 0; i < k; ++i) hit += kept[byrv[i]];
    return (double)hit / k;
  }

  // Screened search, or the plain asci_search when the network is not
  // trusted (untrained, or validation recall below min_recall).
  std::vector<wfn_t<N>> search(const ASCISettings& s, size_t ndets_max,
                               wavefunction_iterator_t<N> cbeg, wavefunction_iterator_t<N> cend,
                               double E0, const std::vector<double>& C, size_t norb) {
    used_fallback_ = !trained_ || recall_ < min_recall_;
    if(used_fallback_)
      return asci_search(s, ndets_max, cbeg, cend, E0, C, norb, gen_.T(), gen_.G_red(),
                         gen_.V_red(), gen_.G(), gen_.V(), gen_);
    auto& ctx = gen_.context();
    const size_t nc = std::distance(cbeg, cend);
    auto dc = hip::upload_dets<N>(ctx, cbeg, cend);
    hip::DeviceBuffer<double> dC(ctx, C.data(), std::max<size_t>(nc, 1));
    hip::DeviceBuffer<uint64_t> out(ctx, std::max<size_t>(ndets_max, 1) * W);
    const mhip_asci_settings ms = to_c(s);
    int64_t nnew = 0;
    hip::check(mhip_asci_search_screened(ctx.get(), W, &ms, mlp_, keep_, ndets_max, dc.data(), nc,
                                         E0, dC.data(), norb, out.data(), &nnew));
    return hip::download_dets<N>(out, nnew);
  }

  bool used_fallback() const { return used_fallback_; }
  double recall() const { return recall_; }
  mhip_mlp* network() const { return mlp_; }

 private:
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
  std::vector<uint64_t> last_dets_;
  std::vector<double> last_logrv_;
  double mu_ = 0., sd_ = 1., recall_ = 0.;
  float clip_ = 0.f;
  bool trained_ = false, used_fallback_ = true;
};

}  // namespace macis
