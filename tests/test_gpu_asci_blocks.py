"""GPU parity of the block screen's kernels against the oracle:

- mhip_asci_blocks: block determinants bit-exact, targets (FP32 logs of the
  largest |h / (E - H_kk)|) to FP32 rounding, -inf where the oracle has no
  candidate;
- mhip_asci_search_blocks with the same block scores: the same kept blocks,
  so the top-k equals the oracle's block search up to the tie band at the
  k-th score;
- keep_fraction 1: the GPU block search returns the plain GPU search's
  output, bit for bit (the contributions to one determinant come from
  distinct core determinants, in core order either way).
"""
import numpy as np
import pytest

from macis_amd import _capi
from macis_amd import dets as D
from macis_amd.synthetic import seeded_integrals, selected_ci_dets
from oracle import oracle as O

from conftest import announce
from test_gpu_asci import check_topk

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h2o_case(h2o, h2o_oracle):
    hf = D.canonical_hf_determinant(1, 5, 5)
    E0 = h2o_oracle.matrix_element(1, hf, hf)
    s = O.ASCISettings.defaults()
    E1, w, X = h2o_oracle.asci_iter(1, s, 100, E0, hf, np.array([1.0]))
    E2, w, X = h2o_oracle.asci_iter(1, s, 800, E1, w, X)
    order = np.argsort(-np.abs(X), kind="stable")
    norb, _, T, V = h2o
    return 1, norb, T, V, h2o_oracle, w[order[:60]], X[order[:60]], E2


@pytest.fixture(scope="module")
def wide_case():
    norb = 64
    T, V = seeded_integrals(norb, 31)
    orc = O.Ham(T, V)
    core = selected_ci_dets(norb, 20, 20, 6, seed=9, T=T)
    C = np.random.default_rng(11).standard_normal(len(core))
    E = orc.matrix_element(2, core[0], core[0]) - 0.5
    return 2, norb, T, V, orc, core, C / np.linalg.norm(C), E


def _case(request, name):
    return request.getfixturevalue("h2o_case" if name == "h2o" else "wide_case")


@pytest.mark.parametrize("name", ["h2o", "wide"])
def test_blocks_match_oracle(ctx, request, name):
    W, norb, T, V, orc, core, C, E = _case(request, name)
    announce(f"asci blocks vs oracle ({name}, {len(core)} core dets, W={W})")
    ctx.set_integrals(norb, T, V)
    s = _capi.ASCISettings.defaults()
    bd_ref, t_ref = orc.asci_blocks(W, O.ASCISettings.defaults(), core, E)
    d_bd, d_t, n = ctx.asci_blocks(W, s, ctx.to_device(core), len(core), E, norb)
    assert n == len(t_ref)
    assert np.array_equal(d_bd.numpy()[:n], bd_ref)
    t = d_t.numpy()[:n]
    fin = np.isfinite(t_ref)
    assert np.array_equal(np.isfinite(t), fin)
    np.testing.assert_allclose(t[fin], t_ref[fin].astype(np.float32), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("name", ["h2o", "wide"])
@pytest.mark.parametrize("keep", [0.3, 0.05])
def test_search_blocks_match_oracle(ctx, request, name, keep):
    W, norb, T, V, orc, core, C, E = _case(request, name)
    announce(f"asci block search vs oracle ({name}, keep {keep})")
    ctx.set_integrals(norb, T, V)
    s_orc, s_gpu = O.ASCISettings.defaults(), _capi.ASCISettings.defaults()
    _, t_ref = orc.asci_blocks(W, s_orc, core, E)
    scores = np.random.default_rng(5).standard_normal(len(t_ref)).astype(np.float32)
    nd = 3000 if name == "h2o" else 20000
    ref, ref_rv = orc.asci_search_blocks(W, s_orc, scores, keep, nd, core, E, C)
    out = ctx.empty((nd, W), np.uint64)
    n = ctx.asci_search_blocks(W, s_gpu, ctx.to_device(scores), keep, nd, ctx.to_device(core),
                               len(core), E, ctx.to_device(C), norb, out)
    tm = ctx.timings()
    assert tm["asci_nblocks"] == len(t_ref)
    assert tm["asci_nblocks_kept"] == min(len(t_ref), max(1, int(np.ceil(keep * len(t_ref)))))
    check_topk(out.numpy()[:n], ref, ref_rv, len(core))


@pytest.mark.parametrize("name", ["h2o", "wide"])
def test_keep_all_blocks_is_the_plain_search(ctx, request, name):
    W, norb, T, V, orc, core, C, E = _case(request, name)
    announce(f"asci block search, every block kept ({name})")
    ctx.set_integrals(norb, T, V)
    s = _capi.ASCISettings.defaults()
    dcore, dC = ctx.to_device(core), ctx.to_device(C)
    _, _, nblk = ctx.asci_blocks(W, s, dcore, len(core), E, norb, targets=False)
    scores = ctx.to_device(np.zeros(nblk, np.float32))
    for nd in (2000, 40000):
        a = ctx.empty((nd, W), np.uint64)
        b = ctx.empty((nd, W), np.uint64)
        na = ctx.asci_search(W, s, nd, dcore, len(core), E, dC, norb, a)
        npairs = ctx.timings()["asci_npairs"]
        nb = ctx.asci_search_blocks(W, s, scores, 1.0, nd, dcore, len(core), E, dC, norb, b)
        assert ctx.timings()["asci_npairs"] == npairs
        assert na == nb and np.array_equal(a.numpy()[:na], b.numpy()[:nb])
