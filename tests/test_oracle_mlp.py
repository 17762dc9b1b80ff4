"""Pins the MLP oracle's backward pass by central finite differences (the
learned pre-screen has no reference counterpart; the mathematics fixes it)."""
import numpy as np

from oracle import oracle as O


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    npar = O.lib().orc_mlp_num_params()
    assert npar == 512 * 128 + 512 + 512 * 512 + 512 + 512 + 1
    P = (rng.standard_normal(npar) * 0.05).astype(np.float32)
    d = rng.integers(0, 2**63, size=(6, 2), dtype=np.uint64)
    t = rng.standard_normal(6).astype(np.float32)
    L, g, _ = O.mlp_loss_grad(P, 2, d, t, bf16_points=False)
    idx = list(rng.integers(0, npar, 12)) + [npar - 1, npar - 2, 512 * 128 + 3]
    for i in idx:
        P1, P2 = P.copy(), P.copy()
        P1[i] += 1e-3
        P2[i] -= 1e-3
        L1, _, _ = O.mlp_loss_grad(P1, 2, d, t, bf16_points=False, want_grad=False)
        L2, _, _ = O.mlp_loss_grad(P2, 2, d, t, bf16_points=False, want_grad=False)
        fd = (L1 - L2) / (float(P1[i]) - float(P2[i]))
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_mlp_oracle_bf16_points_close_to_exact():
    rng = np.random.default_rng(1)
    npar = O.lib().orc_mlp_num_params()
    P = (rng.standard_normal(npar) * 0.05).astype(np.float32)
    d = rng.integers(0, 2**63, size=(5, 2), dtype=np.uint64)
    _, _, p64 = O.mlp_loss_grad(P, 2, d, None, bf16_points=False, want_grad=False)
    _, _, p16 = O.mlp_loss_grad(P, 2, d, None, bf16_points=True, want_grad=False)
    assert np.allclose(p16, p64, atol=0.05 * np.abs(p64).max())
