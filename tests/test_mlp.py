"""CPU checks of the MLP reference and the weight-file format."""
import numpy as np

import mlp_ref as M


def test_numpy_backward_matches_finite_differences():
    g = np.random.default_rng(0)
    P = M.init_params(1).astype(np.float64)
    P[768:832] = g.normal(scale=0.1, size=64)
    P[4928:4992] = g.normal(scale=0.1, size=64)
    P[5056] = 0.3
    norm = np.concatenate([g.normal(size=12), g.uniform(0.5, 2, size=12)])
    X = g.normal(size=(37, 12)) * norm[12:] + norm[:12]
    T = g.normal(size=37)
    _, grad = M.loss_grad(P, norm, X, T, emulate=False)
    idx = np.concatenate([g.choice(768, 20), 768 + g.choice(64, 5), 832 + g.choice(4096, 20),
                          4928 + g.choice(64, 5), 4992 + g.choice(64, 5), [5056]])
    h = 1e-6
    for k in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        fd = (M.loss_grad(Pp, norm, X, T)[0] - M.loss_grad(Pm, norm, X, T)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, grad[k])


def test_weight_file_roundtrip(tmp_path):
    from fastlinkingnumbers_amd import criterion
    P = np.arange(5057, dtype=np.float32) * 0.5
    nm = np.linspace(1, 2, 24).astype(np.float32)
    p = str(tmp_path / "w.bin")
    criterion.save_weights(p, P, nm)
    P2, n2 = criterion.load_weights(p)
    assert np.array_equal(P, P2) and np.array_equal(nm, n2)


def test_base2_inference_form_is_the_same_network():
    """The inference kernel's base-2 rewrite (log2 e folded into layer 1 and
    the biases, ln 2 into the output weights, activations z / (1 + 2^-z)) is
    the SiLU network itself in exact arithmetic."""
    g = np.random.default_rng(2)
    P = M.init_params(7).astype(np.float64)
    P[768:832] = g.normal(scale=0.3, size=64)
    P[4928:4992] = g.normal(scale=0.3, size=64)
    P[5056] = -1.5
    norm = np.concatenate([g.normal(size=12), g.uniform(0.5, 2, size=12)])
    X = g.normal(size=(50, 12)) * norm[12:] + norm[:12]
    y, _ = M.forward(P, norm, X)
    y2 = M.forward_infer(P, norm, X, rounding=False)
    np.testing.assert_allclose(y2, y, rtol=1e-12, atol=1e-12)
