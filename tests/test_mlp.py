"""CPU checks of the MLP reference and the weight-file format."""
import numpy as np
import pytest

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


def test_weight_file_of_the_logistic_network_is_refused(tmp_path):
    """Files of the earlier SiLU network (magic FLKMLP1) hold weights of a
    different function: loading one is an error, not a silent mismatch."""
    from fastlinkingnumbers_amd import criterion
    p = tmp_path / "old.bin"
    p.write_bytes(b"FLKMLP1\0" + np.zeros(5057 + 24, "<f4").tobytes())
    with pytest.raises(ValueError, match="retrain"):
        criterion.load_weights(str(p))


def test_hard_swish_closed_forms():
    """h(z) = z ReLU6(z + 3) / 6: zero at and below -3, the identity at and
    above 3, h(1) = 2/3, h(-1) = -1/3; the derivative matches central
    differences away from the two kinks."""
    z = np.linspace(-9, 9, 1801)
    relu6 = np.minimum(np.maximum(z + 3, 0), 6)
    np.testing.assert_allclose(M.hswish(z), z * relu6 / 6, rtol=0, atol=1e-15)
    assert np.all(M.hswish(z[z <= -3]) == 0) and np.array_equal(M.hswish(z[z >= 3]), z[z >= 3])
    np.testing.assert_allclose(M.hswish(np.array([1.0, -1.0, 0.0])), [2 / 3, -1 / 3, 0.0], rtol=0, atol=1e-15)
    zs = z[(np.abs(np.abs(z) - 3) > 1e-3)]
    h = 1e-6
    fd = (M.hswish(zs + h) - M.hswish(zs - h)) / (2 * h)
    np.testing.assert_allclose(M.dhswish(zs), fd, rtol=0, atol=1e-8)


def test_fp16_inference_form_tracks_the_network():
    """The inference kernels' FP16 form (operands and activations rounded to
    binary16, FP32 accumulation) stays within 1e-2 (1 + |y|) of the float64
    network; its activation within 2^-9 |z| + 2^-20 of hard-swish (three
    FP16 roundings -- of z, the gate and the product -- and the FP16 slope
    1/6; measured 1.3e-3 |z| for |z| > 0.01, and FP16 subnormals below
    6e-5)."""
    g = np.random.default_rng(2)
    z = g.normal(scale=4.0, size=100000)
    assert np.all(np.abs(M.hswish16(z) - M.hswish(z)) <= 2.0 ** -9 * np.abs(z) + 2.0 ** -20)
    P = M.init_params(7).astype(np.float64)
    P[768:832] = g.normal(scale=0.3, size=64)
    P[4928:4992] = g.normal(scale=0.3, size=64)
    P[5056] = -1.5
    norm = np.concatenate([g.normal(size=12), g.uniform(0.5, 2, size=12)])
    X = g.normal(size=(500, 12)) * norm[12:] + norm[:12]
    y, _ = M.forward(P, norm, X)
    y16 = M.forward_infer(P, norm, X)
    print("max |fp16 form - float64|", np.abs(y16 - y).max())
    assert np.all(np.abs(y16 - y) <= 1e-2 * (1 + np.abs(y)))
