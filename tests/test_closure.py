"""The learned closure (csrc/closure_mlp.hip, include/mercurydpm_closure.h):
the oracle's gradient against central finite differences (CPU), the HIP
forward pass and gradient against the oracle (GPU, 1e-12), and training on
samples of the local rheology mu(I), phi(I): the trained network predicts mu
and phi of held-out states far better than the constant closure it replaces
(the mean of the training targets)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import closure_oracle as O  # noqa: E402

from mercurydpm_fork_amd import closure  # noqa: E402


def samples(n, seed):
    """States of steady dense flow: I log-uniform in [1e-3, 0.5], phi and mu
    from the local rheology with 1% noise; pressure and temperature follow
    (P from depth, T ~ (d gamma_dot)^2); inputs standardised."""
    rng = np.random.default_rng(seed)
    I = np.exp(rng.uniform(np.log(1e-3), np.log(0.5), n))
    P = rng.uniform(0.5, 20.0, n)
    gdot = I * np.sqrt(P)  # d = rho = 1
    T = 0.1 * gdot ** 2 * (1 + 0.1 * rng.standard_normal(n))
    phi = closure.phi_of_I(I) * (1 + 0.01 * rng.standard_normal(n))
    mu = closure.mu_of_I(I) * (1 + 0.01 * rng.standard_normal(n))
    x = np.stack([np.log10(I), phi, np.log10(P), np.log10(np.abs(T) + 1e-12)], axis=1)
    x = (x - np.array([-1.6, 0.55, 0.6, -3.5])) / np.array([0.8, 0.05, 0.4, 1.5])
    t = np.stack([(mu - 0.5) / 0.1, (phi - 0.55) / 0.05], axis=1)
    return x, t


def test_oracle_gradient_against_finite_differences():
    x, t = samples(64, 1)
    p = O.init(3)
    _, g = O.loss_and_gradient(p, x, t)
    rng = np.random.default_rng(0)
    for k in rng.choice(O.N_PARAMS, 40, replace=False):
        e = np.zeros(O.N_PARAMS)
        e[k] = 1e-6
        fd = (O.loss_and_gradient(p + e, x, t)[0] - O.loss_and_gradient(p - e, x, t)[0]) / 2e-6
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])
    assert O.N_PARAMS == closure.N_PARAMS
    np.testing.assert_array_equal(O.init(3), closure.glorot(3))


@pytest.mark.gpu
def test_kernels_against_oracle():
    x, t = samples(1000, 2)
    net = closure.MLP(seed=5)
    p = net.params()
    y = net(x)
    np.testing.assert_allclose(y, O.forward(p, x)[0], rtol=1e-12, atol=1e-12)
    loss, g = net.gradient(x, t)
    lo, go = O.loss_and_gradient(p, x, t)
    assert abs(loss - lo) <= 1e-12 * lo
    np.testing.assert_allclose(g, go, rtol=1e-10, atol=1e-13 * np.abs(go).max())


@pytest.mark.gpu
def test_trained_closure_beats_constant():
    xtr, ttr = samples(4000, 3)
    xte, tte = samples(1000, 4)
    net = closure.MLP(seed=7)
    first = net.gradient(xtr, ttr)[0]
    net.train(xtr, ttr, epochs=1500, lr=3e-3)
    last = net.gradient(xtr, ttr)[0]
    assert last < 0.05 * first
    err_net = np.mean((net(xte) - tte) ** 2, axis=0)
    err_const = np.mean((ttr.mean(0) - tte) ** 2, axis=0)
    assert np.all(err_net < 0.05 * err_const), (err_net, err_const)


def chute_data():
    """Bulk CG states of the engine's own chute flows at 21..27 degrees
    (tests/golden/make_closure_data.py)."""
    from mercurydpm_fork_amd import closure_data as CD
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "closure_chute_cg.npz"))
    return {float(k.split("_")[1]): CD.standardise(d[k]) for k in d.files}


def test_chute_data_follow_a_friction_law():
    """The profiles' effective friction grows with the inertial number, as
    the local rheology says: the bin medians per angle are ordered."""
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "closure_chute_cg.npz"))
    med = np.array([np.median(d[k], axis=0) for k in sorted(d.files, key=lambda k: float(k.split("_")[1]))])
    assert np.all(np.diff(med[:, 0]) > -0.01) and med[-1, 0] > 2 * med[0, 0]  # I grows with the angle
    assert np.corrcoef(np.log(med[:, 0]), med[:, 4])[0, 1] > 0.8  # and mu with I


@pytest.mark.gpu
def test_closure_trained_on_dem_profiles_predicts_held_out_angles():
    """Trained on the CG states of five chute angles (50 full-batch epochs:
    the per-bin states are noisy, and longer training fits the noise), the
    network predicts mu at the two held-out angles (23 and 26 degrees) with
    at most 0.8 of the squared error of the constant closure (the mean
    training mu); the numpy oracle with the same updates gives 0.67."""
    data = chute_data()
    train = [a for a in data if a not in (23.0, 26.0)]
    xtr = np.concatenate([data[a][0] for a in train])
    ttr = np.concatenate([data[a][1] for a in train])
    xte = np.concatenate([data[a][0] for a in (23.0, 26.0)])
    tte = np.concatenate([data[a][1] for a in (23.0, 26.0)])
    net = closure.MLP(seed=11)
    net.train(xtr, ttr, epochs=50, lr=2e-3)
    err_net = np.mean((net(xte)[:, 0] - tte[:, 0]) ** 2)
    err_const = np.mean((ttr[:, 0].mean() - tte[:, 0]) ** 2)
    print(f"held-out mu error: network {err_net:.4g}, constant {err_const:.4g}")
    assert err_net < 0.8 * err_const
