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
    # per-block partial sums added in block order: the same bits every call
    loss2, g2 = net.gradient(x, t)
    assert loss2 == loss
    np.testing.assert_array_equal(g2, g)


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


# ---------------------------------------------------------------------------
# The configured network: 4 -> 3 x 128 GELU -> 2, BF16 weights, matrix cores
# ---------------------------------------------------------------------------

def test_oracle128_gradient_against_finite_differences():
    """The back-propagation of the 128-wide network (float64, no rounding)
    against central finite differences, and its parameter count and
    initialisation against the binding's."""
    rng = np.random.default_rng(0)
    x = rng.standard_normal((40, 4))
    t = rng.standard_normal((40, 2))
    p = O.init128(3).astype(np.float64)
    _, g = O.loss_and_gradient128(p, x, t, quantize=False)
    for k in rng.choice(O.N_PARAMS128, 60, replace=False):
        e = np.zeros(O.N_PARAMS128)
        e[k] = 1e-6
        fd = (O.loss_and_gradient128(p + e, x, t, False)[0] - O.loss_and_gradient128(p - e, x, t, False)[0]) / 2e-6
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])
    assert O.N_PARAMS128 == closure.N_PARAMS128
    np.testing.assert_array_equal(O.init128(3), closure.glorot128(3))


def test_oracle_bf16_rounding():
    """Round to nearest even: exact values, ties to the even mantissa."""
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -2.5e-3, 3.0e38], np.float32)
    # bit patterns: 1.0, tie down to even, up, tie up to even, a negative
    # value, and 3e38 = 1.7627 x 2^127 -> 226/128 x 2^127
    want = np.array([0x3F800000, 0x3F800000, 0x3F810000, 0x3F820000, 0xBB240000, 0x7F620000], np.uint32)
    np.testing.assert_array_equal(O.bf16(v).view(np.uint32), want)


@pytest.mark.gpu
def test_closure128_kernels_against_oracle():
    """Forward pass and gradient of the matrix-core kernels against the
    float32 oracle with the same BF16 roundings.  The FP32 sums run in
    another order, so now and then an activation near a BF16 rounding
    boundary rounds the other way and moves by one BF16 step (2^-8
    relative): 99% of the outputs agree within 1e-5 of their scale and all
    within 1e-3 (measured: 0.6% of them beyond 1e-5, at most 5e-4); the
    gradient within 1e-3 of its largest entry, the loss within 1e-4."""
    x, t = samples(1000, 2)
    net = closure.Closure(seed=5)
    p = net.params()
    y = net(x)
    yo = O.forward128(p, x)[0]
    scale = np.abs(yo).max()
    err = np.abs(y - yo)
    assert np.mean(err <= 1e-5 * scale) >= 0.99, np.mean(err <= 1e-5 * scale)
    assert err.max() <= 1e-3 * scale, err.max() / scale
    loss, g = net.gradient(x, t)
    lo, go = O.loss_and_gradient128(p, x, t)
    assert abs(loss - lo) <= 1e-4 * lo
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-3 * np.abs(go).max())
    # and a sample count that is no multiple of the 16-row tiles
    y7 = net(x[:7])
    np.testing.assert_array_equal(y7, y[:7])
    # chunked weight gradients summed in chunk order: the same bits every call
    loss2, g2 = net.gradient(x, t)
    assert loss2 == loss
    np.testing.assert_array_equal(g2, g)


@pytest.mark.gpu
def test_closure128_trained_beats_constant():
    xtr, ttr = samples(4000, 3)
    xte, tte = samples(1000, 4)
    net = closure.Closure(seed=7)
    first = net.gradient(xtr, ttr)[0]
    net.train(xtr, ttr, epochs=600, lr=1e-3)
    last = net.gradient(xtr, ttr)[0]
    assert last < 0.05 * first
    err_net = np.mean((net(xte) - tte) ** 2, axis=0)
    err_const = np.mean((ttr.mean(0) - tte) ** 2, axis=0)
    print(f"128-wide closure, synthetic rheology: error {err_net} against constant {err_const}")
    assert np.all(err_net < 0.05 * err_const), (err_net, err_const)


def friction_data():
    """{(friction, angle): states} of tests/golden/closure_chute_friction.npz."""
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "closure_chute_friction.npz"))
    out = {}
    for k in d.files:
        mu, angle = k[2:].split("_angle")
        out[(float(mu), float(angle))] = d[k].astype(np.float64)
    return out


def test_friction_data_follow_the_rheology():
    """The chute states over friction and angle (the engine's own flows):
    steady flows carry mu_eff = tan(theta), so the median mu_eff grows with
    the angle at every friction; at one angle the flow of less frictional
    spheres is faster (a larger inertial number)."""
    data = friction_data()
    for fr in (0.3, 0.6):
        angles = sorted(a for f, a in data if f == fr)
        med = np.array([np.median(data[(fr, a)][:, 4]) for a in angles])
        assert np.all(np.diff(med) > 0), (fr, med)
    for a in (22.0, 26.0):
        I = [np.median(data[(fr, a)][:, 0]) for fr in (0.3, 0.45, 0.6)]
        assert I[0] > I[1] > I[2], (a, I)


@pytest.mark.gpu
def test_closure128_on_held_out_angle_and_friction():
    """The configured network trained on the chute states of frictions 0.3
    and 0.6 at 20-28 degrees except 24.  Held-out angle: at 24 degrees (both
    frictions) the steady flow carries mu_eff = tan(24) at every depth (force
    balance), and the network's mean prediction over the run's states must
    give it within 0.01.  Held-out friction (0.45, at 22 and 26 degrees;
    friction is no input of the network, so it must be read from the state):
    per state, a fraction of the squared mu_eff error of the constant closure
    (the mean training mu_eff)."""
    from mercurydpm_fork_amd import closure_data as CD
    data = friction_data()
    train = [k for k in data if k[0] in (0.3, 0.6) and k[1] != 24.0]
    xtr, ttr = CD.standardise(np.concatenate([data[k] for k in train]))
    net = closure.Closure(seed=3)
    net.train(xtr, ttr, epochs=2000, lr=1e-3)
    for key in ((0.3, 24.0), (0.6, 24.0)):
        x, t = CD.standardise(data[key])
        mu_net = float(np.mean(0.5 + 0.1 * net(x)[:, 0]))
        print(f"held-out angle 24, friction {key[0]}: mean predicted mu_eff {mu_net:.4f}, "
              f"DEM median {np.median(data[key][:, 4]):.4f}, tan(24) {np.tan(np.radians(24.0)):.4f}")
        assert abs(mu_net - np.tan(np.radians(24.0))) < 0.01, (key, mu_net)
    xte, tte = CD.standardise(np.concatenate([data[(0.45, 22.0)], data[(0.45, 26.0)]]))
    err_net = np.mean((net(xte)[:, 0] - tte[:, 0]) ** 2)
    err_const = np.mean((ttr[:, 0].mean() - tte[:, 0]) ** 2)
    print(f"held-out friction 0.45: mu_eff error per state {err_net:.4g} against the constant closure's "
          f"{err_const:.4g}")
    assert err_net < 0.5 * err_const, (err_net, err_const)
