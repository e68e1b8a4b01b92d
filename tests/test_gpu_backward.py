"""The differentiation path on the GPU: jmm_pair_observables /
jmm_observables (U, W, dU/dtheta, dW/dtheta, dU/dx) against the FP64 oracle,
against central finite differences at dtheta = 1e-4, through the
torch.autograd binding, and the fluctuation gradient against reweighting.

Configurations for the finite-difference set follow the C4 recipe: N
positions uniform on the ring conditioned on no gap below g0 (exactly the
uniform law on that set: spacings g0 + (L - N g0) * uniform-spacings)."""
import dataclasses
import math

import numpy as np
import pytest

from helpers import oracle_spec

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-11      # relative, FP64 kernel vs FP64 oracle (summation order differs)
TOL_FP32 = 1e-5       # relative to the sum of |pair terms|, FP32 pair arithmetic
TOL_FD = 1e-4         # relative, analytic vs central differences at 1e-4


def ring_configs(n, N, L, g0, seed):
    rng = np.random.default_rng(seed)
    out = np.empty((n, N))
    for c in range(n):
        u = np.sort(rng.random(N - 1))
        sp = np.diff(np.concatenate([[0.0], u, [1.0]]))          # N spacings, sum 1
        gaps = g0 + (L - N * g0) * sp
        x = np.cumsum(gaps) - gaps[0] - L / 2 + rng.random() * gaps[0]
        out[c] = x
    return out


def c4_chain(eps=1.0, sigma=1.0):
    from jmmonedmc_amd.configs import c4
    return c4(1, eps=eps, sigma=sigma)[0]


def c5_chain(N):
    from jmmonedmc_amd.configs import c5
    return dataclasses.replace(c5(1)[0], N=N)


def _device_obs(chain, X, L, precision, theta=None):
    import torch
    from jmmonedmc_amd import autograd as ag
    th = theta if theta is not None else chain.pot_p[:2]
    xt = torch.tensor(X, dtype=torch.float64, device="cuda")
    Lt = torch.tensor(L, dtype=torch.float64, device="cuda")
    obs, grad = ag._observables(xt, Lt, chain, th, precision, want_grad=True)
    torch.cuda.synchronize()
    return obs.cpu().numpy(), grad.cpu().numpy()


def _oracle_obs(oracle, chain, X, L, theta=None):
    ch = chain
    if theta is not None:
        p = list(chain.pot_p)
        p[0], p[1] = theta
        ch = dataclasses.replace(chain, pot_p=tuple(p))
    obs = np.empty((len(X), 6))
    grad = np.empty_like(X)
    for c in range(len(X)):
        s = oracle_spec(oracle, dataclasses.replace(ch, N=X.shape[1], L0=float(L[c])))
        obs[c], grad[c] = oracle.observables_x(s, X[c], float(L[c]))
    return obs, grad


CASES = {
    "c4_lj": lambda: (c4_chain(), 1024, 1024 / 0.6, 0.8),
    "c5_yukawa": lambda: (c5_chain(4096), 4096, 4096 / 0.8, 0.55),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp64_kernel_matches_oracle(gpu, oracle, case):
    chain, N, L0, g0 = CASES[case]()
    print(f"[gpu] backward fp64 {case} N={N}", flush=True)
    X = ring_configs(8, N, L0, g0, seed=4)
    L = np.full(8, L0)
    obs, grad = _device_obs(chain, X, L, 64)
    ro, rg = _oracle_obs(oracle, chain, X, L)
    scale = np.abs(ro).max(axis=0) + 1.0
    assert np.all(np.abs(obs - ro) <= TOL_FP64 * scale * N), (obs - ro)
    assert np.abs(grad - rg).max() <= TOL_FP64 * np.abs(rg).max() * 10


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp32_kernel_within_tolerance(gpu, oracle, case):
    chain, N, L0, g0 = CASES[case]()
    print(f"[gpu] backward fp32 {case} N={N}", flush=True)
    X = ring_configs(8, N, L0, g0, seed=5)
    L = np.full(8, L0)
    obs, grad = _device_obs(chain, X, L, 32)
    ro, rg = _oracle_obs(oracle, chain, X, L)
    # bound per column: relative to the column's magnitude per pair times the pairs
    mag = np.abs(ro).max(axis=0) + N * 1e-3
    assert np.all(np.abs(obs - ro) <= TOL_FP32 * mag), (obs - ro) / mag
    assert np.abs(grad - rg).max() <= 10 * TOL_FP32 * np.abs(rg).max()


def test_configurations_inside_the_hard_core(gpu, oracle):
    """Configurations handed to the backward need not keep the hard core:
    LJ rings with gaps down to 0.3 sigma (the exclusion is 0.5 sigma) put
    up to 10 partners within the cutoff where the chain's window holds 6;
    the FP64 kernel walks every partner to the cutoff and its U equals the
    all-pairs energy (the windowed sum does not)."""
    chain = c4_chain()
    N, g0 = 128, 0.3
    L0 = N * 0.36
    print(f"[gpu] backward below the hard core N={N}", flush=True)
    X = ring_configs(4, N, L0, g0, seed=9)
    L = np.full(4, L0)
    obs, _ = _device_obs(chain, X, L, 64)
    missed = 0
    for c in range(4):
        s = oracle_spec(oracle, dataclasses.replace(chain, N=N, L0=L0))
        ub = oracle.brute_energy_x(s, X[c], L0)
        assert abs(obs[c, 0] - ub) <= 1e-11 * abs(ub), (c, obs[c, 0], ub)
        missed += oracle.full_energy_x(s, X[c], L0) != ub
    assert missed > 0


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("precision", [32, 64])
def test_theta_gradients_match_finite_differences(gpu, case, precision):
    """dU/dtheta and dW/dtheta (FP32 or FP64 kernel) against central
    differences of the FP64 kernel's U and W at dtheta = 1e-4."""
    chain, N, L0, g0 = CASES[case]()
    print(f"[gpu] backward FD theta {case} fp{precision}", flush=True)
    X = ring_configs(8, N, L0, g0, seed=6)
    L = np.full(8, L0)
    th = np.array(chain.pot_p[:2], dtype=float)
    obs, _ = _device_obs(chain, X, L, precision)
    h = 1e-4
    for k in range(2):
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        op, _ = _device_obs(chain, X, L, 64, theta=tp)
        om, _ = _device_obs(chain, X, L, 64, theta=tm)
        fd = (op - om) / (2 * h)
        for col, dcol in ((0, 2 + k), (1, 4 + k)):
            err = np.abs(obs[:, dcol] - fd[:, col])
            assert np.all(err <= TOL_FD * (np.abs(fd[:, col]) + 1e-3 * N)), (case, k, col, err)


def test_position_gradient_matches_finite_differences(gpu):
    chain, N, L0, g0 = CASES["c4_lj"]()
    X = ring_configs(2, N, L0, g0, seed=8)
    L = np.full(2, L0)
    _, grad = _device_obs(chain, X, L, 64)
    rng = np.random.default_rng(1)
    h = 1e-6
    idx = rng.choice(N, 12, replace=False)
    for i in idx:
        Xp, Xm = X.copy(), X.copy()
        Xp[:, i] += h
        Xm[:, i] -= h
        up, _ = _device_obs(chain, Xp, L, 64)
        um, _ = _device_obs(chain, Xm, L, 64)
        fd = (up[:, 0] - um[:, 0]) / (2 * h)
        assert np.all(np.abs(grad[:, i] - fd) <= 1e-6 * (np.abs(fd) + 1.0)), (i, grad[:, i], fd)


def test_autograd_binding(gpu):
    """PairEnergy: forward = U per configuration, backward = dU/dx and the
    summed dU/dtheta, equal to the kernel outputs."""
    import torch
    from jmmonedmc_amd.autograd import PairEnergy
    chain, N, L0, g0 = CASES["c4_lj"]()
    X = ring_configs(4, N, L0, g0, seed=9)
    L = np.full(4, L0)
    obs, grad = _device_obs(chain, X, L, 64)
    x = torch.tensor(X, dtype=torch.float64, device="cuda", requires_grad=True)
    Lt = torch.tensor(L, dtype=torch.float64, device="cuda")
    th = torch.tensor([1.0, 1.0], dtype=torch.float64, requires_grad=True)
    U = PairEnergy.apply(x, Lt, th, chain, 64)
    wts = torch.tensor([1.0, -2.0, 0.5, 3.0], dtype=torch.float64, device="cuda")
    (U * wts).sum().backward()
    w = wts.cpu().numpy()
    assert np.allclose(U.detach().cpu().numpy(), obs[:, 0], rtol=0, atol=1e-12 * N)
    assert np.allclose(x.grad.cpu().numpy(), w[:, None] * grad, rtol=1e-14, atol=1e-14)
    assert np.allclose(th.grad.numpy(), (w[:, None] * obs[:, 2:4]).sum(0), rtol=1e-13)


def test_engine_observables_and_reweighting(gpu, oracle):
    """Configurations sampled by the engine (16 periodic LJ chains, NVT):
    jmm_observables equals the oracle on the same states, and the
    fluctuation gradient of <W> equals the derivative of the reweighted
    average on the same samples (an identity at dtheta -> 0)."""
    import torch
    from jmmonedmc_amd.autograd import fluctuation_gradient
    from jmmonedmc_amd.engine import Chain, Engine
    from helpers import lj
    torch.cuda.init()  # torch's device first (its own HIP runtime), then the engine's
    chains = [Chain(N=256, rho=0.6, T=1.0, seed=3, chain_id=c, **lj()) for c in range(16)]
    snaps_x, snaps_L = [], []
    with Engine(chains, form="batched_fp32", block_sweeps=50) as eng:
        for _ in range(6):
            eng.run(50)
            xs, Ls, _ = eng.states()
            snaps_x.append(np.stack(xs))
            snaps_L.append(Ls.copy())
        obs64 = eng.observables()
        xs, Ls, _ = eng.states()
    ro, _ = _oracle_obs(oracle, chains[0], np.stack(xs), Ls)
    assert np.all(np.abs(obs64 - ro) <= TOL_FP32 * (np.abs(ro).max(axis=0) + 1.0))
    X = np.concatenate(snaps_x)
    L = np.concatenate(snaps_L)
    base, _ = _device_obs(chains[0], X, L, 64)
    beta = 1.0
    fg = fluctuation_gradient(base[:, 1], base[:, 4:6], base[:, 2:4], beta)
    # the identity is exact as h -> 0; central differences leave h^2 terms of
    # order (beta h dU/dtheta)^2, so h is kept small (FP64 U: no cancellation
    # problem at this size)
    h = 1e-5
    for k in range(2):
        vals = []
        for sgn in (1, -1):
            th = np.array([1.0, 1.0])
            th[k] += sgn * h
            o, _ = _device_obs(chains[0], X, L, 64, theta=th)
            wgt = np.exp(-beta * (o[:, 0] - base[:, 0]))
            vals.append((o[:, 1] * wgt).sum() / wgt.sum())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert fg[k] == pytest.approx(fd, rel=1e-5, abs=1e-6), (k, fg[k], fd)
