"""Per-type parameter gradients on the water + ions box (1000 waters, 20 Na+,
20 Cl-), with Ewald (as the benchmark definition sets it) and with P3M
(real-space and ik differentiation).

Reference forces come from the FP64 oracle at parameters perturbed by a
seeded 5% Gaussian; L = sum_i |F_i - F_ref_i|^2.  The gradients of the
adjoint kernels (mm_param_gradients) are checked
  * against central finite differences (h = 1e-4 relative) of the FP64
    oracle's energy and loss: each component within 2e-3 of the largest
    gradient magnitude of its kind (pairs are evaluated in FP32; the
    oracle's LJ tail rounds sigma/epsilon keys to 6 digits, which enters
    the difference quotient at ~1e-6 / 2e-4);
  * against the basis of forward evaluations (mm_param_gradients_basis), an
    independent route to the same quantities: within 1e-5 of the largest
    magnitude of each kind (energy gradients) and 2e-4 (loss gradients,
    whose basis assembly differences forces of FP32 pair sums);
and the adjoint's device time is reported beside the basis's wall time.
A test-only torch.autograd.Function takes theta from torch, sets it in the
engine through the C ABI and returns the adjoint gradient to autograd: the
loss at the reference parameters is zero to FP32 rounding, and one gradient
step lowers the loss.
"""
import time

import numpy as np
import pytest

import oracle_lib as O
from molmodel_amd import waterbox as W
from molmodel_amd.engine import Engine

pytestmark = pytest.mark.gpu


def with_params(s, types, site_type):
    t = W.System(**{k: getattr(s, k) for k in s.__dataclass_fields__})
    t.q = types[site_type, 0].copy()
    t.sig = types[site_type, 1].copy()
    t.eps = types[site_type, 2].copy()
    return t


def params_for(kind):
    s, p = W.config("water1000_ions")
    if kind in ("p3m", "p3m_ik"):
        p = dict(p, kmode=1, mesh=(32, 32, 32), order=5, ik=int(kind == "p3m_ik"))
    return s, p


@pytest.fixture(scope="module", params=["ewald", "p3m", "p3m_ik"])
def setup(request, gpu_lib):
    s, p = params_for(request.param)
    e = Engine(s, p)
    e.set_positions(s.r)
    g0 = e.param_gradients(0)
    types = g0["params"].copy()
    st = g0["site_type"]
    rng = np.random.default_rng(20240)
    pert = types * (1 + 0.05 * rng.standard_normal(types.shape))
    f_ref, _ = O.evaluate(with_params(s, pert, st), p)
    print(f"[adjoint] {request.param}: gradients", flush=True)
    g = e.param_gradients(0, f_ref=f_ref)
    t0 = time.perf_counter()
    gb = e.param_gradients(0, f_ref=f_ref, basis=True)
    tb = time.perf_counter() - t0
    print(f"[adjoint] {request.param}: adjoint passes {g['ms']:.3f} ms on the device, "
          f"basis {1e3 * tb:.1f} ms wall", flush=True)
    return s, p, e, g, gb, types, pert, st, f_ref


def oracle_EL(s, p, types, st, f_ref):
    f, en = O.evaluate(with_params(s, types, st), p)
    return en[3], float(np.sum((f - f_ref) ** 2))


def test_gradients_match_oracle_fd(setup):
    s, p, e, g, gb, types, pert, st, f_ref = setup
    print("types (q, sigma, eps):\n", types)
    h = 1e-4
    fdE = np.zeros_like(types)
    fdL = np.zeros_like(types)
    for t in range(types.shape[0]):
        for c in range(3):
            if types[t, c] == 0:
                continue           # no LJ on this type
            tp, tm = types.copy(), types.copy()
            d = h * abs(types[t, c])
            tp[t, c] += d
            tm[t, c] -= d
            Ep, Lp = oracle_EL(s, p, tp, st, f_ref)
            Em, Lm = oracle_EL(s, p, tm, st, f_ref)
            fdE[t, c] = (Ep - Em) / (2 * d)
            fdL[t, c] = (Lp - Lm) / (2 * d)
    mask = types != 0
    for name, dev, fd in (("dE", g["dE"], fdE), ("dL", g["dL"], fdL)):
        print(name, "adjoint\n", dev, "\nfd\n", fd)
        for c, what in enumerate(("q", "sigma", "eps")):
            m = mask[:, c]
            scale = np.abs(fd[m, c]).max()
            err = np.abs(dev[m, c] - fd[m, c]).max()
            print(name, what, "max err", err, "scale", scale)
            assert err <= 2e-3 * scale, (name, what)


def test_adjoint_matches_basis(setup):
    s, p, e, g, gb, types, pert, st, f_ref = setup
    assert gb["basis_err"] < 1e-5
    mask = types != 0
    for name, tol in (("dE", 1e-5), ("dL", 2e-4)):
        for c, what in enumerate(("q", "sigma", "eps")):
            m = mask[:, c]
            scale = np.abs(gb[name][m, c]).max()
            err = np.abs(g[name][m, c] - gb[name][m, c]).max()
            print(f"{name} {what}: adjoint vs basis {err / scale:.2e} of the largest")
            assert err <= tol * scale, (name, what)
    assert abs(g["loss"] - gb["loss"]) <= 1e-12 * gb["loss"]


def test_loss_matches_oracle(setup):
    s, p, e, g, gb, types, pert, st, f_ref = setup
    _, L = oracle_EL(s, p, types, st, f_ref)
    assert abs(g["loss"] - L) <= 1e-4 * L


def test_engine_state_restored(setup):
    s, p, e, g, gb, types, pert, st, f_ref = setup
    e.compute()
    f_after = e.forces()
    f_oracle, _ = O.evaluate(s, p)
    assert np.linalg.norm(f_after - f_oracle) / np.linalg.norm(f_oracle) < 1e-5


def test_torch_autograd_wrapper(setup):
    torch = pytest.importorskip("torch")
    from torch_params import TypeParamLoss
    s, p, e, g, gb, types, pert, st, f_ref = setup
    theta = torch.tensor(types, dtype=torch.float64, requires_grad=True)
    loss = TypeParamLoss.apply(theta, e, 0, f_ref, st)
    loss.backward()
    assert abs(loss.item() - g["loss"]) <= 1e-12 * g["loss"]
    assert np.allclose(theta.grad.numpy(), g["dL"], rtol=0, atol=1e-12 * np.abs(g["dL"]).max())
    # theta taken from torch: at the reference parameters the loss vanishes
    # to the FP32 pair arithmetic
    at_ref = TypeParamLoss.apply(torch.tensor(pert, dtype=torch.float64), e, 0, f_ref, st)
    print(f"[adjoint] loss at theta {loss.item():.4e}, at the reference parameters {at_ref.item():.4e}")
    assert at_ref.item() <= 1e-8 * loss.item()
    # one small gradient step along -dL lowers the loss
    gnorm = theta.grad / theta.grad.abs().max()
    step = (theta.detach() - 1e-4 * gnorm * theta.detach().abs())
    after = TypeParamLoss.apply(step, e, 0, f_ref, st)
    assert after.item() < loss.item()
    # leave the engine at its own parameters
    TypeParamLoss.apply(torch.tensor(types, dtype=torch.float64), e, 0, f_ref, st)
