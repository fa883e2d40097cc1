"""Closed-form pins of GradientDescent::minimize and Adam::minimize
(src/optimization/optimizer.rs:56-107, 139-223), run against both the
oracle's restatement and the library's host optimiser (lq_gd_minimize /
lq_adam_minimize; no device needed).  Each check is chosen so that a
plausible mistake fails it: Adam without bias correction or with
powi(iteration - 1), the three stopping tests in another order or place,
a missing final gradient re-check, GD's step count off by one."""
import pytest

import oracle
import logosq_amd as lq

TOL = 1e-6  # the reference's default tolerance


class Counted:
    """f and grad_f with call counters."""

    def __init__(self, f, g):
        self.nf = self.ng = 0
        self._f, self._g = f, g

    def f(self, x):
        self.nf += 1
        return self._f(x)

    def g(self, x):
        self.ng += 1
        return self._g(x)


def run(impl, method, fn, x0, lr, max_it, tol=TOL):
    """(x, optimal_value, history, converged) from the oracle or the library."""
    if impl == "oracle":
        m = oracle.adam_minimize if method == "adam" else oracle.gd_minimize
        x, v, h, c = m(fn.f, fn.g, x0, lr, max_it, tol)
        return list(x), v, list(h), c
    opt = (lq.Adam if method == "adam" else lq.GradientDescent)(lr, max_it, tol)
    r = opt.minimize(fn.f, fn.g, x0)
    assert r.num_iterations == len(r.convergence_history)
    return list(r.optimal_parameters), r.optimal_value, list(r.convergence_history), r.converged


IMPLS = ["oracle", "library"]


@pytest.mark.parametrize("impl", IMPLS)
def test_adam_first_step_is_lr_g_over_abs_g_plus_eps(impl):
    """After one update with m0 = v0 = 0 the bias-corrected moments are g and
    g^2, so each coordinate moves by lr * g / (|g| + eps) (0 for g = 0)."""
    a = [3.0, -0.5, 1e-3, 0.0]
    fn = Counted(lambda x: sum(ai * xi for ai, xi in zip(a, x)), lambda x: list(a))
    x0 = [0.25, -1.0, 2.0, 0.5]
    x, v, h, conv = run(impl, "adam", fn, x0, 0.05, 1)
    for xi, x0i, ai in zip(x, x0, a):
        want = x0i - 0.05 * ai / (abs(ai) + 1e-8)
        assert abs(xi - want) <= 1e-15 * max(1.0, abs(want))
    assert len(h) == 1 and not conv
    assert fn.ng == 2 and fn.nf == 2  # one in the loop, one final evaluation each


@pytest.mark.parametrize("impl", IMPLS)
def test_adam_second_step_by_hand(impl):
    """f = x^2, x0 = 1, lr = 0.1, worked with exact decimals:
    step 1: g = 2, m = 0.2, v = 0.004, m_hat = 2, v_hat = 4,
            x1 = 1 - 0.1*2/(2 + 1e-8) = 0.9000000004999999975...
    step 2: g = 1.800000000999999995, m = 0.3600000000999999995,
            v = 0.007236000003599999983, m_hat = m/(1 - 0.9^2) = 1.894736842631578944...,
            v_hat = v/(1 - 0.999^2) = 3.619809906753376679...,
            x2 = x1 - 0.1*m_hat/(sqrt(v_hat) + 1e-8) = 0.8004122286917921452..."""
    fn = Counted(lambda x: x[0] * x[0], lambda x: [2.0 * x[0]])
    x, v, h, conv = run(impl, "adam", fn, [1.0], 0.1, 2)
    x1 = 0.9000000004999999975
    assert abs(h[0] - 1.0) == 0.0
    assert abs(h[1] - x1 * x1) <= 1e-15
    assert abs(x[0] - 0.8004122286917921452) <= 1e-15
    assert abs(v - 0.8004122286917921452 ** 2) <= 1e-15
    assert not conv


@pytest.mark.parametrize("impl", IMPLS)
def test_adam_stop_on_value_change(impl):
    """|f_prev - f| < tol is tested right after f is recorded, before the
    gradient: f constant, gradient 1 -> two recorded values, one update, and
    only the first iteration's gradient plus the final one are evaluated."""
    fn = Counted(lambda x: 5.0, lambda x: [1.0])
    x, v, h, conv = run(impl, "adam", fn, [0.3], 0.01, 50)
    assert h == [5.0, 5.0] and conv
    assert abs(x[0] - (0.3 - 0.01 / (1.0 + 1e-8))) <= 1e-16
    assert fn.ng == 2 and fn.nf == 3


@pytest.mark.parametrize("impl", IMPLS)
def test_adam_stop_on_gradient_norm(impl):
    """f = (x - 1)^2 from 0 with lr = 1: the first step lands at
    1 - 5e-9 (|df| = 1 > tol), where |grad| = 1e-8 < tol stops the loop
    before a second update."""
    fn = Counted(lambda x: (x[0] - 1.0) ** 2, lambda x: [2.0 * (x[0] - 1.0)])
    x, v, h, conv = run(impl, "adam", fn, [0.0], 1.0, 50)
    x1 = 0.0 - 1.0 * (-2.0) / (2.0 + 1e-8)
    assert len(h) == 2 and conv
    assert x[0] == x1
    assert fn.ng == 3 and fn.nf == 3


@pytest.mark.parametrize("impl", IMPLS)
def test_adam_stop_on_step_norm(impl):
    """A step shorter than tol (lr = 1e-7, |g| = 1) stops after the first
    update, with one recorded value."""
    fn = Counted(lambda x: x[0], lambda x: [1.0])
    x, v, h, conv = run(impl, "adam", fn, [2.0], 1e-7, 50)
    assert h == [2.0] and conv
    assert x[0] == 2.0 - 1e-7 * 1.0 / (1.0 + 1e-8)
    assert fn.ng == 2


@pytest.mark.parametrize("impl", IMPLS)
def test_adam_runs_to_max_iterations(impl):
    fn = Counted(lambda x: x[0] * x[0], lambda x: [2.0 * x[0]])
    x, v, h, conv = run(impl, "adam", fn, [1.0], 0.001, 5)
    assert len(h) == 5 and not conv
    assert fn.ng == 6 and fn.nf == 6


@pytest.mark.parametrize("impl", IMPLS)
def test_adam_final_gradient_recheck(impl):
    """One iteration only, landing where |grad| < tol: the loop ends at
    max_iterations, and the final gradient check marks it converged."""
    fn = Counted(lambda x: (x[0] - 1.0) ** 2, lambda x: [2.0 * (x[0] - 1.0)])
    x, v, h, conv = run(impl, "adam", fn, [0.0], 1.0, 1)
    assert len(h) == 1 and conv
    assert v == (x[0] - 1.0) ** 2 and v < 1e-16


@pytest.mark.parametrize("impl", IMPLS)
def test_gd_step_count_halving(impl):
    """f = x^2/2, lr = 0.5 halves x each step: x_k = 2^-k exactly, and the
    first |x_k| < 1e-6 is k = 20, so 21 values are recorded and x = 2^-20."""
    fn = Counted(lambda x: 0.5 * x[0] * x[0], lambda x: [x[0]])
    x, v, h, conv = run(impl, "gd", fn, [1.0], 0.5, 100)
    assert h == [0.5 * 0.25 ** k for k in range(21)]
    assert x[0] == 2.0 ** -20 and conv
    assert v == 0.5 * 2.0 ** -40


@pytest.mark.parametrize("impl", IMPLS)
def test_gd_final_gradient_recheck(impl):
    """max_iterations = 1 with lr = 0.5 on x^2 lands on 0: the loop ends by
    its count and the final gradient check sets converged."""
    fn = Counted(lambda x: x[0] * x[0], lambda x: [2.0 * x[0]])
    x, v, h, conv = run(impl, "gd", fn, [1.0], 0.5, 1)
    assert h == [1.0] and x[0] == 0.0 and v == 0.0 and conv
    assert fn.ng == 2


@pytest.mark.parametrize("impl", IMPLS)
def test_gd_not_converged(impl):
    fn = Counted(lambda x: x[0] * x[0], lambda x: [2.0 * x[0]])
    x, v, h, conv = run(impl, "gd", fn, [1.0], 0.1, 3)
    assert len(h) == 3 and max(abs(a - b) for a, b in zip(h, [1.0, 0.64, 0.4096])) <= 1e-15
    assert abs(x[0] - 0.512) <= 1e-15 and abs(v - 0.262144) <= 1e-15 and not conv
