"""Pins of the oracle's dual time stepping with the DPLUR inner solver
(dual_time_stepping, src/DualTimeStepping.cuh:33-236; DPLUR,
src/DPLUR.cuh:14-327), air.

* compute_jacobian_times_dq is (A + lambda) dq with A the Jacobian of the
  physical Euler flux along the metric row: checked against central
  differences of that flux on random states and directions (a sign or index
  slip in b1, b2 or h fails it);
* a uniform free stream stays uniform through dual time steps;
* at convergence of the inner iterations the state solves the BDF2
  equation R(q^(n+1)) = J (3 q^(n+1) - 4 q^n + q^(n-1)) / (2 dt), R the stage
  residual of the same operators -- the qn* and modified-residual terms, and
  the DPLUR solve that drives them to zero, are pinned by this;
* the inner-iteration count adapts as inner_converged describes: it grows by
  5 (capped at max_inner_iteration) while the change is above 1e-3 and
  drops to the iteration where it falls below.
"""
import numpy as np
import pytest

from _util import O, cases, oracle_for

GAMMA = 1.4


def _flux(q, k):
    """Physical Euler flux along metric row k of conservative q."""
    rho, mu, mv, mw, E = q
    u, v, w = mu / rho, mv / rho, mw / rho
    p = (GAMMA - 1) * (E - 0.5 * rho * (u * u + v * v + w * w))
    U = k[0] * u + k[1] * v + k[2] * w
    return np.array([rho * U, mu * U + k[0] * p, mv * U + k[1] * p, mw * U + k[2] * p, (E + p) * U])


@pytest.mark.parametrize("seed", range(6))
def test_jacobian_times_dq_is_the_flux_jacobian(seed):
    rng = np.random.default_rng(seed)
    rho, T = rng.uniform(0.3, 2.0), rng.uniform(200, 900)
    u, v, w = rng.normal(0, 150, 3)
    p = rho * 8314.462618 / 28.9647 * T
    pv = np.array([rho, u, v, w, p, T])
    q = np.array([rho, rho * u, rho * v, rho * w, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)])
    k = rng.normal(0, 1, 3) * rng.uniform(0.1, 10)
    dq = rng.normal(0, 1, 5) * np.abs(q)
    lam = rng.normal(0, 300)
    got = O.jacobian_times_dq(k, pv, lam, dq)
    eps = 1e-6
    fd = (_flux(q + eps * dq, k) - _flux(q - eps * dq, k)) / (2 * eps)
    scale = np.max(np.abs(fd)) + np.max(np.abs(lam * dq))
    assert np.max(np.abs(got - lam * dq - fd)) <= 1e-7 * scale, (got - lam * dq, fd)


def _dual(case, **kw):
    d = dict(dt=kw.pop("dt"), inner_iteration=kw.pop("inner", 6), max_inner_iteration=kw.pop("max_inner", 30),
             iteration_adjust_step=kw.pop("adjust", 0), dplur_inner_step=kw.pop("dplur", 3), step=0)
    return d


def test_free_stream_is_preserved():
    case = cases.single_block_case(10, perturbation=0.0)
    o = oracle_for(case)
    bv = o.array(0, O.BV)
    g = o.ngg
    # a uniform state: the upper stream everywhere
    for l in range(6):
        bv[l][...] = case["streams"]["upper"][l]
    o.initialize()
    before = o.array(0, O.CV).copy()
    o.dual_time_init(_dual(case, dt=1e-8, inner=5))
    o.dual_time_step(3)
    after = o.array(0, O.CV)
    scale = np.max(np.abs(before), axis=(1, 2, 3))
    scale[1:4] = np.max(scale[1:4])   # momentum: the stream's, whatever its direction
    err = np.max(np.abs(after - before)[:, g:-g, g:-g, g:-g], axis=(1, 2, 3)) / scale
    assert np.all(err <= 1e-14), err


def test_converged_inner_iterations_solve_bdf2():
    case = cases.single_block_case(10)
    o = oracle_for(case)
    o.step(1)  # a developed state
    g = o.ngg
    dt = 4e-8
    o.dual_time_init(_dual(case, dt=dt, inner=300, adjust=0))
    cv_prev = o.array(0, O.CV).copy()   # q^(n-1) = q^n at the first step
    o.dual_time_step(1)
    cv_n = o.array(0, O.CV).copy()
    o.dual_time_step(1)
    cv_np1 = o.array(0, O.CV).copy()
    jac = o.array(0, O.JAC)[0]
    o.rhs(0)
    R = o.array(0, O.DQ)
    bdf = (3 * cv_np1 - 4 * cv_n + cv_prev)[:, g:-g, g:-g, g:-g] * jac[g:-g, g:-g, g:-g] / (2 * dt)
    scale = np.max(np.abs(R), axis=(1, 2, 3))
    err = np.max(np.abs(R - bdf), axis=(1, 2, 3)) / scale
    print("BDF2 residual per variable:", err)
    assert np.all(err <= 1e-9), err
    # and the state moved: the step is not trivial
    assert np.max(np.abs(cv_np1 - cv_n)) > 1e-8 * np.max(np.abs(cv_n))


def test_inner_iteration_count_adapts():
    """Monitored every step: from 2 iterations the count grows by 5 while the
    last iteration still changes the state by more than 1e-3 of the first
    iteration's change; from 20 the loop ends early; either way the count
    becomes the iteration where the change fell below 1e-3, and it never
    exceeds max_inner_iteration."""
    counts = {}
    for start in (2, 20):
        case = cases.single_block_case(10)
        o = oracle_for(case)
        o.step(1)
        o.dual_time_init(_dual(case, dt=4e-8, inner=start, max_inner=40, adjust=1))
        counts[start] = [o.dual_time_step(1) for _ in range(3)]
    print("inner iterations after each step:", counts)
    # both end the first step at the iteration where the change fell below
    # 1e-3: from 2 the loop was extended (2, 7, 12, 17) to reach it, from 20
    # it stopped early and the count dropped to it
    assert 2 < counts[2][0] == counts[20][0] < 20
    assert counts[2] == counts[20]
    assert all(c <= 40 for v in counts.values() for c in v)
