"""The sponge layer of the mixture oracle (src/SpongeLayer.cu:28-172), on
the CPU.

* the source alone (dq after orf_source on a non-reacting state with a zero
  mean) is -sigma cv on every conservative row (-sigma J cv with the
  Jacobian scaling), so sigma is read back point by point: zero inside the
  domain, sigma_d at and beyond the outer end of a side, sigma_d / 2 exactly
  half-way (S(1/2) = 1/2), the smoothstep S(xi) = xi^3 (6 xi^2 - 15 xi + 10)
  elsewhere, and the sum of both sides where two layers overlap;
* the running mean: after k steps the source is -sigma J (cv - m) with m the
  arithmetic mean of the cv at the start of each of the k steps.
"""
import numpy as np
import pytest

from _util import cases, pkg


@pytest.fixture(scope="module")
def T():
    import mech
    return mech.load(*pkg.mixture.H2_AIR_FILES)


def _case(T, sides=(1, 2, 3), jacobian_scaled=0):
    case = cases.reacting_case([15], [19], [7], T, reaction=0, spatial=True, inflow_noise=0.0)
    case["sponge"] = cases.sponge_layer(case, sides=sides, depth=0.3, coefficient=2.0,
                                        jacobian_scaled=jacobian_scaled)
    return case


def _oracle(case, T):
    import oracle_fr as OF
    o = OF.FrOracle(case, T, n_threads=4)
    o.setup_mixing_layer()
    o.initialize()
    return o


def _source_dq(o, b=0):
    L = o.L
    L.orf_zero_dq(o.ctx, b)
    L.orf_source(o.ctx, b)
    import oracle_fr as OF
    return o.array(b, OF.DQ).copy()


def _smoothstep(xi):
    return xi ** 3 * (6 * xi * xi - 15 * xi + 10)


def _expected_sigma(case):
    sp = case["sponge"]
    blk = case["blocks"][0]
    k, j, i = np.meshgrid(np.arange(blk["mz"]), np.arange(blk["my"]), np.arange(blk["mx"]), indexing="ij")
    pos = [blk["x0"] + i * case["dx"], blk["y0"] + j * case["dy"], blk["z0"] + k * case["dz"]]
    sigma = np.zeros(i.shape)
    for side in range(6):
        if not sp["enabled"][side]:
            continue
        x, x1, x2 = pos[side // 2], sp["start"][side], sp["end"][side]
        on = x < x1 if side % 2 == 0 else x > x1
        xi = np.minimum((x - x1) / (x2 - x1), 1.0)
        sigma += np.where(on, sp["sigma"][side] * _smoothstep(np.where(on, xi, 0.0)), 0.0)
    return sigma, pos


def test_source_is_minus_sigma_cv_with_the_ramp(T):
    import oracle_fr as OF
    case = _case(T)
    o = _oracle(case, T)
    g = o.ngg
    dq = _source_dq(o)
    cv = o.array(0, OF.CV)[:, g:-g, g:-g, g:-g]
    sigma_read = -dq[0] / cv[0]
    for l in range(dq.shape[0]):   # every conservative row carries the same sigma
        nz = np.abs(cv[l]) > 1e-300
        np.testing.assert_allclose(-dq[l][nz], sigma_read[nz] * cv[l][nz], rtol=1e-13, atol=0)
    want, pos = _expected_sigma(case)
    np.testing.assert_allclose(sigma_read, want, rtol=1e-13, atol=1e-300)
    sp = case["sponge"]
    x, y = pos[0], pos[1]
    inside = (x <= sp["start"][1]) & (y >= sp["start"][2]) & (y <= sp["start"][3])
    assert np.all(sigma_read[inside] == 0)
    # at the outer boundary of x+ only (y within the core): sigma_1 exactly
    at_end = (x == x.max()) & (y > sp["start"][2]) & (y < sp["start"][3])
    np.testing.assert_allclose(sigma_read[at_end], sp["sigma"][1], rtol=1e-14)
    # both layers overlap in the corners: the sum of the two ramps
    corner = (x > sp["start"][1]) & (y > sp["start"][3])
    assert corner.any() and np.all(sigma_read[corner] > sp["sigma"][1] * _smoothstep(
        np.minimum((x[corner] - sp["start"][1]) / (sp["end"][1] - sp["start"][1]), 1)) * (1 - 1e-14))


def test_half_way_is_half_sigma(T):
    import oracle_fr as OF
    case = _case(T, sides=(1,))
    sp = case["sponge"]
    blk = case["blocks"][0]
    xs = blk["x0"] + np.arange(blk["mx"]) * case["dx"]
    # place the ramp so that one node lies half-way, exactly representable
    i_mid = blk["mx"] - 3
    sp["start"][1], sp["end"][1] = xs[i_mid] - 2 * case["dx"], xs[i_mid] + 2 * case["dx"]
    o = _oracle(case, T)
    g = o.ngg
    dq = _source_dq(o)
    cv = o.array(0, OF.CV)[:, g:-g, g:-g, g:-g]
    xi = (xs[i_mid] - sp["start"][1]) / (sp["end"][1] - sp["start"][1])
    assert xi == 0.5
    np.testing.assert_allclose(-dq[0][:, :, i_mid] / cv[0][:, :, i_mid], 0.5 * sp["sigma"][1], rtol=1e-14)


def test_jacobian_scaling_multiplies_sigma_by_j(T):
    import oracle_fr as OF
    a, b = _oracle(_case(T, jacobian_scaled=0), T), _oracle(_case(T, jacobian_scaled=1), T)
    g = a.ngg
    jac = a.array(0, OF.JAC)[0, g:-g, g:-g, g:-g]
    da, db = _source_dq(a), _source_dq(b)
    cv = a.array(0, OF.CV)[:, g:-g, g:-g, g:-g]
    sa, sb = -da[0] / cv[0], -db[0] / cv[0]
    np.testing.assert_allclose(sb, sa * jac, rtol=1e-14, atol=0)


def test_running_mean_of_the_step_starts(T):
    import oracle_fr as OF
    case = _case(T, jacobian_scaled=1)
    o = _oracle(case, T)
    g = o.ngg
    starts = []
    for _ in range(3):
        starts.append(o.array(0, OF.CV)[:, g:-g, g:-g, g:-g].copy())
        o.step(1)
    mean = np.mean(starts, axis=0)
    cv = o.array(0, OF.CV)[:, g:-g, g:-g, g:-g]
    dq = _source_dq(o)
    want_sigma, _ = _expected_sigma(case)
    jac = o.array(0, OF.JAC)[0, g:-g, g:-g, g:-g]
    want = -(want_sigma * jac)[None] * (cv - mean)
    scale = np.max(np.abs(want), axis=(1, 2, 3), keepdims=True)
    assert np.all(np.max(np.abs(dq - want), axis=(1, 2, 3), keepdims=True) <= 1e-12 * scale)
    assert np.any(want != 0)


def test_air_sponge_source_reads_back_sigma_j():
    from _util import O
    case = cases.mixing_layer_case([15], [19], [7], spatial=True, inflow_noise=0.0)
    case["sponge"] = cases.sponge_layer(case, sides=(1, 2, 3), depth=0.3, coefficient=2.0, jacobian_scaled=1)
    o = O.Oracle(case)
    o.setup_mixing_layer()
    o.initialize()
    g = o.ngg
    L = O.lib()
    L.orc_zero_dq(o.ctx, 0)
    L.orc_source(o.ctx, 0)
    dq = o.array(0, O.DQ)
    cv = o.array(0, O.CV)[:, g:-g, g:-g, g:-g]
    jac = o.array(0, O.JAC)[0, g:-g, g:-g, g:-g]
    want, _ = _expected_sigma(case)
    for l in range(5):
        nz = np.abs(cv[l]) > 0
        np.testing.assert_allclose(-dq[l][nz], (want * jac)[nz] * cv[l][nz], rtol=1e-13, atol=0)
    assert np.any(want > 0) and np.any(want == 0)
