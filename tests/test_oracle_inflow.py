"""The mixing-layer inflow boundary (label 5, apply_inflow,
src/BoundCond.cuh:235-612) in the oracle, on the CPU.

* the profile: at every inflow face point u is (u1+u2)/2 exactly where
  y = 0, tends to the stream speeds beyond a few vorticity thicknesses and is
  monotone in y; rho, v, w, p, T (and the mass fractions) switch from the
  lower to the upper stream at y = 0 (y >= 0 upper); the ghost layers repeat
  the face point;
* the white noise: (u - u_profile) / (intensity |V|) and its v partner are
  standard normal (Kolmogorov-Smirnov against N(0, 1), mean, variance, no
  correlation between u and v or between successive passes), zero outside
  |y| < 4 delta_omega, the same for the same seed and pass;
* the reference seeds its cuRAND states from the wall clock, so no run of it
  is reproducible: the deviates themselves are parity unpinned against the
  reference, pinned by their distribution here.
"""
import numpy as np
import pytest
from scipy import stats

from _util import O, cases, pkg


def _spatial(noise=0.0, ny=41, nz=10, **kw):
    return cases.mixing_layer_case([15], [ny - 1], [nz - 1], spatial=True, inflow_noise=noise, **kw)


def _face(o, case, g):
    """bv on the inflow face (i = 0) and its ghost layers: (planes, layers, k, j)."""
    bv = o.array(0, O.BV)
    return bv[:, :, :, :g + 1][..., ::-1].transpose(0, 3, 1, 2)  # layer 0 = face, then ghosts


def test_profile_and_streams():
    case = _spatial()
    o = O.Oracle(case)
    o.setup_mixing_layer()
    o.initialize()
    g = o.ngg
    blk = case["blocks"][0]
    y = blk["y0"] + np.arange(-g, blk["my"] + g) * case["dy"]
    face = _face(o, case, g)[:, :, g:-g, :]       # interior k
    up, lo = case["streams"]["upper"], case["streams"]["lower"]
    u = face[1, 0, 0]
    sgn = 1 if up[1] > lo[1] else -1
    assert np.all(sgn * np.diff(u) >= 0)                      # monotone (tanh saturates far out)
    mid = np.abs(y[:-1]) < 2 * case["delta_omega"]
    assert np.all(sgn * np.diff(u)[mid] > 0)
    j0 = int(np.argmin(np.abs(y)))
    assert y[j0] == 0.0
    assert u[j0] == 0.5 * (up[1] + lo[1])
    far = np.abs(y) > 6 * case["delta_omega"]
    np.testing.assert_allclose(u[far & (y > 0)], up[1], rtol=1e-8)
    np.testing.assert_allclose(u[far & (y < 0)], lo[1], rtol=1e-8)
    for l in (0, 2, 3, 4, 5):
        assert np.all(face[l, 0, 0][y >= 0] == up[l]) and np.all(face[l, 0, 0][y < 0] == lo[l])
    for layer in range(1, g + 1):   # ghost layers repeat the face point
        np.testing.assert_array_equal(face[:, layer], face[:, 0])


def test_mixture_profile_carries_the_stream_mass_fractions():
    import mech
    import oracle_fr as OF
    T = mech.load(*pkg.mixture.H2_AIR_FILES)
    case = cases.reacting_case([11], [30], [7], T, spatial=True, inflow_noise=0.0)
    o = OF.FrOracle(case, T, n_threads=4)
    o.setup_mixing_layer()
    o.initialize()
    g = o.ngg
    blk = case["blocks"][0]
    y = blk["y0"] + np.arange(-g, blk["my"] + g) * case["dy"]
    sv = o.array(0, OF.SV)
    for gi in range(g + 1):      # face i = 0 at padded index g, ghosts below it
        col = sv[:, g + 2, :, g - gi]
        assert np.all(col[:, y >= 0] == np.asarray(case["y_upper"])[:, None])
        assert np.all(col[:, y < 0] == np.asarray(case["y_lower"])[:, None])


def _noise_samples(intensity=0.05, passes=8, seed_shift=0):
    case = _spatial(noise=intensity, ny=81, nz=31)
    for f in case["boundaries"]:
        if "inflow" in f:
            f["inflow"]["seed"] += seed_shift
    o = O.Oracle(case)
    o.setup_mixing_layer()
    clean = _spatial(noise=0.0, ny=81, nz=31)
    oc = O.Oracle(clean)
    oc.setup_mixing_layer()
    L = O.lib()
    g = o.ngg
    blk = case["blocks"][0]
    y = blk["y0"] + np.arange(-g, blk["my"] + g) * case["dy"]
    L.orc_apply_bc(oc.ctx, 0)
    ref = oc.array(0, O.BV)[:, :, :, g].copy()
    vel = np.sqrt(ref[1] ** 2 + ref[2] ** 2 + ref[3] ** 2)
    assert np.all(vel[:, np.abs(y) > 0] > 0)   # |V| = 0 only on y = 0 when the streams are opposite
    scale = np.where(vel > 0, intensity * vel, 1.0)
    out = []
    for _ in range(passes):
        L.orc_apply_bc(o.ctx, 0)
        cur = o.array(0, O.BV)[:, :, :, g]
        out.append(((cur[1] - ref[1]) / scale, (cur[2] - ref[2]) / scale))
    return out, y, case["delta_omega"]


def test_white_noise_is_standard_normal_inside_the_layer():
    samples, y, delta = _noise_samples()
    inside = (y < 4 * delta) & (y > -4 * delta) & (y != 0)
    n_u = np.concatenate([s[0][:, inside].ravel() for s in samples])
    n_v = np.concatenate([s[1][:, inside].ravel() for s in samples])
    n = n_u.size
    assert n > 2000
    for x in (n_u, n_v):
        assert abs(np.mean(x)) < 4 / np.sqrt(n)
        assert abs(np.var(x) - 1) < 6 * np.sqrt(2 / n)
        assert stats.kstest(x, "norm").pvalue > 1e-3
    assert abs(np.corrcoef(n_u, n_v)[0, 1]) < 4 / np.sqrt(n)
    a, b = samples[0][0][:, inside].ravel(), samples[1][0][:, inside].ravel()
    assert abs(np.corrcoef(a, b)[0, 1]) < 4 / np.sqrt(a.size)       # passes draw afresh
    outside = (y >= 4 * delta) | (y <= -4 * delta)
    for s in samples:                                                 # none outside the layer
        assert np.all(s[0][:, outside] == 0) and np.all(s[1][:, outside] == 0)


def test_white_noise_repeats_for_the_same_seed_and_changes_with_it():
    a, _, _ = _noise_samples(passes=2)
    b, _, _ = _noise_samples(passes=2)
    c, _, _ = _noise_samples(passes=2, seed_shift=1)
    for s, t in zip(a, b):
        np.testing.assert_array_equal(s[0], t[0])
    assert not np.array_equal(a[0][0], c[0][0])


@pytest.mark.parametrize("noise", [0.0, 1e-3])
def test_spatial_layer_keeps_its_inflow_and_stays_finite(noise):
    case = _spatial(noise=noise)
    o = O.Oracle(case)
    o.setup_mixing_layer()
    o.initialize()
    g = o.ngg
    face0 = o.array(0, O.BV)[:, :, :, g].copy()
    o.step(5)
    bv = o.array(0, O.BV)
    assert np.all(np.isfinite(bv))
    face = bv[:, :, :, g]
    if noise == 0.0:
        np.testing.assert_array_equal(face, face0)   # the inflow state is imposed every stage
    else:
        assert not np.array_equal(face[1], face0[1])
        np.testing.assert_array_equal(face[4], face0[4])
