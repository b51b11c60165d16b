"""The two stencils of the mixture characteristic WENO-Z5 (scheme 52).

ch_stencil 0 projects F+ at i-2..i+2 and F- at i-1..i+3 for every
characteristic field.  ch_stencil 1 is the mixture code of the reference
(src/WENO.cu:880-881): the five conservative components of the five
acoustic/entropy fields come from F+ at i-3..i+1 and F- at i-2..i+2, while
their species terms and the species fields keep i-2..i+2 / i-1..i+3
(src/WENO.cu:884-894).  At the block's first face the point i-3 lies beyond
the ghost layers; the outermost ghost point stands in for it.

Pins:
* on a state that varies along x only (uniform grid, periodic box, no
  viscosity or chemistry, so the y and z sweeps cancel exactly), the oracle's
  dq equals an independent numpy restatement of the x sweep: split fluxes
  from the LF point routine, the full left/right characteristic matrices of
  every face, WENO-Z5 per field and the back projection, with the stencil
  of each variant;
* the two variants differ (the shift is not a no-op), both preserve a free
  stream exactly, and the consistent stencil telescopes on a periodic box
  (the reference stencil does not quite: its first face uses the stand-in
  point -3 where the periodic image of that face uses the true point -4).
"""
import numpy as np
import pytest

from _util import pkg
import mech
import oracle as OA
import oracle_fr as OF

R_U = 8314.462618
N = 10


@pytest.fixture(scope="module")
def T():
    return mech.load(*pkg.mixture.H2_AIR_FILES)


def _case(T, stencil):
    c = pkg.cases.reacting_single_block(N, N, N, T, scheme=52, viscous_order=0, reaction=0)
    c["boundaries"] = [dict(block=0, axis=a, side=s, type=8) for a in range(3) for s in (-1, 1)]
    c["ch_stencil"] = stencil
    return c


def _oracle(T, stencil, state):
    o = OF.FrOracle(_case(T, stencil), T, n_threads=2)
    o.array(0, OF.JAC)[...] = 0.125
    met = o.array(0, OF.METRIC)
    met[...] = 0.0
    met[0], met[4], met[8] = 2.0, 2.0, 2.0  # uniform grid of spacing 0.5
    bv, sv = o.array(0, OF.BV), o.array(0, OF.SV)
    state(o, bv, sv)
    mw_inv = np.tensordot(1 / np.array(T.mw[:9]), sv, axes=1)
    bv[4] = bv[0] * R_U * mw_inv * bv[5]
    o.L.orf_cv_from_bv(o.ctx, 0)
    o.L.orf_properties(o.ctx, 0)
    return o


def _x_wave(o, bv, sv):
    """Smooth, periodic in x with period N-1 nodes, uniform in y and z."""
    g = o.ngg
    i = (np.arange(-g, N + g) % (N - 1))[None, None, :]
    ph = 2 * np.pi * i / (N - 1)
    w = [np.sin(ph + a) + 0.3 * np.cos(2 * ph + 2 * a) for a in np.linspace(0.1, 5.0, 14)]
    y = np.array([1.0 + 0.6 * w[q] for q in range(9)])
    sv[...] = np.broadcast_to(y / y.sum(axis=0), sv.shape)
    bv[0] = 0.6 * (1 + 0.25 * w[9])
    bv[1] = 180 * w[10]
    bv[2] = 40 * w[11]
    bv[3] = -30 * w[12]
    bv[5] = 1200 * (1 + 0.3 * w[13])


def _uniform(o, bv, sv):
    sv[...] = 0.0
    sv[0], sv[3], sv[8] = 0.1, 0.2, 0.7
    bv[0], bv[1], bv[2], bv[3], bv[5] = 0.5, 120.0, -40.0, 25.0, 1300.0


def _x_sweep_dq(T, o, stencil):
    """dq of the x sweep along the pencil j = k = 0, restated in numpy."""
    g, nv, ns = o.ngg, 14, 9
    cv = o.array(0, OF.CV)[:, g, g, :]
    p = o.array(0, OF.BV)[4, g, g, :]
    c = o.array(0, OF.C_SOUND)[0, g, g, :]
    k = np.array([2.0, 0.0, 0.0])
    jac = 0.125
    npts = N + 2 * g
    Q = np.concatenate([cv, p[None], c[None]]).T  # [point][nv + 2]
    Fp, Fm = np.zeros((npts, nv)), np.zeros((npts, nv))
    for ps in range(npts):
        Fp[ps], Fm[ps] = OA.lf_split(Q[ps], k, jac)
    case = o.case
    eps_scale = (case["rho_ref"] * case["v_ref"]) ** 2
    kxj = 2 * k * jac
    eps = 1e-40 * eps_scale * 0.25 * float(kxj @ kxj)
    sh = stencil
    fc = np.zeros((N + 1, nv))
    for t in range(-1, N):
        s = t + g
        L, R = OF.char_matrices(T, Q[s], Q[s + 1], k, k, jac, jac)
        fchar = np.zeros(nv)
        for f in range(nv):
            vp, vm = np.zeros(5), np.zeros(5)
            for r in range(5):
                ip = max(s - 2 - (sh if f < 5 else 0) + r, 0)
                im = s - 1 - (sh if f < 5 else 0) + r
                vp[r] = L[f, :5] @ Fp[ip, :5] + L[f, 5:] @ Fp[s - 2 + r, 5:]
                vm[r] = L[f, :5] @ Fm[im, :5] + L[f, 5:] @ Fm[s - 1 + r, 5:]
            fchar[f] = OA.weno5z(vp, vm, eps)
        fc[t + 1] = R @ fchar
    return -(fc[1:] - fc[:-1]).T  # [var][i]


@pytest.mark.parametrize("stencil", [0, 1])
def test_oracle_x_sweep_matches_a_numpy_restatement(T, stencil):
    o = _oracle(T, stencil, _x_wave)
    dq = o.rhs(0, 0.0)
    # the state does not vary along y and z: every pencil carries the same dq
    assert np.all(dq == dq[:, :1, :1, :])
    want = _x_sweep_dq(T, o, stencil)
    got = dq[:, 0, 0, :]
    scale = np.max(np.abs(want), axis=1)
    err = np.max(np.abs(got - want), axis=1) / scale
    assert np.all(err <= 1e-10), err


def test_the_reference_stencil_changes_the_acoustic_fields(T):
    d0 = _oracle(T, 0, _x_wave).rhs(0, 0.0).copy()
    d1 = _oracle(T, 1, _x_wave).rhs(0, 0.0).copy()
    rel = np.max(np.abs(d1 - d0), axis=(1, 2, 3)) / np.max(np.abs(d0), axis=(1, 2, 3))
    assert np.all(rel[:5] > 1e-3), rel


@pytest.mark.parametrize("stencil", [0, 1])
def test_free_stream_and_telescoping(T, stencil):
    o = _oracle(T, stencil, _uniform)
    dq = o.rhs(0, 0.0)
    cv = o.array(0, OF.CV)
    for l in range(14):
        assert np.max(np.abs(dq[l])) <= 2e-12 * np.max(np.abs(cv[l])), l
    if stencil == 1:
        return
    o = _oracle(T, stencil, _x_wave)
    dq = o.rhs(0, 0.0)[:, :-1, :-1, :-1]
    for l in range(14):
        s, a = abs(dq[l].sum()), np.abs(dq[l]).sum()
        assert s <= 1e-12 * a, (l, s, a)
