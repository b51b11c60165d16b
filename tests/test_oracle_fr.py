"""Pins of the finite-rate mixture oracle (CPU only).

Every check here is independent of the oracle's own expressions: literal
values read from the reference's data files, textbook identities of NASA-7
thermodynamics, conservation laws of the chemical source, equilibrium
constants from an independent Gibbs-energy evaluation, Chapman-Enskog
formulas for pure-species viscosity and binary diffusion, a closed form of the
temperature Newton iteration, the inverse pair of characteristic matrices,
and free-stream preservation / conservation of the whole mixture residual.
The two CHEMKIN readers (the library's C++ one, the oracle's Python one) must
agree bitwise.
"""
import ctypes as C
import math

import numpy as np
import pytest

from _util import pkg
import mech
import oracle_fr as OF

MX = pkg.mixture
FILES = MX.H2_AIR_FILES
R_U = 8314.462618
R_C = 1.98720425864083
SPECIES = ["H", "H2", "H2O", "H2O2", "HO2", "O", "O2", "OH", "N2"]
ELEMENTS = {"H": {"H": 1}, "H2": {"H": 2}, "H2O": {"H": 2, "O": 1}, "H2O2": {"H": 2, "O": 2},
            "HO2": {"H": 1, "O": 2}, "O": {"O": 1}, "O2": {"O": 2}, "OH": {"O": 1, "H": 1}, "N2": {"N": 2}}


@pytest.fixture(scope="module")
def T():
    return mech.load(*FILES)


def _arr(t, name):
    v = getattr(t, name)
    return v if isinstance(v, int) else np.ctypeslib.as_array(v)


def _random_mixture(rng, ns=9):
    y = rng.uniform(0.05, 1.0, ns)
    return y / y.sum()


# ------------------------------------------------------------------ readers
def test_readers_agree_bitwise(T):
    tp = MX.load_mechanism(*FILES)
    for name, _ in mech.MechTables._fields_:
        a, b = _arr(tp, name), _arr(T, name)
        if name in ("spec_name", "elem_name"):
            a, b = bytes(getattr(tp, name)), bytes(getattr(T, name))
        assert np.array_equal(np.asarray(a), np.asarray(b)), name


def test_tables_hold_the_file_values(T):
    """Literal values of input/chemistry/{H2_mech.inp, therm.dat, tran.dat}."""
    assert mech.species_names(T) == SPECIES
    assert T.n_reac == 19  # 21 reaction lines, two duplicate pairs merged
    h2 = SPECIES.index("H2")
    # therm.dat, the H2 entry: high-range a1, low-range a1 and a7
    assert T.high_coeff[h2][0] == 3.33727920E+00
    assert T.low_coeff[h2][0] == 2.34433112E+00
    assert T.low_coeff[h2][6] == 6.83010238E-01
    assert (T.t_low[h2], T.t_mid[h2], T.t_high[h2]) == (200.0, 1000.0, 3500.0)
    n2 = SPECIES.index("N2")
    assert (T.t_low[n2], T.t_high[n2]) == (300.0, 5000.0)
    # molecular weights from the element weights 1.008, 15.999, 14.007
    assert T.mw[SPECIES.index("H2O")] == 2 * 1.008 + 15.999
    assert T.mw[n2] == 2 * 14.007
    # tran.dat: H2O epsilon/k = 572.4 K, N2 sigma = 3.621 A
    assert T.lj_potent_inv[SPECIES.index("H2O")] == 1 / 572.4
    assert T.vis_coeff[n2] == 2.6693e-6 * math.sqrt(T.mw[n2]) / (3.621 * 3.621)
    # reaction 1: H+O2=O+OH 3.55E+15 -0.40 1.66E+04, reversible
    assert T.label[0] == 1 and T.A[0] == 3.55e15 and T.b[0] == -0.40 and T.Ea[0] == 1.66e4 / R_C
    assert T.stoi_f[0][SPECIES.index("H")] == 1 and T.stoi_f[0][SPECIES.index("O2")] == 1
    assert T.order[0] == 0
    # reaction 5: H2+M=H+H+M, third body with H2/2.5/ H2O/12.0/
    assert T.label[4] == 4 and T.stoi_b[4][SPECIES.index("H")] == 2 and T.order[4] == 1
    assert T.third_body_coeff[4][h2] == 2.5 and T.third_body_coeff[4][SPECIES.index("H2O")] == 12.0
    assert T.third_body_coeff[4][n2] == 1.0
    # reaction 9: H+O2(+M)=HO2(+M), LOW + TROE(0.8, 1e-30, 1e30), O2/0.78/
    assert T.label[8] == 6 and T.A2[8] == 6.37e20 and T.b2[8] == -1.72 and T.Ea2[8] == 5.25e2 / R_C
    assert (T.troe_alpha[8], T.troe_t3[8], T.troe_t1[8]) == (0.8, 1e-30, 1e30)
    assert T.third_body_coeff[8][SPECIES.index("O2")] == 0.78 and T.order[8] == -1
    # reaction 14: HO2+HO2=H2O2+O2, a duplicate pair merged into one row
    assert T.label[13] == 3 and T.A[13] == 4.20e14 and T.A2[13] == 1.30e11
    assert T.Ea[13] == 1.20e4 / R_C and T.Ea2[13] == -1.63e3 / R_C
    # reaction 19 (last line pair): H2O2+OH=HO2+H2O, merged duplicate
    assert T.label[18] == 3 and T.A[18] == 1.00e12 and T.A2[18] == 5.80e14


# ----------------------------------------------------------- thermodynamics
def _nasa_independent(T, t):
    """h/RT and s/R from the NASA-7 definitions (not the oracle's expressions)."""
    ns = T.n_spec
    h, s = np.zeros(ns), np.zeros(ns)
    for i in range(ns):
        a = T.low_coeff[i] if t < T.t_mid[i] else T.high_coeff[i]
        h[i] = a[0] + a[1] * t / 2 + a[2] * t ** 2 / 3 + a[3] * t ** 3 / 4 + a[4] * t ** 4 / 5 + a[5] / t
        s[i] = a[0] * math.log(t) + a[1] * t + a[2] * t ** 2 / 2 + a[3] * t ** 3 / 3 + a[4] * t ** 4 / 4 + a[6]
    return h, s


@pytest.mark.parametrize("t", [450.0, 999.0, 1500.0, 2600.0])
def test_dh_dt_is_cp(T, t):
    d = 1e-3
    hp, _ = OF.enthalpy_cp(T, t + d)
    hm, _ = OF.enthalpy_cp(T, t - d)
    _, cp = OF.enthalpy_cp(T, t)
    assert np.max(np.abs((hp - hm) / (2 * d) - cp) / cp) < 1e-8


@pytest.mark.parametrize("t", [350.0, 1200.0, 2800.0])
def test_enthalpy_and_gibbs_match_the_nasa_definitions(T, t):
    h, _ = OF.enthalpy_cp(T, t)
    hrt, sr = _nasa_independent(T, t)
    mw = np.array(T.mw[:T.n_spec])
    assert np.max(np.abs(h - hrt * R_U * t / mw) / np.abs(h).max()) < 1e-13
    g = OF.gibbs_rt(T, t)
    assert np.max(np.abs(g - (hrt - sr)) / np.abs(hrt - sr).max()) < 1e-13


def test_enthalpy_is_continuous_at_t_mid(T):
    """The two NASA ranges meet at T_mid to the precision of the 9-digit data."""
    for i in range(T.n_spec):
        tm = T.t_mid[i]
        lo, hi = _nasa_independent(T, tm - 1e-9)[0][i], _nasa_independent(T, tm)[0][i]
        assert abs(lo - hi) <= 1e-4 * abs(hi) + 1e-3


def test_enthalpy_below_t_low_is_the_linear_extension(T):
    n2 = SPECIES.index("N2")  # t_low = 300 K
    h250, _ = OF.enthalpy_cp(T, 250.0)
    h300, cp300 = OF.enthalpy_cp(T, 300.0)
    assert abs(h250[n2] - (h300[n2] - 50.0 * cp300[n2])) <= 1e-12 * abs(h300[n2])


# ---------------------------------------------------------------- chemistry
@pytest.mark.parametrize("seed", range(4))
def test_source_conserves_mass_and_elements(T, seed):
    rng = np.random.default_rng(seed)
    y = _random_mixture(rng)
    t = rng.uniform(900, 2500)
    om, _, _ = OF.chemical_source(T, 0.3, t, y)
    scale = np.abs(om).max()
    assert scale > 0
    assert abs(om.sum()) <= 1e-12 * scale
    mw = np.array(T.mw[:T.n_spec])
    for e in ("H", "O", "N"):
        n = np.array([ELEMENTS[s].get(e, 0) for s in SPECIES])
        assert abs(np.sum(om / mw * n)) <= 1e-12 * np.sum(np.abs(om / mw) * n)


@pytest.mark.parametrize("t", [800.0, 1500.0, 2400.0])
def test_backward_rates_follow_equilibrium_constants(T, t):
    """kf/kb = Kc = (p_atm/(R T))^(sum nu) exp(-dG/RT) with the Gibbs energy
    from the NASA definitions (h/RT - s/R), concentrations in mol/cm^3."""
    rng = np.random.default_rng(1)
    y = _random_mixture(rng)
    _, kf, kb = OF.chemical_source(T, 0.3, t, y)
    hrt, sr = _nasa_independent(T, t)
    g = hrt - sr
    for i in range(T.n_reac):
        nu = np.array(T.stoi_b[i][:9]) - np.array(T.stoi_f[i][:9])
        kc = (101325.0 / (R_U * t) * 1e-3) ** nu.sum() * math.exp(-np.dot(nu, g))
        assert abs(kf[i] / kb[i] - kc) <= 1e-12 * kc, i


def test_rates_vanish_in_pure_inert(T):
    y = np.zeros(9)
    y[SPECIES.index("N2")] = 1.0
    om, _, _ = OF.chemical_source(T, 1.0, 2000.0, y)
    assert np.all(om == 0)


# ---------------------------------------------------------------- transport
def test_pure_species_viscosity_is_chapman_enskog(T):
    """X_k = 1: mu = 5/16 sqrt(pi m k T) / (pi sigma^2 Omega(T*)), Omega from
    the 1.147 T*^-0.145 + (T* + 0.5)^-2 fit; lambda = mu (cp + 5/4 R/W)."""
    kb, na = 1.380649e-23, 6.02214076e26
    t = 1234.0
    for k in (SPECIES.index("H2"), SPECIES.index("O2"), SPECIES.index("N2")):
        y = np.zeros(9)
        y[k] = 1.0
        mu, lam, _ = OF.transport(T, t, 101325.0, 1.0, y)
        m = T.mw[k] / na
        sig = math.sqrt(2.6693e-6 * math.sqrt(T.mw[k]) / T.vis_coeff[k]) * 1e-10
        ts = t * T.lj_potent_inv[k]
        omega = 1.147 * ts ** -0.145 + (ts + 0.5) ** -2
        mu_ce = 5.0 / 16.0 * math.sqrt(math.pi * m * kb * t) / (math.pi * sig * sig * omega)
        assert abs(mu - mu_ce) <= 2e-4 * mu_ce  # 2.6693e-6 carries 5 digits
        _, cp = OF.enthalpy_cp(T, t)
        assert abs(lam - mu * (cp[k] + 1.25 * R_U / T.mw[k])) <= 1e-13 * lam


def test_trace_species_diffuses_with_the_binary_coefficient(T):
    """A trace of H2 in N2: rho D_H2 -> rho D_(H2,N2), the Chapman-Enskog
    binary diffusivity 3/16 sqrt(2 pi (kT)^3 / m_jk) / (p pi sigma^2 Omega_D)."""
    kb, na = 1.380649e-23, 6.02214076e26
    t, p, rho = 900.0, 101325.0, 0.4
    y = np.zeros(9)
    h2, n2 = SPECIES.index("H2"), SPECIES.index("N2")
    y[h2], y[n2] = 1e-9, 1.0 - 1e-9
    _, _, rd = OF.transport(T, t, p, rho, y)
    wjk = T.mw[h2] * T.mw[n2] / (T.mw[h2] + T.mw[n2])
    sig = math.sqrt(0.0188324 / (math.sqrt(wjk) * T.binary_diffusivity_coeff[h2][n2])) * 1e-10
    tr = t * T.kb_over_eps_jk[h2][n2]
    om = 1 / tr ** 0.145 + 1 / (tr + 0.5) ** 2
    d_ce = 3.0 / 16.0 * math.sqrt(2 * math.pi * (kb * t) ** 3 / (wjk / na)) / (p * math.pi * sig * sig * om)
    assert abs(rd[h2] / rho - d_ce) <= 3e-4 * d_ce


# -------------------------------------------------------------- T-Newton
def test_newton_inverts_the_internal_energy(T):
    rng = np.random.default_rng(5)
    mw = np.array(T.mw[:9])
    for _ in range(20):
        y = _random_mixture(rng)
        t_true = rng.uniform(400, 2500)
        h, _ = OF.enthalpy_cp(T, t_true)
        e = np.dot(h, y) - R_U * np.sum(y / mw) * t_true
        t, it = OF.temperature(T, e, y, t_true * rng.uniform(0.9, 1.1))
        assert 1 <= it <= 6
        assert abs(t - t_true) <= 1e-5 * t_true


def test_newton_with_constant_cp_is_exact_after_one_update(T):
    """cp constant (a2..a5 = 0): e(T) is linear, the first Newton update lands
    on the closed form T = (e - a6 R/W) / ((a1 - 1) R/W)."""
    t2 = mech.MechTables()
    C.memmove(C.byref(t2), C.byref(T), C.sizeof(T))
    for i in range(9):
        for rng_ in (t2.low_coeff, t2.high_coeff):
            rng_[i][1] = rng_[i][2] = rng_[i][3] = rng_[i][4] = 0.0
            rng_[i][0], rng_[i][5] = 3.5, -1000.0
        t2.t_low[i] = 10.0
    y = np.zeros(9)
    y[SPECIES.index("N2")] = 1.0
    rw = R_U / t2.mw[SPECIES.index("N2")]
    e = 7.0e5
    t, it = OF.temperature(t2, e, y, 300.0)
    closed = (e - (-1000.0) * rw) / ((3.5 - 1) * rw)
    assert it == 2
    assert abs(t - closed) <= 1e-12 * closed


# ----------------------------------------------------- characteristic WENO
@pytest.mark.parametrize("seed", range(3))
def test_left_and_right_characteristic_matrices_are_inverse(T, seed):
    """At a face whose two sides hold the same state the Roe averages are that
    state, and the left and right eigenvector matrices of the mixture Euler
    Jacobian are exact inverses (L R = I).  Between different states the
    reference's averaged h, c and gamma are not mutually consistent, so the
    identity holds only approximately there (checked loosely)."""
    rng = np.random.default_rng(seed)
    ns = 9

    def state():
        y = _random_mixture(rng)
        rho, u = rng.uniform(0.2, 2), rng.uniform(-300, 300, 3)
        t = rng.uniform(500, 2000)
        mw_inv = np.sum(y / np.array(T.mw[:ns]))
        p = rho * R_U * mw_inv * t
        h, cp = OF.enthalpy_cp(T, t)
        E = rho * (0.5 * u @ u + h @ y) - p
        gam = (cp @ y) / (cp @ y - R_U * mw_inv)
        return np.concatenate([[rho], rho * u, [E], rho * y, [p, math.sqrt(gam * R_U * mw_inv * t)]])

    q = state()
    k = rng.normal(size=3)
    L, R = OF.char_matrices(T, q, q, k, k, 0.7, 0.7)
    assert np.max(np.abs(L @ R - np.eye(5 + ns))) < 1e-12
    q2 = q.copy()
    q2[:5 + ns] *= 1.0 + 1e-6 * rng.normal(size=5 + ns)
    L, R = OF.char_matrices(T, q, q2, k, k * 1.001, 0.7, 0.71)
    assert np.max(np.abs(L @ R - np.eye(5 + ns))) < 1e-4


# ---------------------------------------------------------- whole residual
def _periodic_case(T, n=8, scheme=51, viscous=2, reaction=0):
    c = pkg.cases.reacting_single_block(n, n, n, T, scheme=scheme, viscous_order=viscous, reaction=reaction)
    c["boundaries"] = [dict(block=0, axis=a, side=s, type=8) for a in range(3) for s in (-1, 1)]
    return c


def _load_state(o, rng, uniform=False):
    """A smooth random mixture state (or a uniform one) on the padded domain."""
    bv = o.array(0, OF.BV)
    sv = o.array(0, OF.SV)
    shp = bv.shape[1:]
    if uniform:
        y = _random_mixture(rng)
        sv[...] = y[:, None, None, None]
        bv[0], bv[1], bv[2], bv[3], bv[5] = 0.5, 120.0, -40.0, 25.0, 1300.0
    else:
        # period n - 1 nodes on every axis (node n-1 repeats node 0, as the
        # periodic faces assume)
        g = o.ngg
        n = shp[0] - 2 * g
        per = np.arange(-g, n + g) % (n - 1)
        k, j, i = np.meshgrid(per, per, per, indexing="ij")
        ph = [rng.uniform(0, 2 * np.pi, 3) for _ in range(20)]
        w = lambda q: np.sin(2 * np.pi * i / (n - 1) + ph[q][0]) * np.cos(2 * np.pi * j / (n - 1) + ph[q][1]) + \
            0.5 * np.sin(2 * np.pi * k / (n - 1) + ph[q][2])
        y = np.array([1.0 + 0.5 * w(q) for q in range(9)])
        sv[...] = y / y.sum(axis=0)
        bv[0] = 0.5 * (1 + 0.2 * w(9))
        bv[1], bv[2], bv[3] = 150 * w(10), 80 * w(11), 60 * w(12)
        bv[5] = 1300 * (1 + 0.2 * w(13))
    mw_inv = np.tensordot(1 / np.array(o.T.mw[:9]), sv, axes=1)
    bv[4] = bv[0] * R_U * mw_inv * bv[5]


def _oracle_with(T, case, rng, uniform=False):
    o = OF.FrOracle(case, T, n_threads=4)
    jac, met = o.array(0, OF.JAC), o.array(0, OF.METRIC)
    jac[...] = 0.125
    met[...] = 0.0
    met[0], met[4], met[8] = 2.0, 2.0, 2.0  # uniform grid, spacing 0.5
    _load_state(o, rng, uniform)
    o.L.orf_cv_from_bv(o.ctx, 0)
    o.L.orf_properties(o.ctx, 0)
    return o


@pytest.mark.parametrize("scheme", [51, 52, 71, 72])
def test_free_stream_is_preserved(T, scheme):
    o = _oracle_with(T, _periodic_case(T, scheme=scheme), np.random.default_rng(2), uniform=True)
    dq = o.rhs(0, 1e-9)
    cv = o.array(0, OF.CV)
    for l in range(14):
        assert np.max(np.abs(dq[l])) <= 1e-12 * np.max(np.abs(cv[l])) * 2.0, l


@pytest.mark.parametrize("scheme", [51, 52, 71, 72])
def test_periodic_residual_telescopes(T, scheme):
    """Fully periodic box, chemistry off: dq summed over the distinct nodes
    (node n-1 repeats node 0) is zero for every conserved variable -- the
    face fluxes telescope -- to round-off."""
    o = _oracle_with(T, _periodic_case(T, scheme=scheme), np.random.default_rng(3))
    o.L.orf_apply_bc(o.ctx, 0)
    o.L.orf_properties(o.ctx, 0)
    dq = o.rhs(0, 1e-9)[:, :-1, :-1, :-1]
    for l in range(14):
        s, a = abs(dq[l].sum()), np.abs(dq[l]).sum()
        assert s <= 1e-12 * a, (l, s, a)


def test_species_diffusion_fluxes_sum_to_zero(T):
    """The correction velocity makes the diffusion fluxes sum to zero:
    sum_k dq_visc[5+k] = 0 in every cell (viscous part only)."""
    o = _oracle_with(T, _periodic_case(T, viscous=2), np.random.default_rng(4))
    o.L.orf_apply_bc(o.ctx, 0)
    o.L.orf_properties(o.ctx, 0)
    o.L.orf_zero_dq(o.ctx, 0)
    o.L.orf_viscous(o.ctx, 0)
    dq = o.array(0, OF.DQ)
    s = dq[5:].sum(axis=0)
    assert np.max(np.abs(s)) <= 1e-12 * np.max(np.abs(dq[5:]))
    assert np.all(dq[0] == 0)


# ------------------------------------------ primitive reconstruction path
def _prim(T, rng, u=None, rho=None, p=None, y=None):
    y = _random_mixture(rng) if y is None else y
    u = rng.uniform(-150, 150, 3) if u is None else np.asarray(u, dtype=float)
    return np.concatenate([[rng.uniform(0.3, 1.5) if rho is None else rho], u,
                           [rng.uniform(5e4, 2e5) if p is None else p], y])


def _physical_flux(T, q, k, jac):
    """J (rho U, rho u U + p kx, ..., (E + p) U, rho Y U), U = k . u."""
    rho, u, p, y = q[0], q[1:4], q[4], q[5:]
    mw_inv = np.sum(y / np.array(T.mw[:9]))
    t = p / (rho * R_U * mw_inv)
    h, _ = OF.enthalpy_cp(T, t)
    E = rho * (0.5 * u @ u + h @ y) - p
    U = k @ u
    return jac * np.concatenate([[rho * U], rho * u * U + p * k, [(E + p) * U], rho * y * U])


@pytest.mark.parametrize("scheme", [3, 4])
def test_riemann_flux_is_consistent(T, scheme):
    """F(q, q) is the physical flux (AUSM+: M+ + M- = M, P+ + P- = 1)."""
    rng = np.random.default_rng(scheme)
    for _ in range(10):
        q = _prim(T, rng)
        k = rng.normal(size=3)
        f = OF.riemann(T, scheme, q, q, k, k, 0.8, 0.8)
        ref = _physical_flux(T, q, k, 0.8)
        assert np.max(np.abs(f - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("scheme", [3, 4])
def test_supersonic_flow_takes_the_upwind_flux(T, scheme):
    rng = np.random.default_rng(10 + scheme)
    k = np.array([1.0, 0.0, 0.0])
    ql = _prim(T, rng, u=[2500.0, 10.0, -5.0])
    qr = _prim(T, rng, u=[2300.0, -20.0, 8.0])
    f = OF.riemann(T, scheme, ql, qr, k, k, 1.0, 1.0)
    ref = _physical_flux(T, ql, k, 1.0)
    assert np.max(np.abs(f - ref)) <= 1e-12 * np.max(np.abs(ref))
    f = OF.riemann(T, scheme, qr * np.array([1, -1, -1, -1] + [1] * 10), ql * np.array([1, -1, -1, -1] + [1] * 10),
                   k, k, 1.0, 1.0)
    ref = _physical_flux(T, ql * np.array([1, -1, -1, -1] + [1] * 10), k, 1.0)
    assert np.max(np.abs(f - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_hllc_resolves_an_isolated_contact(T):
    """Same u and p, different rho and Y: the HLLC flux is the upwind
    physical flux (the contact wave is captured exactly)."""
    rng = np.random.default_rng(12)
    k = np.array([0.6, 0.8, 0.0])
    for un in (40.0, -40.0):
        u = un * k
        ql = _prim(T, rng, u=u, p=1e5)
        qr = _prim(T, rng, u=u, p=1e5)
        f = OF.riemann(T, 4, ql, qr, k, k, 1.0, 1.0)
        ref = _physical_flux(T, ql if un > 0 else qr, k, 1.0)
        assert np.max(np.abs(f - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_roe_flux_is_consistent(T):
    """Roe (src/RiemannSolver.cuh:234-417): F(q, q) is the physical flux
    whatever the entropy-fix width (no jump, no dissipation)."""
    rng = np.random.default_rng(21)
    for _ in range(10):
        q = _prim(T, rng)
        k = rng.normal(size=3)
        f = OF.roe(T, q, q, k, k, 0.8, 0.8, rng.uniform(0, 300), rng.uniform(0, 300))
        ref = _physical_flux(T, q, k, 0.8)
        assert np.max(np.abs(f - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_roe_flux_upwinds_weak_supersonic_jumps_to_second_order(T, sign):
    """With every characteristic of one sign, R |Lambda| L dQ = +-A dQ, and
    A dQ = F(Q_r) - F(Q_l) + O(|dQ|^2): the Roe flux is the upwind physical
    flux up to an error that, relative to the flux jump, vanishes with the
    jump.  A sign error in the dissipation, a dropped species term or a
    transposed wave strength leaves an error of the size of the jump."""
    rng = np.random.default_rng(22 if sign > 0 else 23)
    k = np.array([0.9, 0.3, -0.2])
    ql = _prim(T, rng, u=sign * np.array([2600.0, 40.0, -30.0]), p=1e5, rho=0.6)
    d = np.concatenate([[0.05, 90.0, -60.0, 45.0, 4e3], rng.normal(size=9)])
    d[5:] -= d[5:].mean()   # mass fractions still sum to one
    d[5:] *= 0.2
    errs = []
    for eps in (1e-3, 5e-4, 2.5e-4):
        qr = ql + eps * d
        f = OF.roe(T, ql, qr, k, k, 1.0, 1.0, 0.0, 0.0)
        up = ql if sign > 0 else qr
        ref = _physical_flux(T, up, k, 1.0)
        jump = _physical_flux(T, qr, k, 1.0) - _physical_flux(T, ql, k, 1.0)
        errs.append(np.max(np.abs(f - ref) / np.maximum(np.abs(jump), 1e-300 + np.max(np.abs(jump)) * 1e-6)))
    assert all(e <= 10 * eps for e, eps in zip(errs, (1e-3, 5e-4, 2.5e-4))), errs


def test_roe_entropy_fix_only_acts_below_the_width(T):
    """Harten's fix replaces |lambda| by (lambda^2/delta + delta)/2 only
    where |lambda| < delta: a width below every |lambda| changes nothing, a
    wider one adds dissipation (a different flux)."""
    rng = np.random.default_rng(24)
    k = np.array([1.0, 0.0, 0.0])
    ql = _prim(T, rng, u=[30.0, 0.0, 0.0], p=1e5)
    qr = _prim(T, rng, u=[35.0, 5.0, 0.0], p=1.02e5)
    f0 = OF.roe(T, ql, qr, k, k, 1.0, 1.0, 0.0, 0.0)
    f1 = OF.roe(T, ql, qr, k, k, 1.0, 1.0, 1.0, 2.0)    # below |u| ~ 32
    assert np.array_equal(f0, f1)
    f2 = OF.roe(T, ql, qr, k, k, 1.0, 1.0, 80.0, 80.0)  # above |u|, below c
    assert not np.allclose(f0, f2, rtol=1e-8, atol=0)


@pytest.mark.parametrize("recon", [2, 3])
def test_limited_reconstruction_is_exact_on_linear_data(recon):
    pv = np.outer(np.arange(4.0), np.linspace(-2, 3, 14)) + np.linspace(1, 5, 14)
    pl, pr = OF.reconstruct(pv, recon)
    mid = 0.5 * (pv[1] + pv[2])
    assert np.max(np.abs(pl - mid)) <= 1e-14 * np.max(np.abs(mid))
    assert np.max(np.abs(pr - mid)) <= 1e-14 * np.max(np.abs(mid))


@pytest.mark.parametrize("recon", [1, 2, 3])
def test_reconstruction_is_first_order_at_extrema(recon):
    pv = np.array([[1.0], [3.0], [2.0], [0.5]])  # extremum at point 1 and a monotone drop to point 3
    pl, pr = OF.reconstruct(pv, recon)
    assert pl[0] == 3.0  # minmod of opposite-signed slopes is 0
    if recon == 1:
        assert pr[0] == 2.0
    else:
        assert 0.5 <= pr[0] <= 3.0


@pytest.mark.parametrize("scheme", [3, 4])
@pytest.mark.parametrize("recon", [1, 2, 3])
def test_pv_path_preserves_free_stream_and_conserves(T, scheme, recon):
    case = _periodic_case(T, scheme=scheme)
    case["reconstruction"] = recon
    o = _oracle_with(T, case, np.random.default_rng(6), uniform=True)
    dq = o.rhs(0, 1e-9)
    cv = o.array(0, OF.CV)
    for l in range(14):
        assert np.max(np.abs(dq[l])) <= 2e-12 * np.max(np.abs(cv[l])), l
    o = _oracle_with(T, case, np.random.default_rng(7))
    o.L.orf_apply_bc(o.ctx, 0)
    o.L.orf_properties(o.ctx, 0)
    dq = o.rhs(0, 1e-9)[:, :-1, :-1, :-1]
    for l in range(14):
        assert abs(dq[l].sum()) <= 1e-12 * np.abs(dq[l]).sum(), l


def _wave_dq(T, scheme, n):
    """Mixture residual of a small-amplitude plane wave along (1, 2, -1) in
    index space, n - 1 nodes per period, viscous fluxes and chemistry off."""
    o = OF.FrOracle(_periodic_case(T, n=n, scheme=scheme, viscous=0), T, n_threads=4)
    jac, met = o.array(0, OF.JAC), o.array(0, OF.METRIC)
    jac[...] = 0.125
    met[...] = 0.0
    met[0], met[4], met[8] = 2.0, 2.0, 2.0
    g = o.ngg
    per = np.arange(-g, n + g) % (n - 1)
    k, j, i = np.meshgrid(per, per, per, indexing="ij")
    w = 2 * np.pi * (i + 2 * j - k) / (n - 1)
    y = np.array([0.01, 0.02, 0.05, 0.001, 0.002, 0.01, 0.2, 0.02, 0.687])
    sv = o.array(0, OF.SV)
    bv = o.array(0, OF.BV)
    for l in range(9):
        sv[l] = y[l] * (1 + 1e-2 * np.sin(w + 0.1 * l))
    sv[...] /= sv.sum(axis=0)
    bv[0] = 0.5 * (1 + 1e-2 * np.sin(w))
    bv[1], bv[2], bv[3] = 40 + 0.5 * np.cos(w), -20 + 0.3 * np.sin(w + 1), 10 + 0 * w
    bv[5] = 1300 * (1 + 1e-2 * np.sin(w + 0.3))
    mw_inv = np.tensordot(1 / np.array(T.mw[:9]), sv, axes=1)
    bv[4] = bv[0] * R_U * mw_inv * bv[5]
    o.L.orf_cv_from_bv(o.ctx, 0)
    o.L.orf_properties(o.ctx, 0)
    return o.rhs(0, 1e-9).copy()


def test_mixture_characteristic_and_componentwise_weno_converge_together(T):
    """Smooth limit: the characteristic residuals (52, 72) approach the
    component-wise ones (51, 71) at the schemes' order, for the flow rows and
    the species rows; a stencil shifted by one point stays at O(h k)."""
    err = {}
    for n in (20, 30):
        d = {s: _wave_dq(T, s, n) for s in (51, 52, 71, 72)}
        for l in range(14):
            scale = np.max(np.abs(d[71][l]))
            e72 = np.max(np.abs(d[72][l] - d[71][l])) / scale
            e52 = np.max(np.abs(d[52][l] - d[51][l])) / scale
            err[(n, l)] = (e72, e52)
            assert e72 <= 1e-2 and e52 <= 3e-2, (n, l, e72, e52)
            assert np.max(np.abs(np.roll(d[72][l], 1, axis=1) - d[71][l])) / scale > 0.2
    for l in range(14):
        assert err[(30, l)][0] <= 0.1 * err[(20, l)][0], (l, err[(20, l)], err[(30, l)])
        assert err[(30, l)][1] <= 0.3 * err[(20, l)][1], (l, err[(20, l)], err[(30, l)])


def test_mixture_residual_is_the_rms_of_the_primitive_changes(T):
    """compute_residual with species: sqrt(sum(dv^2) / N) of rho, |V|, p, T
    over the interior, against numpy, after two steps of the mixing layer."""
    case = pkg.cases.reacting_case([7, 6], [8, 9], [5, 6], T, scheme=52)
    o = OF.FrOracle(case, T, n_threads=4)
    o.setup_mixing_layer()
    o.initialize()
    o.store_last_step()
    g = o.ngg
    before = [o.array(b, OF.BV)[:, g:-g, g:-g, g:-g].copy() for b in range(len(case["blocks"]))]
    o.step(2)
    acc, n = np.zeros(4), 0
    for b, old in enumerate(before):
        new = o.array(b, OF.BV)[:, g:-g, g:-g, g:-g]
        vel = lambda q: np.sqrt(q[1] ** 2 + q[2] ** 2 + q[3] ** 2)
        acc += [np.sum((new[0] - old[0]) ** 2), np.sum((vel(new) - vel(old)) ** 2),
                np.sum((new[4] - old[4]) ** 2), np.sum((new[5] - old[5]) ** 2)]
        n += new[0].size
    want = np.sqrt(acc / n)
    np.testing.assert_allclose(o.residual(), want, rtol=1e-12)
    assert np.all(want > 0)


# ------------------------------------------------------------ falloff rates
def _copy(T):
    return type(T).from_buffer_copy(T)


def _third_body_conc(T, r, rho, y):
    """[M] of reaction r in mol/cm^3 (c_k = rho Y_k / W_k * 1e-3)."""
    return sum(T.third_body_coeff[r][k] * rho * y[k] / T.mw[k] * 1e-3 for k in range(T.n_spec))


@pytest.mark.parametrize("r", [8, 14])
def test_falloff_rates_reduce_to_their_limits(T, r):
    """Lindemann: 1/kf = 1/(k0 [M]) + 1/k_inf, so two densities give k0 and
    k_inf, which must be the low- and high-pressure Arrhenius laws of the
    tables.  Troe with alpha = 1 (and the file's T3 = 1e-30, T1 = 1e30) has
    F_cent = 1 and equals Lindemann; with the file's alpha, at the reduced
    pressure where log10 Pr = -c = 0.4 + 0.67 log10 F_cent the broadening
    factor is F_cent = alpha exactly."""
    rng = np.random.default_rng(r)
    y = _random_mixture(rng)
    t = 1350.0
    lind = _copy(T)
    lind.label[r] = 5
    kf = {}
    for rho in (0.02, 0.5):
        kf[rho] = OF.chemical_source(lind, rho, t, y)[1][r]
    m1, m2 = _third_body_conc(T, r, 0.02, y), _third_body_conc(T, r, 0.5, y)
    k0 = (1 / m1 - 1 / m2) / (1 / kf[0.02] - 1 / kf[0.5])
    k_inf = 1 / (1 / kf[0.5] - 1 / (k0 * m2))
    arr = lambda A, b, Ea: A * t ** b * math.exp(-Ea / t)
    assert math.isclose(k0, arr(T.A2[r], T.b2[r], T.Ea2[r]), rel_tol=1e-9)
    assert math.isclose(k_inf, arr(T.A[r], T.b[r], T.Ea[r]), rel_tol=1e-9)
    # alpha = 1: Troe == Lindemann, bit for bit (F = 10**0)
    troe1 = _copy(T)
    troe1.troe_alpha[r] = 1.0
    assert T.label[r] == 6
    for rho in (0.02, 0.5):
        assert OF.chemical_source(troe1, rho, t, y)[1][r] == kf[rho]
    # the file's alpha at log10 Pr = -c
    alpha = T.troe_alpha[r]
    c = -0.4 - 0.67 * math.log10(alpha)
    m_star = 10 ** (-c) * k_inf / k0
    rho_star = 0.5 * m_star / m2
    kt = OF.chemical_source(T, rho_star, t, y)[1][r]
    kl = OF.chemical_source(lind, rho_star, t, y)[1][r]
    assert math.isclose(kt / kl, alpha, rel_tol=1e-10), (kt / kl, alpha)
