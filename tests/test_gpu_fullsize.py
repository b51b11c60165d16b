"""Parity at the sizes the benchmark reports (needs a GPU).

* the 24-block batched air case (12.5 M cells): 20 SSP-RK3 steps through the
  per-block path and through the batched fused kernel, primitives within
  1e-9 of the oracle's 20 steps and dt within 1e-12;
* the 512 x 256 x 128 reacting layer (H2/air, 9 species, 19 reactions,
  characteristic WENO-Z5): the first stage residual dq on the full grid,
  flow rows within 1e-12 and species rows within 1e-10 of the oracle;
* a 64 x 32 x 32 box of the same reacting layer: 50 steps, primitives and
  mass fractions within 1e-9;
* the 1024 x 512 x 256 reacting layer in 8 blocks: the GPU initialises the
  whole partition (boundary conditions and in-GPU face exchange), and the
  stage residual of its smallest block (about 10.7 M cells) is checked
  against the oracle evaluated on that block's state, ghosts included.

The oracle runs with OpenMP on every host core it is given; these cases take
minutes of host time.
"""
import numpy as np
import pytest

from _util import O, cases, interior, oracle_for, pkg, rel_err
import mech
import oracle_fr as OF

pytestmark = pytest.mark.gpu

MX = pkg.mixture
FILES = MX.H2_AIR_FILES
C2_STEPS = 20
_C2 = {}


def _c2_oracle():
    if not _C2:
        case = cases.batched_case()
        print(f"  oracle: {len(case['blocks'])} blocks, {case['n_cells']} cells, {C2_STEPS} steps", flush=True)
        o = oracle_for(case)
        for n in range(C2_STEPS):
            dt = o.step(1)
            print(f"  oracle step {n + 1} done", flush=True)
        _C2.update(case=case, dt=dt, bv=[interior(o.array(b, O.BV), o.ngg).copy() for b in range(len(case["blocks"]))])
        del o
    return _C2


@pytest.mark.parametrize("mode", [pkg.MODE_BATCHED, pkg.MODE_PLAIN], ids=["batched", "plain"])
def test_batched_24_blocks_20_steps(mode):
    print(f"[gpu] 24-block case, {C2_STEPS} steps, mode {mode}", flush=True)
    ref = _c2_oracle()
    case = ref["case"]
    s = pkg.Solver(case)
    s.setup_mixing_layer()
    s.initialize()
    dt = s.rk3(C2_STEPS, mode)
    print(f"  dt gpu {dt:.15e} oracle {ref['dt']:.15e}", flush=True)
    assert abs(dt - ref["dt"]) <= 1e-12 * ref["dt"]
    worst = 0.0
    for b in range(len(case["blocks"])):
        e = rel_err(interior(s.get(b, pkg.FIELD_BV), s.ngg), ref["bv"][b])
        worst = max(worst, float(e.max()))
        assert np.all(e <= 1e-9), (b, e)
    print(f"  worst primitive rel err {worst:.2e}", flush=True)
    s.close()


@pytest.fixture(scope="module")
def tables():
    return MX.load_mechanism(*FILES), mech.load(*FILES)


def _check_dq(dq_g, dq_o, label):
    err = rel_err(dq_g, dq_o)
    print(f"  {label} dq per-variable: " + " ".join(f"{e:.1e}" for e in err), flush=True)
    assert np.all(err[:5] <= 1e-12), err
    assert np.all(err[5:] <= 1e-10), err


def test_reacting_512x256x128_first_stage_dq(tables):
    print("[gpu] reacting 512x256x128, first stage dq on the full grid", flush=True)
    tg, to = tables
    case = cases.reacting_case([511], [255], [127], tg, scheme=52, spatial=True)
    s = MX.FrSolver(case, tg)
    s.setup_mixing_layer()
    s.initialize()
    s.compute_rhs(0, 0.0)
    dq_g = s.get(0, MX.FIELD_DQ)
    s.close()
    print("  gpu done; oracle setup", flush=True)
    o = OF.FrOracle(case, to)
    o.setup_mixing_layer()
    o.initialize()
    print("  oracle initialised; oracle rhs", flush=True)
    _check_dq(dq_g, o.rhs(0, 0.0), "512x256x128")


def test_reacting_512x256x128_stencil_difference(tables):
    """How far the reference's stencil of the acoustic/entropy fields
    (ch_stencil 1, src/WENO.cu:880-881) moves the first stage residual of the
    512 x 256 x 128 layer from the consistent stencil (ch_stencil 0): the
    relative difference per variable, and the share of nodes where dq moves
    by more than 1e-12 of its scale."""
    print("[gpu] reacting 512x256x128: dq of the two scheme-52 stencils", flush=True)
    tg, _ = tables
    dq = {}
    for st in (0, 1):
        case = cases.reacting_case([511], [255], [127], tg, scheme=52, spatial=True)
        case["ch_stencil"] = st
        s = MX.FrSolver(case, tg)
        s.setup_mixing_layer()
        s.initialize()
        s.compute_rhs(0, 0.0)
        dq[st] = s.get(0, MX.FIELD_DQ)
        s.close()
    scale = np.max(np.abs(dq[0]), axis=(1, 2, 3))
    diff = np.abs(dq[1] - dq[0])
    live = scale > 0   # species absent from both streams have dq = 0 everywhere
    rel = np.where(live, np.max(diff, axis=(1, 2, 3)) / np.where(live, scale, 1.0), 0.0)
    moved = np.mean(diff > 1e-12 * scale[:, None, None, None], axis=(1, 2, 3))
    print("  max relative difference per variable: " + " ".join(f"{e:.2e}" for e in rel), flush=True)
    print("  share of nodes moved: " + " ".join(f"{m:.3f}" for m in moved), flush=True)
    assert np.all(rel[:5] > 1e-6)


def test_reacting_64x32x32_50_steps(tables):
    print("[gpu] reacting 64x32x32 box of the spatially developing layer, 50 steps", flush=True)
    tg, to = tables
    case = cases.reacting_case([63], [31], [31], tg, scheme=52, spatial=True)
    o = OF.FrOracle(case, to)
    o.setup_mixing_layer()
    o.initialize()
    s = MX.FrSolver(case, tg)
    s.setup_mixing_layer()
    s.initialize()
    dt_g = s.rk3(50, MX.MODE_GRAPH)
    print("  gpu done; 50 oracle steps", flush=True)
    dt_o = o.step(50)
    print(f"  dt oracle {dt_o:.15e} gpu {dt_g:.15e}", flush=True)
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o
    g = s.ngg
    for fo, fg in ((OF.BV, MX.FIELD_BV), (OF.SV, MX.FIELD_SV)):
        e = rel_err(interior(s.get(0, fg), g), interior(o.array(0, fo), g))
        print(f"  field {fo}: worst rel err {e.max():.2e}", flush=True)
        assert np.all(e <= 1e-9), e
    s.close()


def test_reacting_1024x512x256_8_blocks_smallest_block_dq(tables):
    print("[gpu] reacting 1024x512x256 in 8 blocks: dq of the smallest block", flush=True)
    tg, to = tables
    rng = np.random.default_rng(42)   # the cuts of the benchmark's 8-block case
    xs = int(rng.integers(384, 640, size=1)[0])
    ys = int(rng.integers(192, 320, size=1)[0])
    zs = int(rng.integers(96, 160, size=1)[0])
    case = cases.reacting_case([xs, 1023 - xs], [ys, 511 - ys], [zs, 255 - zs], tg, scheme=52)
    sizes = [b["mx"] * b["my"] * b["mz"] for b in case["blocks"]]
    b = int(np.argmin(sizes))
    print(f"  block {b}: {sizes[b]} cells of {sum(sizes)}", flush=True)
    s = MX.FrSolver(case, tg)
    s.setup_mixing_layer()
    s.initialize()
    s.compute_rhs(b, 0.0)
    fields = {f: s.get(b, f) for f in (MX.FIELD_CV, MX.FIELD_BV, MX.FIELD_SV, MX.FIELD_MUL, MX.FIELD_LAMBDA,
                                       MX.FIELD_CP, MX.FIELD_GAMMA, MX.FIELD_C, MX.FIELD_MACH, MX.FIELD_RHOD,
                                       MX.FIELD_JAC, MX.FIELD_METRIC, MX.FIELD_DQ)}
    s.close()
    print("  gpu done; oracle on the block's state", flush=True)
    one = dict(case, blocks=[case["blocks"][b]], boundaries=[], inner_faces=[])
    o = OF.FrOracle(one, to)
    for fg, fo in ((MX.FIELD_CV, OF.CV), (MX.FIELD_BV, OF.BV), (MX.FIELD_SV, OF.SV), (MX.FIELD_MUL, OF.MUL),
                   (MX.FIELD_LAMBDA, OF.LAMBDA), (MX.FIELD_CP, OF.CP), (MX.FIELD_GAMMA, OF.GAMMA),
                   (MX.FIELD_C, OF.C_SOUND), (MX.FIELD_MACH, OF.MACH), (MX.FIELD_RHOD, OF.RHOD),
                   (MX.FIELD_JAC, OF.JAC), (MX.FIELD_METRIC, OF.METRIC)):
        o.array(0, fo)[...] = fields[fg].reshape(o.array(0, fo).shape)
    _check_dq(fields[MX.FIELD_DQ], o.rhs(0, 0.0), f"block {b}")
