"""Edge cases of the HIP path against the oracle: an unrestricted molecule
with an empty beta channel, shells up to h (pure and Cartesian g and h: AO
values, screening masks, densities, derived fields, MO amplitudes and the
electrostatic potential with pair orders up to 10), and the Boys function of
the ESP pair loop across its table, the T = 40 switch and large T."""
import numpy as np
import pytest

from kumo_hip import synthetic
from kumo_hip.api import Backend, DensityComponent as DC, Precision, PropertyKind as K
from support import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from kumo_hip.binding import Context
    c = Context(0)
    yield c
    c.close()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# ------------------------------------------------------------ empty beta channel

NO_BETA_PROPS = [(K.Density, DC.Total, 0, 0), (K.Density, DC.Alpha, 0, 0), (K.Density, DC.Beta, 0, 0),
                 (K.Density, DC.Difference, 0, 0), (K.Gradient, DC.Beta, 0, 0), (K.Laplacian, DC.Beta, 0, 0),
                 (K.Tau, DC.Beta, 0, 0), (K.Elf, DC.Total, 0, 0), (K.Elf, DC.Alpha, 0, 0), (K.Elf, DC.Beta, 0, 0),
                 (K.Lol, DC.Beta, 0, 0)]


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.BF16X6])
@pytest.mark.parametrize("backend", [Backend.AO_MO, Backend.AO_P])
def test_empty_beta_channel(ctx, precision, backend):
    from kumo_hip.binding import Molecule
    sm, grid = synthetic.no_beta_molecule()
    g = Molecule(ctx, data=sm)
    o = oracle.Molecule(sm)
    assert g.info["nbeta"] == 0 and g.info["unrestricted"] == 1
    print(f"[gpu] empty beta channel, backend {backend.name}, {precision.name}", flush=True)
    got, used = g.evaluate(NO_BETA_PROPS, grid=grid, backend=backend, precision=precision)
    ref, rused = o.evaluate(NO_BETA_PROPS, grid=grid, backend=backend)
    assert used == rused == int(backend)
    tol = 1e-11 if precision == Precision.FP64 else 1e-6
    for p, a, b in zip(NO_BETA_PROPS, got, ref):
        for ca, cb in zip(a, b):
            if p[1] == DC.Beta and p[0] != K.Elf:
                # nothing in the channel: exactly zero, as in the oracle
                assert np.all(ca == 0.0) and np.all(cb == 0.0), p
            elif p[0] == K.Elf:
                assert np.max(np.abs(ca - cb)[ref[0][0] >= 1e-4]) <= 1e-5, p
            else:
                assert rel(ca, cb) <= tol, (p, rel(ca, cb))
    # the alpha channel carries everything
    assert np.array_equal(got[0][0], got[1][0])
    g.close()


# ------------------------------------------------------------- shells up to h

@pytest.fixture(scope="module")
def high_l(ctx):
    from kumo_hip.binding import Molecule
    sm, grid = synthetic.high_l_molecule()
    g = Molecule(ctx, data=sm)
    o = oracle.Molecule(sm)
    yield sm, grid, g, o
    g.close()


def test_high_l_blocks(high_l):
    sm, _, g, o = high_l
    assert g.info["max_l"] == 5 and g.info["nbasis"] == sm.nbasis == 74
    for a, b in zip(g.blocks(), o.basis.blocks()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("order", [0, 1, 2])
def test_high_l_ao_values(high_l, order):
    sm, _, g, o = high_l
    rng = np.random.default_rng(40 + order)
    pts = sm.xyz[rng.integers(0, 3, 300)] + rng.normal(size=(300, 3)) * 1.2
    print(f"[gpu] high-l AO values order {order}", flush=True)
    a = g.ao(pts, order)
    b = o.basis.ao(pts, order)
    scale = np.abs(b).max(axis=1, keepdims=True) + 1e-300
    assert np.max(np.abs(a - b) / scale) < 1e-13


@pytest.mark.parametrize("order", [0, 1, 2])
def test_high_l_masks(high_l, order):
    sm, grid, g, o = high_l
    print(f"[gpu] high-l masks order {order}", flush=True)
    for block in (128, 300):
        gm = g.masks(block, order, grid=grid)
        om = o.masks(block, order, grid=grid)
        assert np.array_equal(gm, om), block
    rng = np.random.default_rng(7)
    pts = sm.xyz[rng.integers(0, 3, 1000)] + rng.normal(size=(1000, 3)) * 2.5
    assert np.array_equal(g.masks(64, order, points=pts), o.masks(64, order, points=pts))


HIGH_L_PROPS = [(K.Density, DC.Total, 0, 0), (K.Density, DC.Beta, 0, 0), (K.Gradient, DC.Total, 0, 0),
                (K.Laplacian, DC.Total, 0, 0), (K.Hessian, DC.Alpha, 0, 0), (K.Tau, DC.Total, 0, 0),
                (K.Elf, DC.Total, 0, 0), (K.SignLambda2Density, DC.Total, 0, 0)]


@pytest.mark.parametrize("backend", [Backend.AO_MO, Backend.AO_P])
def test_high_l_fields(high_l, backend):
    _, grid, g, o = high_l
    print(f"[gpu] high-l fields backend {backend.name}", flush=True)
    got, used = g.evaluate(HIGH_L_PROPS, grid=grid, backend=backend)
    ref, rused = o.evaluate(HIGH_L_PROPS, grid=grid, backend=backend)
    assert used == rused
    rho = ref[0][0]
    for p, a, b in zip(HIGH_L_PROPS, got, ref):
        for ca, cb in zip(a, b):
            if p[0] == K.Elf:
                assert np.max(np.abs(ca - cb)[rho >= 1e-6]) <= 1e-8, p
            elif p[0] == K.SignLambda2Density:
                # |sign(lambda2) rho| = rho; the sign may differ only where lambda2 is at rounding level
                assert rel(np.abs(ca), np.abs(cb)) <= 1e-11
                assert np.count_nonzero(np.sign(ca) != np.sign(cb)) <= 2
            else:
                assert rel(ca, cb) <= 1e-11, (p, rel(ca, cb))


def test_high_l_reduced_precision(high_l):
    _, grid, g, o = high_l
    props = [(K.Density, DC.Total, 0, 0), (K.Elf, DC.Total, 0, 0)]
    got, _ = g.evaluate(props, grid=grid, precision=Precision.BF16X6)
    ref, _ = o.evaluate(props, grid=grid)
    rho = ref[0][0]
    sig = rho >= 1e-6
    assert np.max(np.abs(got[0][0] - rho)[sig] / rho[sig]) <= 1e-6
    assert np.max(np.abs(got[1][0] - ref[1][0])[rho >= 1e-4]) <= 1e-5


def test_high_l_orbitals(high_l):
    sm, grid, g, o = high_l
    props = [(K.Orbital, 0, 0, 1), (K.Orbital, 0, 0, 8), (K.Orbital, 0, 1, 7), (K.Orbital, 0, 0, 74)]
    got, _ = g.evaluate(props, grid=grid)
    ref, _ = o.evaluate(props, grid=grid)
    for a, b in zip(got, ref):
        assert rel(a[0], b[0]) <= 1e-12
    pts = sm.xyz[:1] + np.random.default_rng(3).normal(size=(200, 3))
    pp = [(K.OrbitalGradient, 0, 0, 8), (K.OrbitalHessian, 0, 1, 3)]
    got, _ = g.evaluate(pp, points=pts)
    ref, _ = o.evaluate(pp, points=pts)
    for a, b in zip(got, ref):
        for ca, cb in zip(a, b):
            assert rel(ca, cb) <= 1e-11


def test_high_l_esp(high_l):
    sm, _, g, o = high_l
    rng = np.random.default_rng(9)
    pts = sm.xyz[rng.integers(0, 3, 400)] + rng.normal(size=(400, 3)) * 2.0
    print(f"[gpu] high-l ESP, {o.esp_pair_count()} pairs", flush=True)
    got, _ = g.evaluate([(K.ElectrostaticPotential, 0, 0, 0)], points=pts)
    ref = o.esp(pts)
    assert np.max(np.abs(got[0][0] - ref) / np.maximum(np.abs(ref), 1e-3)) < 1e-10


# ------------------------------------------------- the pair loop's Boys function

def boys_arguments():
    k = np.arange(800)
    mids = np.concatenate([k * 0.05 + 0.025, k * 0.05 - 0.025, k * 0.05 + 0.0249999])
    special = [0.0, 1e-300, 1e-10, 1e-3, 39.99, np.nextafter(40.0, 0.0), 40.0, np.nextafter(40.0, 100.0), 40.5,
               55.5, 1e2, 1e3, 1e6, 1e12]
    rng = np.random.default_rng(12)
    return np.concatenate([special, mids[mids >= 0], rng.uniform(0, 45, 3000), 10 ** rng.uniform(1.6, 8, 500)])


@pytest.mark.parametrize("top", list(range(11)))
def test_pair_loop_boys_matches_oracle(ctx, top):
    T = boys_arguments()
    g = ctx.boys_esp(T, top)
    o = np.array([oracle.boys(t, top) for t in T])
    # The same scheme with products by rounded reciprocals, Horner's rule and a
    # refined hardware reciprocal square root. The Taylor sum (or the closed
    # form above T = 40) agrees to about 10 units in the last place; each step
    # of the recursion to the other orders multiplies by a rounded reciprocal
    # where the reference divides (one more rounding per step), and the step
    # errors add up: 12 + 3 x TOP ulp.
    ulp = np.abs(g - o) / np.spacing(np.abs(o))
    worst = np.unravel_index(np.argmax(ulp), ulp.shape)
    print(f"[gpu] boys_esp top {top}: max {ulp.max():.1f} ulp at T = {T[worst[0]]!r}, order {worst[1]}", flush=True)
    assert np.all(np.isfinite(g))
    assert ulp.max() <= 12 + 3 * top
