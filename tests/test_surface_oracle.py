"""Pins of the oracle's molecular surfaces (src/surface.cpp restated), before
it checks the HIP path:

* tests/molecular_surface_test.cpp: the isosurface of |r| (a sphere) is a
  closed, consistently oriented 2-manifold of Euler characteristic 2 with its
  vertices on the sphere within 1e-9 bohr, an inscribed polyhedron whose area
  and volume converge at second order, and the statistics of n.r on it match
  their closed forms (hemisphere averages R/2, variances R^2/12, nu = 1/4,
  Archimedes' band area); a regular octahedron with V = x has exact
  statistics; extraction rejects invalid input and gives an empty mesh when no
  sample reaches the isovalue.
* tests/multiwfn_surface_reference_test.cpp: Multiwfn's quantitative surface
  analyses of the ESP, ALIE, LEAE and LEA of water, O2 and acrolein, with the
  reference test's tolerances.
"""
import math

import numpy as np
import pytest

from kumo_hip.api import PropertyKind as K
from support import oracle
from support.fchk import FchkData
from support.reference_cases import FCHK
from support.surface_cases import MAPPED, SURFACE_CASES, check_surface, read_report

_MOLS = {}


def molecule(name):
    if name not in _MOLS:
        _MOLS[name] = oracle.Molecule(FchkData(FCHK[name]))
    return _MOLS[name]


def cube_grid(half, spacing):
    n = int(math.ceil(2 * half / spacing)) + 1
    return {"origin": np.full(3, -half) + np.array([0.013, 0.007, 0.021]), "steps": np.eye(3) * spacing,
            "counts": [n, n, n]}


def grid_xyz(grid):
    c = grid["counts"]
    i, j, k = np.meshgrid(np.arange(c[0]), np.arange(c[1]), np.arange(c[2]), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(float)
    return grid["origin"] + ijk @ np.asarray(grid["steps"]).T


def closed_euler(v, t):
    """Every directed edge once and its reverse once; V - E + F."""
    edges = {}
    for tri in t:
        for a in range(3):
            e = (int(tri[a]), int(tri[(a + 1) % 3]))
            edges[e] = edges.get(e, 0) + 1
    bad = sum(1 for e, c in edges.items() if c != 1 or (e[1], e[0]) not in edges)
    assert bad == 0, f"{bad} directed edges are not shared by two oppositely oriented triangles"
    return len(v) - len(edges) // 2 + len(t)


def sphere(radius, spacing):
    grid = cube_grid(radius + 1.0, spacing)
    samples = -np.linalg.norm(grid_xyz(grid), axis=1)

    def f(p):
        r = np.linalg.norm(p, axis=1)
        return -r, -p / r[:, None]
    return oracle.isosurface(grid, samples, -radius, f)


def test_sphere():
    radius = 1.7
    direction = np.array([0.3, 0.2, 0.9])
    direction /= np.linalg.norm(direction)
    prev = None
    for spacing in (0.2, 0.1):
        v, t = sphere(radius, spacing)
        assert closed_euler(v, t) == 2
        assert np.abs(np.linalg.norm(v, axis=1) - radius).max() <= 1e-9
        area, volume = oracle.mesh_area(v, t), oracle.mesh_volume(v, t)
        ea = (4 * math.pi * radius ** 2 - area) / (4 * math.pi * radius ** 2)
        ev = (4 / 3 * math.pi * radius ** 3 - volume) / (4 / 3 * math.pi * radius ** 3)
        assert 0 < ea < 3e-3 and 0 < ev < 6e-3
        if prev:
            assert prev[0] / ea > 3 and prev[1] / ev > 3
        prev = (ea, ev)
        s = oracle.surface_statistics(v, t, v @ direction, 0.4)
        rel = 4e-3
        assert s["area"] == pytest.approx(area, rel=1e-12)
        assert s["positive_area"] / s["area"] == pytest.approx(0.5, abs=rel)
        assert abs(s["average"]) < 2e-4 * radius
        for key, want in (("positive_average", radius / 2), ("negative_average", -radius / 2),
                          ("positive_variance", radius ** 2 / 12), ("negative_variance", radius ** 2 / 12),
                          ("charge_separation", radius / 2), ("polarity_index", radius / 2),
                          ("nonpolar_area", 4 * math.pi * radius * 0.4)):
            assert s[key] == pytest.approx(want, rel=rel), key
        assert s["balance"] == pytest.approx(0.25, abs=1e-6)
        assert len(s["local_minima"]) == 1 and len(s["local_maxima"]) == 1
        assert s["local_minima"][0] == s["minimum_vertex"] and s["local_maxima"][0] == s["maximum_vertex"]
        assert np.linalg.norm(v[s["maximum_vertex"]] - radius * direction) < spacing
        assert s["maximum"] == pytest.approx(radius, abs=spacing ** 2)
        assert s["minimum"] == pytest.approx(-radius, abs=spacing ** 2)


def octahedron():
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=float)
    t = []
    for x in (0, 1):
        for y in (2, 3):
            for z in (4, 5):
                tri = [x, y, z]
                if np.dot(np.cross(v[y] - v[x], v[z] - v[x]), v[x]) < 0:
                    tri[1], tri[2] = tri[2], tri[1]
                t.append(tri)
    return v, np.array(t)


def test_octahedron_exact_statistics():
    v, t = octahedron()
    r3 = math.sqrt(3.0)
    assert closed_euler(v, t) == 2
    assert oracle.mesh_area(v, t) == pytest.approx(4 * r3, rel=1e-14)
    assert oracle.mesh_volume(v, t) == pytest.approx(4 / 3, rel=1e-14)
    s = oracle.surface_statistics(v, t, v[:, 0], 0.5)
    tol = 1e-14
    for key, want in (("area", 4 * r3), ("positive_area", 2 * r3), ("negative_area", 2 * r3),
                      ("positive_average", 1 / 3), ("negative_average", -1 / 3), ("positive_variance", 1 / 18),
                      ("negative_variance", 1 / 18), ("total_variance", 1 / 9), ("balance", 0.25),
                      ("charge_separation", 1 / 3), ("polarity_index", 1 / 3), ("nonpolar_area", 3 * r3)):
        assert s[key] == pytest.approx(want, rel=tol), key
    assert abs(s["average"]) <= tol
    assert s["minimum_vertex"] == 1 and s["maximum_vertex"] == 0 and s["minimum"] == -1 and s["maximum"] == 1
    assert list(s["local_minima"]) == [1] and list(s["local_maxima"]) == [0]
    u = oracle.surface_statistics(v, t, v[:, 0] + 0.25)
    assert u["negative_area"] == pytest.approx(9 * r3 / 8, rel=tol)
    assert u["average"] == pytest.approx(0.25, rel=tol)
    assert u["negative_average"] == pytest.approx(-0.25, rel=tol)
    p = oracle.surface_statistics(v, t, v[:, 0] + 2.0)
    assert p["negative_area"] == 0 and math.isnan(p["negative_average"]) and p["negative_variance"] == 0
    assert p["balance"] == 0
    c = oracle.surface_statistics(v, t, np.full(6, 2.0))
    assert c["total_variance"] == 0 and math.isnan(c["balance"]) and c["nonpolar_area"] == 0
    assert len(c["local_minima"]) == 0 and len(c["local_maxima"]) == 0
    with pytest.raises(ValueError, match="finite"):
        bad = v[:, 0].copy()
        bad[2] = np.nan
        oracle.surface_statistics(v, t, bad)
    with pytest.raises(ValueError, match="polar threshold"):
        oracle.surface_statistics(v, t, v[:, 0], -1.0)
    with pytest.raises(ValueError, match="missing vertex"):
        bt = t.copy()
        bt[0, 1] = 6
        oracle.surface_statistics(v, bt, v[:, 0])
    with pytest.raises(ValueError, match="area must be positive"):
        oracle.surface_statistics(np.zeros_like(v), t, v[:, 0])


def test_extraction_errors():
    grid = cube_grid(2.0, 0.5)
    n = int(np.prod(grid["counts"]))
    samples = np.full(n, -1.0)
    samples[n // 2] = 1.0

    def f(p):
        return np.zeros(len(p)), np.zeros((len(p), 3))
    with pytest.raises(ValueError, match="do not match"):
        oracle.isosurface(grid, samples[:-1], 0.0, f)
    bad = samples.copy()
    bad[3] = np.inf
    with pytest.raises(ValueError, match="must be finite"):
        oracle.isosurface(grid, bad, 0.0, f)
    with pytest.raises(ValueError, match="isovalue must be finite"):
        oracle.isosurface(grid, samples, np.nan, f)
    bad = samples.copy()
    bad[-1] = 1.0
    with pytest.raises(ValueError, match="reaches the grid boundary"):
        oracle.isosurface(grid, bad, 0.0, f)
    v, t = oracle.isosurface(grid, np.full(n, -1.0), 0.0, f)
    assert v.shape == (0, 3) and t.shape == (0, 3)


def oracle_surface(name, mode):
    """The oracle's surface for a report: mesh, values at the vertices, stats."""
    mol = molecule(name)
    kind = mode.replace("-no-extrema", "")
    iso = MAPPED[kind][1] if kind in MAPPED else 1e-3
    v, t = oracle.density_isosurface(mol, iso, 0.25)
    prop = {"esp": K.ElectrostaticPotential, "alie": K.AverageLocalIonizationEnergy}.get(kind) or MAPPED[kind][0]
    values = mol.evaluate([(prop, 0, 0, 0)], points=v)[0][0][0]
    polar = 0.0 if kind in MAPPED else 10.0 / 627.5094740631
    stats = oracle.surface_statistics(v, t, values, polar)
    return {"vertices": v, "triangles": t, "values": values, "stats": stats}


@pytest.mark.parametrize("case", SURFACE_CASES, ids=[c[1] for c in SURFACE_CASES])
def test_oracle_surfaces_against_multiwfn(case):
    name, report, mode = case
    s = oracle_surface(name, mode)
    assert closed_euler(s["vertices"], s["triangles"]) == 2
    worst = check_surface(s, mode, read_report(report), report, oracle.mesh_volume(s["vertices"], s["triangles"]))
    assert max(w for w, _ in worst) <= 1.0
