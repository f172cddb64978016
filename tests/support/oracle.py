"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so). Test
infrastructure: used only as the checker of the HIP path."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_lib = None
_ref = None


class Prop(C.Structure):
    _fields_ = [("kind", C.c_int), ("component", C.c_int), ("channel", C.c_int), ("orbital_index", C.c_int)]


class Options(C.Structure):
    _fields_ = [("block_size", C.c_int), ("backend", C.c_int), ("screening_threshold", C.c_double),
                ("fukui_delta", C.c_double), ("potential_tolerance", C.c_double)]


SCALAR_FN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                        C.c_void_p)
STATS = ["area", "positive_area", "negative_area", "nonpolar_area", "minimum", "maximum", "minimum_vertex",
         "maximum_vertex", "average", "positive_average", "negative_average", "positive_variance",
         "negative_variance", "total_variance", "balance", "charge_separation", "polarity_index"]


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"])
        L = C.CDLL(path)
        L.orc_basis_create.restype = C.c_void_p
        L.orc_molecule_create.restype = C.c_void_p
        L.orc_basis_nao.argtypes = [C.c_void_p]
        L.orc_basis_nblocks.argtypes = [C.c_void_p]
        L.orc_basis_blocks.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.orc_ao_eval.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_overlap.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_molecule_eligible.argtypes = [C.c_void_p]
        L.orc_evaluate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                   C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        L.orc_masks.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                C.c_int, C.c_int, C.c_double, C.c_void_p]
        L.orc_boys.argtypes = [C.c_double, C.c_int, C.c_void_p]
        L.orc_esp.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.orc_esp_pair_count.argtypes = [C.c_void_p]
        L.orc_esp_pair_count.restype = C.c_int64
        L.orc_esp_create.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.orc_esp_create.restype = C.c_void_p
        L.orc_esp_destroy.argtypes = [C.c_void_p]
        L.orc_far_plan.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_char_p, C.c_int]
        L.orc_far_plan.restype = C.c_void_p
        L.orc_far_plan_destroy.argtypes = [C.c_void_p]
        L.orc_far_plan_info.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_far_plan_cluster.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p]
        L.orc_far_plan_levels.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_far_cluster.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_char_p, C.c_int]
        L.orc_grid_points.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]
        L.orc_pure_transform.argtypes = [C.c_int, C.c_void_p]
        L.orc_molecule_destroy.argtypes = [C.c_void_p]
        L.orc_molecule_set_energies.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_basis_destroy.argtypes = [C.c_void_p]
        L.orc_isosurface.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, SCALAR_FN,
                                     C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        L.orc_isosurface.restype = C.c_void_p
        L.orc_density_isosurface.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_char_p, C.c_int]
        L.orc_density_isosurface.restype = C.c_void_p
        L.orc_mesh_sizes.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_mesh_data.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_mesh_free.argtypes = [C.c_void_p]
        L.orc_surface_statistics.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_double,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        L.orc_write_field_cube.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int,
                                           C.c_int64, C.c_char_p, C.c_int]
        _lib = L
    return _lib


def ref_boys_lib():
    """The reference's own Boys function, if oracle/_ref was built."""
    global _ref
    if _ref is None:
        path = os.path.join(ORACLE_DIR, "_ref", "libref_boys.so")
        if not os.path.exists(path):
            return None
        _ref = C.CDLL(path)
        _ref.ref_boys.argtypes = [C.c_double, C.c_int, C.c_void_p]
    return _ref


def boys(T, m):
    out = np.zeros(m + 1)
    assert lib().orc_boys(float(T), int(m), _ptr(out)) == 0
    return out


def pure_transform(l):
    n = 2 * l + 1
    nc = (l + 1) * (l + 2) // 2
    out = np.zeros((n, nc))
    assert lib().orc_pure_transform(l, _ptr(out)) == 0
    return out


class Basis:
    def __init__(self, d):
        err = C.create_string_buffer(512)
        self._keep = [np.ascontiguousarray(x) for x in (d.shell_types, d.nprim, d.shell_atom)]
        self._keepf = [np.ascontiguousarray(x, dtype=np.float64) for x in (d.shell_xyz, d.exponents, d.coefficients)]
        pc = None if d.p_coefficients is None else np.ascontiguousarray(d.p_coefficients, dtype=np.float64)
        self._pc = pc
        st, npr, sa = [np.ascontiguousarray(x, dtype=np.int32) for x in (d.shell_types, d.nprim, d.shell_atom)]
        self._ints = (st, npr, sa)
        self.h = lib().orc_basis_create(len(st), _ptr(st, C.c_int), _ptr(npr, C.c_int), _ptr(sa, C.c_int),
                                        _ptr(self._keepf[0]), _ptr(self._keepf[1]), _ptr(self._keepf[2]),
                                        _ptr(pc), err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.nao = lib().orc_basis_nao(self.h)

    def blocks(self):
        nb = lib().orc_basis_nblocks(self.h)
        l, off, cnt = (np.zeros(nb, np.int32) for _ in range(3))
        r2 = np.zeros(nb)
        lib().orc_basis_blocks(self.h, _ptr(l, C.c_int), _ptr(off, C.c_int), _ptr(cnt, C.c_int), _ptr(r2))
        return l, off, cnt, r2

    def ao(self, points, order=0):
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        n = pts.shape[0]
        panels = {0: 1, 1: 4, 2: 10}[order]
        out = np.zeros((panels * n, self.nao))  # column-major nao x (panels*n)
        assert lib().orc_ao_eval(self.h, n, _ptr(pts), order, _ptr(out)) == 0
        return out.T.copy()

    def overlap(self):
        out = np.zeros((self.nao, self.nao))
        lib().orc_overlap(self.h, _ptr(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_basis_destroy(self.h)


class Molecule:
    def __init__(self, d):
        self.data = d
        self.basis = Basis(d)
        n = self.basis.nao
        err = C.create_string_buffer(512)
        self._arrs = []

        def keep(a, t=np.float64):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=t)
            self._arrs.append(a)
            return a
        z = keep(d.atomic_numbers, np.int32)
        ch = keep(d.charges)
        xyz = keep(d.xyz)
        pt = keep(d.p_total)
        ps = keep(d.p_spin)
        ca = keep(d.c_alpha)
        cb = keep(d.c_beta)
        nmo_a = 0 if ca is None else ca.size // n
        nmo_b = 0 if cb is None else cb.size // n
        self.h = lib().orc_molecule_create(C.c_void_p(self.basis.h), len(z), _ptr(z, C.c_int), _ptr(ch), _ptr(xyz),
                                           _ptr(pt), _ptr(ps), nmo_a, _ptr(ca), nmo_b, _ptr(cb),
                                           d.nalpha, d.nbeta, err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        ea = keep(getattr(d, "e_alpha", None))
        eb = keep(getattr(d, "e_beta", None))
        lib().orc_molecule_set_energies(self.h, 0 if ea is None else ea.size, _ptr(ea), 0 if eb is None else eb.size,
                                        _ptr(eb))
        self.eligible = bool(lib().orc_molecule_eligible(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_molecule_destroy(self.h)

    def evaluate(self, props, grid=None, points=None, block_size=128, backend=0, threshold=1e-12, fukui_delta=0.1,
                 potential_tolerance=1e-8):
        """props: list of (kind, component, channel, orbital_index).
        Returns (list of arrays channels x n, backend_used)."""
        from kumo_hip.api import channels
        if grid is not None:
            origin = np.ascontiguousarray(grid["origin"], dtype=np.float64)
            steps = np.ascontiguousarray(np.asarray(grid["steps"], dtype=np.float64).T)  # column-major
            counts = np.ascontiguousarray(grid["counts"], dtype=np.int32)
            n = int(np.prod(counts.astype(np.int64)))
            pts = None
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            n = pts.shape[0]
            origin = steps = counts = None
        P = (Prop * len(props))(*[Prop(*p) for p in props])
        outs = [np.zeros((channels(p[0]), n)) for p in props]
        ptrs = []
        for o in outs:
            for c in range(o.shape[0]):
                ptrs.append(o[c].ctypes.data)
        arr = (C.c_void_p * max(1, len(ptrs)))(*ptrs)
        used = C.c_int(0)
        err = C.create_string_buffer(512)
        opt = Options(block_size, backend, threshold, fukui_delta, potential_tolerance)
        rc = lib().orc_evaluate(self.h, _ptr(origin), _ptr(steps), _ptr(counts, C.c_int), n, _ptr(pts), P,
                                len(props), C.byref(opt), arr, C.byref(used), err, 512)
        if rc == 1:
            raise ValueError(err.value.decode())
        if rc == 2:
            raise IndexError(err.value.decode())
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return outs, used.value

    def masks(self, block_size, ao_order, grid=None, points=None, threshold=1e-12):
        if grid is not None:
            origin = np.ascontiguousarray(grid["origin"], dtype=np.float64)
            steps = np.ascontiguousarray(np.asarray(grid["steps"], dtype=np.float64).T)
            counts = np.ascontiguousarray(grid["counts"], dtype=np.int32)
            n = int(np.prod(counts.astype(np.int64)))
            pts = None
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            n = pts.shape[0]
            origin = steps = counts = None
        nb = (n + block_size - 1) // block_size
        out = np.zeros((nb, self.basis.nao), dtype=np.uint8)
        lib().orc_masks(self.h, _ptr(origin), _ptr(steps), _ptr(counts, C.c_int), n, _ptr(pts), block_size,
                        ao_order, threshold, _ptr(out, C.c_uint8))
        return out

    def esp(self, points):
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(pts.shape[0])
        assert lib().orc_esp(self.h, pts.shape[0], _ptr(pts), _ptr(out)) == 0
        return out

    def esp_pair_count(self):
        return lib().orc_esp_pair_count(self.h)

    def potential(self):
        return Potential(self)


class Potential:
    """The oracle's electrostatic potential of a molecule with its far-field
    plans (src/electrostatic_potential.cpp:363-844)."""

    def __init__(self, mol):
        self.mol = mol  # keeps the molecule alive
        err = C.create_string_buffer(512)
        self.h = lib().orc_esp_create(mol.h, err, 512)
        if not self.h:
            raise ValueError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_esp_destroy(self.h)

    def plan(self, radius, tolerance):
        return FarPlan(self, radius, tolerance)

    def cluster(self, points, centre, plan):
        """evaluate_cluster: (values, (clusters translated, pairs expanded, pairs exact))."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        c = np.ascontiguousarray(centre, dtype=np.float64)
        out = np.zeros(pts.shape[0])
        stats = np.zeros(3, dtype=np.int64)
        err = C.create_string_buffer(512)
        if lib().orc_far_cluster(self.h, plan.h, pts.shape[0], _ptr(pts), _ptr(c), _ptr(out), _ptr(stats, C.c_int64),
                                 err, 512):
            raise ValueError(err.value.decode())
        return out, tuple(int(x) for x in stats)


class FarPlan:
    def __init__(self, pot, radius, tolerance):
        self.pot = pot
        err = C.create_string_buffer(512)
        self.h = lib().orc_far_plan(pot.h, radius, tolerance, err, 512)
        if not self.h:
            raise ValueError(err.value.decode())
        sizes = np.zeros(6, dtype=np.int64)
        vals = np.zeros(2)
        lib().orc_far_plan_info(self.h, _ptr(sizes, C.c_int64), _ptr(vals))
        self.npairs, self.finite_levels, self.nclusters, self.moment_count = (int(x) for x in sizes[:4])
        self.pair_tolerance, self.radius = float(vals[0]), float(vals[1])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_far_plan_destroy(self.h)

    def cluster_info(self, c):
        centre = np.zeros(3)
        radius, mind = C.c_double(0), C.c_double(0)
        npairs = C.c_int64(0)
        mom = np.zeros(self.moment_count)
        assert lib().orc_far_plan_cluster(self.h, c, _ptr(centre), C.byref(radius), C.byref(mind), C.byref(npairs),
                                          _ptr(mom)) == 0
        return {"centre": centre, "radius": radius.value, "min_distance": mind.value, "npairs": npairs.value,
                "moments": mom}

    def levels(self, i):
        out = np.zeros(33, dtype=np.float32)
        assert lib().orc_far_plan_levels(self.h, i, _ptr(out, C.c_float)) == 0
        return out


def grid_points(grid, start=0, n=None):
    origin = np.ascontiguousarray(grid["origin"], dtype=np.float64)
    steps = np.ascontiguousarray(np.asarray(grid["steps"], dtype=np.float64).T)
    counts = np.ascontiguousarray(grid["counts"], dtype=np.int32)
    total = int(np.prod(counts.astype(np.int64)))
    n = total - start if n is None else n
    out = np.zeros((n, 3))
    lib().orc_grid_points(_ptr(origin), _ptr(steps), _ptr(counts, C.c_int), start, n, _ptr(out))
    return out


def write_cube(path, grid, atoms, values, comment=None, orbital_index=None):
    """Field::write_cube / write_orbital_cube restated (src/cube.cpp:70-111,
    253-262). atoms: (atomic numbers, nuclear charges, coordinates natoms x 3).
    Returns (status, message): 0, 1 (invalid argument) or 3 (file error)."""
    orbital = orbital_index is not None
    if comment is None:
        comment = "Molecular orbital amplitude" if orbital else "Electron density"
    origin = np.ascontiguousarray(grid["origin"], dtype=np.float64)
    steps = np.ascontiguousarray(np.asarray(grid["steps"], dtype=np.float64).T)
    counts = np.ascontiguousarray(grid["counts"], dtype=np.int32)
    z = np.ascontiguousarray(atoms[0], dtype=np.int32).ravel()
    q = np.ascontiguousarray(atoms[1], dtype=np.float64).ravel()
    xyz = np.ascontiguousarray(atoms[2], dtype=np.float64).reshape(-1)
    v = np.ascontiguousarray(values, dtype=np.float64).ravel()
    err = C.create_string_buffer(256)
    rc = lib().orc_write_field_cube(os.fsencode(path), comment.encode(), _ptr(origin), _ptr(steps),
                                    _ptr(counts, C.c_int), z.size, _ptr(z, C.c_int), _ptr(q), _ptr(xyz), _ptr(v),
                                    v.size, 1 if orbital else 0, int(orbital_index or 0), err, 256)
    return rc, err.value.decode()


def _raise(status, msg):
    raise (ValueError if status == 1 else RuntimeError)(msg)


def _mesh(h):
    nv, nt = C.c_int64(), C.c_int64()
    lib().orc_mesh_sizes(h, C.byref(nv), C.byref(nt))
    v = np.zeros((nv.value, 3))
    t = np.zeros((nt.value, 3), dtype=np.int64)
    lib().orc_mesh_data(h, _ptr(v), _ptr(t, C.c_int64))
    lib().orc_mesh_free(h)
    return v, t


def isosurface(grid, samples, isovalue, fn):
    """extract_isosurface restated (src/surface.cpp:150-331). fn(points n x 3)
    -> (values n, gradients n x 3). Returns (vertices, triangles)."""
    origin = np.ascontiguousarray(grid["origin"], dtype=np.float64)
    steps = np.ascontiguousarray(np.asarray(grid["steps"], dtype=np.float64).T)
    counts = np.ascontiguousarray(grid["counts"], dtype=np.int32)
    f = np.ascontiguousarray(samples, dtype=np.float64).ravel()
    if f.size != int(np.prod(counts.astype(np.int64))):
        raise ValueError("isosurface samples do not match the grid")

    def call(pts, n, values, grads, user):
        p = np.ctypeslib.as_array(pts, shape=(n, 3))
        v, g = fn(p)
        np.ctypeslib.as_array(values, shape=(n,))[:] = v
        np.ctypeslib.as_array(grads, shape=(n, 3))[:] = g
    cb = SCALAR_FN(call)
    status = C.c_int(0)
    err = C.create_string_buffer(256)
    h = lib().orc_isosurface(_ptr(origin), _ptr(steps), _ptr(counts, C.c_int), _ptr(f), isovalue, cb, None,
                             C.byref(status), err, 256)
    if not h:
        _raise(status.value, err.value.decode())
    return _mesh(h)


def density_isosurface(mol, isovalue=1e-3, spacing=0.25):
    status = C.c_int(0)
    err = C.create_string_buffer(256)
    h = lib().orc_density_isosurface(mol.h, isovalue, spacing, C.byref(status), err, 256)
    if not h:
        _raise(status.value, err.value.decode())
    return _mesh(h)


def surface_statistics(vertices, triangles, values, polar_threshold=10.0 / 627.5094740631):
    v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, dtype=np.int64).reshape(-1, 3)
    w = np.ascontiguousarray(values, dtype=np.float64).ravel()
    if w.size != v.shape[0]:
        raise ValueError("surface values do not match the mesh vertices")
    st = np.zeros(17)
    counts = np.zeros(2, dtype=np.int64)
    err = C.create_string_buffer(256)
    rc = lib().orc_surface_statistics(_ptr(v), v.shape[0], _ptr(t, C.c_int64), t.shape[0], _ptr(w), polar_threshold,
                                      _ptr(st), _ptr(counts, C.c_int64), None, None, err, 256)
    if rc:
        _raise(rc, err.value.decode())
    mn = np.zeros(counts[0], dtype=np.int64)
    mx = np.zeros(counts[1], dtype=np.int64)
    lib().orc_surface_statistics(_ptr(v), v.shape[0], _ptr(t, C.c_int64), t.shape[0], _ptr(w), polar_threshold,
                                 _ptr(st), _ptr(counts, C.c_int64), _ptr(mn, C.c_int64), _ptr(mx, C.c_int64), err, 256)
    d = dict(zip(STATS, st.tolist()))
    d["minimum_vertex"], d["maximum_vertex"] = int(d["minimum_vertex"]), int(d["maximum_vertex"])
    d["local_minima"], d["local_maxima"] = mn, mx
    return d


def mesh_area(v, t):
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    return float(np.sum(0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)))


def mesh_volume(v, t):
    c0 = v.mean(axis=0)
    a, b, c = v[t[:, 0]] - c0, v[t[:, 1]] - c0, v[t[:, 2]] - c0
    return float(np.sum(np.einsum("ij,ij->i", a, np.cross(b, c))) / 6.0)
