"""ctypes binding of the C-ABI in include/kumo_hip.h (libkumo_hip.so, built
in-tree by chem-kumo_amd/Makefile). The library is the product: there is no
CPU fallback, and every entry point raises if the library is missing."""
import ctypes as C
import os

import numpy as np

from .api import Backend, Precision, PropertyKind, channels

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libkumo_hip.so")
_lib = None

STATUS = {1: ValueError, 2: IndexError, 3: RuntimeError, 4: RuntimeError, 5: RuntimeError}
HARTREE_IN_KCAL = 627.5094740631  # include/chem/kumo/surface.hpp:17


class _Property(C.Structure):
    _fields_ = [("kind", C.c_int32), ("component", C.c_int32), ("channel", C.c_int32), ("orbital_index", C.c_int32)]


class _Grid(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("steps", C.c_double * 9), ("counts", C.c_int32 * 3)]


PROGRESS = C.CFUNCTYPE(C.c_int, C.c_size_t, C.c_size_t, C.c_void_p)


class _Options(C.Structure):
    _fields_ = [("threads", C.c_int32), ("block_size", C.c_int32), ("backend", C.c_int32), ("precision", C.c_int32),
                ("screening_threshold", C.c_double), ("potential_tolerance", C.c_double), ("fukui_delta", C.c_double),
                ("device_outputs", C.c_int32), ("tile_points", C.c_int32), ("progress", PROGRESS),
                ("progress_user", C.c_void_p), ("esp_far_field", C.c_int32)]


class _Desc(C.Structure):
    _fields_ = [("natoms", C.c_int32), ("atomic_numbers", C.c_void_p), ("nuclear_charges", C.c_void_p),
                ("coordinates", C.c_void_p), ("nshells", C.c_int32), ("shell_types", C.c_void_p),
                ("primitives_per_shell", C.c_void_p), ("shell_to_atom", C.c_void_p),
                ("shell_coordinates", C.c_void_p), ("exponents", C.c_void_p), ("coefficients", C.c_void_p),
                ("p_coefficients", C.c_void_p), ("nbasis", C.c_int32), ("p_total", C.c_void_p),
                ("p_spin", C.c_void_p), ("nalpha", C.c_int32), ("nbeta", C.c_int32), ("nmo_alpha", C.c_int32),
                ("nmo_beta", C.c_int32), ("c_alpha", C.c_void_p), ("c_beta", C.c_void_p), ("e_alpha", C.c_void_p),
                ("e_beta", C.c_void_p)]


class _Info(C.Structure):
    _fields_ = [("natoms", C.c_int32), ("nbasis", C.c_int32), ("nblocks", C.c_int32), ("max_l", C.c_int32),
                ("nalpha", C.c_int32), ("nbeta", C.c_int32), ("nmo_alpha", C.c_int32), ("nmo_beta", C.c_int32),
                ("unrestricted", C.c_int32), ("ao_mo_eligible", C.c_int32), ("total_relative_error", C.c_double),
                ("spin_relative_error", C.c_double), ("device_bytes", C.c_int64)]


class _SurfaceStats(C.Structure):
    _fields_ = [("area", C.c_double), ("positive_area", C.c_double), ("negative_area", C.c_double),
                ("nonpolar_area", C.c_double), ("minimum", C.c_double), ("maximum", C.c_double),
                ("minimum_vertex", C.c_int64), ("maximum_vertex", C.c_int64), ("average", C.c_double),
                ("positive_average", C.c_double), ("negative_average", C.c_double),
                ("positive_variance", C.c_double), ("negative_variance", C.c_double),
                ("total_variance", C.c_double), ("balance", C.c_double), ("charge_separation", C.c_double),
                ("polarity_index", C.c_double), ("local_minima", C.c_int64), ("local_maxima", C.c_int64)]


class _ServerOptions(C.Structure):
    _fields_ = [("workers", C.c_int32), ("reserved", C.c_int32), ("hbm_budget", C.c_int64),
                ("chunk_points", C.c_int64)]


class _RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("backend_used", C.c_int32), ("queue_ms", C.c_double),
                ("run_ms", C.c_double), ("latency_ms", C.c_double), ("error", C.c_char * 256)]


EXPORTS = ["kumo_hip_init", "kumo_hip_release", "kumo_hip_last_error", "kumo_hip_default_options",
           "kumo_hip_molecule_from_fchk", "kumo_hip_molecule_create", "kumo_hip_molecule_info_get",
           "kumo_hip_molecule_release", "kumo_hip_molecule_blocks", "kumo_hip_molecule_atoms",
           "kumo_hip_evaluate_grid", "kumo_hip_evaluate_points", "kumo_hip_evaluate_points_noalloc", "kumo_hip_ao_eval", "kumo_hip_screening_masks",
           "kumo_hip_boys", "kumo_hip_boys_esp", "kumo_hip_last_kernel_stats", "kumo_hip_last_rescue_stats", "kumo_hip_last_field_stats", "kumo_hip_batch_create",
           "kumo_hip_batch_run", "kumo_hip_batch_last_ms", "kumo_hip_batch_kernel_stats", "kumo_hip_batch_release", "kumo_hip_server_create",
           "kumo_hip_server_submit", "kumo_hip_server_wait", "kumo_hip_server_query", "kumo_hip_server_release",
           "kumo_hip_write_cube", "kumo_hip_write_orbital_cube", "kumo_hip_format_scientific",
           "kumo_hip_last_cube_stats", "kumo_hip_density_isosurface", "kumo_hip_electrostatic_potential_surface",
           "kumo_hip_local_ionization_energy_surface", "kumo_hip_map_on_density_isosurface",
           "kumo_hip_average_local_ionization_energy", "kumo_hip_surface_sizes", "kumo_hip_surface_mesh",
           "kumo_hip_surface_field", "kumo_hip_surface_release", "kumo_hip_surface_statistics",
           "kumo_hip_write_ply"]


def lib():
    """Loads libkumo_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise RuntimeError(f"{LIBRARY} is missing: build it with `make -C chem-kumo_amd` "
                               "(the HIP backend has no CPU fallback)")
        # PyTorch-ROCm ships its own HIP runtime under the soname of the
        # system one (libamdhip64.so.7); the first loaded serves the whole
        # process. Loading torch first lets device tensors and this library
        # share one runtime in any import order.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIBRARY)
        vp = C.c_void_p
        L.kumo_hip_init.argtypes = [C.c_int, C.POINTER(vp)]
        L.kumo_hip_release.argtypes = [vp]
        L.kumo_hip_last_error.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.kumo_hip_default_options.argtypes = [C.POINTER(_Options)]
        L.kumo_hip_default_options.restype = None
        L.kumo_hip_molecule_from_fchk.argtypes = [vp, C.c_char_p, C.POINTER(vp)]
        L.kumo_hip_molecule_create.argtypes = [vp, C.POINTER(_Desc), C.POINTER(vp)]
        L.kumo_hip_molecule_info_get.argtypes = [vp, C.POINTER(_Info)]
        L.kumo_hip_molecule_release.argtypes = [vp]
        L.kumo_hip_molecule_blocks.argtypes = [vp, vp, vp, vp, vp]
        L.kumo_hip_molecule_atoms.argtypes = [vp, vp, vp, vp]
        L.kumo_hip_evaluate_grid.argtypes = [vp, vp, C.POINTER(_Grid), vp, C.c_size_t, C.POINTER(_Options), vp, vp]
        L.kumo_hip_evaluate_points.argtypes = [vp, vp, vp, C.c_size_t, vp, C.c_size_t, C.POINTER(_Options), vp, vp]
        L.kumo_hip_evaluate_points_noalloc.argtypes = [vp, vp, vp, C.c_size_t, vp, C.c_size_t, C.POINTER(_Options), vp,
                                                       vp, vp]
        L.kumo_hip_ao_eval.argtypes = [vp, vp, vp, C.c_size_t, C.c_int32, vp]
        L.kumo_hip_screening_masks.argtypes = [vp, vp, vp, vp, C.c_size_t, C.c_int32, C.c_int32, C.c_double, vp]
        L.kumo_hip_boys.argtypes = [vp, vp, C.c_size_t, C.c_int32, vp]
        L.kumo_hip_boys_esp.argtypes = [vp, vp, C.c_size_t, C.c_int32, vp]
        L.kumo_hip_last_kernel_stats.argtypes = [vp, vp, vp, vp, vp]
        L.kumo_hip_last_rescue_stats.argtypes = [vp, vp, vp, vp]
        L.kumo_hip_last_field_stats.argtypes = [vp, vp, vp, vp]
        L.kumo_hip_batch_create.argtypes = [vp, C.c_size_t, vp, vp, vp, C.c_size_t, C.POINTER(_Options), C.c_int32,
                                            vp, vp]
        L.kumo_hip_batch_run.argtypes = [vp, C.c_int32]
        L.kumo_hip_batch_last_ms.argtypes = [vp, vp]
        L.kumo_hip_batch_kernel_stats.argtypes = [vp] + [vp] * 8
        L.kumo_hip_batch_release.argtypes = [vp]
        L.kumo_hip_server_create.argtypes = [vp, C.POINTER(_ServerOptions), vp]
        L.kumo_hip_server_submit.argtypes = [vp, vp, vp, vp, C.c_size_t, vp, C.c_size_t, C.POINTER(_Options), vp, vp]
        L.kumo_hip_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(_RequestInfo)]
        L.kumo_hip_server_query.argtypes = [vp, C.c_uint64, vp]
        L.kumo_hip_server_release.argtypes = [vp]
        L.kumo_hip_write_cube.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(_Grid), C.c_int32, vp, vp, vp, vp,
                                          C.c_int32]
        L.kumo_hip_write_orbital_cube.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(_Grid), C.c_int32, vp, vp,
                                                  vp, vp, C.c_int32, C.c_size_t]
        L.kumo_hip_format_scientific.argtypes = [vp, vp, C.c_size_t, C.c_int32, C.c_int32, vp]
        L.kumo_hip_last_cube_stats.argtypes = [vp, vp, vp, vp, vp, vp]
        L.kumo_hip_density_isosurface.argtypes = [vp, vp, C.c_double, C.c_double, vp]
        L.kumo_hip_electrostatic_potential_surface.argtypes = [vp, vp, C.c_double, C.c_double, C.c_double, vp]
        L.kumo_hip_local_ionization_energy_surface.argtypes = [vp, vp, C.c_double, C.c_double, vp]
        L.kumo_hip_map_on_density_isosurface.argtypes = [vp, vp, vp, C.c_size_t, C.c_double, C.c_double,
                                                         C.POINTER(_Options), vp]
        L.kumo_hip_average_local_ionization_energy.argtypes = [vp, vp, vp, C.c_size_t, vp]
        L.kumo_hip_surface_sizes.argtypes = [vp, vp, vp, vp, vp]
        L.kumo_hip_surface_mesh.argtypes = [vp, vp, vp]
        L.kumo_hip_surface_field.argtypes = [vp, C.c_int32, vp, C.POINTER(_SurfaceStats), vp, vp]
        L.kumo_hip_surface_release.argtypes = [vp]
        L.kumo_hip_surface_statistics.argtypes = [vp, vp, C.c_int64, vp, C.c_int64, vp, C.c_double,
                                                  C.POINTER(_SurfaceStats), vp, vp, vp, vp]
        L.kumo_hip_write_ply.argtypes = [vp, C.c_char_p, vp, C.c_int64, vp, C.c_int64, C.c_int32, vp, vp,
                                         C.c_int32, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    """One device, one stream (kumo_hip_init)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib().kumo_hip_init(device, C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"kumo_hip_init failed with status {rc}")

    def check(self, rc):
        if rc != 0:
            buf = C.create_string_buffer(1024)
            lib().kumo_hip_last_error(self.h, buf, 1024)
            raise STATUS.get(rc, RuntimeError)(buf.value.decode())

    def close(self):
        if self.h:
            lib().kumo_hip_release(self.h)
            self.h = C.c_void_p()

    def boys(self, T, m):
        T = np.ascontiguousarray(T, dtype=np.float64).ravel()
        out = np.zeros((T.size, m + 1))
        self.check(lib().kumo_hip_boys(self.h, _p(T), T.size, m, _p(out)))
        return out

    def boys_esp(self, T, top):
        """The ESP pair loop's Boys function, orders 0..top (top <= 10)."""
        T = np.ascontiguousarray(T, dtype=np.float64).ravel()
        out = np.zeros((T.size, top + 1))
        self.check(lib().kumo_hip_boys_esp(self.h, _p(T), T.size, top, _p(out)))
        return out

    def kernel_stats(self):
        ms, fl, pfl = C.c_double(), C.c_double(), C.c_double()
        n = C.c_int64()
        lib().kumo_hip_last_kernel_stats(self.h, C.byref(ms), C.byref(n), C.byref(fl), C.byref(pfl))
        rms, rpts, rfl = C.c_double(), C.c_int64(), C.c_int64()
        lib().kumo_hip_last_rescue_stats(self.h, C.byref(rms), C.byref(rpts), C.byref(rfl))
        fms, fpts, fbytes = C.c_double(), C.c_int64(), C.c_int64()
        lib().kumo_hip_last_field_stats(self.h, C.byref(fms), C.byref(fpts), C.byref(fbytes))
        return {"kernel_ms": ms.value, "launches": n.value, "flops": fl.value, "padded_flops": pfl.value,
                "rescue_ms": rms.value, "rescue_points": rpts.value, "rescue_flagged": rfl.value, "fields_ms": fms.value,
                "fields_points": fpts.value, "fields_bytes": fbytes.value}


    def write_cube(self, path, grid, atoms, values, comment=None, orbital_index=None):
        """Field::write_cube / write_orbital_cube (src/cube.cpp:70-111, 253-262).
        atoms: (atomic numbers, nuclear charges or None, coordinates natoms x 3);
        values: a float64 array in Field order (host), or a device tensor /
        integer device pointer (kumo_hip_write_cube with values_on_device)."""
        z = np.ascontiguousarray(atoms[0], dtype=np.int32).ravel()
        q = None if atoms[1] is None else np.ascontiguousarray(atoms[1], dtype=np.float64).ravel()
        xyz = np.ascontiguousarray(atoms[2], dtype=np.float64).reshape(-1)
        if isinstance(values, np.ndarray) or isinstance(values, (list, tuple)):
            v = np.ascontiguousarray(values, dtype=np.float64).ravel()
            ptr, on_device, keep = _p(v), 0, v
        else:
            ptr = C.c_void_p(values if isinstance(values, int) else values.data_ptr())
            on_device, keep = 1, values
        g = grid_struct(grid)
        c = None if comment is None else comment.encode()
        p = os.fsencode(path)
        if orbital_index is None:
            rc = lib().kumo_hip_write_cube(self.h, p, c, C.byref(g), z.size, _p(z), _p(q), _p(xyz), ptr, on_device)
        else:
            rc = lib().kumo_hip_write_orbital_cube(self.h, p, c, C.byref(g), z.size, _p(z), _p(q), _p(xyz), ptr,
                                                   on_device, int(orbital_index))
        del keep
        self.check(rc)

    def format_scientific(self, values, precision, exact_only=False):
        """The GPU formatter: printf("%*.*e", precision + 8, precision, v) per value."""
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        width = precision + 8
        out = np.zeros(v.size * width, dtype=np.uint8)
        self.check(lib().kumo_hip_format_scientific(self.h, _p(v), v.size, precision, 1 if exact_only else 0,
                                                    _p(out)))
        text = out.tobytes().decode()
        return [text[i * width:(i + 1) * width] for i in range(v.size)]

    def cube_stats(self):
        t, k, c = C.c_double(), C.c_double(), C.c_double()
        n, b = C.c_int64(), C.c_int64()
        lib().kumo_hip_last_cube_stats(self.h, C.byref(t), C.byref(k), C.byref(c), C.byref(n), C.byref(b))
        return {"total_ms": t.value, "kernel_ms": k.value, "copy_ms": c.value, "values": n.value,
                "bytes": b.value}


    def surface_statistics(self, vertices, triangles, values, polar_threshold=10.0 / HARTREE_IN_KCAL):
        """surface_statistics (src/surface.cpp:418-529) with the mesh's area and
        volume (TriangleMesh::area / volume)."""
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, dtype=np.int64).reshape(-1, 3)
        x = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if x.size != v.shape[0]:
            raise ValueError("surface values do not match the mesh vertices")
        st = _SurfaceStats()
        area, volume = C.c_double(), C.c_double()
        self.check(lib().kumo_hip_surface_statistics(self.h, _p(v), v.shape[0], _p(t), t.shape[0], _p(x),
                                                     polar_threshold, C.byref(st), None, None, C.byref(area),
                                                     C.byref(volume)))
        mn = np.zeros(st.local_minima, dtype=np.int64)
        mx = np.zeros(st.local_maxima, dtype=np.int64)
        self.check(lib().kumo_hip_surface_statistics(self.h, _p(v), v.shape[0], _p(t), t.shape[0], _p(x),
                                                     polar_threshold, C.byref(st), _p(mn), _p(mx), None, None))
        return _stats_dict(st, mn, mx), area.value, volume.value

    def write_ply(self, path, vertices, triangles, fields=(), comments=()):
        """write_ply (src/surface.cpp:701-720); fields: (name, values) pairs."""
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, dtype=np.int64).reshape(-1, 3)
        vals = [np.ascontiguousarray(x, dtype=np.float64).ravel() for _, x in fields]
        for x in vals:
            if x.size != v.shape[0]:
                raise ValueError("PLY properties need one value per vertex")
        names = (C.c_char_p * max(1, len(vals)))(*[n.encode() for n, _ in fields])
        ptrs = (C.c_void_p * max(1, len(vals)))(*[x.ctypes.data for x in vals])
        cmts = (C.c_char_p * max(1, len(comments)))(*[c.encode() for c in comments])
        self.check(lib().kumo_hip_write_ply(self.h, os.fsencode(path), _p(v), v.shape[0], _p(t), t.shape[0],
                                            len(vals), names, ptrs, len(comments), cmts))


def _stats_dict(st, minima, maxima):
    d = {f: getattr(st, f) for f, _ in _SurfaceStats._fields_}
    d["local_minima"], d["local_maxima"] = minima, maxima
    return d


class Surface:
    """A density isosurface (kumo_hip_surface): vertices (n x 3, bohr),
    triangles (m x 3), and per field its vertex values and statistics."""

    def __init__(self, ctx, handle):
        nv, nt, nf, be = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        L = lib()
        L.kumo_hip_surface_sizes(handle, C.byref(nv), C.byref(nt), C.byref(nf), C.byref(be))
        self.vertices = np.zeros((nv.value, 3))
        self.triangles = np.zeros((nt.value, 3), dtype=np.int64)
        L.kumo_hip_surface_mesh(handle, _p(self.vertices), _p(self.triangles))
        self.values, self.statistics = [], []
        for f in range(nf.value):
            st = _SurfaceStats()
            L.kumo_hip_surface_field(handle, f, None, C.byref(st), None, None)
            vals = np.zeros(nv.value)
            mn = np.zeros(st.local_minima, dtype=np.int64)
            mx = np.zeros(st.local_maxima, dtype=np.int64)
            L.kumo_hip_surface_field(handle, f, _p(vals), C.byref(st), _p(mn), _p(mx))
            self.values.append(vals)
            self.statistics.append(_stats_dict(st, mn, mx))
        self.backend_used = be.value
        L.kumo_hip_surface_release(handle)


def grid_struct(grid):
    g = _Grid()
    o = np.asarray(grid["origin"], dtype=np.float64)
    st = np.asarray(grid["steps"], dtype=np.float64)  # columns are axis steps
    for i in range(3):
        g.origin[i] = o[i]
        g.counts[i] = int(grid["counts"][i])
    colmajor = st.T.ravel()
    for i in range(9):
        g.steps[i] = colmajor[i]
    return g


def options(block_size=128, backend=Backend.AUTO, precision=Precision.FP64, screening_threshold=1e-12,
            potential_tolerance=1e-8, fukui_delta=0.1, device_outputs=False, progress=None, threads=0,
            tile_points=0, esp_far_field=0):
    o = _Options()
    lib().kumo_hip_default_options(C.byref(o))
    o.threads = threads
    o.block_size = block_size
    o.backend = int(backend)
    o.precision = int(precision)
    o.screening_threshold = screening_threshold
    o.potential_tolerance = potential_tolerance
    o.fukui_delta = fukui_delta
    o.device_outputs = 1 if device_outputs else 0
    o.tile_points = tile_points
    o.esp_far_field = int(esp_far_field)
    if progress is not None:
        cb = PROGRESS(lambda d, t, u: int(bool(progress(d, t))))
        o.progress = cb
        o._cb = cb
    return o


class Molecule:
    """A molecule resident in HBM (kumo_hip_molecule_from_fchk / _create)."""

    def __init__(self, ctx, fchk_path=None, data=None):
        self.ctx = ctx
        self.h = C.c_void_p()
        if fchk_path is not None:
            ctx.check(lib().kumo_hip_molecule_from_fchk(ctx.h, os.fsencode(fchk_path), C.byref(self.h)))
        else:
            self._create(data)
        info = _Info()
        lib().kumo_hip_molecule_info_get(self.h, C.byref(info))
        self.info = {f: getattr(info, f) for f, _ in _Info._fields_}
        self.nao = info.nbasis

    def _create(self, d):
        keep = []

        def arr(x, t=np.float64):
            if x is None:
                return None
            a = np.ascontiguousarray(x, dtype=t)
            keep.append(a)
            return a.ctypes.data
        desc = _Desc()
        desc.natoms = len(d.atomic_numbers)
        desc.atomic_numbers = arr(d.atomic_numbers, np.int32)
        desc.nuclear_charges = arr(d.charges)
        desc.coordinates = arr(d.xyz)
        desc.nshells = len(d.shell_types)
        desc.shell_types = arr(d.shell_types, np.int32)
        desc.primitives_per_shell = arr(d.nprim, np.int32)
        desc.shell_to_atom = arr(d.shell_atom, np.int32)
        desc.shell_coordinates = arr(d.shell_xyz)
        desc.exponents = arr(d.exponents)
        desc.coefficients = arr(d.coefficients)
        desc.p_coefficients = arr(d.p_coefficients)
        desc.nbasis = d.nbasis
        desc.p_total = arr(d.p_total)
        desc.p_spin = arr(d.p_spin)
        desc.nalpha, desc.nbeta = d.nalpha, d.nbeta
        desc.c_alpha = arr(d.c_alpha)
        desc.c_beta = arr(d.c_beta)
        desc.nmo_alpha = 0 if d.c_alpha is None else len(d.c_alpha) // d.nbasis
        desc.nmo_beta = 0 if d.c_beta is None else len(d.c_beta) // d.nbasis
        desc.e_alpha = arr(d.e_alpha)
        desc.e_beta = arr(d.e_beta)
        self.ctx.check(lib().kumo_hip_molecule_create(self.ctx.h, C.byref(desc), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().kumo_hip_molecule_release(self.h)
            self.h = C.c_void_p()

    def _surface(self, call):
        h = C.c_void_p()
        self.ctx.check(call(C.byref(h)))
        return Surface(self.ctx, h)

    def density_isosurface(self, isovalue=1e-3, spacing=0.25):
        return self._surface(lambda out: lib().kumo_hip_density_isosurface(self.ctx.h, self.h, isovalue, spacing,
                                                                            out))

    def electrostatic_potential_surface(self, isovalue=1e-3, spacing=0.25, potential_tolerance=1e-8):
        return self._surface(lambda out: lib().kumo_hip_electrostatic_potential_surface(
            self.ctx.h, self.h, isovalue, spacing, potential_tolerance, out))

    def local_ionization_energy_surface(self, isovalue=1e-3, spacing=0.25):
        return self._surface(lambda out: lib().kumo_hip_local_ionization_energy_surface(
            self.ctx.h, self.h, isovalue, spacing, out))

    def map_on_density_isosurface(self, props, isovalue=1e-3, spacing=0.25, **opts):
        P = (_Property * max(1, len(props)))(*[_Property(*map(int, p)) for p in props])
        o = options(**opts)
        return self._surface(lambda out: lib().kumo_hip_map_on_density_isosurface(
            self.ctx.h, self.h, P, len(props), isovalue, spacing, C.byref(o), out))

    def average_local_ionization_energy(self, points):
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(pts.shape[0])
        self.ctx.check(lib().kumo_hip_average_local_ionization_energy(self.ctx.h, self.h, _p(pts), pts.shape[0],
                                                                      _p(out)))
        return out

    def atoms(self):
        """(atomic numbers, nuclear charges, coordinates natoms x 3), as a Field keeps them."""
        n = self.info["natoms"]
        z = np.zeros(n, dtype=np.int32)
        q = np.zeros(n)
        xyz = np.zeros((n, 3))
        self.ctx.check(lib().kumo_hip_molecule_atoms(self.h, _p(z), _p(q), _p(xyz)))
        return z, q, xyz

    def blocks(self):
        nb = self.info["nblocks"]
        l, off, cnt = (np.zeros(nb, np.int32) for _ in range(3))
        r2 = np.zeros(nb)
        lib().kumo_hip_molecule_blocks(self.h, _p(l), _p(off), _p(cnt), _p(r2))
        return l, off, cnt, r2

    def ao(self, points, order=0):
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        panels = {0: 1, 1: 4, 2: 10}[order]
        out = np.zeros((panels * pts.shape[0], self.nao))
        self.ctx.check(lib().kumo_hip_ao_eval(self.ctx.h, self.h, _p(pts), pts.shape[0], order, _p(out)))
        return out.T.copy()

    def masks(self, block_size, ao_order, grid=None, points=None, threshold=1e-12):
        if grid is not None:
            g = grid_struct(grid)
            n = int(np.prod(np.asarray(grid["counts"], dtype=np.int64)))
            nb = (n + block_size - 1) // block_size
            out = np.zeros((nb, self.nao), np.uint8)
            self.ctx.check(lib().kumo_hip_screening_masks(self.ctx.h, self.h, C.byref(g), None, 0, block_size,
                                                          ao_order, threshold, _p(out)))
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            nb = (pts.shape[0] + block_size - 1) // block_size
            out = np.zeros((nb, self.nao), np.uint8)
            self.ctx.check(lib().kumo_hip_screening_masks(self.ctx.h, self.h, None, _p(pts), pts.shape[0],
                                                          block_size, ao_order, threshold, _p(out)))
        return out

    def evaluate(self, props, grid=None, points=None, **opts):
        """props: sequence of (kind, component, channel, orbital_index). Returns
        (list of arrays channels x n, backend_used)."""
        P = (_Property * max(1, len(props)))(*[_Property(*map(int, p)) for p in props])
        o = options(**opts)
        if grid is not None:
            n = int(np.prod(np.asarray(grid["counts"], dtype=np.int64)))
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            n = pts.shape[0]
        outs = [np.zeros((channels(p[0]), n)) for p in props]
        ptrs = [o_[c].ctypes.data for o_ in outs for c in range(o_.shape[0])]
        arr = (C.c_void_p * max(1, len(ptrs)))(*ptrs)
        used = C.c_int32(-1)
        if grid is not None:
            g = grid_struct(grid)
            rc = lib().kumo_hip_evaluate_grid(self.ctx.h, self.h, C.byref(g), P, len(props), C.byref(o), arr,
                                              C.byref(used))
        else:
            rc = lib().kumo_hip_evaluate_points(self.ctx.h, self.h, _p(pts), n, P, len(props), C.byref(o), arr,
                                                C.byref(used))
        self.ctx.check(rc)
        return outs, used.value

    def evaluate_points_noalloc(self, props, points, npoints, outputs, stream=None, **opts):
        """kumo_hip_evaluate_points_noalloc: points (3 * npoints doubles) and
        outputs (one per component) are device pointers; the work is issued on
        `stream` (a raw hipStream_t, None: the context's stream) and not
        waited for. Returns the backend used."""
        P = (_Property * max(1, len(props)))(*[_Property(*map(int, p)) for p in props])
        o = options(device_outputs=True, **opts)
        ptrs = [int(x) for x in outputs]
        arr = (C.c_void_p * max(1, len(ptrs)))(*ptrs)
        used = C.c_int32(-1)
        self.ctx.check(lib().kumo_hip_evaluate_points_noalloc(self.ctx.h, self.h, C.c_void_p(int(points)), int(npoints),
                                                              P, len(props), C.byref(o), arr,
                                                              C.c_void_p(int(stream) if stream else 0),
                                                              C.byref(used)))
        return used.value

    def evaluate_device(self, props, grid, outputs, **opts):
        """Evaluation into device buffers (torch tensors or raw pointers);
        nothing is copied to the host."""
        P = (_Property * max(1, len(props)))(*[_Property(*map(int, p)) for p in props])
        o = options(device_outputs=True, **opts)
        ptrs = [int(x) for x in outputs]
        arr = (C.c_void_p * max(1, len(ptrs)))(*ptrs)
        used = C.c_int32(-1)
        g = grid_struct(grid)
        self.ctx.check(lib().kumo_hip_evaluate_grid(self.ctx.h, self.h, C.byref(g), P, len(props), C.byref(o), arr,
                                                    C.byref(used)))
        return used.value


class Batch:
    """Many (molecule, grid) items with one property list, evaluated together
    into device buffers (kumo_hip_batch_*): outputs[i] lists item i's channel
    pointers (torch tensors or raw device pointers)."""

    def __init__(self, ctx, molecules, grids, props, outputs, streams=4, **opts):
        self.ctx = ctx
        self._mols = list(molecules)
        n = len(self._mols)
        M = (C.c_void_p * max(1, n))(*[m.h.value for m in self._mols])
        G = (_Grid * max(1, n))(*[grid_struct(g) for g in grids])
        P = (_Property * max(1, len(props)))(*[_Property(*map(int, p)) for p in props])
        o = options(device_outputs=True, **opts)
        flat = [int(x) for item in outputs for x in item]
        arr = (C.c_void_p * max(1, len(flat)))(*flat)
        h = C.c_void_p()
        ctx.check(lib().kumo_hip_batch_create(ctx.h, n, M, G, P, len(props), C.byref(o), int(streams), arr,
                                              C.byref(h)))
        self.h = h

    def run(self, graph=True):
        self.ctx.check(lib().kumo_hip_batch_run(self.h, 1 if graph else 0))
        ms = C.c_double()
        lib().kumo_hip_batch_last_ms(self.h, C.byref(ms))
        return ms.value

    def timed_run(self, serial=False):
        """One run without a graph with every contraction, rescue and field
        launch bracketed by HIP events on its stream (serial: the streams
        chained, so no launches overlap); returns the run's device ms and the
        launch statistics (kumo_hip_batch_kernel_stats)."""
        self.ctx.check(lib().kumo_hip_batch_run(self.h, 3 if serial else 2))
        ms = C.c_double()
        lib().kumo_hip_batch_last_ms(self.h, C.byref(ms))
        km, fl, rms, fms = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        n, rp, fb, rf = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        self.ctx.check(lib().kumo_hip_batch_kernel_stats(self.h, C.byref(km), C.byref(n), C.byref(fl), C.byref(rms),
                                                         C.byref(rp), C.byref(fms), C.byref(fb), C.byref(rf)))
        return {"ms": ms.value, "kernel_ms": km.value, "launches": n.value, "flops": fl.value, "rescue_ms": rms.value,
                "rescue_points": rp.value, "rescue_flagged": rf.value, "fields_ms": fms.value,
                "fields_bytes": fb.value}

    def close(self):
        if self.h:
            lib().kumo_hip_batch_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Request:
    """A submitted server request; `outputs` fills in when it completes."""

    def __init__(self, server, rid, outputs, props, keep):
        self.server, self.id, self.outputs, self.props, self._keep = server, rid, outputs, props, keep
        self.info = None

    def wait(self):
        """Blocks until done; returns (results as evaluate() does, info dict)."""
        if self.info is None:
            inf = _RequestInfo()
            rc = lib().kumo_hip_server_wait(self.server.h, self.id, C.byref(inf))
            self.info = {"status": inf.status, "backend": inf.backend_used, "queue_ms": inf.queue_ms,
                         "run_ms": inf.run_ms, "latency_ms": inf.latency_ms, "error": inf.error.decode()}
            if rc != 0:
                raise RuntimeError(f"request failed ({rc}): {self.info['error']}")
        res, c = [], 0
        for p in self.props:
            k = channels(PropertyKind(int(p[0])))
            res.append(self.outputs[c:c + k])
            c += k
        return res, self.info


class Server:
    """Serving loop over resident molecules (kumo_hip_server_*)."""

    def __init__(self, ctx, workers=4, hbm_budget=0, chunk_points=0):
        self.ctx = ctx
        o = _ServerOptions(int(workers), 0, int(hbm_budget), int(chunk_points))
        h = C.c_void_p()
        ctx.check(lib().kumo_hip_server_create(ctx.h, C.byref(o), C.byref(h)))
        self.h = h

    def submit(self, mol, props, grid=None, points=None, **opts):
        P = (_Property * max(1, len(props)))(*[_Property(*map(int, p)) for p in props])
        o = options(**opts)
        if grid is not None:
            g = grid_struct(grid)
            n = int(np.prod(np.asarray(grid["counts"], dtype=np.int64)))
            gp, pp, npts = C.byref(g), None, 0
        else:
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            n = pts.shape[0]
            g, gp, pp, npts = None, None, _p(pts), n
        nch = sum(channels(PropertyKind(int(p[0]))) for p in props)
        out = np.zeros((nch, n))
        arr = (C.c_void_p * max(1, nch))(*[out[c].ctypes.data for c in range(nch)])
        rid = C.c_uint64()
        self.ctx.check(lib().kumo_hip_server_submit(self.h, mol.h, gp, pp, npts, P, len(props), C.byref(o), arr,
                                                    C.byref(rid)))
        return Request(self, rid.value, out, props, (g, P, arr, o))

    def close(self):
        if self.h:
            lib().kumo_hip_server_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
