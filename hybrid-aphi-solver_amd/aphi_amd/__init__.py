"""Python host mirror of the A-Phi drop-in path, over the C-ABI of include/aphi_gpu.h.

Names follow the reference's call surface (include/aphi_solver/assembly.hpp,
factorization.hpp): ``refill``, ``factorize_ldlt``, ``solve``, ``solve_symmetric``,
``backward_error``. Error behaviour follows the reference's conventions:
std::invalid_argument -> ValueError, std::logic_error -> LogicError,
std::runtime_error -> RuntimeError, a pivot at or below the floor -> the call
returns ``ok = False`` with its stats filled.

Device arrays are torch tensors (PyTorch is only the allocator here); every
number is computed by libaphi_amd.so. There is no CPU fallback.
"""
import ctypes as C
import sys
import weakref
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import (APHI_FP32, APHI_FP64, APHI_NATURAL, APHI_ORDER_AMD, APHI_ORDER_NATURAL, APHI_ORDER_RCM,
                    APHI_ROW_SCALED, APHI_SCALED_PHI)

__all__ = [
    "LogicError", "DeviceError", "CubeConfig", "Problem", "Context", "Mesh", "frequency_sweep",
    "APHI_FP32", "APHI_FP64", "APHI_NATURAL", "APHI_ROW_SCALED", "APHI_SCALED_PHI",
    "APHI_ORDER_AMD", "APHI_ORDER_NATURAL", "APHI_ORDER_RCM", "lib", "PrecisionPolicy",
    "recover_gauge_free_potential", "philox4x32_10", "random_normal_host", "random_normal", "random_complex_rhs",
    "random_uniform_host", "batched_cube_configs", "Batch",
]


from .sweep import PrecisionPolicy  # noqa: E402


class LogicError(RuntimeError):
    """std::logic_error of the reference: the symbolic and numeric passes disagree."""


class DeviceError(RuntimeError):
    """A HIP runtime failure."""


def lib():
    return _capi.load()


def _alive():
    """False while the interpreter shuts down (module globals may already be gone)."""
    try:
        return not sys.is_finalizing() and _capi._lib is not None
    except Exception:
        return False


def _destroy(fn_name, h):
    """Calls a destroy entry point unless the interpreter is shutting down."""
    try:
        if h and _alive():
            getattr(_capi._lib, fn_name)(h)
    except Exception:
        pass


def _own(obj, fn_name, h):
    """Gives ``obj`` the handle ``h`` and a finalizer that destroys it exactly once:
    on ``close()``, when ``obj`` is collected, or at interpreter exit (pending
    finalizers then run in reverse order of creation, so a mesh goes before the
    context it was made on)."""
    obj._h = h
    obj._fin = weakref.finalize(obj, _destroy, fn_name, C.c_void_p(h.value))


def _release(obj):
    fin = getattr(obj, "_fin", None)
    if fin is not None:
        fin()
    obj._h = None


def _check(status):
    if status == _capi.APHI_OK:
        return
    msg = lib().aphi_last_error().decode()
    if status == _capi.APHI_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == _capi.APHI_LOGIC_ERROR:
        raise LogicError(msg)
    if status in (_capi.APHI_DEVICE_ERROR, _capi.APHI_OUT_OF_MEMORY):
        raise DeviceError(msg)
    raise RuntimeError(msg)


def philox4x32_10(counter, key):
    """One Philox-4x32-10 block (four 32-bit words) of counter[4] and key[2]."""
    c = (C.c_uint32 * 4)(*[int(v) & 0xffffffff for v in counter])
    k = (C.c_uint32 * 2)(*[int(v) & 0xffffffff for v in key])
    o = (C.c_uint32 * 4)()
    _check(lib().aphi_philox4x32_10(c, k, o))
    return [int(v) for v in o]


def random_normal_host(seed, count, offset=0):
    """Normal numbers offset .. offset + count - 1 of stream ``seed`` (numpy float64)."""
    out = np.zeros(int(count), dtype=np.float64)
    _check(lib().aphi_random_normal_host(int(seed), int(offset), int(count), out.ctypes.data_as(_capi.DP)))
    return out


def random_uniform_host(seed, count, offset=0):
    """Uniform numbers in [0, 1), offset .. offset + count - 1 of stream ``seed``."""
    out = np.zeros(int(count), dtype=np.float64)
    _check(lib().aphi_random_uniform_host(int(seed), int(offset), int(count), out.ctypes.data_as(_capi.DP)))
    return out


def batched_cube_configs(seed, count):
    """The seeded problems of the batched workload: for problem i, uniforms
    8i .. 8i + 7 of stream ``seed`` give the cube side U[0.05, 0.2] m, hexes per
    edge U{4..12}, a conductor block of 1-4 hexes per side (at most n - 1, so air
    remains) spanning z at a uniform position, sigma log-uniform in [1e6, 6e7] S/m and a frequency log-uniform in
    [1 Hz, 1 MHz]. Returns (configs, frequencies)."""
    import math
    u = random_uniform_host(seed, 8 * int(count))
    configs, freqs = [], []
    for i in range(int(count)):
        v = u[8 * i: 8 * i + 8]
        n = 4 + min(int(v[1] * 9), 8)
        # at most n - 1 hexes per side, so that air remains (a body needs tets)
        bx, by = min(1 + min(int(v[2] * 4), 3), n - 1), min(1 + min(int(v[3] * 4), 3), n - 1)
        i0, j0 = min(int(v[4] * (n - bx + 1)), n - bx), min(int(v[5] * (n - by + 1)), n - by)
        sigma = math.exp(math.log(1e6) + (math.log(6e7) - math.log(1e6)) * v[6])
        configs.append(CubeConfig(n=n, side=0.05 + 0.15 * v[0], bar=(i0, i0 + bx, j0, j0 + by), sigma=sigma))
        freqs.append(math.exp(math.log(1e6) * v[7]))
    return configs, np.array(freqs)


def random_normal(ctx, seed, shape, offset=0):
    """The same numbers as random_normal_host, drawn on the context's device into a
    float64 tensor of ``shape``."""
    import torch
    t = torch.empty(shape, dtype=torch.float64, device=f"cuda:{ctx.device}")
    _check(lib().aphi_random_normal(ctx._h, int(seed), int(offset), t.numel(), C.c_void_p(t.data_ptr())))
    ctx.synchronize()
    return t


def random_complex_rhs(ctx, seed, nrhs, n, offset=0):
    """nrhs x n complex right-hand sides with iid N(0,1) real and imaginary parts:
    entry (r, i) takes normals offset + 2(r n + i) (real) and the next (imaginary)."""
    import torch
    return torch.view_as_complex(random_normal(ctx, seed, (nrhs, n, 2), offset))


def frequency_sweep(start, stop, points, logarithmic=True):
    """expand_frequency_sweep (src/input_file.cpp:19-50)."""
    out = np.zeros(points, dtype=np.float64)
    _check(lib().aphi_frequency_sweep(start, stop, points, int(logarithmic),
                                      out.ctypes.data_as(_capi.DP)))
    return out


@dataclass
class CubeConfig:
    """A Kuhn cube with a conductor bar spanning z and voltage ports on its ends."""
    n: int = 6
    side: float = 0.1
    bar: tuple = (2, 4, 2, 4)          # hex columns [i0, i1) x [j0, j1)
    sigma: float = 5.8e7
    eps_r: float = 1.0
    mu_r: float = 1.0
    v_top: complex = 1.0
    v_bottom: complex = 0.0
    i_top: complex = None               # a current (A) driving the z = side port instead of v_top
    ordering: int = APHI_ORDER_AMD
    relax_small: int = 0
    relax_fill: float = 0.0

    def spec(self):
        s = _capi.CubeSpec()
        s.n, s.side = self.n, self.side
        s.bar_i0, s.bar_i1, s.bar_j0, s.bar_j1 = self.bar
        s.sigma, s.eps_r, s.mu_r = self.sigma, self.eps_r, self.mu_r
        s.air_sigma, s.air_eps_r, s.air_mu_r = 0.0, 1.0, 1.0
        vt, vb = complex(self.v_top), complex(self.v_bottom)
        s.v_top_re, s.v_top_im, s.v_bottom_re, s.v_bottom_im = vt.real, vt.imag, vb.real, vb.imag
        s.ordering, s.relax_small, s.relax_fill = self.ordering, self.relax_small, self.relax_fill
        if self.i_top is not None:
            it = complex(self.i_top)
            s.top_current, s.i_top_re, s.i_top_im = 1, it.real, it.imag
        return s


class Problem:
    """Host symbolic problem: mesh, binding, gauge, DOF map, pattern, scatter map,
    ordering, elimination tree and supernodal plan (all integer, bit-exact)."""

    def __init__(self, config: CubeConfig):
        self.config = config
        h = C.c_void_p()
        spec = config.spec()
        _check(lib().aphi_problem_create_cube(C.byref(spec), C.byref(h)))
        _own(self, "aphi_problem_destroy", h)
        self._read_counts()

    @classmethod
    def cubes(cls, configs, threads=0):
        """Many cube problems built on up to ``threads`` host threads (0: one per
        hardware thread), each the same as Problem(config)."""
        configs = list(configs)
        specs = (_capi.CubeSpec * max(len(configs), 1))(*[c.spec() for c in configs])
        hs = (C.c_void_p * max(len(configs), 1))()
        _check(lib().aphi_problem_create_cubes(specs, len(configs), int(threads), hs))
        out = []
        for c, h in zip(configs, hs):
            p = cls.__new__(cls)
            p.config = c
            _own(p, "aphi_problem_destroy", C.c_void_p(h))
            p._read_counts()
            out.append(p)
        return out

    @classmethod
    def from_csr(cls, n, row_ptr, col_index, ordering=APHI_ORDER_AMD, max_factor_nnz=0, relax_small=0,
                 relax_fill=0.0):
        """A problem from a matrix pattern alone: SparseSymmetric::from_pattern
        (include/aphi_solver/sparse_symmetric.hpp:38-83) + analyze(pattern, ordering,
        max_factor_nnz) (include/aphi_solver/symbolic.hpp:104). Its device mesh takes
        values through ``Mesh.upload_system``; ``refill`` refuses it."""
        self = cls.__new__(cls)
        self.config = None
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col_index, dtype=np.int32)
        if rp.size != int(n) + 1:
            raise ValueError("SparseSymmetric::from_pattern: row_ptr must have rows + 1 entries")
        if rp[-1] != ci.size:
            raise ValueError("SparseSymmetric::from_pattern: row_ptr must run from 0 to col_index.size()")
        h = C.c_void_p()
        _check(lib().aphi_problem_create_csr(int(n), rp.ctypes.data_as(_capi.I32P), ci.ctypes.data_as(_capi.I32P),
                                             int(ordering), int(max_factor_nnz), int(relax_small), float(relax_fill),
                                             C.byref(h)))
        _own(self, "aphi_problem_destroy", h)
        self._read_counts()
        return self

    @classmethod
    def gauge_free(cls, cube, ordering=APHI_ORDER_AMD):
        """The gauge-free potential's P1 system on the cube problem's mesh
        (src/gauge_free_potential.cpp:40-164): vertex map, pattern, ordering and
        supernodal plan, and the per-tet term plan of its assembly."""
        self = cls.__new__(cls)
        self.config = None
        self.cube = cube
        h = C.c_void_p()
        _check(lib().aphi_problem_create_gauge_free(cube._h, int(ordering), C.byref(h)))
        _own(self, "aphi_problem_destroy", h)
        self._read_counts()
        return self

    @classmethod
    def from_bound(cls, arrays, ordering=APHI_ORDER_AMD, relax_small=0, relax_fill=0.0):
        """A problem from the reference host's bound objects, flattened
        (aphi_problem_create_bound): ``arrays`` maps the aphi_bound_desc field
        names to numpy arrays (the scatter_* arrays may be left out: the map is
        then built from the pattern)."""
        self = cls.__new__(cls)
        self.config = None
        dtypes = {_capi.I32P: np.int32, _capi.DP: np.float64, _capi.I64P: np.int64}
        keep, d = {}, _capi.BoundDesc()
        for f, t in _capi.BoundDesc._fields_:
            if t is C.c_int32:
                continue
            if f not in arrays:
                if f.startswith("scatter_"):
                    continue  # built from the pattern
                raise ValueError(f"from_bound: missing array {f}")
            a = np.ascontiguousarray(arrays[f], dtype=dtypes.get(t, np.int8))
            keep[f] = a
            setattr(d, f, a.ctypes.data_as(t))
        d.num_nodes = keep["coords"].size // 3
        d.num_tets = keep["tet_nodes"].size // 4
        d.num_bodies = keep["body_sigma"].size
        d.num_total = keep["row_ptr"].size - 1
        d.num_ports = keep["port_index"].size
        h = C.c_void_p()
        _check(lib().aphi_problem_create_bound(C.byref(d), int(ordering), int(relax_small), float(relax_fill),
                                               C.byref(h)))
        _own(self, "aphi_problem_destroy", h)
        self._read_counts()
        return self

    def bound_arrays(self):
        """This problem's bound objects flattened as aphi_problem_create_bound
        takes them (what a reference host would hand over for its mesh)."""
        nt = int(self.counts["tets"])
        td = self.ints("tet_dofs").reshape(nt, 16)
        bp = self.doubles("body_param").reshape(-1, 3)
        fixed = self.ints("port_is_fixed")
        amp = np.where(np.repeat(fixed, 2) == 1, self.doubles("port_value"), self.doubles("port_amplitude"))
        return dict(coords=self.doubles("coords"), tet_nodes=self.ints("tet_nodes"),
                    tet_edge_signs=self.ints("tet_edge_signs"), body_sigma=bp[:, 0], body_eps_r=bp[:, 1],
                    body_mu_r=bp[:, 2], body_ptr=self.ints("body_ptr"), body_tets=self.ints("body_tets"),
                    edge_index=td[:, :6], edge_coeff=self.doubles("tet_edge_coeff"), phi_index=td[:, 6:],
                    phi_coeff=self.doubles("tet_phi_coeff"), phi_fixed=self.doubles("tet_phi_fixed"),
                    port_index=self.ints("port_index"), port_current=1 - fixed, port_amplitude=amp,
                    row_ptr=self.ints("row_ptr"), col_index=self.ints("col_index"),
                    scatter_live=self.ints("scatter_live"), scatter_pos=self.ints("scatter_pos"),
                    scatter_offset=self.ints("scatter_offset"), scatter_slot=self.ints("scatter_slot"))

    def describe_gauge_free(self):
        d = _capi.GaugeFreeDesc()
        _check(lib().aphi_problem_describe_gauge_free(self._h, C.byref(d)))
        return d

    def _read_counts(self):
        c = _capi.ProblemCounts()
        _check(lib().aphi_problem_counts_get(self._h, C.byref(c)))
        self.counts = {k: getattr(c, k) for k, _ in _capi.ProblemCounts._fields_}

    def close(self):
        _release(self)

    @property
    def n(self):
        return int(self.counts["num_total"])

    def ints(self, name):
        cnt = C.c_int64()
        _check(lib().aphi_problem_get_int(self._h, name.encode(), None, 0, C.byref(cnt)))
        out = np.zeros(cnt.value, dtype=np.int64)
        _check(lib().aphi_problem_get_int(self._h, name.encode(), out.ctypes.data_as(_capi.I64P), cnt.value,
                                          C.byref(cnt)))
        return out

    def doubles(self, name):
        cnt = C.c_int64()
        _check(lib().aphi_problem_get_double(self._h, name.encode(), None, 0, C.byref(cnt)))
        out = np.zeros(cnt.value, dtype=np.float64)
        _check(lib().aphi_problem_get_double(self._h, name.encode(), out.ctypes.data_as(_capi.DP), cnt.value,
                                             C.byref(cnt)))
        return out

    def write_potential(self, path, x, omega, conditioning=APHI_ROW_SCALED, vtk=False):
        """potential_at_nodes + write_potential (or write_vtk without field data) of
        the solution x (numpy or device tensor, n complex); returns the byte count."""
        if hasattr(x, "cpu"):
            x = x.reshape(-1).cpu().numpy()
        xx = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
        if xx.size != self.n:
            raise ValueError("write_potential: x must hold n values")
        nb = C.c_int64()
        _check(lib().aphi_write_potential(self._h, str(path).encode(), xx.view(np.float64).ctypes.data_as(_capi.DP),
                                          float(omega), int(conditioning), 1 if vtk else 0, C.byref(nb)))
        return nb.value

    def describe(self):
        d = _capi.MeshDesc()
        _check(lib().aphi_problem_describe(self._h, C.byref(d)))
        return d

    def describe_fields(self):
        """compute_fields' node plan (aphi_problem_describe_fields)."""
        d = _capi.FieldsDesc()
        _check(lib().aphi_problem_describe_fields(self._h, C.byref(d)))
        return d

    def write_vtk_fields(self, path, x, omega, fields, conditioning=APHI_ROW_SCALED):
        """write_vtk with compute_fields' data (FieldSet::All, no run information) of
        the solution x and the FieldOutput ``fields`` (Mesh.compute_fields);
        returns the byte count."""
        if hasattr(x, "cpu"):
            x = x.reshape(-1).cpu().numpy()
        xx = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
        if xx.size != self.n:
            raise ValueError("write_vtk_fields: x must hold n values")
        f, keep = _fields_struct(fields)
        nb = C.c_int64()
        _check(lib().aphi_write_vtk_fields(self._h, str(path).encode(), xx.view(np.float64).ctypes.data_as(_capi.DP),
                                           float(omega), int(conditioning), C.byref(f), C.byref(nb)))
        return nb.value


class Context:
    """One HIP device and one stream (default priority)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().aphi_context_create(device, C.byref(h)))
        _own(self, "aphi_context_destroy", h)
        self.device = device
        self._meshes = weakref.WeakSet()

    def synchronize(self):
        _check(lib().aphi_context_synchronize(self._h))

    def close(self):
        """Destroys the context; meshes created on it are destroyed first."""
        for m in list(getattr(self, "_meshes", ())):
            m.close()
        _release(self)


_FIELD_ARRAYS = ("a_tet", "b_tet", "e_tet", "j_tet", "a_node", "b_node", "h_node", "e_node", "j_node")


def _fields_struct(fields):
    """An aphi_fields over the arrays of a field dict (complex (count, 3) arrays,
    sigma_tet, vertex_weight, on_material_interface); returns it and the arrays it
    points into."""
    keep = {}
    f = _capi.Fields()
    for k in _FIELD_ARRAYS:
        a = np.ascontiguousarray(fields[k], dtype=np.complex128)
        keep[k] = a
        setattr(f, k, a.view(np.float64).ctypes.data_as(_capi.DP))
    for k in ("sigma_tet", "vertex_weight"):
        a = np.ascontiguousarray(fields[k], dtype=np.float64)
        keep[k] = a
        setattr(f, k, a.ctypes.data_as(_capi.DP))
    a = np.ascontiguousarray(fields["on_material_interface"], dtype=np.int8)
    keep["on_material_interface"] = a
    f.on_material_interface = a.ctypes.data_as(C.POINTER(C.c_int8))
    return f, keep


def _dev_ptr(t):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.complex128):
        raise ValueError("expected a complex128 device tensor")
    return C.c_void_p(t.data_ptr())


class Mesh:
    """Device-resident per-mesh plan and numeric state (what the reference host
    binds once per mesh and reuses at every frequency)."""

    def __init__(self, ctx: Context, problem: Problem):
        self.ctx = ctx
        self.problem = problem  # keeps the described host arrays alive during upload
        desc = problem.describe()
        h = C.c_void_p()
        if not ctx._h:
            raise ValueError("Mesh: the context is closed")
        _check(lib().aphi_mesh_create(ctx._h, C.byref(desc), C.byref(h)))
        _own(self, "aphi_mesh_destroy", h)
        ctx._meshes.add(self)
        self.n = problem.n
        self.nnz = int(problem.counts["nnz"])

    def close(self):
        _release(self)

    def set_solve_graphs(self, on=True):
        """Graph capture and replay of the solve's level loops (off for meshes used
        beside other threads' GPU work)."""
        _check(lib().aphi_mesh_set_solve_graphs(self._h, int(on)))

    def set_update_tiles(self, mode="auto"):
        """Complex-float trailing-update tiling: "auto" (64 x 64 tiles for regions of
        at most 8192 rows or 512 columns), "small" or "large"; the factor is bitwise
        the same."""
        modes = {"auto": 0, "small": 1, "large": 2}
        if mode not in modes:
            raise ValueError(f"update tiles must be one of {sorted(modes)}")
        _check(lib().aphi_mesh_set_update_tiles(self._h, modes[mode]))

    def tune_update_tiles(self, repeats=2):
        """Autotunes the complex-float update tiles per launch class on this mesh's
        assembled system (aphi_mesh_tune_update_tiles); returns (table, ms with the
        default rule, ms with the tuned table). The table (8 x 8, 0 default, 1
        small, 2 large) is kept only if it is faster."""
        tab = np.zeros(64, dtype=np.int8)
        a, b = C.c_double(), C.c_double()
        _check(lib().aphi_mesh_tune_update_tiles(self._h, int(repeats), tab.ctypes.data_as(C.POINTER(C.c_int8)),
                                                 C.byref(a), C.byref(b)))
        return tab.reshape(8, 8), a.value, b.value

    def update_tile_table(self):
        tab = np.zeros(64, dtype=np.int8)
        _check(lib().aphi_mesh_get_update_tile_table(self._h, tab.ctypes.data_as(C.POINTER(C.c_int8))))
        return tab.reshape(8, 8)

    def set_update_tile_table(self, table):
        tab = np.ascontiguousarray(np.asarray(table, dtype=np.int8).reshape(64))
        _check(lib().aphi_mesh_set_update_tile_table(self._h, tab.ctypes.data_as(C.POINTER(C.c_int8))))

    def set_profiling(self, on=True, serial=False):
        """Per-launch timing of the trailing update (``factorize_ldlt`` then reports
        ``update_ms``). ``serial``: every launch on one stream, timed alone, instead
        of the look-ahead schedule whose overlapping launches share the device."""
        _check(lib().aphi_mesh_set_profiling(self._h, (2 if serial else 1) if on else 0))

    def device_bytes(self):
        return int(lib().aphi_mesh_device_bytes(self._h))

    def release_workspace(self):
        """Frees the per-frequency state and keeps the plan (aphi_mesh_release_workspace)."""
        _check(lib().aphi_mesh_release_workspace(self._h))

    # refill(SymmetricSystem&, ..., omega, conditioning) -- assembly.hpp:197
    def refill(self, omega, conditioning=APHI_ROW_SCALED):
        _check(lib().aphi_refill(self._h, float(omega), int(conditioning)))

    def system(self):
        vals = np.zeros(self.nnz, dtype=np.complex128)
        rhs = np.zeros(self.n, dtype=np.complex128)
        _check(lib().aphi_system_download(self._h, vals.ctypes.data_as(_capi.DP), rhs.ctypes.data_as(_capi.DP)))
        return vals, rhs

    def upload_system(self, values, rhs):
        v = np.ascontiguousarray(values, dtype=np.complex128)
        r = np.ascontiguousarray(rhs, dtype=np.complex128)
        if v.size != self.nnz or r.size != self.n:
            raise ValueError("upload_system: wrong sizes")
        _check(lib().aphi_system_upload(self._h, v.ctypes.data_as(_capi.DP), r.ctypes.data_as(_capi.DP)))

    def rhs_device(self):
        """The assembled right-hand side as a device tensor (a device-side copy)."""
        import torch
        out = torch.empty(self.n, dtype=torch.complex128, device=f"cuda:{self.ctx.device}")
        _check(lib().aphi_rhs_copy(self._h, C.c_void_p(out.data_ptr())))
        return out

    # compute_fields' e_tet (src/postprocess.cpp:311-391)
    def tet_fields(self, x, omega, conditioning=APHI_ROW_SCALED):
        """E at every tet's centroid for one solution x: (num_tets, 3) complex device tensor."""
        import torch
        nt = int(self.problem.counts["tets"])
        xx = x.reshape(-1)
        if xx.numel() != self.n:
            raise ValueError("tet_fields: x must hold n values")
        e = torch.empty((nt, 3), dtype=torch.complex128, device=xx.device)
        _check(lib().aphi_tet_fields(self._h, _dev_ptr(xx.contiguous()), float(omega), int(conditioning), _dev_ptr(e)))
        self.ctx.synchronize()
        return e

    def attach_gauge_free(self, problem):
        """Hands a mesh made from Problem.gauge_free(...) its assembly plan."""
        self._gf_desc = problem.describe_gauge_free()
        _check(lib().aphi_gauge_free_attach(self._h, C.byref(self._gf_desc)))
        self.num_vertices = int(self._gf_desc.num_nodes)
        self.num_ports = int(self._gf_desc.num_ports)

    def gauge_free_refill(self, e_tet, omega):
        """Values and rhs of the gauge-free system from e_tet (device, (num_tets, 3))."""
        _check(lib().aphi_gauge_free_refill(self._h, _dev_ptr(e_tet.contiguous()), float(omega)))

    def gauge_free_potential(self, x):
        """(phi_vertex device tensor, port_potential numpy) from the system's solution x."""
        import torch
        phi = torch.empty(self.num_vertices, dtype=torch.complex128, device=x.device)
        ports = np.zeros(max(self.num_ports, 1), dtype=np.complex128)
        _check(lib().aphi_gauge_free_potential(self._h, _dev_ptr(x.reshape(-1).contiguous()), _dev_ptr(phi),
                                               ports.view(np.float64).ctypes.data_as(_capi.DP)))
        return phi, ports[:self.num_ports]

    def equilibrate(self, iterations=10):
        d = np.zeros(self.n, dtype=np.float64)
        _check(lib().aphi_equilibrate(self._h, int(iterations), d.ctypes.data_as(_capi.DP)))
        return d

    # factorize_ldlt -- factorization.hpp:82
    def factorize_ldlt(self, precision=APHI_FP64, pivot_floor=0.0):
        st = _capi.FactorStats()
        status = lib().aphi_factorize(self._h, int(precision), float(pivot_floor), C.byref(st))
        if status != _capi.APHI_PIVOT_FAILED:
            _check(status)
        return {k: getattr(st, k) for k, _ in _capi.FactorStats._fields_}

    def factor_diagonal(self):
        """D in factorization order (see aphi_factor_diagonal)."""
        out = np.zeros(self.n, dtype=np.complex128)
        _check(lib().aphi_factor_diagonal(self._h, out.ctypes.data_as(_capi.DP)))
        return out

    # solve -- factorization.hpp:110 (device tensors, n x nrhs column-major = (nrhs, n) row-major)
    def solve(self, b, x=None):
        import torch
        b2 = b.reshape(-1, self.n)
        if x is None:
            x = torch.empty_like(b2)
        _check(lib().aphi_solve(self._h, b2.shape[0], _dev_ptr(b2), self.n, _dev_ptr(x), self.n))
        return x.reshape(b.shape)

    # solve_symmetric -- factorization.hpp:206, with FP64 refinement
    def solve_symmetric(self, b, x=None, precision=APHI_FP64, equilibration_iterations=10, pivot_floor=0.0,
                        max_refinements=0, tolerance=1e-12):
        import torch
        b2 = b.reshape(-1, self.n)
        if x is None:
            x = torch.empty_like(b2)
        opt = _capi.SolveOptions(int(precision), int(equilibration_iterations), float(pivot_floor),
                                 int(max_refinements), float(tolerance))
        rep = _capi.SolveReport()
        be = np.zeros(b2.shape[0], dtype=np.float64)
        status = lib().aphi_solve_symmetric(self._h, b2.shape[0], _dev_ptr(b2), self.n, _dev_ptr(x), self.n,
                                            C.byref(opt), C.byref(rep), be.ctypes.data_as(_capi.DP))
        if status != _capi.APHI_PIVOT_FAILED:
            _check(status)
        report = {k: getattr(rep, k) for k, _ in _capi.SolveReport._fields_}
        report["backward_errors"] = be
        return x.reshape(b.shape), report

    # terminal_current + impedance -- postprocess.cpp:526-547, postprocess.hpp:356-358
    def terminal_current(self, x, omega, length, axis=(0.0, 0.0, 1.0), voltage=1.0, conditioning=APHI_ROW_SCALED):
        """Port current I (complex), R = Re(V/I) and L = Im(V/I)/omega of each
        solution in x (device complex128, n per row)."""
        x2 = x.reshape(-1, self.n)
        k = x2.shape[0]
        cur = np.zeros(2 * k)
        res = np.zeros(k)
        ind = np.zeros(k)
        ax = np.ascontiguousarray(axis, dtype=np.float64)
        _check(lib().aphi_terminal_current(self._h, k, _dev_ptr(x2), self.n, float(omega), int(conditioning),
                                           ax.ctypes.data_as(_capi.DP), float(length), float(voltage),
                                           cur.ctypes.data_as(_capi.DP), res.ctypes.data_as(_capi.DP),
                                           ind.ctypes.data_as(_capi.DP)))
        return cur.view(np.complex128), res, ind

    # compute_fields -- postprocess.hpp (FieldOutput), src/postprocess.cpp:223-481
    def compute_fields(self, x, omega, conditioning=APHI_ROW_SCALED):
        """compute_fields of the solution x (device complex128, n) on the device:
        a dict of the FieldOutput arrays (per tet a_tet, b_tet, e_tet, j_tet of
        shape (num_tets, 3) and sigma_tet; per P2 node a_node ... j_node of shape
        (num_p2, 3), vertex_weight, on_material_interface) and the interface and
        orphan counts."""
        if not getattr(self, "_fields_attached", False):
            d = self.problem.describe_fields()
            _check(lib().aphi_fields_attach(self._h, C.byref(d)))
            self._fields_np = int(d.num_p2)
            self._fields_attached = True
        nt, npn = int(self.problem.counts["tets"]), self._fields_np
        out = {k: np.zeros((nt if k.endswith("_tet") else npn, 3), dtype=np.complex128) for k in _FIELD_ARRAYS}
        out["sigma_tet"] = np.zeros(nt)
        out["vertex_weight"] = np.zeros(npn)
        out["on_material_interface"] = np.zeros(npn, dtype=np.int8)
        f, keep = _fields_struct(out)
        _check(lib().aphi_compute_fields(self._h, _dev_ptr(x.reshape(-1).contiguous()), float(omega),
                                         int(conditioning), C.byref(f)))
        out.update(keep)
        out["num_interface_nodes"] = int(f.num_interface_nodes)
        out["num_orphan_vertices"] = int(f.num_orphan_vertices)
        return out

    # backward_error / relative_residual -- factorization.hpp:140, :130
    def backward_error(self, x, b):
        x2 = x.reshape(-1, self.n)
        b2 = b.reshape(-1, self.n)
        be = np.zeros(b2.shape[0])
        rr = np.zeros(b2.shape[0])
        _check(lib().aphi_backward_error(self._h, b2.shape[0], _dev_ptr(x2), self.n, _dev_ptr(b2), self.n,
                                         be.ctypes.data_as(_capi.DP), rr.ctypes.data_as(_capi.DP)))
        return be, rr


class Batch(Mesh):
    """Many independent problems in one device plan (aphi_batch_create): the
    block-diagonal union of their meshes, assembled, factored and solved with
    shared launches, each problem measured on its own. Vectors span every
    problem's unknowns: problem p's are rows[p]:rows[p + 1]."""

    def __init__(self, ctx: Context, problems):
        self.ctx = ctx
        self.problems = list(problems)
        descs = (_capi.MeshDesc * len(self.problems))(*[p.describe() for p in self.problems])
        h = C.c_void_p()
        if not ctx._h:
            raise ValueError("Batch: the context is closed")
        _check(lib().aphi_batch_create(ctx._h, descs, len(self.problems), C.byref(h)))
        _own(self, "aphi_mesh_destroy", h)
        ctx._meshes.add(self)
        self.count = len(self.problems)
        self.rows = np.zeros(self.count + 1, dtype=np.int64)
        self.tets = np.zeros(self.count + 1, dtype=np.int64)
        _check(lib().aphi_batch_info(self._h, None, self.rows.ctypes.data_as(_capi.I64P),
                                     self.tets.ctypes.data_as(_capi.I64P)))
        self.n = int(self.rows[-1])
        self.nnz = int(sum(int(p.counts["nnz"]) for p in self.problems))
        self.problem = None

    def part(self, v, p):
        """Problem p's unknowns of a vector (last axis) spanning the batch."""
        return v[..., int(self.rows[p]):int(self.rows[p + 1])]

    # refill per problem -- assembly.hpp:197 at each problem's own frequency
    def refill(self, omegas, conditioning=APHI_ROW_SCALED):
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(omegas, dtype=np.float64), (self.count,)))
        _check(lib().aphi_batch_refill(self._h, w.ctypes.data_as(_capi.DP), int(conditioning)))

    def solve_symmetric(self, b, x=None, precision=APHI_FP64, equilibration_iterations=10, pivot_floor=0.0,
                        max_refinements=0, tolerance=1e-12):
        """solve_symmetric of every problem; returns (x, reports): one report per
        problem with its own backward errors (shape (nrhs,))."""
        import torch
        b2 = b.reshape(-1, self.n)
        if x is None:
            x = torch.empty_like(b2)
        k = b2.shape[0]
        opt = _capi.SolveOptions(int(precision), int(equilibration_iterations), float(pivot_floor),
                                 int(max_refinements), float(tolerance))
        reps = (_capi.SolveReport * self.count)()
        be = np.zeros(self.count * k, dtype=np.float64)
        status = lib().aphi_batch_solve_symmetric(self._h, k, _dev_ptr(b2), self.n, _dev_ptr(x), self.n, C.byref(opt),
                                                  reps, be.ctypes.data_as(_capi.DP))
        if status != _capi.APHI_PIVOT_FAILED:
            _check(status)
        out = []
        for p in range(self.count):
            d = {f: getattr(reps[p], f) for f, _ in _capi.SolveReport._fields_}
            d["backward_errors"] = be[p * k:(p + 1) * k].copy()
            out.append(d)
        return x.reshape(b.shape), out

    def backward_error(self, x, b):
        """(backward errors, relative residuals), each of shape (count, nrhs)."""
        x2 = x.reshape(-1, self.n)
        b2 = b.reshape(-1, self.n)
        k = b2.shape[0]
        be = np.zeros(self.count * k)
        rr = np.zeros(self.count * k)
        _check(lib().aphi_batch_backward_error(self._h, k, _dev_ptr(x2), self.n, _dev_ptr(b2), self.n,
                                               be.ctypes.data_as(_capi.DP), rr.ctypes.data_as(_capi.DP)))
        return be.reshape(self.count, k), rr.reshape(self.count, k)

    def terminal_current(self, x, omegas, lengths, axis=(0.0, 0.0, 1.0), voltages=1.0, conditioning=APHI_ROW_SCALED):
        """(I, R, L) per problem and right-hand side, each of shape (count, nrhs)."""
        x2 = x.reshape(-1, self.n)
        k = x2.shape[0]
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(omegas, dtype=np.float64), (self.count,)))
        ln = np.ascontiguousarray(np.broadcast_to(np.asarray(lengths, dtype=np.float64), (self.count,)))
        vv = np.ascontiguousarray(np.broadcast_to(np.asarray(voltages, dtype=np.float64), (self.count,)))
        ax = np.ascontiguousarray(axis, dtype=np.float64)
        cur = np.zeros(2 * self.count * k)
        res = np.zeros(self.count * k)
        ind = np.zeros(self.count * k)
        _check(lib().aphi_batch_terminal_current(self._h, k, _dev_ptr(x2), self.n, w.ctypes.data_as(_capi.DP),
                                                 int(conditioning), ax.ctypes.data_as(_capi.DP),
                                                 ln.ctypes.data_as(_capi.DP), vv.ctypes.data_as(_capi.DP),
                                                 cur.ctypes.data_as(_capi.DP), res.ctypes.data_as(_capi.DP),
                                                 ind.ctypes.data_as(_capi.DP)))
        return cur.view(np.complex128).reshape(self.count, k), res.reshape(self.count, k), ind.reshape(self.count, k)


def recover_gauge_free_potential(source, gf, x, omega, conditioning=APHI_ROW_SCALED, precision=APHI_FP64,
                                 max_refinements=0, tolerance=1e-12):
    """recover_gauge_free_potential (src/gauge_free_potential.cpp:87-177) on the
    device: E of every tet from the A-Phi solution x of mesh ``source``
    (compute_fields), the P1 system assembled into ``gf`` (a mesh of
    Problem.gauge_free(...) with its plan attached), solve_symmetric with 10
    equilibration passes (the reference's options; FP32 + refinement on request),
    and the vertex / port potentials. Returns the fields of GaugeFreePotential."""
    import time
    t0 = time.perf_counter()
    e = source.tet_fields(x, omega, conditioning)
    gf.gauge_free_refill(e, omega)
    _, rhs = gf.system()
    import torch
    b = torch.from_numpy(rhs).to(x.device)
    t1 = time.perf_counter()
    xs, rep = gf.solve_symmetric(b, precision=precision, equilibration_iterations=10, max_refinements=max_refinements,
                                 tolerance=tolerance)
    t2 = time.perf_counter()
    phi, ports = gf.gauge_free_potential(xs)
    return dict(phi_vertex=phi, port_potential=ports, unknowns=gf.n, converged=bool(rep["ok"]),
                residual=rep["residual"], backward_error=rep["backward_error"], assemble_ms=(t1 - t0) * 1e3,
                solve_ms=(t2 - t1) * 1e3, report=rep)
