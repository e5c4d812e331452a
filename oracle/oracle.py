"""ctypes wrapper of oracle/oracle.c -- TEST INFRASTRUCTURE ONLY (the checker).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never
by the product path. Inputs are the host problem's exported arrays, so the oracle
sees the same mesh, DOF map, pattern and ordering as the device path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    """gcc -O2 -ffp-contract=off: the reference's rounding on x86-64."""
    cmd = ["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-fPIC", "-shared", "-Wall",
           "-o", LIB + ".tmp", os.path.join(HERE, "oracle.c"), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def element(coords12, nu, alpha, beta):
    """196 complex element entries of one tet (36 AA, 60 APhi, 100 PhiPhi)."""
    out = np.zeros(392)
    ok = lib().oracle_element(_p(_f64(coords12), C.c_double), C.c_double(nu),
                              _p(_f64([alpha.real, alpha.imag]), C.c_double),
                              _p(_f64([beta.real, beta.imag]), C.c_double), _p(out, C.c_double))
    if not ok:
        raise RuntimeError("compute_tet_geometry: degenerate (zero-volume) tet")
    return out.view(np.complex128)


def refill(problem, omega, conditioning=1):
    """values (nnz complex) and rhs (n complex) of the upper-triangle system."""
    n = problem.n
    rp, ci = _i32(problem.ints("row_ptr")), _i32(problem.ints("col_index"))
    vals = np.zeros(2 * len(ci))
    rhs = np.zeros(2 * n)
    port_row = _i32(problem.ints("port_index"))
    nports = len(port_row)
    fixed = problem.ints("port_is_fixed")
    port_cur = _i32(1 - fixed)  # a current-driven port's row takes row * I, a voltage port's V / column
    args = [
        C.c_int(n), C.c_int64(len(ci)), _p(rp, C.c_int32), _p(ci, C.c_int32),
        _p(_f64(problem.doubles("coords")), C.c_double), _p(_i32(problem.ints("tet_nodes")), C.c_int32),
    ]
    body_ptr = _i32(problem.ints("body_ptr"))
    keep = [rp, ci, body_ptr]
    arrays = [body_ptr, _i32(problem.ints("body_tets")), _f64(problem.doubles("body_param")),
              _i32(problem.ints("tet_dofs")), _f64(problem.ints("tet_edge_signs")),
              _f64(problem.doubles("tet_phi_fixed")), _i32(problem.ints("scatter_live")),
              _i32(problem.ints("scatter_pos")), np.ascontiguousarray(problem.ints("scatter_offset"), dtype=np.int64),
              _i32(problem.ints("scatter_slot"))]
    keep += arrays
    types = [C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32,
             C.c_int64, C.c_int32]
    args += [C.c_int(len(body_ptr) - 1)] + [_p(a, t) for a, t in zip(arrays, types)]
    pv = _f64(np.where(np.repeat(fixed, 2) == 1, problem.doubles("port_value"), problem.doubles("port_amplitude")))
    keep += [port_row, port_cur, pv]
    args += [C.c_double(omega), C.c_int(conditioning), C.c_int(nports), _p(port_row, C.c_int32),
             _p(port_cur, C.c_int32), _p(pv, C.c_double), _p(vals, C.c_double), _p(rhs, C.c_double)]
    st = lib().oracle_refill(*args)
    if st != 0:
        raise RuntimeError(f"oracle_refill failed ({st})")
    return vals.view(np.complex128), rhs.view(np.complex128)


def equilibrate(problem, values, iterations=10):
    n = problem.n
    rp, ci = _i32(problem.ints("row_ptr")), _i32(problem.ints("col_index"))
    v = np.ascontiguousarray(values, dtype=np.complex128)
    d = np.zeros(n)
    lib().oracle_equilibrate(C.c_int(n), _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v.view(np.float64), C.c_double),
                             C.c_int(iterations), _p(d, C.c_double))
    return d


def solve_symmetric(problem, values, b, perm=None, iterations=10, pivot_floor=0.0):
    """solve_symmetric of the reference for b of shape (nrhs, n); returns (x, D, stats)."""
    n = problem.n
    rp, ci = _i32(problem.ints("row_ptr")), _i32(problem.ints("col_index"))
    v = np.ascontiguousarray(values, dtype=np.complex128)
    B = np.ascontiguousarray(np.atleast_2d(b), dtype=np.complex128)
    X = np.zeros_like(B)
    D = np.zeros(n, dtype=np.complex128)
    pm = _i32(problem.ints("perm") if perm is None else perm)
    st = np.zeros(8)
    lib().oracle_solve_symmetric(C.c_int(n), _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v.view(np.float64), C.c_double),
                                 C.c_int(B.shape[0]), _p(B.view(np.float64), C.c_double), _p(pm, C.c_int32),
                                 C.c_int(iterations), C.c_double(pivot_floor), _p(X.view(np.float64), C.c_double),
                                 _p(D.view(np.float64), C.c_double), _p(st, C.c_double))
    stats = dict(smallest_pivot=st[0], largest_pivot=st[1], largest_multiplier=st[2], columns_done=int(st[3]),
                 ok=bool(st[4]), backward_error=st[5], equilibration_min=st[6], equilibration_max=st[7])
    return X.reshape(np.shape(b)), D, stats


def backward_error(problem, values, x, b):
    n = problem.n
    rp, ci = _i32(problem.ints("row_ptr")), _i32(problem.ints("col_index"))
    v = np.ascontiguousarray(values, dtype=np.complex128)
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    bb = np.ascontiguousarray(b, dtype=np.complex128)
    out = np.zeros(2)
    lib().oracle_backward_error(C.c_int(n), _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v.view(np.float64), C.c_double),
                                _p(xx.view(np.float64), C.c_double), _p(bb.view(np.float64), C.c_double),
                                _p(out, C.c_double))
    return out[0], out[1]


def residual(problem, values, x, b):
    """r = b - A x on the original system (the matvec of backward_error)."""
    n = problem.n
    rp, ci = _i32(problem.ints("row_ptr")), _i32(problem.ints("col_index"))
    v = np.ascontiguousarray(values, dtype=np.complex128)
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    bb = np.ascontiguousarray(b, dtype=np.complex128)
    r = np.zeros(n, dtype=np.complex128)
    lib().oracle_residual(C.c_int(n), _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v.view(np.float64), C.c_double),
                          _p(xx.view(np.float64), C.c_double), _p(bb.view(np.float64), C.c_double),
                          _p(r.view(np.float64), C.c_double))
    return r


def csr_solve_symmetric_raw(n, row_ptr, col, values, b, perm, iterations=0, pivot_floor=0.0):
    """solve_symmetric on a raw upper-triangle CSR (for hand-built test matrices)."""
    rp, ci = _i32(row_ptr), _i32(col)
    v = np.ascontiguousarray(values, dtype=np.complex128)
    B = np.ascontiguousarray(np.atleast_2d(b), dtype=np.complex128)
    X = np.zeros_like(B)
    D = np.zeros(n, dtype=np.complex128)
    pm = _i32(perm)
    st = np.zeros(8)
    lib().oracle_solve_symmetric(C.c_int(n), _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v.view(np.float64), C.c_double),
                                 C.c_int(B.shape[0]), _p(B.view(np.float64), C.c_double), _p(pm, C.c_int32),
                                 C.c_int(iterations), C.c_double(pivot_floor), _p(X.view(np.float64), C.c_double),
                                 _p(D.view(np.float64), C.c_double), _p(st, C.c_double))
    return X.reshape(np.shape(b)), D, dict(ok=bool(st[4]), columns_done=int(st[3]), smallest_pivot=st[0],
                                           largest_pivot=st[1], largest_multiplier=st[2], backward_error=st[5])


def reconstruction_error(n, row_ptr, col, values, perm):
    lib().oracle_reconstruction_error.restype = C.c_double
    v = np.ascontiguousarray(values, dtype=np.complex128)
    rp, ci, pm = _i32(row_ptr), _i32(col), _i32(perm)
    return lib().oracle_reconstruction_error(C.c_int(n), _p(rp, C.c_int32), _p(ci, C.c_int32),
                                             _p(v.view(np.float64), C.c_double), _p(pm, C.c_int32))


def factor_sample(problem, values, budget_s):
    """The reference's scalar up-looking factorization of the equilibrated system,
    run for at most budget_s seconds (the CPU baseline sample)."""
    n = problem.n
    rp, ci = _i32(problem.ints("row_ptr")), _i32(problem.ints("col_index"))
    v = np.ascontiguousarray(values, dtype=np.complex128)
    pm = _i32(problem.ints("perm"))
    out = np.zeros(5)
    lib().oracle_factor_sample(C.c_int(n), _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v.view(np.float64), C.c_double),
                               _p(pm, C.c_int32), C.c_double(budget_s), _p(out, C.c_double))
    return dict(seconds=out[0], rows=int(out[1]), flops_done=out[2], flops_total=out[3], equilibrate_s=out[4])


def terminal_current(problem, x, omega, length, axis=(0.0, 0.0, 1.0), voltage=1.0, conditioning=1, terms=False):
    """Port current I and R, L of one solution x (n complex); with terms=True also
    the per-tet flux terms (complex, zero outside the conductor)."""
    nt = int(problem.counts["tets"])
    bp = _f64(problem.doubles("body_param")).reshape(-1, 3)
    sig = np.ascontiguousarray(bp[:, 0])
    arrays = [_f64(problem.doubles("coords")), _i32(problem.ints("tet_nodes")), _i32(problem.ints("tet_body")), sig,
              _i32(problem.ints("tet_dofs")), _f64(problem.ints("tet_edge_signs")),
              _f64(problem.doubles("tet_phi_coeff")), _f64(problem.doubles("tet_phi_fixed"))]
    types = [C.c_double, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_double, C.c_double]
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    ax = _f64(axis)
    out = np.zeros(4)
    tr = np.zeros(2 * nt) if terms else None
    ok = lib().oracle_terminal_current(C.c_int(nt), *[_p(a, t) for a, t in zip(arrays, types)],
                                       _p(xx.view(np.float64), C.c_double), C.c_double(omega), C.c_int(conditioning),
                                       _p(ax, C.c_double), C.c_double(length), C.c_double(voltage),
                                       _p(out, C.c_double), _p(tr, C.c_double) if terms else None)
    if not ok:
        raise ValueError("terminal_current: length must be positive (or a tet is degenerate)")
    I = complex(out[0], out[1])
    if terms:
        return I, out[2], out[3], tr.view(np.complex128)
    return I, out[2], out[3]


def tet_fields(problem, x, omega, conditioning=1):
    """compute_fields' e_tet for one solution x: (num_tets, 3) complex."""
    nt = int(problem.counts["tets"])
    arrays = [_f64(problem.doubles("coords")), _i32(problem.ints("tet_nodes")), _i32(problem.ints("tet_dofs")),
              _f64(problem.ints("tet_edge_signs")), _f64(problem.doubles("tet_phi_coeff")),
              _f64(problem.doubles("tet_phi_fixed"))]
    types = [C.c_double, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double]
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(6 * nt)
    ok = lib().oracle_tet_fields(C.c_int(nt), *[_p(a, t) for a, t in zip(arrays, types)],
                                 _p(xx.view(np.float64), C.c_double), C.c_double(omega), C.c_int(conditioning),
                                 _p(out, C.c_double))
    if not ok:
        raise ValueError("tet_fields: a tet is degenerate")
    return out.view(np.complex128).reshape(nt, 3)


def compute_fields(problem, x, omega, conditioning=1):
    """compute_fields (src/postprocess.cpp:223-481) of one solution x, without
    phi_node: a dict with the same keys as Mesh.compute_fields."""
    nt = int(problem.counts["tets"])
    tet_p2 = _i32(problem.ints("tet_p2"))
    npn = int(tet_p2.max()) + 1 if tet_p2.size else 0
    npn = max(npn, int(problem.counts["nodes"]) + int(problem.counts["edges"]))
    bp = _f64(problem.doubles("body_param")).reshape(-1, 3)
    sig, mu = np.ascontiguousarray(bp[:, 0]), np.ascontiguousarray(bp[:, 2])
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    out = {k: np.zeros((nt if k.endswith("_tet") else npn, 3), dtype=np.complex128)
           for k in ("a_tet", "b_tet", "e_tet", "j_tet", "a_node", "b_node", "h_node", "e_node", "j_node")}
    out["sigma_tet"] = np.zeros(nt)
    out["vertex_weight"] = np.zeros(npn)
    out["on_material_interface"] = np.zeros(npn, dtype=np.int8)
    counts = np.zeros(2, dtype=np.int32)
    d = lambda k: _p(out[k].view(np.float64), C.c_double)  # noqa: E731
    ok = lib().oracle_compute_fields(
        C.c_int(nt), C.c_int(npn), _p(_f64(problem.doubles("coords")), C.c_double),
        _p(_i32(problem.ints("tet_nodes")), C.c_int32), _p(tet_p2, C.c_int32),
        _p(_i32(problem.ints("tet_body")), C.c_int32), _p(sig, C.c_double), _p(mu, C.c_double),
        _p(_i32(problem.ints("tet_dofs")), C.c_int32), _p(_f64(problem.ints("tet_edge_signs")), C.c_double),
        _p(_f64(problem.doubles("tet_phi_coeff")), C.c_double), _p(_f64(problem.doubles("tet_phi_fixed")), C.c_double),
        _p(xx.view(np.float64), C.c_double), C.c_double(omega), C.c_int(conditioning),
        d("a_tet"), d("b_tet"), d("e_tet"), d("j_tet"), _p(out["sigma_tet"], C.c_double),
        d("a_node"), d("b_node"), d("h_node"), d("e_node"), d("j_node"), _p(out["vertex_weight"], C.c_double),
        _p(out["on_material_interface"], C.c_int8), _p(counts, C.c_int32))
    if not ok:
        raise ValueError("compute_fields: a tet is degenerate")
    out["num_interface_nodes"] = int(counts[0])
    out["num_orphan_vertices"] = int(counts[1])
    return out


def gauge_free_system(problem, e_tet, omega):
    """The vertex map and P1 system of recover_gauge_free_potential for the cube
    problem's mesh and the field e_tet ((num_tets, 3) complex). Returns a dict
    with unknowns, column, prescribed, port_column, row_ptr, col, values, rhs."""
    nv, nt = int(problem.counts["nodes"]), int(problem.counts["tets"])
    bp = _f64(problem.doubles("body_param")).reshape(-1, 3)
    sig, eps = np.ascontiguousarray(bp[:, 0]), np.ascontiguousarray(bp[:, 1])
    state = _i32(problem.ints("phi_state")[:nv])
    port = _i32(problem.ints("phi_port")[:nv])
    fixed = _i32(problem.ints("port_is_fixed"))
    pval = _f64(problem.doubles("port_value"))
    ee = np.ascontiguousarray(e_tet, dtype=np.complex128).reshape(-1)
    cap = 64 * nv + 64
    unk = np.zeros(1, dtype=np.int32)
    column = np.zeros(nv, dtype=np.int32)
    prescribed = np.zeros(2 * nv)
    port_column = np.zeros(max(fixed.size, 1), dtype=np.int32)
    row_ptr = np.zeros(nv + 1, dtype=np.int32)
    col = np.zeros(cap, dtype=np.int32)
    values = np.zeros(2 * cap)
    rhs = np.zeros(2 * nv)
    f = lib().oracle_gauge_free_system
    f.restype = C.c_int64
    nnz = f(C.c_int(nv), C.c_int(nt), _p(_f64(problem.doubles("coords")), C.c_double),
            _p(_i32(problem.ints("tet_nodes")), C.c_int32), _p(_i32(problem.ints("tet_body")), C.c_int32),
            _p(sig, C.c_double), _p(eps, C.c_double), _p(state, C.c_int32), _p(port, C.c_int32),
            C.c_int(fixed.size), _p(fixed, C.c_int32), _p(pval, C.c_double), _p(ee.view(np.float64), C.c_double),
            C.c_double(omega), C.c_int64(cap), _p(unk, C.c_int32), _p(column, C.c_int32), _p(prescribed, C.c_double),
            _p(port_column, C.c_int32), _p(row_ptr, C.c_int32), _p(col, C.c_int32), _p(values, C.c_double),
            _p(rhs, C.c_double))
    if nnz < 0:
        raise ValueError("gauge_free_system: a row exceeds the oracle's capacity")
    n = int(unk[0])
    return dict(unknowns=n, column=column, prescribed=prescribed.view(np.complex128),
                port_column=port_column[:fixed.size], row_ptr=row_ptr[:n + 1], col=col[:nnz],
                values=values[:2 * nnz].view(np.complex128).copy(), rhs=rhs[:2 * n].view(np.complex128).copy())
