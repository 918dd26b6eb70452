"""CPU checks of the boundary and the host-side pieces (no GPU calls)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import scipy.sparse as sp

from hazmath_amd import capi, problems

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in {"if", "while", "for", "return", "sizeof"}:
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    lib = capi.load_library()
    names = declared_symbols()
    assert "linear_solver_dcsr_krylov_amg" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_agree_between_header_and_ctypes(tmp_path):
    """Compile the public header with gcc and compare every field offset and
    struct size with the ctypes mirror used by the tests and the benchmark."""
    import subprocess
    structs = {"dCSRmat": capi.dCSRmat, "dvector": capi.dvector,
               "linear_itsolver_param": capi.linear_itsolver_param, "AMG_param": capi.AMG_param,
               "hz_options": capi.hz_options, "hz_solve_info": capi.hz_solve_info,
               "hz_hierarchy": capi.hz_hierarchy, "hz_kernel_times": capi.hz_kernel_times,
               "precond": capi.precond, "hz_batch_stats": capi.hz_batch_stats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hazmath_mi355x.h"',
             'int main(void){']
    for s, cls in structs.items():
        lines.append(f'printf("{s} sizeof %zu\\n", sizeof({s}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{s} {fname} %zu\\n", offsetof({s}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        s, field, val = line.split()
        cls = structs[s]
        got = C.sizeof(cls) if field == "sizeof" else getattr(cls, field).offset
        assert got == int(val), (s, field)
    # the reference ABI: 3 ints + 3 pointers; REAL tol after SHORT/SHORT/INT
    assert C.sizeof(capi.dCSRmat) == 40 and capi.AMG_param.tol.offset == 8


def test_param_defaults_match_reference():
    lib = capi.load_library()
    amg = capi.default_amg_param(lib)
    assert (amg.AMG_type, amg.maxit, amg.max_levels, amg.coarse_dof, amg.cycle_type,
            amg.smoother, amg.coarse_solver, amg.aggregation_type, amg.max_aggregation) == \
        (1, 1, 10, 200, 2, 2, 0, 4, 20)
    assert amg.tol == 1e-6 and amg.strong_coupled == 0.04 and amg.tentative_smooth == 0.67
    it = capi.default_itsolver_param(lib)
    assert (it.linear_itsolver_type, it.linear_precond_type, it.linear_stop_type,
            it.linear_maxit, it.linear_restart) == (4, 0, 1, 500, 100)
    assert it.linear_tol == 1e-6


def test_param_defaults_agree_with_oracle():
    import oracle
    lib, orc = capi.load_library(), oracle.load()
    a, b = capi.AMG_param(), capi.AMG_param()
    lib.param_amg_init(C.byref(a))
    orc.orc_param_amg_init(C.byref(b))
    assert bytes(a) == bytes(b)


def test_coo_reader_matches_fixture(tmp_path):
    lib = capi.load_library()
    d = np.load(ROOT / "tests" / "golden" / "solvers_A_b.npz")
    ia, ja, va = d["ia"], d["ja"], d["val"]
    n = ia.size - 1
    rows = np.repeat(np.arange(n), np.diff(ia))
    p = tmp_path / "A.dat"
    with open(p, "w") as f:
        f.write(f"{n} {n} {ia[-1]}\n")
        for r, c, v in zip(rows, ja, va):
            f.write(f"{r} {c} {v:.15e}\n")
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    lib.dcoo_read_dcsr_p.restype = C.POINTER(capi.dCSRmat)
    lib.dcoo_read_dcsr_p.argtypes = [C.c_void_p]
    fp = libc.fopen(str(p).encode(), b"r")
    A = lib.dcoo_read_dcsr_p(fp)
    libc.fclose(fp)
    gia, gja, gva = capi.csr_to_numpy(A.contents)
    assert np.array_equal(gia, ia) and np.array_equal(gja, ja) and np.array_equal(gva, va)


def test_coo_reader_unsorted_input_is_row_stable(tmp_path):
    """dcoo_2_dcsr_p (format.c:785-827) is a counting sort by row: entries of a
    row keep the order they have in the file (duplicates stay separate).  The
    expected CSR is numpy's stable argsort of the row indices."""
    lib = capi.load_library()
    rng = np.random.default_rng(3)
    n, nnz = 57, 400
    rows = rng.integers(0, n, nnz)
    cols = rng.integers(0, n, nnz)
    rows[:5] = 7
    cols[:5] = [3, 3, 9, 1, 3]          # a repeated entry and a descending run in one row
    vals = rng.standard_normal(nnz)
    p = tmp_path / "A.dat"
    with open(p, "w") as f:
        f.write(f"{n} {n} {nnz}\n")
        for r, c, v in zip(rows, cols, vals):
            f.write(f"{r} {c} {v:.17e}\n")
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    lib.dcoo_read_dcsr_p.restype = C.POINTER(capi.dCSRmat)
    lib.dcoo_read_dcsr_p.argtypes = [C.c_void_p]
    fp = libc.fopen(str(p).encode(), b"r")
    A = lib.dcoo_read_dcsr_p(fp)
    libc.fclose(fp)
    gia, gja, gva = capi.csr_to_numpy(A.contents)
    order = np.argsort(rows, kind="stable")
    eia = np.zeros(n + 1, np.int64)
    eia[1:] = np.cumsum(np.bincount(rows, minlength=n))
    assert np.array_equal(gia, eia)
    assert np.array_equal(gja, cols[order]) and np.array_equal(gva, vals[order])


def test_device_entry_fails_loudly_without_gpu():
    lib = capi.load_library()
    if lib.hz_device_count() > 0:
        return
    it, amg = capi.solver_params(lib)
    ia, ja, va = problems.poisson2d_p1(8)
    A = capi.HostCSR(ia, ja, va)
    b = capi.HostVec(np.ones(ia.size - 1))
    x = capi.HostVec(np.zeros(ia.size - 1))
    assert lib.linear_solver_dcsr_krylov_amg(A.ptr(), b.ptr(), x.ptr(), C.byref(it),
                                             C.byref(amg)) == capi.ERROR_DEVICE


def test_poisson2d_generator():
    ia, ja, va = problems.poisson2d_p1(128)
    assert ia.size - 1 == 16129 and ia[-1] == 80137
    k = 127
    T = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(k, k))
    L = (sp.kron(sp.eye(k), T) + sp.kron(T, sp.eye(k))).tocsr()
    assert abs(sp.csr_matrix((va, ja, ia)) - L).max() == 0
    assert all(np.all(np.diff(ja[ia[i]:ia[i + 1]]) > 0) for i in range(0, 16129, 997))


def test_kuhn_generator_sizes_and_stencil():
    ia, ja, va = problems.poisson3d_kuhn(16)
    N = 15
    assert ia.size - 1 == N ** 3
    assert ia[-1] == N ** 3 + 6 * N * N * (N - 1) + 6 * N * (N - 1) ** 2 + 2 * (N - 1) ** 3
    A = sp.csr_matrix((va, ja, ia))
    # symmetric, row sums zero in the interior, 15 stored entries in interior rows
    assert abs(A - A.T).max() == 0
    i = 7 + 7 * N + 7 * N * N
    assert ia[i + 1] - ia[i] == 15
    assert abs(A[i].sum()) < 1e-15
    assert A[i, i] == 6.0 / 16


def test_kuhn_stencil_from_element_assembly():
    """Assemble the six Kuhn tetrahedra of a cube; the interior stencil must be
    6 / -1 on the axes and zero on the Kuhn diagonals (the stored zeros)."""
    K = problems.kuhn_element_stiffness_check()
    st = {}
    for c in range(8):
        pc = np.array([(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1])
        for v in range(8):
            pv = np.array([(v >> 0) & 1, (v >> 1) & 1, (v >> 2) & 1])
            d = tuple(int(t) for t in pv - pc)
            st[d] = st.get(d, 0.0) + K[c, v]
    assert abs(st[(0, 0, 0)] - 6.0) < 1e-12
    for ax in [(1, 0, 0), (0, 1, 0), (0, 0, 1)]:
        assert abs(st[ax] + 1.0) < 1e-12
        assert abs(st[tuple(-a for a in ax)] + 1.0) < 1e-12
    for dzyx in problems.KUHN_DIAGONALS:
        dxyz = tuple(reversed(dzyx))
        assert abs(st[dxyz]) < 1e-12


def test_uniform_rhs_is_counter_based():
    a = problems.uniform_rhs(1000, 3)
    b = problems.uniform_rhs(2000, 3)
    assert np.array_equal(a, b[:1000])
    assert a.min() >= -1 and a.max() < 1
    assert abs(a.mean()) < 0.1
    assert not np.array_equal(a, problems.uniform_rhs(1000, 4))


def test_public_structs_match_reference_layout():
    """Every field offset and size of the public structs in
    include/hazmath_mi355x.h equals the reference's (layout compiled from the
    reference headers into tests/golden/abi_layout_reference.json)."""
    import json
    import sys
    sys_path_golden = Path(__file__).parent / "golden"
    sys.path.insert(0, str(sys_path_golden))
    import make_abi_layout as mal
    ref = json.loads((sys_path_golden / "abi_layout_reference.json").read_text())
    mine = mal.compile_and_run(mal.layout_source(["hazmath_mi355x.h"]), ROOT / "include")
    assert mine == ref


INPUT_DAT = """%----------------------------------------------%
% parameters for the parser test                 %
%----------------------------------------------%
print_level   = 3;   % trailing text is skipped
gridfile      = ../grids/3D/cube.haz
[section] anything here is a comment
linear_itsolver_type    = 1
linear_itsolver_maxit   = 77
linear_itsolver_tol     = 2.5e-9
linear_stop_type        = 2
linear_precond_type     = 2
AMG_type                = UA
AMG_cycle_type          = w
AMG_levels              = 12
AMG_tol                 = 1e-7
AMG_maxit               = 3
AMG_smoother            = L1DIAG
AMG_coarse_scaling      = oN
AMG_smooth_filter       = OFF
AMG_presmooth_iter      = 2
AMG_postsmooth_iter     = 3
AMG_coarse_dof          = 150
AMG_coarse_solver       = 32
AMG_aggregation_type    = 5
AMG_strong_coupled      = 0.125
AMG_max_aggregation     = 9
not_a_keyword           = 5
fas_smooth_tol          = 7.9
BSR_omega               = 0.5
"""


def _parse(lib, text, tmp_path):
    f = tmp_path / "input.dat"
    f.write_text(text)
    p = capi.input_param()
    lib.param_input(str(f).encode(), C.byref(p))
    return p


def test_param_input_reads_reference_format(tmp_path, capfd):
    """param_input (input.c:42-830): keyword = value lines, '%'/'['/'|'
    comments, names mapped to the macro.h codes, rest of line ignored, an
    unknown keyword only warns."""
    lib = capi.load_library()
    p = _parse(lib, INPUT_DAT, tmp_path)
    out = capfd.readouterr().out
    assert "Unknown input keyword not_a_keyword" in out
    assert p.print_level == 3 and p.gridfile == b"../grids/3D/cube.haz"
    assert p.linear_itsolver_type == 1 and p.linear_itsolver_maxit == 77
    assert p.linear_itsolver_tol == 2.5e-9 and p.linear_stop_type == 2 and p.linear_precond_type == 2
    assert p.AMG_type == 1 and p.AMG_cycle_type == 2 and p.AMG_levels == 12
    assert p.AMG_tol == 1e-7 and p.AMG_maxit == 3 and p.AMG_smoother == 10
    assert p.AMG_coarse_scaling == 1 and p.AMG_smooth_filter == 0
    assert (p.AMG_presmooth_iter, p.AMG_postsmooth_iter, p.AMG_coarse_dof) == (2, 3, 150)
    assert (p.AMG_coarse_solver, p.AMG_aggregation_type, p.AMG_max_aggregation) == (32, 5, 9)
    assert p.AMG_strong_coupled == 0.125 and p.fas_smooth_tol == 7 and p.BSR_omega == 0.5
    # untouched fields keep param_input_init's defaults (parameters.c:39-141)
    assert p.AMG_relaxation == 1.2 and p.linear_restart == 25 and p.Schwarz_mmsize == 200
    assert p.inifile == str(tmp_path / "input.dat").encode()


def test_param_input_stops_at_malformed_entry(tmp_path):
    lib = capi.load_library()
    p = _parse(lib, "AMG_levels = 7\nAMG_smoother = bogus\nAMG_coarse_dof = 99\n", tmp_path)
    assert p.AMG_levels == 7
    assert p.AMG_smoother == 2          # default GS: the bad value ended the reading
    assert p.AMG_coarse_dof == 200      # never reached
    p = _parse(lib, "AMG_levels 7\n", tmp_path)     # missing '='
    assert p.AMG_levels == 10


def test_param_setters_copy_like_reference(tmp_path):
    """param_linear_solver_set / param_amg_set (parameters.c:263-348),
    including the SOLVER_AMG swap of (tol, maxit) between the two."""
    lib = capi.load_library()
    p = _parse(lib, INPUT_DAT, tmp_path)
    it, amg = capi.linear_itsolver_param(), capi.AMG_param()
    lib.param_linear_solver_init(C.byref(it))
    lib.param_amg_init(C.byref(amg))
    lib.param_linear_solver_set(C.byref(it), C.byref(p))
    lib.param_amg_set(C.byref(amg), C.byref(p))
    assert (it.linear_itsolver_type, it.linear_maxit, it.linear_tol) == (1, 77, 2.5e-9)
    assert (amg.maxit, amg.tol, amg.cycle_type, amg.smoother) == (3, 1e-7, 2, 10)
    assert (amg.aggregation_type, amg.strong_coupled, amg.coarse_scaling) == (5, 0.125, 1)
    p2 = _parse(lib, "linear_itsolver_type = 21\nlinear_itsolver_maxit = 40\nlinear_itsolver_tol = 1e-5\n"
                "AMG_tol = 1e-3\nAMG_maxit = 4\n", tmp_path)
    lib.param_linear_solver_set(C.byref(it), C.byref(p2))
    lib.param_amg_set(C.byref(amg), C.byref(p2))
    assert (it.linear_maxit, it.linear_tol) == (4, 1e-3)
    assert (amg.maxit, amg.tol) == (40, 1e-5)


def test_mlp_init_layout():
    """hz_mlp_init: deterministic, padding and biases zero, Glorot range."""
    from oracle import mlp
    lib = capi.load_library()
    a = np.zeros(capi.MLP_PARAMS, np.float32)
    b = np.zeros(capi.MLP_PARAMS, np.float32)
    lib.hz_mlp_init(capi.fptr(a), 7)
    lib.hz_mlp_init(capi.fptr(b), 7)
    assert np.array_equal(a, b)
    assert capi.MLP_PARAMS == mlp.N_PARAMS
    assert np.all(a[mlp.OFF_W1:mlp.OFF_B1].reshape(16, 64)[12:] == 0)
    assert np.all(a[mlp.OFF_W3:mlp.OFF_B3].reshape(64, 16)[:, 2:] == 0)
    assert np.all(a[mlp.OFF_B1:mlp.OFF_W2] == 0) and np.all(a[mlp.OFF_B3:] == 0)
    w1 = a[mlp.OFF_W1:mlp.OFF_B1].reshape(16, 64)[:12]
    assert np.abs(w1).max() <= np.sqrt(6.0 / 76) and np.abs(w1).std() > 0.05


def test_committed_model_loads():
    from hazmath_amd import learned
    theta = learned.load_model()
    assert theta.dtype == np.float32 and theta.shape == (capi.MLP_PARAMS,)
    assert np.isfinite(theta).all()


def test_batch_rejects_settings_outside_its_scope():
    """The batched driver refuses, before touching the device, every setting it
    does not implement (ERROR_INPUT_PAR = -13)."""
    import ctypes as C
    lib = capi.load_library()
    ia, ja, va = problems.aniso2d_p1(8, 0.2, 0.1)
    b = np.ones((8, ia.size - 1))
    x = np.zeros((8, ia.size - 1))
    for field, value in (("aggregation_type", capi.HEC), ("smoother", capi.SMOOTHER_GS), ("cycle_type", 2),
                         ("coarse_solver", 0), ("maxit", 2), ("AMG_type", capi.SA_AMG)):
        it, amg = capi.solver_params(lib, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG)
        setattr(amg, field, value)
        r, _ = capi.batch_solve(lib, [(ia, ja, va)], [b], [x], it, amg)
        assert r == -13, field
    it, amg = capi.solver_params(lib, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG)
    it.linear_stop_type = 2
    assert capi.batch_solve(lib, [(ia, ja, va)], [b], [x], it, amg)[0] == -13
    # nrhs beyond HZ_BATCH_RHS
    it, amg = capi.solver_params(lib, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG)
    r, _ = capi.batch_solve(lib, [(ia, ja, va)], [np.ones((9, ia.size - 1))], [np.zeros((9, ia.size - 1))], it, amg)
    assert r == -13
