"""CPU tests of the product library's host side and C ABI (no GPU needed).

Host-only plans build every table the kernels read; they are compared with the
oracle's restatement of BoxTree (Morton sort, relevant cones per level) and of
the RP precomputation (folded correction rows).
"""
import ctypes
import os
import re

import numpy as np
import pytest

import ifgf2d_amd as I
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ifgf2d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ifgf_\w+)\s*\(", text)) - {"ifgf_apply_fn"})


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(I.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(I.EXPORTS)


CASES = [
    ("circle", I.circle(), 40, 10.0, 3, 2, 5, 5),
    ("kite", I.kite(), 400, 40.0, 6, 2, 5, 5),
    ("star", I.star(a=[0, 0, 0, 0, 0, 0.3]), 800, 60.0, 7, 2, 5, 5),
    ("star-slp-5x7", I.star(a=[0, 0, 0, 0, 0, 0.3]), 800, 60.0, 7, 0, 5, 7),
    ("kite-dlp", I.kite(), 400, 40.0, 6, 1, 3, 5),
]


def pair(curve, npatch, k, nlev, F, Ps, Pang):
    lims = I.equal_parameter_lims(npatch)
    delta = I.default_delta(curve, lims)
    fm = I.ForwardMap(curve, lims, k, nlev, formulation=F, delta=delta, p=6.0, Ps=Ps, Pang=Pang,
                      host_only=True)
    om = O.ForwardMap(curve[0], curve[1], lims, formulation=F, k=k, eta=k, delta=delta, p=6.0,
                      Ps=Ps, Pang=Pang, nlevels=nlev)
    return fm, om


@pytest.mark.parametrize("name,curve,npatch,k,nlev,F,Ps,Pang", CASES, ids=[c[0] for c in CASES])
def test_tables_match_oracle(name, curve, npatch, k, nlev, F, Ps, Pang):
    fm, om = pair(curve, npatch, k, nlev, F, Ps, Pang)
    # geometry
    for a, b in zip(fm.points(), om.points()):
        assert np.max(np.abs(a - b)) <= 4e-16 * max(1.0, np.max(np.abs(b)))
    # Morton permutation (BoxTree::SortBox) is an exact integer match
    assert np.array_equal(fm.sorting(), om.sorting())
    # relevant cones per level, canonical order
    for level in range(2, nlev):
        b1, q1 = fm.level_cones(level)
        b2, q2 = om.level_cones(level)
        assert np.array_equal(b1, b2) and np.array_equal(q1, q2), level
    # work counts of one application, against the oracle's counters
    om.apply(np.ones(om.N, dtype=np.complex128))
    so, sp_ = om.stats(), fm.stats()
    for key in ("near", "leaf", "prop", "interp", "cones"):
        assert so[key] == sp_[key], key


@pytest.mark.parametrize("name,curve,npatch,k,nlev,F,Ps,Pang", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_correction_matches_oracle(name, curve, npatch, k, nlev, F, Ps, Pang):
    fm, om = pair(curve, npatch, k, nlev, F, Ps, Pang)
    Np = 10
    counts, idx, _ = om.near_singular()
    off = np.concatenate([[0], np.cumsum(counts)])
    fn = np.cos(np.pi * (2 * np.arange(Np) + 1) / (2 * Np))
    T = np.cos(np.outer(np.arange(Np), np.arccos(fn)))
    fold = (np.where(np.arange(Np) == 0, 1.0, 2.0) / Np)[:, None] * T
    ptr, patch, W = fm.correction()
    worst = 0.0
    for p in range(fm.npatches):
        Wo = om.precomputation(p, Np + counts[p]) @ fold   # [column][source node]
        ns = list(idx[off[p]:off[p + 1]])
        for col, i in enumerate(list(range(p * Np, (p + 1) * Np)) + ns):
            e = [e for e in range(ptr[i], ptr[i + 1]) if patch[e] == p]
            assert len(e) == 1
            worst = max(worst, np.abs(W[e[0]] - Wo[col]).max() / np.abs(Wo[col]).max())
    assert worst < 5e-12


def test_error_paths():
    lims = I.equal_parameter_lims(40)
    with pytest.raises(I.IFGFError, match="at least 3 levels"):
        I.ForwardMap(I.circle(), lims, 10.0, 2, host_only=True)
    with pytest.raises(I.IFGFError, match="Ps and Pang"):
        I.ForwardMap(I.circle(), lims, 10.0, 3, Ps=1, host_only=True)
    with pytest.raises(I.IFGFError, match="curve"):
        I.ForwardMap((7, [1.0]), lims, 10.0, 3, host_only=True)
    # too many levels: a patch's singular points leave the leaf neighbourhood
    with pytest.raises(I.IFGFError, match="neighbourhood"):
        I.ForwardMap(I.circle(), lims, 10.0, 9, host_only=True)
    # mixed precision needs the stored operators
    with pytest.raises(I.IFGFError, match="mixed precision needs the stored operators"):
        I.ForwardMap(I.circle(), lims, 10.0, 3, precision="mixed", on_the_fly=True, host_only=True)
    # the device sort validates its arguments before touching the GPU
    with pytest.raises(I.IFGFError, match="at least 3 levels"):
        I.morton_sort(np.zeros(4), np.zeros(4), 2)
    with pytest.raises(I.IFGFError, match="above 16"):
        I.morton_sort(np.zeros(4), np.zeros(4), 17)
    # compute on a host-only plan fails loudly (no CPU fallback)
    fm = I.ForwardMap(I.circle(), lims, 10.0, 3, host_only=True)
    with pytest.raises(I.IFGFError, match="host-only"):
        fm.compute_Ax_acc(np.ones(fm.N))


FIELD_TARGETS = {
    "ring": lambda rng, x, y: (1.5 * np.cos(rng.uniform(0, 2 * np.pi, 300)),
                               1.5 * np.sin(rng.uniform(0, 2 * np.pi, 300))),
    "grid": lambda rng, x, y: tuple(a.ravel() for a in np.meshgrid(np.linspace(-2, 2, 21),
                                                                    np.linspace(-2, 2, 21))),
    "points": lambda rng, x, y: (x, y),
    "far": lambda rng, x, y: (np.array([0.1, 9.0, -7.5]), np.array([0.2, 3.0, -8.0])),
}


@pytest.mark.parametrize("which", sorted(FIELD_TARGETS))
def test_field_tables_match_oracle(which):
    """The separate-target tree of ifgf_field_create (host tables of a host-only
    plan) against the oracle's: identical work counts per target set, and the
    default level count keeps the leaf size when the box grows."""
    fm = I.ForwardMap(I.circle(), I.equal_parameter_lims(80), 10.0, 4, host_only=True)
    x, y, nx, ny, w = fm.points()
    tx, ty = FIELD_TARGETS[which](np.random.default_rng(4), x, y)
    f = fm.field(tx, ty)
    st = f.stats()
    _, so = O.field_solve(x, y, nx, ny, np.ones(fm.N), tx, ty, f.nlevels, 10.0, O.COMBINED, 10.0)
    for key in ("near", "leaf", "prop", "interp", "cones"):
        assert st[key] == so[key], key
    assert st["targets"] == tx.size
    assert f.nlevels == {"ring": 5, "grid": 5, "points": 4, "far": 7}[which]
    with pytest.raises(I.IFGFError, match="host-only"):
        f.evaluate(np.ones(fm.N))


def test_field_error_paths():
    fm = I.ForwardMap(I.circle(), I.equal_parameter_lims(40), 10.0, 3, host_only=True)
    with pytest.raises(I.IFGFError, match="target count"):
        fm.field(np.zeros(0), np.zeros(0))
    with pytest.raises(I.IFGFError, match="at least 3 levels"):
        fm.field(np.ones(3), np.ones(3), nlevels=2)
    with pytest.raises(I.IFGFError, match="differ in length"):
        fm.field(np.ones(3), np.ones(2))
    with pytest.raises(I.IFGFError, match="Ps and Pang"):
        fm.field(np.ones(3), np.ones(3), Ps=17)
    with pytest.raises(KeyError):
        fm.field(np.ones(3), np.ones(3), precision="fp16")


def test_host_gmres_through_linear_operator():
    """ifgf_gmres_host: the reference's Householder GMRES driving a callback
    LinearOperator (LinearOperator.cpp:14-23), against a dense solve."""
    rng = np.random.default_rng(3)
    n = 80
    A = np.eye(n) * 2 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(4 * n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)

    def mapping(x, out):
        out[:] = A @ x

    op = I.LinearOperator(n, n, mapping)
    x, it, err, ok = I.gmres(op, b, restart=15, tol=1e-12, maxit=400)
    assert ok and err < 1e-12
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(x) < 1e-10
    # same iterates as the oracle's restatement (same algorithm, same data)
    xo, ito, erro, rc = O.gmres_dense(A, b, restart=15, tol=1e-12, maxit=400)
    assert ito == it and np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-12


def test_c_abi_from_plain_c(tmp_path):
    """The header compiles as C99 and the library links and runs from a C
    program (host-only plan, stats, sorting, correction rows, error codes)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = tmp_path / "c_abi_check"
    libdir = os.path.join(ROOT, "ifgf2d_amd")
    subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-O1", os.path.join(ROOT, "tests", "c_abi_check.c"),
                    "-I", os.path.join(ROOT, "include"), "-L", libdir, "-lifgf2d", f"-Wl,-rpath,{libdir}", "-lm",
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stdout + out.stderr
