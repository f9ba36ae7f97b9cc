"""The C++ adapter (include/lpm_gpu.hpp) keeps LPM-CPP's class names over the C
ABI: it compiles as C++17 against the header (CPU), and on the GPU it yields the
same displacements as the same problem driven through the Python binding --
through SolverStatic::solveProblem, as two concurrent requests to the serving
runtime (lpm_gpu::Server), and through SolverFatigue::solveProblemCyclic."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "sc_cube_adapter.cpp")
FATIGUE_SRC = os.path.join(ROOT, "tests", "cpp", "fatigue_adapter.cpp")
PKG = os.path.join(ROOT, "lpm-cpp_amd")


def _compile(out, src=SRC):
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           src, "-o", out, "-L", PKG, "-llpmg", "-Wl,-rpath," + PKG])


def test_adapter_compiles(tmp_path):
    import lpmg

    if not os.path.exists(lpmg.LIB_PATH):
        subprocess.check_call(["make", "-j8", "-C", lpmg.CSRC])
    _compile(str(tmp_path / "adapter"))
    _compile(str(tmp_path / "fatigue_adapter"), FATIGUE_SRC)


@pytest.mark.gpu
def test_adapter_matches_python_binding(tmp_path):
    import lpmg
    from lpmg import problems as P

    exe = str(tmp_path / "adapter")
    _compile(exe)
    out = subprocess.check_output([exe, "6", "3"], text=True)
    u_cpp = np.array([[float(v) for v in line.split()] for line in out.strip().splitlines()])
    prob = P.sc_cube(n=6, steps=3, total=0.05, E=70e3, nu=0.33)
    s, gids = P.build(lpmg.System, prob)
    s.forward()
    P.solve_problem(s, gids, P.Problem(**{**prob.__dict__, "tol": 1e-11}),
                    dict(precond=1, rtol=1e-24, atol=0.0))
    u_py = (s.get("XYZ") - s.get("XYZ_INITIAL")).reshape(-1, 3)
    assert np.linalg.norm(u_cpp - u_py) <= 1e-12 * np.linalg.norm(u_py)


@pytest.mark.gpu
def test_adapter_server_matches_python_binding(tmp_path):
    import lpmg
    from lpmg import problems as P

    exe = str(tmp_path / "adapter")
    _compile(exe)
    out = subprocess.check_output([exe, "6", "3", "serve"], text=True)
    u_srv = np.array([[float(v) for v in line.split()] for line in out.strip().splitlines()])
    prob = P.sc_cube(n=6, steps=3, total=0.05, E=70e3, nu=0.33)
    s, gids = P.build(lpmg.System, prob)
    s.forward()
    P.solve_problem(s, gids, P.Problem(**{**prob.__dict__, "tol": 1e-11}),
                    dict(precond=1, rtol=1e-24, atol=0.0))
    u_py = (s.get("XYZ") - s.get("XYZ_INITIAL")).reshape(-1, 3)
    np.testing.assert_array_equal(u_srv, u_py)


@pytest.mark.gpu
def test_adapter_fatigue_matches_python_binding(tmp_path):
    import lpmg
    from lpmg import fatigue as F
    from lpmg import problems as P

    exe = str(tmp_path / "fatigue_adapter")
    _compile(exe, FATIGUE_SRC)
    n_sim, tau = 3, 1e-4
    prob = P.fatigue_gage(hole=1.0, nonlocal_L=0.6, tau=tau, n_cycles=n_sim * 10000)
    prob.tol = 1e-11
    inp = tmp_path / "gage.txt"
    with open(inp, "w") as f:
        f.write(f"{len(prob.xyz)}\n")
        for p in prob.xyz:
            f.write(" ".join(repr(float(v)) for v in p) + "\n")
        f.write(" ".join(str(int(t)) for t in prob.types) + "\n")
        for name in ("left", "top", "bottom"):
            idx = prob.groups[name]
            f.write(f"{len(idx)} " + " ".join(str(int(i)) for i in idx) + "\n")
        f.write(f"{prob.undamaged} 0.6 {float(prob.extra['nonlocal_ratio'])!r} {tau!r} 0.043927 {n_sim}\n")
    out = subprocess.check_output([exe, str(inp)], text=True)
    res = np.array([[float(v) for v in line.split()] for line in out.strip().splitlines()])

    s, gids = P.build(lpmg.System, prob)
    s.search_nonlocal(prob.extra["nonlocal_ratio"])
    s.forward()
    fz = prob.extra["fatigue"]
    sf = F.SolverFatigue(s, gids, prob.undamaged, tau=tau, max_iter=30, tol=1e-11,
                         opts=dict(precond=1, rtol=1e-24, atol=0.0))
    sf.solve_static(prob.steps)
    sf.solve_cyclic(fz["ltype"], fz["spectrum"], 1, fz["groups"])
    u = (s.get("XYZ") - s.get("XYZ_INITIAL")).reshape(-1, 3)
    np.testing.assert_array_equal(res[:, :2], u[:, :2])
    np.testing.assert_array_equal(res[:, 2], s.get("DAMAGE"))
