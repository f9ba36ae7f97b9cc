"""The C++ adapter (include/lpm_gpu.hpp) keeps LPM-CPP's class names over the C
ABI: it compiles as C++17 against the header (CPU), and on the GPU it yields the
same displacements as the same problem driven through the Python binding --
through SolverStatic::solveProblem and as two concurrent requests to the
serving runtime (lpm_gpu::Server)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "sc_cube_adapter.cpp")
PKG = os.path.join(ROOT, "lpm-cpp_amd")


def _compile(out):
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           SRC, "-o", out, "-L", PKG, "-llpmg", "-Wl,-rpath," + PKG])


def test_adapter_compiles(tmp_path):
    import lpmg

    if not os.path.exists(lpmg.LIB_PATH):
        subprocess.check_call(["make", "-j8", "-C", lpmg.CSRC])
    _compile(str(tmp_path / "adapter"))


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
