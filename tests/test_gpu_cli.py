"""The drop-in driver (sc23ad_run) against the library and the reference's
log format.

The reference's drivers print, per run (Work/main.cpp:101,176-180;
Library/solver/gmres.hpp:253-259): one "# [j] <implicit relres>" line per
Arnoldi step, "[Res] : <implicit relres>" on convergence, "[Time] : <s>",
"[Iter] : <iterations, 0 when not converged>" and "# [Res] : <explicit
relres>".  Work/Artwork/table.py:5-21 reads them by whitespace-split fields:
a line whose fields 0 and 1 are "#" and "[Res]" gives the explicit residual
from field 3; "[Time]" gives the time from field 2; "[Iter]" the iteration
count from field 2; "[Res]" the implicit residual from field 2.  The parser
below restates those rules; the values must equal what the library returns
for the same system and options.  The argv contract is Work/main.cpp:35-47
(<matrix> <smoother> <solver> <m>) and Work/sub.cpp:38-39,89 (<matrix>
<smoother> <m> <bit>, here behind --sub).
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import sc23ad as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "sc23-ad_amd", "bin", "sc23ad_run")


def parse_log(text: str):
    """The fields table.py's reader takes from a log (time, iterations,
    implicit residual, explicit residual), plus the per-step history."""
    time = itr = imp = exp = None
    hist = []
    for line in text.splitlines():
        li = line.split()
        if not li:
            continue
        if li[0] == "#" and li[1] == "[Res]":
            exp = float(li[3])
        elif li[0] == "#" and li[1].startswith("[") and li[1].endswith("]"):
            hist.append((int(li[1][1:-1]), float(li[2])))
        if li[0] == "[Time]":
            time = float(li[2])
        elif li[0] == "[Iter]":
            itr = int(li[2])
        elif li[0] == "[Res]":
            imp = float(li[2])
    return time, itr, imp, exp, hist


def run_cli(*args, timeout=300):
    print("[cli]", " ".join(args), flush=True)
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


def _library_solve(kind, params, seed, smoother, solver, m, eps, bit=None):
    sy = S.make_system(kind, params, seed)
    prm = S.amg_param(smoother)
    prec = solver
    if prec == S.INT:
        hi, hd = S.Hierarchy.build_multi(sy.a, [S.INT, S.FP64], prm)
        if bit is None:
            bit = S.Amg(hd).default_frac_bits(sy.b, sy.nrm_b)
        sol = S.Solver(S.INT, sy.a, S.Amg(hi), m=m, eps=eps, bit=bit)
    else:
        h = S.Hierarchy(sy.a, prec, prm)
        sol = S.Solver(prec, sy.a, S.Amg(h), m=m, eps=eps)
    x, cv = sol.solve(sy.b, sy.nrm_b)
    return cv, sol.history(), sol.relres(sy.b, sy.nrm_b, x), bit


def _check(out, cv, hist, rel, m):
    assert out.returncode == 0, out.stderr
    t, itr, imp, exp, lines = parse_log(out.stdout)
    assert t is not None and t > 0
    assert itr == (cv.iter if cv.is_converged else 0)
    assert exp == float(f"{rel:e}")
    if cv.is_converged:
        assert imp == float(f"{cv.implicit_relres:e}")
    assert len(lines) == len(hist)
    # "# [j]" counts 1..m inside each restart cycle
    assert [j for j, _ in lines] == [k % m + 1 for k in range(len(hist))]
    assert [v for _, v in lines] == [float(f"{h:e}") for h in hist]


@pytest.mark.parametrize("smoother,solver", [(1, 2), (0, 2), (1, 0), (1, 1)],
                         ids=["blocksor_int", "sor_int", "blocksor_fp64", "blocksor_fp32"])
def test_main_driver_lines_match_library(smoother, solver):
    out = run_cli("gen:laplace3d:16", str(smoother), str(solver), "30", "--eps", "1e-10")
    cv, hist, rel, bit = _library_solve(S.GEN_LAPLACE3D, (16,), 0, smoother, solver, 30, 1e-10)
    _check(out, cv, hist, rel, 30)
    if solver == 2:
        assert f"[Bit] : {bit}" in out.stdout.splitlines()


def test_sub_driver_takes_the_bit_from_argv():
    out = run_cli("--sub", "gen:poisson2d:48", "1", "10", "27", "--eps", "1e-9")
    cv, hist, rel, _ = _library_solve(S.GEN_POISSON2D, (48,), 0, 1, S.INT, 10, 1e-9, bit=27)
    _check(out, cv, hist, rel, 10)
    assert "[Bit] : 27" in out.stdout.splitlines()


def test_not_converged_prints_zero_iterations():
    # FP64 FGMRES(2) cannot reach 1e-300: once the residual stops changing
    # Restarted gives up (irsolver.hpp:35), so [Iter] is 0 and no "[Res] :"
    # line is printed
    out = run_cli("gen:laplace3d:12", "1", "0", "2", "--eps", "1e-300")
    assert out.returncode == 0, out.stderr
    _, itr, imp, exp, _ = parse_log(out.stdout)
    assert itr == 0 and imp is None and exp is not None


@pytest.mark.parametrize("argv", [[], ["gen:laplace3d:8", "1", "2"],
                                  ["gen:laplace3d:8", "7", "2", "30"],
                                  ["gen:laplace3d:8", "1", "2", "30", "--bogus"]],
                         ids=["none", "three", "bad_smoother", "bad_option"])
def test_bad_argv_prints_usage(argv):
    out = run_cli(*argv)
    assert out.returncode == 2 and "usage" in out.stderr


def test_unreadable_matrix_fails_cleanly():
    out = run_cli("/nonexistent/matrix.mtx", "1", "2", "30")
    assert out.returncode == 1 and "reading the matrix" in out.stderr
    out = run_cli("gen:nosuchkind:8", "1", "2", "30")
    assert out.returncode == 1 and "generator" in out.stderr


def test_parser_restates_table_rules():
    """The restated parse on a hand-written log of the reference's format."""
    log = ("[Bit] : 29\n# [1] 1.000000e-01\n# [2] 1.000000e-05\n[Res] : 5.000000e-11\n"
           "[Time] : 0.125000000000000\n[Iter] : 2\n# [Res] : 4.500000e-11\n")
    assert parse_log(log) == (0.125, 2, 5e-11, 4.5e-11, [(1, 0.1), (2, 1e-5)])
    assert np.isfinite(parse_log(log)[0])
