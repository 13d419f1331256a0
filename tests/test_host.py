"""Host-side input generation (geometry, grid) and the C ABI surface, on CPU."""
import os
import re
import subprocess

import numpy as np
import pytest

from aero2d_amd import _abi, growth_ratio, naca4, naca4_params, ogrid, self_similar_spacing

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_naca4_closed_form_properties():
    af = naca4("0012", 65)
    assert af.n == 129
    assert af.x[0] == 1.0 and af.x[-1] == 1.0 and af.y[0] == 0.0 and af.y[-1] == 0.0
    assert abs(af.x[64]) < 1e-15 and abs(af.y[64]) < 1e-15          # leading edge
    assert np.allclose(af.y[:64], -af.y[::-1][:64], atol=1e-15)    # symmetric section
    # max thickness of a 4-digit section: 0.12 at x ~ 0.30 chord (NACA TR-460)
    t = 2 * np.max(af.y)
    assert abs(t - 0.12) < 1.5e-3
    cam = naca4("2412", 101)
    assert np.max((cam.y[:101][::-1] + cam.y[100:]) / 2) == pytest.approx(0.02, abs=1e-3)


def test_naca4_params_reduces_to_digit_code():
    for code in ("0012", "2412", "4415"):
        a = naca4(code, 41)
        b = naca4_params(int(code[0]) / 100, int(code[1]) / 10, int(code[2:]) / 100, 41)
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_growth_ratio_closed_form():
    for ds0, tot, n in ((2e-3, 40.0, 48), (4.4965e-3, 20.0, 64), (1e-3, 20.0, 256)):
        r = growth_ratio(ds0, tot, n)
        assert abs(ds0 * (r ** n - 1) / (r - 1) - tot) < 1e-9 * tot


def test_self_similar_spacing_table():
    assert self_similar_spacing(64) == pytest.approx(4.4965e-3, rel=1e-4)
    assert self_similar_spacing(128) == pytest.approx(2.1947e-3, rel=1e-4)
    assert self_similar_spacing(256) == pytest.approx(1.0843e-3, rel=1e-4)


def test_ogrid_invariants_and_inverted_cells():
    g = ogrid(naca4("0012", 65), nj=49, ds0=2e-3, radius=40.0)
    assert g.shape == (128, 49)
    assert g.cell_areas().min() > 0
    r = np.hypot(g.x[:, -1] - 0.5, g.y[:, -1])
    assert np.ptp(r) < 1e-12 * r.mean()
    d1 = np.hypot(g.x[:, 1] - g.x[:, 0], g.y[:, 1] - g.y[:, 0])
    assert abs(np.median(d1) / 2e-3 - 1) < 0.1
    with pytest.raises(RuntimeError):
        ogrid(naca4("0012", 33), nj=129, ds0=1e-5, radius=20.0)


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "aero2d_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|int64_t|void|const char\*)\s+(a2d_\w+)\(", hdr, re.M))
    assert len(declared) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (a2d_\w+)", out))
    assert declared <= exported, declared - exported
    assert declared == set(_abi.SIGNATURES), declared ^ set(_abi.SIGNATURES)
    _abi.load()


def test_knobs_struct_matches_the_header():
    """The ctypes mirror of a2d_knobs has the header's fields, types and order
    (a field added on one side only would shift every later knob)."""
    hdr = open(os.path.join(ROOT, "include", "aero2d_hip.h")).read()
    body = re.search(r"typedef struct a2d_knobs \{(.*?)\} a2d_knobs;", hdr, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"^\s*(double|int)\s+(\w+);", body, re.M)
    ctypes_of = {"double": "c_double", "int": "c_int"}
    mirror = [(n, t.__name__) for n, t in _abi.Knobs._fields_]
    assert mirror == [(n, ctypes_of[t]) for t, n in fields]


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "aero2d_amd")
    for f in os.listdir(pkg):
        if f.endswith(".py"):
            src = open(os.path.join(pkg, f)).read()
            assert not re.search(r"^\s*(from|import)\s+oracle", src, re.M), f


def test_case_footprint_scales_with_size_and_precision():
    a64 = _abi.case_bytes(256, 128, 0)
    a32 = _abi.case_bytes(256, 128, 1)
    assert 0.45 < a32 / a64 < 0.65
    assert _abi.case_bytes(512, 256, 0) > 3.9 * a64


def test_newton_loop_reports_each_case_first_failure_once():
    """The mixed form lowers a case's step limit after its first rejected step:
    the host loop reports that case once, again only after a reset."""
    from aero2d_amd.implicit import NewtonLoop
    loop = NewtonLoop(3, 10.0, 300.0, fallback=True)
    out = np.zeros((3, 10))
    out[:, 0], out[:, 1], out[:, 2] = 1.0, 1.0, 1.0       # rn, omega, accepted
    out[1, 0] = np.nan                                    # case 1 rejected
    loop.update(np.arange(3), out)
    assert loop.newly_failed.tolist() == [1]
    assert loop.cfl[1] == 1.0 and loop.fails[1] == 1
    loop.update(np.arange(3), out)
    assert loop.newly_failed.tolist() == [] and loop.fails[1] == 2
    loop.reset([1])
    loop.update(np.arange(3), out)
    assert loop.newly_failed.tolist() == [1]
    plain = NewtonLoop(3, 10.0, 300.0)
    plain.update(np.arange(3), out)
    assert plain.newly_failed.tolist() == []


def test_censored_latency_counts_unfinished_requests():
    """Latency percentiles include unfinished requests at wall - t_arrive and
    flag a percentile that a censored value reaches."""
    from types import SimpleNamespace
    from aero2d_amd.serve import censored_latency
    done = [SimpleNamespace(latency=v) for v in (1.0, 2.0, 3.0)]
    out = censored_latency(done, [5.0, 9.0], wall=10.0)       # censored at 5 s and 1 s
    assert out["released"] == 5 and out["completed"] == 3
    assert out["completed_fraction"] == 0.6
    assert np.isclose(out["p50_ms"], 2000.0)                  # sorted 1, 1c, 2, 3, 5c
    assert out["censored_at"] == [50, 99]
    out = censored_latency(done, [], wall=10.0)
    assert out["censored_at"] == [] and np.isclose(out["p50_ms"], 2000.0)
    out = censored_latency(done, [0.0], wall=10.0)            # censored at 10 s: above p50
    assert out["censored_at"] == [99]


def test_rans_cfl_formula():
    """The RANS CFL expression (cfl_formula): the reference's default grows
    5 x 1.2^n to 1000; variables, functions, where(c, a, b) (and its name in
    the reference) and '^' evaluate; unknown names, attribute access, other
    calls, syntax errors and non-positive values raise ValueError; a formula
    reading cfl or dres is recursive."""
    import math
    from aero2d_amd.rans import cfl_formula, cfl_formula_recursive
    f = "min(5*1.2**n, 1000)"
    assert cfl_formula(f, n=0) == 5.0 and abs(cfl_formula(f, n=3) - 5 * 1.2 ** 3) < 1e-12
    assert cfl_formula(f, n=100) == 1000.0
    assert abs(cfl_formula("10^orders", orders=2.5) - 10 ** 2.5) < 1e-9
    assert cfl_formula("where(n < 10, 5, 50)", n=3) == 5 and cfl_formula("where(n < 10, 5, 50)", n=30) == 50
    assert cfl_formula("\u0433\u0434\u0435(res > 0.1, 2, 20)", res=0.01) == 20
    v = cfl_formula("max(cfl * min(2, 1/dres), 1) + sqrt(4) + log10(100) + tanh(0) + pi - e", cfl=10, dres=0.5)
    assert abs(v - (20 + 2 + 2 + 0 + math.pi - math.e)) < 1e-12
    for bad in ("n +", "foo(n)", "n.real", "__import__('os')", "(lambda: 1)()", "[1][0]", "'a'", "-n"):
        try:
            cfl_formula(bad, n=1)
        except ValueError:
            continue
        raise AssertionError(bad)
    assert cfl_formula_recursive("cfl*1.1") and cfl_formula_recursive("min(1/dres, 2)")
    assert not cfl_formula_recursive(f) and not cfl_formula_recursive("n +")
