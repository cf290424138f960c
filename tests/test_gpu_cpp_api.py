"""The C++ class API (include/pdmk/periodic_dmk.hpp) driven by a C++ caller
(tests/cpp/api_check.cpp, built by __graft_entry__.build()), on the reference's own
end-to-end cases (cpp/tests/test_periodic_dmk.cpp:20-68: a cubic cell of side 3 with 200
points and a triclinic cell with 100 points, n_per_leaf 20, nz 4, neutral Gaussian charges,
rel-L2 against the Ewald sum below 1e-4).  The C++ API runs the same plan as the C ABI, so
their potentials agree bit for bit; evaluate_timed runs the stages one after another and
agrees to rounding; the reference's exception cases hold.
"""
import ctypes as C
import os

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HELPER = os.path.join(REPO, "tests", "cpp", "_build", "libapi_check.so")


def load_helper():
    if not os.path.exists(HELPER):
        raise RuntimeError(f"{HELPER} is not built; run __graft_entry__.build()")
    h = C.CDLL(HELPER)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    h.api_check.restype = C.c_int
    h.api_check.argtypes = [dp, C.c_int, dp, dp, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, ip]
    return h


def test_helper_exports_entry(pdmk):
    """The C++ caller compiles against include/ and links libpdmk.so (built here if missing)."""
    if not os.path.exists(HELPER):
        import __graft_entry__
        __graft_entry__.build_api_check()
    assert hasattr(load_helper(), "api_check")


def reference_case(name):
    rng = np.random.default_rng(42 if name == "cubic" else 77)
    if name == "cubic":
        cell = np.eye(3) * 3.0
        pos = rng.uniform(0.03 * 3, 0.97 * 3, (200, 3))
    else:  # a1 = (5, 0, 0), a2 = (1, 4.5, 0), a3 = (0.5, 0.8, 4)
        cell = np.array([[5.0, 1.0, 0.5], [0.0, 4.5, 0.8], [0.0, 0.0, 4.0]])
        pos = rng.uniform(0.01, 0.99, (100, 3)) @ cell.T
    q = rng.normal(size=pos.shape[0])
    q[0] -= q.sum()
    return cell, pos, q


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cubic", "triclinic"])
@pytest.mark.parametrize("digits", [6, 9])
def test_cpp_api_reference_cases(pdmk, cuda, name, digits):
    h = load_helper()
    cell, pos, q = reference_case(name)
    n = pos.shape[0]
    cellc = np.ascontiguousarray(cell.T)  # column j = lattice vector j
    posc = np.ascontiguousarray(pos)
    tree, timed, auto, t3 = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(3)
    flags = np.zeros(5, dtype=np.int32)
    d = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    rc = h.api_check(d(cellc), n, d(posc), d(q), digits, 20, 4, d(tree), d(timed), d(auto), d(t3),
                     flags.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc == 0
    ew, _ = pdmk.ewald(cell, pos, q, 5.0)
    err = np.linalg.norm(tree - ew) / np.linalg.norm(ew)
    assert err < (1e-4 if digits == 6 else 1e-7), err
    # the C ABI on the same inputs runs the same plan
    s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=20, nz=4,
                        precision="float" if digits <= 6 else "double")
    np.testing.assert_array_equal(s.evaluate(q).astype(np.float64), tree)
    s.close()
    np.testing.assert_array_equal(auto, tree)  # make_periodic_dmk_tree picks the same precision
    tol = 1e-5 if digits <= 6 else 1e-12
    assert np.linalg.norm(timed - tree) / np.linalg.norm(tree) < tol
    assert np.all(t3 >= 0) and t3.sum() > 0
    # n_points, FP32 for <= 6 digits, invalid_argument (size), logic_error (evaluate_complex
    # without bloch), logic_error (evaluate with bloch)
    assert list(flags) == [n, 1 if digits <= 6 else 0, 1, 1, 1]
