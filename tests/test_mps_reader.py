"""MPS input (parsers::read_mps, reference src/parsers/mps_reader.cpp:65-355): the reference's
two fixtures (tests/unit/test_mps_reader.cpp:10-95) as known answers, a bitwise round trip of a
generated LP, every bound type, the vector-selection rules and the error cases.  Host code
only: runs without a GPU."""
import numpy as np
import pytest

INF = float("inf")

FIXTURE_1 = """NAME          TESTPROB
ROWS
 N  COST
 L  LIM1
 G  LIM2
 E  MYEQN
COLUMNS
    XONE      COST                 1   LIM1                 1
    XONE      LIM2                 1
    YTWO      COST                 4   LIM1                 1
    YTWO      MYEQN               -1
    ZTHREE    COST                 9   LIM2                 1
    ZTHREE    MYEQN                1
RHS
    RHS1      LIM1                 5   LIM2                10
    RHS1      MYEQN                7
BOUNDS
 UP BND1      XONE                 4
 LO BND1      YTWO                -1
 UP BND1      YTWO                 1
ENDATA
"""

FIXTURE_2 = """NAME INTPROB
ROWS
 N  OBJ
 L  C1
COLUMNS
    MARK0000  'MARKER'                 'INTORG'
    X1        OBJ                  1   C1                   1
    MARK0001  'MARKER'                 'INTEND'
    X2        OBJ                  2   C1                   2
RHS
    RHS1      C1                   5
BOUNDS
ENDATA
"""


def _read(native, tmp_path, text, name="t.mps"):
    p = tmp_path / name
    p.write_text(text)
    return native.read_mps(p)


def _dense(lp):
    return lp.to_scipy().toarray()


def test_reference_fixture_constraints_and_bounds(native, tmp_path):
    lp = _read(native, tmp_path, FIXTURE_1)
    assert (lp.m, lp.n) == (3, 5)                      # 3 structural + 2 slack columns
    np.testing.assert_array_equal(lp.c, [1, 4, 9, 0, 0])
    np.testing.assert_array_equal(lp.b, [5, 10, 7])
    np.testing.assert_array_equal(lp.meta["lb"], [0, -1, 0, 0, 0])
    np.testing.assert_array_equal(lp.meta["ub"], [4, 1, INF, INF, INF])
    np.testing.assert_array_equal(_dense(lp), [[1, 1, 0, 1, 0],       # LIM1 + slack
                                               [1, 0, 1, 0, -1],      # LIM2 - surplus
                                               [0, -1, 1, 0, 0]])     # MYEQN
    np.testing.assert_array_equal(lp.meta["vtype"], [0, 0, 0, 0, 0])


def test_reference_fixture_integer_markers(native, tmp_path):
    lp = _read(native, tmp_path, FIXTURE_2)
    np.testing.assert_array_equal(lp.meta["vtype"], [1, 0, 0])        # X1 integer, X2, slack
    np.testing.assert_array_equal(_dense(lp), [[1, 2, 1]])


def _write_mps(lp, rhs_name="RHS"):
    A = lp.to_scipy().tocsc()
    out = ["NAME GEN", "ROWS", " N  OBJ"] + [f" E  R{i}" for i in range(lp.m)] + ["COLUMNS"]
    for j in range(lp.n):
        out.append(f"    X{j}  OBJ  {float(lp.c[j])!r}")
        for p in range(A.indptr[j], A.indptr[j + 1]):
            out.append(f"    X{j}  R{A.indices[p]}  {float(A.data[p])!r}")
    out.append("RHS")
    out += [f"    {rhs_name}  R{i}  {float(lp.b[i])!r}" for i in range(lp.m)]
    out.append("ENDATA")
    return "\n".join(out) + "\n"


def test_generated_lp_round_trip_is_bitwise(native, tmp_path):
    from lakshya import gen
    lp = gen.random_sparse(30, 80, 4, seed=3)
    back = _read(native, tmp_path, _write_mps(lp))
    assert (back.m, back.n) == (lp.m, lp.n)
    A = lp.to_scipy().tocsc()
    A.sort_indices()
    np.testing.assert_array_equal(back.indptr, A.indptr)
    np.testing.assert_array_equal(back.indices, A.indices)
    assert np.array_equal(back.data, A.data) and np.array_equal(back.b, lp.b) and np.array_equal(back.c, lp.c)


def test_bound_types(native, tmp_path):
    text = """NAME B
ROWS
 N  OBJ
 E  R
COLUMNS
    A  R  1
    B  R  1
    C  R  1
    D  R  1
    E  R  1
    F  R  1
    G  R  1
    H  R  1
RHS
    RHS  R  1
BOUNDS
 FR BND  A
 MI BND  B
 UP BND  B  3
 PL BND  C
 BV BND  D
 LI BND  E  2
 UI BND  F  7
 FX BND  G  2.5
 LO BND  H  -4
ENDATA
"""
    lp = _read(native, tmp_path, text)
    np.testing.assert_array_equal(lp.meta["lb"], [-INF, -INF, 0, 0, 2, 0, 2.5, -4])
    np.testing.assert_array_equal(lp.meta["ub"], [INF, 3, INF, 1, INF, 7, 2.5, INF])
    np.testing.assert_array_equal(lp.meta["vtype"], [0, 0, 0, 2, 1, 1, 0, 0])


def test_first_vectors_only_extra_objective_rows_dropped_last_cost_wins(native, tmp_path):
    text = """NAME V
ROWS
 N  OBJ
 N  FREE
 G  R1
COLUMNS
    X  OBJ  1  R1  2
    X  FREE  9
    X  OBJ  5
RHS
    RHSA  R1  3
    RHSB  R1  8
BOUNDS
 UP BA  X  4
 UP BB  X  6
ENDATA
"""
    lp = _read(native, tmp_path, text)
    np.testing.assert_array_equal(lp.c, [5, 0])
    np.testing.assert_array_equal(lp.b, [3])
    np.testing.assert_array_equal(lp.meta["ub"], [4, INF])
    np.testing.assert_array_equal(_dense(lp), [[2, -1]])


@pytest.mark.parametrize("text,needle", [
    ("NAME X\nROWS\n N OBJ\n E R\nRANGES\n    RNG  R  1\nENDATA\n", "unsupported section"),
    ("NAME X\nROWS\n N OBJ\n X R\nENDATA\n", "unknown row type"),
    ("NAME X\nROWS\n N OBJ\n E R\nCOLUMNS\n    X  Q  1\nENDATA\n", "unknown row"),
    ("NAME X\nROWS\n N OBJ\n E R\nCOLUMNS\n    X  R  1  OBJ\nENDATA\n", "missing value"),
    ("NAME X\nROWS\n N OBJ\n E R\nCOLUMNS\n    X  R  abc\nENDATA\n", "invalid numeric"),
    ("NAME X\nROWS\n N OBJ\n E R\nCOLUMNS\n    X  R  1\nRHS\n    RHS  OBJ  2\nENDATA\n", "objective offset"),
    ("NAME X\nROWS\n N OBJ\n E R\nCOLUMNS\n    X  R  1\nBOUNDS\n XX BND  X  1\nENDATA\n", "unsupported bound type"),
    ("NAME X\nROWS\n N OBJ\n E R\nCOLUMNS\n    X  R  1\nBOUNDS\n UP BND  Y  1\nENDATA\n", "unknown column"),
])
def test_errors(native, tmp_path, text, needle):
    with pytest.raises(RuntimeError, match=needle):
        _read(native, tmp_path, text)


def test_missing_file(native, tmp_path):
    with pytest.raises(RuntimeError, match="could not open"):
        native.read_mps(tmp_path / "absent.mps")
