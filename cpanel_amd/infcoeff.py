"""The reference's `.infCoeff` influence-coefficient cache (text).

Layout, as `geometry::writeInfCoeff` writes it and `geometry::readInfCoeff`
reads it (`geometry.cpp:1125-1202`):

    n                                   number of body panels
    A(i,0) \\t A(i,1) \\t ... A(i,n-1) \\t   one line per row i (n lines)
    B(i,0) \\t ...                        one line per row i (n lines)
    [vortex-particle geometries only]
    n_w2                                panels of the second buffer-wake row
    C(i,0) \\t ... C(i,n_w2-1) \\t         one line per row i (n lines)

Values are written the way a default `std::ofstream` writes a double: six
significant digits, fixed or exponent form whichever is shorter (printf
`%g`), so a file written at the default precision holds A and B to about
5e-7 relative; `precision` raises that (17 digits round-trip exactly).
`geometry::infCoeffFileExists` (`geometry.cpp:1106-1123`) accepts a file whose
first number equals the panel count: `matches` is that test.

One deviation: the reference's loops over C run the column index up to n
rather than n_w2 (`geometry.cpp:1158-1164, 1192-1198`), so the C section of its
vortex-particle files is not n x n_w2 data; this module writes and reads C as
n rows of n_w2 values, so steady files interoperate and vortex-particle files
do so in their A and B sections.

`Geometry.load_matrices` (C ABI `cp_load_matrices`) puts A and B read here
on the device in place of the FP64 assembly.
"""
import numpy as np


def _write_rows(f, M, fmt):
    np.savetxt(f, np.asarray(M, dtype=np.float64), fmt=fmt, delimiter="\t", newline="\t\n")


def write_infcoeff(path, A, B, C=None, precision=6):
    """Write A and B (and C for a vortex-particle geometry) in the reference's
    layout; precision = significant digits (6: what the reference writes)."""
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    n = A.shape[0]
    if A.shape != (n, n) or B.shape != (n, n):
        raise ValueError("A and B must be square and of one size")
    fmt = "%%.%dg" % int(precision)
    with open(path, "w") as f:
        f.write("%d\n" % n)
        _write_rows(f, A, fmt)
        _write_rows(f, B, fmt)
        if C is not None:
            C = np.asarray(C, dtype=np.float64)
            if C.ndim != 2 or C.shape[0] != n:
                raise ValueError("C must have n rows")
            f.write("%d\n" % C.shape[1])
            _write_rows(f, C, fmt)


def matches(path, n_body):
    """`geometry::infCoeffFileExists`: the file's first number is the number
    of body panels."""
    try:
        with open(path) as f:
            first = f.readline().split()
        return bool(first) and int(first[0]) == int(n_body)
    except (OSError, ValueError):
        return False


def read_infcoeff(path, vp=False):
    """Read (A, B, C) from a .infCoeff file; C is None unless vp (the file then
    carries the second buffer-wake row's C after B)."""
    vals = np.fromfile(path, sep=" ")  # any whitespace separates values
    if vals.size == 0:
        raise ValueError("%s: empty influence-coefficient file" % path)
    n = int(vals[0])
    need = 1 + 2 * n * n
    if n <= 0 or vals.size < need:
        raise ValueError("%s: %d values, expected at least %d for n = %d" % (path, vals.size, need, n))
    A = vals[1:1 + n * n].reshape(n, n)
    B = vals[1 + n * n:need].reshape(n, n)
    C = None
    if vp:
        if vals.size <= need:
            raise ValueError("%s: no C section" % path)
        m = int(vals[need])
        if vals.size < need + 1 + n * m:
            raise ValueError("%s: C section shorter than %d x %d" % (path, n, m))
        C = vals[need + 1:need + 1 + n * m].reshape(n, m)
    elif vals.size != need:
        raise ValueError("%s: %d values after the panel count, expected %d (a vortex-particle file?)"
                         % (path, vals.size - 1, 2 * n * n))
    return np.ascontiguousarray(A), np.ascontiguousarray(B), C
