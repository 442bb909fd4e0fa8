"""Local operator tables for the oracle (test infrastructure only).

Spin matrices are built from the angular-momentum formulas (basis ordered
m = S, S-1, ..., -S, i.e. "Up" first), fermion matrices from the occupation
basis, both returned in the *standard* (out-row, in-column) convention.  The
reference stores the transpose ((in,out), ``src/dmrgpy/pyitensor/sites/base.py:1-16``);
``stored(name)`` gives that form.  Entries match
``src/dmrgpy/pyitensor/sites/spin.py:15-129`` and ``sites/fermion.py:18-78``.
Site type codes follow ``src/dmrgpy/pyitensor/sites/siteset.py:20-41``.
"""

import numpy as np


def _spin_ops(twice_s):
    s = twice_s / 2.0
    dim = twice_s + 1
    m = s - np.arange(dim)            # m values, index 0 = Up
    sz = np.diag(m).astype(complex)
    sp = np.zeros((dim, dim), complex)  # sp[out, in]
    for k in range(1, dim):           # in state m[k] -> out m[k-1]
        sp[k - 1, k] = np.sqrt(s * (s + 1) - m[k] * (m[k] + 1))
    sm = sp.T.copy()
    sx = 0.5 * (sp + sm)
    sy = -0.5j * (sp - sm)
    ops = {"Sz": sz, "Sx": sx, "Sy": sy, "ISy": 1j * sy,
           "Sp": sp, "S+": sp, "Sm": sm, "S-": sm}
    return dim, ops


def _fermion_ops():
    # basis: Emp, Occ
    a = np.array([[0, 1], [0, 0]], complex)      # a[out, in]: Occ -> Emp
    ops = {"A": a, "C": a, "Adag": a.T.copy(), "Cdag": a.T.copy(),
           "N": np.diag([0, 1]).astype(complex), "n": np.diag([0, 1]).astype(complex),
           "F": np.diag([1, -1]).astype(complex)}
    return 2, ops


def _electron_ops():
    # basis: Emp, Up, Dn, UpDn  (indices 0..3)
    def mat(entries):
        m = np.zeros((4, 4), complex)
        for i_in, i_out, v in entries:
            m[i_out, i_in] = v
        return m
    aup = mat([(1, 0, 1), (3, 2, 1)])
    adn = mat([(2, 0, 1), (3, 1, 1)])
    cdn = mat([(2, 0, 1), (3, 1, -1)])
    ops = {
        "Aup": aup, "Cup": aup, "Adagup": aup.T.copy(), "Cdagup": aup.T.copy(),
        "Adn": adn, "Adagdn": adn.T.copy(), "Cdn": cdn, "Cdagdn": cdn.T.copy(),
        "Nup": np.diag([0, 1, 0, 1]).astype(complex),
        "Ndn": np.diag([0, 0, 1, 1]).astype(complex),
        "Nupdn": np.diag([0, 0, 0, 1]).astype(complex),
        "Ntot": np.diag([0, 1, 1, 2]).astype(complex),
        "F": np.diag([1, -1, -1, 1]).astype(complex),
        "Fup": np.diag([1, -1, 1, -1]).astype(complex),
        "Fdn": np.diag([1, 1, -1, -1]).astype(complex),
        "Sz": np.diag([0, 0.5, -0.5, 0]).astype(complex),
        "Sp": mat([(2, 1, 1)]), "S+": mat([(2, 1, 1)]),
        "Sm": mat([(1, 2, 1)]), "S-": mat([(1, 2, 1)]),
    }
    return 4, ops


_TABLES = {2: _spin_ops(1), 3: _spin_ops(2), 4: _spin_ops(3), 5: _spin_ops(4),
           6: _spin_ops(5), 0: _fermion_ops(), 1: _electron_ops()}


def dim(code):
    return _TABLES[code][0]


def op(code, name):
    """Standard-convention matrix op[out, in] of operator `name` on a site of
    type `code` (``SiteType.matrix``, ``sites/base.py:41-57``: "Id" and a
    generic identity "F" on non-fermionic sites)."""
    d, ops = _TABLES[code]
    if name == "Id":
        return np.eye(d, dtype=complex)
    if name in ops:
        return ops[name]
    if name in ("F", "FermiPhase"):
        return np.eye(d, dtype=complex)
    raise ValueError("operator {!r} not defined for site type {}".format(name, code))


def stored(code, name):
    """The reference's (in,out) storage form."""
    return op(code, name).T


def is_fermionic(name):
    """``sites/base.py:59-74``: names starting with 'C' are fermionic."""
    return bool(name) and name[0] == "C"
