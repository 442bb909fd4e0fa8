"""Local Hilbert spaces and operator matrices of the engine's site types.

Type codes follow the session protocol (``Chain(site_types)``,
``src/dmrgpy/pyitensor/sites/siteset.py:20-41``): 2, 3, 4, 5, 6 = spin 1/2, 1,
3/2, 2, 5/2; 0 = spinless fermion (Emp, Occ); 1 = electron (Emp, Up, Dn,
UpDn).  Matrices are returned in the standard (out-row, in-column)
convention; spin matrices come from the angular-momentum formulas in the
basis m = S, S-1, ..., -S, fermion matrices from the occupation basis, with
the same entries and signs as the reference tables
(``sites/spin.py:15-129``, ``sites/fermion.py:18-78``).
"""

import numpy as np

SPIN_CODES = {2: 1, 3: 2, 4: 3, 5: 4, 6: 5}   # code -> 2S


def _spin(twice_s):
    s = twice_s / 2.0
    d = twice_s + 1
    m = s - np.arange(d)
    sp = np.zeros((d, d))
    for k in range(1, d):
        sp[k - 1, k] = np.sqrt(s * (s + 1) - m[k] * (m[k] + 1))
    sm = sp.T.copy()
    sy = -0.5j * (sp - sm)
    return d, {"Sz": np.diag(m).astype(complex), "Sx": (0.5 * (sp + sm)).astype(complex),
               "Sy": sy, "ISy": 1j * sy, "Sp": sp.astype(complex), "S+": sp.astype(complex),
               "Sm": sm.astype(complex), "S-": sm.astype(complex)}


def _spinless():
    a = np.array([[0, 1], [0, 0]], dtype=complex)
    n = np.diag([0.0, 1.0]).astype(complex)
    return 2, {"A": a, "C": a, "Adag": a.T.copy(), "Cdag": a.T.copy(), "N": n, "n": n,
               "F": np.diag([1.0, -1.0]).astype(complex)}


def _electron():
    def m(entries):
        out = np.zeros((4, 4), dtype=complex)
        for i_in, i_out, v in entries:
            out[i_out, i_in] = v
        return out
    aup = m([(1, 0, 1), (3, 2, 1)])
    adn = m([(2, 0, 1), (3, 1, 1)])
    cdn = m([(2, 0, 1), (3, 1, -1)])
    d = lambda *v: np.diag(v).astype(complex)
    return 4, {
        "Aup": aup, "Cup": aup, "Adagup": aup.T.copy(), "Cdagup": aup.T.copy(),
        "Adn": adn, "Adagdn": adn.T.copy(), "Cdn": cdn, "Cdagdn": cdn.T.copy(),
        "Nup": d(0, 1, 0, 1), "Ndn": d(0, 0, 1, 1), "Nupdn": d(0, 0, 0, 1), "Ntot": d(0, 1, 1, 2),
        "F": d(1, -1, -1, 1), "Fup": d(1, -1, 1, -1), "Fdn": d(1, 1, -1, -1),
        "Sz": d(0, 0.5, -0.5, 0), "Sp": m([(2, 1, 1)]), "S+": m([(2, 1, 1)]),
        "Sm": m([(1, 2, 1)]), "S-": m([(1, 2, 1)]),
    }


_TABLE = {code: _spin(t) for code, t in SPIN_CODES.items()}
_TABLE[0] = _spinless()
_TABLE[1] = _electron()


def dim(code):
    if code not in _TABLE:
        raise ValueError("site type code {} is not supported by the MI355X backend".format(code))
    return _TABLE[code][0]


def matrix(code, name):
    """Standard-convention matrix [out, in] of operator ``name``."""
    d, ops = _TABLE[code]
    if name == "Id":
        return np.eye(d, dtype=complex)
    if name in ops:
        return ops[name]
    if name in ("F", "FermiPhase"):       # identity string on non-fermionic sites
        return np.eye(d, dtype=complex)
    raise ValueError("operator '{}' not recognized for site type {}".format(name, code))


def is_fermionic(name):
    """Names starting with 'C' carry a Jordan-Wigner string
    (``sites/base.py:59-74``)."""
    return bool(name) and name[0] == "C"
