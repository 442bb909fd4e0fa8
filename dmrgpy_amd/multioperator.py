"""Operator algebra of the front end: sums of products of named local
operators, and their wire format for the session.

``MultiOperator`` supports +, -, scalar and operator products, dagger, and
``to_terms()`` -> ``[(coef, [(name, site_1based), ...]), ...]`` with the
Jordan-Wigner transform applied and |coef| <= 1e-8 dropped
(src/dmrgpy/multioperator.py:22-33, 158-184, 331-386).  Sites are 0-based
on the Python side as in the reference.

Jordan-Wigner (src/dmrgpy/multioperatortk/jordanwigner.py,
jordanwigner_spinful.py): every fermion factor is dressed on its own,
c_i = (prod_{k<i} F_k) a_i for spinless "C"/"Cdag", and for the native
spinful site the down flavour also carries the on-site "Fup"; all other
factors pass through.
"""

import numbers

import numpy as np

CLEAN_THRESHOLD = 1e-8

_SPINLESS = {"C": "A", "Cdag": "Adag"}
_SPINFUL = {"Cup": ("Aup", 0), "Cdagup": ("Adagup", 0), "Cdn": ("Adn", 1), "Cdagdn": ("Adagdn", 1)}
_DAGGER = {"C": "Cdag", "Cdag": "C", "A": "Adag", "Adag": "A", "Cup": "Cdagup", "Cdagup": "Cup",
           "Cdn": "Cdagdn", "Cdagdn": "Cdn", "Aup": "Adagup", "Adagup": "Aup", "Adn": "Adagdn",
           "Adagdn": "Adn", "Sp": "Sm", "S+": "S-", "Sm": "Sp", "S-": "S+"}


def _isnum(x):
    return isinstance(x, numbers.Number) or isinstance(x, np.number)


class MultiOperator:
    __slots__ = ("terms",)

    def __init__(self, terms=None):
        self.terms = list(terms) if terms else []     # [(coef, ((name, site0), ...)), ...]

    def copy(self):
        return MultiOperator(self.terms)

    def __add__(self, other):
        if other is None:
            return self.copy()
        if _isnum(other):
            return self + other * identity()
        if isinstance(other, MultiOperator):
            return MultiOperator(self.terms + other.terms)
        return NotImplemented

    __radd__ = __add__

    def __neg__(self):
        return (-1) * self

    def __sub__(self, other):
        return self + (-1) * other

    def __rsub__(self, other):
        return (-1) * self + other

    def __mul__(self, other):
        if _isnum(other):
            return MultiOperator([(c * other, f) for c, f in self.terms])
        if isinstance(other, MultiOperator):
            return MultiOperator([(c1 * c2, f1 + f2) for c1, f1 in self.terms for c2, f2 in other.terms])
        return NotImplemented

    def __rmul__(self, other):
        if _isnum(other):
            return self * other
        return NotImplemented

    def __truediv__(self, other):
        return self * (1.0 / other)

    def get_dagger(self):
        out = []
        for c, f in self.terms:
            fd = []
            for name, site in reversed(f):
                if name in ("Sy",):
                    fd.append((name, site))
                else:
                    fd.append((_DAGGER.get(name, name), site))
            out.append((np.conj(c), tuple(fd)))
        return MultiOperator(out)

    def to_terms(self, jordan_wigner_transform=True):
        out = []
        for c, f in self.terms:
            if abs(c) <= CLEAN_THRESHOLD:
                continue
            facs = []
            for name, site in f:
                if jordan_wigner_transform and name in _SPINLESS:
                    facs += [("F", k) for k in range(site)] + [(_SPINLESS[name], site)]
                elif jordan_wigner_transform and name in _SPINFUL:
                    bare, spin = _SPINFUL[name]
                    facs += [("F", k) for k in range(site)]
                    if spin == 1:
                        facs.append(("Fup", site))
                    facs.append((bare, site))
                else:
                    facs.append((name, site))
            out.append((complex(c), [(n, s + 1) for n, s in facs]))
        return out


def obj2MO(a):
    """[[name, site], ...] (0-based sites) -> MultiOperator with one term."""
    if isinstance(a, MultiOperator):
        return a
    if _isnum(a):
        return a * identity()
    return MultiOperator([(1.0, tuple((str(n), int(s)) for n, s in a))])


def identity():
    return MultiOperator([(1.0, (("Id", 0),))])


def msum(ops):
    terms = []
    for o in ops:
        if o is None or (_isnum(o) and o == 0):
            continue
        if _isnum(o):
            o = o * identity()
        terms.extend(o.terms)
    return MultiOperator(terms)


# ---------------------------------------------------------------------------
# model helpers (terms for the benchmark / test configurations)
# ---------------------------------------------------------------------------
def heisenberg_terms(n, J=1.0, Jz=None):
    """sum_i J_i (Sx Sx + Sy Sy) + Jz_i Sz Sz on sites 1..n (session format)."""
    J = np.broadcast_to(np.asarray(J, dtype=float), (n - 1,))
    Jz = J if Jz is None else np.broadcast_to(np.asarray(Jz, dtype=float), (n - 1,))
    t = []
    for i in range(1, n):
        t.append((complex(J[i - 1]), [("Sx", i), ("Sx", i + 1)]))
        t.append((complex(J[i - 1]), [("Sy", i), ("Sy", i + 1)]))
        t.append((complex(Jz[i - 1]), [("Sz", i), ("Sz", i + 1)]))
    return t


def hubbard_terms(n, t=1.0, U=4.0):
    """-t sum_sigma (c^dag_i c_{i+1} + h.c.) + U sum (n_up - 1/2)(n_dn - 1/2) on
    native spinful sites, Jordan-Wigner transformed."""
    H = MultiOperator()
    for i in range(n - 1):
        for up, dn in (("Cdagup", "Cup"), ("Cdagdn", "Cdn")):
            H = H + (-t) * (obj2MO([[up, i], [dn, i + 1]]) + obj2MO([[up, i + 1], [dn, i]]))
    for i in range(n):
        nu = obj2MO([["Nup", i]]) - 0.5 * obj2MO([["Id", i]])
        nd = obj2MO([["Ndn", i]]) - 0.5 * obj2MO([["Id", i]])
        H = H + U * nu * nd
    return H.to_terms()
