"""System construction: lattice, symmetry, cluster orbits, site bases and the
flattened cluster-expansion tables consumed by the C-ABI library.

This module plays the role of the reference's ``System`` construction
(``src/casm/clexmonte/system/System.cc:41-72``, JSON parse
``src/casm/clexmonte/system/io/json/System_json_io.cc:453-690``) together with
the parts of the generated Clexulator that define the basis functions
(``tests/unit/clexmonte/data/FCC_binary_vacancy/basis_sets/bset.default/
FCC_binary_vacancy_Clexulator_default.cc:306-455``).  The reference compiles a
generated C++ Clexulator per basis set; here the same information (orbits,
cluster functions, neighbourhoods) is produced as flat arrays:

* ``orbit tables``  CT[orbit][tau][function] = sum_d coeff_d prod_i phi_{d_i}(tau_i)
  used for full correlations (the Clexulator's ``eval_bfunc_*``);
* ``flower``        per sublattice, every cluster instance containing a site
  of that sublattice in the origin cell, with a table
  V[tau_others][f] = sum_phi (J_phi/m) sum_{d: d_s=f} coeff prod_{i!=s} phi(tau_i)
  so that dE(s: o->n) = sum_f (phi_f(n)-phi_f(o)) sum_e V_e[tau_e][f]
  (the Clexulator's ``site_deval_bfunc_*`` contracted with the ECIs).

Conventions (stated because the reference's are in non-vendored CASM code):
* linear site index ``l = b * n_cells + cell`` with
  ``cell = i + L1 * (j + L2 * k)`` for a diagonal supercell ``diag(L1,L2,L3)``;
* cluster functions are symmetrised over the prototype's invariant group and
  normalised so each decoration orbit carries coefficient ``1/sqrt(|orbit|)``
  (this reproduces the ``0.707107`` / ``\\sqrt{1/2}`` terms of the fixture);
* "chebychev" site functions are the discrete orthonormal polynomials over
  occupant index with uniform weights (binary: phi(A)=-1, phi(B)=+1).
"""
from __future__ import annotations

import itertools
import json
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# Boltzmann constant in eV/K, the value CASM uses for CASM::KB.
KB = 8.6173303e-05

TOL = 1e-5


# ---------------------------------------------------------------------------
# Prim
# ---------------------------------------------------------------------------
class Prim:
    """Primitive crystal structure.

    ``lattice`` rows are the lattice vectors (as in CASM ``prim.json``),
    ``basis`` holds fractional coordinates, ``occupants`` the allowed species
    on each basis site.
    """

    def __init__(self, lattice, basis, occupants, title="prim"):
        self.lattice = np.array(lattice, dtype=float).reshape(3, 3)
        self.basis = np.array(basis, dtype=float).reshape(-1, 3)
        self.occupants = [list(o) for o in occupants]
        self.title = title
        if len(self.occupants) != len(self.basis):
            raise ValueError("Prim: basis and occupants size mismatch")
        # column matrix: cart = A @ frac
        self.A = self.lattice.T.copy()
        self.Ainv = np.linalg.inv(self.A)

    @property
    def n_sublat(self) -> int:
        return len(self.basis)

    def n_occ(self, b: int) -> int:
        return len(self.occupants[b])

    def variable(self, b: int) -> bool:
        return len(self.occupants[b]) >= 2

    def cart(self, b: int, cell) -> np.ndarray:
        return self.A @ (self.basis[b] + np.asarray(cell, dtype=float))

    def components(self) -> List[str]:
        out: List[str] = []
        for occ in self.occupants:
            for s in occ:
                if s not in out:
                    out.append(s)
        return out

    @classmethod
    def from_dict(cls, d: dict) -> "Prim":
        """Read a CASM ``prim.json`` dict (Fractional coordinates only)."""
        mode = d.get("coordinate_mode", "Fractional")
        lat = np.array(d["lattice_vectors"], dtype=float)
        basis = []
        occ = []
        for site in d["basis"]:
            c = np.array(site["coordinate"], dtype=float)
            if mode.lower().startswith("cart"):
                c = np.linalg.solve(lat.T, c)
            elif not mode.lower().startswith("frac"):
                raise ValueError(f"unsupported coordinate_mode {mode}")
            basis.append(c)
            occ.append(site["occupants"])
        return cls(lat, basis, occ, d.get("title", "prim"))

    def to_dict(self) -> dict:
        return {
            "basis": [{"coordinate": list(map(float, c)), "occupants": o}
                      for c, o in zip(self.basis, self.occupants)],
            "coordinate_mode": "Fractional",
            "lattice_vectors": self.lattice.tolist(),
            "title": self.title,
        }


def fcc_prim(a: float = 1.0, occupants=("A", "B")) -> Prim:
    h = a / 2.0
    return Prim([[0, h, h], [h, 0, h], [h, h, 0]], [[0, 0, 0]], [list(occupants)], "FCC")


def bcc_prim(a: float = 1.0, occupants=("A", "B")) -> Prim:
    h = a / 2.0
    return Prim([[-h, h, h], [h, -h, h], [h, h, -h]], [[0, 0, 0]], [list(occupants)], "BCC")


def hcp_prim(a: float = 1.0, c: Optional[float] = None, occupants=("A", "B")) -> Prim:
    if c is None:
        c = a * math.sqrt(8.0 / 3.0)  # ideal c/a
    lat = [[a, 0, 0], [-a / 2.0, a * math.sqrt(3.0) / 2.0, 0], [0, 0, c]]
    basis = [[1.0 / 3.0, 2.0 / 3.0, 0.25], [2.0 / 3.0, 1.0 / 3.0, 0.75]]
    return Prim(lat, basis, [list(occupants), list(occupants)], "HCP")


# ---------------------------------------------------------------------------
# Symmetry
# ---------------------------------------------------------------------------
@dataclass
class SymOp:
    R: np.ndarray          # integer 3x3, acts on fractional coordinates
    tau: np.ndarray        # fractional translation
    perm: List[int]        # basis site b -> perm[b]
    shift: List[Tuple[int, int, int]]  # R f_b + tau = f_perm[b] + shift[b]

    def apply(self, site):
        b, n = site[0], site[1:]
        n2 = self.R @ np.array(n, dtype=int) + np.array(self.shift[b], dtype=int)
        return (self.perm[b], int(n2[0]), int(n2[1]), int(n2[2]))


def factor_group(prim: Prim) -> List[SymOp]:
    """Space-group operations modulo lattice translations (brute force over
    integer fractional rotations with entries in {-1,0,1})."""
    A, Ainv = prim.A, prim.Ainv
    ops: List[SymOp] = []
    vals = (-1, 0, 1)
    for entries in itertools.product(vals, repeat=9):
        R = np.array(entries, dtype=int).reshape(3, 3)
        d = round(np.linalg.det(R))
        if d not in (1, -1):
            continue
        Q = A @ R @ Ainv
        if not np.allclose(Q.T @ Q, np.eye(3), atol=1e-6):
            continue
        # find translations
        nb = prim.n_sublat
        f0 = prim.basis[0]
        for bt in range(nb):
            if sorted(prim.occupants[bt]) != sorted(prim.occupants[0]):
                continue
            tau = prim.basis[bt] - R @ f0
            tau = tau - np.floor(tau + TOL)
            perm, shift, ok = [], [], True
            for b in range(nb):
                img = R @ prim.basis[b] + tau
                found = False
                for b2 in range(nb):
                    diff = img - prim.basis[b2]
                    if np.allclose(diff, np.round(diff), atol=1e-4) and \
                            sorted(prim.occupants[b2]) == sorted(prim.occupants[b]):
                        perm.append(b2)
                        shift.append(tuple(int(x) for x in np.round(diff)))
                        found = True
                        break
                if not found:
                    ok = False
                    break
            if ok:
                ops.append(SymOp(R, tau, perm, shift))
    # dedupe (same R, same tau mod 1)
    uniq: List[SymOp] = []
    for op in ops:
        dup = False
        for u in uniq:
            if np.array_equal(u.R, op.R):
                dt = u.tau - op.tau
                if np.allclose(dt, np.round(dt), atol=1e-4):
                    dup = True
                    break
        if not dup:
            uniq.append(op)
    return uniq


# ---------------------------------------------------------------------------
# Clusters and orbits
# ---------------------------------------------------------------------------
Site = Tuple[int, int, int, int]  # (b, i, j, k)


def canon(sites: Sequence[Site]) -> Tuple[Site, ...]:
    """Translation-canonical form: sorted, first site in the origin cell."""
    s = sorted(sites)
    if not s:
        return tuple()
    _, i0, j0, k0 = s[0]
    return tuple((b, i - i0, j - j0, k - k0) for (b, i, j, k) in s)


def site_dist(prim: Prim, s: Site, t: Site) -> float:
    return float(np.linalg.norm(prim.cart(s[0], s[1:]) - prim.cart(t[0], t[1:])))


def cluster_lengths(prim: Prim, sites: Sequence[Site]) -> Tuple[float, float]:
    if len(sites) < 2:
        return 0.0, 0.0
    d = [site_dist(prim, a, b) for a, b in itertools.combinations(sites, 2)]
    return max(d), min(d)


@dataclass
class ClusterFunction:
    orbit: int
    terms: List[Tuple[float, Tuple[int, ...]]]  # (coeff, decoration per prototype site)


@dataclass
class Orbit:
    index: int
    prototype: Tuple[Site, ...]
    elements: List[Tuple[Site, ...]]     # each in prototype site order, translation-canonical
    invariant_perms: List[Tuple[int, ...]]
    max_length: float
    min_length: float
    functions: List[int] = field(default_factory=list)  # linear function indices

    @property
    def mult(self) -> int:
        return len(self.elements)

    @property
    def size(self) -> int:
        return len(self.prototype)


def site_basis_functions(n_occ: int, kind: str) -> np.ndarray:
    """phi[f][occ] for f in [0, n_occ-1).

    ``occupation``: phi_f(occ) = 1 if occ == f+1 else 0 (the fixture's
    ``m_occ_func_0_f``).  ``chebychev``: discrete orthonormal polynomials in
    the occupant index, uniform weights, constant removed, sign chosen so the
    last occupant has a positive value.
    """
    if n_occ < 2:
        return np.zeros((0, n_occ))
    if kind == "occupation":
        phi = np.zeros((n_occ - 1, n_occ))
        for f in range(n_occ - 1):
            phi[f, f + 1] = 1.0
        return phi
    if kind == "chebychev":
        x = np.arange(n_occ, dtype=float)
        basis = [np.ones(n_occ)]
        out = []
        for p in range(1, n_occ):
            v = x ** p
            for u in basis:
                v = v - (np.mean(v * u) / np.mean(u * u)) * u
            v = v / math.sqrt(np.mean(v * v))
            if v[-1] < 0:
                v = -v
            basis.append(v)
            out.append(v)
        return np.array(out)
    raise ValueError(f"unknown site basis '{kind}'")


class BasisSet:
    """Cluster orbits and symmetrised cluster functions.

    ``max_length[n]`` bounds the maximum site-to-site distance of
    ``n``-site clusters (CASM "periodic_max_length" cluster specs,
    ``basis_sets/bset.default/bspecs.json`` in the reference fixtures).
    """

    def __init__(self, prim: Prim, max_length: Dict[int, float], site_basis: str = "chebychev",
                 ops: Optional[List[SymOp]] = None):
        self.prim = prim
        self.max_length = {int(k): float(v) for k, v in max_length.items()}
        self.site_basis = site_basis
        self.ops = ops if ops is not None else factor_group(prim)
        self.phi = [site_basis_functions(prim.n_occ(b), site_basis) for b in range(prim.n_sublat)]
        self.orbits: List[Orbit] = []
        self.functions: List[ClusterFunction] = []
        self._build_orbits()
        self._build_functions()

    # -- orbits ----------------------------------------------------------
    def _orbit_of(self, sites: Sequence[Site]):
        elems = {}
        for op in self.ops:
            img = [op.apply(s) for s in sites]
            c = canon(img)
            if c not in elems:
                elems[c] = img
        return elems

    def _ordered_element(self, proto: Tuple[Site, ...], img: List[Site]) -> Tuple[Site, ...]:
        # translate image (in prototype site order) so its canonical first site is at origin
        if not img:
            return tuple()
        s = sorted(img)
        _, i0, j0, k0 = s[0]
        return tuple((b, i - i0, j - j0, k - k0) for (b, i, j, k) in img)

    def _build_orbits(self):
        prim = self.prim
        var = [b for b in range(prim.n_sublat) if prim.variable(b)]
        branches: Dict[int, List[Tuple[Site, ...]]] = {}
        found: Dict[Tuple[Site, ...], bool] = {}
        protos_by_branch: Dict[int, List[Tuple[Tuple[Site, ...], dict]]] = {}

        # branch 0 (empty) and 1 (points)
        protos_by_branch[0] = [(tuple(), {tuple(): []})]
        pts = []
        for b in var:
            s = [(b, 0, 0, 0)]
            c = canon(s)
            if c in found:
                continue
            el = self._orbit_of(s)
            for e in el:
                found[e] = True
            pts.append((min(el.keys()), el))
        protos_by_branch[1] = pts

        maxb = max(self.max_length.keys()) if self.max_length else 1
        rmax = max([v for k, v in self.max_length.items() if k >= 2] + [0.0])
        # candidate neighbourhood sites around origin cell
        cand: List[Site] = []
        if rmax > 0:
            lat_len = min(np.linalg.norm(prim.A[:, i]) for i in range(3))
            # conservative cell range
            span = int(math.ceil(rmax / 1e-9 if lat_len == 0 else rmax * np.abs(prim.Ainv).sum(axis=1).max())) + 2
            rng = range(-span, span + 1)
            for b in var:
                for i, j, k in itertools.product(rng, rng, rng):
                    cand.append((b, i, j, k))

        for n in range(2, maxb + 1):
            L = self.max_length.get(n, 0.0)
            out = []
            if L <= 0:
                protos_by_branch[n] = out
                continue
            for proto, _ in protos_by_branch[n - 1]:
                if cluster_lengths(prim, list(proto))[0] > L + TOL:
                    continue
                for s in cand:
                    if s in proto:
                        continue
                    ok = True
                    for t in proto:
                        if site_dist(prim, s, t) > L + TOL:
                            ok = False
                            break
                    if not ok:
                        continue
                    sites = list(proto) + [s]
                    c = canon(sites)
                    if c in found:
                        continue
                    el = self._orbit_of(sites)
                    for e in el:
                        found[e] = True
                    out.append((min(el.keys()), el))
            protos_by_branch[n] = out

        # order: branch size, then max length, then min length, then prototype
        idx = 0
        for n in sorted(protos_by_branch.keys()):
            items = []
            for proto, el in protos_by_branch[n]:
                mx, mn = cluster_lengths(prim, list(proto))
                items.append((round(mx, 6), round(mn, 6), proto, el))
            items.sort(key=lambda t: (t[0], t[1], t[2]))
            for mx, mn, proto, el in items:
                proto_l = list(proto)
                elements = []
                for e in sorted(el.keys()):
                    # find an op mapping prototype to this element and keep site order
                    img = None
                    for op in self.ops:
                        im = [op.apply(x) for x in proto_l]
                        if canon(im) == e:
                            img = im
                            break
                    elements.append(self._ordered_element(proto, img))
                perms = self._invariant_perms(proto)
                self.orbits.append(Orbit(idx, proto, elements, perms, mx, mn))
                idx += 1

    def _invariant_perms(self, proto: Tuple[Site, ...]) -> List[Tuple[int, ...]]:
        k = len(proto)
        if k == 0:
            return [tuple()]
        base = canon(proto)
        perms = set()
        for op in self.ops:
            img = [op.apply(s) for s in proto]
            if canon(img) != base:
                continue
            s = sorted(img)
            _, i0, j0, k0 = s[0]
            p0 = sorted(proto)
            _, a0, b0, c0 = p0[0]
            # translate img so that it coincides with proto as a set
            tr = [(b, i - i0 + a0, j - j0 + b0, kk - k0 + c0) for (b, i, j, kk) in img]
            perm = tuple(proto.index(x) for x in tr)  # site j of proto maps to site perm[j]
            perms.add(perm)
        return sorted(perms)

    # -- functions -------------------------------------------------------
    def _build_functions(self):
        lfi = 0
        for orb in self.orbits:
            k = orb.size
            if k == 0:
                self.functions.append(ClusterFunction(orb.index, [(1.0, tuple())]))
                orb.functions.append(lfi)
                lfi += 1
                continue
            nf = [self.phi[s[0]].shape[0] for s in orb.prototype]
            seen = set()
            decos = sorted(itertools.product(*[range(x) for x in nf]))
            for d in decos:
                if d in seen:
                    continue
                members = set()
                for p in orb.invariant_perms:
                    # site j -> p[j]; decoration moves with the site
                    nd = [0] * k
                    for j in range(k):
                        nd[p[j]] = d[j]
                    members.add(tuple(nd))
                members = sorted(members)
                for m in members:
                    seen.add(m)
                c = 1.0 / math.sqrt(len(members))
                self.functions.append(ClusterFunction(orb.index, [(c, m) for m in members]))
                orb.functions.append(lfi)
                lfi += 1

    @property
    def n_functions(self) -> int:
        return len(self.functions)

    def orbit_table(self, orbit: Orbit) -> np.ndarray:
        """CT[tau][func_local] with tau the mixed-radix occupant index of the
        prototype sites (site 0 fastest)."""
        k = orbit.size
        nocc = [self.prim.n_occ(s[0]) for s in orbit.prototype]
        nt = int(np.prod(nocc)) if k else 1
        out = np.zeros((nt, len(orbit.functions)))
        for t in range(nt):
            tau = []
            x = t
            for j in range(k):
                tau.append(x % nocc[j])
                x //= nocc[j]
            for fl, fi in enumerate(orbit.functions):
                v = 0.0
                for coeff, d in self.functions[fi].terms:
                    p = coeff
                    for j in range(k):
                        p *= self.phi[orbit.prototype[j][0]][d[j], tau[j]]
                    v += p
                out[t, fl] = v
        return out

    def to_dict(self) -> dict:
        """A CASM ``clust.json``-like summary (orbits with prototype sites,
        multiplicity and function terms)."""
        orbits = []
        for orb in self.orbits:
            orbits.append({
                "linear_orbit_index": orb.index,
                "mult": orb.mult,
                "prototype": {"sites": [list(s) for s in orb.prototype],
                              "max_length": orb.max_length, "min_length": orb.min_length},
                "cluster_functions": [
                    {"linear_function_index": fi,
                     "terms": [[c, list(d)] for c, d in self.functions[fi].terms]}
                    for fi in orb.functions],
            })
        return {"orbits": orbits, "site_functions": [p.tolist() for p in self.phi],
                "site_basis": self.site_basis}


# ---------------------------------------------------------------------------
# Cluster expansion, supercells, flattening
# ---------------------------------------------------------------------------
class ClusterExpansion:
    """A basis set together with a dense ECI vector (per linear function
    index), as in the reference's ``ClexData`` (``system_data.hh:42-46``)."""

    def __init__(self, basis_set: BasisSet, eci):
        self.basis_set = basis_set
        eci = np.asarray(eci, dtype=float)
        if eci.shape != (basis_set.n_functions,):
            raise ValueError(f"ECI size {eci.shape} != n_functions {basis_set.n_functions}")
        self.eci = eci

    @property
    def prim(self) -> Prim:
        return self.basis_set.prim

    def eci_json(self) -> dict:
        return {"linear_function_index": list(range(len(self.eci))),
                "eci": self.eci.tolist()}


@dataclass
class Supercell:
    prim: Prim
    dims: Tuple[int, int, int]

    @property
    def n_cells(self) -> int:
        return self.dims[0] * self.dims[1] * self.dims[2]

    @property
    def n_sites(self) -> int:
        return self.prim.n_sublat * self.n_cells

    def linear_index(self, b, i, j, k) -> int:
        L1, L2, L3 = self.dims
        return b * self.n_cells + (i % L1) + L1 * ((j % L2) + L2 * (k % L3))

    def site(self, l: int) -> Site:
        nc = self.n_cells
        b, c = divmod(l, nc)
        L1, L2, _ = self.dims
        i = c % L1
        j = (c // L1) % L2
        k = c // (L1 * L2)
        return (b, i, j, k)


def decode_tau(t: int, radices: Sequence[int]) -> List[int]:
    out = []
    for r in radices:
        out.append(t % r)
        t //= r
    return out


class FlatCE:
    """Flattened arrays describing a cluster expansion for the C-ABI.

    Arrays (all numpy, C-contiguous):

    * ``site_func[b, f, occ]`` (FP64) site basis functions;
    * correlations: per orbit ``corr_orbit_begin`` into ``corr_elem_sites``
      (elements x k sites of 4 ints), ``orbit_k``, ``orbit_mult``,
      ``orbit_table_offset`` into ``orbit_tables`` (FP64 CT[tau][func_local]),
      ``orbit_func_begin`` (global linear function index of first function),
      ``orbit_nfunc``;
    * flower: per sublattice ``flower_begin[b]`` into entries; per entry
      ``entry_orbit``, ``entry_k`` (other sites), ``entry_site_begin`` into
      ``entry_sites`` (n x 4 ints: b, di, dj, dk), ``entry_table`` offset into
      ``flower_tables`` (FP64 V[tau_others][f]), ``entry_radix_begin`` etc.
    """

    def __init__(self, ce: ClusterExpansion):
        bs = ce.basis_set
        prim = bs.prim
        self.ce = ce
        self.n_sublat = prim.n_sublat
        self.n_occ = np.array([prim.n_occ(b) for b in range(prim.n_sublat)], dtype=np.int32)
        self.n_occ_max = int(self.n_occ.max())
        self.n_func = np.array([bs.phi[b].shape[0] for b in range(prim.n_sublat)], dtype=np.int32)
        self.n_func_max = max(1, int(self.n_func.max()))
        self.site_func = np.zeros((prim.n_sublat, self.n_func_max, self.n_occ_max))
        for b in range(prim.n_sublat):
            p = bs.phi[b]
            self.site_func[b, :p.shape[0], :p.shape[1]] = p
        self.n_corr = bs.n_functions
        self.eci = ce.eci.copy()
        # asymmetric unit of each sublattice: sublattices mapped onto each other by
        # the factor group share one (canonical swaps may exchange occupants
        # between them, as the reference's OccSwap candidates per asym unit do)
        asym = list(range(prim.n_sublat))
        for op in bs.ops:
            for b in range(prim.n_sublat):
                a, c = asym[b], asym[op.perm[b]]
                if a != c:
                    lo, hi = min(a, c), max(a, c)
                    asym = [lo if x == hi else x for x in asym]
        relabel = {u: i for i, u in enumerate(sorted(set(asym)))}
        self.asym = np.array([relabel[u] for u in asym], dtype=np.int32)

        # --- correlations (orbit element lists) ---
        orbit_k, orbit_mult, orbit_nfunc, orbit_func_begin = [], [], [], []
        corr_orbit_begin = [0]
        elem_sites = []
        tables = []
        table_off = []
        toff = 0
        for orb in bs.orbits:
            orbit_k.append(orb.size)
            orbit_mult.append(orb.mult)
            orbit_nfunc.append(len(orb.functions))
            orbit_func_begin.append(orb.functions[0])
            for e in orb.elements:
                for s in e:
                    elem_sites.append(s)
            corr_orbit_begin.append(corr_orbit_begin[-1] + len(orb.elements))
            ct = bs.orbit_table(orb)
            table_off.append(toff)
            tables.append(ct.ravel())
            toff += ct.size
        self.n_orbits = len(bs.orbits)
        self.orbit_k = np.array(orbit_k, dtype=np.int32)
        self.orbit_mult = np.array(orbit_mult, dtype=np.int32)
        self.orbit_nfunc = np.array(orbit_nfunc, dtype=np.int32)
        self.orbit_func_begin = np.array(orbit_func_begin, dtype=np.int32)
        self.corr_orbit_begin = np.array(corr_orbit_begin, dtype=np.int32)
        # element site offsets: per orbit, elements have orbit_k sites each
        self.corr_elem_begin = np.zeros(self.n_orbits + 1, dtype=np.int32)
        acc = 0
        for a, orb in enumerate(bs.orbits):
            self.corr_elem_begin[a] = acc
            acc += len(orb.elements) * orb.size
        self.corr_elem_begin[self.n_orbits] = acc
        self.corr_elem_sites = np.array(elem_sites, dtype=np.int32).reshape(-1, 4) if elem_sites \
            else np.zeros((0, 4), dtype=np.int32)
        self.orbit_table_offset = np.array(table_off, dtype=np.int32)
        self.orbit_tables = np.concatenate(tables) if tables else np.zeros(0)

        # --- flower tables ---
        self._build_flower(bs)
        self._build_binary(bs)

    def _build_flower(self, bs: BasisSet):
        prim = bs.prim
        nb = prim.n_sublat
        eci = self.eci
        per_b: List[List[dict]] = [[] for _ in range(nb)]
        for orb in bs.orbits:
            k = orb.size
            if k == 0:
                continue
            for e in orb.elements:
                for j in range(k):
                    b = e[j][0]
                    _, i0, j0, k0 = e[j]
                    others = [(e[i][0], e[i][1] - i0, e[i][2] - j0, e[i][3] - k0) for i in range(k) if i != j]
                    oidx = [i for i in range(k) if i != j]
                    nf_b = bs.phi[b].shape[0]
                    rad = [prim.n_occ(e[i][0]) for i in oidx]
                    nt = int(np.prod(rad)) if rad else 1
                    V = np.zeros((nt, nf_b))
                    for t in range(nt):
                        tau = decode_tau(t, rad)
                        for fi in orb.functions:
                            J = eci[fi]
                            if J == 0.0:
                                continue
                            for coeff, d in bs.functions[fi].terms:
                                p = coeff * J / orb.mult
                                for q, i in enumerate(oidx):
                                    p *= bs.phi[e[i][0]][d[i], tau[q]]
                                V[t, d[j]] += p
                    per_b[b].append({"orbit": orb.index, "others": others, "V": V, "rad": rad})
        # dedupe tables up to permutation of the other sites
        tables: List[np.ndarray] = []
        table_rad: List[List[int]] = []
        flower_begin = [0]
        entry_orbit, entry_k, entry_site_begin, entry_table = [], [], [0], []
        entry_sites = []
        for b in range(nb):
            for ent in per_b[b]:
                others, V, rad = ent["others"], ent["V"], ent["rad"]
                kk = len(others)
                chosen = None
                for perm in itertools.permutations(range(kk)):
                    prad = [rad[p] for p in perm]
                    # V permuted: new order q -> old perm[q]
                    Vp = self._permute_table(V, rad, perm)
                    for ti, (T, trad) in enumerate(zip(tables, table_rad)):
                        if trad == prad and T.shape == Vp.shape and np.allclose(T, Vp, atol=1e-14, rtol=1e-12):
                            chosen = (ti, perm)
                            break
                    if chosen:
                        break
                if chosen is None:
                    tables.append(V.copy())
                    table_rad.append(list(rad))
                    chosen = (len(tables) - 1, tuple(range(kk)))
                ti, perm = chosen
                entry_orbit.append(ent["orbit"])
                entry_k.append(kk)
                for p in perm:
                    entry_sites.append(others[p])
                entry_site_begin.append(entry_site_begin[-1] + kk)
                entry_table.append(ti)
            flower_begin.append(len(entry_orbit))
        offs = []
        off = 0
        for T in tables:
            offs.append(off)
            off += T.size
        self.flower_begin = np.array(flower_begin, dtype=np.int32)
        self.entry_orbit = np.array(entry_orbit, dtype=np.int32)
        self.entry_k = np.array(entry_k, dtype=np.int32)
        self.entry_site_begin = np.array(entry_site_begin, dtype=np.int32)
        self.entry_sites = np.array(entry_sites, dtype=np.int32).reshape(-1, 4) if entry_sites \
            else np.zeros((0, 4), dtype=np.int32)
        self.n_tables = len(tables)
        self.table_offset = np.array(offs, dtype=np.int32)
        self.table_nf = np.array([T.shape[1] for T in tables], dtype=np.int32)
        self.table_k = np.array([len(r) for r in table_rad], dtype=np.int32)
        self.entry_table = np.array([offs[t] for t in entry_table], dtype=np.int32)
        self.entry_table_id = np.array(entry_table, dtype=np.int32)
        self.flower_tables = np.concatenate([T.ravel() for T in tables]) if tables else np.zeros(0)
        self.n_entries = len(entry_orbit)
        # halo width
        self.halo = int(np.abs(self.entry_sites[:, 1:]).max()) if len(self.entry_sites) else 0

    @staticmethod
    def _permute_table(V, rad, perm):
        kk = len(rad)
        if kk == 0:
            return V.copy()
        nt = V.shape[0]
        prad = [rad[p] for p in perm]
        out = np.zeros_like(V)
        for t in range(nt):
            tau = decode_tau(t, rad)
            ptau = [tau[p] for p in perm]
            tt = 0
            mul = 1
            for q in range(kk):
                tt += ptau[q] * mul
                mul *= prad[q]
            out[tt] = V[t]
        return out

    def _build_binary(self, bs: BasisSet):
        """Integer fast path data when every variable site is binary with
        phi = (-1, +1): dE = dphi(s) * sum_a (J_a/m_a) * K_a(s) where
        K_a(s) = sum over flower instances of orbit a of prod phi(others)
        = (-1)^(k-1) (n_a - 2 X_a), X_a = sum of XOR of the other sites'
        occupant bits."""
        prim = bs.prim
        ok = True
        for b in range(prim.n_sublat):
            if prim.variable(b):
                if prim.n_occ(b) != 2:
                    ok = False
                elif not np.allclose(bs.phi[b], [[-1.0, 1.0]]):
                    ok = False
        for orb in bs.orbits:
            if len(orb.functions) != 1:
                ok = False
        self.binary = ok
        if not ok:
            return
        # entries are grouped by orbit within each sublattice in build order
        nb = prim.n_sublat
        seg_begin = []  # per sublattice: list of (orbit, entry_begin, entry_end)
        self.bin_seg_begin = [0]
        segs = []
        for b in range(nb):
            eb, ee = self.flower_begin[b], self.flower_begin[b + 1]
            cur = None
            for e in range(eb, ee):
                o = int(self.entry_orbit[e])
                if cur is None or cur[0] != o:
                    if cur is not None:
                        segs.append(cur)
                    cur = [o, e, e + 1]
                else:
                    cur[2] = e + 1
            if cur is not None:
                segs.append(cur)
            self.bin_seg_begin.append(len(segs))
        self.bin_seg_begin = np.array(self.bin_seg_begin, dtype=np.int32)
        self.bin_segs = np.array(segs, dtype=np.int32).reshape(-1, 3)
        # weights per orbit: J/m ; count correlation index per orbit
        self.bin_orbit_weight = np.array(
            [self.eci[o.functions[0]] / o.mult for o in bs.orbits], dtype=float)


def flatten(ce: ClusterExpansion) -> FlatCE:
    return FlatCE(ce)


# ---------------------------------------------------------------------------
# Interaction graph, colourings (independent-set partitions)
# ---------------------------------------------------------------------------
def interaction_offsets(flat: FlatCE) -> Dict[int, List[Tuple[int, int, int, int]]]:
    """For each sublattice b: the set of sites (b', di, dj, dk) that share a
    cluster with (b,0,0,0), including pairs of 'other' sites of an instance
    (two sites conflict if some cluster contains both)."""
    nb = flat.n_sublat
    out: Dict[int, set] = {b: set() for b in range(nb)}
    for b in range(nb):
        for e in range(flat.flower_begin[b], flat.flower_begin[b + 1]):
            s0, s1 = flat.entry_site_begin[e], flat.entry_site_begin[e + 1]
            for q in range(s0, s1):
                out[b].add(tuple(int(x) for x in flat.entry_sites[q]))
    return {b: sorted(v) for b, v in out.items()}


# largest independent set (sites per class) the periodic colourings may use:
# canonical sweeps keep one or more FP32 fields per class site in LDS
MAX_CLASS_SITES = 4096


def nn_parities(flat: FlatCE) -> set:
    """Parities (di, dj, dk) mod 2 of the cell offsets of same-sublattice
    neighbours in the shortest pair orbit that has them: the translations
    across which nearest-neighbour-driven orderings put unlike occupants."""
    for a in range(flat.n_orbits):
        if flat.orbit_k[a] != 2:
            continue
        par = set()
        for b in range(flat.n_sublat):
            for e in range(flat.flower_begin[b], flat.flower_begin[b + 1]):
                if flat.entry_orbit[e] != a:
                    continue
                q = flat.entry_site_begin[e]
                b2, di, dj, dk = (int(x) for x in flat.entry_sites[q])
                if b2 == b:
                    par.add((di % 2, dj % 2, dk % 2))
        if par:
            return par
    return set()


def periodic_colouring(flat: FlatCE, dims, offs=None):
    """The periodic colouring of the cells used for the independent sets:
    ("residue", (m1, m2, m3)), class = (i mod m1, j mod m2, k mod m3), or
    ("linear", (M, a, c)), class = (i + a j + c k) mod M; None if neither is
    conflict-free (no interaction offset within one sublattice joins two
    cells of a class).  Candidates hold at most MAX_CLASS_SITES cells per
    class.  A colouring's score is the number of parities of nearest-
    neighbour offsets (nn_parities) among its class translations: with none,
    a period-2 ordering driven by nearest-neighbour pairs (L1_2 or L1_0 on
    FCC, B2 on BCC) can make every class pure, and swaps inside a class then
    never create an antisite.  Choice: among the colourings with a positive
    score (all, if none has one), the fewest classes, except that one holding
    every parity is kept when it needs at most 25 % more classes (a pass
    costs about one barrier-separated phase per class).  FCC 24^3 keeps
    3 x 3 x 3 residues (27 classes, every parity); FCC 32^3 takes
    (i + 3 j + 7 k) mod 16, whose classes hold every L1_2 sublattice; HCP
    40 x 40 x 24 takes 10 linear classes per sublattice instead of 50
    residue classes."""
    if offs is None:
        offs = interaction_offsets(flat)
    same = [o[1:] for b, lst in offs.items() for o in lst if o[0] == b]
    L1, L2, L3 = (int(x) for x in dims)
    n_cells = L1 * L2 * L3
    P = nn_parities(flat)
    cands = []  # (score, n_classes, key)
    divs = [[m for m in range(1, d + 1) if d % m == 0] for d in (L1, L2, L3)]
    for m in itertools.product(*divs):
        n = m[0] * m[1] * m[2]
        if n_cells // n > MAX_CLASS_SITES:
            continue
        if all(d[0] % m[0] or d[1] % m[1] or d[2] % m[2] for d in same):
            odd = [x % 2 == 1 for x in m]
            score = sum(1 for v in P if all(odd[t] or v[t] == 0 for t in range(3)))
            cands.append((score, n, ("residue", tuple(m))))
    for M in range(2, L1 + 1):
        if L1 % M or n_cells // M > MAX_CLASS_SITES:
            continue
        for a in range(M):
            if (a * L2) % M:
                continue
            for c in range(M):
                if (c * L3) % M:
                    continue
                if all((d[0] + a * d[1] + c * d[2]) % M for d in same):
                    score = sum(1 for v in P if M % 2 == 1 or (v[0] + a * v[1] + c * v[2]) % 2 == 0)
                    cands.append((score, M, ("linear", (M, a, c))))
    if not cands:
        return None
    mixing = [c for c in cands if c[0] > 0] or cands
    n_min = min(c[1] for c in mixing)
    full = [c for c in mixing if c[0] == len(P) and c[1] * 4 <= n_min * 5]
    pool = full or mixing
    return min(pool, key=lambda c: (c[1], -c[0]))[2]


def color_classes(flat: FlatCE, sc: Supercell) -> List[np.ndarray]:
    """Partition of the variable sites into independent sets (no cluster
    contains two sites of one class): the periodic colouring of
    periodic_colouring when one exists (equal class sizes, each class a
    coset of a lattice of cell translations), else a greedy colouring of the
    conflict graph.  Each class holds sites of a single sublattice, in
    linear-index order."""
    prim = sc.prim
    offs = interaction_offsets(flat)
    col = periodic_colouring(flat, sc.dims, offs)
    classes: List[np.ndarray] = []
    if col is not None:
        L1, L2, L3 = sc.dims
        cells = np.arange(sc.n_cells)
        i, j, k = cells % L1, (cells // L1) % L2, cells // (L1 * L2)
        kind, p = col
        if kind == "residue":
            lab = (i % p[0]) + p[0] * ((j % p[1]) + p[1] * (k % p[2]))
            n_lab = p[0] * p[1] * p[2]
        else:
            lab = (i + p[1] * j + p[2] * k) % p[0]
            n_lab = p[0]
        for b in range(prim.n_sublat):
            if not prim.variable(b):
                continue
            for r in range(n_lab):
                classes.append((b * sc.n_cells + np.flatnonzero(lab == r)).astype(np.int32))
        return classes
    # greedy colouring
    N = sc.n_sites
    nbrs = [set() for _ in range(N)]
    for l in range(N):
        b, i, j, k = sc.site(l)
        if not prim.variable(b):
            continue
        for (b2, di, dj, dk) in offs[b]:
            t = sc.linear_index(b2, i + di, j + dj, k + dk)
            if t != l:
                nbrs[l].add(t)
                nbrs[t].add(l)
    color = -np.ones(N, dtype=int)
    for l in range(N):
        if not prim.variable(sc.site(l)[0]):
            continue
        used = {color[t] for t in nbrs[l]}
        c = 0
        while c in used:
            c += 1
        color[l] = c
    for b in range(prim.n_sublat):
        if not prim.variable(b):
            continue
        for c in range(color.max() + 1):
            sites = [l for l in range(N) if color[l] == c and sc.site(l)[0] == b]
            if sites:
                classes.append(np.array(sites, dtype=np.int32))
    return classes


def check_supercell(flat: FlatCE, sc: Supercell) -> None:
    """Raise if some cluster instance maps two of its sites onto the same
    supercell site (the flower dE formula assumes distinct sites)."""
    for e in range(flat.n_entries):
        s0, s1 = flat.entry_site_begin[e], flat.entry_site_begin[e + 1]
        b = int(np.searchsorted(flat.flower_begin, e, side="right") - 1)
        sites = [sc.linear_index(b, 0, 0, 0)]
        for q in range(s0, s1):
            bb, di, dj, dk = (int(x) for x in flat.entry_sites[q])
            sites.append(sc.linear_index(bb, di, dj, dk))
        if len(set(sites)) != len(sites):
            raise ValueError(f"supercell {sc.dims} too small for the cluster expansion: a cluster "
                             f"wraps onto itself")


# ---------------------------------------------------------------------------
# Reading the reference's on-disk basis set format
# ---------------------------------------------------------------------------
def parse_casm_function_latex(s: str) -> List[Tuple[float, Dict[int, int]]]:
    """Parse CASM basis.json function strings such as
    ``\\sqrt{1/2}(\\phi_{0,0}(s_{0})\\phi_{0,1}(s_{1})+\\phi_{0,1}(s_{0})\\phi_{0,0}(s_{1}))``
    into (coefficient, {site: function}) terms.  Only the forms produced for
    occupation basis sets with unit-weight products are supported."""
    import re
    s = s.strip()
    if s == "1":
        return [(1.0, {})]
    coeff = 1.0
    m = re.match(r"^\\sqrt\{(\d+)/(\d+)\}\((.*)\)$", s)
    if m:
        coeff = math.sqrt(float(m.group(1)) / float(m.group(2)))
        s = m.group(3)
    terms = []
    for part in s.split("+"):
        d = {}
        for fm in re.finditer(r"\\phi_\{(\d+),(\d+)\}\(s_\{(\d+)\}\)", part):
            d[int(fm.group(3))] = int(fm.group(2))
        terms.append((coeff, d))
    return terms


# ---------------------------------------------------------------------------
# Composition axes and the flat file read by the C++ calculators
# ---------------------------------------------------------------------------
class CompositionAxes:
    """Standard parametric composition axes: origin = every site holding its
    first occupant, end member i = every variable site holding component
    i+1 (where allowed).  x = R^+ (n - n0) with n the number of each
    component per unit cell (the role of the reference's
    CompositionConverter, call sites SemiGrandCanonicalCalculator.cc:144,
    172-178)."""

    def __init__(self, prim: Prim):
        self.components = prim.components()
        nc = len(self.components)
        self.n_sublat = prim.n_sublat
        self.occ_component = []
        for b in range(prim.n_sublat):
            self.occ_component.append([self.components.index(s) for s in prim.occupants[b]])
        origin = np.zeros(nc)
        for b in range(prim.n_sublat):
            origin[self.occ_component[b][0]] += 1
        ends = []
        for c in range(1, nc):
            e = np.zeros(nc)
            for b in range(prim.n_sublat):
                if c in self.occ_component[b]:
                    e[c] += 1
                else:
                    e[self.occ_component[b][0]] += 1
            if not np.allclose(e, origin):
                ends.append(e)
        self.origin = origin
        self.R = np.array(ends).T if ends else np.zeros((nc, 0))  # columns: end - origin
        self.R = self.R - origin[:, None] if ends else self.R
        self.Rpinv = np.linalg.pinv(self.R) if ends else np.zeros((0, nc))

    @property
    def n_independent(self) -> int:
        return self.R.shape[1]

    def mol_composition(self, occ: np.ndarray, n_cells: int) -> np.ndarray:
        n = np.zeros(len(self.components))
        occ = np.asarray(occ)
        for b in range(self.n_sublat):
            sub = occ[b * n_cells:(b + 1) * n_cells]
            cnt = np.bincount(sub, minlength=len(self.occ_component[b]))
            for o, c in enumerate(self.occ_component[b]):
                n[c] += cnt[o]
        return n / n_cells

    def param_composition(self, mol: np.ndarray) -> np.ndarray:
        return self.Rpinv @ (np.asarray(mol) - self.origin)

    def chem_weights(self, param_chem_pot) -> np.ndarray:
        """Per (sublattice, occupant) weights w such that the semi-grand
        potential change of an event is dE - sum (w[new] - w[old]):
        w_s = (mu^T R^+)_component(s)."""
        mu = np.asarray(param_chem_pot, dtype=float)
        wc = mu @ self.Rpinv if self.n_independent else np.zeros(len(self.components))
        nom = max(len(o) for o in self.occ_component)
        w = np.zeros((self.n_sublat, nom))
        for b in range(self.n_sublat):
            for o, c in enumerate(self.occ_component[b]):
                w[b, o] = wc[c]
        return w


def semigrand_canonical_swaps(flat: FlatCE) -> List[Tuple[int, int, int]]:
    """Single semi-grand canonical changes (asymmetric unit, occupant from,
    occupant to) of every unit with mutable sites, as the system's default
    semigrand_canonical_swaps (src/casm/clexmonte/system/System.cc:57-58)."""
    out = []
    for u in sorted({int(flat.asym[b]) for b in range(flat.n_sublat) if flat.n_occ[b] >= 2}):
        n = max(int(flat.n_occ[b]) for b in range(flat.n_sublat) if flat.asym[b] == u)
        out += [(u, a, c) for a in range(n) for c in range(n) if a != c]
    return out


def make_multiswaps(flat: FlatCE, max_total_count: int, keep=None, single_swaps=None) -> List[List[Tuple]]:
    """Multiswap types from single semi-grand canonical swaps: every multiset
    of 2 .. max_total_count single swaps (the system's integer
    "semigrand_canonical_multiswaps" input, System_json_io.cc:974-983),
    each a list of parts (unit, from, to, count).  keep(type) -> bool, if
    given, filters the types (for example charge neutrality); a kept type's
    reverse must be kept too.  Multisets that change one site twice
    (a part and its reverse: a unit's two different occupants exchanged) are
    kept: they are swaps."""
    import itertools
    singles = single_swaps if single_swaps is not None else semigrand_canonical_swaps(flat)
    types = []
    for n in range(2, max_total_count + 1):
        for combo in itertools.combinations_with_replacement(range(len(singles)), n):
            cnt: Dict[int, int] = {}
            for i in combo:
                cnt[i] = cnt.get(i, 0) + 1
            if len(cnt) > 4 or n > 8:
                continue
            t = [(singles[i][0], singles[i][1], singles[i][2], c) for i, c in sorted(cnt.items())]
            if keep is None or keep(t):
                types.append(t)
    return types


def save_flat(flat: FlatCE, path: str) -> None:
    """Write the flattened cluster expansion and composition axes as text
    lines ``name count v0 v1 ...`` (read by the C++ System)."""
    prim = flat.ce.prim
    axes = CompositionAxes(prim)
    nom = flat.n_occ_max
    occ_comp = np.full((prim.n_sublat, nom), -1, dtype=int)
    for b in range(prim.n_sublat):
        occ_comp[b, :len(axes.occ_component[b])] = axes.occ_component[b]
    items = [
        ("n_sublat", [flat.n_sublat]), ("n_occ", flat.n_occ), ("n_occ_max", [flat.n_occ_max]),
        ("n_func_max", [flat.n_func_max]), ("site_func", flat.site_func.ravel()),
        ("n_orbits", [flat.n_orbits]), ("orbit_k", flat.orbit_k), ("orbit_mult", flat.orbit_mult),
        ("orbit_nfunc", flat.orbit_nfunc), ("orbit_func_begin", flat.orbit_func_begin),
        ("corr_orbit_begin", flat.corr_orbit_begin), ("corr_elem_begin", flat.corr_elem_begin),
        ("corr_elem_sites", flat.corr_elem_sites.ravel()), ("orbit_table_offset", flat.orbit_table_offset),
        ("orbit_tables", flat.orbit_tables), ("n_corr", [flat.n_corr]), ("eci", flat.eci),
        ("flower_begin", flat.flower_begin), ("entry_orbit", flat.entry_orbit), ("entry_k", flat.entry_k),
        ("entry_site_begin", flat.entry_site_begin), ("entry_sites", flat.entry_sites.ravel()),
        ("entry_table", flat.entry_table), ("flower_tables", flat.flower_tables),
        ("binary", [1 if flat.binary else 0]), ("asym", flat.asym),
        ("n_components", [len(axes.components)]), ("occ_component", occ_comp.ravel()),
        ("comp_origin", axes.origin), ("comp_R", axes.R.T.ravel()), ("n_independent", [axes.n_independent]),
    ]
    if flat.binary:
        items += [("bin_seg_begin", flat.bin_seg_begin), ("bin_segs", flat.bin_segs.ravel()),
                  ("bin_orbit_weight", flat.bin_orbit_weight)]
    with open(path, "w") as f:
        for name, vals in items:
            vals = np.asarray(vals).ravel()
            if vals.dtype.kind == "f":
                body = " ".join(repr(float(v)) for v in vals)
            else:
                body = " ".join(str(int(v)) for v in vals)
            f.write(f"{name} {vals.size} {body}\n")
