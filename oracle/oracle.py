"""ctypes face of the FP64 CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / the timed CPU baseline.  The
product path (casmcode_clexmonte_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i32p = C.POINTER(C.c_int)
f64p = C.POINTER(C.c_double)
i8p = C.POINTER(C.c_int8)


class orc_system(C.Structure):
    _fields_ = [
        ("n_sublat", C.c_int), ("n_occ", i32p), ("asym", i32p), ("n_orbits", C.c_int),
        ("orbit_k", i32p), ("orbit_mult", i32p), ("orbit_nfunc", i32p), ("orbit_func_begin", i32p),
        ("corr_orbit_begin", i32p), ("corr_elem_begin", i32p), ("corr_elem_sites", i32p),
        ("orbit_table_offset", i32p), ("orbit_tables", f64p), ("n_corr", C.c_int),
        ("eci", f64p), ("L", C.c_int * 3),
    ]


class orc_run_params(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("temperature", C.c_double), ("kb", C.c_double), ("chem", f64p),
        ("n_occ_max", C.c_int), ("n_equil_passes", C.c_longlong), ("n_passes", C.c_longlong),
        ("full_recompute_samplers", C.c_int), ("seed", C.c_uint64),
    ]


class orc_run_result(C.Structure):
    _fields_ = [("n_attempt", C.c_longlong), ("n_accept", C.c_longlong), ("E_final", C.c_double),
                ("elapsed_s", C.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_clexmc.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_energy.restype = C.c_double
        L.orc_delta.restype = C.c_double
        L.orc_mt19937_64_nth.restype = C.c_uint64
        L.orc_mt19937_64_nth.argtypes = [C.c_uint64, C.c_longlong]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


class OracleSystem:
    """Holds numpy buffers alive and the C struct pointing at them."""

    def __init__(self, flat, dims, asym: Optional[Sequence[int]] = None):
        self.flat = flat
        self.dims = tuple(int(x) for x in dims)
        f = flat
        self._keep = []

        def k(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a

        n_occ = k(f.n_occ, np.int32)
        if asym is None:
            asym = getattr(f, "asym", None)
        if asym is None:
            asym = list(range(f.n_sublat))
        asym_a = k(asym, np.int32)
        s = orc_system()
        s.n_sublat = f.n_sublat
        s.n_occ = _p(n_occ, i32p)
        s.asym = _p(asym_a, i32p)
        s.n_orbits = f.n_orbits
        s.orbit_k = _p(k(f.orbit_k, np.int32), i32p)
        s.orbit_mult = _p(k(f.orbit_mult, np.int32), i32p)
        s.orbit_nfunc = _p(k(f.orbit_nfunc, np.int32), i32p)
        s.orbit_func_begin = _p(k(f.orbit_func_begin, np.int32), i32p)
        s.corr_orbit_begin = _p(k(f.corr_orbit_begin, np.int32), i32p)
        s.corr_elem_begin = _p(k(f.corr_elem_begin, np.int32), i32p)
        s.corr_elem_sites = _p(k(f.corr_elem_sites.ravel(), np.int32), i32p)
        s.orbit_table_offset = _p(k(f.orbit_table_offset, np.int32), i32p)
        s.orbit_tables = _p(k(f.orbit_tables, np.float64), f64p)
        s.n_corr = f.n_corr
        s.eci = _p(k(f.eci, np.float64), f64p)
        s.L[0], s.L[1], s.L[2] = self.dims
        self.s = s
        self.n_sites = f.n_sublat * self.dims[0] * self.dims[1] * self.dims[2]

    # -- evaluation -------------------------------------------------------
    def corr(self, occ) -> np.ndarray:
        occ = np.ascontiguousarray(occ, dtype=np.int8)
        out = np.zeros(self.flat.n_corr)
        lib().orc_corr(C.byref(self.s), _p(occ, i8p), _p(out, f64p))
        return out

    def energy(self, occ) -> float:
        occ = np.ascontiguousarray(occ, dtype=np.int8)
        return float(lib().orc_energy(C.byref(self.s), _p(occ, i8p)))

    def delta(self, occ, sites, new_occ) -> float:
        occ = np.ascontiguousarray(occ, dtype=np.int8).copy()
        sites = np.ascontiguousarray(sites, dtype=np.int32)
        nw = np.ascontiguousarray(new_occ, dtype=np.int32)
        return float(lib().orc_delta(C.byref(self.s), _p(occ, i8p), C.c_int(len(sites)),
                                     _p(sites, i32p), _p(nw, i32p)))

    def delta_corr(self, occ, sites, new_occ) -> np.ndarray:
        occ = np.ascontiguousarray(occ, dtype=np.int8).copy()
        sites = np.ascontiguousarray(sites, dtype=np.int32)
        nw = np.ascontiguousarray(new_occ, dtype=np.int32)
        out = np.zeros(self.flat.n_corr)
        lib().orc_delta_corr(C.byref(self.s), _p(occ, i8p), C.c_int(len(sites)), _p(sites, i32p),
                             _p(nw, i32p), _p(out, f64p))
        return out

    # -- Metropolis ---------------------------------------------------------
    def run(self, occ, kind: str, temperature: float, n_equil: int, n_passes: int, seed: int,
            chem=None, kb: float = 8.6173303e-05, full_recompute_samplers: bool = True):
        occ = np.ascontiguousarray(occ, dtype=np.int8).copy()
        p = orc_run_params()
        p.kind = 0 if kind == "canonical" else 1
        p.temperature = temperature
        p.kb = kb
        nom = int(self.flat.n_occ_max)
        p.n_occ_max = nom
        if chem is not None:
            chem_a = np.ascontiguousarray(chem, dtype=np.float64).reshape(self.flat.n_sublat * nom)
            self._chem = chem_a
            p.chem = _p(chem_a, f64p)
        else:
            p.chem = f64p()
        p.n_equil_passes = n_equil
        p.n_passes = n_passes
        p.full_recompute_samplers = 1 if full_recompute_samplers else 0
        p.seed = seed
        sE = np.zeros(n_passes)
        sc = np.zeros((n_passes, nom), dtype=np.int32)
        r = orc_run_result()
        lib().orc_run(C.byref(self.s), _p(occ, i8p), C.byref(p), _p(sE, f64p), _p(sc, i32p),
                      C.byref(r))
        return {"occ": occ, "E": sE, "counts": sc, "n_attempt": r.n_attempt, "n_accept": r.n_accept,
                "E_final": r.E_final, "elapsed_s": r.elapsed_s}

    def run_batch(self, occ, kind: str, temperatures, chem, n_passes: int, seeds, n_threads: int,
                  kb: float = 8.6173303e-05, full_recompute_samplers: bool = True):
        occ = np.ascontiguousarray(occ, dtype=np.int8).copy()
        nchain = occ.shape[0]
        p = orc_run_params()
        p.kind = 0 if kind == "canonical" else 1
        p.kb = kb
        p.n_occ_max = int(self.flat.n_occ_max)
        p.n_equil_passes = 0
        p.n_passes = n_passes
        p.full_recompute_samplers = 1 if full_recompute_samplers else 0
        p.chem = f64p()
        T = np.ascontiguousarray(temperatures, dtype=np.float64)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        chem_p = f64p()
        if chem is not None:
            chem_a = np.ascontiguousarray(chem, dtype=np.float64)
            chem_p = _p(chem_a, f64p)
        natt = C.c_longlong(0)
        el = C.c_double(0)
        lib().orc_run_batch(C.byref(self.s), C.c_int(nchain), _p(occ, i8p), _p(T, f64p), chem_p,
                            C.byref(p), seeds.ctypes.data_as(C.POINTER(C.c_uint64)),
                            C.c_int(n_threads), C.byref(natt), C.byref(el))
        return {"occ": occ, "n_attempt": natt.value, "elapsed_s": el.value}

    def enumerate(self, kind: str, temperature: float, chem=None, fixed_counts=None,
                  kb: float = 8.6173303e-05, want_probs: bool = False):
        nom = int(self.flat.n_occ_max)
        avg = np.zeros(6)
        chem_p = f64p()
        if chem is not None:
            chem_a = np.ascontiguousarray(chem, dtype=np.float64).reshape(self.flat.n_sublat * nom)
            chem_p = _p(chem_a, f64p)
        fc = np.zeros(16, dtype=np.int32)
        if fixed_counts is not None:
            fc[:len(fixed_counts)] = fixed_counts
        nst = 1
        for b in range(self.flat.n_sublat):
            nst *= int(self.flat.n_occ[b]) ** (self.dims[0] * self.dims[1] * self.dims[2])
        probs = np.zeros(nst) if want_probs else None
        rc = lib().orc_enumerate(C.byref(self.s), C.c_int(0 if kind == "canonical" else 1),
                                 C.c_double(temperature), C.c_double(kb), chem_p, C.c_int(nom),
                                 _p(fc, i32p), _p(avg, f64p),
                                 _p(probs, f64p) if probs is not None else f64p(),
                                 C.c_longlong(nst if probs is not None else 0))
        if rc != 0:
            raise RuntimeError(f"orc_enumerate failed rc={rc}")
        out = {"E": avg[0], "E2": avg[1], "n1": avg[2], "n12": avg[3], "En1": avg[4]}
        if probs is not None:
            out["probs"] = probs
        return out

    def state_energies(self, fixed_counts=None, n_threads: int = 8) -> np.ndarray:
        """Energy of every state in enumerate()'s order (site 0 fastest);
        NaN for states off fixed_counts."""
        nom = int(self.flat.n_occ_max)
        nst = 1
        for b in range(self.flat.n_sublat):
            nst *= int(self.flat.n_occ[b]) ** (self.dims[0] * self.dims[1] * self.dims[2])
        E = np.zeros(nst)
        fc_p = i32p()
        if fixed_counts is not None:
            fc = np.zeros(16, dtype=np.int32)
            fc[:len(fixed_counts)] = fixed_counts
            self._fc = fc
            fc_p = _p(fc, i32p)
        n_chunks = 16
        bounds = [nst * i // n_chunks for i in range(n_chunks + 1)]

        def run(i):
            return lib().orc_enumerate_energies(C.byref(self.s), C.c_int(nom), fc_p, _p(E, f64p), C.c_longlong(nst),
                                                C.c_longlong(bounds[i]), C.c_longlong(bounds[i + 1]))

        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max_workers=n_threads) as ex:
            rcs = list(ex.map(run, range(n_chunks)))
        if any(rcs):
            raise RuntimeError(f"orc_enumerate_energies failed rc={rcs}")
        return E


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def mt19937_64_nth(seed: int, n: int) -> int:
    return int(lib().orc_mt19937_64_nth(C.c_uint64(seed), C.c_longlong(n)))
