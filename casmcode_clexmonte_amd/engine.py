"""ctypes binding of the clexmc C ABI (include/clexmc.h).

The library ``libclexmc.so`` is built in-tree by ``__graft_entry__.build()``
(hipcc, gfx950).  There is no CPU fallback: if the library or a GPU is
missing, constructing an :class:`Engine` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

from .system import FlatCE, Supercell, color_classes, check_supercell, interaction_offsets

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclexmc.so")

i32p = C.POINTER(C.c_int)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)
i8p = C.POINTER(C.c_int8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i64p = C.POINTER(C.c_longlong)

CLEXMC_OK = 0
CLEXMC_ERR_NO_DEVICE = 5


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_sublat", C.c_int), ("n_occ", i32p), ("n_occ_max", C.c_int), ("n_func_max", C.c_int),
        ("site_func", f64p),
        ("n_orbits", C.c_int), ("orbit_k", i32p), ("orbit_mult", i32p), ("orbit_nfunc", i32p),
        ("orbit_func_begin", i32p), ("corr_orbit_begin", i32p), ("corr_elem_begin", i32p),
        ("corr_elem_sites", i32p), ("n_corr_elem_sites", C.c_int), ("orbit_table_offset", i32p),
        ("orbit_tables", f64p), ("n_orbit_tables", C.c_int), ("n_corr", C.c_int), ("eci", f64p),
        ("flower_begin", i32p), ("n_entries", C.c_int), ("entry_orbit", i32p), ("entry_k", i32p),
        ("entry_site_begin", i32p), ("entry_sites", i32p), ("n_entry_sites", C.c_int),
        ("entry_table", i32p), ("flower_tables", f64p), ("n_flower_tables", C.c_int),
        ("binary", C.c_int), ("bin_seg_begin", i32p), ("bin_segs", i32p), ("n_bin_segs", C.c_int),
        ("bin_orbit_weight", f64p), ("asym", i32p),
    ]


class SupercellDesc(C.Structure):
    _fields_ = [
        ("L", C.c_int * 3), ("n_classes", C.c_int), ("class_begin", i32p), ("class_sites", i32p),
        ("n_pair_phases", C.c_int), ("pair_phase_begin", i32p), ("pair_sites", i32p),
    ]


_lib = None


def lib():
    """Load the in-tree HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"clexmc HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.clexmc_last_error.restype = C.c_char_p
        L.clexmc_version.restype = C.c_char_p
        L.clexmc_create.argtypes = [C.POINTER(SystemDesc), C.c_int, C.POINTER(C.c_void_p)]
        for name in ["clexmc_add_supercell", "clexmc_load_chains", "clexmc_delta_energy_batch",
                     "clexmc_delta_energy_batch_noalloc", "clexmc_set_multiswaps",
                     "clexmc_correlations", "clexmc_run_passes", "clexmc_read_samples",
                     "clexmc_clear_samples", "clexmc_read_state", "clexmc_read_binary_counts",
                     "clexmc_last_run_ms", "clexmc_set_block_size", "clexmc_philox_device",
                     "clexmc_set_proposal", "clexmc_mlp_info", "clexmc_mlp_set_params", "clexmc_mlp_get_params",
                     "clexmc_mlp_forward", "clexmc_mlp_gradients", "clexmc_mlp_collect", "clexmc_mlp_train",
                     "clexmc_wave_pick_device", "clexmc_set_image_packing", "clexmc_set_global_swaps",
                     "clexmc_set_lds_row_pitch", "clexmc_nfold_run",
                     "clexmc_reanchor_energies", "clexmc_graph_capture", "clexmc_graph_replay",
                     "clexmc_graph_release", "clexmc_autotune", "clexmc_set_specialization",
                     "clexmc_last_run_info", "clexmc_specialization_log", "clexmc_lds_probe",
                     "clexmc_set_lanes_per_site", "clexmc_last_run_lanes", "clexmc_set_slabs", "clexmc_last_run_slabs",
                     "clexmc_set_fused_launch", "clexmc_last_run_fused", "clexmc_nfold_run_moments",
                     "clexmc_nfold_run_time", "clexmc_set_register_cube", "clexmc_last_run_cube",
                     "clexmc_set_nfold_block_size"]:
            getattr(L, name).restype = C.c_int
        L.clexmc_last_error.argtypes = [C.c_void_p]
        L.clexmc_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def exported_symbols() -> List[str]:
    """Names of the C-ABI entry points declared in include/clexmc.h."""
    return ["clexmc_create", "clexmc_add_supercell", "clexmc_load_chains",
            "clexmc_delta_energy_batch", "clexmc_delta_energy_batch_noalloc", "clexmc_set_multiswaps",
            "clexmc_correlations",
            "clexmc_run_passes",
            "clexmc_read_samples", "clexmc_clear_samples", "clexmc_read_state",
            "clexmc_read_binary_counts", "clexmc_last_run_ms", "clexmc_set_block_size",
            "clexmc_philox_device", "clexmc_last_error", "clexmc_destroy", "clexmc_version",
            "clexmc_set_proposal", "clexmc_mlp_info", "clexmc_mlp_set_params", "clexmc_mlp_get_params",
            "clexmc_mlp_forward", "clexmc_mlp_gradients", "clexmc_mlp_collect", "clexmc_mlp_train",
            "clexmc_wave_pick_device", "clexmc_set_image_packing", "clexmc_set_global_swaps",
            "clexmc_set_lds_row_pitch", "clexmc_nfold_run",
            "clexmc_reanchor_energies", "clexmc_graph_capture", "clexmc_graph_replay",
            "clexmc_graph_release", "clexmc_autotune", "clexmc_set_specialization",
            "clexmc_last_run_info", "clexmc_specialization_log", "clexmc_lds_probe",
            "clexmc_set_lanes_per_site", "clexmc_last_run_lanes", "clexmc_set_slabs", "clexmc_last_run_slabs",
            "clexmc_set_fused_launch", "clexmc_last_run_fused", "clexmc_nfold_run_moments",
            "clexmc_nfold_run_time", "clexmc_set_register_cube", "clexmc_last_run_cube",
                     "clexmc_set_nfold_block_size"]


# layout of clexmc_nfold_run_moments' records
NFOLD_MOMENTS = 21
NFOLD_PAIRS = [(o, q) for o in range(4) for q in range(o, 4)]  # index t of sum w N_o N_q is 7 + t


def _p(a, t):
    return a.ctypes.data_as(t)


def _check_cint(name: str, v) -> None:
    """Counts cross the C ABI as C ints: refuse values that would wrap."""
    if not 0 <= int(v) < 2 ** 31:
        raise ValueError(f"{name} = {v} is outside [0, 2^31)")


class ClexmcError(RuntimeError):
    pass


def pair_phases(flat: FlatCE, sc: Supercell, n_phases: int, seed: int = 0,
                max_tries: int = 64) -> List[np.ndarray]:
    """Canonical cross-class mixing phases: each phase is a set of disjoint
    site pairs (s on sublattice b, t on sublattice b2 in the cell translated
    by v) for random b, b2 (with the same number of allowed occupants, so
    that swaps can also move atoms between sublattices) and a random lattice
    translation v, chosen greedily so that no cluster contains sites of two
    different pairs and the two sites of a pair share no cluster.  Selection
    depends on positions only, never on occupations, so every pair move is a
    symmetric proposal."""
    prim = sc.prim
    offs = interaction_offsets(flat)
    rng = np.random.default_rng(seed)
    N = sc.n_sites
    L1, L2, L3 = sc.dims
    nc = sc.n_cells
    # neighbour index lists per site (interaction graph)
    nbr = {}
    for b in range(prim.n_sublat):
        if not prim.variable(b):
            continue
        o = np.array(offs[b], dtype=np.int64).reshape(-1, 4)
        cells = np.arange(nc)
        i = cells % L1
        j = (cells // L1) % L2
        k = cells // (L1 * L2)
        if len(o):
            ii = (i[:, None] + o[None, :, 1]) % L1
            jj = (j[:, None] + o[None, :, 2]) % L2
            kk = (k[:, None] + o[None, :, 3]) % L3
            nbr[b] = o[None, :, 0] * nc + ii + L1 * (jj + L2 * kk)
        else:
            nbr[b] = np.zeros((nc, 0), dtype=np.int64)
    phases: List[np.ndarray] = []
    tries = 0
    variable = [b for b in range(prim.n_sublat) if prim.variable(b)]
    while len(phases) < n_phases and tries < max_tries * max(1, n_phases):
        tries += 1
        b = variable[int(rng.integers(len(variable)))]
        b2 = variable[int(rng.integers(len(variable)))]
        if len(prim.occupants[b2]) != len(prim.occupants[b]):
            continue
        v = (int(rng.integers(L1)), int(rng.integers(L2)), int(rng.integers(L3)))
        if b2 == b and v == (0, 0, 0):
            continue
        # the two sites of a pair must not interact
        bad = False
        for (bo, di, dj, dk) in offs[b]:
            if bo == b2 and (di - v[0]) % L1 == 0 and (dj - v[1]) % L2 == 0 and (dk - v[2]) % L3 == 0:
                bad = True
                break
        if bad:
            continue
        blocked = np.zeros(N, dtype=bool)
        order = rng.permutation(nc)
        pairs = []
        for c in order:
            s = b * nc + int(c)
            i, j, k = c % L1, (c // L1) % L2, c // (L1 * L2)
            t = sc.linear_index(b2, i + v[0], j + v[1], k + v[2])
            if blocked[s] or blocked[t]:
                continue
            pairs.append((s, t))
            blocked[s] = blocked[t] = True
            blocked[nbr[b][s - b * nc]] = True
            blocked[nbr[b2][t - b2 * nc]] = True
        if pairs:
            phases.append(np.array(pairs, dtype=np.int32))
    return phases


class Engine:
    """One clexmc handle: a cluster expansion, supercells and loaded chains."""

    def __init__(self, flat: FlatCE, device: int = 0):
        self.flat = flat
        self._keep = []
        L = lib()
        d = SystemDesc()
        f = flat

        def k(a, dt):
            a = np.ascontiguousarray(a, dtype=dt).ravel()
            if a.size == 0:
                a = np.zeros(1, dtype=dt)
            self._keep.append(a)
            return a

        d.n_sublat = f.n_sublat
        d.n_occ = _p(k(f.n_occ, np.int32), i32p)
        d.n_occ_max = f.n_occ_max
        d.n_func_max = f.n_func_max
        d.site_func = _p(k(f.site_func, np.float64), f64p)
        d.n_orbits = f.n_orbits
        d.orbit_k = _p(k(f.orbit_k, np.int32), i32p)
        d.orbit_mult = _p(k(f.orbit_mult, np.int32), i32p)
        d.orbit_nfunc = _p(k(f.orbit_nfunc, np.int32), i32p)
        d.orbit_func_begin = _p(k(f.orbit_func_begin, np.int32), i32p)
        d.corr_orbit_begin = _p(k(f.corr_orbit_begin, np.int32), i32p)
        d.corr_elem_begin = _p(k(f.corr_elem_begin, np.int32), i32p)
        d.corr_elem_sites = _p(k(f.corr_elem_sites, np.int32), i32p)
        d.n_corr_elem_sites = len(f.corr_elem_sites)
        d.orbit_table_offset = _p(k(f.orbit_table_offset, np.int32), i32p)
        d.orbit_tables = _p(k(f.orbit_tables, np.float64), f64p)
        d.n_orbit_tables = f.orbit_tables.size
        d.n_corr = f.n_corr
        d.eci = _p(k(f.eci, np.float64), f64p)
        d.flower_begin = _p(k(f.flower_begin, np.int32), i32p)
        d.n_entries = f.n_entries
        d.entry_orbit = _p(k(f.entry_orbit, np.int32), i32p)
        d.entry_k = _p(k(f.entry_k, np.int32), i32p)
        d.entry_site_begin = _p(k(f.entry_site_begin, np.int32), i32p)
        d.entry_sites = _p(k(f.entry_sites, np.int32), i32p)
        d.n_entry_sites = len(f.entry_sites)
        d.entry_table = _p(k(f.entry_table, np.int32), i32p)
        d.flower_tables = _p(k(f.flower_tables, np.float64), f64p)
        d.n_flower_tables = f.flower_tables.size
        d.binary = 1 if f.binary else 0
        if f.binary:
            d.bin_seg_begin = _p(k(f.bin_seg_begin, np.int32), i32p)
            d.bin_segs = _p(k(f.bin_segs, np.int32), i32p)
            d.n_bin_segs = len(f.bin_segs)
            d.bin_orbit_weight = _p(k(f.bin_orbit_weight, np.float64), f64p)
        if getattr(f, "asym", None) is not None:
            d.asym = _p(k(f.asym, np.int32), i32p)
        self._desc = d
        h = C.c_void_p()
        rc = L.clexmc_create(C.byref(d), C.c_int(device), C.byref(h))
        if rc == CLEXMC_ERR_NO_DEVICE:
            raise ClexmcError("clexmc_create: no HIP device available")
        self.h = h
        if rc != CLEXMC_OK:
            msg = L.clexmc_last_error(h).decode() if h.value else "clexmc_create failed"
            self.close()
            raise ClexmcError(f"clexmc_create: {msg}")
        self.supercells: List[Supercell] = []
        self.n_chains = 0
        self.chain_sc: List[int] = []

    def _check(self, rc, what):
        if rc != CLEXMC_OK:
            msg = lib().clexmc_last_error(self.h).decode()
            raise ClexmcError(f"{what} failed (status {rc}): {msg}")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().clexmc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- supercells ---------------------------------------------------------
    def add_supercell(self, sc: Supercell, n_pair_phases: int = 8, seed: int = 0) -> int:
        check_supercell(self.flat, sc)
        classes = color_classes(self.flat, sc)
        cb = np.zeros(len(classes) + 1, dtype=np.int32)
        for i, c in enumerate(classes):
            cb[i + 1] = cb[i] + len(c)
        cs = np.concatenate(classes).astype(np.int32)
        phases = pair_phases(self.flat, sc, n_pair_phases, seed) if n_pair_phases > 0 else []
        pb = np.zeros(len(phases) + 1, dtype=np.int32)
        for i, p in enumerate(phases):
            pb[i + 1] = pb[i] + len(p)
        ps = np.concatenate(phases).astype(np.int32).ravel() if phases else np.zeros(2, np.int32)
        d = SupercellDesc()
        d.L[0], d.L[1], d.L[2] = sc.dims
        d.n_classes = len(classes)
        d.class_begin = _p(cb, i32p)
        d.class_sites = _p(cs, i32p)
        d.n_pair_phases = len(phases)
        d.pair_phase_begin = _p(pb, i32p)
        d.pair_sites = _p(ps, i32p)
        sid = C.c_int()
        self._check(lib().clexmc_add_supercell(self.h, C.byref(d), C.byref(sid)), "clexmc_add_supercell")
        self.supercells.append(sc)
        self.classes = classes
        return sid.value

    # -- chains ---------------------------------------------------------------
    def load_chains(self, supercell_ids: Sequence[int], occs: Sequence[np.ndarray], temperatures,
                    seeds, chem=None, kb: float = 8.6173303e-05):
        n = len(supercell_ids)
        sid = np.ascontiguousarray(supercell_ids, dtype=np.int32)
        occ = np.ascontiguousarray(np.concatenate([np.asarray(o, dtype=np.int8).ravel() for o in occs]))
        T = np.ascontiguousarray(temperatures, dtype=np.float64)
        sd = np.ascontiguousarray(seeds, dtype=np.uint64)
        if chem is not None:
            ch = np.ascontiguousarray(chem, dtype=np.float64).reshape(n, self.flat.n_sublat * self.flat.n_occ_max)
            chp = _p(ch, f64p)
        else:
            chp = f64p()
        self._check(lib().clexmc_load_chains(self.h, C.c_int(n), _p(sid, i32p), _p(occ, i8p), _p(T, f64p),
                                             chp, _p(sd, u64p), C.c_double(kb)), "clexmc_load_chains")
        self.n_chains = n
        self.chain_sc = list(int(x) for x in sid)

    def delta_energy_batch(self, chain, sites, new_occ, include_chem: bool = False) -> np.ndarray:
        sites = np.ascontiguousarray(sites, dtype=np.int32)
        if sites.ndim == 1:
            sites = sites[:, None]
        nw = np.ascontiguousarray(new_occ, dtype=np.int32).reshape(sites.shape)
        ch = np.ascontiguousarray(np.broadcast_to(np.asarray(chain, dtype=np.int32), (sites.shape[0],)))
        out = np.zeros(sites.shape[0], dtype=np.float32)
        self._check(lib().clexmc_delta_energy_batch(self.h, C.c_int(sites.shape[0]), _p(ch, i32p), _p(sites, i32p),
                                                    _p(nw, i32p), C.c_int(sites.shape[1]),
                                                    C.c_int(1 if include_chem else 0), _p(out, f32p)),
                    "clexmc_delta_energy_batch")
        return out

    def correlations(self, chain: int) -> np.ndarray:
        out = np.zeros(self.flat.n_corr)
        self._check(lib().clexmc_correlations(self.h, C.c_int(chain), _p(out, f64p)), "clexmc_correlations")
        return out

    def run_passes(self, kind: str, n_passes: int, sample_period: int = 1, canonical_rounds: int = 1) -> float:
        """Run passes on every loaded chain; returns the kernel time in ms.
        canonical_rounds: random matchings per independent set and pass, or
        0 for the automatic count (about one attempt between different
        occupants per mutable site and pass, as the reference's passes)."""
        k = 0 if kind == "canonical" else 1
        _check_cint("n_passes", n_passes)
        _check_cint("sample_period", sample_period)
        self._check(lib().clexmc_run_passes(self.h, C.c_int(k), C.c_int(n_passes), C.c_int(sample_period),
                                            C.c_int(canonical_rounds)), "clexmc_run_passes")
        ms = C.c_double()
        self._check(lib().clexmc_last_run_ms(self.h, C.byref(ms)), "clexmc_last_run_ms")
        return ms.value

    def set_block_size(self, threads: int):
        self._check(lib().clexmc_set_block_size(self.h, C.c_int(threads)), "clexmc_set_block_size")

    def set_nfold_block_size(self, threads: int):
        """Threads per chain of the N-fold kernel: 64, 128, 256, or 0 (default)
        for the smallest multiple of 64 covering the neighbour list."""
        self._check(lib().clexmc_set_nfold_block_size(self.h, C.c_int(threads)), "clexmc_set_nfold_block_size")

    def read_samples(self):
        ns = C.c_int(0)
        self._check(lib().clexmc_read_samples(self.h, C.byref(ns), f64p(), C.POINTER(C.c_int32)()),
                    "clexmc_read_samples")
        n = ns.value
        E = np.zeros((self.n_chains, max(n, 1)))
        cnt = np.zeros((self.n_chains, max(n, 1), 4), dtype=np.int32)
        ns2 = C.c_int(max(n, 1))
        self._check(lib().clexmc_read_samples(self.h, C.byref(ns2), _p(E, f64p), _p(cnt, C.POINTER(C.c_int32))),
                    "clexmc_read_samples")
        return E[:, :n], cnt[:, :n, :]

    def clear_samples(self):
        self._check(lib().clexmc_clear_samples(self.h), "clexmc_clear_samples")

    def read_state(self, chain: int):
        sc = self.supercells[self.chain_sc[chain]]
        occ = np.zeros(sc.n_sites, dtype=np.int8)
        E = C.c_double()
        na = C.c_longlong()
        nacc = C.c_longlong()
        self._check(lib().clexmc_read_state(self.h, C.c_int(chain), _p(occ, i8p), C.byref(E), C.byref(na),
                                            C.byref(nacc)), "clexmc_read_state")
        return {"occ": occ, "E": E.value, "n_attempt": na.value, "n_accept": nacc.value}

    def read_binary_counts(self, chain: int) -> np.ndarray:
        out = np.zeros(self.flat.n_orbits, dtype=np.int64)
        self._check(lib().clexmc_read_binary_counts(self.h, C.c_int(chain), _p(out, i64p)),
                    "clexmc_read_binary_counts")
        return out

    def philox(self, ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
        ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
        key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
        out = np.zeros_like(ctr)
        self._check(lib().clexmc_philox_device(self.h, C.c_int(ctr.shape[0]), _p(ctr, u32p), _p(key, u32p),
                                               _p(out, u32p)), "clexmc_philox_device")
        return out

    # -- learned proposal ------------------------------------------------------
    def set_proposal(self, proposal: str, rounds_per_group: int = 64):
        code = {"plain": 0, "learned": 1, "grouped_uniform": 2, "informed": 3}[proposal]
        self._check(lib().clexmc_set_proposal(self.h, C.c_int(code), C.c_int(rounds_per_group)),
                    "clexmc_set_proposal")

    def mlp_info(self):
        n_env = C.c_int()
        n_params = C.c_int()
        env = np.zeros((self.flat.n_sublat, 48, 4), dtype=np.int32)
        self._check(lib().clexmc_mlp_info(self.h, C.byref(n_env), C.byref(n_params), _p(env, i32p)),
                    "clexmc_mlp_info")
        return n_env.value, n_params.value, env

    def mlp_set_params(self, params: np.ndarray, eps: float = 0.05):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(lib().clexmc_mlp_set_params(self.h, _p(p, f32p), C.c_float(eps)), "clexmc_mlp_set_params")

    def mlp_get_params(self) -> np.ndarray:
        _, n_params, _ = self.mlp_info()
        p = np.zeros(n_params, dtype=np.float32)
        self._check(lib().clexmc_mlp_get_params(self.h, _p(p, f32p)), "clexmc_mlp_get_params")
        return p

    def mlp_forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.int8).reshape(-1, 48)
        out = np.zeros((X.shape[0], 2), dtype=np.float32)
        self._check(lib().clexmc_mlp_forward(self.h, C.c_int(X.shape[0]), _p(X, i8p), _p(out, f32p)),
                    "clexmc_mlp_forward")
        return out

    def mlp_gradients(self, X: np.ndarray, y: np.ndarray, t: Optional[np.ndarray] = None):
        X = np.ascontiguousarray(X, dtype=np.int8).reshape(-1, 48)
        y = np.ascontiguousarray(y, dtype=np.float32)
        tp = f32p()
        if t is not None:
            t = np.ascontiguousarray(t, dtype=np.float32)
            tp = _p(t, f32p)
        _, n_params, _ = self.mlp_info()
        g = np.zeros(n_params, dtype=np.float32)
        loss = C.c_float()
        self._check(lib().clexmc_mlp_gradients(self.h, C.c_int(X.shape[0]), _p(X, i8p), _p(y, f32p), tp,
                                               _p(g, f32p), C.byref(loss)), "clexmc_mlp_gradients")
        return g, loss.value

    def mlp_collect(self, n_per_chain: int, seed: int = 0, include_chem: bool = False):
        n = n_per_chain * self.n_chains
        X = np.zeros((n, 48), dtype=np.int8)
        y = np.zeros(n, dtype=np.float32)
        t = np.zeros(n, dtype=np.float32)
        self._check(lib().clexmc_mlp_collect(self.h, C.c_int(n_per_chain), C.c_uint32(seed),
                                             C.c_int(1 if include_chem else 0), _p(X, i8p), _p(y, f32p),
                                             _p(t, f32p)), "clexmc_mlp_collect")
        return X, y, t

    def mlp_train(self, n_per_chain: int, n_steps: int, lr: float = 1e-3, seed: int = 0,
                  include_chem: bool = False) -> np.ndarray:
        loss = np.zeros(n_steps, dtype=np.float32)
        self._check(lib().clexmc_mlp_train(self.h, C.c_int(n_per_chain), C.c_int(n_steps), C.c_float(lr),
                                           C.c_uint32(seed), C.c_int(1 if include_chem else 0), _p(loss, f32p)),
                    "clexmc_mlp_train")
        return loss

    def wave_pick(self, v: np.ndarray, u: np.ndarray):
        v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 64)
        u = np.ascontiguousarray(u, dtype=np.float32)
        scan = np.zeros_like(v)
        pick = np.zeros(v.shape[0], dtype=np.int32)
        self._check(lib().clexmc_wave_pick_device(self.h, C.c_int(v.shape[0]), _p(v, f32p), _p(u, f32p),
                                                  _p(scan, f32p), _p(pick, i32p)), "clexmc_wave_pick_device")
        return scan, pick

    def set_image_packing(self, mode: str):
        code = {"auto": 0, "int8": 1, "packed2": 2}[mode]
        self._check(lib().clexmc_set_image_packing(self.h, C.c_int(code)), "clexmc_set_image_packing")

    def nfold_run(self, n_events: int, events_per_block: int) -> np.ndarray:
        """N-fold way semi-grand canonical events on every chain.
        Returns [n_chains, n_events // events_per_block, 6]: per block of
        events (sum w, sum w E, sum w N1, sum w E^2, sum w N1^2, sum w E N1),
        w the expected residence time of each visited state in Metropolis
        steps, E the formation energy per supercell, N1 the sites holding
        occupant 1."""
        _check_cint("n_events", n_events)
        _check_cint("events_per_block", events_per_block)
        n_blocks = n_events // events_per_block
        out = np.zeros((self.n_chains, n_blocks, 6))
        self._check(lib().clexmc_nfold_run(self.h, C.c_int(n_events), C.c_int(events_per_block), _p(out, f64p)),
                    "clexmc_nfold_run")
        return out

    def nfold_run_moments(self, n_events: int, events_per_block: int) -> np.ndarray:
        """The same run with the moments of every occupant count: returns
        [n_chains, n_events // events_per_block, 21] with per block
        [0] sum w, [1] sum w E, [2] sum w E^2, [3 + o] sum w N_o,
        [7 + t] sum w N_o N_q (o <= q in row order), [17 + o] sum w E N_o."""
        _check_cint("n_events", n_events)
        _check_cint("events_per_block", events_per_block)
        n_blocks = n_events // events_per_block
        out = np.zeros((self.n_chains, n_blocks, NFOLD_MOMENTS))
        self._check(lib().clexmc_nfold_run_moments(self.h, C.c_int(n_events), C.c_int(events_per_block),
                                                   _p(out, f64p)), "clexmc_nfold_run_moments")
        return out

    def nfold_run_time(self, n_samples: int, period: float, max_events: int):
        """N-fold way sampled by time: every chain records n_samples samples
        (energy per supercell and occupant counts, read with read_samples)
        at times 0, period, 2 period, ... in passes (expected Metropolis steps
        per mutable site), each the state occupied at that time; at most
        max_events events per chain."""
        _check_cint("n_samples", n_samples)
        _check_cint("max_events", max_events)
        self._check(lib().clexmc_nfold_run_time(self.h, C.c_int(n_samples), C.c_double(period), C.c_int(max_events)),
                    "clexmc_nfold_run_time")

    def set_lds_row_pitch(self, extra: int = -1):
        """Padded-image row pitch of supercells added later: -1 automatic
        (bank-conflict model), or a fixed number of extra bytes per row."""
        self._check(lib().clexmc_set_lds_row_pitch(self.h, C.c_int(extra)), "clexmc_set_lds_row_pitch")

    def set_slabs(self, slabs: int):
        """Workgroups per chain: 0 automatic, 1 off, >= 2 requested."""
        self._check(lib().clexmc_set_slabs(self.h, C.c_int(slabs)), "clexmc_set_slabs")

    def last_run_slabs(self) -> int:
        v = C.c_int()
        self._check(lib().clexmc_last_run_slabs(self.h, C.byref(v)), "clexmc_last_run_slabs")
        return v.value

    def set_register_cube(self, mode: int = 1):
        """Register-cube sweep of the specialised semi-grand canonical binary
        kernel where the supercell allows it (1, default) or the class loop (0);
        bit-identical results."""
        self._check(lib().clexmc_set_register_cube(self.h, C.c_int(mode)), "clexmc_set_register_cube")

    def last_run_cube(self) -> bool:
        v = C.c_int()
        self._check(lib().clexmc_last_run_cube(self.h, C.byref(v)), "clexmc_last_run_cube")
        return bool(v.value)

    def set_fused_launch(self, mode: int = 1):
        """Chains on several supercell shapes: 1 one fused launch (default),
        0 one launch per shape on concurrent streams."""
        self._check(lib().clexmc_set_fused_launch(self.h, C.c_int(mode)), "clexmc_set_fused_launch")

    def last_run_fused(self) -> bool:
        v = C.c_int()
        self._check(lib().clexmc_last_run_fused(self.h, C.byref(v)), "clexmc_last_run_fused")
        return bool(v.value)

    def set_multiswaps(self, types, per_pass: int = 0):
        """Semi-grand canonical multiswaps: `types` is a list of multiswap
        types, each a list of parts (asymmetric unit, occupant from, occupant
        to, count); the list must hold the type undoing each type
        (make_multiswaps builds such lists).  An empty list restores
        single-site changes.  per_pass: attempts per pass (0: one per
        mutable site)."""
        types = [list(t) for t in types]
        begin = np.zeros(len(types) + 1, dtype=np.int32)
        for i, t in enumerate(types):
            begin[i + 1] = begin[i] + len(t)
        parts = np.array([list(p) for t in types for p in t], dtype=np.int32).reshape(-1, 4)
        if parts.size == 0:
            parts = np.zeros((1, 4), dtype=np.int32)
        self._check(lib().clexmc_set_multiswaps(self.h, C.c_int(len(types)), _p(begin, i32p), _p(parts, i32p),
                                                C.c_int(per_pass)), "clexmc_set_multiswaps")

    def set_global_swaps(self, per_pass: int):
        self._check(lib().clexmc_set_global_swaps(self.h, C.c_int(per_pass)), "clexmc_set_global_swaps")

    def reanchor_energies(self) -> float:
        d = C.c_double()
        self._check(lib().clexmc_reanchor_energies(self.h, C.byref(d)), "clexmc_reanchor_energies")
        return d.value

    # -- HIP graph of a block of passes ------------------------------------------
    def graph_capture(self, kind: str, passes_per_replay: int, max_replays: int, sample_period: int = 1,
                      canonical_rounds: int = 2):
        k = 0 if kind == "canonical" else 1
        self._check(lib().clexmc_graph_capture(self.h, C.c_int(k), C.c_int(passes_per_replay), C.c_int(max_replays),
                                               C.c_int(sample_period), C.c_int(canonical_rounds)),
                    "clexmc_graph_capture")

    def graph_replay(self, n_replays: int = 1) -> float:
        self._check(lib().clexmc_graph_replay(self.h, C.c_int(n_replays)), "clexmc_graph_replay")
        ms = C.c_double()
        lib().clexmc_last_run_ms(self.h, C.byref(ms))
        return ms.value

    def graph_release(self):
        self._check(lib().clexmc_graph_release(self.h), "clexmc_graph_release")

    def autotune(self, kind: str, trial_passes: int = 20, canonical_rounds: int = 2):
        k = 0 if kind == "canonical" else 1
        t = C.c_int()
        p = C.c_int()
        ln = C.c_int()
        ms = C.c_double()
        self._check(lib().clexmc_autotune(self.h, C.c_int(k), C.c_int(trial_passes), C.c_int(canonical_rounds),
                                          C.byref(t), C.byref(p), C.byref(ln), C.byref(ms)), "clexmc_autotune")
        return {"threads": t.value, "packing": {1: "int8", 2: "packed2"}.get(p.value, "auto"),
                "lanes_per_site": ln.value, "ms_per_pass": ms.value}

    def set_specialization(self, on: bool):
        self._check(lib().clexmc_set_specialization(self.h, C.c_int(1 if on else 0)), "clexmc_set_specialization")

    def last_run_info(self):
        sp, pk, th, ln = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._check(lib().clexmc_last_run_info(self.h, C.byref(sp), C.byref(pk), C.byref(th)), "clexmc_last_run_info")
        self._check(lib().clexmc_last_run_lanes(self.h, C.byref(ln)), "clexmc_last_run_lanes")
        return {"specialized": bool(sp.value), "packed": bool(pk.value), "threads": th.value,
                "lanes_per_site": ln.value}

    def set_lanes_per_site(self, lanes: int):
        self._check(lib().clexmc_set_lanes_per_site(self.h, C.c_int(lanes)), "clexmc_set_lanes_per_site")

    def specialization_log(self) -> str:
        buf = C.create_string_buffer(1 << 16)
        self._check(lib().clexmc_specialization_log(self.h, buf, C.c_int(1 << 16)), "clexmc_specialization_log")
        return buf.value.decode(errors="replace")

    def lds_probe(self, mode: str) -> float:
        g = C.c_double()
        self._check(lib().clexmc_lds_probe(self.h, C.c_int({"u8": 0, "b128": 1}[mode]), C.byref(g)),
                    "clexmc_lds_probe")
        return g.value
