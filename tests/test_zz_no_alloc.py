"""clexmc_delta_energy_batch_noalloc (the entry point that allocates no
device memory) against clexmc_delta_energy_batch, through the C ABI: two
calls in a row on one handle, the second smaller (and with other proposal
shapes), each bit-identical to the allocating entry point (same kernel, same
inputs); requests above the staging array's maximum are refused.  Also a
few results against the oracle (FP32 dE within 1e-5 of the batch scale +
1e-6 eV).  Runs alone in a fresh process as well as after other tests."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd.engine import Engine, lib
from casmcode_clexmonte_amd.system import FlatCE, Supercell
from casmcode_clexmonte_amd.workloads import fcc_binary_ce, random_occupation

pytestmark = pytest.mark.gpu

i32p = C.POINTER(C.c_int32)
f32p = C.POINTER(C.c_float)
MAX_PROPOSALS = 65536


def _call(fn, eng, chain, sites, new_occ, include_chem):
    n, m = sites.shape
    out = np.full(n, np.nan, dtype=np.float32)
    ch = np.ascontiguousarray(chain, dtype=np.int32)
    st = np.ascontiguousarray(sites, dtype=np.int32)
    nw = np.ascontiguousarray(new_occ, dtype=np.int32)
    rc = fn(eng.h, C.c_int(n), ch.ctypes.data_as(i32p), st.ctypes.data_as(i32p), nw.ctypes.data_as(i32p),
            C.c_int(m), C.c_int(include_chem), out.ctypes.data_as(f32p))
    return rc, out


def test_noalloc_delta_batch_matches_allocating_entry_point():
    print("[case] no-alloc delta batch", flush=True)
    L = lib()
    L.clexmc_delta_energy_batch_noalloc.restype = C.c_int
    flat = FlatCE(fcc_binary_ce())
    sc = Supercell(flat.ce.prim, (12, 12, 12))
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    nch = 4
    occs = [random_occupation(sc.n_sites, 0.4 + 0.05 * c, 70 + c) for c in range(nch)]
    eng.load_chains([sid] * nch, occs, [600.0] * nch, 100 + np.arange(nch),
                    chem=[[0.0, 0.01 * c] for c in range(nch)])
    rng = np.random.default_rng(4)
    cases = [(40000, 2, 0), (300, 1, 1)]   # the second call smaller, other shape, with chem
    osys = O.OracleSystem(flat, sc.dims)
    for n, m, chem in cases:
        chain = rng.integers(0, nch, n)
        sites = np.stack([rng.choice(sc.n_sites, m, replace=False) for _ in range(n)])
        new_occ = np.array([[1 - occs[c][s] for s in row] for c, row in zip(chain, sites)])
        rc0, ref = _call(L.clexmc_delta_energy_batch, eng, chain, sites, new_occ, chem)
        rc1, got = _call(L.clexmc_delta_energy_batch_noalloc, eng, chain, sites, new_occ, chem)
        assert rc0 == 0 and rc1 == 0, L.clexmc_last_error(eng.h)
        assert np.array_equal(got, ref), (n, m, np.abs(got - ref).max())
        # a few against the oracle
        o = []
        for p in range(0, n, max(1, n // 50)):
            c = int(chain[p])
            d = osys.delta(occs[c], list(map(int, sites[p])), list(map(int, new_occ[p])))
            if chem:
                d -= 0.01 * c * float(np.sum(new_occ[p] - occs[c][sites[p]]))
            o.append((got[p], d))
        o = np.array(o)
        assert np.all(np.abs(o[:, 0] - o[:, 1]) <= 1e-5 * np.abs(o[:, 1]).max() + 1e-6)
    # beyond the maximum: refused with a message, nothing launched
    n = MAX_PROPOSALS + 1
    big = np.zeros((n, 1), dtype=np.int32)
    rc, out = _call(L.clexmc_delta_energy_batch_noalloc, eng, np.zeros(n), big, big, 0)
    assert rc == 1 and "exceed the maximum" in L.clexmc_last_error(eng.h).decode()
    assert np.all(np.isnan(out))
    eng.close()
