"""The slab decomposition's plane sets, checked by brute force on the CPU.

A chain cut into S slabs along k runs on S workgroups; slab s of a launch
with boundary offset k0 updates the planes [k0 + s0, k0 + s0 + U) (mod L3),
s0 = floor(s L3 / S), U = floor((s + 1) L3 / S) - s0 - kreach, and reads the
planes within kreach of them.  For every supercell height, reach, slab count
the library may choose (at most L3 / (kreach + 1)) and offset:
* two sites on updated planes of different slabs are more than kreach planes
  apart (cyclically), so no cluster holds both;
* every plane a slab reads is either its own updated plane or a plane that no
  slab updates in that launch (frozen);
* over the offsets, every plane is updated by some launch.
The reaches are those of the workloads' cluster expansions (largest |dk| of
a flower site), computed here from their interaction offsets.
"""
import pytest

from casmcode_clexmonte_amd.system import FlatCE, interaction_offsets
from casmcode_clexmonte_amd.workloads import bcc_quaternary_ce, fcc_binary_ce, hcp_ternary_ce


def slab_planes(L3, S, kreach, k0, s):
    s0 = (s * L3) // S
    s1 = ((s + 1) * L3) // S
    U = s1 - s0 - kreach
    return [(k0 + s0 + d) % L3 for d in range(U)]


def cyc(a, b, L3):
    d = abs(a - b) % L3
    return min(d, L3 - d)


@pytest.mark.parametrize("name,ce_fn", [("fcc", fcc_binary_ce), ("hcp", hcp_ternary_ce), ("bcc", bcc_quaternary_ce)])
def test_slab_planes_independent_and_covering(name, ce_fn):
    offs = interaction_offsets(FlatCE(ce_fn()))
    kreach = max(abs(o[3]) for b in offs for o in offs[b])
    assert kreach == {"fcc": 2, "hcp": 1, "bcc": 3}[name]
    for L3 in (8, 10, 12, 16, 24, 32, 64):
        for S in range(2, L3 // (kreach + 1) + 1):
            covered = set()
            for k0 in range(L3):
                upd = [slab_planes(L3, S, kreach, k0, s) for s in range(S)]
                owner = {}
                for s, planes in enumerate(upd):
                    assert planes, (L3, S, s)
                    for k in planes:
                        assert k not in owner
                        owner[k] = s
                covered |= set(owner)
                for s, planes in enumerate(upd):
                    for t in range(S):
                        if t == s:
                            continue
                        for a in planes:
                            for b in upd[t]:
                                assert cyc(a, b, L3) > kreach, (L3, S, k0, s, t, a, b)
                    # reads of slab s: its own planes or frozen ones
                    for a in planes:
                        for d in range(-kreach, kreach + 1):
                            k = (a + d) % L3
                            assert owner.get(k, s) == s, (L3, S, k0, s, k)
            assert covered == set(range(L3))
