"""The serving operator cache (host only): on one structure (N, m, lmax) the
Ekman number enters A affinely and B not at all, so after two full
assemblies every later request is one axpy.  Checked against fresh
assemblies of the same requests; a structure change is a miss."""
import dataclasses

import numpy as np
import scipy.sparse as ss

from kore_amd import serving


def _req(rid, Ek, N=24, m=1, lmax=40):
    return serving.Request(rid=rid, kind="forced", arrival=0.0, Ek=Ek, N=N, m=m, lmax=lmax, omega=0.3, seed=rid)


def test_affine_in_ekman_matches_fresh_assembly():
    cache = serving.OperatorCache()
    es = [3e-5, 7e-6, 1.3e-5, 2.2e-6]
    for i, e in enumerate(es):
        par, A, B = cache.assemble(_req(i, e))
        _, A2, B2 = serving.assemble_request(_req(i, e))
        A2 = ss.csr_matrix(A2)
        # relative to the largest entry of the fresh operator: the affine
        # reconstruction differs from a fresh assembly by rounding only
        assert abs(A - A2).max() <= 1e-14 * abs(A2).max(), (i, e)
        assert abs(ss.csr_matrix(B) - ss.csr_matrix(B2)).max() == 0
        assert par == _req(i, e).params()
    assert cache.misses == 2 and cache.hits == 2


def test_other_structure_is_a_miss_and_lru_bound():
    cache = serving.OperatorCache(max_entries=1)
    cache.assemble(_req(0, 1e-5))
    cache.assemble(_req(1, 2e-5))                  # second of structure 1: affine entry made
    cache.assemble(_req(2, 1e-5, m=2, lmax=41))   # another structure: miss
    cache.assemble(_req(3, 2e-5, m=2, lmax=41))   # its affine entry evicts structure 1's
    cache.assemble(_req(4, 3e-5))                  # structure 1 again: miss (evicted)
    assert cache.hits == 0 and cache.misses == 5
    cache.assemble(_req(5, 4e-5, m=2, lmax=41))   # structure 2 still cached: hit
    assert cache.hits == 1


def test_planted_nonaffine_change_is_caught():
    """The parity check above fails for a matrix that does not follow the
    affine law: a change of another parameter (here the inner radius) under
    the same key is far outside its bar."""
    cache = serving.OperatorCache()
    r0, r1 = _req(0, 1e-5), _req(1, 2e-5)
    cache.assemble(r0)
    cache.assemble(r1)
    par, A, _ = cache.assemble(_req(2, 1.5e-5))
    from kore_amd import assembly
    A_other, _ = assembly.assemble(dataclasses.replace(par, ricb=0.4), which="unscaled")
    assert abs(A - ss.csr_matrix(A_other)).max() > 1e-6 * abs(A).max()


def test_operators_of_a_cached_structure_are_left_affine():
    """operators() gives a cached structure's A as A0, A1 and E (the server
    stages A0 + E A1 without forming A); formed, it is exactly what
    assemble() returns for the same request."""
    cache = serving.OperatorCache()
    cache.operators(_req(0, 3e-5))
    cache.operators(_req(1, 7e-6))
    _, A, B = cache.operators(_req(2, 1.3e-5))
    assert isinstance(A, serving.AffineA) and A.e == 1.3e-5
    F = A.materialize()
    _, G, _ = cache.assemble(_req(3, 1.3e-5))
    assert np.array_equal(F.indptr, G.indptr) and np.array_equal(F.indices, G.indices)
    assert np.array_equal(F.data, G.data)
    assert np.array_equal(F.data, A.a1 * A.e + A.a0)
