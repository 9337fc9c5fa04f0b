"""Large device blocks made of physical chunks.  A context created with 4 MiB
chunks (SHCI_CHUNK_MIB; 1 GiB by default) runs the builds of a growing det
list and then of a smaller one, so its blocks are extended in place, shrunk,
taken apart and mapped again from the context's pool.  Its matrices and
products must equal the oracle's, and those of a context whose blocks at
these sizes are all plain hipMalloc blocks."""
import itertools

import numpy as np
import pytest

import oracle
from shci_amd import integrals as si


def _full_list(n, nu, nd, seed):
    strs_u = [sum(1 << k for k in c) for c in itertools.combinations(range(n), nu)]
    strs_d = [sum(1 << k for k in c) for c in itertools.combinations(range(n), nd)]
    hf = ((1 << nu) - 1, (1 << nd) - 1)
    rest = [(u, d) for u in strs_u for d in strs_d if (u, d) != hf]
    order = np.random.default_rng(seed).permutation(len(rest))
    dets = [hf] + [rest[i] for i in order]
    return (np.array([d[0] for d in dets], dtype=np.uint64), np.array([d[1] for d in dets], dtype=np.uint64))


@pytest.mark.gpu
def test_chunked_blocks_growing_and_shrinking(monkeypatch):
    from shci_amd import hamiltonian as hm

    n, nu, nd = 12, 3, 3
    _, h1, eri, _ = si.synthetic(n, seed=21)
    up, dn = _full_list(n, nu, nd, seed=3)
    plain = hm.Context.get(0)
    monkeypatch.setenv("SHCI_CHUNK_MIB", "4")
    small = hm.Context(0)
    monkeypatch.delenv("SHCI_CHUNK_MIB")
    base_mem = None
    rng = np.random.default_rng(9)
    hams = {}
    try:
        # a growing list (append-only updates), then a new Hamiltonian over a
        # shorter list on the same contexts (every large block shrinks)
        for step, size in enumerate([6000, 20000, len(up), 12000]):
            if step == 0 or step == 3:
                for name, ctx in (("plain", plain), ("small", small)):
                    monkeypatch.setitem(hm.Context._by_device, 0, ctx)
                    hams[name] = hm.Hamiltonian(n, h1, eri, 0.0, store="both", layout="segments")
                if base_mem is None:
                    base_mem = small.memory()  # the integrals
            u, d = up[:size], dn[:size]
            x = rng.uniform(-1, 1, size)
            orc = oracle.System.from_arrays(n, nu, nd, h1, eri, 0.0)
            orc.update(u, d)
            y_ref = orc.mul(x)
            res = {}
            for name, ctx in (("plain", plain), ("small", small)):
                monkeypatch.setitem(hm.Context._by_device, 0, ctx)
                ham = hams[name]
                ham.update(u, d)
                res[name] = (ham.matrix.mul(x), ham.matrix.mul(x, fp32=True), ham.matrix.count_n_elems(),
                             ham.matrix.segment_ranges())
            print(f"{size} dets: {res['small'][2]} elements; small-chunk context holds "
                  f"{small.memory() / 2**20:.0f} MiB", flush=True)
            assert res["small"][2] == res["plain"][2]
            assert res["small"][3] == res["plain"][3]
            np.testing.assert_allclose(res["small"][0], y_ref, rtol=1e-12, atol=1e-11)
            np.testing.assert_allclose(res["plain"][0], y_ref, rtol=1e-12, atol=1e-11)
            scale = np.abs(y_ref).max()
            np.testing.assert_allclose(res["small"][1], res["plain"][1], rtol=0, atol=1e-5 * scale)
            np.testing.assert_allclose(res["small"][1], y_ref, rtol=0, atol=2e-4 * scale)
        assert small.memory() > base_mem + 64 * 2**20  # the store alone is over 16 chunks
        init = np.zeros((1, 12000))
        init[0, 0] = 1.0
        e = {}
        for name, ctx in (("plain", plain), ("small", small)):
            monkeypatch.setitem(hm.Context._by_device, 0, ctx)
            dav = hm.Davidson(1)
            dav.diagonalize(hams[name].matrix, init, 1e-10, max_calls=40)
            e[name] = dav.lowest_eigenvalues[0]
        assert abs(e["small"] - e["plain"]) < 1e-10
        monkeypatch.setitem(hm.Context._by_device, 0, small)
        hams["small"].clear()
        # cached blocks and the chunk pool went back to the driver (a few small
        # working buffers stay)
        assert small.memory() <= base_mem + 16 * 2**20
    finally:
        small.close()
