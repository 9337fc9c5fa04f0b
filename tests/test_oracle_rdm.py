"""Pins the RDMs (src/chem/rdm.cc) by identities and by the Hellmann-Feynman
theorem: tr(gamma) = N, sum_k D_pkks = (N - 1) gamma_ps (rdm.cc:687-699),
E = core + sum h gamma + 1/2 sum D_pqrs (ps|qr) (rdm.cc:731-745) equals the
eigenvalue, and central finite differences of the eigenvalue in the integrals
equal the RDM contractions."""
import numpy as np

import oracle
from bruteforce import dets_all
from shci_amd import integrals as si


def _fci(n, nu, nd, h1, eri, core=0.0):
    dets = dets_all(n, nu, nd)
    hf = ((1 << nu) - 1, (1 << nd) - 1)
    dets.remove(hf)
    dets = [hf] + dets
    up = np.array([d[0] for d in dets], dtype=np.uint64)
    dn = np.array([d[1] for d in dets], dtype=np.uint64)
    orc = oracle.System.from_arrays(n, nu, nd, h1, eri, core)
    orc.update(up, dn)
    vec = np.zeros((1, len(dets)))
    vec[0, 0] = 1.0
    for _ in range(80):
        ev, vec, conv, _, _ = orc.davidson(vec, 1e-13)
        if conv:
            break
    return orc, ev[0], vec


def energy_from_rdm(n, h1, eri, core, r1, r2):
    D = oracle.unpack_2rdm(r2, n)
    v = si.unpack_eri(eri, n)
    # E = core + sum_pq h_pq g_pq + 1/2 sum_pqrs D_pqrs (ps|qr)
    two = 0.5 * np.einsum("pqrs,psqr->", D, v)
    return core + np.sum(h1 * r1) + two


def grad_eri_from_rdm(n, r2):
    """dE/d eri[K] for the packed parameterisation: 1/2 sum over (p,q,r,s)
    with combine4(p,s,q,r) == K of D_pqrs."""
    D = oracle.unpack_2rdm(r2, n)
    p, q, r, s = np.meshgrid(*[np.arange(n)] * 4, indexing="ij")
    keys = si.combine4(p, s, q, r).ravel()
    npair = n * (n + 1) // 2
    return 0.5 * np.bincount(keys, weights=D.ravel(), minlength=npair * (npair + 1) // 2)


def test_rdm_identities_and_energy():
    n, nu, nd = 6, 3, 2
    _, h1, eri, _ = si.synthetic(n, seed=17)
    core = 1.25
    orc, e, vec = _fci(n, nu, nd, h1, eri, core)
    r1, r2 = orc.rdm(vec)
    N = nu + nd
    assert abs(np.trace(r1) - N) < 1e-10
    D = oracle.unpack_2rdm(r2, n)
    np.testing.assert_allclose(np.einsum("pkks->ps", D) / (N - 1), r1, atol=1e-10)
    assert abs(energy_from_rdm(n, h1, eri, core, r1, r2) - e) < 1e-10


def test_hellmann_feynman_finite_differences():
    n, nu, nd = 5, 2, 2
    _, h1, eri, _ = si.synthetic(n, seed=23)
    orc, e0, vec = _fci(n, nu, nd, h1, eri)
    r1, r2 = orc.rdm(vec)
    g_eri = grad_eri_from_rdm(n, r2)
    g_h = r1 + r1.T - np.diag(np.diag(r1))  # symmetric key combine2(p,q)
    rng = np.random.default_rng(3)
    step = 1e-5
    for _ in range(4):
        dh = rng.normal(size=(n, n))
        dh = dh + dh.T
        de = rng.normal(size=eri.shape)
        analytic = np.sum(np.tril(dh) * np.tril(g_h)) + de @ g_eri

        def E(sign):
            o, e, _ = _fci(n, nu, nd, h1 + sign * step * dh, eri + sign * step * de)
            return e

        fd = (E(1) - E(-1)) / (2 * step)
        assert abs(fd - analytic) <= 1e-6 * max(1.0, abs(analytic)), (fd, analytic)
