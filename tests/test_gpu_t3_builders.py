"""The batched T3 builders ``pycc_amd.cctriples.t3c_ijk / t3c_abc / t3d_ijk / t3d_abc``
(``pycc/cctriples.py:27-173``) against the oracle's ``t3c_ijk`` and ``t3d_ijk`` (pinned by the
reference's water (T) energies through ``t_tjl``) and its ``denominator``.

Inputs: a synthetic Hamiltonian with a non-diagonal Fock (f_ov != 0 reaches the t2 f_kc terms), random
t1 and pair-symmetric t2, a frozen-core offset in one case, and the W blocks passed as strided views of
the full <pq|rs> tensor on the device, as ``pycc``'s callers pass ``ERI[v, v, v, o]``.

Bars: every (i, j, k) and every (a, b, c) batch within 1e-12 of the oracle, relative to the batch's
largest element (FP64 MFMA contractions against numpy's einsum); the a, b, c batches are the i, j, k
ones read along the other three indices (the reference gives both the same amplitude); the
WithDenom=True output equals the WithDenom=False output divided by the oracle's denominator bit for
bit (one IEEE division per element, the same left-to-right sums of three)."""
import itertools

import numpy as np
import pytest

from conftest import random_amplitudes
from oracle.triples_ref import denominator, t3c_ijk, t3d_ijk
from pycc_amd.synth import synthetic_df, dirac_eri

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pycc_amd():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycc_amd
    return pycc_amd


def _dev(x):
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


def _close(got, ref):
    got = got.cpu().numpy()
    return np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("no,nv,nfzc,seed", [(3, 7, 0, 31), (4, 5, 1, 32)])
def test_t3_builders_match_oracle(pycc_amd, no, nv, nfzc, seed):
    from pycc_amd import cctriples
    print("T3 builders no=%d nv=%d nfzc=%d" % (no, nv, nfzc), flush=True)
    sd = synthetic_df(nfzc + no, nv, seed=seed, t2max=0.1)
    X = np.random.default_rng(seed + 100).standard_normal(sd.fock.shape) * 0.05
    F = sd.fock + X + X.T
    ERI = dirac_eri(sd.B)
    nmo = F.shape[0]
    o, v = slice(nfzc, nfzc + no), slice(nfzc + no, nmo)
    t1, t2 = random_amplitudes(no, nv, seed, scale=0.1)
    Fo, Fv = np.diag(F)[o], np.diag(F)[v]
    W = {"vvvo": ERI[v, v, v, o], "ovoo": ERI[o, v, o, o], "oovv": ERI[o, o, v, v]}
    Fd, Ed, t1d, t2d = _dev(F), _dev(ERI), _dev(t1), _dev(t2)
    Wvvvo, Wovoo, Woovv = Ed[v, v, v, o], Ed[o, v, o, o], Ed[o, o, v, v]   # strided views
    contract = pycc_amd.DeviceManager('GPU', 'DP').contract

    Tc = np.zeros((no, no, no, nv, nv, nv))
    Td = np.zeros_like(Tc)
    for i, j, k in itertools.product(range(no), repeat=3):
        Tc[i, j, k] = t3c_ijk(i, j, k, t2, W["vvvo"], W["ovoo"])
        Td[i, j, k] = t3d_ijk(i, j, k, t1, t2, W["oovv"], F[o, v])
        D = denominator(Fo, Fv, i=i, j=j, k=k)
        c0 = cctriples.t3c_ijk(o, v, i, j, k, t2d, Wvvvo, Wovoo, Fd, contract, WithDenom=False)
        c1 = cctriples.t3c_ijk(o, v, i, j, k, t2d, Wvvvo, Wovoo, Fd, contract)
        d0 = cctriples.t3d_ijk(o, v, i, j, k, t1d, t2d, Woovv, Fd, contract, WithDenom=False)
        d1 = cctriples.t3d_ijk(o, v, i, j, k, t1d, t2d, Woovv, Fd, contract, WithDenom=True)
        assert _close(c0, Tc[i, j, k]) and _close(d0, Td[i, j, k]), (i, j, k)
        assert np.array_equal(c1.cpu().numpy(), c0.cpu().numpy() / D), (i, j, k)
        assert np.array_equal(d1.cpu().numpy(), d0.cpu().numpy() / D), (i, j, k)
    assert np.abs(Td).max() > 1e-3 and np.abs(Tc).max() > 1e-3          # not vacuous zeros

    for a, b, c in itertools.product(range(nv), repeat=3):
        D = denominator(Fo, Fv, a=a, b=b, c=c)
        c0 = cctriples.t3c_abc(o, v, a, b, c, t2d, Wvvvo, Wovoo, Fd, contract, WithDenom=False)
        c1 = cctriples.t3c_abc(o, v, a, b, c, t2d, Wvvvo, Wovoo, Fd, contract)
        d0 = cctriples.t3d_abc(o, v, a, b, c, t1d, t2d, Woovv, Fd, contract, WithDenom=False)
        d1 = cctriples.t3d_abc(o, v, a, b, c, t1d, t2d, Woovv, Fd, contract)
        assert _close(c0, Tc[:, :, :, a, b, c]) and _close(d0, Td[:, :, :, a, b, c]), (a, b, c)
        assert np.array_equal(c1.cpu().numpy(), c0.cpu().numpy() / D), (a, b, c)
        assert np.array_equal(d1.cpu().numpy(), d0.cpu().numpy() / D), (a, b, c)


def test_t3_builders_reject_bad_arguments(pycc_amd):
    from pycc_amd import cctriples
    no, nv = 2, 3
    sd = synthetic_df(no, nv, seed=33)
    F, E = _dev(sd.fock), _dev(dirac_eri(sd.B))
    o, v = slice(0, no), slice(no, no + nv)
    t2 = torch.zeros((no, no, nv, nv), dtype=torch.float64, device="cuda")
    contract = pycc_amd.DeviceManager('GPU', 'DP').contract
    with pytest.raises(pycc_amd.PyCCError):                              # i outside the occupied range
        cctriples.t3c_ijk(o, v, 0, 0, no, t2, E[v, v, v, o], E[o, v, o, o], F, contract)
    with pytest.raises(pycc_amd.PyCCError):                              # not the device backend
        cctriples.t3c_ijk(o, v, 0, 0, 0, t2, E[v, v, v, o], E[o, v, o, o], F, np.einsum)
    with pytest.raises(pycc_amd.PyCCError):                              # float32 amplitudes
        cctriples.t3c_ijk(o, v, 0, 0, 0, t2.float(), E[v, v, v, o], E[o, v, o, o], F, contract)
