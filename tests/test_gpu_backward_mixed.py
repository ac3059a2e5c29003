"""HBAR, the Lambda residual, the vJp and the Lambda solve in a MIXED_F16X3 context (the FP64
contractions of at least 1e11 flops on a full tile grid as FP16x3 split products, the Lambda ladder
against the context's split ladder rows), against the same quantities of a DP context on the same
Hamiltonian and amplitudes (``pycc/cchbar.py:128-823``, ``pycc/cclambda.py:345-369, :468-496``).

o=16, v=256 is the smallest size at which the ring-shaped contractions (M = N = o v = 4096) reach
the split route, with both of its output forms (C written as a matrix, and C scattered through its
index map); at o=4, v=24 only the ladder is split. Tolerances: each split product carries about
2^-22 relative error, so the residuals agree to 1e-6 relative (measured 1e-9 at o=4, v=24 and
2e-7 to 5e-7 at o=16, v=256) and the Lambda pseudoenergy to 1e-7 relative.
"""
import numpy as np
import pytest

from conftest import random_amplitudes
from pycc_amd.synth import synthetic_df

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


def _relnorm(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def _pair(pycc_amd, no, nv, seed):
    sd = synthetic_df(no, nv, seed=seed)
    dp = pycc_amd.CCwfn(sd, no=no, precision='DP', verbose=False)
    mx = pycc_amd.CCwfn(sd, no=no, precision='MIXED_F16X3', verbose=False)
    t1, t2 = random_amplitudes(no, nv, seed + 1, scale=0.02)
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    return dp, mx, _dev(t1), _dev(t2)


@pytest.mark.parametrize("no,nv,seed", [(4, 24, 41), (16, 256, 42)])
def test_lambda_residual_and_vjp_mixed_vs_dp(pycc_amd, no, nv, seed):
    print("mixed Lambda residual / vJp vs DP no=%d nv=%d" % (no, nv), flush=True)
    dp, mx, t1, t2 = _pair(pycc_amd, no, nv, seed)
    rng = np.random.default_rng(seed)
    l1 = _dev(rng.standard_normal((no, nv)) * 0.05)
    l2 = rng.standard_normal((no, no, nv, nv)) * 0.02
    l2 = _dev(l2 + l2.transpose(1, 0, 3, 2))
    r = {}
    for name, cc in (("dp", dp), ("mx", mx)):
        cc._hbar_for(cc.H.F, t1, t2)
        r[name] = cc._lambda_residual(l1, l2, with_source=True)
        torch.cuda.synchronize()
    e1, e2 = _relnorm(r["mx"][0], r["dp"][0]), _relnorm(r["mx"][1], r["dp"][1])
    print("Lambda residual relative difference r1 %.2e r2 %.2e" % (e1, e2), flush=True)
    assert e1 < 1e-6 and e2 < 1e-6, (e1, e2)
    c1, c2 = l1.clone(), l2.clone()
    g = {name: cc.vjp(t1, t2, c1, c2) for name, cc in (("dp", dp), ("mx", mx))}
    e1, e2 = _relnorm(g["mx"][0], g["dp"][0]), _relnorm(g["mx"][1], g["dp"][1])
    print("vJp relative difference g1 %.2e g2 %.2e" % (e1, e2), flush=True)
    assert e1 < 1e-6 and e2 < 1e-6, (e1, e2)


def test_lambda_solve_mixed_vs_dp(pycc_amd):
    no, nv, seed = 16, 256, 43
    print("mixed Lambda solve vs DP no=%d nv=%d" % (no, nv), flush=True)
    sd = synthetic_df(no, nv, seed=seed)
    dp = pycc_amd.CCwfn(sd, no=no, precision='DP', verbose=False)
    assert dp.solve_cc(1e-10, 1e-10, 100) is not None
    mx = pycc_amd.CCwfn(sd, no=no, precision='MIXED_F16X3', verbose=False)
    mx.t1.copy_(dp.t1)
    mx.t2.copy_(dp.t2)
    # the mixed Lambda residual carries ~1e-7 relative rounding, so its solve stops at r_conv 1e-7
    lam = {}
    for name, cc, rc in (("dp", dp, 1e-8), ("mx", mx, 1e-7)):
        solver = pycc_amd.cclambda(cc, pycc_amd.cchbar(cc))
        lam[name] = (solver.solve_lambda(1e-10, rc, 200), solver)
        print("%s: pseudoenergy %s after %d iterations" % (name, lam[name][0], len(solver.history)), flush=True)
    (ed, sd_), (em, sm) = lam["dp"], lam["mx"]
    assert ed is not None and em is not None
    dl2 = _relnorm(sm.l2, sd_.l2)
    print("pseudoenergy difference %.2e relative, l2 %.2e" % ((em - ed) / abs(ed), dl2), flush=True)
    assert abs(em - ed) < 1e-7 * abs(ed), (em, ed)
    assert dl2 < 1e-5


def test_mixed_context_refuses_the_stored_vvvv_block_and_the_jacobian(pycc_amd):
    sd = synthetic_df(2, 6, seed=44)
    mx = pycc_amd.CCwfn(sd, no=2, precision='MIXED_F16X3', verbose=False)
    with pytest.raises(pycc_amd.PyCCError):
        pycc_amd.cchbar(mx, store_vvvv=True)
    with pytest.raises(pycc_amd.PyCCError):
        mx.jvp(mx.t1, mx.t2, mx.t1, mx.t2)
