"""``helper_diis`` (``pycc/utils.py:257-361``) pinned by what the mathematics fixes, on the oracle.

The converged CCSD energy does not depend on DIIS (it only changes the path to the fixed point),
so the reference's energy constants cannot catch a slip in it. Three properties do:

* **Coefficients.** Pulay's bordered system is the KKT system of: minimise ||sum_k c_k e_k|| subject
  to sum_k c_k = 1. Its solution is computed here independently, by an unconstrained least-squares
  solve on the raw (un-normalised) error vectors (c_last = 1 - sum of the others), from error
  vectors this test records itself (e_n = t_n - the previous output, ``utils.py:287-291``). A
  normalisation that is not a uniform rescaling of B changes the minimiser and fails this check.
* **Pairing.** On a linear fixed-point map t <- M t + b of dimension n, DIIS with a window of at
  least n + 1 vectors is Anderson acceleration, which is equivalent to GMRES on (I - M) t = b
  (Walker & Ni, SIAM J. Numer. Anal. 49, 1715 (2011)): the extrapolated t is the exact solution once
  n + 1 error vectors span the space. That holds only if each stored amplitude vector is combined
  with the coefficient of its own error vector (``vals[num + 1]`` with ``errors[num]``,
  ``utils.py:351-355``).
* **Drop rule.** The window keeps the newest vectors: extrapolate() drops the oldest one, one per
  call, only when more than ``max_diis`` are stored (``utils.py:317-320``). With ``start_diis > 1``
  vectors accumulate before the first extrapolation, so the window can hold more than ``max_diis``.

Each property is shown to catch a planted slip (a wrong ``vals`` offset, dropping the newest
vector, a B normalised only on its diagonal): the slips are planted in subclasses of the oracle.
"""
import numpy as np
import pytest

from oracle.ccsd_ref import OracleDIIS

N1, N2 = 2, 4          # t1 holds 2 numbers, t2 holds 4: the oracle concatenates them (utils.py:289-291)
WIDE = (6, 36)         # for random windows of up to 8 vectors, which must stay linearly independent


def _split(x, n1=N1):
    return x[:n1].reshape(1, n1).copy(), x[n1:].reshape(1, 1, 1, x.size - n1).copy()


def _join(t1, t2):
    return np.concatenate([np.ravel(t1), np.ravel(t2)])


def constrained_lsq(errors):
    """argmin_c ||sum_k c_k e_k|| subject to sum_k c_k = 1, by eliminating the last coefficient."""
    E = np.stack(errors, axis=1)
    if E.shape[1] == 1:
        return np.ones(1)
    A = E[:, :-1] - E[:, -1:]
    z, *_ = np.linalg.lstsq(A, -E[:, -1], rcond=None)
    return np.append(z, 1.0 - z.sum())


def run_window_check(diis_cls, max_diis, start_diis, nsteps=10, seed=0):
    """Drive a DIIS object like solve_cc does (add after every step, extrapolate from start_diis on)
    with random steps, keeping an independent record of every (t, e) pair. Returns the largest
    deviation of the object's coefficients and output from the constrained least-squares solution
    over the newest vectors the drop rule keeps."""
    rng = np.random.default_rng(seed)
    n1 = WIDE[0]
    x = rng.standard_normal(sum(WIDE))
    d = diis_cls(*_split(x, n1), max_diis)
    prev = x.copy()
    vals, errs = [], []
    window = 0
    worst = 0.0
    for niter in range(1, nsteps + 1):
        x = x + rng.standard_normal(x.size) * (0.5 ** niter + 0.05)
        vals.append(x.copy())
        errs.append(x - prev)
        d.add_error_vector(*_split(x, n1))
        prev = x.copy()
        window += 1
        if niter >= start_diis:
            if window > max_diis:
                window -= 1                      # one per call, the oldest
            c = constrained_lsq(errs[-window:])
            t_expected = sum(ck * vk for ck, vk in zip(c, vals[-window:]))
            try:
                t1, t2 = d.extrapolate(*_split(x, n1))
            except np.linalg.LinAlgError:
                return np.inf
            out = _join(t1, t2)
            if len(d.last_coefficients) != window:
                return np.inf
            worst = max(worst, float(np.abs(d.last_coefficients - c).max()),
                        float(np.abs(out - t_expected).max()))
            x = out
            prev = out.copy()
    return worst


def linear_fixed_point_error(diis_cls, n_extra=3, seed=1):
    """max |t - t*| over the iterations after the (n + 1)-th, for t <- M t + b with DIIS."""
    rng = np.random.default_rng(seed)
    n = N1 + N2
    M = rng.standard_normal((n, n))
    M *= 0.6 / np.abs(np.linalg.eigvals(M)).max()
    b = rng.standard_normal(n)
    exact = np.linalg.solve(np.eye(n) - M, b)
    x = np.zeros(n)
    d = diis_cls(*_split(x), n + 2)
    errors = []
    for it in range(1, n + 2 + n_extra):
        x = M @ x + b                                     # the Jacobi step of the CC loop
        d.add_error_vector(*_split(x))
        try:
            x = _join(*d.extrapolate(*_split(x)))
        except np.linalg.LinAlgError:                     # a singular Pulay system fails the check
            return np.inf
        if it >= n + 1:
            errors.append(float(np.abs(x - exact).max()))
    return max(errors)


# ---- slips planted in the oracle, to show that the checks catch them --------------------------
class _ValsOffsetSlip(OracleDIIS):
    """t = sum c_k vals[k] instead of vals[k + 1]: amplitudes paired with the wrong error vector."""

    def extrapolate(self, t1, t2):
        self.diis_vals_t1.insert(1, self.diis_vals_t1[0])
        self.diis_vals_t2.insert(1, self.diis_vals_t2[0])
        try:
            return super().extrapolate(t1, t2)
        finally:
            del self.diis_vals_t1[1]
            del self.diis_vals_t2[1]


class _DropNewestSlip(OracleDIIS):
    """Trims the window from the wrong end."""

    def extrapolate(self, t1, t2):
        if len(self.diis_errors) > self.max_diis:
            del self.diis_vals_t1[-1]
            del self.diis_vals_t2[-1]
            del self.diis_errors[-1]
        return super().extrapolate(t1, t2)


class _DiagonalNormSlip(OracleDIIS):
    """Normalises the overlap matrix by its diagonal (a symmetric scaling, not a uniform one)."""

    def extrapolate(self, t1, t2):
        if self.max_diis and len(self.diis_errors) > self.max_diis:
            del self.diis_vals_t1[0]
            del self.diis_vals_t2[0]
            del self.diis_errors[0]
        errs = self.diis_errors
        scale = [np.sqrt(np.dot(e, e)) for e in errs]
        self.diis_errors = [e / s for e, s in zip(errs, scale)]
        try:
            return super().extrapolate(t1, t2)
        finally:
            self.diis_errors = errs


@pytest.mark.parametrize("max_diis,start_diis", [(8, 1), (3, 1), (1, 1), (3, 2), (3, 6), (2, 8)])
def test_coefficients_are_the_constrained_least_squares_solution(max_diis, start_diis):
    assert run_window_check(OracleDIIS, max_diis, start_diis) < 1e-9


def test_linear_fixed_point_is_solved_exactly():
    assert linear_fixed_point_error(OracleDIIS) < 1e-10


def test_window_keeps_more_than_max_diis_after_a_late_start():
    """start_diis = 6, max_diis = 3: six vectors are stored at the first extrapolation, which drops
    only the oldest, so five take part; each later call adds one and drops one."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal(N1 + N2)
    d = OracleDIIS(*_split(x), 3)
    sizes = []
    for niter in range(1, 10):
        x = x + rng.standard_normal(x.size)
        d.add_error_vector(*_split(x))
        if niter >= 6:
            x = _join(*d.extrapolate(*_split(x)))
            sizes.append(len(d.last_coefficients))
    assert sizes == [5, 5, 5, 5]


@pytest.mark.parametrize("slip", [_ValsOffsetSlip, _DropNewestSlip, _DiagonalNormSlip])
def test_planted_slips_are_caught(slip):
    caught = []
    caught.append(linear_fixed_point_error(slip) > 1e-6)
    for max_diis, start_diis in [(3, 1), (3, 6), (8, 1)]:
        caught.append(run_window_check(slip, max_diis, start_diis) > 1e-6)
    assert any(caught), slip.__doc__
    if slip is _ValsOffsetSlip:
        assert linear_fixed_point_error(slip) > 1e-6
    if slip is _DropNewestSlip:
        assert run_window_check(slip, 3, 1) > 1e-6 and run_window_check(slip, 3, 6) > 1e-6
