"""The exact mode's division sequences and the fast HCR-2's square root, run
on the GPU, against IEEE division and sqrt.

The exact-mode kernels replace the reference's divisions
(Local_code/level-set_MPI_3D/include/weno5.h:39-81, rk3.h:31-81) by three
sequences (csrc/ls_crdiv.hpp), each claimed to return the IEEE quotient bit for
bit inside a stated operand range:
  0. rdiv_u(a, b): the compiler's division sequence without its range scaling
     and fix-up (quotients with a divisor of their own: the WENO weights'
     numerators over (eps + beta)^2, and 1/s, 1/|grad G|);
  1. qdiv_u(a, b, rdiv_u(1, b)): quotients sharing a divisor;
  2. div6_u(a): a / 6.
ls_div_check runs each on the device; the expected values are numpy's IEEE
double quotients on the host.  The operands cover each range log-uniformly
with random significands, plus significands next to powers of two and all-ones
significands, signed zeros, the range ends, and the operand sets the stencils
produce (numerators 0.1, 0.3, 0.6, 1 over (eps + beta)^2 in [2^-40, 2^214] and
weight sums in [2^-218, 2^42]).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1 << 21


def _mag(rng, lo, hi, n):
    """|x| with exponent uniform in [lo, hi) and a significand that is random,
    1, next above 1, or all ones."""
    e = rng.integers(lo, hi, n)
    m = 1.0 + rng.random(n)
    k = rng.integers(0, 8, n)
    m = np.where(k == 0, 1.0, m)
    m = np.where(k == 1, np.nextafter(1.0, 2.0), m)
    m = np.where(k == 2, np.nextafter(2.0, 0.0), m)
    return np.ldexp(m, e)


def _signed(rng, x):
    return np.where(rng.random(x.size) < 0.5, -x, x)


def _run(ls, cuda, a, b, kind):
    a = np.ascontiguousarray(a, dtype=np.float64)
    da = cuda.from_numpy(a).cuda()
    db = None if b is None else cuda.from_numpy(np.ascontiguousarray(b, dtype=np.float64)).cuda()
    q = ls.div_check(da, db, kind)
    cuda.cuda.synchronize()
    return q.cpu().numpy()


def _check(got, want, a, b):
    bad = got.view(np.int64) != want.view(np.int64)
    if bad.any():
        i = int(np.flatnonzero(bad)[0])
        bi = None if b is None else repr(b[i])
        pytest.fail("%d of %d quotients differ; first: a=%r b=%s got %r want %r"
                    % (int(bad.sum()), bad.size, a[i], bi, got[i], want[i]))


def _pairs(rng, n, b_lo, b_hi, q_lo, q_hi):
    """Operands with b in [2^b_lo, 2^b_hi), a = 0 or |a| in [2^-960, 2^960),
    |a|/b in [2^q_lo, 2^q_hi]."""
    b = _mag(rng, b_lo, b_hi, n)
    ea = np.clip(np.floor(np.log2(b)).astype(np.int64) + rng.integers(q_lo + 1, q_hi, n), -960, 959)
    a = _signed(rng, _mag(rng, 0, 1, n) * np.ldexp(1.0, ea))
    q = np.abs(a) / b
    keep = (q >= 2.0 ** q_lo) & (q <= 2.0 ** q_hi) & (np.abs(a) >= 2.0 ** -960) & (np.abs(a) < 2.0 ** 960)
    a, b = a[keep], b[keep]
    # signed zeros and the range ends
    ends_b = np.array([2.0 ** b_lo, np.nextafter(2.0 ** b_hi, 0.0), 1.0, 3.0, 0.75])
    a = np.concatenate([a, [0.0, -0.0, 0.0, -0.0, 1.0]])
    b = np.concatenate([b, ends_b])
    return a, b


def test_rdiv_u_is_ieee_division(ls, cuda):
    rng = np.random.default_rng(11)
    a, b = _pairs(rng, N, -1000, 1000, -1000, 700)
    # the operands of the exact-mode stencils
    t = _mag(rng, -40, 214, N // 4)
    s = _mag(rng, -218, 42, N // 4)
    g = _mag(rng, -100, 117, N // 8)
    num = rng.choice(np.array([0.1, 0.3, 0.6]), t.size)
    a = np.concatenate([a, num, np.ones(s.size), np.ones(g.size)])
    b = np.concatenate([b, t, s, g])
    got = _run(ls, cuda, a, b, 0)
    _check(got, a / b, a, b)


def test_qdiv_u_on_rdiv_reciprocal_is_ieee_division(ls, cuda):
    rng = np.random.default_rng(12)
    a, b = _pairs(rng, N, -300, 300, -1000, 1000)
    # the weights: a_k / s with a_k <= s, both in [2^-218, 2^42]
    s = _mag(rng, -218, 42, N // 4)
    ak = s * rng.random(s.size)
    ak = np.where(ak >= 2.0 ** -260, ak, s)
    a = np.concatenate([a, ak])
    b = np.concatenate([b, s])
    got = _run(ls, cuda, a, b, 1)
    _check(got, a / b, a, b)


def test_div6_u_is_ieee_division(ls, cuda):
    rng = np.random.default_rng(13)
    a = _signed(rng, _mag(rng, -960, 1000, N))
    # integer combinations of stencil values, as the candidate polynomials produce
    v = rng.standard_normal((3, N // 4)) * np.ldexp(1.0, rng.integers(-40, 40, N // 4))
    a = np.concatenate([a, (2.0 * v[0] - 7.0 * v[1]) + 11.0 * v[2], [0.0, -0.0, 6.0, -6.0, 2.0 ** -960]])
    got = _run(ls, cuda, a, None, 2)
    _check(got, a / 6.0, a, None)


def test_sqrt_u_is_ieee_sqrt(ls, cuda):
    """sqrt_u (the fast HCR-2's square root: the compiler's sequence without
    its range scaling and fix-up) against the host's IEEE sqrt on
    [2^-767, 2^1000], including both ends and the squared-gradient sums a
    Godunov |grad phi| produces (sums of three squares of differences / h)."""
    rng = np.random.default_rng(14)
    a = _mag(rng, -767, 1000, N)
    d = rng.standard_normal((3, N // 4)) * np.ldexp(1.0, rng.integers(-30, 30, (3, N // 4)))
    sums = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
    sums = sums[sums >= 2.0 ** -767]
    a = np.concatenate([a, sums, [2.0 ** -767, 2.0 ** 1000, 1.0, 2.0, 4.0, np.nextafter(1.0, 2.0)]])
    got = _run(ls, cuda, a, None, 3)
    _check(got, np.sqrt(a), a, None)


def test_sqrt_u_term_is_ieee_below_the_range(ls, cuda):
    """sqrt_u(max(s2, 2^-767)) - 1 equals sqrt(s2) - 1 for every s2 in
    [0, 2^1000]: below 2^-767 both square roots are under 2^-54, so the
    difference rounds to -1."""
    rng = np.random.default_rng(15)
    lo = _mag(rng, -1074, -767, N // 4)
    hi = _mag(rng, -767, 1000, N // 4)
    a = np.concatenate([lo, hi, [0.0, 5e-324, 2.0 ** -767, np.nextafter(2.0 ** -767, 0.0), 1.0, 2.0 ** -108]])
    got = _run(ls, cuda, a, None, 4)
    _check(got, np.sqrt(a) - 1.0, a, None)


def test_div_check_rejects_bad_arguments(ls, cuda):
    import ctypes
    lib = ls.lib()
    assert lib.ls_div_check(None, None, None, 4, 0, None) != 0
    x = cuda.zeros(4, dtype=cuda.float64, device="cuda")
    p = ctypes.c_void_p(x.data_ptr())
    assert lib.ls_div_check(p, p, p, -1, 0, None) != 0
    assert lib.ls_div_check(p, p, p, 4, 5, None) != 0
    assert lib.ls_div_check(p, None, p, 4, 1, None) != 0
    assert lib.ls_div_check(p, None, p, 4, 2, None) == 0
    assert lib.ls_div_check(p, None, p, 4, 3, None) == 0
    cuda.cuda.synchronize()
