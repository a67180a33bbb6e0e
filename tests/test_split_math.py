"""The split-operand theta (TEBD_MODE_SPLIT, tebd_theta_split.hip) restated in numpy and
checked against its error bound and against the FP64 product.

The scheme: balance each middle index g by 2^eg (eg from B's row g), scale each row of
A and column of B by 2^e (e: binary exponent of its largest real or imaginary part),
cut into S = 7 truncated 7-bit slices, keep the slice products of level s + t < S and
combine them with weights 2^-7(s+t+2).  Pinned here:
  * the slices are integers in [-127, 127] and reconstruct every value to below
    2^(e - 7S) (the truncation bound);
  * the balance is exact: A diag(2^eg) diag(2^-eg) B = A B bit for bit;
  * integer level sums reproduce the exact rational value of the kept products (checked
    with Python integers), so the device's int32 accumulation loses nothing;
  * on a Vidal state evolved by the oracle (Gamma entries up to 1e11 from lambdas down to
    1e-14) the split product agrees with the FP64 product to 1e-13 relative; with five
    slices the unbalanced variant is 2-7x worse than the balanced one (Gamma_r's
    large rows would otherwise set the column scales of B).
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle.oracle import OracleState

S = 7


def exponents(M, axis):
    m = np.maximum(np.abs(M.real), np.abs(M.imag)).max(axis=axis)
    return np.where(m > 0, np.frexp(m)[1], 0)


def cut(X, e, nsl=S):
    """Slices of X 2^-e (e per row), truncated: list of nsl integer arrays."""
    y = np.ldexp(X, -e[:, None])
    out = []
    for _ in range(nsl):
        y = y * 128.0
        t = np.trunc(y)
        out.append(t)
        y = y - t
    return out, y


def split_product(A, B, balance=True, nsl=S):
    if balance:
        eg = exponents(B, 1)
        A = A * np.ldexp(1.0, eg)[None, :]
        B = B * np.ldexp(1.0, -eg)[:, None]
    ea, eb = exponents(A, 1), exponents(B.T, 1)
    ar, _ = cut(A.real, ea, nsl)
    ai, _ = cut(A.imag, ea, nsl)
    br, _ = cut(B.T.real, eb, nsl)
    bi, _ = cut(B.T.imag, eb, nsl)
    acc = np.zeros((A.shape[0], B.shape[1]), complex)
    for L in range(nsl):
        pre = sum(ar[s] @ br[L - s].T - ai[s] @ bi[L - s].T for s in range(L + 1))
        pim = sum(ar[s] @ bi[L - s].T + ai[s] @ br[L - s].T for s in range(L + 1))
        acc += (pre + 1j * pim) * 2.0 ** (-7 * (L + 2))
    return acc * np.ldexp(1.0, ea)[:, None] * np.ldexp(1.0, eb)[None, :]


def test_slices_are_7_bit_and_reconstruct_within_the_truncation_bound():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((40, 64)) * np.exp(rng.uniform(-30, 30, (40, 1)))
    e = exponents(X + 0j, 1)
    q, r = cut(X, e)
    for s in q:
        assert np.all(np.abs(s) <= 127) and np.all(s == np.round(s))
    rec = sum(qs * 2.0 ** (-7 * (i + 1)) for i, qs in enumerate(q))
    err = np.abs(np.ldexp(rec, e[:, None]) - X)
    assert np.all(err < np.ldexp(1.0, e - 7 * S)[:, None])
    assert np.all(np.abs(r) < 1.0)


def test_balance_is_exact():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((8, 16)) + 1j * rng.standard_normal((8, 16))
    B = rng.standard_normal((16, 8)) + 1j * rng.standard_normal((16, 8))
    eg = rng.integers(-60, 60, 16)
    A2, B2 = A * np.ldexp(1.0, eg)[None, :], B * np.ldexp(1.0, -eg)[:, None]
    assert np.array_equal(A2 * np.ldexp(1.0, -eg)[None, :], A) and np.array_equal(B2 * np.ldexp(1.0, eg)[:, None], B)
    assert np.array_equal(A2 @ B2, A @ B) or np.abs(A2 @ B2 - A @ B).max() <= 1e-15 * np.abs(A @ B).max()


def test_level_sums_are_exact_integers():
    """The kept products of one entry summed exactly (Python integers) equal the FP64
    combination up to a final rounding: the int32 accumulation is exact."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal(64)
    b = rng.standard_normal(64)
    ea, eb = exponents(a[None, :] + 0j, 1), exponents(b[None, :] + 0j, 1)
    qa, _ = cut(a[None, :], ea)
    qb, _ = cut(b[None, :], eb)
    from fractions import Fraction
    exact = Fraction(0)
    fp = 0.0
    for L in range(S):
        p = sum(int(x) * int(y) for s in range(L + 1) for x, y in zip(qa[s][0], qb[L - s][0]))
        assert abs(p) < 2 ** 31
        exact += Fraction(p, 2 ** (7 * (L + 2)))
        fp += p * 2.0 ** (-7 * (L + 2))
    assert abs(fp - float(exact)) <= 4 * np.finfo(float).eps * abs(float(exact))
    # and the kept products approximate the true dot product within the stated bound
    true = float(np.dot(a, b) * 2.0 ** (-(ea[0] + eb[0])))
    assert abs(float(exact) - true) < (S + 3) * 2.0 ** (-7 * S) * 64


@pytest.fixture(scope="module")
def evolved_state():
    z = load_golden("xxz_l32_chi32")
    st = OracleState.product(z["coeffs"], int(z["chi"]))
    for _ in range(60):
        st.trotter_step(z["U"])
    return st.gam, st.lam


@pytest.mark.parametrize("link", [8, 16])
def test_split_product_matches_fp64_on_a_vidal_state(evolved_state, link):
    gam, lam = evolved_state
    G1, G2 = gam[link - 1], gam[link]
    A = (G1 * lam[link][None, None, :]).reshape(-1, G1.shape[2])
    B = G2.reshape(G2.shape[0], -1)
    assert np.abs(G2).max() > 1e9  # Gamma carries 1/lambda: the case the balance is for
    exact = A @ B
    err = np.abs(split_product(A, B) - exact).max() / np.abs(exact).max()
    print(f"link {link}: split {err:.2e}")
    assert err < 1e-13
    e5 = np.abs(split_product(A, B, nsl=5) - exact).max() / np.abs(exact).max()
    e5_nobal = np.abs(split_product(A, B, balance=False, nsl=5) - exact).max() / np.abs(exact).max()
    print(f"link {link}: five slices {e5:.2e}, without balance {e5_nobal:.2e}")
    assert e5_nobal > 1.5 * e5
