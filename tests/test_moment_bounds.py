"""The identities the GPP moment form rests on, checked in numpy at sizes and ratios where
the truncation is far above FP64 rounding (the kernels run where it is far below):

* the Moebius expansion of one pole, 1/(x - w) = (c - e)/((x - c)(w - e)) sum_m q^m with
  q = phi(w)/phi(x), phi(z) = (z - c)/(z - e), and the exact tail q^M/(x - w);
* the truncation bound rho^M sum_g |I_g/(x - w_g)|, rho = max|phi(w)| h/D, of a chunk's sum;
* the second-order expansion 1/(x - w)^2 = K^2 sum_m (m + 1) q^m with K = (c - e)/((x - c)(w - e)),
  its exact tail q^M ((M + 1) - M q)/(x - w)^2 and the bound rho^M (M + 1 + M rho) that the
  adjoint checks per pole, and the adjoint's regrouping of sum_poles s H_iw into
  moments B_m = sum_poles s g^2 phi^m;
* the re-expansion of a chunk's moments about another centre (phi' = b + a phi,
  b = (c - c')/(c - e), a = 1 - b) by the Taylor shift, against the moments computed about
  the new centre directly;
* the Horner evaluation of the forward's far-field value and of the adjoint's band sums,
  against the direct sums.
No GPU: these pin the algebra, the GPU tests pin the kernels against the oracle."""
import numpy as np

M = 16


def _phi(z, c, e):
    return (z - c) / (z - e)


def _cloud(rng, n, lo, hi, im=0.1):
    return rng.uniform(lo, hi, n) + 1j * rng.uniform(-im, 0, n)


def test_single_pole_tail_is_exact():
    rng = np.random.default_rng(1)
    e, c = 10.0, 1.7 - 0.04j
    w = _cloud(rng, 200, 0.5, 3.0)
    x = rng.uniform(8, 12, 200)
    q = _phi(w, c, e) / _phi(x, c, e)
    K = (c - e) / ((x - c) * (w - e))
    trunc = K * sum(q**m for m in range(M))
    exact = 1 / (x - w)
    tail = q**M / (x - w)
    assert np.max(np.abs((exact - trunc) - tail) / np.abs(exact)) < 1e-12


def test_chunk_truncation_bound():
    # a wide cloud and window: rho near 0.5, truncation ~1e-6, far above rounding
    rng = np.random.default_rng(2)
    xmin, xmax = 6.0, 14.0
    e, h = 0.5 * (xmin + xmax), 0.5 * (xmax - xmin)
    w = _cloud(rng, 2048, -1.0, 4.5, 0.3)
    I = rng.uniform(0, 0.5, 2048) + 1j * rng.uniform(-0.05, 0.05, 2048)
    c = 0.5 * (w.real.min() + w.real.max()) + 0.5j * (w.imag.min() + w.imag.max())
    phi = _phi(w, c, e)
    P = np.abs(phi).max()
    D = xmin - c.real if c.real < xmin else 0.0
    D = np.hypot(D, c.imag)
    rho = P * h / D
    assert 0.3 < rho < 0.8
    nu = np.array([np.sum(I * phi**m / (w - e)) for m in range(M)])
    for x in np.linspace(xmin, xmax, 9):
        psi = (x - e) / (x - c)
        far = (c - e) / (x - c) * np.polyval(nu[::-1], psi)
        direct = np.sum(I / (x - w))
        mag = np.sum(np.abs(I / (x - w)))
        bound = rho**M * mag
        err = abs(far - direct)
        assert err <= bound * (1 + 1e-9) + 1e-13 * mag
        if x == xmin:
            assert err > 1e3 * 1e-13 * mag  # truncation, not rounding, is what is measured


def test_second_order_tail_is_exact():
    rng = np.random.default_rng(5)
    e, c = 10.0, 1.7 - 0.04j
    w = _cloud(rng, 200, 0.5, 3.0)
    x = rng.uniform(8, 12, 200)
    q = _phi(w, c, e) / _phi(x, c, e)
    K = (c - e) / ((x - c) * (w - e))
    trunc = K**2 * sum((m + 1) * q**m for m in range(M))
    exact = 1 / (x - w) ** 2
    tail = q**M * ((M + 1) - M * q) / (x - w) ** 2
    assert np.max(np.abs((exact - trunc) - tail) / np.abs(exact)) < 1e-12


def test_adjoint_band_sums_and_bound():
    # the adjoint's expansion about a row centre c, for poles w and band frequencies x:
    #   G = sum_iw F_iw = (c - e) g P_A(phi),  H_iw = (c - e)^2 g^2 P_Q_iw(phi),
    #   A_m = sum c_n psi^m/(x - c),  Q_iw,m = (m + 1) sum c_n psi^m/(x - c)^2,
    # with truncation at most rho^M and rho^M (M + 1 + M rho) of the sums of magnitudes;
    # a wide window and cloud put rho near 0.4, so truncation dominates rounding
    rng = np.random.default_rng(3)
    xmin, xmax = 6.0, 14.0
    e, h = 0.5 * (xmin + xmax), 0.5 * (xmax - xmin)
    x = rng.uniform(xmin, xmax, (3, 200))
    cn = rng.uniform(-1, 1, 200) + 1j * rng.uniform(-1, 1, 200)
    c = 1.5 - 0.05j
    D = np.hypot(xmin - c.real, c.imag)
    psi = (x - e) / (x - c)
    A = np.array([np.sum(cn * psi**m / (x - c)) for m in range(M)])
    Q = np.array([[(m + 1) * np.sum(cn * psi[iw] ** m / (x[iw] - c) ** 2) for m in range(M)] for iw in range(3)])
    seen = 0.0
    for w in _cloud(rng, 40, 0.0, 3.0, 0.3):
        g = 1 / (w - e)
        phi = (w - c) * g
        G = (c - e) * g * np.polyval(A[::-1], phi)
        rho = abs(phi) * h / D
        bF, bH = rho**M, rho**M * (M + 1 + M * rho)
        Gd = np.sum(cn / (x - w))
        sG = np.sum(np.abs(cn / (x - w)))
        assert abs(G - Gd) <= bF * sG * (1 + 1e-9) + 1e-13 * sG
        for iw in range(3):
            H = (c - e) ** 2 * g**2 * np.polyval(Q[iw][::-1], phi)
            Hd = np.sum(cn / (x[iw] - w) ** 2)
            sH = np.sum(np.abs(cn / (x[iw] - w) ** 2))
            assert abs(H - Hd) <= bH * sH * (1 + 1e-9) + 1e-13 * sH
            seen = max(seen, abs(H - Hd) / (bH * sH))
    assert seen > 1e-3  # the bound is not vacuous here


def test_adjoint_dsigma_regrouping():
    # sum_poles s_j H_iw(w_j) = (c - e)^2 sum_m Q_iw,m B_m,  B_m = sum_j s_j g_j^2 phi_j^m
    rng = np.random.default_rng(6)
    xmin, xmax = 8.0, 12.0
    e = 0.5 * (xmin + xmax)
    x = rng.uniform(xmin, xmax, (3, 64))
    cn = rng.uniform(-1, 1, 64) + 1j * rng.uniform(-1, 1, 64)
    c = 1.8 - 0.05j
    w = _cloud(rng, 512, 0.6, 2.9)
    s = rng.uniform(-1, 1, 512) + 1j * rng.uniform(-1, 1, 512)
    psi = (x - e) / (x - c)
    g = 1 / (w - e)
    phi = (w - c) * g
    B = np.array([np.sum(s * g**2 * phi**m) for m in range(M)])
    for iw in range(3):
        Q = np.array([(m + 1) * np.sum(cn * psi[iw] ** m / (x[iw] - c) ** 2) for m in range(M)])
        got = (c - e) ** 2 * np.sum(Q * B)
        ref = np.sum(s * np.array([np.sum(cn / (x[iw] - wj) ** 2) for wj in w]))
        # rho < 0.06 here: the truncation is far below rounding, so only rounding separates them
        mag = np.sum(np.abs(s)[None, :] * np.abs(cn[:, None] / (x[iw][:, None] - w[None, :]) ** 2))
        assert abs(got - ref) <= 1e-13 * mag


def test_taylor_shift_reexpansion():
    rng = np.random.default_rng(4)
    e = 10.0
    w = _cloud(rng, 2048, 0.6, 2.9)
    I = rng.uniform(0, 0.5, 2048) + 1j * rng.uniform(-0.05, 0.05, 2048)
    c, c2 = 1.74 - 0.049j, 1.81 - 0.052j
    nu = np.array([np.sum(I * _phi(w, c, e) ** m / (w - e)) for m in range(M)])
    b = (c - c2) / (c - e)
    a = 1 - b
    mu = np.array([a**m * nu[m] for m in range(M)])
    for k in range(1, M):
        mu[k:] = mu[k:] + b * mu[k - 1:-1].copy()  # old values, as the lanes' shuffles
    ref = np.array([np.sum(I * _phi(w, c2, e) ** m / (w - e)) for m in range(M)])
    assert np.max(np.abs(mu - ref) / np.abs(ref)) < 1e-12
    # and the bound of the re-expanded disc: |phi'| <= |b| + |a| max|phi|
    assert np.abs(_phi(w, c2, e)).max() <= abs(b) + abs(a) * np.abs(_phi(w, c, e)).max() + 1e-15
