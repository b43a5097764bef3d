"""Host-side logic of the device path against the oracle (no GPU): the O(N)
and O(P) tables the kernels consume."""
import numpy as np
import pytest

from oracle import fp64 as o
from pyatmosphere_amd import engine


def _spec(n=64, delta=0.01, wvl=808e-9, L=10e3, S=5, P=32, **kw):
    base = engine.log_polar_base(P, 1 / 15000, 2 / 6e-3)
    psd = engine.annulus_psd(base, 2 * np.pi / wvl, L / S, lambda k: engine.mvk_psd_n(k, 1e-15, 6e-3, 1e3))
    return engine.ChannelSpec(n=n, delta=delta, wvl=wvl, w0=kw.get("w0", 0.05), F0=kw.get("F0", np.inf),
                              length=L, positions=o.path_positions(L, S), points=P, base=base,
                              psd=psd.astype(np.float32))


def test_axis_and_base_match_oracle():
    assert np.array_equal(engine.axis_coords(64, 0.01), o.axis_coords(64, 0.01))
    assert np.array_equal(engine.log_polar_base(128, 1e-3, 300.0), o.log_polar_base(128, 1e-3, 300.0))


def test_psd_matches_oracle():
    base = o.log_polar_base(256, 1 / 15000, 2 / 6e-3)
    k = 2 * np.pi / 808e-9
    a = engine.annulus_psd(base, k, 1e4, lambda kap: engine.mvk_psd_n(kap, 5e-16, 6e-3, 1e3))
    b = o.annulus_psd(base, k, 1e4, 5e-16, 6e-3, 1e3)
    assert np.allclose(a, b, rtol=1e-14, atol=0)


def test_step_lengths():
    s = _spec(L=50e3, S=5)
    assert s.step_lengths() == o.step_lengths(50e3, o.path_positions(50e3, 5))


@pytest.mark.parametrize("F0", [np.inf, 2e4, -6e3])
def test_source_factors_outer_product(F0):
    s = _spec(n=128, delta=0.004, w0=0.07, F0=F0)
    gy, gx = s.source_factors()
    ref = o.gaussian_source(128, 0.004, s.wvl, 0.07, F0)
    got = np.outer(gy.astype(np.complex128), gx.astype(np.complex128))
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 3e-7


def test_transfer_factors_reproduce_oracle_step():
    """ifft2(H fft2 u) with the separable complex64 factors equals the oracle step."""
    s = _spec(n=64, delta=0.02, L=10e3, S=2)
    H = s.transfer_factors(3)
    rng = np.random.default_rng(0)
    u = rng.normal(size=(64, 64)) + 1j * rng.normal(size=(64, 64))
    for i, d in enumerate(s.step_lengths()):
        Hf = np.outer(H[i, 0], H[i, 1]).astype(np.complex128) * 64 ** 2
        got = np.fft.ifft2(Hf * np.fft.fft2(u))
        ref = o.vacuum_step(u, 0.02, s.wvl, d)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-6


@pytest.mark.parametrize("F0", [np.inf, 3e3])
def test_separable_first_step_identity(F0):
    """The separable first step as the kernels take it (pa_source_fresnel):
    uy = IFFT(h_0[0] FFT(gy)), ux = IFFT(h_0[1] FFT(gx)) with the unnormalised
    transforms and the package's complex64 factors; uy ux^T is the oracle's
    vacuum step of the Gaussian source over the first step length (the field at
    the first screen plane)."""
    s = _spec(n=128, delta=0.004, w0=0.07, F0=F0, L=8e3, S=3)
    gy, gx = (g.astype(np.complex128) for g in s.source_factors())
    H = s.transfer_factors(s.nscreens + 1).astype(np.complex128)
    n = s.n
    uy = n * np.fft.ifft(H[0, 0] * np.fft.fft(gy))
    ux = n * np.fft.ifft(H[0, 1] * np.fft.fft(gx))
    d0 = s.step_lengths()[0]
    ref = o.vacuum_step(o.gaussian_source(n, s.delta, s.wvl, s.w0, F0), s.delta, s.wvl, d0)
    assert d0 > 0
    assert np.linalg.norm(np.outer(uy, ux) - ref) / np.linalg.norm(ref) < 1e-6


def test_screen_decomposition_identity():
    """phi = phi0 + g(c) + h(r) + Re[(a q) p^T] exactly (the split the device uses)."""
    n, d = 32, 0.01
    base = o.log_polar_base(24, 0.05, 40.0)
    sp = o.draw_spectrum_numpy(np.random.RandomState(2), base, np.logspace(2, -4, 24).astype(np.float32))
    fx = sp.rho.astype(float) * np.cos(sp.theta.astype(float))
    fy = sp.rho.astype(float) * np.sin(sp.theta.astype(float))
    x = o.axis_coords(n, d)
    p = np.expm1(1j * 2 * np.pi * np.outer(fx, x))      # (P, N) columns
    q = np.expm1(1j * 2 * np.pi * np.outer(x, fy))      # (N, P) rows
    a = sp.value.astype(np.complex128)
    phi = a.real.sum() + (a[:, None] * p).sum(0).real[None, :] + (q * a[None, :]).sum(1).real[:, None] \
        + ((q * a[None, :]) @ p).real
    assert np.abs(phi - o.screen_phase(n, d, sp)).max() < 1e-10


@pytest.mark.parametrize("Z", [10.0, 14.0])
def test_polynomial_block_identity(Z):
    """The low-frequency points' share of phi2 equals
    sum_{n,m<32} T_n(s) T_m(t) Re G[n][m] (G[n][m] = sum_k a_k Q_kn P_km with the
    Jacobi-Anger coefficients e^{i w shift} 2 i^m J_m(w X) of e^{i w (shift + X u)} - 1
    in Chebyshev polynomials of u = offset / X) to float64 rounding, against the
    direct plane-wave sum, on the 2048^2 / delta 1.5 mm grid with screen shifts and
    the points whose phase turns by at most Z rad between the grid centre and edge
    (Z = 14, the device's threshold, and 10).  The bound is the 32-term
    truncation of each point, sum_k |a_k| (2 T(z_x) + 2 T(z_y) + T(z_x) T(z_y))
    with T(z) = 2 sum_{m>=32} |J_m(z)| (|P|, |Q| <= 2 on [-1, 1]; T(14) = 2.4e-9),
    or 1e-12 of the sum, whichever is larger."""
    from scipy.special import jv
    n, d, P, TL = 2048, 1.5e-3, 1024, 32
    tail = lambda z: 2 * np.abs(jv(np.arange(TL, 120)[None, :], np.abs(z)[:, None])).sum(axis=1)
    base = o.log_polar_base(P, 1 / 15000, 2 / 6e-3).astype(np.float64)
    sp = o.draw_spectrum_numpy(np.random.RandomState(5), base.astype(np.float32),
                               (base ** (-11 / 6) * 1e-3).astype(np.float32))
    fx = sp.rho.astype(float) * np.cos(sp.theta.astype(float))
    fy = sp.rho.astype(float) * np.sin(sp.theta.astype(float))
    a = sp.value.astype(np.complex128)
    sx, sy, X = 0.37, -1.9, n / 2 * d
    al, be = 2 * np.pi * fx, 2 * np.pi * fy
    ok = (np.abs(al) * X <= Z) & (np.abs(be) * X <= Z)
    klow = (int(np.argmin(ok)) // 16) * 16
    assert klow >= 600
    k = np.arange(klow)
    rows, cols = np.arange(0, n, 37), np.arange(0, n, 41)
    x, y = (cols - n / 2) * d + sx, (rows - n / 2) * d + sy
    direct = ((np.expm1(1j * np.outer(y, be[k])) * a[k]) @ np.expm1(1j * np.outer(x, al[k])).T).real

    def coef(w, shift):
        m = np.arange(TL)
        c = np.exp(1j * w * shift)[:, None] * 2 * (1j) ** m[None, :] * jv(m[None, :], (w * X)[:, None])
        c[:, 0] = c[:, 0] / 2 - 1
        return c
    G = np.einsum("k,kn,km->nm", a[k], coef(be[k], sy), coef(al[k], sx)).real
    s, t = (rows - n / 2) / (n / 2), (cols - n / 2) / (n / 2)
    Ts, Tt = (np.polynomial.chebyshev.chebvander(v, TL - 1) for v in (s, t))
    block = (Ts @ G) @ Tt.T
    tx, ty = tail(al[k] * X), tail(be[k] * X)
    bound = (np.abs(a[k]) * (2 * tx + 2 * ty + tx * ty)).sum()
    assert np.abs(block - direct).max() < max(1e-12 * np.abs(direct).max(), bound)
    # their shares of the float64 1-D parts h(r) = Re sum a q, g(c) = Re sum a p
    h_direct = (np.expm1(1j * np.outer(y, be[k])) @ a[k]).real
    g_direct = (np.expm1(1j * np.outer(x, al[k])) @ a[k]).real
    h_cheb = Ts @ (a[k] @ coef(be[k], sy)).real
    g_cheb = Tt @ (a[k] @ coef(al[k], sx)).real
    assert np.abs(h_cheb - h_direct).max() < 1e-12 * np.abs(h_direct).max()
    assert np.abs(g_cheb - g_direct).max() < 1e-12 * np.abs(g_direct).max()


def test_expmi_polynomial_accuracy():
    """The device's e^{-ix} (float32 Cody-Waite reduction with a three-part pi/2
    applied by FMA, Taylor polynomials to r^9/r^10), emulated with float32
    rounding, stays within 1.2e-7 of the float64 value for |x| <= 4000."""
    f = np.float32

    def fma(a, b, c):
        return (np.float64(a) * np.float64(b) + np.float64(c)).astype(np.float32)

    x = np.linspace(-4000, 4000, 2000001).astype(np.float32)
    xd = x.astype(np.float64)
    q = np.rint(x * f(0.63661977236758134308)).astype(np.float32)
    r = fma(-q, f(1.57079637050628662109375), x)
    r = fma(-q, f(-4.37113900018624283e-8), r)
    r = fma(-q, f(-1.7763568394002504647e-15), r)
    r2 = r * r
    s = fma(r2, f(2.7557319e-6), f(-1.9841270e-4)); s = fma(r2, s, f(8.3333333e-3))
    s = fma(r2, s, f(-1.6666667e-1)); s = fma(r2 * r, s, r)
    c = fma(r2, f(-2.7557319e-7), f(2.4801587e-5)); c = fma(r2, c, f(-1.3888889e-3))
    c = fma(r2, c, f(4.1666668e-2)); c = fma(r2, c, f(-0.5)); c = fma(r2, c, f(1.0))
    qi = q.astype(np.int64) & 3
    sn = np.select([qi == 0, qi == 1, qi == 2], [s, c, -s], -c)
    cs = np.select([qi == 0, qi == 1, qi == 2], [c, -s, -c], s)
    assert np.abs(sn - np.sin(xd)).max() < 1.2e-7
    assert np.abs(cs - np.cos(xd)).max() < 1.2e-7


def test_table_recurrence_over_32_rows():
    """The plane-wave tables advance w = a e^{i t} down a block's rows by the
    float64 product w <- w e^{i step} from one sincos of the first angle and one
    of the step.  Over 32 rows (the largest block), on the 2048^2 / 1.5 mm grid's
    extreme frequencies and offsets: the 31 products drift from the exact powers
    of the rounded start and step by < 1e-14 |a|, and the entries w - a differ
    from a (e^{i t_j} - 1) in extended precision by < 2e-12 |a|, the float64
    rounding of arguments up to ~3000 rad (far below the split operands'
    2^-22)."""
    rng = np.random.default_rng(3)
    n, d = 2048, 1.5e-3
    f = rng.uniform(-2 / 6e-3, 2 / 6e-3, 4000)
    i0 = rng.integers(0, n // 32, 4000) * 32
    a = rng.normal(size=4000) + 1j * rng.normal(size=4000)
    ld = np.longdouble
    t0, ts = 2 * np.pi * f * ((i0 - n / 2) * d), 2 * np.pi * f * d     # float64 arguments, as in the kernel
    w0, step = a * np.exp(1j * t0), np.exp(1j * ts)
    w = w0.copy()
    drift = exact = 0.0
    fl, il = f.astype(ld), i0.astype(ld)
    for j in range(32):
        # the exact j-th power of the rounded start and step, in extended precision
        ang = np.angle(w0).astype(ld) + ld(j) * np.angle(step).astype(ld)
        pw = np.abs(w0).astype(ld) * (np.cos(ang) + 1j * np.sin(ang))
        drift = max(drift, float(np.max(np.abs(w - pw.astype(np.complex128)) / np.abs(a))))
        tj = 2 * ld(np.pi) * fl * ((il + j - n // 2) * ld(d))
        ref = a.astype(np.clongdouble) * (np.cos(tj) - 1 + 1j * np.sin(tj))
        exact = max(exact, float(np.max(np.abs((w - a) - ref.astype(np.complex128)) / np.abs(a))))
        w = w * step
    assert drift < 1e-14 and exact < 2e-12

@pytest.mark.parametrize("rt", [4, 8, 16, 32])
def test_halving_exchange_reduction(rt):
    """Host emulation of the tables kernel's block reduction of RT row sums per
    lane over a 64-lane wave: at offset 32 >> s a lane keeps the half of its
    sums selected by that bit and adds its partner's copy, down to one sum per
    lane; xor shuffles over the remaining offsets finish it, and lane l holds
    the total of row jr(l) (bits 5, 4, ... of l, most significant first)."""
    rng = np.random.default_rng(rt)
    acc = rng.normal(size=(64, rt))
    lg = {4: 2, 8: 3, 16: 4, 32: 5}[rt]
    cur = acc.copy()
    for s in range(lg):
        half, off = rt >> (s + 1), 32 >> s
        nxt = cur.copy()
        for lane in range(64):
            up = bool(lane & off)
            for t in range(half):
                keep = cur[lane, t + half] if up else cur[lane, t]
                partner = lane ^ off
                send = cur[partner, t] if not up else cur[partner, t + half]   # partner's copy of our half
                nxt[lane, t] = keep + send
        cur = nxt
    tot = cur[:, 0].copy()
    off = 32 >> lg
    while off > 0:
        tot = tot + tot[np.arange(64) ^ off]
        off >>= 1
    for lane in range(64):
        jr = 0
        for s in range(lg):
            jr |= ((lane >> (5 - s)) & 1) << (lg - 1 - s)
        assert np.isclose(tot[lane], acc[:, jr].sum(), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n,delta,w,x0", [(512, 0.002, 0.004, 0.0), (512, 0.002, 0.04, 0.05),
                                          (2048, 0.0015, 0.003, 0.0), (2048, 0.0015, 0.2, 0.01),
                                          (4096, 0.001, 0.002, 0.0)])
def test_row_pass_moment_weights(n, delta, w, x0):
    """Host emulation of the last row pass's per-thread x-moments: a thread
    holds the columns c = l + L e (L = n / E), so with xl = l delta and
    Ld = L delta, sum I x = xl S0 + Ld S1 and sum I x^2 = xl^2 S0 + 2 xl Ld S1 +
    Ld^2 S2 from S_k = sum (e - E/2)^k I.  Against the direct float64 sums over
    x = (c - n/2) delta (measures.py:7-38) on narrow and wide noisy beams: the
    cancellation of the three terms stays below 1e-11 relative (a wrong centre
    E/2 or weight would be off by O(1))."""
    E = {512: 8, 2048: 16, 4096: 16}[n]          # points per thread of pa_fft.h's plans
    L, H = n // E, E // 2
    x = (np.arange(n) - n // 2) * delta
    rng = np.random.default_rng(n)
    u = (np.exp(-((x[None, :] - x0) ** 2 + x[:, None] ** 2) / w ** 2) * (1 + 0.1 * rng.normal(size=(n, n)))
         ).astype(np.complex64)
    I = u.real.astype(np.float64) ** 2 + u.imag.astype(np.float64) ** 2
    Ie = I.reshape(n, E, L)                      # [row][e][l]: column e L + l
    e = np.arange(E)[None, :, None] - H
    s0, s1, s2 = Ie.sum(1), (e * Ie).sum(1), (e * e * Ie).sum(1)
    xl, Ld = np.arange(L)[None, :] * delta, L * delta
    sx = (xl * s0 + Ld * s1).sum()
    sxx = (xl * xl * s0 + 2.0 * xl * Ld * s1 + Ld * Ld * s2).sum()
    tot = I.sum()
    assert abs(sx - (I * x).sum()) <= 1e-11 * (I * np.abs(x)).sum()
    assert abs(sxx - (I * x * x).sum()) <= 1e-11 * (I * x * x).sum()
    assert np.isclose(s0.sum(), tot, rtol=1e-14)


@pytest.mark.parametrize("Z", [10.0, 14.0])
def test_column_smooth_points_identity(Z):
    """Points whose plane wave turns by at most Z rad along the columns only
    (|alpha| X <= Z < |beta| X; Z = 14 on the device) share the polynomial
    block's B columns T_m(t_c):
    Re sum_k (a q)_k(r) p_k(c) = sum_m T_m(t_c) Re V_m(r) with V_m(r) = sum_k
    (a q)_k(r) P_km (P_km the Jacobi-Anger coefficients of p_k), against the
    direct plane-wave sum on the 2048^2 / 1.5 mm grid, and their share of g(c)
    as sum_m T_m(t_c) Re sum_k a_k P_km (bound: each point's 32-term truncation,
    as in test_polynomial_block_identity)."""
    from scipy.special import jv
    n, d, P, TL = 2048, 1.5e-3, 1024, 32
    tail = lambda z: 2 * np.abs(jv(np.arange(TL, 120)[None, :], np.abs(z)[:, None])).sum(axis=1)
    base = o.log_polar_base(P, 1 / 15000, 2 / 6e-3).astype(np.float64)
    sp = o.draw_spectrum_numpy(np.random.RandomState(9), base.astype(np.float32),
                               (base ** (-11 / 6) * 1e-3).astype(np.float32))
    fx = sp.rho.astype(float) * np.cos(sp.theta.astype(float))
    fy = sp.rho.astype(float) * np.sin(sp.theta.astype(float))
    a = sp.value.astype(np.complex128)
    X = n / 2 * d
    al, be = 2 * np.pi * fx, 2 * np.pi * fy
    k = np.flatnonzero((np.abs(al) * X <= Z) & (np.abs(be) * X > Z))
    assert 20 <= len(k) <= 80
    rows, cols = np.arange(0, n, 37), np.arange(0, n, 41)
    x, y = (cols - n / 2) * d, (rows - n / 2) * d
    aq = np.expm1(1j * np.outer(y, be[k])) * a[k]                 # (rows, K)
    direct = (aq @ np.expm1(1j * np.outer(x, al[k])).T).real
    m = np.arange(TL)
    Pm = 2 * (1j) ** m[None, :] * jv(m[None, :], (al[k] * X)[:, None])
    Pm[:, 0] = Pm[:, 0] / 2 - 1
    Tt = np.polynomial.chebyshev.chebvander((cols - n / 2) / (n / 2), TL - 1)
    block = (aq @ Pm).real @ Tt.T
    tol = (np.abs(a[k]) * tail(al[k] * X)).sum()          # |q| <= 2 below
    assert np.abs(block - direct).max() < max(1e-12 * np.abs(direct).max(), 2 * tol)
    g_direct = (np.expm1(1j * np.outer(x, al[k])) @ a[k]).real
    assert np.abs(Tt @ (a[k] @ Pm).real - g_direct).max() < max(1e-12 * np.abs(g_direct).max(), tol)


# (E, radices) of the Stockham plans in pa_fft.h
_PLANS = {64: (8, (8, 8)), 128: (16, (16, 8)), 256: (16, (16, 16)), 512: (8, (8, 8, 8)),
          1024: (16, (16, 8, 8)), 2048: (16, (16, 16, 8)), 4096: (16, (16, 16, 16))}


@pytest.mark.parametrize("n", sorted(_PLANS))
def test_padded_exchange_offsets(n):
    """The Stockham exchange addresses its padded LDS line (i + i // 16) with
    one register per thread and constant offsets: for every pass of every plan
    and every thread, lpad(base + r NS) = lpad(base) + lpad(r NS) for the
    stores (base = (j // NS) NS R + j % NS) and lpad(l + L e) = lpad(l) +
    lpad(L e) for the loads.  Checked against the direct addresses, which must
    also cover the line exactly once per pass."""
    lpad = lambda i: i + (i >> 4)
    e_pt, radices = _PLANS[n]
    L = n // e_pt
    ns = 1
    for p, r in enumerate(radices):
        last = p == len(radices) - 1
        m_groups = e_pt // r
        if not last:
            direct, split = [], []
            for l in range(L):
                for m in range(m_groups):
                    j = l + L * m
                    base = (j // ns) * ns * r + (j & (ns - 1))
                    direct += [lpad(base + q * ns) for q in range(r)]
                    split += [lpad(base) + lpad(q * ns) for q in range(r)]
            assert split == direct
            assert sorted(direct) == [lpad(i) for i in range(n)]
        ns *= r
    loads = [[lpad(l) + lpad(L * e) for e in range(e_pt)] for l in range(L)]
    assert loads == [[lpad(l + L * e) for e in range(e_pt)] for l in range(L)]
