"""The identity behind the packed PSPI compat migration kernel, checked in numpy (no GPU).

pspiop (src/Imaging/c/wem.c:380-471) keeps Re(dref[x][iref1]) where vref2 - vref1 > 10 and the
full complex value elsewhere.  With z = s_R + i a r_R and e = s_I - i a r_I, L even in k:
    Re P_src(s) + i a Re P_rec(r) = IFFT(Re L FFT(z) + Im L FFT(e))
    Im P_src(s) - i a Im P_rec(r) = IFFT(Re L FFT(e) - Im L FFT(z))
(P_src uses conj(L), P_rec uses L).  The test runs several compat steps both ways on random
fields, phases and selections, and compares wavefields and image rows."""
import numpy as np
import pytest


def even_phase(rng, n):
    half = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
    k = np.arange(n)
    return half[np.minimum(k, n - k)]


def direct_step(u, L, iref1, cut, conj):
    A = np.fft.fft(u)
    out = np.zeros_like(u)
    for ir in range(L.shape[0]):
        y = np.fft.ifft(A * (np.conj(L[ir]) if conj else L[ir])) * len(u)
        sel = iref1 == ir
        out[sel] = np.where(cut[sel], y[sel].real, y[sel])
    return out


def packed_step(z, e, L, iref1, cut, cr):
    Z, E = np.fft.fft(z), np.fft.fft(e)
    n = len(z)
    out = np.zeros_like(z)
    for ir in range(L.shape[0]):
        w = np.fft.ifft(L[ir].real * Z + L[ir].imag * E) * n
        out[iref1 == ir] = w[iref1 == ir]
    enew = np.zeros_like(e)
    if cr >= 0:
        v = np.fft.ifft(L[cr].real * E - L[cr].imag * Z) * n
        enew[~cut] = v[~cut]
    return out, enew


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_packed_compat_steps_equal_the_reference_steps(seed):
    rng = np.random.default_rng(seed)
    n, nref, nsteps = 64, 4, 6
    s = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r = 30.0 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    alpha = np.linalg.norm(s) / np.linalg.norm(r)
    z, e = s.real + 1j * alpha * r.real, s.imag - 1j * alpha * r.imag
    for step in range(nsteps):
        L = np.stack([even_phase(rng, n) for _ in range(nref)])
        if step == 3:  # a depth with vmax <= vmin + 10: every point uncut, vref 0
            iref1, cut, cr = np.zeros(n, int), np.zeros(n, bool), 0
        else:
            iref1 = rng.integers(0, nref - 1, n)
            cut = np.ones(n, bool)
            top = rng.choice(n, size=1 + step % 3, replace=False)  # points at the depth's vmax
            iref1[top] = nref - 1
            cut[top] = False
            cr = nref - 1
        s = direct_step(s, L, iref1, cut, conj=True)
        r = direct_step(r, L, iref1, cut, conj=False)
        z, e = packed_step(z, e, L, iref1, cut, cr)
        scale = np.linalg.norm(s)
        np.testing.assert_allclose(z.real + 1j * e.real, s, atol=1e-9 * scale)
        np.testing.assert_allclose((z.imag - 1j * e.imag) / alpha, r, atol=1e-9 * np.linalg.norm(r))
        image = (r * np.conj(s)).real
        np.testing.assert_allclose((z.real * z.imag - e.real * e.imag) / alpha, image,
                                   atol=1e-9 * np.abs(image).max())


def test_sparse_sums_match_transforms_in_the_strided_mapping():
    """Depths with few uncut points replace FFT(e) and the inverse transform of the new e by
    direct sums: E(k) = sum_j e_j w^(k x_j) and e_new(x_j) = sum_k U(k) conj(w^(k x_j)),
    w = exp(-2 pi i / N), evaluated per thread t over k = t + q T with the recurrence
    w^(k x) = w^(t x) (w^(T x))^q from two twiddle-table entries."""
    rng = np.random.default_rng(9)
    n, P = 256, 16
    T = n // P
    tw = np.exp(-2j * np.pi * np.arange(n) / n)
    xs = np.array([3, 77, 200])
    ev = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    e = np.zeros(n, complex)
    e[xs] = ev
    E = np.zeros(n, complex)
    for t in range(T):
        for j, x in enumerate(xs):
            w, st = tw[(t * x) % n], tw[(T * x) % n]
            for q in range(P):
                E[t + q * T] += ev[j] * w
                w *= st
    np.testing.assert_allclose(E, np.fft.fft(e), atol=1e-12)
    U = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.ifft(U) * n
    for x in xs:
        acc = 0j
        for t in range(T):
            w, st = tw[(t * x) % n], tw[(T * x) % n]
            for q in range(P):
                acc += U[t + q * T] * np.conj(w)
                w *= st
        assert abs(acc - ref[x]) < 1e-9 * np.abs(ref).max()
