"""Pins of the oracle's SPH primitives (Source/Functions.cpp) against mathematics.

None of these retype a reference formula: each checks a property the formula must have
(normalisation, derivative relations, inverse functions, spectral identities), chosen so a
dropped term, wrong constant, wrong sign or transposed operand fails.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import pyoracle as O

KERNELS = [(0, 2.0), (1, 2.0), (2, 3.0)]       # (kernel type, support in units of h)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kt,support", KERNELS)
def test_kernel_normalisation(dim, kt, support):
    """int W dV = 1 (Functions.cpp:25-63); the 3-D Wendland ("Quintic", KT=1) constant
    7/(8 pi h^3) integrates to 2/3 (the reference's constant, kept as is)."""
    h = 0.013
    R = support*h
    if dim == 2:
        f = lambda r: O.kernel(2, kt, r/h, h)*2.0*math.pi*r
    else:
        f = lambda r: O.kernel(3, kt, r/h, h)*4.0*math.pi*r*r
    pts = [k*h for k in range(1, int(support))]
    val, err = integrate.quad(f, 0.0, R, points=pts, epsabs=1e-13, epsrel=1e-12, limit=200)
    expect = 2.0/3.0 if (dim == 3 and kt == 1) else 1.0
    assert abs(val - expect) < 1e-9, val


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kt,support", KERNELS)
def test_kernel_support_and_positivity(dim, kt, support):
    h = 0.02
    qs = np.linspace(0.0, support + 0.5, 401)
    w = np.array([O.kernel(dim, kt, q, h) for q in qs])
    assert np.all(w[qs >= support] == 0.0)
    assert np.all(w[qs < support - 1e-9] > 0.0)
    assert np.all(np.diff(w[qs < support]) <= 1e-12)     # monotone decreasing in q


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kt,support", KERNELS)
def test_grad_kernel_is_dWdr_over_r(dim, kt, support):
    """GradKernel returns (1/r) dW/dr (Functions.cpp:65-106): central differences of Kernel."""
    h = 0.017
    for q in np.linspace(0.05, support - 0.05, 37):
        if any(abs(q - b) < 0.02 for b in range(1, int(support) + 1)):
            continue
        r = q*h
        d = 1e-7*h
        dw = (O.kernel(dim, kt, (r+d)/h, h) - O.kernel(dim, kt, (r-d)/h, h))/(2*d)
        gk = O.grad_kernel(dim, kt, q, h)
        assert abs(gk - dw/r) <= 1e-6*abs(dw/r) + 1e-9*abs(O.grad_kernel(dim, kt, 0.05, h)), (q, gk, dw/r)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kt,support", KERNELS)
def test_grad_kernel_limit_at_zero(dim, kt, support):
    """The q == 0 branch is the r -> 0 limit of (1/r) dW/dr."""
    h = 0.01
    g0 = O.grad_kernel(dim, kt, 0.0, h)
    g1 = O.grad_kernel(dim, kt, 1e-7, h)
    assert abs(g0 - g1) <= 1e-5*abs(g0)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kt,support", KERNELS)
def test_second_derivative_and_laplacian(dim, kt, support):
    """SecDerivativeKernel = W''(r) and LaplaceKernel = W'' + (D-1)/r W' (Functions.cpp:108-186)."""
    h = 0.02
    for q in np.linspace(0.1, support - 0.1, 29):
        if any(abs(q - b) < 0.03 for b in range(1, int(support) + 1)):
            continue
        r = q*h
        d = 1e-5*h
        w = lambda rr: O.kernel(dim, kt, rr/h, h)
        d2 = (w(r+d) - 2*w(r) + w(r-d))/(d*d)
        d1 = (w(r+d) - w(r-d))/(2*d)
        s2 = O.secderiv_kernel(dim, kt, q, h)
        lap = O.laplace_kernel(dim, kt, q, h)
        scale = abs(O.secderiv_kernel(dim, kt, 0.1, h)) + abs(d1/r)
        assert abs(s2 - d2) <= 1e-4*scale, (q, s2, d2)
        assert abs(lap - (d2 + (dim-1)/r*d1)) <= 1e-4*scale, (q, lap)


@pytest.mark.parametrize("eq", [0, 1, 2])
def test_eos_inverse(eq):
    """DensitySolid inverts EOS for all three equations (Functions.cpp:188-264)."""
    cs, p0, rho0 = 62.64, 13.0, 1000.0
    for rho in [950.0, 999.0, 1000.0, 1001.7, 1050.0]:
        P = O.eos(eq, cs, p0, rho, rho0)
        if eq == 2:
            back = O.density_solid(eq, cs, p0, P, rho0)
        else:
            back = O.density_solid(eq, cs, p0, P, rho0)
        assert abs(back - rho) <= 1e-12*rho


def test_eos_tait_values_and_sound_speed():
    """Tait: P(rho0) = P0, dP/drho = c(rho)^2 (SoundSpeed squared)."""
    cs, p0, rho0 = 62.64, 0.0, 1000.0
    assert O.eos(1, cs, p0, rho0, rho0) == 0.0
    for rho in [990.0, 1000.0, 1020.0]:
        d = 1e-4
        dp = (O.eos(1, cs, p0, rho+d, rho0) - O.eos(1, cs, p0, rho-d, rho0))/(2*d)
        c = O.sound_speed(1, cs, rho, rho0)
        assert abs(dp - c*c) <= 1e-6*c*c
    assert O.sound_speed(0, cs, 1234.0, rho0) == cs
    assert O.sound_speed(2, cs, 1234.0, rho0) == cs
    # linear EOS slope is c0^2
    assert abs((O.eos(0, cs, p0, 1001.0, rho0) - O.eos(0, cs, p0, 1000.0, rho0)) - cs*cs) < 1e-9*cs*cs


def test_viscous_laws_agree_for_equal_densities():
    """Morris 2 mu/(di dj) and Shao 8 mu/(di+dj)^2 coincide when di = dj (Functions.cpp:300-311);
    both are linear in vab and along it."""
    vab = np.array([0.3, -0.2, 0.1])
    xij = np.array([0.004, 0.007, -0.002])
    vij = np.array([0.1, 0.2, 0.3])
    rij = float(np.linalg.norm(xij))
    a = O.viscous_force(0, 1e-3, 1000.0, 1000.0, -5.0, vab, 3, 0, rij, 0.013, xij, vij)
    b = O.viscous_force(1, 1e-3, 1000.0, 1000.0, -5.0, vab, 3, 0, rij, 0.013, xij, vij)
    np.testing.assert_allclose(a, b, rtol=1e-14)
    np.testing.assert_allclose(a, 2e-3/1e6*(-5.0)*vab, rtol=1e-14)
    c = O.viscous_force(0, 1e-3, 990.0, 1010.0, -5.0, 2*vab, 3, 0, rij, 0.013, xij, vij)
    d = O.viscous_force(1, 1e-3, 990.0, 1010.0, -5.0, 2*vab, 3, 0, rij, 0.013, xij, vij)
    assert np.all(np.abs(d) < np.abs(c))     # (di+dj)^2/4 > di dj for di != dj


def test_viscous_incompressible_uses_laplacian():
    """VisEq 2: -mu/(di dj) Lap(W) vab; VisEq 3 adds the Takeda correction, which vanishes
    for vij = 0 only in its GK vij term (Functions.cpp:313-320)."""
    h = 0.013
    xij = np.array([0.009, 0.004, 0.0])
    rij = float(np.linalg.norm(xij))
    vab = np.array([1.0, 0.0, 0.0])
    lap = O.laplace_kernel(2, 0, rij/h, h)
    a = O.viscous_force(2, 2e-3, 1000.0, 1000.0, 0.0, vab, 2, 0, rij, h, xij, np.zeros(3))
    np.testing.assert_allclose(a, -2e-3/1e6*lap*vab, rtol=1e-14)
    b = O.viscous_force(3, 2e-3, 1000.0, 1000.0, 0.0, vab, 2, 0, rij, h, xij, np.zeros(3))
    np.testing.assert_allclose(b, a, rtol=1e-14)


def _random_sym(rng, scale=1.0):
    m = rng.normal(size=(3, 3))*scale
    return 0.5*(m + m.T)


def test_eig_sym3_matches_numpy():
    rng = np.random.default_rng(3)
    for _ in range(200):
        m = _random_sym(rng, 1e4)
        vals, vecs = O.eig_sym3(m)
        ref = np.linalg.eigvalsh(m)
        np.testing.assert_allclose(vals, ref, rtol=0, atol=1e-12*np.abs(ref).max())
        assert np.all(np.diff(vals) >= 0)
        np.testing.assert_allclose(vecs @ np.diag(vals) @ vecs.T, m, atol=1e-11*np.abs(m).max())
        np.testing.assert_allclose(vecs.T @ vecs, np.eye(3), atol=1e-13)
    # degenerate and diagonal inputs
    for m in [np.diag([3.0, 3.0, 3.0]), np.diag([-1.0, 2.0, 0.5]), np.zeros((3, 3))]:
        vals, vecs = O.eig_sym3(m)
        np.testing.assert_allclose(vals, np.sort(np.diag(m)))


def _tir_numpy(sig, ti, rho):
    w, v = np.linalg.eigh(sig)
    g = np.where(w > 0, -ti*w/(rho*rho), 0.0)
    return v @ np.diag(g) @ v.T


def test_tir_3d_is_spectral_function():
    """TIR = sum_k g(lambda_k) v_k v_k^T, g = -TI lambda/rho^2 for lambda > 0 (Domain.cpp:977-987)."""
    rng = np.random.default_rng(5)
    for _ in range(100):
        s = _random_sym(rng, 1e4)
        t = O.tir(3, s, 0.5, 2650.0)
        np.testing.assert_allclose(t, _tir_numpy(s, 0.5, 2650.0), atol=1e-12*np.abs(s).max()/2650.0**2)


def test_tir_2d_closed_form_matches_plane_spectral():
    """The 2-D atan/cos/sin form (Domain.cpp:958-976) equals the spectral function of the
    in-plane 2x2 block."""
    rng = np.random.default_rng(6)
    for _ in range(100):
        s = np.zeros((3, 3))
        b = rng.normal(size=(2, 2))*1e4
        s[:2, :2] = 0.5*(b + b.T)
        t = O.tir(2, s, 0.5, 2000.0)
        w, v = np.linalg.eigh(s[:2, :2])
        g = np.where(w > 0, -0.5*w/(2000.0**2), 0.0)
        ref = v @ np.diag(g) @ v.T
        np.testing.assert_allclose(t[:2, :2], ref, atol=1e-10*np.abs(s).max()/2000.0**2)
