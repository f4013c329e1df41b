"""Pins the CPU oracle (oracle/nsp_oracle.cpp) against mathematics that fixes
each of its parts independently of how the oracle computes them:
polynomial exactness of the finite-difference tables, numpy's FFTs, dense
solves, numpy's 4x4 inverse, analytic Helmholtz/Poisson solutions, the
discrete Poiseuille fixed point, the closed-form decay of an m=0 swirl
eigenvector, a brute-force pseudo-spectral evaluation of the nonlinear
term, and, for the remesh tool, parity-polynomial exactness of its radial
interpolation, the identity on an unchanged grid and a spectral round trip.  The reference ships no tests or golden vectors of its own
(its repository holds none), so these are the pins.
"""
import numpy as np
import pytest

import oracle_lib
from oracle_lib import Oracle

IW = 3
TINY = dict(nr=8, nt=4, nz=5, lz=2 * np.pi, re=300.0, dt=0.01)
SMALL = dict(nr=16, nt=8, nz=9, lz=2 * np.pi, re=500.0, dt=0.01)
CFG1 = dict(nr=32, nt=64, nz=65, lz=4 * np.pi, re=2000.0, dt=0.005)


@pytest.fixture(scope="module")
def small():
    return Oracle(**SMALL)


@pytest.fixture(scope="module")
def cfg1():
    return Oracle(**CFG1)


def stencil(nr, i):
    """Columns of row i and the table slot offset (mod_radialFD.cu:10-30)."""
    left, right = max(0, i - IW), min(i + IW, nr - 1)
    return left, right, i + IW - right


def npoints(nr, i):
    """Number of Fornberg points behind row i (ghost points included)."""
    return 7 if i + IW <= nr - 1 else nr + IW - i


def apply_row(tab, p, i, f, nr):
    left, right, i0 = stencil(nr, i)
    return sum(tab[p, i, j + i0] * f[left + j] for j in range(right + 1 - left))


def mode_index(o, m, l):
    it = m if m >= 0 else m + o.nt
    return it, l


# ----------------------------------------------------------------- tables
@pytest.mark.parametrize("cfg", [SMALL, CFG1])
def test_grid_and_quadrature(cfg):
    o = Oracle(**cfg)
    T = o.tables()
    r, rdr = T["r"], T["rdr"]
    assert np.all(np.diff(r) > 0) and r[0] > 0
    assert abs(r[-1] - 1.0) < 1e-14
    # sum(rdr) telescopes to r[NR-1]*r[NR] - r[0]*r[-1] = 1
    assert abs(rdr.sum() - 1.0) < 1e-14
    # second-order quadrature of 2*int_0^1 f r dr
    nr = cfg["nr"]
    assert abs((rdr * (1 - r ** 2)).sum() - 0.5) < 2.0 / nr ** 2


@pytest.mark.parametrize("cfg", [SMALL, CFG1])
def test_fornberg_exactness_on_monomials(cfg):
    o = Oracle(**cfg)
    T = o.tables()
    r, D, Lu, Lp = T["r"], T["D"], T["Lu"], T["Lp"]
    nr = cfg["nr"]
    for p in (0, 1):  # 0: even functions, 1: odd functions
        for i in range(nr):
            npt = npoints(nr, i)
            for q in range(npt):
                if i < IW and q % 2 != p:
                    continue  # folded ghost points are exact only for the table's parity
                f = r ** q
                d_exact = q * r[i] ** (q - 1) if q > 0 else 0.0
                scale = sum(abs(D[p, i, :])) * max(1.0, r.max() ** q)
                assert abs(apply_row(D, p, i, f, nr) - d_exact) < 1e-12 * scale, (p, i, q)
                if i < nr - 1:
                    d2 = q * (q - 1) * r[i] ** (q - 2) if q > 1 else 0.0
                    lap = d2 + d_exact / r[i]
                    scale2 = sum(abs(Lu[p, i, :]))
                    assert abs(apply_row(Lu, p, i, f, nr) - lap) < 1e-11 * scale2, (p, i, q)
                    assert abs(apply_row(Lp, p, i, f, nr) - lap) < 1e-11 * scale2, (p, i, q)
        # wall rows: Lu = identity (Dirichlet), Lp = D (Neumann)
        assert np.array_equal(Lu[p, nr - 1], np.array([0, 0, 0, 0, 0, 0, 1.0]))
        assert np.array_equal(Lp[p, nr - 1], D[p, nr - 1])
    assert np.array_equal(T["dr1"], D[0, nr - 1])


def test_dr0_interpolates_to_axis(cfg1):
    T = cfg1.tables()
    r, dr0 = T["r"], T["dr0"]
    for q in range(4):
        assert abs(np.dot(dr0, r[:4] ** q) - (1.0 if q == 0 else 0.0)) < 1e-12


# ----------------------------------------------------------------- FFT
@pytest.mark.parametrize("cfg", [TINY, SMALL, CFG1])
def test_fft2_matches_numpy(cfg):
    o = Oracle(**cfg)
    rng = np.random.default_rng(1)
    spec = rng.standard_normal((o.ntp, o.nzp)) + 1j * rng.standard_normal((o.ntp, o.nzp))
    phys = o.fft2_inverse(spec)
    ref = np.fft.irfft2(spec, s=(o.ntp, o.nzf)) * (o.ntp * o.nzf)
    assert np.abs(phys - ref).max() < 1e-13 * np.abs(ref).max()
    x = rng.standard_normal((o.ntp, o.nzf))
    sp = o.fft2_forward(x)
    refs = np.fft.rfft2(x)
    assert np.abs(sp - refs).max() < 1e-13 * np.abs(refs).max()


# ----------------------------------------------------------------- LU solves
def dense_operator(o, T, kind, m, l):
    """Assembled operator of LUsolve_{up,um,uz,p}_k (mod_linear.cu:55-646)."""
    nr = o.nr
    r = T["r"]
    d_im = 0.51
    tab = T["Lp"] if kind == 3 else T["Lu"]
    fp = 1 if kind <= 1 else 0
    p = abs(abs(m) % 2 - fp)
    kz = 2 * np.pi / o.lz * l
    kt = float(m)
    A = np.zeros((nr, nr))
    for i in range(nr):
        left, right, i0 = stencil(nr, i)
        for j in range(right + 1 - left):
            A[i, left + j] = tab[p, i, j + i0]
        if i < nr - 1:
            r2 = r[i] ** 2
            dterm = -kz * kz - kt * kt / r2
            if kind == 0:
                dterm += -1 / r2 - 2 * kt / r2
            elif kind == 1:
                dterm += -1 / r2 + 2 * kt / r2
            A[i, i] += dterm
            if kind < 3:
                A[i, :] = -(d_im / o.re) * A[i, :]
                A[i, i] += 1.0 / o.dt
        elif kind == 3 and m == 0 and l == 0:
            A[i, :] = 0.0
            A[i, i] = 1.0
    return A


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_lusolve_matches_dense_solve(small, kind):
    o = small
    T = o.tables()
    rng = np.random.default_rng(kind)
    rhs = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
    sol = o.lusolve(kind, rhs)
    for it in range(o.nt):
        m = it if it < o.nt // 2 else it - o.nt
        for l in range(o.nz):
            A = dense_operator(o, T, kind, m, l)
            x = np.linalg.solve(A, rhs[:, it, l])
            scale = np.abs(x).max() * np.linalg.cond(A)
            assert np.abs(sol[:, it, l] - x).max() < 1e-14 * scale, (kind, m, l)


def test_lusolve_analytic_helmholtz(cfg1):
    """Polynomial solutions inside the stencils' exactness: round-off recovery."""
    o = cfg1
    T = o.tables()
    r = T["r"]
    d_im, re, dt = 0.51, o.re, o.dt
    for m, l in [(0, 0), (0, 3), (1, 2), (2, 1), (-2, 5)]:
        it, _ = mode_index(o, m, l)
        kz = 2 * np.pi / o.lz * l
        # uz: x = r^|m| (1 - r^2), parity of |m|
        q = abs(m)
        x = r ** q * (1 - r ** 2)
        lap = (q + 2) ** 2 * r ** q * 0 + (q * q * r ** (q - 2) if q else 0) - (q + 2) ** 2 * r ** q
        lap = lap - m * m * x / r ** 2 - kz * kz * x
        f = x / dt - d_im / re * lap
        f[-1] = 0.0
        rhs = o.zeros(); rhs[:, it, l] = f
        sol = o.lusolve(2, rhs)[:, it, l]
        assert np.abs(sol - x).max() < 1e-11, (m, l)
        # u+: x = r^s (1 - r^2), s of parity |m|+1 (exact for degree <= 4 near the wall)
        s = 0 if abs(m) % 2 == 1 else 1
        x = r ** s * (1 - r ** 2)
        d2x = (s * (s - 1) * r ** (s - 2) if s > 1 else 0) - (s + 2) * (s + 1) * r ** s
        d1x = (s * r ** (s - 1) if s > 0 else 0) - (s + 2) * r ** (s + 1)
        for kind, sg in [(0, -1), (1, +1)]:
            lap = d2x + d1x / r + (-1 + sg * 2 * m - m * m) * x / r ** 2 - kz * kz * x
            f = x / dt - d_im / re * lap
            f[-1] = 0.0
            rhs = o.zeros(); rhs[:, it, l] = f
            sol = o.lusolve(kind, rhs)[:, it, l]
            assert np.abs(sol - x).max() < 1e-11, (kind, m, l)
    # pressure with Neumann wall row: m=1, p = r - r^3/3 (p'(1)=0)
    m, l = 1, 2
    it, _ = mode_index(o, m, l)
    kz = 2 * np.pi / o.lz * l
    p = r - r ** 3 / 3
    f = -8 * r / 3 - kz * kz * p
    f[-1] = 0.0
    rhs = o.zeros(); rhs[:, it, l] = f
    sol = o.lusolve(3, rhs)[:, it, l]
    assert np.abs(sol - p).max() < 1e-9


# ----------------------------------------------------------------- radial operators
def test_derivr_parity_selection(cfg1):
    o = cfg1
    r = o.tables()["r"]
    u = o.zeros()
    # flag 1 (ur/ut): even m -> odd function, odd m -> even function
    cases = {(0, 0): r, (1, 1): r ** 2, (2, 3): r ** 3, (-1, 2): 1 + 0 * r}
    dexp = {(0, 0): 1 + 0 * r, (1, 1): 2 * r, (2, 3): 3 * r ** 2, (-1, 2): 0 * r}
    for (m, l), f in cases.items():
        it, _ = mode_index(o, m, l)
        u[:, it, l] = f * (1 + 2j)
    du = o.derivr(u, 1)
    for (m, l), d in dexp.items():
        it, _ = mode_index(o, m, l)
        assert np.abs(du[:, it, l] - d * (1 + 2j)).max() < 1e-10
    # flag 0 (uz/p): parity of |m|
    u = o.zeros()
    it, _ = mode_index(o, 3, 1)
    u[:, it, 1] = r ** 3
    it0, _ = mode_index(o, 2, 1)
    u[:, it0, 1] = r ** 2
    du = o.derivr(u, 0)
    assert np.abs(du[:, it, 1] - 3 * r ** 2).max() < 1e-10
    assert np.abs(du[:, it0, 1] - 2 * r).max() < 1e-10


def test_divergence_and_addgrad_polynomial(cfg1):
    o = cfg1
    r = o.tables()["r"]
    m, l = 1, 2
    it, _ = mode_index(o, m, l)
    kz = 2 * np.pi / o.lz * l
    ur, ut, uz = o.zeros(), o.zeros(), o.zeros()
    ur[:, it, l] = 1 - r ** 2
    ut[:, it, l] = 0.5
    uz[:, it, l] = 2 * r
    dv = o.divergence(ur, ut, uz)[:, it, l]
    exp = -2 * r + (1 - r ** 2) / r + 1j * m * 0.5 / r + 1j * kz * 2 * r
    assert np.abs(dv - exp).max() < 1e-10
    # wall divergence uses ur (not ur/r) at r = 1
    dw = o.divwall(ur, ut, uz)[it, l]
    assert abs(dw - exp[-1]) < 1e-10
    # addGrad subtracts (dp/dr, i m p / r, i kz p); p of parity |m| (odd): p = r
    p = o.zeros(); p[:, it, l] = r
    rr, rt, rz = o.addgrad(o.zeros(), o.zeros(), o.zeros(), p)
    assert np.abs(rr[:, it, l] + 1).max() < 1e-10
    assert np.abs(rt[:, it, l] + 1j * m * r / r).max() < 1e-12
    assert np.abs(rz[:, it, l] + 1j * kz * r).max() < 1e-12


def test_addlap_polynomial(cfg1):
    o = cfg1
    r = o.tables()["r"]
    d_im, re, dt = 0.51, o.re, o.dt
    m, l = 2, 1
    it, _ = mode_index(o, m, l)
    kz = 2 * np.pi / o.lz * l
    u = o.zeros()
    x = r * (1 - r ** 2)  # odd: u+- at even m
    u[:, it, l] = x
    d1 = 1 - 3 * r ** 2
    d2 = -6 * r
    for flag, sg in [(0, -1), (1, +1)]:
        out = o.addlap(o.zeros(), u, flag)[:, it, l]
        lap = d2 + d1 / r + (-1 + sg * 2 * m - m * m) * x / r ** 2 - kz * kz * x
        exp = x / dt + (1 - d_im) / re * lap
        assert np.abs(out[:-1] - exp[:-1]).max() < 1e-9
    u = o.zeros()
    x = r ** 2 * (1 - r ** 2)
    u[:, it, l] = x
    out = o.addlap(o.zeros(), u, 2)[:, it, l]
    lap = (4 - 16 * r ** 2) - m * m * x / r ** 2 - kz * kz * x
    assert np.abs(out[:-1] - (x / dt + (1 - d_im) / re * lap)[:-1]).max() < 1e-9


# ----------------------------------------------------------------- influence matrix
def to_rt(up, um):
    return 0.5 * (up + um), (up - um) / 2j


def test_influence_inverse_matches_numpy(small):
    o = small
    col, (uhp, uhm, uhz, uPp, uPm, uPz) = o.influence()
    zero = o.zeros()
    d1 = o.divwall(*to_rt(uhp, zero), zero)
    d2 = o.divwall(*to_rt(zero, uhm), zero)
    d3 = o.divwall(zero, zero, uhz)
    d4 = o.divwall(*to_rt(uPp, uPm), uPz)
    for it in range(o.nt):
        for l in range(o.nz):
            M = np.array([[1, 0, 0, uPp[-1, it, l]],
                          [0, 1, 0, uPm[-1, it, l]],
                          [0, 0, 1j, uPz[-1, it, l]],
                          [d1[it, l], d2[it, l], d3[it, l], d4[it, l]]], np.complex128)
            inv = np.linalg.inv(M)
            k = it * o.nz + l
            assert np.abs(col[k] - inv[:, 3]).max() < 1e-12 * np.abs(inv).max(), (it, l)
    # homogeneous velocity solutions take unit wall values (mod_boundary.cu:226-232)
    assert np.allclose(uhp[-1], 1.0) and np.allclose(uhm[-1], 1.0) and np.allclose(uhz[-1], 1j)


def test_corbcon_enforces_noslip_and_wall_divergence(cfg1):
    o = cfg1
    rng = np.random.default_rng(5)
    f = [rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape) for _ in range(3)]
    for x in f:
        x[-1] = 0.0  # wall rows as written by write_bcon before the solve
    up, um, uz = o.corbcon(*f)
    ur, ut = to_rt(up, um)
    scale = max(np.abs(x).max() for x in f)
    for x in (ur, ut, uz):
        assert np.abs(x[-1]).max() < 1e-12 * scale
    dw = o.divwall(ur, ut, uz)
    assert np.abs(dw).max() < 1e-9 * scale


def test_adj_flux_sets_bulk_velocity(cfg1):
    o = cfg1
    T = o.tables()
    rng = np.random.default_rng(2)
    uz = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
    out = o.adj_flux(uz)
    assert abs(np.dot(T["rdr"], out[:, 0, 0].real) - 0.5) < 1e-14
    # only mode (0,0) is touched
    diff = out - uz
    diff[:, 0, 0] = 0
    assert np.abs(diff).max() == 0.0
    qc, uc = o.flux_qc()
    assert abs(qc - np.dot(T["rdr"], uc)) < 1e-15


# ----------------------------------------------------------------- whole step
def test_poiseuille_is_discrete_fixed_point():
    o = Oracle(nr=16, nt=8, nz=9, lz=2 * np.pi, re=800.0, dt=0.01)
    T = o.tables()
    r, rdr = T["r"], T["rdr"]
    alpha = 0.5 / np.dot(rdr, 1 - r ** 2)
    ur, ut, uz = o.zeros(), o.zeros(), o.zeros()
    uz[:, 0, 0] = alpha * (1 - r ** 2)
    o.set_state(ur, ut, uz)
    n_it, err = o.step(100)
    a, b, c = o.get_state()
    assert np.all(n_it == 1)
    assert np.abs(c - uz).max() < 1e-13
    assert np.abs(a).max() < 1e-13 and np.abs(b).max() < 1e-13


def test_swirl_eigenvector_decays_by_closed_form_factor():
    o = Oracle(nr=32, nt=8, nz=9, lz=2 * np.pi, re=1000.0, dt=0.01)
    T = o.tables()
    r, rdr, Lu = T["r"], T["rdr"], T["Lu"]
    nr = o.nr
    # A = Lu (odd table, m = 0) - 1/r^2 on the interior rows, wall value 0
    A = np.zeros((nr, nr))
    for i in range(nr):
        left, right, i0 = stencil(nr, i)
        for j in range(right + 1 - left):
            A[i, left + j] = Lu[1, i, j + i0]
        A[i, i] -= 1 / r[i] ** 2
    Ai = A[: nr - 1, : nr - 1]
    mu, V = np.linalg.eig(Ai)
    k = np.argmax(mu.real)
    assert abs(mu[k].imag) < 1e-10
    mu = mu[k].real
    v = np.concatenate([V[:, k].real, [0.0]])
    v /= np.abs(v).max()
    eps = 1e-2
    alpha = 0.5 / np.dot(rdr, 1 - r ** 2)
    ur, ut, uz = o.zeros(), o.zeros(), o.zeros()
    ut[:, 0, 0] = eps * v
    uz[:, 0, 0] = alpha * (1 - r ** 2)
    o.set_state(ur, ut, uz)
    nsteps = 50
    n_it, _ = o.step(nsteps)
    d_im, re, dt = 0.51, o.re, o.dt
    g = (1 + (1 - d_im) * mu * dt / re) / (1 - d_im * mu * dt / re)
    _, b, _ = o.get_state()
    assert np.abs(b[:, 0, 0] - g ** nsteps * eps * v).max() < 1e-12 * eps
    # continuum limit: mu -> -j_{1,1}^2 = -14.6819...
    assert abs(mu + 14.681970642123893) < 1e-3 * 14.68


# ----------------------------------------------------------------- nonlinear term
def test_nonlinear_brute_force():
    """Direct Fourier sums on the 3/2 grid against the oracle's FFT pipeline
    (mod_io_nonlinear.cu:380-409)."""
    o = Oracle(**TINY)
    nr, nt, nz, ntp, nzf = o.nr, o.nt, o.nz, o.ntp, o.nzf
    T = o.tables()
    r, D = T["r"], T["D"]
    rng = np.random.default_rng(11)
    u = [rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape) for _ in range(3)]
    out = o.nonlinear(*u)
    ms = np.array([it if it < nt // 2 else it - nt for it in range(nt)])
    kz = 2 * np.pi / o.lz * np.arange(nz)
    th = 2 * np.pi * np.arange(ntp) / ntp
    z = o.lz * np.arange(nzf) / nzf

    def dr(f, flag):
        g = np.zeros_like(f)
        for it in range(nt):
            p = abs(abs(ms[it]) % 2 - flag)
            for i in range(nr):
                left, right, i0 = stencil(nr, i)
                g[i, it, :] = sum(D[p, i, j + i0] * f[left + j, it, :] for j in range(right + 1 - left))
        return g

    def phys(c):  # c: (nt, nz) level spectrum -> (ntp, nzf) real field
        E = np.exp(1j * (ms[:, None, None, None] * th[None, None, :, None]
                         + kz[None, :, None, None] * z[None, None, None, :]))
        terms = c[:, :, None, None] * E
        f = terms[:, 0].sum(axis=0).real
        f = f + 2 * terms[:, 1:].sum(axis=(0, 1)).real
        return f

    dur = [dr(u[0], 1), dr(u[1], 1), dr(u[2], 0)]
    ref = [o.zeros() for _ in range(3)]
    for i in range(nr):
        U = [phys(x[i]) for x in u]
        R = [phys(x[i]) for x in dur]
        Th = [phys(1j * ms[:, None] * x[i] / r[i]) for x in u]
        Z = [phys(1j * kz[None, :] * x[i]) for x in u]
        ur, ut, uz = U
        N = [-(ur * R[0] + ut * Th[0] + uz * Z[0] - ut * ut / r[i]),
             -(ur * R[1] + ut * Th[1] + uz * Z[1] + ur * ut / r[i]),
             -(ur * R[2] + ut * Th[2] + uz * Z[2])]
        Einv = np.exp(-1j * (ms[:, None, None, None] * th[None, None, :, None]
                             + kz[None, :, None, None] * z[None, None, None, :]))
        for c in range(3):
            co = (N[c][None, None] * Einv).sum(axis=(2, 3)) / (ntp * nzf)
            co[nt // 2, :] = 0.0
            co[:, nz - 1] = 0.0
            ref[c][i] = co
    # mean pressure gradient on mode (0,0) (mod_radialFD.cu:100-111)
    dw = sum(D[0, nr - 1, 3 + j] * u[2][nr - 4 + j, 0, 0].real for j in range(4))
    ref[2][:, 0, 0] -= 2.0 / o.re * dw
    for c in range(3):
        assert np.abs(out[c] - ref[c]).max() < 1e-12 * np.abs(ref[c]).max(), c


# ---------------------------------------------------------------- remesh tool
REMESH_PAIRS = [(8, 8), (48, 64), (64, 48), (32, 96), (16, 12), (96, 160)]


def remesh_grid(nr):
    """Radial grid of the remesh tool (tools/Re_Mesh/src/mod_remesh.cu:61-79)."""
    n_p = nr + int(np.sqrt(nr))
    r = np.zeros(nr)
    for i in range(n_p - nr, n_p):
        ir = n_p - i - 1
        r[i - n_p + nr] = 0.5 * (1.0 + np.cos(np.pi * ir / n_p))
    for _ in range(10):
        dr = 1.5 * r[0] - 0.5 * r[1]
        r = r * (1.0 + dr) - dr
    return r


@pytest.mark.parametrize("nro,nrn", REMESH_PAIRS)
def test_remesh_mirror_weights_interpolate_parity_polynomials(nro, nrn):
    """Six-point Lagrange interpolation through the axis is exact for
    polynomials of degree <= 5 of the field's parity: half 0 (mirrored with +)
    for even powers of r, half 1 (mirrored with -) for odd powers."""
    RM = oracle_lib.remesh_weights(nro, nrn, 1)
    ro, rn = remesh_grid(nro), remesh_grid(nrn)
    for p in range(3):
        np.testing.assert_allclose(RM[0] @ ro ** (2 * p), rn ** (2 * p), rtol=0, atol=1e-8)
        np.testing.assert_allclose(RM[1] @ ro ** (2 * p + 1), rn ** (2 * p + 1), rtol=0, atol=1e-8)
    assert RM[0, -1, -1] == 1.0 and RM[1, -1, -1] == 1.0 and np.count_nonzero(RM[:, -1]) == 2


@pytest.mark.parametrize("nro,nrn", REMESH_PAIRS)
def test_remesh_reference_assignment_differs_only_where_stencil_starts_inside(nro, nrn):
    """The reference maps a mirrored point at stencil position j to column
    IW-1-j, which is its own column only when the stencil starts at the
    outermost mirrored point.  That happens unless ro[0] < rn[i] <= ro[IW-1]
    (stencil start rau[1] or rau[2]): exactly those rows differ from the mirror
    assignment, and the others are identical."""
    ref = oracle_lib.remesh_weights(nro, nrn, 0)
    mir = oracle_lib.remesh_weights(nro, nrn, 1)
    ro, rn = remesh_grid(nro), remesh_grid(nrn)
    expect = {i for i in range(nrn - 1) if ro[0] < rn[i] <= ro[IW - 1]}
    differ = {i for i in range(nrn) if not np.array_equal(ref[:, i], mir[:, i])}
    assert differ <= expect
    if nro != nrn:  # on an unchanged grid every row hits a node: no mirrored weight
        assert differ == expect
    # both assignments keep each row's weights summing to one on even fields
    np.testing.assert_allclose(ref[0].sum(axis=1), 1.0, atol=1e-12)


def test_remesh_same_grid_is_identity():
    RM = oracle_lib.remesh_weights(24, 24, 0)
    np.testing.assert_allclose(RM[0], np.eye(24), atol=1e-14)
    np.testing.assert_allclose(RM[1], np.eye(24), atol=1e-14)


def test_remesh_spectral_pad_then_truncate_round_trip():
    """Padding to more azimuthal/axial modes and truncating back recovers the
    field (same NR), apart from the (0,0) adjustment of modeZZ_k."""
    rng = np.random.default_rng(11)
    shape = (8, 8, 9)
    f = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    big = oracle_lib.remesh(f, (8, 12, 17))
    back = oracle_lib.remesh(big, shape)
    for c in range(3):
        g = f[c].copy()
        if c == 0:
            g[:, 0, 0] = 0.0
        else:
            g[:, 0, 0] = g[:, 0, 0].real
        np.testing.assert_allclose(back[c], g, rtol=0, atol=1e-14)
        # new modes are zero: m = +-4, +-5 rows and l >= 9 of the padded field
        assert np.all(big[c][:, 4:8, :] == 0) and np.all(big[c][:, :, 9:] == 0)


# ------------------------------------------------ physical fields and cross sections
def brute_physical(f, nt, nz, ntp, nzf):
    """u(theta_j, z_n) of one radial level by direct summation: the theta sum
    over every stored m (row it holds m = it for it < NT/2, it - NT otherwise),
    then the real axial synthesis Re Y(0) + 2 sum_{l>=1} Re(Y(l) e^{i 2 pi l n/NZF})."""
    th = 2 * np.pi * np.arange(ntp) / ntp
    m = np.array([it if it < nt // 2 else it - nt for it in range(nt)])
    Y = np.exp(1j * np.outer(th, m)) @ f  # (ntp, nz)
    n = np.arange(nzf)
    E = np.exp(2j * np.pi * np.outer(np.arange(1, nz), n) / nzf)  # (nz-1, nzf)
    return Y[:, :1].real + 2 * np.real(Y[:, 1:] @ E)


def test_physical_field_brute_force():
    rng = np.random.default_rng(5)
    o = Oracle(**TINY)
    shape = (TINY["nr"], TINY["nt"], TINY["nz"])
    f = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    o.set_state(*f)
    for c in range(3):
        ph = o.physical(c)
        assert ph.shape == (o.nr, o.ntp, o.nzf + 2)
        assert np.all(ph[:, :, o.nzf:] == 0)
        for i in range(o.nr):
            want = brute_physical(f[c][i], o.nt, o.nz, o.ntp, o.nzf)
            np.testing.assert_allclose(ph[i, :, :o.nzf], want, rtol=0, atol=1e-12)


def test_cross_sections_closed_form():
    """One mode (m=2, l=1) in ur and ut, one (m=1, l=2) in uz: the theta mean of
    the square of 2|c| cos(phase) is 2|c|^2 at every z, and the periodic
    three-point rule sums to the plain mean, so q(z) = sum_i rdr_i 2(|a_i|^2 +
    |b_i|^2) and u2(z) = sum_i rdr_i 2|c_i|^2."""
    o = Oracle(**SMALL)
    rng = np.random.default_rng(3)
    shape = (o.nr, o.nt, o.nz)
    ur, ut, uz = (np.zeros(shape, np.complex128) for _ in range(3))
    a = rng.standard_normal(o.nr) + 1j * rng.standard_normal(o.nr)
    b = rng.standard_normal(o.nr) + 1j * rng.standard_normal(o.nr)
    c = rng.standard_normal(o.nr) + 1j * rng.standard_normal(o.nr)
    ur[:, 2, 1] = a
    ut[:, 2, 1] = b
    uz[:, 1, 2] = c
    o.set_state(ur, ut, uz)
    rdr = o.tables()["rdr"]
    q, u2 = o.cross()
    np.testing.assert_allclose(q, np.sum(rdr * 2 * (np.abs(a) ** 2 + np.abs(b) ** 2)), rtol=1e-13)
    np.testing.assert_allclose(u2, np.sum(rdr * 2 * np.abs(c) ** 2), rtol=1e-13)


# ----------------------------------------------------------------- diagnostics
def test_diagnostics_on_poiseuille_closed_form():
    """On uz00 = alpha (1 - r^2) the centre-line velocity is alpha (dr0 is
    exact to degree 3), u_tau = sqrt(2 alpha / Re) (the wall row of the even D
    table is exact to degree 6), the bulk flux is alpha sum rdr (1 - r^2), the
    field is divergence-free and no-slip (wrt_frc_k, mod_io_nonlinear.cu:40-52)."""
    o = Oracle(**CFG1)
    T = o.tables()
    r, rdr = T["r"], T["rdr"]
    alpha = 1.7
    ur, ut, uz = o.zeros(), o.zeros(), o.zeros()
    uz[:, 0, 0] = alpha * (1 - r ** 2)
    o.set_state(ur, ut, uz)
    d = o.diagnostics()
    assert abs(d["u_cl"] - alpha) < 1e-14 * alpha
    assert abs(d["u_tau"] - np.sqrt(2 * alpha / o.re)) < 1e-12 * np.sqrt(2 * alpha / o.re)
    assert abs(d["flux"] - alpha * np.dot(rdr, 1 - r ** 2)) < 1e-15
    assert d["div_l2"] < 1e-12 and d["wall_div_max"] < 1e-12
    assert d["wall_max"] < 1e-15
    # KE of a single radial profile: 1/2 sum rdr uz^2 (mode (0,0) has weight 1)
    assert abs(d["ke"] - 0.5 * np.dot(rdr, (alpha * (1 - r ** 2)) ** 2)) < 1e-14


def hermitian_field(o, rng, scale=1.0):
    """Random spectral field whose l = 0 column is the spectrum of a real
    function of theta (u(-m, 0) = conj u(m, 0), u(0, 0) real, no m = -NT/2),
    so that the stored modes are exactly those of a real physical field."""
    f = scale * (rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape))
    nt = o.nt
    f[:, 0, 0] = f[:, 0, 0].real
    f[:, nt // 2, 0] = 0.0
    for it in range(1, nt // 2):
        f[:, nt - it, 0] = np.conj(f[:, it, 0])
    return f


def test_diagnostics_kinetic_energy_is_the_volume_mean_of_physical_samples():
    """KE = 1/2 sum_i rdr_i <u.u>(r_i), <.> the mean over the 3/2-padded
    (theta, z) grid of the physical samples (nspo_physical, pinned by brute
    force): Parseval gives the mode sum with weight 1 for l = 0, 2 otherwise."""
    o = Oracle(**SMALL)
    rng = np.random.default_rng(21)
    f = [hermitian_field(o, rng) for _ in range(3)]
    o.set_state(*f)
    rdr = o.tables()["rdr"]
    ms = 0.0
    for c in range(3):
        ph = o.physical(c)[:, :, :o.nzf]
        ms = ms + (ph ** 2).mean(axis=(1, 2))
    ke_phys = 0.5 * np.dot(rdr, ms)
    d = o.diagnostics()
    assert abs(d["ke"] - ke_phys) < 1e-13 * ke_phys


def test_diagnostics_divergence_and_wall_norms_polynomial():
    """Single-mode polynomial field with an analytic divergence: div_l2 =
    sqrt(sum_i w_l |rdr_i| |div_i|^2) (w_l = 2 for l > 0), the wall norms are
    the analytic wall values."""
    o = Oracle(**CFG1)
    T = o.tables()
    r, rdr = T["r"], T["rdr"]
    m, l = 1, 2
    it, _ = mode_index(o, m, l)
    kz = 2 * np.pi / o.lz * l
    ur, ut, uz = o.zeros(), o.zeros(), o.zeros()
    ur[:, it, l] = 1 - r ** 2 + 0.25 * r ** 2     # even (odd m): wall value 0.25
    ut[:, it, l] = 0.5 * (1 - r ** 4)             # even, wall value 0
    uz[:, it, l] = 2 * r * (1 - r ** 2)           # odd, wall value 0
    o.set_state(ur, ut, uz)
    d = o.diagnostics()
    fr, dfr = 1 - 0.75 * r ** 2, -1.5 * r
    div = dfr + fr / r + 1j * m * 0.5 * (1 - r ** 4) / r + 1j * kz * 2 * r * (1 - r ** 2)
    want = np.sqrt(np.sum(2.0 * np.abs(rdr) * np.abs(div) ** 2))
    assert abs(d["div_l2"] - want) < 1e-10 * want
    assert abs(d["wall_max"] - 0.25) < 1e-14
    assert abs(d["wall_div_max"] - abs(dfr[-1] + fr[-1])) < 1e-10


def test_reference_ic_matches_numpy_transform_of_init_per():
    """initField with restart == 0 (mod_io_nonlinear.cu:424-449): u_z00 = 1 - r^2
    (base_k) plus ur, ut sampled by init_per on the NTP x NZF grid (z scaled to
    [0, 1)), transformed by numpy's rfft2, normalised by NTP*NZF and truncated
    as padBackward_k does (m = -NT/2 and l = NZ-1 zeroed)."""
    o = Oracle(**SMALL)
    A_P = 0.3
    o.reference_ic(A_P)
    got = o.get_state()
    r = o.tables()["r"]
    nt, nz, ntp, nzf = o.nt, o.nz, o.ntp, o.nzf
    th = np.arange(ntp) / ntp * 2 * np.pi
    z = np.arange(nzf) / nzf
    ux = np.exp(-10.0 * np.sin(np.pi * z) ** 2)
    rows = [it if it < nt // 2 else ntp + (it - nt) for it in range(nt)]
    for i in range(o.nr):
        a = (1 - r[i] ** 2) ** 2
        b = a - 4 * r[i] ** 2 * (1 - r[i] ** 2)
        for c, prof in ((0, a * np.outer(np.sin(th), ux)), (1, b * np.outer(np.cos(th), ux))):
            S = np.fft.rfft2(A_P * prof) / (ntp * nzf)
            want = S[rows, :nz]
            want[nt // 2, :] = 0.0
            want[:, nz - 1] = 0.0
            assert np.abs(got[c][i] - want).max() < 1e-15, (c, i)
    want_z = o.zeros()
    want_z[:, 0, 0] = 1 - r ** 2
    assert np.array_equal(got[2], want_z)


# --------------------------------------------- whole step at m != 0: dense one-mode map
def dense_step_map(o, T, m, l, alpha, n_corr):
    """The time step of mod_integrate.cu:164-228, linearised about the
    Poiseuille profile U = alpha (1 - r^2) and restricted to one Fourier mode
    (m, l), assembled as a dense 3NR x 3NR matrix from the radial tables.

    Built from the operators' definitions, not from the oracle's code: the
    linearised convective term -(U d_z u + u_r U' e_z) (products of mode (0,0)
    with mode (m, l) land in (m, l) exactly on the 3/2 grid; U' = D U),
    addLap_k's explicit part, div_k, the Neumann pressure solve, addGrad_k,
    the +/- decoupling, the Dirichlet Helmholtz solves (numpy.linalg.solve),
    and the wall correction as the 4x4 linear system (numpy) that makes the
    wall velocity and the wall divergence zero with the homogeneous solutions
    of setBCon (mod_boundary.cu:205-255): unit-wall Helmholtz solutions and the
    gradient of the unit-flux pressure solution."""
    nr = o.nr
    r, D, Lu = T["r"], T["D"], T["Lu"]
    d_im, re, dt = 0.51, o.re, o.dt
    kz = 2 * np.pi / o.lz * l
    I = np.eye(nr)

    def band(tab, p):
        A = np.zeros((nr, nr))
        for i in range(nr):
            left, right, i0 = stencil(nr, i)
            for j in range(right + 1 - left):
                A[i, left + j] = tab[p, i, j + i0]
        return A

    pu = abs(abs(m) % 2 - 1)   # parity table of ur, ut, u+-
    pp = abs(abs(m) % 2)       # parity table of uz, p
    Dr_u, Dr_p = band(D, pu), band(D, pp)
    U = alpha * (1 - r ** 2)
    dU = band(D, 0) @ U
    Z = np.zeros((nr, nr))
    # linearised N(u) on (ur, ut, uz) stacked
    NL = np.block([[np.diag(-1j * kz * U), Z, Z],
                   [Z, np.diag(-1j * kz * U), Z],
                   [np.diag(-dU), Z, np.diag(-1j * kz * U)]])
    # r/theta <-> +/- (decoupleForward_k / decoupleBackward_k)
    P = np.block([[I, 1j * I, Z], [I, -1j * I, Z], [Z, Z, I]])
    Pi = np.linalg.inv(P)

    def diag_term(kind):
        de = -kz * kz - m * m / r ** 2
        if kind == 0:
            de = de - 1 / r ** 2 - 2 * m / r ** 2
        elif kind == 1:
            de = de - 1 / r ** 2 + 2 * m / r ** 2
        return de

    def lap(kind):
        return band(Lu, pu if kind < 2 else pp) + np.diag(diag_term(kind))

    Lex = [I / dt + (1 - d_im) / re * lap(c) for c in range(3)]   # addLap_k, every row
    H = []
    for c in range(3):   # LUsolve_u*_k: (1/dt) I - (d_im/Re)(Lu + diag), wall row identity
        A = I / dt - d_im / re * lap(c)
        A[-1] = 0.0
        A[-1, -1] = 1.0
        H.append(A)
    Lp = band(T["Lp"], pp).astype(complex)
    Lp[:-1] += np.diag(-kz * kz - m * m / r ** 2)[:-1]   # Neumann wall row D p = rhs
    Div = np.hstack([Dr_u + np.diag(1 / r), np.diag(1j * m / r), 1j * kz * I])   # div_k
    Grad = np.vstack([Dr_p, np.diag(1j * m / r), 1j * kz * I])                    # addGrad_k
    e = np.zeros(nr); e[-1] = 1.0
    # homogeneous solutions (setBCon)
    hp = np.linalg.solve(H[0], e)
    hm = np.linalg.solve(H[1], e)
    hz = np.linalg.solve(H[2], 1j * e)
    ph = np.linalg.solve(Lp, -e)
    hP = P @ (Grad @ ph)
    h = [np.concatenate([hp, 0 * e, 0 * e]), np.concatenate([0 * e, hm, 0 * e]),
         np.concatenate([0 * e, 0 * e, hz]), hP]

    def divwall(x):  # divWall_k on +/- stacked input
        y = Pi @ x
        urw, utw, uzw = y[:nr], y[nr:2 * nr], y[2 * nr:]
        return Dr_u[-1] @ urw + urw[-1] + 1j * m * utw[-1] + 1j * kz * uzw[-1]

    def solve(rhs_pm):  # steps 4-10 of predictor()/corrector(), +/- in and out
        rt = Pi @ rhs_pm
        dv = Div @ rt
        dv[-1] = 0.0
        p = np.linalg.solve(Lp, dv)
        rt = rt - Grad @ p
        b = P @ rt
        for c in range(3):
            b[(c + 1) * nr - 1] = 0.0
        x = np.concatenate([np.linalg.solve(H[c], b[c * nr:(c + 1) * nr]) for c in range(3)])
        # a0 h+ + a1 h- + a2 hz + a3 hP: zero wall velocity, zero wall divergence
        # (x holds one column per input vector)
        M = np.zeros((4, 4), complex)
        rhs4 = np.zeros((4,) + x.shape[1:], complex)
        for c in range(3):
            M[c] = [hj[(c + 1) * nr - 1] for hj in h]
            rhs4[c] = -x[(c + 1) * nr - 1]
        M[3] = [divwall(hj) for hj in h]
        rhs4[3] = -divwall(x)
        a = np.linalg.solve(M, rhs4)
        return x + sum(np.multiply.outer(h[j], a[j]) for j in range(4))

    def explicit(u_pm):
        return np.concatenate([Lex[c] @ u_pm[c * nr:(c + 1) * nr] for c in range(3)])

    # G maps u^n (r/theta/z stacked) to u^{n+1}
    Nn = P @ NL                                     # N(u^n), +/-
    G = Pi @ solve(Nn + explicit(P))                # predictor
    for _ in range(n_corr):
        G = Pi @ solve(d_im * (P @ NL @ G) + (1 - d_im) * Nn + explicit(P))
    return G


@pytest.mark.parametrize("m,l", [(1, 1), (2, 3), (-3, 2)])
def test_step_at_nonzero_mode_matches_dense_linear_map(m, l):
    """One oracle step of Poiseuille plus a 1e-6 perturbation in mode (m, l):
    the mode's coefficients equal the dense linearised step map applied to the
    perturbation, to round-off (the quadratic terms are 1e-12 of it and fall
    in other modes)."""
    o = Oracle(nr=24, nt=8, nz=9, lz=2 * np.pi, re=600.0, dt=0.01)
    T = o.tables()
    r, rdr = T["r"], T["rdr"]
    alpha = 0.5 / np.dot(rdr, 1 - r ** 2)
    rng = np.random.default_rng(7 + l)
    it, _ = mode_index(o, m, l)
    # a divergence-free-ish no-slip perturbation of the mode's parities
    pu = abs(abs(m) % 2 - 1)
    v = [(rng.standard_normal(6) + 1j * rng.standard_normal(6)) for _ in range(3)]
    prof = lambda c, par: (1 - r ** 2) * sum(c[q] * r ** (2 * q + par) for q in range(6))
    eps = 1e-6
    ur, ut, uz = o.zeros(), o.zeros(), o.zeros()
    ur[:, it, l] = eps * prof(v[0], pu)
    ut[:, it, l] = eps * prof(v[1], pu)
    uz[:, it, l] = eps * prof(v[2], 1 - pu)
    uz[:, 0, 0] = alpha * (1 - r ** 2)
    o.set_state(ur, ut, uz)
    x0 = np.concatenate([ur[:, it, l], ut[:, it, l], uz[:, it, l]])
    n_it, _ = o.step(1)
    assert n_it[0] == 1
    a, b, c = o.get_state()
    got = np.concatenate([a[:, it, l], b[:, it, l], c[:, it, l]])
    G = dense_step_map(o, T, m, l, alpha, int(n_it[0]))
    want = G @ x0
    assert np.linalg.norm(got - want) < 1e-11 * np.linalg.norm(want)
    # every other fluctuation mode stays at the quadratic level
    for f in (a, b, c):
        g = f.copy()
        g[:, it, l] = 0.0
        g[:, 0, 0] = 0.0
        assert np.abs(g).max() < 10 * eps ** 2


def chebyshev_pipe_eigenvalues(m, k, re, alpha, n):
    """Eigenvalues of pipe flow linearised about U = alpha (1 - r^2), mode
    exp(i(m theta + k z) + lambda t), by Chebyshev collocation in primitive
    variables (u_r, u_theta, u_z, p) on the n+1 positive points of an odd
    2n+1-point Chebyshev grid on [-1, 1], folded with each field's parity
    (u_r, u_theta of parity m+1, u_z and p of parity m).  Momentum at the
    interior points, no-slip at r = 1, continuity at every point; a generalised
    eigenproblem solved with scipy.  Independent of the finite-difference
    tables."""
    import scipy.linalg
    N = 2 * n + 1
    x = np.cos(np.pi * np.arange(N + 1) / N)
    c = np.ones(N + 1); c[0] = c[-1] = 2.0
    c = c * (-1.0) ** np.arange(N + 1)
    X = np.tile(x, (N + 1, 1)).T
    dX = X - X.T
    Dm = np.outer(c, 1 / c) / (dX + np.eye(N + 1))
    Dm = Dm - np.diag(Dm.sum(axis=1))
    D2 = Dm @ Dm
    h = n + 1
    idx, mir = np.arange(h), N - np.arange(h)   # x[N - j] = -x[j]

    def fold(A, par):
        s = 1.0 if par == 0 else -1.0
        return A[np.ix_(idx, idx)] + s * A[np.ix_(idx, mir)]

    r = x[:h]
    pv, ps = (m + 1) % 2, m % 2
    D1v, D2v, D1s, D2s = fold(Dm, pv), fold(D2, pv), fold(Dm, ps), fold(D2, ps)
    R, Ri, Ri2 = np.diag(r), np.diag(1 / r), np.diag(1 / r ** 2)
    U = np.diag(alpha * (1 - r ** 2))
    dU = np.diag(-2 * alpha * r)
    I, Z = np.eye(h), np.zeros((h, h))
    lapv = D2v + Ri @ D1v - (m * m) * Ri2 - k * k * I
    laps = D2s + Ri @ D1s - (m * m) * Ri2 - k * k * I
    nu = 1.0 / re
    Arr = -1j * k * U + nu * (lapv - Ri2)
    Art = nu * (-2j * m * Ri2)
    Atr = nu * (2j * m * Ri2)
    Azz = -1j * k * U + nu * laps
    A = np.block([[Arr, Art, Z, -D1s],
                  [Atr, Arr, Z, -1j * m * Ri],
                  [-dU, Z, Azz, -1j * k * I],
                  [D1v + Ri, 1j * m * Ri, 1j * k * I, Z]]).astype(complex)
    B = np.block([[I, Z, Z, Z], [Z, I, Z, Z], [Z, Z, I, Z], [Z, Z, Z, Z]]).astype(complex)
    for q in range(3):   # no-slip rows at r = 1 (point 0)
        row = q * h
        A[row] = 0.0
        B[row] = 0.0
        A[row, q * h] = 1.0
    lam = scipy.linalg.eig(A, B, right=False)
    return lam[np.isfinite(lam) & (np.abs(lam) < 1e3)]


def test_step_map_least_stable_mode_matches_chebyshev_eigenvalue():
    """The dense step map (pinned above to the oracle's step at round-off)
    amplifies its least-stable mode by g per step; log(g)/dt matches the
    least-stable eigenvalue of the continuous linearised problem from an
    independent Chebyshev-collocation solve, to the discretisation accuracy of
    the finite differences (NR = 48) and of the time scheme."""
    o = Oracle(nr=48, nt=8, nz=9, lz=2 * np.pi, re=500.0, dt=0.002)
    T = o.tables()
    r, rdr = T["r"], T["rdr"]
    alpha = 0.5 / np.dot(rdr, 1 - r ** 2)
    m, l = 1, 1
    G = dense_step_map(o, T, m, l, alpha, 1)
    g = np.linalg.eigvals(G)
    g = g[np.argmax(np.abs(g))]
    lam_fd = np.log(g) / o.dt
    k = 2 * np.pi / o.lz * l
    lam = chebyshev_pipe_eigenvalues(m, k, o.re, alpha, 40)
    lam2 = chebyshev_pipe_eigenvalues(m, k, o.re, alpha, 56)
    lead = lam[np.argmax(lam.real)]
    lead2 = lam2[np.argmax(lam2.real)]
    assert abs(lead - lead2) < 1e-8 * abs(lead)        # collocation converged
    assert lead.real < 0                                 # linearly stable pipe flow
    assert abs(lam_fd - lead) < 1e-4 * abs(lead), (lam_fd, lead)
