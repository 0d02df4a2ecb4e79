"""Pins the oracle's building blocks against closed forms, brute force and
invariants that a plausible mistake (wrong sign, index, normalisation or
transposed operand) would break.  CPU only."""
import math

import numpy as np
import pytest

from oracle import flusi_oracle as fo


def _grid(nx, ny, nz, xl=2 * math.pi, yl=2 * math.pi, zl=2 * math.pi, tmp_path=None):
    p = fo.Params(nx=nx, ny=ny, nz=nz, xl=xl, yl=yl, zl=zl)
    return fo.Grid(p)


def test_fft_matches_bruteforce_dft():
    """cof_p3dfft.f90:33-60: R2C with 1/N, output (iz,iy,ix) with kx >= 0."""
    g = _grid(8, 6, 4, 1.3, 2.1, 0.7)
    rng = np.random.default_rng(0)
    f = rng.standard_normal((g.nz, g.ny, g.nx))
    fk = fo.fft(g, f)
    ix, iy, iz = 3, 5, 2
    x = np.arange(g.nx)[None, None, :]
    y = np.arange(g.ny)[None, :, None]
    z = np.arange(g.nz)[:, None, None]
    ph = np.exp(-2j * np.pi * (ix * x / g.nx + iy * y / g.ny + iz * z / g.nz))
    ref = (f * ph).sum() / (g.nx * g.ny * g.nz)
    assert abs(fk[ix, iy, iz] - ref) < 1e-14
    assert fk.shape == (g.nx // 2 + 1, g.ny, g.nz)


def test_ifft_is_unnormalised_c2r_ignoring_dc_nyquist_imag():
    """cof_p3dfft.f90:65-83: unnormalised inverse; after the y,z inverse
    transforms, the imaginary parts of the kx=0 and kx=nx/2 entries of each
    x line do not contribute (FFTW c2r)."""
    g = _grid(8, 4, 6)
    rng = np.random.default_rng(1)
    fk = rng.standard_normal((5, 4, 6)) + 1j * rng.standard_normal((5, 4, 6))
    f = fo.ifft(g, fk)
    a = np.fft.ifft(np.fft.ifft(fk, axis=2), axis=1) * (g.ny * g.nz)   # (kx, y, z)
    x = np.arange(g.nx)
    out = np.zeros((g.nz, g.ny, g.nx))
    for yy in range(g.ny):
        for zz in range(g.nz):
            line = a[0, yy, zz].real + ((-1.0) ** x) * a[4, yy, zz].real
            for k in range(1, 4):
                line = line + 2.0 * (a[k, yy, zz] * np.exp(2j * np.pi * k * x / g.nx)).real
            out[zz, yy, :] = line
    assert np.max(np.abs(f - out)) < 1e-12


def test_fft_unit_test_closed_forms():
    """fft_unit_test.f90:81-253: ifft(fft(u)) = u and spectral derivatives of
    sines are cosines, to 1e-13."""
    g = _grid(32, 24, 16, 2.0, 3.0, 1.5)
    X = g.x[None, None, :]
    Y = g.y[None, :, None]
    Z = g.z[:, None, None]
    u = np.sin(2 * np.pi * X / 2.0) * np.cos(4 * np.pi * Y / 3.0) + np.sin(2 * np.pi * 3 * Z / 1.5) + 0 * X * Y
    assert np.max(np.abs(fo.ifft(g, fo.fft(g, u)) - u)) < 1e-13
    for ax, (K, L, m) in enumerate(((g.KX, 2.0, 1), (g.KY, 3.0, 2), (g.KZ, 1.5, 3))):
        coord = (X, Y, Z)[ax]
        f = np.sin(2 * np.pi * m * coord / L) + 0 * X * Y * Z
        d = fo.ifft(g, 1j * K * fo.fft(g, f))
        exact = (2 * np.pi * m / L) * np.cos(2 * np.pi * m * coord / L) + 0 * X * Y * Z
        assert np.max(np.abs(d - exact)) < 1e-12


def test_wavenumbers_and_nyquist_sign():
    """cof_p3dfft.f90:661-680: ky(iy=ny/2) = -ny/2 * scaley."""
    g = _grid(8, 8, 6, 2.0, 4.0, 3.0)
    assert g.ky1[4] == -4 * (2 * math.pi / 4.0)
    assert g.kz1[3] == -3 * (2 * math.pi / 3.0)
    assert g.kx1[4] == 4 * (2 * math.pi / 2.0)


def test_dealias_predicate_bruteforce():
    """vis.f90:23-53: ellipsoid (kx/kxt)^2+(ky/kyt)^2+(kz/kzt)^2 >= 1 with
    kxt = 2/3 (n/2-1), evaluated mode by mode."""
    g = _grid(32, 24, 16)
    for ix in range(g.nxh):
        for iy in range(g.ny):
            for iz in range(g.nz):
                jy = (iy + g.ny // 2) % g.ny - g.ny // 2
                jz = (iz + g.nz // 2) % g.nz - g.nz // 2
                v = (ix / (2 / 3 * 15)) ** 2 + (jy / (2 / 3 * 11)) ** 2 + (jz / (2 / 3 * 7)) ** 2
                if abs(v - 1.0) > 1e-12:
                    assert g.alias[ix, iy, iz] == (v >= 1.0)
    assert g.alias[g.nx // 2, 0, 0]   # x Nyquist is always removed
    assert g.alias[0, g.ny // 2, 0]


def test_cal_vis_values():
    g = _grid(16, 16, 16)
    vis = fo.cal_vis(g, 1e-3, 0.01)
    assert vis[1, 2, 3] == pytest.approx(math.exp(-1e-3 * 0.01 * (1 + 4 + 9)), rel=1e-15)
    assert vis[8, 0, 0] == 0.0


def test_projection_idempotent_and_solenoidal():
    g = _grid(16, 12, 8, 1.0, 2.0, 3.0)
    rng = np.random.default_rng(2)
    n = rng.standard_normal((3, g.nxh, g.ny, g.nz)) + 1j * rng.standard_normal((3, g.nxh, g.ny, g.nz))
    pn = fo.add_grad_pressure(g, n)
    div = g.KX * pn[0] + g.KY * pn[1] + g.KZ * pn[2]
    assert np.max(np.abs(div)) < 1e-12 * np.max(np.abs(n)) * 10
    assert np.allclose(fo.add_grad_pressure(g, pn), pn, atol=1e-13)
    assert np.array_equal(pn[:, 0, 0, 0], n[:, 0, 0, 0])   # zero mode untouched


def test_curl_of_gradient_vanishes_and_biot_savart_inverts_curl():
    g = _grid(16, 16, 16)
    rng = np.random.default_rng(3)
    phi = rng.standard_normal((g.nxh, g.ny, g.nz)) + 1j * rng.standard_normal((g.nxh, g.ny, g.nz))
    grad = np.stack([1j * g.KX * phi, 1j * g.KY * phi, 1j * g.KZ * phi])
    assert np.max(np.abs(fo.curl(g, grad))) < 1e-12
    u = fo.add_grad_pressure(g, rng.standard_normal((3, g.nxh, g.ny, g.nz)) + 0j)
    u[:, 0, 0, 0] = 0
    back = fo.vorticity2velocity(g, fo.curl(g, u))
    assert np.allclose(back, u, atol=1e-12)


def test_energy_parseval():
    """integrals.f90:522-552 equals the physical-space energy."""
    g = _grid(16, 8, 12, 1.0, 2.0, 3.0)
    rng = np.random.default_rng(4)
    u = rng.standard_normal((3, g.nz, g.ny, g.nx))
    ek = fo.energy_k(g, fo.fft3(g, u))
    ep = fo.compute_energies(g, u)[0]
    assert ek == pytest.approx(ep, rel=1e-12)


def test_truncate_two_significant_digits():
    """fluid_time_step.f90:1051-1060 es7.1 round trip."""
    assert fo.truncate(1.246262e-2) == 1.2e-2
    assert fo.truncate(5.74999e-4) == 5.7e-4
    assert fo.truncate(9.96e-3) == 1.0e-2


def test_steps_profile():
    """mask.f90:68-90: 1 inside, 0 outside, 1/2 at the interface."""
    assert fo.steps(0.0, 1.0, 0.1) == 1.0
    assert fo.steps(1.0, 1.0, 0.1) == pytest.approx(0.5)
    assert fo.steps(1.2, 1.0, 0.1) == 0.0


def _write_ini(path, **kw):
    sec = {
        "Resolution": {"nx": 16, "ny": 16, "nz": 16},
        "Time": {"iTimeMethodFluid": "AB2", "dt_fixed": 1e-3, "Tmax": 1.0, "CFL": 0.1},
        "ReynoldsNumber": {"nu": 0.05},
        "InitialCondition": {"inicond": "quiescent"},
        "Penalization": {"iPenalization": 0},
        "Geometry": {"xl": 6.283185307179586, "yl": 6.283185307179586, "zl": 6.283185307179586},
        "Saving": {"tintegral": 100.0, "itdrag": 99999, "tsave": 100.0},
        "Dealiasing": {"iDealias": 1},
        "MeanFlow": {"ux": 0.0, "uy": 0.0, "uz": 0.0},
    }
    for k, v in kw.items():
        s, key = k.split("__")
        sec.setdefault(s, {})[key] = v
    with open(path, "w") as fh:
        for s, kv in sec.items():
            fh.write(f"[{s}]\n")
            for k, v in kv.items():
                fh.write(f"{k}={v};\n")
    return str(path)


@pytest.mark.parametrize("method", ["AB2", "RK2", "RK4"])
def test_steppers_decay_a_stokes_mode(tmp_path, method):
    """A single shear mode u_x = sin(3z) has u x w = grad(.) (projected away), so
    the integrating-factor steppers reproduce exp(-nu k^2 t) exactly and RK4
    (explicit diffusion) its 4th-order amplification polynomial."""
    ini = _write_ini(tmp_path / "s.ini", Time__iTimeMethodFluid=method, Time__dt_fixed=2e-3)
    sim = fo.Simulation(ini)
    sim.init_fields()
    g = sim.g
    u = np.zeros((3, g.nz, g.ny, g.nx))
    u[0] = np.sin(3 * g.z)[:, None, None] + 0 * u[0]
    sim.uk = fo.fft3(g, u)
    n = 10
    for _ in range(n):
        sim.step()
    k2nu = 0.05 * 9
    dt = 2e-3
    a = sim.uk[0, 0, 0, 3]
    a0 = fo.fft3(g, u)[0, 0, 0, 3]
    if method == "RK4":
        z = -k2nu * dt
        amp = (1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24) ** n
    elif method == "RK2":
        # first step: stale factor is zero, so the predictor is u*E and the
        # corrector adds dt/2 N1 = -dt/2 nu... N1 has no diffusion (IF) -> pure decay
        amp = math.exp(-k2nu * dt * n)
    else:
        amp = math.exp(-k2nu * dt * n)
    assert abs(a / a0 - amp) < 1e-12


def test_uniform_flow_through_static_sphere_rhs_zero_mode(tmp_path):
    """Uniform u=(U,0,0), static mask: u x w = 0 so N(k=0) = -U <mask>/eta."""
    ini = _write_ini(tmp_path / "s.ini", Penalization__iPenalization=1, Penalization__iMask="sphere",
                     Penalization__eps=1e-2, Geometry__Size=2.0, InitialCondition__inicond="MeanFlow",
                     MeanFlow__ux=0.7)
    sim = fo.Simulation(ini)
    sim.init_fields()
    nlk, _ = sim.cal_nlk(0.0, sim.uk)
    expect = -0.7 * sim.mask.mean() / 1e-2
    assert nlk[0, 0, 0, 0].real == pytest.approx(expect, rel=1e-12)
    assert abs(nlk[1, 0, 0, 0]) < 1e-12 and abs(nlk[2, 0, 0, 0]) < 1e-12


def test_beltrami_flow_has_no_nonlinear_term(tmp_path):
    """ABC flow: w = u (k=1), so u x w = 0 and the unpenalised RHS vanishes."""
    ini = _write_ini(tmp_path / "s.ini")
    sim = fo.Simulation(ini)
    sim.init_fields()
    g = sim.g
    X, Y, Z = np.meshgrid(g.x, g.y, g.z, indexing="xy")
    X = g.x[None, None, :] + 0 * g.z[:, None, None] + 0 * g.y[None, :, None]
    Y = g.y[None, :, None] + 0 * X
    Z = g.z[:, None, None] + 0 * X
    u = np.stack([np.sin(Z) + np.cos(Y), np.sin(X) + np.cos(Z), np.sin(Y) + np.cos(X)])
    nlk, _ = sim.cal_nlk(0.0, fo.fft3(g, u))
    assert np.max(np.abs(nlk)) < 1e-13


def test_insect_fseries_and_rotations():
    """module_helpers.f90:30-55 and rotation_matrices.f90:4-37 (passive)."""
    u, udt = fo.fseries_eval(0.25, 2.0, [1.0, 0.5], [0.25, 0.0])
    assert u == pytest.approx(1.0 + 1.0 * 0.0 + 0.25 * 1.0 + 0.5 * -1.0, abs=1e-15)
    assert udt == pytest.approx(2 * math.pi * (-1.0 * 1.0) + 4 * math.pi * (0.0), abs=1e-12)
    R = fo.Rz(math.pi / 2)
    assert np.allclose(R @ np.array([1.0, 0, 0]), [0, -1, 0])  # passive: axes rotate, not the vector
    for M in (fo.Rx(0.3), fo.Ry(-1.1), fo.Rz(2.0)):
        assert np.allclose(M @ M.T, np.eye(3))
