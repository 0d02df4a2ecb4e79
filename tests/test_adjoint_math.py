"""The transposition identities the adjoint kernels rely on, checked with
numpy FFTs in the reference's layouts (real fields (ix,iy,iz) with x fastest
as C-order (nz,ny,nx); half spectra (kx,ky,kz)): in the real inner product of
the half spectrum, <a, b> = sum Re(conj(a) b),
  F^T mu = (1/N) I(w_x mu),  w_x = 1 at kx = 0, nx/2, else 1/2,
  I^T f  = N c_x F f,        c_x = 1 at kx = 0, nx/2, else 2,
F the forward R2C times 1/N and I the C2R inverse that ignores the imaginary
parts of the kx = 0 and kx = nx/2 lines; and i k x is symmetric."""
import numpy as np

NX, NY, NZ = 8, 6, 4
N = NX * NY * NZ


def F(x):  # physical (nz,ny,nx) -> half spectrum (nx/2+1, ny, nz), times 1/N
    return np.transpose(np.fft.rfftn(x, axes=(0, 1, 2)) / N, (2, 1, 0))


def I(mu):  # half spectrum -> physical: complex y/z inverses, then C2R along x
    a = np.transpose(mu, (2, 1, 0))             # (nz, ny, nxh)
    a = np.fft.ifft(a, axis=0) * NZ
    a = np.fft.ifft(a, axis=1) * NY
    a = a.copy()
    a[..., 0] = a[..., 0].real
    a[..., -1] = a[..., -1].real
    return np.fft.irfft(a, n=NX, axis=2) * NX


def dot(a, b):
    return float(np.sum(np.real(np.conj(a) * b)))


def wx(kind):
    w = np.full(NX // 2 + 1, 0.5 if kind == "w" else 2.0)
    w[0] = w[-1] = 1.0
    return w[:, None, None]


def test_forward_and_inverse_transposes():
    rng = np.random.default_rng(0)
    for _ in range(3):
        x = rng.standard_normal((NZ, NY, NX))
        mu = rng.standard_normal((NX // 2 + 1, NY, NZ)) + 1j * rng.standard_normal((NX // 2 + 1, NY, NZ))
        assert np.isclose(dot(mu, F(x)), np.sum(I(wx("w") * mu) / N * x), rtol=1e-12)
        assert np.isclose(np.sum(x * I(mu)), dot(N * wx("c") * F(x), mu), rtol=1e-12)


def test_curl_is_symmetric():
    rng = np.random.default_rng(1)
    k = [rng.standard_normal((NX // 2 + 1, NY, NZ)) for _ in range(3)]

    def ikx(v):
        return [1j * (k[(c + 1) % 3] * v[(c + 2) % 3] - k[(c + 2) % 3] * v[(c + 1) % 3]) for c in range(3)]

    a = [rng.standard_normal(k[0].shape) + 1j * rng.standard_normal(k[0].shape) for _ in range(3)]
    b = [rng.standard_normal(k[0].shape) + 1j * rng.standard_normal(k[0].shape) for _ in range(3)]
    lhs = sum(dot(a[c], ikx(b)[c]) for c in range(3))
    rhs = sum(dot(ikx(a)[c], b[c]) for c in range(3))
    assert np.isclose(lhs, rhs, rtol=1e-12)
