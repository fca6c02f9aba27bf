"""The oracle's momentum norms of the __K_MOM_ANALYSIS__ build
(nspt.cc:919-944) against numpy's FFT (the same forward, unnormalised
convention as FFTW_FORWARD of Quenched.cc:792-815) at every momentum -- the
norm of a complex field differs between p and -p, so the sign of the
transform is pinned too -- and Parseval's identity per order."""
import numpy as np
import pytest

from oracle import oracle as O


def numpy_knorm(L, nc, N, U):
    V = int(np.prod(L))
    S = 1 + nc * nc * N
    F = np.fft.fftn(U.reshape(tuple(L) + (4, S)), axes=(0, 1, 2, 3)).reshape(V, 4, S)
    return np.stack([(np.abs(F[:, :, 1 + k * nc * nc:1 + (k + 1) * nc * nc]) ** 2).sum(axis=(1, 2))
                     for k in range(N)])


@pytest.mark.parametrize("nc,L,N", [(3, (3, 4, 2, 5), 4), (2, (2, 6, 3, 2), 5), (3, (4, 4, 4, 4), 3)])
def test_knorm_is_the_power_spectrum_of_the_field(nc, L, N):
    U = O.random_field(L, nc, N, seed=1, scale=0.4)
    kn = O.knorm(L, nc, N, U)
    ref = numpy_knorm(L, nc, N, U)
    assert np.abs(kn - ref).max() <= 1e-14 * np.abs(ref).max()
    if nc == 3:
        # not symmetric under p -> -p: the sign of the transform matters (an
        # SU(2) field's norms are: conj(U) = s2 U s2 with the Pauli matrix s2)
        flip = numpy_knorm(L, nc, N, np.conj(U))
        assert np.abs(kn - flip).max() > 1e-3 * np.abs(ref).max()
    # Parseval: sum_p knorm_k(p) = V sum_x sum_mu |U_k(x)|^2
    V = int(np.prod(L))
    for k in range(N):
        e = (np.abs(U[:, 1 + k * nc * nc:1 + (k + 1) * nc * nc]) ** 2).sum()
        assert abs(kn[k].sum() - V * e) <= 1e-13 * V * e
