"""FFT grid setup (fft.f:36-238) and the step function on the grid
(wavefproducts.f:576-590): the oracle's numpy restatement pinned by what
the mathematics fixes, then the library's host functions against it (bit
for bit on the integer outputs).  CPU only: these are host functions."""
import numpy as np
import pytest

from oracle import fftgrid as fg
from project1_amd import spexhip

U_SKEW = np.array([[1, 1, 0], [0, 1, -1], [1, 1, 1]])  # det 1


def test_skew_unimodular():
    assert round(np.linalg.det(U_SKEW)) == 1


def _bcc():
    """Reciprocal basis of an fcc lattice (a bcc lattice), columns."""
    return 2 * np.pi * np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]], float).T


# ------------------------------------------------------------------ oracle --

def test_oracle_imat_orthogonal_lattice_is_identity():
    imat, imati = fg.determine_imat(np.diag([1.0, 1.3, 0.7]))
    assert (imat == np.eye(3)).all() and (imati == np.eye(3)).all()


def test_oracle_imat_undoes_a_skewed_description_of_a_cubic_lattice():
    """rlat = a * U (a simple cubic lattice in a skewed basis): some imat
    makes rlat * imat orthogonal (non-orthogonality 0), and imati = imat^-1."""
    rlat = 1.1 * U_SKEW.astype(float)
    imat, imati = fg.determine_imat(rlat)
    assert round(np.linalg.det(imat)) == 1
    assert (imat @ imati == np.eye(3)).all()
    assert fg.nonorth(rlat @ imat) < 1e-14
    r1 = rlat @ imat
    assert np.allclose(np.abs(r1), 1.1 * np.abs(np.round(r1 / 1.1)))  # a signed permutation


def test_oracle_imat_bcc_is_the_brute_force_minimum():
    """bcc cannot be made orthogonal.  Its primitive basis is already the
    best (1/3) and stays; described in a skewed basis, the choice reaches
    the minimum over all candidates (found by a vectorised search,
    independently of the scan), which is that 1/3 again."""
    imat, _ = fg.determine_imat(_bcc())
    assert (imat == np.eye(3)).all()
    rlat = _bcc() @ U_SKEW
    imat, _ = fg.determine_imat(rlat)
    cand = fg.unimodular_candidates()
    vals = fg.nonorth(np.einsum("ik,nkj->nij", rlat, cand.astype(float)))
    assert fg.nonorth(rlat @ imat) == pytest.approx(vals.min(), abs=1e-12)
    assert fg.nonorth(rlat @ imat) < fg.nonorth(rlat) - 0.1
    assert fg.nonorth(rlat @ imat) == pytest.approx(1 / 3, abs=1e-12)


def test_oracle_gcount_cubic_closed_form():
    """Simple cubic lattice of spacing a: half-width floor(gcutf/a) per axis,
    and the point count equals a brute-force count over a larger box."""
    for a, gcutf in ((1.0, 7.1), (0.8, 5.0), (1.3, 3.95)):
        nfft, cnt = fg.gcount_sphere(a * np.eye(3), np.eye(3, dtype=int), gcutf)
        m = int(np.floor(gcutf / a))
        assert nfft == (fg.smooth_size(2 * m + 1),) * 3
        r = np.arange(-m - 2, m + 3)
        g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3) * a
        assert cnt == int(((g * g).sum(1) <= gcutf**2).sum())


def test_oracle_smooth_size():
    assert [fg.smooth_size(n) for n in (1, 11, 13, 17, 29, 31, 37, 97)] == \
        [1, 12, 14, 18, 30, 32, 40, 98]


def test_oracle_step_grid_equals_direct_series():
    """The FFT form equals the truncated Fourier series summed directly,
    on an even grid (complex theta: +N/2 kept, -N/2 not) with a non-identity
    imat."""
    rlat = 1.1 * U_SKEW.astype(float)
    imat, _ = fg.determine_imat(rlat)
    N = (8, 10, 9)
    cext = (8, 6, 5)  # the extents imat*g reaches inside |G| <= gmax
    cs = fg.stepfunction_coeffs(rlat, 50.0, [1.2, 0.9], [[0, 0, 0], [0.5, 0.25, 0.3]], cext)
    gmax = 4.0
    th = fg.step_grid(N, imat, rlat, gmax, cs, cext)
    pts = [(0, 0, 0), (3, 1, 4), (7, 9, 8), (4, 5, 2)]
    want = fg.step_direct(N, imat, rlat, gmax, cs, cext, pts)
    got = [th[i + N[0] * (j + N[1] * k)] for i, j, k in pts]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_oracle_stepfunction_mean_and_sphere_limit():
    """theta(G=0) = interstitial volume fraction; the grid mean of theta is
    that value (only G = 0 survives the average); cstep(-G) = conj cstep(G)."""
    rlat = np.eye(3) * 2 * np.pi / 4.0
    vol = 64.0
    cext = (3, 3, 3)
    cs = fg.stepfunction_coeffs(rlat, vol, [1.0], [[0.1, 0.2, 0.3]], cext)
    assert cs[3, 3, 3] == pytest.approx(1 - 4 * np.pi / 3 / vol)
    np.testing.assert_allclose(cs, np.conj(cs[::-1, ::-1, ::-1]), atol=1e-15)
    th = fg.step_grid((9, 9, 9), np.eye(3, dtype=int), rlat, 6.0, cs, cext)
    assert th.mean() == pytest.approx(cs[3, 3, 3], abs=1e-14)
    assert np.abs(th.imag).max() < 1e-13  # odd grid: symmetric frequency set


# ------------------------------------------------------------------ library --

@pytest.mark.parametrize("rlat", [np.diag([1.0, 1.3, 0.7]), 1.1 * U_SKEW.astype(float), _bcc(),
                                  _bcc() @ U_SKEW,
                                  np.array([[1.0, 0.3, 0.1], [0.0, 0.9, -0.2], [0.2, 0.0, 1.1]])])
def test_library_determine_imat_matches_oracle(rlat):
    imat, imati = spexhip.determine_imat(rlat)
    want, wanti = fg.determine_imat(rlat)
    np.testing.assert_array_equal(imat, want)
    np.testing.assert_array_equal(imati, wanti)


@pytest.mark.parametrize("rlat,gcutf", [(np.eye(3), 7.1), (1.1 * U_SKEW.astype(float), 4.3),
                                        (_bcc() / (2 * np.pi), 3.3)])
def test_library_fft_grid_matches_oracle(rlat, gcutf):
    imat, _ = fg.determine_imat(rlat)
    got = spexhip.fft_grid(rlat, imat, gcutf)
    assert got == fg.gcount_sphere(rlat, imat, gcutf)
    # with the identity basis change, as fft_init without determine_imat
    assert spexhip.fft_grid(rlat, np.eye(3, dtype=int), gcutf) == \
        fg.gcount_sphere(rlat, np.eye(3, dtype=int), gcutf)


def test_library_step_grid_matches_oracle():
    rlat = 1.1 * U_SKEW.astype(float)
    imat, _ = fg.determine_imat(rlat)
    cext = (8, 6, 5)  # the extents imat*g reaches inside |G| <= gmax
    cs = fg.stepfunction_coeffs(rlat, 50.0, [1.2, 0.9], [[0, 0, 0], [0.5, 0.25, 0.3]], cext)
    for N in ((8, 10, 9), (16, 16, 16)):
        got = spexhip.step_grid(N, imat, rlat, 4.0, cs)
        want = fg.step_grid(N, imat, rlat, 4.0, cs, cext)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_library_step_grid_rejects_frequency_outside_cstep():
    cs = np.zeros((3, 3, 3), complex)
    cs[1, 1, 1] = 1.0
    with pytest.raises(spexhip.SpexError):
        spexhip.step_grid((8, 8, 8), np.eye(3, dtype=int), np.eye(3), 3.0, cs)
