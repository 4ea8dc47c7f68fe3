"""numpy restatement of the reference FFT product path -- TEST INFRASTRUCTURE.

Independent of the C oracle's explicit convolution (or_wfprod_pw); the two
must agree when the FFT grid is alias-free (the reference's "FFT EXACT"
regime).  Only tests/ may import it.
"""
import numpy as np


def pw_products_fft(sys, b1lo, nocc, b2lo, nunocc, step_r=None, imati=None):
    """numpy restatement of the reference FFT path (wavefproducts.f:598-688).

    FFTW BACKWARD (unnormalised, e^{+i}) == N * numpy.ifftn; FORWARD == fftn.
    G vectors go to the grid point imati.G mod nfft (:625-627, :660-662,
    :675-676).  step_r: theta(r) on the grid (x fastest, real or complex),
    multiplied into conj(psi1)/sqrt(vol) (:631-633).  Returns (n, ngptm)
    products; used by the tests as an independent check of the C oracle's
    explicit convolution.
    """
    N = sys.nfft
    NN = N[0] * N[1] * N[2]
    im = np.eye(3, dtype=np.int64) if imati is None else np.asarray(imati, np.int64)
    out = np.zeros((len(sys.k1) * nocc * nunocc, sys.ngptm), np.complex128)

    def grid(k, b):
        a = np.zeros(N[::-1], np.complex128)  # [z, y, x]
        s = sys.gpt[k, : sys.ngpt[k]] @ im.T
        a[s[:, 2] % N[2], s[:, 1] % N[1], s[:, 0] % N[0]] = sys.cpw[: sys.ngpt[k], b, k]
        return np.fft.ifftn(a) * NN  # BACKWARD

    t = 0
    for ik, (ka, kb) in enumerate(zip(sys.k1, sys.k2)):
        gm = (sys.gptm - sys.g2[ik]) @ im.T
        for b1 in range(nocc):
            r1 = np.conj(grid(ka, b1lo + b1) / np.sqrt(sys.vol))
            if step_r is not None:
                r1 = r1 * step_r.reshape(N[::-1])
            for b2 in range(nunocc):
                r2 = grid(kb, b2lo + b2) / NN
                P = np.fft.fftn(r1 * r2)  # FORWARD
                idx = t + b2 + nunocc * b1
                out[idx] = P[gm[:, 2] % N[2], gm[:, 1] % N[1], gm[:, 0] % N[0]]
        t += nocc * nunocc
    return out
