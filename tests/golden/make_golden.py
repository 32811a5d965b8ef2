"""Writes tests/golden/fft_golden.npz: FP64 golden vectors for the oracle.

The expected outputs come from numpy.fft (pocketfft, an independent C
implementation; not the reference, which has no golden vectors of its own).
Inputs are iid Uniform[-1,1) from numpy's PCG64 with fixed seeds.  Sign and
normalisation follow FFTW's convention (forward exp(-2 pi i jk/n), backward
exp(+...), both unnormalised), i.e. numpy's fft and n * ifft.

Run once: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def cplx(rng, shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def main():
    rng = np.random.default_rng(20241203)
    out = {}
    # N-D C2C forward / backward (the reference's test matrix dims D2..D5)
    for name, shape in [("c2c_8x8x8", (8, 8, 8)), ("c2c_16x16x16", (16, 16, 16)),
                        ("c2c_64x64", (64, 64)), ("c2c_8x8x8x8", (8, 8, 8, 8)),
                        ("c2c_4x4x4x4x4", (4, 4, 4, 4, 4)), ("c2c_12x10x6", (12, 10, 6))]:
        x = cplx(rng, shape)
        out[name + "_in"] = x
        out[name + "_fwd"] = np.fft.fftn(x)
        out[name + "_bwd"] = np.fft.ifftn(x) * x.size
    # 1D lengths: powers of two, 2^a3^b5^c, primes, composite with a large prime
    for n in [1, 2, 3, 5, 7, 16, 48, 80, 97, 120, 160, 1000, 1024, 1021]:
        x = cplx(rng, (n,))
        out[f"c2c_1d_{n}_in"] = x
        out[f"c2c_1d_{n}_fwd"] = np.fft.fft(x)
    # R2C half spectra (even and odd last axis) and C2R (unnormalised)
    for name, shape in [("r2c_16x12x10", (16, 12, 10)), ("r2c_9x15", (9, 15)),
                        ("r2c_8x8x8x8", (8, 8, 8, 8)), ("r2c_1d_160", (160,)),
                        ("r2c_1d_81", (81,))]:
        x = rng.uniform(-1, 1, shape)
        X = np.fft.rfftn(x)
        out[name + "_in"] = x
        out[name + "_fwd"] = X
        out[name + "_c2r"] = np.fft.irfftn(X, s=shape, axes=list(range(len(shape)))) * x.size
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
