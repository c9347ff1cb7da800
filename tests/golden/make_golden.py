"""Write the FFT golden vectors in tests/golden/fft_golden.npz.

The reference holds no fixture files for this path (no testdata/). These
vectors pin the FFT conventions (forward exp(-2 pi i .), unnormalised;
fft2D = rows then columns) on the reference's own deterministic inputs,
computed with numpy.fft (pocketfft, FP64) -- a library routine independent of
the oracle:
  - refField(n, 1.3) for the sizes of TestGPUFFT2DMatchesCPU
    (optics/gpu_test.go:12-22, 54), 2-D forward;
  - sin(0.01 i) + i cos(0.017 i) for the sizes of TestGPUFFT1DMatchesCPU
    (optics/gpu_test.go:86-90), 1-D forward;
  - simpleRNG(99) draws for TestBluesteinMatchesNaiveDFT
    (optics/asm_improve_test.go:23-38; simpleRNG at optics/fft_test.go:62-68).
Inputs are regenerated by the tests from the same formulas; only outputs are
stored. Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

SIZES_2D = [8, 16, 32, 64, 12, 27, 100]
SIZES_1D = [1024, 4096, 1000, 2049]
SIZES_RNG = [100, 200, 480]


def ref_field(n, seed):
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    re = np.sin(seed * (i + 1) * 0.7) + 0.5 * np.cos(seed * (j + 1))
    im = np.cos(seed * (i - j) * 0.3) - 0.25 * np.sin(seed * (i * j + 1))
    return re + 1j * im


def signal_1d(n):
    i = np.arange(n)
    return np.sin(i * 0.01) + 1j * np.cos(i * 0.017)


class SimpleRNG:
    def __init__(self, seed):
        self.seed = seed

    def next(self):
        self.seed = (self.seed * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return float(self.seed >> 11) / 9007199254740992.0 - 0.5


def rng_sequences():
    rng = SimpleRNG(99)
    out = {}
    for n in SIZES_RNG:
        a = np.empty(n, dtype=np.complex128)
        for k in range(n):
            re = rng.next()
            im = rng.next()
            a[k] = complex(re, im)
        out[n] = a
    return out


def main():
    data = {}
    for n in SIZES_2D:
        data[f"fft2_refField_{n}"] = np.fft.fft2(ref_field(n, 1.3))
    for n in SIZES_1D:
        data[f"fft1_signal_{n}"] = np.fft.fft(signal_1d(n))
    for n, a in rng_sequences().items():
        data[f"fft1_rng99_{n}"] = np.fft.fft(a)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **data)
    print("wrote", os.path.join(HERE, "fft_golden.npz"), len(data), "arrays")


if __name__ == "__main__":
    main()
