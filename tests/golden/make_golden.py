"""Write tests/golden/conv_golden.npz: FP64 golden vectors for the oracle.

Run in the build container only (numpy's pocketfft, complex128):
    python tests/golden/make_golden.py
Inputs come from the portable splitmix64 generator, restated here in numpy
(bit-identical to oracle/oracle.c orc_fill_uniform / orc_fill_loguniform);
small cases store the inputs themselves, the 32^3 case stores their SHA-256.
Outputs: unnormalised conv = IFFT3(FFT3(sig) * FFT3(filt)) * N^3 (the
reference convention, src/fft_conv3D.cpp:201-233) and the forward FFT3 of sig.
"""
import hashlib
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def mix(seed, idx):
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * GAMMA
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def uniform(npts, seed):
    k = (mix(seed, np.arange(2 * npts)) >> np.uint64(40)).astype(np.int64)
    f = ((k - (1 << 23)).astype(np.float32) / np.float32(8388608.0)).astype(np.float32)
    return f.view(np.complex64)


def loguniform(npts, seed):
    k = (mix(seed, np.arange(2 * npts)) >> np.uint64(40)).astype(np.float64) / 16777216.0
    mag = 10.0 ** (6.0 * k[0::2] - 3.0)
    ph = 2.0 * np.pi * k[1::2]
    out = np.empty(npts, np.complex64)
    out.real = (mag * np.cos(ph)).astype(np.float32)
    out.imag = (mag * np.sin(ph)).astype(np.float32)
    return out


def conv_ref(n, sig, filt):
    s = sig.astype(np.complex128).reshape(n, n, n)
    f = filt.astype(np.complex128).reshape(n, n, n)
    return (np.fft.ifftn(np.fft.fftn(s) * np.fft.fftn(f)) * n ** 3).ravel()


def main():
    out = {}
    cases = [("u8", 8, "uniform", 11, 12), ("u16", 16, "uniform", 21, 22), ("l16", 16, "loguniform", 31, 32),
             ("u32", 32, "uniform", 41, 42)]
    for name, n, dist, s1, s2 in cases:
        gen = uniform if dist == "uniform" else loguniform
        sig, filt = gen(n ** 3, s1), gen(n ** 3, s2)
        out[name + "_meta"] = np.array([n, s1, s2, 0 if dist == "uniform" else 1], np.int64)
        digest = hashlib.sha256(sig.tobytes() + filt.tobytes()).hexdigest()
        out[name + "_sha256"] = np.frombuffer(bytes.fromhex(digest), np.uint8)
        if n <= 16:
            out[name + "_sig"] = sig
            out[name + "_filt"] = filt
            out[name + "_fft"] = np.fft.fftn(sig.astype(np.complex128).reshape(n, n, n)).ravel()
        out[name + "_conv"] = conv_ref(n, sig, filt)
    np.savez_compressed(os.path.join(HERE, "conv_golden.npz"), **out)


if __name__ == "__main__":
    main()
