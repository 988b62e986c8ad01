"""Generate the exponential-of-semicircle kernel table used in hipng.hip (kBetaW, kErr1D).

For each oversampling sigma and support W, scan beta/W and measure the relative L2 error of a
1-D type-1 NUFFT (random frequencies in [-1/2, 1/2) cycles/pixel, random complex strengths,
N = 256 output pixels) gridded with psi(z) = exp(beta (sqrt(1 - z^2) - 1)) and corrected by
the kernel's Fourier transform, against the exact sum. Prints one line per (sigma, W) with the
best beta/W and its error. Run: python rascil_amd/csrc/gen_es_table.py (a few minutes).
"""

import numpy
from numpy.polynomial.legendre import leggauss

SIGMAS = [1.25, 1.5, 1.75, 2.0, 2.5]
WS = range(3, 17)
rng = numpy.random.default_rng(0)


def kernel_ft(xi, W, beta, Q=200):
    z, wq = leggauss(Q)
    psi = numpy.exp(beta * (numpy.sqrt(1 - z * z) - 1))
    return (W / 2) * numpy.cos(numpy.pi * W * numpy.outer(xi, z)) @ (wq * psi)


def nufft1_error(W, sigma, beta, N=256, M=3000):
    G = int(numpy.ceil(sigma * N / 2)) * 2
    u = rng.uniform(-0.5, 0.5, M)
    a = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    t = numpy.arange(-N // 2, N // 2)
    exact = numpy.exp(2j * numpy.pi * numpy.outer(t, u)) @ a
    ug = u * G + G / 2
    k0 = numpy.floor(ug - W / 2).astype(int) + 1
    grid = numpy.zeros(G, complex)
    for i in range(W):
        k = k0 + i
        z = 2 * (k - ug) / W
        numpy.add.at(grid, k % G, a * numpy.exp(beta * (numpy.sqrt(numpy.maximum(1 - z * z, 0)) - 1)))
    kk = numpy.arange(G) - G // 2
    approx = (numpy.exp(2j * numpy.pi * numpy.outer(t, kk) / G) @ grid) / kernel_ft(t / G, W, beta)
    return numpy.linalg.norm(approx - exact) / numpy.linalg.norm(exact)


if __name__ == "__main__":
    for sigma in SIGMAS:
        for W in WS:
            best = min((nufft1_error(W, sigma, bw * W), bw) for bw in numpy.linspace(1.6, 2.6, 41))
            print(f"sigma={sigma} W={W} beta/W={best[1]:.3f} l2err={best[0]:.3e}", flush=True)
