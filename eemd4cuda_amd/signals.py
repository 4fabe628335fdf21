"""Synthetic workloads of the EEMD hot path (inputs only; no decomposition).

Each generator is seeded and returns float64 arrays, so every consumer (tests,
bench, examples) decomposes the same signals.

* :func:`three_tone`        -- one float64 signal, N=4096: three sines of 5, 37
  and 211 cycles per record (CPU-runnable reference case).
* :func:`audio_mix`         -- mono 48 kHz record of 2**20 samples: six
  prime-frequency tones (1/k amplitudes), an AM tone, a linear FM chirp and
  pink noise at -30 dB.
* :func:`uneven_batch`      -- 512 signals with log-uniform lengths in
  [2**12, 2**18], three random sinusoids each plus white noise.
* :func:`chirp_am_batch`    -- signals of 3 FM chirps with sinusoidal AM plus
  1/f noise (learned-boundary training data and the at-scale batch).
* :func:`reference_sweep`   -- the reference's 10-octave log sweep
  (refsignal.py:16-24 / emd_par.py:307-318).
"""
from __future__ import annotations

import numpy as np


def three_tone(N: int = 4096) -> np.ndarray:
    t = np.arange(N, dtype=np.float64)
    return (1.0 * np.sin(2 * np.pi * 5 * t / N) + 0.5 * np.sin(2 * np.pi * 37 * t / N)
            + 0.2 * np.sin(2 * np.pi * 211 * t / N))


def standard_noise(E: int, N: int, seed: int = 0) -> np.ndarray:
    """The reference's noise matrix: default_rng(seed).standard_normal((E, N))
    (emd_ref.py:167-169); scale by eps*std(x) before adding."""
    return np.random.default_rng(seed).standard_normal((E, N))


def _pink(n: int, rng: np.random.Generator, alpha: float = 1.0) -> np.ndarray:
    """Noise with a 1/f**alpha power spectrum (spectral shaping of white
    Gaussian noise), normalised to unit RMS."""
    w = rng.standard_normal(n)
    W = np.fft.rfft(w)
    f = np.fft.rfftfreq(n)
    f[0] = f[1] if n > 1 else 1.0
    W = W / f ** (alpha / 2.0)
    x = np.fft.irfft(W, n)
    x = x - x.mean()
    rms = np.sqrt(np.mean(x * x))
    return x / rms if rms > 0 else x


def audio_mix(N: int = 1 << 20, fs: float = 48000.0, seed: int = 1) -> np.ndarray:
    """Six prime tones (101, 211, 431, 877, 1741, 3499 Hz; amplitude 1/k), an
    AM tone (1 kHz carrier, 3 Hz modulation, depth 0.5), a linear chirp
    200 -> 8000 Hz over the record, plus pink noise whose RMS is 30 dB below
    the RMS of the deterministic part."""
    t = np.arange(N, dtype=np.float64) / fs
    x = np.zeros(N)
    for k, f in enumerate((101.0, 211.0, 431.0, 877.0, 1741.0, 3499.0), start=1):
        x += np.sin(2 * np.pi * f * t) / k
    x += (1.0 + 0.5 * np.sin(2 * np.pi * 3.0 * t)) * np.sin(2 * np.pi * 1000.0 * t)
    T = N / fs
    f0, f1 = 200.0, 8000.0
    x += np.sin(2 * np.pi * (f0 * t + 0.5 * (f1 - f0) / T * t * t))
    rms = np.sqrt(np.mean(x * x))
    rng = np.random.default_rng(seed)
    x += _pink(N, rng) * rms * 10 ** (-30.0 / 20.0)
    return x


def uneven_batch(n_signals: int = 512, lo_log2: int = 12, hi_log2: int = 18, seed: int = 2,
                 noise_std: float = 0.05) -> list:
    """Lengths L = round(2**(lo + (hi-lo)*u)); per signal three sinusoids with
    f ~ U(0.001, 0.2) cycles/sample, A ~ U(0.1, 1), phase ~ U(0, 2 pi), plus
    white noise N(0, noise_std**2)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_signals):
        L = int(round(2.0 ** (lo_log2 + (hi_log2 - lo_log2) * rng.uniform())))
        t = np.arange(L, dtype=np.float64)
        f = rng.uniform(0.001, 0.2, 3)
        A = rng.uniform(0.1, 1.0, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        x = (A[:, None] * np.sin(2 * np.pi * f[:, None] * t[None, :] + ph[:, None])).sum(0)
        x += noise_std * rng.standard_normal(L)
        out.append(x)
    return out


def chirp_am(N: int, rng: np.random.Generator, noise_db: float = -25.0) -> np.ndarray:
    """Three linear FM chirps (start/end frequency ~ U(0.002, 0.1) cycles per
    sample), each with sinusoidal AM (depth ~ U(0, 0.6), rate ~ U(1, 8)
    cycles per record), plus 1/f noise whose RMS is noise_db below the RMS of
    the chirp mixture."""
    t = np.arange(N, dtype=np.float64)
    x = np.zeros(N)
    for _ in range(3):
        fa, fb = rng.uniform(0.002, 0.1, 2)
        ph = rng.uniform(0, 2 * np.pi)
        depth = rng.uniform(0.0, 0.6)
        rate = rng.uniform(1.0, 8.0) / N
        amp = rng.uniform(0.3, 1.0)
        phase = 2 * np.pi * (fa * t + 0.5 * (fb - fa) / N * t * t) + ph
        env = 1.0 + depth * np.sin(2 * np.pi * rate * t + rng.uniform(0, 2 * np.pi))
        x += amp * env * np.sin(phase)
    rms = np.sqrt(np.mean(x * x))
    x += _pink(N, rng) * rms * 10 ** (noise_db / 20.0)
    return x


def chirp_am_batch(n_signals: int, N: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.stack([chirp_am(N, rng) for _ in range(n_signals)])


def reference_sweep(FS: int = 44100, T: float = 10.0, F1: float = 20.0, F2: float = 20000.0,
                    noise_level: float = 0.01, seed: int = 0) -> np.ndarray:
    """10-octave log sine sweep plus white noise (refsignal.py:16-24)."""
    N = int(FS * T)
    t = np.arange(N) / FS
    ratio = F2 / F1
    phi = 2 * np.pi * F1 * T * (ratio ** (t / T) - 1.0) / np.log(ratio)
    rng = np.random.default_rng(seed)
    return np.sin(phi) + noise_level * rng.standard_normal(N)


def chirp2048(seed: int = 1) -> np.ndarray:
    """The 2048-sample chirp + tone + noise of the reference's per-sift test
    (test_emd_gpu.py:89-96)."""
    N = 2048
    t = np.arange(N)
    rng = np.random.default_rng(seed)
    f = 0.01 + 0.05 * t / N
    return (np.sin(2 * np.pi * np.cumsum(f)) + 0.3 * np.sin(2 * np.pi * 0.013 * t)
            + 0.1 * rng.standard_normal(N))
