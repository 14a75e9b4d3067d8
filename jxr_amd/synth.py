"""Seeded synthetic images for the benchmark workloads.

  gradient_noise   smooth per-channel gradients plus N(0, sigma^2) noise,
                   clamped to the sample range (512x512 RGB u8 and the
                   mixed batch)
  pink_noise       1/f amplitude spectrum over a seeded complex Gaussian,
                   min/max scaled to 0..65535 (4096x4096 gray u16)
  hdr_rgba         lognormal(0, 1.5) RGB with alpha uniform in [0, 1] (f32)
"""
import numpy as np


def gradient_noise(width, height, seed, channels=3, depth=8, sigma=8.0):
    rng = np.random.default_rng(seed)
    top = (1 << depth) - 1
    scale = top / 255.0
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    planes = []
    for c in range(channels):
        a, b = rng.uniform(-1.0, 1.0, 2)
        base = rng.uniform(0.2, 0.8) * top
        grad = base + (a * x / max(width - 1, 1) + b * y / max(height - 1, 1)) * 0.4 * top
        planes.append(grad + rng.normal(0.0, sigma * scale, (height, width)))
    img = np.stack(planes, axis=-1)
    dtype = np.uint8 if depth <= 8 else np.uint16
    return np.clip(np.rint(img), 0, top).astype(dtype)


def pink_noise(width, height, seed):
    rng = np.random.default_rng(seed)
    spec = rng.normal(size=(height, width // 2 + 1)) + 1j * rng.normal(size=(height, width // 2 + 1))
    fy = np.fft.fftfreq(height)[:, None]
    fx = np.fft.rfftfreq(width)[None, :]
    f = np.sqrt(fx * fx + fy * fy)
    f[0, 0] = 1.0
    field = np.fft.irfft2(spec / f, s=(height, width))
    lo, hi = field.min(), field.max()
    return np.rint((field - lo) / (hi - lo) * 65535.0).astype(np.uint16)


def hdr_rgba(width, height, seed):
    rng = np.random.default_rng(seed)
    rgb = rng.lognormal(0.0, 1.5, (height, width, 3)).astype(np.float32)
    a = rng.uniform(0.0, 1.0, (height, width, 1)).astype(np.float32)
    return np.concatenate([rgb, a], axis=-1)
