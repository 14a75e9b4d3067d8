"""Seeded synthetic images for the benchmark workloads.

  gradient_noise   smooth per-channel gradients plus N(0, sigma^2) noise,
                   clamped to the sample range (512x512 RGB u8 and the
                   mixed batch)
  gradient_noise_large  the same for RGB u8 sources of tens of megapixels,
                   generated in strips on a thread pool (the 8192x8192
                   region-decode source)
  pink_noise       1/f amplitude spectrum over a seeded complex Gaussian,
                   min/max scaled to 0..65535 (4096x4096 gray u16)
  hdr_rgba         lognormal(0, 1.5) RGB with alpha uniform in [0, 1] (f32)
"""
import os

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


def gradient_noise_large(width, height, seed, sigma=8.0, strip=512):
    """RGB u8 gradient plus Gaussian noise like gradient_noise, generated in
    float32 row strips on a thread pool (one random stream per strip), for
    sources of tens of megapixels."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    coef = [(rng.uniform(-1.0, 1.0, 2), rng.uniform(0.2, 0.8) * 255.0) for _ in range(3)]
    out = np.empty((height, width, 3), np.uint8)
    xs = (np.arange(width, dtype=np.float32) / max(width - 1, 1))[None, :]

    def fill(k):
        y0, y1 = k * strip, min((k + 1) * strip, height)
        r = np.random.default_rng([seed, k])
        ys = (np.arange(y0, y1, dtype=np.float32) / max(height - 1, 1))[:, None]
        for c, ((a, b), base) in enumerate(coef):
            v = base + (np.float32(a) * xs + np.float32(b) * ys) * np.float32(0.4 * 255.0)
            v = v + r.standard_normal((y1 - y0, width), dtype=np.float32) * np.float32(sigma)
            out[y0:y1, :, c] = np.clip(np.rint(v), 0, 255).astype(np.uint8)

    with ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as ex:
        list(ex.map(fill, range((height + strip - 1) // strip)))
    return out


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
