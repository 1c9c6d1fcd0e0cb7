"""Seeded synthetic images for the benchmark and parity configurations.

Each image: uint16 I = clip(Poisson(lam), 0, 4095) with
    lam = 500 + sum_3 200*sin(2*pi*(x cos t + y sin t)/T + phi) + sum_5 A*exp(-r^2 / (2 s^2)),
t ~ U[0, pi), phi ~ U[0, 2 pi), A ~ U[300, 1500], s ~ U[4, 10] px, blob centres
uniform over the image; the grating period T is the class's.  The generator
is numpy PCG64 seeded with (config seed + global image index).
"""
import hashlib

import numpy as np

CONFIGS = {
    # name: (seed base, periods per class, images per class, shape or None for random)
    'short_128': (1001, (8, 20), 10, (128, 128)),
    'long_256': (2002, (6, 12, 24, 48), 100, (256, 256)),
    'tiles_mixed': (4004, (5, 7, 10, 14, 20, 28, 40, 56), 1250, None),
    # differentiation: 8 classes x 500 training images of 512 x 512, scaled to [0, 1] as float32
    'grad_512': (5005, (4, 6, 8, 12, 16, 24, 32, 48), 500, (512, 512)),
}


def image(seed, period, shape):
    rng = np.random.Generator(np.random.PCG64(seed))
    h, w = shape
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    lam = np.full((h, w), 500.0)
    for _ in range(3):
        t = rng.uniform(0, np.pi)
        phi = rng.uniform(0, 2 * np.pi)
        lam += 200.0 * np.sin(2 * np.pi * (xx * np.cos(t) + yy * np.sin(t)) / period + phi)
    for _ in range(5):
        a = rng.uniform(300, 1500)
        s = rng.uniform(4, 10)
        cy, cx = rng.uniform(0, h), rng.uniform(0, w)
        lam += a * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
    img = rng.poisson(np.clip(lam, 0, None))
    return np.clip(img, 0, 4095).astype(np.uint16)


def config_images(name, limit=None):
    """(images, classes 1..C) of a configuration, in global-index order."""
    base, periods, per_class, shape = CONFIGS[name]
    imgs, cls = [], []
    k = 0
    for c, T in enumerate(periods):
        for _ in range(per_class):
            if limit is not None and k >= limit:
                return imgs, np.array(cls, np.int32)
            if shape is None:
                r = np.random.Generator(np.random.PCG64(base * 1000003 + k))
                shp = (int(r.integers(64, 1025)), int(r.integers(64, 1025)))
            else:
                shp = shape
            imgs.append(image(base + k, T, shp))
            cls.append(c + 1)
            k += 1
    return imgs, np.array(cls, np.int32)


def digest(imgs):
    h = hashlib.sha256()
    for im in imgs:
        h.update(np.ascontiguousarray(im).tobytes())
    return h.hexdigest()
