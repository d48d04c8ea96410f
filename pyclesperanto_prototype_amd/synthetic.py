"""Seeded synthetic microscopy images for benchmarks and tests.

blobs_2d     2D uint8: background 10, isotropic Gaussian spots, Gaussian noise,
             clipped to [0, 255] (the 512 x 512 / 60-spot plain-method image).
nuclei_3d    3D uint16 nuclei stack: non-overlapping randomly oriented ellipsoids of
             uniform intensity on a constant background, then Poisson shot noise and
             Gaussian read noise, rounded and clipped to uint16.  Rendered on the
             device for the large volumes.
mixed_batch  many uint16 nuclei images of mixed sizes (70 % 2D log-uniform 64^2 ..
             1024^2, 30 % 3D log-uniform 32x128x128 .. 96x512x512), densities uniform
             in 1e-4 .. 1e-3 nuclei per voxel (overlap allowed), per-image spot sigma
             uniform in 1 .. 5 and outline sigma in 0.5 .. 2.
"""
from __future__ import annotations

import math

import numpy as np


def blobs_2d(shape=(512, 512), n_spots=60, seed=1, sigma=(4.0, 8.0), amplitude=(100.0, 200.0),
             background=10.0, noise=5.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    H, W = shape
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    img = np.full(shape, background)
    for _ in range(n_spots):
        cy, cx = rng.uniform(0, H), rng.uniform(0, W)
        s = rng.uniform(*sigma)
        a = rng.uniform(*amplitude)
        img += a * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
    img += rng.normal(0, noise, shape)
    return np.clip(img, 0, 255).astype(np.uint8)


def _random_rotations(rng, n):
    """Uniform random rotation matrices from unit quaternions (Shoemake)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def place_nuclei(shape, n, semi=(4.0, 10.0), seed=2, touching_fraction=0.0, max_tries=50):
    """Non-overlapping placement by bounding spheres on a hash grid.

    Returns centres (n, 3) in (z, y, x), semi-axes (n, 3), rotations (n, 3, 3).  With
    touching_fraction > 0 that fraction of the nuclei is placed as pairs whose
    bounding spheres touch (the second nucleus of a pair sits at distance r1 + r2).
    """
    rng = np.random.default_rng(seed)
    D, H, W = shape
    rmax = semi[1]
    cell = 2 * rmax
    grid = {}
    centres, radii, axes = [], [], []

    def free(c, r):
        k = tuple((c // cell).astype(int))
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    for j in grid.get((k[0] + dz, k[1] + dy, k[2] + dx), ()):
                        if np.sum((centres[j] - c) ** 2) < (radii[j] + r) ** 2 - 1e-9:
                            return False
        return True

    def add(c, r, a):
        k = tuple((c // cell).astype(int))
        grid.setdefault(k, []).append(len(centres))
        centres.append(c)
        radii.append(r)
        axes.append(a)

    n_pairs = int(round(n * touching_fraction / 2))
    target = n
    tries = 0
    while len(centres) < target and tries < target * max_tries:
        tries += 1
        a = rng.uniform(semi[0], semi[1], 3)
        r = a.max()
        c = np.array([rng.uniform(r, D - r) if D > 2 * r else D / 2, rng.uniform(r, H - r), rng.uniform(r, W - r)])
        if not free(c, r):
            continue
        if n_pairs > 0 and len(centres) < 2 * n_pairs and len(centres) + 1 < target:
            a2 = rng.uniform(semi[0], semi[1], 3)
            r2 = a2.max()
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            c2 = c + d * (r + r2)
            inside = np.all(c2 >= r2) and c2[0] <= D - r2 and c2[1] <= H - r2 and c2[2] <= W - r2
            if not inside or not free(c2, r2):
                continue
            add(c, r, a)
            add(c2, r2, a2)
            continue
        add(c, r, a)
    m = len(centres)
    return np.array(centres).reshape(m, 3), np.array(axes).reshape(m, 3), _random_rotations(rng, m)


def nuclei_3d(shape=(256, 1024, 1024), n_nuclei=5000, semi=(4.0, 10.0), intensity=(800.0, 3000.0),
              background=200.0, read_noise=20.0, seed=2, touching_fraction=0.0, device=None, chunk=1024):
    """uint16 nuclei stack as a torch tensor on `device` (default: the current GPU).

    Returns (tensor, info) where info holds the realised nucleus count.
    """
    import torch

    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    D, H, W = shape
    cen, ax, rot = place_nuclei(shape, n_nuclei, semi, seed, touching_fraction)
    n = len(cen)
    rng = np.random.default_rng(seed + 1000)
    val = rng.uniform(*intensity, n)
    clean = torch.full(shape, float(background), dtype=torch.float32, device=device)
    R = int(math.ceil(semi[1]))
    o = torch.arange(-R, R + 1, device=device, dtype=torch.float32)
    oz, oy, ox = torch.meshgrid(o, o, o, indexing="ij")
    offs = torch.stack([oz.reshape(-1), oy.reshape(-1), ox.reshape(-1)], 1)  # (P, 3)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        c = torch.tensor(cen[s:e], dtype=torch.float32, device=device)
        a = torch.tensor(ax[s:e], dtype=torch.float32, device=device)
        rt = torch.tensor(rot[s:e], dtype=torch.float32, device=device)
        v = torch.tensor(val[s:e], dtype=torch.float32, device=device)
        ci = torch.round(c)
        pos = ci[:, None, :] + offs[None]                   # (m, P, 3) voxel coordinates
        rel = pos - c[:, None, :]
        loc = torch.einsum("mij,mpj->mpi", rt.transpose(1, 2).contiguous(), rel)
        inside = ((loc / a[:, None, :]) ** 2).sum(-1) <= 1.0
        inb = (pos[..., 0] >= 0) & (pos[..., 0] < D) & (pos[..., 1] >= 0) & (pos[..., 1] < H) & \
              (pos[..., 2] >= 0) & (pos[..., 2] < W)
        m = inside & inb
        p = pos[m].long()
        lin = (p[:, 0] * H + p[:, 1]) * W + p[:, 2]
        clean.view(-1)[lin] = v[:, None].expand(-1, offs.shape[0])[m]
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    noisy = torch.poisson(clean, generator=gen)
    noisy += torch.randn(shape, generator=gen, device=device) * read_noise
    # int32 -> int16 keeps the low 16 bits; viewed as uint16 they are the clipped values
    out = torch.clamp(torch.round(noisy), 0, 65535).to(torch.int32).to(torch.int16).view(torch.uint16)
    del clean, noisy
    return out, {"nuclei": n}


def _render_nuclei(shape, centres, axes, rot, values, background, read_noise, seed, device, chunk=1024):
    """Ellipsoids (ellipses for a 2-axis shape) of uniform values on a background, then
    Poisson shot noise and Gaussian read noise, as a uint16 tensor on `device`."""
    import torch

    nd = len(shape)
    clean = torch.full(shape, float(background), dtype=torch.float32, device=device)
    R = int(math.ceil(float(np.max(axes)) if len(axes) else 1.0))
    o = torch.arange(-R, R + 1, device=device, dtype=torch.float32)
    grids = torch.meshgrid(*([o] * nd), indexing="ij")
    offs = torch.stack([g.reshape(-1) for g in grids], 1)  # (P, nd)
    n = len(centres)
    for s0 in range(0, n, chunk):
        e = min(n, s0 + chunk)
        c = torch.tensor(centres[s0:e], dtype=torch.float32, device=device)
        a = torch.tensor(axes[s0:e], dtype=torch.float32, device=device)
        rt = torch.tensor(rot[s0:e], dtype=torch.float32, device=device)
        v = torch.tensor(values[s0:e], dtype=torch.float32, device=device)
        pos = torch.round(c)[:, None, :] + offs[None]
        rel = pos - c[:, None, :]
        loc = torch.matmul(rel, rt.contiguous())  # rel @ R = (R^T rel^T)^T: coordinates in the body frame
        inside = ((loc / a[:, None, :]) ** 2).sum(-1) <= 1.0
        inb = torch.ones_like(inside)
        for d in range(nd):
            inb &= (pos[..., d] >= 0) & (pos[..., d] < shape[d])
        m = inside & inb
        p = pos[m].long()
        lin = p[:, 0]
        for d in range(1, nd):
            lin = lin * shape[d] + p[:, d]
        vals = v[:, None].expand(-1, offs.shape[0])[m]
        # overlaps: the later nucleus wins, decided before the write (a device write
        # with repeated indices would keep an arbitrary one): sort by (voxel, nucleus)
        # and keep the last entry of every voxel
        ids = torch.arange(s0, e, device=device)[:, None].expand(-1, offs.shape[0])[m]
        order = torch.argsort(lin * n + ids)
        lin, vals = lin[order], vals[order]
        last = torch.ones_like(lin, dtype=torch.bool)
        last[:-1] = lin[1:] != lin[:-1]
        clean.view(-1)[lin[last]] = vals[last]
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    noisy = torch.poisson(clean, generator=gen)
    noisy += torch.randn(shape, generator=gen, device=device) * read_noise
    out = torch.clamp(torch.round(noisy), 0, 65535).to(torch.int32).to(torch.int16).view(torch.uint16)
    del clean, noisy
    return out


def _random_rotations_2d(rng, n):
    t = rng.uniform(0, 2 * np.pi, n)
    R = np.empty((n, 2, 2))
    R[:, 0, 0] = np.cos(t); R[:, 0, 1] = -np.sin(t); R[:, 1, 0] = np.sin(t); R[:, 1, 1] = np.cos(t)
    return R


def nuclei_overlapping(shape, density, semi=(4.0, 10.0), intensity=(800.0, 3000.0), background=200.0,
                       read_noise=20.0, seed=4, device=None):
    """uint16 nuclei image (2D or 3D) with `density` nuclei per voxel placed uniformly
    at random, overlaps allowed (later nuclei overwrite earlier ones)."""
    import torch

    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    rng = np.random.default_rng(seed)
    nd = len(shape)
    n = int(round(density * float(np.prod(shape))))
    centres = rng.uniform(0, 1, (n, nd)) * np.array(shape, np.float64)
    axes = rng.uniform(semi[0], semi[1], (n, nd))
    rot = _random_rotations(rng, n) if nd == 3 else _random_rotations_2d(rng, n)
    values = rng.uniform(*intensity, n)
    return _render_nuclei(tuple(shape), centres, axes, rot, values, background, read_noise, seed, device), n


def mixed_batch(n_images=2000, seed=4, device=None, levels=8):
    """The mixed-size batch: returns (images, spot_sigmas, outline_sigmas, info).

    Images are crops of one source image per density level (2D sources 1024 x 1024,
    3D sources 96 x 512 x 512, `levels` densities evenly spaced over 1e-4 .. 1e-3), so
    the whole set is rendered in a few seconds; each image gets its own size, crop
    offset and sigmas.
    """
    import torch

    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    rng = np.random.default_rng(seed)
    dens = np.linspace(1e-4, 1e-3, levels)
    src2 = [nuclei_overlapping((1024, 1024), d, seed=seed * 100 + k, device=device)[0] for k, d in enumerate(dens)]
    src3 = [nuclei_overlapping((96, 512, 512), d, seed=seed * 100 + 50 + k, device=device)[0]
            for k, d in enumerate(dens)]
    n2 = int(round(0.7 * n_images))
    kinds = np.array([2] * n2 + [3] * (n_images - n2))
    rng.shuffle(kinds)
    images, ss, so, nvox = [], [], [], 0
    for kind in kinds:
        d = rng.uniform(1e-4, 1e-3)
        lvl = int(np.argmin(np.abs(dens - d)))
        if kind == 2:
            side = int(round(math.exp(rng.uniform(math.log(64), math.log(1024)))))
            y0, x0 = rng.integers(0, 1024 - side + 1, 2)
            img = src2[lvl][y0:y0 + side, x0:x0 + side].contiguous()
        else:
            t = rng.uniform(0, 1)
            zs, xy = int(round(32 * 3 ** t)), int(round(128 * 4 ** t))
            z0 = rng.integers(0, 96 - zs + 1)
            y0, x0 = rng.integers(0, 512 - xy + 1, 2)
            img = src3[lvl][z0:z0 + zs, y0:y0 + xy, x0:x0 + xy].contiguous()
        images.append(img)
        nvox += img.numel()
        ss.append(float(rng.uniform(1.0, 5.0)))
        so.append(float(rng.uniform(0.5, 2.0)))
    del src2, src3
    return images, np.array(ss, np.float32), np.array(so, np.float32), {
        "images": n_images, "voxels": nvox, "2d": int(n2), "3d": int(n_images - n2), "density_levels": levels}
