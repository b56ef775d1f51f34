"""Seeded synthetic MRI scans (phantom, coil sensitivities, noise, trajectories).

All generation is FP64 with a counter-based SplitMix64 stream, so a scan is a
pure function of its seed.  Seeds follow seed = 0x4B53504A ^ (config << 32) ^ item.
"""
from __future__ import annotations

import numpy as np

_GOLDEN_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def scan_seed(config: int, item: int = 0) -> int:
    return (0x4B53504A ^ (config << 32) ^ item) & 0xFFFFFFFFFFFFFFFF


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Element i is SplitMix64 output number offset+i+1 of the stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN_GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """U[0, 1) doubles from the top 53 bits."""
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def gaussian(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Standard normal doubles by Box-Muller over consecutive uniform pairs."""
    n2 = (count + 1) // 2
    u = uniform(seed, 2 * n2, 2 * offset)
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:count]


# Modified Shepp-Logan (Toft) ellipses: intensity, semi-axes a, b, centre x, y, angle (deg).
_ELLIPSES = np.array([
    [1.0, 0.6900, 0.9200, 0.00, 0.0000, 0.0],
    [-0.8, 0.6624, 0.8740, 0.00, -0.0184, 0.0],
    [-0.2, 0.1100, 0.3100, 0.22, 0.0000, -18.0],
    [-0.2, 0.1600, 0.4100, -0.22, 0.0000, 18.0],
    [0.1, 0.2100, 0.2500, 0.00, 0.3500, 0.0],
    [0.1, 0.0460, 0.0460, 0.00, 0.1000, 0.0],
    [0.1, 0.0460, 0.0460, 0.00, -0.1000, 0.0],
    [0.1, 0.0460, 0.0230, -0.08, -0.6050, 0.0],
    [0.1, 0.0230, 0.0230, 0.00, -0.6060, 0.0],
    [0.1, 0.0230, 0.0460, 0.06, -0.6050, 0.0],
])


def phantom2(n: int, seed: int) -> np.ndarray:
    """n x n float64 10-ellipse phantom (values in [0, 1]); each ellipse centre is
    offset by U(-2, 2) pixels per axis."""
    off = (uniform(seed, 20, offset=1000) * 4.0 - 2.0).reshape(10, 2)
    half = n / 2.0
    c = (np.arange(n) - (n - 1) / 2.0) / half
    x = c[None, :]
    y = -c[:, None]
    img = np.zeros((n, n))
    for k, (amp, a, b, x0, y0, ang) in enumerate(_ELLIPSES):
        x0 = x0 + off[k, 0] / half
        y0 = y0 + off[k, 1] / half
        t = np.deg2rad(ang)
        xr = (x - x0) * np.cos(t) + (y - y0) * np.sin(t)
        yr = -(x - x0) * np.sin(t) + (y - y0) * np.cos(t)
        img += amp * ((xr / a) ** 2 + (yr / b) ** 2 <= 1.0)
    return np.clip(img, 0.0, 1.0)


def coil_maps2(n: int, coils: int, seed: int, sigma: float | None = None, radius: float | None = None) -> np.ndarray:
    """[coils, n, n] complex128 Gaussian sensitivities (sigma = 90 px, ring radius 140 px
    at n = 256, both scaled with n) with a seeded random phase per coil."""
    sigma = 90.0 * n / 256.0 if sigma is None else sigma
    radius = 140.0 * n / 256.0 if radius is None else radius
    phase = uniform(seed, coils, offset=2000) * 2.0 * np.pi
    rr, cc = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    centre = (n - 1) / 2.0
    maps = np.empty((coils, n, n), np.complex128)
    for c in range(coils):
        th = 2.0 * np.pi * c / coils
        r0, c0 = centre + radius * np.sin(th), centre + radius * np.cos(th)
        g = np.exp(-((rr - r0) ** 2 + (cc - c0) ** 2) / (2.0 * sigma ** 2))
        maps[c] = g * np.exp(1j * phase[c])
    return maps


def coil_images2(n: int, coils: int, seed: int) -> np.ndarray:
    return phantom2(n, seed)[None] * coil_maps2(n, coils, seed)


def add_noise(k: np.ndarray, seed: int, rel: float = 1e-3) -> np.ndarray:
    """Complex Gaussian noise with per-component sigma = rel * max|k|."""
    s = rel * float(np.abs(k).max())
    g = gaussian(seed, 2 * k.size, offset=1 << 20).reshape(2, *k.shape)
    return k + s * (g[0] + 1j * g[1])


def cartesian_kspace2(n: int, coils: int, seed: int, centred: bool = True, noise: bool = True) -> np.ndarray:
    """[coils, n, n] complex64 k-space: FP64 FFT of phantom x coils (+ noise).  centred
    k-space is fftshift(fft2(ifftshift(x))), so the centred inverse recovers x."""
    x = coil_images2(n, coils, seed)
    if centred:
        k = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x, axes=(-2, -1))), axes=(-2, -1))
    else:
        k = np.fft.fft2(x)
    return (add_noise(k, seed) if noise else k).astype(np.complex64)


def golden_angle2(spokes: int, readout: int) -> np.ndarray:
    """[spokes*readout, 2] float64 (ky, kx) radians/pixel; spoke p at p*pi/phi,
    k_j = (j - M/2)/M cycles/pixel."""
    golden = np.pi * (np.sqrt(5.0) - 1.0) / 2.0
    ang = np.arange(spokes, dtype=np.float64) * golden
    k = 2.0 * np.pi * (np.arange(readout, dtype=np.float64) - readout // 2) / readout
    ky = (k[None, :] * np.sin(ang)[:, None]).reshape(-1)
    kx = (k[None, :] * np.cos(ang)[:, None]).reshape(-1)
    return np.stack([ky, kx], axis=1)


UNEVEN_MATRICES = (128, 192, 256, 320)


def uneven_items2(count: int, config: int = 3) -> list[tuple[int, int, int]]:
    """Shapes of an uneven radial batch: item i draws, from seed scan_seed(config, i),
    a matrix from {128, 192, 256, 320}, coils uniform in 8..32 and golden-angle spokes
    uniform in 101..603 (readout = 2 x matrix).  Returns [(matrix, coils, spokes)]."""
    out = []
    for i in range(count):
        u = uniform(scan_seed(config, i), 3, offset=1 << 24)
        out.append((UNEVEN_MATRICES[int(u[0] * 4)], 8 + int(u[1] * 25), 101 + int(u[2] * 503)))
    return out


# 3-D ellipsoid phantom: intensity, semi-axes a (x), b (y), c (z), centre x, y, z, angle about
# z (deg); the 2-D table above extended along z (outer shells, ventricles, small features).
_ELLIPSOIDS = np.array([
    [1.0, 0.6900, 0.9200, 0.8100, 0.00, 0.0000, 0.00, 0.0],
    [-0.8, 0.6624, 0.8740, 0.7800, 0.00, -0.0184, 0.00, 0.0],
    [-0.2, 0.1100, 0.3100, 0.2200, 0.22, 0.0000, 0.00, -18.0],
    [-0.2, 0.1600, 0.4100, 0.2800, -0.22, 0.0000, 0.00, 18.0],
    [0.1, 0.2100, 0.2500, 0.4100, 0.00, 0.3500, -0.15, 0.0],
    [0.1, 0.0460, 0.0460, 0.0500, 0.00, 0.1000, 0.25, 0.0],
    [0.1, 0.0460, 0.0460, 0.0500, 0.00, -0.1000, 0.25, 0.0],
    [0.1, 0.0460, 0.0230, 0.0500, -0.08, -0.6050, 0.00, 0.0],
    [0.1, 0.0230, 0.0230, 0.0200, 0.00, -0.6060, 0.00, 0.0],
    [0.1, 0.0230, 0.0460, 0.0200, 0.06, -0.6050, 0.00, 0.0],
])


def phantom3(n: int, seed: int, torch, device="cuda"):
    """[n, n, n] (z, y, x) float64 torch tensor: the 10-ellipsoid phantom (values clipped to
    [0, 1]), each centre offset by U(-2, 2) voxels per axis."""
    off = (uniform(seed, 30, offset=3000) * 4.0 - 2.0).reshape(10, 3)
    half = n / 2.0
    c = (torch.arange(n, dtype=torch.float64, device=device) - (n - 1) / 2.0) / half
    z = c.view(n, 1, 1)
    y = -c.view(1, n, 1)
    x = c.view(1, 1, n)
    img = torch.zeros((n, n, n), dtype=torch.float64, device=device)
    for k, (amp, a, b, cc, x0, y0, z0, ang) in enumerate(_ELLIPSOIDS):
        x0 = x0 + off[k, 0] / half
        y0 = y0 + off[k, 1] / half
        z0 = z0 + off[k, 2] / half
        t = np.deg2rad(ang)
        xr = (x - x0) * np.cos(t) + (y - y0) * np.sin(t)
        yr = -(x - x0) * np.sin(t) + (y - y0) * np.cos(t)
        img += amp * (((xr / a) ** 2 + (yr / b) ** 2 + ((z - z0) / cc) ** 2) <= 1.0).to(torch.float64)
    return img.clamp_(0.0, 1.0)


def coil_map3(n: int, coil: int, coils: int, seed: int, torch, device="cuda", sigma: float | None = None,
              radius: float | None = None):
    """[n, n, n] complex128 torch tensor: coil `coil` of `coils` 3-D Gaussian sensitivities
    (sigma = 90 voxels, centres on a sphere of radius 140 voxels at n = 256, both scaled with
    n; centres on a Fibonacci sphere) with a seeded random phase."""
    sigma = 90.0 * n / 256.0 if sigma is None else sigma
    radius = 140.0 * n / 256.0 if radius is None else radius
    phase = float(uniform(seed, coils, offset=4000)[coil] * 2.0 * np.pi)
    golden = np.pi * (3.0 - np.sqrt(5.0))
    cz = 1.0 - 2.0 * (coil + 0.5) / coils
    r = np.sqrt(max(0.0, 1.0 - cz * cz))
    cy, cx = r * np.sin(golden * coil), r * np.cos(golden * coil)
    centre = (n - 1) / 2.0
    g = torch.arange(n, dtype=torch.float64, device=device)
    dz = (g - (centre + radius * cz)).view(n, 1, 1) ** 2
    dy = (g - (centre + radius * cy)).view(1, n, 1) ** 2
    dx = (g - (centre + radius * cx)).view(1, 1, n) ** 2
    mag = torch.exp(-(dz + dy + dx) / (2.0 * sigma ** 2))
    return torch.polar(mag, torch.full_like(mag, phase))


def kooshball_kspace3(ops, torch, n: int, coils: int, traj, seed: int, width: int = 6, noise: float = 1e-3):
    """[coils, S] complex64 k-space of the 3-D phantom times each coil's sensitivity,
    simulated by the FP64 Kaiser-Bessel forward NUFFT of `width` on the 2x grid, plus complex
    Gaussian noise of per-component sigma = noise x max|k| (torch's generator, seeded)."""
    obj = phantom3(n, seed, torch)
    out = torch.empty((coils, traj.shape[0]), dtype=torch.complex128, device=traj.device)
    for c in range(coils):
        img = obj * coil_map3(n, c, coils, seed, torch)
        out[c] = ops.nufft3_forward_f64(img.contiguous(), traj, width=width)
        del img
    s = noise * float(out.abs().max())
    g = torch.Generator(device=traj.device).manual_seed(seed & 0x7FFFFFFFFFFFFFFF)
    out += torch.complex(torch.randn(out.shape, dtype=torch.float64, device=traj.device, generator=g),
                         torch.randn(out.shape, dtype=torch.float64, device=traj.device, generator=g)) * s
    return out.to(torch.complex64)
