"""Seeded synthetic inputs for the five benchmark/parity configurations.

Every generator runs on any torch device and produces the same bits on the
CPU and on the GPU: uniforms come from an integer hash evaluated in int64
without overflow, and the Gaussian blobs are products of 1-D f64 profiles
(computed once on the host) rounded to f32, so golden hashes computed on one
machine hold on another.

  c1: Image<f32,2> 512x512, U[0,1) from seed 0, sigma 1.5
  c2: Image<f32,3> 512^3, U[0,1) + 32 Gaussian blobs (amplitude U[1,5],
      blob sigma U[4,40] voxels, centre uniform in the volume) from seed 1,
      sigma 2.0; the blob value is A*exp(-0.5*sum(((x_d-c_d)/s)^2)) as in
      GaussianImageSource (src/filters/sources.rs:78-87), evaluated inside
      +-6 s of the centre (zero beyond)
  c3: Image<VecPixel<f32,3>,3> 256x256x160, spacing 0.7/0.7/2.5, sigma 3.0;
      each component i.i.d. N(0,1) from seed 2, smoothed at sigma 8 voxels
      (periodic FFT smoothing), scaled so max|v| = 5
  c4: 64 whole f32 volumes from seed 3: extents U{64..320} per axis, sigma
      U[0.5,4], one isotropic spacing U[0.5,2] per volume, voxels U[0,1)
  c5: Image<f32,3> 256^3, sigma 1.8, x ~ U[0,1), upstream gradient ~ N(0,1),
      seed 4
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np
import torch

_MASK32 = 0xFFFFFFFF


def _mul32(x: torch.Tensor, c: int) -> torch.Tensor:
    """(x * c) mod 2^32 for 0 <= x < 2^32 without int64 overflow."""
    lo, hi = c & 0xFFFF, c >> 16
    return (x * lo + (((x * hi) & 0xFFFF) << 16)) & _MASK32


def _mix32(x: torch.Tensor) -> torch.Tensor:
    """32-bit avalanche hash (xor-shift / multiply rounds)."""
    x = x ^ (x >> 16)
    x = _mul32(x, 0x7FEB352D)
    x = x ^ (x >> 15)
    x = _mul32(x, 0x846CA68B)
    x = x ^ (x >> 16)
    return x


def _key(seed: int, stream: int) -> int:
    t = torch.tensor([(seed * 0x9E3779B9 + stream * 0x85EBCA6B + 0x27D4EB2F) & _MASK32],
                     dtype=torch.int64)
    return int(_mix32(t)[0])


def hash_u32(seed: int, stream: int, start: int, n: int, device) -> torch.Tensor:
    key = _key(seed, stream)
    i = torch.arange(start, start + n, dtype=torch.int64, device=device)
    return _mix32(_mix32((i ^ key) & _MASK32) ^ key)


def uniform(seed: int, stream: int, shape, device, chunk: int = 1 << 24) -> torch.Tensor:
    """U[0,1) float32: top 24 hash bits times 2^-24 (exact)."""
    n = int(np.prod(shape))
    out = torch.empty(n, dtype=torch.float32, device=device)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        h = hash_u32(seed, stream, s, m, device)
        out[s:s + m] = (h >> 8).to(torch.float32) * (1.0 / (1 << 24))
    return out.view(*shape)


def normal(seed: int, stream: int, shape, device) -> torch.Tensor:
    """N(0,1) float32 by Box-Muller from two hash streams (f64 math)."""
    n = int(np.prod(shape))
    u1 = (hash_u32(seed, stream, 0, n, device).to(torch.float64) + 1.0) / 4294967296.0
    u2 = hash_u32(seed, stream + 1, 0, n, device).to(torch.float64) / 4294967296.0
    z = torch.sqrt(-2.0 * torch.log(u1)) * torch.cos(2.0 * math.pi * u2)
    return z.to(torch.float32).view(*shape)


def _host_uniform(seed: int, stream: int, n: int) -> np.ndarray:
    return uniform(seed, stream, (n,), "cpu").numpy().astype(np.float64)


# ---------------------------------------------------------------------------

def c1(device) -> Tuple[torch.Tensor, float, Tuple[float, ...]]:
    return uniform(0, 0, (512, 512), device), 1.5, (1.0, 1.0)


def add_blobs(vol: torch.Tensor, seed: int, nblobs: int = 32) -> None:
    """Adds the blob field to a (z, y, x) float32 volume in place."""
    nz, ny, nx = vol.shape
    p = _host_uniform(seed, 100, nblobs * 5).reshape(nblobs, 5)
    for b in range(nblobs):
        amp = 1.0 + 4.0 * p[b, 0]
        s = 4.0 + 36.0 * p[b, 1]
        cx, cy, cz = p[b, 2] * nx, p[b, 3] * ny, p[b, 4] * nz
        box = []
        prof = []
        for c, n in ((cz, nz), (cy, ny), (cx, nx)):
            lo = max(0, int(math.floor(c - 6.0 * s)))
            hi = min(n, int(math.ceil(c + 6.0 * s)) + 1)
            x = np.arange(lo, hi, dtype=np.float64)
            d = (x - c) / s
            prof.append(np.exp(-0.5 * d * d))
            box.append((lo, hi))
        gz, gy, gx = (torch.from_numpy(q).to(vol.device) for q in prof)
        field = (amp * gz[:, None, None]) * (gy[None, :, None] * gx[None, None, :])
        (z0, z1), (y0, y1), (x0, x1) = box
        vol[z0:z1, y0:y1, x0:x1] += field.to(torch.float32)


def c2(device, shape=(512, 512, 512)):
    vol = uniform(1, 0, shape, device)
    add_blobs(vol, 1)
    return vol, 2.0, (1.0, 1.0, 1.0)


def c3(device, shape=(160, 256, 256)):
    """Returns a (z, y, x, 3) float32 displacement field."""
    comps = []
    for c in range(3):
        z = normal(2, 10 * c, shape, "cpu").numpy().astype(np.float64)
        f = np.fft.rfftn(z)
        k = [np.fft.fftfreq(n) for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1])]
        s = 8.0
        tr = 1.0
        for ax, kk in enumerate(k):
            g = np.exp(-2.0 * (math.pi * s * kk) ** 2)
            sh = [1] * 3
            sh[ax] = len(kk)
            tr = tr * g.reshape(sh)
        sm = np.fft.irfftn(f * tr, s=shape, axes=(0, 1, 2))
        sm *= 5.0 / np.abs(sm).max()
        comps.append(sm.astype(np.float32))
    v = torch.from_numpy(np.stack(comps, axis=-1)).to(device)
    return v, 3.0, (0.7, 0.7, 2.5)


def c4_params(nvol: int = 64, lo: int = 64, hi: int = 320, seed: int = 3):
    """[(shape_zyx, sigma, spacing)] for the batched configuration."""
    p = _host_uniform(seed, 200, nvol * 5).reshape(nvol, 5)
    out = []
    for v in range(nvol):
        ext = [lo + min(int(p[v, d] * (hi - lo + 1)), hi - lo) for d in range(3)]
        sigma = 0.5 + 3.5 * p[v, 3]
        sp = 0.5 + 1.5 * p[v, 4]
        out.append(((ext[2], ext[1], ext[0]), float(sigma), (float(sp),) * 3))
    return out


def c4(device, nvol: int = 64, lo: int = 64, hi: int = 320, seed: int = 3):
    vols = []
    params = c4_params(nvol, lo, hi, seed)
    for v, (shape, sigma, sp) in enumerate(params):
        vols.append(uniform(seed, 1000 + v, shape, device))
    return vols, params


def c5(device, shape=(256, 256, 256)):
    x = uniform(4, 0, shape, device)
    g = normal(4, 1, shape, device)
    return x, g, 1.8, (1.0, 1.0, 1.0)


def c5_batch(device, nvol: int = 16, lo: int = 48, hi: int = 192, seed: int = 4):
    """Batched adjoint: nvol upstream gradients N(0,1) with extents drawn per
    axis uniformly in [lo, hi], sigma 1.8, spacing 1.  Returns (grads, params)
    with params[v] = (shape_zyx, sigma, spacing)."""
    p = _host_uniform(seed, 300, nvol * 3).reshape(nvol, 3)
    grads, params = [], []
    for v in range(nvol):
        ext = [lo + min(int(p[v, d] * (hi - lo + 1)), hi - lo) for d in range(3)]
        shape = (ext[2], ext[1], ext[0])
        grads.append(normal(seed, 2000 + v, shape, device))
        params.append((shape, 1.8, (1.0, 1.0, 1.0)))
    return grads, params
