"""itk-rs_amd -- MI355X (gfx950) separable Gaussian for the itk-rs pipeline.

Python side of the drop-in.  It mirrors the reference's pipeline protocol
(src/source.rs:13-29) -- ``Region``, ``Image``, ``ImageSource``,
``GaussianFilter``, ``stream_full``, ``split_region`` -- with device-resident
images (torch tensors on the GPU) and calls the C ABI of ``libitkg.so``
(include/itkg.h) for every pixel of work.  torch supplies device memory and
streams only.

There is no CPU fallback: if ``libitkg.so`` is missing, importing this module
raises.  Build it with ``make -C itk-rs_amd`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import threading
from dataclasses import dataclass, field
from typing import Iterator, List, Optional, Sequence, Tuple

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libitkg.so")

MAX_DIM = 4
MAX_RADIUS = 127

OK, ERR_INVALID_ARG, ERR_OOM, ERR_HIP, ERR_UNSUPPORTED, ERR_NO_DEVICE = 0, -1, -2, -3, -4, -5
F32, F64 = 0, 1
EXACT, FAST = 0, 1
PATH_AUTO, PATH_PER_AXIS, PATH_FUSED = 0, 1, 2
BATCH_FORWARD, BATCH_BACKWARD = 0, 1
PT_U8, PT_U16, PT_I16, PT_I32, PT_F32, PT_F64 = 0, 1, 2, 3, 4, 5
CONVERT_AS, CONVERT_ROUND = 0, 1

# Every symbol include/itkg.h declares (checked by the test suite).
EXPORTS = (
    "itkg_kernel_radius", "itkg_build_kernel", "itkg_input_region_for_output",
    "itkg_split_region", "itkg_gaussian_generate_region", "itkg_gaussian_generate_region_noalloc",
    "itkg_gaussian_generate_region_host",
    "itkg_gaussian_backward", "itkg_batch_create", "itkg_batch_run", "itkg_batch_autotune",
    "itkg_batch_voxels",
    "itkg_batch_destroy", "itkg_last_error", "itkg_set_path", "itkg_version",
    "itkg_box_kernel", "itkg_separable_input_region_for_output",
    "itkg_separable_generate_region", "itkg_separable_generate_region_host",
    "itkg_index_gaussian_kernel", "itkg_unsharp_mask_generate_region",
    "itkg_unsharp_mask_generate_region_host", "itkg_yvv_coefficients",
    "itkg_recursive_gaussian_generate_region", "itkg_recursive_gaussian_generate_region_host",
    "itkg_canny_generate_region", "itkg_canny_generate_region_host", "itkg_trim_scratch",
    "itkg_convert_pixels", "itkg_cast_gaussian_cast_host", "itkg_host_staging_info",
    "itkg_gaussian_backward_host", "itkg_canny_direction_bins", "itkg_gaussian_generate_regions",
)


class ItkgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"itkg error {code}: {msg}")
        self.code = code


class c_region(ctypes.Structure):
    _fields_ = [("index", ctypes.c_int64 * MAX_DIM), ("size", ctypes.c_int64 * MAX_DIM)]


class c_view(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("region", c_region),
                ("stride", ctypes.c_int64 * MAX_DIM)]


class c_volume(ctypes.Structure):
    _fields_ = [("input", ctypes.c_void_p), ("output", ctypes.c_void_p),
                ("size", ctypes.c_int64 * 3), ("sigma", ctypes.c_double),
                ("spacing", ctypes.c_double * 3)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C {_HERE}`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    L.itkg_kernel_radius.restype = ctypes.c_uint64
    L.itkg_kernel_radius.argtypes = [ctypes.c_double, ctypes.c_double]
    L.itkg_build_kernel.argtypes = [ctypes.c_double, ctypes.c_uint64, P(ctypes.c_double),
                                    ctypes.c_uint64]
    L.itkg_input_region_for_output.argtypes = [ctypes.c_int32, ctypes.c_double,
                                               P(ctypes.c_double), P(c_region), P(c_region),
                                               P(c_region)]
    L.itkg_split_region.argtypes = [ctypes.c_int32, P(c_region), ctypes.c_uint64, P(c_region),
                                    ctypes.c_uint64, P(ctypes.c_uint64)]
    L.itkg_gaussian_generate_region.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_int32, ctypes.c_double, P(ctypes.c_double),
        P(c_region), P(c_view), P(c_view), ctypes.c_int, ctypes.c_void_p]
    L.itkg_gaussian_generate_region_noalloc.argtypes = L.itkg_gaussian_generate_region.argtypes
    L.itkg_gaussian_generate_regions.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_int32, ctypes.c_double, P(ctypes.c_double),
        P(c_region), P(c_view), ctypes.c_int32, P(c_view), ctypes.c_int, ctypes.c_void_p]
    L.itkg_gaussian_generate_region_host.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_int32, ctypes.c_double, P(ctypes.c_double),
        P(c_region), ctypes.c_void_p, P(c_region), ctypes.c_void_p, ctypes.c_int]
    L.itkg_box_kernel.argtypes = [ctypes.c_uint64, P(ctypes.c_double), ctypes.c_uint64]
    L.itkg_separable_input_region_for_output.argtypes = [
        ctypes.c_int32, P(ctypes.c_uint64), P(c_region), P(c_region), P(c_region)]
    L.itkg_separable_generate_region.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_int32, P(ctypes.c_uint64), P(ctypes.c_double),
        P(c_region), P(c_view), P(c_view), ctypes.c_int, ctypes.c_void_p]
    L.itkg_separable_generate_region_host.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_int32, P(ctypes.c_uint64), P(ctypes.c_double),
        P(c_region), ctypes.c_void_p, P(c_region), ctypes.c_void_p, ctypes.c_int]
    L.itkg_index_gaussian_kernel.argtypes = [ctypes.c_double, P(ctypes.c_double), ctypes.c_uint64,
                                             P(ctypes.c_uint64)]
    L.itkg_unsharp_mask_generate_region.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
        P(c_region), P(c_view), P(c_view), ctypes.c_int, ctypes.c_void_p]
    L.itkg_unsharp_mask_generate_region_host.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
        P(c_region), ctypes.c_void_p, P(c_region), ctypes.c_void_p, ctypes.c_int]
    L.itkg_yvv_coefficients.argtypes = [ctypes.c_double, P(ctypes.c_double)]
    L.itkg_recursive_gaussian_generate_region.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_double, P(ctypes.c_double), ctypes.c_int32,
        P(c_region), P(c_view), P(c_view), ctypes.c_void_p]
    L.itkg_recursive_gaussian_generate_region_host.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_double, P(ctypes.c_double), ctypes.c_int32,
        P(c_region), ctypes.c_void_p, P(c_region), ctypes.c_void_p]
    L.itkg_canny_generate_region.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
        P(c_region), P(c_view), P(c_view), ctypes.c_int, ctypes.c_void_p]
    L.itkg_gaussian_backward.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_int32, ctypes.c_double, P(ctypes.c_double),
        P(c_region), P(c_view), P(c_view), ctypes.c_int32, ctypes.c_void_p]
    L.itkg_batch_create.argtypes = [ctypes.c_int32, P(c_volume), ctypes.c_int, ctypes.c_int,
                                    P(ctypes.c_void_p)]
    L.itkg_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.itkg_batch_autotune.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.itkg_batch_voxels.restype = ctypes.c_int64
    L.itkg_batch_voxels.argtypes = [ctypes.c_void_p]
    L.itkg_batch_destroy.argtypes = [ctypes.c_void_p]
    L.itkg_last_error.restype = ctypes.c_char_p
    L.itkg_set_path.argtypes = [ctypes.c_int]
    L.itkg_trim_scratch.argtypes = [ctypes.c_uint64]
    L.itkg_host_staging_info.argtypes = [P(ctypes.c_int32)]
    L.itkg_canny_direction_bins.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                            ctypes.c_void_p, ctypes.c_void_p]
    L.itkg_gaussian_backward_host.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_int32, ctypes.c_double, P(ctypes.c_double),
        P(c_region), ctypes.c_void_p, P(c_region), ctypes.c_void_p, P(c_region), ctypes.c_int32]
    L.itkg_convert_pixels.argtypes = [ctypes.c_int32, ctypes.c_int, ctypes.c_int, ctypes.c_int32,
                                      P(c_view), P(c_view), ctypes.c_int, ctypes.c_void_p]
    L.itkg_cast_gaussian_cast_host.argtypes = [
        ctypes.c_int32, ctypes.c_int, ctypes.c_int, ctypes.c_double, P(ctypes.c_double),
        P(c_region), ctypes.c_void_p, P(c_region), ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.itkg_version.restype = ctypes.c_char_p
    return L


lib = _load()


def _check(rc: int) -> None:
    if rc != OK:
        raise ItkgError(rc, lib.itkg_last_error().decode())


def set_path(path: int) -> None:
    """Kernel selection: PATH_AUTO, PATH_PER_AXIS or PATH_FUSED."""
    _check(lib.itkg_set_path(path))


def trim_scratch(keep_bytes: int = 0, device=None) -> None:
    """Hands the library's idle scratch-pool memory on `device` (default: the
    current one) beyond `keep_bytes` back to the driver, e.g. before torch
    allocates large tensors (torch's allocator cannot see that pool)."""
    with _on_device(device if device is not None else torch.cuda.current_device()):
        _check(lib.itkg_trim_scratch(int(keep_bytes)))


def host_staging_pinned() -> bool:
    """Whether the last staged *_host call went through the library's pinned
    staging ring (False: it copied from and to pageable memory)."""
    v = ctypes.c_int32()
    _check(lib.itkg_host_staging_info(ctypes.byref(v)))
    return bool(v.value)


# ---------------------------------------------------------------------------
# Region (src/image.rs:9-77)
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class Region:
    index: Tuple[int, ...]
    size: Tuple[int, ...]

    def __post_init__(self):
        object.__setattr__(self, "index", tuple(int(i) for i in self.index))
        object.__setattr__(self, "size", tuple(int(s) for s in self.size))
        if len(self.index) != len(self.size):
            raise ValueError("index and size differ in length")
        if any(s < 0 for s in self.size):
            raise ValueError("negative size")

    @property
    def ndim(self) -> int:
        return len(self.size)

    def linear_len(self) -> int:
        n = 1
        for s in self.size:
            n *= s
        return n

    def padded_per_axis(self, radii: Sequence[int]) -> "Region":
        return Region([i - r for i, r in zip(self.index, radii)],
                      [s + 2 * r for s, r in zip(self.size, radii)])

    def padded(self, radius: int) -> "Region":
        return self.padded_per_axis([radius] * self.ndim)

    def clipped_to(self, bounds: "Region") -> "Region":
        idx, sz = [], []
        for i, s, bi, bs in zip(self.index, self.size, bounds.index, bounds.size):
            st, en = max(i, bi), min(i + s, bi + bs)
            idx.append(st)
            sz.append(max(en - st, 0))
        return Region(idx, sz)

    def contains(self, idx: Sequence[int]) -> bool:
        return all(i <= x < i + s for x, i, s in zip(idx, self.index, self.size))

    def to_c(self) -> c_region:
        r = c_region()
        for d in range(self.ndim):
            r.index[d] = self.index[d]
            r.size[d] = self.size[d]
        return r


def _doubles(v: Sequence[float]):
    return (ctypes.c_double * len(v))(*[float(x) for x in v])


def kernel_radius(sigma: float, spacing: float) -> int:
    """GaussianFilter::kernel_radius (src/filters/gaussian.rs:24-27)."""
    return int(lib.itkg_kernel_radius(float(sigma), float(spacing)))


def build_kernel(sigma_px: float, radius: int) -> List[float]:
    """GaussianFilter::build_kernel (src/filters/gaussian.rs:29-40): f64 weights."""
    n = 2 * int(radius) + 1
    out = (ctypes.c_double * n)()
    _check(lib.itkg_build_kernel(float(sigma_px), int(radius), out, n))
    return list(out)


def split_region(region: Region, num_pieces: int) -> Iterator[Region]:
    """split_region (src/region_splitter.rs:8-39)."""
    cap = max(int(num_pieces), 1)
    pieces = (c_region * cap)()
    count = ctypes.c_uint64(0)
    _check(lib.itkg_split_region(region.ndim, ctypes.byref(region.to_c()), int(num_pieces),
                                 pieces, cap, ctypes.byref(count)))
    D = region.ndim
    for p in pieces[:count.value]:
        yield Region(list(p.index)[:D], list(p.size)[:D])


# ---------------------------------------------------------------------------
# Image (src/image.rs:79-141) with a device-resident torch buffer
# ---------------------------------------------------------------------------

def _pixel_info(t: torch.Tensor, ndim: int) -> Tuple[int, int]:
    if t.dtype == torch.float32:
        dt = F32
    elif t.dtype == torch.float64:
        dt = F64
    else:
        raise TypeError(f"pixel dtype {t.dtype} has no GPU path (f32 and f64 only)")
    if t.dim() == ndim:
        return dt, 1
    if t.dim() == ndim + 1:
        return dt, int(t.shape[-1])
    raise ValueError(f"tensor of rank {t.dim()} does not hold a {ndim}-D image")


@dataclass
class Image:
    """Image<P, D>: `data` is a tensor of shape reversed(region.size) (+ (N,)
    for VecPixel<T, N>), axis 0 fastest in memory as flat_index defines."""
    region: Region
    spacing: Tuple[float, ...]
    origin: Tuple[float, ...]
    data: torch.Tensor

    @staticmethod
    def allocate(region: Region, spacing, origin, fill=0.0, ncomp: int = 1,
                 dtype=torch.float32, device="cuda") -> "Image":
        shape = tuple(reversed(region.size)) + ((ncomp,) if ncomp > 1 else ())
        return Image(region, tuple(spacing), tuple(origin),
                     torch.full(shape, float(fill), dtype=dtype, device=device))

    @staticmethod
    def empty(region: Region, spacing, origin, ncomp: int = 1, dtype=torch.float32,
              device="cuda") -> "Image":
        """An image whose pixels a kernel is about to write in full (no fill)."""
        shape = tuple(reversed(region.size)) + ((ncomp,) if ncomp > 1 else ())
        return Image(region, tuple(spacing), tuple(origin),
                     torch.empty(shape, dtype=dtype, device=device))

    @property
    def ndim(self) -> int:
        return self.region.ndim

    def _slices(self, r: Region):
        D = self.ndim
        return tuple(slice(r.index[d] - self.region.index[d],
                           r.index[d] - self.region.index[d] + r.size[d])
                     for d in reversed(range(D)))

    def view(self, r: Optional[Region] = None) -> c_view:
        """Strided view of this image's pixels (whole buffer by default); the
        last one built is kept while the buffer, its layout and the region
        stay the same."""
        t = self.data
        key = (t.data_ptr(), t.shape, t.stride(), self.region) if t is not None else None
        cached = self.__dict__.get("_view_cache")
        if cached is not None and key is not None and cached[0] == key:
            return cached[1]
        v = self._make_view()
        if key is not None:
            self.__dict__["_view_cache"] = (key, v)
        return v

    def _make_view(self) -> c_view:
        t = self.data
        D = self.ndim
        if t.dim() not in (D, D + 1):
            raise ValueError(f"tensor of rank {t.dim()} does not hold a {D}-D image")
        nc = 1 if t.dim() == D else int(t.shape[-1])
        if nc > 1 and t.stride(-1) != 1:
            raise ValueError("vector pixels must have contiguous components")
        v = c_view()
        v.data = t.data_ptr() if t.numel() else None
        v.region = self.region.to_c()
        for d in range(D):
            st = t.stride(D - 1 - d)
            if nc > 1:
                if st % nc:
                    raise ValueError("pixel stride is not a whole number of pixels")
                st //= nc
            v.stride[d] = st
        return v

    def get_pixel(self, idx):
        return self.data[tuple(i - o for i, o in zip(reversed(idx), reversed(self.region.index)))]

    def set_pixel(self, idx, val):
        self.data[tuple(i - o for i, o in zip(reversed(idx), reversed(self.region.index)))] = \
            torch.as_tensor(val, dtype=self.data.dtype)

    def copy_from(self, src: "Image") -> None:
        """src/image.rs:133-140 (src.region must lie inside self.region)."""
        self.data[self._slices(src.region)].copy_(src.data)

    # --- ImageSource for Image (src/source.rs:32-59): the leaf
    def largest_region(self) -> Region:
        return self.region

    def input_region_for_output(self, output_region: Region) -> Region:
        return output_region

    def generate_region(self, requested: Region) -> "Image":
        """Leaf: the requested region clipped to the buffer, as a view (no copy)."""
        clipped = requested.clipped_to(self.region)
        return Image(clipped, self.spacing, self.origin, self.data[self._slices(clipped)])


class ImageSource:
    """The pipeline protocol of src/source.rs:13-29."""

    def largest_region(self) -> Region:
        raise NotImplementedError

    @property
    def spacing(self):
        raise NotImplementedError

    @property
    def origin(self):
        raise NotImplementedError

    def input_region_for_output(self, output_region: Region) -> Region:
        return output_region

    def generate_region(self, requested: Region) -> Image:
        raise NotImplementedError


_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_NO_SWITCH = contextlib.nullcontext()


def _device_index(device) -> int:
    d = torch.device(device) if not isinstance(device, torch.device) else device
    return d.index if d.index is not None else torch.cuda.current_device()


def _stream_ptr(device) -> ctypes.c_void_p:
    """torch's current stream on `device` (the raw handle, without building a
    Stream object where torch offers that)."""
    if _RAW_STREAM is not None:
        return ctypes.c_void_p(_RAW_STREAM(_device_index(device)))
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _on_device(device):
    """Makes `device` current for the block (no switch when it already is)."""
    idx = _device_index(device)
    return _NO_SWITCH if idx == torch.cuda.current_device() else torch.cuda.device(idx)


class GaussianFilter(ImageSource):
    """GaussianFilter<S> (src/filters/gaussian.rs:13-91) on the GPU.

    Same two public fields as the reference (`source`, `sigma`); the
    arithmetic mode is a module-wide choice (`GaussianFilter.arith`), EXACT by
    default (bit-identical to the reference)."""

    arith = EXACT

    def __init__(self, source, sigma: float):
        self.source = source
        self.sigma = float(sigma)

    def largest_region(self) -> Region:
        return self.source.largest_region()

    @property
    def spacing(self):
        return tuple(self.source.spacing)

    @property
    def origin(self):
        return tuple(self.source.origin)

    def _radii(self):
        return [kernel_radius(self.sigma, sp) for sp in self.spacing]

    def input_region_for_output(self, output_region: Region) -> Region:
        """src/filters/gaussian.rs:52-57."""
        out = c_region()
        D = output_region.ndim
        _check(lib.itkg_input_region_for_output(
            D, self.sigma, _doubles(self.spacing), ctypes.byref(self.largest_region().to_c()),
            ctypes.byref(output_region.to_c()), ctypes.byref(out)))
        return Region(list(out.index)[:D], list(out.size)[:D])

    def generate_region(self, requested: Region) -> Image:
        """src/filters/gaussian.rs:59-90: pull the padded input, run the passes."""
        input_region = requested.padded_per_axis(self._radii()).clipped_to(self.largest_region())
        inp = self.source.generate_region(input_region)
        dt, nc = _pixel_info(inp.data, requested.ndim)
        out = Image.empty(requested, self.spacing, self.origin, nc, inp.data.dtype,
                          inp.data.device)
        self.generate_into(inp, out)
        return out

    def generate_into(self, inp: Image, out: Image) -> None:
        """Write generate_region(out.region) into `out` from an upstream image
        `inp` that covers input_region_for_output(out.region)."""
        D = out.ndim
        dt, nc = _pixel_info(inp.data, D)
        with _on_device(inp.data.device):
            _check(lib.itkg_gaussian_generate_region(
                D, dt, nc, self.sigma, _doubles(self.spacing),
                ctypes.byref(self.largest_region().to_c()), ctypes.byref(inp.view()),
                ctypes.byref(out.view()), self.arith, _stream_ptr(inp.data.device)))

    def backward(self, grad_output: Image, accumulate_into: Optional[Image] = None) -> Image:
        """Adjoint of generate_region(grad_output.region): returns the gradient
        on input_region_for_output(grad_output.region), or adds it into
        `accumulate_into` (whose region must contain that region)."""
        D = grad_output.ndim
        dt, nc = _pixel_info(grad_output.data, D)
        if accumulate_into is None:
            in_reg = self.input_region_for_output(grad_output.region)
            gin = Image.allocate(in_reg, self.spacing, self.origin, 0.0, nc,
                                 grad_output.data.dtype, grad_output.data.device)
            self.backward_into(grad_output, gin, False)
        else:
            gin = accumulate_into
            self.backward_into(grad_output, gin, True)
        return gin

    def backward_into(self, grad_output: Image, grad_input: Image, accumulate: bool) -> None:
        """Writes (or adds) the adjoint of generate_region(grad_output.region)
        into `grad_input`, whose region must contain input_region_for_output."""
        D = grad_output.ndim
        dt, nc = _pixel_info(grad_output.data, D)
        with _on_device(grad_output.data.device):
            _check(lib.itkg_gaussian_backward(
                D, dt, nc, self.sigma, _doubles(self.spacing),
                ctypes.byref(self.largest_region().to_c()), ctypes.byref(grad_output.view()),
                ctypes.byref(grad_input.view()), int(accumulate),
                _stream_ptr(grad_output.data.device)))


def stream_full(source, num_pieces: int) -> Image:
    """stream_full (src/source.rs:81-97): allocate the full output, fill it
    piece by piece.  A GaussianFilter piece is written straight into its slab
    of the output (same values as generate_region + copy_from)."""
    full = source.largest_region()
    first = source.generate_region(Region(full.index, [0] * full.ndim))
    nc = first.data.shape[-1] if first.data.dim() > full.ndim else 1
    if isinstance(source, GaussianFilter):
        # the pieces cover the region and each writes its whole slab
        out = Image.empty(full, source.spacing, source.origin, nc, first.data.dtype,
                          first.data.device)
    else:
        out = Image.allocate(full, source.spacing, source.origin, 0.0, nc, first.data.dtype,
                             first.data.device)
    if isinstance(source, GaussianFilter) and isinstance(source.source, Image) and \
            source.source.data.is_cuda:
        # every piece from the resident leaf in one call (one persistent
        # launch over all pieces' tiles where the fused kernel serves them;
        # each piece still reads its own halo, with the same results)
        leaf = source.source
        slabs = [Image(piece, out.spacing, out.origin, out.data[out._slices(piece)])
                 for piece in split_region(full, num_pieces)]
        views = (c_view * len(slabs))(*[sl.view() for sl in slabs])
        D = full.ndim
        dt, nc = _pixel_info(leaf.data, D)
        with _on_device(leaf.data.device):
            _check(lib.itkg_gaussian_generate_regions(
                D, dt, nc, source.sigma, _doubles(source.spacing),
                ctypes.byref(source.largest_region().to_c()), ctypes.byref(leaf.view()),
                len(slabs), views, source.arith, _stream_ptr(leaf.data.device)))
        return out
    for piece in split_region(full, num_pieces):
        if isinstance(source, GaussianFilter):
            in_reg = source.input_region_for_output(piece)
            inp = source.source.generate_region(in_reg)
            slab = Image(piece, out.spacing, out.origin, out.data[out._slices(piece)])
            source.generate_into(inp, slab)
        else:
            out.copy_from(source.generate_region(piece))
    return out


# ---------------------------------------------------------------------------
# the other convolve_axis filters (src/filters/mean.rs, binomial_blur.rs)
# ---------------------------------------------------------------------------

def box_kernel(radius: int) -> List[float]:
    """box_kernel (src/filters/mean.rs:15-18): 2r+1 weights of 1/(2r+1)."""
    n = 2 * int(radius) + 1
    out = (ctypes.c_double * n)()
    _check(lib.itkg_box_kernel(int(radius), out, n))
    return list(out)


BINOMIAL_KERNEL = (0.25, 0.5, 0.25)  # src/filters/binomial_blur.rs:35


def _separable_input_region(radii: Sequence[int], largest: Region, output_region: Region) -> Region:
    D = output_region.ndim
    out = c_region()
    _check(lib.itkg_separable_input_region_for_output(
        D, (ctypes.c_uint64 * D)(*[int(r) for r in radii]), ctypes.byref(largest.to_c()),
        ctypes.byref(output_region.to_c()), ctypes.byref(out)))
    return Region(list(out.index)[:D], list(out.size)[:D])


def _separable_into(radii: Sequence[int], weights: Sequence[Sequence[float]], largest: Region,
                    inp: Image, out: Image, arith: int) -> None:
    """One generate_region of a separable clamp-to-edge filter (per-axis
    radii and weights) from `inp` into `out` (itkg_separable_generate_region)."""
    D = out.ndim
    dt, nc = _pixel_info(inp.data, D)
    flat = [float(w) for ws in weights for w in ws]
    with _on_device(inp.data.device):
        _check(lib.itkg_separable_generate_region(
            D, dt, nc, (ctypes.c_uint64 * D)(*[int(r) for r in radii]),
            (ctypes.c_double * len(flat))(*flat), ctypes.byref(largest.to_c()),
            ctypes.byref(inp.view()), ctypes.byref(out.view()), arith,
            _stream_ptr(inp.data.device)))


class _SeparableFilter(ImageSource):
    """A separable filter with fixed per-axis radii and weights: the shape of
    MeanFilter and BoxMeanFilterN (src/filters/mean.rs:42-83, 115-144)."""

    arith = EXACT

    def largest_region(self) -> Region:
        return self.source.largest_region()

    @property
    def spacing(self):
        return tuple(self.source.spacing)

    @property
    def origin(self):
        return tuple(self.source.origin)

    def _radii(self, D: int) -> List[int]:
        raise NotImplementedError

    def input_region_for_output(self, output_region: Region) -> Region:
        return _separable_input_region(self._radii(output_region.ndim), self.largest_region(),
                                       output_region)

    def generate_region(self, requested: Region) -> Image:
        radii = self._radii(requested.ndim)
        input_region = requested.padded_per_axis(radii).clipped_to(self.largest_region())
        inp = self.source.generate_region(input_region)
        _, nc = _pixel_info(inp.data, requested.ndim)
        out = Image.empty(requested, self.spacing, self.origin, nc, inp.data.dtype,
                          inp.data.device)
        self.generate_into(inp, out)
        return out

    def generate_into(self, inp: Image, out: Image) -> None:
        """Write generate_region(out.region) into `out` from an upstream image
        covering input_region_for_output(out.region)."""
        radii = self._radii(out.ndim)
        _separable_into(radii, [box_kernel(r) for r in radii], self.largest_region(), inp, out,
                        self.arith)


class MeanFilter(_SeparableFilter):
    """MeanFilter<S> (src/filters/mean.rs:28-84): box mean of half-width
    `radius` voxels on every axis, in index space; the same two public fields."""

    def __init__(self, source, radius: int):
        self.source = source
        self.radius = int(radius)

    def _radii(self, D: int) -> List[int]:
        return [self.radius] * D


class BoxMeanFilterN(_SeparableFilter):
    """BoxMeanFilterN<S, D> (src/filters/mean.rs:102-145): box mean with a
    half-width per axis."""

    def __init__(self, source, radii: Sequence[int]):
        self.source = source
        self.radii = [int(r) for r in radii]

    def _radii(self, D: int) -> List[int]:
        if len(self.radii) != D:
            raise ValueError(f"{len(self.radii)} radii for a {D}-D image")
        return list(self.radii)


class BinomialBlurFilter(ImageSource):
    """BinomialBlurFilter<S> (src/filters/binomial_blur.rs:17-90): `repetitions`
    rounds of the [1, 2, 1] / 4 kernel on every axis.  Each round is one
    radius-1 separable generate_region whose requested region is the final
    request padded by the rounds still to come (the per-pass regions of
    binomial_blur.rs:70-85); the rounds before the last ping-pong between two
    device buffers the filter keeps between calls."""

    arith = EXACT

    def __init__(self, source, repetitions: int):
        self.source = source
        self.repetitions = int(repetitions)

    @staticmethod
    def effective_sigma_pixels(repetitions: int) -> float:
        """binomial_blur.rs:29-31: sqrt(n / 4)."""
        return (repetitions / 4.0) ** 0.5

    def largest_region(self) -> Region:
        return self.source.largest_region()

    @property
    def spacing(self):
        return tuple(self.source.spacing)

    @property
    def origin(self):
        return tuple(self.source.origin)

    def input_region_for_output(self, output_region: Region) -> Region:
        return _separable_input_region([self.repetitions] * output_region.ndim,
                                       self.largest_region(), output_region)

    def _scratch(self, k: int, numel: int, dtype, device) -> torch.Tensor:
        """Intermediate buffer k (0 or 1) of at least `numel` pixels, kept
        across calls; a call on another stream first waits for the last one."""
        bufs = self.__dict__.setdefault("_bufs", {})
        t = bufs.get(k)
        if t is None or t.numel() < numel or t.dtype != dtype or t.device != device:
            t = torch.empty(max(numel, 1), dtype=dtype, device=device)
            bufs[k] = t
        return t[:numel]

    def generate_into(self, inp: Image, out: Image) -> None:
        """Write generate_region(out.region) into `out` from an upstream image
        covering input_region_for_output(out.region); the rounds before the
        last run in two buffers the filter keeps between calls."""
        requested = out.region
        D = requested.ndim
        bounds = self.largest_region()
        n = self.repetitions
        _, nc = _pixel_info(inp.data, D)
        if n == 0:
            out.data.copy_(inp.data[inp._slices(requested)])
            return
        stream = torch.cuda.current_stream(inp.data.device)
        last = self.__dict__.get("_stream")
        if last is not None and last != stream:
            stream.wait_stream(last)
        self._stream = stream
        cur = inp
        for rep in range(n):
            remaining = n - rep - 1
            if remaining == 0:
                nxt = out
            else:
                reg = requested.padded(remaining).clipped_to(bounds)
                shape = tuple(reversed(reg.size)) + ((nc,) if nc > 1 else ())
                numel = 1
                for v in shape:
                    numel *= v
                buf = self._scratch(rep % 2, numel, inp.data.dtype, inp.data.device)
                nxt = Image(reg, self.spacing, self.origin, buf.view(shape))
            _separable_into([1] * D, [BINOMIAL_KERNEL] * D, bounds, cur, nxt, self.arith)
            cur = nxt

    def generate_region(self, requested: Region) -> Image:
        D = requested.ndim
        inp = self.source.generate_region(self.input_region_for_output(requested))
        _, nc = _pixel_info(inp.data, D)
        out = Image.empty(requested, self.spacing, self.origin, nc, inp.data.dtype, inp.data.device)
        self.generate_into(inp, out)
        return out


# ---------------------------------------------------------------------------
# UnaryFilter casts (src/filters/mod.rs:47-80) on the GPU
# ---------------------------------------------------------------------------

_PTYPES = {torch.uint8: PT_U8, torch.int16: PT_I16, torch.int32: PT_I32,
           torch.float32: PT_F32, torch.float64: PT_F64}
if hasattr(torch, "uint16"):
    _PTYPES[torch.uint16] = PT_U16
_NP_PTYPES = {np.dtype(np.uint8): PT_U8, np.dtype(np.uint16): PT_U16, np.dtype(np.int16): PT_I16,
              np.dtype(np.int32): PT_I32, np.dtype(np.float32): PT_F32,
              np.dtype(np.float64): PT_F64}


def _ptype(dtype) -> int:
    if dtype not in _PTYPES:
        raise TypeError(f"no GPU conversion for {dtype}")
    return _PTYPES[dtype]


class CastFilter(ImageSource):
    """UnaryFilter<S, F> (src/filters/mod.rs:47-80) whose closure is a cast,
    on the GPU: `mode` CONVERT_AS is Rust's `p as T` (e.g. `|p: u8| p as
    f32`), CONVERT_ROUND is `p.round() ... as T` (e.g. `|p: f32|
    p.round().clamp(0.0, 255.0) as u8`, examples/blur_png.rs:35-37).  Same
    region, spacing and origin as the upstream; a chain of these around a
    GaussianFilter stays resident on the device."""

    def __init__(self, source, dtype: torch.dtype, mode: int = CONVERT_AS):
        self.source = source
        self.dtype = dtype
        self.mode = int(mode)

    def largest_region(self) -> Region:
        return self.source.largest_region()

    @property
    def spacing(self):
        return tuple(self.source.spacing)

    @property
    def origin(self):
        return tuple(self.source.origin)

    def input_region_for_output(self, output_region: Region) -> Region:
        return output_region

    def generate_into(self, inp: Image, out: Image) -> None:
        D = out.ndim
        nc = 1 if inp.data.dim() == D else int(inp.data.shape[-1])
        with _on_device(inp.data.device):
            _check(lib.itkg_convert_pixels(D, _ptype(inp.data.dtype), _ptype(out.data.dtype), nc,
                                           ctypes.byref(inp.view()), ctypes.byref(out.view()),
                                           self.mode, _stream_ptr(inp.data.device)))

    def generate_region(self, requested: Region) -> Image:
        inp = self.source.generate_region(requested)
        D = requested.ndim
        nc = 1 if inp.data.dim() == D else int(inp.data.shape[-1])
        out = Image.empty(inp.region, self.spacing, self.origin, nc, self.dtype, inp.data.device)
        self.generate_into(inp, out)
        return out


def cast_gaussian_cast_host(x: np.ndarray, sigma: float, out_dtype, spacing=None,
                            arith: int = EXACT, out_mode: int = CONVERT_ROUND) -> np.ndarray:
    """UnaryFilter(x as f32) -> GaussianFilter(sigma) -> UnaryFilter(cast to
    out_dtype) over a whole host image in one GPU round trip
    (itkg_cast_gaussian_cast_host): PCIe carries the pixels in their own
    types."""
    x = np.ascontiguousarray(x)
    D = x.ndim
    sp = tuple(spacing) if spacing is not None else (1.0,) * D
    reg = Region([0] * D, list(reversed(x.shape))).to_c()
    out = np.empty(x.shape, dtype=out_dtype)
    _check(lib.itkg_cast_gaussian_cast_host(
        D, _NP_PTYPES[x.dtype], _NP_PTYPES[out.dtype], float(sigma), _doubles(sp),
        ctypes.byref(reg), x.ctypes.data, ctypes.byref(reg), out.ctypes.data, arith, out_mode))
    return out


# ---------------------------------------------------------------------------
# Gaussian consumers with an index-space sigma (src/filters/edges.rs)
# ---------------------------------------------------------------------------

def index_gaussian_kernel(sigma: float) -> Tuple[int, List[float]]:
    """The kernel of UnsharpMaskFilter / CannyEdgeDetectionFilter
    (src/filters/edges.rs:414-423): (radius, f64 weights)."""
    cap = 2 * MAX_RADIUS + 1
    out = (ctypes.c_double * cap)()
    r = ctypes.c_uint64()
    _check(lib.itkg_index_gaussian_kernel(float(sigma), out, cap, ctypes.byref(r)))
    return int(r.value), list(out)[:2 * int(r.value) + 1]


class UnsharpMaskFilter(ImageSource):
    """UnsharpMaskFilter<S> (src/filters/edges.rs:376-452): I + amount *
    (I - blur(I)) where |I - blur(I)| > threshold; the same public fields
    (`source`, `sigma` in pixels, `amount` = 0.5, `threshold` = 0).  Scalar
    f32 / f64 pixels.  The blur's arithmetic is `arith` (EXACT by default)."""

    arith = EXACT

    def __init__(self, source, sigma: float):
        self.source = source
        self.sigma = float(sigma)
        self.amount = 0.5
        self.threshold = 0.0

    def with_amount(self, a: float) -> "UnsharpMaskFilter":
        self.amount = float(a)
        return self

    def largest_region(self) -> Region:
        return self.source.largest_region()

    @property
    def spacing(self):
        return tuple(self.source.spacing)

    @property
    def origin(self):
        return tuple(self.source.origin)

    def input_region_for_output(self, output_region: Region) -> Region:
        """edges.rs:401-404: the request padded by ceil(3 sigma), clipped."""
        return _separable_input_region([kernel_radius(self.sigma, 1.0)] * output_region.ndim,
                                       self.largest_region(), output_region)

    def generate_into(self, inp: Image, out: Image) -> None:
        D = out.ndim
        dt, nc = _pixel_info(inp.data, D)
        if nc != 1:
            raise ItkgError(ERR_UNSUPPORTED, "unsharp mask takes scalar pixels")
        with _on_device(inp.data.device):
            _check(lib.itkg_unsharp_mask_generate_region(
                D, dt, self.sigma, self.amount, self.threshold,
                ctypes.byref(self.largest_region().to_c()), ctypes.byref(inp.view()),
                ctypes.byref(out.view()), self.arith, _stream_ptr(inp.data.device)))

    def generate_region(self, requested: Region) -> Image:
        inp = self.source.generate_region(self.input_region_for_output(requested))
        _, nc = _pixel_info(inp.data, requested.ndim)
        out = Image.empty(requested, self.spacing, self.origin, nc, inp.data.dtype, inp.data.device)
        self.generate_into(inp, out)
        return out


class CannyEdgeDetectionFilter(ImageSource):
    """CannyEdgeDetectionFilter<S> (src/filters/edges.rs:527-665), 2-D scalar
    f32 / f64: smoothing with the index-space kernel, gradient, non-maximum
    suppression and hysteresis on the GPU; output 1 / 0.  The same public
    fields (`source`, `sigma`, `lower_threshold`, `upper_threshold`)."""

    arith = EXACT

    def __init__(self, source, sigma: float, lower: float, upper: float):
        self.source = source
        self.sigma = float(sigma)
        self.lower_threshold = float(lower)
        self.upper_threshold = float(upper)

    def largest_region(self) -> Region:
        return self.source.largest_region()

    @property
    def spacing(self):
        return tuple(self.source.spacing)

    @property
    def origin(self):
        return tuple(self.source.origin)

    def input_region_for_output(self, output_region: Region) -> Region:
        """edges.rs:548-550: the whole image."""
        return self.largest_region()

    def generate_into(self, inp: Image, out: Image) -> None:
        D = out.ndim
        dt, nc = _pixel_info(inp.data, D)
        if nc != 1:
            raise ItkgError(ERR_UNSUPPORTED, "Canny takes scalar pixels")
        with _on_device(inp.data.device):
            _check(lib.itkg_canny_generate_region(
                D, dt, self.sigma, self.lower_threshold, self.upper_threshold,
                ctypes.byref(self.largest_region().to_c()), ctypes.byref(inp.view()),
                ctypes.byref(out.view()), self.arith, _stream_ptr(inp.data.device)))

    def generate_region(self, requested: Region) -> Image:
        inp = self.source.generate_region(self.largest_region())
        _, nc = _pixel_info(inp.data, requested.ndim)
        out = Image.empty(requested, self.spacing, self.origin, nc, inp.data.dtype, inp.data.device)
        self.generate_into(inp, out)
        return out


# ---------------------------------------------------------------------------
# recursive (IIR) Gaussian (src/filters/recursive_gaussian.rs)
# ---------------------------------------------------------------------------

def yvv_coefficients(sigma_px: float) -> Tuple[float, float, float, float]:
    """yvv_coeffs (src/filters/recursive_gaussian.rs:33-47): (B, a1, a2, a3)."""
    c = (ctypes.c_double * 4)()
    _check(lib.itkg_yvv_coefficients(float(sigma_px), c))
    return tuple(c)


class RecursiveGaussianFilter(ImageSource):
    """RecursiveGaussianFilter<S> (src/filters/recursive_gaussian.rs:129-177):
    the Young - van Vliet recursive Gaussian along every axis (physical
    sigma), f64 inside, on the whole image; scalar f32 / f64 pixels."""

    _axis = -1

    def __init__(self, source, sigma: float):
        self.source = source
        self.sigma = float(sigma)

    def largest_region(self) -> Region:
        return self.source.largest_region()

    @property
    def spacing(self):
        return tuple(self.source.spacing)

    @property
    def origin(self):
        return tuple(self.source.origin)

    def input_region_for_output(self, output_region: Region) -> Region:
        """recursive_gaussian.rs:156-158: the whole image (infinite support)."""
        return self.largest_region()

    def generate_into(self, inp: Image, out: Image) -> None:
        D = out.ndim
        dt, nc = _pixel_info(inp.data, D)
        if nc != 1:
            raise ItkgError(ERR_UNSUPPORTED, "the recursive Gaussian takes scalar pixels")
        with _on_device(inp.data.device):
            _check(lib.itkg_recursive_gaussian_generate_region(
                D, dt, self.sigma, _doubles(self.spacing), self._axis,
                ctypes.byref(self.largest_region().to_c()), ctypes.byref(inp.view()),
                ctypes.byref(out.view()), _stream_ptr(inp.data.device)))

    def generate_region(self, requested: Region) -> Image:
        inp = self.source.generate_region(self.largest_region())
        _, nc = _pixel_info(inp.data, requested.ndim)
        out = Image.empty(requested, self.spacing, self.origin, nc, inp.data.dtype, inp.data.device)
        self.generate_into(inp, out)
        return out


class RecursiveGaussianAxisFilter(RecursiveGaussianFilter):
    """RecursiveGaussianAxisFilter<S> (src/filters/recursive_gaussian.rs:181-241):
    the same along one axis (`axis`, 0 = x)."""

    def __init__(self, source, sigma: float, axis: int):
        super().__init__(source, sigma)
        self.axis = int(axis)

    @property
    def _axis(self):
        return self.axis


# ---------------------------------------------------------------------------
# in-crate GaussianFilter callers: the adapters whose whole work is the blur
# ---------------------------------------------------------------------------

class BSplineDownsampleFilter(ImageSource):
    """BSplineDownsampleFilter<S, P> (src/filters/spatial.rs:1128-1170):
    GaussianFilter with sigma 1.0 over the source's whole largest region,
    then every second pixel along each axis from the region's first index.
    Like the reference it returns the whole output region whatever the
    request."""

    def __init__(self, source):
        self.source = source

    def largest_region(self) -> Region:
        r = self.source.largest_region()
        return Region(list(r.index), [(s + 1) // 2 for s in r.size])

    @property
    def spacing(self):
        return tuple(2.0 * s for s in self.source.spacing)

    @property
    def origin(self):
        return tuple(self.source.origin)

    def generate_region(self, requested: Region) -> Image:
        smoothed = GaussianFilter(self.source, 1.0).generate_region(self.source.largest_region())
        D = smoothed.ndim
        pick = smoothed.data[(slice(None, None, 2),) * D]
        return Image(self.largest_region(), self.spacing, self.origin, pick.contiguous())


class SliceBySliceFilter(ImageSource):
    """SliceBySliceFilter<S, P> (src/filters/spatial.rs:1284-1340): a 2-D
    GaussianFilter (sigma over the slice's x and y spacing) on every z slice
    of a 3-D image, each slice's buffer being the pulled input's.  Here one
    separable call over the volume with z radius 0 (the weight 1.0, which
    leaves each value as the y pass wrote it)."""

    arith = EXACT

    def __init__(self, source, sigma: float):
        self.source = source
        self.sigma = float(sigma)

    def largest_region(self) -> Region:
        return self.source.largest_region()

    @property
    def spacing(self):
        return tuple(self.source.spacing)

    @property
    def origin(self):
        return tuple(self.source.origin)

    def generate_region(self, requested: Region) -> Image:
        if requested.ndim != 3:
            raise ValueError("SliceBySliceFilter takes 3-D images")
        inp = self.source.generate_region(requested)
        radii = [kernel_radius(self.sigma, sp) for sp in inp.spacing[:2]] + [0]
        weights = [build_kernel(self.sigma / sp, r) for sp, r in zip(inp.spacing[:2], radii)]
        weights.append([1.0])
        _, nc = _pixel_info(inp.data, 3)
        out = Image.empty(inp.region, inp.spacing, inp.origin, nc, inp.data.dtype,
                          inp.data.device)
        _separable_into(radii, weights, inp.region, inp, out, self.arith)
        return out


class DisplacementFieldToBSplineFilter(ImageSource):
    """DisplacementFieldToBSplineFilter<S> (src/filters/spatial.rs:1489-1525):
    a 2-D field of VecPixel<f32, 2>, each component smoothed by GaussianFilter
    (sigma 2.0 by default) over the pulled input's own region.  One call with
    two interleaved components: each component's arithmetic is the scalar
    filter's."""

    def __init__(self, source, sigma: float = 2.0):
        self.source = source
        self.sigma = float(sigma)

    def largest_region(self) -> Region:
        return self.source.largest_region()

    @property
    def spacing(self):
        return tuple(self.source.spacing)

    @property
    def origin(self):
        return tuple(self.source.origin)

    def generate_region(self, requested: Region) -> Image:
        inp = self.source.generate_region(requested)
        if inp.ndim != 2 or _pixel_info(inp.data, 2)[1] != 2:
            raise ValueError("DisplacementFieldToBSplineFilter takes 2-D fields of 2 components")
        return GaussianFilter(inp, self.sigma).generate_region(inp.region)


# ---------------------------------------------------------------------------
# functional / autograd entry points
# ---------------------------------------------------------------------------

def _full_image(t: torch.Tensor, ndim: int, spacing) -> Image:
    size = list(reversed(t.shape[:ndim]))
    return Image(Region([0] * ndim, size), tuple(spacing), (0.0,) * ndim, t)


def gaussian(x: torch.Tensor, sigma: float, spacing=None, ndim: Optional[int] = None,
             arith: int = EXACT) -> torch.Tensor:
    """Whole-image GaussianFilter{source: x, sigma}.generate_region(largest)."""
    nd = ndim if ndim is not None else x.dim()
    sp = tuple(spacing) if spacing is not None else (1.0,) * nd
    if not x.is_contiguous():
        x = x.contiguous()
    dt, nc = _pixel_info(x, nd)
    out = torch.empty_like(x)
    vin, vout = _whole_views(x, out, nd, nc)
    with _on_device(x.device):
        _check(lib.itkg_gaussian_generate_region(
            nd, dt, nc, float(sigma), _spacing_array(sp), ctypes.byref(vin.region), ctypes.byref(vin),
            ctypes.byref(vout), arith, _stream_ptr(x.device)))
    return out


# Per-thread view structs of gaussian(), by layout: only the data pointers
# change from call to call (building the structs field by field costs a
# small image's call more host time than its kernel takes).  The library
# reads them during the call only.
_VIEWS = threading.local()
_SPACINGS: dict = {}


def _whole_views(x: torch.Tensor, out: torch.Tensor, nd: int, nc: int) -> Tuple[c_view, c_view]:
    cache = getattr(_VIEWS, "d", None)
    if cache is None:
        cache = _VIEWS.d = {}
    key = (x.shape, x.stride(), nd, nc)
    pair = cache.get(key)
    if pair is None:
        if len(cache) >= 64:
            cache.clear()
        pair = cache[key] = (_whole_view(x, nd, nc), _whole_view(out, nd, nc))
    vin, vout = pair
    vin.data = x.data_ptr() if x.numel() else None
    vout.data = out.data_ptr() if out.numel() else None
    return vin, vout


def _spacing_array(sp: Tuple[float, ...]):
    """The spacing as a C double array, one per distinct spacing (read only)."""
    a = _SPACINGS.get(sp)
    if a is None:
        if len(_SPACINGS) >= 256:
            _SPACINGS.clear()
        a = _SPACINGS[sp] = _doubles(sp)
    return a


def _whole_view(t: torch.Tensor, nd: int, nc: int) -> c_view:
    """The view of a contiguous tensor as a whole image at index 0 (what
    Image(Region([0] * nd, reversed shape), ..., t).view() gives)."""
    v = c_view()
    v.data = t.data_ptr() if t.numel() else None
    st = t.stride()
    for d in range(nd):
        v.region.index[d] = 0
        v.region.size[d] = t.shape[nd - 1 - d]
        v.stride[d] = st[nd - 1 - d] // nc
    return v


def gaussian_backward(g: torch.Tensor, sigma: float, spacing=None, ndim: Optional[int] = None,
                      out: Optional[torch.Tensor] = None, accumulate: bool = False) -> torch.Tensor:
    """Adjoint of the whole-image Gaussian: gradient w.r.t. the input."""
    nd = ndim if ndim is not None else g.dim()
    sp = tuple(spacing) if spacing is not None else (1.0,) * nd
    g = g.contiguous()
    gimg = _full_image(g, nd, sp)
    if out is None:
        out = torch.empty_like(g)
        accumulate = False
    oimg = _full_image(out, nd, sp)
    dt, nc = _pixel_info(g, nd)
    with _on_device(g.device):
        _check(lib.itkg_gaussian_backward(
            nd, dt, nc, float(sigma), _doubles(sp), ctypes.byref(gimg.region.to_c()),
            ctypes.byref(gimg.view()), ctypes.byref(oimg.view()), int(accumulate),
            _stream_ptr(g.device)))
    return out


def mean_filter(x: torch.Tensor, radius, ndim: Optional[int] = None,
                arith: int = EXACT) -> torch.Tensor:
    """Whole-image MeanFilter{source: x, radius} (an int) or BoxMeanFilterN
    {source: x, radii} (a sequence) .generate_region(largest)."""
    nd = ndim if ndim is not None else x.dim()
    src = _full_image(x.contiguous(), nd, (1.0,) * nd)
    f = BoxMeanFilterN(src, radius) if isinstance(radius, (list, tuple)) else MeanFilter(src, radius)
    f.arith = arith
    out = Image.empty(src.region, src.spacing, src.origin,
                      x.shape[-1] if x.dim() > nd else 1, x.dtype, x.device)
    f.generate_into(src, out)
    return out.data


def binomial_blur(x: torch.Tensor, repetitions: int, ndim: Optional[int] = None,
                  arith: int = EXACT) -> torch.Tensor:
    """Whole-image BinomialBlurFilter{source: x, repetitions}.generate_region(largest)."""
    nd = ndim if ndim is not None else x.dim()
    src = _full_image(x.contiguous(), nd, (1.0,) * nd)
    f = BinomialBlurFilter(src, repetitions)
    f.arith = arith
    return f.generate_region(src.region).data


class GaussianFunction(torch.autograd.Function):
    """Differentiable whole-image Gaussian; backward is the hand-written adjoint."""

    @staticmethod
    def forward(ctx, x, sigma, spacing, ndim, arith):
        ctx.sigma, ctx.spacing, ctx.ndim = sigma, spacing, ndim
        return gaussian(x, sigma, spacing, ndim, arith)

    @staticmethod
    def backward(ctx, g):
        return gaussian_backward(g, ctx.sigma, ctx.spacing, ctx.ndim), None, None, None, None


def gaussian_autograd(x, sigma, spacing=None, ndim=None, arith=EXACT):
    nd = ndim if ndim is not None else x.dim()
    sp = tuple(spacing) if spacing is not None else (1.0,) * nd
    return GaussianFunction.apply(x, float(sigma), sp, nd, arith)


# ---------------------------------------------------------------------------
# batched many-volume form
# ---------------------------------------------------------------------------

class BatchPlan:
    """One launch over many whole 3-D f32 volumes (itkg_batch_*)."""

    def __init__(self, inputs: Sequence[torch.Tensor], outputs: Sequence[torch.Tensor],
                 sigmas: Sequence[float], spacings: Sequence[Sequence[float]],
                 direction: int = BATCH_FORWARD, arith: int = EXACT):
        n = len(inputs)
        vols = (c_volume * max(n, 1))()
        self._keep = (list(inputs), list(outputs))
        for i, (x, y) in enumerate(zip(inputs, outputs)):
            if x.dtype != torch.float32 or y.dtype != torch.float32 or x.dim() != 3:
                raise TypeError("batched volumes are 3-D float32")
            if not (x.is_contiguous() and y.is_contiguous()) or x.shape != y.shape:
                raise ValueError("batched volumes must be contiguous and shape-matched")
            vols[i].input = x.data_ptr()
            vols[i].output = y.data_ptr()
            for d in range(3):
                vols[i].size[d] = x.shape[2 - d]
                vols[i].spacing[d] = float(spacings[i][d])
            vols[i].sigma = float(sigmas[i])
        h = ctypes.c_void_p()
        _check(lib.itkg_batch_create(n, vols, direction, arith, ctypes.byref(h)))
        self.handle = h
        self.voxels = int(lib.itkg_batch_voxels(h))
        self.device = inputs[0].device if n else torch.device("cuda")

    def run(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        _check(lib.itkg_batch_run(self.handle, ctypes.c_void_p(s.cuda_stream)))

    def autotune(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        """Pick the fastest tile z-run per radius group (writes valid outputs)."""
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        _check(lib.itkg_batch_autotune(self.handle, ctypes.c_void_p(s.cuda_stream)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.itkg_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
