"""PyTorch-facing wrappers over the C ABI (include/ksj_gpu.h).

torch supplies device memory and the current HIP stream only; every operation
runs in libksj_gpu.so.  Inputs must already be on the GPU with the documented
dtype and layout; nothing is converted or copied implicitly.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import call, load

LAYOUTS = {"coil_last": _lib.LAYOUT_COIL_LAST, "coil_major": _lib.LAYOUT_COIL_MAJOR}


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def _need_all(*specs, host: bool = False) -> None:
    """Checks caller tensors before any pointer of them reaches a kernel.  Each spec is
    (tensor, dtype, name[, shape]); every dtype, exact shape (when given) and contiguity
    is checked first, then every device (GPU, or CPU when `host`).  A wrongly sized
    buffer would otherwise become an out-of-bounds device read or write."""
    for spec in specs:
        t, dtype, name = spec[:3]
        shape = spec[3] if len(spec) > 3 else None
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"{name} must be a tensor")
        if t.dtype != dtype:
            raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
        if shape is not None and tuple(t.shape) != tuple(int(x) for x in shape):
            raise ValueError(f"{name} must have shape {tuple(int(x) for x in shape)}, got {tuple(t.shape)}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
    for spec in specs:
        if host and spec[0].is_cuda:
            raise ValueError(f"{spec[2]} must be a CPU tensor (a host-memory request keeps every buffer on the host)")
        if not host and not spec[0].is_cuda:
            raise ValueError(f"{spec[2]} must be a GPU tensor")


def _need(t: torch.Tensor, dtype: torch.dtype, name: str, shape=None) -> None:
    _need_all((t, dtype, name, shape))


def _image_shape(rows: int, cols: int, channels: int, layout: str):
    if layout not in LAYOUTS:
        raise ValueError(f"unknown layout {layout!r}")
    return (rows, cols, channels) if layout == "coil_last" else (channels, rows, cols)


def _workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def init(device: int = 0) -> None:
    call("ksj_gpu_init", device)


def tune_set(key: str, value: int) -> None:
    """Sets a process-wide launch parameter (see ksj_gpu_tune_set)."""
    call("ksj_gpu_tune_set", key.encode(), int(value))


def tune_get(key: str) -> int:
    v = C.c_int64(0)
    call("ksj_gpu_tune_get", key.encode(), C.byref(v))
    return int(v.value)


def chain_faults() -> int:
    """Launches of the uneven batch's work-queue form (multi_fused = 2) on the current
    device whose bounded waits ran out; 0 in correct operation."""
    v = C.c_uint64(0)
    call("ksj_gpu_chain_faults", C.byref(v))
    return int(v.value)


def profile_enable(on: bool) -> None:
    call("ksj_gpu_profile_enable", 1 if on else 0)


def profile_reset() -> None:
    call("ksj_gpu_profile_reset")


def profile_read(stage: str):
    """(total ms, launches) recorded for a stage since the last reset."""
    tot, cnt = C.c_double(0.0), C.c_uint64(0)
    call("ksj_gpu_profile_read", stage.encode(), C.byref(tot), C.byref(cnt))
    return float(tot.value), int(cnt.value)


def ifft2_coil_images(kspace: torch.Tensor, centred: bool = False, layout: str = "coil_last") -> torch.Tensor:
    """[C, R, Cols] complex64 channel-major k-space -> coil images (1/R/Cols scaling)."""
    if kspace.dim() != 3:
        raise ValueError("kspace must be [C, R, Cols]")
    _need(kspace, torch.complex64, "kspace")
    ch, rows, cols = kspace.shape
    shape = _image_shape(rows, cols, ch, layout)
    out = torch.empty(shape, dtype=torch.complex64, device=kspace.device)
    ws = _workspace(load().ksj_gpu_ifft2_workspace_bytes(ch, rows, cols), kspace.device)
    call("ksj_gpu_ifft2_coil_images", _ptr(kspace), _ptr(out), ch, rows, cols, int(centred), LAYOUTS[layout],
         _ptr(ws), ws.numel(), _stream())
    return out


def rss(coil_images: torch.Tensor) -> torch.Tensor:
    """[..., C] complex64 coil-last -> [...] float32 root sum of squares."""
    _need(coil_images, torch.complex64, "coil_images")
    ch = coil_images.shape[-1]
    out = torch.empty(coil_images.shape[:-1], dtype=torch.float32, device=coil_images.device)
    call("ksj_gpu_rss", _ptr(coil_images), _ptr(out), out.numel(), ch, _stream())
    return out


def sense_combine(coil_images: torch.Tensor, sensitivities: torch.Tensor) -> torch.Tensor:
    _need(coil_images, torch.complex64, "coil_images")
    _need(sensitivities, torch.complex64, "sensitivities")
    if coil_images.shape != sensitivities.shape:
        raise ValueError("coil images and sensitivities must have the same shape")
    ch = coil_images.shape[-1]
    out = torch.empty(coil_images.shape[:-1], dtype=torch.complex64, device=coil_images.device)
    call("ksj_gpu_sense_combine", _ptr(coil_images), _ptr(sensitivities), _ptr(out), out.numel(), ch, _stream())
    return out


def ramp_dcf2(traj: torch.Tensor) -> torch.Tensor:
    if traj.dim() != 2 or traj.shape[1] != 2:
        raise ValueError("traj must be [S, 2]")
    _need(traj, torch.float32, "traj")
    out = torch.empty(traj.shape[0], dtype=torch.float32, device=traj.device)
    call("ksj_gpu_ramp_dcf2", _ptr(traj), _ptr(out), traj.shape[0], _stream())
    return out


def nudft2_adjoint(traj: torch.Tensor, samples: torch.Tensor, rows: int, cols: int, row_origin: float,
                   col_origin: float, layout: str = "coil_last") -> torch.Tensor:
    """Exact adjoint NUDFT (float), samples [C, S] -> coil images."""
    if samples.dim() != 2:
        raise ValueError("samples must be [C, S]")
    ch, S = samples.shape
    _need_all((samples, torch.complex64, "samples"), (traj, torch.float32, "traj", (S, 2)))
    shape = _image_shape(rows, cols, ch, layout)
    out = torch.empty(shape, dtype=torch.complex64, device=samples.device)
    ws = _workspace(load().ksj_gpu_nudft2_adjoint_workspace_bytes(ch, rows, cols), samples.device)
    call("ksj_gpu_nudft2_adjoint", _ptr(traj), _ptr(samples), _ptr(out), ch, S, rows, cols, float(row_origin),
         float(col_origin), LAYOUTS[layout], _ptr(ws), ws.numel(), _stream())
    return out


def nudft2_forward_f64(traj: torch.Tensor, image: torch.Tensor, row_origin: float, col_origin: float) -> torch.Tensor:
    """Exact FP64 forward NUDFT: image [C, R, Cols] complex128 -> samples [C, S] complex128."""
    if image.dim() != 3 or traj.dim() != 2 or traj.shape[1] != 2:
        raise ValueError("image must be [C, R, Cols] and traj [S, 2]")
    _need(traj, torch.float32, "traj")
    _need(image, torch.complex128, "image")
    ch, rows, cols = image.shape
    S = traj.shape[0]
    out = torch.empty((ch, S), dtype=torch.complex128, device=image.device)
    call("ksj_gpu_nudft2_forward_f64", _ptr(traj), _ptr(image), _ptr(out), ch, S, rows, cols, float(row_origin),
         float(col_origin), _stream())
    return out


def golden_angle2(spokes: int, readout: int, device="cuda") -> torch.Tensor:
    out = torch.empty((spokes * readout, 2), dtype=torch.float32, device=device)
    call("ksj_gpu_golden_angle2", _ptr(out), spokes, readout, _stream())
    return out


class Plan2:
    """Device-resident gridding plan for one trajectory (see ksj_gpu_plan2_* in include/ksj_gpu.h)."""

    def __init__(self, image_rows: int, image_cols: int, samples: int, oversampling: float = 2.0,
                 kernel: str = "kaiser_bessel", width: int = 4, density: str = "ramp", centred: bool = True,
                 item_visits: int = 0, beta: float = 0.0, row_origin: float | None = None,
                 col_origin: float | None = None, grid_rows: int | None = None, grid_cols: int | None = None):
        d = _lib.Plan2Desc()
        d.struct_size = C.sizeof(_lib.Plan2Desc)
        d.image_rows, d.image_cols = image_rows, image_cols
        d.grid_rows = grid_rows if grid_rows is not None else int(round(image_rows * oversampling))
        d.grid_cols = grid_cols if grid_cols is not None else int(round(image_cols * oversampling))
        d.kernel = {"linear": _lib.KERNEL_LINEAR, "kaiser_bessel": _lib.KERNEL_KAISER_BESSEL}[kernel]
        d.kernel_width = width
        d.density = {"none": _lib.DENSITY_NONE, "ramp": _lib.DENSITY_RAMP, "given": _lib.DENSITY_GIVEN}[density]
        d.centred = int(centred)
        d.samples = samples
        d.item_visits = item_visits
        d.beta = beta
        d.row_origin = (image_rows - 1) / 2.0 if row_origin is None else row_origin
        d.col_origin = (image_cols - 1) / 2.0 if col_origin is None else col_origin
        self.desc = d
        self._handle = C.c_void_p()
        call("ksj_gpu_plan2_create", C.byref(d), C.byref(self._handle))
        self._lib = load()

    def set_trajectory(self, traj: torch.Tensor, dcf: torch.Tensor | None = None) -> "Plan2":
        _need(traj, torch.float32, "traj")
        if traj.shape != (self.desc.samples, 2):
            raise ValueError("trajectory shape does not match the plan")
        call("ksj_gpu_plan2_set_trajectory", self._handle, _ptr(traj), 0 if dcf is None else _ptr(dcf), _stream())
        return self

    def info(self) -> dict:
        i = _lib.Plan2Info()
        i.struct_size = C.sizeof(_lib.Plan2Info)
        call("ksj_gpu_plan2_get_info", self._handle, C.byref(i))
        return {"items": i.items, "visits": i.visits, "device_bytes": i.device_bytes, "beta": i.beta}

    def export_bins(self, device="cuda"):
        nb = self.desc.grid_rows * self.desc.grid_cols
        start = torch.empty(nb + 1, dtype=torch.int32, device=device)
        perm = torch.empty(self.desc.samples, dtype=torch.int32, device=device)
        call("ksj_gpu_plan2_export_bins", self._handle, _ptr(start), _ptr(perm), _stream())
        return start, perm

    def workspace_bytes(self, channels: int) -> int:
        return int(self._lib.ksj_gpu_adjoint2_workspace_bytes(self._handle, channels))

    def adjoint(self, kspace: torch.Tensor, images: bool = True, rss: bool = False, layout: str = "coil_last",
                workspace: torch.Tensor | None = None):
        if kspace.dim() != 2 or kspace.shape[1] != self.desc.samples:
            raise ValueError("kspace must be [C, S] with the plan's sample count")
        _need(kspace, torch.complex64, "kspace")
        ch, S = kspace.shape
        r, c = self.desc.image_rows, self.desc.image_cols
        _image_shape(r, c, ch, layout)
        img = None
        if images:
            img = torch.empty((r, c, ch) if layout == "coil_last" else (ch, r, c), dtype=torch.complex64,
                              device=kspace.device)
        out_rss = torch.empty((r, c), dtype=torch.float32, device=kspace.device) if rss else None
        need = self.workspace_bytes(ch)
        if workspace is None or workspace.numel() < need:
            workspace = _workspace(need, kspace.device)
        call("ksj_gpu_adjoint2_execute", self._handle, _ptr(kspace), ch, 0 if img is None else _ptr(img),
             LAYOUTS[layout], 0 if out_rss is None else _ptr(out_rss), _ptr(workspace), workspace.numel(), _stream())
        return img, out_rss

    def batch_workspace_bytes(self, channels: int, scans: int) -> int:
        return int(self._lib.ksj_gpu_adjoint2_batch_workspace_bytes(self._handle, channels, scans))

    def adjoint_batch(self, kspaces, images: bool = False, rss: bool = True, layout: str = "coil_last",
                      workspace: torch.Tensor | None = None):
        """Scans sharing this plan's trajectory, stacked along the coil axis (up to 64 coils per pass)."""
        n = len(kspaces)
        ch = kspaces[0].shape[0]
        for k in kspaces:
            _need(k, torch.complex64, "kspace")
            if tuple(k.shape) != (ch, self.desc.samples):
                raise ValueError("every scan needs the same [C, S] shape")
        r, c = self.desc.image_rows, self.desc.image_cols
        dev = kspaces[0].device
        imgs = [torch.empty((r, c, ch) if layout == "coil_last" else (ch, r, c), dtype=torch.complex64, device=dev)
                for _ in range(n)] if images else None
        rsss = [torch.empty((r, c), dtype=torch.float32, device=dev) for _ in range(n)] if rss else None
        need = self.batch_workspace_bytes(ch, n)
        if workspace is None or workspace.numel() < need:
            workspace = _workspace(need, dev)
        P = C.c_void_p * n
        kp = P(*[_ptr(k) for k in kspaces])
        ip = P(*[_ptr(t) for t in imgs]) if imgs else None
        rp = P(*[_ptr(t) for t in rsss]) if rsss else None
        call("ksj_gpu_adjoint2_execute_batch", self._handle, n, kp, ch, ip, LAYOUTS[layout], rp, _ptr(workspace),
             workspace.numel(), _stream())
        return imgs, rsss

    def close(self) -> None:
        if getattr(self, "_handle", None) and self._handle.value:
            self._lib.ksj_gpu_plan2_destroy(self._handle)
            self._handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server (ksj_gpu_serve_*): submit scans, wait for tickets.
    A 2-D request whose tensors are all on the CPU is a host-memory request (staged
    through the server's pinned staging of `host_staging_mib` MiB, 0 = 1024; submit
    blocks while it is full); otherwise every tensor must be on the GPU.  Buffers passed
    to submit() must stay alive until their ticket completes; the server keeps
    references to them until then."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, lanes: int = 2, max_batch_scans: int = 8,
                 plan_cache_entries: int = 16, host_staging_mib: int = 0):
        self._lib = load()
        cfg = _lib.ServeConfig(C.sizeof(_lib.ServeConfig), device, hbm_budget_bytes, lanes, max_batch_scans,
                               plan_cache_entries, host_staging_mib)
        self._handle = C.c_void_p()
        call("ksj_gpu_serve_create", C.byref(cfg), C.byref(self._handle))
        self._live = {}

    def submit_radial(self, kspace: torch.Tensor, traj: torch.Tensor, n: int, images: torch.Tensor | None = None,
                      rss: torch.Tensor | None = None, width: int = 4, density: str = "ramp",
                      trajectory_key: int = 0) -> int:
        if kspace.dim() != 2:
            raise ValueError("kspace must be [C, S]")
        ch, S = kspace.shape
        host = not kspace.is_cuda
        self._check_request((kspace, torch.complex64, "kspace"), (traj, torch.float32, "traj", (S, 2)),
                            images, (n, n, ch), rss, (n, n), host=host)
        r = _lib.ServeRequest(C.sizeof(_lib.ServeRequest), _lib.SERVE_RADIAL2D, n, n, ch, S, width,
                              _lib.DENSITY_RAMP if density == "ramp" else _lib.DENSITY_NONE, 0, 0, trajectory_key,
                              _ptr(traj), _ptr(kspace), _ptr(images) if images is not None else None,
                              _ptr(rss) if rss is not None else None,
                              _lib.SERVE_HOST_MEMORY if host else _lib.SERVE_DEVICE_MEMORY, 0)
        return self._submit(r, (kspace, traj, images, rss), host)

    def submit_radial3d(self, kspace: torch.Tensor, traj: torch.Tensor, shape, images: torch.Tensor | None = None,
                        rss: torch.Tensor | None = None, width: int = 4, density: str = "ramp_squared",
                        trajectory_key: int = 0) -> int:
        """A 3-D volume: kspace [C, S], traj [S, 3] (z, y, x) rad/px, shape (N_z, N_y, N_x);
        images coil-last [N_z, N_y, N_x, C] and/or rss [N_z, N_y, N_x]."""
        if kspace.dim() != 2:
            raise ValueError("kspace must be [C, S]")
        ch, S = kspace.shape
        nz, ny, nx = shape
        self._check_request((kspace, torch.complex64, "kspace"), (traj, torch.float32, "traj", (S, 3)),
                            images, (nz, ny, nx, ch), rss, (nz, ny, nx))
        dens = _lib.DENSITY_RAMP_SQUARED if density == "ramp_squared" else _lib.DENSITY_NONE
        r = _lib.ServeRequest(C.sizeof(_lib.ServeRequest), _lib.SERVE_RADIAL3D, ny, nx, ch, S, width, dens, 0, nz,
                              trajectory_key, _ptr(traj), _ptr(kspace), _ptr(images) if images is not None else None,
                              _ptr(rss) if rss is not None else None)
        return self._submit(r, (kspace, traj, images, rss), False)

    def submit_cartesian(self, kspace: torch.Tensor, images: torch.Tensor | None = None,
                         rss: torch.Tensor | None = None, centred: bool = False) -> int:
        if kspace.dim() != 3:
            raise ValueError("kspace must be [C, R, Cols]")
        ch, rows, cols = kspace.shape
        host = not kspace.is_cuda
        self._check_request((kspace, torch.complex64, "kspace"), None, images, (rows, cols, ch), rss, (rows, cols),
                            host=host)
        r = _lib.ServeRequest(C.sizeof(_lib.ServeRequest), _lib.SERVE_CARTESIAN2D, rows, cols, ch, 0, 0, 0,
                              int(centred), 0, 0, None, _ptr(kspace), _ptr(images) if images is not None else None,
                              _ptr(rss) if rss is not None else None,
                              _lib.SERVE_HOST_MEMORY if host else _lib.SERVE_DEVICE_MEMORY, 0)
        return self._submit(r, (kspace, images, rss), host)

    @staticmethod
    def _check_request(kspace_spec, traj_spec, images, image_shape, rss, rss_shape, host: bool = False) -> None:
        if images is None and rss is None:
            raise ValueError("a request needs images and/or rss")
        specs = [kspace_spec] + ([traj_spec] if traj_spec is not None else [])
        if images is not None:
            specs.append((images, torch.complex64, "images", image_shape))
        if rss is not None:
            specs.append((rss, torch.float32, "rss", rss_shape))
        _need_all(*specs, host=host)

    def _submit(self, req, keep, host: bool) -> int:
        # Work the caller queued on torch's stream (e.g. filling k-space) must land first.
        if not host:
            torch.cuda.current_stream().synchronize()
        t = C.c_uint64(0)
        call("ksj_gpu_serve_submit", self._handle, C.byref(req), C.byref(t))
        self._live[t.value] = keep
        return t.value

    def wait(self, ticket: int) -> float:
        """Blocks until the ticket completes; returns its latency in ms (raises on a failed request)."""
        lat = C.c_double(0)
        rc = self._lib.ksj_gpu_serve_wait(self._handle, ticket, C.byref(lat))
        self._live.pop(ticket, None)
        _lib.check(rc)
        return float(lat.value)

    def drain(self) -> None:
        call("ksj_gpu_serve_drain", self._handle)

    def hold(self, on: bool) -> None:
        """While held the server launches nothing new; queued requests batch on release."""
        call("ksj_gpu_serve_hold", self._handle, int(bool(on)))

    def stats(self) -> dict:
        s = _lib.ServeStats()
        s.struct_size = C.sizeof(_lib.ServeStats)
        call("ksj_gpu_serve_get_stats", self._handle, C.byref(s))
        return {name: int(getattr(s, name)) for name, _ in _lib.ServeStats._fields_[2:]}

    def close(self) -> None:
        if getattr(self, "_handle", None) and self._handle.value:
            self._lib.ksj_gpu_serve_destroy(self._handle)
            self._handle = C.c_void_p()
            self._live.clear()

    def __del__(self):
        self.close()


TRAJ_UNITS = {"unspecified": 0, "cycles_per_fov": 1, "radians_per_pixel": 2, "encoded_matrix_index": 3}


def normalize_trajectory2(raw: torch.Tensor, units: str, encoded_rows: int = 0, encoded_cols: int = 0) -> torch.Tensor:
    """Raw [S, 2] (kx, ky) in `units` -> [S, 2] (ky, kx) radians per pixel (ksj_gpu_normalize_trajectory2)."""
    _need(raw, torch.float32, "raw trajectory")
    out = torch.empty_like(raw)
    call("ksj_gpu_normalize_trajectory2", _ptr(raw), raw.shape[0], TRAJ_UNITS[units], encoded_rows, encoded_cols,
         _ptr(out), _stream())
    return out


def assemble_acquisitions(packed: torch.Tensor, samples_per_acquisition, channels: int) -> torch.Tensor:
    """ISMRMRD acquisitions packed one after another, each [channels][ns_a] complex64 (flat), -> the
    route's k-space frame [channels][S] (ksj_gpu_assemble_acquisitions)."""
    counts = [int(n) for n in samples_per_acquisition]
    if not counts or min(counts) < 1 or max(counts) > 0xFFFFFFFF:
        raise ValueError("every acquisition needs between 1 and 2^32 - 1 samples")
    if not 1 <= channels <= 64:
        raise ValueError("acquisitions must have [1, 64] active channels")
    total = sum(counts)
    _need(packed, torch.complex64, "acquisitions", (channels * total,))
    out = torch.empty((channels, total), dtype=torch.complex64, device=packed.device)
    ws = _workspace(load().ksj_gpu_assemble_workspace_bytes(len(counts)), packed.device)
    host_counts = (C.c_uint32 * len(counts))(*counts)
    call("ksj_gpu_assemble_acquisitions", _ptr(packed), C.addressof(host_counts), len(counts), channels, _ptr(out),
         _ptr(ws), ws.numel(), _stream())
    return out


TRAJ_DIFFERS, TRAJ_INVALID = 1, 2


def trajectory_check(traj: torch.Tensor, resident: "torch.Tensor | None", bound: float) -> int:
    """ksj_gpu_trajectory_check: TRAJ_DIFFERS when `traj` differs bitwise from `resident`
    (always when it is None), TRAJ_INVALID when a value is not finite or outside
    [-bound, bound]."""
    _need(traj, torch.float32, "trajectory")
    if resident is not None:
        _need(resident, torch.float32, "resident trajectory")
        if resident.numel() != traj.numel():
            raise ValueError("resident trajectory size differs")
    flags = torch.empty(1, dtype=torch.int32, device=traj.device)
    call("ksj_gpu_trajectory_check", _ptr(traj), _ptr(resident) if resident is not None else None, traj.numel(),
         bound, _ptr(flags), _stream())
    return int(flags.item())


def channel_mix(matrix: torch.Tensor, kspace: torch.Tensor) -> torch.Tensor:
    """complex64 matrix [out, in] @ channel-major k-space [in, ...] (ksj_gpu_channel_mix):
    noise prewhitening with a whitening matrix, coil compression with a basis."""
    if matrix.dim() != 2 or kspace.dim() < 2:
        raise ValueError("matrix must be [out, in] and k-space [in, ...]")
    _need(matrix, torch.complex64, "matrix")
    _need(kspace, torch.complex64, "k-space")
    out_ch, in_ch = matrix.shape
    if kspace.shape[0] != in_ch:
        raise ValueError("k-space channels differ from the matrix columns")
    out = torch.empty((out_ch,) + tuple(kspace.shape[1:]), dtype=torch.complex64, device=kspace.device)
    call("ksj_gpu_channel_mix", _ptr(matrix), _ptr(kspace), _ptr(out), out_ch, in_ch, kspace[0].numel(), _stream())
    return out


def phase_correct(kspace: torch.Tensor, phase: torch.Tensor) -> torch.Tensor:
    """complex64 k-space [C, R, K] times phase [C, K] per line (ksj_gpu_phase_correct)."""
    if kspace.dim() != 3:
        raise ValueError("k-space must be [C, R, K]")
    c, r, k = kspace.shape
    _need_all((kspace, torch.complex64, "k-space"), (phase, torch.complex64, "phase", (c, k)))
    out = torch.empty_like(kspace)
    call("ksj_gpu_phase_correct", _ptr(kspace), _ptr(phase), _ptr(out), c, r, k, _stream())
    return out


def readout_crop(kspace: torch.Tensor, out_cols: int, offset: int) -> torch.Tensor:
    """complex64 k-space [..., in_cols] -> [..., offset:offset + out_cols] (ksj_gpu_readout_crop)."""
    in_cols = kspace.shape[-1]
    if out_cols < 0 or offset < 0 or offset + out_cols > in_cols:
        raise ValueError("crop window must lie inside the readout")
    _need(kspace, torch.complex64, "k-space")
    out = torch.empty(tuple(kspace.shape[:-1]) + (out_cols,), dtype=torch.complex64, device=kspace.device)
    call("ksj_gpu_readout_crop", _ptr(kspace), _ptr(out), kspace.numel() // in_cols, in_cols, out_cols, offset,
         _stream())
    return out


def golden_means3(spokes: int, readout: int, device="cuda") -> torch.Tensor:
    """Golden-means kooshball [spokes * readout, 3] (z, y, x) in radians per pixel."""
    out = torch.empty((spokes * readout, 3), dtype=torch.float32, device=device)
    call("ksj_gpu_golden_means3", _ptr(out), spokes, readout, _stream())
    return out


class Plan3:
    """Device-resident 3-D gridding plan for one trajectory (ksj_gpu_plan3_* in include/ksj_gpu.h).
    shape = (N_z, N_y, N_x); trajectories are [S, 3] float32 (z, y, x) radians per pixel."""

    def __init__(self, shape, samples: int, oversampling: float = 2.0, width: int = 4, density: str = "ramp_squared",
                 item_visits: int = 0, beta: float = 0.0, origin=None, grid=None):
        d = _lib.Plan3Desc()
        d.struct_size = C.sizeof(_lib.Plan3Desc)
        d.image_z, d.image_y, d.image_x = shape
        g = grid if grid is not None else tuple(int(round(n * oversampling)) for n in shape)
        d.grid_z, d.grid_y, d.grid_x = g
        d.kernel_width = width
        d.density = {"none": _lib.DENSITY_NONE, "ramp_squared": _lib.DENSITY_RAMP_SQUARED,
                     "given": _lib.DENSITY_GIVEN}[density]
        d.samples = samples
        d.item_visits = item_visits
        d.beta = beta
        o = tuple((n - 1) / 2.0 for n in shape) if origin is None else origin
        d.origin_z, d.origin_y, d.origin_x = o
        self.desc = d
        self.shape = tuple(shape)
        self._handle = C.c_void_p()
        call("ksj_gpu_plan3_create", C.byref(d), C.byref(self._handle))
        self._lib = load()

    def set_trajectory(self, traj: torch.Tensor, dcf: torch.Tensor | None = None) -> "Plan3":
        _need(traj, torch.float32, "traj")
        if traj.shape != (self.desc.samples, 3):
            raise ValueError("trajectory shape does not match the plan")
        call("ksj_gpu_plan3_set_trajectory", self._handle, _ptr(traj), 0 if dcf is None else _ptr(dcf), _stream())
        return self

    def info(self) -> dict:
        i = _lib.Plan3Info()
        i.struct_size = C.sizeof(_lib.Plan3Info)
        call("ksj_gpu_plan3_get_info", self._handle, C.byref(i))
        return {"items": i.items, "visits": i.visits, "device_bytes": i.device_bytes, "beta": i.beta}

    def export_bins(self, device="cuda"):
        nb = self.desc.grid_z * self.desc.grid_y * self.desc.grid_x
        start = torch.empty(nb + 1, dtype=torch.int32, device=device)
        perm = torch.empty(self.desc.samples, dtype=torch.int32, device=device)
        call("ksj_gpu_plan3_export_bins", self._handle, _ptr(start), _ptr(perm), _stream())
        return start, perm

    def workspace_bytes(self, channels: int, coils_per_pass: int = 16) -> int:
        return int(self._lib.ksj_gpu_adjoint3_workspace_bytes(self._handle, channels, coils_per_pass))

    def adjoint(self, kspace: torch.Tensor, images: bool = True, rss: bool = False, coils_per_pass: int = 16,
                workspace: torch.Tensor | None = None):
        """[C, S] complex64 -> (coil-last [N_z, N_y, N_x, C] complex64 or None, RSS [N_z, N_y, N_x] or None)."""
        if kspace.dim() != 2 or kspace.shape[1] != self.desc.samples:
            raise ValueError("kspace must be [C, S] with the plan's sample count")
        _need(kspace, torch.complex64, "kspace")
        ch, S = kspace.shape
        img = torch.empty((*self.shape, ch), dtype=torch.complex64, device=kspace.device) if images else None
        out_rss = torch.empty(self.shape, dtype=torch.float32, device=kspace.device) if rss else None
        need = self.workspace_bytes(ch, coils_per_pass)
        if workspace is None or workspace.numel() < need:
            workspace = _workspace(need, kspace.device)
        call("ksj_gpu_adjoint3_execute", self._handle, _ptr(kspace), ch, coils_per_pass,
             0 if img is None else _ptr(img), 0 if out_rss is None else _ptr(out_rss), _ptr(workspace),
             workspace.numel(), _stream())
        return img, out_rss

    def close(self) -> None:
        if getattr(self, "_handle", None) and self._handle.value:
            self._lib.ksj_gpu_plan3_destroy(self._handle)
            self._handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def adjoint_multi_workspace_bytes(plans, channels) -> int:
    n = len(plans)
    P = C.c_void_p * n
    U = C.c_uint32 * n
    return int(load().ksj_gpu_adjoint2_multi_workspace_bytes(n, P(*[p._handle.value for p in plans]),
                                                             U(*channels)))


def adjoint_multi(plans, kspaces, rss=None, images=None, workspace: torch.Tensor | None = None,
                  layout: str = "coil_last") -> None:
    """Uneven batch (ksj_gpu_adjoint2_execute_multi): item i = plans[i] over kspaces[i] ([C_i, S_i])
    into rss[i] ([N, N] float32) and/or images[i] (coil-last).  Outputs and the workspace are
    supplied by the caller, so the call allocates nothing and can be captured into a Graph."""
    n = len(plans)
    if len(kspaces) != n:
        raise ValueError("one k-space per plan")
    for name, outs in (("images", images), ("rss", rss)):
        if outs is not None and len(outs) != n:
            raise ValueError(f"{name} needs one entry per plan")
    specs = []
    for i, (p, k) in enumerate(zip(plans, kspaces)):
        if k.dim() != 2:
            raise ValueError(f"kspace {i} must be [C, S]")
        specs.append((k, torch.complex64, f"kspace {i}", (k.shape[0], p.desc.samples)))
        rows, cols, ch = p.desc.image_rows, p.desc.image_cols, int(k.shape[0])
        img = images[i] if images is not None else None
        out = rss[i] if rss is not None else None
        if img is None and out is None:
            raise ValueError(f"item {i} needs images and/or rss")
        if img is not None:
            specs.append((img, torch.complex64, f"images {i}", _image_shape(rows, cols, ch, layout)))
        if out is not None:
            specs.append((out, torch.float32, f"rss {i}", (rows, cols)))
    _need_all(*specs)
    channels = [int(k.shape[0]) for k in kspaces]
    need = adjoint_multi_workspace_bytes(plans, channels)
    if workspace is None or workspace.numel() < need:
        raise ValueError(f"adjoint_multi needs a workspace of {need} bytes")
    P = C.c_void_p * n
    U = C.c_uint32 * n
    ip = P(*[_ptr(t) if t is not None else None for t in images]) if images is not None else None
    rp = P(*[_ptr(t) if t is not None else None for t in rss]) if rss is not None else None
    call("ksj_gpu_adjoint2_execute_multi", n, P(*[p._handle.value for p in plans]), P(*[_ptr(k) for k in kspaces]),
         U(*channels), ip, LAYOUTS[layout], rp, _ptr(workspace), workspace.numel(), _stream())


class Graph:
    """HIP graph of the work enqueued on a torch stream inside `with Graph(stream) as g:`
    (ksj_gpu_graph_*).  Nothing may allocate while capturing.  g.replay() relaunches it
    on the same stream."""

    def __init__(self, stream: torch.cuda.Stream):
        if stream.cuda_stream == 0:
            raise ValueError("capture needs a stream other than the null stream")
        self.stream = stream
        self._lib = load()
        self._handle = C.c_void_p()
        self._ctx = None

    def __enter__(self):
        self._ctx = torch.cuda.stream(self.stream)
        self._ctx.__enter__()
        call("ksj_gpu_graph_begin", self.stream.cuda_stream)
        return self

    def __exit__(self, *exc):
        try:
            call("ksj_gpu_graph_end", self.stream.cuda_stream, C.byref(self._handle))
        finally:
            self._ctx.__exit__(*exc)
        return False

    def replay(self) -> None:
        call("ksj_gpu_graph_launch", self._handle, self.stream.cuda_stream)

    def close(self) -> None:
        if self._handle.value:
            self._lib.ksj_gpu_graph_destroy(self._handle)
            self._handle = C.c_void_p()

    def __del__(self):
        self.close()


# ---------------------------------------------------------------- calibration

def _calibration_workspace(channels: int, samples: int, device) -> torch.Tensor:
    return _workspace(load().ksj_gpu_calibration_workspace_bytes(channels, samples), device)


def channel_covariance(samples: torch.Tensor, centred: bool) -> torch.Tensor:
    """[C, S] complex64 -> [C, C] complex128 covariance (ksj_gpu_channel_covariance)."""
    if samples.dim() != 2:
        raise ValueError("samples must be [C, S]")
    _need(samples, torch.complex64, "samples")
    C, S = samples.shape
    out = torch.empty((C, C), dtype=torch.complex128, device=samples.device)
    ws = _calibration_workspace(C, S, samples.device)
    call("ksj_gpu_channel_covariance", _ptr(samples), C, S, int(bool(centred)), _ptr(out), _ptr(ws), ws.numel(),
         _stream())
    return out


def hermitian_eigen(matrix: torch.Tensor):
    """Hermitian [n, n] complex128 -> (eigenvalues ascending float64 [n], eigenvectors in columns
    complex128 [n, n]); raises when the decomposition fails (ksj_gpu_hermitian_eigen)."""
    if matrix.dim() != 2 or matrix.shape[0] != matrix.shape[1]:
        raise ValueError("matrix must be square")
    _need(matrix, torch.complex128, "matrix")
    n = matrix.shape[0]
    values = torch.empty(n, dtype=torch.float64, device=matrix.device)
    vectors = torch.empty((n, n), dtype=torch.complex128, device=matrix.device)
    status = torch.empty(1, dtype=torch.int32, device=matrix.device)
    call("ksj_gpu_hermitian_eigen", _ptr(matrix), n, _ptr(values), _ptr(vectors), _ptr(status), _stream())
    code = int(status.item())
    if code != 0:
        raise RuntimeError("eigen-decomposition failed: " + ("non-finite input" if code == 1 else "no convergence"))
    return values, vectors


def noise_model_estimate(samples: torch.Tensor, eigenvalue_floor: float = 1.0e-6) -> torch.Tensor:
    """Noise-only samples [C, S] complex64 -> whitening matrix [C, C] complex64 (ksj_gpu_noise_model_estimate)."""
    if samples.dim() != 2:
        raise ValueError("samples must be [C, S]")
    _need(samples, torch.complex64, "samples")
    C, S = samples.shape
    out = torch.empty((C, C), dtype=torch.complex64, device=samples.device)
    ws = _calibration_workspace(C, S, samples.device)
    call("ksj_gpu_noise_model_estimate", _ptr(samples), C, S, float(eigenvalue_floor), _ptr(out), _ptr(ws),
         ws.numel(), _stream())
    return out


def coil_compression_basis_estimate(kspace: torch.Tensor, virtual_channels: int) -> torch.Tensor:
    """Channel-major k-space [P, ...] complex64 -> basis [V, P] complex64
    (ksj_gpu_coil_compression_basis_estimate)."""
    if kspace.dim() < 2:
        raise ValueError("k-space must be [P, ...]")
    _need(kspace, torch.complex64, "k-space")
    P = kspace.shape[0]
    S = kspace[0].numel()
    if not 1 <= virtual_channels <= P:
        raise ValueError("virtual channels must be in [1, physical]")
    out = torch.empty((virtual_channels, P), dtype=torch.complex64, device=kspace.device)
    ws = _calibration_workspace(P, S, kspace.device)
    call("ksj_gpu_coil_compression_basis_estimate", _ptr(kspace), P, virtual_channels, S, _ptr(out), _ptr(ws),
         ws.numel(), _stream())
    return out


def phase_correction_estimate(reference: torch.Tensor) -> torch.Tensor:
    """Phase-reference frame [C, lines, K] complex64 -> phase model [C, K] complex64
    (ksj_gpu_phase_correction_estimate)."""
    if reference.dim() != 3:
        raise ValueError("reference must be [C, lines, K]")
    _need(reference, torch.complex64, "reference")
    C, L, K = reference.shape
    out = torch.empty((C, K), dtype=torch.complex64, device=reference.device)
    call("ksj_gpu_phase_correction_estimate", _ptr(reference), C, L, K, _ptr(out), _stream())
    return out


def sensitivity_maps_walsh(coil_images: torch.Tensor, window: int = 5, iterations: int = 8) -> torch.Tensor:
    """Coil-last coil images [rows, cols, C] complex64 -> sensitivity maps of the same shape
    (Walsh's adaptive method; see ksj_gpu_sensitivity_maps_walsh)."""
    _need(coil_images, torch.complex64, "coil_images")
    if coil_images.dim() != 3:
        raise ValueError("coil_images must be [rows, cols, channels]")
    rows, cols, c = coil_images.shape
    maps = torch.empty_like(coil_images)
    call("ksj_gpu_sensitivity_maps_walsh", _ptr(coil_images), rows, cols, c, int(window), int(iterations), _ptr(maps),
         _stream())
    return maps


def nufft3_forward_f64(image: torch.Tensor, traj: torch.Tensor, origin=None, width: int = 6) -> torch.Tensor:
    """FP64 Kaiser-Bessel forward NUFFT of a volume [Nz, Ny, Nx] complex128 at a trajectory
    [S, 3] float32 (z, y, x) rad/px: [S] complex128 ~ sum_v x_v exp(-i w . (v - origin)),
    origin (N - 1) / 2 per axis by default (see ksj_gpu_nufft3_forward_f64)."""
    if image.dim() != 3:
        raise ValueError("image must be [Nz, Ny, Nx]")
    _need(image, torch.complex128, "image")
    if traj.dim() != 2 or traj.shape[1] != 3:
        raise ValueError("traj must be [S, 3]")
    _need(traj, torch.float32, "traj")
    nz, ny, nx = image.shape
    if origin is None:
        origin = ((nz - 1) / 2, (ny - 1) / 2, (nx - 1) / 2)
    dims = (C.c_uint32 * 3)(nz, ny, nx)
    org = (C.c_double * 3)(*[float(o) for o in origin])
    out = torch.empty(traj.shape[0], dtype=torch.complex128, device=image.device)
    ws = _workspace(load().ksj_gpu_nufft3_forward_f64_workspace_bytes(dims), image.device)
    call("ksj_gpu_nufft3_forward_f64", _ptr(image), dims, org, _ptr(traj), traj.shape[0], int(width), _ptr(out),
         _ptr(ws), ws.numel(), _stream())
    return out
