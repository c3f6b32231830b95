"""Batched PR76 PT flash on the GPU (device tensors in, device tensors out).

Mirrors the reference's backend boundary: ``PtFlashBackend::solve`` of one
``PtFlashRequest{pressure_pa, temperature_k, feed}``
(modules/flash/include/mpmc/flash/pt_flash_backend.hpp:111-195), applied to every
cell of a batch. Results come back in the library's structure-of-arrays layout.

torch provides device memory and streams only; the arithmetic is the HIP library.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import native

# status word decoding (include/mpmc_hnu_amd.h)
SET_ACCEPTED, SET_PHASE_SET_UNSTABLE, SET_INDETERMINATE = 0, 1, 2
MAX3_NAMES = ("single_phase", "two_phase", "three_phase", "phase_boundary_unresolved",
              "higher_phase_count_or_wrong_candidate", "indeterminate")
FLAG_REJECTED = 0x2000


def _torch():
    import torch

    return torch


class Mixture:
    """A prepared PR76 mixture resident on the current device (PrParameterSet +
    Pr76Phase::from_parameters, pr_parameters.hpp:53-154)."""

    def __init__(self, tc, pc, omega, kij):
        lib = native.library()
        tc = np.ascontiguousarray(tc, dtype=np.float64)
        pc = np.ascontiguousarray(pc, dtype=np.float64)
        om = np.ascontiguousarray(omega, dtype=np.float64)
        kij = np.ascontiguousarray(kij, dtype=np.float64)
        nc = tc.size
        if pc.size != nc or om.size != nc or kij.size != nc * nc:
            raise ValueError("Mixture: Tc, Pc, omega must have nc entries and kij nc*nc")
        self.nc = nc
        self.params = {"tc": tc, "pc": pc, "omega": om, "kij": kij.reshape(nc, nc)}
        h = ctypes.c_void_p()
        P = ctypes.c_void_p
        native.check(lib.mpmc_mixture_create(nc, tc.ctypes.data_as(P), pc.ctypes.data_as(P), om.ctypes.data_as(P),
                                             kij.ctypes.data_as(P), ctypes.byref(h)), "mixture_create")
        self._h = h

    @classmethod
    def from_dict(cls, d):
        return cls(d["tc"], d["pc"], d["omega"], d["kij"])

    @property
    def handle(self):
        if self._h is None:
            raise native.MpmcError("mixture was destroyed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            native.library().mpmc_mixture_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream_ptr(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _dev(a, dtype, device):
    torch = _torch()
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)


def allocate_outputs(nc: int, ncell: int, device, counters: bool = True) -> dict:
    torch = _torch()
    f64, i32 = torch.float64, torch.int32
    out = {
        "status": torch.empty(ncell, dtype=i32, device=device),
        "phase_count": torch.empty(ncell, dtype=i32, device=device),
        "beta": torch.empty((3, ncell), dtype=f64, device=device),
        "zfactor": torch.empty((3, ncell), dtype=f64, device=device),
        "branch": torch.empty((3, ncell), dtype=i32, device=device),
        "x": torch.empty((3, nc, ncell), dtype=f64, device=device),
        "ln_phi": torch.empty((3, nc, ncell), dtype=f64, device=device),
    }
    out["counters"] = torch.empty((4, ncell), dtype=torch.int32, device=device) if counters else None
    return out


def flash_batch(mixture: Mixture, p, t, z, fugacity_tolerance: float = 1e-11, max_phases: int = 3,
                precision: int = 0, stream=None, out: dict | None = None, counters: bool = True,
                schedule: int = 0, workspace=None, hints=None, final_hints=None) -> dict:
    """Flash every cell (p[ncell] Pa, t[ncell] K, z[nc, ncell]); returns device tensors.

    ``hints`` ([nhint, nc, ncell], nhint <= 6; NaN in component 0 marks an absent hint)
    are warm-start compositions appended to each cell's initial stability starts
    (mpmc_flash_batch_hinted); ``final_hints`` (same layout) are extra starts of the
    common-tangent review (mpmc_flash_batch_warm).

    schedule 0 runs the staged batch (one lane per stability trial / split attempt),
    schedule 1 one lane per cell. ``workspace`` (a uint8 device tensor of at least
    workspace_bytes(nc, ncell)) makes the call use caller-owned scratch; otherwise the
    mixture handle's own workspace is used. Asynchronous on ``stream`` (default:
    torch's current stream)."""
    torch = _torch()
    lib = native.library()
    device = torch.device("cuda", torch.cuda.current_device())
    p = _dev(p, torch.float64, device)
    t = _dev(t, torch.float64, device)
    z = _dev(z, torch.float64, device)
    if z.dim() != 2 or z.shape[0] != mixture.nc or z.shape[1] != p.numel() or t.numel() != p.numel():
        raise ValueError("flash_batch: expected p[ncell], t[ncell], z[nc, ncell]")
    n = p.numel()
    if out is None:
        out = allocate_outputs(mixture.nc, n, device, counters)
    opts = native.FlashOptions()
    lib.mpmc_flash_default_options(ctypes.byref(opts))
    opts.fugacity_tolerance = float(fugacity_tolerance)
    opts.max_phases = int(max_phases)
    opts.precision = int(precision)
    opts.schedule = int(schedule)
    o = native.FlashOutputs(
        out["status"].data_ptr(), out["phase_count"].data_ptr(), out["beta"].data_ptr(), out["zfactor"].data_ptr(),
        out["branch"].data_ptr(), out["x"].data_ptr(), out["ln_phi"].data_ptr(),
        out["counters"].data_ptr() if out.get("counters") is not None else None)
    V = ctypes.c_void_p
    if final_hints is not None:
        hs = []
        for h in (hints, final_hints):
            if h is None:
                hs.append(None)
                continue
            h = _dev(h, torch.float64, device)
            if h.dim() != 3 or h.shape[1:] != (mixture.nc, n) or h.shape[0] > 6:
                raise ValueError("flash_batch: expected hints[nhint <= 6, nc, ncell]")
            hs.append(h)
        hints, final_hints = hs
        ws, wsb = (V(workspace.data_ptr()), workspace.numel() * workspace.element_size()) if workspace is not None \
            else (None, 0)
        nh = 0 if hints is None else int(hints.shape[0])
        rc = lib.mpmc_flash_batch_warm(mixture.handle, n, V(p.data_ptr()), V(t.data_ptr()), V(z.data_ptr()), nh,
                                       V(hints.data_ptr()) if nh else None, int(final_hints.shape[0]),
                                       V(final_hints.data_ptr()), ctypes.byref(opts), ctypes.byref(o), ws, wsb,
                                       _stream_ptr(stream))
        native.check(rc, "flash_batch_warm")
        out["_inputs"] = (p, t, z, hints, final_hints)
        return out
    if hints is not None:
        hints = _dev(hints, torch.float64, device)
        if hints.dim() != 3 or hints.shape[1:] != (mixture.nc, n) or hints.shape[0] > 6:
            raise ValueError("flash_batch: expected hints[nhint <= 6, nc, ncell]")
        ws, wsb = (V(workspace.data_ptr()), workspace.numel() * workspace.element_size()) if workspace is not None \
            else (None, 0)
        rc = lib.mpmc_flash_batch_hinted(mixture.handle, n, V(p.data_ptr()), V(t.data_ptr()), V(z.data_ptr()),
                                         int(hints.shape[0]), V(hints.data_ptr()), ctypes.byref(opts), ctypes.byref(o),
                                         ws, wsb, _stream_ptr(stream))
        native.check(rc, "flash_batch_hinted")
        out["_inputs"] = (p, t, z, hints)
        return out
    if workspace is None:
        rc = lib.mpmc_flash_batch(mixture.handle, n, V(p.data_ptr()), V(t.data_ptr()), V(z.data_ptr()),
                                  ctypes.byref(opts), ctypes.byref(o), _stream_ptr(stream))
    else:
        rc = lib.mpmc_flash_batch_ws(mixture.handle, n, V(p.data_ptr()), V(t.data_ptr()), V(z.data_ptr()),
                                     ctypes.byref(opts), ctypes.byref(o), V(workspace.data_ptr()),
                                     workspace.numel() * workspace.element_size(), _stream_ptr(stream))
    native.check(rc, "flash_batch")
    out["_inputs"] = (p, t, z)  # keep inputs alive until the stream consumes them
    return out


def pt_continuation(mixture: Mixture, p, t, z, carry_compositions: bool = True, carry_three_phase: bool = True,
                    fallback_initial=None, fallback_final=None, start=None, seed=(1.0 / 3.0, 1.0 / 3.0),
                    fugacity_tolerance: float = 1e-11, precision: int = 0, stream=None, counters: bool = True,
                    workspace=None) -> dict:
    """Ordered-path PT continuation (mpmc_pt_continuation; solve_pr76_pt_continuation,
    pr76_pt_continuation.hpp:194-310) of npath paths: p, t [npoint, npath] ordered states,
    z [nc, npath] path feeds. fallback_initial / fallback_final ([k <= 3, nc]) are the
    fallback stability starts shared by every path, ``start`` ([3, nc]) a fixed three-phase
    start with fraction seed ``seed``. Returns device tensors with a leading point axis
    (status [npoint, npath], beta [npoint, 3, npath], ...) plus ``info`` [npoint, npath]
    = incoming hint phase count | carried start count << 4 | carried three-phase start << 8."""
    torch = _torch()
    lib = native.library()
    device = torch.device("cuda", torch.cuda.current_device())
    p = _dev(p, torch.float64, device)
    t = _dev(t, torch.float64, device)
    z = _dev(z, torch.float64, device)
    nc = mixture.nc
    if p.dim() != 2 or t.shape != p.shape or z.dim() != 2 or z.shape != (nc, p.shape[1]):
        raise ValueError("pt_continuation: expected p[npoint, npath], t[npoint, npath], z[nc, npath]")
    m, n = p.shape
    co = native.ContinuationOptions()
    lib.mpmc_continuation_default_options(ctypes.byref(co))
    co.carry_compositions = int(bool(carry_compositions))
    co.carry_three_phase = int(bool(carry_three_phase))
    for name, arr in (("fallback_initial", fallback_initial), ("fallback_final", fallback_final)):
        if arr is None:
            continue
        arr = np.asarray(arr, dtype=np.float64)
        if arr.ndim != 2 or arr.shape[1] != nc or arr.shape[0] > 3:
            raise ValueError(f"pt_continuation: {name} must be [k <= 3, nc]")
        setattr(co, "n" + name, arr.shape[0])
        dst = getattr(co, name)
        for k in range(arr.shape[0]):
            for i in range(nc):
                dst[k][i] = float(arr[k, i])
    if start is not None:
        st = np.asarray(start, dtype=np.float64)
        if st.shape != (3, nc):
            raise ValueError("pt_continuation: start must be [3, nc]")
        co.three_phase_start = 1
        for k in range(3):
            for i in range(nc):
                co.start[k][i] = float(st[k, i])
        co.start_seed[0], co.start_seed[1] = float(seed[0]), float(seed[1])
    f64, i32 = torch.float64, torch.int32
    out = {
        "status": torch.empty((m, n), dtype=i32, device=device),
        "phase_count": torch.empty((m, n), dtype=i32, device=device),
        "beta": torch.empty((m, 3, n), dtype=f64, device=device),
        "zfactor": torch.empty((m, 3, n), dtype=f64, device=device),
        "branch": torch.empty((m, 3, n), dtype=i32, device=device),
        "x": torch.empty((m, 3, nc, n), dtype=f64, device=device),
        "ln_phi": torch.empty((m, 3, nc, n), dtype=f64, device=device),
        "counters": torch.empty((m, 4, n), dtype=i32, device=device) if counters else None,
        "info": torch.empty((m, n), dtype=i32, device=device),
    }
    opts = _options(fugacity_tolerance, 3, precision)
    o = _outputs_struct(out)
    V = ctypes.c_void_p
    ws, wsb = (V(workspace.data_ptr()), workspace.numel() * workspace.element_size()) if workspace is not None \
        else (None, 0)
    rc = lib.mpmc_pt_continuation(mixture.handle, n, m, V(p.data_ptr()), V(t.data_ptr()), V(z.data_ptr()),
                                  ctypes.byref(co), ctypes.byref(opts), ctypes.byref(o), V(out["info"].data_ptr()), ws,
                                  wsb, _stream_ptr(stream))
    native.check(rc, "pt_continuation")
    out["_inputs"] = (p, t, z)
    return out


def continuation_workspace_bytes(nc: int, npath: int) -> int:
    return int(native.library().mpmc_pt_continuation_workspace_bytes(nc, npath))


def workspace_bytes(nc: int, ncell: int) -> int:
    return int(native.library().mpmc_flash_workspace_bytes(nc, ncell))


def _options(fugacity_tolerance, max_phases, precision) -> native.FlashOptions:
    opts = native.FlashOptions()
    native.library().mpmc_flash_default_options(ctypes.byref(opts))
    opts.fugacity_tolerance = float(fugacity_tolerance)
    opts.max_phases = int(max_phases)
    opts.precision = int(precision)
    return opts


def _outputs_struct(out: dict) -> native.FlashOutputs:
    return native.FlashOutputs(
        out["status"].data_ptr(), out["phase_count"].data_ptr(), out["beta"].data_ptr(), out["zfactor"].data_ptr(),
        out["branch"].data_ptr(), out["x"].data_ptr(), out["ln_phi"].data_ptr(),
        out["counters"].data_ptr() if out.get("counters") is not None else None)


def multi_max_mixtures(nc: int) -> int:
    return int(native.library().mpmc_flash_multi_max_mixtures(nc))


def flash_multi(mixtures, cell_offset, p, t, z, fugacity_tolerance: float = 1e-11, max_phases: int = 3,
                precision: int = 0, stream=None, out: dict | None = None, counters: bool = True,
                workspace=None) -> dict:
    """Ragged multi-mixture batch (mpmc_flash_multi): mixtures[k] owns cells
    [cell_offset[k], cell_offset[k+1]) of p[ncell], t[ncell], z[nc, ncell]; all mixtures
    have the same nc. Results equal flash_batch on each mixture's cells."""
    torch = _torch()
    lib = native.library()
    device = torch.device("cuda", torch.cuda.current_device())
    nc = mixtures[0].nc
    p = _dev(p, torch.float64, device)
    t = _dev(t, torch.float64, device)
    z = _dev(z, torch.float64, device)
    n = p.numel()
    off = np.ascontiguousarray(cell_offset, dtype=np.int64)
    if off.size != len(mixtures) + 1 or off[-1] != n or z.shape != (nc, n):
        raise ValueError("flash_multi: expected len(mixtures) + 1 offsets ending at ncell and z[nc, ncell]")
    if out is None:
        out = allocate_outputs(nc, n, device, counters)
    need = int(lib.mpmc_flash_multi_workspace_bytes(nc, len(mixtures), n))
    if workspace is None or workspace.numel() * workspace.element_size() < need:
        workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=device)
    handles = (ctypes.c_void_p * len(mixtures))(*[m.handle.value if isinstance(m.handle, ctypes.c_void_p) else m.handle
                                                   for m in mixtures])
    opts = _options(fugacity_tolerance, max_phases, precision)
    o = _outputs_struct(out)
    V = ctypes.c_void_p
    native.check(lib.mpmc_flash_multi(len(mixtures), handles, off.ctypes.data_as(V), V(p.data_ptr()),
                                      V(t.data_ptr()), V(z.data_ptr()), ctypes.byref(opts), ctypes.byref(o),
                                      V(workspace.data_ptr()), workspace.numel() * workspace.element_size(),
                                      _stream_ptr(stream)), "flash_multi")
    out["_inputs"] = (p, t, z, workspace)
    return out


class Plan:
    """A prepared uneven batch (mpmc_flash_plan_*): ``segments`` is a list of
    (mixture, p, t, z) with device tensors; outputs are allocated per segment. ``launch``
    replays the captured graph over the segments' current input contents."""

    def __init__(self, segments, fugacity_tolerance: float = 1e-11, max_phases: int = 3, precision: int = 0,
                 counters: bool = True, group_mixtures: int = 0):
        torch = _torch()
        lib = native.library()
        device = torch.device("cuda", torch.cuda.current_device())
        self.inputs, self.outputs = [], []
        segs = (native.FlashSegment * len(segments))()
        for k, (mix, p, t, z) in enumerate(segments):
            p = _dev(p, torch.float64, device)
            t = _dev(t, torch.float64, device)
            z = _dev(z, torch.float64, device)
            n = p.numel()
            if z.shape != (mix.nc, n) or t.numel() != n:
                raise ValueError("Plan: expected p[n], t[n], z[nc, n] per segment")
            out = allocate_outputs(mix.nc, n, device, counters)
            self.inputs.append((mix, p, t, z))
            self.outputs.append(out)
            h = mix.handle.value if isinstance(mix.handle, ctypes.c_void_p) else mix.handle
            segs[k] = native.FlashSegment(h, n, p.data_ptr(), t.data_ptr(), z.data_ptr(), _outputs_struct(out))
        opts = _options(fugacity_tolerance, max_phases, precision)
        opts.group_mixtures = int(group_mixtures)
        self.group_mixtures = int(group_mixtures)
        self.handle = ctypes.c_void_p()
        native.check(lib.mpmc_flash_plan_create(len(segments), segs, ctypes.byref(opts), ctypes.byref(self.handle)),
                     "flash_plan_create")
        g, b = ctypes.c_int32(0), ctypes.c_size_t(0)
        native.check(lib.mpmc_flash_plan_info(self.handle, ctypes.byref(g), ctypes.byref(b)), "flash_plan_info")
        self.groups, self.device_bytes = int(g.value), int(b.value)

    def launch(self, stream=None):
        native.check(native.library().mpmc_flash_plan_launch(self.handle, _stream_ptr(stream)), "flash_plan_launch")

    def close(self):
        if self.handle:
            native.library().mpmc_flash_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_plan(segments, candidates=(0, 16, 8), replays: int = 1, stream=None, **plan_kwargs):
    """Launch-group autotuner for an uneven batch: builds the plan once per candidate
    ``group_mixtures`` (0: the LDS-table maximum per group), times ``replays`` graph replays
    of each with HIP events on ``stream``, and keeps the fastest. Small groups trade more
    launches for smaller LDS mixture tables (more resident workgroups per CU) and shorter
    per-launch tails. Every candidate gives the same records (the grouping does not change
    any cell's computation). Returns (plan, {candidate: ms per replay})."""
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    timings, best, best_ms = {}, None, float("inf")
    for cand in candidates:
        plan = Plan(segments, group_mixtures=cand, **plan_kwargs)
        plan.launch(s)  # first replay outside the timing
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(replays):
            plan.launch(s)
        e1.record(s)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / replays
        timings[int(cand)] = ms
        if ms < best_ms:
            if best is not None:
                best.close()
            best, best_ms = plan, ms
        else:
            plan.close()
    return best, timings


def phase_batch(mixture: Mixture, p, t, x, stream=None) -> dict:
    """All roots and the ln(phi) of each admissible root at x[nc, ncell]."""
    torch = _torch()
    lib = native.library()
    device = torch.device("cuda", torch.cuda.current_device())
    p = _dev(p, torch.float64, device)
    t = _dev(t, torch.float64, device)
    x = _dev(x, torch.float64, device)
    n = p.numel()
    nc = mixture.nc
    out = {
        "root_status": torch.empty(n, dtype=torch.int32, device=device),
        "root_count": torch.empty(n, dtype=torch.int32, device=device),
        "zroot": torch.empty((3, n), dtype=torch.float64, device=device),
        "root_flags": torch.empty((3, n), dtype=torch.int32, device=device),
        "ln_phi": torch.empty((3, nc, n), dtype=torch.float64, device=device),
    }
    V = ctypes.c_void_p
    native.check(lib.mpmc_phase_batch(mixture.handle, n, V(p.data_ptr()), V(t.data_ptr()), V(x.data_ptr()),
                                      V(out["root_status"].data_ptr()), V(out["root_count"].data_ptr()),
                                      V(out["zroot"].data_ptr()), V(out["root_flags"].data_ptr()),
                                      V(out["ln_phi"].data_ptr()), _stream_ptr(stream)), "phase_batch")
    out["_inputs"] = (p, t, x)
    return out


def rachford_rice_batch(z, log_k, guess=None, stream=None) -> dict:
    """Rachford-Rice at fixed log K for z[nc, n], log_k[nc, n] (device tensors out)."""
    torch = _torch()
    lib = native.library()
    device = torch.device("cuda", torch.cuda.current_device())
    z = _dev(z, torch.float64, device)
    lk = _dev(log_k, torch.float64, device)
    nc, n = z.shape
    g = _dev(guess, torch.float64, device) if guess is not None else None
    out = {
        "status": torch.empty(n, dtype=torch.int32, device=device),
        "beta": torch.empty(n, dtype=torch.float64, device=device),
        "x": torch.empty((nc, n), dtype=torch.float64, device=device),
        "y": torch.empty((nc, n), dtype=torch.float64, device=device),
        "iterations": torch.empty(n, dtype=torch.int32, device=device),
    }
    V = ctypes.c_void_p
    native.check(lib.mpmc_rr_batch(nc, n, V(z.data_ptr()), V(lk.data_ptr()), V(g.data_ptr()) if g is not None else None,
                                   V(out["status"].data_ptr()), V(out["beta"].data_ptr()), V(out["x"].data_ptr()),
                                   V(out["y"].data_ptr()), V(out["iterations"].data_ptr()), _stream_ptr(stream)),
                 "rr_batch")
    out["_inputs"] = (z, lk, g)
    return out


def _flash_struct(fl: dict) -> native.FlashOutputs:
    return native.FlashOutputs(
        fl["status"].data_ptr(), fl["phase_count"].data_ptr(), fl["beta"].data_ptr(), fl["zfactor"].data_ptr(),
        fl["branch"].data_ptr(), fl["x"].data_ptr(), fl["ln_phi"].data_ptr() if fl.get("ln_phi") is not None else None,
        None)


def closure_batch(mixture: Mixture, p, t, z, flash: dict, fugacity_tolerance: float = 1e-11,
                  min_fraction: float = 1.490116119384765625e-8, min_rcond: float = 0.0, log_k: bool = True,
                  diag: bool = True, stream=None, out: dict | None = None) -> dict:
    """Thermodynamic closure (build_pr76_pt_vle_closure, pr76_thermodynamic_closure.hpp:148-381)
    of every cell of a flash result ``flash`` computed on the same p, t, z.

    Returns status [n] (primal | linearization << 4 | reason << 8), primal [4, n]
    (Z_L, Z_V, rho_L, rho_V), jac [2nc+5, nc+1, n] over q = (p, T, z_0..z_{nc-2}) with
    rows beta_V, x_L[nc], y_V[nc], Z_L, Z_V, rho_L, rho_V, and optionally log_k_jac
    [nc, nc+1, n] and diag [3, n] (rcond, backward error, residual norm)."""
    torch = _torch()
    lib = native.library()
    device = torch.device("cuda", torch.cuda.current_device())
    p = _dev(p, torch.float64, device)
    t = _dev(t, torch.float64, device)
    z = _dev(z, torch.float64, device)
    nc, n = mixture.nc, p.numel()
    if z.shape != (nc, n) or t.numel() != n or flash["status"].numel() != n:
        raise ValueError("closure_batch: expected p[n], t[n], z[nc, n] and a flash result of n cells")
    q = nc + 1
    if out is None:
        f64 = torch.float64
        out = {
            "status": torch.empty(n, dtype=torch.int32, device=device),
            "primal": torch.empty((4, n), dtype=f64, device=device),
            "jac": torch.empty((2 * nc + 5, q, n), dtype=f64, device=device),
            "log_k_jac": torch.empty((nc, q, n), dtype=f64, device=device) if log_k else None,
            "diag": torch.empty((3, n), dtype=f64, device=device) if diag else None,
        }
    opts = native.ClosureOptions()
    lib.mpmc_closure_default_options(ctypes.byref(opts))
    opts.minimum_derivative_phase_fraction = float(min_fraction)
    opts.minimum_reciprocal_condition = float(min_rcond)
    opts.fugacity_tolerance = float(fugacity_tolerance)
    o = native.ClosureOutputs(out["status"].data_ptr(), out["primal"].data_ptr(), out["jac"].data_ptr(),
                              out["log_k_jac"].data_ptr() if out.get("log_k_jac") is not None else None,
                              out["diag"].data_ptr() if out.get("diag") is not None else None)
    fo = _flash_struct(flash)
    V = ctypes.c_void_p
    native.check(lib.mpmc_closure_batch(mixture.handle, n, V(p.data_ptr()), V(t.data_ptr()), V(z.data_ptr()),
                                        ctypes.byref(fo), ctypes.byref(opts), ctypes.byref(o), _stream_ptr(stream)),
                 "closure_batch")
    out["_inputs"] = (p, t, z, flash)
    return out


def to_host(res: dict) -> dict:
    """Synchronise and copy a result dict to numpy (counters as uint32)."""
    torch = _torch()
    torch.cuda.synchronize()
    h = {}
    for k, v in res.items():
        if k.startswith("_") or v is None:
            continue
        a = v.cpu().numpy()
        if k == "counters":
            a = a.view(np.uint32)
        h[k] = a
    return h
