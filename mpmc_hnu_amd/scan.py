"""Post-SNES PT-flash phase-transition scan on the GPU (mpmc_transition_scan,
mpmc_transition_scan_2p in include/mpmc_hnu_amd.h): the simulator-side batch caller of the
flash, after the reference's scanner
(modules/flow/discretization/petsc/include/mpmc/flow_discretization_petsc/
post_snes_pt_flash_phase_transition_scanner.hpp).

Per cell the scan flashes the cell's current PT feed and compares the accepted phase count
with the cell's current one: SAME (no proposal), PROPOSAL (another accepted count, with the
target phases' molar densities p / (Z R T)), INDETERMINATE, ERROR. The summary is the
reference's stop-at-first-cell outcome over the batch in cell order.

density_mode 0 reads Z from the flash record, as the reference's scanner test resolver does,
so a proposal to a single phase (whose record publishes no Z) is indeterminate; 1 gives that
phase the Z of its provider root at the feed (an extension with no reference counterpart)."""
from __future__ import annotations

import ctypes

from . import native
from .assembly import Grid
from .flash import Mixture, _dev, _outputs_struct, _stream_ptr, _torch, allocate_outputs

V = ctypes.c_void_p

SAME, PROPOSAL, INDETERMINATE, ERROR = 0, 1, 2, 3
COMPLETE, STATUS_INDETERMINATE = 0, 1


def _scan_buffers(n: int, device) -> dict:
    torch = _torch()
    return {
        "code": torch.empty(n, dtype=torch.int32, device=device),
        "target_count": torch.empty(n, dtype=torch.int32, device=device),
        "target_density": torch.empty((3, n), dtype=torch.float64, device=device),
    }


def _options(fugacity_tolerance: float, max_phases: int, precision: int) -> native.FlashOptions:
    opts = native.FlashOptions()
    native.library().mpmc_flash_default_options(ctypes.byref(opts))
    opts.fugacity_tolerance = float(fugacity_tolerance)
    opts.max_phases = int(max_phases)
    opts.precision = int(precision)
    return opts


def _finish(rc: int, what: str, res: dict, summary: native.ScanSummary) -> dict:
    res["status"] = int(summary.status)
    res["first_stop"] = int(summary.first_stop)
    res["proposals"] = int(summary.proposals)
    res["rc"] = int(rc)
    if rc != 0 and not (rc == 1 and summary.first_stop >= 0):
        native.check(rc, what)
    if rc == 1:
        # the first stopping cell raises invalid_argument in the reference; the codes stay
        # readable, the caller decides
        res["error"] = native.library().mpmc_last_error().decode(errors="replace")
    return res


def transition_scan(mixture: Mixture, p, t, z, source_phase_count, fugacity_tolerance: float = 1e-11,
                    max_phases: int = 3, precision: int = 0, density_mode: int = 0, stream=None) -> dict:
    """Scan caller-given sources: p[ncell] Pa, t[ncell] K, z[nc, ncell], the cells' current
    phase counts source_phase_count[ncell]. Returns the scan codes, target counts and
    densities, the flash record ("flash") and the summary (status, first_stop, proposals,
    rc: 0, or 1 when the first stopping cell is an error)."""
    torch = _torch()
    lib = native.library()
    dev = torch.device("cuda", torch.cuda.current_device())
    p = _dev(p, torch.float64, dev)
    t = _dev(t, torch.float64, dev)
    z = _dev(z, torch.float64, dev)
    cnt = _dev(source_phase_count, torch.int32, dev)
    n = p.numel()
    if z.dim() != 2 or z.shape != (mixture.nc, n) or t.numel() != n or cnt.numel() != n:
        raise ValueError("transition_scan: expected p[ncell], t[ncell], z[nc, ncell], source_phase_count[ncell]")
    out = allocate_outputs(mixture.nc, n, dev, counters=False)
    res = _scan_buffers(n, dev)
    opts = _options(fugacity_tolerance, max_phases, precision)
    so = native.ScanOutputs(res["code"].data_ptr(), res["target_count"].data_ptr(), res["target_density"].data_ptr())
    summary = native.ScanSummary(0, -1, 0)
    fo = _outputs_struct(out)
    rc = lib.mpmc_transition_scan(mixture.handle, n, V(p.data_ptr()), V(t.data_ptr()), V(z.data_ptr()),
                                  V(cnt.data_ptr()), ctypes.byref(opts), int(density_mode), ctypes.byref(fo),
                                  ctypes.byref(so),
                                  ctypes.byref(summary), _stream_ptr(stream))
    res["flash"] = out
    return _finish(rc, "transition_scan", res, summary)


def transition_scan_2p(mixture: Mixture, grid: Grid, model: native.AssemblyModel, q, branch,
                       fugacity_tolerance: float = 1e-11, max_phases: int = 3, precision: int = 0,
                       density_mode: int = 0, stream=None) -> dict:
    """Scan a converged two-phase state q[2nc+1, ncell] (with the phase root branches of the
    assembly) on `grid`: each cell's feed is its normalised pore-volume inventory, its
    current phase count 2. Returns the scan result of transition_scan plus "inventory"
    [nc+1, ncell] and "feed" [nc, ncell]."""
    torch = _torch()
    lib = native.library()
    dev = torch.device("cuda", torch.cuda.current_device())
    q = _dev(q, torch.float64, dev)
    branch = _dev(branch, torch.int32, dev)
    n = grid.ncell
    nc = mixture.nc
    if q.shape != (2 * nc + 1, n):
        raise ValueError("transition_scan_2p: expected q[2nc+1, ncell]")
    out = allocate_outputs(nc, n, dev, counters=False)
    res = _scan_buffers(n, dev)
    res["inventory"] = torch.empty((nc + 1, n), dtype=torch.float64, device=dev)
    res["feed"] = torch.empty((nc, n), dtype=torch.float64, device=dev)
    opts = _options(fugacity_tolerance, max_phases, precision)
    so = native.ScanOutputs(res["code"].data_ptr(), res["target_count"].data_ptr(), res["target_density"].data_ptr())
    summary = native.ScanSummary(0, -1, 0)
    fo = _outputs_struct(out)
    rc = lib.mpmc_transition_scan_2p(mixture.handle, grid.h, ctypes.byref(model), V(q.data_ptr()),
                                     V(branch.data_ptr()), V(res["inventory"].data_ptr()), V(res["feed"].data_ptr()),
                                     ctypes.byref(opts), int(density_mode), ctypes.byref(fo), ctypes.byref(so),
                                     ctypes.byref(summary), _stream_ptr(stream))
    res["flash"] = out
    return _finish(rc, "transition_scan_2p", res, summary)
