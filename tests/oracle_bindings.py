"""ctypes bindings of the checkers: the oracle (oracle/libcpuref.so) and, where it
was built, the reference compiled from its own headers (oracle/_ref/libmpmc_ref.so).

Both return results in the same structure-of-arrays layout as the HIP library.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
CPUREF = ROOT / "oracle" / "libcpuref.so"
REFLIB = ROOT / "oracle" / "_ref" / "libmpmc_ref.so"

_P = ctypes.c_void_p
_FLASH_ARGS = [ctypes.c_int, _P, _P, _P, _P, ctypes.c_int64, _P, _P, _P, ctypes.c_double, ctypes.c_int,
               ctypes.c_int] + [_P] * 8
_PHASE_ARGS = [ctypes.c_int, _P, _P, _P, _P, ctypes.c_int64, _P, _P, _P] + [_P] * 5
_FLASH_HINT_ARGS = [ctypes.c_int, _P, _P, _P, _P, ctypes.c_int64, _P, _P, _P, ctypes.c_int, _P, ctypes.c_double,
                    ctypes.c_int, ctypes.c_int] + [_P] * 8
_CONT_ARGS = [ctypes.c_int, _P, _P, _P, _P, ctypes.c_int64, ctypes.c_int, _P, _P, _P, ctypes.c_int, ctypes.c_int, _P,
              ctypes.c_int, _P, ctypes.c_int, _P, _P, ctypes.c_double, ctypes.c_int] + [_P] * 9
_FLASH_WARM_ARGS = [ctypes.c_int, _P, _P, _P, _P, ctypes.c_int64, _P, _P, _P, ctypes.c_int, _P, ctypes.c_int, _P,
                    ctypes.c_double, ctypes.c_int, ctypes.c_int] + [_P] * 8


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def host_threads(threads: int = 0) -> int:
    """Worker threads of a checker call: the given count, else the CPUs this process may
    run on, at most 16 (a GPU box gives one GPU's jobs a 16-CPU share whatever the
    machine reports)."""
    if threads > 0:
        return int(threads)
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def _load(path: Path, prefix: str):
    if not path.exists():
        return None
    lib = ctypes.CDLL(str(path))
    getattr(lib, prefix + "_flash_batch").argtypes = _FLASH_ARGS
    getattr(lib, prefix + "_flash_batch_hinted").argtypes = _FLASH_HINT_ARGS
    getattr(lib, prefix + "_flash_batch_warm").argtypes = _FLASH_WARM_ARGS
    getattr(lib, prefix + "_phase_batch").argtypes = _PHASE_ARGS
    getattr(lib, prefix + "_pt_continuation").argtypes = _CONT_ARGS
    return lib


_libs: dict = {}


def lib(which: str = "cpuref"):
    """which: 'cpuref' (the oracle) or 'ref' (reference build; None when absent)."""
    if which not in _libs:
        _libs[which] = _load(CPUREF, "cpuref") if which == "cpuref" else _load(REFLIB, "ref")
    return _libs[which]


def have_ref() -> bool:
    return lib("ref") is not None


def _mix_arrays(mix):
    return (np.ascontiguousarray(mix["tc"], dtype=np.float64), np.ascontiguousarray(mix["pc"], dtype=np.float64),
            np.ascontiguousarray(mix["omega"], dtype=np.float64),
            np.ascontiguousarray(mix["kij"], dtype=np.float64))


def flash(mix, p, t, z, fugacity_tolerance=1e-11, max_phases=3, threads=0, which="cpuref", hints=None,
          final_hints=None):
    """Run the checker flash on host arrays; returns a dict of SoA arrays. ``hints``
    ([nhint <= 6, nc, ncell], NaN in component 0 = absent) are warm-start compositions
    appended to the initial stability starts; ``final_hints`` (same layout) are the extra
    starts of the common-tangent review."""
    L = lib(which)
    if L is None:
        raise RuntimeError(f"checker library '{which}' is not built")
    tc, pc, om, kij = _mix_arrays(mix)
    p = np.ascontiguousarray(p, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    nc, n = z.shape
    out = {
        "status": np.zeros(n, np.int32),
        "phase_count": np.zeros(n, np.int32),
        "beta": np.zeros((3, n)),
        "zfactor": np.zeros((3, n)),
        "branch": np.zeros((3, n), np.int32),
        "x": np.zeros((3, nc, n)),
        "ln_phi": np.zeros((3, nc, n)),
        "counters": np.zeros((4, n), np.uint32),
    }
    if final_hints is not None:
        hints = np.zeros((0, nc, n)) if hints is None else np.ascontiguousarray(hints, dtype=np.float64)
        final_hints = np.ascontiguousarray(final_hints, dtype=np.float64)
        for h in (hints, final_hints):
            assert h.ndim == 3 and h.shape[1:] == (nc, n) and h.shape[0] <= 6
        fn = getattr(L, ("cpuref" if which == "cpuref" else "ref") + "_flash_batch_warm")
        rc = fn(nc, _ptr(tc), _ptr(pc), _ptr(om), _ptr(kij), n, _ptr(p), _ptr(t), _ptr(z), int(hints.shape[0]),
                _ptr(hints), int(final_hints.shape[0]), _ptr(final_hints), float(fugacity_tolerance), int(max_phases),
                host_threads(threads), _ptr(out["status"]), _ptr(out["phase_count"]), _ptr(out["beta"]),
                _ptr(out["zfactor"]), _ptr(out["branch"]), _ptr(out["x"]), _ptr(out["ln_phi"]),
                _ptr(out["counters"]))
    elif hints is not None:
        hints = np.ascontiguousarray(hints, dtype=np.float64)
        assert hints.ndim == 3 and hints.shape[1:] == (nc, n) and hints.shape[0] <= 6
        fn = getattr(L, ("cpuref" if which == "cpuref" else "ref") + "_flash_batch_hinted")
        rc = fn(nc, _ptr(tc), _ptr(pc), _ptr(om), _ptr(kij), n, _ptr(p), _ptr(t), _ptr(z), int(hints.shape[0]),
                _ptr(hints), float(fugacity_tolerance), int(max_phases), host_threads(threads), _ptr(out["status"]),
                _ptr(out["phase_count"]), _ptr(out["beta"]), _ptr(out["zfactor"]), _ptr(out["branch"]),
                _ptr(out["x"]), _ptr(out["ln_phi"]), _ptr(out["counters"]))
    else:
        fn = getattr(L, ("cpuref" if which == "cpuref" else "ref") + "_flash_batch")
        rc = fn(nc, _ptr(tc), _ptr(pc), _ptr(om), _ptr(kij), n, _ptr(p), _ptr(t), _ptr(z),
                float(fugacity_tolerance), int(max_phases), host_threads(threads), _ptr(out["status"]),
                _ptr(out["phase_count"]), _ptr(out["beta"]), _ptr(out["zfactor"]), _ptr(out["branch"]),
                _ptr(out["x"]), _ptr(out["ln_phi"]), _ptr(out["counters"]))
    if rc != 0:
        raise RuntimeError(f"{which}_flash_batch failed rc={rc}")
    return out


def phase(mix, p, t, x, which="cpuref"):
    L = lib(which)
    tc, pc, om, kij = _mix_arrays(mix)
    p = np.ascontiguousarray(p, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    nc, n = x.shape
    out = {
        "root_status": np.zeros(n, np.int32),
        "root_count": np.zeros(n, np.int32),
        "zroot": np.zeros((3, n)),
        "root_flags": np.zeros((3, n), np.int32),
        "ln_phi": np.zeros((3, nc, n)),
    }
    fn = getattr(L, ("cpuref" if which == "cpuref" else "ref") + "_phase_batch")
    rc = fn(nc, _ptr(tc), _ptr(pc), _ptr(om), _ptr(kij), n, _ptr(p), _ptr(t), _ptr(x), _ptr(out["root_status"]),
            _ptr(out["root_count"]), _ptr(out["zroot"]), _ptr(out["root_flags"]), _ptr(out["ln_phi"]))
    if rc != 0:
        raise RuntimeError(f"{which}_phase_batch failed rc={rc}")
    return out


def rachford_rice(z, log_k):
    L = lib("cpuref")
    z = np.ascontiguousarray(z, dtype=np.float64)
    lk = np.ascontiguousarray(log_k, dtype=np.float64)
    n = z.size
    beta = ctypes.c_double()
    x = np.zeros(n)
    y = np.zeros(n)
    it = ctypes.c_int32()
    st = L.cpuref_rachford_rice(n, _ptr(z), _ptr(lk), ctypes.byref(beta), _ptr(x), _ptr(y), ctypes.byref(it))
    return st, beta.value, x, y, it.value


_CLOSURE_ARGS = [ctypes.c_int, _P, _P, _P, _P, ctypes.c_int64, _P, _P, _P, ctypes.c_double, ctypes.c_double,
                 ctypes.c_double, ctypes.c_int] + [_P] * 5


def closure(mix, p, t, z, fugacity_tolerance=1e-11, min_fraction=1.490116119384765625e-8, min_rcond=0.0,
            threads=0, which="cpuref"):
    """Two-phase (VLE) flash of each cell followed by its thermodynamic closure.

    Returns status word (primal | linearization << 4 | reason << 8), primal [4][n]
    (Z_L, Z_V, rho_L, rho_V), jac [2nc+5][nc+1][n] (d of beta_V, x_L, y_V, Z_L, Z_V,
    rho_L, rho_V over q = (p, T, z_0..z_{nc-2})), log_k_jac [nc][nc+1][n], diag [3][n]."""
    L = lib(which)
    if L is None:
        raise RuntimeError(f"checker library '{which}' is not built")
    tc, pc, om, kij = _mix_arrays(mix)
    p = np.ascontiguousarray(p, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    nc, n = z.shape
    q = nc + 1
    out = {
        "status": np.zeros(n, np.int32),
        "primal": np.zeros((4, n)),
        "jac": np.zeros((2 * nc + 5, q, n)),
        "log_k_jac": np.zeros((nc, q, n)),
        "diag": np.zeros((3, n)),
    }
    prefix = "cpuref" if which == "cpuref" else "ref"
    fn = getattr(L, prefix + "_closure_batch")
    fn.argtypes = _CLOSURE_ARGS
    rc = fn(nc, _ptr(tc), _ptr(pc), _ptr(om), _ptr(kij), n, _ptr(p), _ptr(t), _ptr(z), float(fugacity_tolerance),
            float(min_fraction), float(min_rcond), host_threads(threads), _ptr(out["status"]), _ptr(out["primal"]),
            _ptr(out["jac"]), _ptr(out["log_k_jac"]), _ptr(out["diag"]))
    if rc != 0:
        raise RuntimeError(f"{prefix}_closure_batch failed rc={rc}")
    return out


def assembly(mix, model, q, branch, grid, prev, dt, threads=0, which="cpuref"):
    """The two-phase assembly of the oracle (cpuref_assembly) or of the reference's own
    building blocks (ref_assembly): returns rc, cell_status, residual[Q][n], diag[Q*Q][n],
    offdiag[2][(nc+1)*Q][nface] (side 0: owner row x neighbour column)."""
    L = lib(which)
    if L is None:
        raise RuntimeError(f"checker library '{which}' is not built")
    tc, pc, om, kij = _mix_arrays(mix)
    q = np.ascontiguousarray(q, dtype=np.float64)
    nc = (q.shape[0] - 1) // 2
    Q = q.shape[0]
    n = q.shape[1]
    nf = grid["owner"].size
    br = np.ascontiguousarray(branch, dtype=np.int32)
    model = np.ascontiguousarray(model, dtype=np.float64)
    prev = np.ascontiguousarray(prev, dtype=np.float64)
    g = {k: np.ascontiguousarray(grid[k]) for k in ("bulk", "poro", "owner", "neighbour", "trans", "gdot", "gth")}
    out = {"cell_status": np.zeros(n, np.int32), "residual": np.zeros((Q, n)), "diag": np.zeros((Q * Q, n)),
           "offdiag": np.zeros((2, (nc + 1) * Q, max(1, nf)))}
    fn = getattr(L, ("cpuref" if which == "cpuref" else "ref") + "_assembly")
    fn.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_int64, _P, _P, _P, _P, _P, ctypes.c_double,
                   ctypes.c_int64, _P, _P, _P, _P, _P, ctypes.c_int, _P, _P, _P, _P]
    rc = fn(nc, _ptr(tc), _ptr(pc), _ptr(om), _ptr(kij), _ptr(model), n, _ptr(q), _ptr(br), _ptr(g["bulk"]),
            _ptr(g["poro"]), _ptr(prev), float(dt), nf, _ptr(g["owner"]), _ptr(g["neighbour"]), _ptr(g["trans"]),
            _ptr(g["gdot"]), _ptr(g["gth"]), host_threads(threads), _ptr(out["cell_status"]), _ptr(out["residual"]),
            _ptr(out["diag"]), _ptr(out["offdiag"]))
    out["rc"] = rc
    return out


def inventory(mix, model, q, branch, poro, threads=0):
    L = lib("cpuref")
    tc, pc, om, kij = _mix_arrays(mix)
    q = np.ascontiguousarray(q, dtype=np.float64)
    nc = (q.shape[0] - 1) // 2
    n = q.shape[1]
    acc = np.zeros((nc + 1, n))
    fn = L.cpuref_accumulation
    fn.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_int64, _P, _P, _P, ctypes.c_int, _P]
    rc = fn(nc, _ptr(tc), _ptr(pc), _ptr(om), _ptr(kij), _ptr(np.ascontiguousarray(model, dtype=np.float64)), n,
            _ptr(q), _ptr(np.ascontiguousarray(branch, dtype=np.int32)), _ptr(np.ascontiguousarray(poro)),
            host_threads(threads), _ptr(acc))
    if rc != 0:
        raise RuntimeError(f"cpuref_accumulation rc={rc}")
    return acc


def phase_roots(mix):
    """roots(p, t, x) -> (root_count, root_flags) through the oracle, for state generators."""
    def roots(p, t, x):
        r = phase(mix, p, t, x)
        return r["root_count"], r["root_flags"]
    return roots


def pt_continuation(mix, p, t, z, carry=3, fugacity_tolerance=1e-11, threads=0, which="cpuref",
                    fallback_initial=None, fallback_final=None, start=None, seed=(1.0 / 3.0, 1.0 / 3.0)):
    """Ordered-path continuation (solve_pr76_pt_continuation) of npath paths: p, t are
    [npoint, npath] states, z [nc, npath] feeds. Returns the flash record arrays with a
    leading point axis plus ``info`` [npoint, npath] = incoming hint phase count | carried
    start count << 4 | carried three-phase start << 8. carry: bit 0 carries phase
    compositions, bit 1 the three-phase state; fallback_initial / fallback_final: [k <= 3, nc]
    starts shared by every path; start: a fixed three-phase start [3, nc] with ``seed``."""
    L = lib(which)
    if L is None:
        raise RuntimeError(f"checker library '{which}' is not built")
    tc, pc, om, kij = _mix_arrays(mix)
    p = np.ascontiguousarray(p, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    nc, n = z.shape
    m = p.shape[0]
    assert p.shape == (m, n) and t.shape == (m, n)
    fi = np.zeros((0, nc)) if fallback_initial is None else np.ascontiguousarray(fallback_initial, dtype=np.float64)
    ff = np.zeros((0, nc)) if fallback_final is None else np.ascontiguousarray(fallback_final, dtype=np.float64)
    st = np.zeros((3, nc)) if start is None else np.ascontiguousarray(start, dtype=np.float64)
    sd = np.ascontiguousarray(seed, dtype=np.float64)
    assert fi.shape[1:] == (nc,) and ff.shape[1:] == (nc,) and st.shape == (3, nc) and sd.shape == (2,)
    out = {
        "status": np.zeros((m, n), np.int32),
        "phase_count": np.zeros((m, n), np.int32),
        "beta": np.zeros((m, 3, n)),
        "zfactor": np.zeros((m, 3, n)),
        "branch": np.zeros((m, 3, n), np.int32),
        "x": np.zeros((m, 3, nc, n)),
        "ln_phi": np.zeros((m, 3, nc, n)),
        "counters": np.zeros((m, 4, n), np.uint32),
        "info": np.zeros((m, n), np.int32),
    }
    fn = getattr(L, ("cpuref" if which == "cpuref" else "ref") + "_pt_continuation")
    rc = fn(nc, _ptr(tc), _ptr(pc), _ptr(om), _ptr(kij), n, m, _ptr(p), _ptr(t), _ptr(z), int(carry), fi.shape[0],
            _ptr(fi), ff.shape[0], _ptr(ff), int(start is not None), _ptr(st), _ptr(sd), float(fugacity_tolerance),
            host_threads(threads), _ptr(out["status"]), _ptr(out["phase_count"]), _ptr(out["beta"]),
            _ptr(out["zfactor"]), _ptr(out["branch"]), _ptr(out["x"]), _ptr(out["ln_phi"]), _ptr(out["counters"]),
            _ptr(out["info"]))
    if rc != 0:
        raise RuntimeError(f"{which}_pt_continuation failed ({rc})")
    return out
