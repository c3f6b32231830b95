"""Two-phase natural-variable residual/Jacobian assembly on the GPU (mpmc_tpfa_grid_*,
mpmc_assembly_2p, mpmc_inventory_2p in include/mpmc_hnu_amd.h): the per-Newton-iteration
work of the reference's TwoPhaseSnesAssemblyContext3D::evaluate_complete_assembly
(modules/flow/discretization/petsc/include/mpmc/flow_discretization_petsc/
two_phase_snes_assembly.hpp:550-1090); PETSc stays on the host."""
from __future__ import annotations

import ctypes

import numpy as np

from . import native
from .flash import Mixture, _dev, _stream_ptr, _torch

V = ctypes.c_void_p


def default_model(nc: int, molar_mass=None) -> native.AssemblyModel:
    md = native.AssemblyModel()
    native.library().mpmc_assembly_default_model(nc, ctypes.byref(md))
    if molar_mass is not None:
        for i, v in enumerate(molar_mass):
            md.molar_mass[i] = float(v)
    return md


def model_array(md: native.AssemblyModel, nc: int) -> np.ndarray:
    """The model as the oracle takes it: mu_a0, mu_a1, cp, t0, cv_rock, molar_mass[nc]."""
    return np.array([md.mu_a0, md.mu_a1, md.cp, md.t0, md.cv_rock] + [md.molar_mass[i] for i in range(nc)])


class Grid:
    """Device-resident TPFA grid (faces, transmissibilities, bulk volume, porosity)."""

    def __init__(self, g: dict):
        torch = _torch()
        lib = native.library()
        dev = torch.device("cuda", torch.cuda.current_device())
        self.ncell, self.nface = int(g["ncell"]), int(g["owner"].size)
        self.t = {k: _dev(g[k], torch.float64, dev) for k in ("bulk", "poro", "trans", "gdot", "gth")}
        self.t["owner"] = _dev(g["owner"], torch.int32, dev)
        self.t["neighbour"] = _dev(g["neighbour"], torch.int32, dev)
        self.h = V()
        t = self.t
        native.check(lib.mpmc_tpfa_grid_create(self.ncell, self.nface, V(t["bulk"].data_ptr()), V(t["poro"].data_ptr()),
                                               V(t["owner"].data_ptr()), V(t["neighbour"].data_ptr()),
                                               V(t["trans"].data_ptr()), V(t["gdot"].data_ptr()),
                                               V(t["gth"].data_ptr()), ctypes.byref(self.h)), "tpfa_grid_create")
        deg = ctypes.c_int32(0)
        native.check(lib.mpmc_tpfa_grid_max_degree(self.h, ctypes.byref(deg)), "tpfa_grid_max_degree")
        self.max_degree = int(deg.value)

    def close(self):
        if self.h:
            native.library().mpmc_tpfa_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def adjacency_host(owner, neighbour, ncell: int):
    """The cell -> face CSR the library builds (faces of a cell in face order, each face
    listed once under its owner and once under its neighbour)."""
    owner = np.asarray(owner)
    neighbour = np.asarray(neighbour)
    off = np.zeros(ncell + 1, np.int64)
    np.add.at(off, owner + 1, 1)
    np.add.at(off, neighbour + 1, 1)
    off = np.cumsum(off)
    fill = off[:-1].copy()
    adj = np.zeros(2 * owner.size, np.int64)
    for f in range(owner.size):
        adj[fill[owner[f]]] = f
        fill[owner[f]] += 1
        adj[fill[neighbour[f]]] = f
        fill[neighbour[f]] += 1
    return off, adj


def ell_to_csr(offdiag, off, nface: int):
    """Off-diagonal blocks from the library's (slot, cell) layout [rows*Q][max_degree][ncell]
    to CSR positions [rows*Q][2 nface]: slot j of cell c is CSR position off[c] + j."""
    offdiag = np.asarray(offdiag)
    ncell = off.size - 1
    deg = np.diff(off)
    out = np.zeros((offdiag.shape[0], 2 * nface))
    for j in range(offdiag.shape[1]):
        cells = np.nonzero(deg > j)[0]
        out[:, off[cells] + j] = offdiag[:, j, cells]
    return out


def allocate(nc: int, grid: Grid, device) -> dict:
    torch = _torch()
    Q = 2 * nc + 1
    return {
        "cell_status": torch.empty(grid.ncell, dtype=torch.int32, device=device),
        "residual": torch.empty((Q, grid.ncell), dtype=torch.float64, device=device),
        "diag": torch.empty((Q * Q, grid.ncell), dtype=torch.float64, device=device),
        "offdiag": torch.empty(((nc + 1) * Q, max(1, grid.max_degree), grid.ncell), dtype=torch.float64,
                               device=device),
    }


def assemble(mixture: Mixture, grid: Grid, model, q, branch, prev, dt: float, out: dict | None = None,
             stream=None, check_domain: bool = True):
    torch = _torch()
    lib = native.library()
    dev = torch.device("cuda", torch.cuda.current_device())
    q = _dev(q, torch.float64, dev)
    branch = _dev(branch, torch.int32, dev)
    prev = _dev(prev, torch.float64, dev)
    if out is None:
        out = allocate(mixture.nc, grid, dev)
    o = native.AssemblyOutputs(out["cell_status"].data_ptr(), out["residual"].data_ptr(), out["diag"].data_ptr(),
                               out["offdiag"].data_ptr())
    flag = ctypes.c_int32(0)
    native.check(lib.mpmc_assembly_2p(mixture.handle, grid.h, ctypes.byref(model), V(q.data_ptr()),
                                      V(branch.data_ptr()), V(prev.data_ptr()), float(dt), ctypes.byref(o),
                                      ctypes.byref(flag) if check_domain else None, _stream_ptr(stream)),
                 "assembly_2p")
    out["_inputs"] = (q, branch, prev)
    out["domain_error"] = int(flag.value) if check_domain else None
    return out


def default_secondary(nc: int) -> list:
    """The secondary q columns of mpmc_assembly_reduce_2p's default split: x_0 of component 0
    and the nc-1 independent compositions of phase 1."""
    return [3] + [3 + (nc - 1) + k for k in range(nc - 1)]


def reduce(grid: Grid, nc: int, assembled: dict, secondary=None, out: dict | None = None, stream=None) -> dict:
    """Schur reduction of an assembled two-phase system to (nc+1) x (nc+1) blocks
    (mpmc_assembly_reduce_2p): residual [nc+1, n], diag [(nc+1)^2, n], offdiag
    [(nc+1)^2, max_degree, n], back [nc, nc+2, n] (dx_S = -(M_P dx_P + M_r)), status [n]."""
    torch = _torch()
    lib = native.library()
    dev = assembled["residual"].device
    n = grid.ncell
    P = nc + 1
    if out is None:
        out = {
            "status": torch.empty(n, dtype=torch.int32, device=dev),
            "residual": torch.empty((P, n), dtype=torch.float64, device=dev),
            "diag": torch.empty((P * P, n), dtype=torch.float64, device=dev),
            "offdiag": torch.empty((P * P, max(1, grid.max_degree), n), dtype=torch.float64, device=dev),
            "back": torch.empty((nc, nc + 2, n), dtype=torch.float64, device=dev),
        }
    i = native.AssemblyOutputs(assembled["cell_status"].data_ptr(), assembled["residual"].data_ptr(),
                               assembled["diag"].data_ptr(), assembled["offdiag"].data_ptr())
    o = native.ReducedOutputs(out["status"].data_ptr(), out["residual"].data_ptr(), out["diag"].data_ptr(),
                              out["offdiag"].data_ptr(), out["back"].data_ptr())
    sec = None
    if secondary is not None:
        arr = (ctypes.c_int32 * nc)(*[int(v) for v in secondary])
        sec = ctypes.cast(arr, ctypes.c_void_p)
    native.check(lib.mpmc_assembly_reduce_2p(grid.h, int(nc), sec, ctypes.byref(i), ctypes.byref(o),
                                             _stream_ptr(stream)), "assembly_reduce_2p")
    return out


class NewtonPlan:
    """One Newton evaluation (mpmc_assembly_2p) on fixed device buffers, captured once into
    a HIP graph (mpmc_assembly_plan_*). Write the next iterate into ``q`` (and ``branch``,
    ``prev``) in place, then ``launch``; the outputs land in ``out``."""

    def __init__(self, mixture: Mixture, grid: Grid, model, q, branch, prev, dt: float, out: dict | None = None):
        torch = _torch()
        lib = native.library()
        dev = torch.device("cuda", torch.cuda.current_device())
        self.q = _dev(q, torch.float64, dev)
        self.branch = _dev(branch, torch.int32, dev)
        self.prev = _dev(prev, torch.float64, dev)
        self.out = out if out is not None else allocate(mixture.nc, grid, dev)
        o = native.AssemblyOutputs(self.out["cell_status"].data_ptr(), self.out["residual"].data_ptr(),
                                   self.out["diag"].data_ptr(), self.out["offdiag"].data_ptr())
        h = ctypes.c_void_p()
        native.check(lib.mpmc_assembly_plan_create(mixture.handle, grid.h, ctypes.byref(model),
                                                   V(self.q.data_ptr()), V(self.branch.data_ptr()),
                                                   V(self.prev.data_ptr()), float(dt), ctypes.byref(o),
                                                   ctypes.byref(h)), "assembly_plan_create")
        self.h = h

    def launch(self, stream=None):
        native.check(native.library().mpmc_assembly_plan_launch(self.h, _stream_ptr(stream)), "assembly_plan_launch")
        return self.out

    def domain_error(self, stream=None) -> int:
        flag = ctypes.c_int32(0)
        native.check(native.library().mpmc_assembly_plan_domain_error(self.h, _stream_ptr(stream), ctypes.byref(flag)),
                     "assembly_plan_domain_error")
        return int(flag.value)

    def close(self):
        if getattr(self, "h", None):
            native.library().mpmc_assembly_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def inventory(mixture: Mixture, grid: Grid, model, q, branch, stream=None):
    torch = _torch()
    lib = native.library()
    dev = torch.device("cuda", torch.cuda.current_device())
    q = _dev(q, torch.float64, dev)
    branch = _dev(branch, torch.int32, dev)
    status = torch.empty(grid.ncell, dtype=torch.int32, device=dev)
    inv = torch.empty((mixture.nc + 1, grid.ncell), dtype=torch.float64, device=dev)
    flag = ctypes.c_int32(0)
    native.check(lib.mpmc_inventory_2p(mixture.handle, grid.h, ctypes.byref(model), V(q.data_ptr()),
                                       V(branch.data_ptr()), V(status.data_ptr()), V(inv.data_ptr()),
                                       ctypes.byref(flag), _stream_ptr(stream)), "inventory_2p")
    return inv, status, int(flag.value)
