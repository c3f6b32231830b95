"""mpmc_flash_batch_static, the flash entry point that allocates nothing (its workspace is
a static device array): through the C ABI, two calls in a row (the second on fewer cells of
another workload) give records bit-identical to mpmc_flash_batch on the same inputs, and
a batch beyond the static scratch is rejected before any device work."""
import ctypes

import numpy as np
import pytest

from mpmc_hnu_amd import workloads as W

KEYS = ("status", "phase_count", "beta", "zfactor", "branch", "x", "ln_phi", "counters")


def _static(m, p, t, z, precision):
    import torch

    from mpmc_hnu_amd import flash, native

    lib = native.library()
    dev = torch.device("cuda", 0)
    pd, td, zd = (torch.as_tensor(np.ascontiguousarray(a), device=dev) for a in (p, t, z))
    out = flash.allocate_outputs(m.nc, p.size, dev)
    opts = flash._options(1e-11, 3, precision)
    o = flash._outputs_struct(out)
    V = ctypes.c_void_p
    rc = lib.mpmc_flash_batch_static(m.handle, p.size, V(pd.data_ptr()), V(td.data_ptr()), V(zd.data_ptr()),
                                     ctypes.byref(opts), ctypes.byref(o),
                                     V(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, native.library().mpmc_last_error()
    torch.cuda.synchronize()
    return flash.to_host(out)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_static_scratch_flash_matches_flash_batch(gpu, precision):
    from mpmc_hnu_amd import flash, native

    mix = W.mixture_linspaced(6)
    m = flash.Mixture.from_dict(mix)
    try:
        cases = [("C2 grid prefix", *W.cells_c2(ncell=32768)), ("uniform", *W.cells_uniform(6, 5000, seed=91))]
        for name, p, t, z in cases:
            print(f"[case] static-scratch flash, {name}, {p.size} cells, precision {precision}", flush=True)
            a = _static(m, p, t, z, precision)
            b = flash.to_host(flash.flash_batch(m, p, t, z, precision=precision))
            for k in KEYS:
                assert np.array_equal(a[k], b[k], equal_nan=True), (name, k)
        # beyond the static scratch: rejected with an argument error, no device work
        lib = native.library()
        n_big = 1
        while lib.mpmc_flash_workspace_bytes(6, n_big) <= 256 * 1024 * 1024:
            n_big *= 2
        opts = flash._options(1e-11, 3, precision)
        rc = lib.mpmc_flash_batch_static(m.handle, n_big, None, None, None, ctypes.byref(opts),
                                         ctypes.byref(native.FlashOutputs()), None)
        assert rc == 1 and b"static scratch" in lib.mpmc_last_error(), (rc, lib.mpmc_last_error())
    finally:
        m.close()
