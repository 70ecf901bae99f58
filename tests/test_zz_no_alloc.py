"""bf_force_eval_device_noalloc (no device allocation; scratch in a static device array)
against bf_force_eval_device through the C ABI: bit for bit, in every precision, on two
calls in a row, the second with a smaller mesh, and a mesh beyond the static maximum is
refused with BF_ERR_UNSUPPORTED.  Collected last: it runs after every other GPU test and
alone in a fresh process alike."""
import ctypes as C

import numpy as np
import pytest
import torch

from biofem_amd import _lib as L
from biofem_amd import configs
from biofem_amd.fem import ActStrainFE

pytestmark = pytest.mark.gpu

vp = C.c_void_p


def _eval(fe, entry, bufs, prec, stream):
    bx, by, bz, f = bufs
    rc = entry(fe.ctx, vp(bx.data_ptr()), vp(by.data_ptr()), vp(bz.data_ptr()), vp(f.data_ptr()),
               0, prec, vp(stream.cuda_stream))
    assert rc == L.BF_OK, rc
    stream.synchronize()
    return f.cpu().numpy().copy()


def _bufs(x, y, z, dev):
    t = [torch.as_tensor(a, device=dev) for a in (x, y, z)]
    return (*t, torch.zeros(3 * len(x), dtype=torch.float64, device=dev))


def test_noalloc_matches_device_entry_point():
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    big = configs.membrane_large()                       # 98 304 elements, noisy state
    m2, p2, _ = configs.membrane_small()
    small = (m2, p2, configs.smooth_displacement(m2, 0.01))   # 1 536 elements
    lib = L.lib()
    for name, (m, p, (x, y, z)) in (("256x192", big), ("32x24", small)):
        print(f"case no-alloc {name}", flush=True)
        fe = ActStrainFE(m, p)
        try:
            bufs = _bufs(x, y, z, dev)
            for prec in (L.BF_PREC_FP64, L.BF_PREC_FP64_FAST, L.BF_PREC_MIXED):
                ref = _eval(fe, lib.bf_force_eval_device, bufs, prec, stream)
                n_ref = fe.fallback_count()
                c33_ref = fe.c33()
                got = _eval(fe, lib.bf_force_eval_device_noalloc, bufs, prec, stream)
                assert np.array_equal(got, ref), (name, prec)
                assert fe.fallback_count() == n_ref
                assert np.array_equal(fe.c33(), c33_ref)
                again = _eval(fe, lib.bf_force_eval_device_noalloc, bufs, prec, stream)
                assert np.array_equal(again, ref), (name, prec)
        finally:
            fe.close()


def test_noalloc_refuses_meshes_beyond_the_static_maximum():
    m, p, (x, y, z) = configs.ellipsoid_shell(levels=6)    # 184 320 elements
    print("case no-alloc too large", flush=True)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    fe = ActStrainFE(m, p)
    try:
        bx, by, bz, f = _bufs(x, y, z, dev)
        rc = L.lib().bf_force_eval_device_noalloc(fe.ctx, vp(bx.data_ptr()), vp(by.data_ptr()),
                                                  vp(bz.data_ptr()), vp(f.data_ptr()), 0,
                                                  L.BF_PREC_FP64_FAST, vp(stream.cuda_stream))
        assert rc == L.BF_ERR_UNSUPPORTED
    finally:
        fe.close()
