"""Device builds of include/mtx_fmath.h return the host build's bits."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_device_matches_host(orc):
    import torch
    assert torch.cuda.is_available()
    from mutracer_amd import abi
    rng = np.random.default_rng(5)
    n = 1 << 20
    xs = {
        "sin": (rng.uniform(-50, 50, n), None), "cos": (rng.uniform(-50, 50, n), None),
        "pow": (rng.uniform(0, 4, n), rng.uniform(-8, 8, n)), "atan2": (rng.uniform(-3, 3, n), rng.uniform(-3, 3, n)),
        "acos": (rng.uniform(-1, 1, n), None), "exp": (rng.uniform(-100, 88, n), None),
        "sincos.s": (rng.uniform(-50, 50, n), None), "sincos.c": (rng.uniform(-50, 50, n), None)}
    for kind, (x, y) in xs.items():
        x = x.astype(np.float32)
        y = np.zeros_like(x) if y is None else y.astype(np.float32)
        x[:4] = [0.0, -0.0, 1.0, -1.0]
        host = orc.math_eval(kind, x, y)
        dx, dy = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
        out = torch.empty_like(dx)
        k = ["sin", "cos", "pow", "atan2", "acos", "exp", "sincos.s", "sincos.c"].index(kind)
        abi.check(abi.lib().mtx_math_eval(k, C.c_void_p(dx.data_ptr()), C.c_void_p(dy.data_ptr()),
                                          C.c_void_p(out.data_ptr()), n, None))
        torch.cuda.synchronize()
        dev = out.cpu().numpy()
        assert np.array_equal(dev.view(np.uint32), host.view(np.uint32)), kind
