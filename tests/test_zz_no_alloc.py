"""The allocation-free stage epilogue (`sttopt_stage_epilogue_noalloc`) against the
entry point that takes a caller scratch buffer (`sttopt_stage_epilogue`), through the C
ABI: the same compliances and adjoints, bit for bit, on two calls in a row with the
second on a smaller grid (the static scratch keeps the first call's partials)."""

import ctypes

import numpy as np
import pytest
import torch


def _lib():
    from sttopt_amd import _lib

    return _lib


def _case(nelx, nely, nstage, seed, dev):
    g = torch.Generator(device="cpu").manual_seed(seed)
    nel, ndof = nelx * nely, 2 * (nelx + 1) * (nely + 1)
    rnd = lambda *s: torch.rand(*s, generator=g, dtype=torch.float64)  # noqa: E731
    return {
        "nelx": nelx, "nely": nely, "ns": nstage,
        "x": (0.05 + 0.95 * rnd(nel)).to(dev),
        "t": rnd(nel).to(dev),
        "times": torch.arange(1, nstage + 1, dtype=torch.float64).div(nstage).to(dev),
        "fext": torch.randn(ndof, generator=g, dtype=torch.float64).to(dev),
        "u": torch.randn(1 + nstage, ndof, generator=g, dtype=torch.float64).to(dev),
        "g": (rnd(1 + nstage) + 0.5).to(dev),
    }


def _run(c, noalloc):
    L = _lib()
    lib = L.load()
    dev = c["x"].device
    stream = torch.cuda.current_stream(dev).cuda_stream
    nelx, nely, ns = c["nelx"], c["nely"], c["ns"]
    nel, ndof = nelx * nely, 2 * (nelx + 1) * (nely + 1)
    beta, emin, emax, penal, fe = 16.0, 1e-9, 1.0, 3.0, 1.0 / nel
    density = torch.empty(1 + ns, nel, dtype=torch.float64, device=dev)
    load = torch.empty(1 + ns, ndof, dtype=torch.float64, device=dev)
    L.check(lib.sttopt_stage_prologue(nelx, nely, ns, c["x"].data_ptr(), c["t"].data_ptr(),
                                      c["times"].data_ptr(), beta, emin, emax, penal, fe,
                                      c["fext"].data_ptr(), density.data_ptr(), load.data_ptr(),
                                      stream), "sttopt_stage_prologue")
    from sttopt_amd import fem

    ke = L.ke_array(torch.as_tensor(fem.plane_stress_KE(0.3)))
    comp = torch.full((1 + ns,), float("nan"), dtype=torch.float64, device=dev)
    dx = torch.full((nel,), float("nan"), dtype=torch.float64, device=dev)
    dt = torch.full((nel,), float("nan"), dtype=torch.float64, device=dev)
    common = (nelx, nely, ns, c["x"].data_ptr(), c["t"].data_ptr(), c["times"].data_ptr(),
              beta, emin, emax, penal, fe, ke, c["u"].data_ptr())
    if noalloc:
        st = lib.sttopt_stage_epilogue_noalloc(*common, density.data_ptr(), None, comp.data_ptr(),
                                               None, None, stream)
        L.check(st, "sttopt_stage_epilogue_noalloc")
        st = lib.sttopt_stage_epilogue_noalloc(*common, None, c["g"].data_ptr(), None,
                                               dx.data_ptr(), dt.data_ptr(), stream)
        L.check(st, "sttopt_stage_epilogue_noalloc")
    else:
        scratch = torch.empty(lib.sttopt_stage_scratch_count(nelx, nely, ns),
                              dtype=torch.float64, device=dev)
        L.check(lib.sttopt_stage_epilogue(*common, density.data_ptr(), None, comp.data_ptr(),
                                          None, None, scratch.data_ptr(), stream),
                "sttopt_stage_epilogue")
        L.check(lib.sttopt_stage_epilogue(*common, None, c["g"].data_ptr(), None,
                                          dx.data_ptr(), dt.data_ptr(), None, stream),
                "sttopt_stage_epilogue")
    torch.cuda.synchronize(dev)
    return comp.cpu().numpy(), dx.cpu().numpy(), dt.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_epilogue_equals_the_scratch_buffer_entry_point(dev):
    # two calls in a row, the second on a smaller grid with fewer stages
    for nelx, nely, ns, seed in ((300, 160, 9, 1), (40, 24, 3, 2)):
        print(f"noalloc epilogue {nelx}x{nely}, {ns} stages", flush=True)
        c = _case(nelx, nely, ns, seed, dev)
        want = _run(c, noalloc=False)
        got = _run(c, noalloc=True)
        for a, b in zip(got, want):
            assert np.all(np.isfinite(a))
            np.testing.assert_array_equal(a, b)
        # the compliances are the sums they claim to be (a check independent of both)
        assert np.all(want[0] > 0)


def test_noalloc_epilogue_refuses_sizes_beyond_its_static_scratch():
    """Refused on the host before any device call (the pointers are never touched)."""
    L = _lib()
    lib = L.load()
    ke = L.ke_array(torch.eye(8, dtype=torch.float64))
    cap = 262144
    nelx, nely, ns = 2048, 1024, 32  # 33 rows x 8192 blocks > cap
    assert lib.sttopt_stage_scratch_count(nelx, nely, ns) > cap
    st = lib.sttopt_stage_epilogue_noalloc(nelx, nely, ns, 4096, 8192, 12288, 16.0, 1e-9, 1.0,
                                           3.0, 1.0, ke, 16384, 20480, None, 24576, None, None,
                                           None)
    assert st == L.EINVAL and "static buffer holds 262144" in L.last_error()
    assert lib.sttopt_stage_scratch_count(nelx, nely, 31) <= cap
