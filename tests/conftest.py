import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libintc kernels)")


def have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(autouse=True)
def _stack_dump_if_stuck(request):
    """A GPU test still running after 100 s prints every thread's Python stack
    once (names the call it waits in: a kernel, a copy or the oracle)."""
    if request.node.get_closest_marker("gpu") is None:
        yield
        return
    import faulthandler
    faulthandler.dump_traceback_later(100, repeat=False)
    try:
        yield
    finally:
        faulthandler.cancel_dump_traceback_later()


@pytest.fixture(scope="session")
def gpu():
    if not have_gpu():
        pytest.skip("no GPU")
    from intception_amd import lib
    L = lib.load()
    lib.check(L.intc_init())
    return L


def assert_parity(got, ref, rtol=1e-12, atol=1e-14, what=""):
    """|got - ref| <= atol + rtol |ref| elementwise (the FP64 parity bar)."""
    import numpy as np
    got = np.asarray(got); ref = np.asarray(ref)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    err = np.abs(got - ref)
    lim = atol + rtol * np.abs(ref)
    bad = np.nonzero(~(err <= lim))[0]
    if len(bad):
        i = bad[np.argmax(err[bad] / lim[bad])]
        raise AssertionError("%s: %d/%d elements outside |d| <= %g + %g|ref|; worst at %d: got %r ref %r"
                             % (what, len(bad), len(ref), atol, rtol, i, got[i], ref[i]))
