"""tecno_audit_state_noalloc (the allocation-free audit) through the C ABI
against tecno_audit_state on the same states: every field bit for bit, on
two calls in a row with a smaller grid the second time (the device array
keeps the first call's rows), for the FP64 2D Euler state and the FP32
scalar state."""
import numpy as np
import pytest

from tecno import FP32, Solver, make_cfg

pytestmark = pytest.mark.gpu

FIELDS = ("total_entropy", "kinetic_energy", "mass", "energy", "max_speed", "n_nonfinite",
          "n_sign_fallback")


def _euler(nx, ny, seed):
    rng = np.random.default_rng(seed)
    s = Solver(make_cfg(n=(nx, ny, 2), dims=2, recon="sp_weno", bc="periodic"))
    prim = np.stack([rng.uniform(0.7, 1.3, nx * ny), rng.uniform(-0.3, 0.3, nx * ny),
                     rng.uniform(-0.3, 0.3, nx * ny), rng.uniform(0.7, 1.3, nx * ny)])
    s.set_state(prim)
    s.run_steps(2)
    return s


def _scalar(n, seed):
    rng = np.random.default_rng(seed)
    s = Solver(make_cfg(system="scalar", flux="lin_tecno4", recon="sp_weno", n=(n, 1, 1), dims=1,
                        lo=(-np.pi, 0, 0), hi=(np.pi, 1, 1), precision=FP32))
    s.set_state(rng.normal(size=n))
    s.run_steps(2)
    return s


def _same(a, b):
    for f in FIELDS:
        assert getattr(a, f) == getattr(b, f), (f, getattr(a, f), getattr(b, f))
    assert list(a.momentum) == list(b.momentum)


@pytest.mark.parametrize("make,sizes", [(_euler, [(300, 200), (40, 30)]), (_scalar, [(5000,), (200,)])])
def test_noalloc_audit_matches_audit(make, sizes):
    for k, size in enumerate(sizes):
        s = make(*size, seed=k)
        a = s.audit()
        b = s.audit_noalloc()
        _same(a, b)
        # and again in the opposite order on the same state
        _same(s.audit_noalloc(), s.audit())
        s.close()
