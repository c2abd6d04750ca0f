"""GPU tests of the mixed form's allocation variants (MIXED_RECON64,
MIXED_FLUX64, MIXED_FLUX64_NARROW): each is the mixed step with one FP32
part in FP64, so on a short Orszag-Tang run each lies within FP32 rounding of
FP64; the variant whose flux and differences are both FP64 is the closest;
the configurations they do not run are refused."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(mhd, prec, nx=96, ny=80, steps=40):
    s = mhd.Solver(mhd.SolverConfig(nx=nx, ny=ny, precision=prec))
    s.init_problem("ot")
    st = s.advance(0.5, steps)
    u = s.get_state(0)
    s.close()
    assert st.code == 0 and st.steps == steps
    return u


def test_variants_near_fp64(mhd):
    from mhdglm import compare

    u64 = run(mhd, mhd.PREC_FP64)
    errs = {}
    for name in ("mixed", "mixed_recon64", "mixed_flux64", "mixed_flux64_narrow"):
        u = run(mhd, mhd.PRECISIONS[name])
        e = compare.rel_l2_state(u64, u, ("rho", "p", "|B|"))
        errs[name] = max(e.values())
        say(f"{name}: {e}")
        assert 0 < errs[name] < 1e-6, name
    assert errs["mixed_flux64"] < errs["mixed"]
    assert errs["mixed_flux64"] < errs["mixed_flux64_narrow"]


@pytest.mark.parametrize("kw", [{"muscl": False}, {"hlld": False}, {"tile": 0}, {"fuse": 2}])
def test_variants_refuse_other_configurations(mhd, kw):
    cfg = mhd.SolverConfig(nx=32, ny=32, precision=mhd.PREC_MIXED_FLUX64, **kw)
    with pytest.raises(mhd.MhdError):
        mhd.Solver(cfg)
    with pytest.raises(mhd.MhdError):
        mhd.Ensemble(mhd.SolverConfig(nx=2, ny=2, precision=mhd.PREC_MIXED_FLUX64), [(32, 32)])
