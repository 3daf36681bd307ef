"""The direct-sum kernel's arctangent polynomials (fastlinkingnumbers_amd/
csrc/ds.hip) against their stated error bounds, re-derived on the CPU.

The coefficients are read from the kernel source itself; their maximum
relative error against numpy's arctangent is measured on a dense grid in
FP64 arithmetic and with the kernel's FP32 Horner evaluation (fused
multiply-adds emulated in FP64), and the committed fitting script
(csrc/fit_atan_poly.py, a Remez exchange) must reproduce a polynomial at
least as good."""
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fastlinkingnumbers_amd", "csrc")
sys.path.insert(0, CSRC)
import fit_atan_poly as F  # noqa: E402

NUM = r"(-?\d+\.\d+(?:e-?\d+)?)f"


def _kernel_coefficients():
    src = open(os.path.join(CSRC, "ds.hip")).read()
    body = src[src.index("float atan_unit("):]
    body = body[: body.index("return t * p;")]
    unit = [float(v) for v in re.findall(NUM, body)]
    short = src[src.index("const float tt = zi[c]"):]
    short = short[: short.index("accf = fmaf(tt, p, accf);")]
    small = [float(v) for v in re.findall(NUM, short)]
    # Horner form: p = fma(s2, c4, c3); p = fma(s2, p, c2); ... ; last c0 = 1.0f
    return unit, small


def test_kernel_coefficients_meet_their_stated_bounds():
    unit, small = _kernel_coefficients()
    assert len(unit) == 8 and len(small) == 5 and small[-1] == 1.0
    e, bias = F.max_rel_error(unit, 1.0)
    e32, _ = F.max_rel_error(unit, 1.0, fp32=True)
    print("atan_unit", e, e32, bias)
    assert e < 1.3e-7 and e32 < 2.5e-7 and abs(bias) < 0.25 * e
    e, bias = F.max_rel_error(small, 16.0 / 63.0)
    e32, bias32 = F.max_rel_error(small, 16.0 / 63.0, fp32=True)
    print("small-angle", e, e32, bias)
    assert e < 1.2e-9 and e32 < 1e-7
    # the series' own bias is far below the FP32 rounding of one evaluation
    assert abs(bias) < 4e-10 and abs(bias32) < 1e-9


def test_fitting_script_reproduces_the_polynomials():
    unit, small = _kernel_coefficients()
    c = F.remez(1.0, 7, grid=40001)
    fit = [float(np.float32(v)) for v in c[::-1]]
    np.testing.assert_allclose(fit, unit, rtol=2e-5, atol=2e-7)
    assert F.max_rel_error(fit, 1.0)[0] < 1.3e-7
    c = F.remez(16.0 / 63.0, 4, grid=40001)
    assert F.max_rel_error(list(c[::-1]), 16.0 / 63.0)[0] < 3e-10
    assert F.max_rel_error([float(np.float32(v)) for v in c[::-1]], 16.0 / 63.0)[0] < 1.2e-9
