"""The correctly rounded logarithm of the FP64 step (fesom-c_amd/csrc/
fcg_logcr.h, used by d3_end's bottom boundary value), compiled for the host
from the same header, and the oracle's own C restatement of it: both equal
to the double nearest the exact logarithm (Python's decimal at 60 digits) on
arguments over the whole double range and in the range d3_end meets
(delta_b/z0b_min between 1 and a few thousand), and to glibc's log wherever
glibc rounds correctly (all but a few in 10^4 there)."""
import math
import os
import random
import subprocess
from decimal import Decimal, getcontext

import numpy as np

from conftest import ROOT

DRIVER = r"""
#include <cstdio>
#include "fcg_logcr.h"
int main(int argc, char **argv) {
    FILE *f = std::fopen(argv[1], "rb"), *g = std::fopen(argv[2], "wb");
    double x;
    while (std::fread(&x, 8, 1, f) == 1) { double y = fcglog::log_cr(x); std::fwrite(&y, 8, 1, g); }
    std::fclose(f); std::fclose(g);
    return 0;
}
"""


def test_log_is_correctly_rounded(tmp_path):
    src = tmp_path / "drv.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "drv"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-I", os.path.join(ROOT, "fesom-c_amd", "csrc"),
                    str(src), "-o", str(exe)], check=True)
    rng = random.Random(7)
    xs = [rng.uniform(1.0, 4000.0) for _ in range(40000)]
    xs += [math.exp(rng.uniform(-700.0, 700.0)) for _ in range(20000)]
    xs += [1.0 + rng.uniform(-1e-3, 1e-3) for _ in range(5000)]
    xs += [1.0, 2.0, 0.5, 10.0, 1e300, 3e-300, math.sqrt(2.0), math.sqrt(0.5)]
    a = np.array(xs)
    a.tofile(tmp_path / "x.bin")
    subprocess.run([str(exe), str(tmp_path / "x.bin"), str(tmp_path / "y.bin")], check=True)
    y = np.fromfile(tmp_path / "y.bin")
    getcontext().prec = 60
    exact = np.array([float(Decimal(v).ln()) for v in xs])   # decimal -> float rounds to nearest
    assert np.array_equal(y, exact), int((y != exact).sum())
    from oracle.pyoracle import lib
    orc = np.array([lib().orc_log_cr(v) for v in xs])
    assert np.array_equal(orc, exact), int((orc != exact).sum())
    glibc = np.array([math.log(v) for v in xs])
    agree = glibc == exact
    assert np.array_equal(y[agree], glibc[agree])
    assert agree.mean() > 0.999
