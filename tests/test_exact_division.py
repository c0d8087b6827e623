"""The corner gradients' divisions by 2 dx on the device are computed from
the correctly rounded reciprocal with one FMA correction (div_by in
tmc_amd/csrc/dmc_extract_kernels.h).  tests/cpp/exact_div_check.c restates
it in C and checks it against the IEEE division bit for bit: differences of
float pairs, moderate values, arbitrary bit patterns, and the special
values the device hands to the division itself, for fixed divisors and for
random ones (the normals' division by their length uses it too; CPU
only)."""
import os
import subprocess

from conftest import ROOT


def test_reciprocal_division_is_exact(tmp_path):
    src = os.path.join(ROOT, "tests", "cpp", "exact_div_check.c")
    exe = str(tmp_path / "exact_div_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
    # the spacings of the benchmark fields (2 dx) and others
    cs = ["0.049087385212340517", "0.0078125", "0.0046875", "0.046875", "0.1", "2.0", "1.7", "3e-7", "12345.678",
          "rand", "rand"]
    r = subprocess.run([exe, "2000000", "11", *cs], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("0 of "), r.stdout
