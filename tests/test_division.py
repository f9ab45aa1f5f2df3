"""Arithmetic shortcuts of the kernels, checked exhaustively on the CPU (they are plain
IEEE single-precision identities, so the CPU's results are the GPU's):
  * the guided filter's first pass divides the covariance by var + eps (in [2^-7, 2)) with the steps of the
    IEEE division sequence without its range scaling, which must equal IEEE division from
    the correctly rounded reciprocal, and from its neighbours everywhere but at the 8 values
    of b named in tests/div_mid_check.c (the GPU's estimate there is checked on the GPU);
  * the forward DWT converts |c| < 2^30 to a multiple of 2^-32 with a float multiply, add and
    truncation, which must equal the general exponent/mantissa conversion
    (tests/q32_check.c)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_short_q32_conversion_matches_general(tmp_path):
    exe = str(tmp_path / "q32_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "q32_check.c"),
                    "-lm"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "0 mismatches" in out.stdout


def test_covariance_division_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "div_mid_check")
    subprocess.run(["gcc", "-O2", "-march=native", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "div_mid_check.c"), "-lm"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    assert "mismatches from RN(1/b) 0," in out.stdout and "(0 at other b)" in out.stdout
