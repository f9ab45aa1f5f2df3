"""Arithmetic shortcuts of the kernels, checked exhaustively on the CPU (they are plain
IEEE single-precision identities, so the CPU's results are the GPU's):
  * the guided-filter kernels divide window sums by their count as a * RN(1/c) plus one
    exact-residual correction, which must equal IEEE division for every count the 9x9
    window can have (tests/div_check.c);
  * the forward DWT converts |c| < 2^30 to a multiple of 2^-32 with a float multiply, add and
    truncation, which must equal the general exponent/mantissa conversion
    (tests/q32_check.c)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_count_division_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "div_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "div_check.c"),
                    "-lm"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "0 mismatches" in out.stdout


def test_short_q32_conversion_matches_general(tmp_path):
    exe = str(tmp_path / "q32_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "q32_check.c"),
                    "-lm"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "0 mismatches" in out.stdout
