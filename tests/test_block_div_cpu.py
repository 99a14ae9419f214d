"""Block coordinates of the block kernels come from multiply-shift quotients by the block counts
(mpm_math.hpp: block_div_magic, block_div), built here for the host from the same header. Checked against
integer division for every block count up to 65536 and a spread of larger ones, at the numerators where the
quotient changes (q d - 1, q d, q d + 1), at 0 and at 2^31 - 1, together with the multiplier's bounds
m d - 2^(31+l) in [0, d) and d <= 2^l."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "block_div_check.cpp")
CLANG = "/opt/rocm/llvm/bin/clang++"


def test_block_div_matches_integer_division(tmp_path):
    exe = str(tmp_path / "block_div_check")
    subprocess.run([CLANG, "-O2", "-std=c++17", SRC, "-o", exe], check=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout[-2000:]
    assert "fails 0" in res.stdout
