"""The correctly rounded square root and reciprocal of the HIP kernels (csrc/cr_math.hpp) rest on two
claims about their correction steps, checked exhaustively on the CPU by tests/c/rn_math_check.c:
from every start within 1 ulp (what v_sqrt_f32 / v_rcp_f32 guarantee), for every float mantissa, the
neighbour selection returns RN(sqrt(x)) and the Newton step with its all-ones fix-up returns RN(1/s)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_correction_steps_are_exact(tmp_path):
    exe = str(tmp_path / "rn_math_check")
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", os.path.join(HERE, "c", "rn_math_check.c"),
                    "-lm", "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "reciprocal: " in r.stdout and ", 0 misses" in r.stdout.splitlines()[-2]
    assert r.stdout.splitlines()[-1].endswith(", 0 misses")
