"""The blend uses fh_expf_nonpos, a range-specialised exp; it must equal the
oracle's fh_expf bit for bit on the blend's argument range. (Run with stride 1
this covers all 1,115,684,865 floats in [-64, 0]; the suite samples every 7th.)"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_exp_fast_path_matches_expf(tmp_path):
    exe = tmp_path / "exp_check"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fno-fast-math",
                    os.path.join(HERE, "exp_fastpath_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "7"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
