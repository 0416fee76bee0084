"""bench.py on the GPU: the roofline's blend evaluations come from the timed
stereo frame (left + right eye), whether or not the plain leg, which renders
the eyes one at a time into the same handle slots, runs after it."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _line(*extra):
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--gaussians", "400000", "--steps", "2", "--warmup", "1",
           "--latency-frames", "5", "--no-cpu-baseline", "--no-configs", "--no-batched", "--no-serving", *extra]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_bench_evaluations_are_the_stereo_frames():
    alone = _line("--no-plain")
    full = _line()
    assert alone["blend_evaluations_per_eye"] == full["blend_evaluations_per_eye"]
    left, right = full["blend_evaluations_per_eye"]
    assert left > 0 and right > 0 and left != right
    if full["roofline"]["kernel"] == "blend_kernel":
        assert full["roofline"]["evaluations"] == left + right
        assert full["roofline"]["bound"] == "valu"
    assert full["latency_frames"] == 5 and full["plain_fps"] > 0
