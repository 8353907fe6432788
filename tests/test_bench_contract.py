"""bench.py's contract on the CPU: it measures one GPU only and says so for any other
count, and every traffic figure a bench line quotes comes from a committed rocprofv3
summary that states it."""
import ast
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("gpus", ["2", "8"])
def test_more_than_one_gpu_is_refused_with_a_message(gpus):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", gpus], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == int(gpus) and "one GPU" in line["error"]


def _traffic_names():
    """The profile names bench.py passes to read_traffic, literal or formatted."""
    tree = ast.parse(open(os.path.join(ROOT, "bench.py")).read())
    names = []
    for node in ast.walk(tree):
        if isinstance(node, ast.Call) and getattr(node.func, "id", "") == "read_traffic" and node.args:
            a = node.args[0]
            if isinstance(a, ast.Constant):
                names.append(a.value)
            elif isinstance(a, ast.JoinedStr):  # f-strings: the workload sizes bench runs by default
                names.append(a)
    return names


def test_quoted_traffic_profiles_exist_and_state_bytes_per_launch():
    names = _traffic_names()
    assert names
    # the formatted names at the sizes the bench lines quote
    literal = [n for n in names if isinstance(n, str)]
    literal += ["c2c_large_2e20", "c2c_large_2e22", "c2c_n1024_b64"]
    for name in literal:
        path = os.path.join(ROOT, "profiles", f"{name}_traffic.txt")
        assert os.path.exists(path), name
        vals = [l for l in open(path) if l.startswith("traffic_bytes_per_launch:")]
        assert len(vals) == 1, name
        assert float(vals[0].split(":", 1)[1]) > 0, name
