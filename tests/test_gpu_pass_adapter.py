"""The C++ PathTracingPass adapter (bisemutum-engine_amd/pass/) driven like the engine's
renderer, frame after frame with accumulation; its demo checks a closed-form image."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "bisemutum-engine_amd", "pass_demo")

pytestmark = pytest.mark.gpu


def test_pass_adapter_closed_form():
    print("case: pass_demo (C++ adapter, 8 accumulated frames)", flush=True)
    assert os.path.exists(DEMO), "pass_demo not built (run __graft_entry__.build())"
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "samples accumulated 512" in r.stdout
