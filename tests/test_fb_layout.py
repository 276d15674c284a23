"""Host logic of the forward-backward checkpoint layout (no GPU): every launch
shape the run driver, the batch entry points, the tuner and a GH_FB_SHAPE pin
can produce passes the host guard and writes no more checkpoint floats than
the buffers sized for the run hold; launches that would overrun are refused."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_checkpoint_layout_guard(tmp_path):
    exe = str(tmp_path / "fb_layout_check")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "glimpse_amd", "csrc"),
                           "-o", exe, os.path.join(ROOT, "tests", "native", "fb_layout_check.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "failures 0" in out.stdout
