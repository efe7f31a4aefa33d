"""The cell order of a box slot (a5_streamer_amd/csrc/a5g_layout.hpp), which
the kernels and the host conversion share, checked on the host: a bijection
onto the slot, colour halves, interior cells first."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_box_layout(tmp_path):
    exe = str(tmp_path / "layout_check")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-o", exe,
                           os.path.join(HERE, "layout_check.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
