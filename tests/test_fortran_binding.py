"""The Fortran binding (a5_streamer_amd/fortran/m_a5g.f90) and the 2D driver
built on it: every C ABI entry point is bound, failures stop the program
with the library's message, and on the GPU the Fortran driver's run is
bit-identical to the Python driver's over the same library."""
import os
import re
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "a5_streamer_amd", "fortran")
EXE = os.path.join(FDIR, "streamer_2d_a5g")


def _abi_names():
    with open(os.path.join(ROOT, "include", "a5g.h")) as f:
        return set(re.findall(r"^(?:int|const char \*)\s*(a5g_\w+)\(", f.read(), re.M))


def test_binding_covers_every_entry_point():
    with open(os.path.join(FDIR, "m_a5g.f90")) as f:
        bound = set(re.findall(r'bind\(C, name="(a5g_\w+)"\)', f.read()))
    names = _abi_names()
    assert len(names) >= 40
    assert bound == names


def _gpu_visible():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_driver_stops_with_message_without_gpu():
    if _gpu_visible():
        pytest.skip("a GPU is visible")
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built")
    p = subprocess.run([EXE, "5", "1"], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
    assert "a5g error in set_cc" in p.stdout + p.stderr
    assert "no GPU" in p.stdout + p.stderr


def _hex(x):
    return "%016X" % struct.unpack("<Q", struct.pack("<d", x))[0]


@pytest.mark.gpu
@pytest.mark.parametrize("lvls_max,n_blocks,fld", [(7, 4, 2.5e6), (8, 3, 1.0e7)])
def test_fortran_driver_matches_python_driver(lvls_max, n_blocks, fld):
    import a5_streamer_amd as A
    from a5_streamer_amd import streamer as S
    print("fortran driver", lvls_max, n_blocks, fld, flush=True)
    p = subprocess.run([EXE, str(lvls_max), str(n_blocks), repr(fld)], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    lines = [ln.split() for ln in p.stdout.splitlines() if ln.strip()]
    assert len(lines) == n_blocks + 1

    run = S.StreamerRun(A.Tree, S.Config(lvls_max=lvls_max, applied_fld=fld))
    run.initialise()

    def row(blk, n_add, n_rm):
        T = run.T
        return [str(blk), _hex(run.time), _hex(run.dt), str(T.info()["max_id"]), str(n_add),
                str(n_rm), _hex(T.tree_minmax(S.I_ELEC)[1]), _hex(T.tree_minmax(S.I_FLD)[0]),
                _hex(T.tree_minmax(S.I_FLD)[1])]

    assert lines[0] == row(0, 0, 0)
    for blk in range(1, n_blocks + 1):
        n_add, n_rm = run.step_block()
        assert lines[blk] == row(blk, n_add, n_rm), blk
