"""The Fortran binding (a5_streamer_amd/fortran/m_a5g.f90) and the 2D, 3D and
cylindrical drivers built on it: every C ABI entry point is bound, failures
stop the program with the library's message, and on the GPU each Fortran
driver's run is bit-identical to the Python driver's over the same
library."""
import os
import re
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "a5_streamer_amd", "fortran")
EXE = os.path.join(FDIR, "streamer_2d_a5g")
EXE3 = os.path.join(FDIR, "streamer_3d_a5g")
EXECYL = os.path.join(FDIR, "streamer_cyl_a5g")


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


@pytest.mark.parametrize("exe", [EXE3, EXECYL])
def test_other_drivers_stop_with_message_without_gpu(exe):
    if _gpu_visible():
        pytest.skip("a GPU is visible")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    p = subprocess.run([exe, "4", "1"], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
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


def _compare_runs(exe, args, cfg, n_blocks):
    import a5_streamer_amd as A
    from a5_streamer_amd import streamer as S
    p = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    lines = [ln.split() for ln in p.stdout.splitlines() if ln.strip()]
    assert len(lines) == n_blocks + 1
    run = S.StreamerRun(A.Tree, cfg)
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
    run.T.close()


@pytest.mark.gpu
def test_fortran_3d_driver_matches_python_driver():
    """The 3D driver (src/streamer_3d.f90's step order): a 16 mm cube of one
    base box, 2.0 MV/m, the seed at (8, 8, 14) mm, six levels, 2 blocks."""
    from a5_streamer_amd import streamer as S
    print("fortran 3D driver", flush=True)
    L = 16e-3
    cfg = S.Config(variant="3d", domain_len=L, n_base=1, lvls_max=6, applied_fld=2.0e6,
                   seed_r0=[[0.5 * L, 0.5 * L, 0.875 * L]], seed_r1=[[0.5 * L, 0.5 * L, 0.875 * L]],
                   n_boxes=5000)
    _compare_runs(EXE3, [6, 2, repr(2.0e6), repr(L)], cfg, 2)


@pytest.mark.gpu
def test_fortran_cyl_driver_matches_python_driver():
    """The cylindrical driver (src/streamer_cyl.f90's step order, time inside
    the substeps, the top face's rule over the whole ghost row): a 16 mm
    square of one base box, seven levels, 4 blocks."""
    from a5_streamer_amd import streamer as S
    print("fortran cylindrical driver", flush=True)
    cfg = S.Config(variant="cyl", domain_len=16e-3, n_base=1, lvls_max=7, applied_fld=2.5e6,
                   seed_r0=[[0.0, 12e-3]], seed_r1=[[0.0, 12e-3]])
    _compare_runs(EXECYL, [7, 4, repr(2.5e6)], cfg, 4)
