"""Host logic of the autotuner's wisdom file (no GPU): keys, persistence and the options
a stored choice sets."""
import json
import os

from omx import tuning


class _G:
    Nx, Ny = 256, 192


def test_wisdom_round_trip(tmp_path):
    path = str(tmp_path / "sub" / "wisdom.json")
    assert tuning.load(path) == {}
    choice = {k: i for i, k in enumerate(tuning.KEYS)}
    tuning.save({tuning._key(_G(), 8, 0): choice}, path)
    w = tuning.load(path)
    assert w == {"dev0:256x192x1:nev8": choice}
    assert json.load(open(path)) == w


def test_wisdom_keys_separate_grids_nev_and_devices():
    class G3:
        Nx, Ny, Nz = 128, 96, 24
    keys = {tuning._key(_G(), 8, 0), tuning._key(_G(), 4, 0), tuning._key(_G(), 8, 1), tuning._key(G3(), 8, 0)}
    assert len(keys) == 4
    assert tuning._key(G3(), 4, 0) == "dev0:128x96x24:nev4"


def test_choice_keys_are_solver_options():
    # every stored key is an option omx_set_option accepts, in the order omx_autotune's
    # choice array reports them (the header lists that order)
    assert tuning.KEYS == ("reg_ops", "fft_variant32", "tile_rows32", "tile_cols32", "gram_blocks",
                           "combine_blocks", "slab3d", "reg_rows32", "reg_nt2")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = open(os.path.join(root, "optimode.jl_amd", "csrc", "solver.hip")).read()
    hdr = open(os.path.join(root, "include", "omx.h")).read()
    for k in tuning.KEYS:
        assert f'"{k}"' in src and f'"{k}"' in hdr, k
    assert "9 ints in that order" in hdr
