"""Output path and command line (no device): the PNG writer against a numpy
restatement of saveImage (src/app/save_image.cpp:32-84, src/image.cpp:23-43),
and the CLI's argument rules (src/main.cpp:71-147)."""
import ctypes as C
import os

import numpy as np
import pytest

from luxion_amd import _ffi as F
from luxion_amd import cli


def _expected(acc, spp):
    """float32 restatement: / spp, ACES, pow 0.45, clamp * 255, truncate,
    column x -> w - 1 - x."""
    f = np.float32
    p = (acc / f(spp)).astype(np.float32)
    a, b, c, d, e = f(2.51), f(0.03), f(2.43), f(0.59), f(0.14)
    num = p * (a * p + b)
    den = p * (c * p + d) + e
    v = np.clip((num / den).astype(np.float32), f(0), f(1))
    v = np.power(v, f(0.45)).astype(np.float32)
    v = (np.clip(v, f(0), f(1)) * f(255.0)).astype(np.float32)
    return v.astype(np.uint8)[:, ::-1, :]


def test_png_output_matches_restatement(tmp_path):
    from PIL import Image
    lib = F.load()
    rng = np.random.default_rng(3)
    h, w, spp = 45, 61, 7
    acc = (np.exp(rng.normal(-1, 1.5, (h, w, 3))) * spp).astype(np.float32)
    acc[3, 0] = [1e6, 0, 0]  # one saturated red pixel in column 0
    path = str(tmp_path / "o.png")
    assert lib.lx_debug_write_png(path.encode(), acc.ctypes.data, w, h, spp) == 0
    got = np.asarray(Image.open(path).convert("RGB"))
    want = _expected(acc, spp)
    assert got.shape == (h, w, 3)
    d = np.abs(got.astype(int) - want.astype(int))
    assert d.max() <= 1 and (d > 0).mean() < 1e-3  # powf vs numpy float32 pow: rare last-bit differences
    assert tuple(got[3, w - 1]) == (255, 0, 0)  # mirrored horizontally


def test_png_writer_rejects_bad_path(tmp_path):
    lib = F.load()
    acc = np.ones((2, 2, 3), np.float32)
    assert lib.lx_debug_write_png(str(tmp_path / "no" / "dir" / "x.png").encode(), acc.ctypes.data, 2, 2, 1) == 4


@pytest.mark.parametrize("argv,msg", [
    (["luxion"], "Usage: luxion SCENEFILE"),
    (["luxion", "a.glb", "-e"], "Error: -e requires a path"),
    (["luxion", "a.glb", "--iterations"], "Error: --iterations requires a number"),
    (["luxion", "a.glb", "-i", "0"], "Error: invalid iteration count '0'"),
    (["luxion", "a.glb", "-i", "-4"], "Error: invalid iteration count '-4'"),
    (["luxion", "a.glb", "-i", "12x"], "Error: invalid iteration count '12x'"),
    (["luxion", "a.glb", "-o"], "Error: -o requires a name"),
    (["luxion", "a.glb", "b.glb"], "Error: unexpected argument 'b.glb'"),
])
def test_cli_argument_errors(argv, msg):
    with pytest.raises(cli.UsageError) as e:
        cli.parse_args(argv)
    assert str(e.value).startswith(msg)


def test_cli_arguments_and_defaults():
    a = cli.parse_args(["luxion", "-dmis", "s.glb", "--envmap", "e.exr", "-i", "64", "-emis", "-lock", "-o", "out"])
    assert a == dict(scene="s.glb", envmap="e.exr", iterations=64, output="out", dmis=True, emis=True, lock=True)
    b = cli.parse_args(["luxion", "s.glb"])
    assert b["iterations"] == 5000 and b["output"] is None and not b["dmis"]
    assert cli.default_output("room", "2026-01-02_03-04-05z", 5000) == "img/room.2026-01-02_03-04-05z.5000samp"
    assert cli.default_output("GLTF_DEFAULT_NAME", "t", 1000000) == "img/GLTF_DEFAULT_NAME.t.1e+06samp"


def test_cli_rejects_json_and_other_scenes(capsys):
    assert cli.main(["luxion", "old.json"]) == 1
    assert "deprecated" in capsys.readouterr().out
    assert cli.main(["luxion", "scene.obj"]) == 1
