"""The headless executables end to end on the GPU: render_path_tracing,
render_hair_msnn and render_nrc (the reference's three entry points,
render_path_tracing.cu:817-841, render_hair_msnn.cu:1146-1175,
render_nrc.cu:1080-1100) load a scene file, render, and write the EXR and
the stats JSON; the EXR holds exactly the image the library renders for the
same scene in this process."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.test_cli_exr import read_exr_uncompressed

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "hairmsnn_amd" / "bin"
SCENES = ROOT / "hairmsnn_amd" / "scenes"


def write_scene(tmp_path, name):
    cfg = json.loads((SCENES / "c0_hairball.json").read_text())
    cfg["integrator"].update({"width": 48, "height": 40, "image_output": str(tmp_path / f"{name}.png"),
                              "stats_output": str(tmp_path / f"{name}_stats.json")})
    cfg["hairmsnn"] = {"train_records": 512, "warmup_steps": 4, "learning_rate": 0.01}
    cfg["tcnn"] = {"config": "unused.json", "init_train": True}
    p = tmp_path / f"{name}.json"
    p.write_text(json.dumps(cfg))
    return p


def run(args):
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.parametrize("exe,extra", [("render_path_tracing", ["3"]), ("render_hair_msnn", ["1", "3"]),
                                       ("render_nrc", ["3"])])
def test_cli_writes_the_library_image(gpu, tmp_path, exe, extra):
    scene = write_scene(tmp_path, exe)
    run([BIN / exe, scene, *extra])
    exr = tmp_path / f"{exe}.exr"
    img = read_exr_uncompressed(exr)[::-1]  # file row 0 is the image's top row
    stats = json.loads((tmp_path / f"{exe}_stats.json").read_text())
    assert stats["width"] == 48 and stats["height"] == 40
    with gpu.Scene(path=str(scene)) as sc:
        if exe == "render_path_tracing":
            ref = sc.render_path_tracing(3)
        elif exe == "render_hair_msnn":
            ref = sc.render_hair_msnn(BETA=1, spp=3)
        else:
            ref = sc.render_nrc(3)
    assert img.shape == ref.shape
    np.testing.assert_array_equal(img[..., :3], ref[..., :3])
