"""The full-size scene configurations on the GPU against the oracle: plain
path tracing of the C1 scene (50,000 curly strands x 64 segments, the
procedural environment with importance sampling, the sphere head)
converging to the FP64 oracle, and HairMSNN's truncated paths on that
geometry with 2^16 training records for BETA in {1, 2, 5, 10}."""
import json
from pathlib import Path

import numpy as np
import pytest

from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu
SCENES = Path(__file__).resolve().parent.parent / "hairmsnn_amd" / "scenes"


def c1_cfg(w, h):
    cfg = json.loads((SCENES / "c1_curly.json").read_text())
    cfg["integrator"].update({"width": w, "height": h})
    return cfg


def oracle_for(sc):
    return ol.OracleScene(sc.info, sc.segments(), sc.env(), mesh=sc.mesh())


def test_c1_crop_converges_to_fp64_oracle(gpu):
    """A 24 x 24 crop of the C1 frame at 256 x 256 (hair over the head and
    over the environment): the per-pixel mean of 256 GPU passes against 256
    FP64 oracle passes on other seeds, z-scores from both sample variances."""
    n, W = 256, 256
    x0, y0, cw = 100, 120, 24
    with gpu.Scene(json_text=c1_cfg(W, W)) as sc:
        assert int(sc.info["num_segments"]) == 3_200_000
        sc.pt_reset()
        passes = []
        for _ in range(n):
            sc.pt_passes(1)
            passes.append(sc.last_pass()[y0:y0 + cw, x0:x0 + cw, :3].astype(np.float64))
        orc = oracle_for(sc)
    g = np.stack(passes)
    o = np.stack([orc.render(x0, y0, cw, cw, spp=1, accum_id0=70_000 + k, prec=0)[0] for k in range(n)])
    _, _, hit = orc.render(x0, y0, cw, cw, spp=1, accum_id0=0, prec=1)
    print(f"crop: hair {np.mean(hit == 1):.2f}, head {np.mean(hit == 3):.2f}, env {np.mean(hit == 0):.2f}")
    assert np.mean(hit == 1) > 0.3
    diff = g.mean(0) - o.mean(0)
    se = np.sqrt(g.var(0, ddof=1) / n + o.var(0, ddof=1) / n) + 1e-12
    zz = (diff / se)[se > 1e-9]
    print(f"z: mean {zz.mean():.3f} std {zz.std():.3f} |z|>4: {(np.abs(zz) > 4).mean():.4f}")
    assert abs(zz.mean()) < 0.15
    assert 0.8 < zz.std() < 1.25
    assert (np.abs(zz) > 4).mean() < 0.01
    se_m = np.sqrt(g.mean((1, 2, 3)).var(ddof=1) / n + o.mean((1, 2, 3)).var(ddof=1) / n)
    assert abs(g.mean() - o.mean()) < 4 * se_m


@pytest.mark.parametrize("beta,need", [(1, 0.80), (2, 0.76), (5, 0.72), (10, 0.70)])
def test_c2_truncated_paths_on_c1_geometry(gpu, beta, need):
    """HairMSNN's G_BUFFER pass on the C1 scene at 512 x 512 with 2^16
    training records (one training pixel in four): every pixel that is not a
    training pixel carries the truncated path, which is a plain path-tracing
    pass limited to BETA + 1 vertices bit for bit (the kernel break after
    vertex BETA, hair_msnn.cu:166-179); a training pixel carries its full
    path (quirk Q4). Against the oracle's msnnPathTrace / msnnTrainingPath
    on a 32 x 32 crop, per pixel."""
    W = 512
    cfg = c1_cfg(W, W)
    cfg["hairmsnn"].update({"train_records": 1 << 16, "warmup_steps": 2})
    cfg_t = json.loads(json.dumps(cfg))
    cfg_t["integrator"]["path_v2"] = beta + 1
    with gpu.Scene(json_text=cfg_t) as st:
        st.pt_reset()
        st.pt_passes(1)
        trunc = st.last_pass()[..., :3].astype(np.float64)
    with gpu.Scene(json_text=cfg) as sc:
        sc.pt_reset()
        sc.pt_passes(1)
        full = sc.last_pass()[..., :3].astype(np.float64)
        sc.msnn_begin(BETA=beta)
        sc.msnn_frames(1)
        short = sc.msnn_image(1)[..., :3].astype(np.float64)
        z = sc.msnn_sizes()
        pix = sc.msnn_buffer("train_pixel")
        orc = oracle_for(sc)
    assert z["train_records"] == 1 << 16 and z["every_nth"] == 4
    train = np.zeros(W * W, bool)
    train[pix] = True
    train = train.reshape(W, W)
    np.testing.assert_array_equal(short[~train], trunc[~train])
    np.testing.assert_array_equal(short[train], full[train])
    x0, y0, cw = 240, 240, 32
    o_s, _, _ = orc.render(x0, y0, cw, cw, spp=1, accum_id0=0, mode=1, kbeta=beta - 1, prec=1)
    o_f, _, _ = orc.render(x0, y0, cw, cw, spp=1, accum_id0=0, mode=2, kbeta=beta - 1, prec=1)
    g = short[y0:y0 + cw, x0:x0 + cw]
    tc = train[y0:y0 + cw, x0:x0 + cw]
    ref = np.where(tc[..., None], o_f, o_s)
    close = (np.abs(g - ref) <= 1e-3 * np.maximum(np.abs(ref), 1e-2)).all(-1)
    print(f"BETA {beta}: per-pixel agreement with the oracle {close.mean():.4f} "
          f"(training pixels {tc.mean():.3f})")
    assert close.mean() >= need
