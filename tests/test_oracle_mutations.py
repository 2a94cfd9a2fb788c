"""Each oracle pin must catch a plausible slip in the line it pins: a copy
of the oracle with one planted mistake (a dropped term, a swapped sign, a
wrong index) is built on the CPU and the pin tests are run against it, and
they must fail; the unmodified copy must pass. The mistakes are planted in
temporary copies only."""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
ORC = ROOT / "oracle"

# (name, file, text, replacement, occurrence, pytest -k expression that must fail)
MUTANTS = [
    ("env bilinear without the half-texel offset", "orc_scene.cpp",
     "const R xb = x * (R)e.W - (R)0.5, yb = y * (R)e.H - (R)0.5;",
     "const R xb = x * (R)e.W, yb = y * (R)e.H - (R)0.5;", 0, "env_radiance"),
    ("env pdf without sin(theta)", "orc_scene.cpp",
     "const R denom = ((R)2 * Pi<R>() * Pi<R>() * std::sin(theta)) * last_u * last_v;",
     "const R denom = ((R)2 * Pi<R>() * Pi<R>()) * last_u * last_v;", 0, "constant_env or integrates"),
    ("env pdf reads the next row's marginal", "orc_scene.cpp",
     "cv = e.mpdf[iv];", "cv = e.mpdf[std::min(iv + 1, e.H - 1)];", 0, "env_pdf_integrates or env_sampling"),
    ("env sample at the row centre", "orc_scene.cpp",
     "const R v = ((R)iv + (cnv - (R)uy) / (cnv - cv)) / (R)e.H;",
     "const R v = ((R)iv + (R)0.5) / (R)e.H;", 0, "env_sampling"),
    ("GGX D without pi", "orc_math.h",
     "return (R)1 / (Pi<R>() * a * a * std::pow(", "return (R)1 / (a * a * std::pow(", 0, "ggx_ndf"),
    ("GGX separable masking-shadowing", "orc_math.h",
     "return (R)1 / ((R)1 + ggxLambda(a, V) + ggxLambda(a, L));", "return ggxG1(a, V) * ggxG1(a, L);", 0,
     "ggx_sample_weight"),
    ("hair frame Y = X x wo", "orc_scene.cpp",
     "const Vec<R> Y = nrm(cross(si.wo, X));", "const Vec<R> Y = nrm(cross(X, si.wo));", 0, "hair_hit_geometry"),
    ("hit point not moved onto the fibre surface", "orc_scene.cpp",
     "o1 = (r / len(o1)) * o1;", "o1 = ((R)0.999 * r / len(o1)) * o1;", 1, "hair_hit_geometry"),
    ("curve: Catmull-Rom u^2 coefficient with -4 q1 (FP32)", "orc_curve.cpp",
     "k.b = ((2.f * q0 - 5.f * q1) + (4.f * q2 - q3)) * 0.5f;",
     "k.b = ((2.f * q0 - 4.f * q1) + (4.f * q2 - q3)) * 0.5f;", 0, "catmull_rom"),
    ("curve: Catmull-Rom tangent (q2 - q1) / 2 (FP64)", "orc_curve.cpp",
     "0.5 * (q2 - q0), q1};", "0.5 * (q2 - q1), q1};", 0, "catmull_rom"),
    ("curve: closest approach to the chord with its sign flipped", "orc_curve.cpp",
     "(dd * ew - de * dw) / den", "(de * dw - dd * ew) / den", 0, "hair_like_strands"),
    ("curve: secant step away from the root", "orc_curve.cpp",
     "du = -f * (u - up) / (f - fp);", "du = f * (u - up) / (f - fp);", 0, "hair_like_strands or cylinder"),
    ("curve: normal without the radius derivative's sign", "orc_scene.cpp",
     "n = nrm(dd * o1 - (dr * r) * dv);", "n = nrm(dd * o1 + (dr * r) * dv);", 0, "hit_frame"),
    ("curve: hit point not moved onto the swept circle", "orc_scene.cpp",
     "o1 = (r / len(o1)) * o1;", "o1 = ((R)0.999 * r / len(o1)) * o1;", 0, "hit_frame"),
    ("R lobe tilted by +2 alpha", "orc_math.h",
     "sOp = sinThetaO * hs.cos2k[1] - cosThetaO * hs.sin2k[1];",
     "sOp = sinThetaO * hs.cos2k[1] + cosThetaO * hs.sin2k[1];", 0, "lobe_tilts"),
    ("TT lobe tilted by 2 alpha", "orc_math.h",
     "sOp = sinThetaO * hs.cos2k[0] + cosThetaO * hs.sin2k[0];",
     "sOp = sinThetaO * hs.cos2k[1] + cosThetaO * hs.sin2k[1];", 0, "lobe_tilts"),
    ("MIS light pdf without the light-pick probability (Q1)", "orc_scene.cpp",
     "const R lightPdf = env_pdf<R>(c.s, si.wi) * (R)1 / (R)(c.s.ndl + (c.s.has_env ? 1 : 0));",
     "const R lightPdf = env_pdf<R>(c.s, si.wi);", 0, "q1_q2"),
    ("BSDF sample without the light sample's push-through (Q2)", "orc_scene.cpp",
     "const Vec<R> ls = sample_lights(c, si, rng, isDelta, r1);",
     "Inter<R> si_l = si;\n    const Vec<R> ls = sample_lights(c, si_l, rng, isDelta, r1);", 0, "q1_q2"),
]


def _plant(dst: Path, fname: str, old: str, new: str, occurrence: int) -> None:
    p = dst / fname
    s = p.read_text()
    k = -1
    for _ in range(occurrence + 1):
        k = s.index(old, k + 1)
    p.write_text(s[:k] + new + s[k + len(old):])


def _build(tmp: Path, i: int, mutant) -> Path:
    d = tmp / f"m{i}"
    d.mkdir()
    for f in ORC.iterdir():
        if f.is_file():
            shutil.copy(f, d / f.name)
    if mutant is not None:
        _plant(d, *mutant[1:5])
    subprocess.run(["make", "-s", "-C", str(d)], check=True, stdout=subprocess.DEVNULL)
    return d / "_build" / "liborc.so"


def _run_pins(lib: Path, kexpr: str | None) -> subprocess.CompletedProcess:
    env = dict(os.environ, HMSNN_ORACLE_LIB=str(lib), OMP_NUM_THREADS="2")
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", str(ROOT / "tests" / "test_oracle_pins.py"),
           str(ROOT / "tests" / "test_oracle_curves.py")]
    if kexpr:
        cmd += ["-k", kexpr]
    return subprocess.run(cmd, cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=600)


def test_pins_catch_planted_mistakes(tmp_path):
    if shutil.which("make") is None or shutil.which("g++") is None:
        pytest.skip("no C++ toolchain")
    jobs = [None, *MUTANTS]
    with ThreadPoolExecutor(min(8, os.cpu_count() or 4)) as ex:
        libs = list(ex.map(lambda a: _build(tmp_path, *a), enumerate(jobs)))
    with ThreadPoolExecutor(4) as ex:
        runs = list(ex.map(lambda a: _run_pins(a[0], a[1][5] if a[1] else None), zip(libs, jobs)))
    base = runs[0]
    assert base.returncode == 0, "unmodified copy must pass:\n" + base.stdout[-3000:]
    missed = [m[0] for m, r in zip(MUTANTS, runs[1:]) if r.returncode == 0]
    assert not missed, f"planted mistakes not caught: {missed}"
    # every failure is a test failure, not a crash or a collection error
    for m, r in zip(MUTANTS, runs[1:]):
        assert r.returncode == 1, (m[0], r.stdout[-2000:])
