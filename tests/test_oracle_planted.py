"""Planted mistakes: each one is written into a copy of the oracle
(oracle/refcpu.cpp), compiled on the CPU, and the pins that guard that part
of the restatement must fail against it (and pass against the real oracle,
which the rest of the suite checks).  This shows the closed-form pins of the
specular and transmissive transport, camera, envmap and roulette code are
sharp enough to catch a plausible slip, not only that the oracle agrees with
itself.  Nothing here runs on a GPU."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "oracle" / "refcpu.cpp"
SPEC = "tests/test_oracle_specular.py"

# (name, text in refcpu.cpp, replacement, occurrence (0-based), pins that must fail)
MUTANTS = [
    ("refract: inverted eta", "f32 eta = e1 / e2;", "f32 eta = e2 / e1;", 0,
     "test_refraction_into_glass_follows_snell or test_leaving_glass"),
    ("refract: TIR threshold", "if (st2 > 1.0f) {", "if (st2 > 1.05f) {", 0, "test_leaving_glass"),
    ("ior stack: far side read at the top", "return (s->cur == 0) ? AIR_IOR : s->ior[s->cur - 1];",
     "return s->ior[s->cur];", 0, "test_leaving_glass or test_ior_stack"),
    ("ior stack: no pop on exit", "        if (s->cur > 0)\n            s->cur--;", "", 0, "test_leaving_glass"),
    ("dielectric Fresnel: r0 not squared", "return fresnel(q * q, i, n);", "return fresnel(phx_fabsf(q), i, n);", 0,
     "test_refraction_into_glass_follows_snell"),
    ("vndf: alpha = roughness", "f32 alpha = roughness * roughness;\n    V4 t, b;", "f32 alpha = roughness;\n    V4 t, b;",
     0, "test_ggx_vndf_moments"),
    ("vndf: stretch sign", "V4 nl = normalize(mk(alpha * nh.x, alpha * nh.y,", "V4 nl = normalize(mk(-alpha * nh.x, alpha * nh.y,",
     0, "test_ggx_vndf_moments"),
    ("metallic weight: G1 of the view", "s.throughput = fresnel4(c.base_color, -view, h) * g1;",
     "s.throughput = fresnel4(c.base_color, -view, h) * smith_g1_ggx(dot(view, sn), alpha);", 0,
     "test_metallic_weight"),
    ("transmit weight: no base colour", "s.throughput = c.base_color * g1;", "s.throughput = mk(g1, g1, g1, 0.0f);", 0,
     "test_fresnel_and_transmit_weights or test_refraction_into_glass_follows_snell"),
    ("specular NEE: one G1 term", "f32 sm = smith_g1_ggx(f.n_dot_dir, alpha) * smith_g1_ggx(ndv, alpha);",
     "f32 sm = smith_g1_ggx(f.n_dot_dir, alpha);", 0, "test_specular_nee"),
    ("specular NEE: D with alpha = roughness", "f32 ggx = ggx_ndf(ndh, alpha);", "f32 ggx = ggx_ndf(ndh, roughness);", 0,
     "test_specular_nee"),
    ("trace: FRESNEL NEE with the base colour as F0", "V4 f0 = mk(r0, r0, r0, 0.0f);", "V4 f0 = c.base_color;", 0,
     "test_specular_nee"),
    ("area light NEE: no emitter cosine", "rad = ef * ndd * cl / (PI * dist_sq);", "rad = ef * ndd / (PI * dist_sq);", 0,
     "test_area_light"),
    ("envmap: y up", "f32 theta = phx_acosf(phx_clampf(dir.z, -1.0f, 1.0f));\n    f32 phi = phx_atan2f(dir.y, dir.x);",
     "f32 theta = phx_acosf(phx_clampf(dir.y, -1.0f, 1.0f));\n    f32 phi = phx_atan2f(dir.z, dir.x);", 0,
     "test_envmap"),
    ("envmap: rows flipped", "f32 v = phx_clampf(1.0f - theta / PI, 0.0f, 1.0f);",
     "f32 v = phx_clampf(theta / PI, 0.0f, 1.0f);", 0, "test_envmap"),
    ("camera: row 0 at the bottom", "f32 t = 1.0f - ((f32)py + 0.5f + v) / (f32)height;",
     "f32 t = ((f32)py + 0.5f + v) / (f32)height;", 0, "test_camera_rays"),
    ("camera: lens radius without sqrt", "f32 r1 = phx_sqrtf(random_float(rng));", "f32 r1 = random_float(rng);", 0,
     "test_camera_rays"),
    ("photon roulette: no 1/q boost", "            power = power * iq;\n", "", 0, "test_russian_roulette_is_unbiased"),
    ("photon roulette: mean instead of max",
     "f32 q = phx_fmaxf(throughput.x, phx_fmaxf(throughput.y, throughput.z));",
     "f32 q = (throughput.x + throughput.y + throughput.z) / 3.0f;", 1, "test_russian_roulette_never_culls"),
    ("camera roulette: no 1/q boost", "throughput = throughput * (1.0f / q);", "throughput = throughput * 1.0f;", 0,
     "test_camera_path_roulette"),
    ("camera roulette: from depth 1", "if ((u32)depth >= RR_MIN_RAY_DEPTH) {", "if ((u32)depth >= 1u) {", 0,
     "test_camera_path_roulette"),
]


def _mutate(text: str, old: str, new: str, occurrence: int) -> str:
    pos = -1
    for _ in range(occurrence + 1):
        pos = text.find(old, pos + 1)
        assert pos >= 0, f"mutation site not found: {old!r}"
    return text[:pos] + new + text[pos + len(old):]


def _compiler():
    return os.environ.get("CXX") or shutil.which("g++")


def _run_mutant(i: int, tmp: Path):
    name, old, new, occ, pins = MUTANTS[i]
    src = tmp / f"mutant{i}.cpp"
    lib = tmp / f"libmutant{i}.so"
    src.write_text(_mutate(SRC.read_text(), old, new, occ))
    cc = subprocess.run([_compiler(), "-O1", "-ffp-contract=off", "-fno-fast-math", "-fno-math-errno", "-fopenmp",
                         "-fPIC", "-std=c++17", "-shared", "-iquote", str(ROOT / "oracle"), "-o", str(lib), str(src)],
                        capture_output=True, text=True)
    if cc.returncode != 0:
        return name, "compile", cc.stderr[-2000:]
    env = dict(os.environ, PHOSPHOR_ORACLE_LIB=str(lib), OMP_NUM_THREADS="2")
    res = subprocess.run([sys.executable, "-m", "pytest", SPEC, "-x", "-q", "-p", "no:cacheprovider", "-k", pins],
                         cwd=ROOT, env=env, capture_output=True, text=True)
    return name, res.returncode, res.stdout[-1500:]


@pytest.mark.skipif(_compiler() is None, reason="no C++ compiler")
def test_planted_mistakes_fail_their_pins(tmp_path):
    for _, old, new, occ, _ in MUTANTS:               # every site exists in the oracle as written
        _mutate(SRC.read_text(), old, new, occ)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        results = list(pool.map(lambda i: _run_mutant(i, tmp_path), range(len(MUTANTS))))
    survivors = [(name, rc, out) for name, rc, out in results if rc != 1]
    # pytest exits 1 when tests ran and failed; 0 would mean the mistake went unnoticed
    assert not survivors, "\n\n".join(f"{n}: exit {rc}\n{out}" for n, rc, out in survivors)
