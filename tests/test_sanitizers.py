"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU code (no GPU needed): the
oracle restatement (oracle/linon_oracle.cpp) under its own pin tests and golden fixtures, and the
library's host code, the WGSL translator (wgsl.cpp) and the host grid build (grid.cpp), driven
over the benchmark field bodies (and every prefix of them) and over several meshes.  Any
sanitizer report fails the run (-fno-sanitize-recover=all)."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from linon_amd import configs, scenes

ROOT = Path(__file__).resolve().parent.parent
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _runtime(name: str) -> str:
    p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True, check=True).stdout.strip()
    if not os.path.isabs(p):
        pytest.skip(f"{name} not found")
    return p


def _write_mesh(path: Path, v: np.ndarray, f: np.ndarray) -> None:
    v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(f, np.uint32).reshape(-1, 3)
    path.write_bytes(np.array([len(v), len(f)], np.uint32).tobytes() + v.tobytes() + f.tobytes())


def test_host_code_under_sanitizers(tmp_path):
    exe = tmp_path / "host_driver"
    csrc = ROOT / "linon_amd" / "csrc"
    # the host code includes the HIP runtime header only for __host__ __device__ qualifiers
    subprocess.run(["g++", "-std=c++17", "-O1", *SAN, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{ROOT / 'include'}", f"-I{csrc}", str(ROOT / "tests" / "sanitize" / "host_driver.cpp"),
                    str(csrc / "wgsl.cpp"), str(csrc / "grid.cpp"), "-pthread", "-o", str(exe)], check=True)
    args = []
    bodies = list(configs.FIELD_WGSL.values()) + [
        "var acc = vec3<f32>(0.0);\nfor (var i = 0; i < 4; i++) { if i == 2 { continue; } acc += p.zyx * f32(i); }\n"
        "var k = 0u; while (k < 2u) { k += 1u; }\nlet q = i32(t) / 3 % 2;\n"
        "if (t > 1.0) { return acc; } else { return v * f32(q); }",
    ]
    for i, b in enumerate(bodies):
        p = tmp_path / f"body{i}.wgsl"
        p.write_text(b)
        args.append(f"wgsl:{p}")
    rng = np.random.default_rng(0)
    tri = rng.normal(size=(60, 3)).astype(np.float32)
    tri[10:20] = tri[0]                                   # degenerate (repeated) vertices
    meshes = [scenes.cornell_box_centered(), scenes.benchmark_scene(level=2),
              (tri, rng.integers(0, 60, size=(80, 3)).astype(np.uint32)),
              (np.zeros((3, 3), np.float32), np.array([[0, 1, 2]], np.uint32))]   # one zero-area face
    for i, (v, f) in enumerate(meshes):
        p = tmp_path / f"mesh{i}.bin"
        _write_mesh(p, v, f)
        args.append(f"mesh:{p}")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "translated" in r.stdout and "grids 16" in r.stdout, r.stdout


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize"], check=True)
    env = dict(os.environ, LINON_ORACLE_SANITIZE="1",
               LD_PRELOAD=f"{_runtime('libasan.so')}:{_runtime('libubsan.so')}",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "tests/test_oracle_pins.py", "tests/test_golden_fixtures.py"],
                       cwd=ROOT, capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert " passed" in r.stdout
