"""Drop-in checks of the C++ API (include/gpu/gpu_bvh.h).

The reference's own consumers of gpu/gpu_bvh.h -- examples/main.cpp,
examples/dynamic_scene.cpp and bench/gpu_{,query_,cloud_,gauss_}bench.cpp -- are
compiled unchanged against this repository's header and linked against
libquarry_hip.so by `make -C oracle dropin` (outputs in oracle/_ref/dropin/).
Their built-in --verify passes compare every GPU result with the reference's CPU
BVH / kd-tree / Gaussian BVH and exit non-zero on the first mismatch, so running
them on the GPU is parity checked by the reference's own test code.

CPU tests: the API surface (every function the reference header declares exists
here) and that the programs were built.  GPU tests: run the programs.
"""
from __future__ import annotations

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpu", "gpu_bvh.h")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin")
# the reference's examples/CMakeLists.txt built unchanged through find_package(quarry)
# against this repository's installed CMake package (make -C oracle cmake_examples)
CMAKE_EXAMPLES = os.path.join(ROOT, "oracle", "_ref", "cmake", "examples")
REF_HEADER = "/root/reference/src/gpu/gpu_bvh.h"


def _functions(text: str) -> set[str]:
    # free functions returning a value or void at the start of a line
    pat = re.compile(r"^(?:inline\s+)?[A-Za-z_][\w:<>,\s\*&]*?\b([a-z]\w*GPU)\s*\(", re.M)
    return set(pat.findall(text))


@pytest.mark.skipif(not os.path.exists(REF_HEADER), reason="reference sources not present")
def test_api_surface_matches_reference():
    ref = _functions(open(REF_HEADER, encoding="utf-8").read())
    ours = _functions(open(HEADER, encoding="utf-8").read())
    assert len(ref) >= 40, sorted(ref)
    missing = sorted(ref - ours)
    assert not missing, f"functions of the reference API missing here: {missing}"


def test_header_declares_reference_types():
    text = open(HEADER, encoding="utf-8").read()
    for name in ("struct GpuBVH", "struct GpuBuildTimings", "struct GpuRayBuffer", "struct GpuRefitTimings",
                 "struct GpuTopoEdit", "struct GpuTopoUpdateTimings", "enum class RayOrder",
                 "struct GpuTraversalStats", "struct GpuClosestResult", "struct GpuPointBuffer",
                 "struct GpuQueryStats", "enum class QueryOrder", "struct GpuRadiusBuffer"):
        assert name in text, name


@pytest.mark.skipif(not os.path.exists(REF_HEADER), reason="reference sources not present")
def test_dropin_programs_built():
    for prog in ("example_main", "example_dynamic_scene", "gpu_bench", "gpu_query_bench",
                 "gpu_cloud_bench", "gpu_gauss_bench"):
        assert os.access(os.path.join(DROPIN, prog), os.X_OK), prog


def _run(prog: str, *args: str, timeout: int = 120, where: str = DROPIN) -> str:
    path = os.path.join(where, prog)
    if not os.access(path, os.X_OK):
        pytest.skip(f"{prog} not built (make -C oracle dropin needs the reference sources)")
    print(f"[dropin] {prog} {' '.join(args)}", flush=True)
    r = subprocess.run([path, *args], capture_output=True, text=True, timeout=timeout)
    out = r.stdout + r.stderr
    print(out[-3000:], flush=True)
    assert r.returncode == 0, f"{prog} exited {r.returncode}:\n{out[-2000:]}"
    return out


@pytest.mark.gpu
def test_example_main():
    out = _run("example_main")
    assert "MISMATCH" not in out
    assert "matches cpu" in out


@pytest.mark.gpu
def test_example_dynamic_scene():
    out = _run("example_dynamic_scene")
    assert "checks ok" in out
    counts = re.findall(r"mismatch(?:es)? (\d+)/\d+", out)
    assert counts and all(c == "0" for c in counts), counts
    assert "validate failures 0" in out


@pytest.mark.gpu
def test_cmake_package_examples():
    """The reference's examples as its own CMake project builds them (find_package(quarry)
    + quarry::gpu), linked against the installed package's libquarry_hip."""
    out = _run("quarry_example", where=CMAKE_EXAMPLES)
    assert "MISMATCH" not in out and "matches cpu" in out
    out = _run("dynamic_scene", where=CMAKE_EXAMPLES)
    assert "checks ok" in out and "validate failures 0" in out


FAST = ("--runs", "1", "--warmup-ms", "0")


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [(), ("--wide",), ("--ray-sort",), ("--wide", "--ray-sort-dir"),
                                   ("--scene", "city", "--buildings", "4096", "--wide")])
def test_gpu_bench_rays(extra):
    out = _run("gpu_bench", "--grid", "128", "--rays", "50000", *FAST, "--verify", *extra)
    assert "hit_mismatch=0 t_mismatch=0" in out


@pytest.mark.gpu
@pytest.mark.parametrize("query,extra", [
    ("closest", ()), ("closest", ("--wide", "--query-sort")), ("knn", ("--k", "16")),
    ("knn", ("--k", "32", "--wide")), ("radius", ("--radius", "2.0")),
    ("radius", ("--radius", "3.0", "--wide", "--query-sort")),
    ("closest", ("--scene", "mixed", "--random-tris", "2000")),
    ("closest", ("--refit",)), ("knn", ("--refit", "--wide", "--k", "8")),
    ("closest", ("--topo", "--topo-rounds", "2")),
])
def test_gpu_query_bench(query, extra):
    _run("gpu_query_bench", "--grid", "96", "--points", "4000", "--query", query, *FAST, "--verify", *extra)


@pytest.mark.gpu
@pytest.mark.parametrize("query,extra", [("closest", ()), ("knn", ("--k", "8", "--wide")),
                                         ("radius", ("--radius", "1.0")), ("closest", ("--refit",))])
def test_gpu_cloud_bench(query, extra):
    _run("gpu_cloud_bench", "--points", "60000", "--queries", "3000", "--query", query, "--runs", "1",
         "--verify", *extra)


@pytest.mark.gpu
@pytest.mark.parametrize("query,extra", [("closest", ()), ("knn", ("--k", "8", "--wide")),
                                         ("radius", ("--radius", "0.5")), ("closest", ("--refit",))])
def test_gpu_gauss_bench(query, extra):
    _run("gpu_gauss_bench", "--gaussians", "30000", "--queries", "2000", "--query", query, "--runs", "1",
         "--verify", *extra)
