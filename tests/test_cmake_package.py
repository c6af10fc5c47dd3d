"""The CMake package (CMakeLists.txt at the repository root): configure, build and
install quarry::gpu / quarry::core into a temporary prefix, then configure and
build the reference's examples/CMakeLists.txt unchanged against it through
find_package(quarry 1.0 CONFIG REQUIRED).  Needs the reference sources (for
quarry::core and the examples); the GPU test in test_dropin.py runs the examples."""
import os
import shutil
import sys
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


def _cmake(*args, cwd=None):
    r = subprocess.run(["cmake", *args], capture_output=True, text=True, cwd=cwd)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r.stdout


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "examples")), reason="reference sources not present")
@pytest.mark.skipif(shutil.which("cmake") is None, reason="cmake not installed")
def test_install_and_find_package(tmp_path):
    build, prefix, ex = tmp_path / "build", tmp_path / "prefix", tmp_path / "examples"
    _cmake("-S", ROOT, "-B", str(build), f"-DQUARRY_CORE_SOURCE_DIR={REF}", "-DCMAKE_BUILD_TYPE=Release")
    _cmake("--build", str(build), "-j8")
    _cmake("--install", str(build), "--prefix", str(prefix))
    cfg = prefix / "lib" / "cmake" / "quarry"
    for f in ("quarryConfig.cmake", "quarryConfigVersion.cmake", "quarryTargets.cmake"):
        assert (cfg / f).exists(), f
    for h in ("gpu/gpu_bvh.h", "quarry/quarry_abi.h", "core/geometry.h", "cpu/bvh.h"):
        assert (prefix / "include" / "quarry" / h).exists(), h
    out = _cmake("-S", os.path.join(REF, "examples"), "-B", str(ex), f"-DCMAKE_PREFIX_PATH={prefix}")
    assert "GPU backend" in out  # quarry::gpu was found
    _cmake("--build", str(ex), "-j8")
    for prog in ("quarry_example", "dynamic_scene"):
        assert os.access(ex / prog, os.X_OK), prog
    ldd = subprocess.run(["ldd", str(ex / "quarry_example")], capture_output=True, text=True).stdout
    assert str(prefix / "lib" / "libquarry_hip.so") in ldd


@pytest.mark.skipif(shutil.which("cmake") is None, reason="cmake not installed")
def test_wheel_installs_the_python_module(tmp_path):
    """pip builds a wheel of the quarry module (scikit-build-core over CMakeLists.txt,
    the library compiled for gfx950), and the installed module loads its own copy."""
    wheels, site = tmp_path / "wheels", tmp_path / "site"
    r = subprocess.run([sys.executable, "-m", "pip", "wheel", ROOT, "--no-deps", "--no-build-isolation",
                        "-w", str(wheels)], capture_output=True, text=True,
                       env={**os.environ, "SKBUILD_BUILD_DIR": str(tmp_path / "build")})
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    whl = next(wheels.glob("quarry_gpu-*.whl"))
    r = subprocess.run([sys.executable, "-m", "pip", "install", "--no-deps", "--target", str(site), str(whl)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    code = ("import quarry, quarry._abi as A; print(A.LIB_PATH); "
            "print(sorted(n for n in A.exported_symbols() if n.startswith('qr_'))[:3])")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=str(tmp_path),
                       env={**os.environ, "PYTHONPATH": str(site)})
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stdout.splitlines()[0] == str(site / "quarry" / "libquarry_hip.so")
