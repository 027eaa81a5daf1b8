"""The C++ drop-in header compiles against the C ABI alone (no HIP headers)
and the reference README's call sequence behaves as documented."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "readme_example")


def build_example():
    lib_dir = os.path.join(ROOT, "gmapf_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "readme_example.cpp"), "-L", lib_dir,
                    "-lgmapf_amd", f"-Wl,-rpath,{lib_dir}", "-o", EXE], check=True)


def test_readme_example_host_calls():
    build_example()
    out = subprocess.run([EXE, "host"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "host ok" in out.stdout


@pytest.mark.gpu
def test_readme_example_runs_on_gpu():
    if not os.path.exists(EXE):
        build_example()
    out = subprocess.run([EXE, "gpu"], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "gpu ok" in out.stdout
