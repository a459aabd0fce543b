"""The C ABI from C and C++, and the C++ interface.

tests/abi/capi_consumer (built by tests/abi/Makefile from
__graft_entry__.build()) includes include/mnl.h, links libmnl.so and the
oracle, and calls the entry points the way the reference's C++ core would
(neighbour_list_gpu_device, the request form, the plan API).
tests/abi/cpp_api_test runs the scenarios of the reference's GPU test
(tests/test_neighbour_list_gpu.cc:95-233: dense, few large cells, larger
cutoff, Morton order, a cluster in a 40x vacuum box on the hashed grid, the
K = 80 matrix and coordination) through include/mnl.hpp, the reference's C++
GPU API (neighbour_list_gpu, neighbour_list_gpu_device, ...).

CPU: the headers compile on their own (mnl.h as C99 and C++11, mnl.hpp as
C++17), and the checks that need no device pass. GPU: every build equals the
oracle bit for bit (the reference compares sorted multisets to 1e-12)."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "tests", "abi", "capi_consumer")
CPP_TEST = os.path.join(ROOT, "tests", "abi", "cpp_api_test")


def _need_binary(path=BINARY):
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built: run __graft_entry__.build() "
                    "(make -C tests/abi) first")


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_compiles_alone(tmp_path, lang):
    cc = "gcc" if lang == "c" else "g++"
    if shutil.which(cc) is None:
        pytest.skip(f"{cc} not available")
    src = tmp_path / ("h.c" if lang == "c" else "h.cc")
    src.write_text('#include "mnl.h"\nint main(void) { return MNL_SUCCESS; }\n')
    std = "-std=c99" if lang == "c" else "-std=c++11"
    r = subprocess.run([cc, std, "-Wall", "-Wextra", "-pedantic", "-Werror",
                        "-I", os.path.join(ROOT, "include"), "-fsyntax-only", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_header_compiles_alone(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "h.cc"
    src.write_text('#include "mnl.hpp"\nint main() { mnl::NeighbourListDevice d; '
                   'return d.npairs; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-pedantic", "-Werror",
                        "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                        "-D__HIP_PLATFORM_AMD__", "-fsyntax-only", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_api_without_device():
    _need_binary(CPP_TEST)
    r = subprocess.run([CPP_TEST, "--no-gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_api_on_device(gpu):
    _need_binary(CPP_TEST)
    print("running tests/abi/cpp_api_test", flush=True)
    r = subprocess.run([CPP_TEST], capture_output=True, text=True, timeout=180)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout


def test_consumer_without_device():
    _need_binary()
    r = subprocess.run([BINARY, "--no-gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "checks passed" in r.stdout


@pytest.mark.gpu
def test_consumer_on_device(gpu):
    _need_binary()
    print("running tests/abi/capi_consumer", flush=True)
    r = subprocess.run([BINARY], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "checks passed" in r.stdout
