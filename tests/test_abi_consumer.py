"""The C ABI from C and C++: tests/abi/capi_consumer (built by
tests/abi/Makefile from __graft_entry__.build()) includes include/mnl.h, links
libmnl.so and the oracle, and calls the entry points the way the reference's
C++ core would (neighbour_list_gpu_device, the request form, the plan API).

CPU: the header compiles as C99 and C++11 on its own, and the consumer's
checks that need no device pass. GPU: the one-shot, request (Morton order,
own stream), plan and asynchronous plan builds of two systems equal the
oracle bit for bit, including a capacity short of the pair total."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "tests", "abi", "capi_consumer")


def _need_binary():
    if not os.path.exists(BINARY):
        pytest.fail(f"{BINARY} is not built: run __graft_entry__.build() "
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
