"""The C++ renderer plugin (WavefrontRenderer behind the reference's Renderer interface,
include/renderers/renderer.h:14-35) driven by a compiled C++ program: RenderRecord events,
getSpp, reset, getRadianceBuffer, the pixel buffer, a renderer switch on one DeviceGlobals,
and bit-equality with the same render driven through the C ABI."""
import os
import subprocess

import pytest

import pathed_capi as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "tests", "plugin", "renderer_plugin_test")


def _run():
    if not os.path.exists(PROGRAM):
        pytest.fail(f"{PROGRAM} is not built; run __graft_entry__.build()")
    return subprocess.run([PROGRAM], capture_output=True, text=True, timeout=300)


def test_plugin_program_without_a_device_runs_nothing():
    if P.device_count() > 0:
        pytest.skip("a GPU is visible: the program renders (see the gpu test)")
    res = _run()
    assert res.returncode == 77 and "no HIP device" in res.stdout


@pytest.mark.gpu
def test_wavefront_renderer_through_the_plugin_interface():
    print("case: C++ renderer plugin program", flush=True)
    res = _run()
    print(res.stdout + res.stderr, flush=True)
    assert res.returncode == 0 and "PASSED" in res.stdout
    assert res.stdout.count("ok  ") >= 10
