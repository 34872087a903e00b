"""The script front end (gpugi_render) without a device: command parsing and
the errors the reference's GlobalConfig raises (globalconfig.cpp:19-76) for
unknown parameters and malformed arguments.  Nothing here creates a context."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "gpugi_amd", "bin", "gpugi_render")


def run(*lines, timeout=30):
    args = [APP]
    for l in lines:
        args += ["-e", l]
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(APP):
        pytest.fail("gpugi_render is not built; run __graft_entry__.build()")


def test_help_lists_every_command():
    r = run("help")
    assert r.returncode == 0
    for name in ("sceneFilename", "cameraPos", "cameraLookAt", "cameraFOV", "resolution", "pathLength",
                 "renderer", "traversal", "addPointLight", "waitIterations", "namedscreenshot",
                 "referenceImage", "mseCheckInterval", "waitMSE", "envMap", "exit"):
        assert name in r.stdout


def test_comments_blank_lines_and_exit_stop_processing():
    # everything after `exit` is ignored, so the bogus command never runs
    r = run("# a comment", "   ", "// another", "help // trailing comment", "exit", "bogus = 1")
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("line,msg", [
    ("bogus = 1", "unknown parameter 'bogus'"),
    ("renderer = 7", "not part of this build"),
    ("renderer = x", "not an integer"),
    ("cameraPos = 1, 2", "expects 3 argument"),
    ("resolution = 64", "expects 2 argument"),
    ("bvhType = 2", "0 (axis-aligned boxes) or 1 (oriented boxes)"),
    ("cameraFOV = abc", "not a number"),
    ("mseCheckInterval = x", "not an integer"),
    ("waitMSE = 1, 2", "expects 1 argument"),
    ("envMap = 4, a, b", "expects 7 argument"),
])
def test_argument_errors(line, msg):
    r = run(line)
    assert r.returncode == 2
    assert msg in r.stderr and "line 1" in r.stderr


def test_missing_script_file():
    r = subprocess.run([APP, os.path.join(ROOT, "no_such_script.txt")], capture_output=True, text=True, timeout=30)
    assert r.returncode == 2 and "cannot open script" in r.stderr


def test_point_light_needs_a_scene():
    r = run("addPointLight = 0, 1, 0, 1, 1, 1")
    assert r.returncode == 2 and "no scene loaded" in r.stderr


def test_wait_mse_needs_a_check_interval():
    # application.cpp:181-182: logged, and the script goes on (no context is created)
    r = run("waitMSE = 0.5", "help")
    assert r.returncode == 0
    assert "mseCheckInterval is 0" in r.stderr and "sceneFilename" in r.stdout
