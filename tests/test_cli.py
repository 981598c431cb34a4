"""The lrt command's argument handling, which needs no GPU: usage and help,
unknown options, the reference's options this build does not read (point
files, streamed .lrtc clouds) and --lod outside the rasteriser refused with
status 1 before anything runs."""
import os
import subprocess

import pytest

LRT = os.path.join(os.path.dirname(__file__), "..", "lucabrtrender_amd", "lrt")


def _run(*args):
    p = subprocess.run([LRT, *args], capture_output=True, text=True, timeout=60)
    return p.returncode, p.stdout + p.stderr


def test_the_command_is_built():
    assert os.access(LRT, os.X_OK), "run __graft_entry__.build() first"


@pytest.mark.parametrize("command", ["render", "bench"])
def test_help_lists_the_reference_options(command):
    code, out = _run(command, "--help")
    assert code == 0
    for option in ("--splats", "--technique", "--eye", "--target", "--up", "--rotate-x", "--scale", "--focal",
                   "--size", "--near", "--degree", "--no-antialias", "--lod"):
        assert option in out
    assert ("--repeat" in out and "--stages" in out) if command == "bench" else "--output" in out


@pytest.mark.parametrize("args", [[], ["draw"], ["render", "--bogus"], ["render", "--eye", "1", "2"],
                                  ["bench", "--points", "p.ply"], ["render", "--lod", "-1"],
                                  ["render", "--splats", "x.ply", "--lod", "2", "--technique", "reference"],
                                  ["bench", "--stream-budget", "1000"],
                                  ["render", "--repeat", "3"], ["bench", "-o", "x.exr"]])
def test_bad_commands_and_options_exit_with_status_one(args):
    code, out = _run(*args)
    assert code == 1 and out
