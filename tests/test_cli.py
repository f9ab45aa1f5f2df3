"""The command-line front end (wema_amd/wema): the reference CLI's option set and checks
(src/main.c:63-218) plus --levels and --device, and the stream probe, without a GPU. FFmpeg
is absent from this image: tests/cli/ holds stand-ins for ffprobe and ffmpeg (raw RGB24
files in and out). CPU only: the run ends where the device session would open."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "wema_amd", "wema")
FAKES = os.path.join(ROOT, "tests", "cli")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="wema CLI not built")


def run(args, env=None, timeout=60):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=timeout, env=e)


def test_help_lists_the_reference_options_and_the_new_ones():
    r = run(["--help"])
    assert r.returncode == 0
    for opt in ("--input", "--output", "--amp", "--fl", "--fh", "--ff-codec", "--ff-option",
                "--temporal-window", "--no-edge-aware", "--bilateral-filter", "--color", "--batch",
                "--threads", "--verbose", "--levels", "--device"):
        assert opt in r.stderr, opt


@pytest.mark.parametrize("args,msg", [
    ([], "Input file required"),
    (["-i"], "-i requires an argument"),
    (["-i", "x", "--bogus"], "Unknown option: --bogus"),
    (["-i", "x", "-a", "0"], "amplification factor must be positive"),
    (["-i", "x", "--fl", "-1"], "low frequency must be non-negative"),
    (["-i", "x", "--fh", "0"], "high frequency must be positive"),
    (["-i", "x", "--fl", "3", "--fh", "2"], "Low frequency must be less than high frequency"),
    (["-i", "x", "--temporal-window", "3"], "temporal window must be at least 4"),
    (["-i", "x", "--temporal-window", "257"], "temporal window cannot exceed 256"),
    (["-i", "x", "--batch", "0"], "batch size must be at least 1"),
    (["-i", "x", "--batch", "257"], "batch size cannot exceed 256"),
    (["-i", "x", "--threads", "-1"], "thread count must be non-negative"),
    (["-i", "x", "--levels", "0"], "levels must be in 1..12"),
    (["-i", "x", "--levels", "13"], "levels must be in 1..12"),
    (["-i", "x", "--device", "-1"], "device must be non-negative"),
])
def test_option_errors(args, msg):
    r = run(args)
    assert r.returncode == 1
    assert msg in r.stderr


def test_unreadable_input_is_reported(tmp_path):
    env = {"PATH": FAKES + os.pathsep + os.environ["PATH"], "FAKE_W": "96", "FAKE_H": "64", "FAKE_FPS": "30/1"}
    r = run(["-i", str(tmp_path / "missing.raw")], env)
    assert r.returncode == 1 and "Cannot open input" in r.stderr


@pytest.mark.parametrize("extra,levels", [([], "3 (automatic)"), (["--levels", "5"], "5")])
def test_probe_and_settings(tmp_path, extra, levels):
    src = tmp_path / "clip.raw"
    src.write_bytes(bytes(96 * 64 * 3 * 2))
    env = {"PATH": FAKES + os.pathsep + os.environ["PATH"], "FAKE_W": "96", "FAKE_H": "64",
           "FAKE_FPS": "30000/1001", "FAKE_N": "2"}
    r = run(["-i", str(src), "-v", "--fl", "0.6"] + extra, env)
    # the probe is read and the settings printed; without a GPU the session does not open
    assert "Size: 96x64" in r.stderr and "FPS: 29.970" in r.stderr and "Frames: 2" in r.stderr
    assert "Output: %s" % (tmp_path / "clip.wema.mkv") in r.stderr
    assert "Levels: %s" % levels in r.stderr
    assert "Frequency band: 0.60 - 14.98 Hz" in r.stderr     # --fh clamped to Nyquist
    if r.returncode != 0:
        assert "Failed to initialize WEMA" in r.stderr


def test_option_paths_under_address_and_undefined_sanitizers(tmp_path):
    """The CLI's host code (option parsing, the stream probe, the error paths, the set-up up
    to the device session) built with AddressSanitizer and UndefinedBehaviorSanitizer, leak
    detection on: no report on any of the option and probe cases above."""
    exe = tmp_path / "wema_asan"
    lib = os.path.join(ROOT, "wema_amd")
    b = subprocess.run(["gcc", "-std=c11", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                        "-fno-sanitize-recover=undefined", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(lib, "host", "wema_cli.c"), "-L" + lib, "-lwema_dev", "-Wl,-rpath," + lib,
                        "-o", str(exe)], capture_output=True, text=True)
    if b.returncode != 0:
        pytest.skip("sanitizer build unavailable: " + b.stderr[-200:])
    src = tmp_path / "clip.raw"
    src.write_bytes(bytes(96 * 64 * 3 * 2))
    env = dict(os.environ, PATH=FAKES + os.pathsep + os.environ["PATH"], FAKE_W="96", FAKE_H="64",
               FAKE_FPS="30/1", FAKE_N="2", ASAN_OPTIONS="detect_leaks=1:exitcode=97", UBSAN_OPTIONS="exitcode=98")
    cases = [["--help"], [], ["-i"], ["-i", "x", "--bogus"], ["-i", "x", "--levels", "13"],
             ["-i", str(tmp_path / "missing.raw")], ["-i", str(src), "-v", "--levels", "5", "--fl", "0.6"],
             ["-i", str(src), "--color", "--haar", "--bilateral-filter", "--temporal-window", "16"]]
    for args in cases:
        r = subprocess.run([str(exe)] + args, capture_output=True, text=True, timeout=120, env=env)
        assert r.returncode not in (97, 98), (args, r.stderr[-2000:])
        assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, (args, r.stderr[-2000:])
