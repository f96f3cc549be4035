"""The host-only C-ABI code (case-file reader, status and grid.json texts)
under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU: valid
case files, every invalid case of the reader's tests, every truncation of a
valid file, and random bytes.  A sanitizer report aborts the harness."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("asan")
    exe = d / "harness"
    subprocess.run(["g++", "-std=c++17", "-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "host_asan_harness.cpp"),
                    os.path.join(ROOT, "fen_amd", "csrc", "case_file.cpp"), "-o", str(exe)], check=True)
    return exe


def _run(exe, files):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")
    r = subprocess.run([str(exe)] + [str(f) for f in files], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return int(r.stdout.strip())


def test_reader_and_texts_are_clean(harness, tmp_path):
    from tests.test_case_file import FULL
    good = tmp_path / "good.json"
    good.write_text(json.dumps(FULL, indent=1))
    files = [good]
    text = good.read_bytes()
    for cut in range(len(text)):   # every truncation
        p = tmp_path / f"cut{cut}.json"
        p.write_bytes(text[:cut])
        files.append(p)
    rng = np.random.default_rng(5)
    for n in range(40):   # random bytes, and random bytes spliced into the valid file
        p = tmp_path / f"rnd{n}.json"
        noise = rng.integers(0, 256, size=int(rng.integers(1, 300)), dtype=np.uint8).tobytes()
        at = int(rng.integers(0, len(text)))
        p.write_bytes(noise if n % 2 else text[:at] + noise + text[at:])
        files.append(p)
    deep = tmp_path / "deep.json"   # nesting is bounded before the recursion can exhaust the stack
    deep.write_text("[" * 200000 + "]" * 200000)
    files.append(deep)
    assert _run(harness, files) == 1   # only the untouched file parses
