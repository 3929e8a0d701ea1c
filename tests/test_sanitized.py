"""Host code under AddressSanitizer and UBSan (CPU only).

- The oracle (oracle/desalm_oracle.c) over the cases whose array extents
  depend on the faces: inflow and characteristic-outlet faces, both case
  hooks, the actuator line, long i-lines (tests/sanitized_cases.py).
- The library's file readers (des-alm_amd/csrc/io_files.hip, host-only)
  on valid files and on every truncation, random byte flips and an empty
  first record: each is read or refused, never an invalid access.
Both are built by `make -C oracle sanitized`; a sanitizer report fails the
child process with exit code 86."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
ASAN_DIR = os.path.join(ROOT, "oracle", "_asan")
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:exitcode=86:abort_on_error=0",
           "UBSAN_OPTIONS": "halt_on_error=1:exitcode=86:print_stacktrace=1"}


def asan_runtime():
    try:
        p = subprocess.check_output(["gcc", "-print-file-name=libasan.so"], text=True).strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def sanitized():
    if asan_runtime() is None:
        pytest.skip("no AddressSanitizer runtime for gcc")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitized"])
    return ASAN_DIR


def test_oracle_face_cases_are_clean(sanitized):
    env = dict(os.environ, DESALM_ORACLE_LIB=os.path.join(sanitized, "liboracle.so"),
               LD_PRELOAD=asan_runtime(), **SAN_ENV)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "sanitized_cases.py")], env=env,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ALL OK" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def record(b):
    return struct.pack("<i", len(b)) + b + struct.pack("<i", len(b))


def io_check(sanitized, *args):
    r = subprocess.run([os.path.join(sanitized, "io_check")] + [str(a) for a in args],
                       env=dict(os.environ, **SAN_ENV), capture_output=True, text=True, timeout=60)
    assert r.returncode in (0, 1), (args, r.returncode, r.stderr[-3000:])
    return r.returncode


def test_file_readers_are_clean_on_damaged_files(sanitized, tmp_path):
    dims = (4, 3, 3)
    n = 36
    x = np.arange(n, dtype=np.float32)
    grid = b"".join(record((x + c).tobytes()) for c in range(4))
    f = tmp_path / "g"
    f.write_bytes(grid)
    assert io_check(sanitized, "grid", f, *dims, 1) == 0
    for cut in range(0, len(grid), 5):
        f.write_bytes(grid[:cut])
        assert io_check(sanitized, "grid", f, *dims, 1) == 1
    rng = np.random.default_rng(3)
    for _ in range(60):
        b = bytearray(grid)
        for p in rng.integers(0, len(b), size=3):
            b[p] = int(rng.integers(0, 256))
        f.write_bytes(bytes(b))
        io_check(sanitized, "grid", f, *dims, 1)
    for turb in (0, 1):
        f.write_bytes(record(b"") + grid)
        assert io_check(sanitized, "grid", f, *dims, turb) == 1
    solu = b"".join(record((x + c).astype(np.float64).tobytes()) for c in range(11))
    f.write_bytes(solu)
    assert io_check(sanitized, "solu", f, *dims, 1) == 0
    f.write_bytes(record(b"") + solu)
    assert io_check(sanitized, "solu", f, *dims, 1) == 1
    for cut in range(0, len(solu), 23):
        f.write_bytes(solu[:cut])
        assert io_check(sanitized, "solu", f, *dims, 1) == 1


def test_deck_and_filestat_readers_are_clean(sanitized, tmp_path):
    deck = os.path.join(ROOT, "tests", "golden", "ind3dmg_sample.dat")
    assert io_check(sanitized, "deck", deck, 3) == 0
    text = open(deck).read()
    f = tmp_path / "d"
    for cut in range(0, len(text), 11):
        f.write_text(text[:cut])
        io_check(sanitized, "deck", f, 3)
    f.write_text("0.5 12\n1.0\n")
    assert io_check(sanitized, "filestat", f, 1) == 0
    for t in ("", "1 2 3 4\n", "x\n", "0.5 12\n"):
        f.write_text(t)
        io_check(sanitized, "filestat", f, 2)
