"""AddressSanitizer over the library's host code (no GPU).

Builds every csrc/*.hip with AddressSanitizer and UndefinedBehaviorSanitizer
on the host side only (-Xarch_host; the device code is built as usual and
never launched), links tests/asan/host_checks.cpp against the objects and
runs it with leak detection (the HIP runtime's own allocations suppressed):
plans
and tile queries with tuning overrides, workspace queries of every path, the
host builds of the Green's-function and interpolation code, the batch
planner's dry recording and launch merge, and the argument checks.  Any
sanitizer report or failed check fails the test.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "fourierd3_amd" / "csrc"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not available")
def test_host_code_under_address_sanitizer(tmp_path):
    common = ["--offload-arch=gfx950", "-O1", "-Xarch_host", "-g", "-std=c++17", "-fPIC",
              f"-I{ROOT / 'include'}", "-Wno-unused-result"] + SAN

    def compile_one(src):
        obj = tmp_path / (src.stem + ".o")
        r = subprocess.run([HIPCC, *common, "-c", str(src), "-o", str(obj)], capture_output=True, text=True)
        assert r.returncode == 0, f"{src.name}:\n{r.stderr[-4000:]}"
        return obj

    srcs = sorted(CSRC.glob("*.hip")) + [ROOT / "tests" / "asan" / "host_checks.cpp"]
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(compile_one, srcs))
    exe = tmp_path / "host_checks"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", *SAN, "-o", str(exe), *map(str, objs)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1", HIP_VISIBLE_DEVICES="")
    lsan = tmp_path / "lsan.supp"
    # the HIP runtime's own allocations at process exit are not the library's
    lsan.write_text("leak:libamdhip64\nleak:libhsa-runtime64\nleak:librocprofiler\n")
    env["LSAN_OPTIONS"] = f"suppressions={lsan}"
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out
    assert r.returncode == 0 and "0 failure(s)" in out
