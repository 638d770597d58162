"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

The host-side C and C++ of this repository is rebuilt with both sanitizers
(errors fatal) into a temporary folder, and the CPU tests that drive it run
again against those builds in a child process with the sanitizer runtime
preloaded:

- libatmos_io (the netCDF classic-format library, atm_nc.cpp) and the oracle
  (atm_oracle.c): gcc's runtime; test_ncio and the oracle's tests;
- the host builds of the device numerics shared with the HIP kernels: the
  linearised step (lin_host.cpp over lin_*.hpp, atm_common.hpp) and the
  remap column and pow_kappa (remap_column_host.cpp over remap_column.hpp):
  clang's runtime; their tests, including the frozen-branch differences.

A bounds error, use after free, misaligned or null access, integer overflow
or invalid conversion in any of them ends the child with a report.
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "atmosphere-model_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
# hipcc: each sanitizer flag for the host compilation only
HIP_SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
           "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize-recover=all", "-shared-libasan"]


def _clang_runtime():
    found = sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return found[-1] if found else None


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (cmd, r.stdout[-3000:], r.stderr[-3000:])


def _pytest_under(runtime, env_libs, tests, tmp):
    """the tests in a child process with `runtime` preloaded and the loaders
    pointed at the sanitizer builds; returns the child's output"""
    log = os.path.join(tmp, "san")
    env = dict(os.environ, LD_PRELOAD=runtime, ASAN_OPTIONS=f"detect_leaks=0:log_path={log}",
               UBSAN_OPTIONS="print_stacktrace=1", **env_libs)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        *[os.path.join(ROOT, "tests", t) for t in tests]],
                       env=env, cwd=ROOT, capture_output=True, text=True, timeout=900)
    reports = "".join(open(f).read() for f in glob.glob(log + "*"))
    assert r.returncode == 0 and not reports, (r.stdout[-3000:], r.stderr[-3000:], reports[-5000:])
    # the loaders took the sanitizer builds
    probe = ("import os, sys; sys.path[:0] = [%r, %r, %r]\n"
             "import oracle, lin_host\nfrom atmos import ncio\n"
             "print(ncio.LIB_PATH); print(lin_host.LIB)\n"
             "print(os.environ.get('ORACLE_LIB') and oracle.lib()._name)\n") % (
        ROOT, os.path.join(ROOT, "atmosphere-model_amd"), os.path.join(ROOT, "tests"))
    p = subprocess.run([sys.executable, "-c", probe], env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    return r.stdout, p.stdout


@pytest.mark.skipif(not os.path.exists("/usr/bin/gcc"), reason="gcc missing")
def test_io_library_and_oracle_under_gcc_sanitizers(tmp_path):
    tmp = str(tmp_path)
    io = os.path.join(tmp, "libatmos_io.so")
    orc = os.path.join(tmp, "liboracle.so")
    _run(["g++", "-std=c++17", "-fPIC", "-shared", *SAN, "-o", io, os.path.join(CSRC, "atm_nc.cpp")])
    _run(["gcc", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", *SAN, "-o", orc,
          os.path.join(ROOT, "oracle", "atm_oracle.c"), "-lquadmath", "-lm"])
    runtime = subprocess.check_output(["gcc", "-print-file-name=libasan.so"], text=True).strip()
    out, probe = _pytest_under(runtime, {"ATM_IO_LIB": io, "ORACLE_LIB": orc},
                               ["test_ncio.py", "test_oracle_pins.py", "test_oracle_pins_closed_forms.py",
                                "test_oracle_physics.py", "test_oracle_output.py", "test_oracle_twostream.py"], tmp)
    print(out.strip().splitlines()[-1])
    assert io in probe and orc in probe


@pytest.mark.skipif(not os.path.exists(HIPCC) or _clang_runtime() is None, reason="hipcc or its ASan runtime missing")
def test_device_numerics_host_builds_under_clang_sanitizers(tmp_path):
    tmp = str(tmp_path)
    lin = os.path.join(tmp, "libatmos_lin_host.so")
    col = os.path.join(tmp, "libatmos_column_host.so")
    base = [HIPCC, "-x", "hip", "--cuda-host-only", "-no-hip-rt", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", *HIP_SAN]
    _run(base + ["-o", lin, os.path.join(CSRC, "lin_host.cpp")])
    _run(base + ["-o", col, os.path.join(CSRC, "remap_column_host.cpp")])
    out, probe = _pytest_under(_clang_runtime(), {"ATM_LIN_HOST_LIB": lin, "ATM_COLUMN_HOST_LIB": col},
                               ["test_lin_host.py", "test_lin_frozen_fd.py", "test_remap_column_host.py",
                                "test_pow_kappa.py"], tmp)
    print(out.strip().splitlines()[-1])
    assert lin in probe
