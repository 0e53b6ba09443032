"""CPU: the library's untrusted-input host code under AddressSanitizer and
UndefinedBehaviorSanitizer.

tests/host_sanitize/driver.cpp is built with g++ together with the netlist
reader, the krets.toml reader, the symbolic pass and the run-time source
generators (the parts that parse text and build index plans), and run over
krets' netlists and spec files plus seeded mutations of each. Any memory
error, leak or undefined behaviour aborts the driver."""
import glob
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

CSRC = os.path.join(ROOT, "krets_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("san") / "driver")
    srcs = [os.path.join(ROOT, "tests", "host_sanitize", "driver.cpp")] + [
        os.path.join(CSRC, f) for f in ("netlist.cpp", "spec.cpp", "symbolic.cpp", "jit.cpp", "jit_tran.cpp")]
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
           "-I" + CSRC] + srcs + ["-o", out, "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lhiprtc", "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def inputs():
    files = sorted(glob.glob(os.path.join(GOLDEN, "circuits", "**", "*.cir"), recursive=True))
    files += sorted(glob.glob(os.path.join(GOLDEN, "circuits", "**", "krets.toml"), recursive=True))
    return files


def test_inputs_are_present():
    names = [os.path.basename(f) for f in inputs()]
    assert len(names) >= 16 and "krets.toml" in names and "resistor_ladder_1000.cir" in names


@pytest.mark.parametrize("seed", [1, 2])
def test_host_parsers_and_plans_under_sanitizers(driver, seed, tmp_path):
    extra = tmp_path / "controlled.cir"
    from conftest import CS_NETS, active_netlist
    extra.write_text("".join(CS_NETS.values()) + active_netlist())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([driver, "150", str(seed)] + inputs() + [str(extra)], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    parsed, rejected, plans, sources = (int(v) for v in r.stdout.split()[1::2])
    assert parsed > 500 and rejected > 100 and plans > 500 and sources > 100, r.stdout
