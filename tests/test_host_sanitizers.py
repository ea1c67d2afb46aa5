"""The host-side model parsers (JSON, YAML, the L3 / L4 coupling tables)
built with g++ under AddressSanitizer and UndefinedBehaviorSanitizer
(tests/host_sanitize.cpp): sample model files parse, and seeded byte-level
mutations of them either parse or are refused with a library error, with no
sanitizer report.  CPU only; skipped where g++ or the sanitizer runtimes are
missing."""
import json
import os
import shutil
import subprocess

import pytest
import yaml

from conftest import ROOT
from jse_amd import model as M
from jse_amd import workloads as W

CSRC = os.path.join(ROOT, "jse_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the ROCm headers are missing")
    out = tmp_path_factory.mktemp("asan") / "host_sanitize"
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           os.path.join(ROOT, "tests", "host_sanitize.cpp"),
           *(os.path.join(CSRC, f) for f in ("model_host.cpp", "yaml_json.cpp", "sh_coupling.cpp")), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "asan" in (r.stderr + r.stdout).lower():
        pytest.skip("sanitizer runtimes missing: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr[-2000:]
    return str(out)


def _models(tmp):
    merge = M.random_merge_model(
        ["Cu", "Zr", "Al"],
        [{"type": "spherical_chebyshev", "nmax": 3, "lmax": 4, "l3max": 3, "l4max": 2, "rcut": 5.0, "wtype": "full"},
         {"type": "chebyshev", "nmax": 4, "rcut": 4.5}], (8, 8), 9)
    merge["models"][1] = {"symbol": "Zr", "basis": {"type": "mirror", "mirror": 1}}
    cu = M.random_model(["Cu"], nmax=3, lmax=4, rcut=5.0, hidden=(8, 8), seed=4)
    files = {"merge.json": json.dumps(merge), "merge.yaml": yaml.safe_dump(merge, default_flow_style=False),
             "cu.json": json.dumps(cu, indent=1), "cu.yml": yaml.safe_dump(cu, default_flow_style=True, width=70)}
    paths = []
    for name, text in files.items():
        p = tmp / name
        p.write_text(text)
        paths.append(str(p))
    return paths


def test_parsers_under_sanitizers(driver, tmp_path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([driver, "80", "7", *_models(tmp_path)], capture_output=True, text=True, env=env,
                       timeout=600)
    print(r.stdout, r.stderr[-3000:])
    assert r.returncode == 0
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert "mutations parsed" in r.stdout
