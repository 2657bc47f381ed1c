"""The reference-side binding of libsnc_hip, compiled into the reference.

INTEGRATION.md holds the binding a maintainer adds to the reference tree: a
patch to include/snc/runtime.hpp and src/runtime/cpu_backend.cpp (enum
values, names, the hip:: hook and forward()'s fallback branch,
runtime.hpp:21-69, cpu_backend.cpp:14-37 and 195-207) and three new files
(src/runtime/hip_backend.cpp, include/snc/hip_trainer.hpp mirroring
cuda_trainer.hpp:24-57, src/training/hip_trainer.cpp). This test takes them
from the document as written, applies the patch to copies of the two files in
a scratch directory, compiles the reference's own CPU sources where they lie
together with the binding and tests/integration/binding_check.cpp with g++,
links libsnc_hip.so and runs the check: enum numbers against SNC_VARIANT_*,
the InputEvent/snc_event layout, the names, and the fallback contract (no
device: forward() returns the CPU result as Backend::Cpu and hiptrain
declines, so the caller keeps the CPU Trainer).

Skipped where the reference tree is absent.
"""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")
LIB = ROOT / "snc_amd" / "libsnc_hip.so"
DOC = ROOT / "INTEGRATION.md"

REF_SRCS = ["src/graph/snn_graph.cpp", "src/runtime/cuda_stub.cpp",
            "src/training/encoders.cpp", "src/training/datasets.cpp",
            "src/training/trainer.cpp", "src/training/cuda_trainer_stub.cpp"]
PATCHED = ["include/snc/runtime.hpp", "src/runtime/cpu_backend.cpp"]
NEW_FILES = ["src/runtime/hip_backend.cpp", "include/snc/hip_trainer.hpp",
             "src/training/hip_trainer.cpp"]


def binding_from_doc():
    """(patch text, {path: source}) from the document's fenced blocks."""
    text = DOC.read_text()
    diffs = re.findall(r"```diff\n(.*?)```", text, re.S)
    files = {}
    for body in re.findall(r"```cpp\n(.*?)```", text, re.S):
        first = body.split("\n", 1)[0]
        m = re.fullmatch(r"// ((?:src|include)/\S+)", first)
        if m:
            files[m.group(1)] = body
    return diffs, files


def test_document_holds_the_whole_binding():
    diffs, files = binding_from_doc()
    assert len(diffs) == 1
    for p in PATCHED:
        assert f"--- a/{p}" in diffs[0] and f"+++ b/{p}" in diffs[0]
    assert sorted(files) == sorted(NEW_FILES)


@pytest.mark.skipif(not (REF / "include/snc/runtime.hpp").exists(),
                    reason="reference tree absent")
@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("patch") is None,
                    reason="g++ or patch missing")
def test_binding_compiles_into_the_reference_and_keeps_its_contract(tmp_path):
    if not LIB.exists():
        pytest.skip("libsnc_hip.so not built")
    diffs, files = binding_from_doc()
    for p in PATCHED:
        (tmp_path / p).parent.mkdir(parents=True, exist_ok=True)
        shutil.copyfile(REF / p, tmp_path / p)
    subprocess.run(["patch", "-p1", "--forward", "-s", "-d", str(tmp_path)],
                   input=diffs[0], text=True, check=True)
    for p, body in files.items():
        (tmp_path / p).parent.mkdir(parents=True, exist_ok=True)
        (tmp_path / p).write_text(body)
    exe = tmp_path / "binding_check"
    cmd = (["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-fopenmp",
            "-Wall", "-Wextra", "-Werror=return-type",
            # the patched copy of runtime.hpp shadows the reference's own
            f"-I{tmp_path / 'include'}", f"-I{REF / 'include'}", f"-I{ROOT / 'include'}",
            str(tmp_path / "src/runtime/cpu_backend.cpp"),
            str(tmp_path / "src/runtime/hip_backend.cpp"),
            str(tmp_path / "src/training/hip_trainer.cpp")]
           + [str(REF / s) for s in REF_SRCS]
           + [str(ROOT / "tests/integration/binding_check.cpp"),
              f"-L{LIB.parent}", "-lsnc_hip", f"-Wl,-rpath,{LIB.parent}", "-o", str(exe)])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    warn = [l for l in r.stderr.splitlines()
            if "warning" in l and str(tmp_path) in l and "/src/runtime/cpu_backend.cpp" not in l]
    assert not warn, "\n".join(warn)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "binding ok" in out.stdout
