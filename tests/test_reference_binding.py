"""The reference-side binding (lsdtt3_amd/binding/lsdtt_lem_gpu_dispatch.hpp:
run_lem_step's GPU dispatch, options and every step statistic mapped)
compiles against the reference's own headers (src/cpu/lsdtt_lem_cpu.hpp and
what it includes) and this repository's include/lsdtt3_gpu.h, and every
entry point it calls is exported by the library.  Skipped when the reference
tree or g++ is absent."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
HDR = os.path.join(ROOT, "lsdtt3_amd", "binding", "lsdtt_lem_gpu_dispatch.hpp")


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "cpu")) or shutil.which("g++") is None,
                    reason="reference tree or g++ absent")
def test_binding_compiles_against_reference_headers(tmp_path):
    src = tmp_path / "check.cpp"
    src.write_text('#include "lsdtt_lem_gpu_dispatch.hpp"\n'
                   "// instantiate the dispatch as run_lem_step would call it\n"
                   "bool probe(const lsdtt::lsdtt_raster_cpu& dem, const lsdtt::lsdtt_lem_options& o,\n"
                   "           const std::vector<std::int64_t>& bf) {\n"
                   "    static thread_local lsdtt_mi355x::gpu_lem_context ctx;\n"
                   "    std::pair<lsdtt::lsdtt_raster_cpu, lsdtt::lsdtt_lem_step_stats> out;\n"
                   "    return lsdtt_mi355x::try_run_lem_step(ctx, dem, o, bf, &out);\n"
                   "}\n")
    cmd = ["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.dirname(HDR),
           "-I" + os.path.join(REF, "core"), "-I" + os.path.join(REF, "utils"), "-I" + os.path.join(REF, "cpu"),
           str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_binding_calls_only_declared_entry_points():
    text = open(HDR).read()
    called = set(re.findall(r"\b(lsdtt_gpu_[a-z_]+)\s*\(", text))
    header = open(os.path.join(ROOT, "include", "lsdtt3_gpu.h")).read()
    declared = set(re.findall(r"\b(lsdtt_gpu_[a-z_]+)\s*\(", header))
    assert called and called <= declared, called - declared
