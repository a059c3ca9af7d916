"""The drop-in include directory against the reference harness's own code (CPU).

The reference's harness examples/hpxfft/shared_loop_2d.cpp opens with three
includes (:1-3) -- hpxfft/shared/loop.hpp, hpxfft/util/create_dir.hpp and
hpxfft/util/print_vector.hpp -- and drives the class with :33-62
(vector_2d, ramp input, initialize, fft_2d_r2c_par/_seq, print_vector_2d).
These lines are read from the reference tree at test time (nothing of it is
stored here), placed in a translation unit whose HPX-only parts (the hpx_main
signature, the HPX timer) are replaced by stand-ins, and compiled with this
repository's include directory ahead of the reference's own
(-I include -I <reference>/core/include), then linked with -lhpxfft_hip.
Every header must resolve to this repository: the reference's need FFTW and
HPX.
With the reference's directory first the same unit does not compile, which
shows the check has teeth.  Skipped where the reference tree is absent.
"""
import os
import subprocess

import pytest

from conftest import REPO

REF = "/root/reference"
EXAMPLE = os.path.join(REF, "examples", "hpxfft", "shared_loop_2d.cpp")
INCLUDE = os.path.join(REPO, "include")
LIB_DIR = os.path.join(REPO, "hpx-fft_amd", "lib")

pytestmark = pytest.mark.skipif(not os.path.exists(EXAMPLE), reason="reference tree not present")


def _harness_unit() -> str:
    lines = open(EXAMPLE).read().splitlines()
    includes = lines[0:3]
    assert [ln.split('"')[1] for ln in includes] == ["hpxfft/shared/loop.hpp", "hpxfft/util/create_dir.hpp",
                                                     "hpxfft/util/print_vector.hpp"]
    body = lines[32:62]
    assert body[0].strip().startswith("hpxfft::shared::vector_2d values_vec(")
    assert body[-1].strip() == "}" and "print_vector_2d(values_vec)" in body[-2]
    rt = next(i for i, ln in enumerate(lines) if "create_parent_dir(runtime_file_path)" in ln)
    plan_write = next(ln for ln in lines if "write_plans_to_file(plan_file_path)" in ln)
    return "\n".join(includes + [
        "#include <chrono>",
        "#include <cstdio>",
        "#include <string>",
        "namespace {",
        "// stands in for hpx::chrono::high_resolution_timer: now() in seconds",
        "struct wall_timer {",
        "    double now() const {",
        "        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();",
        "    }",
        "};",
        "}",
        "static int harness(std::size_t dim_c_x, std::size_t dim_r_y, const std::string plan_flag,",
        "                   const std::string run_flag, bool print_result)",
        "{",
        "    auto t = wall_timer();",
        "    const std::size_t dim_c_y = dim_r_y / 2 + 1;",
    ] + body + [
        "    std::printf(\"total %g init %g fft %g\\n\", stop_total - start_total, stop_init - start_total,",
        "                fft_computer.get_measurement(\"total\"));",
    ] + lines[rt - 1:rt + 1] + [
        "    std::string plan_file_path = \"plans/plan_hpx_shared_loop.txt\";",
        "    hpxfft::util::create_parent_dir(plan_file_path);",
        plan_write,
        "    return 0;",
        "}",
        "int main()",
        "{",
        "    try { return harness(4, 6, \"estimate\", \"par\", true); }",
        "    catch (const std::runtime_error &e) { std::printf(\"runtime_error: %s\\n\", e.what()); return 3; }",
        "}",
        "",
    ])


def _compile(tmp_path, include_order, extra=()):
    src = tmp_path / "harness.cpp"
    src.write_text(_harness_unit())
    cmd = ["g++", "-std=c++17", "-H", *[a for d in include_order for a in ("-I", d)], str(src), *extra]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=tmp_path)


def test_reference_harness_builds_against_the_dropin(tmp_path):
    exe = tmp_path / "harness"
    r = _compile(tmp_path, [INCLUDE, os.path.join(REF, "core", "include")],
                 ["-o", str(exe), f"-L{LIB_DIR}", "-lhpxfft_hip", f"-Wl,-rpath,{LIB_DIR}"])
    assert r.returncode == 0, r.stderr[-4000:]
    headers = [ln.lstrip(". ").strip() for ln in r.stderr.splitlines() if ln.startswith(".")]
    ours = [h for h in headers if os.path.realpath(h).startswith(os.path.realpath(INCLUDE))]
    assert not [h for h in headers if "reference" in h], "a reference header was included"
    for name in ("shared/loop.hpp", "util/create_dir.hpp", "util/print_vector.hpp", "util/vector_2d.hpp",
                 "gpu/loop.hpp"):
        assert any(h.endswith(name) for h in ours), name
    # no HIP device in this container: the class reports it as std::runtime_error
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    if run.returncode == 3:
        assert "runtime_error: no HIP device available" in run.stdout
    else:   # a device is visible: the 4 x 6 ramp (x[i][j] = j) has X[0] = (60,0) (-12,20.7846) ...
        assert run.returncode == 0 and "(60 0)" in run.stdout, run.stdout + run.stderr


def test_reference_headers_first_do_not_build(tmp_path):
    """Control: with the reference's include directory first, its headers
    (which need FFTW and HPX) are picked and the unit cannot compile."""
    r = _compile(tmp_path, [os.path.join(REF, "core", "include"), INCLUDE], ["-fsyntax-only"])
    assert r.returncode != 0
    assert "fftw3.h" in r.stderr or "hpx/" in r.stderr


def test_print_vector_and_create_dir(tmp_path):
    """print_vector_2d prints (re im) pairs per row, then a blank line;
    create_parent_dir makes the parent directories once and announces them."""
    src = tmp_path / "u.cpp"
    src.write_text(
        '#include "hpxfft/util/print_vector.hpp"\n'
        '#include "hpxfft/util/create_dir.hpp"\n'
        "int main() {\n"
        "  hpxfft::util::vector_2d<double> v(2, 4);\n"
        "  for (std::size_t i = 0; i < 2; ++i) for (std::size_t j = 0; j < 4; ++j) v(i, j) = 10.0 * i + j;\n"
        "  hpxfft::util::print_vector_2d(v);\n"
        '  hpxfft::util::create_parent_dir("a/b/c.txt");\n'
        '  hpxfft::util::create_parent_dir("a/b/d.txt");\n'
        '  hpxfft::util::create_parent_dir("e.txt");\n'
        "  try { hpxfft::util::create_parent_dir(\"blocker/x/y.txt\"); }\n"
        '  catch (const std::runtime_error &e) { std::cout << "caught: " << e.what() << "\\n"; }\n'
        "}\n")
    (tmp_path / "blocker").write_text("a file where a directory is wanted")
    exe = tmp_path / "u"
    r = subprocess.run(["g++", "-std=c++17", "-I", INCLUDE, str(src), "-o", str(exe)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, cwd=tmp_path).stdout
    assert out.startswith("(0 1) (2 3) \n(10 11) (12 13) \n\n")
    assert out.count("Directory created: ") == 1 and '"a/b"' in out
    assert "caught: Failed to create directory: blocker/x" in out
    assert (tmp_path / "a" / "b").is_dir()
