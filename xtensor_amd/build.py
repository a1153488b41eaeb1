"""Build the gfx950 shared library `xtensor_amd/lib/libxg.so`.

Steps: compile the host-side spec generator with g++ and run it to produce
`_build/specs.inc` (constexpr micro-op tables of the compiled
specialisations), then compile every HIP / C++ source of `csrc/` with hipcc
for gfx950 and link them into one shared library.  Sources are compiled in
parallel; a source is recompiled when it or any header is newer than its
object.

Flags: -ffp-contract=off keeps every floating-point operator individually
rounded (xtensor's functors are separate C++ operations), and packed FP32
instructions are disabled for all device code.
"""

from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libxg.so")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", CSRC,
    "-I", BUILD,
    "-I", INCLUDE,
]


HOST_FLAGS = [
    "-O2",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-D__HIP_PLATFORM_AMD__",
    "-I", "/opt/rocm/include",
    "-I", CSRC,
    "-I", BUILD,
    "-I", INCLUDE,
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("command failed: " + " ".join(cmd))


def _newest(paths: list[str]) -> float:
    return max((os.path.getmtime(p) for p in paths if os.path.exists(p)), default=0.0)


def build(force: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    headers = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(INCLUDE, "*.h"))
    hdr_time = _newest(headers)

    # 1. spec generator (host only) -> specs.inc
    gen = os.path.join(BUILD, "specgen")
    gen_src = [os.path.join(CSRC, f) for f in ("specgen.cpp", "planner.cpp", "reduce_plan.cpp", "scan_plan.cpp")]
    gen_src = [f for f in gen_src if os.path.exists(f)]
    specs = os.path.join(BUILD, "specs.inc")
    if force or not os.path.exists(gen) or _newest(gen_src) > os.path.getmtime(gen) or hdr_time > os.path.getmtime(gen):
        _run(["g++", "-std=c++20", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I", CSRC,
              "-I", INCLUDE, "-o", gen] + gen_src)
    _run([gen, specs + ".tmp"])
    old = open(specs).read() if os.path.exists(specs) else None
    new = open(specs + ".tmp").read()
    if old != new:
        os.replace(specs + ".tmp", specs)
    else:
        os.remove(specs + ".tmp")
    hdr_time = max(hdr_time, os.path.getmtime(specs))

    # 2. device library
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    sources = [s for s in sources if not s.endswith("specgen.cpp")]
    objs = []
    todo = []
    for s in sources:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        objs.append(o)
        if force or not os.path.exists(o) or max(os.path.getmtime(s), hdr_time) > os.path.getmtime(o):
            if s.endswith(".hip"):
                todo.append([HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", s, "-o", o])
            else:  # host-only translation units
                todo.append(["g++"] + HOST_FLAGS + ["-c", s, "-o", o])
    if todo:
        n = jobs or min(8, os.cpu_count() or 4)
        with cf.ThreadPoolExecutor(n) as ex:
            list(ex.map(_run, todo))
    if todo or not os.path.exists(LIB) or _newest(objs) > os.path.getmtime(LIB):
        tmp = LIB + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs +
             ["-L/opt/rocm/lib", "-lhiprtc", "-ldl", "-Wl,-rpath,/opt/rocm/lib"])
        os.replace(tmp, LIB)

    # 3. the C++ front end's test program (host code against libxg.so)
    cpp_test = os.path.join(ROOT, "tests", "cpp", "test_xdevice.cpp")
    exe = os.path.join(BUILD, "test_xdevice")
    cpp_hdr = os.path.join(HERE, "cpp", "xdevice.hpp")
    if os.path.exists(cpp_test) and (force or not os.path.exists(exe) or
                                     _newest([cpp_test, cpp_hdr, LIB] + headers) > os.path.getmtime(exe)):
        _run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I", INCLUDE,
              "-I", os.path.join(HERE, "cpp"), cpp_test, "-o", exe, "-L", LIBDIR, "-lxg",
              "-L", "/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN/../lib",
              "-Wl,-rpath,/opt/rocm/lib"])

    # 4. the C plan-loop timer (a C caller of xg_plan_execute, used by the bench)
    loop_src = os.path.join(HERE, "cpp", "plan_loop.cpp")
    loop_lib = os.path.join(LIBDIR, "libxgloop.so")
    if os.path.exists(loop_src) and (force or not os.path.exists(loop_lib) or
                                     _newest([loop_src, LIB] + headers) > os.path.getmtime(loop_lib)):
        _run(["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include",
              "-I", INCLUDE, loop_src, "-o", loop_lib, "-L", LIBDIR, "-lxg", "-L", "/opt/rocm/lib", "-lamdhip64",
              "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
