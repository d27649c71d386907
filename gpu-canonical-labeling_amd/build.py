"""Build recipe of libgcl (HIP device engine + host C++), the CLI and the
CPU oracle.  Everything is compiled in-tree for gfx950:

    python gpu-canonical-labeling_amd/build.py

produces
    gpu-canonical-labeling_amd/libgcl.so        C ABI of include/gcl.h
    gpu-canonical-labeling_amd/gpu_canon_label  CLI (reference main.cpp surface)
    oracle/liboracle.so                         CPU oracle (test checker)
    hbm_probe                                   1 GiB copy kernel that calibrates
                                                the traffic counters
    oracle/_ref/gpu_canon_label_ref             reference sequential binary,
                                                only where the reference
                                                sources are present
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
INC = os.path.join(REPO, "include")
ORACLE = os.path.join(REPO, "oracle")
REFERENCE = "/root/reference"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

CXXFLAGS = ["-O3", "-g", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
            "-I" + INC, "-I" + CSRC, "-pthread"]
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I" + INC, "-I" + CSRC, "-pthread"]

HOST_SOURCES = ["gcl_api.cpp", "gcl_cpu.cpp", "gcl_graph.cpp", "gcl_server.cpp", "gcl_workload.cpp"]
DEVICE_SOURCES = ["gcl_device.hip"]


def _run_all(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((c, text))
        elif text.strip():
            sys.stderr.write(text)
    if failed:
        for c, text in failed:
            sys.stderr.write("FAILED: " + " ".join(c) + "\n" + text + "\n")
        raise RuntimeError("build failed")


def build(with_reference=True):
    os.makedirs(OBJ, exist_ok=True)
    cmds = []
    objs = []
    for src in HOST_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        cmds.append([CXX] + CXXFLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in DEVICE_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        cmds.append([HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    cmds.append(["make", "-s", "-B", "-C", ORACLE, "liboracle.so"])
    if with_reference and os.path.isdir(REFERENCE):
        cmds.append(["make", "-s", "-B", "-C", ORACLE, "ref", "REF=" + REFERENCE])
    _run_all(cmds)
    lib = os.path.join(PKG, "libgcl.so")
    _run_all([[HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", lib] + objs])
    # the copy kernel that calibrates the HBM traffic counters (beside bench.py)
    probe = os.path.join(REPO, "hbm_probe")
    _run_all([[HIPCC] + HIPFLAGS + [os.path.join(REPO, "hbm_probe.hip"), "-o", probe]])
    cli = os.path.join(PKG, "gpu_canon_label")
    _run_all([[CXX] + CXXFLAGS + [os.path.join(CSRC, "gpu_canon_label.cpp"), "-o", cli,
                                  "-L" + PKG, "-lgcl", "-Wl,-rpath,$ORIGIN"]])
    return lib


if __name__ == "__main__":
    build()
    print("built", os.path.join(PKG, "libgcl.so"))
