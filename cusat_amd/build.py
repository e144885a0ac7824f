#!/usr/bin/env python3
"""Build libcusat_fec.so (gfx950 HIP kernels + C ABI + host code) in-tree.

    python -m cusat_amd.build        # or: python cusat_amd/build.py

Every call recompiles.  Objects go to cusat_amd/_build/, the library to
cusat_amd/libcusat_fec.so.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libcusat_fec.so")
ARCH = os.environ.get("CUSAT_OFFLOAD_ARCH", "gfx950")

SOURCES = ["codes.cc", "modulation.cc", "bch.cc", "gen.cc", "capi.cc",
           "ldpc_decode.hip", "ldpc_inst_a.hip", "ldpc_inst_b.hip", "ldpc_inst_c.hip",
           "ldpc_inst_d.hip", "ldpc_inst_s.hip", "demap.hip", "bch_check.hip", "learn.hip", "ldpc_stream.hip", "bb.hip",
           "gen_dev.hip"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


# generated header -> (symbol, data file) pairs it embeds
DATA_HEADERS = {
    "data_inc.h": (("k_ldpc_tables_txt", "ldpc_tables.txt"), ("k_ldpc_schedules_txt", "ldpc_schedules.txt")),
    "apsk256_inc.h": (("k_apsk256_points_txt", "apsk256_points.txt"),),
}


def gen_data_header():
    out_dir = os.path.join(CSRC, "_gen")
    os.makedirs(out_dir, exist_ok=True)
    for header, items in DATA_HEADERS.items():
        parts = []
        for sym, fname in items:
            data = open(os.path.join(PKG, "data", fname), "rb").read()
            body = ",".join(str(b) for b in data)
            parts.append("static const char %s[] = {%s,0};\n" % (sym, body))
        path = os.path.join(out_dir, header)
        text = "// generated by cusat_amd/build.py from cusat_amd/data/*.txt\n#pragma once\n" + "".join(parts)
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as f:
                f.write(text)


def compile_one(src):
    cc = hipcc()
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    common = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
              "-I", os.path.join(ROOT, "include"), "-I", CSRC]
    if src.endswith(".hip"):
        cmd = [cc, "-x", "hip", "--offload-arch=" + ARCH, "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-munsafe-fp-atomics"] + common + ["-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [cc, "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__"] + common + \
              ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


HOST = os.path.join(PKG, "host")
HOST_SOURCES = ["ldpc_decode.cc", "rx_fec_chain.cc", "rx_fec_pool.cc", "rx_fec_stream_pool.cc"]
HOST_LIB = os.path.join(PKG, "libcusat_dsp.so")
CPP_TESTS = os.path.join(ROOT, "tests", "cpp")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed: %s\n%s\n%s" % (what, " ".join(cmd), r.stdout, r.stderr))


def build_host(verbose=True):
    """The C++ cusat_dsp API (ldpc_decode, rx_fec_chain, rx_fec_pool) over the C ABI:
    libcusat_dsp.so next to libcusat_fec.so, and the C++ test programs of
    tests/cpp (tests/cpp/_build/<name>)."""
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
             "-I", os.path.join(ROCM, "include")]
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, "host_" + src + ".o")
        _run(["g++"] + flags + ["-c", os.path.join(HOST, src), "-o", obj], "compile " + src)
        objs.append(obj)
    tmp = HOST_LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp] + objs + ["-L", PKG, "-l:libcusat_fec.so", "-L", os.path.join(ROCM, "lib"),
                                                 "-lamdhip64", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")],
         "link libcusat_dsp.so")
    os.replace(tmp, HOST_LIB)
    # the pool throughput benchmark (cusat_amd/pool_bench, next to the libraries)
    bench = os.path.join(PKG, "pool_bench")
    _run(["g++"] + flags + [os.path.join(HOST, "pool_bench.cc"), "-o", bench + ".tmp", "-L", PKG,
                            "-l:libcusat_dsp.so", "-l:libcusat_fec.so", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
                            "-lpthread", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")],
         "compile pool_bench.cc")
    os.replace(bench + ".tmp", bench)
    out = os.path.join(CPP_TESTS, "_build")
    os.makedirs(out, exist_ok=True)
    for fn in sorted(os.listdir(CPP_TESTS)):
        if not fn.endswith(".cc"):
            continue
        exe = os.path.join(out, fn[:-3])
        _run(["g++"] + flags + [os.path.join(CPP_TESTS, fn), "-o", exe + ".tmp", "-L", PKG, "-l:libcusat_dsp.so",
                                "-l:libcusat_fec.so", "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-lpthread",
                                "-Wl,-rpath,$ORIGIN/../../../cusat_amd", "-Wl,-rpath," + os.path.join(ROCM, "lib")],
             "compile " + fn)
        os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", HOST_LIB)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    gen_data_header()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
