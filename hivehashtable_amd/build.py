"""Build libhive.so (HIP, gfx950) in-tree with hipcc.

    python -m hivehashtable_amd.build          # or __graft_entry__.build()

Each .hip translation unit is compiled in parallel, then linked into
hivehashtable_amd/libhive.so.  Packed-FP32 instructions are disabled for all
device code (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libhive.so")
SOURCES = ["hive_capi.hip", "hive_compat.hip", "hive_stream.hip", "hive_serve.hip", "hive_sequence.hip",
           "hive_window.hip"]
ARCH = os.environ.get("HIVE_OFFLOAD_ARCH", "gfx950")


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found")
    return p


def flags():
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-command-line-argument", f"-I{os.path.join(ROOT, 'include')}"]


def _compile(src, obj):
    cmd = [hipcc()] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


CLI_SRC = os.path.join(PKG, "cli", "hive_cli.cpp")
CLI_BIN = os.path.join(PKG, "bin")
CLI_NAMES = {1: "hive_hash_table_benchmark", 2: "hive_hash_table_ycsb", 3: "hive_hash_table_lookup_only_workload"}


def build_cli(verbose=True):
    """The reference's three benchmark binaries over libhive.so (host code, g++)."""
    os.makedirs(CLI_BIN, exist_ok=True)
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(hipcc()))), "lib")
    for mode, name in CLI_NAMES.items():
        out = os.path.join(CLI_BIN, name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-DHIVE_CLI_MODE={mode}", f"-I{os.path.join(ROOT, 'include')}",
               CLI_SRC, "-o", out + ".tmp", f"-L{PKG}", "-lhive", f"-L{rocm_lib}", "-Wl,-rpath,$ORIGIN/..",
               f"-Wl,-rpath,{rocm_lib}"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {name}:\n{r.stderr}")
        os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {CLI_BIN}/{{{','.join(CLI_NAMES.values())}}}")


def build(verbose=True):
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = [os.path.join(bdir, s.replace(".hip", ".o")) for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        list(ex.map(_compile, SOURCES, objs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    if verbose:
        print(f"built {LIB}")
    build_cli(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
