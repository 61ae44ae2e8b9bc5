"""Builds libbi_amd/lib/libbihip.so for gfx950 with hipcc.

Every HIP source is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); objects go to
libbi_amd/_build/ and the shared library to libbi_amd/lib/.  The LibBi model
files in libbi_amd/bi/ are turned into device functors first
(libbi_amd/dsl) and compiled in, as the reference compiles a program per
model file.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libbihip.so")
PROBE_LIB = os.path.join(LIBDIR, "libbiprobe.so")
MODELS = os.path.join(HERE, "bi")  # LibBi model files compiled into the library
GEN = os.path.join(OBJ, "gen")

SOURCES = ["transition.hip", "reduce.hip", "scan.hip", "resample.hip", "batch.hip", "capi.cpp",
           "batch_capi.cpp", "smc2.cpp",
           "mlp.hip", "mlp_capi.cpp", "surrogate.hip",
           "guided.hip", "netcdf.cpp", "prior.cpp", "noalloc.hip", "ancestry.hip", "bank.cpp"]
ARCH = os.environ.get("BIHIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(ROOT, "include"), "-I", GEN,
]


def generate_models():
    """Writes the functors of the model files in libbi_amd/bi into the build
    folder (generated_models.hpp, included by the kernels) and the priors of
    their parameter blocks (generated_priors.hpp, host data)."""
    sys.path.insert(0, ROOT)
    from libbi_amd.dsl import codegen
    paths = glob.glob(os.path.join(MODELS, "*.bi"))
    models = codegen.load_models(paths)
    os.makedirs(GEN, exist_ok=True)
    for name, text in (("generated_models.hpp", codegen.generate(models)),
                       ("generated_priors.hpp", codegen.generate_priors(models))):
        out = os.path.join(GEN, name)
        old = open(out).read() if os.path.exists(out) else None
        if old != text:
            with open(out, "w") as f:
                f.write(text)
    return GEN


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    generate_models()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC] + CFLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # the host pass does not know the packed-fp32 feature; that warning is harmless
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln]
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write("\n".join(lines) + "\n")
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # the VALU issue-rate probe that bench.py --workload valu times (not part
    # of the C ABI of bihip.h)
    tmp = PROBE_LIB + ".tmp"
    p = subprocess.run([HIPCC] + CFLAGS + ["-x", "hip", "-shared", os.path.join(CSRC, "probe.hip"),
                                           "-o", tmp], capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("hipcc failed for probe.hip")
    os.replace(tmp, PROBE_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
