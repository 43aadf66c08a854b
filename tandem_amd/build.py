"""Build libtandem_dg.so (host C++ + HIP kernels for gfx950) in-tree under tandem_amd/_build/.

    python -m tandem_amd.build        # or python tandem_amd/build.py

Every call recompiles (objects whose source and flags are unchanged are reused).
"""
import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libtandem_dg.so")
RULES = os.path.join(HERE, "data", "simplex_rules.txt")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TDG_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-variable", "-Wno-sign-compare",
          "-I" + os.path.join(ROOT, "include")]
# PETSc (optional): with PETSC_DIR set, the MATSHELL adapter (csrc/petsc) is compiled against its
# headers and the library links libpetsc; without it that translation unit is empty.
PETSC_DIR = os.environ.get("PETSC_DIR")
if PETSC_DIR:
    COMMON += ["-I" + os.path.join(PETSC_DIR, "include")]
    _arch = os.environ.get("PETSC_ARCH")
    if _arch:
        COMMON += ["-I" + os.path.join(PETSC_DIR, _arch, "include")]
# Packed FP32 instructions are disabled for all device code on this platform.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources():
    out = []
    for d, _, files in os.walk(CSRC):
        for f in sorted(files):
            if f.endswith((".cpp", ".hip")):
                out.append(os.path.join(d, f))
    return sorted(out)


def write_rule_blob():
    """Embed the quadrature tables as a C string so the library has no data-file dependency."""
    path = os.path.join(BUILD, "rules_blob.cpp")
    with open(RULES) as f:
        text = f.read()
    chunks = [text[i:i + 4000] for i in range(0, len(text), 4000)]
    body = "\n".join('R"TDGRULES(%s)TDGRULES"' % c for c in chunks)
    src = ("namespace tdg {\nchar const* rule_table_text() {\n    static char const text[] =\n%s;\n"
           "    return text;\n}\n}\n" % body)
    old = open(path).read() if os.path.exists(path) else None
    if old != src:
        with open(path, "w") as f:
            f.write(src)
    return path


def compile_one(src):
    rel = os.path.relpath(src, HERE).replace(os.sep, "_")
    flags = COMMON + (DEVICE if src.endswith(".hip") else ["-D__HIP_PLATFORM_AMD__"])
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + flags + ["-c", src]
    else:
        cmd = [HIPCC] + flags + ["-c", src]
    key = hashlib.sha1((" ".join(cmd) + open(src, "rb").read().decode("latin1")).encode("latin1"))
    for dep in _headers():
        key.update(open(dep, "rb").read())
    obj = os.path.join(BUILD, "obj", rel + "." + key.hexdigest()[:12] + ".o")
    if not os.path.exists(obj):
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        r = subprocess.run(cmd + ["-o", obj + ".tmp"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
        os.replace(obj + ".tmp", obj)
    return obj


_HDRS = None


def _headers():
    global _HDRS
    if _HDRS is None:
        _HDRS = []
        for d, _, files in os.walk(CSRC):
            _HDRS += [os.path.join(d, f) for f in files if f.endswith(".h")]
        _HDRS.append(os.path.join(ROOT, "include", "tandem_dg.h"))
        _HDRS.sort()
    return _HDRS


def build(verbose=True, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources() + [write_rule_blob()]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    cmd = [HIPCC, "-shared", "-fopenmp", "--offload-arch=" + ARCH, "-o", LIB + ".tmp"] + objs + [
        "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib/llvm/lib"]
    if PETSC_DIR:
        libdir = os.path.join(PETSC_DIR, os.environ.get("PETSC_ARCH", ""), "lib")
        cmd += ["-L" + libdir, "-Wl,-rpath," + libdir, "-lpetsc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(LIB + ".tmp", LIB)
    # drop stale objects
    keep = set(objs)
    objdir = os.path.join(BUILD, "obj")
    for f in os.listdir(objdir):
        p = os.path.join(objdir, f)
        if p not in keep:
            os.remove(p)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
