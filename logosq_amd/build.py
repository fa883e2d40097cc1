"""Builds liblogosq_hip.so (HIP kernels + C ABI) for gfx950 with hipcc.

Every translation unit is compiled separately (parallel) into
logosq_amd/_build/ and linked into logosq_amd/liblogosq_hip.so, next to the
Python package, so the in-tree library is the one that loads.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblogosq_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops off: hipcc must not form v_pk_* f32 instructions.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++20", "-ffp-contract=fast",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


# Per-file flags.  lq_fused.hip: LLVM's common-code sinking would merge the
# per-position copies of the register subcube code into one block with a
# computed array index, which sends the subcube to scratch memory.
EXTRA = {"lq_fused.hip": ["-mllvm", "-simplifycfg-sink-common=false"], "lq_jit.hip": ["-I", BUILD]}

# Headers the run-time specialised kernels are compiled from (hipRTC), embedded
# in the library as text.
JIT_HEADERS = [("kTileTypesSrc", "lq_tile_types.h"), ("kTileOpsSrc", "lq_tile_ops.h")]
JIT_INC = os.path.join(BUILD, "lq_jit_src.inc")


def write_jit_sources():
    """Write _build/lq_jit_src.inc: each header as a C++ raw string literal
    (rewritten only when the text changes, so incremental builds hold)."""
    out = []
    for var, name in JIT_HEADERS:
        text = open(os.path.join(CSRC, name)).read()
        if ')LQSRC"' in text:
            raise RuntimeError("%s holds the raw-string delimiter" % name)
        out.append('static const char %s[] = R"LQSRC(%s)LQSRC";\n' % (var, text))
    text = "// generated by build.py from csrc headers; do not edit\n" + "".join(out)
    if not os.path.exists(JIT_INC) or open(JIT_INC).read() != text:
        with open(JIT_INC, "w") as f:
            f.write(text)


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer variant
# (build(sanitize=True) -> liblogosq_hip_san.so): only the host pass is
# instrumented (-Xarch_host before each option; device code is built as
# usual), for the C ABI's argument checks, planners and kernel-text
# generators.  Run the host-only tests against it with sanitize_check().
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]
SAN_BUILD = os.path.join(BUILD, "san")
SAN_LIB = os.path.join(HERE, "liblogosq_hip_san.so")


def _compile(src, incremental=False, sanitize=False):
    obj = os.path.join(SAN_BUILD if sanitize else BUILD, os.path.basename(src) + ".o")
    if incremental and os.path.exists(obj):
        deps = [src] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
        deps.append(os.path.join(os.path.dirname(HERE), "include", "logosq_hip.h"))
        if os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps if os.path.exists(d)):
            return obj
    cmd = [HIPCC, *HIP_FLAGS, *(SAN_HOST if sanitize else []), *EXTRA.get(os.path.basename(src), []), "-c", src,
           "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def build(verbose=True, incremental=False, sanitize=False):
    """Compile every source (incremental=True: only sources newer than their
    object file or than any header) and link the library (sanitize=True: the
    host-sanitized variant, liblogosq_hip_san.so)."""
    os.makedirs(BUILD, exist_ok=True)
    if sanitize:
        os.makedirs(SAN_BUILD, exist_ok=True)
    write_jit_sources()
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental, sanitize), srcs))
    out = SAN_LIB if sanitize else LIB
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


# Tests that use the library's host code only (no device).
SAN_TESTS = ["tests/test_abi.py", "tests/test_planner.py", "tests/test_expect_plan.py", "tests/test_jit_build.py",
             "tests/test_expect_jit_emul.py", "tests/test_circuit_api.py", "tests/test_optimizer_pins.py"]


def sanitize_check():
    """Run the host-only tests against liblogosq_hip_san.so with the clang
    sanitizer runtimes preloaded (the interpreter is not instrumented; leak
    checking is off).  Returns the pytest exit status."""
    clang = os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "lib", "llvm", "bin", "clang")
    if not os.path.exists(clang):
        clang = "/opt/rocm/llvm/bin/clang"
    rt = subprocess.run([clang, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                        text=True).stdout.strip()
    env = dict(os.environ, LOGOSQ_HIP_LIB=SAN_LIB, LD_PRELOAD=rt,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    root = os.path.dirname(HERE)
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", *SAN_TESTS], cwd=root,
                          env=env).returncode


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        build(incremental="--incremental" in sys.argv, sanitize=True)
        sys.exit(sanitize_check() if "--check" in sys.argv else 0)
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
