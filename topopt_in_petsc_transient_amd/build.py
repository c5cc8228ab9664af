"""Build libtopopt_gpu.so (HIP kernels + C-ABI) and the topopt driver for gfx950.

Compiles every source under csrc/ with hipcc in parallel, in-tree, so that the
built shared library travels with the source tree.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtopopt_gpu.so")
DRIVER = os.path.join(HERE, "topopt_gpu")
ARCH = os.environ.get("TT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

LIB_SOURCES = ["tt_kernels.hip", "tt_kernels2.hip", "tt_element.cpp", "tt_solver.cpp", "tt_transient.cpp", "tt_reduced.cpp",
               "tt_capi.cpp", "tt_noalloc.hip", "tt_tables.cpp"]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
             # no packed FP32 instructions (v_pk_*_f32) in any generated code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # matrix-core results straight into VGPRs (no accumulation-register copies in the tile loop)
             "-mllvm", "-amdgpu-mfma-vgpr-form=1"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, f"--offload-arch={ARCH}", *CXX_FLAGS, "-I", os.path.join(HERE, "..", "include"), "-c",
           os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd.insert(1, "-xhip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    drv_src = os.path.join(CSRC, "topopt_main.cpp")
    if os.path.exists(drv_src):
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-xhip", "-I", os.path.join(HERE, "..", "include"),
               "-I", CSRC, drv_src, os.path.join(CSRC, "topopt_classes.cpp"), "-o", DRIVER + ".tmp",
               "-L", HERE, "-ltopopt_gpu", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


# AddressSanitizer + UndefinedBehaviorSanitizer on the host code (the option parser, the PETSc binary
# reader and writer, the set-up and the C-ABI): -Xarch_host puts the sanitizers on the host compilation
# only; GPU code is built as usual.  The library leaves the sanitizer runtime unresolved: the process
# that loads it preloads clang's shared runtime (tests/test_sanitizers_cpu.py).
SAN_HOST_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                  "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
                  "-Xarch_host", "-g"]


def build_sanitized(out_dir, jobs=None):
    """The library with sanitized host code, as out_dir/libtopopt_gpu.so (not the product library)."""
    os.makedirs(out_dir, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)

    def comp(src):
        obj = os.path.join(out_dir, os.path.basename(src) + ".o")
        cmd = [HIPCC, f"--offload-arch={ARCH}", *CXX_FLAGS, *SAN_HOST_FLAGS, "-I", os.path.join(HERE, "..", "include"),
               "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd.insert(1, "-xhip")
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        return obj

    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(comp, LIB_SOURCES))
    lib = os.path.join(out_dir, "libtopopt_gpu.so")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib, *objs, "-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return lib


if __name__ == "__main__":
    build(verbose=True)
