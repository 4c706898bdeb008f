"""Build step of the package: compiles every native piece in-tree.

  lib/libtflows_gpu.so     HIP kernels + C ABI (hipcc, gfx950)
  lib/libtflows_cases.so   synthetic problem generators (gcc)
  fortran/                 iso_c_binding module + driver (amdflang, optional)
  ../oracle/liboracle*.so  CPU restatement used as the checker (gcc, make)

Run `python t-flows_amd/build_native.py` or `__graft_entry__.build()`.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["csrc/tfg_api.hip", "csrc/tfg_kernels.hip", "csrc/tfg_ir.hip", "csrc/tfg_ir_ns.hip",
               "csrc/tfg_nonsym.hip", "csrc/tfg_batch.hip", "csrc/tfg_icc.hip",
               "csrc/tfg_comm.hip", "csrc/tfg_assembly.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # every product/update is rounded on its own, as in the reference's
    # FP64 arithmetic; no fused multiply-add contraction
    "-ffp-contract=off",
    # packed FP32 VALU ops are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd):
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}")


def build_gpu():
    os.makedirs(os.path.join(HERE, "lib"), exist_ok=True)
    out = "lib/libtflows_gpu.so"
    _run([HIPCC, *HIP_FLAGS, f"-I{ROOT}/include", "-Icsrc", "-o", out + ".tmp", *HIP_SOURCES], HERE)
    os.replace(os.path.join(HERE, out + ".tmp"), os.path.join(HERE, out))


def build_cases():
    os.makedirs(os.path.join(HERE, "lib"), exist_ok=True)
    _run(["gcc", "-O3", "-fPIC", "-std=gnu11", "-fopenmp", "-Wall", "-shared", "-o",
          "lib/libtflows_cases.so", "csrc/cases.c", "-lm"], HERE)


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], ROOT)


def build_fortran():
    fc = shutil.which("amdflang") or shutil.which("flang")
    if not fc:
        print("amdflang not found: Fortran binding not built", flush=True)
        return False
    fdir = os.path.join(HERE, "fortran")
    src = os.path.join(fdir, "tflows_gpu_mod.f90")
    if not os.path.exists(src):
        return False
    _run([fc, "-O2", "-c", "-fPIC", "-module-dir", fdir, "-o", "tflows_gpu_mod.o",
          "tflows_gpu_mod.f90"], fdir)
    drv = os.path.join(fdir, "solve_driver.f90")
    if os.path.exists(drv):
        _run([fc, "-O2", "-I", fdir, "-o", "solve_driver", "solve_driver.f90", "tflows_gpu_mod.o",
              f"-L{HERE}/lib", "-ltflows_gpu", f"-Wl,-rpath,{HERE}/lib",
              "-Wl,-rpath,$ORIGIN/../lib"], fdir)
    return True


def build_all():
    build_cases()
    build_oracle()
    build_gpu()
    build_fortran()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "gpu": build_gpu, "cases": build_cases, "oracle": build_oracle,
         "fortran": build_fortran}[w]()
