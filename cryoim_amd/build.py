"""Build the native parts in-tree.

* ``cryoim_amd/lib/libcryoim_hip.so`` -- every ``csrc/*.hip`` compiled for
  gfx950 with hipcc (separate objects, parallel, then one shared library);
* ``oracle/liboracle{64,32}.so`` -- the C oracle (test infrastructure);
* ``oracle/_ref/`` -- the reference's own Fortran programs, only where the
  reference sources are present (never on the GPU box);
* ``cryoim_amd/bin/icosprocess_hip``, ``merge_icosprocess_subrec_hip``,
  ``local_reconstruct_hip``, ``local_box_hip`` and ``local_refine_hip`` -- the
  Fortran drivers over the C ABI, when amdflang is available.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcryoim_hip.so")
OBJDIR = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("CIM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-munsafe-fp-atomics",
    # packed FP32 instructions stay off: hipcc forms them from scalar code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def build_hip(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    headers.append(os.path.join(ROOT, "include", "cryoim.h"))
    hdr_time = max(os.path.getmtime(h) for h in headers)

    def one(src):
        s = os.path.join(CSRC, src)
        o = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        cmd = [HIPCC, *HIP_FLAGS, "-c", s, "-o", o]
        out = _run(cmd)
        if verbose and out.strip():
            print(out)
        return o

    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(one, srcs))
    _ = hdr_time
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_reference() -> bool:
    ref = os.environ.get("CIM_REFERENCE", "/root/reference")
    if not os.path.isdir(os.path.join(ref, "icosprocess", "SRC")):
        return False
    if shutil.which("amdflang") is None and not os.path.exists(os.path.join(ROCM, "bin", "amdflang")):
        return False
    env = dict(os.environ)
    env.setdefault("FC", shutil.which("amdflang") or os.path.join(ROCM, "bin", "amdflang"))
    r = subprocess.run([os.path.join(ROOT, "oracle", "build_ref.sh")], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("reference build failed:\n" + r.stdout)
    return True


def build_fortran() -> bool:
    fc = shutil.which("amdflang") or os.path.join(ROCM, "bin", "amdflang")
    fdir = os.path.join(HERE, "fortran")
    if not os.path.exists(fc) or not os.path.isdir(fdir):
        return False
    srcs = [os.path.join(fdir, f) for f in ("cim_iface.f90", "icosprocess_hip.f90")
            if os.path.exists(os.path.join(fdir, f))]
    if len(srcs) < 2:
        return False
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    modd = os.path.join(OBJDIR, "fmod")
    os.makedirs(modd, exist_ok=True)
    link = [f"-L{LIBDIR}", "-lcryoim_hip", f"-Wl,-rpath,{LIBDIR}", f"-Wl,-rpath,{ROCM}/lib", f"-L{ROCM}/lib",
            "-lamdhip64"]
    iface_o = os.path.join(modd, "cim_iface.o")
    _run([fc, "-O2", "-J", modd, "-c", srcs[0], "-o", iface_o])
    for prog in ("icosprocess_hip", "merge_icosprocess_subrec_hip", "local_reconstruct_hip", "local_box_hip",
                 "local_refine_hip"):
        src = os.path.join(fdir, prog + ".f90")
        if os.path.exists(src):
            _run([fc, "-O2", "-J", modd, "-o", os.path.join(bindir, prog), src, iface_o, *link])
    return True


def build_all(verbose: bool = False) -> None:
    build_hip(verbose)
    build_oracle()
    build_reference()
    build_fortran()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"all": build_all, "hip": build_hip, "oracle": build_oracle, "ref": build_reference,
     "fortran": build_fortran}[what]()
    print("built", what)
