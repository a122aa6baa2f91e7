"""Build the gfx950 transport library in-tree: mohid_amd/libmohid_gpu.so.

    python -m mohid_amd.build

Flags: -ffp-contract=off keeps the reference's operation order (the FP64
path is bit-for-bit with the reference's arithmetic); packed FP32 is switched
off for every HIP source.
"""
from __future__ import annotations

import glob
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmohid_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MGPU_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    # per-kernel register/LDS/scratch report, checked below
    "-Rpass-analysis=kernel-resource-usage",
]


def check_resources(text: str) -> None:
    """Fail the build if a kernel spills to scratch memory: every kernel here
    is sized to keep its state in registers, and a spill (for instance an
    array the compiler could not keep in registers) turns into hidden HBM
    traffic.  SGPR spills go to VGPR lanes and VGPR spills with no scratch
    bytes go to the accumulation registers (no memory traffic either); a
    kernel with SGPR spills also reports the register scavenger's reserved
    emergency slot as scratch, which is accepted while no VGPR spills."""
    name = None
    info = {}
    for line in text.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            info[name] = {}
            continue
        m = re.search(r"(ScratchSize \[bytes/lane\]|SGPRs Spill|VGPRs Spill): (\d+)", line)
        if m and name:
            info[name][m.group(1).split()[0]] = int(m.group(2))
    bad = []
    for name, r in info.items():
        scratch, sspill, vspill = r.get("ScratchSize", 0), r.get("SGPRs", 0), r.get("VGPRs", 0)
        if (vspill > 0 and scratch > 0) or (scratch > 0 and sspill == 0):
            bad.append(f"{name}: {scratch} bytes/lane scratch, {vspill} VGPRs spilled")
    if bad:
        raise RuntimeError("kernels spill to scratch memory:\n  " + "\n  ".join(bad))


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FDIR, "_build")
DRIVER = os.path.join(FDIR, "adgpu_driver")
PRODUCERS_DRIVER = os.path.join(FDIR, "producers_driver")
# the reference's production build is default-real-8 (compile_mohid.sh:185)
FFLAGS = ["-cpp", "-O2", "-fdefault-real-8", "-fdefault-double-8"]


def build_fortran(verbose: bool = False) -> str:
    """The ISO_C_BINDING module (ModuleAdvectionDiffusionGPU) and the test
    driver program, linked against the HIP library."""
    os.makedirs(FBUILD, exist_ok=True)
    mod_o = os.path.join(FBUILD, "ModuleAdvectionDiffusionGPU.o")
    drv_o = os.path.join(FBUILD, "adgpu_driver.o")
    prd_o = os.path.join(FBUILD, "producers_driver.o")
    steps = [
        [FC, *FFLAGS, "-c", os.path.join(FDIR, "ModuleAdvectionDiffusionGPU.F90"), "-o", mod_o,
         "-J", FBUILD],
        [FC, *FFLAGS, "-c", os.path.join(FDIR, "adgpu_driver.F90"), "-o", drv_o, "-J", FBUILD],
        [FC, "-o", DRIVER + ".tmp", drv_o, mod_o, f"-L{HERE}", "-lmohid_gpu",
         "-Wl,-rpath,$ORIGIN/.."],
        [FC, *FFLAGS, "-c", os.path.join(FDIR, "producers_driver.F90"), "-o", prd_o, "-J",
         FBUILD],
        [FC, "-o", PRODUCERS_DRIVER + ".tmp", prd_o, mod_o, f"-L{HERE}", "-lmohid_gpu",
         "-Wl,-rpath,$ORIGIN/.."],
    ]
    for cmd in steps:
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"{' '.join(cmd)} failed:\n{res.stdout}{res.stderr}")
        if verbose and (res.stdout or res.stderr).strip():
            print(res.stdout + res.stderr)
    os.replace(DRIVER + ".tmp", DRIVER)
    os.replace(PRODUCERS_DRIVER + ".tmp", PRODUCERS_DRIVER)
    return DRIVER


OBJ = os.path.join(HERE, "_build")


def _filtered(text: str) -> str:
    return "\n".join(l for l in text.splitlines()
                     if ("error" in l or "warning" in l) and "packed-fp32-ops" not in l)


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compile every HIP source to an object (in parallel: the transport
    kernels are split over several translation units), check the kernels'
    resource use, and link the shared library."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1), 16)
    objs = [os.path.join(OBJ, os.path.basename(f)[:-4] + ".o") for f in srcs]
    procs = []
    pending = list(zip(srcs, objs))
    stderr = []
    while pending or procs:
        while pending and len(procs) < jobs:
            src, obj = pending.pop(0)
            cmd = [HIPCC, *FLAGS, "-c", "-o", obj, src]
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.PIPE, text=True)))
        cmd, p = procs.pop(0)
        out, err = p.communicate()
        if p.returncode != 0:
            for _, q in procs:
                q.kill()
            raise RuntimeError(f"hipcc failed ({p.returncode}):\n{' '.join(cmd)}\n"
                               f"{_filtered(out + err)}")
        stderr.append(err)
    check_resources("\n".join(stderr))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed ({res.returncode}):\n{' '.join(cmd)}\n"
                           f"{_filtered(res.stdout + res.stderr)}")
    if verbose:
        text = _filtered("\n".join(stderr))
        if text:
            print(text)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
    if os.path.exists(FC):
        print(build_fortran(verbose=True))
    sys.exit(0)
