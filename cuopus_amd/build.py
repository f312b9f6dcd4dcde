"""Build steps: the HIP library (gfx950), the same-source host gate and the
CPU oracle.

* ``build_lib()``      hipcc -> cuopus_amd/libcuopus.so (the product).
* ``build_hostgate()`` the device decoder source compiled for the host CPU
                       (tests only) -> tests/hostgate/libhostgate.so.
* ``build_oracle()``   make -C oracle (libopus 1.5.2 from the reference tree,
                       tests / CPU baseline only) -> oracle/_ref/.
* ``build_twins()``    the library with the serial twins of the lane-parallel
                       stages on the GPU (tests only) ->
                       tests/hostgate/libcuopus_twins.so.
"""
import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cuopus_amd")
CSRC = os.path.join(PKG, "csrc")
REFERENCE = "/root/reference"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", "-Wno-unused-function",
             # packed fp32 instructions stay off: some of their forms have given
             # wrong results on this hardware without any error
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), r.stdout[-6000:]))
    return r.stdout


def build_lib(arch="gfx950", out=None, defines=()):
    """defines: extra preprocessor flags, e.g. ("CUO_PROF",) for the
    cycle-accounting build (written to `out`, never over the product)."""
    out = out or os.path.join(PKG, "libcuopus.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + arch] + HIP_FLAGS + ["-D" + d for d in defines] + [
        "-shared", "-I" + CSRC,
        os.path.join(CSRC, "libcuopus.hip"), os.path.join(CSRC, "stereo.hip"), os.path.join(CSRC, "host", "demux.cpp"),
        "-o", tmp]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_twins():
    """tests/hostgate/libcuopus_twins.so: the library with CUO_SERIAL_TWINS,
    i.e. the serial twins of the wavefront-parallel decoder stages (the code
    the host gate and its sanitizer build run) compiled for the GPU.  Test
    infrastructure: tests/test_gpu_twins.py compares it with the product."""
    return build_lib(out=os.path.join(ROOT, "tests", "hostgate", "libcuopus_twins.so"), defines=("CUO_SERIAL_TWINS",))


def build_cli(out=None):
    """cuopus_amd/cuopus_decode: the command-line front-end (host code over
    the C ABI, linked to libcuopus.so beside it)."""
    out = out or os.path.join(PKG, "cuopus_decode")
    cmd = [HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           os.path.join(CSRC, "cli", "cuopus_decode.cpp"), "-L" + PKG, "-lcuopus",
           "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_hostgate():
    """tests/hostgate/libhostgate.so (mono build) and libhostgate_st.so (the
    stereo-capable build, CUO_MAXCH=2)."""
    src = os.path.join(ROOT, "tests", "hostgate", "hostgate.cpp")
    outs = []
    for name, defs in (("libhostgate.so", []), ("libhostgate_st.so", ["-DCUO_MAXCH=2", "-DCUO_NS=cuo_st"])):
        out = os.path.join(ROOT, "tests", "hostgate", name)
        cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-fPIC", "-shared",
               "-Wno-unused-function"] + defs + [src, "-o", out + ".tmp"]
        _run(cmd)
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_fuzz_gate():
    """tests/hostgate/fuzz_gate and fuzz_gate_st (the stereo build): the
    decoder source on the CPU under AddressSanitizer +
    UndefinedBehaviorSanitizer (host code only; the sanitizer flags go to the
    host compilation), linked to the oracle."""
    ref = os.path.join(ROOT, "oracle", "_ref")
    if not os.path.exists(os.path.join(ref, "liboracle.so")):
        return None
    src = os.path.join(ROOT, "tests", "hostgate", "fuzz_gate.cpp")
    outs = []
    for name, defs in (("fuzz_gate", []), ("fuzz_gate_st", ["-DCUO_MAXCH=2", "-DCUO_NS=cuo_st"])):
        out = os.path.join(ROOT, "tests", "hostgate", name)
        cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-O1", "-g", "-std=c++17", "-Wno-unused-function",
               "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
               "-Xarch_host", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"] + defs + [
               "-I" + os.path.join(ROOT, "tests", "hostgate"), src, "-L" + ref, "-loracle", "-Wl,-rpath," + ref,
               "-o", out + ".tmp"]
        _run(cmd)
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_oracle(jobs=8):
    """Compile the oracle when the reference tree is present; otherwise keep
    whatever prebuilt oracle/_ref is there (the GPU box has no reference)."""
    opus = os.path.join(REFERENCE, "vendor", "opus-1.5.2")
    if not os.path.isdir(opus):
        return None
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-j%d" % jobs, "OPUS=" + opus])
    return os.path.join(ROOT, "oracle", "_ref")


def build_all():
    build_oracle()
    build_lib()
    build_cli()
    build_hostgate()
    build_fuzz_gate()
    build_twins()
