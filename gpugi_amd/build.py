"""Builds the gfx950 shared library libgpugi_rt.so in-tree.

Every translation unit (host C++ and HIP) is compiled by hipcc with
-ffp-contract=off: the integrators and the triangle test must round exactly
like the restated GLSL, on the host and on the device.
"""
import concurrent.futures
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpugi_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "lib", "libgpugi_rt.so")
# script-driven front end over the C++ renderer layer (links libgpugi_rt.so)
APP = os.path.join(PKG, "bin", "gpugi_render")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
DEVICE = [
    "--offload-arch=gfx950", "-munsafe-fp-atomics",
    # no packed-FP32 instructions in any generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    host = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host")) if f.endswith(".cpp"))
    dev = sorted(os.path.join(CSRC, "device", f) for f in os.listdir(os.path.join(CSRC, "device")) if f.endswith(".hip"))
    return host, dev


def app_sources():
    d = os.path.join(CSRC, "renderer")
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(".cpp"))


def _compile(src, is_device):
    obj = os.path.join(BUILD, os.path.relpath(src, CSRC).replace(os.sep, "_") + ".o")
    cmd = [HIPCC] + COMMON
    if is_device:
        cmd += DEVICE + ["-x", "hip"]
    else:
        cmd += ["-x", "c++"]
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj, r.stderr


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    host, dev = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, False) for s in host] + [ex.submit(_compile, s, True) for s in dev]
        for f in futs:
            obj, warn = f.result()
            objs.append(obj)
            if verbose and warn.strip():
                sys.stderr.write(warn)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)

    # the renderer layer uses nothing but the C ABI: plain host C++ linked against the library
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        app_objs = [o for o, _ in ex.map(lambda s: _compile(s, False), app_sources())]
    os.makedirs(os.path.dirname(APP), exist_ok=True)
    tmp = APP + ".tmp"
    cmd = [HIPCC, "-o", tmp] + app_objs + ["-L" + os.path.dirname(LIB), "-lgpugi_rt",
                                          "-Wl,-rpath,$ORIGIN/../lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, APP)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
