"""Reads the per-kernel resource metadata (register counts, spills, scratch)
and the function symbols of every gfx950 code object in a built HIP shared
library: the .hip_fatbin section holds one clang offload bundle per
translation unit, each with an AMDGPU ELF whose NT_AMDGPU_METADATA note
lists the kernels.  Uses llvm-objcopy and llvm-readelf from the ROCm
toolchain; no GPU is needed."""

import os
import re
import struct
import subprocess
import tempfile

LLVM_BIN = os.environ.get("ROCM_LLVM_BIN", "/opt/rocm/llvm/bin")
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def tools_available():
    return all(os.access(os.path.join(LLVM_BIN, t), os.X_OK) for t in ("llvm-objcopy", "llvm-readelf"))


def _bundles(fatbin):
    """Yields (triple, bytes) of every entry of every offload bundle."""
    pos = fatbin.find(MAGIC)
    while pos >= 0:
        (count,) = struct.unpack_from("<Q", fatbin, pos + len(MAGIC))
        p = pos + len(MAGIC) + 8
        for _ in range(count):
            off, size, tlen = struct.unpack_from("<QQQ", fatbin, p)
            triple = fatbin[p + 24:p + 24 + tlen].decode()
            p += 24 + tlen
            yield triple, fatbin[pos + off:pos + off + size]
        pos = fatbin.find(MAGIC, pos + len(MAGIC))


def _parse_notes(text):
    kernels = {}
    cur = None
    for line in text.splitlines():
        m = re.match(r"\s*(?:- )?\.(\w+):\s*(\S*)\s*$", line)
        if not m:
            continue
        key, val = m.group(1), m.group(2)
        if line.lstrip().startswith("- .agpr_count"):
            cur = {}
        if cur is None:
            continue
        if key == "name":
            kernels[val] = cur
        elif re.fullmatch(r"-?\d+", val or ""):
            cur[key] = int(val)
    return kernels


def _functions(text):
    """Names of FUNC symbols (kernels and called device functions)."""
    out = set()
    for line in text.splitlines():
        f = line.split()
        if len(f) >= 8 and f[3] == "FUNC":
            out.add(f[7])
    return out


def code_objects(lib_path):
    """Per translation unit: {"kernels": {mangled: metadata}, "functions": set}."""
    with tempfile.TemporaryDirectory() as tmp:
        fat = os.path.join(tmp, "fat.bin")
        subprocess.run([os.path.join(LLVM_BIN, "llvm-objcopy"), "--dump-section", ".hip_fatbin=" + fat, lib_path,
                        os.path.join(tmp, "lib.copy")], check=True, capture_output=True)
        data = open(fat, "rb").read()
        result = []
        for k, (triple, blob) in enumerate(_bundles(data)):
            if "gfx950" not in triple:
                continue
            path = os.path.join(tmp, "co%d.o" % k)
            with open(path, "wb") as f:
                f.write(blob)
            notes = subprocess.run([os.path.join(LLVM_BIN, "llvm-readelf"), "--notes", path], check=True,
                                   capture_output=True, text=True).stdout
            syms = subprocess.run([os.path.join(LLVM_BIN, "llvm-readelf"), "--symbols", path], check=True,
                                  capture_output=True, text=True).stdout
            result.append({"kernels": _parse_notes(notes), "functions": _functions(syms)})
        return result


def demangle(names):
    p = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True)
    return p.stdout.splitlines() if p.returncode == 0 else list(names)


if __name__ == "__main__":
    import sys

    lib = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.dirname(
        os.path.abspath(__file__))), "blast_amd", "libblast_hip.so")
    for co in code_objects(lib):
        calls = sorted(set(co["functions"]) - set(co["kernels"]))
        names = list(co["kernels"])
        for mangled, pretty in zip(names, demangle(names)):
            m = co["kernels"][mangled]
            print("%-100s sgpr=%3d vgpr=%3d agpr=%3d scratch=%4d sspill=%3d vspill=%3d" % (
                pretty[:100], m.get("sgpr_count", -1), m.get("vgpr_count", -1), m.get("agpr_count", -1),
                m.get("private_segment_fixed_size", -1), m.get("sgpr_spill_count", -1),
                m.get("vgpr_spill_count", -1)))
        if calls:
            print("   called functions:", ", ".join(demangle(calls)))
