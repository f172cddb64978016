"""Build-time check of the gfx950 code objects: no kernel may call a
non-inlined function or use a dynamic stack.

A device function that is not inlined is reached through s_swappc_b64 and
keeps its frame in the caller's private segment; one such callee once wrote
through flat accesses beyond the segment it was given and faulted the GPU.
Every device function here is meant to be inlined (always_inline), so a call
instruction in the disassembly, or a kernel whose metadata reports a dynamic
stack, fails the build. Usage: python3 check_kernels.py OBJ.o [...] (objects
compiled by hipcc for gfx950)."""
import os
import re
import subprocess
import sys
import tempfile

LLVM = "/opt/rocm/llvm/bin"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def code_object(obj, tmp):
    fatbin = os.path.join(tmp, os.path.basename(obj) + ".fatbin")
    co = os.path.join(tmp, os.path.basename(obj) + ".co")
    subprocess.check_call([f"{LLVM}/llvm-objcopy", f"--dump-section=.hip_fatbin={fatbin}", obj, os.devnull])
    subprocess.check_call([f"{LLVM}/clang-offload-bundler", "--type=o", f"--targets={TARGET}", f"--input={fatbin}",
                           f"--output={co}", "--unbundle"])
    return co


def check(obj, tmp):
    co = code_object(obj, tmp)
    asm = subprocess.run([f"{LLVM}/llvm-objdump", "-d", co], capture_output=True, text=True, check=True).stdout
    problems = []
    fn = None
    for line in asm.splitlines():
        m = re.match(r"^[0-9a-f]+ <(\S+)>:", line)
        if m:
            fn = m.group(1)
        elif "s_swappc_b64" in line:
            problems.append(f"{obj}: {fn} calls a non-inlined function")
    notes = subprocess.run([f"{LLVM}/llvm-readelf", "--notes", co], capture_output=True, text=True,
                           check=True).stdout
    name = None
    for line in notes.splitlines():
        m = re.search(r"\.name:\s+(\S+)", line)
        if m:
            name = m.group(1)
        if re.search(r"\.uses_dynamic_stack:\s+true", line):
            problems.append(f"{obj}: {name} uses a dynamic stack")
    return problems


def main(objs):
    problems = []
    with tempfile.TemporaryDirectory() as tmp:
        for obj in objs:
            problems += check(obj, tmp)
    for p in problems:
        print(p, file=sys.stderr)
    if problems:
        sys.exit(1)
    print(f"check_kernels: {len(objs)} objects, no calls and no dynamic stacks")


if __name__ == "__main__":
    main(sys.argv[1:])
