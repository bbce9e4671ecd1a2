"""Build check: the gfx950 code objects of libzkgpu.so contain no device function calls.

Every device function of this library is force-inlined. An out-of-line device function that is
large enough for relaxed long branches (s_getpc/s_setpc) can have its return address (s[30:31])
reused by those branches, so it never returns. This check disassembles each compiled object's
gfx950 code object (the library's own build output) and fails if any `s_swappc_b64` (a call)
appears. Usage: python check_code_objects.py OBJ.o [OBJ.o ...]
"""
import os
import subprocess
import sys
import tempfile

LLVM = "/opt/rocm/llvm/bin"


def code_object(obj: str, tmp: str) -> str:
    fat = os.path.join(tmp, os.path.basename(obj) + ".fatbin")
    subprocess.run([f"{LLVM}/llvm-objcopy", f"--dump-section=.hip_fatbin={fat}", obj, os.devnull],
                   check=True, capture_output=True)
    targets = subprocess.run([f"{LLVM}/clang-offload-bundler", "--type=o", "--list", f"--input={fat}"],
                             check=True, capture_output=True, text=True).stdout.split()
    target = [t for t in targets if "gfx950" in t][0]
    co = fat + ".co"
    subprocess.run([f"{LLVM}/clang-offload-bundler", "--type=o", f"--targets={target}", f"--input={fat}",
                    f"--output={co}", "--unbundle"], check=True, capture_output=True)
    return co


def calls_in(obj: str):
    with tempfile.TemporaryDirectory() as tmp:
        try:
            co = code_object(obj, tmp)
        except (subprocess.CalledProcessError, IndexError):
            return None  # no device code in this object
        dis = subprocess.run([f"{LLVM}/llvm-objdump", "-d", "--no-show-raw-insn", co], check=True,
                             capture_output=True, text=True).stdout
    fn, found = None, []
    for line in dis.splitlines():
        if line.endswith(">:"):
            fn = line.split("<", 1)[1][:-2]
        elif "s_swappc_b64" in line:
            found.append(fn)
    return found


def main(objs):
    bad = {}
    for o in objs:
        found = calls_in(o)
        if found:
            bad[o] = sorted(set(found))
    if bad:
        for o, fns in bad.items():
            print(f"{o}: device function calls in {fns}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
