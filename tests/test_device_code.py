"""Checks on the gfx950 machine code of the built library (CPU only).

The library's device code is disassembled with the ROCm LLVM tools (our own
build output). The check guards a hazard of out-of-line device functions: a
function larger than the short-branch range gets long branches of the form
`s_getpc_b64 s[30:31]; ...; s_setpc_b64 s[30:31]`, but in a called function
s[30:31] holds the return address, so the function can no longer return (the
wave spins on its exit block). The tile functions are forced inline for this
reason; this test fails if such a sequence appears anywhere.
"""
import os
import re
import shutil
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "visar_amd", "libvisar_amd.so")
LLVM = "/opt/rocm/llvm/bin"


def _code_objects(tmp_path):
    fatbin = tmp_path / "fatbin.bin"
    subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", LIB, str(fatbin)], check=True)
    data = fatbin.read_bytes()
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    out = []
    for m in re.finditer(re.escape(magic), data):
        o = m.start()
        (num,) = struct.unpack_from("<Q", data, o + 24)
        p = o + 32
        for _ in range(num):
            eo, es, il = struct.unpack_from("<QQQ", data, p)
            p += 24
            tid = data[p:p + il].decode()
            p += il
            if "gfx950" in tid:
                f = tmp_path / f"co{len(out)}.co"
                f.write_bytes(data[o + eo:o + eo + es])
                out.append(f)
    return out


@pytest.mark.skipif(not os.path.exists(LIB) or not os.path.exists(f"{LLVM}/llvm-objdump")
                    or shutil.which("objcopy") is None, reason="needs the built library and ROCm LLVM tools")
def test_no_long_branch_through_return_address(tmp_path):
    cos = _code_objects(tmp_path)
    assert cos, "no gfx950 code object in the library"
    bad = []
    for co in cos:
        asm = subprocess.run([f"{LLVM}/llvm-objdump", "-d", str(co)], check=True, capture_output=True,
                             text=True).stdout.splitlines()
        func = "?"
        for i, line in enumerate(asm):
            if line.endswith(">:"):
                func = line
            if "s_setpc_b64 s[30:31]" in line and any("s_getpc_b64 s[30:31]" in x for x in asm[max(0, i - 6):i]):
                bad.append(func)
    assert not bad, f"long branch through the return address in {sorted(set(bad))[:5]}"
