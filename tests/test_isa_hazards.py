"""CPU checks of the built gfx950 code.  First: no buffer store of more than
64 bits uses an SGPR soffset.

Such a store reads its data VGPRs after issue; the compiler inserts the
wait state before a VALU write to those VGPRs only when soffset is not an
SGPR, so with an SGPR soffset the next instruction may overwrite the data
before it is read.  On the GPU this showed up as corrupted imaginary parts of
FP64 Dslash output whenever other kernels shared the compute units (solver
results then depended on what else ran).  The Dslash output stores use one
buffer resource per spin plane with soffset 0 to stay inside the compiler's
check; this test disassembles every code object in the library and fails if
the pattern comes back.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "thirring-rhmc_amd", "libthirring_hip.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"

STORE = re.compile(r"^\s*buffer_store_dword(x3|x4)\s+v\[(\d+):(\d+)\],\s*v\d+,\s*s\[\d+:\d+\],\s*(\S+)")
VALU_DST = re.compile(r"^\s*v_\S+\s+v\[?(\d+)(?::(\d+))?\]?")


def _tools():
    objcopy = shutil.which("objcopy")
    bundler = os.path.join(LLVM, "clang-offload-bundler")
    objdump = os.path.join(LLVM, "llvm-objdump")
    if not (objcopy and os.path.exists(bundler) and os.path.exists(objdump)):
        pytest.skip("binutils objcopy / ROCm LLVM tools not available")
    return objcopy, bundler, objdump


def _disassemble(tmp_path):
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    objcopy, bundler, objdump = _tools()
    fat = tmp_path / "fatbin.bin"
    # -O binary with an output file: objcopy never rewrites the library itself
    subprocess.run([objcopy, "-O", "binary", "--only-section=.hip_fatbin", LIB, str(fat)], check=True)
    data = fat.read_bytes()
    # the linked library holds one offload bundle per translation unit
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    assert starts, "no offload bundle in the library"
    texts = []
    for k, s in enumerate(starts):
        e = starts[k + 1] if k + 1 < len(starts) else len(data)
        part = tmp_path / f"b{k}.bin"
        part.write_bytes(data[s:e])
        co = tmp_path / f"b{k}.o"
        r = subprocess.run([bundler, "--unbundle", "--type=o", f"--targets={TARGET}",
                            f"--input={part}", f"--output={co}"], capture_output=True)
        if r.returncode != 0 or not co.exists() or co.stat().st_size == 0:
            continue
        out = subprocess.run([objdump, "-d", "--mcpu=gfx950", str(co)], check=True,
                             capture_output=True, text=True).stdout
        texts.append(out)
    assert texts, "no gfx950 code object found"
    return texts


def _instructions(text):
    for line in text.splitlines():
        code = line.split("//")[0].rstrip()
        if code.startswith(("\t", " ")) and code.strip():
            yield code


def test_no_wide_buffer_store_with_sgpr_soffset(tmp_path):
    texts = _disassemble(tmp_path)
    wide, hazards, total = [], [], 0
    for text in texts:
        ins = list(_instructions(text))
        for i, l in enumerate(ins):
            m = STORE.match(l)
            if not m:
                continue
            total += 1
            lo, hi, soff = int(m.group(2)), int(m.group(3)), m.group(4)
            if soff.startswith("s"):
                wide.append(l.strip())
                v = VALU_DST.match(ins[i + 1]) if i + 1 < len(ins) else None
                if v and int(v.group(1)) <= hi and int(v.group(2) or v.group(1)) >= lo:
                    hazards.append((l.strip(), ins[i + 1].strip()))
    assert total > 0, "expected the FP64 Dslash output stores in the code"
    assert not hazards, hazards[:5]
    assert not wide, wide[:5]


def test_f16_dslash_reads_sites_whole(tmp_path):
    """FP16 spinors keep the 4 spins of a site in 16 bytes: every FP16 Dslash
    kernel must read the centre, the two t neighbours and the four x/y
    neighbours with one 128-bit load each.  (A bit_cast of an ext-vector
    element lvalue once compiled to a 32-bit load of spin 0 only.)"""
    label = re.compile(r"^[0-9a-f]+ <(.*)>:$")
    counts, cur = {}, None
    for text in _disassemble(tmp_path):
        for line in text.splitlines():
            m = label.match(line)
            if m:
                cur = m.group(1)
                counts[cur] = 0
            elif cur is not None and "buffer_load_dwordx4" in line:
                counts[cur] += 1
    f16 = {k: v for k, v in counts.items() if "8k_dslashILi2E" in k}
    assert f16, "no FP16 Dslash kernels found"
    short = {k: v for k, v in f16.items() if v < 7}
    assert not short, list(short.items())[:5]


def test_f32_dslash_reads_sites_whole(tmp_path):
    """FP32 spinors keep the 4 spins of a site in 32 bytes: every FP32 Dslash
    kernel reads the centre, the two t neighbours and the four x/y neighbours
    with two 128-bit loads each (14 at least), and stores its output sites
    with 128-bit stores."""
    label = re.compile(r"^[0-9a-f]+ <(.*)>:$")
    loads, stores, cur = {}, {}, None
    for text in _disassemble(tmp_path):
        for line in text.splitlines():
            m = label.match(line)
            if m:
                cur = m.group(1)
                loads[cur] = stores[cur] = 0
            elif cur is not None and "buffer_load_dwordx4" in line:
                loads[cur] += 1
            elif cur is not None and "buffer_store_dwordx4" in line:
                stores[cur] += 1
    f32 = [k for k in loads if "8k_dslashILi1E" in k]
    assert f32, "no FP32 Dslash kernels found"
    short = [(k, loads[k]) for k in f32 if loads[k] < 14]
    assert not short, short[:5]
    assert all(stores[k] >= 2 for k in f32), [(k, stores[k]) for k in f32][:5]
