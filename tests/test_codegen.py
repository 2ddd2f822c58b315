"""Checks on the gfx950 machine code of the built library (CPU only: the
code objects are taken out of lib/libpsdns_gpu.so and disassembled with the
ROCm LLVM tools).

A vector-memory store of more than 8 bytes reads its data registers after it
issues.  The compiler inserts the wait state that protects them from the next
VALU write only when the buffer store's scalar offset is not a register; on
gfx950 an FP64 buffer_store_dwordx4 with a register offset stored the value
the following v_mov wrote into one of its data registers (run-to-run
differences of the FP64 step).  The kernels therefore pass the offset of
16-byte stores in the vector offset."""
import pathlib
import re
import shutil
import subprocess

import pytest

REPO = pathlib.Path(__file__).resolve().parent.parent
LIB = REPO / "nasa-bluewater-psdns-benchmark-minsky_amd" / "lib" / "libpsdns_gpu.so"
LLVM = pathlib.Path("/opt/rocm/llvm/bin")
TOOLS = [LLVM / t for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-objdump")]
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"

WIDE_SGPR_SOFFSET = re.compile(r"buffer_store_dwordx[34] v\[\d+:\d+\], v\d+, s\[\d+:\d+\], s\d+")
WIDE_STORE = re.compile(r"^\s+(?:buffer_store_dwordx[34] v\[(\d+):(\d+)\]|global_store_dwordx[34] v(?:\d+|\[\d+:\d+\]), "
                        r"v\[(\d+):(\d+)\])")
VALU_DST = re.compile(r"^\s+v_\w+ (?:v(\d+)|v\[(\d+):(\d+)\])")


@pytest.fixture(scope="module")
def disassembly(tmp_path_factory):
    if not LIB.exists():
        pytest.skip("library not built")
    if not all(t.exists() for t in TOOLS) or shutil.which("true") is None:
        pytest.skip("ROCm LLVM tools not found")
    d = tmp_path_factory.mktemp("codeobj")
    fat = d / "fatbin.bin"
    subprocess.run([str(TOOLS[0]), f"--dump-section=.hip_fatbin={fat}", str(LIB), str(d / "stripped.so")], check=True)
    data = fat.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    assert starts, "no offload bundle in .hip_fatbin"
    text = []
    for k, o in enumerate(starts):
        b = d / f"b{k}.bin"
        b.write_bytes(data[o:starts[k + 1] if k + 1 < len(starts) else len(data)])
        co = d / f"b{k}.co"
        subprocess.run([str(TOOLS[1]), "--unbundle", "--type=o", f"--input={b}",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True)
        r = subprocess.run([str(TOOLS[2]), "-d", "--mcpu=gfx950", str(co)], check=True, capture_output=True, text=True)
        text.append(r.stdout)
    return "\n".join(text)


def test_wide_buffer_stores_take_no_register_offset(disassembly):
    assert "buffer_store_dwordx4" in disassembly
    bad = WIDE_SGPR_SOFFSET.findall(disassembly)
    assert not bad, f"{len(bad)} wide buffer stores with a register scalar offset, e.g. {bad[0]}"


def test_no_wide_store_data_overwritten_by_next_instruction(disassembly):
    lines = [l for l in disassembly.splitlines() if l.startswith("\t")]
    bad = []
    for i, l in enumerate(lines[:-1]):
        m = WIDE_STORE.match(l)
        if not m:
            continue
        lo, hi = (int(m.group(1)), int(m.group(2))) if m.group(1) else (int(m.group(3)), int(m.group(4)))
        n = VALU_DST.match(lines[i + 1])
        if n:
            a = int(n.group(1) or n.group(2))
            b = int(n.group(1) or n.group(3))
            if a <= hi and b >= lo:
                bad.append((l.strip(), lines[i + 1].strip()))
    assert not bad, f"{len(bad)} wide stores whose data the next instruction writes, e.g. {bad[0]}"
