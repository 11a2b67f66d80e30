"""Machine-code checks on the engine as built for gfx950 (CPU only).

On CDNA a FLAT (generic) instruction picks LDS, scratch or global memory from
the high bits of its address register alone; the immediate offset is added
afterwards.  A generic pointer into LDS whose register value sits below the
LDS aperture (a negative intermediate index with a positive immediate folded
back in) is then decoded as an address outside every aperture and the wave
takes a memory violation.  The engine therefore never addresses LDS or
scratch through generic pointers: every kernel reaches LDS with ds_*
instructions only, and no kernel has a private (scratch) segment.  A kernel
that builds a generic LDS pointer has to materialise the aperture base
(src_shared_base), which is what these tests look for in the disassembly of
libluda.so's own code objects.
"""
import os
import re
import shutil
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "luda_amd", "libluda.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _tool(name):
    p = os.path.join(LLVM, name)
    return p if os.path.exists(p) else shutil.which(name)


def _code_objects(lib):
    """gfx950 ELF code objects of every translation unit in lib's .hip_fatbin."""
    objcopy = _tool("llvm-objcopy")
    tmp = lib + ".fatbin.tmp"
    subprocess.check_call([objcopy, "--dump-section", f".hip_fatbin={tmp}", lib, os.devnull])
    try:
        data = open(tmp, "rb").read()
    finally:
        os.unlink(tmp)
    out = []
    pos = data.find(MAGIC)
    while pos >= 0:
        n = struct.unpack_from("<Q", data, pos + len(MAGIC))[0]
        h = pos + len(MAGIC) + 8
        for _ in range(n):
            off, size, idlen = struct.unpack_from("<QQQ", data, h)
            ident = data[h + 24:h + 24 + idlen].decode()
            h += 24 + idlen
            if "gfx950" in ident:
                out.append(data[pos + off:pos + off + size])
        pos = data.find(MAGIC, pos + len(MAGIC))
    return out


def _disassemble(blobs, tmpdir):
    objdump = _tool("llvm-objdump")
    texts = []
    for i, blob in enumerate(blobs):
        p = os.path.join(tmpdir, f"co{i}.o")
        with open(p, "wb") as f:
            f.write(blob)
        texts.append(subprocess.run([objdump, "-d", "--mcpu=gfx950", p], capture_output=True, text=True,
                                    check=True).stdout)
    return texts


def _kernels(text):
    """kernel symbol -> its instruction lines."""
    ks, cur = {}, None
    for line in text.splitlines():
        m = re.match(r"^[0-9a-f]+ <(.+)>:$", line)
        if m:
            cur = m.group(1)
            ks[cur] = []
        elif cur and line.startswith("\t"):
            ks[cur].append(line.strip())
    return ks


@pytest.fixture(scope="module")
def kernels(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built; run __graft_entry__.build()")
    if not _tool("llvm-objdump") or not _tool("llvm-objcopy"):
        pytest.skip("llvm-objdump / llvm-objcopy not found")
    blobs = _code_objects(LIB)
    assert blobs, "no gfx950 code object in libluda.so"
    ks = {}
    for text in _disassemble(blobs, str(tmp_path_factory.mktemp("isa"))):
        ks.update(_kernels(text))
    return ks


def test_every_hot_kernel_is_present(kernels):
    names = " ".join(kernels)
    for k in ("k_decode", "k_merge", "k_drop", "k_gather", "k_write_blocks", "k_write_filters", "k_write_tail",
              "k_orbit_chunk", "k_entry_sizes", "k_bnext"):
        assert k in names, k


def test_no_generic_pointer_reaches_lds(kernels):
    bad = {k: [l for l in ins if "src_shared_base" in l] for k, ins in kernels.items()}
    bad = {k: v for k, v in bad.items() if v}
    assert not bad, f"kernels forming generic LDS pointers: {sorted(bad)}"


def test_no_scratch(kernels):
    bad = {k: [l for l in ins if "src_private_base" in l or l.startswith("scratch_")] for k, ins in kernels.items()}
    bad = {k: v for k, v in bad.items() if v}
    assert not bad, f"kernels using scratch memory: {sorted(bad)}"
