"""The decode and query kernels are built without scratch (register spills): a build of
the segment kernel that spilled decoded Elf wrongly in workgroups sharing a
CU with an earlier one (the GPU test test_decode_grid_shapes covers that
case).  Compiles the decode kernels for gfx950 with hipcc's resource-usage
remarks (no GPU needed) and checks VGPRs, scratch and occupancy."""
import re
import subprocess

import pytest

from ebi_amd import build as B

KERNELS = {  # mangled-name prefix -> minimum occupancy (waves per SIMD)
    "_ZN3ebi8k_decodeILj0EE": 8,  # Uncompressed tile jobs
    "_ZN3ebi8k_decodeILj1EE": 6,  # RLE
    "_ZN3ebi8k_decodeILj2EE": 5,  # BUFF
    "_ZN3ebi8k_decodeILj8EE": 5,  # DeltaSprintz
    "_ZN3ebi12k_decode_segE": 4,
    "_ZN3ebi12k_decode_mixE": 4,  # both job lists in one launch
    "_ZN3ebi11k_query_segILj3EE": 5,  # the segment decode with the fused queries: Gorilla
    "_ZN3ebi11k_query_segILj4EE": 5,  # Chimp
    "_ZN3ebi11k_query_segILj5EE": 4,  # Chimp128 (LDS: four blocks per CU)
    "_ZN3ebi11k_query_segILj6EE": 4,  # ElfOnChimp
    "_ZN3ebi11k_query_segILj7EE": 4,  # Elf
    "_ZN3ebi13k_group_tilesILi0E": 6,  # query spans on BUFF codes
    "_ZN3ebi13k_group_tilesILi1E": 6,  # query spans on DeltaSprintz codes
    "_ZN3ebi13k_group_tilesILi2E": 7,  # query spans on the bitmask / decoded records
    "_ZN3ebi13k_group_tilesILi3E": 6,  # query spans on RLE runs
    "_ZN3ebi14k_group_finishE": 4,  # the per-plan fold
}


@pytest.fixture(scope="module")
def usage(tmp_path_factory):
    src = B.PKG / "csrc" / "kernels" / "decode.hip"
    out = tmp_path_factory.mktemp("res") / "decode.o"
    cmd = [B.hipcc(), *B.COMMON, *B.DEVICE, "--offload-device-only", "-c", str(src), "-o", str(out),
           "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    found, name = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            found[name] = {}
            continue
        m = re.search(r"remark:\s+(VGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)", line)
        if m and name:
            found[name][m.group(1).split()[0]] = int(m.group(2))
    return found


@pytest.mark.parametrize("prefix", list(KERNELS))
def test_decode_kernel_needs_no_scratch(usage, prefix):
    names = [n for n in usage if n.startswith(prefix)]
    assert names, f"{prefix} not found in the remarks"
    for n in names:
        u = usage[n]
        assert u.get("ScratchSize") == 0, f"{n} spills: {u}"
        assert u.get("Occupancy", 0) >= KERNELS[prefix], f"{n}: {u}"
