"""Host-run native checks of device helpers (tests/native/*.hip), built by
ebi_amd.build into ebi_amd/_build/ (built here if missing or stale)."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRCS = sorted((ROOT / "tests" / "native").glob("*.hip"))
HEADERS = list((ROOT / "ebi_amd" / "csrc").rglob("*.hpp"))


def exe_for(src: Path) -> Path:
    exe = ROOT / "ebi_amd" / "_build" / src.stem
    newest = max(p.stat().st_mtime for p in [src, *HEADERS])
    if not exe.exists() or exe.stat().st_mtime < newest:
        from ebi_amd import build as b

        b.build_native_tests()
    return exe


@pytest.mark.parametrize("src", SRCS, ids=[s.stem for s in SRCS])
def test_native(src):
    r = subprocess.run([str(exe_for(src))], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
