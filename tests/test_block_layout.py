"""Block-strategy layout invariants, checked on the host (g++ build of the layout
builder plus tests/cpp/check_block_layout.cpp; no GPU)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("blk") / "check_block_layout"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{ROOT / 'rlmusclev2_amd' / 'csrc'}",
                    str(ROOT / "tests" / "cpp" / "check_block_layout.cpp"),
                    str(ROOT / "rlmusclev2_amd" / "csrc" / "xpbd_layout.cpp"), "-o", str(exe)], check=True)
    return exe


@pytest.mark.parametrize("n,tb", [(7, 64), (12, 256), (20, 1024), (30, 4096)])
def test_block_layout_sums_match_flat(checker, n, tb):
    r = subprocess.run([str(checker), str(n), str(tb)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad 0" in r.stdout


def test_attach_records_and_duplicate_rejection(tmp_path):
    """ATTACH records follow their vertex through the renumbering; a second
    ATTACH on one vertex is rejected with XPBD_EUNSUPPORTED"""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "check_attach_layout"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{ROOT / 'rlmusclev2_amd' / 'csrc'}",
                    str(ROOT / "tests" / "cpp" / "check_attach_layout.cpp"),
                    str(ROOT / "rlmusclev2_amd" / "csrc" / "xpbd_layout.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad 0" in r.stdout


@pytest.fixture(scope="module")
def parts_checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("parts") / "check_parts_layout"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{ROOT / 'rlmusclev2_amd' / 'csrc'}",
                    str(ROOT / "tests" / "cpp" / "check_parts_layout.cpp"),
                    str(ROOT / "rlmusclev2_amd" / "csrc" / "xpbd_layout.cpp"), "-o", str(exe)], check=True)
    return exe


@pytest.mark.parametrize("n,tb,r,target", [(7, 64, 1, 0), (12, 256, 3, 0), (20, 256, 4, 0), (20, 128, 2, 0),
                                           (30, 256, 3, 0), (12, 64, 64, 0), (20, 128, 1, 7), (30, 256, 1, 23),
                                           (12, 64, 1, 1000)])
def test_resident_parts_equal_block_sums(parts_checker, n, tb, r, target):
    """parts of r blocks (or a target count of parts, blocks spread evenly):
    cover, write-back once, symmetric neighbour lists, corner ids, and FP32 sums
    bit-identical to the block strategy's"""
    res = subprocess.run([str(parts_checker), str(n), str(tb), str(r), str(target)], capture_output=True,
                         text=True)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "bad 0" in res.stdout
