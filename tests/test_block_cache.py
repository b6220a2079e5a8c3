"""The device block cache's host logic (edk_blockcache.h) under threads, built
with the host compiler and its address/undefined-behaviour and thread
sanitizers, over a budgeted host heap (tests/cache_stress.cpp): a block is
never handed out twice or released while held, cache releases after an
allocation failure free only cached blocks, a second release of a block is
refused, and releasing everything leaves nothing alive."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "edkit.jl_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_block_cache_under_threads(tmp_path, san):
    exe = str(tmp_path / "cache_stress")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", f"-fsanitize={san}", "-I", CSRC,
                    os.path.join(HERE, "cache_stress.cpp"), "-o", exe], check=True)
    iters = "4000" if san == "thread" else "20000"
    r = subprocess.run([exe, "8", iters, "48"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "errors 0" in r.stdout
