"""pr.exe on the GPU (nwgr_amd.pr): the reference driver's options, its
Times table, the --verify print and --log rows (needs the GPU)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import nwgr_amd
import oracle
from tests.conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu


def test_pr_driver_versions_verify_and_log(tmp_path):
    mtx = os.path.join(GOLDEN, "mtx", "karate.mtx")
    log = tmp_path / "pr.log"
    cmd = [sys.executable, "-m", "nwgr_amd.pr", "-f", mtx, "-t", "1e-7", "-i", "100", "--version", "11",
           "--version", "13", "-n", "2", "-v", "--log", str(log), "--log-header", "4"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert out.returncode == 0, out.stderr
    text = out.stdout
    # two verification prints (one per version), ten "id: rank" lines each
    blocks = text.split("Verifying\n")[1:]
    assert len(blocks) == 2
    el = nwgr_amd.read_mm(mtx)
    g = nwgr_amd.adjacency_from_edge_list(el)
    deg = nwgr_amd.build_degrees(g)
    r11, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-7, 100)
    r13, _, _ = oracle.delta(g.offsets, g.indices, deg, 0.85, float(np.float32(1e-7)), 100, np.float32)
    for block, ref in zip(blocks, (r11, r13)):
        lines = [ln for ln in block.splitlines() if ":" in ln][:10]
        ids = [int(ln.split(":")[0]) for ln in lines]
        assert ids == oracle.top_k(ref, 10).tolist()
    # Times table: a header and one row per version, threads label 4
    table = [ln.split() for ln in text.splitlines() if ln.startswith(mtx) or ln.startswith("File")]
    assert table[0] == ["File", "Version", "Threads", "Min", "Avg", "Max"]
    assert [row[1:3] for row in table[1:]] == [["11", "4"], ["13", "4"]]
    # log: header + 2 versions x 2 trials
    rows = log.read_text().splitlines()
    assert len(rows) == 5 and rows[0].split()[-2:] == ["Time(s)", "Tolerance"]
    assert sorted(r.split()[12] for r in rows[1:]) == ["v11", "v11", "v13", "v13"]
    assert all(r.split()[14] == "karate" for r in rows[1:])
    # life_timer lines of the phases (util/timer.hpp:60-70); the device upload
    # is inside the first timed call, as the reference's setup is in page_rank
    assert not any("device setup" in ln for ln in text.splitlines())
    for phase in ("read mm", "build adjacency", "degrees"):
        assert any(ln.startswith(f"# [ {phase} ]: ") and ln.endswith(" ms") for ln in text.splitlines()), phase


def test_pr_driver_rejects_unknown_version():
    mtx = os.path.join(GOLDEN, "mtx", "karate.mtx")
    out = subprocess.run([sys.executable, "-m", "nwgr_amd.pr", "-f", mtx, "--version", "7"], capture_output=True,
                         text=True, timeout=120, cwd=ROOT)
    assert out.returncode != 0 and "Unknown version id 7" in out.stderr
