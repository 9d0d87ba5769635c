"""Result reporting in the reference benchmark's formats (nwgr_amd.report):
the Times table of bench/common.hpp:205-231 and the Log rows of
bench/Log.hpp:102-160, checked column by column."""
import io

from nwgr_amd.report import LIBRARY, Times, life_timer, log_rows


def test_times_table_layout():
    t = Times()
    t.append("graphs/rmat_s16.mtx", 11, 8, 0.5)
    t.append("graphs/rmat_s16.mtx", 11, 8, 1.5)
    t.append("graphs/rmat_s16.mtx", 13, 8, 0.25)
    out = io.StringIO()
    t.print(out)
    lines = out.getvalue().splitlines()
    n = len("graphs/rmat_s16.mtx")
    assert lines[0] == f"{'File':<{n + 2}}{'Version':<10}{'Threads':<10}{'Min':<20}{'Avg':<20}{'Max':<20}"
    assert lines[1] == (f"{'graphs/rmat_s16.mtx':<{n + 2}}{'11':<10}{'8':<10}{'0.500000':<20}{'1.000000':<20}"
                        f"{'1.500000':<20}")
    assert lines[2].split() == ["graphs/rmat_s16.mtx", "13", "8", "0.250000", "0.250000", "0.250000"]


def test_log_rows_columns(tmp_path):
    t = Times()
    t.append("/data/road_usa_sym.el", 11, 16, 0.125, 1e-4)
    t.append("/data/road_usa_sym.el", 11, 16, 0.25, 1e-4)
    path = tmp_path / "logs" / "pr.log"
    log_rows(str(path), "pr", t, True, "Time(s)", "Tolerance", run_id="abcdefghijklmnopqrstuvwx")
    log_rows(str(path), "pr", t, False, run_id="abcdefghijklmnopqrstuvwx")  # appends
    lines = path.read_text().splitlines()
    assert len(lines) == 1 + 2 + 2
    head = lines[0].split()
    assert head[:15] == ["UUID", "Library", "Branch", "Revision", "CXX", "CXX_ID", "CXX_VER", "Build", "TBBMALLOC",
                         "Date", "Host", "Benchmark", "Version", "Threads", "Graph"]
    assert head[15:] == ["Time(s)", "Tolerance"]
    row = lines[1]
    assert row[:26] == "abcdefghijklmnopqrstuvwx  "
    assert row[26:36] == f"{LIBRARY:<10}"
    f = row.split()
    assert f[0] == "abcdefghijklmnopqrstuvwx" and f[11:15] == ["pr", "v11", "16", "road"]
    assert f[15:] == ["1.25000000e-01", "1.00000000e-04"]
    assert lines[2].split()[15] == "2.50000000e-01"
    # the header and every row line up: the Graph column starts at one offset
    g0 = lines[0].index("Graph")
    assert all(ln[g0:g0 + 4] == "road" for ln in lines[1:])
    out = io.StringIO()
    import sys
    old = sys.stdout
    try:
        sys.stdout = out
        log_rows("-", "pr", t, False, run_id="x" * 24)
    finally:
        sys.stdout = old
    assert len(out.getvalue().splitlines()) == 2


def test_life_timer_line():
    out = io.StringIO()
    with life_timer("build adjacency", out):
        pass
    line = out.getvalue()
    assert line.startswith("# [ build adjacency ]: ") and line.endswith(" ms\n")
    assert int(line.split(": ")[1].split()[0]) >= 0
    out = io.StringIO()
    with life_timer("", out):
        pass
    assert out.getvalue() == ""
