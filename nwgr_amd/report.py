"""Result reporting in the reference benchmark's formats.

``life_timer`` prints the lifetime of a block as nw::util::life_timer does.

``Times`` collects per-trial seconds keyed by (file, version, threads) and
prints the min/avg/max table of ``Times::print`` (bench/common.hpp:175-231).
``log_rows`` appends fixed-width result rows laid out as ``Log``
(bench/Log.hpp:102-160): one row per trial with a 24-letter run id, library,
git branch and revision, compiler columns, date, host, benchmark, "v<id>",
threads, the graph file's stem up to its first '_', then the values, so GPU
results merge into tooling that reads the reference's logs.  The library
column reads NWGR_AMD, the compiler columns describe the HIP build.
"""
from __future__ import annotations

import datetime
import os
import random
import socket
import string
import subprocess
import sys
from typing import IO

LIBRARY = "NWGR_AMD"
CXX, CXX_ID, CXX_VER, BUILD, TBBMALLOC = "hipcc", "Clang", "rocm-7.2", "Release", "OFF"


class life_timer:
    """Prints "# [ name ]: <ms> ms" when the block ends, whole milliseconds, as
    nw::util::life_timer does (util/timer.hpp:60-70)."""

    def __init__(self, name: str, out: IO[str] | None = None):
        self.name, self.out = name, out

    def __enter__(self):
        import time
        self._t0 = time.perf_counter()
        return self

    def __exit__(self, *exc):
        import time
        if self.name:
            ms = int((time.perf_counter() - self._t0) * 1e3)
            (self.out or sys.stdout).write(f"# [ {self.name} ]: {ms} ms\n")
        return False


class Times:
    """Trial times by (file, version, threads), as bench/common.hpp:175-250."""

    def __init__(self):
        self._samples: dict[tuple[str, int, int], list[tuple]] = {}

    def append(self, file: str, version: int, threads: int, seconds: float, *extra) -> None:
        self._samples.setdefault((file, int(version), int(threads)), []).append((float(seconds), *extra))

    def items(self):
        return sorted(self._samples.items())

    def print(self, out: IO[str] = sys.stdout) -> None:
        n = max([4] + [len(k[0]) for k in self._samples])
        out.write(f"{'File':<{n + 2}}{'Version':<10}{'Threads':<10}{'Min':<20}{'Avg':<20}{'Max':<20}\n")
        for (file, version, threads), samples in self.items():
            t = [s[0] for s in samples]
            out.write(f"{file:<{n + 2}}{version:<10}{threads:<10}{min(t):<20.6f}{sum(t) / len(t):<20.6f}"
                      f"{max(t):<20.6f}\n")


def _git(args: list[str]) -> str:
    try:
        r = subprocess.run(["git", *args], capture_output=True, text=True, timeout=10,
                           cwd=os.path.dirname(os.path.abspath(__file__)))
        return r.stdout.strip() if r.returncode == 0 and r.stdout.strip() else "-"
    except (OSError, subprocess.SubprocessError):
        return "-"


def _stem(path: str) -> str:
    stem = os.path.splitext(os.path.basename(path))[0]
    return stem.split("_", 1)[0]


def _widths() -> dict[str, int]:
    # the reference sizes these columns from its compiler strings (with the
    # terminating NUL) or the header word, plus one
    def w(value, head):
        return max(len(value) + 1, len(head) + 1) + 1
    return {"uuid": 26, "cxx": w(CXX, "CXX"), "cxx_id": w(CXX_ID, "CXX_ID"), "cxx_ver": w(CXX_VER, "CXX_VER"),
            "build": w(BUILD, "Build"), "tbb": w(TBBMALLOC, "TBBMALLOC"), "host": 20}


def log_rows(path: str, benchmark: str, times: Times, header: bool, *headers: str, run_id: str | None = None) -> None:
    """Append Log.hpp-style rows for every trial in `times` to `path` ("-" =
    stdout), with a header line first when `header`."""
    W = _widths()
    uid = run_id or "".join(random.choice(string.ascii_lowercase) for _ in range(24))
    host = socket.gethostname()[:19]
    date = datetime.date.today().isoformat()
    branch = _git(["rev-parse", "--abbrev-ref", "HEAD"])
    rev = _git(["log", "--pretty=format:%h", "-n", "1"])
    lines = []
    if header:
        cols = [("UUID", W["uuid"]), ("Library", 10), ("Branch", 10), ("Revision", 10), ("CXX", W["cxx"]),
                ("CXX_ID", W["cxx_id"]), ("CXX_VER", W["cxx_ver"]), ("Build", W["build"]),
                ("TBBMALLOC", W["tbb"]), ("Date", 15), ("Host", W["host"]), ("Benchmark", 10), ("Version", 10),
                ("Threads", 10), ("Graph", 32)] + [(h, 20) for h in headers]
        lines.append("".join(f"{c:<{w}}" for c, w in cols))
    for (file, version, threads), samples in times.items():
        for sample in samples:
            head = [(uid, W["uuid"]), (LIBRARY, 10), (branch, 10), (rev, 10), (CXX, W["cxx"]), (CXX_ID, W["cxx_id"]),
                    (CXX_VER, W["cxx_ver"]), (BUILD, W["build"]), (TBBMALLOC, W["tbb"]), (date, 15),
                    (host, W["host"]), (benchmark, 10), (f"v{version}", 10), (str(threads), 10), (_stem(file), 32)]
            vals = "".join(f"{(f'{v:.8e}' if isinstance(v, float) else str(v)):<20}" for v in sample)
            lines.append("".join(f"{c:<{w}}" for c, w in head) + vals)
    text = "\n".join(lines) + "\n"
    if path == "-":
        sys.stdout.write(text)
    else:
        if os.path.dirname(path):
            os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "a") as f:
            f.write(text)
