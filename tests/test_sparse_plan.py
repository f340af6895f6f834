"""Host planning of the D^T g pieces (moquimc_amd/csrc/mqi_plan.h,
cut_pieces), compiled with g++ and checked against the properties that fix
the cut: the pieces tile every row in order (an empty row has one empty
piece), none is longer than the piece length or spans two rows, every
in-row gap starts a piece, and a piece that starts neither a row nor at a gap
starts exactly one piece length after the previous one."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cstdio>
#include <vector>
#include "mqi_plan.h"
int main() {
    long long rows, ng, plen;
    while (std::scanf("%lld %lld %lld", &rows, &ng, &plen) == 3) {
        std::vector<int64_t> ip(rows + 1), gaps(ng), pk;
        std::vector<int32_t> rp;
        for (auto& v : ip)
            if (std::scanf("%lld", (long long*) &v) != 1) return 1;
        for (auto& v : gaps)
            if (std::scanf("%lld", (long long*) &v) != 1) return 1;
        if (!cut_pieces(ip, gaps, plen, pk, rp)) { std::printf("overflow\n"); continue; }
        std::printf("%zu", pk.size());
        for (auto v : pk) std::printf(" %lld", (long long) v);
        for (auto v : rp) std::printf(" %d", v);
        std::printf("\n");
    }
    return 0;
}
"""


def _cases(rng):
    out = []
    for rows, plen in ((1, 4), (7, 5), (40, 16), (3, 1), (0, 8), (25, 1000)):
        lens = rng.integers(0, 60, rows)
        lens[rng.random(rows) < 0.2] = 0
        ip = np.r_[0, np.cumsum(lens)].astype(np.int64)
        nnz = int(ip[-1])
        gaps = np.unique(rng.integers(0, max(nnz, 1), rng.integers(0, 12))) if nnz else np.zeros(0, np.int64)
        gaps = np.unique(np.r_[gaps, ip[:-1][lens > 0][:2]])   # gap positions at row starts are ignored
        out.append((ip, gaps.astype(np.int64), plen))
    return out


def test_cut_pieces_tile_rows(tmp_path):
    src = tmp_path / "plan.cpp"
    src.write_text(DRIVER)
    exe = str(tmp_path / "plan")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "moquimc_amd", "csrc"),
                    str(src), "-o", exe], check=True)
    cases = _cases(np.random.default_rng(5))
    inp = "".join("%d %d %d %s %s\n" % (len(ip) - 1, len(g), plen, " ".join(map(str, ip)), " ".join(map(str, g)))
                  for ip, g, plen in cases)
    lines = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for (ip, gaps, plen), line in zip(cases, lines):
        v = [int(x) for x in line.split()]
        n = v[0]
        pk = np.array(v[1:1 + n])
        rp = np.array(v[1 + n:])
        rows = len(ip) - 1
        assert len(rp) == rows + 1 and pk[-1] == ip[-1] and rp[-1] == n - 1
        assert np.all(np.diff(pk) >= 0)
        assert np.all(np.diff(rp) >= 1)                          # every row has a piece
        starts = set(ip[:-1][np.diff(ip) > 0].tolist())
        inner_gaps = {int(g) for g in gaps
                      for r in range(rows) if ip[r] < g < ip[r + 1]}
        for r in range(rows):
            ps = pk[rp[r]:rp[r + 1] + 1] if rp[r + 1] > rp[r] else np.array([], np.int64)
            if ip[r + 1] == ip[r]:
                assert rp[r + 1] == rp[r] + 1 and pk[rp[r]] == ip[r]   # one empty piece
                continue
            assert ps[0] == ip[r] and ps[-1] == ip[r + 1]       # the row's pieces tile it
            assert np.all(np.diff(ps) <= plen)
            for a, b in zip(ps[:-1], ps[1:]):
                assert not any(a < g < b for g in inner_gaps)   # a gap inside a piece
        empty = set(rp[:-1][np.diff(ip) == 0].tolist())
        for k in range(n - 1):   # the last entry is the sentinel
            s = int(pk[k])
            if s in starts or s in inner_gaps or k in empty:
                continue
            assert s - pk[k - 1] == plen


CHECK_DRIVER = r"""
#include <cstdio>
#include <vector>
#include "mqi_plan.h"
int main() {
    long long rows, cols, nnz;
    while (std::scanf("%lld %lld %lld", &rows, &cols, &nnz) == 3) {
        std::vector<int64_t> ip(rows + 1);
        std::vector<int32_t> idx(nnz);
        for (auto& v : ip)
            if (std::scanf("%lld", (long long*) &v) != 1) return 1;
        for (auto& v : idx)
            if (std::scanf("%d", &v) != 1) return 1;
        int64_t bad = -1;
        const int e = check_csr(rows, cols, ip.data(), idx.data(), &bad);
        std::printf("%d %lld\n", e, (long long) bad);
    }
    return 0;
}
"""


def test_check_csr_codes(tmp_path):
    """The validation of a caller's CSR (check_csr): each kind of defect is
    reported with its code and the first row or entry that has it."""
    src = tmp_path / "check.cpp"
    src.write_text(CHECK_DRIVER)
    exe = str(tmp_path / "check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "moquimc_amd", "csrc"),
                    str(src), "-o", exe], check=True)
    cases = [  # rows, cols, indptr, indices, (code, bad)
        (2, 10, [0, 2, 4], [1, 5, 0, 9], (0, -1)),
        (3, 10, [0, 0, 0, 0], [], (0, -1)),
        (2, 10, [1, 2, 4], [1, 5, 0, 9], (1, 0)),
        (2, 10, [0, 3, 2], [1, 5, 7], (1, 1)),
        (2, 10, [0, 2, 4], [1, 5, 0, 10], (2, 3)),
        (2, 10, [0, 2, 4], [-1, 5, 0, 9], (2, 0)),
        (2, 10, [0, 2, 4], [5, 5, 0, 9], (3, 1)),
        (2, 10, [0, 2, 4], [1, 5, 9, 0], (3, 3)),
        (1, 3, [0, 3], [0, 1, 2], (0, -1)),
    ]
    inp = "".join("%d %d %d %s %s\n" % (r, c, len(ix), " ".join(map(str, ip)), " ".join(map(str, ix)))
                  for r, c, ip, ix, _ in cases)
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split()
    got = [(int(out[2 * k]), int(out[2 * k + 1])) for k in range(len(cases))]
    assert got == [want for *_, want in cases]
