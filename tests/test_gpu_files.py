"""Query files through the per-query traversal on the GPU, against the oracle: the handler's
query-file entry point (QueryHandler.cu:168-193 performQuery(path, ...)), the CSV row the CLI
appends (QueryEngine.cu:254-305; main.cu:152-161) and the reference-style C++ caller."""
import csv
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import tblaeq

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "t-blaeq_amd", "tblaeq-cli")
CALLER = os.path.join(ROOT, "tests", "cpp", "reference_caller")
GEN = dict(N=20000, D=4, vmin=1.0, vmax=100.0, is_int=False, ratios=(20, 10), seed=7, sigma_div=3.0)
HEADER = ["Dataset", "Size", "Dim", "Query Type", "Query Parameter", "Query Count", "Total Time (ms)", "Avg Time (ms)",
          "Memory Policy", "Median Log Volume", "Avg Range Volume", "Avg Fine Mesh"]


def _rows(a):
    return "\n".join(" ".join(repr(float(v)) for v in row) for row in a) + "\n"


def _boxes(data, n, seed, half):
    rng = np.random.default_rng(seed)
    c = data[rng.integers(data.shape[0], size=n)]
    h = half * rng.uniform(0.5, 1.0, (n, 1))
    return c - h, c + h


@pytest.fixture(scope="module")
def pair():
    g = GEN
    h = tblaeq.Handler.random(g["N"], g["D"], g["vmin"], g["vmax"], g["is_int"], g["ratios"], g["seed"],
                              g["sigma_div"])
    o = O.Index.generate(g["N"], g["D"], g["vmin"], g["vmax"], g["is_int"], g["ratios"], g["seed"], g["sigma_div"])
    h.prepare("L0")
    return h, o


def _write_ranges(path, lo, hi):
    with open(path, "w") as f:
        f.write("%d %d\n" % lo.shape + _rows(np.concatenate([lo, hi], 1)))
    return str(path)


def test_range_file_matches_oracle(pair, tmp_path):
    h, o = pair
    lo, hi = _boxes(h.data(), 7, 11, 12.0)
    p = _write_ranges(tmp_path / "boxes_20-4.txt", lo, hi)
    r = h.query_file(p, "range", max_queries=5, save_fine_mesh=True)
    assert r.nq == 5
    for i in range(5):
        ro = o.range(lo[i], hi[i])
        ids, vals = r.fine_mesh(i)
        oids, ovals = ro.grid()
        assert np.array_equal(ids, oids) and np.array_equal(vals.view(np.uint64), ovals.view(np.uint64)), i
        assert np.array_equal(r.levels(i)[0], ro.levels()[1]), i
    # a query file runs the traversal only (QueryEngine.cu:62-143 keeps no refined results)
    with pytest.raises(tblaeq.TblaeqError, match="without the exact refine"):
        r.refined(0)
    r.close()


def test_knn_file_matches_oracle(pair, tmp_path):
    h, o = pair
    q = h.data()[np.random.default_rng(12).integers(GEN["N"], size=4)] + 0.25
    p = str(tmp_path / "pts_k.txt")
    with open(p, "w") as f:
        f.write("4 4\n" + _rows(q))
    r = h.query_file(p, "knn", K=9, save_fine_mesh=True)
    assert r.nq == 4
    for i in range(4):
        ids, vals = r.fine_mesh(i)
        oids, ovals = o.knn(q[i], 9).grid()
        assert np.array_equal(ids, oids) and np.array_equal(vals.view(np.uint64), ovals.view(np.uint64)), i
    r.close()
    # without save_fine_mesh only the sizes are kept
    r = h.query_file(p, "knn", K=9, max_queries=2)
    assert r.nq == 2 and r.info(1)["fine"] == o.knn(q[1], 9).grid()[0].size
    with pytest.raises(tblaeq.TblaeqError, match="save_fine_mesh off"):
        r.fine_mesh(0)
    r.close()


def test_query_file_dimension_mismatch(pair, tmp_path):
    h, _ = pair
    p = str(tmp_path / "pts_k.txt")
    with open(p, "w") as f:
        f.write("1 3\n1 2 3\n")
    with pytest.raises(tblaeq.TblaeqError, match="dimension"):
        h.query_file(p, "knn", K=3)


def test_cli_query_appends_csv_rows(tmp_path):
    idx = str(tmp_path / "synthetic")
    g = GEN
    r = subprocess.run([CLI, "--gen-index", "--gen-N", str(g["N"]), "--gen-D", str(g["D"]), "--gen-seed",
                        str(g["seed"]), "--height", "3", "--ratios", "20,10", "-i", idx], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    o = O.Index.generate(g["N"], g["D"], g["vmin"], g["vmax"], g["is_int"], g["ratios"], g["seed"], g["sigma_div"])
    lo, hi = _boxes(o.data(), 6, 13, 10.0)
    qf = _write_ranges(tmp_path / "boxes_10-4.txt", lo, hi)
    out = str(tmp_path / "result.csv")
    for policy in ("L0", "L2"):
        r = subprocess.run([CLI, "--test-query", "-i", idx, "-f", qf, "-r", out, "-q", "6", "--force-policy",
                            policy], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
    with open(out) as f:
        rows = list(csv.reader(f))
    assert rows[0] == HEADER and len(rows) == 3
    fine = np.mean([o.range(lo[i], hi[i]).grid()[0].size for i in range(6)])
    # median log volume of the 6 boxes, and e^median as m.mmE+x (QueryEngine.cu:35-60)
    med = np.median(np.log(hi - lo).sum(1))
    ex = int(np.floor(med / np.log(10.0)))
    vol = "%.2fE%s%d" % (10.0 ** (med / np.log(10.0) - ex), "+" if ex >= 0 else "", ex)
    for row, policy in zip(rows[1:], ("L0", "L2")):
        assert row[0] == "synthetic" and row[1] == str(g["N"]) and row[2] == str(g["D"]), row
        assert row[3] == "RANGE" and row[4] == "10%^4=0.01%", row
        assert row[5] == "4", row  # the fastest and slowest of 6 queries are dropped
        assert row[8] == policy + "-" + policy, row
        assert row[11] == "%.1f" % fine, row
        assert row[9] == "%.3f" % med and row[10] == vol, row
        assert float(row[6]) >= 0 and float(row[7]) >= 0


def test_reference_style_cpp_caller_runs(tmp_path):
    assert os.path.exists(CALLER), "tests/cpp/reference_caller is built by __graft_entry__.build()"
    g = GEN
    o = O.Index.generate(g["N"], g["D"], g["vmin"], g["vmax"], g["is_int"], g["ratios"], g["seed"], g["sigma_div"])
    lo, hi = _boxes(o.data(), 3, 14, 10.0)
    qf = _write_ranges(tmp_path / "q_7.txt", lo, hi)
    out = str(tmp_path / "caller.csv")
    r = subprocess.run([CALLER, qf, out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    first = o.range(lo[0], hi[0]).grid()[0].size
    assert ("queries 1 fine mesh of the first %d" % first) in r.stdout, r.stdout
    with open(out) as f:
        rows = list(csv.reader(f))
    assert rows[0] == HEADER and rows[1][4] == "7%" and rows[1][0] == "Gen20000D4"
