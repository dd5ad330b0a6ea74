"""Text formats, the command-line front end and the C++ drop-in on the CPU (no kernels run):
the dataset and query-file loaders of the reference's format (File.cu:8-216) with its
tolerances and errors, the CSV label of a range query file's name (Utils.cu:18-35), the CLI's
index generation and CPU build byte-identical to the oracle's, its argument checks
(main.cu:11-29,113-165), and a reference-style C++ caller that compiles and links against the
public host header and the shared library."""
import filecmp
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import tblaeq

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "t-blaeq_amd", "tblaeq-cli")


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return str(path)


def _rows(a):
    return "\n".join(" ".join(repr(float(v)) for v in row) for row in a) + "\n"


# ---------------------------------------------------------------- loaders

def test_dataset_loader_reads_rows_exactly(tmp_path):
    a = np.random.default_rng(1).uniform(-5, 5, (37, 4))
    p = _write(tmp_path / "pts_a.txt", "37 4\n" + _rows(a))
    assert np.array_equal(tblaeq.load_dataset(p).view(np.uint64), a.view(np.uint64))


def test_dataset_loader_tolerances(tmp_path):
    # extra numbers on a row are ignored and rows past the count are not read (File.cu:30-45)
    p = _write(tmp_path / "extra_a.txt", "2 2\n1 2 99\n3 4\n5 6\n")
    assert tblaeq.load_dataset(p).tolist() == [[1.0, 2.0], [3.0, 4.0]]


@pytest.mark.parametrize("text", ["", "x 2\n1 2\n", "0 2\n", "2 0\n", "2 2\n1 2\n3\n", "3 2\n1 2\n3 4\n",
                                  "2 2\n1 2\n\n3 4\n"])
def test_dataset_loader_errors(tmp_path, text):
    p = _write(tmp_path / "bad_a.txt", text)
    with pytest.raises(tblaeq.TblaeqError, match="index file error"):
        tblaeq.load_dataset(p)


def test_dataset_loader_missing_file(tmp_path):
    with pytest.raises(tblaeq.TblaeqError, match="index file error"):
        tblaeq.load_dataset(str(tmp_path / "none_a.txt"))


def test_query_file_loaders(tmp_path):
    lo = np.random.default_rng(2).uniform(0, 10, (5, 3))
    hi = lo + 1.5
    p = _write(tmp_path / "boxes_10-3.txt", "5 3\n" + _rows(np.concatenate([lo, hi], 1)))
    rows, info = tblaeq.load_query_file(p, "range")
    assert np.array_equal(rows[:, :3], lo) and np.array_equal(rows[:, 3:], hi)
    # "P-E" labels: P%^E = (P/100)^E * 100 % with two decimals (Utils.cu:18-35)
    assert info == "10%^3=0.10%"
    pts = np.random.default_rng(3).normal(size=(4, 3))
    p2 = _write(tmp_path / "pts_k.txt", "4 3\n" + _rows(pts))
    rows2, _ = tblaeq.load_query_file(p2, "knn")
    assert np.array_equal(rows2, pts)
    # a point file read as ranges needs 2D numbers per row
    with pytest.raises(tblaeq.TblaeqError, match="index file error"):
        tblaeq.load_query_file(p2, "range")


@pytest.mark.parametrize("name,label", [("q_5.txt", "5%"), ("set_50-2.txt", "50%^2=25.00%"),
                                        ("a_b_0.1.txt", "0%"), ("x_25-1.dat", "25%^1=25.00%")])
def test_range_file_labels(tmp_path, name, label):
    p = _write(tmp_path / name, "1 1\n0 1\n")
    assert tblaeq.load_query_file(p, "range")[1] == label


# ---------------------------------------------------------------- CLI

def _cli(*args, cwd=None):
    return subprocess.run([CLI, *args], capture_output=True, text=True, cwd=cwd, timeout=120)


def test_cli_gen_index_identical_to_oracle(tmp_path):
    out = str(tmp_path / "gen")
    r = _cli("--gen-index", "--gen-N", "12000", "--gen-D", "3", "--gen-min", "1", "--gen-max", "50", "--gen-int",
             "--gen-seed", "77", "--gen-sigma-divisor", "2.5", "--height", "3", "--ratios", "20,10", "-i", out)
    assert r.returncode == 0, r.stderr
    ref = str(tmp_path / "oracle")
    O.Index.generate(12000, 3, 1.0, 50.0, True, (20, 10), 77, 2.5).save(ref)
    for f in sorted(os.listdir(ref)):
        assert filecmp.cmp(os.path.join(out, f), os.path.join(ref, f), shallow=False), f


def test_cli_cpu_build_identical_to_oracle(tmp_path):
    data = np.random.default_rng(4).uniform(1, 100, (2500, 3))
    ds = _write(tmp_path / "cloud_a.txt", "2500 3\n" + _rows(data))
    out = str(tmp_path / "built")
    r = _cli("--build-index", "-d", ds, "--force-cpu", "true", "--height", "3", "--ratios", "10,5", "-i", out)
    assert r.returncode == 0, r.stderr
    ref = str(tmp_path / "oracle")
    O.Index.build(data, (10, 5), 7).save(ref)
    for f in sorted(os.listdir(ref)):
        assert filecmp.cmp(os.path.join(out, f), os.path.join(ref, f), shallow=False), f


@pytest.mark.parametrize("args,msg", [
    ((), "exactly one of"),
    (("--gen-index", "--test-query"), "exactly one of"),
    (("--gen-index", "--height", "1"), "--height must be at least 2"),
    (("--gen-index", "--ratios", "100,,20"), "empty entry"),
    (("--gen-index", "--ratios", "100,0,20"), "positive integers"),
    (("--gen-index", "--height", "3"), "height - 1 entries"),
    (("--gen-index", "--gen-sigma-divisor", "0"), "must be positive"),
    (("--test-query", "-t", "2"), "--query-type must be 0"),
    (("--gen-index", "--bogus"), "unknown argument"),
    (("--gen-index", "--force-cpu", "maybe"), "not a boolean"),
])
def test_cli_argument_errors(tmp_path, args, msg):
    r = _cli(*args, "-i", str(tmp_path / "unused"))
    assert r.returncode != 0 and msg in r.stderr, (r.returncode, r.stderr)


def test_cli_help():
    r = _cli("--help")
    assert r.returncode == 0 and "--build-index" in r.stdout and "--force-policy" in r.stdout


# ---------------------------------------------------------------- C++ drop-in

def test_reference_style_cpp_caller_compiles_and_links(tmp_path):
    """The C++ call site of the reference's main.cu, written against include/tblaeq_host.hpp,
    compiles with the host compiler and links against libtblaeq.so."""
    exe = str(tmp_path / "caller")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                        "-D__HIP_PLATFORM_AMD__", os.path.join(ROOT, "tests", "cpp", "reference_caller.cpp"), "-o",
                        exe, "-L", os.path.join(ROOT, "t-blaeq_amd"), "-ltblaeq",
                        "-Wl,-rpath," + os.path.join(ROOT, "t-blaeq_amd")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert os.path.getsize(exe) > 0
