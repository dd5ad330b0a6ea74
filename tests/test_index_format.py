"""The on-disk index format pinned against the reference's writer, IndexSerializer.cu:69-259,
read as a specification: a parser and a writer of that layout written here from the
reference's write calls alone (not from this repository's serializer), then

* the oracle's files (which the library reproduces byte for byte, test_host.py) are consumed by
  the parser with no bytes left over, their headers agree with metadata.bin and with each other,
  and their structure is what the reference's readers assume (CSC pointers, maps as
  permutations, one radius per centroid);
* the finest points rebuilt from the parsed files alone by the reference's query rules
  (child = P value * parent, per dimension, SpMSpV.cuh:64; child id = map[row id],
  Refactor.cu:10-27 with L0Strategy.cu:61-184) equal the oracle's reconstruction bitwise;
* the writer reproduces every file byte for byte from the parsed arrays, and a directory
  written by it (with the rows of every P column given in a different order) loads in the
  library and answers range queries exactly as the original index.
"""
import os

import numpy as np
import pytest

import oracle_py as O
import tblaeq

U64, F64 = np.dtype("<u8"), np.dtype("<f8")


class _Reader:
    def __init__(self, path):
        self.raw = open(path, "rb").read()
        self.pos = 0

    def u64(self, n=None):
        k = 1 if n is None else n
        a = np.frombuffer(self.raw, U64, k, self.pos)
        self.pos += 8 * k
        return int(a[0]) if n is None else a.copy()

    def f64(self, n):
        a = np.frombuffer(self.raw, F64, n, self.pos)
        self.pos += 8 * n
        return a.copy()

    def byte(self):
        b = self.raw[self.pos]
        self.pos += 1
        return b

    def done(self):
        return self.pos == len(self.raw)


def read_index(d):
    """IndexSerializer::load's view of a directory (IndexSerializer.cu:34-66, 90-259)."""
    r = _Reader(os.path.join(d, "metadata.bin"))  # saveMetadata, :69-88
    ix = {"D": r.u64(), "N": r.u64(), "height": r.u64(), "intervals": r.u64(), "isAos": r.byte()}
    ix["ratios"] = r.u64(ix["intervals"])
    ix["meshSizes"] = r.u64(r.u64())
    assert r.done(), "metadata.bin: trailing bytes"
    L = ix["intervals"]
    ix["P"] = []
    for i in range(L):  # saveSparseTensorCsc, :116-138
        r = _Reader(os.path.join(d, "P_Tensor_%d.bin" % i))
        t = {"D": r.u64(), "rows": r.u64(), "cols": r.u64(), "nnz": r.u64(), "isAos": r.byte()}
        t["rowIds"] = r.u64(t["nnz"])
        t["nnzPerCol"] = r.u64(t["cols"])
        t["colRes"] = r.u64(t["cols"] + 1)
        t["vals"] = r.f64(t["nnz"] * t["D"])
        assert r.done(), "P_Tensor_%d.bin: trailing bytes" % i
        ix["P"].append(t)
    ix["radius"] = []
    for i in range(L):  # saveMaxRadius, :178-185 (count = meshSizes[intervals - i], :20-23)
        r = _Reader(os.path.join(d, "MaxRadius_%d.bin" % i))
        ix["radius"].append(r.f64(r.u64()))
        assert r.done()
    ix["maps"] = {}
    for i in range(1, ix["height"]):  # saveMap, :198-205 (maps 1..height-1, :25-29)
        p = os.path.join(d, "Map_%d.bin" % i)
        if os.path.exists(p):
            r = _Reader(p)
            ix["maps"][i] = r.u64(r.u64())
            assert r.done()
    r = _Reader(os.path.join(d, "Coreast_Mesh.bin"))  # saveCoarsestMesh, :218-235
    m = {"rows": r.u64(), "dim": r.u64(), "nnz": r.u64(), "isAos": r.byte()}
    m["ids"] = r.u64(m["nnz"])
    m["vals"] = r.f64(m["nnz"] * m["dim"])
    assert r.done()
    ix["coarsest"] = m
    return ix


def write_index(ix, d):
    """IndexSerializer::save of the parsed arrays (IndexSerializer.cu:11-32)."""
    os.makedirs(d, exist_ok=True)

    def u64(*v):
        return np.asarray(v, U64).tobytes()

    with open(os.path.join(d, "metadata.bin"), "wb") as f:
        f.write(u64(ix["D"], ix["N"], ix["height"], ix["intervals"]) + bytes([ix["isAos"]]))
        f.write(np.asarray(ix["ratios"], U64).tobytes() + u64(len(ix["meshSizes"])))
        f.write(np.asarray(ix["meshSizes"], U64).tobytes())
    for i, t in enumerate(ix["P"]):
        with open(os.path.join(d, "P_Tensor_%d.bin" % i), "wb") as f:
            f.write(u64(t["D"], t["rows"], t["cols"], t["nnz"]) + bytes([t["isAos"]]))
            for a, dt in ((t["rowIds"], U64), (t["nnzPerCol"], U64), (t["colRes"], U64), (t["vals"], F64)):
                f.write(np.asarray(a, dt).tobytes())
    for i, rad in enumerate(ix["radius"]):
        with open(os.path.join(d, "MaxRadius_%d.bin" % i), "wb") as f:
            f.write(u64(len(rad)) + np.asarray(rad, F64).tobytes())
    for i, mp in ix["maps"].items():
        with open(os.path.join(d, "Map_%d.bin" % i), "wb") as f:
            f.write(u64(len(mp)) + np.asarray(mp, U64).tobytes())
    m = ix["coarsest"]
    with open(os.path.join(d, "Coreast_Mesh.bin"), "wb") as f:
        f.write(u64(m["rows"], m["dim"], m["nnz"]) + bytes([m["isAos"]]))
        f.write(np.asarray(m["ids"], U64).tobytes() + np.asarray(m["vals"], F64).tobytes())


def rebuild_finest(ix):
    """Every point of every level from the files alone, by the reference's query rules: the
    coarsest mesh gives level 0 by id; a column c of P_l spreads parent c to its rows j, child
    value = vals[j] * parent value per dimension (AoS), child id = Map_{l+1}[rowIds[j]]."""
    D = ix["D"]
    m = ix["coarsest"]
    X = np.zeros((ix["meshSizes"][-1], D))
    X[m["ids"].astype(np.int64)] = m["vals"].reshape(-1, D)
    for l, t in enumerate(ix["P"]):
        mp = ix["maps"].get(l + 1)
        cols = np.repeat(np.arange(t["cols"]), np.diff(t["colRes"]).astype(np.int64))
        child = t["vals"].reshape(-1, D) * X[cols]
        ids = t["rowIds"] if mp is None else mp[t["rowIds"].astype(np.int64)]
        nxt = np.zeros((t["nnz"], D))
        nxt[ids.astype(np.int64)] = child
        X = nxt
    return X


CASES = [
    (6000, 3, 1.0, 100.0, False, (10, 5), 4, 3.0),
    (5000, 5, 1.0, 50.0, True, (10, 5), 99, 2.2),
    (20000, 4, 1.0, 100.0, False, (20, 10, 5), 7, 3.0),
]


@pytest.mark.parametrize("args", CASES)
def test_files_follow_the_reference_layout(tmp_path, args):
    o = O.Index.generate(*args)
    d = str(tmp_path / "ix")
    o.save(d)
    ix = read_index(d)
    L, D = ix["intervals"], ix["D"]
    ms = ix["meshSizes"]
    assert (D, ix["N"], ix["height"], L) == (args[1], args[0], L + 1, len(args[5]))
    assert len(ms) == L + 1 and ms[0] == ix["N"]
    for i, t in enumerate(ix["P"]):
        # P_i maps the centroids of mesh level L - i (columns) to the points of level L - i - 1
        assert t["D"] == D and t["isAos"] == ix["isAos"]
        assert t["cols"] == ms[L - i] and t["nnz"] == ms[L - i - 1]
        assert t["colRes"][0] == 0 and t["colRes"][-1] == t["nnz"]
        assert np.array_equal(np.diff(t["colRes"]), t["nnzPerCol"])
        assert np.array_equal(np.sort(t["rowIds"]), np.arange(t["nnz"]))
        assert len(ix["radius"][i]) == ms[L - i] and (ix["radius"][i] >= 0).all()
    for i, mp in ix["maps"].items():
        assert len(mp) == ms[L - i] and np.array_equal(np.sort(mp), np.arange(len(mp)))
    m = ix["coarsest"]
    assert m["dim"] == D and m["nnz"] == ms[L] and np.array_equal(np.sort(m["ids"]), np.arange(ms[L]))


@pytest.mark.parametrize("args", CASES)
def test_rebuilt_points_equal_the_oracle_reconstruction(tmp_path, args):
    o = O.Index.generate(*args)
    d = str(tmp_path / "ix")
    o.save(d)
    X = rebuild_finest(read_index(d))
    R = o.recon()
    assert np.array_equal(X.view(np.uint64), R.view(np.uint64))


def test_radius_file_bounds_the_rebuilt_children(tmp_path):
    """MaxRadius_i[c] is the largest distance of centroid c's children (from the original
    coordinates, Setup.cuh:148-171): the rebuilt children lie within it up to rounding."""
    o = O.Index.generate(*CASES[0])
    d = str(tmp_path / "ix")
    o.save(d)
    ix = read_index(d)
    D = ix["D"]
    m = ix["coarsest"]
    parent = np.zeros((ix["meshSizes"][-1], D))
    parent[m["ids"].astype(np.int64)] = m["vals"].reshape(-1, D)
    t = ix["P"][0]
    cols = np.repeat(np.arange(t["cols"]), np.diff(t["colRes"]).astype(np.int64))
    child = t["vals"].reshape(-1, D) * parent[cols]
    dist = np.sqrt(((child - parent[cols]) ** 2).sum(1))
    far = np.zeros(t["cols"])
    np.maximum.at(far, cols, dist)
    assert np.allclose(far, ix["radius"][0], rtol=1e-12, atol=1e-12)


def test_writer_reproduces_the_files_and_the_library_reads_them(tmp_path):
    o = O.Index.generate(*CASES[2])
    a, b, c = (str(tmp_path / x) for x in "abc")
    o.save(a)
    ix = read_index(a)
    write_index(ix, b)
    for f in sorted(os.listdir(a)):
        assert open(os.path.join(a, f), "rb").read() == open(os.path.join(b, f), "rb").read(), f
    # the rows of every column in reverse order (a valid CSC the reference could write; the
    # values and ids move with their rows): the library reads it and answers as before
    for t in ix["P"]:
        cr = t["colRes"].astype(np.int64)
        perm = np.concatenate([np.arange(cr[k + 1] - 1, cr[k] - 1, -1) for k in range(t["cols"])])
        t["rowIds"] = t["rowIds"][perm]
        t["vals"] = t["vals"].reshape(-1, ix["D"])[perm].reshape(-1)
    write_index(ix, c)
    assert np.array_equal(rebuild_finest(ix).view(np.uint64), o.recon().view(np.uint64))
    h = tblaeq.Handler.load(c)
    try:
        assert h.mesh_sizes == ix["meshSizes"].tolist()
        h.save(str(tmp_path / "d"))
    finally:
        h.close()
    for f in sorted(os.listdir(c)):
        assert open(os.path.join(c, f), "rb").read() == open(os.path.join(str(tmp_path / "d"), f), "rb").read(), f


@pytest.mark.gpu
def test_reordered_columns_answer_as_the_original(tmp_path):
    """The directory with every column's rows reversed answers range and KNN queries on the GPU
    exactly as the original index (the candidate order may differ; the refined answers may not)."""
    o = O.Index.generate(*CASES[2])
    a, c = str(tmp_path / "a"), str(tmp_path / "c")
    o.save(a)
    ix = read_index(a)
    for t in ix["P"]:
        cr = t["colRes"].astype(np.int64)
        perm = np.concatenate([np.arange(cr[k + 1] - 1, cr[k] - 1, -1) for k in range(t["cols"])])
        t["rowIds"] = t["rowIds"][perm]
        t["vals"] = t["vals"].reshape(-1, ix["D"])[perm].reshape(-1)
    write_index(ix, c)
    h1, h2 = tblaeq.Handler.load(a), tblaeq.Handler.load(c)
    try:
        for h in (h1, h2):
            h.prepare("L0")
        R = o.recon()
        rng = np.random.default_rng(3)
        for _ in range(20):
            x = R[rng.integers(len(R))]
            lo, hi = x - 8.0, x + 8.0
            r1, r2 = h1.range(lo, hi), h2.range(lo, hi)
            assert np.array_equal(r1.refined()[0], r2.refined()[0])
            assert np.array_equal(r1.refined()[0], o.range(lo, hi).refine_range(lo, hi))
            k1, k2 = h1.knn(x + 0.5, 10), h2.knn(x + 0.5, 10)
            assert np.array_equal(k1.refined()[0], k2.refined()[0])
            assert np.array_equal(k1.refined()[1].view(np.uint64), k2.refined()[1].view(np.uint64))
            for r in (r1, r2, k1, k2):
                r.close()
    finally:
        h1.close()
        h2.close()
