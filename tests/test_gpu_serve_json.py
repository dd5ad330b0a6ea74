"""The reference's POST /query contract (HTTPServer.cu:414-553) over the serving loop: default
requests answer as the reference does (fine-mesh ids in traversal order, per-level sizes) and
equal the oracle's traversal; "exact" requests equal the oracle's refine; malformed requests get
the reference's 400 checks; JSON requests mixed with batched traffic keep their answers."""
import json
import threading

import numpy as np
import pytest

import oracle_py as O
import tblaeq

pytestmark = pytest.mark.gpu

CFG = dict(N=30000, D=5, vmin=1.0, vmax=100.0, is_int=False, ratios=(20, 10, 5), seed=19, sigma_div=3.0)
NAME = "Gen30000D5"


@pytest.fixture(scope="module")
def served():
    h = tblaeq.Handler.random(CFG["N"], CFG["D"], CFG["vmin"], CFG["vmax"], CFG["is_int"], CFG["ratios"],
                              CFG["seed"], CFG["sigma_div"])
    o = O.Index.generate(CFG["N"], CFG["D"], CFG["vmin"], CFG["vmax"], CFG["is_int"], CFG["ratios"], CFG["seed"],
                         CFG["sigma_div"])
    h.prepare("L0")
    srv = tblaeq.Server(h, max_batch=256)
    yield h, o, srv
    srv.close()
    h.close()


def _box(data, seed, half):
    rng = np.random.default_rng(seed)
    c = data[rng.integers(data.shape[0])]
    return c - half, c + half


def _check_common(body, qtype, h):
    assert body["success"] is True and body["dataset_name"] == NAME and body["query_type"] == qtype
    assert body["result_count"] == len(body["result"])
    assert body["memory_policy"] == "-".join(["L0"] * h.intervals)
    # mesh sizes of the levels' outputs, coarse to fine: the last is N (QueryEngine.cu:136-139)
    assert body["level_original_sizes"] == [h.mesh_sizes[h.intervals - 1 - l] for l in range(h.intervals)]
    assert body["level_original_sizes"][-1] == CFG["N"]
    rt = body["query_runtime"]
    assert set(rt) == {"query_time_us", "query_time_ms", "total_time_us", "avg_time_us", "query_count"}
    assert rt["query_count"] == 1 and rt["query_time_ms"] == pytest.approx(rt["query_time_us"] / 1000.0)


@pytest.mark.parametrize("half", [3.0, 12.0])
def test_range_request_matches_reference_traversal(served, half):
    h, o, srv = served
    lo, hi = _box(h.data(), int(half), half)
    status, body = srv.query_json({"dataset_name": NAME, "query_type": "range", "lower_bound": lo.tolist(),
                                   "upper_bound": hi.tolist()})
    assert status == 200, body
    _check_common(body, "RANGE", h)
    ro = o.range(lo, hi)
    assert body["result"] == ro.grid()[0].tolist()  # fine mesh, traversal order
    assert body["level_pruned_sizes"] == ro.levels()[2].tolist()


@pytest.mark.parametrize("k", [1, 10, 300])
def test_knn_request_matches_reference_traversal(served, k):
    h, o, srv = served
    q = h.data()[k * 7] + 0.3
    req = {"dataset_name": NAME, "query_type": "KNN", "query_point": q.tolist(), "k": k}
    status, body = srv.query_json(json.dumps(req))
    assert status == 200, body
    _check_common(body, "KNN", h)
    ro = o.knn(q, k)
    assert body["result"] == ro.grid()[0].tolist()
    assert body["level_pruned_sizes"] == ro.levels()[2].tolist()


def test_knn_default_k_is_10(served):
    h, o, srv = served
    q = h.data()[5] - 0.2
    status, body = srv.query_json({"dataset_name": NAME, "query_type": "knn", "query_point": q.tolist()})
    assert status == 200 and body["result"] == o.knn(q, 10).grid()[0].tolist()


@pytest.mark.parametrize("k", [4, 100, 700])
def test_exact_requests_match_oracle_refine(served, k):
    h, o, srv = served
    q = h.data()[k] + 0.1
    status, body = srv.query_json({"dataset_name": NAME, "query_type": "KNN", "query_point": q.tolist(), "k": k,
                                   "exact": True})
    assert status == 200, body
    assert body["exact"] is True and "level_pruned_sizes" not in body
    eid, ed2 = o.knn(q, k).refine_knn(q, k)
    assert body["result"] == eid.tolist()
    assert np.array_equal(np.array(body["squared_distances"]).view(np.uint64), ed2.view(np.uint64))
    lo, hi = _box(h.data(), k, 8.0)
    status, body = srv.query_json({"dataset_name": NAME, "query_type": "RANGE", "lower_bound": lo.tolist(),
                                   "upper_bound": hi.tolist(), "exact": True})
    assert status == 200, body
    assert body["result"] == o.range(lo, hi).refine_range(lo, hi).tolist()


BAD = [
    ("{not json", "not valid JSON"),
    ("[1, 2]", "JSON object"),
    ({"query_type": "KNN"}, "'dataset_name' is required"),
    ({"dataset_name": 5, "query_type": "KNN"}, "'dataset_name' is required"),
    ({"dataset_name": "a/b", "query_type": "KNN"}, "may hold only"),
    ({"dataset_name": "", "query_type": "KNN"}, "may hold only"),
    ({"dataset_name": "other", "query_type": "KNN"}, "not served here"),
    ({"dataset_name": NAME}, "'query_type' is required"),
    ({"dataset_name": NAME, "query_type": "SCAN"}, "must be 'KNN' or 'RANGE'"),
    ({"dataset_name": NAME, "query_type": "KNN"}, "'query_point' is required"),
    ({"dataset_name": NAME, "query_type": "KNN", "query_point": [1, "2", 3, 4, 5]}, "array of numbers"),
    ({"dataset_name": NAME, "query_type": "KNN", "query_point": [1, 2, 3]}, "must have 5 numbers"),
    ({"dataset_name": NAME, "query_type": "KNN", "query_point": [1, 2, 3, 4, 5], "k": 0}, "positive integer"),
    ({"dataset_name": NAME, "query_type": "KNN", "query_point": [1, 2, 3, 4, 5], "k": 2.5}, "positive integer"),
    ({"dataset_name": NAME, "query_type": "KNN", "query_point": [1, 2, 3, 4, 5], "k": "3"}, "positive integer"),
    ({"dataset_name": NAME, "query_type": "RANGE", "lower_bound": [0] * 5}, "are required"),
    ({"dataset_name": NAME, "query_type": "RANGE", "lower_bound": [0] * 5, "upper_bound": [1] * 4}, "5 numbers"),
    ({"dataset_name": NAME, "query_type": "RANGE", "lower_bound": [0, 0, 2, 0, 0], "upper_bound": [1] * 5},
     "lower_bound[2] > upper_bound[2]"),
    ({"dataset_name": NAME, "query_type": "RANGE", "lower_bound": [0] * 5, "upper_bound": [1] * 5, "exact": 1},
     "'exact' must be a boolean"),
]


@pytest.mark.parametrize("req,msg", BAD)
def test_rejected_requests(served, req, msg):
    _, _, srv = served
    status, body = srv.query_json(req)
    assert status == 400 and body["success"] is False and msg in body["message"], body


def test_dataset_name_suffix_accepted(served):
    # the reference names an index by its directory up to the first '_'
    h, o, srv = served
    q = h.data()[11]
    status, body = srv.query_json({"dataset_name": NAME + "_v2", "query_type": "KNN", "query_point": q.tolist(),
                                   "k": 3})
    assert status == 200 and body["dataset_name"] == NAME + "_v2"


def test_json_requests_in_a_batched_mix(served):
    h, o, srv = served
    data = h.data()
    rng = np.random.default_rng(3)
    expected, got, errors = {}, {}, []
    jobs = []
    for i in range(24):
        q = data[rng.integers(data.shape[0])] + rng.normal(0, 0.5, CFG["D"])
        k = int(rng.choice([1, 10, 50]))
        exact = bool(i % 2)
        jobs.append((i, q, k, exact))
        expected[i] = (o.knn(q, k).refine_knn(q, k)[0] if exact else o.knn(q, k).grid()[0]).tolist()

    def json_worker(part):
        try:
            for i, q, k, exact in jobs[part::3]:
                st, body = srv.query_json({"dataset_name": NAME, "query_type": "KNN", "query_point": q.tolist(),
                                           "k": k, "exact": exact})
                got[i] = (st, body["result"])
        except Exception as e:  # surfaced below
            errors.append(e)

    def batch_traffic():
        try:
            for _ in range(4):
                tickets = [srv.submit("range", np.concatenate(_box(data, int(s), 5.0))) for s in range(40)]
                for t in tickets:
                    srv.wait(t, 60.0)
                    srv.release(t)
        except Exception as e:
            errors.append(e)

    threads = [threading.Thread(target=json_worker, args=(p,)) for p in range(3)]
    threads.append(threading.Thread(target=batch_traffic))
    for t in threads:
        t.start()
    for t in threads:
        t.join(120)
    assert not errors, errors
    for i in expected:
        assert got[i] == (200, expected[i]), i
