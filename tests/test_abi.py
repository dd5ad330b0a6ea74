"""The C-ABI library loads (no GPU needed) and exports every symbol the public
headers declare; the Python binding declares each of them."""
import os
import re
import subprocess

import tblaeq

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def declared():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(tblaeq_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_headers_declare_entry_points():
    names = declared()
    for must in ("tblaeq_range_prune_compact", "tblaeq_knn_lower_bound_sort_step", "tblaeq_prolong_refactor",
                 "tblaeq_refine_range", "tblaeq_refine_knn_topk", "tblaeq_sort_pairs", "tblaeq_batch_range",
                 "tblaeq_batch_knn", "tblaeq_handler_query", "tblaeq_ctx_create"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = tblaeq.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", tblaeq.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    for n in declared():
        getattr(lib, n)


def test_python_binding_covers_headers():
    assert declared() <= set(tblaeq.SIGNATURES), sorted(declared() - set(tblaeq.SIGNATURES))


def test_status_strings():
    lib = tblaeq.lib()
    for code in range(7):
        assert lib.tblaeq_status_string(code)
