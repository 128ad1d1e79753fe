"""The base-graph tables every decoder is built from (code_tables.inc, which
the product and the oracle both compile in), pinned against the
reference's data (CPU only).

tests/golden/bg_tables.json is the reference's data/bg_tables.json (3GPP TS
38.212 Tables 5.3.2-2 and 5.3.2-3, consumed there by build.rs:26-92), kept
verbatim as a fixture.  For both base graphs and all 51 lifting sizes, the
code's layered layout must be the JSON's entries grouped by row block in
the JSON's own order, with shift = V[iLS] mod Z (QcLdpcParams::new and
entry_col_shift, qc_ldpc.rs:134-140, 281-366).  Where the reference tree is
mounted, regenerating code_tables.inc from it must reproduce the committed
file byte for byte (802.11 tables included)."""
import json
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF = "/root/reference"

# 3GPP TS 38.212 Table 5.3.2-1 (qc_ldpc.rs:103-125)
SETS = [[2, 4, 8, 16, 32, 64, 128, 256], [3, 6, 12, 24, 48, 96, 192, 384], [5, 10, 20, 40, 80, 160, 320],
        [7, 14, 28, 56, 112, 224], [9, 18, 36, 72, 144, 288], [11, 22, 44, 88, 176, 352], [13, 26, 52, 104, 208],
        [15, 30, 60, 120, 240]]


def test_every_lifting_size_of_both_graphs_matches_the_reference_json(oracle):
    tables = json.load(open(os.path.join(GOLDEN, "bg_tables.json")))
    assert sum(len(s) for s in SETS) == 51
    for bg, key in ((0, "bg1"), (1, "bg2")):
        t = tables[key]
        rows = [[] for _ in range(t["rows"])]
        for e in t["entries"]:
            rows[e["r"]].append((e["c"], e["v"]))
        for ils, zs in enumerate(SETS):
            for z in zs:
                c = oracle.code_3gpp(bg, z)
                assert (c.m_b, c.n_b, c.n_edges, c.z) == (t["rows"], t["cols"], len(t["entries"]), z)
                e = 0
                for r, row in enumerate(rows):
                    assert c.layer_off[r] == e, (key, z, r)
                    for col, v in row:
                        assert (c.cols[e], c.shifts[e]) == (col, v[ils] % z), (key, z, r, col)
                        e += 1
                assert c.layer_off[t["rows"]] == e
                assert c.max_deg == max(len(r) for r in rows)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "data", "bg_tables.json")),
                    reason="the reference tree is not mounted here")
def test_regenerated_tables_equal_the_committed_file(tmp_path):
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "gen_code_tables", os.path.join(ROOT, "syndrome_amd", "csrc", "gen_code_tables.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    out = tmp_path / "code_tables.inc"
    gen.main(REF, str(out))
    committed = open(os.path.join(ROOT, "syndrome_amd", "csrc", "code_tables.inc")).read()
    assert out.read_text() == committed


def test_fixture_is_the_reference_file():
    path = os.path.join(REF, "data", "bg_tables.json")
    if not os.path.exists(path):
        pytest.skip("the reference tree is not mounted here")
    assert open(path, "rb").read() == open(os.path.join(GOLDEN, "bg_tables.json"), "rb").read()
