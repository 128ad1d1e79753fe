#!/usr/bin/env python3
"""Regenerate ``code_tables.inc`` — the base-graph tables every decoder here
is built from.

The data are standard tables, not code:

* 3GPP TS 38.212 Table 5.3.2-2 (BG1, 46 x 68, 316 non-null entries) and
  Table 5.3.2-3 (BG2, 42 x 52, 197 entries): for each non-null entry its
  row block, column block and the eight shift coefficients V_{i,j}, one per
  lifting-size set iLS = 0..7.  Read from the reference's
  ``data/bg_tables.json`` (consumed there by ``build.rs:26-92``).
* IEEE 802.11 LDPC parity-check matrices for Z in {27, 54, 81} and rates
  {1/2, 2/3, 3/4, 5/6}: absolute per-Z shifts
  (``src/wifi_ldpc_tables.rs:109-436``).

Entries are emitted grouped by row block, and within a row in the
reference table's own order.  That order is observable: the decoder's
check-to-variable buffer is laid out edge by edge in it
(``src/qc_ldpc.rs:281-318``), and it is *not* ascending-column everywhere
(BG1 row 15 lists column 0 last; BG2 row 16 lists column 26 first).

Run once, in a checkout where the reference tree is mounted:

    python3 syndrome_amd/csrc/gen_code_tables.py /root/reference
"""
import json
import os
import re
import sys


def bg_rows(bg):
    """Stable group-by-row of a base graph's entry list."""
    rows = [[] for _ in range(bg["rows"])]
    for e in bg["entries"]:
        rows[e["r"]].append((e["c"], list(e["v"])))
    return rows


def wifi_matrices(src):
    arrays = {}
    for m in re.finditer(r"const (WIFI_Z\d+_R\d_\d_ENTRIES): \[\(u8, u8, i16\); (\d+)\] = \[(.*?)\];", src, re.S):
        name, count, body = m.group(1), int(m.group(2)), m.group(3)
        triples = [tuple(int(x) for x in t) for t in re.findall(r"\((\d+),\s*(\d+),\s*(\d+)\)", body)]
        assert len(triples) == count, (name, len(triples), count)
        arrays[name] = triples
    mats = []
    for m in re.finditer(
        r"WifiLdpcMatrix \{\s*z: (\d+),\s*rate_num: (\d+),\s*rate_den: (\d+),\s*row_blocks: (\d+),\s*"
        r"col_blocks: (\d+),\s*entries: &(\w+),\s*\}",
        src,
    ):
        z, rn, rd, mb, nb, name = m.groups()
        mats.append((int(z), int(rn), int(rd), int(mb), int(nb), arrays[name]))
    assert len(mats) == 12, len(mats)
    return mats


def main(ref_root, dest=None):
    with open(os.path.join(ref_root, "data", "bg_tables.json")) as f:
        tables = json.load(f)
    with open(os.path.join(ref_root, "src", "wifi_ldpc_tables.rs")) as f:
        wifi_src = f.read()

    out = []
    out.append("/* Generated by gen_code_tables.py -- base-graph tables (data, not code).\n"
               " * 3GPP TS 38.212 Tables 5.3.2-2 / 5.3.2-3 and IEEE 802.11 LDPC matrices.\n"
               " * Each row block's entries keep the source table's within-row order. */\n"
               "#ifndef SYN_TABLE_QUAL\n#define SYN_TABLE_QUAL static const\n#endif\n")
    for key, tag in (("bg1", "BG1"), ("bg2", "BG2")):
        bg = tables[key]
        rows = bg_rows(bg)
        n = sum(len(r) for r in rows)
        out.append(f"#define SYN_{tag}_ROWS {bg['rows']}\n#define SYN_{tag}_COLS {bg['cols']}\n"
                   f"#define SYN_{tag}_EDGES {n}\n")
        out.append(f"SYN_TABLE_QUAL unsigned char SYN_{tag}_ROW_DEG[{bg['rows']}] = {{"
                   + ", ".join(str(len(r)) for r in rows) + "};\n")
        out.append(f"SYN_TABLE_QUAL unsigned char SYN_{tag}_EDGE_COL[{n}] = {{\n")
        for r, ents in enumerate(rows):
            out.append("  " + ", ".join(str(c) for c, _ in ents) + f",  /* row {r} */\n")
        out.append("};\n")
        out.append(f"SYN_TABLE_QUAL short SYN_{tag}_EDGE_V[{n}][8] = {{\n")
        for r, ents in enumerate(rows):
            out.append("  " + ", ".join("{" + ",".join(str(x) for x in v) + "}" for _, v in ents)
                       + f",  /* row {r} */\n")
        out.append("};\n\n")

    mats = wifi_matrices(wifi_src)
    out.append(f"#define SYN_WIFI_COUNT {len(mats)}\n")
    for i, (z, rn, rd, mb, nb, ents) in enumerate(mats):
        rows = [[] for _ in range(mb)]
        for r, c, s in ents:
            rows[r].append((c, s))
        flat = [(r, c, s) for r in range(mb) for c, s in rows[r]]
        out.append(f"SYN_TABLE_QUAL short SYN_WIFI_{i}_EDGES[{len(flat)}][3] = {{\n  "
                   + ", ".join("{%d,%d,%d}" % t for t in flat) + "\n};\n")
    out.append("typedef struct { int z, rate_num, rate_den, row_blocks, col_blocks, n_edges;"
               " const short (*edges)[3]; } syn_wifi_matrix_t;\n")
    out.append("SYN_TABLE_QUAL syn_wifi_matrix_t SYN_WIFI_MATRICES[SYN_WIFI_COUNT] = {\n")
    for i, (z, rn, rd, mb, nb, ents) in enumerate(mats):
        out.append(f"  {{{z}, {rn}, {rd}, {mb}, {nb}, {len(ents)}, SYN_WIFI_{i}_EDGES}},\n")
    out.append("};\n")

    if dest is None:
        dest = os.path.join(os.path.dirname(os.path.abspath(__file__)), "code_tables.inc")
    with open(dest, "w") as f:
        f.write("".join(out))
    print("wrote", dest)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
