"""Extract the reference's golden vectors and known-answer values into JSON.

Run in a container where the reference tree is mounted:

    python tests/golden/make_golden.py /root/reference

It reads the reference's Rust sources as text and pulls out only data: the
Backblaze LOG table and known products (src/galois.rs:119-151), the tower
structural constants (src/galois16.rs:19, :195-196), the GFNI identity qword
(src/gf/gfni.rs:85), the RS16(3,2) frozen wire vector
(src/reedsolomon16.rs:1087-1101) and the per-shape register / op / multiply
counts of the generated FFT programs (src/fft_programs.rs header lines).  The
output, reference_kats.json, is committed; tests never read the reference.
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import sys


def _read(root: str, rel: str) -> str:
    with open(os.path.join(root, rel), encoding="utf-8") as f:
        return f.read()


def extract(root: str) -> dict:
    out: dict = {}

    galois = _read(root, "src/galois.rs")
    m = re.search(r"BACKBLAZE_LOG_TABLE: \[u8; 256\] = \[(.*?)\];", galois, re.S)
    log = [int(x) for x in re.findall(r"\d+", m.group(1))]
    assert len(log) == 256
    out["gf8_log_table"] = {"source": "src/galois.rs:119-134", "values": log}

    prods = [
        {"a": int(a), "b": int(b), "want": int(w)}
        for a, b, w in re.findall(r"assert_eq!\(mul\((\d+), (\d+)\), (\d+)\)", galois)
    ]
    exps = [
        {"a": int(a), "n": int(n), "want": int(w)}
        for a, n, w in re.findall(r"assert_eq!\(exp\((\d+), (\d+)\), (\d+)\)", galois)
    ]
    out["gf8_known_products"] = {"source": "src/galois.rs:144-151", "mul": prods, "exp": exps}

    g16 = _read(root, "src/galois16.rs")
    tl = int(re.search(r"pub const TOWER_L: u8 = (\d+);", g16).group(1))
    out["tower_l"] = {"source": "src/galois16.rs:19", "value": tl}
    cm = re.findall(r"assert_eq!\(const_matrix\((0x[0-9a-fA-F]+|\d+)\), \[([^\]]*)\]", g16)
    consts = []
    for c, vals in cm:
        items = [v.strip() for v in vals.split(",")]
        consts.append({"c": int(c, 0),
                       "want": [tl if v == "TOWER_L" else int(v, 0) for v in items]})
    out["gf16_const_matrix"] = {"source": "src/galois16.rs:195-196", "cases": consts}

    gfni = _read(root, "src/gf/gfni.rs")
    ident = re.search(r"gfni_matrix\(1\) as u64, (0x[0-9a-fA-F_]+)u64", gfni).group(1)
    out["gfni_identity"] = {"source": "src/gf/gfni.rs:85",
                            "c": 1, "qword": int(ident.replace("_", ""), 16)}

    rs16 = _read(root, "src/reedsolomon16.rs")
    body = rs16[rs16.index("fn golden_wire_vector"):]
    body = body[: body.index("\n    }\n")]
    shape = re.search(r"ReedSolomon16::new\((\d+), (\d+)\)", body)
    vecs = re.findall(r"vec!\[(0x[0-9a-f, x]+)\]", body)
    data = [[int(v, 16) for v in re.findall(r"0x[0-9a-f]+", s)] for s in vecs]
    want = re.findall(r"assert_eq!\(shards\[(\d+)\], \[([\d, ]+)\]", body)
    out["rs16_wire_vector"] = {
        "source": "src/reedsolomon16.rs:1087-1101",
        "k": int(shape.group(1)),
        "m": int(shape.group(2)),
        "data": data,
        "parity": {int(i): [int(x) for x in v.split(",")] for i, v in want},
    }

    progs = _read(root, "src/fft_programs.rs")
    shapes = []
    for k, m, regs, ops, mults in re.findall(
        r"// shape \((\d+),(\d+)\): (\d+) registers, (\d+) ops, (\d+) multiplies", progs
    ):
        # The op list of the shape's macro, one "op a b [c]" line per token
        # group, hashed: the reference's builder output, byte for byte.
        body = progs[progs.index(f"macro_rules! fft_program_{k}_{m} "):]
        body = body[: body.index("pub(crate) use")]
        lines = [" ".join(t.split()) for t in re.findall(r"^\s+((?:ld|mx|xo|mc|cp|st)[^;]*);",
                                                         body, re.M)]
        digest = hashlib.sha256("\n".join(lines).encode()).hexdigest()
        shapes.append({"k": int(k), "m": int(m), "registers": int(regs), "ops": int(ops),
                       "multiplies": int(mults), "op_lines": len(lines), "op_sha256": digest})
    out["fft_generated_programs"] = {"source": "src/fft_programs.rs:19,133,254,451,622,787",
                                     "shapes": shapes}
    return out


def main() -> None:
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = extract(root)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")
    with open(dst, "w", encoding="utf-8") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", dst)


if __name__ == "__main__":
    main()
