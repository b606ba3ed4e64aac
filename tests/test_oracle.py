"""Pins for the CPU oracle (oracle/oracle.c) -- CPU only.

Each check compares the oracle with something other than itself: the
reference's golden vectors (tests/golden/reference_kats.json, extracted from
the reference's own tests), independent textbook restatements written here in
numpy (carry-less polynomial products, a polynomial-basis tower multiply, a
separate Gauss-Jordan), or mathematical invariants (field axioms, inverses,
V * V^-1 == I, MDS, round trips).
"""
import json
import os
import itertools

import numpy as np
import pytest

import oracle as O

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


# ---------------------------------------------------------------- independent helpers

def clmul_mod_11d(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Textbook GF(2^8) product: carry-less multiply then reduce by x^8+x^4+x^3+x^2+1."""
    a = a.astype(np.uint32)
    b = b.astype(np.uint32)
    p = np.zeros(np.broadcast(a, b).shape, np.uint32)
    for bit in range(8):
        p ^= np.where((b >> bit) & 1, a << bit, 0).astype(np.uint32)
    for bit in range(14, 7, -1):
        p ^= np.where((p >> bit) & 1, 0x11D << (bit - 8), 0).astype(np.uint32)
    return p.astype(np.uint8)


MUL = None


def mul_table() -> np.ndarray:
    global MUL
    if MUL is None:
        a = np.arange(256)
        MUL = clmul_mod_11d(a[:, None], a[None, :])
    return MUL


def tower_mul_poly(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """(ah*y+al)(bh*y+bl) mod y^2+y+L, written from the polynomial definition with
    the textbook GF(2^8) table above (not the oracle's log/exp tables)."""
    T = mul_table()
    L = KATS["tower_l"]["value"]
    ah, al, bh, bl = a >> 8, a & 0xFF, b >> 8, b & 0xFF
    y2 = T[ah, bh]
    y1 = T[ah, bl] ^ T[al, bh]
    y0 = T[al, bl]
    # y^2 = y + L
    hi = y1 ^ y2
    lo = y0 ^ T[y2, L]
    return (hi.astype(np.uint16) << 8) | lo


def gauss_jordan_np(m: np.ndarray, mul, inv):
    """Independent inversion (partial pivot by any nonzero row), returns None if singular."""
    n = m.shape[0]
    a = [list(map(int, row)) + [1 if i == j else 0 for j in range(n)] for i, row in enumerate(m)]
    for c in range(n):
        p = next((r for r in range(c, n) if a[r][c]), None)
        if p is None:
            return None
        a[c], a[p] = a[p], a[c]
        s = inv(a[c][c])
        a[c] = [mul(s, x) for x in a[c]]
        for r in range(n):
            if r != c and a[r][c]:
                f = a[r][c]
                a[r] = [x ^ mul(f, y) for x, y in zip(a[r], a[c])]
    return np.array([row[n:] for row in a])


# ---------------------------------------------------------------- GF(2^8)

def test_log_table_matches_backblaze_golden():
    log, exp = O.gf8_tables()
    assert list(log) == KATS["gf8_log_table"]["values"]
    # doubled EXP: exp[i] == exp[i + 255]
    assert (exp[:255] == exp[255:510]).all()


def test_known_products_golden():
    for p in KATS["gf8_known_products"]["mul"]:
        assert O.gf8_mul(p["a"], p["b"]) == p["want"]
    for p in KATS["gf8_known_products"]["exp"]:
        assert O.gf8_exp(p["a"], p["n"]) == p["want"]


def test_mul_table_equals_clmul_exhaustive():
    assert (O.gf8_mul_table() == mul_table()).all()


def test_div_is_inverse_of_mul():
    T = mul_table()
    for a in range(1, 256):
        assert T[a, O.gf8_div(1, a)] == 1
        b = (a * 37 + 11) % 255 + 1
        assert T[O.gf8_div(a, b), b] == a


def test_gfni_matrix_identity_and_model():
    assert O.gfni_matrix(1) == KATS["gfni_identity"]["qword"]
    assert O.gfni_matrix(0) == 0
    T = mul_table()
    v = np.arange(256, dtype=np.uint8)
    for c in range(256):
        q = O.gfni_matrix(c)
        rows = [(q >> (8 * r)) & 0xFF for r in range(8)]
        out = np.zeros(256, np.uint8)
        for i in range(8):
            par = np.unpackbits((v & rows[7 - i]).astype(np.uint8)[:, None], axis=1).sum(1) & 1
            out |= (par << i).astype(np.uint8)
        assert (out == T[c]).all(), c


# ---------------------------------------------------------------- GF((2^8)^2)

def _trace8(x: int) -> int:
    acc = 0
    for _ in range(8):
        acc ^= x
        x = int(mul_table()[x, x])
    return acc


def test_tower_l_is_smallest_trace_one():
    L = KATS["tower_l"]["value"]
    assert all(_trace8(c) == 0 for c in range(L))
    assert _trace8(L) == 1


def test_const_matrix_golden_and_exhaustive():
    for case in KATS["gf16_const_matrix"]["cases"]:
        assert O.gf16_const_matrix(case["c"]) == case["want"]
    T = mul_table()
    rng = np.random.default_rng(5)
    a = rng.integers(0, 65536, 64, dtype=np.uint16)
    for c in list(range(0, 65536, 97)) + [0xFFFF, 0x0100, 0x1234]:
        m00, m01, m10, m11 = O.gf16_const_matrix(c)
        ah, al = a >> 8, a & 0xFF
        hi = T[m00, ah] ^ T[m01, al]
        lo = T[m10, ah] ^ T[m11, al]
        got = (hi.astype(np.uint16) << 8) | lo
        want = tower_mul_poly(np.full_like(a, c), a)
        assert (got == want).all(), c


def test_tower_mul_matches_polynomial_definition():
    rng = np.random.default_rng(7)
    a = rng.integers(0, 65536, 4000, dtype=np.uint16)
    b = rng.integers(0, 65536, 4000, dtype=np.uint16)
    want = tower_mul_poly(a, b)
    got = np.array([O.gf16_mul(int(x), int(y)) for x, y in zip(a, b)], np.uint16)
    fast = np.array([O.gf16_mul_fast(int(x), int(y)) for x, y in zip(a, b)], np.uint16)
    assert (got == want).all()
    assert (fast == want).all()


def test_tower_inverse_every_element():
    a = np.arange(1, 65536, dtype=np.uint16)
    inv = np.array([O.gf16_inv(int(x)) for x in a], np.uint16)
    assert (tower_mul_poly(a, inv) == 1).all()


def test_tower_axioms_sampled():
    rng = np.random.default_rng(9)
    a, b, c = (rng.integers(0, 65536, 20000, dtype=np.uint16) for _ in range(3))
    assert (tower_mul_poly(tower_mul_poly(a, b), c) == tower_mul_poly(a, tower_mul_poly(b, c))).all()
    assert (tower_mul_poly(a, b ^ c) == (tower_mul_poly(a, b) ^ tower_mul_poly(a, c))).all()


# ---------------------------------------------------------------- matrices

@pytest.mark.parametrize("field,size", [(8, 20), (16, 86)])
def test_vandermonde_inverse_is_identity(field, size):
    v = O.vandermonde(field, size, size)
    rc, inv = O.invert(field, v)
    assert rc == 0
    prod = O.multiply(field, v, inv)
    assert (prod == np.eye(size, dtype=prod.dtype)).all()


def test_singular_matrix_rejected():
    rc, _ = O.invert(16, np.array([[1, 1], [1, 1]], np.uint16))
    assert rc == 14  # Error::SingularMatrix


def test_invert_matches_independent_gauss_jordan():
    T = mul_table()
    mul = lambda x, y: int(T[x, y])
    inv = lambda x: O.gf8_div(1, x)
    rng = np.random.default_rng(3)
    for n in (3, 6, 10):
        m = rng.integers(0, 256, (n, n)).astype(np.uint8)
        want = gauss_jordan_np(m, mul, inv)
        rc, got = O.invert(8, m)
        if want is None:
            assert rc == 14
        else:
            assert rc == 0 and (got == want).all()


@pytest.mark.parametrize("k,m", [(4, 2), (5, 3), (3, 4)])
def test_generator_systematic_and_mds(k, m):
    G = O.build_matrix(8, k, k + m)
    assert (G[:k] == np.eye(k, dtype=np.uint8)).all()
    for rows in itertools.combinations(range(k + m), k):
        rc, _ = O.invert(8, G[list(rows)])
        assert rc == 0, rows


def test_generator_matches_independent_numpy_construction():
    T = mul_table()
    k, n = 10, 14
    V = np.zeros((n, k), np.uint8)
    for r in range(n):
        p = 1
        for c in range(k):
            V[r, c] = p
            p = T[p, r]
    top_inv = gauss_jordan_np(V[:k], lambda x, y: int(T[x, y]), lambda x: O.gf8_div(1, x))
    G = np.zeros((n, k), np.uint8)
    for r in range(n):
        for c in range(k):
            acc = 0
            for i in range(k):
                acc ^= int(T[V[r, i], top_inv[i, c]])
            G[r, c] = acc
    assert (O.build_matrix(8, k, n) == G).all()


# ---------------------------------------------------------------- codec ops

def test_encode_matches_numpy_matmul():
    T = mul_table()
    k, m, L = 10, 4, 777
    G = O.build_matrix(8, k, k + m)
    data = np.stack([O.fill(100 + i, L) for i in range(k)])
    par = O.rs8_encode(G, k, data)
    want = np.zeros((m, L), np.uint8)
    for j in range(m):
        for i in range(k):
            want[j] ^= T[G[k + j, i], data[i]]
    assert (par == want).all()


@pytest.mark.parametrize("k,m", [(4, 2), (7, 13), (10, 10), (6, 4), (10, 4)])
@pytest.mark.parametrize("L", [15, 100, 1024])
def test_rs8_reconstruct_patterns_round_trip(k, m, L):
    # pattern list follows src/reedsolomon.rs:1048-1084
    G = O.build_matrix(8, k, k + m)
    data = np.stack([O.fill(7 * k + i, L) for i in range(k)])
    orig = np.concatenate([data, O.rs8_encode(G, k, data)])
    patterns = [[0], [k - 1], [k], list(range(min(m, k))), list(range(k, k + m)),
                [(e * 2) % (k + m) for e in range(m)]]
    for erased in patterns:
        shards = orig.copy()
        shards[erased] = 0
        present = [i not in erased for i in range(k + m)]
        assert O.rs8_reconstruct(G, k, m, shards, present) == 0
        assert (shards == orig).all(), erased


def test_rs8_data_only_leaves_parity_untouched():
    k, m, L = 6, 4, 128
    G = O.build_matrix(8, k, k + m)
    data = np.stack([O.fill(i, L) for i in range(k)])
    orig = np.concatenate([data, O.rs8_encode(G, k, data)])
    shards = orig.copy()
    shards[[2, k + 1]] = 0
    present = [i not in (2, k + 1) for i in range(k + m)]
    assert O.rs8_reconstruct(G, k, m, shards, present, data_only=True) == 0
    assert (shards[2] == orig[2]).all()
    assert (shards[k + 1] == 0).all()


def test_rs8_too_few_present():
    k, m, L = 4, 2, 16
    G = O.build_matrix(8, k, k + m)
    shards = np.zeros((k + m, L), np.uint8)
    assert O.rs8_reconstruct(G, k, m, shards, [True] * 3 + [False] * 3) == 10


def test_rs16_golden_wire_vector():
    w = KATS["rs16_wire_vector"]
    k, m = w["k"], w["m"]
    G = O.build_matrix(16, k, k + m)
    data = np.array(w["data"], np.uint8)
    par = O.rs16_apply(G[k:], list(data), data.shape[1])
    for idx, want in w["parity"].items():
        assert par[int(idx) - k].tolist() == want


@pytest.mark.parametrize("k,m", [(4, 2), (7, 13), (17, 33)])
@pytest.mark.parametrize("L", [2, 34, 200, 2048])
def test_rs16_apply_matches_symbolwise(k, m, L):
    G = O.build_matrix(16, k, k + m)
    data = [O.fill(31 + i, L) for i in range(k)]
    assert (O.rs16_apply(G[k:], data, L) == O.rs16_symbolwise(G[k:], data, L)).all()


def test_rs16_reconstruct_round_trip_past_256():
    k, m, L = 100, 200, 34
    G = O.build_matrix(16, k, k + m)
    data = np.stack([O.fill(900 + i, L) for i in range(k)])
    orig = np.concatenate([data, O.rs16_apply(G[k:], list(data), L)])
    for erased in ([0], list(range(k, k + m)), [(e * 3) % (k + m) for e in range(m)]):
        shards = orig.copy()
        shards[erased] = 0
        present = [i not in erased for i in range(k + m)]
        assert O.rs16_reconstruct(G, k, m, shards, present) == 0
        assert (shards == orig).all()


def test_reference_strip_mining_diverges_past_subfield():
    """Finding about the reference: ReedSolomon16's matrix route strip-mines
    shards longer than 16 KiB and re-splits every strip at its midpoint
    (src/reedsolomon16.rs:364-385 -> :143-168).  That equals the split-plane
    wire only while every coefficient lies in the GF(2^8) subfield (parity rows
    below index 256).  This build follows the wire (symbol-wise) definition,
    which the reference's FFT route and every shard <= 16 KiB also follow."""
    k, m, L = 200, 100, 32768
    G = O.build_matrix(16, k, k + m)
    data = [O.fill(7 + i, L) for i in range(k)]
    sub = G[k:k + 2]          # points < 256: subfield rows
    full = G[k + 96:k + 98]   # points >= 256: full tower rows
    assert int(sub.max()) < 256 and int(full.max()) >= 256
    assert (O.rs16_apply(sub, data, L, strip_bytes=16384) == O.rs16_symbolwise(sub, data, L)).all()
    assert (O.rs16_apply(full, data, L) == O.rs16_symbolwise(full, data, L)).all()
    assert not (O.rs16_apply(full, data, L, strip_bytes=16384)
                == O.rs16_symbolwise(full, data, L)).all()


def test_fill_is_splitmix64_stream():
    def mix(z):
        M = (1 << 64) - 1
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    seed = 0x7A9E5EED00000001
    want = b"".join(mix((seed + (i + 1) * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)).to_bytes(8, "little")
                    for i in range(5))
    assert bytes(O.fill(seed, 37)) == want[:37]
