/*
 * CPU oracle for the Reed-Solomon bulk coding path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference crate's scalar arithmetic (tape-reed-solomon 0.2.2).  It exists
 * so tests, the smoke check and the CPU-baseline leg of the benchmark have an
 * independent checker.  The shipped library never links, loads or calls it.
 *
 * Every function cites the reference file:line it follows.  Pinning (what
 * checks each function against something other than itself) is listed in the
 * header comment of oracle.py and in tests/test_oracle.py.
 *
 * Field elements of both fields are held as uint16_t in the matrix helpers;
 * `field` selects GF(2^8) (8) or the GF((2^8)^2) tower (16).
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* GF(2^8), poly 0x11d, generator 2            src/galois.rs:3-111     */
/* ------------------------------------------------------------------ */

static uint8_t LOG8[256];
static uint8_t EXP8[510];
static uint8_t MUL8[256][256];
static int tables_ready = 0;

/* GF((2^8)^2) log/antilog tables, used only to speed up the oracle's big
 * matrix work.  They are derived from orc_gf16_mul (the restatement of
 * galois16::mul) and checked against it by the tests. */
static uint16_t LOG16[65536];
static uint16_t EXP16[2 * 65535];
static int tables16_ready = 0;

/* src/galois.rs:9-22 (gen_log_table) and :24-35 (gen_exp_table) */
static void build_tables8(void) {
    if (tables_ready) return;
    memset(LOG8, 0, sizeof LOG8);
    unsigned b = 1;
    for (unsigned log = 0; log < 255; ++log) {
        LOG8[b] = (uint8_t)log;
        b <<= 1;
        if (b >= 256) b = (b - 256) ^ 0x1d;
    }
    memset(EXP8, 0, sizeof EXP8);
    for (unsigned i = 1; i < 256; ++i) {
        unsigned l = LOG8[i];
        EXP8[l] = (uint8_t)i;
        EXP8[l + 255] = (uint8_t)i;
    }
    /* src/galois.rs:37-57 (gen_mul_table) */
    for (unsigned a = 0; a < 256; ++a)
        for (unsigned c = 0; c < 256; ++c)
            MUL8[a][c] = (a == 0 || c == 0) ? 0 : EXP8[LOG8[a] + LOG8[c]];
    tables_ready = 1;
}

void orc_gf8_tables(uint8_t *log_out, uint8_t *exp_out) {
    build_tables8();
    memcpy(log_out, LOG8, 256);
    memcpy(exp_out, EXP8, 510);
}

/* src/galois.rs:74-76 */
uint8_t orc_gf8_mul(uint8_t a, uint8_t b) {
    build_tables8();
    return MUL8[a][b];
}

/* src/galois.rs:80-94; returns 0 for b == 0 where the reference panics */
uint8_t orc_gf8_div(uint8_t a, uint8_t b) {
    build_tables8();
    if (a == 0 || b == 0) return 0;
    int l = (int)LOG8[a] - (int)LOG8[b];
    if (l < 0) l += 255;
    return EXP8[l];
}

/* src/galois.rs:98-111 */
uint8_t orc_gf8_exp(uint8_t a, size_t n) {
    build_tables8();
    if (n == 0) return 1;
    if (a == 0) return 0;
    size_t l = (size_t)LOG8[a] * n;
    while (l >= 255) l -= 255;
    return EXP8[l];
}

/* src/gf/gfni.rs:47-64: A.byte[r].bit[j] = bit_(7-r)(c * 2^j) */
uint64_t orc_gfni_matrix(uint8_t c) {
    build_tables8();
    uint64_t q = 0;
    for (int r = 0; r < 8; ++r) {
        uint8_t byte = 0;
        for (int j = 0; j < 8; ++j)
            if ((MUL8[c][1u << j] >> (7 - r)) & 1) byte |= (uint8_t)(1u << j);
        q |= (uint64_t)byte << (8 * r);
    }
    return q;
}

/* ------------------------------------------------------------------ */
/* GF((2^8)^2) tower, y^2 + y + 32             src/galois16.rs:16-125  */
/* ------------------------------------------------------------------ */

#define TOWER_L 32u

/* src/galois16.rs:45-59 */
uint16_t orc_gf16_mul(uint16_t a, uint16_t b) {
    build_tables8();
    uint8_t ah = a >> 8, al = a & 0xff, bh = b >> 8, bl = b & 0xff;
    uint8_t p2 = MUL8[ah][bh];
    uint8_t p1 = MUL8[ah][bl] ^ MUL8[al][bh];
    uint8_t p0 = MUL8[al][bl];
    uint8_t hi = p1 ^ p2;
    uint8_t lo = p0 ^ MUL8[p2][TOWER_L];
    return (uint16_t)((hi << 8) | lo);
}

/* src/galois16.rs:110-125 */
uint16_t orc_gf16_inv(uint16_t a) {
    build_tables8();
    if (a == 0) return 0;
    uint8_t ah = a >> 8, al = a & 0xff;
    uint16_t conj = (uint16_t)((ah << 8) | (uint8_t)(ah ^ al));
    uint16_t norm = orc_gf16_mul(a, conj);
    uint8_t ninv = orc_gf8_div(1, norm & 0xff);
    return (uint16_t)((MUL8[conj >> 8][ninv] << 8) | MUL8[conj & 0xff][ninv]);
}

/* src/galois16.rs:76-84: [m00, m01, m10, m11] = [h^l, h, h*L, l] */
void orc_gf16_const_matrix(uint16_t c, uint8_t *out4) {
    build_tables8();
    uint8_t h = c >> 8, l = c & 0xff;
    out4[0] = h ^ l;
    out4[1] = h;
    out4[2] = MUL8[h][TOWER_L];
    out4[3] = l;
}

/* Log tables for the tower: pick the smallest element of full order 65535
 * (the same search as galois16.rs:202-209) and walk its powers. */
static void build_tables16(void) {
    if (tables16_ready) return;
    build_tables8();
    static const unsigned primes[4] = {3, 5, 17, 257};
    uint16_t g = 0;
    for (unsigned c = 2; c < 65536 && !g; ++c) {
        int full = 1;
        for (int p = 0; p < 4 && full; ++p) {
            /* c^(65535/p) by square-and-multiply */
            unsigned e = 65535 / primes[p];
            uint16_t r = 1, f = (uint16_t)c;
            while (e) {
                if (e & 1) r = orc_gf16_mul(r, f);
                f = orc_gf16_mul(f, f);
                e >>= 1;
            }
            if (r == 1) full = 0;
        }
        if (full) g = (uint16_t)c;
    }
    uint16_t x = 1;
    for (unsigned i = 0; i < 65535; ++i) {
        EXP16[i] = x;
        EXP16[i + 65535] = x;
        LOG16[x] = (uint16_t)i;
        x = orc_gf16_mul(x, g);
    }
    LOG16[0] = 0;
    tables16_ready = 1;
}

static inline uint16_t fmul16(uint16_t a, uint16_t b) {
    if (a == 0 || b == 0) return 0;
    return EXP16[(unsigned)LOG16[a] + LOG16[b]];
}

/* Element-wise helpers for the generic matrix code (src/field.rs:35-88). */
static inline uint16_t fmul(int field, uint16_t a, uint16_t b) {
    return field == 8 ? MUL8[a & 0xff][b & 0xff] : fmul16(a, b);
}
static inline uint16_t finv(int field, uint16_t a) {
    if (field == 8) return orc_gf8_div(1, (uint8_t)a);
    return EXP16[(65535 - LOG16[a]) % 65535];
}

uint16_t orc_gf16_mul_fast(uint16_t a, uint16_t b) {
    build_tables16();
    return fmul16(a, b);
}

/* ------------------------------------------------------------------ */
/* Dense matrices                               src/matrix.rs:19-196   */
/* ------------------------------------------------------------------ */

static void prep(int field) {
    build_tables8();
    if (field == 16) build_tables16();
}

/* src/matrix.rs:183-196; evaluation point of row r is element r
 * (src/field.rs:56-61, :84-87) */
void orc_vandermonde(int field, uint16_t *out, size_t rows, size_t cols) {
    prep(field);
    for (size_t r = 0; r < rows; ++r) {
        uint16_t point = (uint16_t)r, power = 1;
        for (size_t c = 0; c < cols; ++c) {
            out[r * cols + c] = power;
            power = fmul(field, power, point);
        }
    }
}

/* src/matrix.rs:51-77 */
void orc_multiply(int field, const uint16_t *a, const uint16_t *b, uint16_t *out,
                  size_t rows, size_t inner, size_t cols) {
    prep(field);
    memset(out, 0, rows * cols * sizeof(uint16_t));
    for (size_t r = 0; r < rows; ++r)
        for (size_t i = 0; i < inner; ++i) {
            uint16_t s = a[r * inner + i];
            if (!s) continue;
            const uint16_t *brow = b + i * cols;
            uint16_t *orow = out + r * cols;
            if (field == 8) {
                const uint8_t *mrow = MUL8[s];
                for (size_t c = 0; c < cols; ++c) orow[c] ^= mrow[brow[c] & 0xff];
            } else {
                unsigned ls = LOG16[s];
                for (size_t c = 0; c < cols; ++c)
                    if (brow[c]) orow[c] ^= EXP16[ls + LOG16[brow[c]]];
            }
        }
}

/* src/matrix.rs:123-181: augment with I, Gauss-Jordan with the pivot taken as
 * the first nonzero entry at or below the diagonal, then read the right half.
 * Returns 0, or 14 (Error::SingularMatrix, src/errors.rs:20). */
int orc_invert(int field, const uint16_t *in, uint16_t *out, size_t n) {
    prep(field);
    size_t w = 2 * n;
    uint16_t *a = (uint16_t *)calloc(n * w, sizeof(uint16_t));
    if (!a) return -2;
    for (size_t r = 0; r < n; ++r) {
        memcpy(a + r * w, in + r * n, n * sizeof(uint16_t));
        a[r * w + n + r] = 1;
    }
    int rc = 0;
    for (size_t r = 0; r < n && !rc; ++r) {
        if (a[r * w + r] == 0) {
            for (size_t rb = r + 1; rb < n; ++rb)
                if (a[rb * w + r]) {
                    for (size_t c = 0; c < w; ++c) {
                        uint16_t t = a[r * w + c];
                        a[r * w + c] = a[rb * w + c];
                        a[rb * w + c] = t;
                    }
                    break;
                }
        }
        if (a[r * w + r] == 0) { rc = 14; break; }
        if (a[r * w + r] != 1) {
            uint16_t s = finv(field, a[r * w + r]);
            for (size_t c = 0; c < w; ++c) a[r * w + c] = fmul(field, s, a[r * w + c]);
        }
        for (size_t rb = r + 1; rb < n; ++rb) {
            uint16_t s = a[rb * w + r];
            if (!s) continue;
            for (size_t c = 0; c < w; ++c) a[rb * w + c] ^= fmul(field, s, a[r * w + c]);
        }
    }
    if (!rc) {
        for (size_t d = 0; d < n; ++d)
            for (size_t ra = 0; ra < d; ++ra) {
                uint16_t s = a[ra * w + d];
                if (!s) continue;
                for (size_t c = 0; c < w; ++c) a[ra * w + c] ^= fmul(field, s, a[d * w + c]);
            }
        for (size_t r = 0; r < n; ++r) memcpy(out + r * n, a + r * w + n, n * sizeof(uint16_t));
    }
    free(a);
    return rc;
}

/* src/reedsolomon.rs:154-159 and src/reedsolomon16.rs:260-265:
 * G = V(n, k) * inv(V[0..k]); rows k..n are the parity rows. */
int orc_build_matrix(int field, size_t k, size_t n, uint16_t *out) {
    prep(field);
    uint16_t *v = (uint16_t *)malloc(n * k * sizeof(uint16_t));
    uint16_t *top_inv = (uint16_t *)malloc(k * k * sizeof(uint16_t));
    if (!v || !top_inv) { free(v); free(top_inv); return -2; }
    orc_vandermonde(field, v, n, k);
    int rc = orc_invert(field, v, top_inv, k); /* top k rows of v, row-major */
    if (!rc) orc_multiply(field, v, top_inv, out, n, k, k);
    free(v);
    free(top_inv);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Slice kernels and the GF(2^8) row apply                             */
/* ------------------------------------------------------------------ */

/* src/gf/scalar.rs:7-23 */
void orc_mul_slice(uint8_t *out, const uint8_t *in, size_t len, uint8_t c, int xor_into) {
    build_tables8();
    const uint8_t *row = MUL8[c];
    if (xor_into)
        for (size_t i = 0; i < len; ++i) out[i] ^= row[in[i]];
    else
        for (size_t i = 0; i < len; ++i) out[i] = row[in[i]];
}

/* RowTables::apply semantics (src/gf/tables.rs:162-219, the per-coefficient
 * fallback at :202-218): every output overwritten with sum_i rows[j][i]*in[i]. */
void orc_rs8_apply(const uint8_t *rows, size_t nout, size_t k,
                   const uint8_t *const *in, uint8_t *const *out, size_t len) {
    for (size_t j = 0; j < nout; ++j)
        for (size_t i = 0; i < k; ++i)
            orc_mul_slice(out[j], in[i], len, rows[j * k + i], i != 0);
}

/* ------------------------------------------------------------------ */
/* GF(2^8) reconstruct                                                 */
/* src/reedsolomon.rs:537-585 (plan) and :829-891 (apply order)        */
/* ------------------------------------------------------------------ */

/* G is the n x k generator (uint8), shards holds n buffers of len bytes,
 * present[i] != 0 marks a valid shard.  Missing shards are overwritten.
 * Returns 0, 10 (TooFewShardsPresent) or 14 (SingularMatrix). */
int orc_rs8_reconstruct(const uint8_t *G, size_t k, size_t m, uint8_t *const *shards,
                        const uint8_t *present, size_t len, int data_only) {
    build_tables8();
    size_t n = k + m, npresent = 0;
    for (size_t i = 0; i < n; ++i) npresent += present[i] != 0;
    if (npresent == n) return 0;
    if (npresent < k) return 10;
    size_t *valid = (size_t *)malloc(k * sizeof(size_t));
    size_t *mdata = (size_t *)malloc(n * sizeof(size_t));
    size_t *mpar = (size_t *)malloc(n * sizeof(size_t));
    uint16_t *sub = (uint16_t *)malloc(k * k * sizeof(uint16_t));
    uint16_t *inv = (uint16_t *)malloc(k * k * sizeof(uint16_t));
    size_t nv = 0, nmd = 0, nmp = 0;
    for (size_t i = 0; i < n; ++i) {
        if (present[i]) {
            if (nv < k) valid[nv++] = i;
        } else if (i < k) {
            mdata[nmd++] = i;
        } else {
            mpar[nmp++] = i;
        }
    }
    for (size_t r = 0; r < k; ++r)
        for (size_t c = 0; c < k; ++c) sub[r * k + c] = G[valid[r] * k + c];
    int rc = orc_invert(8, sub, inv, k);
    if (!rc) {
        const uint8_t **surv = (const uint8_t **)malloc(k * sizeof(void *));
        for (size_t r = 0; r < k; ++r) surv[r] = shards[valid[r]];
        uint8_t *row = (uint8_t *)malloc(k);
        for (size_t d = 0; d < nmd; ++d) {
            for (size_t c = 0; c < k; ++c) row[c] = (uint8_t)inv[mdata[d] * k + c];
            uint8_t *o = shards[mdata[d]];
            orc_rs8_apply(row, 1, k, surv, &o, len);
        }
        if (!data_only && nmp) {
            const uint8_t **all = (const uint8_t **)malloc(k * sizeof(void *));
            for (size_t c = 0; c < k; ++c) all[c] = shards[c];
            for (size_t p = 0; p < nmp; ++p) {
                for (size_t c = 0; c < k; ++c) row[c] = G[mpar[p] * k + c];
                uint8_t *o = shards[mpar[p]];
                orc_rs8_apply(row, 1, k, all, &o, len);
            }
            free(all);
        }
        free(row);
        free(surv);
    }
    free(valid); free(mdata); free(mpar); free(sub); free(inv);
    return rc;
}

/* ------------------------------------------------------------------ */
/* GF((2^8)^2) row apply over split planes                             */
/* src/reedsolomon16.rs:96-168 (dense 2m x 2k expansion, plane order)  */
/* ------------------------------------------------------------------ */

/* RowTables16::apply (src/reedsolomon16.rs:143-168): each shard of `len`
 * bytes is split at len/2 into [low | high] planes; outputs overwritten. */
void orc_rs16_apply(const uint16_t *rows, size_t nout, size_t k,
                    const uint8_t *const *in, uint8_t *const *out, size_t len) {
    build_tables8();
    size_t pl = len / 2;
    for (size_t j = 0; j < nout; ++j) {
        uint8_t *olo = out[j], *ohi = out[j] + pl;
        memset(olo, 0, pl);
        memset(ohi, 0, pl);
        for (size_t i = 0; i < k; ++i) {
            uint8_t mm[4];
            orc_gf16_const_matrix(rows[j * k + i], mm);
            const uint8_t *ilo = in[i], *ihi = in[i] + pl;
            /* out_high = m00*a_high + m01*a_low ; out_low = m10*a_high + m11*a_low */
            const uint8_t *t00 = MUL8[mm[0]], *t01 = MUL8[mm[1]];
            const uint8_t *t10 = MUL8[mm[2]], *t11 = MUL8[mm[3]];
            for (size_t s = 0; s < pl; ++s) {
                ohi[s] ^= t00[ihi[s]] ^ t01[ilo[s]];
                olo[s] ^= t10[ihi[s]] ^ t11[ilo[s]];
            }
        }
    }
}

/* Symbol-wise parity through the scalar tower multiply
 * (src/reedsolomon16.rs:919-936): the wire definition of the code. */
void orc_rs16_symbolwise(const uint16_t *rows, size_t nout, size_t k,
                         const uint8_t *const *in, uint8_t *const *out, size_t len) {
    size_t sc = len / 2;
    for (size_t j = 0; j < nout; ++j)
        for (size_t s = 0; s < sc; ++s) {
            uint16_t acc = 0;
            for (size_t i = 0; i < k; ++i) {
                uint16_t sym = (uint16_t)((in[i][sc + s] << 8) | in[i][s]);
                acc ^= orc_gf16_mul(rows[j * k + i], sym);
            }
            out[j][s] = acc & 0xff;
            out[j][sc + s] = acc >> 8;
        }
}

/* ReedSolomon16's matrix-route encode_into (src/reedsolomon16.rs:351-386):
 * shards longer than strip_bytes are applied one byte strip at a time and
 * each strip goes through RowTables16::apply.  strip_bytes == 0 applies the
 * whole shard at once. */
void orc_rs16_apply_strips(const uint16_t *rows, size_t nout, size_t k,
                           const uint8_t *const *in, uint8_t *const *out, size_t len,
                           size_t strip_bytes) {
    if (strip_bytes == 0 || len <= strip_bytes) {
        orc_rs16_apply(rows, nout, k, in, out, len);
        return;
    }
    const uint8_t **si = (const uint8_t **)malloc(k * sizeof(void *));
    uint8_t **so = (uint8_t **)malloc(nout * sizeof(void *));
    for (size_t off = 0; off < len;) {
        size_t w = len - off < strip_bytes ? len - off : strip_bytes;
        for (size_t i = 0; i < k; ++i) si[i] = in[i] + off;
        for (size_t j = 0; j < nout; ++j) so[j] = out[j] + off;
        orc_rs16_apply(rows, nout, k, si, so, w);
        off += w;
    }
    free(si);
    free(so);
}

/* GF(2^16) reconstruct, src/reedsolomon16.rs:577-612 (plan) and :806-889
 * (apply order).  strip_bytes as in orc_rs16_apply_strips. */
int orc_rs16_reconstruct(const uint16_t *G, size_t k, size_t m, uint8_t *const *shards,
                         const uint8_t *present, size_t len, int data_only,
                         size_t strip_bytes) {
    prep(16);
    size_t n = k + m, npresent = 0;
    for (size_t i = 0; i < n; ++i) npresent += present[i] != 0;
    if (npresent == n) return 0;
    if (npresent < k) return 10;
    size_t *valid = (size_t *)malloc(k * sizeof(size_t));
    size_t *mdata = (size_t *)malloc(n * sizeof(size_t));
    size_t *mpar = (size_t *)malloc(n * sizeof(size_t));
    uint16_t *sub = (uint16_t *)malloc(k * k * sizeof(uint16_t));
    uint16_t *inv = (uint16_t *)malloc(k * k * sizeof(uint16_t));
    size_t nv = 0, nmd = 0, nmp = 0;
    for (size_t i = 0; i < n; ++i) {
        if (present[i]) {
            if (nv < k) valid[nv++] = i;
        } else if (i < k) {
            mdata[nmd++] = i;
        } else {
            mpar[nmp++] = i;
        }
    }
    for (size_t r = 0; r < k; ++r)
        memcpy(sub + r * k, G + valid[r] * k, k * sizeof(uint16_t));
    int rc = orc_invert(16, sub, inv, k);
    if (!rc) {
        uint16_t *drows = (uint16_t *)malloc((nmd + 1) * k * sizeof(uint16_t));
        uint16_t *prows = (uint16_t *)malloc((nmp + 1) * k * sizeof(uint16_t));
        for (size_t d = 0; d < nmd; ++d) memcpy(drows + d * k, inv + mdata[d] * k, k * 2);
        for (size_t p = 0; p < nmp; ++p) memcpy(prows + p * k, G + mpar[p] * k, k * 2);
        const uint8_t **surv = (const uint8_t **)malloc(k * sizeof(void *));
        const uint8_t **all = (const uint8_t **)malloc(k * sizeof(void *));
        uint8_t **od = (uint8_t **)malloc((nmd + 1) * sizeof(void *));
        uint8_t **op = (uint8_t **)malloc((nmp + 1) * sizeof(void *));
        size_t sw = strip_bytes ? strip_bytes : len;
        for (size_t off = 0; off < len;) {
            size_t w = len - off < sw ? len - off : sw;
            for (size_t r = 0; r < k; ++r) surv[r] = shards[valid[r]] + off;
            for (size_t d = 0; d < nmd; ++d) od[d] = shards[mdata[d]] + off;
            if (nmd) orc_rs16_apply(drows, nmd, k, surv, od, w);
            if (!data_only && nmp) {
                for (size_t c = 0; c < k; ++c) all[c] = shards[c] + off;
                for (size_t p = 0; p < nmp; ++p) op[p] = shards[mpar[p]] + off;
                orc_rs16_apply(prows, nmp, k, all, op, w);
            }
            off += w;
        }
        free(drows); free(prows); free(surv); free(all); free(od); free(op);
    }
    free(valid); free(mdata); free(mpar); free(sub); free(inv);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Seeded data: splitmix64 as a counter-based stream.                  */
/* Word i of a stream is mix(seed + (i + 1) * 0x9E3779B97F4A7C15); its  */
/* eight bytes are emitted little-endian.                              */
/* ------------------------------------------------------------------ */

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_fill(uint64_t seed, uint8_t *buf, size_t len) {
    size_t words = len / 8;
    for (size_t i = 0; i < words; ++i) {
        uint64_t v = mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
        memcpy(buf + 8 * i, &v, 8);
    }
    if (len % 8) {
        uint64_t v = mix64(seed + (words + 1) * 0x9E3779B97F4A7C15ull);
        memcpy(buf + 8 * words, &v, len % 8);
    }
}
