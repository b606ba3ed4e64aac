/* CPU baseline for bench.py: the reference's x86 GFNI / AVX-512 algorithms
 * restated in C.  TEST INFRASTRUCTURE, like oracle.c: only the tests and the
 * bench's CPU-baseline leg load it, and it is never linked by the library.
 *
 * The reference ships no C; these functions follow the algorithms of
 *   src/gf/x86_fused.rs:13-130   (dot_prod_gfni: fused multi-output apply,
 *                                 one gf2p8affine per coefficient, tiles of
 *                                 at most 6 outputs, 128-byte blocks)
 *   src/gf/fft_x86.rs:22-165     (encode_generated: the straight-line LCH
 *                                 programs over byte strips)
 *   src/reedsolomon.rs:829-891   (reconstruct: missing data from the first k
 *                                 present shards, then missing parity from
 *                                 the rebuilt data)
 * with the affine matrices of src/gf/gfni.rs:41-64 (orc_gfni_matrix in
 * oracle.c).  Every function is checked byte for byte against the scalar
 * oracle (tests/test_simd_ref.py) before the bench times it.
 */
#include <immintrin.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

uint64_t orc_gfni_matrix(uint8_t c);
int orc_invert(int field, const uint16_t *in, uint16_t *out, size_t n);

#define SIMD_FN __attribute__((target("gfni,avx512f,avx512bw,avx512vl")))

int orc_simd_available(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("gfni") && __builtin_cpu_supports("avx512f") &&
           __builtin_cpu_supports("avx512bw") && __builtin_cpu_supports("avx512vl");
}

static inline __mmask64 tail_mask(size_t rem) {
    return rem >= 64 ? ~(__mmask64)0 : (((__mmask64)1 << rem) - 1);
}

/* One tile of NT outputs (NT a constant after inlining): mat[i * NT + j] is
 * the affine matrix of output j's coefficient for input i. */
static inline __attribute__((always_inline)) SIMD_FN void
apply_tile(const int NT, const uint64_t *mat, size_t k, const uint8_t *const *in,
           uint8_t *const *out, size_t len) {
    size_t off = 0;
    for (; off + 128 <= len; off += 128) {
        __m512i acc0[6], acc1[6];
        for (int j = 0; j < NT; ++j) acc0[j] = acc1[j] = _mm512_setzero_si512();
        for (size_t i = 0; i < k; ++i) {
            const __m512i x0 = _mm512_loadu_si512((const void *)(in[i] + off));
            const __m512i x1 = _mm512_loadu_si512((const void *)(in[i] + off + 64));
            for (int j = 0; j < NT; ++j) {
                const __m512i a = _mm512_set1_epi64((long long)mat[i * NT + j]);
                acc0[j] = _mm512_xor_si512(acc0[j], _mm512_gf2p8affine_epi64_epi8(x0, a, 0));
                acc1[j] = _mm512_xor_si512(acc1[j], _mm512_gf2p8affine_epi64_epi8(x1, a, 0));
            }
        }
        for (int j = 0; j < NT; ++j) {
            _mm512_storeu_si512((void *)(out[j] + off), acc0[j]);
            _mm512_storeu_si512((void *)(out[j] + off + 64), acc1[j]);
        }
    }
    for (; off < len; off += 64) {  /* the last < 128 bytes, masked */
        const __mmask64 mk = tail_mask(len - off);
        __m512i acc[6];
        for (int j = 0; j < NT; ++j) acc[j] = _mm512_setzero_si512();
        for (size_t i = 0; i < k; ++i) {
            const __m512i x = _mm512_maskz_loadu_epi8(mk, (const void *)(in[i] + off));
            for (int j = 0; j < NT; ++j) {
                const __m512i a = _mm512_set1_epi64((long long)mat[i * NT + j]);
                acc[j] = _mm512_xor_si512(acc[j], _mm512_gf2p8affine_epi64_epi8(x, a, 0));
            }
        }
        for (int j = 0; j < NT; ++j) _mm512_mask_storeu_epi8((void *)(out[j] + off), mk, acc[j]);
    }
}

/* RowTables::apply (src/gf/tables.rs:162-219) on the fused GFNI kernel:
 * out[j] = sum_i rows[j * k + i] * in[i], every output overwritten. */
SIMD_FN void orc_simd_apply(const uint8_t *rows, size_t nout, size_t k,
                            const uint8_t *const *in, uint8_t *const *out, size_t len) {
    uint64_t *mat = (uint64_t *)malloc(k * 6 * sizeof(uint64_t));
    for (size_t j0 = 0; j0 < nout; j0 += 6) {
        const int nt = (int)(nout - j0 < 6 ? nout - j0 : 6);
        for (size_t i = 0; i < k; ++i)
            for (int j = 0; j < nt; ++j) mat[i * nt + j] = orc_gfni_matrix(rows[(j0 + j) * k + i]);
        switch (nt) {
            case 1: apply_tile(1, mat, k, in, out + j0, len); break;
            case 2: apply_tile(2, mat, k, in, out + j0, len); break;
            case 3: apply_tile(3, mat, k, in, out + j0, len); break;
            case 4: apply_tile(4, mat, k, in, out + j0, len); break;
            case 5: apply_tile(5, mat, k, in, out + j0, len); break;
            default: apply_tile(6, mat, k, in, out + j0, len); break;
        }
    }
    free(mat);
}

/* An LCH FFT encode program (the op lists of src/fft_programs.rs, as
 * reed-solomon_amd/fftgen.py regenerates them) run over the shards: ops are
 * int32 quadruples (kind, dst, src, c) with kinds 0 xor (dst ^= src), 1 copy,
 * 2 multiply-xor (dst ^= c * src), 3 multiply-copy, 4 zero; registers 0..k-1
 * start as the data shards and parity j is register parity_regs[j] at the
 * end.  The reference runs the programs as straight-line code over 32 / 64
 * byte strips; here each register holds a 256-byte strip (four vectors), so
 * the per-op dispatch is amortised over four affine or XOR instructions. */
SIMD_FN void orc_simd_fft_encode(const int32_t *ops, size_t nops, size_t nreg, size_t k,
                                 size_t m, const int32_t *parity_regs,
                                 const uint8_t *const *in, uint8_t *const *out, size_t len) {
    enum { V = 4 };
    __m512i *R = (__m512i *)_mm_malloc(nreg * V * sizeof(__m512i), 64);
    uint64_t *mat = (uint64_t *)malloc((nops ? nops : 1) * sizeof(uint64_t));
    for (size_t o = 0; o < nops; ++o)
        mat[o] = (ops[4 * o] == 2 || ops[4 * o] == 3) ? orc_gfni_matrix((uint8_t)ops[4 * o + 3]) : 0;
    for (size_t off = 0; off < len; off += 64 * V) {
        __mmask64 mk[V];
        for (int v = 0; v < V; ++v)
            mk[v] = off + 64 * v < len ? tail_mask(len - off - 64 * v) : 0;
        for (size_t i = 0; i < k; ++i)
            for (int v = 0; v < V; ++v)
                R[i * V + v] = _mm512_maskz_loadu_epi8(mk[v], (const void *)(in[i] + off + 64 * v));
        for (size_t o = 0; o < nops; ++o) {
            const int32_t kind = ops[4 * o];
            __m512i *d = R + (size_t)ops[4 * o + 1] * V;
            const __m512i *s = R + (size_t)ops[4 * o + 2] * V;
            switch (kind) {
                case 0:
                    for (int v = 0; v < V; ++v) d[v] = _mm512_xor_si512(d[v], s[v]);
                    break;
                case 1:
                    for (int v = 0; v < V; ++v) d[v] = s[v];
                    break;
                case 2: {
                    const __m512i a = _mm512_set1_epi64((long long)mat[o]);
                    for (int v = 0; v < V; ++v)
                        d[v] = _mm512_xor_si512(d[v], _mm512_gf2p8affine_epi64_epi8(s[v], a, 0));
                    break;
                }
                case 3: {
                    const __m512i a = _mm512_set1_epi64((long long)mat[o]);
                    for (int v = 0; v < V; ++v) d[v] = _mm512_gf2p8affine_epi64_epi8(s[v], a, 0);
                    break;
                }
                default:
                    for (int v = 0; v < V; ++v) d[v] = _mm512_setzero_si512();
                    break;
            }
        }
        for (size_t j = 0; j < m; ++j)
            for (int v = 0; v < V; ++v)
                _mm512_mask_storeu_epi8((void *)(out[j] + off + 64 * v), mk[v],
                                        R[(size_t)parity_regs[j] * V + v]);
    }
    free(mat);
    _mm_free(R);
}

/* ReedSolomon::reconstruct (src/reedsolomon.rs:829-891) with the fused
 * apply: the first k present shards, the inverse of their generator rows,
 * one apply for every missing data shard, then one apply of the generator's
 * parity rows for every missing parity shard (skipped when data_only).
 * Returns 0, 10 (TooFewShardsPresent) or 14 (SingularMatrix). */
SIMD_FN int orc_simd_rs8_reconstruct(const uint8_t *G, size_t k, size_t m, uint8_t *const *shards,
                                     const uint8_t *present, size_t len, int data_only) {
    const size_t n = k + m;
    size_t npresent = 0;
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
        uint8_t *rows = (uint8_t *)malloc((n ? n : 1) * k);
        const uint8_t **src = (const uint8_t **)malloc(k * sizeof(void *));
        uint8_t **dst = (uint8_t **)malloc((n ? n : 1) * sizeof(void *));
        if (nmd) {
            for (size_t r = 0; r < k; ++r) src[r] = shards[valid[r]];
            for (size_t d = 0; d < nmd; ++d) {
                for (size_t c = 0; c < k; ++c) rows[d * k + c] = (uint8_t)inv[mdata[d] * k + c];
                dst[d] = shards[mdata[d]];
            }
            orc_simd_apply(rows, nmd, k, src, dst, len);
        }
        if (!data_only && nmp) {
            for (size_t c = 0; c < k; ++c) src[c] = shards[c];
            for (size_t p = 0; p < nmp; ++p) {
                memcpy(rows + p * k, G + mpar[p] * k, k);
                dst[p] = shards[mpar[p]];
            }
            orc_simd_apply(rows, nmp, k, src, dst, len);
        }
        free(rows);
        free(src);
        free(dst);
    }
    free(valid); free(mdata); free(mpar); free(sub); free(inv);
    return rc;
}
