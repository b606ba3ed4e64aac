/*
 * trs_gpu.h -- C ABI of the MI355X (gfx950) Reed-Solomon bulk coding library.
 *
 * Drop-in boundary for tape-reed-solomon 0.2.2's bulk coding path.  Each entry
 * point names the reference method it replaces (file:line in the reference
 * crate).  Conventions shared by every call:
 *
 *  - Shard buffers are DEVICE pointers (HBM), except in the *_host entry
 *    points, which take the same arguments over HOST memory (the reference's
 *    slices).  A shard set is a pointer array plus a parallel array of buffer
 *    lengths in bytes, so the reference's length checks (and their error
 *    order) are kept.
 *  - `stream` is a hipStream_t passed as void*; NULL means the default stream.
 *    Work is enqueued on that stream and the call returns; results are ready
 *    once the stream reaches that point.  trs_*_verify waits for its answer.
 *  - Return value: 0 = OK, 1..14 = the reference's Error variants in
 *    declaration order (src/errors.rs:6-21), negative = device / runtime
 *    failures.  trs_status_message() describes every code.
 *  - Handles are thread-safe: calls may share one codec across threads.  The
 *    decode-plan cache keeps the reference's 64-entry most-recent-first policy
 *    (src/reedsolomon.rs:24, :592-613) and lives in device memory.
 */
#ifndef TRS_GPU_H
#define TRS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: src/errors.rs:6-21, in order -------------------------- */
enum trs_status {
    TRS_OK = 0,
    TRS_TOO_FEW_SHARDS = 1,
    TRS_TOO_MANY_SHARDS = 2,
    TRS_TOO_FEW_DATA_SHARDS = 3,
    TRS_TOO_MANY_DATA_SHARDS = 4,
    TRS_TOO_FEW_PARITY_SHARDS = 5,
    TRS_TOO_MANY_PARITY_SHARDS = 6,
    TRS_TOO_FEW_BUFFER_SHARDS = 7,
    TRS_TOO_MANY_BUFFER_SHARDS = 8,
    TRS_INCORRECT_SHARD_SIZE = 9,
    TRS_TOO_FEW_SHARDS_PRESENT = 10,
    TRS_EMPTY_SHARD = 11,
    TRS_INVALID_SHARD_FLAGS = 12,
    TRS_INVALID_INDEX = 13,
    TRS_SINGULAR_MATRIX = 14,
    /* device / runtime failures (no reference counterpart) */
    TRS_ERR_DEVICE = -1,      /* a HIP runtime call failed */
    TRS_ERR_OUT_OF_MEMORY = -2,
    TRS_ERR_INVALID_ARGUMENT = -3, /* where the reference panics (length asserts) */
    TRS_ERR_UNSUPPORTED = -4,
    TRS_ERR_SCRATCH_LIMIT = -5 /* beyond the allocation-free entry point's scratch */
};

/* Human-readable text for a status code (src/errors.rs:24-41 for 1..14). */
const char *trs_status_message(int status);

/* Library version string. */
const char *trs_version(void);

/* ---- slice API: src/gf/mod.rs:69-85 (gf::mul_slice / gf::mul_slice_xor) -- */
/* out[i] = c * in[i] (xor_into == 0) or out[i] ^= c * in[i] (xor_into != 0).
 * Unequal lengths return TRS_ERR_INVALID_ARGUMENT (the reference panics,
 * src/gf/scalar.rs:8). */
int trs_gf8_mul_slice(uint8_t *out, size_t out_len, const uint8_t *in, size_t in_len,
                      uint8_t c, int xor_into, void *stream);

/* ===================== ReedSolomon over GF(2^8) ========================== */
typedef struct trs_rs8 trs_rs8;
typedef struct trs_dec8 trs_dec8;

/* ReedSolomon::new, src/reedsolomon.rs:166-196. */
int trs_rs8_new(size_t data_shards, size_t parity_shards, trs_rs8 **out);
void trs_rs8_free(trs_rs8 *rs);
/* src/reedsolomon.rs:255-265 */
size_t trs_rs8_data_shard_count(const trs_rs8 *rs);
size_t trs_rs8_parity_shard_count(const trs_rs8 *rs);
size_t trs_rs8_total_shard_count(const trs_rs8 *rs);
/* ReedSolomon::encode_route, src/reedsolomon.rs:239-253: the device route
 * this codec's encode takes at the given shard length. */
const char *trs_rs8_encode_route(const trs_rs8 *rs, size_t shard_len);
/* ReedSolomon::encode_route as the reference's x86 build with GFNI answers it
 * (src/reedsolomon.rs:239-253, src/gf/fft_x86.rs:239-250): "fft-generated",
 * "fft-staged" or "fused-matrix".  Host only. */
const char *trs_rs8_reference_route(const trs_rs8 *rs, size_t shard_len);
/* Host copy of the n x k generator matrix (row-major, n*k bytes). */
int trs_rs8_generator(const trs_rs8 *rs, uint8_t *out, size_t out_len);

/* ReedSolomon::encode, src/reedsolomon.rs:271-279: shards = k data then m
 * parity buffers; parity overwritten. */
int trs_rs8_encode(const trs_rs8 *rs, uint8_t *const *shards, const size_t *lens,
                   size_t n_shards, void *stream);
/* ReedSolomon::encode_sep, src/reedsolomon.rs:284-309. */
int trs_rs8_encode_sep(const trs_rs8 *rs, const uint8_t *const *data, const size_t *data_lens,
                       size_t n_data, uint8_t *const *parity, const size_t *parity_lens,
                       size_t n_parity, void *stream);
/* Encode through a named route (0 = matrix apply, 1 = bit-sliced constant
 * matrix, 2 = bit-sliced LCH FFT, 3 = the staged FFT program of
 * src/fft.rs:393-592 on the byte-plane LDS interpreter); the bench / test
 * hook that ReedSolomon::encode_forced is in the reference
 * (src/reedsolomon.rs:365-398).  TRS_ERR_UNSUPPORTED when the shape has no
 * program for that route.  Automatic encodes take route 3 when k > 32: at any
 * length when the program has fewer than 0.13 k*m multiplies (k = 64, 128)
 * or 0.078 k*m (other k), else up to 1 MiB (k = 64, 128) or 256 KiB
 * (measured crossovers against the matrix route; the reference's rule,
 * src/reedsolomon.rs:204-234, is trs_rs8_reference_route). */
int trs_rs8_encode_sep_route(const trs_rs8 *rs, int route, const uint8_t *const *data,
                             const size_t *data_lens, size_t n_data, uint8_t *const *parity,
                             const size_t *parity_lens, size_t n_parity, void *stream);
/* ReedSolomon::encode_rows_into, src/reedsolomon.rs:667-682: data holds k
 * rows of row_len bytes back to back, parity m rows. */
int trs_rs8_encode_rows_into(const trs_rs8 *rs, const uint8_t *data, size_t data_len,
                             size_t row_len, uint8_t *parity, size_t parity_len, void *stream);
/* ReedSolomon::verify, src/reedsolomon.rs:430-463: *ok = 1 when every parity
 * shard matches the data, else 0.  Synchronises with `stream`. */
int trs_rs8_verify(const trs_rs8 *rs, const uint8_t *const *shards, const size_t *lens,
                   size_t n_shards, int *ok, void *stream);
/* ReedSolomon::reconstruct / reconstruct_data, src/reedsolomon.rs:470-477,
 * :615-641, :829-891, for (buffer, present) shards: present[i] != 0 marks a
 * valid shard; missing shards are overwritten (missing parity left untouched
 * when data_only != 0). lens[i] is every buffer's length. */
int trs_rs8_reconstruct(const trs_rs8 *rs, uint8_t *const *shards, const size_t *lens,
                        const uint8_t *present, size_t n_shards, int data_only, void *stream);
/* ReedSolomon::reconstruct_rows, src/reedsolomon.rs:514-535. */
int trs_rs8_reconstruct_rows(const trs_rs8 *rs, uint8_t *rows, size_t rows_len, size_t row_len,
                             const uint8_t *present, size_t n_present, void *stream);
/* Matrix-apply route of this codec (reconstruct, and encode of shapes without
 * a generated program; also decoders and batch requests created after the
 * call): 0 = automatic, 1 = fused table kernel, 2 = int8 MFMA on the
 * bit-expanded rows (k <= 32; TRS_ERR_UNSUPPORTED otherwise).  Automatic
 * takes MFMA for single calls with k <= 32 and shards under 512 KiB (about
 * twice as fast there, measured), otherwise the bit-sliced planes kernel from
 * 128 KiB and the table kernel below; batch work lists use the planes
 * kernel.  Every route is byte-identical; no
 * reference counterpart (the reference has one CPU kernel per ISA). */
int trs_rs8_set_apply_route(trs_rs8 *rs, int route);
/* Reconstruct route of this codec (single calls and decoders prepared after
 * the call): 0 = automatic, 1 = matrix plans (the reference's decode,
 * src/reedsolomon.rs:829-891), 2 = the additive-FFT erasure decoder over
 * GF(2^8) on the byte-plane interpreter (formal-derivative method; no
 * reference counterpart).  Automatic takes the decoder for k > 32 below
 * 128 KiB shards, up to 256 KiB when 100 N < 170 k + 60 e, and beyond when
 * k e > 7 N log2 N (N the power of two >= k + m, e the rebuilt shards;
 * measured crossovers), else the matrix plans.  Byte-identical either way. */
int trs_rs8_set_decode_route(trs_rs8 *rs, int route);
/* ReedSolomon::prepare_decode, src/reedsolomon.rs:484-507. */
int trs_rs8_prepare_decode(const trs_rs8 *rs, const uint8_t *present, size_t n_present,
                           trs_dec8 **out);
/* PreparedDecoder::{reconstruct, reconstruct_data}, src/reedsolomon.rs:762-797. */
int trs_dec8_reconstruct(const trs_dec8 *dec, uint8_t *const *shards, const size_t *lens,
                         const uint8_t *present, size_t n_shards, int data_only, void *stream);
void trs_dec8_free(trs_dec8 *dec);

/* ================= ReedSolomon16 over GF((2^8)^2) ======================== */
typedef struct trs_rs16 trs_rs16;
typedef struct trs_dec16 trs_dec16;

/* ReedSolomon16::new, src/reedsolomon16.rs:281-319 (k + m <= 65536). */
int trs_rs16_new(size_t data_shards, size_t parity_shards, trs_rs16 **out);
void trs_rs16_free(trs_rs16 *rs);
size_t trs_rs16_data_shard_count(const trs_rs16 *rs);
size_t trs_rs16_parity_shard_count(const trs_rs16 *rs);
size_t trs_rs16_total_shard_count(const trs_rs16 *rs);
/* ReedSolomon16::encode_route, src/reedsolomon16.rs:338-346: "fft" or
 * "matrix".  The reference routes per host (src/fft16.rs:70-106; both routes
 * are byte-identical); here the route is "fft" whenever the shape has a staged
 * program (2 <= k <= 1024, k + m <= 4096) whose register file fits a
 * workgroup's LDS.  The answer takes no shard length, as the reference's:
 * for short data (k <= 16) the automatic encode still runs a matrix kernel
 * where it measured faster (the Karatsuba kernel from 4 MiB shards, int8
 * MFMA for m <= 4 from 512 KiB to 4 MiB). */
const char *trs_rs16_encode_route(const trs_rs16 *rs);
/* Host copy of the n x k tower generator (row-major, n*k uint16). */
int trs_rs16_generator(const trs_rs16 *rs, uint16_t *out, size_t out_len);
/* ReedSolomon16::encode, src/reedsolomon16.rs:412-418 (split-plane shards of
 * equal even length: N low bytes then N high bytes). */
int trs_rs16_encode(const trs_rs16 *rs, uint8_t *const *shards, const size_t *lens,
                    size_t n_shards, void *stream);
/* ReedSolomon16::encode_sep, src/reedsolomon16.rs:422-447. */
int trs_rs16_encode_sep(const trs_rs16 *rs, const uint8_t *const *data, const size_t *data_lens,
                        size_t n_data, uint8_t *const *parity, const size_t *parity_lens,
                        size_t n_parity, void *stream);
/* ReedSolomon16::encode_sep_matrix, src/reedsolomon16.rs:324-336: the matrix
 * route regardless of routing. */
int trs_rs16_encode_sep_matrix(const trs_rs16 *rs, const uint8_t *const *data,
                               const size_t *data_lens, size_t n_data, uint8_t *const *parity,
                               const size_t *parity_lens, size_t n_parity, void *stream);
/* The tower-FFT route regardless of routing: the staged program of
 * src/fft16.rs:357-376 run by the GPU interpreter (fft16::encode,
 * src/fft16.rs:887-930, is the reference path it replaces).  Available for
 * 2 <= k <= 1024 and k + m <= 4096 whose program fits one workgroup's LDS;
 * otherwise TRS_ERR_UNSUPPORTED. */
int trs_rs16_encode_sep_fft(const trs_rs16 *rs, const uint8_t *const *data,
                            const size_t *data_lens, size_t n_data, uint8_t *const *parity,
                            const size_t *parity_lens, size_t n_parity, void *stream);
/* Host-only introspection of the staged program for shape (k, m) (the
 * fields of StagedProgram16, src/fft16.rs:151-174, as the dump_stats test
 * src/fft16.rs:1035-1074 prints them) and of its device lowering.  out[0..n):
 * register_count, mult_count, glue ops, stages, zero registers, profitable
 * (routing rule src/fft16.rs:85-106), device registers, device steps, LDS
 * bytes per 32-column group.  TRS_ERR_UNSUPPORTED outside 2 <= k <= 1024,
 * k + m <= 4096.  Touches no device. */
int trs_fft16_program_info(size_t k, size_t m, size_t *out, size_t n);
/* Host-only dump of the device lowering of shape (k, m)'s program, the data
 * the GPU interpreter runs (no reference counterpart; tests replay it on the
 * CPU).  counts[0..5): u32 words of steps (8 per step: kind, count, a, b, c,
 * d, 0, 0), constant words, op words (2 per op), parity registers, and the
 * register count.  Buffers too small for any count: only counts are filled. */
int trs_fft16_lowered_program(size_t k, size_t m, uint32_t *steps, size_t steps_cap,
                              uint32_t *consts, size_t consts_cap, uint32_t *ops, size_t ops_cap,
                              uint32_t *parity, size_t parity_cap, size_t *counts);
/* The same for the GF(2^8) staged program, src/fft.rs:393-592
 * (build_staged_program, lower_staged; fields of StagedProgram :134-155).
 * out[0..n): register_count, mult_count, glue ops, stages, zero registers
 * (all of src/fft.rs's lowering), reference route (1 when the reference's
 * x86 build routes the shape to this program, src/reedsolomon.rs:204-234),
 * device registers, device steps, LDS bytes per 32-byte column group (8 bit
 * planes per register), and the multiply count after the device lowering's
 * liveness pass.
 * TRS_ERR_UNSUPPORTED unless k >= 1, m >= 1, k + m <= 256.  Touches no
 * device. */
int trs_fft8_program_info(size_t k, size_t m, size_t *out, size_t n);
/* Host-only dump of its device lowering, as trs_fft16_lowered_program.  The
 * constants are bytes c of the subfield, written as their tower Karatsuba
 * words [c, c, 0]; the byte-plane interpreter reads the low byte. */
int trs_fft8_lowered_program(size_t k, size_t m, uint32_t *steps, size_t steps_cap,
                             uint32_t *consts, size_t consts_cap, uint32_t *ops, size_t ops_cap,
                             uint32_t *parity, size_t parity_cap, size_t *counts);
/* Encode of `stripes` stripes of the codec's shape in one call: shards holds
 * stripes * (k + m) pointers (stripe s: data then parity), every shard `len`
 * bytes.  route 0: the codec's route, 1: matrix, 2: tower FFT.  The reference
 * encodes stripe by stripe (src/reedsolomon16.rs:412-418); batching only
 * merges the launches. */
int trs_rs16_encode_stripes(const trs_rs16 *rs, uint8_t *const *shards, size_t stripes,
                            size_t len, int route, void *stream);
/* trs_rs16_encode_stripes on the tower-FFT route without any device
 * allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The per-call shard-pointer table goes into
 * a fixed 1 MiB device array that each GPU holds for the life of the process
 * (the codec's FFT program is uploaded by trs_rs16_new), so stripes * (k + m)
 * may not exceed 131072 pointers (e.g. 64 stripes of (1024,1024)):
 * TRS_ERR_SCRATCH_LIMIT beyond.  TRS_ERR_UNSUPPORTED for a codec without a
 * staged program that fits the device (trs_rs16_encode_route "matrix").  All
 * calls on one GPU share that array and its last call's contents: calls that
 * overlap on one GPU (other host threads, other streams) race; calls ordered
 * on one stream do not. */
int trs_rs16_encode_stripes_noalloc(const trs_rs16 *rs, uint8_t *const *shards, size_t stripes,
                                    size_t len, void *stream);
/* Host-only: the fused decode rows a reconstruct with flags `present` (n
 * entries) applies -- rows for the missing data shards, then for the missing
 * parity shards, k coefficients each (uint8 for field 8, uint16 for field 16),
 * over the first k present shards.  These are the rows of the reference's
 * DecodePlan (src/reedsolomon.rs:537-613, src/reedsolomon16.rs:595-640) with
 * its parity pass (src/reedsolomon.rs:866-889) folded in.  *rows_out gets the
 * row count; out may be NULL to query it. */
int trs_decode_matrix(int field, size_t k, size_t n, const uint8_t *present, void *out,
                      size_t out_len, size_t *rows_out);
/* Host-only dump of the additive-FFT erasure decoder the tower codec runs for
 * presence flags `present` (n entries, k data shards); no reference
 * counterpart, tests replay it on the CPU.  Layout as
 * trs_fft16_lowered_program; counts[0..6): step words, constant words, op
 * words, loaded registers (pointer slots 0..), stored registers, register
 * count.  TRS_ERR_UNSUPPORTED past 2048 shards. */
int trs_fft16_decode_program(size_t k, size_t n, const uint8_t *present, uint32_t *steps,
                             size_t steps_cap, uint32_t *consts, size_t consts_cap, uint32_t *ops,
                             size_t ops_cap, uint32_t *load, size_t load_cap, uint32_t *store,
                             size_t store_cap, size_t *counts);
/* The same for the GF(2^8) codec's FFT erasure decoder (byte field, n <= 256;
 * constants lowered as the Karatsuba words [c, c, 0] of subfield bytes).
 * TRS_ERR_UNSUPPORTED past 256 shards. */
int trs_fft8_decode_program(size_t k, size_t n, const uint8_t *present, uint32_t *steps,
                            size_t steps_cap, uint32_t *consts, size_t consts_cap, uint32_t *ops,
                            size_t ops_cap, uint32_t *load, size_t load_cap, uint32_t *store,
                            size_t store_cap, size_t *counts);
/* Reconstruct route of the tower codec: 0 auto, 1 matrix (the reference's
 * decode plans, src/reedsolomon16.rs:595-640), 2 additive-FFT erasure decoder
 * (no reference counterpart; byte-identical, since the erased symbols are
 * unique).  Auto takes the FFT when k * (shards rebuilt) >= N log2 N, N the
 * power of two >= k + m, and k + m <= 2048.  Route 2 past 2048 shards:
 * TRS_ERR_UNSUPPORTED.  Prepared decoders keep the route they were made with. */
int trs_rs16_set_decode_route(trs_rs16 *rs, int route);
/* Matrix-apply route of the tower codec (matrix-route encode incl.
 * trs_rs16_encode_sep_matrix, matrix-plan reconstruct, and decoders prepared
 * after the call): 0 = automatic, 1 = the Karatsuba table kernel, 2 = int8
 * MFMA (k <= 16 and 2 (k + m) <= 256; TRS_ERR_UNSUPPORTED otherwise).  The
 * MFMA form treats each split-plane shard as two byte shards (its low and
 * high planes) and multiplies by the 16m x 16k GF(2) expansion of the tower
 * matrix, keeping the low bit of each exact int32 sum.  Every route is
 * byte-identical; no reference counterpart (its tower kernels are CPU code,
 * src/gf/tower_x86.rs:33-133). */
int trs_rs16_set_apply_route(trs_rs16 *rs, int route);
/* Reconstruct of `stripes` stripes of the codec's shape in one call, each with
 * its own presence flags: shards holds stripes * (k + m) pointers, present
 * stripes * (k + m) flags, every shard `len` bytes.  Per stripe this is
 * ReedSolomon16::reconstruct / reconstruct_data (src/reedsolomon16.rs:508-515)
 * with the same checks; every stripe is validated and planned before any is
 * rebuilt, and stripes on the FFT decode route run in one launch. */
int trs_rs16_reconstruct_stripes(const trs_rs16 *rs, uint8_t *const *shards,
                                 const uint8_t *present, size_t stripes, size_t len, int data_only,
                                 void *stream);
/* ReedSolomon16::verify, src/reedsolomon16.rs:453-505. */
int trs_rs16_verify(const trs_rs16 *rs, const uint8_t *const *shards, const size_t *lens,
                    size_t n_shards, int *ok, void *stream);
/* ReedSolomon16::reconstruct / reconstruct_data, src/reedsolomon16.rs:508-515. */
int trs_rs16_reconstruct(const trs_rs16 *rs, uint8_t *const *shards, const size_t *lens,
                         const uint8_t *present, size_t n_shards, int data_only, void *stream);
/* ReedSolomon16::reconstruct_rows, src/reedsolomon16.rs:546-569. */
int trs_rs16_reconstruct_rows(const trs_rs16 *rs, uint8_t *rows, size_t rows_len,
                              size_t row_len, const uint8_t *present, size_t n_present,
                              void *stream);
/* ReedSolomon16::prepare_decode, src/reedsolomon16.rs:518-543. */
int trs_rs16_prepare_decode(const trs_rs16 *rs, const uint8_t *present, size_t n_present,
                            trs_dec16 **out);
/* PreparedDecoder16::{reconstruct, reconstruct_data}, src/reedsolomon16.rs:722-753. */
int trs_dec16_reconstruct(const trs_dec16 *dec, uint8_t *const *shards, const size_t *lens,
                          const uint8_t *present, size_t n_shards, int data_only, void *stream);
void trs_dec16_free(trs_dec16 *dec);

/* ================= batched GF(2^8) coding (no reference counterpart) ====== */
/* The reference is one-call-one-stripe and leaves batching to its callers
 * (BENCH-RESULTS.md:98-100).  A batch collects many encode / reconstruct
 * requests on GF(2^8) codecs, of mixed shapes, lengths and erasure patterns,
 * and runs them with one kernel launch per device work list (generated
 * program, MFMA, runtime-coefficient tile, interpreter).
 * Each request is validated exactly as the single-stripe call would be, and
 * each result is byte-identical to that call's. */
typedef struct trs_batch8 trs_batch8;

int trs_batch8_new(trs_batch8 **out);
void trs_batch8_free(trs_batch8 *b);
/* Drop every queued request (keeps device buffers for reuse). */
void trs_batch8_clear(trs_batch8 *b);
/* Queue rs.encode(shards) (same checks as trs_rs8_encode). */
int trs_batch8_add_encode(trs_batch8 *b, const trs_rs8 *rs, uint8_t *const *shards,
                          const size_t *lens, size_t n_shards);
/* Queue rs.reconstruct(shards) / reconstruct_data (same checks as
 * trs_rs8_reconstruct; the decode plan comes from rs's device plan cache). */
int trs_batch8_add_reconstruct(trs_batch8 *b, const trs_rs8 *rs, uint8_t *const *shards,
                               const size_t *lens, const uint8_t *present, size_t n_shards,
                               int data_only);
/* Upload the work list for the queued requests (once per batch shape).
 * Requests whose shape has a generated program run it (one kernel per
 * program); the others run the bit-sliced runtime-coefficient kernel
 * (8 KiB tiles), or the packed-byte table kernel when the environment sets
 * TRS_BATCH8_TILE=table.  TRS_BATCH_GRID_PER_CU / TRS_BATCH_BS_GRID_PER_CU /
 * TRS_BATCH_MFMA_GRID_PER_CU cap each list's grid at that many workgroups per
 * CU (default 128 / 128 / 2; the first two amount to one workgroup per tile).
 * Prepare first waits for earlier runs of this batch to finish, since they
 * read the device tables it rewrites. */
int trs_batch8_prepare(trs_batch8 *b);
/* Run every queued request on `stream`: one kernel per non-empty work list. */
int trs_batch8_run(trs_batch8 *b, void *stream);
/* Graph mode: the first run after prepare captures the batch's launches into
 * a hipGraph (kernel nodes only, on a private stream) and every run replays
 * it on `stream`; prepare and clear drop the graph.  0 turns it off. */
int trs_batch8_set_graph(trs_batch8 *b, int enable);
/* Algorithmic bytes of the queued work: k reads + outputs written, per stripe. */
uint64_t trs_batch8_bytes(const trs_batch8 *b);
size_t trs_batch8_size(const trs_batch8 *b);

/* ================= serving loop (no reference counterpart) ================= */
/* Continuous batching of encode / reconstruct requests over device-resident
 * shards.  Submissions are validated like the single calls (same error codes)
 * and return a ticket; a dispatcher packs pending requests, in arrival order,
 * into the batch of a free slot (slot = stream + batch) while the bytes of
 * in-flight work stay within hbm_budget_bytes (0 = unlimited), and launches
 * it, so requests arriving while others run start as soon as a slot and budget
 * free up.  A GF(2^8) request that is alone in its slot runs as its plain
 * per-call launch (no batch tables to build and upload).  window_us: how long the dispatcher waits for more arrivals before
 * launching a partial batch while other slots have work in flight (with every
 * slot idle a request launches at once).  Shard buffers must stay valid until
 * the ticket completes. */
typedef struct trs_server trs_server;
int trs_server_new(uint32_t slots, uint64_t hbm_budget_bytes, uint32_t max_batch,
                   uint32_t window_us, trs_server **out);
void trs_server_free(trs_server *s);   /* drains queued and in-flight work first */
int trs_server_submit_encode8(trs_server *s, const trs_rs8 *rs, uint8_t *const *shards,
                              const size_t *lens, size_t n_shards, uint64_t *ticket);
int trs_server_submit_reconstruct8(trs_server *s, const trs_rs8 *rs, uint8_t *const *shards,
                                   const size_t *lens, const uint8_t *present, size_t n_shards,
                                   int data_only, uint64_t *ticket);
int trs_server_submit_encode16(trs_server *s, const trs_rs16 *rs, uint8_t *const *shards,
                               const size_t *lens, size_t n_shards, uint64_t *ticket);
int trs_server_submit_reconstruct16(trs_server *s, const trs_rs16 *rs, uint8_t *const *shards,
                                    const size_t *lens, const uint8_t *present, size_t n_shards,
                                    int data_only, uint64_t *ticket);
int trs_server_poll(trs_server *s, uint64_t ticket, int *done, int *status);
int trs_server_wait(trs_server *s, uint64_t ticket, int *status);
int trs_server_drain(trs_server *s);
/* Submit-to-completion latency (microseconds), algorithmic bytes and status of
 * a finished ticket; the ticket is forgotten afterwards. */
int trs_server_take(trs_server *s, uint64_t ticket, double *latency_us, uint64_t *bytes,
                    int *status);
/* trs_server_take plus where the latency went: phases_us[0] queued (submit to
 * taken into a slot's batch), [1] host launch (batch tables built, uploaded
 * and kernels enqueued), [2] device and completion (enqueued to seen done).
 * phases_us may be NULL; a request with nothing to do reports zeros. */
int trs_server_take_detail(trs_server *s, uint64_t ticket, double *latency_us,
                           double *phases_us, uint64_t *bytes, int *status);
/* Device memory the server's own batch tables may not take (default 256 MiB).
 * Before a slot's table buffer grows, the dispatcher asks hipMemGetInfo for
 * free HBM; a batch whose tables need more than free less this reserve is
 * split in two (the later half returns to the head of the queue) until it
 * fits.  A GF(2^8) request alone in a slot runs as its plain per-call launch
 * (shard pointers in the kernel arguments), which needs no tables, so
 * splitting always ends; should an allocation still fail, the slot's requests
 * complete with TRS_ERR_OUT_OF_MEMORY and the server goes on. */
int trs_server_set_hbm_reserve(trs_server *s, uint64_t reserve_bytes);
/* Microseconds the dispatcher (waiting for a submission) and the completion
 * threads (waiting for a batch) poll before they block (default 200, at most
 * 100000; 0 = block at once).  Polling trades a busy host thread for not
 * paying a thread wake-up on each request at low load. */
int trs_server_set_spin(trs_server *s, uint32_t spin_us);
/* Counters since creation: out4[0] requests completed through a slot, [1]
 * batches launched, [2] batches split because their tables did not fit free
 * HBM less the reserve, [3] requests failed with TRS_ERR_OUT_OF_MEMORY. */
int trs_server_stats(trs_server *s, uint64_t *out4);

/* ============ host-resident shards: the reference's slice API ============== */
/* The reference's methods take host slices (src/reedsolomon.rs:271-535,
 * src/reedsolomon16.rs:412-565).  Each _host entry point below takes the same
 * arguments as its device counterpart above (minus the stream), keeps the same
 * checks, error order and codes, and takes HOST pointers: pageable memory, or
 * pinned memory from hipHostMalloc.  The shards stream through device memory
 * in column chunks (trs_set_host_chunk): the host-to-device copy of one chunk,
 * the kernels of the previous one and the device-to-host copy of the one
 * before run on three library streams at once.  Only the shards a call reads
 * go to the device (for a reconstruct: the first k present) and only the
 * shards it writes come back.  Every call blocks until its outputs are in host
 * memory.  A device pointer among the shards returns TRS_ERR_INVALID_ARGUMENT.
 * RS16 shards keep the split-plane wire layout: each chunk carries the same
 * column range of the low and the high plane. */
/* ReedSolomon::encode, src/reedsolomon.rs:271-279. */
int trs_rs8_encode_host(const trs_rs8 *rs, uint8_t *const *shards, const size_t *lens,
                        size_t n_shards);
/* ReedSolomon::encode_sep, src/reedsolomon.rs:284-309. */
int trs_rs8_encode_sep_host(const trs_rs8 *rs, const uint8_t *const *data, const size_t *data_lens,
                            size_t n_data, uint8_t *const *parity, const size_t *parity_lens,
                            size_t n_parity);
/* ReedSolomon::encode_rows_into, src/reedsolomon.rs:667-682. */
int trs_rs8_encode_rows_into_host(const trs_rs8 *rs, const uint8_t *data, size_t data_len,
                                  size_t row_len, uint8_t *parity, size_t parity_len);
/* ReedSolomon::verify, src/reedsolomon.rs:430-463. */
int trs_rs8_verify_host(const trs_rs8 *rs, const uint8_t *const *shards, const size_t *lens,
                        size_t n_shards, int *ok);
/* ReedSolomon::reconstruct / reconstruct_data, src/reedsolomon.rs:470-477. */
int trs_rs8_reconstruct_host(const trs_rs8 *rs, uint8_t *const *shards, const size_t *lens,
                             const uint8_t *present, size_t n_shards, int data_only);
/* ReedSolomon::reconstruct_rows, src/reedsolomon.rs:514-535. */
int trs_rs8_reconstruct_rows_host(const trs_rs8 *rs, uint8_t *rows, size_t rows_len,
                                  size_t row_len, const uint8_t *present, size_t n_present);
/* PreparedDecoder::{reconstruct, reconstruct_data}, src/reedsolomon.rs:762-797. */
int trs_dec8_reconstruct_host(const trs_dec8 *dec, uint8_t *const *shards, const size_t *lens,
                              const uint8_t *present, size_t n_shards, int data_only);
/* ReedSolomon16::encode, src/reedsolomon16.rs:412-418. */
int trs_rs16_encode_host(const trs_rs16 *rs, uint8_t *const *shards, const size_t *lens,
                         size_t n_shards);
/* ReedSolomon16::encode_sep, src/reedsolomon16.rs:422-447. */
int trs_rs16_encode_sep_host(const trs_rs16 *rs, const uint8_t *const *data,
                             const size_t *data_lens, size_t n_data, uint8_t *const *parity,
                             const size_t *parity_lens, size_t n_parity);
/* ReedSolomon16::verify, src/reedsolomon16.rs:453-505. */
int trs_rs16_verify_host(const trs_rs16 *rs, const uint8_t *const *shards, const size_t *lens,
                         size_t n_shards, int *ok);
/* ReedSolomon16::reconstruct / reconstruct_data, src/reedsolomon16.rs:508-515. */
int trs_rs16_reconstruct_host(const trs_rs16 *rs, uint8_t *const *shards, const size_t *lens,
                              const uint8_t *present, size_t n_shards, int data_only);
/* ReedSolomon16::reconstruct_rows, src/reedsolomon16.rs:546-569. */
int trs_rs16_reconstruct_rows_host(const trs_rs16 *rs, uint8_t *rows, size_t rows_len,
                                   size_t row_len, const uint8_t *present, size_t n_present);
/* PreparedDecoder16::{reconstruct, reconstruct_data}, src/reedsolomon16.rs:722-753. */
int trs_dec16_reconstruct_host(const trs_dec16 *dec, uint8_t *const *shards, const size_t *lens,
                               const uint8_t *present, size_t n_shards, int data_only);
/* Bytes of each shard per pipeline chunk of the _host calls (0 = 1 MiB, the
 * default); a chunk of all of a call's shards is further capped at 64 MiB.
 * Process-wide.  No reference counterpart (the reference tunes strips by hand,
 * src/reedsolomon16.rs:796-804). */
int trs_set_host_chunk(size_t bytes);

/* Clone, src/reedsolomon.rs:725-739 and src/reedsolomon16.rs:685: a new codec
 * of the same shape and routes whose plan cache starts as a copy of this
 * one's (the cached plans are shared, they never change).  Free with
 * trs_rs8_free / trs_rs16_free. */
int trs_rs8_clone(const trs_rs8 *rs, trs_rs8 **out);
int trs_rs16_clone(const trs_rs16 *rs, trs_rs16 **out);

/* ============ device memory for callers without a HIP runtime of their own == */
/* No reference counterpart (the reference's shards are host memory). */
int trs_device_alloc(size_t bytes, void **out);
int trs_device_free(void *p);
/* Stream-ordered copy between any host and device buffers (direction taken
 * from the pointers); pageable host memory makes it synchronous. */
int trs_memcpy(void *dst, const void *src, size_t bytes, void *stream);
int trs_stream_synchronize(void *stream);

/* ============ wire formats and Option<T> shards (no reference counterpart) === */
/* ReedSolomon::reconstruct with Option<T> shards (src/reedsolomon.rs:88-101,
 * :842-847): a NULL entry of `shards` is a missing shard that the codec
 * allocates in device memory (the present shards' length) and rebuilds; on
 * return the entry points to it and lens[i] holds its length.  With data_only
 * missing parity entries stay NULL, as the reference leaves them None.  Free
 * the buffers with trs_device_free.  On an error nothing stays allocated. */
int trs_rs8_reconstruct_alloc(const trs_rs8 *rs, uint8_t **shards, size_t *lens, size_t n,
                              int data_only, void *stream);
/* Wire formats of a GF(2^16) shard (src/reedsolomon16.rs:4-7): n_symbols
 * interleaved little-endian u16 symbols (value (high << 8) | low) <-> the
 * split-plane layout [low n_symbols | high n_symbols] that the RS16 calls
 * take.  Device buffers, stream-ordered; in and out must not overlap. */
int trs_u16_to_split(const uint16_t *symbols, size_t n_symbols, uint8_t *shard, void *stream);
int trs_split_to_u16(const uint8_t *shard, size_t n_symbols, uint16_t *symbols, void *stream);

/* ================= autotuning (no reference counterpart) ================== */
/* Kernel variants of one apply of k inputs to nout outputs (the autotuner's
 * and the learned router's candidates; every variant is byte-identical):
 *   0 table tile (4 KiB column chunks, rows by nout)
 *   1 / 2 bit-sliced planes tile with 4 / 8 output rows per workgroup
 *   3 int8 MFMA on the bit-expanded matrix (GF(2^8), k <= 32)
 *   4 / 5 / 6 the shape's FFT program (encode): the generated bit-sliced
 *     program where the GF(2^8) shape has one (variant 5 only), else the
 *     staged program on the LDS interpreter -- GF(2^8) byte planes
 *     (src/fft.rs:393-592) or the tower (src/fft16.rs:357-376) -- sized for
 *     16 / 32 / 64 KiB of LDS per workgroup; targets that give the same launch
 *     are one variant.
 * The reference tunes by hand: MULT_MARGIN (src/reedsolomon.rs:226-232),
 * STRIP_BYTES (src/reedsolomon16.rs:796-804), 9 mults < 2 k m
 * (src/fft16.rs:105). */
/* *mask: bit v set when variant v can encode shape (k, nout) over GF(field)
 * at this shard length (the tower's len is the split-plane shard length). */
int trs_tune_feasible(int field, uint32_t k, uint32_t nout, uint64_t len, uint32_t *mask);
/* The variant the codec's rules pick for that encode (no model). */
int trs_tune_rule(int field, uint32_t k, uint32_t nout, uint64_t len, int *variant);
/* The cost model's 12 inputs for (shape, length, variant): field, log2 k,
 * log2 nout, log2 len, log2 k*nout, log2 bytes moved, four kernel-kind flags,
 * log2 of the tile parameter (rows or LDS KiB), k / (k + nout). */
int trs_tune_features(int field, uint32_t k, uint32_t nout, uint64_t len, int variant,
                      float *out, size_t n);
/* Times variant `variant` on synthetic device data (random nonzero
 * coefficients; kernel time does not depend on their values): *ms_out = mean
 * milliseconds of `reps` applies after one warm-up, or the warm-up's time when
 * that alone exceeds limit_ms (> 0; a candidate already far behind).
 * TRS_ERR_UNSUPPORTED for an infeasible variant. */
int trs_tune_apply(int field, uint32_t k, uint32_t nout, uint64_t len, int variant, int reps,
                   double limit_ms, double *ms_out);
/* The learned router.  Installs a trained cost model (the trs_mlp parameters
 * and the feature standardisation: 12 means then 12 standard deviations):
 * from then on an automatic encode (both codecs) or matrix-plan reconstruct
 * (GF(2^8)) whose (field, op, k, nout, half-octave length bucket) has no
 * cached choice ranks its feasible variants with the model on the GPU, runs
 * the top `top_n` on the call's own buffers (a warm-up and a timed run each;
 * the call then synchronises its stream once) and caches the fastest; later
 * calls of that key run the cached variant.  Calls with a forced route
 * (trs_rs8_set_apply_route) and the FFT erasure decoders keep their rules.
 * Process-wide; installing again or removing drops the cache. */
int trs_tuner_install(const float *params, size_t n_params, const float *norm, size_t n_norm,
                      int top_n);
int trs_tuner_remove(void);
/* The cached choice for a key (op 0 encode, 1 reconstruct), -1 when none. */
int trs_tuner_lookup(int field, int op, size_t k, size_t nout, size_t len, int *variant);

/* FP32 MLP cost model (12 -> 64 -> 64 -> 1, ReLU), trained and evaluated by
 * hand-written HIP kernels; it predicts log kernel time from shape / route /
 * strip features so the autotuner measures only the top candidates.  Inputs
 * are host float arrays: X is n x 12 row-major, T and Y are n.  Training is
 * full-batch Adam on the mean squared error; gradients are summed in a fixed
 * order (deterministic).  Parameter layout: W1[64][12], b1[64], W2[64][64],
 * b2[64], W3[64], b3. */
typedef struct trs_mlp trs_mlp;
int trs_mlp_param_count(void);
int trs_mlp_new(uint64_t seed, trs_mlp **out);
void trs_mlp_free(trs_mlp *m);
int trs_mlp_get_params(const trs_mlp *m, float *out, size_t n);
int trs_mlp_set_params(trs_mlp *m, const float *in, size_t n);
/* Loss and gradient of the current parameters (for gradient checks). */
int trs_mlp_gradients(trs_mlp *m, const float *X, const float *T, size_t n, float *grad_out,
                      float *loss_out);
int trs_mlp_train(trs_mlp *m, const float *X, const float *T, size_t n, int steps, float lr,
                  float *loss_out);
int trs_mlp_predict(const trs_mlp *m, const float *X, size_t n, float *Y);

/* ================= utilities (bench / test helpers) ======================= */
/* Fill a device buffer with the seeded splitmix64 byte stream: byte j of
 * 8-byte word w is byte j (little-endian) of mix64(seed + (w+1)*0x9E3779B97F4A7C15). */
int trs_fill_random(uint8_t *dst, size_t len, uint64_t seed, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TRS_GPU_H */
