/*
 * FRSZ2 oracle: a plain-C restatement of the reference CPU codec.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (frsz_amd/, include/)
 * links, loads or calls this code.  It is used by tests/ as the bit-exact
 * checker, by __graft_entry__.smoke() as the checker of one small GPU call,
 * and by bench.py's CPU leg (cpu_leg) as the checker of the benchmark's GPU
 * outputs and the "port" CPU baseline.
 *
 * Every function follows a reference function in include/frsz2.hpp of
 * robertu94/frsz; the citation is given at each definition in
 * frsz2_oracle.c.  Parity status of each function (what pins it) is listed
 * in the header comment of frsz2_oracle.c.
 *
 * dtype: 0 = float32, 1 = float64.  bits b and exponent block size B are
 * runtime values.  Buffers are host memory.
 */
#ifndef FRSZ2_ORACLE_H
#define FRSZ2_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int fo_storage_bits(int bits);
uint64_t fo_exponent_region_bytes(int bits, uint64_t B, uint64_t n);
uint64_t fo_compressed_bytes(int bits, uint64_t B, uint64_t n);
uint64_t fo_block_payload_bytes(int bits, uint64_t B);

/* floating-point decomposition (float32 / float64) */
int32_t fo_exponent_f32(float x);
uint32_t fo_significand_f32(float x);
int fo_sign_f32(float x);
int32_t fo_exponent_f64(double x);
uint64_t fo_significand_f64(double x);
int fo_sign_f64(double x);
uint32_t fo_floating_to_fixed_f32(float x, int16_t block_exponent);
uint64_t fo_floating_to_fixed_f64(double x, int16_t block_exponent);
float fo_fixed_to_floating_f32(uint32_t fixed, int64_t block_exponent);
double fo_fixed_to_floating_f64(uint64_t fixed, int64_t block_exponent);

/* value-level batch: codes[i] = top `bits` bits of floating_to_fixed(x_i, E),
 * values[i] = fixed_to_floating(code_i << (W-bits), E) (bit patterns) */
void fo_roundtrip_f32(const uint32_t* u, uint64_t n, int16_t E, int bits, uint32_t* codes, uint32_t* values);
void fo_roundtrip_f64(const uint64_t* u, uint64_t n, int16_t E, int bits, uint64_t* codes, uint64_t* values);

/* whole-array codec; returns 0, 1 (bad dtype) or 2 (bad bits / block size).
 * compress writes exactly the bytes the reference writes: the exponent
 * array, full blocks and the touched words of a partial last block.  The
 * exponent pad and the untouched tail words are left as they were. */
int fo_compress(int dtype, int bits, uint64_t B, const void* in, uint64_t n, uint8_t* out);
int fo_decompress(int dtype, int bits, uint64_t B, const uint8_t* in, uint64_t n, void* out);
int fo_gather(int dtype, int bits, uint64_t B, const uint8_t* in, uint64_t n, const uint64_t* idx,
              uint64_t m, void* out);

/* one block (compress_cpu_block / decompress_cpu_block of a view over n
 * values): vals holds the block's values; compress uses the first len
 * (1..B, k*B + len <= n) and writes exponent k and the words its codes
 * touch.  Returns 0, 1, 2, or 3 (block index / length out of range). */
int fo_compress_block(int dtype, int bits, uint64_t B, uint64_t n, uint64_t k, const void* vals, uint64_t len,
                      uint8_t* out);
int fo_decompress_block(int dtype, int bits, uint64_t B, uint64_t n, uint64_t k, const uint8_t* in, void* vals);

/* block-parallel variants over the same per-block routine (OpenMP, static
 * schedule); byte-identical to the serial ones. threads <= 0: all cores. */
int fo_compress_mt(int dtype, int bits, uint64_t B, const void* in, uint64_t n, uint8_t* out, int threads);
int fo_decompress_mt(int dtype, int bits, uint64_t B, const uint8_t* in, uint64_t n, void* out, int threads);
int fo_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif
