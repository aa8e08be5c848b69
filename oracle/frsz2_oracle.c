/*
 * FRSZ2 oracle: plain-C restatement of the reference CPU codec
 * (robertu94/frsz, include/frsz2.hpp).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * CPU leg of bench.py (cpu_leg, which yields cpu_baseline) load this
 * library, and only as the checker (or, for bench.py, also as the timed
 * "port" CPU baseline).  The product never links it.
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   fo_exponent_* / fo_significand_* / fo_sign_*  -- the reference's known-
 *       answer vectors (test/test_frsz.cc:529-564, test/test_frsz_gpu.cu:213-229)
 *       and the compiled reference header on sampled bit patterns.
 *   fo_floating_to_fixed_* / fo_fixed_to_floating_* -- known answers of
 *       test/test_frsz_gpu.cu:276-301 and test/test_frsz.cc:561-563, the
 *       compiled reference on sampled patterns, and the closed forms
 *       (b = W, B = 1 is lossless; zero code -> +-2^(E-31)).
 *   fo_compressed_bytes / fo_exponent_region_bytes -- closed form of
 *       include/frsz2.hpp:534-548 and the compiled reference.
 *   fo_compress / fo_decompress / fo_gather (+ _mt) -- golden vectors written
 *       by the compiled reference header (tests/golden/), the reference's own
 *       round-trip cases (test/test_frsz_gpu.cu:382-441), and invariants
 *       (lossless at b = W, B = 1; truncation error bound; block locality).
 *   fo_compress_block / fo_decompress_block -- the compiled reference's own
 *       compress_cpu_block / decompress_cpu_block on random blocks, lengths
 *       (partial blocks inside the array included) and buffer contents.
 *   No function is left unpinned.
 *
 * Behaviour that is undefined in the reference (shift counts >= width, which
 * arise only for inf/NaN inputs with doubles, or a block exponent below an
 * element exponent) is made deterministic here by masking the shift count;
 * it is not part of any parity claim.
 */
#include "frsz2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct
{
  int W;    /* bits of the float type */
  int SB;   /* significand bits */
  int EB;   /* exponent bits */
  int bias; /* exponent bias */
} fo_fmt;

static const fo_fmt FMT32 = { 32, 23, 8, 127 };
static const fo_fmt FMT64 = { 64, 52, 11, 1023 };

static inline uint64_t
wmask(int w)
{
  return w >= 64 ? ~UINT64_C(0) : ((UINT64_C(1) << w) - 1);
}

/* storage_t<bits>: include/frsz2.hpp:258-262 */
int
fo_storage_bits(int bits)
{
  return bits <= 8 ? 8 : bits <= 16 ? 16 : bits <= 32 ? 32 : 64;
}

static inline uint64_t
ceildiv_u64(uint64_t a, uint64_t b)
{
  return a / b + (a % b != 0);
}

/* compressed_block_size_byte: include/frsz2.hpp:524-529 */
uint64_t
fo_block_payload_bytes(int bits, uint64_t B)
{
  const uint64_t U = (uint64_t)fo_storage_bits(bits);
  return ceildiv_u64(B * (uint64_t)bits, U) * (U / 8);
}

/* compute_exponent_block_size_byte: include/frsz2.hpp:534-540 (always adds 1..256 B) */
uint64_t
fo_exponent_region_bytes(int bits, uint64_t B, uint64_t n)
{
  const uint64_t x = ceildiv_u64(n, B) * (uint64_t)(fo_storage_bits(bits) / 8);
  return x + 256 - x % 256;
}

/* compute_compressed_memory_size_byte: include/frsz2.hpp:542-548 */
uint64_t
fo_compressed_bytes(int bits, uint64_t B, uint64_t n)
{
  return fo_exponent_region_bytes(bits, B, n) + ceildiv_u64(n, B) * fo_block_payload_bytes(bits, B);
}

/* ---- floating-point decomposition, include/frsz2.hpp:312-338 ---- */

static inline uint64_t
biased_of(uint64_t u, const fo_fmt* f)
{
  return (u >> f->SB) & ((UINT64_C(1) << f->EB) - 1);
}

/* exponent(): biased - bias, returned as exponent_t which is int8 for float
 * (include/frsz2.hpp:280, so 255-127 wraps to -128) and int16 for double. */
static inline int32_t
exp_of(uint64_t u, const fo_fmt* f)
{
  const uint64_t biased = biased_of(u, f);
  if (f->W == 32) {
    return (int8_t)(uint8_t)(biased - 127u);
  }
  return (int16_t)(biased - 1023u);
}

/* significand(): fraction plus the implicit bit unless exponent() == -bias */
static inline uint64_t
signif_of(uint64_t u, const fo_fmt* f)
{
  uint64_t s = u & wmask(f->SB);
  if (exp_of(u, f) != -f->bias) {
    s |= UINT64_C(1) << f->SB;
  }
  return s;
}

/* floating_to_fixed(): include/frsz2.hpp:352-376 */
static inline uint64_t
to_fixed(uint64_t u, int block_exponent, const fo_fmt* f)
{
  const uint64_t s = (u >> (f->W - 1)) & 1u;
  const int e = exp_of(u, f);
  const uint64_t m = signif_of(u, f);
  const int shift = (f->EB - 1) - (block_exponent - e);
  uint64_t r;
  if (-f->SB > shift) {
    r = 0;
  } else if (shift > 0) {
    r = (m << (shift & (f->W - 1))) & wmask(f->W);
  } else {
    r = m >> ((-shift) & (f->W - 1));
  }
  if (s) {
    r |= UINT64_C(1) << (f->W - 1);
  }
  return r;
}

/* countl_zero as the reference's host build on LP64 Linux evaluates it
 * (include/frsz2.hpp:170-192): zero gives 32 for both uint32 and uint64,
 * because the unsigned long overload returns CHAR_BIT*sizeof(unsigned int). */
static inline int
clz_ref(uint64_t v, int W)
{
  if (v == 0) {
    return 32;
  }
  return W == 32 ? __builtin_clz((uint32_t)v) : __builtin_clzll(v);
}

/* fixed_to_floating(): include/frsz2.hpp:378-399 */
static inline uint64_t
from_fixed(uint64_t fixed, int64_t block_exponent, const fo_fmt* f)
{
  const uint64_t W = wmask(f->W);
  const uint64_t mag = fixed & (W >> 1);
  const int z = clz_ref(mag, f->W) - 1;
  const int eb = (int)(block_exponent + f->bias);
  const int shift = -(z < eb ? z : eb);
  const uint64_t s = fixed & (UINT64_C(1) << (f->W - 1));
  const uint64_t e = ((uint64_t)((int64_t)shift + block_exponent + f->bias) << f->SB) & W;
  const int ms = f->EB - 1 + shift;
  uint64_t m;
  if (ms > 0) {
    m = mag >> (ms & (f->W - 1));
  } else {
    m = (mag << ((-ms) & (f->W - 1))) & W;
  }
  m &= ~(UINT64_C(1) << f->SB);
  return (s | e | m) & W;
}

static inline uint64_t
load_bits(const void* p, uint64_t i, int W)
{
  if (W == 32) {
    uint32_t v;
    memcpy(&v, (const uint8_t*)p + 4 * i, 4);
    return v;
  }
  uint64_t v;
  memcpy(&v, (const uint8_t*)p + 8 * i, 8);
  return v;
}

static inline void
store_bits(void* p, uint64_t i, int W, uint64_t v)
{
  if (W == 32) {
    uint32_t t = (uint32_t)v;
    memcpy((uint8_t*)p + 4 * i, &t, 4);
  } else {
    memcpy((uint8_t*)p + 8 * i, &v, 8);
  }
}

/* little-endian U-bit word access */
static inline uint64_t
load_word(const uint8_t* p, int ubytes)
{
  uint64_t v = 0;
  memcpy(&v, p, (size_t)ubytes);
  return v;
}

static inline void
store_word(uint8_t* p, int ubytes, uint64_t v)
{
  memcpy(p, &v, (size_t)ubytes);
}

/* exp_type is a signed U-bit integer (include/frsz2.hpp:510): the exponent()
 * value is converted to it (int8 truncation for double with bits <= 8). */
static inline int64_t
to_exp_type(int64_t e, int U)
{
  switch (U) {
    case 8:
      return (int8_t)(uint8_t)(uint64_t)e;
    case 16:
      return (int16_t)(uint16_t)(uint64_t)e;
    case 32:
      return (int32_t)(uint32_t)(uint64_t)e;
    default:
      return e;
  }
}

static inline int64_t
load_exp(const uint8_t* exps, uint64_t k, int U)
{
  const uint64_t raw = load_word(exps + k * (uint64_t)(U / 8), U / 8);
  switch (U) {
    case 8:
      return (int8_t)(uint8_t)raw;
    case 16:
      return (int16_t)(uint16_t)raw;
    case 32:
      return (int32_t)(uint32_t)raw;
    default:
      return (int64_t)raw;
  }
}

typedef struct
{
  const fo_fmt* f;
  int bits;
  int U;
  uint64_t B;
  uint64_t CB; /* payload bytes per block */
  uint64_t R;  /* exponent region bytes incl. pad */
  uint64_t n;
  uint64_t nb;
} fo_layout;

static int
make_layout(int dtype, int bits, uint64_t B, uint64_t n, fo_layout* L)
{
  if (dtype != 0 && dtype != 1) {
    return 1;
  }
  L->f = dtype == 0 ? &FMT32 : &FMT64;
  /* include/frsz2.hpp:511-517: 4 <= bits <= min(64, W), B > 0 */
  if (bits < 4 || bits > L->f->W || B == 0) {
    return 2;
  }
  L->bits = bits;
  L->U = fo_storage_bits(bits);
  L->B = B;
  L->CB = fo_block_payload_bytes(bits, B);
  L->R = fo_exponent_region_bytes(bits, B, n);
  L->n = n;
  L->nb = ceildiv_u64(n, B);
  return 0;
}

/* compress_cpu_block(exp_block_id = k, data = vals, exp_block_elements =
 * len): include/frsz2.hpp:1108-1179.  vals points at the block's first value.
 * words: scratch of CB/(U/8) entries. */
static void
compress_block(const fo_layout* L, uint64_t k, const void* vals, uint64_t len, uint8_t* out, uint64_t* words)
{
  const fo_fmt* f = L->f;
  const int U = L->U;
  const int ub = U / 8;

  /* max |x| with std::max(in_max, std::abs(x)) from 0 (include/frsz2.hpp:1121-1124):
   * in_max is replaced only when in_max < |x|, so NaN never enters. */
  uint64_t maxbits;
  if (f->W == 32) {
    float in_max = 0.0f;
    for (uint64_t i = 0; i < len; ++i) {
      float x;
      memcpy(&x, (const uint8_t*)vals + 4 * i, 4);
      const float a = fabsf(x);
      if (in_max < a) {
        in_max = a;
      }
    }
    uint32_t t;
    memcpy(&t, &in_max, 4);
    maxbits = t;
  } else {
    double in_max = 0.0;
    for (uint64_t i = 0; i < len; ++i) {
      double x;
      memcpy(&x, (const uint8_t*)vals + 8 * i, 8);
      const double a = fabs(x);
      if (in_max < a) {
        in_max = a;
      }
    }
    memcpy(&maxbits, &in_max, 8);
  }
  const int64_t E = to_exp_type(exp_of(maxbits, f), U);
  store_word(out + k * (uint64_t)ub, ub, (uint64_t)E);

  /* floating_to_fixed takes the block exponent as int16 (include/frsz2.hpp:354) */
  const int E16 = (int16_t)(uint16_t)(uint64_t)E;
  const uint64_t nwords = L->CB / (uint64_t)ub;
  memset(words, 0, nwords * sizeof(uint64_t));
  const uint64_t umask = wmask(U);
  for (uint64_t j = 0; j < len; ++j) {
    const uint64_t fixed = to_fixed(load_bits(vals, j, f->W), E16, f);
    /* uint_fp_to_compressed: top `bits` bits (include/frsz2.hpp:555-558) */
    const uint64_t code = (fixed >> (f->W - L->bits)) & umask;
    /* write_shift_value: include/frsz2.hpp:629-644 */
    const uint64_t p = j * (uint64_t)L->bits;
    const uint64_t w = p / (uint64_t)U;
    const int o = (int)(p % (uint64_t)U);
    words[w] |= (code << o) & umask;
    if (o + L->bits > U) {
      words[w + 1] |= (code >> (U - o)) & umask;
    }
  }
  /* the reference stores exactly words [0, ceil(len*bits/U)) of the block
   * (loop stores at include/frsz2.hpp:1155-1163, tail word at :1166-1178) */
  const uint64_t touched = ceildiv_u64(len * (uint64_t)L->bits, (uint64_t)U);
  uint8_t* blk = out + L->R + k * L->CB;
  for (uint64_t w = 0; w < touched; ++w) {
    store_word(blk + w * (uint64_t)ub, ub, words[w]);
  }
}

/* retrieve_shift_value + compressed_to_uint_fp: include/frsz2.hpp:550-553, 615-627 */
static inline uint64_t
fetch_fixed(const fo_layout* L, const uint8_t* blk, uint64_t j)
{
  const int U = L->U;
  const int ub = U / 8;
  const uint64_t p = j * (uint64_t)L->bits;
  const uint64_t w = p / (uint64_t)U;
  const int o = (int)(p % (uint64_t)U);
  uint64_t res = load_word(blk + w * (uint64_t)ub, ub) >> o;
  if (o + L->bits > U) {
    res |= load_word(blk + (w + 1) * (uint64_t)ub, ub) << (U - o);
  }
  res &= wmask(U);
  return (res << (L->f->W - L->bits)) & wmask(L->f->W);
}

/* decompress_cpu_block(exp_block_id = k, output = vals): include/frsz2.hpp:1217-1240.
 * vals receives the block's min(B, n - k*B) values. */
static void
decompress_block(const fo_layout* L, uint64_t k, const uint8_t* in, void* vals)
{
  const uint64_t first = k * L->B;
  const uint64_t len = (L->n - first) < L->B ? (L->n - first) : L->B;
  const int64_t E = load_exp(in, k, L->U);
  const uint8_t* blk = in + L->R + k * L->CB;
  for (uint64_t j = 0; j < len; ++j) {
    store_bits(vals, j, L->f->W, from_fixed(fetch_fixed(L, blk, j), E, L->f));
  }
}

/* compress_cpu_block(exp_block_id, data, exp_block_elements) of a view over
 * n values (include/frsz2.hpp:1108-1179): the reference asserts
 * exp_block_id * B + exp_block_elements <= n (:1115); here that is checked
 * (returns 3) and exp_block_elements must also be 1..B. */
int
fo_compress_block(int dtype, int bits, uint64_t B, uint64_t n, uint64_t k, const void* vals, uint64_t len,
                  uint8_t* out)
{
  fo_layout L;
  const int rc = make_layout(dtype, bits, B, n, &L);
  if (rc) {
    return rc;
  }
  if (len == 0 || len > B || k >= L.nb || k * B + len > n) {
    return 3;
  }
  uint64_t* words = (uint64_t*)malloc((L.CB / (uint64_t)(L.U / 8)) * sizeof(uint64_t));
  if (!words) {
    return 4;
  }
  compress_block(&L, k, vals, len, out, words);
  free(words);
  return 0;
}

/* decompress_cpu_block(exp_block_id, output): include/frsz2.hpp:1217-1240;
 * writes min(B, n - k*B) values. */
int
fo_decompress_block(int dtype, int bits, uint64_t B, uint64_t n, uint64_t k, const uint8_t* in, void* vals)
{
  fo_layout L;
  const int rc = make_layout(dtype, bits, B, n, &L);
  if (rc) {
    return rc;
  }
  if (k >= L.nb) {
    return 3;
  }
  decompress_block(&L, k, in, vals);
  return 0;
}

int
fo_compress(int dtype, int bits, uint64_t B, const void* in, uint64_t n, uint8_t* out)
{
  return fo_compress_mt(dtype, bits, B, in, n, out, 1);
}

int
fo_decompress(int dtype, int bits, uint64_t B, const uint8_t* in, uint64_t n, void* out)
{
  return fo_decompress_mt(dtype, bits, B, in, n, out, 1);
}

/* compress_cpu_impl (include/frsz2.hpp:1181-1193) is a serial loop over
 * blocks; blocks are independent, so a static block partition gives the
 * same bytes. */
int
fo_compress_mt(int dtype, int bits, uint64_t B, const void* in, uint64_t n, uint8_t* out, int threads)
{
  fo_layout L;
  const int rc = make_layout(dtype, bits, B, n, &L);
  if (rc) {
    return rc;
  }
  const uint64_t nwords = L.CB / (uint64_t)(L.U / 8);
  int nt = threads;
#ifdef _OPENMP
  if (nt <= 0) {
    nt = omp_get_max_threads();
  }
#else
  nt = 1;
#endif
  int failed = 0;
#pragma omp parallel num_threads(nt)
  {
    uint64_t* words = (uint64_t*)malloc((nwords ? nwords : 1) * sizeof(uint64_t));
    if (!words) {
#pragma omp atomic write
      failed = 1;
    } else {
#pragma omp for schedule(static)
      for (int64_t k = 0; k < (int64_t)L.nb; ++k) {
        const uint64_t first = (uint64_t)k * B;
        const uint64_t len = (n - first) < B ? (n - first) : B;
        compress_block(&L, (uint64_t)k, (const uint8_t*)in + first * (uint64_t)(L.f->W / 8), len, out, words);
      }
    }
    free(words);
  }
  return failed ? 3 : 0;
}

/* decompress_cpu_impl: include/frsz2.hpp:1242-1249 */
int
fo_decompress_mt(int dtype, int bits, uint64_t B, const uint8_t* in, uint64_t n, void* out, int threads)
{
  fo_layout L;
  const int rc = make_layout(dtype, bits, B, n, &L);
  if (rc) {
    return rc;
  }
  int nt = threads;
#ifdef _OPENMP
  if (nt <= 0) {
    nt = omp_get_max_threads();
  }
#else
  nt = 1;
#endif
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t k = 0; k < (int64_t)L.nb; ++k) {
    decompress_block(&L, (uint64_t)k, in, (uint8_t*)out + (uint64_t)k * B * (uint64_t)(L.f->W / 8));
  }
  return 0;
}

/* decompress_cpu_element(): include/frsz2.hpp:1199-1214 */
int
fo_gather(int dtype, int bits, uint64_t B, const uint8_t* in, uint64_t n, const uint64_t* idx, uint64_t m,
          void* out)
{
  fo_layout L;
  const int rc = make_layout(dtype, bits, B, n, &L);
  if (rc) {
    return rc;
  }
  for (uint64_t i = 0; i < m; ++i) {
    const uint64_t k = idx[i] / L.B;
    const uint64_t j = idx[i] % L.B;
    const int64_t E = load_exp(in, k, L.U);
    const uint8_t* blk = in + L.R + k * L.CB;
    store_bits(out, i, L.f->W, from_fixed(fetch_fixed(&L, blk, j), E, L.f));
  }
  return 0;
}

int
fo_max_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---- exported single-value helpers (known-answer tests) ---- */

static uint64_t
f32_bits(float x)
{
  uint32_t u;
  memcpy(&u, &x, 4);
  return u;
}

static uint64_t
f64_bits(double x)
{
  uint64_t u;
  memcpy(&u, &x, 8);
  return u;
}

int32_t
fo_exponent_f32(float x)
{
  return exp_of(f32_bits(x), &FMT32);
}

uint32_t
fo_significand_f32(float x)
{
  return (uint32_t)signif_of(f32_bits(x), &FMT32);
}

int
fo_sign_f32(float x)
{
  return (int)(f32_bits(x) >> 31);
}

int32_t
fo_exponent_f64(double x)
{
  return exp_of(f64_bits(x), &FMT64);
}

uint64_t
fo_significand_f64(double x)
{
  return signif_of(f64_bits(x), &FMT64);
}

int
fo_sign_f64(double x)
{
  return (int)(f64_bits(x) >> 63);
}

uint32_t
fo_floating_to_fixed_f32(float x, int16_t block_exponent)
{
  return (uint32_t)to_fixed(f32_bits(x), block_exponent, &FMT32);
}

uint64_t
fo_floating_to_fixed_f64(double x, int16_t block_exponent)
{
  return to_fixed(f64_bits(x), block_exponent, &FMT64);
}

float
fo_fixed_to_floating_f32(uint32_t fixed, int64_t block_exponent)
{
  const uint32_t r = (uint32_t)from_fixed(fixed, block_exponent, &FMT32);
  float x;
  memcpy(&x, &r, 4);
  return x;
}

double
fo_fixed_to_floating_f64(uint64_t fixed, int64_t block_exponent)
{
  const uint64_t r = from_fixed(fixed, block_exponent, &FMT64);
  double x;
  memcpy(&x, &r, 8);
  return x;
}

void
fo_roundtrip_f32(const uint32_t* u, uint64_t n, int16_t E, int bits, uint32_t* codes, uint32_t* values)
{
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t code = to_fixed(u[i], E, &FMT32) >> (32 - bits);
    codes[i] = (uint32_t)code;
    values[i] = (uint32_t)from_fixed((code << (32 - bits)) & 0xffffffffu, E, &FMT32);
  }
}

void
fo_roundtrip_f64(const uint64_t* u, uint64_t n, int16_t E, int bits, uint64_t* codes, uint64_t* values)
{
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t code = to_fixed(u[i], E, &FMT64) >> (64 - bits);
    codes[i] = code;
    values[i] = from_fixed(bits == 64 ? code : (code << (64 - bits)), E, &FMT64);
  }
}
