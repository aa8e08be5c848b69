// Single entry points over the split reference translation units built from
// oracle/ref_driver.cc (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY.
// dtype: 0 = float32, 1 = float64.  Returns 0, 1 (dtype) or 2 (bits / block
// size not instantiated).
#include <cstdint>

#define REF_DECL(tag)                                                                                      \
  int ref_compressed_size_##tag(int, int, std::uint64_t, std::uint64_t*);                                  \
  int ref_compress_##tag(int, int, const void*, std::uint64_t, std::uint8_t*);                             \
  int ref_decompress_##tag(int, int, const std::uint8_t*, std::uint64_t, void*);                           \
  int ref_gather_##tag(int, int, const std::uint8_t*, std::uint64_t, const std::uint64_t*, std::uint64_t, void*); \
  int ref_compress_block_##tag(int, int, std::uint64_t, std::uint64_t, const void*, std::uint64_t, std::uint8_t*); \
  int ref_decompress_block_##tag(int, int, std::uint64_t, std::uint64_t, const std::uint8_t*, void*);

extern "C" {
REF_DECL(f32a)
REF_DECL(f32b)
REF_DECL(f64a)
REF_DECL(f64b)
REF_DECL(f64c)
REF_DECL(f64d)

#define REF_TRY(call, tag)                                                                                 \
  do {                                                                                                     \
    if (call##_##tag(ARGS) == 0)                                                                           \
      return 0;                                                                                            \
  } while (0)

#define REF_CHAIN(call)                                                                                    \
  if (dtype == 0) {                                                                                        \
    REF_TRY(call, f32a);                                                                                   \
    REF_TRY(call, f32b);                                                                                   \
    return 2;                                                                                              \
  }                                                                                                        \
  if (dtype == 1) {                                                                                        \
    REF_TRY(call, f64a);                                                                                   \
    REF_TRY(call, f64b);                                                                                   \
    REF_TRY(call, f64c);                                                                                   \
    REF_TRY(call, f64d);                                                                                   \
    return 2;                                                                                              \
  }                                                                                                        \
  return 1

int
ref_compressed_size(int dtype, int bits, int B, std::uint64_t n, std::uint64_t* out)
{
#define ARGS bits, B, n, out
  REF_CHAIN(ref_compressed_size);
#undef ARGS
}

int
ref_compress(int dtype, int bits, int B, const void* in, std::uint64_t n, std::uint8_t* out)
{
#define ARGS bits, B, in, n, out
  REF_CHAIN(ref_compress);
#undef ARGS
}

int
ref_decompress(int dtype, int bits, int B, const std::uint8_t* in, std::uint64_t n, void* out)
{
#define ARGS bits, B, in, n, out
  REF_CHAIN(ref_decompress);
#undef ARGS
}

int
ref_gather(int dtype,
           int bits,
           int B,
           const std::uint8_t* in,
           std::uint64_t n,
           const std::uint64_t* idx,
           std::uint64_t m,
           void* out)
{
#define ARGS bits, B, in, n, idx, m, out
  REF_CHAIN(ref_gather);
#undef ARGS
}

int
ref_compress_block(int dtype, int bits, int B, std::uint64_t n, std::uint64_t k, const void* vals, std::uint64_t len,
                   std::uint8_t* out)
{
#define ARGS bits, B, n, k, vals, len, out
  REF_CHAIN(ref_compress_block);
#undef ARGS
}

int
ref_decompress_block(int dtype, int bits, int B, std::uint64_t n, std::uint64_t k, const std::uint8_t* in, void* vals)
{
#define ARGS bits, B, n, k, in, vals
  REF_CHAIN(ref_decompress_block);
#undef ARGS
}
}
