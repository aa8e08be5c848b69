// Host-only checks of the drop-in header (include/frsz2.hpp), compiled with
// g++ (no hipcc): the header must serve plain C++ host code, as the
// reference's header does.  Test infrastructure, built into
// tests/native/libheader_host.so by tests/native/build.py; tests/test_header_host.py
// runs it on every golden vector.
//
// hh_check(dtype, bits, B, stream, n, y) checks, for the view over a golden
// stream of n values whose decode is y:
//   1. decompress_cpu_element(i) == y[i] for every i (bit patterns), and 0
//      for i == n;
//   2. the public static helpers: for every i, the code that
//      retrieve_shift_value reads at value i's word and bit offset (both
//      overloads), taken back to the code by uint_fp_to_compressed and
//      decoded with the block's exponent, is y[i]; and re-packing every code
//      with compressed_to_uint_fp + write_shift_value into zeroed words gives
//      the stream's payload bytes;
//   3. decompress_gpu_value(exponent, code) and
//      decompress_gpu_element_impl(payload words, exponents, i) are y[i].
// Returns 0 or the number of the failed check (1-5), with the index of the
// first failing value in *where.
#include "frsz2.hpp"

#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

namespace {

template<int bits, int B, class T>
int
check_case(const std::uint8_t* stream, std::uint64_t n, const void* y_raw, std::uint64_t* where)
{
  using C = frsz::frsz2_compressor<bits, B, T>;
  using uw = typename frsz2::fmt<T>::uw;
  using word_t = typename C::uint_compressed_type;
  constexpr int U = C::uint_compressed_size_bit;
  const uw* y = static_cast<const uw*>(y_raw);
  std::vector<std::uint8_t> buf(stream, stream + C::compute_compressed_memory_size_byte(n));
  const C view(buf.data(), n);

  for (std::uint64_t i = 0; i < n; ++i) {
    const T e = view.decompress_cpu_element(i);
    if (std::memcmp(&e, &y[i], sizeof(T)) != 0) {
      *where = i;
      return 1;
    }
  }
  const T past = view.decompress_cpu_element(n);
  if (past != T(0)) {
    *where = n;
    return 1;
  }

  const std::uint64_t nb = n / B + (n % B != 0);
  const std::uint64_t wpb = C::compressed_block_size_byte / sizeof(word_t);
  std::vector<word_t> repacked(nb * wpb, word_t{ 0 });
  const word_t* words = view.get_const_compressed_data();
  const auto* exps = view.get_const_exponent_data();
  for (std::uint64_t i = 0; i < n; ++i) {
    const std::uint64_t k = i / B;
    const std::uint64_t p = (i % B) * static_cast<std::uint64_t>(bits);
    const std::uint64_t w = k * wpb + p / U;
    const int o = static_cast<int>(p % U);
    const bool spills = o + bits > U;
    const typename C::uint_fp_type a = C::retrieve_shift_value(o, words + w);
    const typename C::uint_fp_type b = C::retrieve_shift_value(o, words[w], spills ? words[w + 1] : word_t{ 0 });
    if (a != b) {
      *where = i;
      return 2;
    }
    const word_t code = C::uint_fp_to_compressed(a);
    const uw v = frsz2::decode_value<T>(static_cast<uw>(code), static_cast<int>(exps[k]), bits);
    if (v != y[i]) {
      *where = i;
      return 3;
    }
    const T gv = C::decompress_gpu_value(exps[k], code);
    const T ge = C::decompress_gpu_element_impl(words, exps, i);
    if (std::memcmp(&gv, &y[i], sizeof(T)) != 0 || std::memcmp(&ge, &y[i], sizeof(T)) != 0) {
      *where = i;
      return 5;
    }
    word_t first{ 0 };
    word_t second{ 0 };
    C::write_shift_value(o, C::compressed_to_uint_fp(code), first, second);
    repacked[w] |= first;
    if (spills) {
      repacked[w + 1] |= second;
    }
  }
  if (std::memcmp(repacked.data(), words, repacked.size() * sizeof(word_t)) != 0) {
    *where = 0;
    return 4;
  }
  return 0;
}

// the (bits, B) pairs of the golden vectors: bits from these lists, B from
// {1, 2, 3, 4, 5, 7, 8, 9, 16, ..., 1024}
template<int... V>
struct ints
{};
using blocks = ints<1, 2, 3, 4, 5, 7, 8, 9, 16, 32, 64, 128, 256, 512, 1024>;
using bits32 = ints<4, 5, 7, 8, 9, 12, 15, 16, 17, 21, 24, 31, 32>;
using bits64 = ints<4, 5, 7, 8, 9, 12, 15, 16, 17, 21, 24, 31, 32, 33, 40, 47, 52, 53, 60, 63, 64>;

template<class T, int bits, int... Bs>
int
by_block(ints<Bs...>, int B, const std::uint8_t* s, std::uint64_t n, const void* y, std::uint64_t* where)
{
  int rc = -1;
  using expand = int[];
  (void)expand{ 0, (B == Bs ? (rc = check_case<bits, Bs, T>(s, n, y, where), 0) : 0)... };
  return rc;
}

template<class T, int... Bits>
int
by_bits(ints<Bits...>, int bits, int B, const std::uint8_t* s, std::uint64_t n, const void* y, std::uint64_t* where)
{
  int rc = -1;
  using expand = int[];
  (void)expand{ 0, (bits == Bits ? (rc = by_block<T, Bits>(blocks{}, B, s, n, y, where), 0) : 0)... };
  return rc;
}

} // namespace

extern "C" {

// -1: (dtype, bits, B) not instantiated here
int
hh_check(int dtype, int bits, int B, const std::uint8_t* stream, std::uint64_t n, const void* y, std::uint64_t* where)
{
  return dtype == 0 ? by_bits<float>(bits32{}, bits, B, stream, n, y, where)
                    : by_bits<double>(bits64{}, bits, B, stream, n, y, where);
}

// compute_compressed_memory_size_byte and the constants, host-evaluated
std::uint64_t
hh_size_f32_b16_B32(std::uint64_t n)
{
  return frsz::frsz2_compressor<16, 32, float>::compute_compressed_memory_size_byte(n);
}

} // extern "C"
