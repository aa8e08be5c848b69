// FRSZ2 value-level arithmetic for HIP device code (gfx950), plus a
// device-callable random-access element decoder for consumer kernels.
//
// Stream layout (little-endian, no header; n, dtype, bits and the exponent
// block size B travel out of band), as include/frsz2.hpp:534-548 and
// :1108-1249 of robertu94/frsz define it:
//   [nb exponents, signed U-bit each][zero pad to a 256-byte boundary, 1..256 B]
//   [nb blocks of CB bytes]
// with U = 8/16/32/64 the storage width for `bits`, nb = ceil(n/B),
// CB = ceil(B*bits/U)*U/8.  Inside a block the codes form an LSB-first
// bitstream: code j occupies bits [j*bits, (j+1)*bits).
//
// Everything here is integer arithmetic on the IEEE bit patterns; no
// floating-point operation takes part, so results are bit-exact.
#pragma once

#include <cstdint>

// Every function here is plain integer code on bit patterns and pointers; it
// compiles for the device under hipcc and for the host with any C++17
// compiler (the drop-in header decodes single elements of host streams with
// it).
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FRSZ2_HD __host__ __device__
#define FRSZ2_HDI __host__ __device__ __forceinline__
#else
#define FRSZ2_HD
#define FRSZ2_HDI inline __attribute__((always_inline))
#endif

namespace frsz2 {

template<typename T>
struct fmt;

template<>
struct fmt<float>
{
  using uw = std::uint32_t;
  static constexpr int W = 32;
  static constexpr int SB = 23;
  static constexpr int EB = 8;
  static constexpr int BIAS = 127;
  static constexpr uw ABS = 0x7fffffffu;
  static constexpr uw INF = 0x7f800000u;
};

template<>
struct fmt<double>
{
  using uw = std::uint64_t;
  static constexpr int W = 64;
  static constexpr int SB = 52;
  static constexpr int EB = 11;
  static constexpr int BIAS = 1023;
  static constexpr uw ABS = 0x7fffffffffffffffull;
  static constexpr uw INF = 0x7ff0000000000000ull;
};

// storage_t<bits> of include/frsz2.hpp:258-262
FRSZ2_HD constexpr int
storage_bits(int bits)
{
  return bits <= 8 ? 8 : bits <= 16 ? 16 : bits <= 32 ? 32 : 64;
}

// Block-max key of one value: the biased exponent field of |x|, with NaN
// mapped to 0.  For finite values the biased exponent is monotone in |x|, so
// max(key) is the biased exponent of max|x| that include/frsz2.hpp:1121-1125
// takes with std::max from 0 (which never lets a NaN in).
template<typename T>
FRSZ2_HDI std::uint32_t
max_key(typename fmt<T>::uw u)
{
  using F = fmt<T>;
  if constexpr (sizeof(T) == 4) {
    // NaN as an unordered float compare (one v_cmp_u_f32), the exponent
    // field with one v_bfe
    const float f = __builtin_bit_cast(float, u);
    return f != f ? 0u : static_cast<std::uint32_t>((u >> F::SB) & 0xffu);
  } else {
    const typename F::uw a = u & F::ABS;
    return a > F::INF ? 0u : static_cast<std::uint32_t>(a >> F::SB);
  }
}

// Block exponent as stored, from the block's max key: exponent() returns
// int8 for float (include/frsz2.hpp:280, so 128 wraps to -128) and int16 for
// double; it is then converted to the signed U-bit exp_type (:510), which
// truncates for double with U = 8.
template<typename T, int U>
FRSZ2_HDI int
block_exponent_from_key(std::uint32_t key)
{
  int e = static_cast<int>(key) - fmt<T>::BIAS;
  if (sizeof(T) == 4) {
    e = static_cast<std::int8_t>(static_cast<std::uint8_t>(e));
  }
  if (U == 8) {
    e = static_cast<std::int8_t>(static_cast<std::uint8_t>(e));
  }
  return e;
}

// Code of one value for block exponent E (floating_to_fixed followed by
// uint_fp_to_compressed, include/frsz2.hpp:352-376 and :555-558): the
// significand with its implicit bit, aligned so that a value of exponent E
// has its leading one at bit W-2, truncated to the top `bits` bits, sign in
// bit bits-1.  The two shifts of the reference are folded into one right
// shift by (W - bits) - shift, clamped to W-1 (the clamp only ever turns a
// result that is zero anyway into zero).  Valid for E >= exponent(x), which
// holds for every value of a block whose exponent was taken from its max.
template<typename T>
FRSZ2_HDI typename fmt<T>::uw
encode_value(typename fmt<T>::uw u, int E, int bits)
{
  using F = fmt<T>;
  using uw = typename F::uw;
  const std::uint32_t biased = static_cast<std::uint32_t>((u & F::ABS) >> F::SB);
  int e = static_cast<int>(biased) - F::BIAS;
  if (sizeof(T) == 4) {
    e = static_cast<std::int8_t>(static_cast<std::uint8_t>(e));
  }
  const uw m = (u & ((uw{ 1 } << F::SB) - 1)) | (biased != 0 ? (uw{ 1 } << F::SB) : uw{ 0 });
  const int shift = (F::EB - 1) - (E - e);
  const int tot = (F::W - bits) - shift;
  uw mag;
  if (tot >= 0) {
    mag = m >> (tot < F::W - 1 ? tot : F::W - 1);
  } else {
    mag = m << (-tot);
  }
  const uw s = u >> (F::W - 1);
  return mag | (s << (bits - 1));
}

// encode_value for bits <= SB + 1 (float32 up to 24 bits, float64 up to 53):
// the folded shift (W - bits) - shift is then never negative for a value of
// the block (E >= exponent(x)), so only the right shift and its clamp
// remain.  Same code as encode_value for every such value, NaN included.
template<typename T>
FRSZ2_HDI typename fmt<T>::uw
encode_value_short(typename fmt<T>::uw u, int E, int bits)
{
  using F = fmt<T>;
  using uw = typename F::uw;
  const std::uint32_t biased = static_cast<std::uint32_t>((u & F::ABS) >> F::SB);
  int e = static_cast<int>(biased) - F::BIAS;
  if (sizeof(T) == 4) {
    e = static_cast<std::int8_t>(static_cast<std::uint8_t>(e));
  }
  uw m;
  if constexpr (sizeof(T) == 4) {
    // the implicit bit unless |x| is zero or subnormal, as one float compare
    // (NaN compares false and keeps the bit; a compare that flushes
    // subnormals to zero gives the same answer)
    const bool low = __builtin_fabsf(__builtin_bit_cast(float, u)) < 0x1p-126f;
    m = (u & ((uw{ 1 } << F::SB) - 1)) | (low ? uw{ 0 } : (uw{ 1 } << F::SB));
  } else {
    m = (u & ((uw{ 1 } << F::SB) - 1)) | (biased != 0 ? (uw{ 1 } << F::SB) : uw{ 0 });
  }
  // the per-block part of the folded shift first, so that one subtraction
  // per value remains
  const int k = (F::W - bits) - (F::EB - 1) + E;
  const int tot = k - e;
  const uw mag = m >> (tot < F::W - 1 ? tot : F::W - 1);
  const uw s = u >> (F::W - 1);
  return mag | (s << (bits - 1));
}

// count of leading zeros of a W-bit value known to be non-zero
FRSZ2_HDI int
clz_nz(std::uint32_t v)
{
  return __builtin_clz(v);
}
FRSZ2_HDI int
clz_nz(std::uint64_t v)
{
  return __builtin_clzll(v);
}

// Value of one code for block exponent E (compressed_to_uint_fp followed by
// fixed_to_floating, include/frsz2.hpp:550-553 and :378-399).  The
// reference's host countl_zero returns 32 for a zero magnitude for both
// float and double (include/frsz2.hpp:180-192 on LP64), so a zero code in a
// block with E + bias >= 31 decodes to +-2^(E-31); that is reproduced here.
template<typename T>
FRSZ2_HDI typename fmt<T>::uw
decode_value(typename fmt<T>::uw code, int E, int bits)
{
  using F = fmt<T>;
  using uw = typename F::uw;
  const uw cm = code & ((uw{ 1 } << (bits - 1)) - 1);
  const uw s = (code >> (bits - 1)) & 1u;
  // z = countl_zero(cm << (W-bits)) - 1, with countl_zero(0) == 32
  const int z = cm == 0 ? 31 : clz_nz(cm) - (F::W - bits) - 1;
  const int eb = E + F::BIAS;
  const int sh = -(z < eb ? z : eb);
  const uw ex = static_cast<uw>(static_cast<typename F::uw>(static_cast<std::int64_t>(sh + eb))) << F::SB;
  // mantissa = (cm << (W-bits)) shifted right by EB-1+sh, as one shift of cm
  const int lsh = (F::W - bits) - (F::EB - 1 + sh);
  uw mant;
  if (lsh >= 0) {
    mant = cm << (lsh < F::W - 1 ? lsh : F::W - 1);
  } else {
    mant = cm >> (-lsh);
  }
  mant &= ~(uw{ 1 } << F::SB);
  return (s << (F::W - 1)) | ex | mant;
}

// countl_zero(x) for x != 0 and 31 for x == 0: the reference's host
// countl_zero(0) == 32 less one, for both widths.  For 32 bits this is
// min(v_ffbh, 31) (v_ffbh gives all ones for zero).
FRSZ2_HDI std::uint32_t
clz_zero31(std::uint32_t x)
{
  const std::uint32_t f = x == 0 ? 32u : static_cast<std::uint32_t>(__builtin_clz(x));
  return f < 31u ? f : 31u;
}
FRSZ2_HDI std::uint32_t
clz_zero31(std::uint64_t x)
{
  return x == 0 ? 31u : static_cast<std::uint32_t>(__builtin_clzll(x));
}

// decode_value for bits <= SB + 1 (float32 up to 24 bits, float64 up to 53)
// and blocks with E + bias >= 0 (every block a compress writes).  The
// magnitude is left-aligned at bit W-1 (the sign and any bits of `code`
// above `bits` shift out) and shifted left by the renormalising count z,
// capped at E + bias; the SB bits below bit W-1 are then the mantissa: the
// leading one lands on bit W-1 and drops out (normal) or stays below it
// (denormal).  Sign and exponent are joined before one shift into place.
// The result is decode_value's bit pattern, in about half the operations.
template<typename T>
FRSZ2_HDI typename fmt<T>::uw
decode_value_short(typename fmt<T>::uw code, int E, int bits)
{
  using F = fmt<T>;
  using uw = typename F::uw;
  const uw top = code << (F::W + 1 - bits);
  const std::uint32_t eb = static_cast<std::uint32_t>(E + F::BIAS);
  // z = countl_zero(magnitude << (W - bits)) - 1, 31 for a zero magnitude
  const std::uint32_t z = clz_zero31(top);
  const std::uint32_t mz = z < eb ? z : eb;
  const uw mant = ((top << mz) >> F::EB) & ((uw{ 1 } << F::SB) - 1);
  const uw sign = (code >> (bits - 1)) & 1u;
  const uw se = static_cast<uw>(eb - mz) | (sign << (F::W - 1 - F::SB));
  return (se << F::SB) | mant;
}

// Sign-extended exponent of block k (exponents start at the stream base).
template<int U>
FRSZ2_HDI int
load_block_exponent(const std::uint8_t* __restrict__ stream, std::uint64_t k)
{
  if (U == 8) {
    return static_cast<std::int8_t>(stream[k]);
  } else if (U == 16) {
    return reinterpret_cast<const std::int16_t*>(stream)[k];
  } else if (U == 32) {
    return reinterpret_cast<const std::int32_t*>(stream)[k];
  } else {
    return static_cast<int>(reinterpret_cast<const std::int64_t*>(stream)[k]);
  }
}

// signed U-bit integer: the stored block exponent (exp_type, include/frsz2.hpp:510)
template<int U>
struct signed_word;
template<>
struct signed_word<8>
{
  using type = std::int8_t;
};
template<>
struct signed_word<16>
{
  using type = std::int16_t;
};
template<>
struct signed_word<32>
{
  using type = std::int32_t;
};
template<>
struct signed_word<64>
{
  using type = std::int64_t;
};

template<int U>
struct word;
template<>
struct word<8>
{
  using type = std::uint8_t;
};
template<>
struct word<16>
{
  using type = std::uint16_t;
};
template<>
struct word<32>
{
  using type = std::uint32_t;
};
template<>
struct word<64>
{
  using type = std::uint64_t;
};

// *p, on the device with the non-temporal hint when NT (data read once at a
// random place: keeping it out of the caches leaves them to data that is
// read again)
template<bool NT, typename V>
FRSZ2_HDI V
load_hinted(const V* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (NT) {
    return __builtin_nontemporal_load(p);
  }
#endif
  return *p;
}

// Code j of a block whose payload starts at blk: the U-bit word holding its
// first bit, plus the next word only when the code straddles the boundary
// (retrieve_shift_value, include/frsz2.hpp:615-627).  Never reads outside the
// block.  NT: the word loads carry the non-temporal hint.
template<int U, bool NT = false>
FRSZ2_HDI std::uint64_t
load_code(const std::uint8_t* __restrict__ blk, std::uint64_t j, int bits)
{
  using wt = typename word<U>::type;
  const std::uint64_t p = j * static_cast<std::uint64_t>(bits);
  const std::uint64_t w = p / U;
  const int o = static_cast<int>(p % U);
  const wt* words = reinterpret_cast<const wt*>(blk);
  std::uint64_t r = static_cast<std::uint64_t>(load_hinted<NT>(words + w)) >> o;
  if (o + bits > U) {
    r |= static_cast<std::uint64_t>(load_hinted<NT>(words + w + 1)) << (U - o);
  }
  return bits == 64 ? r : (r & ((std::uint64_t{ 1 } << bits) - 1));
}

// Layout of one stream; host and device.
struct layout
{
  std::uint64_t n;   // values
  std::uint64_t B;   // exponent block size
  std::uint64_t nb;  // blocks = ceil(n / B)
  std::uint64_t R;   // exponent region incl. pad (multiple of 256)
  std::uint64_t CB;  // payload bytes per block
  int bits;
  int U;

  FRSZ2_HD static layout make(int bits, std::uint64_t B, std::uint64_t n)
  {
    layout L{};
    L.n = n;
    L.B = B;
    L.bits = bits;
    L.U = storage_bits(bits);
    L.nb = n / B + (n % B != 0);
    const std::uint64_t x = L.nb * static_cast<std::uint64_t>(L.U / 8);
    L.R = x + 256 - x % 256;
    const std::uint64_t bb = B * static_cast<std::uint64_t>(bits);
    L.CB = (bb / L.U + (bb % L.U != 0)) * static_cast<std::uint64_t>(L.U / 8);
    return L;
  }
  FRSZ2_HD std::uint64_t total_bytes() const { return R + nb * CB; }
};

// Exponent slot k of a stream: the block exponent as a signed U-bit integer
// (exp_type of include/frsz2.hpp:510, written at :1128).
template<int U>
FRSZ2_HDI void
store_block_exponent(std::uint8_t* __restrict__ stream, std::uint64_t k, int e)
{
  using wt = typename word<U>::type;
  reinterpret_cast<wt*>(stream)[k] = static_cast<wt>(static_cast<std::int64_t>(e));
}

// In-kernel compress of exponent block k by one thread (compress_cpu_block,
// include/frsz2.hpp:1108-1179): `values` holds the block's first `count`
// inputs (count <= B; the stream's last block has n - k*B).  Writes exponent
// slot k and all CB payload bytes of block k; codes past `count` are zero.
// The exponent pad is not touched.  T = float/double, U = storage_bits(bits).
template<typename T, int U>
FRSZ2_HD inline void
encode_block(std::uint8_t* __restrict__ stream, const layout& L, std::uint64_t k, const T* __restrict__ values,
             std::uint64_t count)
{
  using uw = typename fmt<T>::uw;
  using wt = typename word<U>::type;
  std::uint32_t m = 0;
  for (std::uint64_t i = 0; i < count; ++i) {
    const std::uint32_t key = max_key<T>(__builtin_bit_cast(uw, values[i]));
    m = m > key ? m : key;
  }
  const int E = block_exponent_from_key<T, U>(m);
  store_block_exponent<U>(stream, k, E);
  wt* words = reinterpret_cast<wt*>(stream + L.R + k * L.CB);
  const std::uint64_t nwords = L.CB / (U / 8);
  const std::uint64_t bits = static_cast<std::uint64_t>(L.bits);
  for (std::uint64_t w = 0; w < nwords; ++w) {
    // the codes that overlap bits [w*U, w*U + U) of the block
    const std::uint64_t lo = w * U;
    const std::uint64_t j0 = lo / bits;
    std::uint64_t j1 = (lo + U + bits - 1) / bits;
    j1 = j1 < count ? j1 : count;
    std::uint64_t acc = 0;
    for (std::uint64_t j = j0; j < j1; ++j) {
      const std::uint64_t c = static_cast<std::uint64_t>(encode_value<T>(__builtin_bit_cast(uw, values[j]), E, L.bits));
      const std::int64_t sh = static_cast<std::int64_t>(j * bits) - static_cast<std::int64_t>(lo);
      acc |= sh >= 0 ? (c << sh) : (c >> (-sh));
    }
    words[w] = static_cast<wt>(acc);
  }
}

// In-kernel decode of the first `count` values of exponent block k by one
// thread (decompress_cpu_block, include/frsz2.hpp:1217-1240).
template<typename T, int U>
FRSZ2_HD inline void
decode_block(const std::uint8_t* __restrict__ stream, const layout& L, std::uint64_t k, T* __restrict__ out,
             std::uint64_t count)
{
  const int E = load_block_exponent<U>(stream, k);
  const std::uint8_t* blk = stream + L.R + k * L.CB;
  for (std::uint64_t j = 0; j < count; ++j) {
    const std::uint64_t code = load_code<U>(blk, j, L.bits);
    out[j] = __builtin_bit_cast(T, decode_value<T>(static_cast<typename fmt<T>::uw>(code), E, L.bits));
  }
}

// Random-access decode of element idx of a stream (decompress_cpu_element,
// include/frsz2.hpp:1199-1214), callable from any kernel.  T = float/double,
// U = storage_bits(L.bits).  NT: the code word is loaded with the
// non-temporal hint (for gathers over streams much larger than the caches:
// the codes, touched once each, then do not evict the block exponents,
// which neighbouring indices share).
template<typename T, int U, bool NT = false>
FRSZ2_HDI T
decode_element(const std::uint8_t* __restrict__ stream, const layout& L, std::uint64_t idx)
{
  // a shift for power-of-two B (the branch is uniform when L is), else a
  // 64-bit division
  const std::uint64_t k = (L.B & (L.B - 1)) == 0 ? idx >> __builtin_ctzll(L.B) : idx / L.B;
  const std::uint64_t j = idx - k * L.B;
  const int E = load_block_exponent<U>(stream, k);
  const std::uint64_t code = load_code<U, NT>(stream + L.R + k * L.CB, j, L.bits);
  const auto u = decode_value<T>(static_cast<typename fmt<T>::uw>(code), E, L.bits);
  return __builtin_bit_cast(T, u);
}

} // namespace frsz2
