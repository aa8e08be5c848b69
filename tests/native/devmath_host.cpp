// Host build of the value-level device arithmetic (include/frsz2_device.hpp)
// so tests can compare it with the oracle exhaustively on the CPU.  Test
// infrastructure: built into tests/native/libdevmath_host.so by
// __graft_entry__.build(); the product does not use it.
#include "frsz2_device.hpp"

#include <cstdint>

extern "C" {

// codes[i] = encode(u[i], E, bits); values[i] = decode(codes[i], E, bits)
void
dm_roundtrip_f32(const std::uint32_t* u, std::uint64_t n, int E, int bits, std::uint32_t* codes,
                 std::uint32_t* values)
{
  for (std::uint64_t i = 0; i < n; ++i) {
    codes[i] = frsz2::encode_value<float>(u[i], E, bits);
    values[i] = frsz2::decode_value<float>(codes[i], E, bits);
  }
}

void
dm_roundtrip_f64(const std::uint64_t* u, std::uint64_t n, int E, int bits, std::uint64_t* codes,
                 std::uint64_t* values)
{
  for (std::uint64_t i = 0; i < n; ++i) {
    codes[i] = frsz2::encode_value<double>(u[i], E, bits);
    values[i] = frsz2::decode_value<double>(codes[i], E, bits);
  }
}

// out[i] = encode_value_short(u[i], E, bits) / encode_value (the narrow-width encode)
void
dm_encode_short_f32(const std::uint32_t* u, std::uint64_t n, int E, int bits, std::uint32_t* out)
{
  for (std::uint64_t i = 0; i < n; ++i) {
    out[i] = frsz2::encode_value_short<float>(u[i], E, bits);
  }
}

void
dm_encode_f32(const std::uint32_t* u, std::uint64_t n, int E, int bits, std::uint32_t* out)
{
  for (std::uint64_t i = 0; i < n; ++i) {
    out[i] = frsz2::encode_value<float>(u[i], E, bits);
  }
}

void
dm_encode_short_f64(const std::uint64_t* u, std::uint64_t n, int E, int bits, std::uint64_t* out)
{
  for (std::uint64_t i = 0; i < n; ++i) {
    out[i] = frsz2::encode_value_short<double>(u[i], E, bits);
  }
}

void
dm_encode_f64(const std::uint64_t* u, std::uint64_t n, int E, int bits, std::uint64_t* out)
{
  for (std::uint64_t i = 0; i < n; ++i) {
    out[i] = frsz2::encode_value<double>(u[i], E, bits);
  }
}

// out[i] = decode_value_short(codes[i], E, bits) (the narrow-width decode)
void
dm_decode_short_f32(const std::uint32_t* codes, std::uint64_t n, int E, int bits, std::uint32_t* out)
{
  for (std::uint64_t i = 0; i < n; ++i) {
    out[i] = frsz2::decode_value_short<float>(codes[i], E, bits);
  }
}

void
dm_decode_f32(const std::uint32_t* codes, std::uint64_t n, int E, int bits, std::uint32_t* out)
{
  for (std::uint64_t i = 0; i < n; ++i) {
    out[i] = frsz2::decode_value<float>(codes[i], E, bits);
  }
}

void
dm_decode_short_f64(const std::uint64_t* codes, std::uint64_t n, int E, int bits, std::uint64_t* out)
{
  for (std::uint64_t i = 0; i < n; ++i) {
    out[i] = frsz2::decode_value_short<double>(codes[i], E, bits);
  }
}

void
dm_decode_f64(const std::uint64_t* codes, std::uint64_t n, int E, int bits, std::uint64_t* out)
{
  for (std::uint64_t i = 0; i < n; ++i) {
    out[i] = frsz2::decode_value<double>(codes[i], E, bits);
  }
}

// block exponent (as stored) of a block whose values are u[0..n)
int
dm_block_exponent_f32(const std::uint32_t* u, std::uint64_t n, int U)
{
  std::uint32_t m = 0;
  for (std::uint64_t i = 0; i < n; ++i) {
    const std::uint32_t k = frsz2::max_key<float>(u[i]);
    m = m > k ? m : k;
  }
  return U == 8 ? frsz2::block_exponent_from_key<float, 8>(m) : frsz2::block_exponent_from_key<float, 16>(m);
}

int
dm_block_exponent_f64(const std::uint64_t* u, std::uint64_t n, int U)
{
  std::uint32_t m = 0;
  for (std::uint64_t i = 0; i < n; ++i) {
    const std::uint32_t k = frsz2::max_key<double>(u[i]);
    m = m > k ? m : k;
  }
  return U == 8 ? frsz2::block_exponent_from_key<double, 8>(m) : frsz2::block_exponent_from_key<double, 16>(m);
}

} // extern "C"
