// Self-describing container around an unchanged FRSZ2 stream.
//
// The reference stream has no header: n, dtype, bits and B travel out of
// band (src/frsz2.cc:152-186; the decompressor takes n and the dtype from
// its output buffer, :174-179).  The container prepends one 64-byte header
// and leaves the stream bytes exactly as frsz2_compress writes them:
//
//   [0, 8)    magic "FRSZ2CT\0"
//   [8, 12)   u32 header version (1)
//   [12, 16)  u32 dtype (0 float32, 1 float64)
//   [16, 20)  u32 bits
//   [20, 24)  u32 header bytes (64): offset of the stream
//   [24, 32)  u64 exponent block size B
//   [32, 40)  u64 n
//   [40, 48)  u64 stream bytes (compute_compressed_memory_size_byte(n))
//   [48, 64)  zero
//
// all little-endian.  The stream starts 64-byte aligned when the container
// does, so the tile kernels apply to it unchanged.
#include "frsz2_capi.h"
#include "frsz2_device.hpp"
#include "frsz2_internal.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

namespace frsz2 {

namespace {

constexpr std::uint64_t HDR = FRSZ2_CONTAINER_HEADER_BYTES;
constexpr char MAGIC[8] = { 'F', 'R', 'S', 'Z', '2', 'C', 'T', '\0' };
constexpr std::uint32_t VERSION = 1;

struct header_words
{
  std::uint64_t w[8];
};

header_words
make_header(int dtype, int bits, std::uint64_t block, std::uint64_t n, std::uint64_t stream_bytes)
{
  header_words h{};
  std::memcpy(&h.w[0], MAGIC, 8);
  h.w[1] = static_cast<std::uint64_t>(VERSION) | (static_cast<std::uint64_t>(dtype) << 32);
  h.w[2] = static_cast<std::uint64_t>(bits) | (HDR << 32);
  h.w[3] = block;
  h.w[4] = n;
  h.w[5] = stream_bytes;
  return h;
}

// one thread writes the 64 header bytes (kernel arguments, so the write is
// stream-ordered and capturable without a host copy)
__global__ void
k_write_header(header_words h, std::uint64_t* __restrict__ out)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      out[i] = h.w[i];
    }
  }
}

inline std::uint32_t
rd32(const std::uint8_t* p)
{
  std::uint32_t v;
  std::memcpy(&v, p, 4);
  return v;
}

inline std::uint64_t
rd64(const std::uint8_t* p)
{
  std::uint64_t v;
  std::memcpy(&v, p, 8);
  return v;
}

} // namespace

} // namespace frsz2

extern "C" {

int
frsz2_container_size(int dtype, int bits, uint64_t block, uint64_t n, uint64_t* bytes)
{
  using namespace frsz2;
  clear_error();
  const int rc = check_params(dtype, bits, block);
  if (rc) {
    return rc;
  }
  if (!bytes) {
    return set_error(FRSZ2_EARG, "null output pointer");
  }
  *bytes = HDR + layout::make(bits, block, n).total_bytes();
  return FRSZ2_OK;
}

int
frsz2_container_parse(const void* h_container, uint64_t bytes, frsz2_container_info* info)
{
  using namespace frsz2;
  clear_error();
  if (!h_container || !info) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  if (bytes < HDR) {
    return set_error(FRSZ2_EFORMAT, "container shorter than its 64-byte header");
  }
  const auto* p = static_cast<const std::uint8_t*>(h_container);
  if (std::memcmp(p, MAGIC, 8) != 0) {
    return set_error(FRSZ2_EFORMAT, "not an FRSZ2 container (bad magic)");
  }
  const std::uint32_t version = rd32(p + 8);
  if (version != VERSION || rd32(p + 20) != HDR || rd64(p + 48) != 0 || rd64(p + 56) != 0) {
    return set_error(FRSZ2_EFORMAT, "unsupported container version " + std::to_string(version));
  }
  const int dtype = static_cast<int>(rd32(p + 12));
  const int bits = static_cast<int>(rd32(p + 16));
  const std::uint64_t block = rd64(p + 24);
  const int rc = check_params(dtype, bits, block);
  if (rc) {
    return rc;
  }
  const std::uint64_t n = rd64(p + 32);
  const std::uint64_t stream_bytes = rd64(p + 40);
  if (stream_bytes != layout::make(bits, block, n).total_bytes()) {
    return set_error(FRSZ2_EFORMAT, "container header: stream size does not match n, bits and block");
  }
  info->dtype = dtype;
  info->bits = bits;
  info->block = block;
  info->n = n;
  info->stream_bytes = stream_bytes;
  info->header_bytes = HDR;
  return FRSZ2_OK;
}

int
frsz2_compress_container(int dtype, int bits, uint64_t block, const void* d_in, uint64_t n, void* d_out,
                         uint64_t out_capacity, frsz2_stream_t stream)
{
  using namespace frsz2;
  clear_error();
  const int rc = check_params(dtype, bits, block);
  if (rc) {
    return rc;
  }
  const std::uint64_t total = layout::make(bits, block, n).total_bytes();
  if (out_capacity < HDR + total) {
    return set_error(FRSZ2_ECAPACITY, "output capacity " + std::to_string(out_capacity) + " < container size " +
                                        std::to_string(HDR + total));
  }
  if (!d_out || (reinterpret_cast<std::uintptr_t>(d_out) & 7u) != 0) {
    return set_error(FRSZ2_EARG, "null or misaligned pointer");
  }
  const hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_write_header, dim3(1), dim3(64), 0, s, make_header(dtype, bits, block, n, total),
                     static_cast<std::uint64_t*>(d_out));
  const int hrc = hip_status(hipGetLastError(), "container header launch");
  if (hrc) {
    return hrc;
  }
  return compress_device(dtype, bits, block, d_in, n, static_cast<std::uint8_t*>(d_out) + HDR, total, s);
}

int
frsz2_decompress_container(const frsz2_container_info* info, const void* d_in, uint64_t in_bytes, void* d_out,
                           uint64_t out_values, frsz2_stream_t stream)
{
  using namespace frsz2;
  clear_error();
  if (!info) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  if (in_bytes < HDR + info->stream_bytes) {
    return set_error(FRSZ2_ECAPACITY, "container of " + std::to_string(in_bytes) + " bytes < header + stream " +
                                        std::to_string(HDR + info->stream_bytes));
  }
  if (out_values < info->n) {
    return set_error(FRSZ2_ECAPACITY, "output holds " + std::to_string(out_values) + " values < n = " +
                                        std::to_string(info->n));
  }
  return decompress_device(info->dtype, info->bits, info->block, static_cast<const std::uint8_t*>(d_in) + HDR,
                           in_bytes - HDR, info->n, d_out, static_cast<hipStream_t>(stream));
}

int
frsz2_compress_container_host(int dtype, int bits, uint64_t block, const void* h_in, uint64_t n, void* h_out,
                              uint64_t out_capacity)
{
  using namespace frsz2;
  clear_error();
  const int rc = check_params(dtype, bits, block);
  if (rc) {
    return rc;
  }
  const std::uint64_t total = layout::make(bits, block, n).total_bytes();
  if (out_capacity < HDR + total) {
    return set_error(FRSZ2_ECAPACITY, "output capacity " + std::to_string(out_capacity) + " < container size " +
                                        std::to_string(HDR + total));
  }
  if (!h_out) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  const int crc =
    frsz2_compress_host(dtype, bits, block, h_in, n, static_cast<std::uint8_t*>(h_out) + HDR, out_capacity - HDR);
  if (crc) {
    return crc;
  }
  const header_words h = make_header(dtype, bits, block, n, total);
  std::memcpy(h_out, h.w, HDR);
  return FRSZ2_OK;
}

int
frsz2_decompress_container_host(const void* h_in, uint64_t in_bytes, void* h_out, uint64_t out_values)
{
  using namespace frsz2;
  frsz2_container_info info{};
  const int rc = frsz2_container_parse(h_in, in_bytes, &info);
  if (rc) {
    return rc;
  }
  if (in_bytes < HDR + info.stream_bytes) {
    return set_error(FRSZ2_ECAPACITY, "container of " + std::to_string(in_bytes) + " bytes < header + stream " +
                                        std::to_string(HDR + info.stream_bytes));
  }
  if (out_values < info.n) {
    return set_error(FRSZ2_ECAPACITY,
                     "output holds " + std::to_string(out_values) + " values < n = " + std::to_string(info.n));
  }
  return frsz2_decompress_host(info.dtype, info.bits, info.block, static_cast<const std::uint8_t*>(h_in) + HDR,
                               in_bytes - HDR, info.n, h_out);
}

} // extern "C"
