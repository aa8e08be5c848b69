// Internal declarations shared by the translation units of libfrsz2_hip.so.
#pragma once

#include "frsz2_device.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace frsz2 {

// thread-local last error (frsz2_last_error)
int set_error(int code, const std::string& msg);
void clear_error();

int check_params(int dtype, int bits, std::uint64_t block);

// kernel variant of the tiled plain kernels: values per thread (4, 8, 16),
// whether global accesses carry the non-temporal hint, and (decompress)
// whether the output goes out through an LDS staging buffer
enum
{
  OP_COMPRESS = 0,
  OP_DECOMPRESS = 1
};
struct variant
{
  int vpt;
  int nt;
  int stage;
  int linear; // tiles in workgroup order instead of one contiguous range per XCD
  int small;  // 64-thread workgroups (one wave) instead of 256
};
variant get_variant(int op, int dtype, int bits);
int set_variant(int op, int dtype, int bits, int vpt, int nt, int stage, int linear, int small);
int hip_status(hipError_t e, const char* what);

// A run of exponent blocks [first, first + nb) of the stream of an n-value
// array as a plan of its own: L.n = the run's values, L.nb = its blocks,
// L.R = the distance from its first exponent slot (stream + exp_off) to its
// first payload block (stream + pay_off), zpad = the exponent-pad bytes it
// clears (the whole pad when the run ends at the stream's last block).
struct run_plan
{
  layout L;
  std::uint64_t exp_off;
  std::uint64_t pay_off;
  std::uint32_t zpad;
};
run_plan make_run(int bits, std::uint64_t block, std::uint64_t n, std::uint64_t first, std::uint64_t count);
int check_run(int dtype, int bits, std::uint64_t block, std::uint64_t n, std::uint64_t first, std::uint64_t count);
// launch on a run: exponents at d_out, payload at d_out + L.R
int compress_run(int dtype, const void* d_values, void* d_out, const layout& L, std::uint32_t zpad, hipStream_t s);
int decompress_run(int dtype, const void* d_in, void* d_values, const layout& L, hipStream_t s);

int compress_blocks_device(int dtype, int bits, std::uint64_t block, std::uint64_t n, std::uint64_t first,
                           const void* d_values, std::uint64_t count, void* d_stream, std::uint64_t stream_bytes,
                           hipStream_t s);
int decompress_blocks_device(int dtype, int bits, std::uint64_t block, std::uint64_t n, std::uint64_t first,
                             const void* d_stream, std::uint64_t stream_bytes, std::uint64_t count, void* d_values,
                             hipStream_t s);
int compress_device(int dtype, int bits, std::uint64_t block, const void* d_in, std::uint64_t n, void* d_out,
                    std::uint64_t out_capacity, hipStream_t s);
int decompress_device(int dtype, int bits, std::uint64_t block, const void* d_in, std::uint64_t in_bytes,
                      std::uint64_t n, void* d_out, hipStream_t s);
int gather_device(int dtype, int bits, std::uint64_t block, const void* d_in, std::uint64_t in_bytes,
                  std::uint64_t n, const std::uint64_t* d_idx, std::uint64_t m, void* d_out, hipStream_t s);

} // namespace frsz2
