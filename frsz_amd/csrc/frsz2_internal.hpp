// Internal declarations shared by the translation units of libfrsz2_hip.so.
#pragma once

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
};
variant get_variant(int op, int dtype, int bits);
int set_variant(int op, int dtype, int bits, int vpt, int nt, int stage, int linear);
int hip_status(hipError_t e, const char* what);

int compress_device(int dtype, int bits, std::uint64_t block, const void* d_in, std::uint64_t n, void* d_out,
                    std::uint64_t out_capacity, hipStream_t s);
int decompress_device(int dtype, int bits, std::uint64_t block, const void* d_in, std::uint64_t in_bytes,
                      std::uint64_t n, void* d_out, hipStream_t s);
int gather_device(int dtype, int bits, std::uint64_t block, const void* d_in, std::uint64_t in_bytes,
                  std::uint64_t n, const std::uint64_t* d_idx, std::uint64_t m, void* d_out, hipStream_t s);

} // namespace frsz2
