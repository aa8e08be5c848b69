// Internal interface of the persistent batch kernel, shared by the batch
// plans (frsz2_batch.hip) and the serving runtime (frsz2_serve.hip).
#pragma once

#include "frsz2_capi.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace frsz2 {

// device descriptor of one job of the persistent kernel (96 bytes)
struct bjob
{
  const void* in;           // values (compress) or stream (decompress, gather)
  void* out;                // stream (compress) or values
  const std::uint64_t* idx; // gather: indices
  std::uint64_t n;          // values of the array / stream
  std::uint64_t nb;
  std::uint64_t R;
  std::uint64_t tile0;      // first global tile of the job
  std::uint64_t m;          // gather: indices
  std::uint64_t B;
  std::uint32_t CB;
  int bits;
  int lb;                   // log2 B, -1 if B is not a power of two (gather only)
  int cls;
  std::uint64_t pad_;
};
static_assert(sizeof(bjob) == 96, "descriptor layout");

struct batch_shape
{
  int tile_shape;      // 0 narrow, 1 balanced, 2 wide, 3 grouped (one launch per class)
  int workgroups;      // persistent grid
  int tiles_per_claim; // 0 = static round-robin
  int ops;             // directions present (1 compress, 2 decompress, 4 gather); 0 = any
};

// checks one job like the single-array entry points do
int validate_job(const frsz2_job& j, int idx);

// Appends the tile-path descriptors of jobs[0..n) to desc (grouped by kernel
// class, job order within a class) with global tile numbers starting at 0,
// and sets tiled[i] = 1 for the jobs they cover; returns the total tile
// count.  The other jobs must run through run_plain.
std::uint64_t build_descriptors(const frsz2_job* jobs, std::size_t n, int tile_shape, std::vector<bjob>& desc,
                                std::vector<char>& tiled);

// Descriptors a persistent launch can carry in its kernel arguments.
constexpr int INLINE_JOBS = 8;

// True when launch_persistent passes the descriptors in the kernel arguments
// (at most INLINE_JOBS descriptors, and a grid of one workgroup per tile:
// tile shape 3, or a persistent grid no larger than the tile count): d_desc
// is then not read and need not hold them.
bool descriptors_inline(const batch_shape& shape, int ndesc, std::uint64_t total_tiles);

// Counter reset (when claiming) + persistent kernel over ndesc descriptors in
// device memory (d_desc; h_desc is the host copy, which is what the kernel
// reads when descriptors_inline holds), on stream s.  Tile shape 3 launches
// one grouped kernel per class instead.  Each tile's job is read from d_jtab
// (tile_table of the descriptors for this tile shape, in device memory) or,
// when that is null, found by a binary search.
hipError_t launch_persistent(const bjob* d_desc, const bjob* h_desc, int ndesc, std::uint64_t total_tiles,
                             unsigned long long* d_counter, const batch_shape& shape, hipStream_t s,
                             const std::uint32_t* d_jtab = nullptr);

// Entry t is the index of the descriptor that owns global tile t: within its
// class's run of descriptors (tile shape 3, whose kernels see one class's
// descriptors) or among all of them (the persistent shapes).
std::vector<std::uint32_t> tile_table(const std::vector<bjob>& desc, std::uint64_t total_tiles, bool within_class);

// One job as a plain single-array launch on s.
int run_plain(const frsz2_job& j, hipStream_t s);

// compute units of the current device (256 if the query fails)
int device_cus();

} // namespace frsz2
