// Internal interface of the grouped batch kernels, shared by the batch plans
// (frsz2_batch.hip) and the serving runtime (frsz2_serve.hip).
#pragma once

#include "frsz2_capi.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace frsz2 {

// device descriptor of one job of the grouped kernels (96 bytes)
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

// tile schedule of the grouped launches (frsz2_batch_config)
struct batch_schedule
{
  int persistent; // 0: one workgroup per tile; 1: per class, a persistent grid striding over its tiles
  int workgroups; // persistent: workgroups per CU (0 = as many as the device holds)
  int xcd_order;  // 1: one contiguous range of a class's tiles per XCD; 0: workgroup order
};

// Tile shape of each compress / decompress kernel class (values per thread
// x threads per workgroup; frsz2_batch_config.values_per_thread and
// .workgroup_size, 0 = the class's built-in value), for the
// one-workgroup-per-tile schedule.  A null class_shapes pointer, wherever
// one is taken, means the built-in shapes; the descriptors and the launch
// of one batch must be given the same shapes.
struct class_shapes
{
  std::uint8_t vpt[FRSZ2_BATCH_CLASSES];
  std::uint16_t wgs[FRSZ2_BATCH_CLASSES];
};

// checks one job like the single-array entry points do
int validate_job(const frsz2_job& j, int idx);

// Checks a shape table (values per thread 0, 4, 8, 16, or 2 for float64
// classes with U >= 16; workgroup size 0, 64 or 256).
int validate_shapes(const class_shapes& sh);

// Kernel class of a job (compress and decompress: 0..27, gather: 28..34).
int job_class(const frsz2_job& j);

// Values per thread and workgroup size of class cls under shapes sh.
void class_shape(int cls, const class_shapes* sh, int& vpt, int& wgs);

// Appends the tile-path descriptors of jobs[0..n) to desc (grouped by kernel
// class, job order within a class) with global tile numbers starting at 0,
// and sets tiled[i] = 1 for the jobs they cover; returns the total tile
// count.  The other jobs must run through run_plain.
std::uint64_t build_descriptors(const frsz2_job* jobs, std::size_t n, std::vector<bjob>& desc,
                                std::vector<char>& tiled, const class_shapes* sh = nullptr);

// Descriptors a launch can carry in its kernel arguments.
constexpr int INLINE_JOBS = 8;

// True when launch_batch passes the descriptors in the kernel arguments (at
// most INLINE_JOBS descriptors, one workgroup per tile, built-in shapes):
// d_desc is then not read and need not hold them.
bool descriptors_inline(const batch_schedule& sc, int ndesc, std::uint64_t total_tiles,
                        const class_shapes* sh = nullptr);

// One grouped kernel per class present over the ndesc descriptors (in device
// memory at d_desc; h_desc is the host copy, which is what the kernels read
// when descriptors_inline holds), on stream s, with the tile schedule sc.
// Each tile's job is read from d_jtab (tile_table of the descriptors, in
// device memory) or, when that is null, found by a binary search.  sh: the
// tile shapes the descriptors were built with (honoured by the
// one-workgroup-per-tile schedule only; pass null with the persistent one).
hipError_t launch_batch(const bjob* d_desc, const bjob* h_desc, int ndesc, std::uint64_t total_tiles,
                        const batch_schedule& sc, hipStream_t s, const std::uint32_t* d_jtab = nullptr,
                        const class_shapes* sh = nullptr);

// Entry t is the index of the descriptor that owns global tile t: within its
// class's run of descriptors (within_class: the grouped kernels see one
// class's descriptors) or among all of them.
std::vector<std::uint32_t> tile_table(const std::vector<bjob>& desc, std::uint64_t total_tiles, bool within_class);

// One job as a plain single-array launch on s.
int run_plain(const frsz2_job& j, hipStream_t s);

// compute units of the current device (256 if the query fails)
int device_cus();

} // namespace frsz2
