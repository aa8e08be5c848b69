// Host-memory entry points of libfrsz2_hip.so: a pipelined H2D -> kernel ->
// D2H path over runs of exponent blocks.
//
// The reference's CPU codec and LibPressio plugin (include/frsz2.hpp:1108-1249,
// src/frsz2.cc:152-186) take and return host buffers.  Here such a call is
// cut into chunks of whole exponent blocks (about 32 MiB of values each).
// Each chunk goes through one of three slots, each with its own stream,
// device buffers and pinned host buffers:
//
//   CPU: pageable input -> pinned slot buffer (a pool of copy threads)
//   GPU: H2D -> compress / decompress of the chunk as a run of blocks -> D2H
//   CPU: pinned slot buffer -> pageable output (when the slot is reused)
//
// so the CPU copies of one chunk, the transfers in both directions and the
// kernels of the others overlap.  A chunk's stream bytes are staged as a
// compact sub-stream (its exponent slots, then its payload blocks at a
// 256-byte aligned offset) and copied to their places in the caller's
// stream, so any run of blocks takes the tile kernels.  Slots stay allocated
// for reuse until the process ends; concurrent callers get separate slot
// sets.
#include "copy_pool.hpp"
#include "frsz2_capi.h"
#include "frsz2_device.hpp"
#include "frsz2_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace frsz2 {

namespace {

constexpr std::uint64_t CHUNK_BYTES = std::uint64_t{ 32 } << 20; // values per chunk, in bytes
constexpr int SLOTS = 3;

struct slot
{
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  void* d_vals = nullptr;
  void* d_sub = nullptr;
  void* h_vals = nullptr; // pinned
  void* h_sub = nullptr;  // pinned
  std::uint64_t vals_cap = 0, sub_cap = 0;
  // output of the chunk in flight: up to two pieces, copied from the pinned
  // buffers to the caller's when the chunk is retired
  struct piece
  {
    std::uint8_t* to = nullptr;
    const std::uint8_t* from = nullptr;
    std::uint64_t bytes = 0;
  } out[2];
  bool busy = false;
};

struct pipeline
{
  int device = -1;
  slot slots[SLOTS];
  copy_pool pool;

  explicit pipeline(int workers) : pool(workers) {}

  hipError_t init(int dev)
  {
    device = dev;
    for (slot& q : slots) {
      hipError_t e = hipStreamCreateWithFlags(&q.s, hipStreamNonBlocking);
      if (e == hipSuccess) {
        e = hipEventCreateWithFlags(&q.done, hipEventDisableTiming);
      }
      if (e != hipSuccess) {
        return e;
      }
    }
    return hipSuccess;
  }

  hipError_t reserve(slot& q, std::uint64_t vals, std::uint64_t sub)
  {
    if (vals > q.vals_cap) {
      if (q.d_vals) {
        (void)hipFree(q.d_vals);
        (void)hipHostFree(q.h_vals);
        q.d_vals = q.h_vals = nullptr;
        q.vals_cap = 0;
      }
      hipError_t e = hipMalloc(&q.d_vals, vals);
      if (e == hipSuccess) {
        e = hipHostMalloc(&q.h_vals, vals, hipHostMallocDefault);
      }
      if (e != hipSuccess) {
        return e;
      }
      q.vals_cap = vals;
    }
    if (sub > q.sub_cap) {
      if (q.d_sub) {
        (void)hipFree(q.d_sub);
        (void)hipHostFree(q.h_sub);
        q.d_sub = q.h_sub = nullptr;
        q.sub_cap = 0;
      }
      hipError_t e = hipMalloc(&q.d_sub, sub);
      if (e == hipSuccess) {
        e = hipHostMalloc(&q.h_sub, sub, hipHostMallocDefault);
      }
      if (e != hipSuccess) {
        return e;
      }
      q.sub_cap = sub;
    }
    return hipSuccess;
  }

  // wait for the slot's chunk and copy its output to the caller's buffer
  hipError_t retire(slot& q)
  {
    if (!q.busy) {
      return hipSuccess;
    }
    q.busy = false;
    const hipError_t e = hipEventSynchronize(q.done);
    if (e != hipSuccess) {
      return e;
    }
    for (const slot::piece& c : q.out) {
      if (c.bytes) {
        pool.copy(c.to, c.from, c.bytes);
      }
    }
    return hipSuccess;
  }

  // wait for every slot without copying (error path)
  void drain()
  {
    for (slot& q : slots) {
      if (q.busy) {
        (void)hipStreamSynchronize(q.s);
        q.busy = false;
      }
    }
  }
};

std::mutex g_pipes_mu;
std::vector<pipeline*> g_free_pipes;

int
copy_workers()
{
  const unsigned hc = std::thread::hardware_concurrency();
  return static_cast<int>(std::min(7u, hc > 1 ? hc - 1 : 0u));
}

// one pipeline for the current device, exclusively held until released
struct pipe_lease
{
  pipeline* p = nullptr;
  hipError_t err = hipSuccess;

  pipe_lease()
  {
    int dev = 0;
    err = hipGetDevice(&dev);
    if (err != hipSuccess) {
      return;
    }
    {
      std::lock_guard<std::mutex> g(g_pipes_mu);
      for (std::size_t i = 0; i < g_free_pipes.size(); ++i) {
        if (g_free_pipes[i]->device == dev) {
          p = g_free_pipes[i];
          g_free_pipes.erase(g_free_pipes.begin() + static_cast<std::ptrdiff_t>(i));
          return;
        }
      }
    }
    p = new pipeline(copy_workers());
    err = p->init(dev);
  }
  ~pipe_lease()
  {
    if (p && err != hipSuccess) {
      delete p; // its streams and events are not reused
    } else if (p) {
      p->drain();
      std::lock_guard<std::mutex> g(g_pipes_mu);
      g_free_pipes.push_back(p);
    }
  }
};

std::uint64_t
round_up(std::uint64_t x, std::uint64_t a)
{
  return (x + a - 1) / a * a;
}

// blocks per chunk: about CHUNK_BYTES of values, at least one block (a
// chunk is staged as a compact sub-stream, so any block count keeps the
// tile kernels' alignment)
std::uint64_t
chunk_blocks(std::uint64_t block, std::uint64_t vs)
{
  return std::max<std::uint64_t>(1, CHUNK_BYTES / (block * vs));
}

#define FRSZ2_PIPE_TRY(expr, what)                                                                            \
  do {                                                                                                       \
    const hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) {                                                                                  \
      lease.p->drain();                                                                                      \
      return hip_status(e_, what);                                                                           \
    }                                                                                                        \
  } while (0)

// Compress values [first*B, first*B + count) of an n-value array from host
// memory into the host stream h_stream, chunk by chunk.
int
compress_run_host(int dtype, int bits, std::uint64_t block, std::uint64_t n, std::uint64_t first,
                  const void* h_values, std::uint64_t count, void* h_stream)
{
  const std::uint64_t vs = dtype == FRSZ2_FLOAT32 ? 4u : 8u;
  const std::uint64_t C = chunk_blocks(block, vs);
  const std::uint64_t nblk = count / block + (count % block != 0);
  if (nblk == 0 && make_run(bits, block, n, first, 0).zpad == 0) {
    return FRSZ2_OK; // nothing to write
  }
  const std::uint64_t chunks = nblk == 0 ? 1 : (nblk + C - 1) / C;
  pipe_lease lease;
  if (lease.err != hipSuccess) {
    return hip_status(lease.err, "pipeline setup");
  }
  pipeline& P = *lease.p;
  auto* out = static_cast<std::uint8_t*>(h_stream);
  const auto* in = static_cast<const std::uint8_t*>(h_values);
  for (std::uint64_t c = 0; c < chunks; ++c) {
    slot& q = P.slots[c % SLOTS];
    FRSZ2_PIPE_TRY(P.retire(q), "hipEventSynchronize");
    const std::uint64_t b0 = first + c * C;
    const std::uint64_t v0 = c * C * block;
    const std::uint64_t cnt = std::min(count - v0, C * block);
    const run_plan r = make_run(bits, block, n, b0, cnt);
    const std::uint64_t es = static_cast<std::uint64_t>(r.L.U / 8);
    const std::uint64_t head = r.L.nb * es + r.zpad;
    const std::uint64_t P0 = round_up(head, 256);
    const std::uint64_t sub = P0 + r.L.nb * r.L.CB;
    FRSZ2_PIPE_TRY(P.reserve(q, std::max<std::uint64_t>(cnt * vs, 16), sub), "staging allocation");
    if (cnt) {
      P.pool.copy(q.h_vals, in + v0 * vs, cnt * vs);
      FRSZ2_PIPE_TRY(hipMemcpyAsync(q.d_vals, q.h_vals, cnt * vs, hipMemcpyHostToDevice, q.s), "hipMemcpyAsync");
    }
    layout Lc = r.L;
    Lc.R = P0;
    int rc = compress_run(dtype, q.d_vals, q.d_sub, Lc, r.zpad, q.s);
    if (rc) {
      P.drain();
      return rc;
    }
    FRSZ2_PIPE_TRY(hipMemcpyAsync(q.h_sub, q.d_sub, sub, hipMemcpyDeviceToHost, q.s), "hipMemcpyAsync");
    FRSZ2_PIPE_TRY(hipEventRecord(q.done, q.s), "hipEventRecord");
    q.busy = true;
    const auto* hs = static_cast<const std::uint8_t*>(q.h_sub);
    q.out[0] = { out + r.exp_off, hs, head };                  // exponent slots (+ pad)
    q.out[1] = { out + r.pay_off, hs + P0, r.L.nb * r.L.CB }; // payload blocks
  }
  for (std::uint64_t c = chunks; c < chunks + SLOTS; ++c) {
    FRSZ2_PIPE_TRY(P.retire(P.slots[c % SLOTS]), "hipEventSynchronize");
  }
  return FRSZ2_OK;
}

// Decompress values [first*B, first*B + count) of an n-value array from the
// host stream h_stream into host memory.
int
decompress_run_host(int dtype, int bits, std::uint64_t block, std::uint64_t n, std::uint64_t first,
                    const void* h_stream, std::uint64_t count, void* h_values)
{
  if (count == 0) {
    return FRSZ2_OK;
  }
  const std::uint64_t vs = dtype == FRSZ2_FLOAT32 ? 4u : 8u;
  const std::uint64_t C = chunk_blocks(block, vs);
  const std::uint64_t nblk = count / block + (count % block != 0);
  const std::uint64_t chunks = (nblk + C - 1) / C;
  pipe_lease lease;
  if (lease.err != hipSuccess) {
    return hip_status(lease.err, "pipeline setup");
  }
  pipeline& P = *lease.p;
  const auto* in = static_cast<const std::uint8_t*>(h_stream);
  auto* out = static_cast<std::uint8_t*>(h_values);
  for (std::uint64_t c = 0; c < chunks; ++c) {
    slot& q = P.slots[c % SLOTS];
    FRSZ2_PIPE_TRY(P.retire(q), "hipEventSynchronize");
    const std::uint64_t b0 = first + c * C;
    const std::uint64_t v0 = c * C * block;
    const std::uint64_t cnt = std::min(count - v0, C * block);
    const run_plan r = make_run(bits, block, n, b0, cnt);
    const std::uint64_t es = static_cast<std::uint64_t>(r.L.U / 8);
    const std::uint64_t P0 = round_up(r.L.nb * es, 256);
    const std::uint64_t sub = P0 + r.L.nb * r.L.CB;
    FRSZ2_PIPE_TRY(P.reserve(q, cnt * vs, sub), "staging allocation");
    auto* hs = static_cast<std::uint8_t*>(q.h_sub);
    P.pool.copy(hs, in + r.exp_off, r.L.nb * es);
    P.pool.copy(hs + P0, in + r.pay_off, r.L.nb * r.L.CB);
    FRSZ2_PIPE_TRY(hipMemcpyAsync(q.d_sub, q.h_sub, sub, hipMemcpyHostToDevice, q.s), "hipMemcpyAsync");
    layout Lc = r.L;
    Lc.R = P0;
    int rc = decompress_run(dtype, q.d_sub, q.d_vals, Lc, q.s);
    if (rc) {
      P.drain();
      return rc;
    }
    FRSZ2_PIPE_TRY(hipMemcpyAsync(q.h_vals, q.d_vals, cnt * vs, hipMemcpyDeviceToHost, q.s), "hipMemcpyAsync");
    FRSZ2_PIPE_TRY(hipEventRecord(q.done, q.s), "hipEventRecord");
    q.busy = true;
    q.out[0] = { out + v0 * vs, static_cast<const std::uint8_t*>(q.h_vals), cnt * vs };
    q.out[1] = {};
  }
  for (std::uint64_t c = chunks; c < chunks + SLOTS; ++c) {
    FRSZ2_PIPE_TRY(P.retire(P.slots[c % SLOTS]), "hipEventSynchronize");
  }
  return FRSZ2_OK;
}

#undef FRSZ2_PIPE_TRY

} // namespace

} // namespace frsz2

extern "C" {

int
frsz2_compress_host(int dtype, int bits, uint64_t block, const void* h_in, uint64_t n, void* h_out,
                    uint64_t out_capacity)
{
  using namespace frsz2;
  clear_error();
  const int rc = check_params(dtype, bits, block);
  if (rc) {
    return rc;
  }
  const std::uint64_t total = layout::make(bits, block, n).total_bytes();
  if (out_capacity < total) {
    return set_error(FRSZ2_ECAPACITY,
                     "output capacity " + std::to_string(out_capacity) + " < stream size " + std::to_string(total));
  }
  if (!h_out || (n > 0 && !h_in)) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  return compress_run_host(dtype, bits, block, n, 0, h_in, n, h_out);
}

int
frsz2_decompress_host(int dtype, int bits, uint64_t block, const void* h_in, uint64_t in_bytes, uint64_t n,
                      void* h_out)
{
  using namespace frsz2;
  clear_error();
  const int rc = check_params(dtype, bits, block);
  if (rc) {
    return rc;
  }
  const std::uint64_t total = layout::make(bits, block, n).total_bytes();
  if (in_bytes < total) {
    return set_error(FRSZ2_ECAPACITY,
                     "input stream of " + std::to_string(in_bytes) + " bytes < stream size " + std::to_string(total));
  }
  if (!h_in || (n > 0 && !h_out)) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  return decompress_run_host(dtype, bits, block, n, 0, h_in, n, h_out);
}

int
frsz2_compress_blocks_host(int dtype, int bits, uint64_t block, uint64_t n, uint64_t first_block,
                           const void* h_values, uint64_t count, void* h_stream, uint64_t stream_bytes)
{
  using namespace frsz2;
  clear_error();
  const int rc = check_run(dtype, bits, block, n, first_block, count);
  if (rc) {
    return rc;
  }
  const std::uint64_t total = layout::make(bits, block, n).total_bytes();
  if (stream_bytes < total) {
    return set_error(FRSZ2_ECAPACITY,
                     "output capacity " + std::to_string(stream_bytes) + " < stream size " + std::to_string(total));
  }
  if (!h_stream || (count > 0 && !h_values)) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  return compress_run_host(dtype, bits, block, n, first_block, h_values, count, h_stream);
}

int
frsz2_decompress_blocks_host(int dtype, int bits, uint64_t block, uint64_t n, uint64_t first_block,
                             const void* h_stream, uint64_t stream_bytes, uint64_t count, void* h_values)
{
  using namespace frsz2;
  clear_error();
  const int rc = check_run(dtype, bits, block, n, first_block, count);
  if (rc) {
    return rc;
  }
  const std::uint64_t total = layout::make(bits, block, n).total_bytes();
  if (stream_bytes < total) {
    return set_error(FRSZ2_ECAPACITY,
                     "input stream of " + std::to_string(stream_bytes) + " bytes < stream size " +
                       std::to_string(total));
  }
  if (!h_stream || (count > 0 && !h_values)) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  return decompress_run_host(dtype, bits, block, n, first_block, h_stream, count, h_values);
}

} // extern "C"
