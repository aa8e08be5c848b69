// Error state, the kernel-variant table and the host-memory gather of
// libfrsz2_hip.so (the pipelined host compress / decompress live in
// frsz2_pipeline.hip).
//
// frsz2_gather_host keeps the host-in / host-out meaning of the reference's
// decompress_cpu_element (include/frsz2.hpp:1199-1214): it copies the stream
// and the indices to device memory, runs the gather kernel on a private
// stream, and copies the values back before returning.
#include "frsz2_capi.h"
#include "frsz2_device.hpp"
#include "frsz2_internal.hpp"
#include "frsz2_tuning.hpp"

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <cstdint>
#include <string>

namespace frsz2 {

namespace {
thread_local std::string g_error;

// tuning table of the tiled kernels: [op][dtype][U = 8,16,32,64][bits < U],
// each entry vpt * 8 + linear * 4 + stage * 2 + nt.  Defaults come from measurements on MI355X;
// frsz2_tune_set (the autotuner) overrides them per process.
std::atomic<int> g_variant[2][2][4][2];
std::once_flag g_variant_once;

int
u_index(int bits)
{
  const int U = storage_bits(bits);
  return U == 8 ? 0 : U == 16 ? 1 : U == 32 ? 2 : 3;
}

// Defaults measured on MI355X (include/frsz2_tuning.hpp): [op][dtype][U index][packed].
void
fill_variants()
{
  for (int op = 0; op < 2; ++op)
    for (int d = 0; d < 2; ++d)
      for (int u = 0; u < 4; ++u)
        for (int pk = 0; pk < 2; ++pk)
          g_variant[op][d][u][pk].store(default_variant(op, d, u, pk), std::memory_order_relaxed);
}

void
init_variants()
{
  std::call_once(g_variant_once, fill_variants);
}
} // namespace

variant
get_variant(int op, int dtype, int bits)
{
  init_variants();
  const int v = g_variant[op][dtype][u_index(bits)][bits < storage_bits(bits) ? 1 : 0].load(std::memory_order_relaxed);
  return variant{ venc_vpt(v), venc_nt(v), venc_stage(v), venc_linear(v), venc_small(v) };
}

int
set_variant(int op, int dtype, int bits, int vpt, int nt, int stage, int linear, int small)
{
  init_variants();
  g_variant[op][dtype][u_index(bits)][bits < storage_bits(bits) ? 1 : 0].store(
    venc(vpt, stage ? 1 : 0, nt ? 1 : 0, linear ? 1 : 0, small ? 1 : 0), std::memory_order_relaxed);
  return 0;
}

int
set_error(int code, const std::string& msg)
{
  g_error = msg;
  return code;
}

void
clear_error()
{
  g_error.clear();
}

namespace {

// device buffers + a non-blocking stream released on scope exit
struct staging
{
  hipStream_t s = nullptr;
  void* bufs[3] = { nullptr, nullptr, nullptr };
  ~staging()
  {
    for (void* b : bufs) {
      if (b) {
        (void)hipFree(b);
      }
    }
    if (s) {
      (void)hipStreamDestroy(s);
    }
  }
  hipError_t init() { return hipStreamCreateWithFlags(&s, hipStreamNonBlocking); }
  hipError_t alloc(int i, std::uint64_t bytes) { return hipMalloc(&bufs[i], bytes ? bytes : 16); }
};

#define FRSZ2_HIP_TRY(expr, what)                                                                            \
  do {                                                                                                       \
    const hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess)                                                                                    \
      return hip_status(e_, what);                                                                           \
  } while (0)

} // namespace

} // namespace frsz2

extern "C" {

const char*
frsz2_last_error(void)
{
  return frsz2::g_error.c_str();
}

int
frsz2_tune_set(int op, int dtype, int bits, int values_per_thread, int flags)
{
  using namespace frsz2;
  clear_error();
  const int rc = check_params(dtype, bits, 1);
  if (rc) {
    return rc;
  }
  if ((op != OP_COMPRESS && op != OP_DECOMPRESS) ||
      (values_per_thread != 2 && values_per_thread != 4 && values_per_thread != 8 && values_per_thread != 16)) {
    return set_error(FRSZ2_EARG, "op must be 0/1 and values_per_thread 2, 4, 8 or 16");
  }
  const int small = (flags & FRSZ2_TUNE_SMALL_WORKGROUPS) != 0;
  // 64-thread workgroups are built with the non-temporal hint only
  return set_variant(op, dtype, bits, values_per_thread, small || (flags & FRSZ2_TUNE_NONTEMPORAL),
                     op == OP_DECOMPRESS ? (flags & FRSZ2_TUNE_STAGED_STORES) : 0, flags & FRSZ2_TUNE_LINEAR_ORDER,
                     small);
}

int
frsz2_tune_auto(int op, int dtype, int bits, uint64_t block, uint64_t n, const void* d_in, uint64_t in_bytes,
                void* d_out, uint64_t out_bytes, int reps, int* values_per_thread, int* flags, frsz2_stream_t stream)
{
  using namespace frsz2;
  clear_error();
  int rc = check_params(dtype, bits, block);
  if (rc) {
    return rc;
  }
  if (op != OP_COMPRESS && op != OP_DECOMPRESS) {
    return set_error(FRSZ2_EARG, "op must be 0 (compress) or 1 (decompress)");
  }
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const std::uint64_t value_bytes = n * (dtype == FRSZ2_FLOAT32 ? 4u : 8u);
  if ((op == OP_COMPRESS ? in_bytes : out_bytes) < value_bytes) {
    return set_error(FRSZ2_ECAPACITY, "value buffer holds fewer than n values");
  }
  auto launch = [&]() {
    return op == OP_COMPRESS ? compress_device(dtype, bits, block, d_in, n, d_out, out_bytes, s)
                             : decompress_device(dtype, bits, block, d_in, in_bytes, n, d_out, s);
  };
  // the argument checks run once, before any variant is timed
  rc = launch();
  if (rc) {
    return rc;
  }
  reps = reps > 0 ? reps : 5;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  FRSZ2_HIP_TRY(hipEventCreate(&e0), "hipEventCreate");
  const hipError_t ee = hipEventCreate(&e1);
  if (ee != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_status(ee, "hipEventCreate");
  }
  const variant before = get_variant(op, dtype, bits);
  variant best = before;
  float best_ms = -1.0f;
  const int U = storage_bits(bits);
  for (int vpt : { 2, 4, 8, 16 }) {
    if (vpt == 2 && !(dtype == FRSZ2_FLOAT64 && U >= 16)) {
      continue;
    }
    for (int stage = 0; stage <= (op == OP_DECOMPRESS ? 1 : 0); ++stage) {
      for (int linear = 0; linear <= 1; ++linear) {
        for (int small = 0; small <= 1; ++small) {
          for (int nt = small; nt <= 1; ++nt) {
            set_variant(op, dtype, bits, vpt, nt, stage, linear, small);
            float ms = 0.0f;
            hipError_t e = hipSuccess;
            rc = launch(); // first launch of a variant loads its code object
            if (!rc) {
              e = hipEventRecord(e0, s);
            }
            for (int r = 0; !rc && e == hipSuccess && r < reps; ++r) {
              rc = launch();
            }
            if (!rc && e == hipSuccess) {
              e = hipEventRecord(e1, s);
            }
            if (!rc && e == hipSuccess) {
              e = hipEventSynchronize(e1);
            }
            if (!rc && e == hipSuccess) {
              e = hipEventElapsedTime(&ms, e0, e1);
            }
            if (rc || e != hipSuccess) {
              set_variant(op, dtype, bits, before.vpt, before.nt, before.stage, before.linear, before.small);
              (void)hipEventDestroy(e0);
              (void)hipEventDestroy(e1);
              return rc ? rc : hip_status(e, "variant timing");
            }
            if (best_ms < 0.0f || ms < best_ms) {
              best_ms = ms;
              best = variant{ vpt, nt, stage, linear, small };
            }
          }
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  set_variant(op, dtype, bits, best.vpt, best.nt, best.stage, best.linear, best.small);
  if (values_per_thread) {
    *values_per_thread = best.vpt;
  }
  if (flags) {
    *flags = (best.nt ? FRSZ2_TUNE_NONTEMPORAL : 0) | (best.stage ? FRSZ2_TUNE_STAGED_STORES : 0) |
             (best.linear ? FRSZ2_TUNE_LINEAR_ORDER : 0) | (best.small ? FRSZ2_TUNE_SMALL_WORKGROUPS : 0);
  }
  return FRSZ2_OK;
}

int
frsz2_tune_get(int op, int dtype, int bits, int* values_per_thread, int* flags)
{
  using namespace frsz2;
  clear_error();
  const int rc = check_params(dtype, bits, 1);
  if (rc) {
    return rc;
  }
  if ((op != OP_COMPRESS && op != OP_DECOMPRESS) || !values_per_thread || !flags) {
    return set_error(FRSZ2_EARG, "bad op or null pointer");
  }
  const variant v = get_variant(op, dtype, bits);
  *values_per_thread = v.vpt;
  *flags = (v.nt ? FRSZ2_TUNE_NONTEMPORAL : 0) | (v.stage ? FRSZ2_TUNE_STAGED_STORES : 0) |
           (v.linear ? FRSZ2_TUNE_LINEAR_ORDER : 0) | (v.small ? FRSZ2_TUNE_SMALL_WORKGROUPS : 0);
  return FRSZ2_OK;
}

int
frsz2_gather_host(int dtype, int bits, uint64_t block, const void* h_in, uint64_t in_bytes, uint64_t n,
                  const uint64_t* h_idx, uint64_t m, void* h_out)
{
  using namespace frsz2;
  clear_error();
  int rc = check_params(dtype, bits, block);
  if (rc) {
    return rc;
  }
  const layout L = layout::make(bits, block, n);
  const std::uint64_t total = L.total_bytes();
  if (in_bytes < total) {
    return set_error(FRSZ2_ECAPACITY,
                     "input stream of " + std::to_string(in_bytes) + " bytes < stream size " + std::to_string(total));
  }
  if (m == 0) {
    return FRSZ2_OK;
  }
  if (!h_in || !h_idx || !h_out) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  const std::uint64_t vs = dtype == FRSZ2_FLOAT32 ? 4u : 8u;
  staging st;
  FRSZ2_HIP_TRY(st.init(), "hipStreamCreate");
  FRSZ2_HIP_TRY(st.alloc(0, total), "hipMalloc");
  FRSZ2_HIP_TRY(st.alloc(1, m * 8), "hipMalloc");
  FRSZ2_HIP_TRY(st.alloc(2, m * vs), "hipMalloc");
  FRSZ2_HIP_TRY(hipMemcpyAsync(st.bufs[0], h_in, total, hipMemcpyHostToDevice, st.s), "hipMemcpyAsync");
  FRSZ2_HIP_TRY(hipMemcpyAsync(st.bufs[1], h_idx, m * 8, hipMemcpyHostToDevice, st.s), "hipMemcpyAsync");
  rc = gather_device(dtype, bits, block, st.bufs[0], total, n, static_cast<const std::uint64_t*>(st.bufs[1]), m,
                     st.bufs[2], st.s);
  if (rc) {
    return rc;
  }
  FRSZ2_HIP_TRY(hipMemcpyAsync(h_out, st.bufs[2], m * vs, hipMemcpyDeviceToHost, st.s), "hipMemcpyAsync");
  FRSZ2_HIP_TRY(hipStreamSynchronize(st.s), "hipStreamSynchronize");
  return FRSZ2_OK;
}

} // extern "C"
