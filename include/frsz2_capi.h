/*
 * C ABI of the MI355X-native FRSZ2 codec (libfrsz2_hip.so).
 *
 * Plain pointers and sizes only.  Device pointers are HIP device memory on
 * the current device; `stream` is a hipStream_t passed as void* (NULL = the
 * default stream).  Every call is stream-ordered and asynchronous unless its
 * comment says otherwise.  The state shared between calls is the tuning
 * table (frsz2_tune_*; every variant it can select writes the same bytes)
 * and a pool of host staging pipelines that each host-memory call leases for
 * its duration; the last-error string is per thread.  Calls on different
 * streams or threads may therefore run concurrently (the reference plugin
 * declares pressio:thread_safe = multiple, src/frsz2.cc:40).
 *
 * dtype: FRSZ2_FLOAT32 or FRSZ2_FLOAT64.
 * bits:  4..32 for float32, 4..64 for float64 (src/frsz2.cc:86-101 lists the
 *        same ranges; include/frsz2.hpp:511-517 asserts 4 <= bits <= W).
 * block: the exponent block size B ("max_exp_block_size"), any value in
 *        1..FRSZ2_MAX_BLOCK.  Powers of two up to 4096 take the tiled
 *        kernels (compress: with a 16-byte aligned payload; decompress:
 *        with a 16-byte aligned payload when bits < U; values need only
 *        their element alignment).  Other cases take slower generic
 *        kernels (compress: one thread per block up to B = 16, else a
 *        block-maximum pass and one thread per payload word; decompress and
 *        gather: one thread per value).
 *
 * Status codes follow the reference plugin: 0 ok, 1 unsupported dtype
 * (src/frsz2.cc:164, :181), 2 unsupported bits / block size with the
 * reference's message text "Unsupported number of bits: N" /
 * "Unsupported exponent block size: N" (include/frsz2.hpp:1300, :1326).
 * Codes 3..6 are new: capacity too small, bad pointer or alignment, HIP
 * runtime error, malformed container header.  frsz2_last_error() returns
 * the message of the calling thread's last failure.
 */
#ifndef FRSZ2_CAPI_H
#define FRSZ2_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* frsz2_stream_t;

enum
{
  FRSZ2_FLOAT32 = 0,
  FRSZ2_FLOAT64 = 1
};

enum
{
  FRSZ2_OK = 0,
  FRSZ2_EDTYPE = 1,
  FRSZ2_EPARAM = 2,
  FRSZ2_ECAPACITY = 3,
  FRSZ2_EARG = 4,
  FRSZ2_EHIP = 5,
  FRSZ2_EFORMAT = 6
};

#define FRSZ2_MAX_BLOCK (1u << 24)

/* Message of the calling thread's last failed call ("" if none). */
const char* frsz2_last_error(void);

/* Library version string. */
const char* frsz2_version(void);

/* 0 if (dtype, bits, block) is supported, else 1 or 2 with the message set.
 * Replaces the supported-list checks of the runtime dispatch
 * (include/frsz2.hpp:1296-1347, src/frsz2.cc:83-134). */
int frsz2_check(int dtype, int bits, uint64_t block);

/* Exact stream size in bytes for n values: exponent region rounded up to the
 * next multiple of 256 (always 1..256 bytes of pad) plus ceil(n/B) blocks.
 * Replaces frsz2_compressor::compute_compressed_memory_size_byte
 * (include/frsz2.hpp:542-548).  Host-only, synchronous. */
int frsz2_compressed_size(int dtype, int bits, uint64_t block, uint64_t n, uint64_t* bytes);

/* Compress n values at d_in into the stream at d_out (capacity
 * out_capacity >= frsz2_compressed_size).  Writes every byte of the stream:
 * exponents, a zeroed pad, all payload words (the words of a partial last
 * block that hold no code are written as zero).  d_in must be aligned to the
 * value size and d_out to 8 bytes.
 * Replaces frsz2_compressor::compress_cpu_impl (include/frsz2.hpp:1181-1193)
 * and the compress_gpu kernel (include/frsz2.hpp:1258-1274). */
int frsz2_compress(int dtype,
                   int bits,
                   uint64_t block,
                   const void* d_in,
                   uint64_t n,
                   void* d_out,
                   uint64_t out_capacity,
                   frsz2_stream_t stream);

/* Decompress the stream at d_in (in_bytes >= frsz2_compressed_size) into n
 * values at d_out.
 * Replaces frsz2_compressor::decompress_cpu_impl (include/frsz2.hpp:1242-1249)
 * and the decompress_gpu kernel (include/frsz2.hpp:1276-1288). */
int frsz2_decompress(int dtype,
                     int bits,
                     uint64_t block,
                     const void* d_in,
                     uint64_t in_bytes,
                     uint64_t n,
                     void* d_out,
                     frsz2_stream_t stream);

/* Random-access decode: d_out[i] = value d_idx[i] of the stream, for i < m.
 * An index >= n yields 0 (the reference does not check; this reads nothing
 * out of bounds instead).
 * Replaces frsz2_compressor::decompress_cpu_element (include/frsz2.hpp:1199-1214)
 * and decompress_gpu_element (include/frsz2.hpp:1096-1104). */
int frsz2_gather(int dtype,
                 int bits,
                 uint64_t block,
                 const void* d_in,
                 uint64_t in_bytes,
                 uint64_t n,
                 const uint64_t* d_idx,
                 uint64_t m,
                 void* d_out,
                 frsz2_stream_t stream);

/* Compress part of an array into its stream: the `count` values at d_values
 * are values [first_block*B, first_block*B + count) of an n-value array;
 * they fill exponent blocks [first_block, first_block + ceil(count/B)) of the
 * array's stream at d_stream (stream_bytes >= frsz2_compressed_size(n)),
 * which is written at exactly those exponent slots and payload blocks, plus
 * the zeroed exponent pad when the run ends at the stream's last block.
 * The run's last block may be short of B values (a partial block inside the
 * array, as exp_block_elements < B gives): its exponent is taken over the
 * values given and its remaining codes are zero.  Runs over disjoint blocks
 * may run concurrently and together give the bytes of frsz2_compress.
 * Replaces frsz2_compressor::compress_cpu_block(exp_block_id, data,
 * exp_block_elements) (include/frsz2.hpp:1108-1179), generalized to a run of
 * blocks. */
int frsz2_compress_blocks(int dtype,
                          int bits,
                          uint64_t block,
                          uint64_t n,
                          uint64_t first_block,
                          const void* d_values,
                          uint64_t count,
                          void* d_stream,
                          uint64_t stream_bytes,
                          frsz2_stream_t stream);

/* Decompress values [first_block*B, first_block*B + count) of the n-value
 * array whose stream is at d_stream into d_values.
 * Replaces frsz2_compressor::decompress_cpu_block(exp_block_id, output)
 * (include/frsz2.hpp:1217-1240), generalized to a run of blocks. */
int frsz2_decompress_blocks(int dtype,
                            int bits,
                            uint64_t block,
                            uint64_t n,
                            uint64_t first_block,
                            const void* d_stream,
                            uint64_t stream_bytes,
                            uint64_t count,
                            void* d_values,
                            frsz2_stream_t stream);

/* Kernel variant of the tiled plain kernels for (op, dtype, bits): op 0 =
 * compress, 1 = decompress; values_per_thread 2 (float64 with bits > 8
 * only; otherwise taken as 4), 4, 8 or 16: a tile is 256 x that many values
 * and is raised to the next size that holds a whole exponent block;
 * flags: FRSZ2_TUNE_NONTEMPORAL puts the non-temporal hint on the streaming
 * accesses, FRSZ2_TUNE_STAGED_STORES (decompress) selects the decode of
 * consecutive values per thread whose tile output goes through LDS with
 * whole-cache-line stores (without it: the strided decode, one value per
 * lane per step, stored straight from registers), FRSZ2_TUNE_LINEAR_ORDER takes
 * tiles in workgroup order (default: one contiguous range of tiles per
 * XCD), FRSZ2_TUNE_SMALL_WORKGROUPS launches 64-thread workgroups (one wave,
 * tiles of 64 x values_per_thread; implies the non-temporal hint) instead of
 * 256-thread ones.  Process-wide.  No counterpart
 * in the reference (its kernels have a fixed launch shape,
 * test/test_frsz_gpu.cu:325-330). */
enum
{
  FRSZ2_TUNE_NONTEMPORAL = 1,
  FRSZ2_TUNE_STAGED_STORES = 2,
  FRSZ2_TUNE_LINEAR_ORDER = 4,
  FRSZ2_TUNE_SMALL_WORKGROUPS = 8
};
int frsz2_tune_set(int op, int dtype, int bits, int values_per_thread, int flags);
int frsz2_tune_get(int op, int dtype, int bits, int* values_per_thread, int* flags);
/* Autotune the variant of (op, dtype, bits) on the caller's data: every
 * variant (values per thread, non-temporal hint, staged stores, tile order,
 * workgroup size) runs `reps` times (5 when reps <= 0) after one untimed
 * launch, timed with events on `stream`, and the fastest is kept in the
 * table (returned in *values_per_thread / *flags when not NULL).  op 0:
 * d_in holds n values (in_bytes) and d_out receives the stream (out_bytes);
 * op 1: d_in is a stream of n values (in_bytes) and d_out receives the
 * values (out_bytes).  d_out is overwritten many times; every variant
 * writes the same bytes.  Synchronous; calls on other threads meanwhile run
 * with whatever variant is being timed (all are bit-exact). */
int frsz2_tune_auto(int op,
                    int dtype,
                    int bits,
                    uint64_t block,
                    uint64_t n,
                    const void* d_in,
                    uint64_t in_bytes,
                    void* d_out,
                    uint64_t out_bytes,
                    int reps,
                    int* values_per_thread,
                    int* flags,
                    frsz2_stream_t stream);

/* ---- batched: many arrays in one graph-captured launch ----
 * The graph holds one grouped kernel per class of jobs present (direction,
 * dtype, storage width, packed or not), each with the plain kernels' variant
 * for that class; the tile schedule is picked by frsz2_batch_autotune on the
 * plan's own jobs.  No counterpart in the reference: its kernels take one
 * array per launch (include/frsz2.hpp:1258-1288).  Each job has the meaning
 * of one frsz2_compress / frsz2_decompress / frsz2_gather call (same checks,
 * same bytes). */

enum
{
  FRSZ2_OP_COMPRESS = 0,
  FRSZ2_OP_DECOMPRESS = 1,
  FRSZ2_OP_GATHER = 2
};

typedef struct frsz2_job
{
  int op;               /* FRSZ2_OP_COMPRESS, _DECOMPRESS or _GATHER */
  int dtype;            /* FRSZ2_FLOAT32 / FRSZ2_FLOAT64 */
  int bits;
  int reserved;
  uint64_t block;       /* exponent block size B */
  uint64_t n;           /* values of the array / stream */
  const void* in;       /* compress: values; decompress, gather: stream (device memory) */
  void* out;            /* compress: stream; decompress: n values; gather: m values */
  uint64_t in_bytes;    /* decompress, gather: bytes available at in */
  uint64_t out_bytes;   /* bytes available at out: compress >= stream size,
                           decompress >= n values, gather >= m values */
  const uint64_t* idx;  /* gather: m indices (device memory) */
  uint64_t m;           /* gather: index count */
} frsz2_job;

/* Tile schedules of the batched launch (frsz2_batch_config.schedule) */
enum
{
  FRSZ2_SCHEDULE_TILES = 0,     /* one workgroup per tile (the default) */
  FRSZ2_SCHEDULE_PERSISTENT = 1 /* per class, a grid of what the device holds at
                                   once striding over the class's tiles */
};

/* Kernel classes of the tiled compress and decompress jobs of a batch:
 * class = op * 14 + slot, op = FRSZ2_OP_COMPRESS (0) or FRSZ2_OP_DECOMPRESS
 * (1); slot = 2 * u + packed for float32 (u = 0, 1, 2 for the storage width
 * U = 8, 16, 32) and 6 + 2 * u + packed for float64 (u = 0..3 for U = 8..64),
 * where U is the smallest of 8, 16, 32, 64 bits that holds `bits` and
 * packed = (bits < U).  Each class runs one grouped kernel. */
#define FRSZ2_BATCH_CLASSES 28

typedef struct frsz2_batch_config
{
  int schedule;        /* FRSZ2_SCHEDULE_TILES or FRSZ2_SCHEDULE_PERSISTENT
                          (other values = FRSZ2_SCHEDULE_TILES) */
  int workgroups;      /* persistent schedule: workgroups per CU (0, or more
                          than fit = as many as the device holds) */
  int xcd_order;       /* 1: each of the 8 XCDs takes one contiguous range of
                          a class's tiles; 0: workgroup order */
  int use_graph;       /* 1 = launch through a captured hipGraph */
  int plain_launches;  /* 1 = no grouped kernels: one plain frsz2_compress /
                          frsz2_decompress / frsz2_gather launch per job, in
                          job order (the per-array baseline the batch is
                          measured against) */
  /* Tile shape of each kernel class under FRSZ2_SCHEDULE_TILES (the
     persistent schedule runs the built-in shapes): values per thread (4, 8
     or 16; 2 as well for float64 with U >= 16) and threads per workgroup
     (64 or 256); 0 = the class's built-in value (the plain kernels'
     measured variant).  A job whose block size B exceeds its class's tile
     (threads x values per thread) runs as a plain launch.  The shape changes
     only how the work is split, never the bytes written. */
  unsigned char values_per_thread[FRSZ2_BATCH_CLASSES];
  unsigned short workgroup_size[FRSZ2_BATCH_CLASSES];
} frsz2_batch_config;

typedef struct frsz2_batch_plan frsz2_batch_plan;

/* Validate the jobs, upload their descriptors and tile table, and capture
 * the launch.  cfg NULL = defaults (one workgroup per tile, workgroup order,
 * graph on).  Jobs must not overlap in memory. */
int frsz2_batch_create(const frsz2_job* jobs, size_t njobs, const frsz2_batch_config* cfg,
                       frsz2_batch_plan** plan);
/* Run every job of the plan, stream-ordered.  Launches of one plan that
 * overlap in time write the same outputs concurrently, so order them. */
int frsz2_batch_launch(frsz2_batch_plan* plan, frsz2_stream_t stream);
/* Autotune the plan on its own jobs (each candidate launched reps times on
 * stream, outputs overwritten with identical bytes): first the tile
 * schedule (one workgroup per tile and persistent, each in workgroup and XCD
 * order, built-in tile shapes); then, when one workgroup per tile wins, each
 * class's tile shape (values per thread x workgroup size, every shape whose
 * tile holds the class's blocks), timed on that class's jobs alone.  Keeps
 * the fastest and re-captures; the choice is written to *chosen if not
 * NULL. */
int frsz2_batch_autotune(frsz2_batch_plan* plan, frsz2_stream_t stream, int reps, frsz2_batch_config* chosen);
int frsz2_batch_get_config(const frsz2_batch_plan* plan, frsz2_batch_config* cfg);
int frsz2_batch_set_config(frsz2_batch_plan* plan, const frsz2_batch_config* cfg);
int frsz2_batch_destroy(frsz2_batch_plan* plan);

/* Run the jobs once, as frsz2_batch_create(jobs, njobs, NULL) +
 * frsz2_batch_launch + frsz2_batch_destroy would (grouped launches, same
 * bytes), without allocating any memory: the descriptors of the jobs the
 * tiled kernels take go to a 1 MiB file-scope device array of the library
 * (at most FRSZ2_NOALLOC_MAX_JOBS such jobs; more fail with
 * FRSZ2_ECAPACITY before anything runs), written by stream-ordered kernels
 * on `stream`.  No counterpart in the reference: this is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Every call on a device shares
 * that device's array, so calls on one device whose launches overlap in time
 * (other streams, other threads) race; calls ordered on one stream do not.
 * The array stays allocated on each device the library has run this call
 * on until the process ends. */
#define FRSZ2_NOALLOC_MAX_JOBS 10922u
int frsz2_batch_run_noalloc(const frsz2_job* jobs, size_t njobs, frsz2_stream_t stream);

/* ---- serving: resident streams, HBM-budget admission, continuous batching ----
 * No counterpart in the reference (its API is one synchronous call per
 * array).  A server owns a device pool of hbm_budget bytes from which every
 * request's output is allocated; a scheduler thread admits requests in
 * arrival order while their outputs fit, and runs each group of admitted
 * requests as one batch launch (the grouped kernels of frsz2_batch_launch) on one of
 * `slots` streams, so requests arriving while others run are admitted into the next
 * launch.  A request computes exactly what frsz2_compress / frsz2_decompress
 * / frsz2_gather compute. */

typedef struct frsz2_server frsz2_server;

typedef struct frsz2_server_config
{
  uint64_t hbm_budget; /* bytes of the device pool (0 = 64 GiB) */
  int slots;           /* concurrent launches (0 = 4) */
  int max_batch;       /* requests per launch (0 = 1024) */
  int schedule;        /* tile schedule of the batch launches (see
                          frsz2_batch_config; 0 = one workgroup per tile) */
  int workgroups;      /* persistent schedule: workgroups per CU (0 = as many
                          as the device holds) */
  int xcd_order;       /* 1 = XCD-contiguous tile order */
  int stall_ms;        /* once nothing has run for this long and no queued
                          request can be admitted (the pool is held by
                          unreleased outputs), every queued request that has
                          itself waited this long fails with FRSZ2_ECAPACITY;
                          younger ones fail as they reach it, until something
                          is admitted (0 = 10000; negative = never) */
  uint64_t large_bytes; /* a request that moves at least this many HBM bytes
                           (values + stream; a gather: its elements) is large:
                           it is batched only with other large requests, and
                           launches of large requests hold at most slots - 1
                           slots at once, so small requests always find one
                           (0 = 32 MiB; UINT64_MAX = no request is large) */
} frsz2_server_config;

enum
{
  FRSZ2_REQ_DISCARD = 1 /* free the output as soon as it is no longer needed */
};

typedef struct frsz2_request
{
  int op;              /* FRSZ2_OP_COMPRESS, _DECOMPRESS or _GATHER */
  int dtype;
  int bits;
  int flags;           /* FRSZ2_REQ_DISCARD */
  uint64_t block;
  uint64_t n;          /* values of the array / of the stream read */
  const void* values;  /* compress: n input values (device memory) */
  uint64_t source;     /* decompress, gather: ticket of the compress request whose
                          stream is read (replay: index of that request in the trace) */
  const uint64_t* idx; /* gather: m indices (device memory) */
  uint64_t m;
  double arrival;      /* replay: release time in seconds after the replay starts */
} frsz2_request;

typedef struct frsz2_result
{
  int status;
  int reserved;
  double t_arrival, t_admit, t_launch, t_done; /* seconds (server epoch; replay: replay start) */
  void* out;           /* the output (stream or values) in the pool, while held */
  uint64_t out_bytes;
} frsz2_result;

typedef struct frsz2_server_stats
{
  uint64_t launches;
  uint64_t launched_requests;
  uint64_t completed;
  uint64_t pool_used;
  uint64_t pool_peak;
  uint64_t max_queue;
} frsz2_server_stats;

int frsz2_server_create(const frsz2_server_config* cfg, frsz2_server** server);
/* Queue a request (thread-safe).  Its output stays in the pool until
 * frsz2_server_release, unless FRSZ2_REQ_DISCARD; a compress stream also
 * stays while queued requests read it. */
int frsz2_server_submit(frsz2_server* server, const frsz2_request* req, uint64_t* ticket);
/* Block until the request completes; fills *res. */
int frsz2_server_wait(frsz2_server* server, uint64_t ticket, frsz2_result* res);
/* Drop the client's hold on a request's output. */
int frsz2_server_release(frsz2_server* server, uint64_t ticket);
/* Open-loop replay of a trace: request i is queued reqs[i].arrival seconds
 * after the call; readers name their compress request by trace index; each
 * stream stays resident until the last request of the trace reading it has
 * completed; blocks until every request has completed.  tickets (optional)
 * receives the ticket of every request whose output is still held (0 for
 * FRSZ2_REQ_DISCARD). */
int frsz2_server_replay(frsz2_server* server, const frsz2_request* reqs, size_t count, frsz2_result* results,
                        uint64_t* tickets);
int frsz2_server_get_stats(frsz2_server* server, frsz2_server_stats* stats);
int frsz2_server_destroy(frsz2_server* server);

/* ---- self-describing container ----
 * The reference stream carries no header: n, dtype, bits and B travel out of
 * band (src/frsz2.cc:152-186).  A container is a 64-byte header (magic
 * "FRSZ2CT", version 1, dtype, bits, B, n, stream bytes; little-endian)
 * followed by the unchanged stream, so container[64:] is exactly what
 * frsz2_compress writes.  No counterpart in the reference. */
#define FRSZ2_CONTAINER_HEADER_BYTES 64u

typedef struct frsz2_container_info
{
  int dtype;
  int bits;
  uint64_t block;
  uint64_t n;
  uint64_t stream_bytes;
  uint64_t header_bytes; /* offset of the stream in the container (64) */
} frsz2_container_info;

/* Header + stream bytes for n values.  Host-only. */
int frsz2_container_size(int dtype, int bits, uint64_t block, uint64_t n, uint64_t* bytes);
/* Parse and check a header in host memory (at least its first 64 bytes;
 * `bytes` = how many are there): magic, version, supported (dtype, bits, B)
 * and a stream size that matches them; 6 if malformed.  Host-only. */
int frsz2_container_parse(const void* h_container, uint64_t bytes, frsz2_container_info* info);
/* Write header + stream of n values at d_in into d_out (8-byte aligned;
 * 64-byte aligned keeps the stream on the tiled kernels). */
int frsz2_compress_container(int dtype,
                             int bits,
                             uint64_t block,
                             const void* d_in,
                             uint64_t n,
                             void* d_out,
                             uint64_t out_capacity,
                             frsz2_stream_t stream);
/* Decompress the container at d_in, described by `info` (frsz2_container_parse
 * of a host copy of its header), into d_out (room for out_values >= n). */
int frsz2_decompress_container(const frsz2_container_info* info,
                               const void* d_in,
                               uint64_t in_bytes,
                               void* d_out,
                               uint64_t out_values,
                               frsz2_stream_t stream);

/* Host-memory variants; synchronous.  They keep the host-in / host-out
 * meaning of the reference's CPU entry points and of the LibPressio plugin
 * (src/frsz2.cc:152-186), which dereference host buffers.  Compress and
 * decompress run as a pipeline over chunks of about 32 MiB of values (whole
 * exponent blocks): CPU copies into pinned staging, H2D, kernel and D2H of
 * different chunks overlap on three streams; staging stays allocated for
 * reuse until the process ends.  Gather copies the stream and the indices
 * to the device, runs the gather kernel and copies the values back. */
int frsz2_compress_host(int dtype, int bits, uint64_t block, const void* h_in, uint64_t n, void* h_out,
                        uint64_t out_capacity);
int frsz2_decompress_host(int dtype, int bits, uint64_t block, const void* h_in, uint64_t in_bytes, uint64_t n,
                          void* h_out);
int frsz2_gather_host(int dtype, int bits, uint64_t block, const void* h_in, uint64_t in_bytes, uint64_t n,
                      const uint64_t* h_idx, uint64_t m, void* h_out);
/* Host-memory variants of frsz2_compress_blocks / frsz2_decompress_blocks:
 * h_stream is the whole host stream of the n-value array; only the run's
 * exponent slots, payload blocks and (last run) exponent pad are copied. */
/* Container in host memory (e.g. a file image): compress writes header +
 * stream; decompress parses its own header.  Synchronous. */
int frsz2_compress_container_host(int dtype, int bits, uint64_t block, const void* h_in, uint64_t n, void* h_out,
                                  uint64_t out_capacity);
int frsz2_decompress_container_host(const void* h_in, uint64_t in_bytes, void* h_out, uint64_t out_values);
int frsz2_compress_blocks_host(int dtype, int bits, uint64_t block, uint64_t n, uint64_t first_block,
                               const void* h_values, uint64_t count, void* h_stream, uint64_t stream_bytes);
int frsz2_decompress_blocks_host(int dtype, int bits, uint64_t block, uint64_t n, uint64_t first_block,
                                 const void* h_stream, uint64_t stream_bytes, uint64_t count, void* h_values);

#ifdef __cplusplus
}
#endif

#endif
