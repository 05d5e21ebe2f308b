/*
 * hptt_amd.h -- C ABI of the MI355X (gfx950) HPTT transposition library.
 *
 *   B[perm(i)] = alpha * A[i] + beta * B[perm(i)]
 *
 * Column-major tensors (dim 0 is stride-1), ranks 2..6, element types float,
 * double, complex<float> and complex<double> (interleaved re/im).  A and B are
 * DEVICE pointers; exec enqueues on a HIP stream and returns without waiting.
 *
 * Every entry point returns an hpttx status code; hpttx_last_error() gives the
 * reason for the last failure on the calling thread.  Plans are immutable
 * after creation except for hpttx_plan_reset_data / hpttx_plan_set_stream;
 * exec of one plan from several host threads is safe (no per-exec host state).
 *
 * Each declaration names the reference interface it replaces
 * (tongsucn/hptt, file:line).
 */
#ifndef HPTT_AMD_H_
#define HPTT_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPTTX_VERSION 100

/* Status codes.  The reference reports failure only as a null plan
 * (inc/hptt/hptt.h:71-83); these refine that null. */
#define HPTTX_OK 0
#define HPTTX_EINVAL 1       /* invalid argument (src/hptt.cc:25-56 cases) */
#define HPTTX_ENOMEM 2       /* host or device allocation failed */
#define HPTTX_EHIP 3         /* a HIP runtime call failed */
#define HPTTX_EUNSUPPORTED 4 /* rank > 6 or extents beyond the kernel limits */

/* dtype codes used by the batch and serving interfaces */
#define HPTTX_F32 0
#define HPTTX_F64 1
#define HPTTX_C64 2
#define HPTTX_C128 3

typedef struct hpttx_plan hpttx_plan;

/*
 * Plan creation.  Replaces the C++-mangled per-dtype factories
 * create_trans_plan_impl_{s,d,c,z} (src/impl/hptt_trans_impl_s.cc:8-20, and
 * the d/c/z twins) that hptt::create_plan<T> dlsyms (src/hptt.cc:66-92), with
 * the validation of src/hptt.cc:25-56 folded in.
 *
 *   in, out          device pointers to A and B (non-owning; see reset_data)
 *   rank             tensor order, 2..6 (the reference takes it from perm.size())
 *   in_size[rank]    A's extents
 *   perm[rank]       B dim d holds A dim perm[d]; out_size[d] = in_size[perm[d]]
 *   alpha, beta      scaling; B is read iff fabs(beta) >= 1e-16
 *                    (inc/hptt/impl/hptt_trans_impl.h:75-78)
 *   num_threads      accepted and ignored (the reference's OpenMP thread count)
 *   tuning_timeout   0: model pick; > 0: autotune at most this many seconds;
 *                    < 0: exhaustive autotune (inc/hptt/hptt.h:46-57).
 *                    Tuning runs on scratch buffers, never on `out`.
 *   in_outer_size    NULL or [rank]: A's padded extents (strides)
 *   out_outer_size   NULL or [rank]: B's padded extents
 *   stream           hipStream_t the plan enqueues on (NULL = null stream)
 *   plan             receives the plan; *plan = NULL on failure
 */
int hpttx_plan_create_s(const float *in, float *out, uint32_t rank,
                        const uint32_t *in_size, const uint32_t *perm,
                        float alpha, float beta, uint32_t num_threads,
                        double tuning_timeout, const uint32_t *in_outer_size,
                        const uint32_t *out_outer_size, void *stream,
                        hpttx_plan **plan);
int hpttx_plan_create_d(const double *in, double *out, uint32_t rank,
                        const uint32_t *in_size, const uint32_t *perm,
                        double alpha, double beta, uint32_t num_threads,
                        double tuning_timeout, const uint32_t *in_outer_size,
                        const uint32_t *out_outer_size, void *stream,
                        hpttx_plan **plan);
/* complex<float>; alpha/beta real, applied to re and im alike
 * (inc/hptt/types.h:29-54) -- replaces create_trans_plan_impl_c */
int hpttx_plan_create_c(const void *in, void *out, uint32_t rank,
                        const uint32_t *in_size, const uint32_t *perm,
                        float alpha, float beta, uint32_t num_threads,
                        double tuning_timeout, const uint32_t *in_outer_size,
                        const uint32_t *out_outer_size, void *stream,
                        hpttx_plan **plan);
/* complex<double>; replaces create_trans_plan_impl_z */
int hpttx_plan_create_z(const void *in, void *out, uint32_t rank,
                        const uint32_t *in_size, const uint32_t *perm,
                        double alpha, double beta, uint32_t num_threads,
                        double tuning_timeout, const uint32_t *in_outer_size,
                        const uint32_t *out_outer_size, void *stream,
                        hpttx_plan **plan);
/* Complex coefficients (an extension: the reference's coefficients are real,
 * inc/hptt/types.h:29-54, inc/hptt/hptt.h:89-90).  Order:
 *   a*x = (fl(fl(ar*xr) - fl(ai*xi)), fl(fl(ar*xi) + fl(ai*xr))), b*y alike,
 *   r = (fl(by.re + ax.re), fl(by.im + ax.im)); B read iff |br| or |bi| >= 1e-16.
 * With ai == bi == 0 these dispatch to the real-coefficient plans above. */
int hpttx_plan_create_c_cplx(const void *in, void *out, uint32_t rank,
                             const uint32_t *in_size, const uint32_t *perm,
                             float alpha_re, float alpha_im, float beta_re,
                             float beta_im, uint32_t num_threads,
                             double tuning_timeout,
                             const uint32_t *in_outer_size,
                             const uint32_t *out_outer_size, void *stream,
                             hpttx_plan **plan);
int hpttx_plan_create_z_cplx(const void *in, void *out, uint32_t rank,
                             const uint32_t *in_size, const uint32_t *perm,
                             double alpha_re, double alpha_im, double beta_re,
                             double beta_im, uint32_t num_threads,
                             double tuning_timeout,
                             const uint32_t *in_outer_size,
                             const uint32_t *out_outer_size, void *stream,
                             hpttx_plan **plan);

/* Enqueue the transposition on the plan's stream.  Replaces
 * CGraphTransPackBase::exec / operator() (inc/hptt/impl/hptt_trans.h:27-28),
 * which run synchronously; here B is final once the stream reaches this point
 * (hpttx_plan_sync). */
int hpttx_plan_exec(hpttx_plan *plan);
/* Same, on another stream (does not change the plan's stream). */
int hpttx_plan_exec_on(hpttx_plan *plan, void *stream);
/* Wait for the plan's stream. */
int hpttx_plan_sync(hpttx_plan *plan);
/* Swap the data pointers; sizes, perm, alpha and beta stay fixed.  Replaces
 * CGraphTransPackBase::reset_data (inc/hptt/impl/hptt_trans.h:30). */
int hpttx_plan_reset_data(hpttx_plan *plan, const void *in, void *out);
int hpttx_plan_set_stream(hpttx_plan *plan, void *stream);
/* Human-readable plan (fused dims, path, tile, grid) into buf (NUL-terminated,
 * truncated to len).  Returns the full length needed.  Replaces
 * CGraphTransPackBase::print_plan (inc/hptt/impl/hptt_trans.h:29), which
 * prints loop order and thread split. */
int hpttx_plan_describe(const hpttx_plan *plan, char *buf, size_t len);

typedef struct hpttx_plan_info {
  uint64_t elements;        /* N */
  uint32_t elem_bytes;      /* 4, 8, 8, 16 */
  int32_t update;           /* 1 iff B is read */
  int32_t path;             /* 0 tile, 1 tile-shared, 2 wide-copy, 3 axpby, 4 gather */
  int32_t fused_rank;
  uint64_t algorithmic_bytes; /* N*elem*(2 + update) */
  uint32_t grid;            /* blocks per launch */
  uint32_t lds_bytes;       /* LDS per block */
  uint32_t tile_r, tile_c, tile_o, vec_r, vec_w;
  uint32_t unaligned_r, unaligned_w; /* 1: that side's runs use vectors at any
                                        element offset (partial last vectors) */
  uint32_t threads;         /* threads per block (256, or 512 for tiles over 32 KiB of LDS) */
} hpttx_plan_info;
int hpttx_plan_get_info(const hpttx_plan *plan, hpttx_plan_info *info);

/* Release the plan (the reference's shared_ptr deleter, src/hptt.cc:89-92). */
void hpttx_plan_destroy(hpttx_plan *plan);

/*
 * Batched transpositions: many independent plans whose block tiles form one
 * queue, walked by work-queue kernels (one launch per element type and
 * kernel variant, or per element type when HPTTX_BATCH_MODE=merged; one block
 * per tile, or a persistent grid with HPTTX_BATCH_GRID=persistent), captured
 * once in a hipGraph and replayed by hpttx_batch_exec.  No counterpart in the
 * reference, whose only way to run several transpositions is one
 * plan->exec() after another
 * (src/hptt.cc:16-93, inc/hptt/cgraph/cgraph_trans.tcc:27-48); each item has
 * the meaning of one hpttx_plan_create_* call.  Tile shapes are chosen per
 * (dtype, fused rank, path, size class); with tuning_timeout != 0 each class
 * is autotuned on scratch buffers.
 */
typedef struct hpttx_batch hpttx_batch;

typedef struct hpttx_batch_item {
  int32_t dtype;               /* HPTTX_F32 / F64 / C64 / C128 */
  uint32_t rank;               /* 2..6 */
  const void *in;              /* device A */
  void *out;                   /* device B */
  uint32_t in_size[6];
  uint32_t perm[6];
  uint32_t in_outer[6];        /* all zero = unpadded */
  uint32_t out_outer[6];       /* all zero = unpadded */
  double alpha_re, alpha_im;   /* real dtypes use the _re parts */
  double beta_re, beta_im;
  int32_t complex_coef;        /* nonzero: complex alpha/beta (C64/C128) */
  int32_t reserved_;
} hpttx_batch_item;

int hpttx_batch_create(uint32_t count, const hpttx_batch_item *items,
                       double tuning_timeout, void *stream,
                       hpttx_batch **batch);
/* The same with a launch layout.  flags = 0: as hpttx_batch_create.
 * HPTTX_BATCH_ONE_LAUNCH: one persistent launch per element type (the kernel
 * dispatches each item's variant at run time), whose resident blocks take
 * chunks of tiles, largest items first, off a device counter (a dynamic work
 * queue, cleared by a kernel node before each launch); captured in the
 * hipGraph like the default layout. */
#define HPTTX_BATCH_ONE_LAUNCH 1u
int hpttx_batch_create_ex(uint32_t count, const hpttx_batch_item *items,
                          double tuning_timeout, void *stream, uint32_t flags,
                          hpttx_batch **batch);
/* Enqueue every item of the batch (graph replay) on the batch's stream. */
int hpttx_batch_exec(hpttx_batch *batch);
int hpttx_batch_exec_on(hpttx_batch *batch, void *stream);
int hpttx_batch_sync(hpttx_batch *batch);
/* Re-point item `index` (creation order) at new buffers with the same
 * 16-byte alignment; takes effect at the next exec. */
int hpttx_batch_set_data(hpttx_batch *batch, uint32_t index, const void *in,
                         void *out);
/* 1 = replay a captured hipGraph (default), 0 = launch directly. */
int hpttx_batch_use_graph(hpttx_batch *batch, int enable);

typedef struct hpttx_batch_info {
  uint32_t count;              /* items */
  uint32_t groups;             /* (element type, variant) groups = launches per exec */
  uint64_t algorithmic_bytes;  /* sum over items of N*elem*(2 + update) */
  uint64_t tiles;              /* block tiles over all groups */
  uint32_t grid[4];            /* blocks of the first four groups' launches */
  uint32_t lds_bytes[4];
} hpttx_batch_info;
int hpttx_batch_get_info(const hpttx_batch *batch, hpttx_batch_info *info);
/* Where item `index` (creation order) runs: its launch group, the group's
 * kernel (a variant code, or 128 for the instance that dispatches variants
 * at run time), the item's own variant code (bits: 1 rows, 2 exact read
 * vectors, 4 exact write vectors (rows: unaligned vectors), 8 update,
 * 16 shared innermost group, 32 complex coefficients, 64 streaming), the LDS
 * its tiles need and its tile count. */
typedef struct hpttx_batch_item_info {
  uint32_t group;
  uint32_t kernel;
  uint32_t variant;
  uint32_t lds_bytes;
  uint32_t tiles;
  uint32_t reserved_;
} hpttx_batch_item_info;
int hpttx_batch_get_item_info(const hpttx_batch *batch, uint32_t index,
                              hpttx_batch_item_info *info);
/*
 * A batch run that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It plans
 * `items` (as hpttx_batch_create with tuning_timeout 0), writes their tables
 * into a static device array of HPTTX_NOALLOC_SCRATCH_BYTES (one per GPU,
 * held until the process ends) and enqueues the one-launch layout on
 * `stream` (one persistent launch per element type over a dynamic queue).
 * Returns once the table upload has been read (it waits for `stream` to
 * reach it); the transpositions themselves run asynchronously.  A batch
 * whose tables do not fit the array is refused with HPTTX_EINVAL (about
 * 9000 items fit).  Every call on one GPU shares that GPU's array: calls
 * whose work overlaps on one GPU (another stream, or another thread before
 * the previous call's launches finished) race on it.
 */
#define HPTTX_NOALLOC_SCRATCH_BYTES (8u << 20)
int hpttx_batch_run_noalloc(uint32_t count, const hpttx_batch_item *items,
                            void *stream);
/* Path of item `index` ("tile", "tile-shared", "wide-copy", "axpby",
 * "gather") into buf. */
int hpttx_batch_item_path(const hpttx_batch *batch, uint32_t index, char *buf,
                          size_t len);
void hpttx_batch_destroy(hpttx_batch *batch);

/*
 * Serving: a continuous-batching transposition server.  No counterpart in
 * the reference (one synchronous plan->exec() per transposition,
 * inc/hptt/impl/hptt_trans.h:27).  Requests (hpttx_batch_item, one
 * transposition each) are queued by hpttx_server_submit and admitted in
 * arrival order while the algorithmic bytes in flight stay within the budget
 * (default: a quarter of the HBM free at that admission round, so memory
 * allocated outside the server shrinks it) and the output arena has room;
 * the admitted requests go into the next launches of the batch kernels on
 * one of `slots` streams, so requests arriving while others run join the
 * next launch and launches overlap on the device.  Plans are
 * cached by shape.  A request whose `out` is NULL gets B allocated in the
 * server's arena (beta must then be below 1e-16); its address is reported at
 * completion and the buffer is held until hpttx_server_release.
 */
typedef struct hpttx_server hpttx_server;

typedef struct hpttx_server_config {
  uint64_t inflight_bytes;   /* 0: a quarter of the HBM free at each admission round */
  uint64_t arena_bytes;      /* bytes for server-allocated outputs (0: none) */
  uint32_t slots;            /* launches in flight (streams), 1..8; 0: 3 */
  uint32_t max_batch;        /* requests per launch; 0: 256 */
} hpttx_server_config;

typedef struct hpttx_completion {
  uint64_t ticket;
  int32_t state;             /* 2 done, 3 failed */
  int32_t status;            /* HPTTX_* when failed; hpttx_server_poll / _wait then
                                set hpttx_last_error() to the reason */
  void *out;                 /* B (the arena slot for server-allocated outputs) */
  double t_submit, t_launch, t_done;   /* seconds, one steady clock */
  uint64_t bytes;            /* algorithmic bytes */
} hpttx_completion;

typedef struct hpttx_server_stats {
  uint64_t submitted, completed, failed, launches;
  uint64_t bytes_done;
  double first_submit, last_done;      /* seconds, same clock as completions */
  double p50_latency, p99_latency, mean_latency;   /* submit -> done, seconds */
  uint64_t inflight_budget;  /* budget of the last admission round */
} hpttx_server_stats;

int hpttx_server_create(const hpttx_server_config *config, hpttx_server **server);
int hpttx_server_submit(hpttx_server *server, const hpttx_batch_item *request,
                        uint64_t *ticket);
/* Submit `count` requests at once, in order: they are planned and queued in
 * chunks of doubling size (128, 256, ...), the dispatcher woken after each,
 * so the first launches run while the rest are planned.  tickets[i]
 * receives request i's ticket (0 if not queued) and
 * *queued how many were queued.  An argument error queues none; a request
 * that fails to plan stops the submission there (the requests before it are
 * queued).  No counterpart in the reference; for callers that hand over many
 * transpositions at a time (one call instead of one per request). */
int hpttx_server_submit_many(hpttx_server *server, uint32_t count,
                             const hpttx_batch_item *requests, uint64_t *tickets,
                             uint32_t *queued);
/* Up to max finished requests not reported before; *n receives the count. */
int hpttx_server_poll(hpttx_server *server, hpttx_completion *out,
                      uint32_t max, uint32_t *n);
/* Block until `ticket` has finished (timeout_s < 0: no limit). */
int hpttx_server_wait(hpttx_server *server, uint64_t ticket, double timeout_s,
                      hpttx_completion *out);
/* Block until every submitted request has finished. */
int hpttx_server_drain(hpttx_server *server);
/* Forget a finished request and free its arena output, if any. */
int hpttx_server_release(hpttx_server *server, uint64_t ticket);
int hpttx_server_get_stats(hpttx_server *server, hpttx_server_stats *stats);
/* Seconds on the clock the completion times use. */
double hpttx_server_now(hpttx_server *server);
void hpttx_server_destroy(hpttx_server *server);

/* Reason for the last failure on this thread ("" if none). */
const char *hpttx_last_error(void);
int hpttx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* HPTT_AMD_H_ */
