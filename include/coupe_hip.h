/*
 * coupe_hip.h -- device-resident entry points of the MI355X RCB library.
 *
 * coupe.h keeps coupe's host-memory C ABI.  This header adds the calls a
 * GPU application needs: a context that owns the device work buffers and
 * entry points whose inputs and outputs already live in device memory (HBM),
 * so no PCIe copy sits on the path.  Results are identical to coupe_rcb.
 */
#ifndef COUPE_HIP_H
#define COUPE_HIP_H

#include <stdint.h>
#include "coupe.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * A per-device context: work buffers (grown on demand, reused across calls),
 * kept until coupe_hip_ctx_destroy.  Calls on one context must not overlap.
 * Returns NULL when the device cannot be used.
 * No counterpart in coupe (rayon's global pool plays this role there).
 */
typedef struct coupe_hip_ctx coupe_hip_ctx;
coupe_hip_ctx *coupe_hip_ctx_create(int device);
void coupe_hip_ctx_destroy(coupe_hip_ctx *ctx);

/*
 * RCB on device-resident data.
 *
 *   d_points    N*dimension doubles (AoS, like PointND<D>), device memory
 *   d_weights   N values of `weight_type` (int, int64_t or double), device
 *   d_partition N uint64 part ids, device memory (written)
 *   stream      hipStream_t to run on (NULL: the null stream)
 *
 * Same semantics, errors and ids as coupe_rcb.  Returns when the ids are
 * written (the stream is synchronised).  A null array with n > 0 returns
 * COUPE_ERR_CRASH before any device work (so do the batched entries below
 * when their instances hold points).
 * Replaces: Rcb::partition, coupe/src/algorithms/recursive_bisection.rs:805-811
 * (through rcb(), :644-705), for data already in device memory.
 */
enum coupe_err coupe_hip_rcb_device(coupe_hip_ctx *ctx, void *stream,
		uint64_t *d_partition, uintptr_t n, uintptr_t dimension,
		const double *d_points, enum coupe_type weight_type, const void *d_weights,
		uintptr_t iter_count, double tolerance);

/*
 * coupe_hip_rcb_device for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in coupe.
 *
 * Same arguments, semantics, ids and errors as coupe_hip_rcb_device on the
 * current device, without a context: every work buffer is a carve of one
 * static device array of COUPE_HIP_NOALLOC_SCRATCH_BYTES per GPU, which the
 * GPU holds from the library's load until the process ends.  The call
 * allocates no device or pinned memory (its host bookkeeping uses a few
 * small heap vectors).  Limits: n <= COUPE_HIP_NOALLOC_MAX_POINTS and
 * iter_count <= COUPE_HIP_NOALLOC_MAX_ITER_COUNT, else COUPE_ERR_ALLOC with
 * nothing written.  Calls on one GPU share its array: they are serialised
 * by a lock, and work that overlaps a call on that GPU from outside the
 * lock (another process, device code racing the call) races on the array.
 */
#define COUPE_HIP_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define COUPE_HIP_NOALLOC_MAX_POINTS (1ull << 19)
#define COUPE_HIP_NOALLOC_MAX_ITER_COUNT 12
enum coupe_err coupe_hip_rcb_device_noalloc(void *stream, uint64_t *d_partition, uintptr_t n,
		uintptr_t dimension, const double *d_points, enum coupe_type weight_type, const void *d_weights,
		uintptr_t iter_count, double tolerance);

/*
 * Batched RCB: many independent instances in one call.
 *
 * Instance q holds instances[q].n points; the instances are packed one after
 * the other in d_points (AoS, `dimension` doubles per point), d_weights and
 * d_partition, in the order given.  Instance q is partitioned exactly as
 * coupe_hip_rcb_device would partition it alone, with its own iter_count and
 * tolerance: its ids start at 0 and are written at its own offsets.
 * All instances share the dimension and the weight type (a batch of both
 * dimensions: coupe_hip_rcb_batch_mixed).  Every instance of a batch of more than one
 * needs iter_count <= 24 (else COUPE_ERR_BAD_TYPE).
 * No counterpart in coupe: it replaces a loop of Rcb::partition calls
 * (coupe/src/algorithms/recursive_bisection.rs:805-811), one per instance.
 */
struct coupe_hip_instance {
	uintptr_t n;            /* points of this instance */
	uintptr_t iter_count;   /* Rcb::iter_count */
	double tolerance;       /* Rcb::tolerance */
};
enum coupe_err coupe_hip_rcb_batch(coupe_hip_ctx *ctx, void *stream, uint64_t *d_partition,
		uintptr_t dimension, const double *d_points, enum coupe_type weight_type,
		const void *d_weights, const struct coupe_hip_instance *instances, uintptr_t n_instances);

/*
 * Batched RCB of instances of dimension 2 and 3 in one call: instance q has
 * instances[q].n points of instances[q].dimension doubles each, packed one
 * instance after the other in d_points (so the 2D and 3D instances may
 * alternate), its weights and ids at the same point positions of d_weights
 * and d_partition.  Each instance is partitioned exactly as
 * coupe_hip_rcb_device would partition it alone.  A call of more than one
 * instance needs every iter_count <= 24 (else COUPE_ERR_BAD_TYPE).
 * No counterpart in coupe (a loop of Rcb::partition calls over instances of
 * both dimensions).
 */
struct coupe_hip_instance_nd {
	uintptr_t n;            /* points of this instance */
	uintptr_t dimension;    /* 2 or 3 */
	uintptr_t iter_count;   /* Rcb::iter_count */
	double tolerance;       /* Rcb::tolerance */
};
enum coupe_err coupe_hip_rcb_batch_mixed(coupe_hip_ctx *ctx, void *stream, uint64_t *d_partition,
		const double *d_points, enum coupe_type weight_type, const void *d_weights,
		const struct coupe_hip_instance_nd *instances, uintptr_t n_instances);

/*
 * The engine's switches a call runs with.  None changes the ids, only the
 * work done: bins per part of the later split-search passes (16..256, a
 * power of two: log2(bins) bisection steps per pass), workgroups per CU of
 * the persistent kernels (0: as many as fit; fewer gives each workgroup
 * more tiles), split-search passes queued per level before the host looks
 * for a shape not run before (2..8), and the last pair of levels' first pass with every candidate where
 * most searches outlast their first candidate (0/1).
 */
struct coupe_hip_tuning {
	uint32_t bins;
	uint32_t workgroups_per_cu;
	uint32_t search_passes;
	uint32_t full_first;
	uint32_t graphs;        /* 1: a call of a shape seen before replays a captured HIP graph */
	uint32_t search_chunk;  /* consecutive tiles a workgroup takes at a time in the later search passes */
};

/*
 * Autotuner: runs the batch of coupe_hip_rcb_batch_mixed (same arguments)
 * under candidate switches, one switch at a time from the current ones
 * (graphs, bins, workgroups per CU, the full first pass, the search chunk), each candidate timed with HIP events over
 * `trials` calls after a
 * warm-up, and keeps the fastest for the batch's class: its dimension, the
 * log2 of its point count, its deepest iter_count and the log4 of its
 * instance count.  Later calls of that class on the context run with those
 * switches.  The ids are written as by coupe_hip_rcb_batch_mixed; `chosen`
 * (may be NULL) receives the switches kept and `best_ms` the fastest mean
 * time.  No counterpart in coupe.
 */
enum coupe_err coupe_hip_autotune(coupe_hip_ctx *ctx, void *stream, uint64_t *d_partition, const double *d_points,
		enum coupe_type weight_type, const void *d_weights, const struct coupe_hip_instance_nd *instances,
		uintptr_t n_instances, uint32_t trials, struct coupe_hip_tuning *chosen, double *best_ms);

/*
 * coupe::Rib on device-resident data: same arguments and errors as
 * coupe_hip_rcb_device.  The points are rotated so that the first axis is
 * their inertia axis, then partitioned by RCB; the rotation is applied on the
 * fly (the points are not modified and no rotated copy is stored).
 * Replaces: coupe::Rib::partition, coupe/src/algorithms/recursive_bisection.rs:829-851, :911-930.
 */
enum coupe_err coupe_hip_rib_device(coupe_hip_ctx *ctx, void *stream, uint64_t *d_partition, uintptr_t n,
		uintptr_t dimension, const double *d_points, enum coupe_type weight_type,
		const void *d_weights, uintptr_t iter_count, double tolerance);

/*
 * The rotation of the context's last Rib call.  Matrices are row-major
 * dimension x dimension, packed at the start of their arrays.
 * centroid / inertia: geometry.rs:256-267; axis: geometry.rs:269-288 (unit
 * eigenvector of the largest eigenvalue); aabb_to_obb: the Householder
 * reflection, geometry.rs:290-305; obb_to_aabb: its inverse, the map applied
 * to every point (geometry.rs:217-231).
 */
struct coupe_hip_rib_info {
	uintptr_t dimension;
	double centroid[3];
	double inertia[9];
	double axis[3];
	double aabb_to_obb[9];
	double obb_to_aabb[9];
};
enum coupe_err coupe_hip_last_rib(coupe_hip_ctx *ctx, struct coupe_hip_rib_info *info);

/*
 * RCB of a cartesian grid (coupe::Grid::rcb): `size` holds the cell counts of
 * the dimension (2 or 3) axes; the weights and the ids are in memory order,
 * x fastest (cell (x, y[, z]) at x + size[0] * (y + size[1] * z)).  Splits
 * start on axis 1 and cycle; ids are the split paths (left 0, right 1),
 * most significant first.  `threads` is the rayon pool size whose chunking
 * coupe's weighted median follows (0: RAYON_NUM_THREADS, else the number of
 * logical CPUs, as rayon's global pool).  Integer weights give coupe's ids
 * bit for bit.  An axis of 0 or more than 2^32 - 1 cells, a cell count past
 * 64 bits or iter_count above 62 returns COUPE_ERR_CRASH before any device
 * work.
 * Replaces: Grid::new_2d / new_3d + Grid::rcb, coupe/src/cartesian/mod.rs:127-180,
 * coupe/src/cartesian/rcb.rs:52-281.
 */
enum coupe_err coupe_hip_grid_rcb(coupe_hip_ctx *ctx, void *stream, uintptr_t dimension, const uintptr_t *size,
		enum coupe_type weight_type, const void *d_weights, uintptr_t iter_count, uint32_t threads,
		uint64_t *d_partition);

/*
 * Barycentres of mesh elements of one type, in device memory: element e has
 * the nodes element_nodes[e*k .. e*k+k) (0-based indices into the node
 * array, k = nodes_per_element); out[e*dimension + a] = (sum over its nodes
 * of coords[node*dimension + a], in node order, from 0.0) / k.  A node index
 * outside [0, n_nodes) returns COUPE_ERR_CRASH (coupe's indexing panics).
 * Replaces: the per-element loop of barycentres(), tools/lib/lib.rs:511-539
 * (the caller chooses the element blocks, as that function filters them).
 */
enum coupe_err coupe_hip_barycentres(coupe_hip_ctx *ctx, void *stream, const double *d_coords,
		uintptr_t n_nodes, uintptr_t dimension, const int64_t *d_element_nodes, uintptr_t n_elements,
		uintptr_t nodes_per_element, double *d_out);

/*
 * Per-part loads of a partition held in device memory: loads[p] = sum of the
 * weights of the points with id p, in the weight type (int sums wrap like
 * Rust's release arithmetic; int64 exact; double in device order).
 * `loads` is a host array of num_parts values of weight_type.  Ids >=
 * num_parts return COUPE_ERR_CRASH (coupe would panic on the index).
 * Replaces: coupe::imbalance::compute_parts_load, coupe/src/imbalance.rs:14-39.
 */
enum coupe_err coupe_hip_parts_load(coupe_hip_ctx *ctx, void *stream, const uint64_t *d_partition,
		uintptr_t n, enum coupe_type weight_type, const void *d_weights, uintptr_t num_parts,
		void *loads);
/*
 * Imbalance of a partition: the largest (load - ideal) / ideal over the parts,
 * ideal = total / num_parts; 0 when num_parts or the ideal load is 0.
 * Replaces: coupe::imbalance::imbalance, coupe/src/imbalance.rs:41-77.
 */
enum coupe_err coupe_hip_imbalance(coupe_hip_ctx *ctx, void *stream, const uint64_t *d_partition,
		uintptr_t n, enum coupe_type weight_type, const void *d_weights, uintptr_t num_parts,
		double *imbalance);

/*
 * Serving: a request server with continuous batching and resident meshes.
 *
 * `slots` workers each own a context and a stream.  A free worker takes the
 * oldest queued request that can run and, unless it must run alone, the
 * queued requests of the same dimension and weight type (in arrival order)
 * that fit its batch limits and the memory budget; it packs them, runs them
 * as one batch, and copies every request's ids to its output.  Requests that
 * arrive while batches run are admitted as workers and memory free up, and
 * batches of different workers overlap on the device.  A request's ids are
 * exactly those of coupe_hip_rcb_device on the same input.
 *
 * Runs alone: a request on a resident mesh, and a request with
 * iter_count > 24 (a batch of several instances needs iter_count <= 24).
 *
 * Device memory is admitted against `memory_budget` from live figures: the
 * buffers every worker holds now plus the resident meshes; a batch that
 * does not fit waits, and idle workers give their buffers back meanwhile.
 *
 * Inputs may be host or device pointers.  Device inputs written by work on
 * a stream are ordered by passing that stream in the request: the worker
 * waits for that stream's work up to the submit (NULL: the caller has
 * synchronised).  Inputs and outputs must stay valid until the request is
 * waited for.  No counterpart in coupe.
 */
typedef struct coupe_hip_server coupe_hip_server;
struct coupe_hip_server_config {
	int device;
	uint32_t slots;               /* concurrent batches (streams), >= 1 */
	uint64_t memory_budget;       /* device bytes the workers and the meshes may hold together */
	uint64_t max_batch_points;    /* points packed into one batch */
	uint32_t max_batch_requests;  /* requests packed into one batch */
};
struct coupe_hip_request {
	const double *points;         /* n * dimension doubles */
	const void *weights;          /* n values of weight_type */
	enum coupe_type weight_type;
	uintptr_t n;
	uintptr_t dimension;
	uintptr_t iter_count;
	double tolerance;
	uint64_t *partition;          /* n ids (may be NULL for a mesh request: ids kept on the mesh only) */
	int device_input;             /* points and weights are device pointers */
	int device_output;            /* partition is a device pointer */
	void *stream;                 /* device inputs: the stream that produced them, or NULL */
	uint64_t mesh;                /* 0, or a resident mesh: points, weights, n, dimension and
	                                 weight_type are then the mesh's */
};
struct coupe_hip_request_stats {
	double queue_ms;              /* submit -> batch start */
	double run_ms;                /* batch start -> ids written */
	double total_ms;              /* submit -> ids written */
	uint32_t batch_requests;      /* requests in its batch */
	uint64_t batch_points;        /* points in its batch */
};
coupe_hip_server *coupe_hip_server_create(const struct coupe_hip_server_config *config);
/* Runs every submitted request, waits for the waiters to leave, then frees the server. */
void coupe_hip_server_destroy(coupe_hip_server *server);
/* COUPE_ERR_ALLOC when the request alone cannot fit the budget next to the
 * resident meshes or exceeds max_batch_points; COUPE_ERR_NOT_FOUND for an
 * unknown mesh. */
enum coupe_err coupe_hip_server_submit(coupe_hip_server *server, const struct coupe_hip_request *request,
		uint64_t *ticket);
/* Blocks until the request is done; returns its status.  Each ticket is
 * waited for once: another wait on it returns COUPE_ERR_NOT_FOUND. */
enum coupe_err coupe_hip_server_wait(coupe_hip_server *server, uint64_t ticket,
		struct coupe_hip_request_stats *stats);
/* *done = 1 when the request is done (it must still be waited for). */
enum coupe_err coupe_hip_server_poll(coupe_hip_server *server, uint64_t ticket, int *done);
/* A request still queued is removed (*cancelled = 1; its ticket is then
 * unknown); one that started completes as usual (*cancelled = 0). */
enum coupe_err coupe_hip_server_cancel(coupe_hip_server *server, uint64_t ticket, int *cancelled);
struct coupe_hip_server_stats {
	uint64_t memory_budget;
	uint64_t device_bytes;        /* held now by the workers and the meshes */
	uint64_t mesh_bytes;          /* held by the resident meshes */
	uint64_t peak_device_bytes;   /* largest device_bytes seen */
	uint32_t queued, running, meshes;
	uint64_t completed, cancelled;
};
enum coupe_err coupe_hip_server_stats(coupe_hip_server *server, struct coupe_hip_server_stats *stats);

/*
 * Resident meshes.  A mesh keeps its working copy in HBM: the f32
 * coordinates of its points (`as f32`, as coupe's rcb() takes them) and its
 * weights, in the caller's point order, plus the ids of its last partition.
 * A request with `mesh` set partitions that working copy with no input
 * copy, starting from what the mesh's previous partition taught the engine
 * (weight mode, split-search passes per level); its ids equal those of a
 * fresh coupe_hip_rcb_device call on the mesh's points.  Updates apply after
 * the mesh's queued and running requests and before later ones.
 */
/* A mesh of n points (f64 AoS) and weights, from host or device memory. */
enum coupe_err coupe_hip_mesh_create(coupe_hip_server *server, const double *points, const void *weights,
		enum coupe_type weight_type, uintptr_t n, uintptr_t dimension, int device_input, void *stream,
		uint64_t *mesh);
/* Points [first, first + count) and/or their weights replaced in place
 * (either may be NULL). */
enum coupe_err coupe_hip_mesh_update(coupe_hip_server *server, uint64_t mesh, uintptr_t first, uintptr_t count,
		const double *points, const void *weights, int device_input, void *stream);
/* The points of the mesh's last partition grouped by part (the order of
 * coupe's permuted working set after rcb(), recursive_bisection.rs:122-181,
 * made stable): perm (n entries) lists part 0's point indices in increasing
 * order, then part 1's, ...; offsets (num_parts + 1 entries) bound each
 * part's range.  Kept with the mesh until its next partition or update.
 * perm / offsets may be NULL (then only *num_parts is returned);
 * COUPE_ERR_LEN_MISMATCH when offsets_len < num_parts + 1. */
enum coupe_err coupe_hip_mesh_parts(coupe_hip_server *server, uint64_t mesh, uint32_t *perm, uint64_t *offsets,
		uintptr_t offsets_len, uint64_t *num_parts, int device_output);
/* Waits for the mesh's requests, then frees it. */
enum coupe_err coupe_hip_mesh_destroy(coupe_hip_server *server, uint64_t mesh);

/*
 * Counters of the last call on a context (for benchmarks and tests).
 */
struct coupe_hip_stats {
	uint64_t points;          /* N of the last call */
	uint32_t levels;          /* bisection levels run */
	uint32_t passes;          /* histogram passes (all levels) */
	uint32_t max_passes;      /* largest pass count of one level */
	uint32_t weight_mode;     /* 0 int64 exact sums, 1 f64 sums */
	uint32_t weight_bytes;    /* bytes per stored weight (2, 4 or 8) */
	uint32_t bins;            /* histogram bins per part per pass */
	uint32_t reruns;          /* 1 when a split search outran the queued passes and the call ran again */
	uint64_t unit_points;     /* points per work unit */
	uint64_t device_bytes;    /* device memory held by the context */
	uint64_t first_read;      /* points read by the children's first passes (two levels per move) */
	uint32_t forwarded;       /* parts whose children's first pass came with the move that filled them */
	uint32_t full_first;      /* levels whose children's first pass took every candidate (last pair, adaptive) */
	uint32_t graph;           /* 0: launched kernel by kernel; 1: one launch of the graph captured from an
	                             earlier call of the same shape; 2: captured by this call, then launched */
	uint32_t tuned;           /* 1 when the call ran with switches the autotuner chose for its class */
	struct coupe_hip_tuning tuning;   /* the switches the call ran with */
};
enum coupe_err coupe_hip_last_stats(const coupe_hip_ctx *ctx, struct coupe_hip_stats *out);

/*
 * Per-kernel timing of the last call, measured with HIP events recorded on
 * the call's stream around every launch when profiling is enabled, and the
 * algorithmic bytes each kernel class must move (its required reads and
 * writes, counted once).  Profiling adds two event records per launch.
 */
enum coupe_hip_kernel {
	COUPE_HIP_K_PROLOGUE,   /* AoS f64 -> SoA f32, bounding box, weight sums */
	COUPE_HIP_K_HIST,       /* split-candidate histogram pass */
	COUPE_HIP_K_DECIDE,     /* split decision replay */
	COUPE_HIP_K_PARTITION,  /* stable partition of all parts of a level */
	COUPE_HIP_K_IDS,        /* part ids through the original index */
	COUPE_HIP_K_OTHER,      /* tables, scans, weight conversion */
	COUPE_HIP_K_MOMENTS,    /* Rib: centroid and inertia-matrix passes */
	COUPE_HIP_K_COUNT
};
struct coupe_hip_kernel_time {
	uint32_t launches;
	double total_ms;          /* sum of launch durations */
	double algo_bytes;        /* algorithmic bytes, summed over launches */
};
enum coupe_err coupe_hip_set_profiling(coupe_hip_ctx *ctx, int enable);
enum coupe_err coupe_hip_kernel_times(const coupe_hip_ctx *ctx, struct coupe_hip_kernel_time *out,
		int count);

#ifdef __cplusplus
}
#endif

#endif /* COUPE_HIP_H */
