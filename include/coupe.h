/*
 * coupe.h -- C ABI of the MI355X RCB library (libcoupe_hip.so).
 *
 * These are the entry points, enums and values of coupe's C binding for the
 * RCB path, so that a C program written against coupe links unchanged
 * against this library.  Each declaration names the reference interface it
 * replaces (paths relative to the coupe repository root).
 *
 * Computation runs on the current HIP device; results are written to host
 * memory exactly as coupe writes them.
 */
#ifndef COUPE_H
#define COUPE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Status codes.  Same names and values as coupe-ffi/include/coupe.h:16-57
 * (coupe-ffi/src/lib.rs:35-46).  Device out-of-memory maps to
 * COUPE_ERR_ALLOC; any other HIP runtime failure maps to COUPE_ERR_CRASH.
 */
enum coupe_err {
	COUPE_ERR_OK,            /* 0: success */
	COUPE_ERR_ALLOC,         /* 1: host or device allocation failed */
	COUPE_ERR_CRASH,         /* 2: internal failure (HIP error) */
	COUPE_ERR_BAD_DIMENSION, /* 3: dimension is not supported (RCB: 2 or 3) */
	COUPE_ERR_BAD_TYPE,      /* 4: value type not supported */
	COUPE_ERR_BIPART_ONLY,   /* 5: kept for ABI compatibility */
	COUPE_ERR_LEN_MISMATCH,  /* 6: data sets have different lengths */
	COUPE_ERR_NOT_FOUND,     /* 7: kept for ABI compatibility */
	COUPE_ERR_NEG_VALUES,    /* 8: kept for ABI compatibility */
};

/*
 * Static, NUL-terminated description of a status code; the same strings as
 * coupe_strerror in coupe-ffi/src/lib.rs:69-110.
 * Replaces: coupe-ffi/include/coupe.h:65 (coupe_strerror).
 */
const char *coupe_strerror(enum coupe_err err);

/*
 * A data set feeding values to an algorithm (array, constant or callback).
 * Replaces: coupe-ffi/include/coupe.h:77 (typedef struct coupe_data).
 */
typedef struct coupe_data coupe_data;

/*
 * Element types of a data set; same values as coupe-ffi/include/coupe.h:86-93.
 */
enum coupe_type {
	COUPE_INT,    /* 0: int */
	COUPE_INT64,  /* 1: int64_t */
	COUPE_DOUBLE, /* 2: double */
};

/*
 * Frees a data set (never the memory it refers to).  NULL is accepted.
 * Replaces: coupe-ffi/include/coupe.h:105 (coupe_data_free),
 * coupe-ffi/src/lib.rs:112-120.
 */
void coupe_data_free(coupe_data *data);

/*
 * Data set over a caller-owned contiguous array of `len` elements of `type`
 * (or, for points, `len` points of `dimension` doubles each).  The array is
 * borrowed, must outlive the data set, and is read from several threads.
 * Returns NULL on allocation failure.
 * Replaces: coupe-ffi/include/coupe.h:127 (coupe_data_array),
 * coupe-ffi/src/lib.rs:122-129.
 */
coupe_data *coupe_data_array(uintptr_t len, enum coupe_type type, const void *data);

/*
 * Data set repeating the value at `value` `len` times (value is borrowed).
 * Returns NULL on allocation failure.
 * Replaces: coupe-ffi/include/coupe.h:147 (coupe_data_constant),
 * coupe-ffi/src/lib.rs:131-138.
 */
coupe_data *coupe_data_constant(uintptr_t len, enum coupe_type type, const void *value);

/*
 * Data set whose element i is read through `i_th(context, i)`.  The callback
 * may be called from several threads, in any order, more than once or not
 * at all for some indices.  Returns NULL on allocation failure.
 * Replaces: coupe-ffi/include/coupe.h:185 (coupe_data_fn),
 * coupe-ffi/src/lib.rs:140-157.
 */
coupe_data *coupe_data_fn(const void *context, uintptr_t len, enum coupe_type type,
		const void *(*i_th)(const void *, uintptr_t));

/*
 * Recursive Coordinate Bisection into at most 2^iter_count parts.
 *
 * `points` holds N points of `dimension` doubles each (read as doubles
 * whatever type the data set declares, as coupe does); `weights` holds N
 * values of type int, int64_t or double.  `partition` receives N part ids
 * (coupe's usize ids: the leaf's heap index minus the smallest one used).
 *
 * Errors, checked in coupe's order: COUPE_ERR_LEN_MISMATCH when the data
 * sets differ in length, then COUPE_ERR_BAD_DIMENSION when `dimension` is
 * not 2 or 3.  An empty input returns COUPE_ERR_OK and writes nothing.
 * The ids are those of coupe::Rcb on the same input (bit for bit when the
 * weights are integers or integer-valued doubles).
 *
 * Replaces: coupe-ffi/include/coupe.h:278-280 (coupe_rcb),
 * coupe-ffi/src/lib.rs:272-304.
 */
enum coupe_err coupe_rcb(uintptr_t *partition, uintptr_t dimension,
		const coupe_data *points, const coupe_data *weights,
		uintptr_t iter_count, double tolerance);

/*
 * Recursive Inertial Bisection: coupe_rcb after a change of basis whose first
 * axis is the inertia axis of the points.  Same arguments, data-set rules and
 * errors as coupe_rcb.
 *
 * Replaces: coupe-ffi/include/coupe.h:294-296 (coupe_rib),
 * coupe-ffi/src/lib.rs:331-364.
 */
enum coupe_err coupe_rib(uintptr_t *partition, uintptr_t dimension,
		const coupe_data *points, const coupe_data *weights,
		uintptr_t iter_count, double tolerance);

#ifdef __cplusplus
}
#endif

#endif /* COUPE_H */
