/*
 * Basic types of the AOCL-DA C API, MI355X build.
 * Values and names follow source/include/aoclda_types.h:42-92 of the reference so that
 * callers compiled against the reference headers are ABI-compatible.
 */
#ifndef AOCLDA_TYPES
#define AOCLDA_TYPES

#include <inttypes.h>
#include <stddef.h>
#include <stdint.h>

/* Floating point precision of a handle (aoclda_types.h:42-49). */
enum da_precision_ {
    da_unknown = 255,
    da_double = 0,
    da_single = 1,
};
typedef enum da_precision_ da_precision;

/* Storage order of 2-D arrays (aoclda_types.h:54-60). */
enum da_order_ {
    row_major = 0,
    column_major,
};
typedef enum da_order_ da_order;

/* Integer type: 32-bit by default, 64-bit with AOCLDA_ILP64 (aoclda_types.h:80-92). */
#if defined(AOCLDA_ILP64)
typedef int64_t da_int;
#define DA_INT_MAX INT64_MAX
#define DA_INT_MIN INT64_MIN
#define DA_INT_FMT PRId64
#else
typedef int32_t da_int;
#define DA_INT_MAX INT32_MAX
#define DA_INT_MIN INT32_MIN
#define DA_INT_FMT PRId32
#endif

#endif /* AOCLDA_TYPES */
