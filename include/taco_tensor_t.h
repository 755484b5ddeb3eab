/*
 * taco_tensor_t: the TACO C tensor descriptor, the carrier of every operand
 * at the tentaco drop-in boundary.
 *
 * ABI-identical to the reference's struct (taco_tensor_t.h:4-15 of
 * zhang677/tenTACO): same field order, same types, 64 bytes on LP64 with
 * offsets 0/8/16/24/32/40/48/56. The layout has to match bit for bit so that
 * a caller built against the reference header can hand its structs to this
 * library unchanged; the static asserts below pin it.
 *
 * Field use on the MTTKRP/TTV path (only these are read):
 *   dimensions[m]       extent of mode m (for factors: {rows, R})
 *   indices[m][0]       pos array of level m (int32)
 *   indices[m][1]       crd array of level m (int32)
 *   vals                FP32 values
 * Level 0 of a CSF operand holds pos = {0, n0} and crd[n0] (crd = 0..N0-1
 * for a dense root, as compress_top_level produces in ds.h:426-436).
 * order, csize, mode_ordering, mode_types and vals_size are not read.
 */
#ifndef TACO_TENSOR_T_H
#define TACO_TENSOR_T_H
#include <stddef.h>
#include <stdint.h>

typedef enum { taco_mode_dense, taco_mode_sparse } taco_mode_t;

typedef struct taco_tensor_t {
  int32_t      order;         /* tensor order (number of modes) */
  int32_t*     dimensions;    /* tensor dimensions */
  int32_t      csize;         /* component size */
  int32_t*     mode_ordering; /* mode storage ordering */
  taco_mode_t* mode_types;    /* mode storage types */
  int32_t***   indices;       /* tensor index data (per mode) */
  float*       vals;          /* tensor values */
  int32_t      vals_size;     /* values array size */
} taco_tensor_t;

#if defined(__cplusplus)
static_assert(sizeof(taco_tensor_t) == 64, "taco_tensor_t must be 64 bytes");
static_assert(offsetof(taco_tensor_t, order) == 0, "order @0");
static_assert(offsetof(taco_tensor_t, dimensions) == 8, "dimensions @8");
static_assert(offsetof(taco_tensor_t, csize) == 16, "csize @16");
static_assert(offsetof(taco_tensor_t, mode_ordering) == 24, "mode_ordering @24");
static_assert(offsetof(taco_tensor_t, mode_types) == 32, "mode_types @32");
static_assert(offsetof(taco_tensor_t, indices) == 40, "indices @40");
static_assert(offsetof(taco_tensor_t, vals) == 48, "vals @48");
static_assert(offsetof(taco_tensor_t, vals_size) == 56, "vals_size @56");
#elif defined(__STDC_VERSION__) && __STDC_VERSION__ >= 201112L
_Static_assert(sizeof(taco_tensor_t) == 64, "taco_tensor_t must be 64 bytes");
_Static_assert(offsetof(taco_tensor_t, indices) == 40, "indices @40");
_Static_assert(offsetof(taco_tensor_t, vals) == 48, "vals @48");
#endif

#endif /* TACO_TENSOR_T_H */
