/* The runtime descriptor passed to every generated or hand-written kernel
 * entry point.  Valid C99.
 *
 * Layout is byte-identical to taco's descriptor (reference
 * include/taco/taco_tensor_t.h:11-23): same members, same order, same types,
 * so that a module built by this backend can be handed descriptors built by
 * taco's TensorStorage (src/storage/storage.cpp:127+) and vice versa.
 *
 * Conventions used by this backend:
 *   - csize is the component size in bits (64 = f64, 32 = f32, 16 = bf16);
 *   - indices[k] holds, for a dense mode, one array whose element 0 is the
 *     mode size; for a compressed mode, two arrays: pos (int32, dim+1
 *     entries) and crd (int32, one per stored entry);
 *   - the descriptor and its small arrays (dimensions, mode_ordering,
 *     mode_types, indices) live in host memory; pos/crd/vals live in device
 *     memory (or managed memory when unified memory is enabled).
 */
#ifndef TACO_TENSOR_T_DEFINED
#define TACO_TENSOR_T_DEFINED

#include <stdint.h>

typedef enum { taco_mode_dense, taco_mode_sparse } taco_mode_t;

typedef struct taco_tensor_t {
  int32_t      order;         /* number of modes */
  int32_t*     dimensions;    /* size of each mode */
  int32_t      csize;         /* component size (bits) */
  int32_t*     mode_ordering; /* storage order of the modes */
  taco_mode_t* mode_types;    /* dense or compressed per stored level */
  uint8_t***   indices;       /* per level: {size} or {pos, crd} */
  uint8_t*     vals;          /* component values */
  uint8_t*     fill_value;    /* value of entries that are not stored */
  int32_t      vals_size;     /* capacity of vals, in components */
} taco_tensor_t;

#ifdef __cplusplus
extern "C" {
#endif

/* Allocate a descriptor (reference src/taco_tensor_t.cpp:32-62).  The
 * descriptor and its small arrays are host allocations; fill_ptr, if not
 * NULL, points to csize/8 bytes copied into fill_value. */
taco_tensor_t* init_taco_tensor_t(int32_t order, int32_t csize,
                                  int32_t* dimensions, int32_t* modeOrdering,
                                  taco_mode_t* mode_types, void* fill_ptr);

/* Free a descriptor made by init_taco_tensor_t (reference
 * src/taco_tensor_t.cpp:64-79).  Does not free pos/crd/vals. */
void deinit_taco_tensor_t(taco_tensor_t* t);

#ifdef __cplusplus
}
#endif

#endif
