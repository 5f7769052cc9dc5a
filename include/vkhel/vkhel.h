/*
 * vkhel public C API -- MI355X (gfx950) implementation.
 *
 * Declaration-for-declaration the interface of lolzballs/vkhel
 * include/vkhel/vkhel.h:8-53, so existing callers relink unchanged.  Each
 * entry point below replaces the reference function of the same name:
 *
 *   vkhel_ctx_create / _destroy            src/vkhel.c:4-13
 *   vkhel_ntt_tables_create / _destroy     src/ntt_tables.c:65-87
 *   vkhel_vector_create / create2          src/vector.c:206-229
 *   vkhel_vector_destroy                   src/vector.c:231-235
 *   vkhel_vector_dup                       src/vector.c:249-260
 *   vkhel_vector_copy_from_host            src/vector.c:262-268
 *   vkhel_vector_map / _unmap              src/vector.c:270-296
 *   vkhel_vector_elemfma                   src/vector.c:298-340
 *   vkhel_vector_elemmod                   src/vector.c:342-386
 *   vkhel_vector_elemmul                   src/vector.c:388-427
 *   vkhel_vector_elemgtadd                 src/vector.c:429-468
 *   vkhel_vector_elemgtsub                 src/vector.c:470-511
 *   vkhel_vector_forward_transform         src/vector.c:513-574
 *   vkhel_vector_inverse_transform         src/vector.c:576-657
 *
 * Semantics kept from the reference: every call is synchronous (results are
 * visible to vkhel_vector_map on return); calls on one context are not
 * thread-safe; output vectors are caller-allocated (only dup allocates);
 * `operand == result` is allowed; tables are host objects with no context
 * and are uploaded to device memory on first use per context.
 *
 * Errors: the reference asserts.  This library prints a message and aborts
 * by default; vkhel_ext_set_abort_on_error(0) (vkhel_ext.h) makes calls
 * return early and record the error instead.
 */
#ifndef VKHEL_H
#define VKHEL_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct vkhel_ctx;
struct vkhel_ctx *vkhel_ctx_create();
void vkhel_ctx_destroy(struct vkhel_ctx *);

struct vkhel_ntt_tables;
struct vkhel_ntt_tables *vkhel_ntt_tables_create(
		uint64_t n, uint64_t q, uint64_t w);
void vkhel_ntt_tables_destroy(struct vkhel_ntt_tables *);

struct vkhel_vector;
struct vkhel_vector *vkhel_vector_create(struct vkhel_ctx *, uint64_t length);
struct vkhel_vector *vkhel_vector_create2(struct vkhel_ctx *, uint64_t length,
		bool zero);
void vkhel_vector_destroy(struct vkhel_vector *);
struct vkhel_vector *vkhel_vector_dup(struct vkhel_vector *);
void vkhel_vector_copy_from_host(struct vkhel_vector *, const uint64_t *);
void vkhel_vector_map(struct vkhel_vector *, void **, size_t);
void vkhel_vector_unmap(struct vkhel_vector *);

void vkhel_vector_elemfma(
		const struct vkhel_vector *a,
		const struct vkhel_vector *b,
		struct vkhel_vector *result,
		uint64_t multiplier, uint64_t mod);
void vkhel_vector_elemmod(
		const struct vkhel_vector *a,
		struct vkhel_vector *result, uint64_t mod, uint64_t q);
void vkhel_vector_elemmul(
		const struct vkhel_vector *a,
		const struct vkhel_vector *b,
		struct vkhel_vector *result, uint64_t mod);
void vkhel_vector_elemgtadd(const struct vkhel_vector *operand,
		struct vkhel_vector *result,
		uint64_t bound, uint64_t diff);
void vkhel_vector_elemgtsub(
		const struct vkhel_vector *operand,
		struct vkhel_vector *result,
		uint64_t bound, uint64_t diff, uint64_t mod);
void vkhel_vector_forward_transform(
		const struct vkhel_vector *operand,
		struct vkhel_vector *result,
		struct vkhel_ntt_tables *ntt);
void vkhel_vector_inverse_transform(
		const struct vkhel_vector *operand,
		struct vkhel_vector *result,
		struct vkhel_ntt_tables *ntt);

#ifdef __cplusplus
}
#endif

#endif
