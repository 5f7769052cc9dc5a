/*
 * C caller of the drop-in libvkhel.so: the checks of the reference's
 * test/vector.c:23-346 (same inputs and expected outputs, restated here)
 * through the unchanged public header, built with a plain C compiler.
 * Prints "<name> passed" per check and exits non-zero on the first
 * mismatch.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vkhel/vkhel.h>

static struct vkhel_ctx *ctx;

#define CHECK(cond, name) do { if (!(cond)) { \
	fprintf(stderr, "%s FAILED at %s:%d\n", name, __FILE__, __LINE__); \
	exit(1); } } while (0)

static int equal(struct vkhel_vector *v, const uint64_t *want, size_t len) {
	uint64_t *m = NULL;
	vkhel_vector_map(v, (void **) &m, len * sizeof(uint64_t));
	int ok = memcmp(m, want, len * sizeof(uint64_t)) == 0;
	vkhel_vector_unmap(v);
	return ok;
}

static struct vkhel_vector *from(const uint64_t *e, size_t len) {
	struct vkhel_vector *v = vkhel_vector_create(ctx, len);
	vkhel_vector_copy_from_host(v, e);
	return v;
}

int main(void) {
	ctx = vkhel_ctx_create();
	{
		uint64_t e[10];
		for (int i = 0; i < 10; i++)
			e[i] = (uint64_t) i;
		struct vkhel_vector *v = from(e, 10);
		CHECK(equal(v, e, 10), "copy_from_host");
		struct vkhel_vector *d = vkhel_vector_dup(v);
		CHECK(equal(d, e, 10), "dup");
		vkhel_vector_destroy(v);
		vkhel_vector_destroy(d);
		printf("copy_from_host passed\ndup passed\n");
	}
	{
		const uint64_t a[] = {1, 2, 3, 4, 5, 6, 7, 8};
		const uint64_t b[] = {9, 10, 11, 12, 13, 14, 15, 16};
		const uint64_t m100[] = {109, 210, 311, 412, 513, 614, 715, 47};
		struct vkhel_vector *va = from(a, 8), *vb = from(b, 8);
		struct vkhel_vector *c = vkhel_vector_create(ctx, 8);
		vkhel_vector_elemfma(va, vb, c, 100, 769);
		CHECK(equal(c, m100, 8), "elemfma");
		const uint64_t mod2[] = {1, 0, 1, 0, 1, 0, 1, 0};
		vkhel_vector_elemmod(va, c, 2, UINT64_MAX);
		CHECK(equal(c, mod2, 8), "elemmod");
		const uint64_t gts[] = {1, 2, 3, 4, 5, 9, 0, 1};
		vkhel_vector_elemgtsub(va, c, 5, 7, 10);
		CHECK(equal(c, gts, 8), "elemgtsub");
		const uint64_t gta[] = {1, 2, 5, 6, 7, 8, 9, 10};
		vkhel_vector_elemgtadd(va, c, 2, 2);
		CHECK(equal(c, gta, 8), "elemgtadd");
		vkhel_vector_destroy(va);
		vkhel_vector_destroy(vb);
		vkhel_vector_destroy(c);
		printf("elemfma passed\nelemmod passed\nelemgtsub passed\nelemgtadd passed\n");
	}
	{
		const uint64_t a[] = {706712574074152, 943467560561867, 1};
		const uint64_t b[] = {515910833966633, 96924929169117, 1};
		const uint64_t want[] = {231838787758587, 618753612121218, 1};
		struct vkhel_vector *va = from(a, 3), *vb = from(b, 3);
		struct vkhel_vector *c = vkhel_vector_create(ctx, 3);
		vkhel_vector_elemmul(va, vb, c, 1125891450734593);
		CHECK(equal(c, want, 3), "elemmul");
		vkhel_vector_destroy(va);
		vkhel_vector_destroy(vb);
		vkhel_vector_destroy(c);
		printf("elemmul passed\n");
	}
	{
		const uint64_t x[] = {94, 109, 11, 18};
		const uint64_t want[] = {82, 2, 81, 98};
		struct vkhel_ntt_tables *t = vkhel_ntt_tables_create(4, 113, 18);
		struct vkhel_vector *a = from(x, 4);
		struct vkhel_vector *b = vkhel_vector_create(ctx, 4);
		vkhel_vector_forward_transform(a, b, t);
		CHECK(equal(b, want, 4), "forward_transform");
		vkhel_vector_inverse_transform(b, b, t);
		CHECK(equal(b, x, 4), "inverse_transform");
		vkhel_vector_destroy(a);
		vkhel_vector_destroy(b);
		vkhel_ntt_tables_destroy(t);
		printf("forward_transform passed\ninverse_transform passed\n");
	}
	vkhel_ctx_destroy(ctx);
	return 0;
}
