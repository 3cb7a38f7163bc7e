/*
 * output.h -- field output of the host driver in the reference's on-disk
 * format (src/output.c:63-130, 400-635): per sampled iteration, the raw
 * padded FP64 matrices bin/<iter>/{rho,phi,E_X,E_Y}.bin and an XDMF file
 * xdmf/fields-iter<iter>.xdmf whose HyperSlabs select the nx x ny interior.
 */
#ifndef CPIC_HOST_OUTPUT_H
#define CPIC_HOST_OUTPUT_H

#include <limits.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct sim;

typedef struct output {
	int enabled;
	char path[PATH_MAX];
	int64_t period_field;      /* output.period.field (default 1) */
	int64_t period_particle;   /* output.period.particle (read, unused as in the reference) */
	int64_t max_slices;        /* output.slices (default 1) */
	int64_t alignment;         /* output.alignment (default 512) */
	double *dense;             /* nx*ny staging */
	double *padded;            /* largest padded matrix */
	size_t padded_bytes;
} output_t;

/* output_init (src/output.c:63-130): reads output.path (no path: output
 * disabled), the periods, slices and alignment, and creates the
 * directories.  Returns 0 or -1 with a message. */
int output_init(struct sim *sim, output_t *out);

/* output_fields (src/output.c:601-635) for sim->iter, when enabled and
 * iter is a multiple of output.period.field. */
int output_fields(struct sim *sim);

void output_free(output_t *out);

#ifdef __cplusplus
}
#endif

#endif
