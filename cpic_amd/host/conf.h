/*
 * conf.h -- reader for the configuration-file syntax cpic uses (libconfig's
 * grammar: the .conf files of the reference), written for this driver because
 * libconfig is not available.
 *
 * Supported: settings `name = value` or `name : value` ended by an optional
 * ';' or ','; groups { }, lists ( ), arrays [ ]; integers (decimal, hex,
 * optional L/LL suffix), floats, booleans, strings (with escapes and adjacent
 * literal concatenation); comments '#', '//' and C-style blocks; and
 * `@include "file"` resolved relative to the including file's directory
 * (the reference resolves includes against the conf directory,
 * src/cpic.c:152-155).
 */
#ifndef CPIC_CONF_H
#define CPIC_CONF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum cfg_type {
	CFG_GROUP = 1,
	CFG_LIST,
	CFG_ARRAY,
	CFG_INT,
	CFG_FLOAT,
	CFG_BOOL,
	CFG_STRING,
};

typedef struct cfg_node cfg_node_t;

struct cfg_node {
	char *name;            /* NULL for list/array elements */
	int type;
	int line;
	int64_t i;             /* CFG_INT, CFG_BOOL */
	double f;              /* CFG_FLOAT */
	char *s;               /* CFG_STRING */
	cfg_node_t **child;    /* CFG_GROUP, CFG_LIST, CFG_ARRAY */
	int nchild, capchild;
	cfg_node_t *parent;
};

typedef struct cfg {
	cfg_node_t *root;      /* a group */
	char err[256];         /* parse error message */
	int err_line;
	char path[4096];       /* file name of the top-level file */
} cfg_t;

/* Parse a file or a string.  On failure returns non-zero and fills err. */
int cfg_read_file(cfg_t *c, const char *path);
int cfg_read_string(cfg_t *c, const char *text);
void cfg_destroy(cfg_t *c);

/* Path lookup "a.b.c"; list elements are addressed as "[i]" components
 * ("species.[0].mass").  NULL if absent. */
cfg_node_t *cfg_lookup(const cfg_t *c, const char *path);
cfg_node_t *cfg_member(const cfg_node_t *n, const char *name);
cfg_node_t *cfg_elem(const cfg_node_t *n, int i);
int cfg_length(const cfg_node_t *n);

/* Typed getters: return 0 on success, -1 if absent or of another type.
 * Integers are accepted where a float is asked. */
int cfg_get_int64(const cfg_node_t *n, int64_t *v);
int cfg_get_float(const cfg_node_t *n, double *v);
int cfg_get_bool(const cfg_node_t *n, int *v);
int cfg_get_string(const cfg_node_t *n, const char **v);

int cfg_lookup_int64(const cfg_t *c, const char *path, int64_t *v);
int cfg_lookup_float(const cfg_t *c, const char *path, double *v);
int cfg_lookup_bool(const cfg_t *c, const char *path, int *v);
int cfg_lookup_string(const cfg_t *c, const char *path, const char **v);

/* config_array_float (src/config.c:10-58): an array of exactly `size`
 * numbers, or a single number when size is 1. */
int cfg_array_float(const cfg_node_t *n, double *v, int64_t size);
int cfg_lookup_array_float(const cfg_t *c, const char *path, double *v, int64_t size);
/* config_lookup_array_int (src/config.c:61-72) */
int cfg_lookup_array_int(const cfg_t *c, const char *path, int64_t *v, int64_t size);

#ifdef __cplusplus
}
#endif

#endif
