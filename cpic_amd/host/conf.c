/*
 * conf.c -- recursive-descent reader for cpic's configuration files (the
 * libconfig grammar, see conf.h).
 */
#define _GNU_SOURCE
#include "conf.h"

#include <ctype.h>
#include <errno.h>
#include <libgen.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct lex {
	const char *p;        /* cursor */
	int line;
	cfg_t *c;
	char dir[4096];       /* directory for @include */
	int depth;            /* include depth */
} lex_t;

static int
fail(lex_t *L, const char *fmt, ...)
{
	va_list ap;
	if(L->c->err[0] == 0)
	{
		va_start(ap, fmt);
		vsnprintf(L->c->err, sizeof(L->c->err), fmt, ap);
		va_end(ap);
		L->c->err_line = L->line;
	}
	return -1;
}

static cfg_node_t *
node_new(int type, const char *name, int line)
{
	cfg_node_t *n = calloc(1, sizeof(*n));
	if(!n)
		return NULL;
	n->type = type;
	n->line = line;
	if(name)
		n->name = strdup(name);
	return n;
}

static void
node_free(cfg_node_t *n)
{
	int i;
	if(!n)
		return;
	for(i = 0; i < n->nchild; i++)
		node_free(n->child[i]);
	free(n->child);
	free(n->name);
	free(n->s);
	free(n);
}

static int
node_add(cfg_node_t *parent, cfg_node_t *child)
{
	if(parent->nchild == parent->capchild)
	{
		int cap = parent->capchild ? 2 * parent->capchild : 8;
		cfg_node_t **c = realloc(parent->child, sizeof(*c) * (size_t) cap);
		if(!c)
			return -1;
		parent->child = c;
		parent->capchild = cap;
	}
	parent->child[parent->nchild++] = child;
	child->parent = parent;
	return 0;
}

static void
skip_ws(lex_t *L)
{
	for(;;)
	{
		while(*L->p && isspace((unsigned char) *L->p))
		{
			if(*L->p == '\n')
				L->line++;
			L->p++;
		}
		if(*L->p == '#' || (L->p[0] == '/' && L->p[1] == '/'))
		{
			while(*L->p && *L->p != '\n')
				L->p++;
			continue;
		}
		if(L->p[0] == '/' && L->p[1] == '*')
		{
			L->p += 2;
			while(*L->p && !(L->p[0] == '*' && L->p[1] == '/'))
			{
				if(*L->p == '\n')
					L->line++;
				L->p++;
			}
			if(*L->p)
				L->p += 2;
			continue;
		}
		break;
	}
}

static int
is_name_start(int ch)
{
	return isalpha(ch) || ch == '*';
}

static int
is_name_char(int ch)
{
	return isalnum(ch) || ch == '_' || ch == '-' || ch == '*';
}

static int parse_settings(lex_t *L, cfg_node_t *group, int close);
static int parse_value(lex_t *L, cfg_node_t *n);

static int
parse_string(lex_t *L, char **out)
{
	size_t cap = 64, len = 0;
	char *s = malloc(cap);
	if(!s)
		return fail(L, "out of memory");
	for(;;)
	{
		if(*L->p != '"')
			break;
		L->p++;
		while(*L->p && *L->p != '"')
		{
			char ch = *L->p++;
			if(ch == '\n')
				L->line++;
			if(ch == '\\')
			{
				char e = *L->p++;
				switch(e)
				{
					case 'n': ch = '\n'; break;
					case 't': ch = '\t'; break;
					case 'r': ch = '\r'; break;
					case 'f': ch = '\f'; break;
					case '\\': ch = '\\'; break;
					case '"': ch = '"'; break;
					case 'x':
					{
						char hex[3] = {0, 0, 0};
						if(!isxdigit((unsigned char) L->p[0]) ||
								!isxdigit((unsigned char) L->p[1]))
						{
							free(s);
							return fail(L, "bad \\x escape");
						}
						hex[0] = L->p[0];
						hex[1] = L->p[1];
						L->p += 2;
						ch = (char) strtol(hex, NULL, 16);
						break;
					}
					default:
						free(s);
						return fail(L, "unknown escape \\%c", e);
				}
			}
			if(len + 2 > cap)
			{
				char *t;
				cap *= 2;
				t = realloc(s, cap);
				if(!t)
				{
					free(s);
					return fail(L, "out of memory");
				}
				s = t;
			}
			s[len++] = ch;
		}
		if(*L->p != '"')
		{
			free(s);
			return fail(L, "unterminated string");
		}
		L->p++;
		skip_ws(L);    /* adjacent literals are concatenated */
	}
	s[len] = 0;
	*out = s;
	return 0;
}

static int
parse_scalar(lex_t *L, cfg_node_t *n)
{
	const char *p = L->p;

	if(*p == '"')
	{
		n->type = CFG_STRING;
		return parse_string(L, &n->s);
	}
	if(strncasecmp(p, "true", 4) == 0 && !is_name_char((unsigned char) p[4]))
	{
		n->type = CFG_BOOL;
		n->i = 1;
		L->p += 4;
		return 0;
	}
	if(strncasecmp(p, "false", 5) == 0 && !is_name_char((unsigned char) p[5]))
	{
		n->type = CFG_BOOL;
		n->i = 0;
		L->p += 5;
		return 0;
	}
	{
		const char *q = p;
		int isfloat = 0, ishex = 0;
		char *end;
		if(*q == '+' || *q == '-')
			q++;
		if(q[0] == '0' && (q[1] == 'x' || q[1] == 'X'))
			ishex = 1;
		else
		{
			const char *r = q;
			while(isdigit((unsigned char) *r) || *r == '.' || *r == 'e' || *r == 'E' ||
					((*r == '+' || *r == '-') && (r[-1] == 'e' || r[-1] == 'E')))
			{
				if(*r == '.' || *r == 'e' || *r == 'E')
					isfloat = 1;
				r++;
			}
			if(r == q)
				return fail(L, "expected a value");
		}
		errno = 0;
		if(isfloat)
		{
			n->type = CFG_FLOAT;
			n->f = strtod(p, &end);
		}
		else
		{
			n->type = CFG_INT;
			if(ishex)
				n->i = (int64_t) strtoull(p, &end, 16);
			else
				n->i = strtoll(p, &end, 10);
		}
		if(end == p || errno == ERANGE)
			return fail(L, "bad number");
		L->p = end;
		if(!isfloat)
		{
			if(*L->p == 'L')
				L->p++;
			if(*L->p == 'L')
				L->p++;
		}
		if(is_name_char((unsigned char) *L->p) || *L->p == '.')
			return fail(L, "bad number");
	}
	return 0;
}

static int
parse_elems(lex_t *L, cfg_node_t *n, char close, int scalars_only)
{
	skip_ws(L);
	if(*L->p == close)
	{
		L->p++;
		return 0;
	}
	for(;;)
	{
		cfg_node_t *e = node_new(0, NULL, L->line);
		if(!e)
			return fail(L, "out of memory");
		if(node_add(n, e))
		{
			node_free(e);
			return fail(L, "out of memory");
		}
		skip_ws(L);
		if(scalars_only)
		{
			if(parse_scalar(L, e))
				return -1;
			if(n->nchild > 1 && e->type != n->child[0]->type &&
					!((e->type == CFG_INT || e->type == CFG_FLOAT) &&
					  (n->child[0]->type == CFG_INT || n->child[0]->type == CFG_FLOAT)))
				return fail(L, "array elements must have one type");
		}
		else if(parse_value(L, e))
			return -1;
		skip_ws(L);
		if(*L->p == ',')
		{
			L->p++;
			skip_ws(L);
			if(*L->p == close)   /* trailing comma */
			{
				L->p++;
				return 0;
			}
			continue;
		}
		if(*L->p == close)
		{
			L->p++;
			return 0;
		}
		return fail(L, "expected ',' or '%c'", close);
	}
}

static int
parse_value(lex_t *L, cfg_node_t *n)
{
	skip_ws(L);
	switch(*L->p)
	{
		case '{':
			L->p++;
			n->type = CFG_GROUP;
			return parse_settings(L, n, '}');
		case '(':
			L->p++;
			n->type = CFG_LIST;
			return parse_elems(L, n, ')', 0);
		case '[':
			L->p++;
			n->type = CFG_ARRAY;
			return parse_elems(L, n, ']', 1);
		default:
			return parse_scalar(L, n);
	}
}

static int include_file(lex_t *L, cfg_node_t *group, const char *path);

static int
parse_settings(lex_t *L, cfg_node_t *group, int close)
{
	for(;;)
	{
		char name[256];
		size_t k = 0;
		cfg_node_t *n;

		skip_ws(L);
		if(*L->p == 0)
		{
			if(close)
				return fail(L, "missing '%c'", close);
			return 0;
		}
		if(close && *L->p == close)
		{
			L->p++;
			return 0;
		}
		if(strncmp(L->p, "@include", 8) == 0)
		{
			char *file = NULL;
			char full[8192];
			int r;
			L->p += 8;
			skip_ws(L);
			if(*L->p != '"' || parse_string(L, &file))
				return fail(L, "@include needs a file name");
			if(file[0] == '/')
				r = snprintf(full, sizeof(full), "%s", file);
			else
				r = snprintf(full, sizeof(full), "%s/%s", L->dir, file);
			free(file);
			if(r < 0 || r >= (int) sizeof(full))
				return fail(L, "@include path too long");
			r = include_file(L, group, full);
			if(r)
				return r;
			continue;
		}
		if(!is_name_start((unsigned char) *L->p))
			return fail(L, "expected a setting name");
		while(is_name_char((unsigned char) *L->p))
		{
			if(k + 1 < sizeof(name))
				name[k++] = *L->p;
			L->p++;
		}
		name[k] = 0;
		skip_ws(L);
		if(*L->p != '=' && *L->p != ':')
			return fail(L, "expected '=' or ':' after '%s'", name);
		L->p++;
		if(cfg_member(group, name))
			return fail(L, "duplicate setting '%s'", name);
		n = node_new(0, name, L->line);
		if(!n)
			return fail(L, "out of memory");
		if(node_add(group, n))
		{
			node_free(n);
			return fail(L, "out of memory");
		}
		if(parse_value(L, n))
			return -1;
		skip_ws(L);
		if(*L->p == ';' || *L->p == ',')
			L->p++;
	}
}

static char *
slurp(const char *path)
{
	FILE *f = fopen(path, "rb");
	long n;
	char *buf;
	if(!f)
		return NULL;
	fseek(f, 0, SEEK_END);
	n = ftell(f);
	fseek(f, 0, SEEK_SET);
	buf = malloc((size_t) n + 1);
	if(buf && fread(buf, 1, (size_t) n, f) != (size_t) n)
	{
		free(buf);
		buf = NULL;
	}
	if(buf)
		buf[n] = 0;
	fclose(f);
	return buf;
}

static int
include_file(lex_t *L, cfg_node_t *group, const char *path)
{
	lex_t sub;
	char *text, *tmp;
	int r;
	if(L->depth > 16)
		return fail(L, "@include nested too deep");
	text = slurp(path);
	if(!text)
		return fail(L, "cannot read included file %s", path);
	memset(&sub, 0, sizeof(sub));
	sub.p = text;
	sub.line = 1;
	sub.c = L->c;
	sub.depth = L->depth + 1;
	tmp = strdup(path);
	if(!tmp)
	{
		free(text);
		return fail(L, "out of memory");
	}
	snprintf(sub.dir, sizeof(sub.dir), "%s", dirname(tmp));
	free(tmp);
	r = parse_settings(&sub, group, 0);
	free(text);
	return r;
}

static int
parse_text(cfg_t *c, const char *text, const char *dir)
{
	lex_t L;
	memset(&L, 0, sizeof(L));
	c->err[0] = 0;
	c->err_line = 0;
	c->root = node_new(CFG_GROUP, NULL, 1);
	if(!c->root)
		return -1;
	L.p = text;
	L.line = 1;
	L.c = c;
	snprintf(L.dir, sizeof(L.dir), "%s", dir);
	if(parse_settings(&L, c->root, 0))
	{
		node_free(c->root);
		c->root = NULL;
		return -1;
	}
	return 0;
}

int
cfg_read_string(cfg_t *c, const char *text)
{
	memset(c, 0, sizeof(*c));
	return parse_text(c, text, ".");
}

int
cfg_read_file(cfg_t *c, const char *path)
{
	char *text, *tmp;
	int r;
	memset(c, 0, sizeof(*c));
	snprintf(c->path, sizeof(c->path), "%s", path);
	text = slurp(path);
	if(!text)
	{
		snprintf(c->err, sizeof(c->err), "cannot read %s", path);
		return -1;
	}
	tmp = strdup(path);
	r = parse_text(c, text, dirname(tmp));
	free(tmp);
	free(text);
	return r;
}

void
cfg_destroy(cfg_t *c)
{
	node_free(c->root);
	c->root = NULL;
}

cfg_node_t *
cfg_member(const cfg_node_t *n, const char *name)
{
	int i;
	if(!n || n->type != CFG_GROUP)
		return NULL;
	for(i = 0; i < n->nchild; i++)
		if(n->child[i]->name && strcmp(n->child[i]->name, name) == 0)
			return n->child[i];
	return NULL;
}

cfg_node_t *
cfg_elem(const cfg_node_t *n, int i)
{
	if(!n || (n->type != CFG_LIST && n->type != CFG_ARRAY && n->type != CFG_GROUP))
		return NULL;
	if(i < 0 || i >= n->nchild)
		return NULL;
	return n->child[i];
}

int
cfg_length(const cfg_node_t *n)
{
	if(!n || (n->type != CFG_LIST && n->type != CFG_ARRAY && n->type != CFG_GROUP))
		return 0;
	return n->nchild;
}

cfg_node_t *
cfg_lookup(const cfg_t *c, const char *path)
{
	char buf[1024], *tok, *save = NULL;
	cfg_node_t *n;
	if(!c->root)
		return NULL;
	snprintf(buf, sizeof(buf), "%s", path);
	n = c->root;
	for(tok = strtok_r(buf, ".", &save); tok && n; tok = strtok_r(NULL, ".", &save))
	{
		if(tok[0] == '[')
			n = cfg_elem(n, atoi(tok + 1));
		else
			n = cfg_member(n, tok);
	}
	return n;
}

int
cfg_get_int64(const cfg_node_t *n, int64_t *v)
{
	if(!n || n->type != CFG_INT)
		return -1;
	*v = n->i;
	return 0;
}

int
cfg_get_float(const cfg_node_t *n, double *v)
{
	if(!n)
		return -1;
	if(n->type == CFG_FLOAT)
		*v = n->f;
	else if(n->type == CFG_INT)
		*v = (double) n->i;
	else
		return -1;
	return 0;
}

int
cfg_get_bool(const cfg_node_t *n, int *v)
{
	if(!n || n->type != CFG_BOOL)
		return -1;
	*v = (int) n->i;
	return 0;
}

int
cfg_get_string(const cfg_node_t *n, const char **v)
{
	if(!n || n->type != CFG_STRING)
		return -1;
	*v = n->s;
	return 0;
}

int cfg_lookup_int64(const cfg_t *c, const char *p, int64_t *v) { return cfg_get_int64(cfg_lookup(c, p), v); }
int cfg_lookup_float(const cfg_t *c, const char *p, double *v) { return cfg_get_float(cfg_lookup(c, p), v); }
int cfg_lookup_bool(const cfg_t *c, const char *p, int *v) { return cfg_get_bool(cfg_lookup(c, p), v); }
int cfg_lookup_string(const cfg_t *c, const char *p, const char **v) { return cfg_get_string(cfg_lookup(c, p), v); }

int
cfg_array_float(const cfg_node_t *n, double *v, int64_t size)
{
	int64_t i;
	if(!n)
		return -1;
	if(n->type == CFG_ARRAY)
	{
		if(n->nchild != size)
		{
			fprintf(stderr, "Line %d: The setting %s should have %d dimensions, but %d found.\n",
					n->line, n->name ? n->name : "?", (int) size, n->nchild);
			return 1;
		}
		for(i = 0; i < size; i++)
			if(cfg_get_float(n->child[i], &v[i]))
				return 1;
		return 0;
	}
	if(n->type == CFG_FLOAT || n->type == CFG_INT)
	{
		if(size != 1)
		{
			fprintf(stderr, "Line %d: The setting %s should have %d dimensions, but only one found.\n",
					n->line, n->name ? n->name : "?", (int) size);
			return 1;
		}
		return cfg_get_float(n, &v[0]);
	}
	fprintf(stderr, "Line %d: The setting %s is expected to be and array of %d dimensions.\n",
			n->line, n->name ? n->name : "?", (int) size);
	return 1;
}

int
cfg_lookup_array_float(const cfg_t *c, const char *path, double *v, int64_t size)
{
	return cfg_array_float(cfg_lookup(c, path), v, size);
}

int
cfg_lookup_array_int(const cfg_t *c, const char *path, int64_t *v, int64_t size)
{
	cfg_node_t *n = cfg_lookup(c, path);
	int64_t i;
	if(!n || n->type != CFG_ARRAY || n->nchild < size)
		return -1;
	for(i = 0; i < size; i++)
		if(cfg_get_int64(n->child[i], &v[i]))
			return -1;
	return 0;
}
