/*
 * output.c -- field output in the reference's format (src/output.c).
 *
 * For each sampled iteration the four fields are downloaded from the
 * device and laid out as the reference's padded matrices (one rank,
 * src/field.c:21-135), each written whole, padded with zero bytes to a
 * multiple of output.alignment:
 *   rho  (ny+1) rows x 2(nx/2+1) columns: rows 0..ny-1 the charge density;
 *        the ghost row ny and the padding columns are NaN (the reference's
 *        ghost row keeps the pre-fold contributions, which the device does
 *        not store separately; neither is selected by the XDMF);
 *   phi  (ny+3) rows x 2(nx/2+1) columns: row 0 = phi[ny-1], rows 1..ny =
 *        phi, rows ny+1, ny+2 = phi[0], phi[1] (comm_phi's periodic ghosts,
 *        src/comm_field.c:140-201); padding columns NaN;
 *   E_X, E_Y  (ny+1) rows x nx columns: row ny = row 0 (field_E_compute
 *        fills the ghost row from the periodic phi, src/field.c:359-416).
 * The XDMF text follows write_xdmf_fields / write_field_attribute
 * (src/output.c:400-459) line for line, so readers of the reference's
 * output read this one unchanged.
 */
#define _GNU_SOURCE
#include "output.h"
#include "sim.h"

#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#define err(...) fprintf(stderr, __VA_ARGS__)

static int
make_dir(const char *p)
{
	if(mkdir(p, 0700) && errno != EEXIST)
	{
		perror("mkdir");
		return -1;
	}
	return 0;
}

int
output_init(sim_t *sim, output_t *out)
{
	const char *path;
	char p[PATH_MAX];

	memset(out, 0, sizeof(*out));
	if(cfg_lookup_string(sim->conf, "output.path", &path))
	{
		if(!sim->quiet)
			err("No output path specified, output will not be saved\n");
		return 0;
	}
	if(strlen(path) >= PATH_MAX - 32)
	{
		err("Too long output path\n");
		return -1;
	}
	strcpy(out->path, path);
	if(cfg_lookup_int64(sim->conf, "output.period.field", &out->period_field))
	{
		err("Using default period for fields of 1 sample/iteration\n");
		out->period_field = 1;
	}
	if(cfg_lookup_int64(sim->conf, "output.period.particle", &out->period_particle))
	{
		err("Using default period for particles of 1 sample/iteration\n");
		out->period_particle = 1;
	}
	if(cfg_lookup_int64(sim->conf, "output.slices", &out->max_slices))
	{
		err("Using one slice for field output\n");
		out->max_slices = 1;
	}
	if(cfg_lookup_int64(sim->conf, "output.alignment", &out->alignment))
	{
		err("Using default 512 bytes for alignment\n");
		out->alignment = 512;
	}
	if(out->max_slices < 1 || sim->ntpoints[Y] % out->max_slices)
	{
		err("The number of slices %ld doesn't divide the blocksize in Y of %ld\n",
				(long) out->max_slices, (long) sim->ntpoints[Y]);
		return -1;
	}
	if(out->alignment < 8 || out->alignment % 8)
	{
		err("The output alignment must be a positive multiple of 8 bytes\n");
		return -1;
	}
	if(make_dir(out->path))
		goto fail;
	if(snprintf(p, sizeof(p), "%s/xdmf", out->path) >= (int) sizeof(p) || make_dir(p))
		goto fail;
	if(snprintf(p, sizeof(p), "%s/bin", out->path) >= (int) sizeof(p) || make_dir(p))
		goto fail;
	out->enabled = 1;
	return 0;
fail:
	err("Cannot create output directories\n");
	return -1;
}

void
output_free(output_t *out)
{
	if(!out)
		return;
	free(out->dense);
	free(out->padded);
	out->dense = out->padded = NULL;
}

/* shape of the reference's padded matrix of a field, and the row of the
 * interior's first line in it */
static void
padded_shape(const sim_t *s, const char *name, int64_t *rows, int64_t *cols, int64_t *dy)
{
	int64_t nx = s->ntpoints[X], ny = s->ntpoints[Y];
	int64_t rx = 2 * (nx / 2 + 1) > nx ? 2 * (nx / 2 + 1) : nx;
	*dy = 0;
	if(!strcmp(name, "rho"))
	{
		*rows = ny + 1;
		*cols = rx;
	}
	else if(!strcmp(name, "phi"))
	{
		*rows = ny + 3;
		*cols = rx;
		*dy = 1;
	}
	else
	{
		*rows = ny + 1;
		*cols = nx;
	}
}

/* write_field_attribute (src/output.c:400-420) */
static void
xdmf_attribute(FILE *f, const sim_t *s, const char *name)
{
	int64_t rows, cols, dy, nx = s->ntpoints[X], ny = s->ntpoints[Y];
	padded_shape(s, name, &rows, &cols, &dy);
	fprintf(f, "      <Attribute Center=\"Node\" Name=\"%s\" DataType=\"Scalar\">\n", name);
	fprintf(f, "        <DataItem ItemType=\"HyperSlab\" Dimensions=\"1 %ld %ld\" Type=\"HyperSlab\">\n",
			(long) ny, (long) nx);
	fprintf(f, "          <DataItem Dimensions=\"3 3\" Format=\"XML\">\n");
	fprintf(f, "            0 %ld %ld\n", (long) dy, 0L);
	fprintf(f, "            1 1 1\n");
	fprintf(f, "            1 %ld %ld\n", (long) (dy + ny), (long) nx);
	fprintf(f, "          </DataItem>\n");
	fprintf(f, "          <DataItem Dimensions=\"1 %ld %ld\" DataType=\"Float\" Precision=\"8\" Format=\"Binary\">\n",
			(long) rows, (long) cols);
	fprintf(f, "            ../bin/%ld/%s.bin\n", (long) s->iter, name);
	fprintf(f, "          </DataItem>\n");
	fprintf(f, "        </DataItem>\n");
	fprintf(f, "      </Attribute>\n");
}

/* write_xdmf_fields (src/output.c:422-459) */
static int
write_xdmf(const sim_t *s, const output_t *out)
{
	char file[PATH_MAX];
	FILE *f;
	if(snprintf(file, sizeof(file), "%s/xdmf/fields-iter%ld.xdmf", out->path,
				(long) s->iter) >= (int) sizeof(file))
		return -1;
	f = fopen(file, "w");
	if(!f)
	{
		perror("fopen");
		return -1;
	}
	fprintf(f, "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n");
	fprintf(f, "<Xdmf xmlns:xi=\"http://www.w3.org/2001/XInclude\" Version=\"3.0\">\n");
	fprintf(f, "  <Domain>\n");
	fprintf(f, "    <Grid Name=\"fields\">\n");
	fprintf(f, "      <Topology TopologyType=\"3DCoRectMesh\" NumberOfElements=\"%ld %ld %ld\"/>\n",
			1L, (long) s->ntpoints[Y], (long) s->ntpoints[X]);
	fprintf(f, "      <Geometry Origin=\"\" Type=\"ORIGIN_DXDYDZ\">\n");
	fprintf(f, "        <DataItem Format=\"XML\" Dimensions=\"3\">\n");
	fprintf(f, "            0.0 0.0 0.0\n");
	fprintf(f, "        </DataItem>\n");
	fprintf(f, "        <DataItem Format=\"XML\" Dimensions=\"3\">\n");
	fprintf(f, "            %f %f %f\n", s->dx[Z], s->dx[Y], s->dx[X]);
	fprintf(f, "        </DataItem>\n");
	fprintf(f, "      </Geometry>\n");
	xdmf_attribute(f, s, "phi");
	xdmf_attribute(f, s, "rho");
	xdmf_attribute(f, s, "E_X");
	xdmf_attribute(f, s, "E_Y");
	fprintf(f, "    </Grid>\n");
	fprintf(f, "  </Domain>\n");
	fprintf(f, "</Xdmf>\n");
	return fclose(f) ? -1 : 0;
}

/* write_field (src/output.c:500-598): the whole padded matrix, aligned */
static int
write_matrix(const sim_t *s, output_t *out, const char *name, int which)
{
	int64_t rows, cols, dy, nx = s->ntpoints[X], ny = s->ntpoints[Y], x, y;
	char file[PATH_MAX];
	int ret;

	padded_shape(s, name, &rows, &cols, &dy);
	size_t bytes = sizeof(double) * (size_t) (rows * cols);
	size_t aligned = (bytes + (size_t) out->alignment - 1) / (size_t) out->alignment *
		(size_t) out->alignment;
	if(out->padded_bytes < aligned)
	{
		free(out->padded);
		out->padded = malloc(aligned);
		out->padded_bytes = out->padded ? aligned : 0;
		if(!out->padded)
			return -1;
	}
	if(!out->dense)
	{
		out->dense = malloc(sizeof(double) * (size_t) (nx * ny));
		if(!out->dense)
			return -1;
	}
	if((ret = cpic_dev_download_field(s->dev, which, out->dense)))
	{
		err("field download failed: %s\n", cpic_dev_strerror(ret));
		return -1;
	}
	double *m = out->padded, *d = out->dense;
	memset(m, 0, aligned);
	for(y = 0; y < rows; y++)
	{
		int64_t src;           /* interior row shown in padded row y, -1 = none */
		if(!strcmp(name, "phi"))
			src = (y - 1 + ny) % ny;
		else if(!strcmp(name, "rho"))
			src = y < ny ? y : -1;
		else
			src = y % ny;
		for(x = 0; x < cols; x++)
			m[y * cols + x] = (src >= 0 && x < nx) ? d[src * nx + x] : NAN;
	}
	if(snprintf(file, sizeof(file), "%s/bin/%ld/%s.bin", out->path, (long) s->iter, name)
			>= (int) sizeof(file))
	{
		err("Path exceeds PATH_MAX\n");
		return -1;
	}
	int fd = open(file, O_WRONLY | O_CREAT | O_TRUNC, S_IRUSR | S_IWUSR);
	if(fd < 0)
	{
		perror("open");
		return -1;
	}
	/* output.slices parts, each a whole number of alignment blocks */
	size_t bsize = (size_t) out->alignment, total_blocks = aligned / bsize;
	size_t per = total_blocks / (size_t) out->max_slices, left = total_blocks % (size_t) out->max_slices;
	size_t off = 0;
	for(int64_t i = 0; i < out->max_slices; i++)
	{
		size_t n = per * bsize;
		if(left > 0)
		{
			n += bsize;
			left--;
		}
		size_t done = 0;
		while(done < n)
		{
			ssize_t w = pwrite(fd, (char *) m + off + done, n - done, (off_t) (off + done));
			if(w <= 0)
			{
				perror("pwrite");
				close(fd);
				return -1;
			}
			done += (size_t) w;
		}
		off += n;
	}
	return close(fd) ? -1 : 0;
}

int
output_fields(sim_t *sim)
{
	output_t *out = sim->output;
	char dir[PATH_MAX];

	if(!out || !out->enabled)
		return 0;
	if(out->period_field <= 0 || sim->iter % out->period_field)
		return 0;
	perf_start(&sim->timers[TIMER_OUTPUT_FIELDS]);
	if(write_xdmf(sim, out))
		return -1;
	if(snprintf(dir, sizeof(dir), "%s/bin/%ld", out->path, (long) sim->iter) >= (int) sizeof(dir))
	{
		err("Dir path exceeds PATH_MAX\n");
		return -1;
	}
	if(make_dir(dir))
		return -1;
	if(write_matrix(sim, out, "rho", CPIC_RHO) || write_matrix(sim, out, "phi", CPIC_PHI) ||
			write_matrix(sim, out, "E_X", CPIC_EX) || write_matrix(sim, out, "E_Y", CPIC_EY))
		return -1;
	perf_stop(&sim->timers[TIMER_OUTPUT_FIELDS]);
	return 0;
}
