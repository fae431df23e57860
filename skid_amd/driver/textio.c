/* Parallel, byte-identical line formatting for the per-particle text outputs. */
#include "textio.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define TIO_BLOCK (1L << 22)   /* lines per round (split over the threads) */

typedef struct {
	const int *iv;
	const float *fv;
	const char *fmt;
	int nan_zero;
	long lo, hi;               /* lines [lo, hi) */
	char *buf;
	size_t len, cap;
} Job;

/* "%d\n" without printf: digits of |v| as unsigned, so INT_MIN is exact */
static size_t fmt_int(char *out, int v)
{
	char tmp[16];
	unsigned u = v < 0 ? 0u - (unsigned)v : (unsigned)v;
	size_t k = 0, n = 0;
	do { tmp[k++] = (char)('0' + u % 10u); u /= 10u; } while (u);
	if (v < 0) out[n++] = '-';
	while (k) out[n++] = tmp[--k];
	out[n++] = '\n';
	return n;
}

static void *run_job(void *arg)
{
	Job *j = (Job *)arg;
	long i;
	j->len = 0;
	for (i = j->lo; i < j->hi; ++i) {
		if (j->cap - j->len < 48) {                     /* every line here is shorter */
			size_t nc = j->cap ? 2 * j->cap : 1 << 20;
			char *nb = realloc(j->buf, nc);
			if (!nb) return (void *)1;
			j->buf = nb;
			j->cap = nc;
		}
		if (j->iv) {
			j->len += fmt_int(j->buf + j->len, j->iv[i]);
		} else {
			const float f = j->fv[i];
			int w;
			if (j->nan_zero && isnan(f)) w = snprintf(j->buf + j->len, 48, "0\n");
			else {
				w = snprintf(j->buf + j->len, 46, j->fmt, (double)f);
				j->buf[j->len + (size_t)w] = '\n';
				++w;
			}
			j->len += (size_t)w;
		}
	}
	return NULL;
}

static int write_par(const char *path, long header, long n, const int *iv, const float *fv, const char *fmt,
		int nan_zero, int nthreads)
{
	FILE *fp = fopen(path, "w");
	Job *jobs;
	pthread_t *th;
	long b, t;
	int rc = 0;
	if (!fp) { perror(path); return -1; }
	if (nthreads < 1) nthreads = 1;
	jobs = calloc((size_t)nthreads, sizeof(Job));
	th = calloc((size_t)nthreads, sizeof(pthread_t));
	if (!jobs || !th) { fclose(fp); free(jobs); free(th); return -1; }
	fprintf(fp, "%ld\n", header);
	for (b = 0; b < n && !rc; b += TIO_BLOCK) {
		const long e = b + TIO_BLOCK < n ? b + TIO_BLOCK : n;
		const long per = (e - b + nthreads - 1) / nthreads;
		for (t = 0; t < nthreads; ++t) {
			Job *j = &jobs[t];
			j->iv = iv; j->fv = fv; j->fmt = fmt; j->nan_zero = nan_zero;
			j->lo = b + t * per < e ? b + t * per : e;
			j->hi = j->lo + per < e ? j->lo + per : e;
			if (t == 0 || nthreads == 1) continue;
			if (pthread_create(&th[t], NULL, run_job, j)) { run_job(j); th[t] = 0; }
		}
		if (run_job(&jobs[0])) rc = -1;
		for (t = 1; t < nthreads; ++t) {
			void *r = NULL;
			if (th[t]) { pthread_join(th[t], &r); th[t] = 0; }
			if (r) rc = -1;
		}
		for (t = 0; t < nthreads && !rc; ++t)
			if (jobs[t].len && fwrite(jobs[t].buf, 1, jobs[t].len, fp) != jobs[t].len) rc = -1;
	}
	for (t = 0; t < nthreads; ++t) free(jobs[t].buf);
	free(jobs);
	free(th);
	if (fclose(fp)) rc = -1;
	return rc;
}

int tio_write_ints(const char *path, long n, const int *v, int nthreads)
{
	return write_par(path, n, n, v, NULL, NULL, 0, nthreads);
}

int tio_write_floats(const char *path, long header, long n, const float *v, const char *fmt, int nan_zero,
		int nthreads)
{
	return write_par(path, header, n, NULL, v, fmt, nan_zero, nthreads);
}

int tio_threads(void)
{
	const char *e = getenv("SKID_IO_THREADS");
	long p = e ? atol(e) : sysconf(_SC_NPROCESSORS_ONLN);
	if (p < 1) p = 1;
	if (p > 16) p = 16;
	return (int)p;
}
