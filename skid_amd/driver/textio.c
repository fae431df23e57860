/* Parallel, byte-identical line formatting for the per-particle text outputs,
 * and the parallel readers of the .grp and tipsy inputs. */
#include "textio.h"

#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#define TIO_BLOCK (1L << 22)   /* lines per round (split over the threads) */

typedef struct {
	const int *iv;
	const float *fv;
	const char *fmt;
	int prec;                  /* fmt is "%.<prec>g" or "%g" (prec 6): the formatter above; else 0 */
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

/* printf("%.<P>g", (double)f) for a finite, nonzero float f, without printf and
 * byte for byte (the decimal value rounded once to P significant digits, ties
 * to even).  f = M 2^E exactly (M < 2^24); R = M 2^E 10^k with k = P - 1 - X
 * is formed exactly in 128-bit integers, X the decimal exponent (an estimate
 * from log10, corrected until 10^(P-1) <= R < 10^P: a rounding that carries to
 * 10^P lands there on the next exponent, as printf's does).  Returns the
 * characters written, or 0 where R does not fit 128 bits (|f| below about
 * 1e-29): the caller then uses snprintf. */
typedef unsigned __int128 u128;
static size_t fmt_g(char *out, float f, int P)
{
	static const uint64_t p10[20] = {1ull, 10ull, 100ull, 1000ull, 10000ull, 100000ull, 1000000ull, 10000000ull,
		100000000ull, 1000000000ull, 10000000000ull, 100000000000ull, 1000000000000ull, 10000000000000ull,
		100000000000000ull, 1000000000000000ull, 10000000000000000ull, 100000000000000000ull,
		1000000000000000000ull, 10000000000000000000ull};
	uint32_t bits;
	memcpy(&bits, &f, 4);
	const int neg = (int)(bits >> 31);
	const uint32_t ex = (bits >> 23) & 0xFFu, mt = bits & 0x7FFFFFu;
	const uint64_t M = ex ? (mt | 0x800000u) : mt;
	const int E = ex ? (int)ex - 150 : -149;
	int X = (int)floor(log10(ldexp((double)M, E)));
	uint64_t R = 0;
	for (int tries = 0;; ++tries) {
		if (tries > 4) return 0;
		const int k = P - 1 - X;
		u128 A, D = 1;
		if (k >= 0) {
			if (k > 38) return 0;                        /* 5^38 M < 2^113 */
			A = M;
			for (int i = 0; i < k; ++i) A *= 5u;
			const int s = E + k;                         /* R = A 2^s */
			if (s >= 0) {
				if (s > 60 || (A >> (64 - s)) != 0) { ++X; continue; }   /* R far above 10^P */
				R = (uint64_t)(A << s);
			} else if (-s >= 120) {
				R = 0;
			} else {
				const int sh = -s;
				const u128 q = A >> sh, rem = A - (q << sh), half = (u128)1 << (sh - 1);
				R = (uint64_t)q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u);
			}
		} else {
			const int kk = -k;                           /* R = M 2^(E - kk) / 5^kk */
			if (kk > 38) return 0;
			for (int i = 0; i < kk; ++i) D *= 5u;
			const int s = E - kk;
			if (s >= 0) { A = (u128)M << s; }
			else { A = M; if (-s > 30) return 0; D <<= -s; }            /* 5^38 2^30 < 2^120 */
			const u128 q = A / D, rem = A - q * D;
			R = (uint64_t)q + ((2 * rem > D || (2 * rem == D && (q & 1u))) ? 1u : 0u);
		}
		if (R >= p10[P]) { ++X; continue; }
		if (R < p10[P - 1]) { --X; continue; }
		break;
	}
	char d[20];
	for (int i = P - 1; i >= 0; --i) { d[i] = (char)('0' + R % 10u); R /= 10u; }
	int nd = P;                                          /* significant digits left after trailing zeros */
	while (nd > 1 && d[nd - 1] == '0') --nd;
	size_t n = 0;
	if (neg) out[n++] = '-';
	if (X < -4 || X >= P) {                              /* %e style */
		out[n++] = d[0];
		if (nd > 1) { out[n++] = '.'; for (int i = 1; i < nd; ++i) out[n++] = d[i]; }
		out[n++] = 'e';
		out[n++] = X < 0 ? '-' : '+';
		const int ax = X < 0 ? -X : X;
		if (ax >= 100) out[n++] = (char)('0' + ax / 100);
		out[n++] = (char)('0' + ax / 10 % 10);
		out[n++] = (char)('0' + ax % 10);
	} else if (X >= 0) {                                 /* %f style, X + 1 integer digits */
		for (int i = 0; i <= X; ++i) out[n++] = d[i];
		if (nd > X + 1) { out[n++] = '.'; for (int i = X + 1; i < nd; ++i) out[n++] = d[i]; }
	} else {                                             /* 0.000ddd */
		out[n++] = '0';
		out[n++] = '.';
		for (int i = 0; i < -X - 1; ++i) out[n++] = '0';
		for (int i = 0; i < nd; ++i) out[n++] = d[i];
	}
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
			size_t w = 0;
			if (j->nan_zero && isnan(f)) {
				memcpy(j->buf + j->len, "0\n", 2);
				w = 2;
			} else {
				if (j->prec > 0 && isfinite(f) && f != 0.f) w = fmt_g(j->buf + j->len, f, j->prec);
				if (!w) w = (size_t)snprintf(j->buf + j->len, 46, j->fmt, (double)f);
				j->buf[j->len + w] = '\n';
				++w;
			}
			j->len += w;
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
			j->prec = !fmt ? 0 : !strcmp(fmt, "%.10g") ? 10 : !strcmp(fmt, "%g") ? 6 : 0;
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

/* ------------------------------------------------------------------ reader */
/* kdInGroup (kd.c:920-962) reads a .grp with one fscanf("%d") per value; here the
 * file is read whole and split over threads at whitespace: each thread counts the
 * tokens that start in its share, a prefix sum gives each share its first value's
 * index, then each thread converts its tokens.  A token is a run of non-space bytes
 * (the C locale's isspace set); it must be an optional sign and decimal digits of
 * a value that fits an int, ended by a space or the end of the file. */

static int is_ws(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

/* The token at p (p < end, *p not a space): 0 and *out when it is an int. */
static int tok_int(const char *p, const char *end, int *out)
{
	long long v = 0;
	int neg = 0;
	if (*p == '+' || *p == '-') neg = *p++ == '-';
	if (p >= end || *p < '0' || *p > '9') return -1;
	for (; p < end && *p >= '0' && *p <= '9'; ++p)
		if ((v = 10 * v + (*p - '0')) > 2147483648LL) return -1;
	if (p < end && !is_ws(*p)) return -1;
	if (!neg && v > 2147483647LL) return -1;
	*out = (int)(neg ? -v : v);
	return 0;
}

typedef struct {
	const char *buf, *end;
	size_t base, lo, hi;       /* tokens that start in [lo, hi); base = first byte after the header */
	long first, count, nexp;
	int *v;
	int bad;
} RJob;

static int tok_start(const RJob *j, size_t i) { return !is_ws(j->buf[i]) && (i == j->base || is_ws(j->buf[i - 1])); }

static void *count_job(void *arg)
{
	RJob *j = (RJob *)arg;
	long c = 0;
	for (size_t i = j->lo; i < j->hi; ++i) c += tok_start(j, i);
	j->count = c;
	return NULL;
}

static void *parse_job(void *arg)
{
	RJob *j = (RJob *)arg;
	long k = j->first;
	for (size_t i = j->lo; i < j->hi && k < j->nexp; ++i) {
		if (!tok_start(j, i)) continue;
		if (tok_int(j->buf + i, j->end, &j->v[k++])) { j->bad = 1; break; }
	}
	return NULL;
}

/* The whole file (a regular file or a pipe). */
static char *slurp(const char *path, size_t *len)
{
	FILE *fp = fopen(path, "rb");
	size_t cap = 1 << 20, n = 0, r;
	char *b;
	if (!fp) return NULL;
	if (!fseek(fp, 0, SEEK_END)) {
		long sz = ftell(fp);
		if (sz > 0) cap = (size_t)sz + 1;
		rewind(fp);
	}
	if (!(b = malloc(cap))) { fclose(fp); return NULL; }
	while ((r = fread(b + n, 1, cap - n, fp)) > 0) {
		n += r;
		if (n == cap) {
			char *nb = realloc(b, 2 * cap);
			if (!nb) { free(b); fclose(fp); return NULL; }
			b = nb;
			cap *= 2;
		}
	}
	if (ferror(fp)) { free(b); fclose(fp); return NULL; }
	fclose(fp);
	*len = n;
	return b;
}

static void run_all(RJob *jobs, int nthreads, void *(*fn)(void *))
{
	pthread_t *th = calloc((size_t)nthreads, sizeof(pthread_t));
	for (int t = 1; t < nthreads; ++t)
		if (!th || pthread_create(&th[t], NULL, fn, &jobs[t])) { fn(&jobs[t]); if (th) th[t] = 0; }
	fn(&jobs[0]);
	for (int t = 1; th && t < nthreads; ++t)
		if (th[t]) pthread_join(th[t], NULL);
	free(th);
}

int tio_read_ints(const char *path, long nexp, int *v, long *header, int nthreads)
{
	size_t len = 0, i = 0, base;
	char *b = slurp(path, &len);
	RJob *jobs;
	long total = 0;
	int h, rc = 0;
	if (!b) return -1;
	while (i < len && is_ws(b[i])) ++i;
	if (i == len || tok_int(b + i, b + len, &h)) { free(b); return 2; }
	*header = h;
	if (h != nexp) { free(b); return 1; }
	while (i < len && !is_ws(b[i])) ++i;
	base = i;
	if (nthreads < 1) nthreads = 1;
	if ((size_t)nthreads > len - base + 1) nthreads = (int)(len - base + 1);
	if (!(jobs = calloc((size_t)nthreads, sizeof(RJob)))) { free(b); return -1; }
	for (int t = 0; t < nthreads; ++t) {
		RJob *j = &jobs[t];
		j->buf = b; j->end = b + len; j->base = base; j->nexp = nexp; j->v = v;
		j->lo = base + (len - base) * (size_t)t / (size_t)nthreads;
		j->hi = base + (len - base) * (size_t)(t + 1) / (size_t)nthreads;
	}
	run_all(jobs, nthreads, count_job);
	for (int t = 0; t < nthreads; ++t) { jobs[t].first = total; total += jobs[t].count; }
	if (total < nexp) rc = 2;
	else {
		run_all(jobs, nthreads, parse_job);
		for (int t = 0; t < nthreads; ++t) rc |= jobs[t].bad ? 2 : 0;
	}
	free(jobs);
	free(b);
	return rc;
}

/* ------------------------------------------------------------ binary input */
typedef struct {
	long lo, hi;
	int fd;
	off_t pos;
	char *buf;
	void (*fn)(void *, long, long);
	void *arg;
	int bad;
} PJob;

static void *pread_job(void *arg)
{
	PJob *j = (PJob *)arg;
	long k = j->lo;
	while (k < j->hi) {
		const ssize_t r = pread(j->fd, j->buf + k, (size_t)(j->hi - k), j->pos + (off_t)k);
		if (r < 0 && errno == EINTR) continue;
		if (r <= 0) { j->bad = 1; break; }
		k += r;
	}
	return NULL;
}

static void *for_job(void *arg)
{
	PJob *j = (PJob *)arg;
	if (j->lo < j->hi) j->fn(j->arg, j->lo, j->hi);
	return NULL;
}

static int run_shares(long n, int nthreads, void *(*body)(void *), PJob *proto)
{
	PJob *jobs;
	pthread_t *th;
	int bad = 0;
	if (nthreads < 1) nthreads = 1;
	jobs = calloc((size_t)nthreads, sizeof(PJob));
	th = calloc((size_t)nthreads, sizeof(pthread_t));
	if (!jobs || !th) { free(jobs); free(th); return -1; }
	for (int t = 0; t < nthreads; ++t) {
		jobs[t] = *proto;
		jobs[t].lo = (long)((__int128)n * t / nthreads);
		jobs[t].hi = (long)((__int128)n * (t + 1) / nthreads);
		if (t && pthread_create(&th[t], NULL, body, &jobs[t])) { body(&jobs[t]); th[t] = 0; }
	}
	body(&jobs[0]);
	for (int t = 1; t < nthreads; ++t)
		if (th[t]) pthread_join(th[t], NULL);
	for (int t = 0; t < nthreads; ++t) bad |= jobs[t].bad;
	free(jobs);
	free(th);
	return bad ? -1 : 0;
}

int tio_fread(FILE *fp, void *buf, size_t n, int nthreads)
{
	struct stat st;
	off_t pos;
	PJob p;
	const int fd = fileno(fp);
	if (n < ((size_t)1 << 24) || nthreads < 2 || fd < 0 || fstat(fd, &st) || !S_ISREG(st.st_mode)
			|| (pos = ftello(fp)) < 0 || pos + (off_t)n > st.st_size)
		return fread(buf, 1, n, fp) == n ? 0 : -1;
	memset(&p, 0, sizeof(p));
	p.fd = fd;
	p.pos = pos;
	p.buf = (char *)buf;
	if (run_shares((long)n, nthreads, pread_job, &p)) return -1;
	return fseeko(fp, pos + (off_t)n, SEEK_SET) ? -1 : 0;
}

void tio_parallel_for(long n, int nthreads, void (*fn)(void *arg, long lo, long hi), void *arg)
{
	PJob p;
	memset(&p, 0, sizeof(p));
	p.fn = fn;
	p.arg = arg;
	if (n < 65536 || run_shares(n, nthreads, for_job, &p)) fn(arg, 0, n);
}

/* ------------------------------------------------------------ tipsy input */
static inline float get_f32(const unsigned char *p, int xdr)
{
	uint32_t u;
	float f;
	memcpy(&u, p, 4);
	if (xdr) u = __builtin_bswap32(u);
	memcpy(&f, &u, 4);
	return f;
}

static inline int get_i32(const unsigned char *p, int xdr)
{
	uint32_t u;
	memcpy(&u, p, 4);
	if (xdr) u = __builtin_bswap32(u);
	return (int)u;
}

float tio_get_f32(const unsigned char *p, int xdr) { return get_f32(p, xdr); }
int tio_get_i32(const unsigned char *p, int xdr) { return get_i32(p, xdr); }

/* One record type's decode: the records in buf go to particles i0 + [lo, hi). */
typedef struct {
	TioSnap *s;
	const unsigned char *buf;
	int nrec, xdr;
	long i0;
} Decode;

static const int tipsy_nf[3] = {12, 9, 11};   /* floats per gas, dark, star record */

static void decode_range(void *arg, long lo, long hi)
{
	const Decode *d = (const Decode *)arg;
	TioSnap *s = d->s;
	const int x = d->xdr;
	for (long j = lo; j < hi; ++j) {
		const unsigned char *r = d->buf + 4 * (size_t)tipsy_nf[d->nrec] * j;
		const long i = d->i0 + j;
		s->mass[i] = get_f32(r, x);
		for (int k = 0; k < 3; ++k) {
			s->pos[3*i+k] = get_f32(r + 4 + 4*k, x);
			s->vel[3*i+k] = get_f32(r + 16 + 4*k, x);
		}
		if (d->nrec == 0) { s->temp[i] = get_f32(r + 32, x); s->soft[i] = get_f32(r + 36, x); }
		else if (d->nrec == 1) { s->temp[i] = 0.0f; s->soft[i] = get_f32(r + 28, x); }
		else { s->temp[i] = 0.0f; s->soft[i] = get_f32(r + 36, x); }
	}
}

void tio_free_tipsy(TioSnap *s)
{
	free(s->pos); free(s->vel); free(s->mass); free(s->soft); free(s->temp);
	s->pos = s->vel = s->mass = s->soft = s->temp = NULL;
}

int tio_read_tipsy(FILE *fp, int xdr, TioSnap *s, int nthreads)
{
	unsigned char h[32];
	long i = 0;
	memset(s, 0, sizeof(*s));
	if (fread(h, 1, 32, fp) != 32) return -1;
	if (xdr) {
		uint64_t u;
		memcpy(&u, h, 8);
		u = ((uint64_t)__builtin_bswap32((uint32_t)u) << 32) | __builtin_bswap32((uint32_t)(u >> 32));
		memcpy(&s->time, &u, 8);
	} else memcpy(&s->time, h, 8);
	s->nbodies = get_i32(h + 8, xdr); s->ndim = get_i32(h + 12, xdr);
	s->nsph = get_i32(h + 16, xdr); s->ndark = get_i32(h + 20, xdr); s->nstar = get_i32(h + 24, xdr);
	if (s->nsph < 0 || s->ndark < 0 || s->nstar < 0 || (long)s->nsph + s->ndark + s->nstar > 2147483647L)
		return -1;
	s->n = s->nsph + s->ndark + s->nstar;
	s->pos = malloc(12 * (size_t)s->n + 1); s->vel = malloc(12 * (size_t)s->n + 1);
	s->mass = malloc(4 * (size_t)s->n + 1); s->soft = malloc(4 * (size_t)s->n + 1); s->temp = malloc(4 * (size_t)s->n + 1);
	if (!s->pos || !s->vel || !s->mass || !s->soft || !s->temp) { tio_free_tipsy(s); return -2; }
	for (int nrec = 0; nrec < 3; ++nrec) {
		const long cnt = nrec == 0 ? s->nsph : (nrec == 1 ? s->ndark : s->nstar);
		const size_t bytes = 4 * (size_t)tipsy_nf[nrec] * (size_t)cnt;
		unsigned char *buf = malloc(bytes ? bytes : 1);
		Decode d;
		if (!buf) { tio_free_tipsy(s); return -2; }
		if (tio_fread(fp, buf, bytes, nthreads)) { free(buf); tio_free_tipsy(s); return -1; }
		d.s = s; d.buf = buf; d.nrec = nrec; d.xdr = xdr; d.i0 = i;
		tio_parallel_for(cnt, nthreads, decode_range, &d);
		i += cnt;
		free(buf);
	}
	return 0;
}

/* ----------------------------------------------------------- -ray vectors */
typedef struct {
	const float *moved, *pos;
	float h[3];
	long n;
	float *d3;
} Ray;

static void ray_range(void *arg, long lo, long hi)
{
	const Ray *r = (const Ray *)arg;
	for (int j = 0; j < 3; ++j) {
		const float h = r->h[j];
		float *out = r->d3 + (size_t)j * r->n;
		for (long i = lo; i < hi; ++i) {
			float d = r->moved[3*i+j];
			if (!isnan(d)) {
				d = d - r->pos[3*i+j];
				if (d > h) d -= 2*h;
				if (d <= -h) d += 2*h;
			}
			out[i] = d;
		}
	}
}

void tio_ray_vectors(long n, const float *moved, const float *pos, const float *period, float *d3, int nthreads)
{
	Ray r;
	r.moved = moved; r.pos = pos; r.n = n; r.d3 = d3;
	for (int j = 0; j < 3; ++j) r.h[j] = 0.5f * period[j];
	tio_parallel_for(n, nthreads, ray_range, &r);
}
