/* Text writers for the per-particle output files (.grp, .den, .ray; kd.c:1502-1606)
 * at 10^8 lines: the lines are formatted by several threads into blocks and
 * written in order, so the bytes are those of one fprintf per line.  The .grp
 * reader (kdInGroup, kd.c:920-962) parses on several threads likewise. */
#ifndef SKID_TEXTIO_H
#define SKID_TEXTIO_H

#include <stdio.h>

/* "<n>\n" then one "%d\n" line per value. */
int tio_write_ints(const char *path, long n, const int *v, int nthreads);

/* "<header>\n" then one line per value (n values) in the printf format fmt
 * ("%.10g" or "%g": these two through an exact formatter of the writer's own,
 * any other through snprintf); with nan_zero a NaN value prints as "0"
 * (kdOutVector's particles that did not move). */
int tio_write_floats(const char *path, long header, long n, const float *v, const char *fmt, int nan_zero,
		int nthreads);

/* Reads a file of that form (a count, then whitespace-separated integers) on
 * nthreads threads: 0 when the count equals nexp and the next nexp tokens are
 * ints, stored in v; 1 when the count is not nexp (*header receives it); 2 when
 * the count or one of the nexp values is missing or is not an int token (a sign
 * and digits ended by a space or the end of the file, within int range), in which
 * case the caller's own fscanf loop gives that file's exact behaviour; -1 when the
 * file cannot be read.  Tokens after the nexp-th are not read. */
int tio_read_ints(const char *path, long nexp, int *v, long *header, int nthreads);

/* n bytes from fp's position into buf, leaving fp after them: a regular file
 * of at least 16 MiB is read by nthreads threads (pread at their own offsets),
 * anything else by one fread.  0, or -1 when fewer than n bytes could be read. */
int tio_fread(FILE *fp, void *buf, size_t n, int nthreads);

/* fn(arg, lo, hi) over [0, n) in nthreads contiguous shares, one thread each
 * (one call on the calling thread when n < 65536 or threads cannot start). */
void tio_parallel_for(long n, int nthreads, void (*fn)(void *arg, long lo, long hi), void *arg);

/* A tipsy snapshot's particles in input order (gas, then dark, then star). */
typedef struct {
	double time;
	int nbodies, ndim, nsph, ndark, nstar;
	int n;                                 /* nsph + ndark + nstar */
	float *pos, *vel;                      /* 3 per particle */
	float *mass, *soft, *temp;             /* temp: 0 for dark and star particles */
} TioSnap;

/* The tipsy reader of kd.c:122-222 (kdReadTipsy): a 32-byte header, then gas
 * (12 floats: mass, pos, vel, rho, temp, hsmooth, metals, phi), dark (9: mass, pos,
 * vel, eps, phi) and star (11: mass, pos, vel, metals, tform, eps, phi) records,
 * native byte order or, with xdr, big-endian.  The records are read with
 * tio_fread and decoded on nthreads threads; fp is left after the last record.
 * 0; -1 when the header is bad or the file short; -2 out of memory (the arrays
 * are then freed). */
int tio_read_tipsy(FILE *fp, int xdr, TioSnap *s, int nthreads);
void tio_free_tipsy(TioSnap *s);

/* A 4-byte float / int of a tipsy file (byte-swapped with xdr). */
float tio_get_f32(const unsigned char *p, int xdr);
int tio_get_i32(const unsigned char *p, int xdr);

/* kdOutVector's values (kd.c:1550-1606): d3[j n + i] = moved[3i+j] - pos[3i+j]
 * on axis j, taken once into (-period[j]/2, period[j]/2] (one period subtracted
 * above half, one added at or below minus half), all x, then all y, then all z;
 * a NaN in moved (a particle that did not move, printed as 0) stays NaN.  On
 * nthreads threads. */
void tio_ray_vectors(long n, const float *moved, const float *pos, const float *period, float *d3, int nthreads);

/* Threads to use: SKID_IO_THREADS if set, else the online processors, at most 16. */
int tio_threads(void);

#endif
