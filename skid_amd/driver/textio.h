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

/* Threads to use: SKID_IO_THREADS if set, else the online processors, at most 16. */
int tio_threads(void);

#endif
