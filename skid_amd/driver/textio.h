/* Text writers for the per-particle output files (.grp, .den, .ray; kd.c:1502-1606)
 * at 10^8 lines: the lines are formatted by several threads into blocks and
 * written in order, so the bytes are those of one fprintf per line. */
#ifndef SKID_TEXTIO_H
#define SKID_TEXTIO_H

/* "<n>\n" then one "%d\n" line per value. */
int tio_write_ints(const char *path, long n, const int *v, int nthreads);

/* "<header>\n" then one line per value (n values) in the printf format fmt
 * ("%.10g" or "%g": these two through an exact formatter of the writer's own,
 * any other through snprintf); with nan_zero a NaN value prints as "0"
 * (kdOutVector's particles that did not move). */
int tio_write_floats(const char *path, long header, long n, const float *v, const char *fmt, int nan_zero,
		int nthreads);

/* Threads to use: SKID_IO_THREADS if set, else the online processors, at most 16. */
int tio_threads(void);

#endif
