/*
 * gray_job.h -- the `lux gray` job surface, as a C library (libgray.so).
 *
 * A job is created by gray_mkmod, configured by any number of
 * conf("key=value") calls, then init() and exec() run it, and gray_rmmod
 * frees it: the life cycle of the reference's lux job module
 * (sim-org/gray.c:308-346, LUX_MKMOD / LUX_RMMOD and the conf/init/exec
 * methods of Lux_job).  The object returned by gray_mkmod starts with a
 * gray_job_ops, as the reference's `struct gray` starts with `Lux_job super`
 * (sim-org/gray.h:58-60).
 *
 * Return conventions follow the reference: conf returns 0 when the key was
 * recognised and its value parsed, non-zero otherwise (sim-org/gray.c:88-90);
 * init and exec return EXIT_SUCCESS / EXIT_FAILURE
 * (sim-org/gray.c:135-144, 296-305).
 */
#ifndef GRAY_JOB_H
#define GRAY_JOB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gray_job_ops {
	int (*conf)(void *ego, const char *arg);
	int (*init)(void *ego);
	int (*exec)(void *ego);
} gray_job_ops;

/* LUX_MKMOD (sim-org/gray.c:308-328); opts is unused, as in the reference. */
void *gray_mkmod(const void *opts);
/* LUX_RMMOD (sim-org/gray.c:330-346). */
void  gray_rmmod(void *ego);

/* Convenience wrappers around the ops table. */
int gray_conf(void *ego, const char *arg);
int gray_init(void *ego);
int gray_exec(void *ego);

/* Run a whole job from "key=value" strings: mkmod, conf each, init, exec,
 * rmmod.  Returns EXIT_SUCCESS or EXIT_FAILURE. */
int gray_run(int argc, const char *const *argv);

/* Print every option with its aliases, type, default and current value. */
void gray_print_options(void *ego, void *FILE_stream);

/* Read-only views used by tests and tools. */
size_t gray_n_freq(const void *ego);
size_t gray_n_data(const void *ego);
/* Value of an option as text ("" if unknown); buf of size n. */
int    gray_get_option(const void *ego, const char *name, char *buf, size_t n);

/* Raw dump files (sim-org/io.c:39-72 layout): size_t sz, n_data, w_rays,
 * h_rays, then n_rays*n_data reals (device order: AoS [h][w][n_data] or
 * SoA [n_data][h][w]), then n_rays reals of info. */
int gray_write_raw(const char *path, size_t sz, size_t n_data, size_t w_rays,
                   size_t h_rays, const void *data, const void *info);
/* Reads the header into hdr[4]; data/info (may be NULL) must hold
 * n_rays*n_data*sz and n_rays*sz bytes. */
int gray_read_raw(const char *path, size_t hdr[4], void *data, void *info);

#ifdef __cplusplus
}
#endif
#endif
