/* host_internal.h -- error plumbing shared by the C host files (LogError's role, src/io.c:274-290: a position and a
 * message; the caller decides whether to stop). */
#ifndef ADDA_HOST_INTERNAL_H
#define ADDA_HOST_INTERNAL_H

int host_err(const char *file, int line, const char *fmt, ...);
int host_err_hip(const char *file, int line); /* wraps adda_hip_last_error() */

#endif
