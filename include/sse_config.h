/*
 * sse_config.h -- precision selection for the drop-in Wilson Dslash API.
 *
 * As in the reference (include/sse_config.h:13, :35-41), SSEREAL is float when
 * SSE_PRECISION == 32 and double when SSE_PRECISION == 64, and SSE_PRECISION
 * comes from the per-library configuration header sse_config_internal.h.  One
 * ships beside each library build: include/f32/ (libsse_dslash_f32.so,
 * libsse_dslash_cpu_f32.so) and include/f64/ (the *_f64 libraries), so an
 * unchanged Chroma build line needs only the matching -I.  A -DSSE_PRECISION on
 * the command line takes precedence and skips that header.
 */
#ifndef SSE_CONFIG_H
#define SSE_CONFIG_H

#if !defined(SSE_PRECISION)
#include "sse_config_internal.h"
#endif

#if SSE_PRECISION == 32
typedef float SSEREAL;
#elif SSE_PRECISION == 64
typedef double SSEREAL;
#else
#error "Precision Not supported, define SSE_PRECISION to 32 or 64"
#endif

#endif /* SSE_CONFIG_H */
