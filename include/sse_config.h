/*
 * sse_config.h -- precision selection for the drop-in Wilson Dslash API.
 *
 * As in the reference (include/sse_config.h:35-41), SSEREAL is float when
 * SSE_PRECISION == 32 and double when SSE_PRECISION == 64.  Precision is fixed
 * per library: link libsse_dslash_f32.so with -DSSE_PRECISION=32 and
 * libsse_dslash_f64.so with -DSSE_PRECISION=64.
 */
#ifndef SSE_CONFIG_H
#define SSE_CONFIG_H

#if !defined(SSE_PRECISION)
#error "define SSE_PRECISION to 32 or 64 (it selects the library: libsse_dslash_f32.so / libsse_dslash_f64.so)"
#endif

#if SSE_PRECISION == 32
typedef float SSEREAL;
#elif SSE_PRECISION == 64
typedef double SSEREAL;
#else
#error "Precision Not supported, define SSE_PRECISION to 32 or 64"
#endif

#endif /* SSE_CONFIG_H */
