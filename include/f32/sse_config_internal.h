/*
 * sse_config_internal.h for the single-precision libraries
 * (libsse_dslash_f32.so, libsse_dslash_cpu_f32.so): the configuration the
 * reference's configure step writes for --enable-precision=single
 * (include/sse_config_internal.h.in).  Put this directory on the include path
 * after include/.
 */
#ifndef SSE_CONFIG_INTERNAL_H
#define SSE_CONFIG_INTERNAL_H
#define SSE_PRECISION 32
#define SSE_SCALAR 1
#endif
