/*
 * sse_config_internal.h for the double-precision libraries
 * (libsse_dslash_f64.so, libsse_dslash_cpu_f64.so): the configuration the
 * reference's configure step writes for --enable-precision=double
 * (include/sse_config_internal.h.in).  Put this directory on the include path
 * after include/.
 */
#ifndef SSE_CONFIG_INTERNAL_H
#define SSE_CONFIG_INTERNAL_H
#define SSE_PRECISION 64
#define SSE_SCALAR 1
#endif
