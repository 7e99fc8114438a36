/*
 * Context access for the library's translation units other than
 * ldpc_gpu.hip (not part of the C-ABI).
 */
#ifndef LDPC_INTERNAL_H
#define LDPC_INTERNAL_H

#include "ldpc_gpu.h"

/* device the context runs on */
int lg_ctx_device(ldpc_gpu_ctx* c);
/* record msg as the context's last error; returns code */
int lg_ctx_fail(ldpc_gpu_ctx* c, int code, const char* msg);

#endif
