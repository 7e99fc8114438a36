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

/* the max-log demap kernel (ldpc_demap.hip) of MODCOD pls over nframes
 * frames on stream st (a hipStream_t) of device dev; LDPC_GPU_* code */
int lg_demap_launch(int dev, int pls, int nframes, const float* d_sym, const float* d_n0, int8_t* d_llr,
                    struct ihipStream_t* st);

#endif
