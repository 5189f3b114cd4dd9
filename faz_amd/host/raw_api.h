// Internal entry points shared by the GPU orchestration (csrc/faz_gpu.hip)
// and the stream layer (host/stream.cpp).
#pragma once
#include <stdint.h>
#include <vector>

#include "faz_gpu.h"

extern "C++" {
// Whole engine buffer (header, quantizer, Huffman tree and bits) of one field
// into `raw`, sized exactly.
int faz_compress_raw_vec(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_interp_params *p,
                         std::vector<uint8_t> &raw);
// The same for the 3-D Lorenzo/regression engine (SZGeneralCompressor over
// SZFastFrontend).
int faz_lorenzo_raw_vec(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_lorenzo_params *p,
                        std::vector<uint8_t> &raw);
// The same for the general-frontend engine (any N).
int faz_general_raw_vec(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_general_params *p,
                        std::vector<uint8_t> &raw);
int faz_general_decompress_raw_vec(faz_gpu_ctx *ctx, const uint8_t *raw, size_t n, int N, int lorenzo, int lorenzo2,
                                   int regression,
                                   float *d_out, size_t num);
// Device scratch of the context, grown on demand (n floats).
float *faz_ctx_work_in(faz_gpu_ctx *ctx, size_t n);
hipStream_t faz_ctx_stream(faz_gpu_ctx *ctx);
void faz_ctx_set_error(faz_gpu_ctx *ctx, const char *msg);
}
