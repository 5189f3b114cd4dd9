/* faz_gpu.h — C-ABI of the MI355X (gfx950) FAZ compression hot path.
 *
 * The reference (JLiu-1/FAZ, namespace QoZ) has no FFI: its hot path is the
 * C++ template engine SZInterpolationCompressor<T,N,LinearQuantizer,
 * HuffmanEncoder,Lossless_zstd>.  Each entry point below replaces one piece of
 * that engine and names the reference code it stands in for.  The C++ host
 * layer (include/faz/sz.hpp: QoZ::Config, SZ_compress, SZ_decompress) and the
 * `faz` command line sit on top of these calls.
 *
 * Conventions
 *   - every call returns faz_status (0 = ok, < 0 = error; never exits);
 *     faz_gpu_last_error() gives a message for the last failing call;
 *   - d_ pointers are device (HBM) pointers, h_ pointers host pointers;
 *   - a context owns its device scratch and stream; output device pointers
 *     returned through a context stay valid until the next call on it;
 *   - one context per host thread; calls are stream-ordered and synchronous
 *     at return.
 */
#ifndef FAZ_GPU_H
#define FAZ_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int faz_status;
#define FAZ_OK 0
#define FAZ_ERR_INVALID (-1)     /* bad argument */
#define FAZ_ERR_HIP (-2)         /* HIP runtime error or no device */
#define FAZ_ERR_UNSUPPORTED (-3) /* engine mode outside the GPU path */
#define FAZ_ERR_CAPACITY (-4)    /* caller buffer too small (size written back) */
#define FAZ_ERR_FORMAT (-5)      /* malformed or truncated stream */

typedef struct faz_gpu_ctx faz_gpu_ctx;

/* Interpolation engine parameters: the QoZ::Config fields read by
 * SZInterpolationCompressor::compress (include/QoZ/compressor/
 * SZInterpolationCompressor.hpp:333-345) plus its tuning arguments. */
typedef struct {
    int N;                  /* 1..4 */
    size_t dims[4];         /* slowest first, last dimension contiguous */
    double eb;              /* absolute error bound */
    int interp_algo;        /* 0 linear, 1 cubic (Config::interpAlgo) */
    int direction;          /* Config::interpDirection: permutation index */
    double alpha, beta;     /* level-wise error-bound scaling */
    size_t max_step;        /* anchor stride (Config::maxStep), 0 = none */
    int block_size;         /* Config::interpBlockSize (even) */
    int radius;             /* Config::quantbinCnt / 2, <= 32768 */
    int levelwise;          /* entries used in ops/dirs */
    uint8_t ops[32];        /* Config::interpAlgo_list */
    uint8_t dirs[32];       /* Config::interpDirection_list */
    int tuning;             /* 0 compress, 1 codes + squared error, 2 |prediction error| */
    int start_level;        /* 0 = top level */
    int end_level;          /* 0 = down to level 1 */
} faz_interp_params;

/* Results of faz_gpu_interp_compress_f32 (device pointers owned by ctx). */
typedef struct {
    uint16_t *d_codes;      /* quantization codes in the reference's push order */
    size_t n_codes;
    float *d_unpred;        /* unpredictable values, same order */
    size_t n_unpred;
    double loss;            /* tuning 1: sum of squared errors; 2: sum |pred error| */
    size_t level_counts[64];/* tuning modes: codes pushed after each level */
    int levels;             /* interpolation levels (SIC:820-852) */
} faz_interp_out;

/* Huffman code table built on the host from the GPU histogram. */
typedef struct {
    int32_t offset;         /* smallest symbol (HuffmanEncoder::offset) */
    int32_t n_states;       /* max - offset + 1 entries in code/len */
    const uint64_t *code;   /* MSB-aligned codewords, indexed by symbol - offset */
    const uint8_t *len;     /* codeword lengths (0..64) */
} faz_huff_table;

/* --- context ------------------------------------------------------------ */
faz_status faz_gpu_ctx_create(int device, faz_gpu_ctx **out);
void faz_gpu_ctx_destroy(faz_gpu_ctx *ctx);
const char *faz_gpu_last_error(const faz_gpu_ctx *ctx);
/* Device id the context runs on, or -1. */
int faz_gpu_ctx_device(const faz_gpu_ctx *ctx);

/* --- A1: QoZ::data_range (include/QoZ/utils/Statistic.hpp:11-21) -------- */
faz_status faz_gpu_range_f32(faz_gpu_ctx *ctx, const float *d_in, size_t n, float *h_min, float *h_max);

/* --- A2..A10: SZInterpolationCompressor::compress level/block/pass loop,
 * quantize / quantize_tuning and LinearQuantizer::quantize_and_overwrite
 * (SIC:328-476, 924-974, 1306-1417, 3076-3442; IntegerQuantizer.hpp:69-108).
 * d_in is read only; d_recon (n floats, may equal d_in) receives the values
 * the reference leaves in its data buffer. */
faz_status faz_gpu_interp_compress_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon,
                                       const faz_interp_params *p, faz_interp_out *out);

/* --- A12 input: the frequency loop of HuffmanEncoder::init
 * (include/QoZ/encoder/HuffmanEncoder.hpp:529-538) plus the first sequence
 * index of every symbol, which fixes ska::unordered_map's iteration order.
 * h_freq and h_first have 65536 entries; h_first is UINT32_MAX for absent. */
faz_status faz_gpu_huffman_hist(faz_gpu_ctx *ctx, const uint16_t *d_codes, size_t n, uint32_t *h_freq,
                                uint32_t *h_first);

/* --- A14: HuffmanEncoder::encode (HuffmanEncoder.hpp:150-227).  Writes the
 * MSB-first bitstream (without the size_t prefix) to d_bits (capacity
 * cap_bytes) and its byte count to *nbytes. */
faz_status faz_gpu_huffman_encode(faz_gpu_ctx *ctx, const uint16_t *d_codes, size_t n, const faz_huff_table *t,
                                  uint8_t *d_bits, size_t cap_bytes, size_t *nbytes);

/* --- faz_gpu_huffman_encode without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, output and errors as faz_gpu_huffman_encode;
 * its chunk totals, offsets, table and bit total live in one static device
 * array per GPU (faz_gpu_noalloc_scratch_bytes(), about 5 MB, held until the
 * process ends) instead of context buffers, and it calls no hipMalloc /
 * hipFree.  At most 2^30 codes and 65537 table states (FAZ_ERR_INVALID
 * beyond).  Every call on one GPU shares that GPU's array: calls that overlap
 * on the same GPU (from any context or thread) race. */
faz_status faz_gpu_huffman_encode_noalloc(faz_gpu_ctx *ctx, const uint16_t *d_codes, size_t n,
                                          const faz_huff_table *t, uint8_t *d_bits, size_t cap_bytes,
                                          size_t *nbytes);
size_t faz_gpu_noalloc_scratch_bytes(void);

/* --- A12/A13 host side: HuffmanEncoder::init + save (HuffmanEncoder.hpp:
 * 103-137, 455-633).  From per-symbol counts and first indices (65536 entries
 * each) builds the tree exactly as the reference does and writes the code
 * table (code/len, 65536 entries each, indexed by symbol - *offset) and the
 * serialised tree bytes (capacity tree_cap, size to *tree_size). */
faz_status faz_huffman_build(const uint32_t *h_freq, const uint32_t *h_first, int32_t *offset, int32_t *n_states,
                             uint64_t *code, uint8_t *len, uint8_t *tree, size_t tree_cap, size_t *tree_size);

/* --- A15: HuffmanEncoder::load + decode (include/QoZ/encoder/
 * HuffmanEncoder.hpp:236-296) on the device.  h_buf holds a saved tree
 * (offset, node count, tree arrays) followed by the size_t byte count and the
 * MSB-first bitstream (n bytes in all).  Writes num symbols (each symbol +
 * the tree's offset, which must lie in [0, limit) as an unsigned value unless
 * limit is 0) to d_out and the bytes of the section to *used.  A stream that
 * runs out before num symbols, bits that leave the tree or a symbol out of
 * range fail with FAZ_ERR_FORMAT (the reference reads past its buffer or
 * prints "garbage input"). */
faz_status faz_gpu_huffman_decode(faz_gpu_ctx *ctx, const uint8_t *h_buf, size_t n, size_t num, uint32_t limit,
                                  int32_t *d_out, size_t *used);

/* Host-only validation of a saved Huffman section (HuffmanEncoder::load,
 * HuffmanEncoder.hpp:236-262), the check faz_gpu_huffman_decode makes before
 * any device work: the node arrays parse, every slot reachable from the root
 * is reached once (no shared children or cycles), no path is deeper than 64
 * bits and the bitstream fits in n bytes; FAZ_ERR_FORMAT otherwise.  *nodes
 * and *section_bytes (may be NULL) get the node count and the section's
 * size. */
faz_status faz_huffman_check_section(const uint8_t *buf, size_t n, size_t *nodes, size_t *section_bytes);

/* Host-only read of a stream's Config trailer (Config::load, Config.hpp:
 * 311-388) with every field bounded by the stream: N (1..4), dims[N] and
 * cmprAlgo (QoZ::ALGO); FAZ_ERR_FORMAT for a truncated or corrupt trailer
 * (the reference reads past its buffer). */
faz_status faz_stream_info(const uint8_t *stream, size_t n, int *N, size_t *dims, int *cmpr_algo);

/* --- A17: HuffmanEncoder::decode + SZInterpolationCompressor::decompress
 * (HuffmanEncoder.hpp:236-296; SIC:42-199) on a pre-zstd engine buffer. */
faz_status faz_gpu_interp_decompress_raw(faz_gpu_ctx *ctx, const uint8_t *h_raw, size_t n, int N, float *d_out,
                                         size_t num);
/* The same decompression of an engine buffer already in device memory
 * (d_raw, n bytes, any alignment): the header and the Huffman tree are read
 * back to the host, the unpredictable list and the bitstream stay on the
 * device.  Same SIC:42-199 / HUF:236-296 replacement for callers that keep
 * streams resident in HBM. */
faz_status faz_gpu_interp_decompress_dev(faz_gpu_ctx *ctx, const uint8_t *d_raw, size_t n, int N, float *d_out,
                                         size_t num);

/* --- A11..A16 on one GPU: the whole engine buffer that
 * SZInterpolationCompressor::compress hands to zstd (SIC:485-518): header,
 * LinearQuantizer::save (IntegerQuantizer.hpp:185-201), Huffman tree and
 * bits.  Written to h_raw (capacity cap), size to *raw_size. */
faz_status faz_gpu_interp_raw_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_interp_params *p,
                                  uint8_t *h_raw, size_t cap, size_t *raw_size);

/* --- The same engine buffer in two steps, so the device work can be timed
 * on its own.  d_recon may be NULL: SZ_compress returns no reconstruction,
 * and without one the finest level writes none (the coarse levels still keep
 * theirs in context scratch); with one it receives every reconstructed
 * point.  faz_gpu_compress_f32 runs the value range (rel_mode != 0:
 * eb is relative to max-min, Statistic.hpp:34-41), interpolation +
 * quantization, histogram, outlier compaction, the host tree and the encode
 * (SIC:328-518) and leaves the result in the context; *raw_size gets the
 * engine buffer size.  faz_gpu_fetch_raw copies it to the host. */
faz_status faz_gpu_compress_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_interp_params *p,
                                int rel_mode, double eb, size_t *raw_size);
faz_status faz_gpu_fetch_raw(faz_gpu_ctx *ctx, uint8_t *h_raw, size_t cap);

/* --- SZ_compress / SZ_decompress for ALGO_INTERP streams
 * (include/QoZ/api/sz.hpp:54-205, SZInterp.hpp:387-431,
 * Lossless_zstd.hpp:22-47, Config.hpp:240-388).  `rel_mode` != 0 applies
 * calAbsErrorBound's EB_REL conversion (Statistic.hpp:34-41) to eb.
 * data is a host pointer unless data_on_device.  *out is malloc'ed; free it
 * with faz_free. */
faz_status faz_sz_compress_f32(faz_gpu_ctx *ctx, const faz_interp_params *p, int rel_mode, double eb,
                               const float *data, int data_on_device, uint8_t **out, size_t *out_size);
faz_status faz_sz_decompress_f32(faz_gpu_ctx *ctx, const uint8_t *stream, size_t n, float *out, size_t num,
                                 int out_on_device);
void faz_free(void *p);

/* --- A19: the 3-D Lorenzo/regression engine, SZFastFrontend
 * (include/QoZ/frontend/SZFastFrontend.hpp:176-345, 429-530) as
 * SZ_compress_LorenzoReg composes it for 3-D fields (SZLorenzoReg.hpp:95-114:
 * SZGeneralCompressor, LinearQuantizer, HuffmanEncoder<int>).  The
 * parameters are the QoZ::Config fields the frontend reads (FF:25-31). */
typedef struct {
    size_t dims[3];         /* slowest first, last dimension contiguous */
    double eb;              /* absolute error bound */
    int block_size;         /* Config::blockSize, 1..16 */
    int pred_dim;           /* Config::pred_dim: 3, 2 or 1 */
    int lorenzo;            /* Config::lorenzo */
    int lorenzo2;           /* Config::lorenzo2 */
    int regression;         /* Config::regression */
    int radius;             /* Config::quantbinCnt / 2, <= 32768 */
} faz_lorenzo_params;

/* Results of faz_gpu_lorenzo_compress_f32 (device pointers owned by ctx). */
typedef struct {
    uint16_t *d_codes;      /* quantization codes, block-major (FF:191, 212) */
    size_t n_codes;
    float *d_unpred;        /* unpredictable values in code order */
    size_t n_unpred;
    const uint8_t *d_selector; /* per block: 0 Lorenzo, 1 regression, 2 2-layer Lorenzo */
    size_t n_blocks;
    size_t n_regression;    /* blocks predicted by regression */
} faz_lorenzo_out;

/* SZFastFrontend::compress_3d (FF:176-345).  d_recon (n floats) receives
 * the reconstruction the reference keeps in its prediction buffer. */
faz_status faz_gpu_lorenzo_compress_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon,
                                        const faz_lorenzo_params *p, faz_lorenzo_out *out);
/* The buffer SZGeneralCompressor::compress hands to zstd (SGC:30-71): the
 * frontend's save (FF:50-84: meta_params, bound, selector and coefficient
 * Huffman sections, LinearQuantizer::save) then the codes' tree and bits.
 * d_recon may be NULL. */
faz_status faz_gpu_lorenzo_raw_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_lorenzo_params *p,
                                   uint8_t *h_raw, size_t cap, size_t *raw_size);
/* SZGeneralCompressor::decompress after zstd (SGC:115-140) into d_out. */
faz_status faz_gpu_lorenzo_decompress_raw(faz_gpu_ctx *ctx, const uint8_t *h_raw, size_t n, const size_t *dims,
                                          float *d_out);
/* SZ_compress for cmprAlgo = ALGO_LORENZO_REG on a 3-D field
 * (SZDispatcher.hpp:13-20, SZLorenzoReg.hpp:95-114, sz.hpp:54-91): bound
 * conversion as faz_sz_compress_f32, zstd, Config trailer.  faz_sz_decompress_f32
 * decodes the result. */
faz_status faz_sz_compress_lorenzo_f32(faz_gpu_ctx *ctx, const faz_lorenzo_params *p, int rel_mode, double eb,
                                       const float *data, int data_on_device, uint8_t **out, size_t *out_size);

/* --- A18, the default (SZ3) tuning mode: Tuning with
 * predictorTuningRate = 0 (include/QoZ/api/impl/SZInterp.hpp:2096-2180),
 * what `faz` runs without -q: sampling (utils/Extraction.hpp:17-212), the
 * Lorenzo compressor and the block interpolation compressor
 * (compressor/deprecated/SZBlockInterpolationCompressor.hpp, driven by
 * do_not_use_this_interp_compress_block_test, SZI:558-576) on the sample,
 * and for the Lorenzo choice optimize_quant_invl_3d
 * (utils/QuantOptimization.hpp:89-147) and the tests of SZI:2783-2850. */
typedef struct {
    int sampled;                 /* 0: the field is its own sample (sampling block < 9) */
    size_t sample_dims[4];
    size_t sampling_block;
    size_t lorenzo_bytes;        /* Lorenzo compressor on the sample (blockSize 5, 1- and 2-layer) */
    size_t interp_bytes[3];      /* block interpolation test: linear, cubic (direction 0), best with N!-1 */
    double lorenzo_ratio, interp_ratio;
    int use_interp, interp_algo, direction;
    int quantbin_estimate;       /* optimize_quant_invl_3d (3-D, Lorenzo chosen) */
    size_t lorenzo_bytes_pd2, lorenzo_bytes_16384;
    int quantbin;                /* quantbinCnt of the Lorenzo path */
    double final_lorenzo_ratio;
} faz_sz3_out;

/* sampling<T,N> of a device field (N <= 3) into d_out (cap floats);
 * *sampled = 0 when the field is its own sample (nothing written). */
faz_status faz_gpu_sampling_f32(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims, float *d_out,
                                size_t cap, size_t *sample_dims, size_t *sampling_block, int *sampled);
/* do_not_use_this_interp_compress_block_test (SZI:558-576): the block
 * interpolation compressor on a device array (N <= 3); *bytes = zstd-framed
 * size, h_raw (may be NULL) = the pre-zstd buffer (SBI:150-165). */
faz_status faz_gpu_block_interp_test_f32(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims, double eb,
                                         int algo, int direction, size_t block, uint8_t *h_raw, size_t cap,
                                         size_t *raw_size, size_t *bytes);
/* optimize_quant_invl_3d (QO:89-147) on a 3-D device field. */
faz_status faz_gpu_quant_intervals_f32(faz_gpu_ctx *ctx, const float *d_in, const size_t *dims, double eb,
                                       int *quantbin, size_t *samples);
/* The default-mode tuning decisions for a device field (N <= 3). */
faz_status faz_gpu_sz3_tune_f32(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims, double abs_eb,
                                double rel_eb, faz_sz3_out *out);
/* SZ_compress for ALGO_INTERP_LORENZO in the default mode (no conf.QoZ, no
 * conf.FZ; SZI:2562-2899): tuning, then the interpolation or Lorenzo engine
 * with the tuned Config; zstd and trailer as the reference writes them. */
faz_status faz_sz_compress_sz3_f32(faz_gpu_ctx *ctx, int N, const size_t *dims, int rel_mode, double eb,
                                   const float *data, int data_on_device, uint8_t **out, size_t *out_size);

/* --- Batches of fields of uneven shapes (ALGO_INTERP engine).  No
 * counterpart in the reference, which compresses one field per call: this is
 * SZ_compress (sz.hpp:54-91, SZInterp.hpp:387-414) over many fields at once,
 * each field's stream identical to a single-field call's.  The inputs are
 * bound at creation (device arrays that stay valid and that the caller
 * refills); each field's range kernel sits in one hipGraph and its
 * prediction/quantization/histogram/outlier launches on its own branch of a
 * second, which is replayed while the bounds stay the same and re-captured
 * and updated in place (hipGraphExecUpdate) when they change.  autotune != 0
 * times the launch schedules of every 3-D field on its bound input and keeps
 * the fastest (all produce the same bytes). */
typedef struct faz_batch faz_batch;

typedef struct {
    uint64_t calls, captures, updates, instantiations, replays, tuned_fields;
    double range_ms, pred_ms, encode_ms, total_ms;  /* phases of the last call */
} faz_batch_stats;

/* fields[i]: shape and engine options of field i (eb ignored); d_in[i]: its
 * device input. */
faz_status faz_gpu_batch_create(faz_gpu_ctx *ctx, const faz_interp_params *fields, const float *const *d_in, int n,
                                int autotune, faz_batch **out);
void faz_gpu_batch_destroy(faz_batch *b);
/* Compresses every bound field: rel_mode[i] != 0 takes eb[i] relative to the
 * field's range (calAbsErrorBound, SZ_compress), else absolute.  raw_sizes[i]
 * receives the engine buffer size; abs_eb / rng (may be NULL) the absolute
 * bound and the float range of each field. */
faz_status faz_gpu_batch_compress(faz_batch *b, const int *rel_mode, const double *eb, size_t *raw_sizes,
                                  double *abs_eb, double *rng);
/* Engine buffer of field i from the last faz_gpu_batch_compress. */
faz_status faz_gpu_batch_fetch_raw(faz_batch *b, int i, uint8_t *h_raw, size_t cap);
/* Counters of the batch and (schedules, may be NULL: n ints) the launch
 * schedule the autotuner picked per field. */
faz_status faz_gpu_batch_stats(const faz_batch *b, faz_batch_stats *s, int *schedules);
/* SZ_compress streams of every bound field (zstd + Config trailer per field,
 * as faz_sz_compress_f32 writes them); out[i] is released with faz_free. */
faz_status faz_sz_compress_batch_f32(faz_batch *b, const int *rel_mode, const double *eb, uint8_t **out,
                                     size_t *out_size);
/* Parameters field i was bound with. */
faz_status faz_gpu_batch_field(const faz_batch *b, int i, faz_interp_params *p);

/* --- A20: the general frontend, N = 1..4: SZGeneralFrontend
 * (include/QoZ/frontend/SZGeneralFrontend.hpp:36-116) over
 * ComposedPredictor (ComposedPredictor.hpp:45-93, 203-249) of the enabled
 * predictors in the order LorenzoPredictor<1>, LorenzoPredictor<2>
 * (LorenzoPredictor.hpp:25-247), RegressionPredictor
 * (RegressionPredictor.hpp:17-246), or the one enabled alone, as
 * make_lorenzo_regression_compressor composes them with regression2 off
 * (SZLorenzoReg.hpp:27-91).  This is what SZ_compress_LorenzoReg runs for
 * N != 3 (LR:103-110) and what the tuner's Lorenzo CompressTest runs for
 * every N (SZI:1136-1170). */
typedef struct {
    int N;                  /* 1..4 */
    size_t dims[4];
    double eb;              /* absolute error bound */
    int block_size;         /* Config::blockSize (N=1: <= 2^20, 2: <= 64, 3: <= 16, 4: <= 8) */
    int lorenzo;            /* Config::lorenzo */
    int lorenzo2;           /* Config::lorenzo2 (needs block_size >= 2) */
    int radius;             /* Config::quantbinCnt / 2, <= 65536 */
    int regression;         /* Config::regression (one field per call) */
} faz_general_params;

/* Results of faz_gpu_general_compress_f32 (device pointers owned by ctx). */
typedef struct {
    uint32_t *d_codes;          /* quantization codes, block-major (GF:47-55) */
    size_t n_codes;
    float *d_outval;            /* at every zero code, the unpredictable value */
    const uint8_t *d_selection; /* per block, index into the predictor list; 255 where the
                                   chosen predictor declined the block and none is recorded */
    size_t n_selection;         /* 0 when a single predictor is on */
    size_t n_regression;        /* blocks predicted by the RegressionPredictor */
    int passes;                 /* wavefront passes until the predictor choices settled (1
                                   without the RegressionPredictor) */
} faz_general_out;

/* SZGeneralFrontend::compress (GF:36-71); d_recon (n floats) receives the
 * array the reference leaves overwritten. */
faz_status faz_gpu_general_compress_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon,
                                        const faz_general_params *p, faz_general_out *out);
/* The buffer SZGeneralCompressor::compress hands to zstd (SGC:30-71): the
 * frontend's save (GF:98-103, CP:70-81), LinearQuantizer::save, the codes'
 * tree and bits.  Needs radius <= 32768.  d_recon may be NULL. */
faz_status faz_gpu_general_raw_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_general_params *p,
                                   uint8_t *h_raw, size_t cap, size_t *raw_size);
/* SZGeneralCompressor::decompress after zstd (SGC:115-140, GF:72-96, 105-116,
 * RP:134-156) into d_out (num floats); lorenzo / lorenzo2 / regression come
 * from the stream's Config. */
faz_status faz_gpu_general_decompress_raw(faz_gpu_ctx *ctx, const uint8_t *h_raw, size_t n, int N, int lorenzo,
                                          int lorenzo2, int regression, float *d_out, size_t num);
/* CompressTest with algo = ALGO_LORENZO_REG, useFast = false (SZI:1136-1170,
 * 1254-1256, 1283-1294, 1353-1392): one general compressor (lorenzo +
 * lorenzo2, blockSize 5, radius 65536) over the (sbs+1)^N sample blocks at
 * `starts` (size_t triples) of the N = 2 or 3 field d_in (dims fdims);
 * *bytes = the zstd-framed encoding_lossless size, *square_error = the summed
 * float squared errors.  h_raw (may be NULL) receives the pre-zstd buffer. */
faz_status faz_gpu_compress_test_lorenzo_f32(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *fdims,
                                             const size_t *starts, size_t n_blocks, size_t sbs, double eb,
                                             uint8_t *h_raw, size_t cap, size_t *raw_size, size_t *bytes,
                                             double *square_error);
/* SZ_compress for cmprAlgo = ALGO_LORENZO_REG through the general frontend
 * (N != 3: LR:95-114; sz.hpp:54-91): bound conversion, zstd, Config trailer.
 * faz_sz_decompress_f32 decodes the result. */
faz_status faz_sz_compress_general_f32(faz_gpu_ctx *ctx, const faz_general_params *p, int rel_mode, double eb,
                                       const float *data, int data_on_device, uint8_t **out, size_t *out_size);

/* --- A18: the exhaustive QoZ sampling tuner (`faz -q`): Tuning() with
 * conf.QoZ set, RD target, no wavelets (include/QoZ/api/impl/SZInterp.hpp:
 * 1589-2557): sampleBlocks (SZI:993-1127), per-level interpolator and
 * direction from tuning-mode-2 losses, alpha/beta from CompressTest
 * (SZI:1131-1392) bitrate/PSNR, setABwithRelBound (SZI:1394-1468).  The
 * sample blocks are read in place from d_in (the field on the device). */
typedef struct {
    int N;                  /* 2 or 3 */
    size_t dims[4];
    double abs_eb;          /* Config::absErrorBound after calAbsErrorBound */
    double rel_eb;          /* Config::relErrorBound when Tuning starts */
    double rng;             /* Config::rng (max - min) */
    int interp_block_size;  /* Config::interpBlockSize */
    size_t max_step;        /* Config::maxStep, 0: the QoZ default */
    int levelwise;          /* Config::levelwisePredictionSelection, 0: the QoZ default */
    int sample_block_size;  /* Config::sampleBlockSize, 0: the QoZ default */
    int ab_list;            /* Config::abList: alphas {1..2}, 1: up to 2.5, 2: up to 3 (SZI:883-891) */
    int test_lorenzo;       /* Config::testLorenzo: the Lorenzo CompressTest after the alpha/beta
                               search (SZI:2431-2493); the QoZ preset clears it (SZI:1616), the
                               FZ preset sets it (SZI:1645) */
    /* learned tuner only (faz_gpu_tune_learned_f32; no reference counterpart):
     * candidates whose predicted cost (bits/value at equal PSNR, 6.02 dB per
     * bit) lies within learned_margin of the best are verified by real
     * CompressTests, at most learned_verify of them (0: 4; learned_margin < 0:
     * never); an input row whose standardised value leaves [learned_lo[i],
     * learned_hi[i]] (FAZ_LEARNED_INPUTS each, NULL: no check) falls back to
     * the exhaustive search. */
    double learned_margin;
    int learned_verify;
    const float *learned_lo, *learned_hi;
    /* with test_lorenzo: the Lorenzo candidate and the best interpolation
     * candidate are both verified when their predicted costs lie within this
     * margin of each other (< 0: never) */
    double learned_margin_lorenzo;
} faz_tune_in;

typedef struct {
    double alpha, beta;     /* chosen level-wise bound scaling */
    int levelwise;          /* entries of ops/dirs */
    uint8_t ops[32];        /* interpAlgo_list */
    uint8_t dirs[32];       /* interpDirection_list */
    int interp_algo;        /* Config::interpAlgo / interpDirection as Tuning leaves them */
    int direction;
    size_t max_step;
    int sample_block_size;
    size_t n_blocks;        /* sampled blocks and their starts */
    size_t block_starts[4096][3];
    double pred_loss[32][4];/* [level-1][op * 2 + (direction != 0)] summed mode-2 loss */
    int n_tests;            /* CompressTest calls, in order */
    double test_alpha[128], test_beta[128], test_eb[128], test_bitrate[128], test_psnr[128];
    size_t test_bytes[128]; /* compressed bytes of each CompressTest */
    double ms;              /* device+host time of the tuner */
    int use_interp;         /* 1: interpolation (ALGO_INTERP), 0: the Lorenzo test won (ALGO_LORENZO_REG) */
    int learned_path;       /* learned tuner: 0 the network's choice, 1 chosen among real CompressTests of
                               the closest candidates, 2 exhaustive fallback (input out of the training
                               range); -1 the exhaustive tuner */
    int n_verified;         /* real CompressTests the learned tuner ran */
    double level_res[32];   /* [level-1] mean |mode-2 prediction error| per point of the chosen
                               interpolator and direction (levels above levelwise folded into the
                               top tuned level), in units of abs_eb */
} faz_tune_out;

faz_status faz_gpu_tune_qoz_f32(faz_gpu_ctx *ctx, const float *d_in, const faz_tune_in *in, faz_tune_out *out);

/* One CompressTest (SZI:1131-1392, ALGO_INTERP, RD) on the sample blocks of
 * edge sbs+1 starting at starts[3*b..3*b+2] (b < nblocks) of the device field
 * d_in (dims fdims, p->N of them): every block in tuning mode 1 with p's
 * engine parameters (p->dims is ignored), then encoding_lossless
 * (SIC:784-809).  Returns the pre-zstd buffer (h_raw, capacity cap), the
 * zstd-framed size (*bytes) and the summed squared error. */
faz_status faz_gpu_compress_test_f32(faz_gpu_ctx *ctx, const float *d_in, const size_t *fdims,
                                     const faz_interp_params *p, const size_t *starts, size_t nblocks, size_t sbs,
                                     double rng, uint8_t *h_raw, size_t cap, size_t *raw_size, size_t *bytes,
                                     double *square_error);
/* The same for ncand candidates (same N, maxStep, interpBlockSize) on the
 * same blocks, in one launch: bytes[c] and square_error[c] per candidate. */
faz_status faz_gpu_compress_test_batch_f32(faz_gpu_ctx *ctx, const float *d_in, const size_t *fdims,
                                           const faz_interp_params *ps, size_t ncand, const size_t *starts,
                                           size_t nblocks, size_t sbs, double rng, size_t *bytes,
                                           double *square_error);

/* SZ_compress with cmprAlgo = ALGO_INTERP_LORENZO and conf.QoZ (`faz -q`,
 * SZInterp.hpp:2562-2899): the QoZ tuner above, then the interpolation
 * engine with the tuned Config (stream as faz_sz_compress_f32). */
faz_status faz_sz_compress_qoz_f32(faz_gpu_ctx *ctx, int N, const size_t *dims, int rel_mode, double eb,
                                   const float *data, int data_on_device, uint8_t **out, size_t *out_size);

/* --- Learned tuner: MLP surrogate with BF16 matrix-core kernels (no
 * counterpart in the reference, which tunes only exhaustively).
 * in_dim -> hidden (x hidden_layers, ReLU) -> out_dim (linear); weights
 * He-uniform from `seed`.  Inputs and targets are device fp32 row-major
 * [batch x in_dim] / [batch x out_dim].  Parameters on the host are, per
 * layer, W as [K x N] row-major (padded shapes from faz_mlp_layout) then b[N]. */
typedef struct faz_mlp faz_mlp;
faz_status faz_mlp_create(faz_gpu_ctx *ctx, int in_dim, int hidden, int hidden_layers, int out_dim, uint64_t seed,
                          faz_mlp **out);
void faz_mlp_destroy(faz_mlp *net);
faz_status faz_mlp_layout(const faz_mlp *net, int *n_layers, int *K, int *N);
faz_status faz_mlp_forward(faz_gpu_ctx *ctx, faz_mlp *net, const float *d_x, int batch, float *d_y);
/* Mean-squared-error gradients of every parameter (host, layout above) and
 * the loss, without updating. */
faz_status faz_mlp_gradients(faz_gpu_ctx *ctx, faz_mlp *net, const float *d_x, const float *d_t, int batch,
                             float *h_grad, size_t cap, double *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch. */
faz_status faz_mlp_train_step(faz_gpu_ctx *ctx, faz_mlp *net, const float *d_x, const float *d_t, int batch,
                              float lr, double *loss);
faz_status faz_mlp_get_params(faz_gpu_ctx *ctx, const faz_mlp *net, float *h, size_t cap, size_t *n);
faz_status faz_mlp_set_params(faz_gpu_ctx *ctx, faz_mlp *net, const float *h, size_t n);

/* --- Learned tuner: the QoZ tuner with the alpha/beta CompressTest search
 * replaced by the MLP's predictions of the same bitrate and PSNR (no
 * counterpart in the reference).  Block statistics of the sample blocks
 * (FAZ_LEARNED_FEATURES per block, of u = (v - vmin) / rng):
 *   0 mean, 1 std, 2 min, 3 max; for strides 1, 2, 4 (base 4 + 9 * i): mean
 *   |first difference| per axis (3), its RMS per axis (3), mean |Laplacian|,
 *   RMS Laplacian, max |Laplacian|; 31-33 lag-1 and 34-36 lag-2 correlation
 *   per axis; 37 kurtosis of the stride-1 Laplacian; 38 share of it above
 *   twice its RMS; 39 share of stride-1 differences below eb_norm.
 * The MLP input row (FAZ_LEARNED_INPUTS) of one candidate: those statistics
 * averaged over the blocks (magnitudes as log2 of their ratio to eb_norm =
 * eb / rng), alpha - 1, beta / 4, log2 beta, level-1 interpolator and
 * direction flag, their means over the levels, log10 of the relative bound,
 * log2 of the bound's scale, for levels 1..4 log2 of the level's mean mode-2
 * prediction error over the candidate's bound of that level (level_res in
 * units of the base bound, times min(alpha^(level-1), beta) / scale), and
 * 1 for the Lorenzo candidate (alpha, beta entries 0) or 0.
 * Targets: log2 bitrate and (PSNR + 20 log10(eb_norm * scale)) / 5. */
#define FAZ_LEARNED_FEATURES 40
#define FAZ_LEARNED_INPUTS 54
faz_status faz_gpu_block_features_f32(faz_gpu_ctx *ctx, const float *d_in, int N, const size_t *dims,
                                      const size_t *starts, size_t nblocks, size_t sbs, float vmin, float rng,
                                      double eb_norm, float *h_feats);
faz_status faz_learned_input(const float *block_feats, size_t nblocks, double eb_norm, double rel_eb, double alpha,
                             double beta, double eb_scale, int levelwise, const uint8_t *ops, const uint8_t *dirs,
                             const double *level_res, int lorenzo, float *row);
faz_status faz_learned_target(double bitrate, double psnr, double eb_norm, double eb_scale, float *t);
/* The tuner with the trained network (inputs standardised as (x - mu) / sigma):
 * the network predicts every alpha/beta candidate (and the Lorenzo test when
 * in->test_lorenzo) at once and the lowest predicted bits/value at equal PSNR
 * wins, unless the confidence gates of faz_tune_in send the closest
 * candidates to real CompressTests or the whole search to the exhaustive
 * tuner.  out as faz_gpu_tune_qoz_f32: predicted tests (test_bytes 0), then
 * the real ones. */
faz_status faz_gpu_tune_learned_f32(faz_gpu_ctx *ctx, const float *d_in, const faz_tune_in *in, faz_mlp *net,
                                    const float *mu, const float *sigma, faz_tune_out *out);

/* Timing of the last faz_gpu_interp_raw_f32 / faz_sz_compress_f32 call, in
 * milliseconds, measured with HIP events on the context stream (entries 0-6
 * of the device-resident compressions only while faz_gpu_set_profiling has
 * bit 1 set, else 0):
 * [0] range, [1] interpolation+quantization, [2] histogram+outlier compaction,
 * [3] host tree build, [4] encode, [5] device total, [6] dominant kernel
 * (interpolation pass launches) summed, [8] host time from the histogram's
 * arrival to the Huffman tree, [9] host time enqueueing the encode.
 * faz_gpu_interp_decompress_raw and faz_gpu_interp_decompress_dev set [6] to
 * the host time of their Huffman stage (tree parse, tables, upload or device
 * copy enqueued; the device decode runs on behind it) and [7] to the whole
 * call, host clock. */
faz_status faz_gpu_last_timings(const faz_gpu_ctx *ctx, double *ms, int n);

/* on: bit 0 enables HIP-event timing of the interpolation kernels,
 * accumulated for each kernel family separately (consecutive whole-tile level
 * launches are timed as one span, so the average launch time includes the
 * few idle microseconds between them); faz_gpu_profile_read reports the
 * family that took longest (the dominant kernel): total milliseconds,
 * launches and predicted points, and faz_gpu_profile_kernel which family that
 * is (0 whole-tile 3-D level kernel, 1 per-pass kernel, 2 per-block 3-D level
 * kernel).  Bit 1 enables the phase timings of faz_gpu_last_timings.  Each
 * event record idles the stream for a few microseconds.  Enabling resets the
 * counters. */
faz_status faz_gpu_set_profiling(faz_gpu_ctx *ctx, int on);
faz_status faz_gpu_profile_read(const faz_gpu_ctx *ctx, double *ms, uint64_t *launches, uint64_t *points);
int faz_gpu_profile_kernel(const faz_gpu_ctx *ctx);
/* Algorithmic HBM bytes of the dominant family's launches: 4 per original
 * read and 2 per code written, plus 4 per reconstructed lattice point read
 * from the coarser level and 4 per point of the level's lattice written
 * (when it is written). */
double faz_gpu_profile_bytes(const faz_gpu_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
