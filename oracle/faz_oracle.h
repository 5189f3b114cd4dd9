/* CPU oracle for the FAZ compression hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A scalar restatement of the reference algorithm, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * The product (faz_amd/) never links, loads or calls it.
 *
 * All functions return 0 on success, a negative value on error.
 */
#ifndef FAZ_ORACLE_H
#define FAZ_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the QoZ::Config fields the interpolation engine reads
 * (include/QoZ/utils/Config.hpp:394-499, SIC:333-345). */
typedef struct {
    int N;
    size_t dims[4];
    double eb;            /* absolute error bound (after calAbsErrorBound) */
    int interp_algo;      /* 0 linear, 1 cubic */
    int direction;        /* index into next_permutation order of axes */
    double alpha, beta;
    size_t max_step;      /* anchor stride, 0 = no anchors */
    int block_size;       /* interpBlockSize */
    int radius;           /* quantbinCnt / 2 */
    int levelwise;        /* entries in ops/dirs (0 = none) */
    uint8_t ops[32];
    uint8_t dirs[32];
} oracle_interp_params;

/* SZInterpolationCompressor::compress in tuning mode 0/1/2 (SIC:328-476).
 * data is overwritten with the reconstruction as the reference does.
 * codes: capacity n (mode 2 pushes none); unpred capacity n.
 * level_counts: capacity 64, filled in tuning modes. */
int oracle_interp_compress(const oracle_interp_params *p, float *data, int tuning, int start_level, int end_level,
                           int32_t *codes, size_t *n_codes, float *unpred, size_t *n_unpred,
                           double *loss, size_t *level_counts, int *n_levels);

/* Pre-zstd engine buffer (SIC:485-518): header, quantizer, Huffman tree, bits. */
int oracle_interp_raw(const oracle_interp_params *p, const float *data, uint8_t *out, size_t cap, size_t *n_out);

/* SZInterpolationCompressor::decompress on a pre-zstd buffer (SIC:42-199). */
int oracle_interp_decompress_raw(const uint8_t *raw, size_t n, int N, float *out, size_t num);

/* HuffmanEncoder preprocess_encode + save + encode (HUF:103-227). */
int oracle_huffman_encode(const int32_t *codes, size_t n, uint8_t *out, size_t cap, size_t *n_out);
/* HuffmanEncoder load + decode (HUF:236-296). *n_used = bytes consumed. */
int oracle_huffman_decode(const uint8_t *buf, size_t n, int32_t *out, size_t num, size_t *n_used);

/* Full SZ_compress stream for ALGO_INTERP (sz.hpp:54-91): the error bound
 * mode is applied first (rel_mode != 0 → eb is relative, Statistic.hpp:34-66). */
int oracle_compress_stream(const oracle_interp_params *p, int rel_mode, double eb_in, const float *data,
                           uint8_t *out, size_t cap, size_t *n_out);
/* SZ_decompress of an ALGO_INTERP or (3-D) ALGO_LORENZO_REG stream. */
int oracle_decompress_stream(const uint8_t *stream, size_t n, float *out, size_t num);

/* data_range (Statistic.hpp:11-21). */
int oracle_range(const float *data, size_t n, float *mn, float *mx);

/* The block-order sequence index of every point (where its code lands in
 * the code stream), -1 for none. Used by tests to pin the closed form. */
int oracle_interp_seq_index(const oracle_interp_params *p, int64_t *seq);

/* Config fields the 3-D Lorenzo/regression engine reads: SZFastFrontend
 * (include/QoZ/frontend/SZFastFrontend.hpp:25-31) and the quantizer radius
 * (SZLorenzoReg.hpp:102). */
typedef struct {
    size_t dims[3];
    double eb;            /* absolute error bound (after calAbsErrorBound) */
    int block_size;       /* conf.blockSize */
    int pred_dim;         /* conf.pred_dim: 3, 2 or 1 */
    int lorenzo, lorenzo2, regression;
    int radius;           /* quantbinCnt / 2 */
} oracle_lr_params;

/* SZFastFrontend::compress_3d (FF:176-345). Every output is optional (NULL).
 * codes, recon: n points (codes block-major); indicator: one selector per
 * block; raw_coeffs: 4 per block, the unquantized regression coefficients
 * (0 where regression is off for the block); reg_codes / reg_coeffs: 4 per
 * regression block, quantized; reg_unpred and unpred: capacity n. */
int oracle_lr_compress(const oracle_lr_params *p, const float *data, int32_t *codes, float *recon,
                       int32_t *indicator, size_t *n_blocks, float *raw_coeffs,
                       int32_t *reg_codes, float *reg_coeffs, size_t *n_reg,
                       float *reg_unpred, size_t *n_reg_unpred, float *unpred, size_t *n_unpred);

/* Pre-zstd buffer of SZGeneralCompressor<SZFastFrontend> (SGC:30-71, FF:50-84). */
int oracle_lr_raw(const oracle_lr_params *p, const float *data, uint8_t *out, size_t cap, size_t *n_out);
/* SGC:115-140 on a pre-zstd buffer. */
int oracle_lr_decompress_raw(const uint8_t *raw, size_t n, const size_t *dims, float *out);

/* Full SZ_compress stream for ALGO_LORENZO_REG (SZDispatcher.hpp:13-20,
 * SZLorenzoReg.hpp:95-114, sz.hpp:54-91); p->eb is replaced by the bound
 * the error-bound mode gives. */
int oracle_lr_compress_stream(const oracle_lr_params *p, int rel_mode, double eb_in, const float *data,
                              uint8_t *out, size_t cap, size_t *n_out);

/* Config fields of the general frontend with the Lorenzo predictors
 * (SZGeneralFrontend.hpp, ComposedPredictor.hpp, LorenzoPredictor.hpp as
 * composed by SZLorenzoReg.hpp:27-91), N = 1..4. */
typedef struct {
    int N;
    size_t dims[4];
    double eb;            /* absolute error bound */
    int block_size;       /* conf.blockSize */
    int lorenzo, lorenzo2;
    int radius;           /* quantbinCnt / 2 */
    int regression;       /* conf.regression: RegressionPredictor in the list */
} oracle_gf_params;

/* SZGeneralFrontend::compress (GF:36-71). Outputs optional: codes (n,
 * block-major), recon (n), selection (one per block when both predictors are
 * on), unpred (capacity n). */
int oracle_gf_compress(const oracle_gf_params *p, const float *data, int32_t *codes, float *recon,
                       int32_t *selection, size_t *n_sel, float *unpred, size_t *n_unpred);
/* Pre-zstd buffer of SZGeneralCompressor<SZGeneralFrontend> (SGC:30-71, GF:98-103). */
int oracle_gf_raw(const oracle_gf_params *p, const float *data, uint8_t *out, size_t cap, size_t *n_out);
/* SGC:115-140 on a pre-zstd buffer; lorenzo/lorenzo2/regression as in the Config. */
int oracle_gf_decompress_raw(const uint8_t *raw, size_t n, int N, int lorenzo, int lorenzo2, int regression,
                             float *out, size_t num);
/* SZ_compress stream for ALGO_LORENZO_REG through the general frontend. */
int oracle_gf_compress_stream(const oracle_gf_params *p, int rel_mode, double eb_in, const float *data,
                              uint8_t *out, size_t cap, size_t *n_out);
/* CompressTest's Lorenzo variant (useFast = false): one general compressor
 * with lorenzo + lorenzo2, blockSize 5, radius 65536 over the sample blocks
 * of (sbs+1)^N at `starts` (int64-free: size_t triples, N = 2 or 3). */
int oracle_compress_test_lorenzo(int N, double eb, const float *field, const size_t *fdims, const size_t *starts,
                                 size_t nblocks, size_t sbs, uint8_t *raw, size_t cap, size_t *raw_size,
                                 size_t *bytes, double *square_error);

/* The default (SZ3) tuning mode: Tuning with predictorTuningRate = 0
 * (SZInterp.hpp:2096-2180) and the Lorenzo branch of
 * SZ_compress_Interp_lorenzo (SZI:2783-2898). */
typedef struct {
    int sampled;                 /* 0: the field is its own sample (sampling block < 9) */
    size_t sample_dims[4];
    size_t sampling_block;
    size_t lorenzo_bytes;        /* SZ_compress_LorenzoReg on the sample (bs 5, lorenzo + lorenzo2) */
    size_t interp_bytes[3];      /* block interpolation test: linear, cubic (direction 0), best with N!-1 */
    double lorenzo_ratio, interp_ratio;
    int use_interp, interp_algo, direction;
    int quantbin_estimate;       /* optimize_quant_invl_3d (N = 3, Lorenzo chosen) */
    size_t lorenzo_bytes_pd2, lorenzo_bytes_16384;
    int quantbin;                /* final quantbinCnt of the Lorenzo path */
    double final_lorenzo_ratio;
} oracle_sz3_out;

/* sampling<T,N> (Extraction.hpp:17-212).  Returns 1 when the field is its
 * own sample (nothing written). */
int oracle_sampling(const float *data, int N, const size_t *dims, float *out, size_t cap, size_t *sample_dims,
                    size_t *sampling_block);
/* do_not_use_this_interp_compress_block_test (SZI:558-576) through the
 * block interpolation compressor (deprecated/SZBlockInterpolationCompressor.hpp):
 * *bytes = zstd-framed size; raw (may be NULL) = the pre-zstd buffer. */
int oracle_block_interp_test(const float *data, int N, const size_t *dims, double eb, int algo, int dir, size_t bs,
                             uint8_t *raw, size_t cap, size_t *raw_size, size_t *bytes);
/* optimize_quant_invl_3d (QuantOptimization.hpp:89-147). */
int oracle_quant_intervals(const float *data, const size_t *dims, double eb, size_t *samples);
/* The SZ3 tuning decisions for a field (N = 1..3). */
int oracle_sz3_tune(const float *data, int N, const size_t *dims, double eb, double rel, oracle_sz3_out *out);
/* SZ_compress for ALGO_INTERP_LORENZO in the default mode (no -q, no FZ). */
int oracle_sz3_compress_stream(const float *data, int N, const size_t *dims, int rel_mode, double eb_in,
                               uint8_t *out, size_t cap, size_t *n_out, oracle_sz3_out *tune);

/* QoZ exhaustive sampling tuner (`faz -q`): Tuning() with conf.QoZ set,
 * no wavelets, tuningTarget RD (include/QoZ/api/impl/SZInterp.hpp:1589-2557),
 * with sampleBlocks (SZI:993-1127), CompressTest (SZI:1131-1392),
 * setABwithRelBound (SZI:1394-1468) and init_alphalist/betalist
 * (SZI:869-933).  N = 2 or 3. */
typedef struct {
    int N;
    size_t dims[4];
    double abs_eb;            /* conf.absErrorBound after calAbsErrorBound */
    double rel_eb;            /* conf.relErrorBound when Tuning starts (> 0) */
    double rng;               /* conf.rng */
    int interp_block_size;    /* conf.interpBlockSize */
    size_t max_step;          /* conf.maxStep (0: QoZ default) */
    int levelwise;            /* conf.levelwisePredictionSelection (0: QoZ default) */
    int sample_block_size;    /* conf.sampleBlockSize (0: QoZ default) */
    int ab_list;              /* conf.abList (SZI:883-891) */
    int test_lorenzo;         /* conf.testLorenzo: the Lorenzo test of SZI:2431-2493 */
} oracle_tune_in;

typedef struct {
    double alpha, beta;       /* chosen */
    int levelwise;
    uint8_t ops[32], dirs[32];
    int interp_algo, direction;   /* conf.interpAlgo / interpDirection left by the predictor loop */
    size_t max_step;
    int sample_block_size;
    size_t n_blocks;          /* sampled blocks */
    size_t block_starts[4096][3];
    double pred_loss[32][4];  /* [level-1][op*2 + (dir != 0)] summed mode-2 loss */
    int n_tests;              /* CompressTest calls in order */
    double test_alpha[128], test_beta[128], test_eb[128], test_bitrate[128], test_psnr[128];
    size_t test_bytes[128];   /* compressed bytes of encoding_lossless */
    int use_interp;           /* 0: the Lorenzo test won (alpha = beta = -1) */
} oracle_tune_out;

int oracle_qoz_tune(const oracle_tune_in *in, const float *data, oracle_tune_out *out);

/* CompressTest (ALGO_INTERP, RD) on the blocks of (sbs+1)^N values starting
 * at starts[3*b..] of a field: the pre-zstd encoding_lossless buffer, the
 * compressed size and the summed squared error. */
int oracle_compress_test(const oracle_interp_params *p, const float *field, const size_t *fdims, const size_t *starts,
                         size_t nblocks, size_t sbs, double rng, uint8_t *raw, size_t cap, size_t *raw_size,
                         size_t *bytes, double *square_error);

#ifdef __cplusplus
}
#endif
#endif
