// Host orchestration of the gfx950 hot path behind the C-ABI (faz_gpu.h).
//
// The interpolation engine runs level by level and, within a level, pass by
// pass: every (level, pass) is one data-parallel launch over all blocks,
// because a pass only reads points fixed by earlier passes or coarser levels
// (in the reference's sequential block order those are exactly the points
// already written).  The host computes, per level, the number of codes each
// block emits so every kernel thread can place its code at the position the
// reference would push it (level, block raster, pass, line, stencil order).
#include "faz_gpu.h"
#include "faz_internal.h"
#include "faz_ctx.h"
#include "../host/huffman_host.h"
#include "../host/engine_format.h"
#include "../host/raw_api.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <memory>
#include <numeric>
#include <vector>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <thread>

using namespace faz;

namespace {

// ---------------------------------------------------------------------------
// Engine plan: SZInterpolationCompressor::init (SIC:820-852), the level-wise
// error bounds and operators (SIC:384-422) and the per-level launches.
// ---------------------------------------------------------------------------
struct LevelPlan {
    unsigned level;
    uint32_t nb[kMaxN];        // blocks per axis in the raster
    uint64_t bs;
    double eb, recip;
    int linear;
    int dir;
    uint64_t base;             // codes before this level
    uint64_t codes;            // codes of this level
    size_t prefix_at;          // offset into Plan::prefix
    std::vector<PassDesc> passes;
};

struct Plan {
    int N;
    uint32_t dims[kMaxN];
    uint64_t offs[kMaxN];
    uint64_t num;
    int levels;                // interpolation_level after init
    bool anchor;
    uint64_t head_codes;       // anchors or the first point
    uint64_t anchor_count;
    uint64_t total_codes;
    std::vector<std::array<int, kMaxN>> perms;
    std::vector<LevelPlan> lv;
    std::vector<uint64_t> prefix;
};

int make_plan(const faz_interp_params *p, Plan &pl, std::string &why) {
    pl.N = p->N;
    if (pl.N < 1 || pl.N > 4) {
        why = "N must be 1..4";
        return FAZ_ERR_INVALID;
    }
    pl.num = 1;
    for (int i = 0; i < pl.N; i++) {
        if (p->dims[i] == 0 || p->dims[i] >= (1ull << 31)) {
            why = "dimension out of range";
            return FAZ_ERR_INVALID;
        }
        pl.dims[i] = (uint32_t) p->dims[i];
        pl.num *= p->dims[i];
    }
    if (pl.num >= (1ull << 32)) {
        why = "more than 2^32-1 points per field";
        return FAZ_ERR_UNSUPPORTED;
    }
    if (p->block_size < 2 || (p->block_size & 1)) {
        why = "interpBlockSize must be even and >= 2";
        return FAZ_ERR_INVALID;
    }
    if (p->radius < 1 || p->radius > 32768) {
        why = "quantbinCnt/2 must be in 1..32768 (16-bit codes)";
        return FAZ_ERR_UNSUPPORTED;
    }
    if (!(p->eb > 0)) {
        why = "error bound must be positive";
        return FAZ_ERR_INVALID;
    }
    if (p->levelwise < 0 || p->levelwise > 32) {
        why = "levelwise list too long";
        return FAZ_ERR_INVALID;
    }
    if (p->tuning != 0) {
        why = "tuning modes run through the batched tuner";
        return FAZ_ERR_UNSUPPORTED;
    }
    if (p->max_step > 0 && pl.N != 2 && pl.N != 3) {
        why = "anchors (maxStep) exist only for 2-D and 3-D fields in the reference";
        return FAZ_ERR_UNSUPPORTED;
    }
    // init(): level count and anchors
    int levels = -1;
    for (int i = 0; i < pl.N; i++)
        if (levels < std::ceil(std::log2((double) pl.dims[i]))) levels = (unsigned) std::ceil(std::log2((double) pl.dims[i]));
    pl.anchor = false;
    if (p->max_step > 0) {
        pl.anchor = true;
        levels += 1;
        int max_level = (unsigned) std::log2((double) p->max_step) + 1;
        if (max_level <= levels) levels = max_level;
    }
    pl.levels = levels;
    pl.offs[pl.N - 1] = 1;
    for (int i = pl.N - 2; i >= 0; i--) pl.offs[i] = pl.offs[i + 1] * pl.dims[i + 1];
    pl.perms.clear();
    std::array<int, kMaxN> seq = {0, 1, 2, 3};
    do {
        pl.perms.push_back(seq);
    } while (std::next_permutation(seq.begin(), seq.begin() + pl.N));
    for (int i = 0; i < p->levelwise; i++)
        if (p->dirs[i] >= pl.perms.size() || p->ops[i] > 1) {
            why = "levelwise operator/direction out of range";
            return FAZ_ERR_INVALID;
        }
    if ((size_t) p->direction >= pl.perms.size() || p->interp_algo < 0 || p->interp_algo > 1) {
        why = "interpolator or direction out of range";
        return FAZ_ERR_INVALID;
    }

    // head: anchors (build_grid, SIC:854-900) or the first point
    unsigned start = (unsigned) levels;
    if (pl.anchor) {
        uint64_t c = 1;
        for (int i = 0; i < pl.N; i++) c *= (pl.dims[i] - 1) / p->max_step + 1;
        pl.anchor_count = c;
        pl.head_codes = c;
        start--;
    } else {
        pl.anchor_count = 0;
        pl.head_codes = 1;
    }
    unsigned stop = (unsigned) std::max(p->end_level, 0);
    if (p->start_level > 0 && p->start_level <= levels) {
        why = "partial level ranges run through the batched tuner";
        return FAZ_ERR_UNSUPPORTED;
    }

    pl.lv.clear();
    pl.prefix.clear();
    uint64_t base = pl.head_codes;
    const double eb = p->eb;
    for (unsigned level = start; level > stop && level <= start; level--) {
        LevelPlan L;
        L.level = level;
        if (p->alpha < 0) {
            L.eb = level >= 3 ? eb * 0.5 : eb;
        } else if (p->alpha >= 1) {
            double r = std::pow(p->alpha, level - 1);
            if (r > p->beta) r = p->beta;
            L.eb = eb / r;
        } else {
            double r = 1 - (level - 1) * p->alpha;
            if (r < p->beta) r = p->beta;
            L.eb = eb * r;
        }
        L.recip = 1.0 / L.eb;
        int interp = p->interp_algo, dir = p->direction;
        if (p->levelwise > 0) {
            int k = (int) level - 1 < p->levelwise ? (int) level - 1 : p->levelwise - 1;
            interp = p->ops[k];
            dir = p->dirs[k];
        }
        L.linear = interp == 0;
        L.dir = dir;
        const auto &perm = pl.perms[dir];
        const uint64_t s = (uint64_t) 1 << (level - 1);
        const uint64_t bs = (uint64_t) p->block_size * s;
        if (bs >= (1ull << 31)) {
            why = "block edge exceeds 2^31";
            return FAZ_ERR_UNSUPPORTED;
        }
        // per-axis block tables
        uint32_t nb[kMaxN] = {1, 1, 1, 1};
        std::vector<uint64_t> tn[kMaxN], tg1[kMaxN], tg2[kMaxN];
        uint64_t nblocks = 1;
        for (int a = 0; a < pl.N; a++) {
            nb[a] = (uint32_t) ((pl.dims[a] - 1) / bs + 1);
            nblocks *= nb[a];
            for (uint32_t k = 0; k < nb[a]; k++) {
                uint64_t b = k * bs;
                uint64_t e = std::min<uint64_t>(b + bs, pl.dims[a] - 1);
                tn[a].push_back((e - b) / s + 1);
                tg1[a].push_back(b ? (e - b) / s : e / s + 1);
                tg2[a].push_back(b ? (e - b) / (2 * s) : e / (2 * s) + 1);
            }
        }
        // codes per block in raster order → exclusive prefix
        L.prefix_at = pl.prefix.size();
        uint64_t run = 0;
        uint32_t k[kMaxN] = {0, 0, 0, 0};
        for (uint64_t blk = 0; blk < nblocks; blk++) {
            pl.prefix.push_back(run);
            uint64_t cnt = 0;
            for (int P = 0; P < pl.N; P++) {
                uint64_t c = tn[perm[P]][k[perm[P]]] / 2;
                for (int Q = 0; Q < pl.N; Q++) {
                    if (Q == P) continue;
                    c *= Q < P ? tg1[perm[Q]][k[perm[Q]]] : tg2[perm[Q]][k[perm[Q]]];
                }
                cnt += c;
            }
            run += cnt;
            for (int a = pl.N - 1; a >= 0; a--) {
                if (++k[a] < nb[a]) break;
                k[a] = 0;
            }
        }
        L.base = base;
        L.codes = run;
        for (int a = 0; a < kMaxN; a++) L.nb[a] = nb[a];
        L.bs = bs;
        // one launch per pass over the pass lattice
        uint64_t lattice_total = 0;
        for (int P = 0; P < pl.N; P++) {
            PassDesc d;
            memset(&d, 0, sizeof(d));
            d.N = pl.N;
            d.P = P;
            d.axis = perm[P];
            for (int a = 0; a < kMaxN; a++) d.perm[a] = a < pl.N ? perm[a] : a;
            for (int Q = 0; Q < kMaxN; Q++) d.invp[d.perm[Q]] = Q;
            uint64_t total = 1;
            for (int a = 0; a < pl.N; a++) {
                d.dims[a] = pl.dims[a];
                d.offs[a] = pl.offs[a];
                d.nb[a] = nb[a];
                int Q = 0;
                while (perm[Q] != a) Q++;
                uint64_t M = (pl.dims[a] - 1) / s;
                if (Q == P) {
                    d.kind[a] = kPassAxis;
                    d.lat[a] = (uint32_t) ((M + 1) / 2);
                } else if (Q < P) {
                    d.kind[a] = kG1;
                    d.lat[a] = (uint32_t) (M + 1);
                } else {
                    d.kind[a] = kG2;
                    d.lat[a] = (uint32_t) ((pl.dims[a] - 1) / (2 * s) + 1);
                }
                total *= d.lat[a];
            }
            d.s = (uint32_t) s;
            d.ls = 0;
            while (((uint64_t) 1 << d.ls) < s) d.ls++;
            d.bs = (uint32_t) bs;
            for (int a = 0; a < pl.N; a++) d.dlat[a] = FastDiv(d.lat[a]);
            d.dbs = FastDiv(d.bs);
            d.total = total;
            d.level_base = base;
            d.eb = L.eb;
            d.recip = L.recip;
            d.radius = p->radius;
            d.linear = L.linear;
            d.tuning = 0;
            lattice_total += total;
            L.passes.push_back(d);
        }
        if (lattice_total != run) {
            why = "internal: pass lattice and block counts disagree";
            return FAZ_ERR_INVALID;
        }
        base += run;
        pl.lv.push_back(L);
    }
    pl.total_codes = base;
    return FAZ_OK;
}

// The plan of the last parameter set a context compressed with: every
// field but the bound keys it; the level bounds are set again from p->eb
// (the same arithmetic as make_plan), so repeated compressions of one shape
// skip the plan and the upload of its block prefix table.
Plan &cached_plan(faz_gpu_ctx *ctx, const faz_interp_params *p, int &rc) {
    std::vector<uint8_t> key;
    auto add = [&](const void *v, size_t n) {
        const uint8_t *b = static_cast<const uint8_t *>(v);
        key.insert(key.end(), b, b + n);
    };
    add(&p->N, sizeof(p->N));
    add(p->dims, sizeof(p->dims));
    add(&p->interp_algo, sizeof(p->interp_algo));
    add(&p->direction, sizeof(p->direction));
    add(&p->alpha, sizeof(p->alpha));
    add(&p->beta, sizeof(p->beta));
    add(&p->max_step, sizeof(p->max_step));
    add(&p->block_size, sizeof(p->block_size));
    add(&p->radius, sizeof(p->radius));
    add(&p->levelwise, sizeof(p->levelwise));
    add(p->ops, sizeof(p->ops));
    add(p->dirs, sizeof(p->dirs));
    add(&p->tuning, sizeof(p->tuning));
    add(&p->start_level, sizeof(p->start_level));
    add(&p->end_level, sizeof(p->end_level));
    Plan *pl = static_cast<Plan *>(ctx->plan_cache.get());
    rc = FAZ_OK;
    if (!pl || key != ctx->plan_key || !(p->eb > 0)) {
        ctx->plan_cache.reset();
        ctx->plan_key.clear();
        std::shared_ptr<Plan> fresh = std::make_shared<Plan>();
        rc = make_plan(p, *fresh, ctx->err);
        ctx->plan_cache = fresh;
        if (!rc) ctx->plan_key = key;
        return *fresh;
    }
    const double eb = p->eb;
    for (auto &L: pl->lv) {
        const unsigned level = L.level;
        if (p->alpha < 0) {
            L.eb = level >= 3 ? eb * 0.5 : eb;
        } else if (p->alpha >= 1) {
            double r = std::pow(p->alpha, level - 1);
            if (r > p->beta) r = p->beta;
            L.eb = eb / r;
        } else {
            double r = 1 - (level - 1) * p->alpha;
            if (r < p->beta) r = p->beta;
            L.eb = eb * r;
        }
        L.recip = 1.0 / L.eb;
        for (auto &d: L.passes) {
            d.eb = L.eb;
            d.recip = L.recip;
        }
    }
    return *pl;
}

double elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

// The (min, max) pair of the last range pass on ctx->stream, through pinned
// memory.
int range_readback(faz_gpu_ctx *ctx, float *h_min, float *h_max) {
    uint32_t *k = ctx->hp_hist.data() + 2 * 65536 + 6;  // pinned
    HIPCHECK(hipMemcpyAsync(k, ctx->minmax.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(spin_wait(ctx->stream));
    auto key2f = [](uint32_t key) {
        uint32_t u = (key & 0x80000000u) ? (key & 0x7FFFFFFFu) : ~key;
        float f;
        memcpy(&f, &u, 4);
        return f;
    };
    *h_min = key2f(k[0]);
    *h_max = key2f(k[1]);
    return FAZ_OK;
}

// Adds the recorded launches of the last compression to the profile buckets
// (their events must have completed).
void flush_profile(faz_gpu_ctx *ctx) {
    for (const auto &r: ctx->prof_pending) {
        auto &b = ctx->prof[r.bucket];
        b.ms += elapsed(ctx->prof_ev[2 * r.ev], ctx->prof_ev[2 * r.ev + 1]);
        b.launches += r.launches;
        b.points += r.points;
        b.bytes += r.bytes;
    }
    ctx->prof_pending.clear();
}

}  // namespace

extern "C" {

faz_status faz_gpu_ctx_create(int device, faz_gpu_ctx **out) {
    if (!out) return FAZ_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return FAZ_ERR_HIP;
    if (device < 0 || device >= count) return FAZ_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return FAZ_ERR_HIP;
    faz_gpu_ctx *ctx = new faz_gpu_ctx();
    ctx->device = device;
    // FAZ_INTERP_KERNEL selects the 3-D level kernels: "pass" = per-pass
    // launches only, "level" = the per-block level kernel without the
    // whole-tile fast path, "tile" = level kernels on every level whatever
    // its block count (the whole-tile kernel where a level qualifies)
    const char *kern = getenv("FAZ_INTERP_KERNEL");
    if (kern && strcmp(kern, "pass") == 0) ctx->fused3d = false;
    if (kern && (strcmp(kern, "pass") == 0 || strcmp(kern, "level") == 0)) ctx->tile3d = false;
    if (kern && strcmp(kern, "tile") == 0) ctx->fused_min_blocks = 1;
    const char *enc = getenv("FAZ_ENCODE");
    if (enc && strcmp(enc, "onepass") == 0) ctx->enc_onepass = true;
    // FAZ_DECODE_RUN=off: the decoder searches the anchors' run of zero codes
    // like any other part of the bitstream
    const char *drun = getenv("FAZ_DECODE_RUN");
    if (drun && strcmp(drun, "off") == 0) ctx->dec_run = false;
    // FAZ_DEVICE_BOUNDS=off: the REL compression reads the value range back
    // and sets the level bounds on the host
    const char *dbnd = getenv("FAZ_DEVICE_BOUNDS");
    if (dbnd && strcmp(dbnd, "off") == 0) ctx->dev_bounds = false;
    // the allocation-free encoder's scratch address on this device (its
    // entry point reports a failure here)
    (void) noalloc_init(device);
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu >= 8)
        ctx->tile_grid = (uint32_t) (ncu / 8 * 8);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return FAZ_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&ctx->ustream, hipStreamNonBlocking) != hipSuccess) {
        (void) hipStreamDestroy(ctx->stream);
        delete ctx;
        return FAZ_ERR_HIP;
    }
    for (auto &e: ctx->dec_ev) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void) hipStreamDestroy(ctx->ustream);
            (void) hipStreamDestroy(ctx->stream);
            delete ctx;
            return FAZ_ERR_HIP;
        }
    }
    for (auto &e: ctx->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return FAZ_ERR_HIP;
        }
    }
    *out = ctx;
    return FAZ_OK;
}

void faz_gpu_ctx_destroy(faz_gpu_ctx *ctx) {
    if (!ctx) return;
    (void) hipSetDevice(ctx->device);
    (void) hipStreamSynchronize(ctx->stream);
    ctx->codes.release();
    ctx->outval.release();
    ctx->unpred.release();
    ctx->work_in.release();
    ctx->work_out.release();
    ctx->lat[0].release();
    ctx->lat[1].release();
    ctx->freq.release();
    ctx->hist_pack.release();
    ctx->first.release();
    ctx->chunk_zeros.release();
    ctx->err_flag.release();
    ctx->minmax.release();
    ctx->words.release();
    ctx->chunk_base.release();
    ctx->scalars.release();
    ctx->prefix.release();
    ctx->chunk_bits.release();
    ctx->chunk_off.release();
    ctx->enc_code.release();
    ctx->tiles.release();
    ctx->lr_ind.release();
    ctx->lr_coef.release();
    ctx->lr_regc.release();
    ctx->lr_flag.release();
    ctx->lr_diag.release();
    ctx->lr_rank.release();
    ctx->lr_tab.release();
    ctx->gf_codes.release();
    ctx->gf_tab.release();
    ctx->gf_diag.release();
    ctx->gf_sel.release();
    ctx->gf_fields.release();
    ctx->gf_recon.release();
    ctx->gf_reg_raw.release();
    ctx->gf_reg_deq.release();
    ctx->gf_reg_code.release();
    ctx->enc_head.release();
    ctx->lvl.release();
    if (ctx->tune_cache_free) ctx->tune_cache_free(ctx->tune_cache);
    ctx->tune_cache = nullptr;
    for (auto &e: ctx->ev)
        if (e) (void) hipEventDestroy(e);
    for (auto &e: ctx->prof_ev) (void) hipEventDestroy(e);
    for (auto &e: ctx->dec_ev)
        if (e) (void) hipEventDestroy(e);
    (void) hipStreamDestroy(ctx->stream);
    if (ctx->ustream) (void) hipStreamDestroy(ctx->ustream);
    delete ctx;
}

const char *faz_gpu_last_error(const faz_gpu_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int faz_gpu_ctx_device(const faz_gpu_ctx *ctx) { return ctx ? ctx->device : -1; }

faz_status faz_gpu_last_timings(const faz_gpu_ctx *ctx, double *ms, int n) {
    if (!ctx || !ms) return FAZ_ERR_INVALID;
    for (int i = 0; i < n && i < 10; i++) ms[i] = ctx->timings[i];
    return FAZ_OK;
}

faz_status faz_gpu_range_f32(faz_gpu_ctx *ctx, const float *d_in, size_t n, float *h_min, float *h_max) {
    if (!ctx || !d_in || !n || !h_min || !h_max) return FAZ_ERR_INVALID;
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(ctx->minmax.ensure(kRangeScratch));
    launch_range(d_in, n, ctx->minmax.p, ctx->stream);
    HIPCHECK(hipGetLastError());
    return range_readback(ctx, h_min, h_max);
}

}  // extern "C"

// Kernel of each level, coarse to fine: 0 per-pass launches, 1 the per-block
// level kernel, 2 the whole-tile kernel (the decoder has no per-block level
// kernel: its levels are 0 or 2).  g3 holds the level launches, with the
// whole-tile levels' lattice plumbing: each reads the coarser lattice and
// writes its own either at full resolution in d_recon or, between two
// whole-tile levels, as a compact array in ctx->lat.
static hipError_t level_kernels(faz_gpu_ctx *ctx, const Plan &pl, const faz_interp_params *p, float *d_recon,
                                bool final_recon, bool decode, std::vector<Level3D> &g3, std::vector<int> &kind) {
    const bool fused = ctx->fused3d && pl.N == 3 && (uint32_t) p->block_size <= kFusedMaxBlock;
    // levels with few blocks run as per-pass launches, which spread a coarse
    // level's points over the whole device
    auto level_fused = [&](const LevelPlan &L) {
        if (!fused) return false;
        uint64_t nbw = 1;
        for (int a = 0; a < 3; a++) nbw *= pl.dims[a] == 1 ? 1 : (pl.dims[a] - 1 + L.bs - 1) / L.bs;
        return nbw >= ctx->fused_min_blocks;
    };
    // kernel per level (coarse to fine): 0 per-pass launches, 1 per-block
    // level kernel, 2 whole-tile kernel
    const size_t nlv = pl.lv.size();
    g3.assign(nlv, Level3D{});
    kind.assign(nlv, 0);
    for (size_t i = 0; i < nlv; i++) {
        const LevelPlan &L = pl.lv[i];
        if (!level_fused(L)) continue;
        Level3D &g = g3[i];
        memset(&g, 0, sizeof(g));
        for (int a = 0; a < 3; a++) {
            g.dims[a] = pl.dims[a];
            g.offs[a] = pl.offs[a];
            g.nb[a] = L.nb[a];
            g.nbw[a] = pl.dims[a] == 1 ? 1 : (uint32_t) ((pl.dims[a] - 1 + L.bs - 1) / L.bs);
            g.p[a] = pl.perms[L.dir][a];
        }
        g.s = (uint32_t) ((uint64_t) 1 << (L.level - 1));
        g.bs = (uint32_t) L.bs;
        g.linear = L.linear;
        g.eb = L.eb;
        g.recip = L.recip;
        g.ebf = float_floor(L.eb);
        g.radius = p->radius;
        g.level_base = L.base;
        g.block_prefix = ctx->prefix.p + L.prefix_at;
        g.code_limit = pl.total_codes;
        kind[i] = ctx->tile3d && tile3d_eligible(g) ? 2 : (decode ? 0 : 1);
    }
    // Whole-tile levels read the coarser lattice and write their own either
    // at full resolution in d_recon or, between two whole-tile levels, as a
    // compact array (lattice index c at c0 m1 m2 + c1 m2 + c2).
    {
        int cur = -1;  // lat[] buffer with the previous level's compact lattice
        for (size_t i = 0; i < nlv; i++) {
            if (kind[i] != 2) {
                cur = -1;
                continue;
            }
            Level3D &g = g3[i];
            const uint64_t s = g.s;
            auto compact = [&](uint64_t step, uint64_t *str, uint64_t &count) {
                uint64_t m[3];
                for (int a = 0; a < 3; a++) m[a] = (pl.dims[a] - 1) / step + 1;
                str[0] = m[1] * m[2];
                str[1] = m[2];
                str[2] = 1;
                count = m[0] * m[1] * m[2];
            };
            uint64_t count = 0;
            if (cur < 0) {
                g.csrc = d_recon;
                for (int a = 0; a < 3; a++) g.cstr[a] = 2 * s * pl.offs[a];
            } else {
                g.csrc = ctx->lat[cur].p;
                compact(2 * s, g.cstr, count);
            }
            const bool last = i + 1 == nlv;
            const bool next_tile = !last && kind[i + 1] == 2;
            if (next_tile) {
                const int nb = cur < 0 ? 0 : 1 - cur;
                compact(s, g.dstr, count);
                {
                    const hipError_t e = ctx->lat[nb].ensure(count);
                    if (e != hipSuccess) return e;
                }
                g.dst = ctx->lat[nb].p;
                cur = nb;
            } else {
                g.dst = last && !final_recon ? nullptr : d_recon;
                for (int a = 0; a < 3; a++) g.dstr[a] = s * pl.offs[a];
                cur = -1;
            }
        }
    }
    return hipSuccess;
}

// The level schedule of a plan as factors of the base bound (the arithmetic
// of make_plan, SIC:384-422), for launch_level_bounds.
static LevelScale level_scale(const Plan &pl, const faz_interp_params *p) {
    LevelScale sc;
    memset(&sc, 0, sizeof(sc));
    sc.n = (int) std::min<size_t>(pl.lv.size(), kMaxLevels);
    for (int i = 0; i < sc.n; i++) {
        const unsigned level = pl.lv[i].level;
        if (p->alpha < 0) {
            sc.div[i] = 0;
            sc.f[i] = level >= 3 ? 0.5 : 1.0;
        } else if (p->alpha >= 1) {
            double r = std::pow(p->alpha, level - 1);
            if (r > p->beta) r = p->beta;
            sc.div[i] = 1;
            sc.f[i] = r;
        } else {
            double r = 1 - (level - 1) * p->alpha;
            if (r < p->beta) r = p->beta;
            sc.div[i] = 0;
            sc.f[i] = r;
        }
    }
    return sc;
}

// Interpolation + quantization of one field.  d_recon is a full-size working
// array that receives every reconstructed point when final_recon is set;
// otherwise it only carries the coarse levels (the compression itself needs
// no reconstruction of the finest level).
// absmax: max |x| of the input when the caller knows it, else negative.
// enqueue: device work only, for stream capture: no host synchronisation or
// transfer and no allocation (a run with the same shape before has sized every
// buffer and uploaded the block prefix table), outliers gathered into a
// buffer of one slot per code, results read by the caller after the stream
// drains (ctx->scalars[0] = outlier count, ctx->err_flag).
// defer: no host synchronisation either, but allocations and the prefix
// upload are allowed; the caller reads the outlier count and the error flag
// after its own wait and computes timings[1..2] from ev[0..2].
// d_lvl: the bounds of the first point (entry 0) and of every level (1 + i)
// on the device (launch_level_bounds); p->eb is then not used.
static int interp_compress(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_interp_params *p,
                           faz_interp_out *out, bool final_recon, double absmax, bool enqueue = false,
                           bool defer = false, const LevelEB *d_lvl = nullptr) {
    int rc = FAZ_OK;
    Plan &pl = cached_plan(ctx, p, rc);
    if (rc) return rc;
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const uint64_t prefix_key = pl.prefix.size() * 0x9E3779B97F4A7C15ull ^
                                std::accumulate(pl.prefix.begin(), pl.prefix.end(), (uint64_t) pl.total_codes);
    if (enqueue) {
        if (ctx->prefix_key != prefix_key || ctx->prefix.cap < pl.prefix.size() || ctx->codes.cap < pl.total_codes ||
            ctx->unpred.cap < pl.total_codes || !(absmax >= 0))
            CTX_FAIL(FAZ_ERR_INVALID, "internal: enqueue without a prepared run of this shape");
    }
    HIPCHECK(ctx->codes.ensure(pl.total_codes));
    HIPCHECK(ctx->outval.ensure(pl.total_codes));
    HIPCHECK(ctx->err_flag.ensure(1));
    HIPCHECK(ctx->prefix.ensure(pl.prefix.size()));
    HIPCHECK(ctx->scalars.ensure(4));
    const uint32_t nchunks = (uint32_t) ((pl.total_codes + kHistChunk - 1) / kHistChunk);
    HIPCHECK(ctx->chunk_zeros.ensure(nchunks));
    HIPCHECK(ctx->chunk_base.ensure(nchunks));
    HIPCHECK(ctx->freq.ensure(kHistSlots));
    HIPCHECK(ctx->first.ensure(65536));
    HIPCHECK(ctx->tiles.ensure(scan_tiles(nchunks)));
    if (!enqueue && (ctx->prefix_key != prefix_key || ctx->prefix_at != ctx->prefix.p)) {
        HIPCHECK(hipMemcpyAsync(ctx->prefix.p, pl.prefix.data(), pl.prefix.size() * 8, hipMemcpyHostToDevice, st));
        ctx->prefix_key = prefix_key;
        ctx->prefix_at = ctx->prefix.p;
    }
    if (!ctx->err_cleared) launch_fill_u32(ctx->err_flag.p, 1, 0u, st);
    ctx->err_cleared = false;

    // phase events: always on the synchronous path, on request when deferred
    const bool phase_ev = !enqueue && (!defer || ctx->phase_events);
    if (phase_ev) HIPCHECK(hipEventRecord(ctx->ev[0], st));
    if (pl.anchor) {
        launch_anchor_grid(d_in, d_recon, ctx->codes.p, ctx->outval.p, pl.N, pl.dims, pl.offs,
                           (uint32_t) p->max_step, pl.anchor_count, st);
    } else {
        double eb0 = p->eb;
        launch_first_point(d_in, d_recon, ctx->codes.p, ctx->outval.p, eb0, 1.0 / eb0, p->radius, ctx->err_flag.p, st,
                           d_lvl);
    }
    size_t npass = 0;
    for (auto &L: pl.lv) npass += L.passes.size();
    const bool profile = ctx->profile && !enqueue;
    if (profile && ctx->prof_ev.size() < 2 * npass) {
        size_t have = ctx->prof_ev.size();
        ctx->prof_ev.resize(2 * npass);
        for (size_t i = have; i < ctx->prof_ev.size(); i++) HIPCHECK(hipEventCreate(&ctx->prof_ev[i]));
    }
    const size_t nlv = pl.lv.size();
    std::vector<Level3D> g3;
    std::vector<int> kind;
    HIPCHECK(level_kernels(ctx, pl, p, d_recon, final_recon, false, g3, kind));
    // The bounded quantizer needs |x - pred| / eb < 2^30 on the level.  Every
    // prediction is a stencil of reconstructed values of magnitude at most
    // max|x| + eb with coefficient sum at most 28/8 (quad_3), so
    // |x - pred| <= 4.5 max|x| + 3.5 eb; half of the range is kept as margin
    // for rounding.  A non-finite or huge max |x| keeps the exact quantizer.
    if (d_lvl) {
        for (size_t i = 0; i < nlv; i++) g3[i].ebp = d_lvl + 1 + i;
    } else if (std::find(kind.begin(), kind.end(), 2) != kind.end()) {
        if (!(absmax >= 0)) {  // never in enqueue mode (checked above)
            float mn = 0, mx = 0;
            int rr = faz_gpu_range_f32(ctx, d_in, pl.num, &mn, &mx);
            if (rr) return rr;
            absmax = std::max(std::fabs((double) mn), std::fabs((double) mx));
        }
        for (size_t i = 0; i < nlv; i++) {
            if (kind[i] != 2) continue;
            const double v = 4.5 * absmax * g3[i].recip + 8.0;
            g3[i].bounded = std::isfinite(absmax) && absmax < 1e37 && v < 1073741824.0;
        }
    }
    // with level launches in the plan the profile times those only: the pass
    // launches of the coarse levels are a few microseconds each, and event
    // records around them would slow their enqueue below the device's pace.
    // A run of consecutive whole-tile launches is timed as one span (an event
    // record between two launches idles the stream for a few microseconds).
    bool any_level = false;
    for (size_t i = 0; i < nlv; i++) any_level |= kind[i] != 0;
    const bool prof_pass = profile && !any_level;
    if (profile) ctx->prof_pending.clear();
    size_t k = 0;
    for (size_t i = 0; i < nlv; i++) {
        LevelPlan &L = pl.lv[i];
        if (kind[i]) {
            const bool opens = kind[i] != 2 || i == 0 || kind[i - 1] != 2;
            const bool closes = kind[i] != 2 || i + 1 == nlv || kind[i + 1] != 2;
            if (profile && opens) {
                HIPCHECK(hipEventRecord(ctx->prof_ev[2 * k], st));
                ctx->prof_pending.push_back({kind[i] == 2 ? 0 : 2, (uint32_t) k, 0, 0.0, 0});
            }
            if (kind[i] == 2)
                launch_interp_tile3d(g3[i], d_in, ctx->codes.p, ctx->outval.p, ctx->err_flag.p, ctx->tile_grid, st);
            else
                launch_interp_level3d(g3[i], d_in, d_recon, ctx->codes.p, ctx->outval.p, ctx->err_flag.p, st);
            if (profile) {
                // read each original and write its code; the reconstruction
                // traffic is the coarser lattice read and this level's
                // lattice written, where that happens
                uint64_t lat = 1, coarse = 1;
                for (int a = 0; a < 3; a++) {
                    lat *= (pl.dims[a] - 1) / g3[i].s + 1;
                    coarse *= (pl.dims[a] - 1) / (2 * g3[i].s) + 1;
                }
                const bool writes = kind[i] == 1 || g3[i].dst != nullptr;
                faz_gpu_ctx::ProfRec &rec = ctx->prof_pending.back();
                rec.points += L.codes;
                rec.bytes += 6.0 * L.codes + 4.0 * coarse + (writes ? 4.0 * lat : 0.0);
                rec.launches++;
                if (closes) {
                    HIPCHECK(hipEventRecord(ctx->prof_ev[2 * k + 1], st));
                    k++;
                }
            }
            continue;
        }
        for (auto &d: L.passes) {
            d.block_prefix = ctx->prefix.p + L.prefix_at;
            d.ebp = d_lvl ? d_lvl + 1 + i : nullptr;
            if (prof_pass) HIPCHECK(hipEventRecord(ctx->prof_ev[2 * k], st));
            launch_interp_pass(d, d_in, d_recon, ctx->codes.p, ctx->outval.p, nullptr, ctx->err_flag.p, st);
            if (prof_pass) {
                HIPCHECK(hipEventRecord(ctx->prof_ev[2 * k + 1], st));
                // original in, reconstruction and code out
                if (d.total) ctx->prof_pending.push_back({1, (uint32_t) k, d.total, 10.0 * d.total, 1});
                k++;
            }
        }
    }
    HIPCHECK(hipGetLastError());
    if (phase_ev) HIPCHECK(hipEventRecord(ctx->ev[1], st));
    // histogram, first positions and outlier compaction
    const bool cleared = ctx->hist_cleared && ctx->freq.cap >= kHistSlots && ctx->first.cap >= 65536;
    ctx->hist_cleared = false;
    launch_hist(ctx->codes.p, pl.total_codes, ctx->freq.p, ctx->first.p, ctx->chunk_zeros.p, p->radius, st, cleared);
    launch_zero_scan(ctx->chunk_zeros.p, ctx->chunk_base.p, nchunks, ctx->tiles.p, ctx->scalars.p, st);
    HIPCHECK(hipGetLastError());
    uint64_t n_unpred = 0;
    if (enqueue || defer) {
        // outliers gathered into one slot per code: no wait for their count
        // (deferred: the caller enqueues the gather after its readback, so
        // it runs while the host builds the tree)
        if (defer) HIPCHECK(ctx->unpred.ensure(pl.total_codes));
        if (enqueue)
            launch_unpred_gather(ctx->codes.p, pl.total_codes, ctx->outval.p, ctx->chunk_zeros.p, ctx->chunk_base.p,
                                 ctx->unpred.p, st);
        HIPCHECK(hipGetLastError());
        if (enqueue) {
            out->d_codes = ctx->codes.p;
            out->n_codes = pl.total_codes;
            out->d_unpred = ctx->unpred.p;
            out->n_unpred = 0;
            out->loss = 0;
            memset(out->level_counts, 0, sizeof(out->level_counts));
            out->levels = pl.levels;
            return FAZ_OK;
        }
        if (phase_ev) HIPCHECK(hipEventRecord(ctx->ev[2], st));
    } else {
        uint32_t errflag = 0;
        HIPCHECK(hipMemcpyAsync(&n_unpred, ctx->scalars.p, 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(&errflag, ctx->err_flag.p, 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (errflag & 2u) CTX_FAIL(FAZ_ERR_INVALID, "internal: a block segment fell outside the code stream");
        if (errflag) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "a quantization code exceeded 16 bits");
        HIPCHECK(ctx->unpred.ensure(n_unpred));
        launch_unpred_gather(ctx->codes.p, pl.total_codes, ctx->outval.p, ctx->chunk_zeros.p, ctx->chunk_base.p,
                             ctx->unpred.p, st);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(ctx->ev[2], st));
        HIPCHECK(hipStreamSynchronize(st));
        ctx->timings[1] = elapsed(ctx->ev[0], ctx->ev[1]);
        ctx->timings[2] = elapsed(ctx->ev[1], ctx->ev[2]);
        ctx->timings[6] = ctx->timings[1];
    }
    // bucket 0: whole-tile level launches, 1: per-pass launches, 2: per-block
    // level launches; the event times are read once the stream has drained
    if (profile && !defer) flush_profile(ctx);  // the stream was drained above
    ctx->hist_codes = ctx->codes.p;
    ctx->hist_n = pl.total_codes;

    out->d_codes = ctx->codes.p;
    out->n_codes = pl.total_codes;
    out->d_unpred = ctx->unpred.p;
    out->n_unpred = n_unpred;
    out->loss = 0;
    memset(out->level_counts, 0, sizeof(out->level_counts));
    out->levels = pl.levels;
    return FAZ_OK;
}

extern "C" {

faz_status faz_gpu_interp_compress_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon,
                                       const faz_interp_params *p, faz_interp_out *out) {
    if (!ctx || !d_in || !d_recon || !p || !out) return FAZ_ERR_INVALID;
    return interp_compress(ctx, d_in, d_recon, p, out, true, -1.0);
}

faz_status faz_gpu_huffman_hist(faz_gpu_ctx *ctx, const uint16_t *d_codes, size_t n, uint32_t *h_freq,
                                uint32_t *h_first) {
    if (!ctx || !d_codes || !n || !h_freq || !h_first) return FAZ_ERR_INVALID;
    if (n >= (1ull << 32)) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "more than 2^32-1 codes");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // the histogram the last compression left is reused only for the
    // context's own code buffer, which nothing else writes
    if (!(d_codes == ctx->codes.p && ctx->hist_codes == d_codes && ctx->hist_n == n)) {
        const uint32_t nchunks = (uint32_t) ((n + kHistChunk - 1) / kHistChunk);
        HIPCHECK(ctx->chunk_zeros.ensure(nchunks));
        HIPCHECK(ctx->freq.ensure(kHistSlots));
        HIPCHECK(ctx->first.ensure(65536));
        launch_hist(d_codes, n, ctx->freq.p, ctx->first.p, ctx->chunk_zeros.p, 32768, st);
        HIPCHECK(hipGetLastError());
        ctx->hist_codes = d_codes == ctx->codes.p ? d_codes : nullptr;
        ctx->hist_n = n;
    }
    HIPCHECK(hipMemcpyAsync(h_freq, ctx->freq.p, 65536 * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(h_first, ctx->first.p, 65536 * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return FAZ_OK;
}

faz_status faz_huffman_build(const uint32_t *h_freq, const uint32_t *h_first, int32_t *offset, int32_t *n_states,
                             uint64_t *code, uint8_t *len, uint8_t *tree, size_t tree_cap, size_t *tree_size) {
    if (!h_freq || !h_first || !offset || !n_states || !code || !len || !tree_size) return FAZ_ERR_INVALID;
    HuffTable t;
    if (build_huffman(h_freq, h_first, 65536, t)) return FAZ_ERR_INVALID;
    *offset = t.offset;
    *n_states = t.n_states;
    memcpy(code, t.code.data(), t.code.size() * 8);
    memcpy(len, t.len.data(), t.len.size());
    *tree_size = t.tree.size();
    if (!tree || t.tree.size() > tree_cap) return FAZ_ERR_CAPACITY;
    memcpy(tree, t.tree.data(), t.tree.size());
    return FAZ_OK;
}

}  // extern "C"

// Packed code table of t in ctx->h_tab (one entry per symbol: codeword
// right-aligned above an 8-bit length) and the LDS window of the encoder.
// The packed code table (ctx->h_tab) and its LDS window.  Only the window
// (the present symbols other than 0) and symbol 0's entry are filled when the
// window fits: the kernels then read nothing else, and *upload_lo/_n name the
// entries to copy.  *maxbits = the longest codeword.
static int encode_table(faz_gpu_ctx *ctx, const faz_huff_table *t, EncodeWindow &ew, uint64_t *maxbits,
                        size_t *upload_lo, size_t *upload_n) {
    const int32_t ns = t->n_states;
    const int64_t zi = (t->offset <= 0 && -t->offset < ns) ? -t->offset : -1;  // symbol 0's index
    int64_t lo = -1, hi = -1;  // table indices of the present symbols other than 0
    uint32_t maxlen = 0;
    int32_t i = 0;
    for (; i + 8 <= ns; i += 8) {  // most lengths are zero: eight at a time
        uint64_t w;
        memcpy(&w, t->len + i, 8);
        if (!w) continue;
        for (int32_t j = i; j < i + 8; j++) {
            const uint32_t L = t->len[j];
            if (!L) continue;
            maxlen = std::max(maxlen, L);
            if (j != zi) {
                if (lo < 0) lo = j;
                hi = j;
            }
        }
    }
    for (; i < ns; i++) {
        const uint32_t L = t->len[i];
        if (!L) continue;
        maxlen = std::max(maxlen, L);
        if (i != zi) {
            if (lo < 0) lo = i;
            hi = i;
        }
    }
    if (maxlen > 56) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "Huffman codeword of %u bits (at most 56 supported)", maxlen);
    *maxbits = maxlen;
    const bool narrow = maxlen <= 32;
    auto pack = [&](int64_t k) { return encode_entry(t->code[k], t->len[k], narrow); };
    ctx->h_tab.ensure((size_t) ns);
    ew = EncodeWindow{0, 0, 0};
    ew.max_len = maxlen;
    if (lo >= 0 && hi - lo + 1 <= (int64_t) encode_lds_entries()) {
        for (int64_t k = lo; k <= hi; k++) ctx->h_tab[k] = pack(k);
        ew.win = (uint32_t) lo;
        ew.nwin = (uint32_t) (hi - lo + 1);
        ew.e0 = zi >= 0 ? pack(zi) : 0;
        *upload_lo = (size_t) lo;
        *upload_n = (size_t) (hi - lo + 1);
    } else {
        for (int64_t k = 0; k < ns; k++) ctx->h_tab[k] = pack(k);
        *upload_lo = 0;
        *upload_n = (size_t) ns;
    }
    return FAZ_OK;
}

__global__ void zero_words_kernel(uint32_t *__restrict__ w, const uint64_t *__restrict__ bits, size_t cap) {
    const size_t words = std::min<size_t>((size_t) ((bits[0] + 31) / 32), cap);
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t) gridDim.x * blockDim.x)
        w[i] = 0u;
}

// Huffman encode of n codes into ctx->words (sized for the longest codeword
// on every code) without a host synchronisation: the bit total stays in
// ctx->scalars[1].  ctx->h_tab must stay untouched until the stream drains.
static int encode_enqueue(faz_gpu_ctx *ctx, const uint16_t *d_codes, size_t n, const faz_huff_table *t) {
    if (n >= (1ull << 32)) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "more than 2^32-1 codes");
    hipStream_t st = ctx->stream;
    const uint32_t nchunks = (uint32_t) ((n + kEncChunk - 1) / kEncChunk);
    EncodeWindow ew;
    uint64_t maxbits = 0;
    size_t up_lo = 0, up_n = 0;
    int rc = encode_table(ctx, t, ew, &maxbits, &up_lo, &up_n);
    if (rc) return rc;
    const size_t cap = (size_t) ((n * maxbits + 31) / 32 + 2);
    HIPCHECK(ctx->words.ensure(cap));
    HIPCHECK(ctx->enc_code.ensure((size_t) t->n_states));
    HIPCHECK(ctx->chunk_bits.ensure(nchunks));
    HIPCHECK(ctx->chunk_off.ensure(nchunks));
    HIPCHECK(ctx->scalars.ensure(4));
    HIPCHECK(ctx->tiles.ensure(scan_tiles(nchunks)));
    HIPCHECK(hipMemcpyAsync(ctx->enc_code.p + up_lo, ctx->h_tab.data() + up_lo, up_n * 8, hipMemcpyHostToDevice, st));
    if (ctx->enc_onepass && maxbits <= 32) {
        // chunk status words start at zero
        HIPCHECK(ctx->enc_head.ensure(nchunks));
        HIPCHECK(ctx->err_flag.ensure(1));
        launch_fill_u32(reinterpret_cast<uint32_t *>(ctx->chunk_bits.p), 2 * (size_t) nchunks, 0u, st);
        launch_encode_onepass(d_codes, n, ctx->enc_code.p, ew, t->offset, ctx->words.p, ctx->chunk_bits.p,
                              ctx->chunk_off.p, ctx->enc_head.p, ctx->scalars.p + 1, ctx->err_flag.p, st);
        HIPCHECK(hipGetLastError());
        return FAZ_OK;
    }
    launch_encode_sizes(d_codes, n, ctx->enc_code.p, ew, t->offset, ctx->chunk_bits.p, st);
    // codewords of at most 32 bits: the packer ORs only into the words the
    // offsets fall in, which the scan zeroes; longer ones OR every word
    if (maxbits <= 32) {
        launch_encode_offsets(ctx->chunk_bits.p, ctx->chunk_off.p, nchunks, ctx->tiles.p, ctx->scalars.p + 1,
                              ctx->words.p, st);
    } else {
        launch_scan_u64(ctx->chunk_bits.p, ctx->chunk_off.p, nchunks, ctx->tiles.p, ctx->scalars.p + 1, st);
        hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned) std::min<size_t>((cap + 255) / 256, 8192)), dim3(256), 0,
                           st, ctx->words.p, ctx->scalars.p + 1, ctx->words.cap);
    }
    launch_encode_pack(d_codes, n, ctx->enc_code.p, ew, t->offset, ctx->chunk_off.p, ctx->words.p, st);
    HIPCHECK(hipGetLastError());
    return FAZ_OK;
}

extern "C" {

faz_status faz_gpu_huffman_encode(faz_gpu_ctx *ctx, const uint16_t *d_codes, size_t n, const faz_huff_table *t,
                                  uint8_t *d_bits, size_t cap_bytes, size_t *nbytes) {
    if (!ctx || !d_codes || !n || !t || !nbytes || !t->code || !t->len || t->n_states <= 0) return FAZ_ERR_INVALID;
    if (n >= (1ull << 32)) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "more than 2^32-1 codes");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const uint32_t nchunks = (uint32_t) ((n + kEncChunk - 1) / kEncChunk);
    // one packed entry per symbol (encode_entry)
    ctx->h_tab.ensure((size_t) t->n_states);
    int64_t lo = -1, hi = -1;  // table indices of the present symbols other than 0
    uint32_t maxlen = 0;
    for (int32_t i = 0; i < t->n_states; i++) {
        const uint32_t L = t->len[i];
        if (L > 56) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "Huffman codeword of %u bits (at most 56 supported)", L);
        maxlen = std::max(maxlen, L);
        if (L && i + t->offset != 0) {
            if (lo < 0) lo = i;
            hi = i;
        }
    }
    for (int32_t i = 0; i < t->n_states; i++) ctx->h_tab[i] = encode_entry(t->code[i], t->len[i], maxlen <= 32);
    EncodeWindow ew{0, 0, 0};
    ew.max_len = maxlen;
    if (lo >= 0 && hi - lo + 1 <= (int64_t) encode_lds_entries()) {
        ew.win = (uint32_t) lo;
        ew.nwin = (uint32_t) (hi - lo + 1);
        ew.e0 = (t->offset <= 0 && -t->offset < t->n_states) ? ctx->h_tab[-t->offset] : 0;
    }
    HIPCHECK(ctx->enc_code.ensure((size_t) t->n_states));
    HIPCHECK(ctx->chunk_bits.ensure(nchunks));
    HIPCHECK(ctx->chunk_off.ensure(nchunks));
    HIPCHECK(ctx->scalars.ensure(4));
    HIPCHECK(hipMemcpyAsync(ctx->enc_code.p, ctx->h_tab.data(), (size_t) t->n_states * 8, hipMemcpyHostToDevice, st));
    launch_encode_sizes(d_codes, n, ctx->enc_code.p, ew, t->offset, ctx->chunk_bits.p, st);
    HIPCHECK(ctx->tiles.ensure(scan_tiles(nchunks)));
    launch_scan_u64(ctx->chunk_bits.p, ctx->chunk_off.p, nchunks, ctx->tiles.p, ctx->scalars.p + 1, st);
    HIPCHECK(hipGetLastError());
    uint64_t bits = 0;
    HIPCHECK(hipMemcpyAsync(&bits, ctx->scalars.p + 1, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const size_t bytes = (size_t) ((bits + 7) / 8);
    const size_t words = (size_t) ((bits + 31) / 32);
    *nbytes = bytes;
    if (!d_bits || cap_bytes < words * 4) CTX_FAIL(FAZ_ERR_CAPACITY, "bit buffer needs %zu bytes", words * 4);
    uint32_t *w = reinterpret_cast<uint32_t *>(d_bits);
    if (((uintptr_t) d_bits & 3) != 0) CTX_FAIL(FAZ_ERR_INVALID, "bit buffer must be 4-byte aligned");
    launch_fill_u32(w, words, 0u, st);
    launch_encode_pack(d_codes, n, ctx->enc_code.p, ew, t->offset, ctx->chunk_off.p, w, st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    return FAZ_OK;
}

}  // extern "C"

float *faz_ctx_work_in(faz_gpu_ctx *ctx, size_t n) {
    if (hipSetDevice(ctx->device) != hipSuccess || ctx->work_in.ensure(n) != hipSuccess) return nullptr;
    return ctx->work_in.p;
}

hipStream_t faz_ctx_stream(faz_gpu_ctx *ctx) { return ctx->stream; }

void faz_ctx_set_error(faz_gpu_ctx *ctx, const char *msg) { ctx->err = msg; }

// Device part of the engine buffer: interpolation + quantization, histogram
// and outliers, host tree, encode.  Results stay in the context until fetched.
// compress_device's answer when a whole-tile level flagged that its
// device-computed bound needs the exact quantizer (not a faz_status).
constexpr int kRetryHostBounds = 1000;

static int compress_device(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_interp_params *p,
                           double absmax = -1.0, const LevelEB *d_lvl = nullptr) {
    if (!ctx || !d_in || !p) return FAZ_ERR_INVALID;
    ctx->pending.valid = false;
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    float *recon = d_recon;
    if (!recon) {
        size_t num = 1;
        for (int i = 0; i < p->N && i < 4; i++) num *= p->dims[i];
        HIPCHECK(ctx->work_out.ensure(num));
        recon = ctx->work_out.p;
    }
    // one wait after the histogram (counts, first positions, outlier count
    // and error flag through pinned memory), one after the encode
    faz_interp_out io;
    int rc = interp_compress(ctx, d_in, recon, p, &io, d_recon != nullptr, absmax, false, true, d_lvl);
    if (rc) return rc;
    // one readback: the histogram's window around the radius, symbol 0, the
    // outside-window flag, the error flag and the outlier count
    // (the pack kernel writes it into the coherent pinned mailbox directly,
    // or into device memory and a copy brings it back), and the base bound
    // computed on the device with it
    uint32_t *hf = ctx->hp_hist.data(), *hs = ctx->hp_hist.data() + 2 * 65536;
    const bool mail = ctx->mailbox.pinned;
    const uint32_t *pk = mail ? ctx->mailbox.data() : hf;
    if (mail) {
        launch_hist_pack(ctx->freq.p, ctx->first.p, p->radius, ctx->scalars.p, ctx->err_flag.p, ctx->mailbox.data(),
                         st, d_lvl ? &d_lvl->eb : nullptr);
    } else {
        HIPCHECK(ctx->hist_pack.ensure(kHistPackWords));
        launch_hist_pack(ctx->freq.p, ctx->first.p, p->radius, ctx->scalars.p, ctx->err_flag.p, ctx->hist_pack.p, st);
        HIPCHECK(hipMemcpyAsync(hf, ctx->hist_pack.p, kHistPackWords * 4, hipMemcpyDeviceToHost, st));
        if (d_lvl) HIPCHECK(hipMemcpyAsync(hf + kHistPackWords, &d_lvl->eb, 8, hipMemcpyDeviceToHost, st));
    }
    HIPCHECK(hipEventRecord(ctx->ev[3], st));
    // the outlier gather runs while the host waits for the counts and builds the tree
    launch_unpred_gather(io.d_codes, io.n_codes, ctx->outval.p, ctx->chunk_zeros.p, ctx->chunk_base.p, ctx->unpred.p,
                         st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(spin_wait_event(ctx->ev[3]));
    const auto th0 = std::chrono::steady_clock::now();
    double eb = p->eb;
    if (d_lvl) {
        memcpy(&eb, pk + kHistPackWords, 8);
        if (!(eb > 0)) CTX_FAIL(FAZ_ERR_INVALID, "error bound must be positive");
    }
    const uint32_t errflag = pk[kHistPackFreq0 + 3];
    // a whole-tile level whose device bound needs the exact quantizer: the
    // caller compresses again with host bounds
    if (errflag & 8u) return kRetryHostBounds;
    if (errflag & 2u) CTX_FAIL(FAZ_ERR_INVALID, "internal: a block segment fell outside the code stream");
    if (errflag) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "a quantization code exceeded 16 bits");
    uint64_t n_unpred;
    memcpy(&n_unpred, pk + kHistPackFreq0 + 4, 8);
    PendingRaw &pr = ctx->pending;
    int brc;
    // a symbol outside the window: the full arrays instead of the window
    const bool outside = pk[kHistPackFreq0 + 2] != 0;
    const uint32_t freq0 = pk[kHistPackFreq0];
    // the code table now, the tree bytes once the encode is enqueued
    if (!outside) {
        brc = build_huffman_window_codes(pk, pk + kHistWin, hist_window_lo(p->radius), kHistWin,
                                         pk[kHistPackFreq0], pk[kHistPackFreq0 + 1], pr.tab);
    } else {  // a symbol outside the window: the whole arrays
        HIPCHECK(hipMemcpyAsync(hf, ctx->freq.p, 65536 * 4, hipMemcpyDeviceToHost, ctx->ustream));
        HIPCHECK(hipMemcpyAsync(hf + 65536, ctx->first.p, 65536 * 4, hipMemcpyDeviceToHost, ctx->ustream));
        HIPCHECK(hipStreamSynchronize(ctx->ustream));
        brc = build_huffman_codes(hf, hf + 65536, 65536, pr.tab);
    }
    if (brc) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "Huffman tree could not be built (codes longer than 64 bits)");
    // the stream's bit total from the histogram and the code lengths: the
    // encoder's own total must agree
    uint64_t want_bits = 0;
    {
        auto add = [&](uint32_t sym, uint64_t cnt) {
            const int64_t k = (int64_t) sym - pr.tab.offset;
            if (cnt && k >= 0 && k < pr.tab.n_states) want_bits += cnt * pr.tab.len[(size_t) k];
        };
        if (!outside) {
            const uint32_t lo = (uint32_t) hist_window_lo(p->radius);
            for (uint32_t i = 0; i < (uint32_t) kHistWin; i++) add(lo + i, pk[i]);
            add(0, freq0);
        } else {
            for (uint32_t i = 0; i < 65536; i++) add(i, hf[i]);
        }
    }
    const auto th1 = std::chrono::steady_clock::now();
    if (ctx->phase_events) HIPCHECK(hipEventRecord(ctx->ev[4], st));
    faz_huff_table ht{pr.tab.offset, pr.tab.n_states, pr.tab.code.data(), pr.tab.len.data()};
    rc = encode_enqueue(ctx, io.d_codes, io.n_codes, &ht);
    if (rc) return rc;
    const auto th2 = std::chrono::steady_clock::now();
    ctx->timings[8] = std::chrono::duration<double, std::milli>(th1 - th0).count();
    ctx->timings[9] = std::chrono::duration<double, std::milli>(th2 - th1).count();
    finish_huffman_tree(pr.tab);  // while the encode runs
    uint32_t *rep = mail ? ctx->mailbox.data() + faz_gpu_ctx::kMailReport : hs + 4;
    if (mail) {
        launch_report(ctx->scalars.p + 1, ctx->err_flag.p, rep, st);
    } else {
        HIPCHECK(hipMemcpyAsync(rep, ctx->scalars.p + 1, 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(rep + 2, ctx->err_flag.p, 4, hipMemcpyDeviceToHost, st));
    }
    if (ctx->phase_events) HIPCHECK(hipEventRecord(ctx->ev[5], st));
    HIPCHECK(spin_wait(st));
    flush_profile(ctx);
    uint64_t bits;
    memcpy(&bits, rep, 8);
    if (rep[2] & 4u) CTX_FAIL(FAZ_ERR_HIP, "internal: an encode look-back wait ran out");
    if (bits != want_bits)
        CTX_FAIL(FAZ_ERR_HIP, "internal: encoded %llu bits, the histogram gives %llu", (unsigned long long) bits,
                 (unsigned long long) want_bits);
    const size_t nbytes = (size_t) ((bits + 7) / 8);
    io.n_unpred = n_unpred;
    if (ctx->phase_events) {
        ctx->timings[1] = elapsed(ctx->ev[0], ctx->ev[1]);
        ctx->timings[2] = elapsed(ctx->ev[1], ctx->ev[2]);
        ctx->timings[6] = ctx->timings[1];
        ctx->timings[3] = elapsed(ctx->ev[2], ctx->ev[4]);
        ctx->timings[4] = elapsed(ctx->ev[4], ctx->ev[5]);
        ctx->timings[5] = elapsed(ctx->ev[0], ctx->ev[5]);
    } else {
        for (int i = 0; i < 7; i++) ctx->timings[i] = 0;
    }
    // engine buffer header (SIC:489-512)
    EngineHeader &h = pr.h;
    h.N = p->N;
    for (int i = 0; i < p->N; i++) h.dims[i] = p->dims[i];
    h.block_size = (unsigned) p->block_size;
    h.interp_algo = p->interp_algo;
    h.direction = p->direction;
    h.alpha = p->alpha;
    h.beta = p->beta;
    h.max_step = p->max_step;
    h.levelwise = p->levelwise;
    h.ops.assign(p->ops, p->ops + p->levelwise);
    h.dirs.assign(p->dirs, p->dirs + p->levelwise);
    pr.eb = eb;
    pr.radius = p->radius;
    pr.n_unpred = io.n_unpred;
    pr.nbytes = nbytes;
    pr.valid = true;
    return FAZ_OK;
}

static size_t pending_size(const faz_gpu_ctx *ctx) {
    const PendingRaw &pr = ctx->pending;
    return engine_raw_size(pr.h, pr.n_unpred, pr.tab.tree.size(), pr.nbytes);
}

// Host copy of the pending engine buffer (SIC:489-518, Q:185-201, HUF:119-227).
static int fetch_raw(faz_gpu_ctx *ctx, uint8_t *out) {
    const PendingRaw &pr = ctx->pending;
    hipStream_t st = ctx->stream;
    uint8_t *w = out;
    write_engine_header(pr.h, w);
    write_quantizer_head(pr.eb, pr.radius, pr.n_unpred, w);
    HIPCHECK(hipMemcpyAsync(w, ctx->unpred.p, pr.n_unpred * sizeof(float), hipMemcpyDeviceToHost, st));
    w += pr.n_unpred * sizeof(float);
    memcpy(w, pr.tab.tree.data(), pr.tab.tree.size());
    w += pr.tab.tree.size();
    uint64_t nb64 = pr.nbytes;
    memcpy(w, &nb64, 8);
    w += 8;
    HIPCHECK(hipMemcpyAsync(w, ctx->words.p, pr.nbytes, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return FAZ_OK;
}

int faz_compress_raw_vec(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_interp_params *p,
                         std::vector<uint8_t> &raw) {
    int rc = compress_device(ctx, d_in, d_recon, p);
    if (rc) return rc;
    raw.resize(pending_size(ctx));
    return fetch_raw(ctx, raw.data());
}

extern "C" {

faz_status faz_gpu_interp_raw_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_interp_params *p,
                                  uint8_t *h_raw, size_t cap, size_t *raw_size) {
    if (!ctx || !d_in || !p || !raw_size) return FAZ_ERR_INVALID;
    std::vector<uint8_t> raw;
    int rc = faz_compress_raw_vec(ctx, d_in, d_recon, p, raw);
    if (rc) return rc;
    *raw_size = raw.size();
    if (!h_raw || cap < raw.size()) CTX_FAIL(FAZ_ERR_CAPACITY, "raw buffer needs %zu bytes", raw.size());
    memcpy(h_raw, raw.data(), raw.size());
    return FAZ_OK;
}

faz_status faz_gpu_compress_f32(faz_gpu_ctx *ctx, const float *d_in, float *d_recon, const faz_interp_params *p,
                                int rel_mode, double eb, size_t *raw_size) {
    if (!ctx || !d_in || !p || !raw_size || p->N < 1 || p->N > 4) return FAZ_ERR_INVALID;
    faz_interp_params q = *p;
    ctx->timings[0] = 0;
    double absmax = -1.0;
    if (rel_mode) {
        size_t num = 1;
        for (int i = 0; i < p->N; i++) num *= p->dims[i];
        HIPCHECK(hipSetDevice(ctx->device));
        if (ctx->phase_events) HIPCHECK(hipEventRecord(ctx->ev[6], ctx->stream));
        float mn, mx;
        HIPCHECK(ctx->minmax.ensure(kRangeScratch));
        HIPCHECK(ctx->freq.ensure(kHistSlots));
        HIPCHECK(ctx->first.ensure(65536));
        if (ctx->dev_bounds) {
            // the level bounds follow from the range on the device: no host
            // round trip before the level launches (the plan does not depend
            // on the bound; 1.0 stands in for it).  The range pass also clears
            // the histogram and the error flag, its fold computes the bounds.
            q.eb = 1.0;
            int rc = FAZ_OK;
            const Plan &pl = cached_plan(ctx, &q, rc);
            if (rc) return rc;
            HIPCHECK(ctx->lvl.ensure(kMaxLevels + 1));
            HIPCHECK(ctx->err_flag.ensure(1));
            launch_range_rel(d_in, num, ctx->minmax.p, ctx->freq.p, ctx->first.p, ctx->err_flag.p, eb,
                             level_scale(pl, &q), ctx->lvl.p, ctx->stream);
            HIPCHECK(hipGetLastError());
            ctx->hist_cleared = true;
            ctx->err_cleared = true;
            if (ctx->phase_events) HIPCHECK(hipEventRecord(ctx->ev[7], ctx->stream));
            rc = compress_device(ctx, d_in, d_recon, &q, -1.0, ctx->lvl.p);
            ctx->hist_cleared = false;
            ctx->err_cleared = false;
            if (rc != kRetryHostBounds) {
                if (rc) return rc;
                if (ctx->phase_events) ctx->timings[0] = elapsed(ctx->ev[6], ctx->ev[7]);  // drained by compress_device
                *raw_size = pending_size(ctx);
                return FAZ_OK;
            }
            // rare: bounds too fine for the bounded quantizer on a tile level
            if (ctx->phase_events) HIPCHECK(hipEventRecord(ctx->ev[6], ctx->stream));
        }
        // the histogram's arrays are cleared while the host waits for the range
        launch_range(d_in, num, ctx->minmax.p, ctx->stream);
        launch_hist_clear(ctx->freq.p, ctx->first.p, ctx->stream);
        HIPCHECK(hipGetLastError());
        int rc = range_readback(ctx, &mn, &mx);
        if (rc) return rc;
        ctx->hist_cleared = true;
        if (ctx->phase_events) HIPCHECK(hipEventRecord(ctx->ev[7], ctx->stream));
        q.eb = eb * (double) (float) (mx - mn);
        absmax = std::max(std::fabs((double) mn), std::fabs((double) mx));
    } else {
        q.eb = eb;
    }
    int rc = compress_device(ctx, d_in, d_recon, &q, absmax);
    ctx->hist_cleared = false;
    if (rc) return rc;
    if (rel_mode && ctx->phase_events) ctx->timings[0] = elapsed(ctx->ev[6], ctx->ev[7]);  // drained by compress_device
    *raw_size = pending_size(ctx);
    return FAZ_OK;
}

faz_status faz_gpu_fetch_raw(faz_gpu_ctx *ctx, uint8_t *h_raw, size_t cap) {
    if (!ctx || !ctx->pending.valid) return FAZ_ERR_INVALID;
    size_t need = pending_size(ctx);
    if (!h_raw || cap < need) CTX_FAIL(FAZ_ERR_CAPACITY, "raw buffer needs %zu bytes", need);
    HIPCHECK(hipSetDevice(ctx->device));
    return fetch_raw(ctx, h_raw);
}

faz_status faz_gpu_set_profiling(faz_gpu_ctx *ctx, int on) {
    if (!ctx) return FAZ_ERR_INVALID;
    ctx->profile = (on & 1) != 0;
    ctx->phase_events = (on & 2) != 0;
    for (auto &b: ctx->prof) b = faz_gpu_ctx::ProfBucket();
    return FAZ_OK;
}

faz_status faz_gpu_profile_read(const faz_gpu_ctx *ctx, double *ms, uint64_t *launches, uint64_t *points) {
    if (!ctx || !ms || !launches || !points) return FAZ_ERR_INVALID;
    // the dominant kernel: whichever launch family took longest
    const auto &b = ctx->prof[faz_gpu_profile_kernel(ctx)];
    *ms = b.ms;
    *launches = b.launches;
    *points = b.points;
    return FAZ_OK;
}

double faz_gpu_profile_bytes(const faz_gpu_ctx *ctx) {
    if (!ctx) return 0;
    return ctx->prof[faz_gpu_profile_kernel(ctx)].bytes;
}

int faz_gpu_profile_kernel(const faz_gpu_ctx *ctx) {
    if (!ctx) return -1;
    int best = 0;
    for (int i = 1; i < 3; i++)
        if (ctx->prof[i].ms > ctx->prof[best].ms) best = i;
    return best;
}

}  // extern "C"

// SZInterpolationCompressor::decompress (SIC:42-199) of a parsed engine
// buffer: the unpredictable list either in host memory (h_unpred) or in
// device memory (d_unpred, any alignment), the Huffman section's tree and byte
// count in host memory (huff, huff_n bytes up to the end of the buffer) and,
// when d_huff is given, the same section in device memory (the bitstream is
// then read there).  Whole-tile levels run the tile decoder, the others one
// recovery launch per pass.
static int interp_decompress(faz_gpu_ctx *ctx, const EngineHeader &h, double eb, int radius, uint64_t U,
                             const float *h_unpred, const uint8_t *d_unpred, const uint8_t *huff, size_t huff_n,
                             const uint8_t *d_huff, int N, float *d_out, size_t num) {
    const auto t0 = std::chrono::steady_clock::now();
    auto ms_since = [](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    };
    hipStream_t st = ctx->stream;
    size_t total = 1;
    for (int i = 0; i < N; i++) total *= h.dims[i];
    if (total != num) CTX_FAIL(FAZ_ERR_INVALID, "stream holds %zu points, caller expects %zu", total, num);
    faz_interp_params p;
    memset(&p, 0, sizeof(p));
    p.N = N;
    for (int i = 0; i < N; i++) p.dims[i] = h.dims[i];
    p.eb = eb;
    p.interp_algo = h.interp_algo;
    p.direction = h.direction;
    p.alpha = h.alpha;
    p.beta = h.beta;
    p.max_step = h.max_step;
    p.block_size = (int) h.block_size;
    p.radius = radius;
    p.levelwise = h.levelwise;
    memcpy(p.ops, h.ops.data(), h.ops.size());
    memcpy(p.dirs, h.dirs.data(), h.dirs.size());
    Plan pl;
    int rc = make_plan(&p, pl, ctx->err);
    if (rc) return rc;
    const uint32_t nchunks = (uint32_t) ((num + kHistChunk - 1) / kHistChunk);
    HIPCHECK(ctx->codes.ensure(num));
    HIPCHECK(ctx->unpred.ensure(U));
    HIPCHECK(ctx->outval.ensure(num));
    HIPCHECK(ctx->prefix.ensure(pl.prefix.size()));
    HIPCHECK(ctx->chunk_zeros.ensure(nchunks));
    HIPCHECK(ctx->chunk_base.ensure(nchunks));
    HIPCHECK(ctx->scalars.ensure(4));
    HIPCHECK(ctx->err_flag.ensure(1));
    // codes decoded on the device (HUF:236-268), then the outliers' ranks
    const auto td = std::chrono::steady_clock::now();
    // with anchors the codes open with anchor_count zeros (SIC:854-900).
    // Every check of the stream (symbol count, tree, zero count, segments)
    // is read back once at the end: the device work is bounded whatever the
    // stream holds, so the whole decompression is enqueued without a wait.
    // the unpredictable list and the block prefixes go first: the device is
    // idle while the host builds the decoder's tables
    if (U) {
        if (d_unpred)
            HIPCHECK(hipMemcpyAsync(ctx->unpred.p, d_unpred, U * 4, hipMemcpyDeviceToDevice, st));
        else
            HIPCHECK(hipMemcpyAsync(ctx->unpred.p, h_unpred, U * 4, hipMemcpyHostToDevice, st));
    }
    HIPCHECK(hipMemcpyAsync(ctx->prefix.p, pl.prefix.data(), pl.prefix.size() * 8, hipMemcpyHostToDevice, st));
    ctx->prefix_key = 0;  // the compression path's table is gone
    DecodeCheck chk;
    int drc = gpu_huffman_decode<uint16_t>(ctx, huff, huff_n, num, 65536u, ctx->codes.p, nullptr, d_huff, 0u,
                                           pl.anchor && ctx->dec_run ? pl.anchor_count : 0, &chk,
                                           ctx->chunk_zeros.p);
    if (drc) return drc;
    ctx->timings[6] = ms_since(td);
    // the write pass counted the zero codes per chunk; the single-symbol
    // stream (no write pass) is counted here
    if (chk.total == nullptr) launch_zero_count_u16(ctx->codes.p, num, ctx->chunk_zeros.p, st);
    HIPCHECK(ctx->tiles.ensure(scan_tiles(nchunks)));
    // the zero count in scalars[1] (scalars[0] holds the decoded symbol count)
    launch_zero_scan(ctx->chunk_zeros.p, ctx->chunk_base.p, nchunks, ctx->tiles.p, ctx->scalars.p + 1, st);
    launch_unpred_scatter(ctx->codes.p, num, ctx->unpred.p, ctx->chunk_zeros.p, ctx->chunk_base.p, ctx->outval.p, st,
                          U);
    ctx->hist_codes = nullptr;
    // head: anchors / first point come straight from the unpredictable list
    if (pl.anchor) {
        // the anchors are the first anchor_count codes (all zero) and take the
        // leading unpredictable values
        if (U < pl.anchor_count) CTX_FAIL(FAZ_ERR_FORMAT, "fewer unpredictable values than anchors");
        launch_anchor_scatter(ctx->unpred.p, d_out, N, pl.dims, pl.offs, (uint32_t) h.max_step, pl.anchor_count, st);
    } else {
        launch_first_value(ctx->codes.p, ctx->unpred.p, U, radius, eb, d_out, st);
    }
    std::vector<Level3D> g3;
    std::vector<int> kind;
    HIPCHECK(level_kernels(ctx, pl, &p, d_out, true, true, g3, kind));
    launch_fill_u32(ctx->err_flag.p, 1, 0u, st);
    for (size_t i = 0; i < pl.lv.size(); i++) {
        LevelPlan &L = pl.lv[i];
        if (kind[i] == 2) {
            launch_recover_tile3d(g3[i], ctx->codes.p, ctx->outval.p, ctx->err_flag.p, ctx->tile_grid, st);
            continue;
        }
        for (auto &d: L.passes) {
            d.block_prefix = ctx->prefix.p + L.prefix_at;
            launch_recover_pass(d, d_out, ctx->codes.p, ctx->outval.p, st);
        }
    }
    HIPCHECK(hipGetLastError());
    uint32_t *rep;
    if (ctx->mailbox.pinned) {
        rep = ctx->mailbox.data() + faz_gpu_ctx::kMailDecode;
        launch_decode_report(chk.total, chk.err, ctx->scalars.p + 1, ctx->err_flag.p, rep, st);
    } else {
        HIPCHECK(ctx->hist_pack.ensure(kHistPackWords));
        rep = ctx->hp_hist.data() + 2 * 65536;
        launch_decode_report(chk.total, chk.err, ctx->scalars.p + 1, ctx->err_flag.p, ctx->hist_pack.p, st);
        HIPCHECK(hipMemcpyAsync(rep, ctx->hist_pack.p, 24, hipMemcpyDeviceToHost, st));
    }
    HIPCHECK(spin_wait(st));
    uint64_t symbols, zeros;
    memcpy(&symbols, rep, 8);
    memcpy(&zeros, rep + 4, 8);
    if (symbols < num)
        CTX_FAIL(FAZ_ERR_FORMAT, "Huffman bitstream holds %llu symbols, %zu expected", (unsigned long long) symbols,
                 num);
    if (rep[2]) CTX_FAIL(FAZ_ERR_FORMAT, "Huffman bitstream leaves the tree or a symbol is out of range");
    if (zeros != U) CTX_FAIL(FAZ_ERR_FORMAT, "unpredictable count %llu does not match %llu zero codes",
                             (unsigned long long) U, (unsigned long long) zeros);
    if (rep[3]) CTX_FAIL(FAZ_ERR_FORMAT, "a block's codes lie outside the stream");
    ctx->timings[7] = ms_since(t0);
    return FAZ_OK;
}

extern "C" {

faz_status faz_gpu_interp_decompress_raw(faz_gpu_ctx *ctx, const uint8_t *h_raw, size_t n, int N, float *d_out,
                                         size_t num) {
    if (!ctx || !h_raw || !d_out || N < 1 || N > 4) return FAZ_ERR_INVALID;
    HIPCHECK(hipSetDevice(ctx->device));
    EngineHeader h;
    const uint8_t *pos = h_raw;
    if (!read_engine_header(pos, n, N, h)) CTX_FAIL(FAZ_ERR_FORMAT, "engine header truncated or unsupported");
    double eb;
    int radius;
    uint64_t U;
    if (!read_quantizer_head(pos, n - (pos - h_raw), eb, radius, U)) CTX_FAIL(FAZ_ERR_FORMAT, "quantizer truncated");
    if ((size_t) (n - (pos - h_raw)) / sizeof(float) < U) CTX_FAIL(FAZ_ERR_FORMAT, "unpredictable list truncated");
    const float *unpred_h = reinterpret_cast<const float *>(pos);
    pos += U * sizeof(float);
    return (faz_status) interp_decompress(ctx, h, eb, radius, U, unpred_h, nullptr, pos, n - (pos - h_raw), nullptr,
                                          N, d_out, num);
}

faz_status faz_gpu_interp_decompress_dev(faz_gpu_ctx *ctx, const uint8_t *d_raw, size_t n, int N, float *d_out,
                                         size_t num) {
    if (!ctx || !d_raw || !d_out || N < 1 || N > 4) return FAZ_ERR_INVALID;
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // the header and the quantizer head lie in the first few hundred bytes
    std::vector<uint8_t> &hb = ctx->h_stage;
    const size_t head = std::min<size_t>(n, 4096);
    hb.resize(head);
    HIPCHECK(hipMemcpyAsync(hb.data(), d_raw, head, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    EngineHeader h;
    const uint8_t *pos = hb.data();
    if (!read_engine_header(pos, head, N, h)) CTX_FAIL(FAZ_ERR_FORMAT, "engine header truncated or unsupported");
    double eb;
    int radius;
    uint64_t U;
    if (!read_quantizer_head(pos, head - (pos - hb.data()), eb, radius, U))
        CTX_FAIL(FAZ_ERR_FORMAT, "quantizer truncated");
    const size_t at_unpred = (size_t) (pos - hb.data());
    if ((n - at_unpred) / sizeof(float) < U) CTX_FAIL(FAZ_ERR_FORMAT, "unpredictable list truncated");
    const size_t at_huff = at_unpred + U * sizeof(float);
    // the Huffman section's tree (node count in its first 12 bytes) and
    // byte count, read in one or two copies
    size_t want = std::min<size_t>(n - at_huff, 1 << 16);
    std::vector<uint8_t> &tb = ctx->h_stage2;
    for (int round = 0; round < 2; round++) {
        tb.resize(want);
        if (want) HIPCHECK(hipMemcpyAsync(tb.data(), d_raw + at_huff, want, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (want < 12) break;
        const unsigned nc = ((unsigned) tb[4] << 24) | ((unsigned) tb[5] << 16) | ((unsigned) tb[6] << 8) | tb[7];
        const size_t w = nc <= 256 ? 1 : (nc <= 65536 ? 2 : 4);
        const size_t need = std::min<size_t>(n - at_huff, 12 + (1 + 2 * (size_t) nc * w + (size_t) nc * 5) + 8);
        if (need <= want) break;
        want = need;
    }
    return (faz_status) interp_decompress(ctx, h, eb, radius, U, nullptr, d_raw + at_unpred, tb.data(),
                                          n - at_huff, d_raw + at_huff, N, d_out, num);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: many fields of uneven shapes compressed together.  Every field has
// a child context (its own buffers and stream).  One hipGraph holds the range
// launches of every field, a second one every field's interpolation,
// quantization, histogram and outlier compaction, one graph branch per field
// (fork/join on the child streams during capture, so small fields run beside
// big ones).  The host then builds each field's Huffman tree and enqueues
// every field's encode before it waits once.  The prediction graph carries the
// fields' bounds as kernel arguments: it is replayed as is while they stay the
// same (absolute bounds, or unchanged ranges) and re-captured and updated in
// place (hipGraphExecUpdate) when they change.  At creation an autotuner picks
// each 3-D field's launch schedule (whole-tile kernel grid, per-block level
// kernel or per-pass launches for the coarse levels) by timing the candidates
// on the bound input; every candidate writes the same codes.
// ---------------------------------------------------------------------------
struct faz_batch {
    faz_gpu_ctx *ctx = nullptr;  // parent: device, capture stream
    int n = 0;
    std::vector<faz_interp_params> prm;
    std::vector<const float *> d_in;
    std::vector<uint64_t> num, ncodes;
    std::vector<faz_gpu_ctx *> kid;
    std::vector<int> schedule;   // index into kSchedules per field
    hipGraph_t g_range = nullptr, g_pred = nullptr;
    hipGraphExec_t x_range = nullptr, x_pred = nullptr;
    std::vector<double> key;     // bound and max |x| of each field inside x_pred
    std::vector<double> eb, rng, absmax;
    hipEvent_t ev[6] = {};
    faz_batch_stats st;
    uint32_t *h_minmax = nullptr;  // pinned, 2 per field
    uint32_t *h_hist = nullptr;    // pinned, freq then first (2 x 65536) per field
    uint64_t *h_scal = nullptr;    // pinned, outlier count, bit total, error flag per field
};

namespace {

// Launch schedules the autotuner tries on 3-D fields: {tile3d, tile_grid
// multiple of the CU count / 256, fused3d, fused_min_blocks}.
struct Schedule {
    bool tile3d;
    uint32_t grid_mul;  // whole-tile grid = CUs * grid_mul
    bool fused3d;
    uint64_t fused_min_blocks;
};
constexpr Schedule kSchedules[] = {
    {true, 1, true, 256},  {true, 2, true, 256}, {true, 1, true, 64},
    {true, 2, true, 1024}, {false, 1, true, 256}, {false, 1, false, 256},
};
constexpr int kNumSchedules = sizeof(kSchedules) / sizeof(kSchedules[0]);

void apply_schedule(faz_gpu_ctx *kid, const faz_gpu_ctx *parent, int s) {
    const Schedule &c = kSchedules[s];
    kid->tile3d = c.tile3d && parent->tile3d;
    kid->fused3d = c.fused3d && parent->fused3d;
    kid->tile_grid = parent->tile_grid * c.grid_mul;
    kid->fused_min_blocks = c.fused_min_blocks;
}

double ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

float key2f(uint32_t key) {
    uint32_t u = (key & 0x80000000u) ? (key & 0x7FFFFFFFu) : ~key;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

#define BFAIL(code, ...) do { char _m[512]; snprintf(_m, sizeof(_m), __VA_ARGS__); b->ctx->err = _m; return code; } while (0)
#define BHIP(x)                                                                                    \
    do {                                                                                           \
        hipError_t _e = (x);                                                                       \
        if (_e != hipSuccess) BFAIL(FAZ_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(_e));      \
    } while (0)

// Every field's range kernel, captured once.
int batch_capture_range(faz_batch *b) {
    hipStream_t cs = b->ctx->stream;
    BHIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    BHIP(hipEventRecord(b->ev[0], cs));
    bool ok = true;
    for (int i = 0; i < b->n && ok; i++) {
        faz_gpu_ctx *k = b->kid[i];
        ok = hipStreamWaitEvent(k->stream, b->ev[0], 0) == hipSuccess;
        if (!ok) break;
        launch_range(b->d_in[i], b->num[i], k->minmax.p, k->stream);
        ok = hipEventRecord(k->ev[6], k->stream) == hipSuccess && hipStreamWaitEvent(cs, k->ev[6], 0) == hipSuccess;
    }
    hipError_t e = hipStreamEndCapture(cs, &b->g_range);
    if (!ok) BFAIL(FAZ_ERR_HIP, "capturing the range graph failed");
    BHIP(e);
    BHIP(hipGraphInstantiate(&b->x_range, b->g_range, nullptr, nullptr, 0));
    return FAZ_OK;
}

// Every field's device pipeline on its own branch, bounds as in b->eb.
int batch_capture_pred(faz_batch *b, hipGraph_t &g) {
    hipStream_t cs = b->ctx->stream;
    BHIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int rc = FAZ_OK;
    BHIP(hipEventRecord(b->ev[0], cs));
    for (int i = 0; i < b->n && !rc; i++) {
        faz_gpu_ctx *k = b->kid[i];
        if (hipStreamWaitEvent(k->stream, b->ev[0], 0) != hipSuccess) {
            rc = FAZ_ERR_HIP;
            break;
        }
        faz_interp_params q = b->prm[i];
        q.eb = b->eb[i];
        faz_interp_out io;
        rc = interp_compress(k, b->d_in[i], k->work_out.p, &q, &io, false, b->absmax[i], true);
        if (rc) {
            b->ctx->err = k->err;
            break;
        }
        // the histogram's window, symbol 0, flags and outlier count in one block
        launch_hist_pack(k->freq.p, k->first.p, q.radius, k->scalars.p, k->err_flag.p, k->hist_pack.p, k->stream);
        if (hipEventRecord(k->ev[7], k->stream) != hipSuccess || hipStreamWaitEvent(cs, k->ev[7], 0) != hipSuccess)
            rc = FAZ_ERR_HIP;
    }
    hipGraph_t tmp = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &tmp);
    if (rc) {
        if (tmp) (void) hipGraphDestroy(tmp);
        if (rc == FAZ_ERR_HIP) BFAIL(FAZ_ERR_HIP, "capturing the batch failed");
        return rc;
    }
    BHIP(e);
    g = tmp;
    return FAZ_OK;
}

// One eager compression of field i with its current schedule: sizes every
// buffer and uploads the block prefix table the captured pipeline relies on;
// returns the interpolation time.
int batch_prepare_field(faz_batch *b, int i, double &ms) {
    faz_gpu_ctx *k = b->kid[i];
    faz_interp_params q = b->prm[i];
    q.eb = b->eb[i];
    int rc = FAZ_OK;
    if (hipSetDevice(k->device) != hipSuccess) BFAIL(FAZ_ERR_HIP, "hipSetDevice failed");
    if (k->work_out.ensure(b->num[i]) != hipSuccess || k->unpred.ensure(b->num[i]) != hipSuccess ||
        k->hist_pack.ensure(kHistPackWords) != hipSuccess)
        BFAIL(FAZ_ERR_HIP, "batch: device allocation failed");
    faz_interp_out io;
    rc = interp_compress(k, b->d_in[i], k->work_out.p, &q, &io, false, b->absmax[i]);
    if (rc) {
        b->ctx->err = k->err;
        return rc;
    }
    ms = k->timings[1];
    b->ncodes[i] = io.n_codes;
    return FAZ_OK;
}

}  // namespace

extern "C" {

faz_status faz_gpu_batch_create(faz_gpu_ctx *ctx, const faz_interp_params *fields, const float *const *d_in, int n,
                                int autotune, faz_batch **out) {
    if (!ctx || !fields || !d_in || n <= 0 || !out) return FAZ_ERR_INVALID;
    *out = nullptr;
    HIPCHECK(hipSetDevice(ctx->device));
    faz_batch *b = new faz_batch();
    b->ctx = ctx;
    b->n = n;
    b->prm.assign(fields, fields + n);
    b->d_in.assign(d_in, d_in + n);
    b->eb.assign(n, 0.0), b->rng.assign(n, 0.0), b->absmax.assign(n, 0.0), b->schedule.assign(n, 0);
    b->ncodes.assign(n, 0);
    memset(&b->st, 0, sizeof(b->st));
    auto fail = [&](int rc) {
        faz_gpu_batch_destroy(b);
        return (faz_status) rc;
    };
    for (int i = 0; i < n; i++) {
        const faz_interp_params &p = fields[i];
        if (!d_in[i] || p.N < 1 || p.N > 4) {
            ctx->err = "batch: field without input or with N outside 1..4";
            return fail(FAZ_ERR_INVALID);
        }
        uint64_t m = 1;
        for (int a = 0; a < p.N; a++) m *= p.dims[a];
        b->num.push_back(m);
        faz_gpu_ctx *k = nullptr;
        int rc = faz_gpu_ctx_create(ctx->device, &k);
        if (rc) return fail(rc);
        k->tile3d = ctx->tile3d, k->fused3d = ctx->fused3d, k->tile_grid = ctx->tile_grid;
        k->fused_min_blocks = ctx->fused_min_blocks;
        // the fields' encodes run on concurrent streams: the one-pass encoder's
        // look-back needs all of its workgroups resident, so the batch keeps
        // the two-pass encoder
        k->enc_onepass = false;
        b->kid.push_back(k);
        if (k->minmax.ensure(kRangeScratch) != hipSuccess) return fail(FAZ_ERR_HIP);
    }
    for (auto &e: b->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(FAZ_ERR_HIP);
    if (hipHostMalloc((void **) &b->h_minmax, 8 * (size_t) n, 0) != hipSuccess ||
        hipHostMalloc((void **) &b->h_hist, 2 * 65536 * 4 * (size_t) n, 0) != hipSuccess ||
        hipHostMalloc((void **) &b->h_scal, 3 * 8 * (size_t) n, 0) != hipSuccess)
        return fail(FAZ_ERR_HIP);
    int rc = batch_capture_range(b);
    if (rc) return fail(rc);
    // ranges of the bound inputs: the eager runs below need a bound and max |x|
    if (hipGraphLaunch(b->x_range, ctx->stream) != hipSuccess) return fail(FAZ_ERR_HIP);
    for (int i = 0; i < n; i++)
        if (hipMemcpyAsync(b->h_minmax + 2 * i, b->kid[i]->minmax.p, 8, hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess)
            return fail(FAZ_ERR_HIP);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(FAZ_ERR_HIP);
    for (int i = 0; i < n; i++) {
        const float mn = key2f(b->h_minmax[2 * i]), mx = key2f(b->h_minmax[2 * i + 1]);
        b->absmax[i] = std::max(std::fabs((double) mn), std::fabs((double) mx));
        const double r = (double) (float) (mx - mn);
        b->eb[i] = r > 0 ? 1e-3 * r : 1e-3;  // a representative bound for preparation and tuning
    }
    // autotune (3-D fields) and prepare every field
    for (int i = 0; i < n; i++) {
        double ms = 0;
        if (autotune && fields[i].N == 3) {
            double best = 1e30;
            int pick = 0;
            for (int s = 0; s < kNumSchedules; s++) {
                apply_schedule(b->kid[i], ctx, s);
                double t = 1e30;
                for (int rep = 0; rep < 2; rep++) {  // first run allocates
                    if ((rc = batch_prepare_field(b, i, ms))) return fail(rc);
                    t = std::min(t, ms);
                }
                if (t < best) best = t, pick = s;
            }
            b->schedule[i] = pick;
            b->st.tuned_fields++;
        }
        apply_schedule(b->kid[i], ctx, b->schedule[i]);
        if ((rc = batch_prepare_field(b, i, ms))) return fail(rc);
    }
    *out = b;
    return FAZ_OK;
}

void faz_gpu_batch_destroy(faz_batch *b) {
    if (!b) return;
    (void) hipSetDevice(b->ctx->device);
    (void) hipStreamSynchronize(b->ctx->stream);
    if (b->x_pred) (void) hipGraphExecDestroy(b->x_pred);
    if (b->g_pred) (void) hipGraphDestroy(b->g_pred);
    if (b->x_range) (void) hipGraphExecDestroy(b->x_range);
    if (b->g_range) (void) hipGraphDestroy(b->g_range);
    for (auto k: b->kid) faz_gpu_ctx_destroy(k);
    for (auto &e: b->ev)
        if (e) (void) hipEventDestroy(e);
    if (b->h_minmax) (void) hipHostFree(b->h_minmax);
    if (b->h_hist) (void) hipHostFree(b->h_hist);
    if (b->h_scal) (void) hipHostFree(b->h_scal);
    delete b;
}

faz_status faz_gpu_batch_compress(faz_batch *b, const int *rel_mode, const double *eb, size_t *raw_sizes,
                                  double *abs_eb, double *rng) {
    if (!b || !rel_mode || !eb || !raw_sizes) return FAZ_ERR_INVALID;
    faz_gpu_ctx *ctx = b->ctx;
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t cs = ctx->stream;
    const int n = b->n;
    char errbuf[512] = {0};
    for (int i = 0; i < n; i++) b->h_scal[3 * i + 2] = 0;
    // 1. ranges of every field (one graph)
    HIPCHECK(hipEventRecord(b->ev[1], cs));
    HIPCHECK(hipGraphLaunch(b->x_range, cs));
    for (int i = 0; i < n; i++)
        HIPCHECK(hipMemcpyAsync(b->h_minmax + 2 * i, b->kid[i]->minmax.p, 8, hipMemcpyDeviceToHost, cs));
    HIPCHECK(hipStreamSynchronize(cs));
    std::vector<double> key(2 * (size_t) n);
    for (int i = 0; i < n; i++) {
        const float mn = key2f(b->h_minmax[2 * i]), mx = key2f(b->h_minmax[2 * i + 1]);
        b->absmax[i] = std::max(std::fabs((double) mn), std::fabs((double) mx));
        b->rng[i] = (double) (float) (mx - mn);
        // calAbsErrorBound (Config.hpp / SZ_compress): REL bound times the float range
        b->eb[i] = rel_mode[i] ? eb[i] * b->rng[i] : eb[i];
        if (!(b->eb[i] > 0)) CTX_FAIL(FAZ_ERR_INVALID, "field %d: error bound not positive", i);
        key[2 * i] = b->eb[i];
        key[2 * i + 1] = b->absmax[i];
    }
    // 2. prediction graph: replay, or re-capture and update in place
    if (!b->x_pred || key != b->key) {
        hipGraph_t g = nullptr;
        int rc = batch_capture_pred(b, g);
        if (rc) return (faz_status) rc;
        bool updated = false;
        if (b->x_pred) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res;
            if (hipGraphExecUpdate(b->x_pred, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
                updated = true;
                b->st.updates++;
            } else {
                (void) hipGetLastError();
                (void) hipGraphExecDestroy(b->x_pred);
                b->x_pred = nullptr;
            }
        }
        if (!updated) {
            HIPCHECK(hipGraphInstantiate(&b->x_pred, g, nullptr, nullptr, 0));
            b->st.instantiations++;
        }
        if (b->g_pred) (void) hipGraphDestroy(b->g_pred);
        b->g_pred = g;
        b->key = key;
        b->st.captures++;
    } else {
        b->st.replays++;
    }
    HIPCHECK(hipEventRecord(b->ev[2], cs));
    HIPCHECK(hipGraphLaunch(b->x_pred, cs));
    HIPCHECK(hipEventRecord(b->ev[3], cs));
    // 3. each field's packed histogram block (window counts and first
    //    positions, symbol 0, flags, outlier count) to pinned host memory;
    //    a pool of host threads builds the trees and enqueues each field's
    //    encode on its own stream as soon as its tree exists; one wait at the
    //    end
    for (int i = 0; i < n; i++) {
        faz_gpu_ctx *k = b->kid[i];
        uint32_t *hh = b->h_hist + (size_t) i * 2 * 65536;
        HIPCHECK(hipMemcpyAsync(hh, k->hist_pack.p, kHistPackWords * 4, hipMemcpyDeviceToHost, cs));
    }
    HIPCHECK(hipStreamSynchronize(cs));
    std::vector<int> field_rc(n, FAZ_OK);
    std::vector<std::string> field_err(n);
    std::atomic<int> next{0};
    auto builder = [&]() {
        for (int i; (i = next.fetch_add(1)) < n;) {
            faz_gpu_ctx *k = b->kid[i];
            PendingRaw &pr = k->pending;
            pr.valid = false;
            uint32_t *hf = b->h_hist + (size_t) i * 2 * 65536;
            int r;
            if (hf[kHistPackFreq0 + 2] == 0) {
                r = build_huffman_window_codes(hf, hf + kHistWin, hist_window_lo(b->prm[i].radius), kHistWin,
                                               hf[kHistPackFreq0], hf[kHistPackFreq0 + 1], pr.tab);
            } else {  // a symbol outside the window: the whole arrays
                if (hipSetDevice(k->device) != hipSuccess ||
                    hipMemcpy(hf, k->freq.p, 65536 * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(hf + 65536, k->first.p, 65536 * 4, hipMemcpyDeviceToHost) != hipSuccess) {
                    field_rc[i] = FAZ_ERR_HIP;
                    field_err[i] = "histogram readback failed";
                    continue;
                }
                r = build_huffman_codes(hf, hf + 65536, 65536, pr.tab);
            }
            if (r) {
                field_rc[i] = FAZ_ERR_UNSUPPORTED;
                field_err[i] = "Huffman tree could not be built";
                continue;
            }
            faz_huff_table ht{pr.tab.offset, pr.tab.n_states, pr.tab.code.data(), pr.tab.len.data()};
            if (hipSetDevice(k->device) != hipSuccess) {
                field_rc[i] = FAZ_ERR_HIP;
                field_err[i] = "hipSetDevice failed";
                continue;
            }
            field_rc[i] = encode_enqueue(k, k->codes.p, b->ncodes[i], &ht);
            if (field_rc[i]) field_err[i] = k->err;
            finish_huffman_tree(pr.tab);  // while the encode runs
        }
    };
    const int nthreads = std::max(1, std::min(n, 8));
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++) pool.emplace_back(builder);
    for (auto &th: pool) th.join();
    for (int i = 0; i < n; i++)
        if (field_rc[i]) CTX_FAIL(field_rc[i], "field %d: %s", i, field_err[i].c_str());
    for (int i = 0; i < n; i++) {
        faz_gpu_ctx *k = b->kid[i];
        HIPCHECK(hipEventRecord(k->ev[6], k->stream));
        HIPCHECK(hipStreamWaitEvent(cs, k->ev[6], 0));
        HIPCHECK(hipMemcpyAsync(b->h_scal + 3 * i, k->scalars.p, 16, hipMemcpyDeviceToHost, cs));
        HIPCHECK(hipMemcpyAsync(b->h_scal + 3 * i + 2, k->err_flag.p, 4, hipMemcpyDeviceToHost, cs));
    }
    HIPCHECK(hipEventRecord(b->ev[5], cs));
    HIPCHECK(hipStreamSynchronize(cs));
    // 4. sizes of every field's engine buffer
    for (int i = 0; i < n; i++) {
        faz_gpu_ctx *k = b->kid[i];
        const uint64_t *sc = b->h_scal + 3 * i;
        const uint32_t errflag = (uint32_t) sc[2];
        if (errflag & 2u) CTX_FAIL(FAZ_ERR_INVALID, "internal: field %d: a block segment fell outside the stream", i);
        if (errflag) CTX_FAIL(FAZ_ERR_UNSUPPORTED, "field %d: a quantization code exceeded 16 bits", i);
        const faz_interp_params &p = b->prm[i];
        PendingRaw &pr = k->pending;
        EngineHeader &h = pr.h;
        h.N = p.N;
        for (int a = 0; a < p.N; a++) h.dims[a] = p.dims[a];
        h.block_size = (unsigned) p.block_size;
        h.interp_algo = p.interp_algo;
        h.direction = p.direction;
        h.alpha = p.alpha;
        h.beta = p.beta;
        h.max_step = p.max_step;
        h.levelwise = p.levelwise;
        h.ops.assign(p.ops, p.ops + p.levelwise);
        h.dirs.assign(p.dirs, p.dirs + p.levelwise);
        pr.eb = b->eb[i];
        pr.radius = p.radius;
        pr.n_unpred = sc[0];
        pr.nbytes = (size_t) ((sc[1] + 7) / 8);
        pr.valid = true;
        raw_sizes[i] = engine_raw_size(pr.h, pr.n_unpred, pr.tab.tree.size(), pr.nbytes);
        if (abs_eb) abs_eb[i] = b->eb[i];
        if (rng) rng[i] = b->rng[i];
    }
    b->st.range_ms = ev_ms(b->ev[1], b->ev[2]);
    b->st.pred_ms = ev_ms(b->ev[2], b->ev[3]);
    b->st.encode_ms = ev_ms(b->ev[3], b->ev[5]);
    b->st.total_ms = ev_ms(b->ev[1], b->ev[5]);
    b->st.calls++;
    return FAZ_OK;
}

faz_status faz_gpu_batch_fetch_raw(faz_batch *b, int i, uint8_t *h_raw, size_t cap) {
    if (!b || i < 0 || i >= b->n) return FAZ_ERR_INVALID;
    faz_gpu_ctx *k = b->kid[i];
    if (!k->pending.valid) return FAZ_ERR_INVALID;
    const size_t need = pending_size(k);
    faz_gpu_ctx *ctx = b->ctx;
    if (!h_raw || cap < need) CTX_FAIL(FAZ_ERR_CAPACITY, "raw buffer needs %zu bytes", need);
    HIPCHECK(hipSetDevice(k->device));
    int rc = fetch_raw(k, h_raw);
    if (rc) ctx->err = k->err;
    return (faz_status) rc;
}

faz_status faz_gpu_batch_field(const faz_batch *b, int i, faz_interp_params *p) {
    if (!b || !p || i < 0 || i >= b->n) return FAZ_ERR_INVALID;
    *p = b->prm[i];
    return FAZ_OK;
}

faz_status faz_gpu_batch_stats(const faz_batch *b, faz_batch_stats *s, int *schedules) {
    if (!b || !s) return FAZ_ERR_INVALID;
    *s = b->st;
    if (schedules)
        for (int i = 0; i < b->n; i++) schedules[i] = b->schedule[i];
    return FAZ_OK;
}

}  // extern "C"
