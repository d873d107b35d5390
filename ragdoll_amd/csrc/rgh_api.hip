// Host side of the C ABI declared in include/ragdoll_hip.h: contexts, device buffers and the
// launch schedule of a P frame (analysis wavefront, deblock wavefront, border extension).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "rgh_device.h"

extern "C" __global__ void rgh_k_analyse_step(const RghJob *jobs, int t);
extern "C" __global__ void rgh_k_deblock_step(const RghJob *jobs, int t);
extern "C" __global__ void rgh_k_extend(const RghJob *jobs);

namespace {

// Standard H.264 tables needed on the host for the per-QP constants.
const int32_t kQuant4Mf[6][3] = {{13107, 5243, 8066}, {11916, 4660, 7490}, {10082, 4194, 6554},
                                 {9362, 3647, 5825},  {8192, 3355, 5243},  {7282, 2893, 4559}};
const uint8_t kChromaQp[52] = {0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16, 17,
                               18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 29, 30, 31, 32, 32, 33,
                               34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};

int pos_type(int i) {
    int r = i >> 2, c = i & 3;
    return (!(r & 1) && !(c & 1)) ? 0 : (((r & 1) && (c & 1)) ? 1 : 2);
}

// enc.rs:1784 zero_sad_bound / :2122 zero_ssd_bound for the inter quantizer at qp.
void zero_bounds(int qp, uint32_t *sad, uint32_t *ssd, uint32_t *probe) {
    static const uint32_t Wt[4] = {1, 2, 1, 2}, Nn[4] = {4, 10, 4, 10};
    uint32_t shift = 15 + (uint32_t)(qp / 6), bias = (1u << shift) / 6u;
    uint32_t room = (1u << shift) - bias - 1, ms = UINT32_MAX, mq = UINT32_MAX;
    for (int i = 0; i < 16; i++) {
        uint32_t mf = (uint32_t)kQuant4Mf[qp % 6][pos_type(i)];
        ms = std::min(ms, room / (mf * Wt[i >> 2] * Wt[i & 3]));
        uint32_t lim = room / mf;
        mq = std::min(mq, lim * lim / (Nn[i >> 2] * Nn[i & 3]));
    }
    *sad = ms;
    *ssd = mq;
    if (probe) *probe = 16 * (((1u << shift) - bias) / (uint32_t)kQuant4Mf[qp % 6][0] + 1);
}

struct Picture {
    uint8_t *base[3] = {nullptr, nullptr, nullptr};  // allocation
    uint8_t *org[3] = {nullptr, nullptr, nullptr};   // sample (0, 0)
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int width = 0, height = 0, mb_w = 0, mb_h = 0, profile = 0, qp = 0, me_iters = 16, qpel_iters = 2;
    int sy = 0, sc = 0, rows_y = 0, rows_c = 0;
    Picture pic[2];
    int ref = 0;  // index of the reference picture; the other is written
    rgh_mb_decision *dec[2] = {nullptr, nullptr};
    int dcur = 0;  // decisions written by the next frame
    int16_t *levels = nullptr;
    uint8_t *src[3] = {nullptr, nullptr, nullptr};  // staging for host sources
    uint8_t *pre[3] = {nullptr, nullptr, nullptr};  // undeblocked copy of the last frame
    RghJob *jobs_dev = nullptr;
    int jobs_cap = 0;
    bool has_ref = false;
    bool analysed = false;
    bool keep_pre = true;
    uint32_t sb[5] = {0, 0, 0, 0, 0};
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};  // analysis start, analysis end, frame end
    bool timing = false, timed = false;
    int timed_launches = 0;
};

rgh_status hip_status(hipError_t e) {
    if (e == hipSuccess) return RGH_OK;
    if (e == hipErrorOutOfMemory) return RGH_ERR_OOM;
    return RGH_ERR_DEVICE;
}
#define HIP_TRY(expr)                                     \
    do {                                                  \
        hipError_t e__ = (expr);                          \
        if (e__ != hipSuccess) return hip_status(e__);    \
    } while (0)

void free_ctx(Ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    for (auto &p : c->pic)
        for (auto &b : p.base)
            if (b) (void)hipFree(b);
    for (auto &d : c->dec)
        if (d) (void)hipFree(d);
    if (c->levels) (void)hipFree(c->levels);
    for (auto &b : c->src)
        if (b) (void)hipFree(b);
    for (auto &b : c->pre)
        if (b) (void)hipFree(b);
    if (c->jobs_dev) (void)hipFree(c->jobs_dev);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

RghJob make_job(Ctx *c, const uint8_t *sy, const uint8_t *su, const uint8_t *sv, int ssy, int ssu, int ssv) {
    RghJob j;
    std::memset(&j, 0, sizeof(j));
    j.src_y = sy;
    j.src_u = su;
    j.src_v = sv;
    j.src_sy = ssy;
    j.src_su = ssu;
    j.src_sv = ssv;
    j.width = c->width;
    j.height = c->height;
    const Picture &r = c->pic[c->ref], &w = c->pic[1 - c->ref];
    j.ref_y = r.org[0];
    j.ref_u = r.org[1];
    j.ref_v = r.org[2];
    j.cur_y = w.org[0];
    j.cur_u = w.org[1];
    j.cur_v = w.org[2];
    j.sy = c->sy;
    j.sc = c->sc;
    j.mb_w = c->mb_w;
    j.mb_h = c->mb_h;
    j.qp = c->qp;
    j.me_iters = c->me_iters;
    j.qpel_iters = c->qpel_iters;
    j.t8x8 = c->profile == RGH_PROFILE_HIGH;
    j.dec = c->dec[c->dcur];
    j.prev = c->dec[1 - c->dcur];
    j.levels = c->levels;
    j.sb_luma = c->sb[0];
    j.sb_luma_ssd = c->sb[1];
    j.sb_chroma = c->sb[2];
    j.sb_chroma_ssd = c->sb[3];
    j.sb_probe = c->sb[4];
    j.steps = (c->mb_w - 1) + 2 * (c->mb_h - 1) + 1;
    j.deblock = c->qp != 0;
    return j;
}

// Launches the wavefront of one P frame for every job, on `stream`.
rgh_status run_jobs(Ctx *const *ctxs, int n, const std::vector<RghJob> &jobs, hipStream_t stream, RghJob *jobs_dev) {
    HIP_TRY(hipMemcpyAsync(jobs_dev, jobs.data(), sizeof(RghJob) * (size_t)n, hipMemcpyHostToDevice, stream));
    int steps = 0, width = 0, rows = 0;
    for (int i = 0; i < n; i++) {
        steps = std::max(steps, jobs[i].steps);
        width = std::max(width, std::min(jobs[i].mb_h, (jobs[i].mb_w + 1) / 2));
        rows = std::max(rows, ctxs[i]->rows_y);
    }
    dim3 grid((unsigned)width, (unsigned)n);
    Ctx *c0 = ctxs[0];
    if (c0->timing) HIP_TRY(hipEventRecord(c0->ev[0], stream));
    for (int t = 0; t < steps; t++) hipLaunchKernelGGL(rgh_k_analyse_step, grid, dim3(64), 0, stream, jobs_dev, t);
    HIP_TRY(hipGetLastError());
    if (c0->timing) HIP_TRY(hipEventRecord(c0->ev[1], stream));
    for (int i = 0; i < n; i++) {
        Ctx *c = ctxs[i];
        if (c->keep_pre) {
            const Picture &w = c->pic[1 - c->ref];
            HIP_TRY(hipMemcpy2DAsync(c->pre[0], (size_t)c->mb_w * 16, w.org[0], (size_t)c->sy, (size_t)c->mb_w * 16,
                                     (size_t)c->mb_h * 16, hipMemcpyDeviceToDevice, stream));
            for (int p = 1; p < 3; p++)
                HIP_TRY(hipMemcpy2DAsync(c->pre[p], (size_t)c->mb_w * 8, w.org[p], (size_t)c->sc, (size_t)c->mb_w * 8,
                                         (size_t)c->mb_h * 8, hipMemcpyDeviceToDevice, stream));
        }
    }
    for (int t = 0; t < steps; t++) hipLaunchKernelGGL(rgh_k_deblock_step, grid, dim3(64), 0, stream, jobs_dev, t);
    hipLaunchKernelGGL(rgh_k_extend, dim3((unsigned)rows, 3, (unsigned)n), dim3(256), 0, stream, jobs_dev);
    HIP_TRY(hipGetLastError());
    if (c0->timing) {
        HIP_TRY(hipEventRecord(c0->ev[2], stream));
        c0->timed = true;
        c0->timed_launches = steps;
    }
    for (int i = 0; i < n; i++) {
        Ctx *c = ctxs[i];
        c->ref = 1 - c->ref;
        c->dcur = 1 - c->dcur;
        c->analysed = true;
    }
    return RGH_OK;
}

rgh_status ensure_jobs(Ctx *c, int n) {
    if (c->jobs_cap >= n) return RGH_OK;
    if (c->jobs_dev) HIP_TRY(hipFree(c->jobs_dev));
    c->jobs_dev = nullptr;
    int cap = std::max(n, 8);
    HIP_TRY(hipMalloc(&c->jobs_dev, sizeof(RghJob) * (size_t)cap));
    c->jobs_cap = cap;
    return RGH_OK;
}

rgh_status check_i420(const rgh_i420 *p, int w, int h) {
    if (!p || !p->y || !p->u || !p->v) return RGH_ERR_INVALID_ARGUMENT;
    if (p->width != w || p->height != h) return RGH_ERR_INVALID_DIMENSIONS;
    if (p->stride_y < w || p->stride_u < (w + 1) / 2 || p->stride_v < (w + 1) / 2) return RGH_ERR_INVALID_ARGUMENT;
    return RGH_OK;
}

}  // namespace

extern "C" {

const char *rgh_status_str(int32_t s) {
    switch (s) {
    case RGH_OK: return "ok";
    case RGH_ERR_INVALID_ARGUMENT: return "invalid argument";
    case RGH_ERR_UNSUPPORTED: return "unsupported";
    case RGH_ERR_INVALID_DIMENSIONS: return "invalid dimensions";
    case RGH_ERR_NO_REFERENCE: return "no reference picture";
    case RGH_ERR_DEVICE: return "HIP device error";
    case RGH_ERR_OOM: return "out of device memory";
    default: return "unknown status";
    }
}

rgh_status rgh_create(const rgh_config *cfg, void **ctx_out) {
    if (!cfg || !ctx_out) return RGH_ERR_INVALID_ARGUMENT;
    *ctx_out = nullptr;
    if (cfg->width <= 0 || cfg->height <= 0 || (cfg->width & 1) || (cfg->height & 1) || cfg->width > 8192 ||
        cfg->height > 8192)
        return RGH_ERR_INVALID_DIMENSIONS;
    if (cfg->profile < RGH_PROFILE_BASELINE || cfg->profile > RGH_PROFILE_HIGH) return RGH_ERR_UNSUPPORTED;
    // QP 0 codes I_PCM macroblocks (enc.rs:2423), which this analysis does not produce.
    if (cfg->qp < 1 || cfg->qp > 51) return RGH_ERR_UNSUPPORTED;
    if (cfg->me_range < 0 || cfg->me_range > 1024 || cfg->qpel_iters < 0 || cfg->qpel_iters > 8)
        return RGH_ERR_INVALID_ARGUMENT;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RGH_ERR_DEVICE;
    if (cfg->device < 0 || cfg->device >= ndev) return RGH_ERR_INVALID_ARGUMENT;
    Ctx *c = new (std::nothrow) Ctx();
    if (!c) return RGH_ERR_OOM;
    c->device = cfg->device;
    c->width = cfg->width;
    c->height = cfg->height;
    c->mb_w = (cfg->width + 15) / 16;
    c->mb_h = (cfg->height + 15) / 16;
    c->profile = cfg->profile;
    c->qp = cfg->qp;
    c->me_iters = cfg->me_range ? cfg->me_range : 16;
    c->qpel_iters = cfg->qpel_iters ? cfg->qpel_iters : 2;
    c->keep_pre = !(cfg->flags & RGH_FLAG_NO_UNDEBLOCKED_COPY);
    const int W = c->mb_w * 16, H = c->mb_h * 16;
    c->sy = (W + 2 * RGH_PAD_Y + 31) & ~31;
    c->sc = (W / 2 + 2 * RGH_PAD_C + 31) & ~31;
    c->rows_y = H + 2 * RGH_PAD_Y;
    c->rows_c = H / 2 + 2 * RGH_PAD_C;
    zero_bounds(c->qp, &c->sb[0], &c->sb[1], &c->sb[4]);
    {
        // enc.rs:2131 SkipBounds::new: the chroma bounds also keep the 2x2 DC level zero.
        int qpc = kChromaQp[c->qp];
        uint32_t cs, css;
        zero_bounds(qpc, &cs, &css, nullptr);
        uint32_t dc_shift = 16 + (uint32_t)qpc / 6;
        uint32_t dc_room = (1u << dc_shift) - (1u << dc_shift) / 3 - 1;
        uint32_t dc = dc_room / (uint32_t)kQuant4Mf[qpc % 6][0];
        c->sb[2] = std::min(cs, dc);
        c->sb[3] = std::min(css, dc * dc / 64);
    }
    auto fail = [&](rgh_status s) {
        free_ctx(c);
        return s;
    };
    if (hipSetDevice(c->device) != hipSuccess) return fail(RGH_ERR_DEVICE);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(RGH_ERR_DEVICE);
    const size_t ny = (size_t)c->sy * c->rows_y + 64, nc = (size_t)c->sc * c->rows_c + 64;
    for (auto &p : c->pic) {
        size_t sizes[3] = {ny, nc, nc};
        for (int k = 0; k < 3; k++) {
            hipError_t e = hipMalloc(&p.base[k], sizes[k]);
            if (e != hipSuccess) return fail(hip_status(e));
            if (hipMemset(p.base[k], 0, sizes[k]) != hipSuccess) return fail(RGH_ERR_DEVICE);
            int pad = k ? RGH_PAD_C : RGH_PAD_Y, s = k ? c->sc : c->sy;
            p.org[k] = p.base[k] + (size_t)pad * s + pad;
        }
    }
    const size_t n = (size_t)c->mb_w * c->mb_h;
    for (auto &d : c->dec) {
        hipError_t e = hipMalloc(&d, n * sizeof(rgh_mb_decision));
        if (e != hipSuccess) return fail(hip_status(e));
        if (hipMemset(d, 0, n * sizeof(rgh_mb_decision)) != hipSuccess) return fail(RGH_ERR_DEVICE);
    }
    {
        hipError_t e = hipMalloc(&c->levels, n * RGH_LEVELS_PER_MB * sizeof(int16_t));
        if (e != hipSuccess) return fail(hip_status(e));
    }
    const size_t srcs[3] = {(size_t)c->width * c->height, (size_t)((c->width + 1) / 2) * ((c->height + 1) / 2),
                            (size_t)((c->width + 1) / 2) * ((c->height + 1) / 2)};
    for (int k = 0; k < 3; k++) {
        hipError_t e = hipMalloc(&c->src[k], srcs[k]);
        if (e != hipSuccess) return fail(hip_status(e));
        e = hipMalloc(&c->pre[k], k ? (size_t)W * H / 4 : (size_t)W * H);
        if (e != hipSuccess) return fail(hip_status(e));
    }
    rgh_status s = ensure_jobs(c, 1);
    if (s != RGH_OK) return fail(s);
    *ctx_out = c;
    return RGH_OK;
}

rgh_status rgh_destroy(void *ctx) {
    if (!ctx) return RGH_ERR_INVALID_ARGUMENT;
    Ctx *c = static_cast<Ctx *>(ctx);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_ctx(c);
    return RGH_OK;
}

rgh_status rgh_set_reference(void *ctx, const rgh_i420 *recon) {
    if (!ctx) return RGH_ERR_INVALID_ARGUMENT;
    Ctx *c = static_cast<Ctx *>(ctx);
    const int W = c->mb_w * 16, H = c->mb_h * 16;
    rgh_status s = check_i420(recon, W, H);
    if (s != RGH_OK) return s;
    HIP_TRY(hipSetDevice(c->device));
    // The picture goes into the reference slot and is padded there.
    Picture &p = c->pic[c->ref];
    HIP_TRY(hipMemcpy2DAsync(p.org[0], (size_t)c->sy, recon->y, (size_t)recon->stride_y, (size_t)W, (size_t)H,
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpy2DAsync(p.org[1], (size_t)c->sc, recon->u, (size_t)recon->stride_u, (size_t)W / 2, (size_t)H / 2,
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpy2DAsync(p.org[2], (size_t)c->sc, recon->v, (size_t)recon->stride_v, (size_t)W / 2, (size_t)H / 2,
                             hipMemcpyHostToDevice, c->stream));
    RghJob j = make_job(c, nullptr, nullptr, nullptr, 0, 0, 0);
    j.cur_y = p.org[0];
    j.cur_u = p.org[1];
    j.cur_v = p.org[2];
    HIP_TRY(hipMemcpyAsync(c->jobs_dev, &j, sizeof(j), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(rgh_k_extend, dim3((unsigned)c->rows_y, 3, 1), dim3(256), 0, c->stream, c->jobs_dev);
    HIP_TRY(hipGetLastError());
    // An IDR picture is all intra: every co-located vector of the next frame is zero.
    HIP_TRY(hipMemsetAsync(c->dec[1 - c->dcur], 0, sizeof(rgh_mb_decision) * (size_t)c->mb_w * c->mb_h, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->has_ref = true;
    return RGH_OK;
}

rgh_status rgh_analyse_batch(void *const *ctxs, int32_t n, const rgh_i420 *src) {
    if (!ctxs || n <= 0 || !src) return RGH_ERR_INVALID_ARGUMENT;
    std::vector<RghJob> jobs((size_t)n);
    std::vector<Ctx *> cs((size_t)n);
    for (int i = 0; i < n; i++) {
        Ctx *c = static_cast<Ctx *>(ctxs[i]);
        if (!c) return RGH_ERR_INVALID_ARGUMENT;
        if (c->device != static_cast<Ctx *>(ctxs[0])->device) return RGH_ERR_INVALID_ARGUMENT;
        if (!c->has_ref) return RGH_ERR_NO_REFERENCE;
        rgh_status s = check_i420(&src[i], c->width, c->height);
        if (s != RGH_OK) return s;
        cs[i] = c;
        jobs[i] = make_job(c, src[i].y, src[i].u, src[i].v, src[i].stride_y, src[i].stride_u, src[i].stride_v);
    }
    Ctx *c0 = cs[0];
    HIP_TRY(hipSetDevice(c0->device));
    rgh_status s = ensure_jobs(c0, n);
    if (s != RGH_OK) return s;
    return run_jobs(cs.data(), n, jobs, c0->stream, c0->jobs_dev);
}

rgh_status rgh_analyse_p_frame(void *ctx, const rgh_i420 *src, rgh_mb_decision *decisions, int16_t *levels,
                               const rgh_i420_out *recon) {
    if (!ctx) return RGH_ERR_INVALID_ARGUMENT;
    Ctx *c = static_cast<Ctx *>(ctx);
    rgh_status s = check_i420(src, c->width, c->height);
    if (s != RGH_OK) return s;
    if (!c->has_ref) return RGH_ERR_NO_REFERENCE;
    HIP_TRY(hipSetDevice(c->device));
    const int cw = (c->width + 1) / 2, ch = (c->height + 1) / 2;
    HIP_TRY(hipMemcpy2DAsync(c->src[0], (size_t)c->width, src->y, (size_t)src->stride_y, (size_t)c->width,
                             (size_t)c->height, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpy2DAsync(c->src[1], (size_t)cw, src->u, (size_t)src->stride_u, (size_t)cw, (size_t)ch,
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpy2DAsync(c->src[2], (size_t)cw, src->v, (size_t)src->stride_v, (size_t)cw, (size_t)ch,
                             hipMemcpyHostToDevice, c->stream));
    std::vector<RghJob> jobs(1, make_job(c, c->src[0], c->src[1], c->src[2], c->width, cw, cw));
    Ctx *cs[1] = {c};
    s = run_jobs(cs, 1, jobs, c->stream, c->jobs_dev);
    if (s != RGH_OK) return s;
    s = rgh_get_outputs(ctx, decisions, levels, recon, 1);
    if (s != RGH_OK) return s;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return RGH_OK;
}

rgh_status rgh_get_outputs(void *ctx, rgh_mb_decision *decisions, int16_t *levels, const rgh_i420_out *recon,
                           int32_t which) {
    if (!ctx) return RGH_ERR_INVALID_ARGUMENT;
    Ctx *c = static_cast<Ctx *>(ctx);
    if (!c->analysed) return RGH_ERR_NO_REFERENCE;
    HIP_TRY(hipSetDevice(c->device));
    const size_t n = (size_t)c->mb_w * c->mb_h;
    // After a frame the written decisions became "previous" and the written picture the reference.
    if (decisions)
        HIP_TRY(hipMemcpyAsync(decisions, c->dec[1 - c->dcur], n * sizeof(rgh_mb_decision), hipMemcpyDeviceToHost,
                               c->stream));
    if (levels)
        HIP_TRY(hipMemcpyAsync(levels, c->levels, n * RGH_LEVELS_PER_MB * sizeof(int16_t), hipMemcpyDeviceToHost,
                               c->stream));
    if (recon) {
        const int W = c->mb_w * 16, H = c->mb_h * 16;
        const uint8_t *sp[3];
        size_t ss[3];
        if (which == 0) {
            if (!c->keep_pre) return RGH_ERR_INVALID_ARGUMENT;
            for (int k = 0; k < 3; k++) sp[k] = c->pre[k];
            ss[0] = (size_t)W;
            ss[1] = ss[2] = (size_t)W / 2;
        } else {
            for (int k = 0; k < 3; k++) sp[k] = c->pic[c->ref].org[k];
            ss[0] = (size_t)c->sy;
            ss[1] = ss[2] = (size_t)c->sc;
        }
        HIP_TRY(hipMemcpy2DAsync(recon->y, (size_t)recon->stride_y, sp[0], ss[0], (size_t)W, (size_t)H,
                                 hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpy2DAsync(recon->u, (size_t)recon->stride_u, sp[1], ss[1], (size_t)W / 2, (size_t)H / 2,
                                 hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpy2DAsync(recon->v, (size_t)recon->stride_v, sp[2], ss[2], (size_t)W / 2, (size_t)H / 2,
                                 hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return RGH_OK;
}

rgh_status rgh_get_mb_dims(void *ctx, int32_t *mb_w, int32_t *mb_h) {
    if (!ctx || !mb_w || !mb_h) return RGH_ERR_INVALID_ARGUMENT;
    Ctx *c = static_cast<Ctx *>(ctx);
    *mb_w = c->mb_w;
    *mb_h = c->mb_h;
    return RGH_OK;
}

rgh_status rgh_synchronize(void *ctx) {
    if (!ctx) return RGH_ERR_INVALID_ARGUMENT;
    Ctx *c = static_cast<Ctx *>(ctx);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return RGH_OK;
}

void *rgh_get_stream(void *ctx) { return ctx ? static_cast<Ctx *>(ctx)->stream : nullptr; }

rgh_status rgh_enable_timing(void *ctx, int32_t enable) {
    if (!ctx) return RGH_ERR_INVALID_ARGUMENT;
    Ctx *c = static_cast<Ctx *>(ctx);
    HIP_TRY(hipSetDevice(c->device));
    if (enable && !c->ev[0])
        for (auto &e : c->ev) HIP_TRY(hipEventCreate(&e));
    c->timing = enable != 0;
    c->timed = false;
    return RGH_OK;
}

rgh_status rgh_get_timing(void *ctx, float *analyse_ms, float *deblock_ms, int32_t *launches) {
    if (!ctx || !analyse_ms || !deblock_ms || !launches) return RGH_ERR_INVALID_ARGUMENT;
    Ctx *c = static_cast<Ctx *>(ctx);
    if (!c->timed) return RGH_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipEventSynchronize(c->ev[2]));
    HIP_TRY(hipEventElapsedTime(analyse_ms, c->ev[0], c->ev[1]));
    HIP_TRY(hipEventElapsedTime(deblock_ms, c->ev[1], c->ev[2]));
    *launches = c->timed_launches;
    return RGH_OK;
}

}  // extern "C"
