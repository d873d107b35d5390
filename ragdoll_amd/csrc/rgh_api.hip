// Host side of the C ABI declared in include/ragdoll_hip.h: contexts, device buffers and the
// launch schedule of a P frame (analysis wavefront, deblock wavefront, border extension).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "rgh_device.h"

extern "C" __global__ void rgh_k_frame_step(const RghJob *jobs, int j0, int launch, uint32_t seq);
extern "C" __global__ void rgh_k_intra_step(const RghJob *jobs, int j0, int launch, uint32_t seq);
extern "C" __global__ void rgh_k_compact_step(const RghJob *jobs, int j0, int launch, uint32_t seq);
extern "C" __global__ void rgh_k_learned_step(const RghJob *jobs, int j0, int launch, uint32_t seq);
extern "C" __global__ void rgh_k_mixed_step(const RghJob *jobs, int j0, int launch, uint32_t seq);
extern "C" __global__ void rgh_k_mixed_learned_step(const RghJob *jobs, int j0, int launch, uint32_t seq);
extern "C" __global__ void rgh_k_persist(const RghJob *jobs, const RghPersist *P, uint32_t seq);
extern "C" __global__ void rgh_k_persist_learned(const RghJob *jobs, const RghPersist *P, uint32_t seq);
extern "C" __global__ void rgh_k_persist_fast(const RghJob *jobs, const RghPersist *P, uint32_t seq);
extern "C" __global__ void rgh_k_intra_fast_step(const RghJob *jobs, int j0, int launch, uint32_t seq);
extern "C" __global__ void rgh_k_fast_step(const RghJob *jobs, int j0, int launch, uint32_t seq);
extern "C" __global__ void rgh_k_mixed_fast_step(const RghJob *jobs, int j0, int launch, uint32_t seq);
extern "C" __global__ void rgh_k_parts_step(const RghJob *jobs, int j0, int launch, uint32_t seq);
extern "C" __global__ void rgh_k_mixed_parts_step(const RghJob *jobs, int j0, int launch, uint32_t seq);
extern "C" __global__ void rgh_k_xcc_probe(uint32_t *mask);
extern "C" __global__ void rgh_k_recount(const RghJob *jobs);
extern "C" __global__ void rgh_k_extend(const RghJob *jobs);
extern "C" __global__ void rgh_k_fill_u64(uint64_t *p, int n, uint64_t v);

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
// fine: the Compact preset's intra-style rounding of inter residuals (enc.rs:2135 with fine).
void zero_bounds(int qp, bool fine, uint32_t *sad, uint32_t *ssd, uint32_t *probe) {
    static const uint32_t Wt[4] = {1, 2, 1, 2}, Nn[4] = {4, 10, 4, 10};
    uint32_t shift = 15 + (uint32_t)(qp / 6), bias = (1u << shift) / (fine ? 3u : 6u);
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

// One launch of a per-step schedule: kernel, grid (gx x gy workgroups of one wave), the group
// whose stream runs it, and its arguments (job descriptors, first job, launch number L, which is
// also its sequence number relative to the call's seq0).
using StepKernel = void (*)(const RghJob *, int, int, uint32_t);
struct LaunchRec {
    StepKernel fn;
    int gx, gy, group;
    RghJob *jobs;
    int j0, L;
    bool operator==(const LaunchRec &o) const {
        return fn == o.fn && gx == o.gx && gy == o.gy && group == o.group && jobs == o.jobs && j0 == o.j0 && L == o.L;
    }
};

struct Picture {
    uint8_t *base[3] = {nullptr, nullptr, nullptr};  // allocation
    uint8_t *org[3] = {nullptr, nullptr, nullptr};   // sample (0, 0)
    uint64_t *ana = nullptr;  // analysis stamps of the frame written here, mb_w x mb_h
    uint64_t *fin = nullptr;  // deblock/pad item stamps, (mb_w + 1) x (mb_h + 1)
    uint32_t serial = 0;      // frame held (0: none)
};

constexpr int kDefaultLag = 20;  // wavefront steps between consecutive frames of a stream
constexpr int kDeblockLag = 4;   // deblock items run this many steps behind the analysis
#ifndef RGH_MAX_DELAY
#define RGH_MAX_DELAY 4
#endif
constexpr int kMaxDelay = RGH_MAX_DELAY;  // launches a frame stays in the planned range past its end

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int width = 0, height = 0, mb_w = 0, mb_h = 0, profile = 0, qp = 0, me_iters = 16, qpel_iters = 2;
    bool compact = false;  // RGH_FLAG_COMPACT
    bool fast = false;     // RGH_FLAG_FAST
    bool parts = false;    // RGH_FLAG_PARTITIONS
    int cqo = 0;           // chroma_qp_index_offset
    int sy = 0, sc = 0, rows_y = 0, rows_c = 0;
    std::vector<Picture> pic;  // ring of reconstructed pictures
    int ref = 0;               // slot of the reference picture (the last frame analysed)
    uint32_t serial = 0;       // serial number of the reference frame
    rgh_mb_decision *dec[2] = {nullptr, nullptr};  // context-owned decision buffers
    int dcur = 0;              // decisions the next single frame writes
    const rgh_mb_decision *prev = nullptr;  // the reference frame's decisions
    int16_t *levels = nullptr;
    const int16_t *last_levels = nullptr;  // levels of the last analysed frame
    uint8_t *src[3] = {nullptr, nullptr, nullptr};  // staging for host sources
    uint8_t *pre[3] = {nullptr, nullptr, nullptr};  // undeblocked copy of the last frame
    RghJob *jobs_dev = nullptr;
    RghFrameCtl *ctl_dev = nullptr;
    RghFrameCtl *ctl_host = nullptr;  // pinned readback
    int jobs_cap = 0;
    uint8_t *pbuf = nullptr;   // persistent schedule: plan tables, counters, deferral marks, status
    uint8_t *phost = nullptr;  // pinned readback of the deferral marks and status
    size_t pbuf_cap = 0, phost_cap = 0;
    bool has_ref = false;
    bool analysed = false;
    bool keep_pre = true;
    bool pre_valid = false;  // pre[] holds the last analysed frame's undeblocked reconstruction
    uint32_t sb[5] = {0, 0, 0, 0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};  // schedule start, schedule end
    hipEvent_t go = nullptr;                // start of a multi-group call (other groups wait on it)
    bool timing = false, timed = false;
    int timed_launches = 0;
    // The planned launches of the last RGH_FRAMES_GRAPH schedule this context owned, captured,
    // and the launches and stream they were captured from (replayed while both are unchanged).
    hipGraphExec_t graph = nullptr;
    std::vector<LaunchRec> graph_sig;
    hipStream_t graph_stream = nullptr;
};

// Numeric settings read from the environment (test hooks and debugging), dflt when unset.
uint32_t env_u32(const char *name, uint32_t dflt) {
    const char *v = std::getenv(name);
    return v && *v ? (uint32_t)std::strtoul(v, nullptr, 10) : dflt;
}

// Launch numbers stamped into completion stamps; strictly increasing over the process.
std::atomic<uint32_t> g_seq{1};

// Description of the last failure on this thread (rgh_last_error).
thread_local char g_last_error[256] = "";

void set_error(const char *what, const char *detail, int line) {
    std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s (rgh_api.hip:%d)", what, detail, line);
}

rgh_status hip_status(hipError_t e) {
    if (e == hipSuccess) return RGH_OK;
    if (e == hipErrorOutOfMemory) return RGH_ERR_OOM;
    return RGH_ERR_DEVICE;
}
#define HIP_TRY(expr)                                            \
    do {                                                         \
        hipError_t e__ = (expr);                                 \
        if (e__ != hipSuccess) {                                 \
            set_error(#expr, hipGetErrorString(e__), __LINE__);  \
            return hip_status(e__);                              \
        }                                                        \
    } while (0)

void free_ctx(Ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    for (auto &p : c->pic) {
        for (auto &b : p.base)
            if (b) (void)hipFree(b);
        if (p.ana) (void)hipFree(p.ana);
        if (p.fin) (void)hipFree(p.fin);
    }
    for (auto &d : c->dec)
        if (d) (void)hipFree(d);
    if (c->levels) (void)hipFree(c->levels);
    for (auto &b : c->src)
        if (b) (void)hipFree(b);
    for (auto &b : c->pre)
        if (b) (void)hipFree(b);
    if (c->jobs_dev) (void)hipFree(c->jobs_dev);
    if (c->ctl_dev) (void)hipFree(c->ctl_dev);
    if (c->ctl_host) (void)hipHostFree(c->ctl_host);
    if (c->pbuf) (void)hipFree(c->pbuf);
    if (c->phost) (void)hipHostFree(c->phost);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->go) (void)hipEventDestroy(c->go);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int frame_steps(const Ctx *c) { return (c->mb_w - 1) + 2 * (c->mb_h - 1) + 1; }

// Frame k of a call on context c: analysed into slot (ref + 1 + k) % ring, with the frame of
// slot (ref + k) % ring as its reference.
RghJob make_job(Ctx *c, int k, const rgh_i420 &src, rgh_mb_decision *dec, const rgh_mb_decision *prev,
                int16_t *levels, int base, RghFrameCtl *ctl, const RghFrameCtl *wait, bool intra = false) {
    RghJob j;
    std::memset(&j, 0, sizeof(j));
    const int R = (int)c->pic.size();
    const Picture &r = c->pic[(c->ref + k) % R], &w = c->pic[(c->ref + 1 + k) % R];
    j.src_y = src.y;
    j.src_u = src.u;
    j.src_v = src.v;
    j.src_sy = src.stride_y;
    j.src_su = src.stride_u;
    j.src_sv = src.stride_v;
    j.width = c->width;
    j.height = c->height;
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
    j.dec = dec;
    j.prev = prev;
    j.levels = levels;
    j.sb_luma = c->sb[0];
    j.sb_luma_ssd = c->sb[1];
    j.sb_chroma = c->sb[2];
    j.sb_chroma_ssd = c->sb[3];
    j.sb_probe = c->sb[4];
    j.steps = frame_steps(c);
    j.deblock = c->qp != 0 && !c->fast;  // Fast slices disable the loop filter (enc.rs:2334)
    j.compact = c->compact ? 1 : 0;
    j.fast = c->fast ? 1 : 0;
    j.parts = c->parts ? 1 : 0;
    j.cqo = c->cqo;
    j.base = base;
    j.serial = c->serial + 1 + (uint32_t)k;
    j.ref_serial = c->serial + (uint32_t)k;
    j.ana = w.ana;
    j.fin = w.fin;
    j.ref_fin = r.fin;
    j.ref_job = j.wait_job = -1;
    if (intra) {  // an IDR picture reads no reference
        j.intra = 1;
        j.ref_fin = nullptr;
        j.prev = dec;
    }
    j.ctl = ctl;
    if (wait) {
        j.wait_analysed = &wait->analysed;
        j.wait_total = c->mb_w * c->mb_h;
    }
    return j;
}

rgh_status ensure_jobs(Ctx *c, int n);

// The launch schedule of `jobs` (frame-major: job k * n + i is frame k of context i, planned to
// start at launch k * lag), on the owner context's stream with its job and progress buffers.
// Launch L runs every frame planned to be active then; a frame that fell behind (deferred
// macroblocks) keeps running in later launches, and after the planned launches the per-frame
// delays are read back and the launches continue until every frame has run all of its steps.
// Several schedules (disjoint contexts) can interleave their launches on their own streams.
struct Schedule {
    Ctx *owner = nullptr;
    // Device job descriptors and progress records of this schedule (nj each), and the host
    // buffer the progress records are read back into: the owner context's, or scratch memory.
    RghJob *jobs_dev = nullptr;
    RghFrameCtl *ctl_dev = nullptr;
    RghFrameCtl *ctl_host = nullptr;
    int n = 0, n_frames = 0, lag = 1, nj = 0, n_intra = 0;
    bool compact = false, fast = false, parts = false, learned = false;
    uint32_t seq0 = 0;  // sequence number of the call's launch 0
    // While the planned launches are listed, launch() appends them here (group `group`) instead
    // of issuing them.
    std::vector<LaunchRec> *out = nullptr;
    int group = 0;
    bool recount = false;  // recount analysed macroblocks from their stamps (after a persistent launch)
    std::vector<RghJob> jobs;
    std::vector<int> span, last;
    int gx = 1, max_span = 0, planned = 0, L = 0, extra_total = 0;

    // Checks the jobs and plans the launches (span of each frame, planned launch count).
    rgh_status plan() {
        nj = n * n_frames;
        // IDR frames: frame-major jobs, every context's frame k of one kind (IDR or P).
        n_intra = 0;
        for (int j = 0; j < nj; j++) n_intra += jobs[(size_t)j].intra != 0;
        for (int j = 0; j < nj; j++)
            if ((jobs[(size_t)j].intra != 0) != (jobs[(size_t)(j - j % n)].intra != 0)) {
                set_error("run_schedule", "a frame index must be IDR in every context or in none", __LINE__);
                return RGH_ERR_INVALID_ARGUMENT;
            }
        // One P kernel per schedule: every job of it uses one preset.
        compact = nj > 0 && jobs[0].compact;
        fast = nj > 0 && jobs[0].fast;
        parts = nj > 0 && jobs[0].parts;
        for (int j = 0; j < nj; j++)
            if ((jobs[(size_t)j].compact != 0) != compact || (jobs[(size_t)j].fast != 0) != fast ||
                (jobs[(size_t)j].parts != 0) != parts) {
                set_error("run_schedule", "contexts of one call must use one preset", __LINE__);
                return RGH_ERR_INVALID_ARGUMENT;
            }
        learned = false;
        for (int j = 0; j < nj; j++)
            if (!jobs[(size_t)j].intra) {
                learned = jobs[(size_t)j].prop != nullptr;
                break;
            }
        for (int j = 0; j < nj; j++)
            if (!jobs[(size_t)j].intra && (jobs[(size_t)j].prop != nullptr) != learned) {
                set_error("run_schedule", "every P frame of a learned call needs proposals", __LINE__);
                return RGH_ERR_INVALID_ARGUMENT;
            }
        span.assign((size_t)nj, 0);
        last.assign((size_t)nj, -1);
        for (int j = 0; j < nj; j++) {
            const int mb_w = jobs[(size_t)j].mb_w, mb_h = jobs[(size_t)j].mb_h;
            // The analysis steps, then the deblock/pad items (their grid reaches 3 steps
            // further) kDeblockLag behind.
            span[(size_t)j] = jobs[(size_t)j].steps + 3 + kDeblockLag;
            max_span = std::max(max_span, span[(size_t)j]);
            // Items per step: analyses min(mb_h, ceil(mb_w / 2)), deblock/pad items one more
            // at most; a grid row is whole groups of the XCD band mapping.
            const int items = std::max(std::min(mb_h, (mb_w + 1) / 2), std::min(mb_h + 1, (mb_w + 2) / 2));
            gx = std::max(gx, (items + RGH_ITEM_GROUP - 1) / RGH_ITEM_GROUP * RGH_ITEM_GROUP);
        }
        planned = (n_frames - 1) * lag + max_span;
        return RGH_OK;
    }
    // Uploads the job descriptors (with the call's seq0) and clears the progress records.
    rgh_status upload() {
        for (auto &j : jobs) j.seq0 = seq0;
        HIP_TRY(hipMemcpyAsync(jobs_dev, jobs.data(), sizeof(RghJob) * (size_t)nj, hipMemcpyHostToDevice,
                               owner->stream));
        // The progress records start at zero (a fill kernel, not a memset node, so that the
        // sequence stays correct under graph capture).
        static_assert(sizeof(RghFrameCtl) == 16, "RghFrameCtl is two 64-bit words");
        const int words = 2 * nj;
        hipLaunchKernelGGL(rgh_k_fill_u64, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, owner->stream,
                           (uint64_t *)ctl_dev, words, (uint64_t)0);
        if (recount) hipLaunchKernelGGL(rgh_k_recount, dim3((unsigned)nj), dim3(64), 0, owner->stream, jobs_dev);
        HIP_TRY(hipGetLastError());
        return RGH_OK;
    }
    // One kernel of launch L over jobs [j0, j0 + nf) (grid gx x 2 nf: the frames' analysis rows,
    // then their deblock/pad rows): listed while the planned launches are listed (relative
    // sequence number L), otherwise issued with a fresh sequence number, above every number of
    // the planned launches.
    void emit(StepKernel fn, int j0, int nf) {
        const int gy = 2 * nf;
        if (out) {
            out->push_back(LaunchRec{fn, gx, gy, group, jobs_dev, j0, L});
            return;
        }
        hipLaunchKernelGGL(fn, dim3((unsigned)gx, (unsigned)gy), dim3(64), 0, owner->stream, jobs_dev, j0, L,
                           g_seq.fetch_add(1) - seq0);
    }
    void launch(int j_lo, int j_hi) {
        // IDR jobs and P jobs of the launch (whole frames, frame-major), in one sequence number: a
        // stamp of one is not final for the other within the launch.
        bool any_i = false, any_p = false;
        for (int f = j_lo / n; f <= (j_hi - 1) / n; f++) (jobs[(size_t)f * n].intra ? any_i : any_p) = true;
        StepKernel pk = learned ? rgh_k_learned_step
                                : (compact ? rgh_k_compact_step
                                           : (fast ? rgh_k_fast_step : (parts ? rgh_k_parts_step : rgh_k_frame_step)));
        StepKernel ik = fast ? rgh_k_intra_fast_step : rgh_k_intra_step;
        if (any_i && any_p && !compact) {
            // IDR and Balanced P jobs together: one kernel, so that neither waits for the other.
            emit(learned ? rgh_k_mixed_learned_step
                         : (fast ? rgh_k_mixed_fast_step : (parts ? rgh_k_mixed_parts_step : rgh_k_mixed_step)),
                 j_lo, j_hi - j_lo);
        } else {
            // One kernel per run of frames of one kind (Compact: its IDR and P kernels).
            for (int a = j_lo; a < j_hi;) {
                const bool intra = jobs[(size_t)a].intra != 0;
                int b = (a / n + 1) * n;  // the end of a's frame, then of the following frames of its kind
                while (b < j_hi && (jobs[(size_t)b].intra != 0) == intra) b += n;
                b = std::min(b, j_hi);
                emit(intra ? ik : pk, a, b - a);
                a = b;
            }
        }
        for (int j = j_lo; j < j_hi; j++) last[(size_t)j] = L;
        L++;
    }
    // The next planned launch (frame k active over [k * lag, k * lag + span + kMaxDelay)).
    void planned_launch() {
        const int k_hi = std::min(n_frames - 1, L / lag);
        int k_lo = 0;
        while (k_lo <= k_hi && k_lo * lag + max_span + kMaxDelay <= L) k_lo++;
        if (k_lo > k_hi) {
            L++;
            return;
        }
        launch(k_lo * n, (k_hi + 1) * n);
    }
    // Completion: frames that fell behind run on until done. A frame's next step after its last
    // launch P is P + 1 - base - delay[(P + 1) & 1]; rejoining at launch L it needs the delay
    // L - base - next.
    rgh_status complete() {
        for (;;) {
            HIP_TRY(hipMemcpyAsync(ctl_host, ctl_dev, sizeof(RghFrameCtl) * (size_t)nj, hipMemcpyDeviceToHost,
                                   owner->stream));
            HIP_TRY(hipStreamSynchronize(owner->stream));
            int lo = nj, hi = -1, need = 0;
            for (int j = 0; j < nj; j++) {
                const int base = jobs[(size_t)j].base, P = last[(size_t)j];
                const int next = P < 0 ? 0 : P + 1 - base - ctl_host[j].delay[(P + 1) & 1];
                if (next >= span[(size_t)j]) continue;
                lo = std::min(lo, j);
                hi = std::max(hi, j);
                need = std::max(need, span[(size_t)j] - next);
                ctl_host[j].delay[0] = ctl_host[j].delay[1] = L - base - next;
            }
            if (hi < 0) return RGH_OK;
            extra_total += need;
            if (extra_total > 64 * max_span + 4096) {  // no progress: a defect
                set_error("run_schedule", "frames stopped making progress", __LINE__);
                return RGH_ERR_DEVICE;
            }
            HIP_TRY(hipMemcpyAsync(ctl_dev + lo, ctl_host + lo, sizeof(RghFrameCtl) * (size_t)(hi - lo + 1),
                                   hipMemcpyHostToDevice, owner->stream));
            for (int e = 0; e < need; e++) launch(lo, hi + 1);
            HIP_TRY(hipGetLastError());
        }
    }
};

// Captures the listed launches of one schedule (on its owner's stream) into an executable graph,
// or returns the graph captured last time from the same launches on the same stream. Kernel
// nodes only: the descriptors and progress records are uploaded and cleared outside the graph
// before each replay. Each schedule has a graph of its own on its own stream, so that the
// schedules' launches run concurrently as the streams run them.
rgh_status schedule_graph(Schedule &sc, const std::vector<LaunchRec> &recs, hipGraphExec_t *exec) {
    Ctx *c = sc.owner;
    hipStream_t st = c->stream;
    if (c->graph && c->graph_sig == recs && c->graph_stream == st) {
        *exec = c->graph;
        return RGH_OK;
    }
    if (c->graph) {
        HIP_TRY(hipGraphExecDestroy(c->graph));
        c->graph = nullptr;
    }
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < recs.size() && e == hipSuccess; i++) {
        const LaunchRec &r = recs[i];
        hipLaunchKernelGGL(r.fn, dim3((unsigned)r.gx, (unsigned)r.gy), dim3(64), 0, st, r.jobs, r.j0, r.L,
                           (uint32_t)r.L);
        e = hipGetLastError();
    }
    hipGraph_t graph = nullptr;
    const hipError_t ee = hipStreamEndCapture(st, &graph);
    if (e == hipSuccess) e = ee;
    if (e == hipSuccess) e = hipGraphInstantiate(&c->graph, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        c->graph = nullptr;
        set_error("graph capture", hipGetErrorString(e), __LINE__);
        return RGH_ERR_DEVICE;
    }
    c->graph_sig = recs;
    c->graph_stream = st;
    *exec = c->graph;
    return RGH_OK;
}

// Runs schedules concurrently (their launches interleaved across their streams) and waits for
// all of them. The planned launches take sequence numbers seq0 + L from one range reserved for
// the call; with `graph` they are issued as one captured graph (replayed while the launch list
// is unchanged), otherwise one by one. Timing, when enabled on `timer`, spans the whole run.
rgh_status run_schedules(std::vector<Schedule> &sch, Ctx *timer, bool record = true, bool graph = false) {
    if (record && timer->timing) HIP_TRY(hipEventRecord(timer->ev[0], timer->stream));
    int most = 0;
    for (auto &s : sch) {
        rgh_status st = s.plan();
        if (st != RGH_OK) return st;
        most = std::max(most, s.planned);
    }
    const uint32_t seq0 = g_seq.fetch_add((uint32_t)most + 1);
    for (auto &s : sch) {
        s.seq0 = seq0;
        rgh_status st = s.upload();
        if (st != RGH_OK) return st;
    }
    std::vector<LaunchRec> recs;
    for (size_t g = 0; g < sch.size(); g++) {
        sch[g].out = &recs;
        sch[g].group = (int)g;
    }
    for (int L = 0; L < most; L++)
        for (auto &s : sch)
            if (L < s.planned) s.planned_launch();
    for (auto &s : sch) s.out = nullptr;
    if (graph) {
        for (size_t g = 0; g < sch.size(); g++) {
            std::vector<LaunchRec> mine;
            for (const LaunchRec &r : recs)
                if (r.group == (int)g) mine.push_back(r);
            hipGraphExec_t exec = nullptr;
            rgh_status st = schedule_graph(sch[g], mine, &exec);
            if (st != RGH_OK) return st;
            HIP_TRY(hipGraphLaunch(exec, sch[g].owner->stream));
        }
    } else {
        for (const LaunchRec &r : recs)
            hipLaunchKernelGGL(r.fn, dim3((unsigned)r.gx, (unsigned)r.gy), dim3(64), 0,
                               sch[(size_t)r.group].owner->stream, r.jobs, r.j0, r.L, (uint32_t)r.L);
    }
    HIP_TRY(hipGetLastError());
    if (env_u32("RGH_TEST_FAIL_AFTER_LAUNCH", 0)) {
        // Test hook: a failure once launches have started (their work completes first).
        for (auto &s : sch) HIP_TRY(hipStreamSynchronize(s.owner->stream));
        set_error("run_schedule", "failure forced after the launches started (RGH_TEST_FAIL_AFTER_LAUNCH)", __LINE__);
        return RGH_ERR_DEVICE;
    }
    int launches = 0;
    for (auto &s : sch) {
        rgh_status st = s.complete();
        if (st != RGH_OK) return st;
        launches = std::max(launches, s.L);
    }
    if (record && timer->timing) {
        HIP_TRY(hipEventRecord(timer->ev[1], timer->stream));
        timer->timed = true;
        timer->timed_launches = launches;
    }
    return RGH_OK;
}

rgh_status run_schedule(Ctx *const *ctxs, int n, int n_frames, std::vector<RghJob> &jobs, int lag) {
    std::vector<Schedule> sch(1);
    sch[0].owner = ctxs[0];
    sch[0].jobs_dev = ctxs[0]->jobs_dev;
    sch[0].ctl_dev = ctxs[0]->ctl_dev;
    sch[0].ctl_host = ctxs[0]->ctl_host;
    sch[0].n = n;
    sch[0].n_frames = n_frames;
    sch[0].lag = lag;
    sch[0].jobs = jobs;
    return run_schedules(sch, ctxs[0]);
}

// ---- The persistent schedule (rgh_k_persist) ----
// One launch runs every item of a call. Each stream (context) is homed on one XCD; the items of
// an XCD's streams form its queue of tickets ordered by planned time. A frame's items span
// times [base, base + steps + 7): analysis step s at base + s, deblock/pad step sd at base + sd +
// 4. Pipelined plan: frame k of a stream starts at k * lag, and not before the frame that last
// read its picture slot has finished its analysis. An item that would have to wait for a later
// item defers (its frame is marked); frames with a mark are then finished by a second launch
// whose plan runs each stream's remaining frames one after another, where nothing can defer.

constexpr int kMaxDevices = 16;
constexpr int kFrameSpanExtra = 7;  // deblock/pad steps reach 3 past the analysis, 4 behind it

struct XccMap {
    std::once_flag once;
    rgh_status st = RGH_OK;
    int n = 0;                    // XCC ids seen
    int xcc[RGH_MAX_QUEUES] = {}; // queue q is served by the waves of XCC xcc[q]
};
XccMap g_xcc[kMaxDevices];

// The XCC ids the device's waves report (probed once per device with a grid large enough to
// reach every XCD).
rgh_status xcc_map(Ctx *c, const XccMap **out) {
    if (c->device < 0 || c->device >= kMaxDevices) {
        set_error("persistent schedule", "device ordinal beyond the XCC table", __LINE__);
        return RGH_ERR_UNSUPPORTED;
    }
    XccMap &m = g_xcc[c->device];
    std::call_once(m.once, [&] {
        uint32_t *d = nullptr, h = 0;
        if (hipMalloc(&d, sizeof(uint32_t)) != hipSuccess) {
            m.st = RGH_ERR_OOM;
            return;
        }
        bool ok = hipMemsetAsync(d, 0, sizeof(uint32_t), c->stream) == hipSuccess;
        if (ok) {
            hipLaunchKernelGGL(rgh_k_xcc_probe, dim3(4096), dim3(64), 0, c->stream, d);
            ok = hipGetLastError() == hipSuccess;
        }
        ok = ok && hipMemcpyAsync(&h, d, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
             hipStreamSynchronize(c->stream) == hipSuccess;
        (void)hipFree(d);
        if (!ok || !h) {
            m.st = RGH_ERR_DEVICE;
            return;
        }
        for (int x = 0; x < RGH_MAX_QUEUES; x++)
            if ((h >> x) & 1) m.xcc[m.n++] = x;
    });
    if (m.st != RGH_OK) {
        set_error("persistent schedule", "XCC probe failed", __LINE__);
        return m.st;
    }
    *out = &m;
    return RGH_OK;
}

int n_analysis_at(int s, int mb_w, int mb_h, int steps) {
    if (s < 0 || s >= steps) return 0;
    const int a = s - mb_w + 1, y0 = a > 0 ? (a + 1) >> 1 : 0;
    return std::min(mb_h - 1, s >> 1) - y0 + 1;
}
int n_items_at(int sd, int mb_w, int mb_h, int steps) {
    if (sd < 0 || sd > steps + 2) return 0;
    const int a = sd - mb_w, y0 = a > 0 ? (a + 1) >> 1 : 0;
    return std::min(mb_h, sd >> 1) - y0 + 1;
}
int items_at(const RghJob &j, int T) {
    const int s = T - j.base;
    return n_analysis_at(s, j.mb_w, j.mb_h, j.steps) + n_items_at(s - 4, j.mb_w, j.mb_h, j.steps);
}

// The device image of a plan: RghPersist, then its tables; after them (not uploaded, cleared on
// the device) the ticket counters, the deferral marks and the status words.
struct PersistPlan {
    std::vector<uint8_t> img;
    size_t mut_off = 0, mut_bytes = 0, deferred_off = 0, total = 0;
    uint32_t max_queue = 0;
};

size_t align8(size_t v) { return (v + 7) & ~(size_t)7; }

void build_plan(const std::vector<RghJob> &lj, const std::vector<int> &queue_of_job, const XccMap &xm, int nq,
                uint8_t *dev, uint32_t spin_limit, uint32_t jitter, PersistPlan &pl) {
    const int nj = (int)lj.size();
    std::vector<int> t0((size_t)nq, INT32_MAX), t1((size_t)nq, INT32_MIN);
    for (int j = 0; j < nj; j++) {
        const int q = queue_of_job[(size_t)j];
        t0[(size_t)q] = std::min(t0[(size_t)q], lj[(size_t)j].base);
        t1[(size_t)q] = std::max(t1[(size_t)q], lj[(size_t)j].base + lj[(size_t)j].steps + kFrameSpanExtra - 1);
    }
    std::vector<uint32_t> slot0((size_t)nq, 0), nslots((size_t)nq, 0);
    size_t S = 0;
    for (int q = 0; q < nq; q++) {
        slot0[(size_t)q] = (uint32_t)S;
        nslots[(size_t)q] = t1[(size_t)q] >= t0[(size_t)q] ? (uint32_t)(t1[(size_t)q] - t0[(size_t)q] + 1) : 0;
        S += nslots[(size_t)q] + 1;
    }
    // Segments (job, time) per slot, in job order within a slot.
    std::vector<uint32_t> slot_seg(S + 1, 0), slot_tl(S, 0);
    for (int j = 0; j < nj; j++) {
        const RghJob &J = lj[(size_t)j];
        const int q = queue_of_job[(size_t)j];
        for (int T = J.base; T < J.base + J.steps + kFrameSpanExtra; T++)
            if (items_at(J, T) > 0) slot_seg[slot0[(size_t)q] + (size_t)(T - t0[(size_t)q]) + 1]++;
    }
    for (size_t i = 0; i < S; i++) slot_seg[i + 1] += slot_seg[i];
    const size_t E = slot_seg[S];
    std::vector<uint32_t> seg_job(E), seg_cnt(E), fillp(slot_seg.begin(), slot_seg.end() - 1);
    for (int j = 0; j < nj; j++) {
        const RghJob &J = lj[(size_t)j];
        const int q = queue_of_job[(size_t)j];
        for (int T = J.base; T < J.base + J.steps + kFrameSpanExtra; T++) {
            const int c = items_at(J, T);
            if (c <= 0) continue;
            const size_t e = fillp[slot0[(size_t)q] + (size_t)(T - t0[(size_t)q])]++;
            seg_job[e] = (uint32_t)j;
            seg_cnt[e] = (uint32_t)c;
        }
    }
    // Tickets: each queue numbers its slots' segments in order.
    std::vector<uint32_t> seg_tl(E);
    RghPersist P;
    std::memset(&P, 0, sizeof(P));
    P.n_queues = nq;
    P.n_xcc = xm.n;
    for (int x = 0; x < RGH_MAX_QUEUES; x++) P.queue_of_xcc[x] = -1;
    for (int q = 0; q < nq; q++) P.queue_of_xcc[xm.xcc[q]] = q;
    P.spin_limit = spin_limit;
    P.jitter = jitter;
    P.debug = env_u32("RGH_PERSIST_DEBUG", 0);
    pl.max_queue = 0;
    uint32_t claims = 0;
    for (int q = 0; q < nq; q++) {
        uint32_t tk = 0;
        const size_t s0 = slot0[(size_t)q];
        for (uint32_t i = 0; i <= nslots[(size_t)q]; i++) {
            slot_tl[s0 + i] = tk;
            if (i == nslots[(size_t)q]) break;
            for (size_t e = slot_seg[s0 + i]; e < slot_seg[s0 + i + 1]; e++) {
                seg_tl[e] = tk;
                tk += seg_cnt[e];
            }
        }
        P.q[q].total = tk;
        P.q[q].t0 = t0[(size_t)q];
        P.q[q].slot0 = (uint32_t)s0;
        P.q[q].n_slots = nslots[(size_t)q];
        P.q[q].claim0 = claims;
        claims += tk;
        pl.max_queue = std::max(pl.max_queue, tk);
    }
    // Layout.
    size_t off = align8(sizeof(RghPersist));
    const size_t o_tl = off;
    off = align8(off + 4 * S);
    const size_t o_sseg = off;
    off = align8(off + 4 * (S + 1));
    const size_t o_sjob = off;
    off = align8(off + 4 * E);
    const size_t o_stl = off;
    off = align8(off + 4 * E);
    pl.mut_off = off;
    off = (off + 255) & ~(size_t)255;
    pl.mut_off = off;
    const size_t o_claim = off;
    off = align8(off + 4 * (size_t)claims);
    pl.deferred_off = off;
    off = align8(off + 4 * (size_t)nj);
    const size_t o_status = off;
    off = align8(off + 4 * 8);
    pl.mut_bytes = off - pl.mut_off;
    pl.total = off;
    P.slot_tl = (const uint32_t *)(dev + o_tl);
    P.slot_seg = (const uint32_t *)(dev + o_sseg);
    P.seg_job = (const uint32_t *)(dev + o_sjob);
    P.seg_tl = (const uint32_t *)(dev + o_stl);
    P.claim = (uint32_t *)(dev + o_claim);
    P.deferred = (int32_t *)(dev + pl.deferred_off);
    P.status = (uint32_t *)(dev + o_status);
    pl.img.assign(pl.mut_off, 0);
    std::memcpy(pl.img.data(), &P, sizeof(P));
    std::memcpy(pl.img.data() + o_tl, slot_tl.data(), 4 * S);
    std::memcpy(pl.img.data() + o_sseg, slot_seg.data(), 4 * (S + 1));
    std::memcpy(pl.img.data() + o_sjob, seg_job.data(), 4 * E);
    std::memcpy(pl.img.data() + o_stl, seg_tl.data(), 4 * E);
}

// One persistent launch over lj (bases, ref_job and wait_job already in lj's numbering) on the
// owner's stream; returns after it finishes, with the deferral marks (per job), the items
// deferred and the give-up status (1: a wait reached its bound, 2: a workgroup ran off its XCD).
rgh_status persistent_launch(Ctx *owner, const XccMap &xm, std::vector<RghJob> &lj, const std::vector<int> &queue_of_job,
                             int nq, bool learned, std::vector<int32_t> &marks, uint32_t &deferred, uint32_t &gave_up) {
    const int nj = (int)lj.size();
    rgh_status st = ensure_jobs(owner, nj);
    if (st != RGH_OK) return st;
    PersistPlan pl;
    // A wait that sleeps this many times (about a second) means the schedule cannot advance.
    const uint32_t spin = env_u32("RGH_PERSIST_SPIN_LIMIT", 1u << 20), jitter = env_u32("RGH_PERSIST_JITTER", 0);
    build_plan(lj, queue_of_job, xm, nq, nullptr, spin, jitter, pl);
    if (owner->pbuf_cap < pl.total) {
        if (owner->pbuf) HIP_TRY(hipFree(owner->pbuf));
        owner->pbuf = nullptr;
        owner->pbuf_cap = 0;
        const size_t cap = pl.total + pl.total / 4 + 4096;
        HIP_TRY(hipMalloc(&owner->pbuf, cap));
        owner->pbuf_cap = cap;
    }
    const size_t rb = pl.total - pl.deferred_off;  // deferral marks, status
    if (owner->phost_cap < rb) {
        if (owner->phost) HIP_TRY(hipHostFree(owner->phost));
        owner->phost = nullptr;
        owner->phost_cap = 0;
        HIP_TRY(hipHostMalloc(&owner->phost, rb + 4096, hipHostMallocDefault));
        owner->phost_cap = rb + 4096;
    }
    build_plan(lj, queue_of_job, xm, nq, owner->pbuf, spin, jitter, pl);  // again, with device addresses
    HIP_TRY(hipMemcpyAsync(owner->jobs_dev, lj.data(), sizeof(RghJob) * (size_t)nj, hipMemcpyHostToDevice,
                           owner->stream));
    HIP_TRY(hipMemcpyAsync(owner->pbuf, pl.img.data(), pl.img.size(), hipMemcpyHostToDevice, owner->stream));
    const int words = (int)(pl.mut_bytes / 8);
    hipLaunchKernelGGL(rgh_k_fill_u64, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, owner->stream,
                       (uint64_t *)(owner->pbuf + pl.mut_off), words, (uint64_t)0);
    const uint32_t seq = g_seq.fetch_add(1);
    // At least max_queue workgroups per XCD (dispatched round-robin over the XCDs).
    const uint64_t grid = (uint64_t)xm.n * pl.max_queue;
    if (grid > 0x7fffffffu) {
        set_error("persistent schedule", "too many items in one call", __LINE__);
        return RGH_ERR_INVALID_ARGUMENT;
    }
    if (grid)
        hipLaunchKernelGGL(learned ? rgh_k_persist_learned : (lj[0].fast ? rgh_k_persist_fast : rgh_k_persist),
                           dim3((unsigned)grid), dim3(64), 0,
                           owner->stream, owner->jobs_dev, (const RghPersist *)owner->pbuf, seq);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(owner->phost, owner->pbuf + pl.deferred_off, rb, hipMemcpyDeviceToHost, owner->stream));
    HIP_TRY(hipStreamSynchronize(owner->stream));
    const int32_t *m = (const int32_t *)owner->phost;
    marks.assign(m, m + nj);
    const uint32_t *status = (const uint32_t *)(owner->phost + (rb - 32));
    deferred = status[1];
    gave_up = status[0];
    if (env_u32("RGH_PERSIST_DEBUG", 0)) {
        RghPersist P;
        std::memcpy(&P, pl.img.data(), sizeof(P));
        uint64_t items = 0;
        for (int q = 0; q < nq; q++) items += P.q[q].total;
        std::fprintf(stderr, "persist: xcc ids %d:", xm.n);
        for (int q = 0; q < xm.n; q++) std::fprintf(stderr, " %d", xm.xcc[q]);
        std::fprintf(stderr, "; queues %d:", nq);
        for (int q = 0; q < nq; q++) std::fprintf(stderr, " %u", P.q[q].total);
        std::fprintf(stderr, "; jobs %d; items %llu run %u deferred %u gave_up %u spins %u; grid %llu; wait/item %.1f us "
                     "run/item %.1f us (analysis %.1f us)\n", nj,
                     (unsigned long long)items, status[2], status[1], status[0], status[3],
                     (unsigned long long)xm.n * pl.max_queue, status[4] * 256.0 / 2400.0 / std::max(1u, status[2]),
                     status[5] * 256.0 / 2400.0 / std::max(1u, status[2]), status[6] * 256.0 / 2400.0 / std::max(1u, status[7]));
    }
    return RGH_OK;
}

// Runs a call's schedule S (jobs frame-major: job k * n + i is frame k of context i, ctl pointers
// into the owner's progress records) with one persistent launch. When an item deferred, or the
// launch gave up, the per-step schedule finishes the call (items already done skip themselves;
// the analysed counts of its picture-slot gate are recounted from the stamps first).
// `ring[i]`: context i's ring size.
rgh_status run_persistent(Schedule &S, int lag, bool learned, const std::vector<int> &ring, Ctx *timer) {
    Ctx *owner = S.owner;
    const int n = S.n, n_frames = S.n_frames, nj = n * n_frames;
    const XccMap *xm = nullptr;
    rgh_status st = xcc_map(owner, &xm);
    if (st != RGH_OK) return st;
    const int nq = std::min(xm->n, n);
    std::vector<RghJob> jobs = S.jobs;
    // Streams homed on queues (XCDs) by a greedy balance of their macroblocks.
    std::vector<int> order((size_t)n), queue_of_stream((size_t)n, 0);
    std::vector<double> load((size_t)nq, 0.0);
    for (int i = 0; i < n; i++) order[(size_t)i] = i;
    auto work = [&](int i) { return (double)jobs[(size_t)i].mb_w * jobs[(size_t)i].mb_h; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    for (int i : order) {
        const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        queue_of_stream[(size_t)i] = q;
        load[(size_t)q] += work(i);
    }
    // The pipelined plan: frame k at k * lag, after the frame gating its picture slot.
    std::vector<int> qj((size_t)nj);
    for (int k = 0; k < n_frames; k++)
        for (int i = 0; i < n; i++) {
            RghJob &J = jobs[(size_t)k * n + i];
            const int wk = k + 1 - ring[(size_t)i];
            J.ref_job = (k > 0 && !J.intra) ? (k - 1) * n + i : -1;
            J.wait_job = wk >= 0 ? wk * n + i : -1;
            J.base = k * lag;
            if (wk >= 0) J.base = std::max(J.base, jobs[(size_t)wk * n + i].base + J.steps);
            qj[(size_t)k * n + i] = queue_of_stream[(size_t)i];
        }
    std::vector<int32_t> marks;
    uint32_t deferred = 0, gave_up = 0;
    st = persistent_launch(owner, *xm, jobs, qj, nq, learned, marks, deferred, gave_up);
    timer->timed_launches = 1;
    if (st != RGH_OK || (deferred == 0 && gave_up == 0)) return st;
    std::vector<Schedule> sch(1, S);
    sch[0].recount = true;
    st = run_schedules(sch, timer, false);
    timer->timed_launches = 1 + sch[0].L;
    return st;
}

// After a call failed once its launches may have started: stamps of this call's serials may be
// written and its picture slots (the reference's included, when the ring is short) overwritten,
// so the context takes no P frame until a new reference or IDR, and later calls use serials
// beyond every one this call could have stamped.
void poison(Ctx *c, int n_frames) {
    c->serial += (uint32_t)n_frames + 1;
    c->has_ref = false;
    c->analysed = false;
    c->pre_valid = false;
}

// After a call analysed n_frames frames of c: the last one is the reference.
void advance(Ctx *c, int n_frames, const rgh_mb_decision *last_dec) {
    const int R = (int)c->pic.size();
    for (int k = 0; k < n_frames; k++) c->pic[(c->ref + 1 + k) % R].serial = c->serial + 1 + (uint32_t)k;
    c->ref = (c->ref + n_frames) % R;
    c->serial += (uint32_t)n_frames;
    c->prev = last_dec;
    c->analysed = true;
}

rgh_status ensure_jobs(Ctx *c, int n) {
    if (c->jobs_cap >= n) return RGH_OK;
    if (c->jobs_dev) HIP_TRY(hipFree(c->jobs_dev));
    if (c->ctl_dev) HIP_TRY(hipFree(c->ctl_dev));
    if (c->ctl_host) HIP_TRY(hipHostFree(c->ctl_host));
    c->jobs_dev = nullptr;
    c->ctl_dev = nullptr;
    c->ctl_host = nullptr;
    c->jobs_cap = 0;
    const int cap = std::max(n, 8);
    HIP_TRY(hipMalloc(&c->jobs_dev, sizeof(RghJob) * (size_t)cap));
    HIP_TRY(hipMalloc(&c->ctl_dev, sizeof(RghFrameCtl) * (size_t)cap));
    HIP_TRY(hipHostMalloc(&c->ctl_host, sizeof(RghFrameCtl) * (size_t)cap, hipHostMallocDefault));
    c->jobs_cap = cap;
    return RGH_OK;
}

// ---- Scratch memory of rgh_analyse_frames_noalloc ----
// The job descriptors and progress records of one call live in this file-scope device array
// instead of buffers the library allocates: RGH_NOALLOC_MAX_FRAMES of each (about 0.5 MB per
// device, held until the process ends). Host code never passes the array's name to the device
// (on the host it is only a stand-in): the address comes from hipGetSymbolAddress, once per
// device with that device current, and every launch, copy and fill uses that address. Every
// call on a device shares its array (and the matching host readback buffer below), so calls
// that overlap on one device race; each call rewrites the descriptors and clears the progress
// records it uses before its first launch.
constexpr int kScratchDevices = 16;
constexpr size_t kScratchJobBytes = (sizeof(RghJob) * RGH_NOALLOC_MAX_FRAMES + 255) & ~(size_t)255;
constexpr size_t kScratchBytes = kScratchJobBytes + sizeof(RghFrameCtl) * RGH_NOALLOC_MAX_FRAMES;

struct ScratchDevice {
    std::once_flag once;
    void *ptr = nullptr;
    hipError_t err = hipSuccess;
};
ScratchDevice g_scratch_dev[kScratchDevices];
// Host readback of the progress records, per device (ordinary host memory: read with a copy
// that the stream synchronizes).
RghFrameCtl g_scratch_ctl_host[kScratchDevices][RGH_NOALLOC_MAX_FRAMES];

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

// The device address of `scratch` on `device` (the current device), looked up on the first call.
rgh_status scratch_address(int device, uint8_t **out) {
    if (device < 0 || device >= kScratchDevices) {
        set_error("rgh_analyse_frames_noalloc", "device ordinal beyond the scratch table", __LINE__);
        return RGH_ERR_UNSUPPORTED;
    }
    ScratchDevice &d = g_scratch_dev[device];
    std::call_once(d.once, [&] { d.err = hipGetSymbolAddress(&d.ptr, HIP_SYMBOL(scratch)); });
    if (d.err != hipSuccess || !d.ptr) {
        set_error("hipGetSymbolAddress(scratch)", hipGetErrorString(d.err), __LINE__);
        return RGH_ERR_DEVICE;
    }
    *out = static_cast<uint8_t *>(d.ptr);
    return RGH_OK;
}

rgh_status check_i420(const rgh_i420 *p, int w, int h) {
    if (!p || !p->y || !p->u || !p->v) return RGH_ERR_INVALID_ARGUMENT;
    if (p->width != w || p->height != h) return RGH_ERR_INVALID_DIMENSIONS;
    if (p->stride_y < w || p->stride_u < (w + 1) / 2 || p->stride_v < (w + 1) / 2) return RGH_ERR_INVALID_ARGUMENT;
    return RGH_OK;
}

}  // namespace

// Records the last failure for rgh_last_error from the library's other translation units.
void rgh_note_error(const char *what, const char *detail, const char *file, int line) {
    std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s (%s:%d)", what, detail, file, line);
}

extern "C" {

const char *rgh_last_error(void) { return g_last_error; }

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
    if (cfg->ring < 0 || cfg->ring == 1 || cfg->ring > 64) return RGH_ERR_INVALID_ARGUMENT;
    // The Compact preset: with CABAC it codes B pictures (h264.rs:3075), not analysed here.
    if ((cfg->flags & RGH_FLAG_COMPACT) && cfg->profile != RGH_PROFILE_BASELINE) return RGH_ERR_UNSUPPORTED;
    if ((cfg->flags & RGH_FLAG_COMPACT) && (cfg->flags & RGH_FLAG_FAST)) return RGH_ERR_INVALID_ARGUMENT;
    if ((cfg->flags & RGH_FLAG_PARTITIONS) && (cfg->flags & (RGH_FLAG_COMPACT | RGH_FLAG_FAST)))
        return RGH_ERR_INVALID_ARGUMENT;
    if (cfg->flags & ~(RGH_FLAG_NO_UNDEBLOCKED_COPY | RGH_FLAG_COMPACT | RGH_FLAG_FAST | RGH_FLAG_PARTITIONS))
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
    c->compact = (cfg->flags & RGH_FLAG_COMPACT) != 0;
    c->fast = (cfg->flags & RGH_FLAG_FAST) != 0;
    c->parts = (cfg->flags & RGH_FLAG_PARTITIONS) != 0;
    c->cqo = c->compact ? 1 : 0;
    const int W = c->mb_w * 16, H = c->mb_h * 16;
    c->sy = (W + 2 * RGH_PAD_Y + 31) & ~31;
    c->sc = (W / 2 + 2 * RGH_PAD_C + 31) & ~31;
    c->rows_y = H + 2 * RGH_PAD_Y;
    c->rows_c = H / 2 + 2 * RGH_PAD_C;
    zero_bounds(c->qp, c->compact, &c->sb[0], &c->sb[1], &c->sb[4]);
    {
        // enc.rs:2131 SkipBounds::new: the chroma bounds also keep the 2x2 DC level zero.
        int qpc = kChromaQp[std::min(c->qp + c->cqo, 51)];
        uint32_t cs, css;
        zero_bounds(qpc, c->compact, &cs, &css, nullptr);
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
    if (hipEventCreateWithFlags(&c->go, hipEventDisableTiming) != hipSuccess) return fail(RGH_ERR_DEVICE);
    const size_t ny = (size_t)c->sy * c->rows_y + 64, nc = (size_t)c->sc * c->rows_c + 64;
    const size_t nmb = (size_t)c->mb_w * c->mb_h, nit = (size_t)(c->mb_w + 1) * (c->mb_h + 1);
    c->pic.resize((size_t)(cfg->ring ? cfg->ring : 2));
    for (auto &p : c->pic) {
        size_t sizes[3] = {ny, nc, nc};
        for (int k = 0; k < 3; k++) {
            hipError_t e = hipMalloc(&p.base[k], sizes[k]);
            if (e != hipSuccess) return fail(hip_status(e));
            if (hipMemset(p.base[k], 0, sizes[k]) != hipSuccess) return fail(RGH_ERR_DEVICE);
            int pad = k ? RGH_PAD_C : RGH_PAD_Y, s = k ? c->sc : c->sy;
            p.org[k] = p.base[k] + (size_t)pad * s + pad;
        }
        hipError_t e = hipMalloc(&p.ana, nmb * sizeof(uint64_t));
        if (e == hipSuccess) e = hipMalloc(&p.fin, nit * sizeof(uint64_t));
        if (e != hipSuccess) return fail(hip_status(e));
        // Serial 0 is never a frame's: zeroed stamps mean "not done".
        if (hipMemset(p.ana, 0, nmb * sizeof(uint64_t)) != hipSuccess ||
            hipMemset(p.fin, 0, nit * sizeof(uint64_t)) != hipSuccess)
            return fail(RGH_ERR_DEVICE);
    }
    for (auto &d : c->dec) {
        hipError_t e = hipMalloc(&d, nmb * sizeof(rgh_mb_decision));
        if (e != hipSuccess) return fail(hip_status(e));
        if (hipMemset(d, 0, nmb * sizeof(rgh_mb_decision)) != hipSuccess) return fail(RGH_ERR_DEVICE);
    }
    {
        hipError_t e = hipMalloc(&c->levels, nmb * RGH_LEVELS_PER_MB * sizeof(int16_t));
        if (e != hipSuccess) return fail(hip_status(e));
    }
    const size_t srcs[3] = {(size_t)c->width * c->height, (size_t)((c->width + 1) / 2) * ((c->height + 1) / 2),
                            (size_t)((c->width + 1) / 2) * ((c->height + 1) / 2)};
    for (int k = 0; k < 3; k++) {
        hipError_t e = hipMalloc(&c->src[k], srcs[k]);
        if (e != hipSuccess) return fail(hip_status(e));
        if (c->keep_pre) {
            e = hipMalloc(&c->pre[k], k ? (size_t)W * H / 4 : (size_t)W * H);
            if (e != hipSuccess) return fail(hip_status(e));
        }
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
    Picture &p = c->pic[(size_t)c->ref];
    HIP_TRY(hipMemcpy2DAsync(p.org[0], (size_t)c->sy, recon->y, (size_t)recon->stride_y, (size_t)W, (size_t)H,
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpy2DAsync(p.org[1], (size_t)c->sc, recon->u, (size_t)recon->stride_u, (size_t)W / 2, (size_t)H / 2,
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpy2DAsync(p.org[2], (size_t)c->sc, recon->v, (size_t)recon->stride_v, (size_t)W / 2, (size_t)H / 2,
                             hipMemcpyHostToDevice, c->stream));
    RghJob j;
    std::memset(&j, 0, sizeof(j));
    j.cur_y = p.org[0];
    j.cur_u = p.org[1];
    j.cur_v = p.org[2];
    j.sy = c->sy;
    j.sc = c->sc;
    j.mb_w = c->mb_w;
    j.mb_h = c->mb_h;
    HIP_TRY(hipMemcpyAsync(c->jobs_dev, &j, sizeof(j), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(rgh_k_extend, dim3((unsigned)c->rows_y, 3, 1), dim3(256), 0, c->stream, c->jobs_dev);
    // A new serial; every item of the picture is final for any later launch.
    c->serial += 1;
    p.serial = c->serial;
    const int nit = (c->mb_w + 1) * (c->mb_h + 1);
    hipLaunchKernelGGL(rgh_k_fill_u64, dim3((unsigned)((nit + 255) / 256)), dim3(256), 0, c->stream, p.fin, nit,
                       (uint64_t)c->serial << 32);
    HIP_TRY(hipGetLastError());
    // An IDR picture is all intra: every co-located vector of the next frame is zero.
    rgh_mb_decision *z = c->dec[1 - c->dcur];
    HIP_TRY(hipMemsetAsync(z, 0, sizeof(rgh_mb_decision) * (size_t)c->mb_w * c->mb_h, c->stream));
    c->prev = z;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->has_ref = true;
    return RGH_OK;
}

// One frame of each context, outputs into the contexts' own buffers.
// The contexts of one call share a preset: each preset's P pictures run their own kernels.
rgh_status check_one_preset(Ctx *const *cs, int n, const char *fn) {
    for (int i = 1; i < n; i++)
        if (cs[i]->compact != cs[0]->compact || cs[i]->fast != cs[0]->fast || cs[i]->parts != cs[0]->parts) {
            set_error(fn, "contexts of one call must use one preset", __LINE__);
            return RGH_ERR_INVALID_ARGUMENT;
        }
    return RGH_OK;
}

rgh_status analyse_single(Ctx *const *cs, int n, const rgh_i420 *src, bool intra = false) {
    Ctx *c0 = cs[0];
    HIP_TRY(hipSetDevice(c0->device));
    rgh_status s = ensure_jobs(c0, n);
    if (s != RGH_OK) return s;
    std::vector<RghJob> jobs((size_t)n);
    for (int i = 0; i < n; i++) {
        Ctx *c = cs[i];
        jobs[(size_t)i] = make_job(c, 0, src[i], c->dec[c->dcur], c->prev, c->levels, 0, c0->ctl_dev + i, nullptr, intra);
        if (c->keep_pre) {
            jobs[(size_t)i].pre_y = c->pre[0];
            jobs[(size_t)i].pre_u = c->pre[1];
            jobs[(size_t)i].pre_v = c->pre[2];
        }
    }
    s = run_schedule(cs, n, 1, jobs, kDefaultLag);
    if (s != RGH_OK) {
        for (int i = 0; i < n; i++) poison(cs[i], 1);
        return s;
    }
    for (int i = 0; i < n; i++) {
        Ctx *c = cs[i];
        advance(c, 1, c->dec[c->dcur]);
        c->last_levels = c->levels;
        c->dcur = 1 - c->dcur;
        c->has_ref = true;
        c->pre_valid = c->keep_pre;
    }
    return RGH_OK;
}

rgh_status rgh_analyse_batch(void *const *ctxs, int32_t n, const rgh_i420 *src) {
    if (!ctxs || n <= 0 || !src) return RGH_ERR_INVALID_ARGUMENT;
    std::vector<Ctx *> cs((size_t)n);
    for (int i = 0; i < n; i++) {
        Ctx *c = static_cast<Ctx *>(ctxs[i]);
        if (!c) return RGH_ERR_INVALID_ARGUMENT;
        if (c->device != static_cast<Ctx *>(ctxs[0])->device) return RGH_ERR_INVALID_ARGUMENT;
        if (!c->has_ref) return RGH_ERR_NO_REFERENCE;
        for (int k = 0; k < i; k++)
            if (cs[(size_t)k] == c) return RGH_ERR_INVALID_ARGUMENT;
        rgh_status s = check_i420(&src[i], c->width, c->height);
        if (s != RGH_OK) return s;
        cs[(size_t)i] = c;
    }
    rgh_status s = check_one_preset(cs.data(), n, "rgh_analyse_batch");
    if (s != RGH_OK) return s;
    return analyse_single(cs.data(), n, src);
}

// noalloc: the schedules' job descriptors and progress records come from the scratch array (no
// device or pinned allocation, no event creation); otherwise from the group owners' buffers.
static rgh_status analyse_frames_impl(void *const *ctxs, int32_t n, int32_t n_frames, const rgh_i420 *src,
                               const rgh_proposal *const *props, const rgh_learned_config *lcfg,
                               rgh_mb_decision *const *decisions, int16_t *const *levels, int32_t lag, int32_t groups,
                               int32_t flags, bool noalloc = false) {
    const int keyint = (flags >> 16) & 0x7FFF;
    if (!ctxs || n <= 0 || n_frames <= 0 || !src || !decisions || !levels || lag < 0 || groups < 0 ||
        (flags & ~(RGH_FRAMES_FIRST_IDR | RGH_FRAMES_PERSISTENT | RGH_FRAMES_GRAPH | RGH_FRAMES_KEYINT(0x7FFF))) ||
        (keyint && !(flags & RGH_FRAMES_FIRST_IDR)))
        return RGH_ERR_INVALID_ARGUMENT;
    if (noalloc && (flags & RGH_FRAMES_GRAPH)) {
        set_error("rgh_analyse_frames_noalloc", "RGH_FRAMES_GRAPH allocates the graph", __LINE__);
        return RGH_ERR_INVALID_ARGUMENT;
    }
    const bool first_idr = (flags & RGH_FRAMES_FIRST_IDR) != 0;
    // Frame k of every context is an IDR picture: the first (RGH_FRAMES_FIRST_IDR), and every
    // keyint-th with RGH_FRAMES_KEYINT.
    auto is_idr = [&](int k) { return first_idr && (k == 0 || (keyint && k % keyint == 0)); };
    rgh_learned_config lc{4, 0, 0, 105, 256};
    if (props) {
        if (lcfg) {
            if (lcfg->iters < 0 || lcfg->iters > 64 || lcfg->skip_gate < 0 || lcfg->skip_gate > 256 ||
                lcfg->intra_gate < 0 || lcfg->intra_gate > 256 || lcfg->fallback_pct < 0 ||
                lcfg->fallback_pct > 100000 || lcfg->intra_first < 0 || lcfg->intra_first > 256)
                return RGH_ERR_INVALID_ARGUMENT;
            if (lcfg->iters) lc.iters = lcfg->iters;
            lc.skip_gate = lcfg->skip_gate;
            lc.intra_gate = lcfg->intra_gate;
            if (lcfg->fallback_pct) lc.fallback_pct = lcfg->fallback_pct;
            if (lcfg->intra_first) lc.intra_first = lcfg->intra_first;
        }
        for (int i = 0; i < n; i++) {
            const Ctx *c = static_cast<Ctx *>(ctxs[i]);
            if (c && (c->compact || c->fast || c->parts)) {
                set_error("rgh_analyse_frames_learned", "learned proposals serve the Balanced preset", __LINE__);
                return RGH_ERR_UNSUPPORTED;
            }
            for (int k = 0; k < n_frames; k++)
                if (!is_idr(k) && !props[(size_t)i * n_frames + k]) return RGH_ERR_INVALID_ARGUMENT;
        }
    }
    if ((int64_t)n * n_frames > (1 << 20)) return RGH_ERR_INVALID_ARGUMENT;
    std::vector<Ctx *> cs((size_t)n);
    for (int i = 0; i < n; i++) {
        Ctx *c = static_cast<Ctx *>(ctxs[i]);
        if (!c) return RGH_ERR_INVALID_ARGUMENT;
        if (c->device != static_cast<Ctx *>(ctxs[0])->device) return RGH_ERR_INVALID_ARGUMENT;
        if (!first_idr && !c->has_ref) return RGH_ERR_NO_REFERENCE;
        for (int k = 0; k < i; k++)
            if (cs[(size_t)k] == c) return RGH_ERR_INVALID_ARGUMENT;
        for (int k = 0; k < n_frames; k++) {
            const size_t q = (size_t)i * n_frames + k;
            rgh_status s = check_i420(&src[q], c->width, c->height);
            if (s != RGH_OK) return s;
            if (!decisions[q] || !levels[q]) return RGH_ERR_INVALID_ARGUMENT;
            // The kernels store records and clear levels as 64-bit words.
            if (((uintptr_t)decisions[q] & 7) || ((uintptr_t)levels[q] & 7)) return RGH_ERR_INVALID_ARGUMENT;
        }
        cs[(size_t)i] = c;
    }
    {
        rgh_status s = check_one_preset(cs.data(), n, "rgh_analyse_frames");
        if (s != RGH_OK) return s;
    }
    if (noalloc && (int64_t)n * n_frames > RGH_NOALLOC_MAX_FRAMES) {
        set_error("rgh_analyse_frames_noalloc", "more frames than RGH_NOALLOC_MAX_FRAMES in one call", __LINE__);
        return RGH_ERR_INVALID_ARGUMENT;
    }
    if (lag == 0) lag = kDefaultLag;
    // The persistent schedule (RGH_FRAMES_PERSISTENT) serves Balanced and IDR pictures; Compact
    // contexts and the no-allocation entry point use the per-step launch schedule.
    bool compact_any = false;  // the Compact preset and partitions have no persistent kernel
    for (int i = 0; i < n; i++) compact_any |= cs[(size_t)i]->compact || cs[(size_t)i]->parts;
    const bool persistent = !noalloc && !compact_any && !keyint && (flags & RGH_FRAMES_PERSISTENT);
    const int G = persistent ? 1 : std::max(1, std::min(groups ? groups : 1, n));
    HIP_TRY(hipSetDevice(cs[0]->device));
    uint8_t *scr = nullptr;
    if (noalloc) {
        rgh_status st = scratch_address(cs[0]->device, &scr);
        if (st != RGH_OK) return st;
    }
    size_t scr_used = 0;  // frames of the call placed in the scratch array so far
    // Context i joins group i % G; each group runs on its first context's stream and buffers.
    std::vector<Schedule> sch((size_t)G);
    for (int g = 0; g < G; g++) {
        Schedule &S = sch[(size_t)g];
        std::vector<int> mine;
        for (int i = g; i < n; i += G) mine.push_back(i);
        const int m = (int)mine.size();
        S.owner = cs[(size_t)mine[0]];
        S.n = m;
        S.n_frames = n_frames;
        S.lag = lag;
        if (noalloc) {
            S.jobs_dev = reinterpret_cast<RghJob *>(scr) + scr_used;
            S.ctl_dev = reinterpret_cast<RghFrameCtl *>(scr + kScratchJobBytes) + scr_used;
            S.ctl_host = g_scratch_ctl_host[cs[0]->device] + scr_used;
            scr_used += (size_t)m * n_frames;
        } else {
            rgh_status s = ensure_jobs(S.owner, m * n_frames);
            if (s != RGH_OK) return s;
            S.jobs_dev = S.owner->jobs_dev;
            S.ctl_dev = S.owner->ctl_dev;
            S.ctl_host = S.owner->ctl_host;
        }
        RghFrameCtl *ctl = S.ctl_dev;
        S.jobs.resize((size_t)m * n_frames);
        for (int k = 0; k < n_frames; k++)
            for (int t = 0; t < m; t++) {
                const int i = mine[(size_t)t];
                Ctx *c = cs[(size_t)i];
                const size_t q = (size_t)i * n_frames + k;
                const int R = (int)c->pic.size();
                // The picture slot of frame k last served as the reference of frame k + 1 - R.
                const int wk = k + 1 - R;
                const RghFrameCtl *wait = wk >= 0 ? ctl + (size_t)wk * m + t : nullptr;
                const rgh_mb_decision *prev = k ? decisions[q - 1] : c->prev;
                RghJob &j = S.jobs[(size_t)k * m + t];
                j = make_job(c, k, src[q], decisions[q], prev, levels[q], k * lag, ctl + (size_t)k * m + t, wait,
                             is_idr(k));
                if (props && !j.intra) {
                    j.prop = props[q];
                    j.learn_iters = std::min(lc.iters, c->me_iters);
                    j.learn_skip_gate = lc.skip_gate;
                    j.learn_intra_gate = lc.intra_gate;
                    j.learn_fallback = lc.fallback_pct;
                    j.learn_intra_first = lc.intra_first;
                }
            }
    }
    rgh_status s;
    if (persistent) {
        std::vector<int> ring((size_t)n);
        for (int i = 0; i < n; i++) ring[(size_t)i] = (int)cs[(size_t)i]->pic.size();
        Ctx *timer = cs[0];
        if (timer->timing) HIP_TRY(hipEventRecord(timer->ev[0], timer->stream));
        s = run_persistent(sch[0], lag, props != nullptr, ring, timer);
        if (s == RGH_OK && timer->timing) {
            HIP_TRY(hipEventRecord(timer->ev[1], timer->stream));
            timer->timed = true;
        }
    } else {
        // The other groups' streams start after everything queued on the first one.
        if (G > 1) {
            HIP_TRY(hipEventRecord(cs[0]->go, cs[0]->stream));
            for (int g = 1; g < G; g++) HIP_TRY(hipStreamWaitEvent(sch[(size_t)g].owner->stream, cs[0]->go, 0));
        }
        s = run_schedules(sch, cs[0], true, (flags & RGH_FRAMES_GRAPH) != 0);
    }
    if (s != RGH_OK) {
        for (int i = 0; i < n; i++) poison(cs[(size_t)i], n_frames);
        return s;
    }
    // The last frame's decisions carry the co-located vectors into the next call.
    for (int i = 0; i < n; i++) {
        Ctx *c = cs[(size_t)i];
        rgh_mb_decision *keep = c->dec[c->dcur];
        HIP_TRY(hipMemcpyAsync(keep, decisions[(size_t)i * n_frames + n_frames - 1],
                               sizeof(rgh_mb_decision) * (size_t)c->mb_w * c->mb_h, hipMemcpyDeviceToDevice,
                               cs[0]->stream));
        advance(c, n_frames, keep);
        c->last_levels = levels[(size_t)i * n_frames + n_frames - 1];
        c->dcur = 1 - c->dcur;
        c->has_ref = true;
        c->pre_valid = false;  // this form keeps no undeblocked copy
    }
    HIP_TRY(hipStreamSynchronize(cs[0]->stream));
    return RGH_OK;
}

rgh_status rgh_analyse_frames(void *const *ctxs, int32_t n, int32_t n_frames, const rgh_i420 *src,
                              rgh_mb_decision *const *decisions, int16_t *const *levels, int32_t lag, int32_t groups,
                              int32_t flags) {
    return analyse_frames_impl(ctxs, n, n_frames, src, nullptr, nullptr, decisions, levels, lag, groups, flags);
}

rgh_status rgh_analyse_frames_noalloc(void *const *ctxs, int32_t n, int32_t n_frames, const rgh_i420 *src,
                                      rgh_mb_decision *const *decisions, int16_t *const *levels, int32_t lag,
                                      int32_t groups, int32_t flags) {
    return analyse_frames_impl(ctxs, n, n_frames, src, nullptr, nullptr, decisions, levels, lag, groups, flags, true);
}

rgh_status rgh_analyse_frames_learned(void *const *ctxs, int32_t n, int32_t n_frames, const rgh_i420 *src,
                                      const rgh_proposal *const *proposals, const rgh_learned_config *config,
                                      rgh_mb_decision *const *decisions, int16_t *const *levels, int32_t lag,
                                      int32_t groups, int32_t flags) {
    if (!proposals) return RGH_ERR_INVALID_ARGUMENT;
    return analyse_frames_impl(ctxs, n, n_frames, src, proposals, config, decisions, levels, lag, groups, flags);
}

rgh_status analyse_host_frame(void *ctx, const rgh_i420 *src, rgh_mb_decision *decisions, int16_t *levels,
                              const rgh_i420_out *recon, bool intra) {
    if (!ctx) return RGH_ERR_INVALID_ARGUMENT;
    Ctx *c = static_cast<Ctx *>(ctx);
    rgh_status s = check_i420(src, c->width, c->height);
    if (s != RGH_OK) return s;
    if (!intra && !c->has_ref) return RGH_ERR_NO_REFERENCE;
    HIP_TRY(hipSetDevice(c->device));
    const int cw = (c->width + 1) / 2, ch = (c->height + 1) / 2;
    HIP_TRY(hipMemcpy2DAsync(c->src[0], (size_t)c->width, src->y, (size_t)src->stride_y, (size_t)c->width,
                             (size_t)c->height, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpy2DAsync(c->src[1], (size_t)cw, src->u, (size_t)src->stride_u, (size_t)cw, (size_t)ch,
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpy2DAsync(c->src[2], (size_t)cw, src->v, (size_t)src->stride_v, (size_t)cw, (size_t)ch,
                             hipMemcpyHostToDevice, c->stream));
    rgh_i420 dsrc = {c->src[0], c->src[1], c->src[2], c->width, cw, cw, c->width, c->height};
    Ctx *cs[1] = {c};
    s = analyse_single(cs, 1, &dsrc, intra);
    if (s != RGH_OK) return s;
    return rgh_get_outputs(ctx, decisions, levels, recon, 1);
}

rgh_status rgh_analyse_p_frame(void *ctx, const rgh_i420 *src, rgh_mb_decision *decisions, int16_t *levels,
                               const rgh_i420_out *recon) {
    return analyse_host_frame(ctx, src, decisions, levels, recon, false);
}

rgh_status rgh_analyse_idr(void *ctx, const rgh_i420 *src, rgh_mb_decision *decisions, int16_t *levels,
                           const rgh_i420_out *recon) {
    return analyse_host_frame(ctx, src, decisions, levels, recon, true);
}

rgh_status rgh_get_outputs(void *ctx, rgh_mb_decision *decisions, int16_t *levels, const rgh_i420_out *recon,
                           int32_t which) {
    if (!ctx) return RGH_ERR_INVALID_ARGUMENT;
    Ctx *c = static_cast<Ctx *>(ctx);
    if (!c->analysed) return RGH_ERR_NO_REFERENCE;
    HIP_TRY(hipSetDevice(c->device));
    const size_t n = (size_t)c->mb_w * c->mb_h;
    // The last analysed frame's decisions are the reference's; its picture is the reference.
    if (decisions)
        HIP_TRY(hipMemcpyAsync(decisions, c->prev, n * sizeof(rgh_mb_decision), hipMemcpyDeviceToHost, c->stream));
    if (levels)
        HIP_TRY(hipMemcpyAsync(levels, c->last_levels, n * RGH_LEVELS_PER_MB * sizeof(int16_t), hipMemcpyDeviceToHost,
                               c->stream));
    if (recon) {
        const int W = c->mb_w * 16, H = c->mb_h * 16;
        const uint8_t *sp[3];
        size_t ss[3];
        if (which == 0) {
            if (!c->pre_valid) {
                set_error("rgh_get_outputs", "no undeblocked copy of the last frame (disabled, or the last call was "
                          "rgh_analyse_frames)", __LINE__);
                return RGH_ERR_INVALID_ARGUMENT;
            }
            for (int k = 0; k < 3; k++) sp[k] = c->pre[k];
            ss[0] = (size_t)W;
            ss[1] = ss[2] = (size_t)W / 2;
        } else {
            for (int k = 0; k < 3; k++) sp[k] = c->pic[(size_t)c->ref].org[k];
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
    HIP_TRY(hipEventSynchronize(c->ev[1]));
    HIP_TRY(hipEventElapsedTime(analyse_ms, c->ev[0], c->ev[1]));
    *deblock_ms = 0.0f;
    *launches = c->timed_launches;
    return RGH_OK;
}

}  // extern "C"
