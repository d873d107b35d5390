// Device-side definitions shared by the analysis and deblocking kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ragdoll_hip.h"
#include "../../include/ragdoll_learned.h"

#define RGH_PAD_Y 48      // luma padding of reconstructed pictures (enc.rs:56)
#define RGH_PAD_C 16      // chroma padding (enc.rs:58)
#define RGH_MV_MARGIN 28  // how far a vector may move a block past the edge (enc.rs:61)

// Progress of one frame of a schedule, in device memory. delay[launch & 1] is the number of
// launches the frame has fallen behind its planned start (its macroblocks wait for reference
// samples that are not final yet); blocks of launch L read delay[L & 1] and raise
// delay[(L + 1) & 1]. analysed counts finished analysis items.
struct RghFrameCtl {
    int32_t delay[2];
    int32_t analysed;
    int32_t pad_;
};

// A completion stamp: frame serial number in the high half, schedule launch number in the low
// half. A stamp counts as final for a reader in launch `seq` only when it was written by an
// earlier launch (kernel boundaries make the data it covers visible).
__device__ __forceinline__ bool stamp_final(uint64_t st, uint32_t serial, uint32_t seq) {
    return (uint32_t)(st >> 32) == serial && (uint32_t)st < seq;
}
__device__ __forceinline__ uint64_t make_stamp(uint32_t serial, uint32_t seq) { return ((uint64_t)serial << 32) | seq; }

#include "rgh_band.h"

// The band phase of a frame: spreads the frames' bands over the XCDs; pictures coded from the
// same source (streams of one content at several rates) share it, and their source reads with it.
__device__ __forceinline__ int band_phase(uint64_t src_y) {
    return (int)(((uint32_t)(src_y >> 12) * 0x9E3779B1u) >> 29);
}

// One P frame of one stream in a schedule. Plane pointers address sample (0, 0) inside the
// padding, so negative offsets reach the padding.
struct RghJob {
    const uint8_t *src_y, *src_u, *src_v;
    int32_t src_sy, src_su, src_sv, width, height;
    const uint8_t *ref_y, *ref_u, *ref_v;
    uint8_t *cur_y, *cur_u, *cur_v;
    uint8_t *pre_y, *pre_u, *pre_v;  // optional copy of the reconstruction before deblocking (W x H)
    int32_t sy, sc;            // padded strides
    int32_t mb_w, mb_h, qp, me_iters, qpel_iters, t8x8;
    rgh_mb_decision *dec;      // this frame's decisions (written)
    const rgh_mb_decision *prev;  // the reference frame's decisions (co-located vectors)
    int16_t *levels;
    uint32_t sb_luma, sb_luma_ssd, sb_chroma, sb_chroma_ssd, sb_probe;
    int32_t steps;             // wavefront steps of the frame: (mb_w - 1) + 2 (mb_h - 1) + 1
    int32_t deblock;           // 0 when the slice disables the loop filter
    int32_t intra;             // an IDR picture (no reference)
    int32_t compact;           // the Compact preset (enc/inter.rs, enc/compact.rs)
    int32_t cqo;               // chroma_qp_index_offset
    // Learned proposals (rgh_k_learned_step): one record per macroblock, and the search effort.
    const rgh_proposal *prop;
    int32_t learn_iters, learn_skip_gate, learn_intra_gate, learn_fallback, learn_intra_first;
    // Schedule.
    int32_t base;              // launch of the frame's first wavefront step
    uint32_t serial, ref_serial;  // this frame's and its reference's serial numbers
    uint64_t *ana;             // analysis stamps, mb_w x mb_h
    uint64_t *fin;             // deblock/pad item stamps of this picture, (mb_w + 1) x (mb_h + 1)
    const uint64_t *ref_fin;   // the same for the reference picture
    RghFrameCtl *ctl;
    const int32_t *wait_analysed;  // analysis count of the frame that last reads the picture slot
    int32_t wait_total;        // ... which must reach this value before the first write
    // Persistent schedule: the jobs of this launch that hold the reference frame and the frame
    // gating the picture slot (-1: not in this launch, so complete before it started).
    int32_t ref_job, wait_job;
    int32_t fast;  // the Fast preset's decisions (RGH_FLAG_FAST)
    int32_t parts; // P partitions on the Balanced search (RGH_FLAG_PARTITIONS)
    // Sequence number of the call's launch 0: a per-step launch's number is seq0 plus the
    // relative number it is given, so that a replayed launch sequence takes fresh numbers.
    uint32_t seq0;
};

// The device-side view of a job: the same fields with its buffers typed as global memory, so
// that every access compiles to global (not flat) loads and stores — flat accesses also wait
// on LDS and scalar traffic.
#define RGH_G __attribute__((address_space(1)))
struct RghJobD {
    const RGH_G uint8_t *src_y, *src_u, *src_v;
    int32_t src_sy, src_su, src_sv, width, height;
    const RGH_G uint8_t *ref_y, *ref_u, *ref_v;
    RGH_G uint8_t *cur_y, *cur_u, *cur_v;
    RGH_G uint8_t *pre_y, *pre_u, *pre_v;
    int32_t sy, sc;
    int32_t mb_w, mb_h, qp, me_iters, qpel_iters, t8x8;
    RGH_G rgh_mb_decision *dec;
    const RGH_G rgh_mb_decision *prev;
    RGH_G int16_t *levels;
    uint32_t sb_luma, sb_luma_ssd, sb_chroma, sb_chroma_ssd, sb_probe;
    int32_t steps, deblock, intra, compact, cqo, base;
    const RGH_G rgh_proposal *prop;
    int32_t learn_iters, learn_skip_gate, learn_intra_gate, learn_fallback, learn_intra_first;
    uint32_t serial, ref_serial;
    RGH_G uint64_t *ana;
    RGH_G uint64_t *fin;
    const RGH_G uint64_t *ref_fin;
    RGH_G RghFrameCtl *ctl;
    const RGH_G int32_t *wait_analysed;
    int32_t wait_total;
    int32_t ref_job, wait_job;
    int32_t fast;
    int32_t parts;
    uint32_t seq0;
};

template <typename T>
__device__ __forceinline__ RGH_G T *gp(T *p) {
    return (RGH_G T *)p;
}

__device__ __forceinline__ RghJobD device_view(const RghJob &j) {
    RghJobD d;
    d.src_y = gp(j.src_y);
    d.src_u = gp(j.src_u);
    d.src_v = gp(j.src_v);
    d.src_sy = j.src_sy;
    d.src_su = j.src_su;
    d.src_sv = j.src_sv;
    d.width = j.width;
    d.height = j.height;
    d.ref_y = gp(j.ref_y);
    d.ref_u = gp(j.ref_u);
    d.ref_v = gp(j.ref_v);
    d.cur_y = gp(j.cur_y);
    d.cur_u = gp(j.cur_u);
    d.cur_v = gp(j.cur_v);
    d.pre_y = gp(j.pre_y);
    d.pre_u = gp(j.pre_u);
    d.pre_v = gp(j.pre_v);
    d.sy = j.sy;
    d.sc = j.sc;
    d.mb_w = j.mb_w;
    d.mb_h = j.mb_h;
    d.qp = j.qp;
    d.me_iters = j.me_iters;
    d.qpel_iters = j.qpel_iters;
    d.t8x8 = j.t8x8;
    d.dec = gp(j.dec);
    d.prev = gp(j.prev);
    d.levels = gp(j.levels);
    d.sb_luma = j.sb_luma;
    d.sb_luma_ssd = j.sb_luma_ssd;
    d.sb_chroma = j.sb_chroma;
    d.sb_chroma_ssd = j.sb_chroma_ssd;
    d.sb_probe = j.sb_probe;
    d.steps = j.steps;
    d.deblock = j.deblock;
    d.intra = j.intra;
    d.compact = j.compact;
    d.cqo = j.cqo;
    d.prop = gp(j.prop);
    d.learn_iters = j.learn_iters;
    d.learn_skip_gate = j.learn_skip_gate;
    d.learn_intra_gate = j.learn_intra_gate;
    d.learn_fallback = j.learn_fallback;
    d.learn_intra_first = j.learn_intra_first;
    d.base = j.base;
    d.serial = j.serial;
    d.ref_serial = j.ref_serial;
    d.ana = gp(j.ana);
    d.fin = gp(j.fin);
    d.ref_fin = gp(j.ref_fin);
    d.ctl = gp(j.ctl);
    d.wait_analysed = gp(j.wait_analysed);
    d.wait_total = j.wait_total;
    d.ref_job = j.ref_job;
    d.wait_job = j.wait_job;
    d.fast = j.fast;
    d.parts = j.parts;
    d.seq0 = j.seq0;
    return d;
}

// ---- Persistent schedule (rgh_k_persist) ----
// One launch runs every item of a call (analysis macroblocks and deblock/pad items of every
// frame). Each stream is homed on one XCD; the items of the streams homed on an XCD form that
// XCD's queue of tickets, ordered by planned time (a frame's base time plus the wavefront step
// x + 2y, deblock/pad items 4 steps later), and only waves running on that XCD take its tickets,
// so every item that reads another's results shares its L2 with it. An item waits only for items
// of earlier time (held by running waves); one that would need a later item defers, and its frame
// is finished by a later launch.
constexpr int RGH_MAX_QUEUES = 16;
struct RghQueue {
    uint32_t total;    // tickets (items) in the queue
    int32_t t0;        // time of the queue's first slot
    uint32_t slot0;    // index of its first entry in the slot tables (n_slots + 1 entries)
    uint32_t n_slots;  // time slots
    uint32_t claim0;   // index of its first ticket's claim word
    uint32_t pad_[3];
};
struct RghPersist {
    int32_t n_queues;
    int32_t n_xcc;                         // XCDs the grid is dealt over (round-robin by workgroup)
    int32_t queue_of_xcc[RGH_MAX_QUEUES];  // queue of the waves on each XCC id (-1: none)
    uint32_t spin_limit;  // sleeps a wait may take before the schedule gives up
    uint32_t jitter;      // test hook: a pseudo-random delay before each item (0: none)
    uint32_t debug;       // count items, spins and cycles into status[2..7]
    RghQueue q[RGH_MAX_QUEUES];
    const uint32_t *slot_tl;   // per slot: its first ticket (the queue's total after the last)
    const uint32_t *slot_seg;  // per slot: its first segment (the end after the last)
    const uint32_t *seg_job;   // per segment: the job whose items it holds
    const uint32_t *seg_tl;    // per segment: its first ticket
    uint32_t *claim;           // per ticket: set by the workgroup that runs it (zero at the launch)
    int32_t *deferred;         // per job: 0x7fffffff - the earliest time an item deferred (0: none)
    uint32_t *status;          // [0] gave up (1: a wait's bound, 2: a ticket claimed twice),
                               // [1] items deferred, [2] items run, [3..7] debug counters
};

// Sum of a 32-bit value over the 64 lanes of the wave, broadcast as a scalar. DPP quad and row
// mirrors reduce each row of 16 lanes; the four row sums are read out with readlane. Every lane
// of the wave must be active.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);  // row_mirror
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 0) + (uint32_t)__builtin_amdgcn_readlane((int)v, 16) +
           (uint32_t)__builtin_amdgcn_readlane((int)v, 32) + (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
}

// N independent wave sums, interleaved so their DPP chains overlap.
template <int N>
__device__ __forceinline__ void wave_sums(const uint32_t *in, uint32_t *out) {
    uint32_t v[N];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = in[k];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[k], 0xB1, 0xF, 0xF, false);
#pragma unroll
    for (int k = 0; k < N; k++) v[k] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[k], 0x4E, 0xF, 0xF, false);
#pragma unroll
    for (int k = 0; k < N; k++) v[k] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[k], 0x141, 0xF, 0xF, false);
#pragma unroll
    for (int k = 0; k < N; k++) v[k] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[k], 0x140, 0xF, 0xF, false);
#pragma unroll
    for (int k = 0; k < N; k++)
        out[k] = (uint32_t)__builtin_amdgcn_readlane((int)v[k], 0) + (uint32_t)__builtin_amdgcn_readlane((int)v[k], 16) +
                 (uint32_t)__builtin_amdgcn_readlane((int)v[k], 32) + (uint32_t)__builtin_amdgcn_readlane((int)v[k], 48);
}
__device__ __forceinline__ void wave_sum4(const uint32_t *in, uint32_t *out) { wave_sums<4>(in, out); }

// Four bytes at any address of global memory, from the two aligned dwords that hold them. The
// second dword must lie inside the allocation (true for every padded-plane read of the kernels).
__device__ __forceinline__ uint32_t ld4u(const RGH_G uint8_t *p) {
    const uint32_t s = (uint32_t)(uintptr_t)p & 3;
    const RGH_G uint32_t *q = (const RGH_G uint32_t *)(p - s);
    return __builtin_amdgcn_alignbyte(q[1], q[0], s);
}

// A decision record from / to global memory, as four 64-bit words.
__device__ __forceinline__ rgh_mb_decision load_dec(const RGH_G rgh_mb_decision *p) {
    const RGH_G uint64_t *q = (const RGH_G uint64_t *)p;
    const uint64_t w[4] = {q[0], q[1], q[2], q[3]};
    rgh_mb_decision d;
    __builtin_memcpy(&d, w, sizeof(d));
    return d;
}
__device__ __forceinline__ void store_dec(RGH_G rgh_mb_decision *p, const rgh_mb_decision &d) {
    uint64_t w[4];
    __builtin_memcpy(w, &d, sizeof(d));
    RGH_G uint64_t *q = (RGH_G uint64_t *)p;
#pragma unroll
    for (int i = 0; i < 4; i++) q[i] = w[i];
}

__device__ __forceinline__ int clip255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }


// clip255(v >> S) (arithmetic shift), computed as a clamp before a logical shift. Written as
// shift-then-clip, two such results packed into one dword are selected by the gfx950 backend as
// v_ashr_pk_u8_i32, after which it assumes bits 31:16 of that register are zero; on MI355X they
// keep the register's previous contents and corrupt the bytes packed above. This form never
// matches that instruction (the build rejects any code object that contains it).
template <int S>
__device__ __forceinline__ uint32_t sat_shr(int v) {
    constexpr int HI = (256 << S) - 1;
    return (uint32_t)(v < 0 ? 0 : (v > HI ? HI : v)) >> S;
}
__device__ __forceinline__ int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }
__device__ __forceinline__ uint32_t byte_of(uint32_t w, int i) { return (w >> (8 * i)) & 0xFF; }

// Column and row of 4x4 block b in decoding order (enc.rs:432 blk_xy).
__device__ __forceinline__ int blk_x(int b) { return (b & 1) + ((b >> 1) & 2); }
__device__ __forceinline__ int blk_y(int b) { return ((b >> 1) & 1) + ((b >> 2) & 2); }
