// k_fused.hpp - forward CLT + pairwise correlation in one pass over the tiles.
//
// The staged path (convert_correct_tiles, then correlate2D) writes the whole
// CLT (12 KB per tile) to memory and reads it back three times per camera.
// Here a workgroup converts a group of GT tiles (4 cameras x 3 colors) into
// LDS and correlates the group's pairs from there; storing the CLT to memory
// is optional.  Same per-tile arithmetic as the staged kernels
// (convert_unit, corr_accumulate, corr_finish), so results are identical.
//
// Workgroup: 3 wavefronts (one per color) of 8 units x 8 lanes, as in
// k_convert_correct_tiles.  Per group of GT tiles:
//   1. convert: each wavefront converts its color for GT x 4 (tile, camera)
//      units (GT / 2 passes of 8 units) into the CLT region of LDS;
//   2. barrier; correlate: 6 GT (tile, pair) slots, 8 per wavefront, each
//      accumulates its bicomplex product from LDS;
//   3. barrier; the CLT region is reused as per-slot scratch for the
//      normalisation, DCT-II back-transform and unfold; rows are stored;
//   4. barrier before the next group overwrites LDS.
//
// LDS layout (bank rules of gfx950: a 16-byte read pass is 16 lanes, which
// must hit 16 distinct 16-byte bank groups, address / 16 mod 16, with lane
// j + 16 of a half-wavefront on the group of lane j; a 4-byte pass is 32
// lanes on 32 banks):
//   - CLT block of (tile, camera, color): 256 floats at
//     ((tile * 4 + camera) * 3 + color) * 256, rows as put_row8 writes them
//     for cameras 0 and 3 (halves of rows 4..7 swapped) and with the halves
//     of rows 0..3 swapped for cameras 1 and 2: a half-row read of the 8
//     rows covers 8 bank groups, the complementary 8 for the other kind.
//   - Correlation slots: per tile pairs 0..3 first (slots 4 tile + pair),
//     then pairs 4, 5 of all tiles (slots 4 GT + 2 tile + pair - 4).  The
//     two slots of each 16-lane group then read blocks of opposite kinds,
//     and lanes j and j + 16 blocks of the same kind, in every one of the
//     48 loads.
//   - Slot scratch (normalisation, back-transform, staged output tile): 296
//     floats per slot, so the 4 slots of a 32-lane pass fall 8 banks apart.
//
// Work items are runs of tiles_per_item consecutive tiles (in the visiting
// order of swizzle_strip) of one frame.  Single-frame launches map one
// workgroup to one item.  Stream launches (QUEUE) are persistent: a fixed
// grid of workgroups takes items from 8 per-XCD queues (queue x holds the
// x-th eighth of each frame's items, frame after frame; idle workgroups
// steal from the other queues) until every frame's items are done; a claim
// is mapped to its frame by binary search in a per-queue prefix table.
#pragma once

#include "k_convert.hpp"
#include "k_correlate.hpp"
#include "k_combine.hpp"

namespace tp {

constexpr int TP_WORK_LINE = 32;    // ints per 128-byte line: one claim counter per line
constexpr int TP_WORK_ITEMS = 256;  // item count, after the 8 counter lines
constexpr int TP_WORK_TABLE = 288;  // then [nframes + 1][8]: per-XCD item prefix over frames
static_assert(TP_WORK_TABLE + 8 == TP_STREAM_WORK_INTS(0), "stream work buffer");

// Items of frame f (n items) in XCD x's queue: the x-th eighth of the frame,
// [n x / 8, n (x + 1) / 8).
__device__ __forceinline__ int eighth_lo(int n, int x) { return (int)((long long)n * x / 8); }

// Binary-identical to tp_frame (include/tp.h).
struct FrameDesc {
    const tp_task *tasks;
    const float *const *images;
    const float *const *kernels;
    const float *const *koffs;
    float *const *clt;  // may be null: CLT not stored
    float *corrs;       // row 6 * task + pair, may be null
    float *combined;    // row task: the six-pair combination, may be null
    int num_tiles;
    int num_items;
};

struct FusedArgs {
    Geo g;
    FrameDesc f0;              // single-frame launch
    const FrameDesc *frames;   // stream launch: device array [nframes]
    int nframes;
    int *work;                 // stream launch: per-XCD claims at [TP_WORK_LINE x], items at [TP_WORK_ITEMS]
    unsigned long long *stamps;  // stream launch: per frame {first start, last end} or null
    int tiles_per_item;        // multiple of GT
    size_t dstride;
    int lpf_mask;
    float scale[3];
    float fat_zero;
    size_t corr_stride;
    int radius;
    size_t combined_stride;
    int pair_mask;            // pairs summed into the combined tile
};

// Occupancy: the FP32-DCT kernel fits 168 VGPRs, i.e. 3 waves per SIMD;
// left to itself the compiler takes a few more registers in some forms and
// drops to 2 waves per SIMD, 1.2-1.6x slower, so the FP32 forms ask for 3.
// The matrix-core forms need more than 168 registers and keep the
// compiler's choice.
template <int GT, bool QUEUE, int COLORS, int DCT>
constexpr int fused_min_waves() {
    return DCT == 0 ? TP_FUSED_QWAVES : TP_FUSED_WAVES;
}

template <int GT, bool QUEUE, int COLORS, int DCT = 0>
__global__ __launch_bounds__(64 * TP_FUSED_WG_WAVES)
__attribute__((amdgpu_waves_per_eu(fused_min_waves<GT, QUEUE, COLORS, DCT>())))
void k_fused(FusedArgs a) {
    static_assert(GT == 2 || GT == 4, "tiles per group");
    constexpr int NSLOT = 6 * GT;
    // [tile][camera][color][4 x 8 x 8]; a unit's slice also serves as its
    // transposition staging (2 x 72 floats) before its CLT is written there
    constexpr int FB = 256, CLT_TILE = 12 * FB, SLOT_STRIDE = 296;
    static_assert(NSLOT * SLOT_STRIDE <= GT * CLT_TILE, "slot scratch fits the CLT region");
    __shared__ __align__(16) float clt_lds[GT * CLT_TILE];
    auto cam_base = [](int c) { return c * 3 * FB; };
    // which row halves a camera's blocks swap: 4 (rows 4..7) or 0 (rows 0..3)
    // is XORed into (row & 4) ^ 4 h; cameras 1 and 2 swap rows 0..3
    auto cam_swap = [](int c) { return ((c ^ (c >> 1)) & 1) << 2; };
    // slot -> (tile, pair) and back
    auto slot_tile = [](int sl) { return sl < 4 * GT ? sl >> 2 : (sl - 4 * GT) >> 1; };
    auto slot_pair = [](int sl) { return sl < 4 * GT ? sl & 3 : 4 + ((sl - 4 * GT) & 1); };
    auto pair_slot = [](int tl, int pr) { return pr < 4 ? 4 * tl + pr : 4 * GT + 2 * tl + pr - 4; };
    __shared__ int s_item, s_frame;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int unit = lane >> 3;
    const int li = lane & 7;
    const int cam = unit & 3;
    const LaneFold lf = lane_fold(li);
    const CvtMfma<DCT> mt = cvt_mfma<DCT>(lane);
    int src_off[4];
    corr_out_offsets(lane, a.radius, src_off);
    const float fz2 = a.fat_zero * a.fat_zero;
    const Geo &g = a.g;
    int item = QUEUE ? 0 : xcd_remap(blockIdx.x, gridDim.x);
    // frame descriptors live in global memory (scalar loads, not flat ones)
    typedef __attribute__((address_space(1))) const FrameDesc gframe;
    gframe *frames = (gframe *)a.frames;
    typedef __attribute__((address_space(1))) const int gint;
    gint *table = (gint *)a.work + TP_WORK_TABLE;

    while (true) {
        int f = 0;
        if (QUEUE) {
            // One queue per XCD: the x-th eighth of every frame, frame after
            // frame.  Neighbouring tiles (which share kernel cells and image
            // rows) are converted under one L2, and all XCDs work on the
            // same frames at a time, so a frame finishes soon after it
            // starts.  A workgroup (on XCD blockIdx.x % 8) drains its own
            // queue first, then the others.
            if (threadIdx.x == 0) {
                int it = -1, fi = 0;
                const int x0 = blockIdx.x & 7;
                for (int k = 0; k < 8 && it < 0; k++) {
                    const int x = (x0 + k) & 7;
                    const int c = atomicAdd(&a.work[x * TP_WORK_LINE], 1);
                    if (c < table[a.nframes * 8 + x]) {
                        // last frame whose prefix for queue x is <= c
                        int lo = 0, hi = a.nframes - 1;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (table[mid * 8 + x] <= c) lo = mid; else hi = mid - 1;
                        }
                        const int n = frames[lo].num_items;
                        it = eighth_lo(n, x) + (c - table[lo * 8 + x]);
                        fi = lo;
                    }
                }
                s_item = it;
                s_frame = fi;
            }
            __syncthreads();
            // the same value in every lane: keep it (and the frame descriptor
            // and pointers derived from it) in scalar registers
            item = __builtin_amdgcn_readfirstlane(s_item);
            f = __builtin_amdgcn_readfirstlane(s_frame);
            __syncthreads();  // s_item is rewritten by the next iteration
            if (item < 0) break;  // every queue is drained
        }
        FrameDesc fr = a.f0;
        if (QUEUE) {
            gframe &d = frames[f];
            fr.tasks = d.tasks;
            fr.images = d.images;
            fr.kernels = d.kernels;
            fr.koffs = d.koffs;
            fr.clt = d.clt;
            fr.corrs = d.corrs;
            fr.combined = d.combined;
            fr.num_tiles = d.num_tiles;
            fr.num_items = d.num_items;
        }
        if (QUEUE && a.stamps && threadIdx.x == 0)
            atomicMin(&a.stamps[2 * f], (unsigned long long)wall_clock64());
        ConvertArgs ca;
        ca.g = g;
        ca.koffs = fr.koffs;
        ca.kernels = fr.kernels;
        ca.images = fr.images;
        ca.tasks = fr.tasks;
        ca.clt = fr.clt;
        ca.dstride = a.dstride;
        ca.num_tiles = fr.num_tiles;
        ca.lpf_mask = a.lpf_mask;
        resolve_cam(ca, cam);
        const int s0 = item * a.tiles_per_item;  // item: index within the frame
        const int s1 = min(fr.num_tiles, s0 + a.tiles_per_item);

        // This wavefront's jobs (color, tile pair) in every group of GT tiles:
        // see (1) below.
        static_assert(TP_FUSED_WG_WAVES == 4, "4 wavefronts per workgroup");
        constexpr bool W4 = GT == 4;
        const int njobs = W4 ? (wave < 2 ? 1 : 2) : wave == 3 ? 0 : GT / 2;
        auto job_color = [&](int job) { return W4 ? (wave < 2 ? 2 : wave - 2) : wave; };
        auto job_pass = [&](int job) { return W4 ? (wave < 2 ? wave : job) : job; };

        for (int gs = s0; gs < s1; gs += GT) {
            // (1) forward CLT of the group into LDS
            // one job converts one color of 2 tiles (8 units per wavefront);
            // kept as a loop so that two jobs' registers are not live
            // together.  G costs about twice R or B.  With 4 wavefronts and
            // GT = 4 (two tile pairs): wavefronts 0 and 1 do G of pair 0 /
            // pair 1, wavefront 2 does R and wavefront 3 B of both pairs.
            // With GT = 2, wavefront w converts color w of the pair and
            // wavefront 3 rests.
#pragma unroll 1
            for (int job = 0; job < njobs; job++) {
                const int color = job_color(job);
                const int pass = job_pass(job);
                const int tl = pass * 2 + (unit >> 2);
                // the unit's task words (its validity and tile come from
                // them too: no separate load of the task record)
                const int st = gs + tl;
                bool valid = st < s1;
                const int task_num = swizzle_strip(valid ? st : s1 - 1, g.TX, fr.num_tiles);
                UnitPre up;
                unit_centre(ca, cam, task_num, up);
                if (color == 0) unit_extra<0>(ca, cam, up);
                else if (color == 1) unit_extra<1>(ca, cam, up);
                else unit_extra<2>(ca, cam, up);
                valid = valid && up.task != 0;
                float q0[8], q1[8], q2[8], q3[8];
                float *slice = clt_lds + tl * CLT_TILE + cam_base(cam) + color * FB;
                if (color == 0)
                    convert_unit<0, DCT>(ca, lf, slice, slice + 72, li, cam, task_num, q0, q1, q2, q3, mt, 3 * FB, &up);
                else if (color == 1)
                    convert_unit<1, DCT>(ca, lf, slice, slice + 72, li, cam, task_num, q0, q1, q2, q3, mt, 3 * FB, &up);
                else
                    convert_unit<2, DCT>(ca, lf, slice, slice + 72, li, cam, task_num, q0, q1, q2, q3, mt, 3 * FB, &up);
                {
                    const int sw = (li & 4) ^ cam_swap(cam);  // offset of the row's first half
                    float *r = slice + li * 8;
                    const float(*qv[4])[8] = {&q0, &q1, &q2, &q3};
#pragma unroll
                    for (int qq = 0; qq < 4; qq++) {
                        const float(&v)[8] = *qv[qq];
                        *(float4 *)(r + qq * 64 + sw) = make_float4(v[0], v[1], v[2], v[3]);
                        *(float4 *)(r + qq * 64 + (sw ^ 4)) = make_float4(v[4], v[5], v[6], v[7]);
                    }
                }
                if (fr.clt) {
                    const int tile = task_tile(g, up.txy);
                    if (valid && tile >= 0)
                        store_clt_rows(ca.l_clt + ((size_t)tile * TP_NUM_COLORS + color) * 256, li,
                                       q0, q1, q2, q3);
                }
            }
            __syncthreads();

            // (2) correlation products from LDS
            const int slot = wave * 8 + unit;
            const bool active = wave * 8 < NSLOT;  // wavefront-uniform
            int row = -1;
            float c0[8], c1[8], c2[8], c3[8];
            if (active) {
                // slots past the last one (GT = 2) read pairs 0 and 1 of
                // tile 0, in the offset pattern of real slots
                const int tl = slot < NSLOT ? slot_tile(slot) : 0;
                const int pair = slot < NSLOT ? slot_pair(slot) : slot & 1;
                const int s = gs + tl;
                if (slot < NSLOT && s < s1) {
                    const int task_num = swizzle_strip(s, g.TX, fr.num_tiles);
                    gtask *tk = gtasks(fr.tasks) + task_num;
                    if (((tk->task >> (TP_TASK_CORR_BITS + pair)) & 1) && task_tile(g, tk->txy) >= 0)
                        row = 6 * task_num + pair;
                }
                const int ca1 = pair_cam1(pair), ca2 = pair_cam2(pair);
                const float *ta = clt_lds + tl * CLT_TILE + cam_base(ca1) + li * 8;
                const float *tb = clt_lds + tl * CLT_TILE + cam_base(ca2) + li * 8;
                const int xa = (li & 4) ^ cam_swap(ca1), xb = (li & 4) ^ cam_swap(ca2);
                auto load = [&](int color, int which, int qq, int h) -> f32x4 {
                    return *(const f32x4 *)((which ? tb : ta) + color * FB + qq * 64 + ((4 * h) ^ (which ? xb : xa)));
                };
                corr_accumulate<COLORS>(load, a.scale, li, c0, c1, c2, c3);
            }
            __syncthreads();

            // (3) normalise, back-transform, unfold in the (now free) CLT region
            if (active) {
                float *L = clt_lds + slot * SLOT_STRIDE;
                corr_finish(c0, c1, c2, c3, L, L, li, fz2);
                wave_lds_sync();
                if (fr.corrs) {
                    for (int u = 0; u < 8; u++) {
                        const int r = __shfl(row, u * 8);
                        if (r < 0) continue;
                        const float *src = clt_lds + (wave * 8 + u) * SLOT_STRIDE;
                        float *dst = fr.corrs + (size_t)r * a.corr_stride;
#pragma unroll
                        for (int t = 0; t < 4; t++)
                            if (src_off[t] >= 0) store1(dst + lane + 64 * t, src[src_off[t]]);
                    }
                }
            }
            // (4) six-pair combination from the staged pair tiles (the
            // mapping of tp_corr_combine, pairs of the task in pair_mask)
            if (fr.combined) {
                __syncthreads();
                const int r = a.radius, n = 2 * r + 1;
                for (int o = threadIdx.x; o < GT * n * n; o += blockDim.x) {
                    const int tl = o / (n * n), oo = o - tl * n * n;
                    const int s = gs + tl;
                    if (s >= s1) continue;
                    const int task_num = swizzle_strip(s, g.TX, fr.num_tiles);
                    gtask *tk = gtasks(fr.tasks) + task_num;
                    if (tk->task == 0 || task_tile(g, tk->txy) < 0) continue;
                    const int cy = oo / n - r, cx = oo % n - r;
                    float sum = 0.f;
#pragma unroll
                    for (int pr = 0; pr < TP_NUM_PAIRS; pr++) {
                        int sy, sx;
                        if (((a.pair_mask >> pr) & 1) && ((tk->task >> (TP_TASK_CORR_BITS + pr)) & 1) &&
                            pair_src(pr, r, cy, cx, sy, sx))
                            sum += clt_lds[pair_slot(tl, pr) * SLOT_STRIDE + (sy + 7) * 15 + sx + 7];
                    }
                    store1(fr.combined + (size_t)task_num * a.combined_stride + oo, sum);
                }
            }
            __syncthreads();
        }
        if (QUEUE && a.stamps && threadIdx.x == 0)
            atomicMax(&a.stamps[2 * f + 1], (unsigned long long)wall_clock64());
        if (!QUEUE) break;
    }
}

// Stream set-up: each frame's item count, the per-XCD item
// prefix table over frames, the queue counters and the per-frame time
// stamps.  One wavefront; lane x < 8 fills column x of the table.
__global__ void k_stream_setup(FrameDesc *frames, int nframes, int tiles_per_item, int *work,
                               unsigned long long *stamps) {
    const int x = threadIdx.x;
    if (x < 8) {
        int acc = 0;
        for (int f = 0; f < nframes; f++) {
            const int n = (max(frames[f].num_tiles, 0) + tiles_per_item - 1) / tiles_per_item;
            if (x == 0) frames[f].num_items = n;
            work[TP_WORK_TABLE + f * 8 + x] = acc;
            acc += eighth_lo(n, x + 1) - eighth_lo(n, x);
        }
        work[TP_WORK_TABLE + nframes * 8 + x] = acc;  // queue x's length
        work[x * TP_WORK_LINE] = 0;                    // claim counter
    }
    if (x == 0) {
        int total = 0;
        for (int f = 0; f < nframes; f++) total += frames[f].num_items;
        work[TP_WORK_ITEMS] = total;  // number of items
    }
    if (stamps)
        for (int i = threadIdx.x; i < nframes; i += blockDim.x) {
            stamps[2 * i] = ~0ull;
            stamps[2 * i + 1] = 0ull;
        }
}

}  // namespace tp
