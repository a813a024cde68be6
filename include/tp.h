/*
 * tp.h - C ABI of the MI355X (gfx950) tile processor, libtp.so.
 *
 * Each entry point is a host function that validates its arguments and
 * enqueues hand-written HIP kernels on the caller's stream.  Names, argument
 * order and argument meaning follow the reference's extern "C" kernels
 * (src/TileProcessor.cuh, mfkiwl/tile_processor_gpu); the compile-time
 * geometry of the reference (IMG_WIDTH, KERNELS_HOR, ...) is passed at run
 * time in a tp_geometry, and every call takes a stream.
 *
 * Pointer conventions (same as the reference):
 *   - "float **" arguments are DEVICE arrays of TP_NUM_CAMS device pointers
 *     (one buffer per camera), exactly what the reference kernels receive.
 *   - every other pointer is a device pointer unless stated otherwise.
 *   - strides are in floats, not bytes.
 * No entry point allocates device memory or synchronises the stream, so all
 * of them can be captured into a hipGraph.
 *
 * Return value: TP_OK (0) or a negative TP_E* code; tp_last_error() gives a
 * text for the last failure on the calling thread.
 */
#ifndef TP_H
#define TP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- fixed constants (src/TileProcessor.cuh:44-71, src/dtt8x8.cuh:48-52) ---- */
#define TP_DTT_SIZE           8   /* 8x8 DCT/DST, 16x16 lapped windows      */
#define TP_NUM_CAMS           4
#define TP_NUM_COLORS         3   /* Bayer R, B, G (in that order)          */
#define TP_NUM_PAIRS          6
#define TP_KERNELS_LSTEP      4   /* aberration kernel cell = 16x16 px      */
#define TP_CORR_NTILE_SHIFT   8   /* list entry = tile << 8 | low bits      */
#define TP_TASK_CORR_BITS     4   /* tp_task.task bits 4..9: pairs          */
#define TP_TASK_TEXTURE_BITS  0xf /* tp_task.task bits 0..3: N,E,S,W        */
#define TP_LIST_TEXTURE_BIT   7   /* texture list entry flag                */
#define TP_MAX_CORR_RADIUS    7   /* 15x15 correlation tiles                */
#define TP_FAT_ZERO_WEIGHT    0.0001 /* added to texture port weights      */

/* Bayer color indices (src/TileProcessor.cuh:155-157); GR/BG mosaic with red
 * at (row 0, col 1). */
#define TP_BAYER_RED   0
#define TP_BAYER_BLUE  1
#define TP_BAYER_GREEN 2

/* Camera pairs of the correlation list, pair index -> (cam1, cam2)
 * (src/TileProcessor.cuh:484-490):
 * 0:(0,1) 1:(2,3) 2:(0,2) 3:(1,3) 4:(0,3) 5:(2,1) */

/* ---- status codes ---- */
#define TP_OK        0
#define TP_EINVAL   -1   /* bad argument, nothing was enqueued  */
#define TP_ELAUNCH  -2   /* kernel launch failed                */
#define TP_EHIP     -3   /* other HIP runtime error             */

/* Tile task, binary identical to the reference (40 bytes,
 * src/TileProcessor.cuh:179-186).
 *   task : bits 0-3 texture neighbours N,E,S,W; bits 4-9 the 6 pairs;
 *          0 = nothing to do.
 *   txy  : tx | ty << 16.
 *   xy   : tile centre in each camera, pixels (image x, y). */
typedef struct tp_task {
    int32_t task;
    int32_t txy;
    float   xy[TP_NUM_CAMS][2];
} tp_task;

/* Per kernel cell and color offsets, binary identical to the reference's
 * CltExtra (src/TileProcessor.cuh:187-196). */
typedef struct tp_clt_extra {
    float data_x;    /* kernel is relative to this displacement X (px) */
    float data_y;
    float center_x;  /* informative only                               */
    float center_y;
    float dxc_dx;    /* d(data_x)/dx per pixel from the cell centre      */
    float dxc_dy;
    float dyc_dx;
    float dyc_dy;
} tp_clt_extra;

/* Run-time replacement for IMG_WIDTH, IMG_HEIGHT, KERNELS_HOR and
 * KERNELS_VERT (src/TileProcessor.cuh:47-50).  Derived:
 *   tilesx = img_width / 8, tilesy = img_height / 8,
 *   tilesya = (tilesy + 3) & ~3.
 * img_width and img_height must be positive multiples of 16. */
typedef struct tp_geometry {
    int32_t img_width;
    int32_t img_height;
    int32_t kernels_hor;
    int32_t kernels_vert;
} tp_geometry;

typedef void *tp_stream_t; /* a hipStream_t; NULL = the null stream */

/* Layouts the reference's buffers and files fix (tp_task 40 bytes, CltExtra
 * 32 bytes), checked at compile time in C and C++. */
#ifdef __cplusplus
#define TP_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define TP_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif
TP_STATIC_ASSERT(sizeof(tp_task) == 40, "tp_task must be 40 bytes (src/TileProcessor.cuh:179-186)");
TP_STATIC_ASSERT(sizeof(tp_clt_extra) == 32, "tp_clt_extra must be 32 bytes (src/TileProcessor.cuh:187-196)");
TP_STATIC_ASSERT(sizeof(tp_geometry) == 16, "tp_geometry: 4 int32");

/* Text of the last error on this thread ("" if none). */
const char *tp_last_error(void);

/* Library version string. */
const char *tp_version(void);

/* Replaces the reference's cudaMemcpyToSymbol on its low-pass tables
 * (src/TileProcessor.cuh:421-476).  which: 0..3 = lpf_data[R,B,G,mono],
 * 4 = lpf_rb_corr, 5 = lpf_corr.  host_lpf: 64 floats in HOST memory,
 * [vertical freq][horizontal freq], or NULL to restore the built-in table.
 * Applies to the current device; not stream-ordered (call it outside any
 * graph capture).  The built-in tables are the closed-form Gaussian
 * responses the reference's tables hold. */
int tp_set_lpf(int which, const float *host_lpf);

/* Forward CLT of every camera and color of each task tile, with per-tile
 * aberration correction (CLT-domain kernel multiply) and fractional-pixel
 * phase rotation.  Replaces convert_correct_tiles
 * (src/TileProcessor.cuh:1859-1937) and convertCorrectTile (:3106-3634).
 *   gpu_kernel_offsets: [cam] -> tp_clt_extra[kernels_vert][kernels_hor][3]
 *   gpu_kernels       : [cam] -> float[kernels_vert][kernels_hor][3][4][8][8]
 *   gpu_images        : [cam] -> float[img_height][dstride] Bayer image
 *   gpu_tasks         : tp_task[num_tiles]; task == 0 tiles are skipped
 *   gpu_clt           : [cam] -> float[tilesy][tilesx][3][4][8][8] (output)
 *   lpf_mask          : bit c applies lpf_data[c] to color c (reference: 0)
 * Tasks whose (tx, ty) lies outside the tile grid are skipped.
 * Border tiles: windows are placed as the reference does; pixels outside
 * the image are read from the nearest pixel of the same Bayer color. */
int tp_convert_correct_tiles(const tp_geometry *geom,
                             float **gpu_kernel_offsets,
                             float **gpu_kernels,
                             float **gpu_images,
                             const tp_task *gpu_tasks,
                             float **gpu_clt,
                             size_t dstride,
                             int num_tiles,
                             int lpf_mask,
                             tp_stream_t stream);

/* Pairwise phase correlation in the CLT domain, back to a pixel-domain
 * (2r+1)^2 tile per list entry.  Replaces correlate2D
 * (src/TileProcessor.cuh:1176-1443).
 *   gpu_corr_indices : int[num_corr_tiles], tile << 8 | pair (pair 0..5)
 *   gpu_corrs        : row i (corr_stride floats apart) receives entry i
 *   corr_radius      : 0..7
 * Entries with pair > 5 or a tile outside the grid are skipped (row left
 * untouched). */
int tp_correlate2D(const tp_geometry *geom,
                   float **gpu_clt,
                   int colors,
                   float scale0,
                   float scale1,
                   float scale2,
                   float fat_zero,
                   size_t num_corr_tiles,
                   const int *gpu_corr_indices,
                   size_t corr_stride,
                   int corr_radius,
                   float *gpu_corrs,
                   tp_stream_t stream);

/* Inverse CLT of one color of one camera, overlap-added into a plane, one of
 * the 4 tile-parity passes.  Replaces imclt_rbg
 * (src/TileProcessor.cuh:2588-2712).
 *   gpu_clt : float[tilesy][tilesx][3][4][8][8] of ONE camera
 *   gpu_rbg : 3 color planes of (img_height + 8) rows x dstride floats;
 *             tile (tx,ty) lands at row 8*ty, column 8*tx of its plane.
 * Pass (v_offset, h_offset) = (0,0) initialises the tiles it covers, the
 * other three accumulate.  The reference's debug writes (tile (0,0) set to
 * 100, tile (161,111) doubled) are not reproduced. */
int tp_imclt_rbg(const tp_geometry *geom,
                 const float *gpu_clt,
                 float *gpu_rbg,
                 int color,
                 int v_offset,
                 int h_offset,
                 size_t dstride,
                 tp_stream_t stream);

/* Whole-frame variant: for each camera and color, the plane that the 4
 * passes of tp_imclt_rbg give (each pixel the sum of the 4 windowed tile
 * outputs that cover it, rounded in another order), in one launch.  Every
 * pixel of the (W + 8) x (H + 8) planes is written (nothing accumulates).
 * gpu_rbg: [cam] -> planes as in tp_imclt_rbg (device pointer array). */
int tp_imclt_rbg_all(const tp_geometry *geom,
                     float **gpu_clt,
                     float **gpu_rbg,
                     size_t dstride,
                     tp_stream_t stream);

/* Textures of a list of tiles.  Replaces textures_accumulate
 * (src/TileProcessor.cuh:2207-2585): per tile the 4 cameras' CLT are
 * low-passed (lpf_data), inverse transformed, debayered with shot-noise
 * companding, combined by the 4-camera consensus, and written as
 *   - non-overlapping tiles (if gpu_texture_tiles && texture_stride):
 *     entry i -> gpu_texture_tiles + i * texture_stride, slices of 256
 *     floats: colors, alpha (+ 4 port weights, colors + 1 rms if
 *     keep_weights); and/or
 *   - the alpha-faded overlap-add into a woi-relative RGBA image (if
 *     gpu_texture_rbg && texture_rbg_stride; entries with texture mask
 *     bits 0..3 == 0 are not added): colors + 1 slices of
 *     (woi[3] + 1) * 8 rows of texture_rbg_stride floats.
 *   gpu_texture_indices : tile << 8 | 1 << 7 | N,E,S,W mask
 *   gpu_port_offsets    : float[4][2] camera port offsets
 *   woi                 : device int[4] x, y, width, height (tiles)
 * Tiles of one call must not overlap in the RGBA image (the reference runs
 * one call per tile parity); is_lwir is unused, as in the reference. */
int tp_textures_accumulate(const tp_geometry *geom,
                           const int *woi,
                           float **gpu_clt,
                           size_t num_texture_tiles,
                           const int *gpu_texture_indices,
                           const float *gpu_port_offsets,
                           int colors,
                           int is_lwir,
                           float min_shot,
                           float scale_shot,
                           float diff_sigma,
                           float diff_threshold,
                           float min_agree,
                           float weight0,
                           float weight1,
                           float weight2,
                           int dust_remove,
                           int keep_weights,
                           size_t texture_rbg_stride,
                           float *gpu_texture_rbg,
                           size_t texture_stride,
                           float *gpu_texture_tiles,
                           tp_stream_t stream);

/* Texture list of a task list.  Replaces prepare_texture_list
 * (src/TileProcessor.cuh:1643-1713) and its child kernels (:1716-1854),
 * which the reference launches with CUDA dynamic parallelism; here they are
 * stream-ordered kernels with device-side counts.
 *   gpu_tasks          : bits 0..3 of each textured task are REPLACED by the
 *                        N,E,S,W neighbours that are textured too (a tile
 *                        with none drops out of the list, as in the reference)
 *   gpu_texture_indices: int[tilesx * tilesya]; 4 parity sub-lists of
 *                        tilesx * tilesya / 4 entries, inner tiles from the
 *                        start, border tiles from the end of each; the
 *                        first tilesx * height entries also hold the tile
 *                        occupancy map while the list is built
 *   num_texture_tiles  : device int[8]: [2 * parity + border]
 *   woi                : device int[4] x, y, width, height of the textured
 *                        tiles
 *   width, height      : extent of the occupancy map to clear (<= tilesx,
 *                        <= tilesy)
 * Entry order within a sub-list is unspecified (as in the reference). */
int tp_prepare_texture_list(const tp_geometry *geom,
                            tp_task *gpu_tasks,
                            int num_tiles,
                            int *gpu_texture_indices,
                            int *num_texture_tiles,
                            int *woi,
                            int width,
                            int height,
                            tp_stream_t stream);

/* Texture list + RGBA image in one stream-ordered sequence.  Replaces
 * generate_RBGA (src/TileProcessor.cuh:1446-1613): prepare_texture_list,
 * then clearing the RGBA image of the woi, then the textures_accumulate
 * passes (one launch per tile parity, taking its inner and its border list;
 * the reference runs the two lists as separate passes, which only changes
 * the order of the overlap-add sums) into gpu_texture_tiles, an RGBA
 * image of colors + 1 slices with (woi height + 1) * 8 rows of
 * texture_rbga_stride floats (allocate (tilesy + 1) * 8 rows per slice).
 * No host synchronisation: pass sizes are read on the device.  keep_weights
 * is ignored (the reference's passes use 0). */
int tp_generate_RBGA(const tp_geometry *geom,
                     tp_task *gpu_tasks,
                     int num_tiles,
                     int *gpu_texture_indices,
                     int *num_texture_tiles,
                     int *woi,
                     int width,
                     int height,
                     float **gpu_clt,
                     const float *gpu_port_offsets,
                     int colors,
                     int is_lwir,
                     float min_shot,
                     float scale_shot,
                     float diff_sigma,
                     float diff_threshold,
                     float min_agree,
                     float weight0,
                     float weight1,
                     float weight2,
                     int dust_remove,
                     int keep_weights,
                     size_t texture_rbga_stride,
                     float *gpu_texture_tiles,
                     tp_stream_t stream);

/* Six-pair combination (no counterpart in the reference: its Java host
 * combines the per-pair tiles of correlate2D, src/TileProcessor.cuh:1421).
 * Each pair's (2r+1)^2 tile is mapped onto the horizontal disparity axis
 * (pairs 0,1 as is; 2,3 transposed; 4 (0,3): out(y,x) = in(x+y, x-y);
 * 5 (2,1): out(y,x) = in(y-x, x+y), offsets from the centre) and the tiles
 * of the pairs in pair_mask are summed into gpu_combined + tile *
 * combined_stride.  Entries of one tile must be adjacent in the list (as in
 * the tile-major list the reference harness builds); tiles without entries
 * are not written.  A residual disparity dd peaks at (r, r + dd). */
int tp_corr_combine(const tp_geometry *geom,
                    size_t num_corr_tiles,
                    const int *gpu_corr_indices,
                    const float *gpu_corrs,
                    size_t corr_stride,
                    int corr_radius,
                    int pair_mask,
                    float *gpu_combined,
                    size_t combined_stride,
                    tp_stream_t stream);

/* Sub-pixel disparity from the centre row of num_tiles combined tiles
 * (no counterpart in the reference): integer argmax, then a 3-point
 * parabola (mode 0) or the centroid of the positive samples within +-2
 * (mode 1).  gpu_out: float[num_tiles][2] = (disparity in px, peak value). */
int tp_corr_disparity(const float *gpu_tiles,
                      size_t stride,
                      int num_tiles,
                      int corr_radius,
                      int mode,
                      float *gpu_out,
                      tp_stream_t stream);

/* ------------------------------------------------------------------------
 * Fused and stream forms (no counterpart in the reference).  They compute
 * what tp_convert_correct_tiles followed by tp_correlate2D computes, with the
 * same per-tile arithmetic, but keep each tile's CLT in on-chip memory
 * between the two steps.  Correlation rows are dense per task: the pair p of
 * task t goes to row 6 t + p (for a full task list with all pairs set, the
 * same order as the tile-major correlation list the reference harness
 * builds, src/dtt8x8.cu:453-472); rows of pairs not set in a task are not
 * written.
 * ---------------------------------------------------------------------- */

/* Launch configuration (NULL: defaults). */
typedef struct tp_fused_config {
    int32_t tiles_per_item;   /* tiles per work item, a multiple of tiles_per_group (default 12) */
    int32_t tiles_per_group;  /* tiles held in LDS at once: 2 or 4 (default 4) */
    int32_t blocks_per_cu;    /* stream form: resident workgroups per CU, 1..8 (default 4) */
    int32_t dct;              /* forward DCT-IV/DST-IV: 0 FP32 butterflies (default); on the matrix
                                 cores with split operands and FP32 accumulation: 1 FP16 hi + lo,
                                 2 BF16 hi + lo (about 1e-4 relative), 3 BF16 hi + mid + lo */
} tp_fused_config;
TP_STATIC_ASSERT(sizeof(tp_fused_config) == 16, "tp_fused_config: 4 int32");

/* One frame of the stream form; an array of these lives in device memory.
 * images / kernels / kernel_offsets / clt are device arrays of TP_NUM_CAMS
 * device pointers as for tp_convert_correct_tiles; clt may be NULL (CLT not
 * stored), corrs may be NULL (pair tiles not stored), combined may be NULL
 * (no six-pair combination).  num_items is written by the library. */
typedef struct tp_frame {
    const tp_task *tasks;
    float **images;
    float **kernels;
    float **kernel_offsets;
    float **clt;
    float *corrs;
    float *combined;
    int32_t num_tiles;
    int32_t num_items;
} tp_frame;
TP_STATIC_ASSERT(sizeof(tp_frame) == 64, "tp_frame: 7 pointers and 2 int32");

/* Forward CLT + correlation of one frame.  Arguments as for
 * tp_convert_correct_tiles and tp_correlate2D; gpu_clt and gpu_corrs may be
 * NULL.  gpu_combined (or NULL): row `task` receives the combination of the
 * task's pairs in pair_mask, as tp_corr_combine computes it
 * (combined_stride floats per row). */
int tp_convert_correlate(const tp_geometry *geom,
                         float **gpu_kernel_offsets,
                         float **gpu_kernels,
                         float **gpu_images,
                         const tp_task *gpu_tasks,
                         float **gpu_clt,
                         size_t dstride,
                         int num_tiles,
                         int lpf_mask,
                         int colors,
                         float scale0,
                         float scale1,
                         float scale2,
                         float fat_zero,
                         size_t corr_stride,
                         int corr_radius,
                         float *gpu_corrs,
                         float *gpu_combined,
                         size_t combined_stride,
                         int pair_mask,
                         const tp_fused_config *config,
                         tp_stream_t stream);

#define TP_STREAM_WORK_INTS(nframes) (296 + 8 * (nframes))

/* Forward CLT + correlation of a stream of frames (each with its own task
 * list, images, kernels and outputs) in one persistent launch: a fixed grid
 * of blocks_per_cu workgroups per CU takes work items from 8 per-XCD queues
 * (queue x: the x-th eighth of every frame's items, frame after frame, so
 * that all XCDs work on the same frames at a time; idle workgroups steal).
 * gpu_work: device int32[TP_STREAM_WORK_INTS(nframes)], written by the call:
 * the 8 queue counters at [32 x] (one 128-byte line each, so that the claims
 * of different XCDs do not contend for one line), the item count at [256],
 * and from [288] a [nframes + 1][8] table of per-queue item prefixes.
 * gpu_stamps: NULL or device uint64[nframes][2], set to each frame's first
 * item start and last item end (GPU wall clock, wall_clock_hz ticks/s). */
int tp_stream_convert_correlate(const tp_geometry *geom,
                                tp_frame *gpu_frames,
                                int nframes,
                                int32_t *gpu_work,
                                uint64_t *gpu_stamps,
                                size_t dstride,
                                int lpf_mask,
                                int colors,
                                float scale0,
                                float scale1,
                                float scale2,
                                float fat_zero,
                                size_t corr_stride,
                                int corr_radius,
                                size_t combined_stride,
                                int pair_mask,
                                const tp_fused_config *config,
                                tp_stream_t stream);

/* tp_stream_convert_correlate without the gpu_work argument.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, hosts that forbid
 * it).  The work buffer is a fixed array of the library, one per GPU, sized
 * for TP_STREAM_NOALLOC_MAX_FRAMES frames; larger nframes is refused with
 * TP_EINVAL.  The call allocates nothing (the first call on a device looks
 * up that device's array).  Every call on one GPU uses the same array:
 * calls that overlap on one GPU (on different streams, or from several
 * threads) race and must be serialised by the caller; consecutive calls on
 * one stream are fine (each call re-initialises the array). */
#define TP_STREAM_NOALLOC_MAX_FRAMES 4096
int tp_stream_convert_correlate_noalloc(const tp_geometry *geom,
                                        tp_frame *gpu_frames,
                                        int nframes,
                                        uint64_t *gpu_stamps,
                                        size_t dstride,
                                        int lpf_mask,
                                        int colors,
                                        float scale0,
                                        float scale1,
                                        float scale2,
                                        float fat_zero,
                                        size_t corr_stride,
                                        int corr_radius,
                                        size_t combined_stride,
                                        int pair_mask,
                                        const tp_fused_config *config,
                                        tp_stream_t stream);

/* The same stream launch captured once into an executable HIP graph
 * (set-up kernel + persistent kernel, no memset nodes); replaying it
 * re-reads every buffer, so new frame data can be uploaded between
 * launches. */
typedef void *tp_graph_t;
int tp_stream_graph_create(const tp_geometry *geom,
                           tp_frame *gpu_frames,
                           int nframes,
                           int32_t *gpu_work,
                           uint64_t *gpu_stamps,
                           size_t dstride,
                           int lpf_mask,
                           int colors,
                           float scale0,
                           float scale1,
                           float scale2,
                           float fat_zero,
                           size_t corr_stride,
                           int corr_radius,
                           size_t combined_stride,
                           int pair_mask,
                           const tp_fused_config *config,
                           tp_graph_t *graph);
int tp_graph_launch(tp_graph_t graph, tp_stream_t stream);
int tp_graph_destroy(tp_graph_t graph);

/* Ticks per second of the clock behind gpu_stamps. */
double tp_wall_clock_hz(void);

/* ------------------------------------------------------------------------
 * Disparity-refinement CNN (no counterpart in the reference: its Java host
 * took a parabolic/centroid argmax of the correlation tiles).  Per tile, 7
 * input planes of 15 x 15 (the combined correlation and the 6 pair tiles
 * mapped onto the horizontal disparity axis), two 3x3 convolutions of 32
 * channels with ReLU, a 3x3 convolution to a logit map and a confidence
 * map; the residual disparity is the soft-argmax column of the logits
 * (px from the centre), the confidence the sigmoid of the confidence map
 * pooled by the same softmax.  Parameters: TP_NN_PARAMS values in the
 * layout W1[32][7][3][3], b1[32], W2[32][32][3][3], b2[32], W3[2][32][3][3],
 * b3[2]; the kernels read a BF16 copy and compute in FP32.
 * ---------------------------------------------------------------------- */
#define TP_NN_PARAMS 11874
#define TP_NN_INPUT 1575          /* floats of input per tile (7 x 15 x 15) */
#define TP_NN_ACTS 14850          /* floats of saved activations per tile */

/* Network input from fused/stream correlation rows (6 task + pair, radius
 * 7): gpu_x[num_tasks][7][15][15]. */
int tp_nn_input(const tp_geometry *geom, const tp_task *gpu_tasks, int num_tasks,
                const float *gpu_corrs, size_t corr_stride, float *gpu_x, tp_stream_t stream);

/* gpu_out[n][2] = (residual disparity px, confidence); gpu_acts (NULL for
 * inference) receives TP_NN_ACTS floats per tile for tp_nn_backward. */
int tp_nn_forward(const uint16_t *gpu_params_bf16, const float *gpu_x, int n, float *gpu_acts,
                  float *gpu_out, tp_stream_t stream);

/* Inference form of tp_nn_forward: BF16 activations, the two 32-channel
 * convolutions on the matrix cores (FP32 accumulation). */
int tp_nn_infer(const uint16_t *gpu_params_bf16, const float *gpu_x, int n, float *gpu_out,
                tp_stream_t stream);

/* tp_nn_input followed by tp_nn_infer in one pass: the network's input is
 * assembled from the correlation rows (6 task + pair, radius 7) as it is
 * read, without the [num_tasks][7][225] buffer; the same results. */
int tp_nn_infer_corrs(const tp_geometry *geom, const uint16_t *gpu_params_bf16, const tp_task *gpu_tasks,
                      int num_tasks, const float *gpu_corrs, size_t corr_stride, float *gpu_out,
                      tp_stream_t stream);

/* Adds the gradient of the mean training loss over the tiles with a finite
 * target (huber(d - target) + conf_weight * BCE(confidence, |d - target| <
 * 0.5)) to gpu_grads[TP_NN_PARAMS] and the loss to gpu_loss[0];
 * loss_scale = 1 / (number of tiles with a finite target). */
int tp_nn_backward(const uint16_t *gpu_params_bf16, const float *gpu_x, const float *gpu_acts,
                   const float *gpu_target, int n, float conf_weight, float huber_delta,
                   float loss_scale, float *gpu_grads, float *gpu_loss, tp_stream_t stream);

/* One Adam step on the FP32 master parameters (step >= 1), refreshing the
 * BF16 copy. */
int tp_nn_adam(float *gpu_params, float *gpu_m, float *gpu_v, const float *gpu_grads, int n,
               float lr, float beta1, float beta2, float eps, int step,
               uint16_t *gpu_params_bf16, tp_stream_t stream);

/* BF16 copy (round to nearest even) of n FP32 values. */
int tp_nn_pack_bf16(const float *gpu_src, int n, uint16_t *gpu_dst, tp_stream_t stream);

/* Aberration kernels in the CLT domain and their CltExtra records, from
 * sampled point-spread functions.  The reference reads both from files its
 * Java host generated (src/dtt8x8.cu:228-238, :364-378, formats
 * [KV][KH][3][4][8][8] and CltExtra[KV][KH][3]); this builds them in those
 * layouts, ready for tp_convert_correct_tiles.
 *   gpu_psfs  : [cam] -> float[kernels_vert][kernels_hor][3][psf_height]
 *               [psf_width]: the spatial kernel of each cell and color,
 *               element [i][j] at (j - ox, i - oy) px from the origin
 *   psf_height, psf_width: 1..32
 *   center    : 0: origin (ox, oy) = the central element ((w-1)/2, (h-1)/2),
 *               data_x = data_y = 0 and zero slopes; 1: origin = the
 *               centroid of the kernel (a kernel summing to 0 keeps the
 *               central element), data_x / data_y = the centroid relative to
 *               the central element, dxc_dx, dxc_dy, dyc_dx, dyc_dy = its
 *               central differences over the kernel grid (16 px per cell,
 *               one-sided at the grid edges)
 *   gpu_kernels        : [cam] -> CC, SC, CS, SS quadrants of
 *                        K(k, l) = sum p e^{i u_k x} e^{j u_l y},
 *                        u_k = pi (2k+1) / 16, [l][k] (output)
 *   gpu_kernel_offsets : [cam] -> tp_clt_extra records; center_x, center_y
 *                        = 16 kx - 8, 16 ky - 8 (output)
 * Sums in FP64, stored as FP32. */
int tp_psf_to_clt(const tp_geometry *geom, float **gpu_psfs, int psf_height, int psf_width,
                  int center, float **gpu_kernels, float **gpu_kernel_offsets, tp_stream_t stream);

/* 2-D 8x8 transform of every 8x8 tile of a width x height image, in place
 * of the tile (row = vertical frequency, column = horizontal frequency).
 * Replaces GPU_DTT24_DRV (src/dtt8x8.cuh:103-161), the reference's test
 * entry for its DCT/DST butterflies; width and height replace the grid its
 * caller launches (32 x 16 pixel blocks) and may be any multiples of 8.
 *   dtt_mode: 0 horizontal DCT-IV, vertical DCT-IV; 1 DST-IV, DCT-IV;
 *             2 DCT-IV, DST-IV; 3 DST-IV, DST-IV; 4-7 the same with DCT-II /
 *             DST-II (all orthonormal)
 *   gpu_src, gpu_dst: float[height][src_stride] (may be the same buffer) */
int tp_dtt8x8(float *gpu_dst, const float *gpu_src, int src_stride, int dtt_mode, int width,
              int height, tp_stream_t stream);

/* gpu_dst[0..n) = value, by a kernel (safe inside captured graphs). */
int tp_fill(float *gpu_dst, size_t n, float value, tp_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TP_H */
