// tp_host.hpp - C++ host side of libtp.so: the work the reference's C++
// harness (src/dtt8x8.cu) does around its kernel launches, over the C ABI of
// include/tp.h.  Header-only; link with -ltp -lamdhip64.
//
//  - device buffers owned by RAII handles; images and row outputs pitched as
//    the harness allocates them (cudaMallocPitch, src/dtt8x8.cu:117-142), the
//    per-camera device pointer arrays that the float ** arguments take
//    (:103-114);
//  - the harness's raw float files (readFloatsFromFile / writeFloatsToFile,
//    :144-165);
//  - the line-scan task list with every pair and texture bit (:417-428), the
//    tile-major correlation list (:453-472) and the texture list (:474-492);
//  - Frame: one 4-camera frame's buffers and the calls of the path.  Every
//    call checks the library's status code and throws tph::Error carrying
//    tp_last_error(); nothing is synchronised except the downloads.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tp.h"

namespace tph {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int rc, const char *what) {
    if (rc != TP_OK) throw Error(std::string(what) + ": " + tp_last_error());
}

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

// Sizes that the reference fixes at compile time (src/TileProcessor.cuh:47-58).
struct Dims {
    int W, H, KH, KV, TX, TY, TYA;
    explicit Dims(const tp_geometry &g)
        : W(g.img_width), H(g.img_height), KH(g.kernels_hor), KV(g.kernels_vert), TX(g.img_width / 8),
          TY(g.img_height / 8), TYA(((g.img_height / 8) + 3) & ~3) {}
    size_t tiles() const { return (size_t)TX * TY; }
    size_t kernel_floats() const { return (size_t)KV * KH * TP_NUM_COLORS * 256; }
    size_t offset_floats() const { return (size_t)KV * KH * TP_NUM_COLORS * 8; }
    size_t clt_floats() const { return tiles() * TP_NUM_COLORS * 256; }
};

// One device allocation of n elements.
template <class T>
class DeviceArray {
  public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) : n_(n) {
        if (n) hip_check(hipMalloc((void **)&p_, n * sizeof(T)), "hipMalloc");
    }
    ~DeviceArray() {
        if (p_) (void)hipFree(p_);
    }
    DeviceArray(DeviceArray &&o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceArray &operator=(DeviceArray &&o) noexcept {
        if (this != &o) {
            if (p_) (void)hipFree(p_);
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;

    T *get() const { return p_; }
    size_t size() const { return n_; }
    void upload(const T *h, size_t n) {
        if (n > n_) throw Error("upload: more elements than the allocation holds");
        hip_check(hipMemcpy(p_, h, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy to device");
    }
    void upload(const std::vector<T> &h) { upload(h.data(), h.size()); }
    std::vector<T> download(size_t n) const {
        if (n > n_) throw Error("download: more elements than the allocation holds");
        std::vector<T> h(n);
        hip_check(hipMemcpy(h.data(), p_, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy to host");
        return h;
    }

  private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

// Pitched array of `height` rows of `width` floats, `stride` floats apart
// (hipMallocPitch's pitch in floats).
class PitchedArray {
  public:
    PitchedArray() = default;
    PitchedArray(int width, int height) : width_(width), height_(height) {
        size_t pitch = 0;
        hip_check(hipMallocPitch((void **)&p_, &pitch, (size_t)width * sizeof(float), (size_t)height),
                  "hipMallocPitch");
        stride_ = pitch / sizeof(float);
    }
    ~PitchedArray() {
        if (p_) (void)hipFree(p_);
    }
    PitchedArray(PitchedArray &&o) noexcept { *this = std::move(o); }
    PitchedArray &operator=(PitchedArray &&o) noexcept {
        if (this != &o) {
            if (p_) (void)hipFree(p_);
            p_ = o.p_;
            stride_ = o.stride_;
            width_ = o.width_;
            height_ = o.height_;
            o.p_ = nullptr;
        }
        return *this;
    }
    PitchedArray(const PitchedArray &) = delete;
    PitchedArray &operator=(const PitchedArray &) = delete;

    float *get() const { return p_; }
    size_t stride() const { return stride_; }  // floats
    int width() const { return width_; }
    int height() const { return height_; }
    // rows of `width` packed floats
    void upload(const std::vector<float> &h) {
        if (h.size() != (size_t)width_ * height_) throw Error("pitched upload: size mismatch");
        hip_check(hipMemcpy2D(p_, stride_ * sizeof(float), h.data(), (size_t)width_ * sizeof(float),
                              (size_t)width_ * sizeof(float), (size_t)height_, hipMemcpyHostToDevice),
                  "hipMemcpy2D to device");
    }
    // the first `rows` rows, packed
    std::vector<float> download(int rows = -1) const {
        if (rows < 0) rows = height_;
        std::vector<float> h((size_t)width_ * rows);
        if (rows)
            hip_check(hipMemcpy2D(h.data(), (size_t)width_ * sizeof(float), p_, stride_ * sizeof(float),
                                  (size_t)width_ * sizeof(float), (size_t)rows, hipMemcpyDeviceToHost),
                      "hipMemcpy2D to host");
        return h;
    }
    void fill(float v, hipStream_t st) { check(tp_fill(p_, stride_ * height_, v, st), "tp_fill"); }

  private:
    float *p_ = nullptr;
    size_t stride_ = 0;
    int width_ = 0, height_ = 0;
};

// ---------------------------------------------------------------- files

// A file of exactly n little-endian floats (the harness's raw files).
inline std::vector<float> read_floats(const std::string &path, size_t n) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error("cannot open " + path);
    std::vector<float> v(n);
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(float)));
    if ((size_t)f.gcount() != n * sizeof(float))
        throw Error(path + ": expected " + std::to_string(n) + " floats");
    if (f.peek() != std::char_traits<char>::eof())
        throw Error(path + ": longer than " + std::to_string(n) + " floats");
    return v;
}

inline void write_floats(const std::string &path, const float *data, size_t n) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw Error("cannot create " + path);
    f.write(reinterpret_cast<const char *>(data), (std::streamsize)(n * sizeof(float)));
    if (!f) throw Error("cannot write " + path);
}

inline void write_floats(const std::string &path, const std::vector<float> &v) {
    write_floats(path, v.data(), v.size());
}

// ---------------------------------------------------------------- lists

// Every tile in line-scan order with the given task bits (default: all 4
// texture neighbours and all 6 pairs); xy[cam] holds the camera's tile
// centres, [tiles][2] (a .portsxy file).
inline std::vector<tp_task> linescan_tasks(
    const tp_geometry &g, const std::array<std::vector<float>, TP_NUM_CAMS> &xy,
    int32_t task = TP_TASK_TEXTURE_BITS | (((1 << TP_NUM_PAIRS) - 1) << TP_TASK_CORR_BITS)) {
    const Dims d(g);
    std::vector<tp_task> t(d.tiles());
    for (size_t nt = 0; nt < t.size(); nt++) {
        const int ty = (int)(nt / d.TX), tx = (int)(nt % d.TX);
        t[nt].task = task;
        t[nt].txy = tx | (ty << 16);
        for (int c = 0; c < TP_NUM_CAMS; c++) {
            if (xy[c].size() != 2 * d.tiles()) throw Error("tile centres: one (x, y) per tile expected");
            t[nt].xy[c][0] = xy[c][2 * nt];
            t[nt].xy[c][1] = xy[c][2 * nt + 1];
        }
    }
    return t;
}

// Correlation list, tile-major and pair-minor: (tx + ty tilesx) << 8 | pair
// for every pair bit of every task.
inline std::vector<int32_t> corr_list(const tp_geometry &g, const std::vector<tp_task> &tasks) {
    const Dims d(g);
    std::vector<int32_t> idx;
    idx.reserve(tasks.size() * TP_NUM_PAIRS);
    for (const tp_task &t : tasks) {
        const int nt = (t.txy & 0xffff) + (t.txy >> 16) * d.TX;
        const int cm = (t.task >> TP_TASK_CORR_BITS) & ((1 << TP_NUM_PAIRS) - 1);
        for (int b = 0; b < TP_NUM_PAIRS; b++)
            if (cm & (1 << b)) idx.push_back((nt << TP_CORR_NTILE_SHIFT) | b);
    }
    return idx;
}

// Texture list: tile << 8 | 1 << 7 for every task with a texture bit.
inline std::vector<int32_t> texture_list(const tp_geometry &g, const std::vector<tp_task> &tasks) {
    const Dims d(g);
    std::vector<int32_t> idx;
    for (const tp_task &t : tasks) {
        if (!(t.task & TP_TASK_TEXTURE_BITS)) continue;
        const int nt = (t.txy & 0xffff) + (t.txy >> 16) * d.TX;
        idx.push_back((nt << TP_CORR_NTILE_SHIFT) | (1 << TP_LIST_TEXTURE_BIT));
    }
    return idx;
}

// ---------------------------------------------------------------- frame

// The harness's texture parameters (src/dtt8x8.cu:791-814).
struct TextureParams {
    int colors = 3;
    int is_lwir = 0;
    float min_shot = 10.0f, scale_shot = 3.0f, diff_sigma = 1.5f, diff_threshold = 10.0f;
    float min_agree = 3.0f;
    float weight0 = 0.294118f, weight1 = 0.117647f, weight2 = 0.588235f;
    int dust_remove = 1;
    int keep_weights = 0;
    int tile_floats() const { return (colors + 1 + (keep_weights ? TP_NUM_CAMS + colors + 1 : 0)) * 256; }
};

// Port offsets of the 4 cameras (src/dtt8x8.cu:268-272).
constexpr float kPortOffsets[TP_NUM_CAMS][2] = {{-0.5f, -0.5f}, {0.5f, -0.5f}, {-0.5f, 0.5f}, {0.5f, 0.5f}};

class Frame {
  public:
    Frame(const tp_geometry &g, hipStream_t stream = nullptr, int corr_radius = TP_MAX_CORR_RADIUS)
        : g_(g), d_(g), st_(stream), radius_(corr_radius) {
        if (g.img_width <= 0 || g.img_height <= 0 || g.img_width % 16 || g.img_height % 16)
            throw Error("geometry: image width and height must be positive multiples of 16");
        std::array<float *, TP_NUM_CAMS> k{}, o{}, im{}, cl{}, rb{};
        for (int c = 0; c < TP_NUM_CAMS; c++) {
            kernels_[c] = DeviceArray<float>(d_.kernel_floats());
            offsets_[c] = DeviceArray<float>(d_.offset_floats());
            clt_[c] = DeviceArray<float>(d_.clt_floats());
            images_[c] = PitchedArray(d_.W, d_.H);
            // the harness's imclt planes: 3 colors of (H + 8) rows of W + 8
            rbg_[c] = PitchedArray(d_.W + TP_DTT_SIZE, TP_NUM_COLORS * (d_.H + TP_DTT_SIZE));
            k[c] = kernels_[c].get();
            o[c] = offsets_[c].get();
            im[c] = images_[c].get();
            cl[c] = clt_[c].get();
            rb[c] = rbg_[c].get();
        }
        for (int c = 1; c < TP_NUM_CAMS; c++)
            if (images_[c].stride() != images_[0].stride() || rbg_[c].stride() != rbg_[0].stride())
                throw Error("pitched allocations of one frame differ in pitch");
        p_kernels_ = pointer_array(k);
        p_offsets_ = pointer_array(o);
        p_images_ = pointer_array(im);
        p_clt_ = pointer_array(cl);
        p_rbg_ = pointer_array(rb);
        ports_ = DeviceArray<float>(2 * TP_NUM_CAMS);
        ports_.upload(&kPortOffsets[0][0], 2 * TP_NUM_CAMS);
        woi_ = DeviceArray<int32_t>(4);
        num_tex_ = DeviceArray<int32_t>(8);
    }

    const tp_geometry &geometry() const { return g_; }
    const Dims &dims() const { return d_; }
    hipStream_t stream() const { return st_; }

    // inputs: a camera's CLT-domain kernels [KV][KH][3][4][8][8] and CltExtra
    // records [KV][KH][3] (a .kernel and a .kernel_offsets file), its Bayer
    // image [H][W] (a .bayer file); the task list
    void set_kernels(int cam, const std::vector<float> &kernels, const std::vector<float> &offsets) {
        if (kernels.size() != d_.kernel_floats() || offsets.size() != d_.offset_floats())
            throw Error("kernels: sizes do not match the kernel grid");
        kernels_.at(cam).upload(kernels);
        offsets_.at(cam).upload(offsets);
    }
    void set_image(int cam, const std::vector<float> &bayer) { images_.at(cam).upload(bayer); }
    void set_tasks(const std::vector<tp_task> &tasks) {
        tasks_host_ = tasks;
        tasks_ = DeviceArray<tp_task>(std::max<size_t>(1, tasks.size()));
        if (!tasks.empty()) tasks_.upload(tasks);
        const std::vector<int32_t> ci = corr_list(g_, tasks), ti = texture_list(g_, tasks);
        num_corr_ = ci.size();
        num_tex_list_ = ti.size();
        corr_idx_ = DeviceArray<int32_t>(std::max<size_t>(1, num_corr_));
        if (num_corr_) corr_idx_.upload(ci);
        // texture indices: room for the 4 parity sub-lists of generate_RBGA
        tex_idx_ = DeviceArray<int32_t>(std::max<size_t>((size_t)d_.TX * d_.TYA, std::max<size_t>(1, num_tex_list_)));
        if (num_tex_list_) tex_idx_.upload(ti);
        const int n = 2 * radius_ + 1;
        corrs_ = PitchedArray(n * n, (int)std::max<size_t>(1, num_corr_));
    }
    // the tasks after generate_RBGA / prepare_texture_list (which write the
    // neighbour bits into them)
    std::vector<tp_task> tasks() const { return tasks_.download(tasks_host_.size()); }

    // ---- the path (src/dtt8x8.cu:563-1050 in order)

    void convert(int lpf_mask = 0) {
        check(tp_convert_correct_tiles(&g_, p_offsets_.get(), p_kernels_.get(), p_images_.get(), tasks_.get(),
                                       p_clt_.get(), images_[0].stride(), (int)tasks_host_.size(), lpf_mask, st_),
              "tp_convert_correct_tiles");
    }

    // the reference's 48 launches: 4 cameras x 3 colors x 4 tile parities
    void imclt_passes() {
        for (int c = 0; c < TP_NUM_CAMS; c++)
            for (int color = 0; color < TP_NUM_COLORS; color++)
                for (int v = 0; v < 2; v++)
                    for (int h = 0; h < 2; h++)
                        check(tp_imclt_rbg(&g_, clt_[c].get(), rbg_[c].get(), color, v, h, rbg_[0].stride(), st_),
                              "tp_imclt_rbg");
    }
    // zero the imclt planes: the 48-launch form's first pass overwrites only
    // the pixels its tiles cover, and the rows and columns past the last
    // even tile take sums from the odd ones on top of what the plane held
    void clear_rbg() {
        for (int c = 0; c < TP_NUM_CAMS; c++) rbg_[c].fill(0.f, st_);
    }
    // the same in one launch
    void imclt_all() {
        check(tp_imclt_rbg_all(&g_, p_clt_.get(), p_rbg_.get(), rbg_[0].stride(), st_), "tp_imclt_rbg_all");
    }

    // forward CLT and pair correlation in one pass (no counterpart in the
    // reference): rows 6 task + pair of corrs_fused(), which for a full
    // line-scan task list are the correlation list's order; the CLT is not
    // stored (store_clt: also into the CLT buffers)
    void convert_correlate(int colors, float scale0, float scale1, float scale2, float fat_zero,
                           bool store_clt = false, const tp_fused_config *config = nullptr) {
        const int n = 2 * radius_ + 1;
        const size_t rows = std::max<size_t>(1, TP_NUM_PAIRS * tasks_host_.size());
        if ((size_t)fused_.height() < rows) fused_ = PitchedArray(n * n, (int)rows);
        check(tp_convert_correlate(&g_, p_offsets_.get(), p_kernels_.get(), p_images_.get(), tasks_.get(),
                                   store_clt ? p_clt_.get() : nullptr, images_[0].stride(),
                                   (int)tasks_host_.size(), 0, colors, scale0, scale1, scale2, fat_zero,
                                   fused_.stride(), radius_, fused_.get(), nullptr, 0, 0x3F, config, st_),
              "tp_convert_correlate");
    }

    void correlate(int colors, float scale0, float scale1, float scale2, float fat_zero) {
        check(tp_correlate2D(&g_, p_clt_.get(), colors, scale0, scale1, scale2, fat_zero, num_corr_,
                             corr_idx_.get(), corrs_.stride(), radius_, corrs_.get(), st_),
              "tp_correlate2D");
    }

    // textures_accumulate of the host-built texture list into non-overlapping
    // tiles (the harness passes no woi and no RGBA image)
    void textures(const TextureParams &p) {
        const int tf = p.tile_floats();
        if (tex_tiles_.width() != tf || (size_t)tex_tiles_.height() < std::max<size_t>(1, num_tex_list_))
            tex_tiles_ = PitchedArray(tf, (int)std::max<size_t>(1, num_tex_list_));
        check(tp_textures_accumulate(&g_, nullptr, p_clt_.get(), num_tex_list_, tex_idx_.get(), ports_.get(),
                                     p.colors, p.is_lwir, p.min_shot, p.scale_shot, p.diff_sigma,
                                     p.diff_threshold, p.min_agree, p.weight0, p.weight1, p.weight2,
                                     p.dust_remove, p.keep_weights, 0, nullptr, tex_tiles_.stride(),
                                     tex_tiles_.get(), st_),
              "tp_textures_accumulate");
    }

    void prepare_texture_list() {
        check(tp_prepare_texture_list(&g_, tasks_.get(), (int)tasks_host_.size(), tex_idx_.get(), num_tex_.get(),
                                      woi_.get(), d_.TX, d_.TY, st_),
              "tp_prepare_texture_list");
    }

    // generate_RBGA into the RGBA image of (tilesx + 1) 8 columns and
    // (tilesy + 1) 8 rows per slice (the reference passes keep_weights = 0)
    void generate_rbga(const TextureParams &p) {
        const int slices = p.colors + 1;
        const int w = (d_.TX + 1) * TP_DTT_SIZE, h = (d_.TY + 1) * TP_DTT_SIZE * slices;
        if (rbga_.width() != w || rbga_.height() != h) rbga_ = PitchedArray(w, h);
        check(tp_generate_RBGA(&g_, tasks_.get(), (int)tasks_host_.size(), tex_idx_.get(), num_tex_.get(), woi_.get(),
                               d_.TX, d_.TY, p_clt_.get(), ports_.get(), p.colors, p.is_lwir, p.min_shot,
                               p.scale_shot, p.diff_sigma, p.diff_threshold, p.min_agree, p.weight0, p.weight1,
                               p.weight2, p.dust_remove, 0, rbga_.stride(), rbga_.get(), st_),
              "tp_generate_RBGA");
    }

    // ---- results (synchronise the stream first)

    void sync() const { hip_check(hipStreamSynchronize(st_), "hipStreamSynchronize"); }
    std::vector<float> clt(int cam) const { return clt_.at(cam).download(d_.clt_floats()); }
    std::vector<float> rbg(int cam) const { return rbg_.at(cam).download(); }
    std::vector<float> corrs() const { return corrs_.download((int)num_corr_); }
    std::vector<float> corrs_fused() const { return fused_.download((int)(TP_NUM_PAIRS * tasks_host_.size())); }
    std::vector<float> texture_tiles() const { return tex_tiles_.download((int)num_tex_list_); }
    std::vector<float> rbga() const { return rbga_.download(); }
    std::array<int32_t, 4> woi() const {
        const std::vector<int32_t> v = woi_.download(4);
        return {v[0], v[1], v[2], v[3]};
    }
    std::array<int32_t, 8> num_texture_tiles() const {
        const std::vector<int32_t> v = num_tex_.download(8);
        std::array<int32_t, 8> a{};
        for (int i = 0; i < 8; i++) a[i] = v[i];
        return a;
    }
    size_t num_corr() const { return num_corr_; }
    size_t num_texture_list() const { return num_tex_list_; }
    int corr_radius() const { return radius_; }

  private:
    static DeviceArray<float *> pointer_array(const std::array<float *, TP_NUM_CAMS> &p) {
        DeviceArray<float *> a(TP_NUM_CAMS);
        a.upload(p.data(), TP_NUM_CAMS);
        return a;
    }

    tp_geometry g_;
    Dims d_;
    hipStream_t st_;
    int radius_;
    std::array<DeviceArray<float>, TP_NUM_CAMS> kernels_, offsets_, clt_;
    std::array<PitchedArray, TP_NUM_CAMS> images_, rbg_;
    DeviceArray<float *> p_kernels_, p_offsets_, p_images_, p_clt_, p_rbg_;
    DeviceArray<float> ports_;
    DeviceArray<int32_t> woi_, num_tex_, corr_idx_, tex_idx_;
    DeviceArray<tp_task> tasks_;
    std::vector<tp_task> tasks_host_;
    size_t num_corr_ = 0, num_tex_list_ = 0;
    PitchedArray corrs_, fused_, tex_tiles_, rbga_;
};

}  // namespace tph
