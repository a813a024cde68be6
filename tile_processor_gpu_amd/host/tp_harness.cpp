// tp_harness.cpp - the reference's C++ test harness (src/dtt8x8.cu main,
// :215-1050) on libtp.so: reads the harness's input files, builds the task,
// correlation and texture lists on the host, runs the path stage by stage
// with `warmup` untimed runs (the reference: one) and `iterations` timed ones
// (HIP events on the frame's stream), and writes the harness's result files.
//
// Inputs (prefix P, as P = .../clt/main in the reference):
//   P_chn{c}_transposed.kernel          [KV][KH][3][4][8][8] floats
//   P_chn{c}_transposed.kernel_offsets  [KV][KH][3] CltExtra (8 floats)
//   P_chn{c}.bayer                      [H][W] floats
//   P_chn{c}.portsxy                    [tilesy * tilesx][2] tile centres
// Results (in the output directory):
//   main_chn{c}.clt          with --save-clt: the forward CLT
//   main_chn{c}.rbg          imclt planes, 3 x (H + 8) rows of W + 8
//   main_corr.corr           (2r+1)^2 floats per correlation-list entry
//   main_corr_fused.corr     with --fused: the same from tp_convert_correlate
//                            (rows 6 tile + pair: the list's order here)
//   texture.rgba             textures_accumulate tiles, one per list entry
//   texture_rgba.rgba        generate_RBGA image, (tilesx + 1) 8 columns,
//                            (tilesy + 1) 8 rows per slice, colors + 1 slices
//   lists.txt                with --lists-only: the host-built lists, and
//                            nothing runs on the GPU
// The reference writes the first (woi_w + 1) 8 x (woi_h + 1) 8 x slices floats
// of its pitched RGBA copy (:1019-1034); here the whole image is written.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "tp_host.hpp"

namespace {

struct Options {
    std::string prefix, out = ".";
    tp_geometry g = {2592, 1936, 164, 123};
    int iterations = 10;
    int warmup = 1;
    int radius = 7;
    float fat_zero = 30.0f;  // the harness's value for raw pixel data (src/dtt8x8.cu:725)
    int colors = 3;
    int keep_weights = 1;    // the harness's textures_accumulate run (:274)
    bool save_clt = false, lists_only = false, all_imclt = false, fused = false;
};

void usage(const char *prog) {
    std::printf(
        "usage: %s --prefix P [--out DIR] [--width W --height H --kernels-hor KH --kernels-vert KV]\n"
        "          [--iterations N] [--warmup N] [--radius R] [--fat-zero F] [--colors 3|1]\n"
        "          [--keep-weights 0|1]\n"
        "          [--save-clt] [--imclt-all] [--fused] [--lists-only]\n"
        "Runs convert_correct_tiles, imclt_rbg (48 launches; --imclt-all: one), correlate2D,\n"
        "textures_accumulate, prepare_texture_list and generate_RBGA of libtp.so on the\n"
        "reference harness's input files and writes its result files.\n",
        prog);
}

bool parse(int argc, char **argv, Options &o) {
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto next = [&](const char *name) -> const char * {
            if (i + 1 >= argc) throw tph::Error(std::string("missing value of ") + name);
            return argv[++i];
        };
        if (a == "--help" || a == "-h") return false;
        else if (a == "--prefix") o.prefix = next("--prefix");
        else if (a == "--out") o.out = next("--out");
        else if (a == "--width") o.g.img_width = std::atoi(next("--width"));
        else if (a == "--height") o.g.img_height = std::atoi(next("--height"));
        else if (a == "--kernels-hor") o.g.kernels_hor = std::atoi(next("--kernels-hor"));
        else if (a == "--kernels-vert") o.g.kernels_vert = std::atoi(next("--kernels-vert"));
        else if (a == "--iterations") o.iterations = std::atoi(next("--iterations"));
        else if (a == "--warmup") o.warmup = std::atoi(next("--warmup"));
        else if (a == "--radius") o.radius = std::atoi(next("--radius"));
        else if (a == "--fat-zero") o.fat_zero = (float)std::atof(next("--fat-zero"));
        else if (a == "--colors") o.colors = std::atoi(next("--colors"));
        else if (a == "--keep-weights") o.keep_weights = std::atoi(next("--keep-weights"));
        else if (a == "--save-clt") o.save_clt = true;
        else if (a == "--imclt-all") o.all_imclt = true;
        else if (a == "--fused") o.fused = true;
        else if (a == "--lists-only") o.lists_only = true;
        else throw tph::Error("unknown argument " + a);
    }
    if (o.prefix.empty()) throw tph::Error("--prefix is required");
    if (o.iterations < 1) throw tph::Error("--iterations must be positive");
    if (o.warmup < 0) throw tph::Error("--warmup must not be negative");
    return true;
}

std::string chn(const std::string &p, int c, const char *suffix) {
    return p + "_chn" + std::to_string(c) + suffix;
}

// One stage: `warm` untimed runs, then `n` timed runs bracketed by events.
template <class F>
float timed(const char *name, int warm, int n, hipStream_t st, F &&run) {
    for (int i = 0; i < warm; i++) run();
    tph::hip_check(hipStreamSynchronize(st), name);
    hipEvent_t a, b;
    tph::hip_check(hipEventCreate(&a), "hipEventCreate");
    tph::hip_check(hipEventCreate(&b), "hipEventCreate");
    tph::hip_check(hipEventRecord(a, st), "hipEventRecord");
    for (int i = 0; i < n; i++) run();
    tph::hip_check(hipEventRecord(b, st), "hipEventRecord");
    tph::hip_check(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0.f;
    tph::hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    ms /= (float)n;
    std::printf("Average %s run time = %f ms\n", name, ms);
    return ms;
}

int run(const Options &o) {
    const tph::Dims d(o.g);
    std::array<std::vector<float>, TP_NUM_CAMS> xy;
    for (int c = 0; c < TP_NUM_CAMS; c++) xy[c] = tph::read_floats(chn(o.prefix, c, ".portsxy"), 2 * d.tiles());
    const std::vector<tp_task> tasks = tph::linescan_tasks(o.g, xy);
    if (o.lists_only) {
        const std::vector<int32_t> ci = tph::corr_list(o.g, tasks), ti = tph::texture_list(o.g, tasks);
        const std::string path = o.out + "/lists.txt";
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) throw tph::Error("cannot create " + path);
        std::fprintf(f, "tasks %zu\n", tasks.size());
        for (const tp_task &t : tasks) std::fprintf(f, "%d %d\n", t.task, t.txy);
        std::fprintf(f, "corr %zu\n", ci.size());
        for (int v : ci) std::fprintf(f, "%d\n", v);
        std::fprintf(f, "texture %zu\n", ti.size());
        for (int v : ti) std::fprintf(f, "%d\n", v);
        std::fclose(f);
        std::printf("%zu tasks, %zu correlations, %zu texture tiles -> %s\n", tasks.size(), ci.size(), ti.size(),
                    path.c_str());
        return 0;
    }

    std::printf("%s\n", tp_version());
    hipStream_t st;
    tph::hip_check(hipStreamCreate(&st), "hipStreamCreate");
    {
        tph::Frame fr(o.g, st, o.radius);
        for (int c = 0; c < TP_NUM_CAMS; c++) {
            fr.set_kernels(c, tph::read_floats(chn(o.prefix, c, "_transposed.kernel"), d.kernel_floats()),
                           tph::read_floats(chn(o.prefix, c, "_transposed.kernel_offsets"), d.offset_floats()));
            fr.set_image(c, tph::read_floats(chn(o.prefix, c, ".bayer"), (size_t)d.W * d.H));
        }
        fr.set_tasks(tasks);
        std::printf("%zu tiles, %zu correlations, %zu texture tiles\n", tasks.size(), fr.num_corr(),
                    fr.num_texture_list());

        timed("convert_correct_tiles", o.warmup, o.iterations, st, [&] { fr.convert(); });
        fr.sync();
        if (o.save_clt)
            for (int c = 0; c < TP_NUM_CAMS; c++) tph::write_floats(o.out + "/main_chn" + std::to_string(c) + ".clt", fr.clt(c));

        if (o.all_imclt)
            timed("imclt_rbg_all", o.warmup, o.iterations, st, [&] { fr.imclt_all(); });
        else
            timed("imclt_rbg (48 launches)", o.warmup, o.iterations, st, [&] { fr.imclt_passes(); });
        // the result files come from one more run on cleared planes (repeated
        // 48-launch runs add up in the margin rows and columns)
        fr.clear_rbg();
        if (o.all_imclt) fr.imclt_all();
        else fr.imclt_passes();
        fr.sync();
        for (int c = 0; c < TP_NUM_CAMS; c++) tph::write_floats(o.out + "/main_chn" + std::to_string(c) + ".rbg", fr.rbg(c));

        timed("correlate2D", o.warmup, o.iterations, st, [&] { fr.correlate(3, 0.25f, 0.25f, 0.5f, o.fat_zero); });
        fr.sync();
        tph::write_floats(o.out + "/main_corr.corr", fr.corrs());
        if (o.fused) {
            timed("convert_correlate (fused)", o.warmup, o.iterations, st,
                  [&] { fr.convert_correlate(3, 0.25f, 0.25f, 0.5f, o.fat_zero); });
            fr.sync();
            tph::write_floats(o.out + "/main_corr_fused.corr", fr.corrs_fused());
        }

        tph::TextureParams tp;
        tp.colors = o.colors;
        tp.is_lwir = o.colors == 1;
        tp.keep_weights = o.keep_weights;
        timed("textures_accumulate", o.warmup, o.iterations, st, [&] { fr.textures(tp); });
        fr.sync();
        tph::write_floats(o.out + "/texture.rgba", fr.texture_tiles());

        timed("prepare_texture_list", o.warmup, o.iterations, st, [&] { fr.prepare_texture_list(); });
        fr.sync();
        std::array<int32_t, 4> woi = fr.woi();
        std::array<int32_t, 8> n = fr.num_texture_tiles();
        std::printf("WOI x=%d, y=%d, width=%d, height=%d\n", woi[0], woi[1], woi[2], woi[3]);
        std::printf("num_texture_tiles=(%d(%d), %d(%d), %d(%d), %d(%d))\n", n[0], n[1], n[2], n[3], n[4], n[5],
                    n[6], n[7]);

        timed("generate_RBGA", o.warmup, o.iterations, st, [&] { fr.generate_rbga(tp); });
        fr.sync();
        woi = fr.woi();
        std::printf("WOI x=%d, y=%d, width=%d, height=%d\n", woi[0], woi[1], woi[2], woi[3]);
        tph::write_floats(o.out + "/texture_rgba.rgba", fr.rbga());
    }
    tph::hip_check(hipStreamDestroy(st), "hipStreamDestroy");
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    Options o;
    try {
        if (!parse(argc, argv, o)) {
            usage(argv[0]);
            return 0;
        }
        return run(o);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "tp_harness: %s\n", e.what());
        return 1;
    }
}
