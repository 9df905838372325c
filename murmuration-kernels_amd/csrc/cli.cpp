// flock_hip: the drop-in command line for the reference step binary
// (cuda/flock.cu:374-523): same flags, same .flk files, same report rows, so a
// parity ladder like scripts/parity.sh can drive it unchanged.
//
//   flock_hip --in FILE [--steps K] [--warmup W] [--out FILE]
//             [--canonical] [--unfused] [--graph] [--cell M] [--block B] [--bpt K]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/flock_hip.h"

static int usage() {
    std::fprintf(stderr,
                 "usage: flock_hip --in FILE [--steps K] [--warmup W] [--out FILE]\n"
                 "       [--canonical] [--unfused] [--graph] [--resident] [--occ] [--spatial]\n"
                 "       [--persistent] [--xcd] [--window] [--flight] [--lazy]\n"
                 "       [--cell M] [--block B] [--bpt K] [--group G] [--shift S]\n"
                 "(the first W steps are not timed)\n");
    return 2;
}

static int fail(const char *what) {
    std::fprintf(stderr, "flock_hip: %s: %s\n", what, fh_last_error());
    return 1;
}

int main(int argc, char **argv) {
    const char *in = nullptr, *out = nullptr;
    long steps = 1, warmup = 0;
    fh_params prm;
    fh_params_default(&prm);
    for (int k = 1; k < argc; ++k) {
        std::string a = argv[k];
        auto val = [&]() -> const char * {
            if (k + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", a.c_str());
                std::exit(2);
            }
            return argv[++k];
        };
        if (a == "--in") in = val();
        else if (a == "--out") out = val();
        else if (a == "--steps") steps = std::atol(val());
        else if (a == "--warmup") warmup = std::atol(val());
        else if (a == "--canonical") prm.flags |= FH_FLAG_CANONICAL;
        else if (a == "--unfused") prm.flags |= FH_FLAG_UNFUSED;
        else if (a == "--graph") prm.flags |= FH_FLAG_GRAPH;
        else if (a == "--cell") prm.cell = (float)std::atof(val());
        else if (a == "--block") prm.decide_block = (uint32_t)std::atol(val());
        else if (a == "--bpt") prm.birds_per_thread = (uint32_t)std::atol(val());
        else if (a == "--group") prm.decide_group = (uint32_t)std::atol(val());
        else if (a == "--shift") prm.bucket_shift = (uint32_t)std::atol(val());
        else if (a == "--occ") prm.flags |= FH_FLAG_OCC;
        else if (a == "--resident") prm.flags |= FH_FLAG_RESIDENT;
        else if (a == "--spatial") prm.flags |= FH_FLAG_SPATIAL;
        else if (a == "--persistent") prm.flags |= FH_FLAG_PERSISTENT;
        else if (a == "--xcd") prm.flags |= FH_FLAG_XCD;
        else if (a == "--window") prm.flags |= FH_FLAG_WINDOW;
        else if (a == "--flight") prm.flags |= FH_FLAG_FLIGHT;
        else if (a == "--lazy") prm.flags |= FH_FLAG_LAZY;
        else return usage();
    }
    if (!in) {
        std::fprintf(stderr, "missing --in\n");
        return 2;
    }
    if (steps < 0 || warmup < 0) return usage();
    if (warmup > steps) warmup = steps;

    uint32_t n = 0;
    float time = 0.0f;
    if (fh_read_flk(in, &n, &time, nullptr, 0)) return fail(in);
    std::vector<float> words((size_t)n * FH_WORDS);
    if (fh_read_flk(in, &n, &time, words.data(), n)) return fail(in);

    char dev[256] = {0};
    int cus = 0;
    size_t hbm = 0;
    if (fh_device_info(dev, sizeof dev, &cus, &hbm)) return fail("device");
    std::printf("# flock_hip: %s, %d CUs, %.0f GiB HBM, %s\n", dev, cus, hbm / 1073741824.0, fh_version());
    std::printf("# %u birds, %ld steps (%ld warm-up, not timed)\n", n, steps, warmup);

    fh_flock *h = nullptr;
    if (fh_create(n, &prm, &h)) return fail("fh_create");
    int vgprs = 0, lds = 0, scratch = 0;
    if (fh_kernel_info("decide", &vgprs, &lds, &scratch) == FH_OK)
        std::printf("# decide: %d VGPRs, %d B static LDS, %d B scratch, block %u x %u birds\n", vgprs, lds,
                    scratch, prm.decide_block, prm.birds_per_thread);
    if (fh_upload_words(h, words.data(), n, time)) return fail("upload");
    fh_pass_ms ms;
    if (fh_step(h, (uint32_t)steps, (uint32_t)warmup, &ms)) return fail("fh_step");
    const long timed = steps - warmup;
    if (timed > 0) {
        std::printf("# %ld timed steps (mean GPU ms per pass, HIP events)\n", timed);
        const char *names[5] = {"grid_histogram", "grid_scan", "grid_scatter", "decide", "integrate"};
        const double vals[5] = {ms.grid_histogram, ms.grid_scan, ms.grid_scatter, ms.decide, ms.integrate};
        for (int p = 0; p < 5; ++p) std::printf("  %-16s %9.4f\n", names[p], vals[p]);
        std::printf("  %-16s %9.4f\n  bird-steps/s     %9.0f\n", "step total", ms.step_total,
                    ms.step_total > 0 ? n / (ms.step_total / 1000.0) : 0.0);
    }
    if (out) {
        if (fh_download_words(h, words.data(), &time)) return fail("download");
        if (fh_write_flk(out, n, time, words.data())) return fail(out);
    }
    fh_destroy(h);
    return 0;
}
