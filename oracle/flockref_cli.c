/*
 * flock-ref: the oracle's command line (test infrastructure only).
 * Restates crates/flock-ref/src/main.rs:1-208: the same sub-commands, flags,
 * output lines and exit codes (0 pass, 1 fail, 2 usage).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "flockref.h"

static const char *USAGE =
    "usage:\n"
    "  flock-ref gen   --birds N [--seed S] --out FILE\n"
    "  flock-ref step  --in FILE --steps K --out FILE [--threads T]   (brute force)\n"
    "  flock-ref diff  A.flk B.flk [--pos-tol X] [--vel-tol X] [--allow-outliers N] [--outlier-max X]\n"
    "  flock-ref stats FILE\n"
    "  flock-ref statdiff A.flk B.flk     (aggregate agreement for long, chaotic runs)\n";

static int usage(const char *why) {
    fprintf(stderr, "flock-ref: %s\n\n%s", why, USAGE);
    return 2;
}

static const char *opt(int argc, char **argv, const char *name) {
    for (int k = 1; k + 1 < argc; ++k)
        if (strcmp(argv[k], name) == 0) return argv[k + 1];
    return NULL;
}

typedef struct {
    uint32_t n;
    float time;
    float *w;
} flock_t;

static int load(const char *path, flock_t *f) {
    int rc = fr_read_flk(path, &f->n, &f->time, NULL, 0);
    if (rc == 0) {
        f->w = (float *)malloc((size_t)(f->n ? f->n : 1) * FR_WORDS * 4);
        rc = fr_read_flk(path, &f->n, &f->time, f->w, f->n);
    }
    if (rc != 0) {
        static const char *why[] = {"", "cannot open", "not a FLK1 file", "unexpected words per bird",
                                    "truncated"};
        fprintf(stderr, "flock-ref: %s: %s\n", path, why[rc]);
    }
    return rc;
}

int main(int argc, char **argv) {
    if (argc < 2) return usage("missing or unknown command");
    const char *cmd = argv[1];
    if (strcmp(cmd, "gen") == 0) {
        const char *nb = opt(argc, argv, "--birds"), *out = opt(argc, argv, "--out");
        const char *sd = opt(argc, argv, "--seed");
        if (!nb) return usage("missing --birds");
        if (!out) return usage("missing --out");
        uint32_t n = (uint32_t)strtoul(nb, NULL, 10);
        uint64_t seed = sd ? strtoull(sd, NULL, 10) : 1;
        float *w = (float *)malloc((size_t)(n ? n : 1) * FR_WORDS * 4);
        fr_scatter(n, seed, w);
        int rc = fr_write_flk(out, n, 0.0f, w);
        free(w);
        return rc ? 1 : 0;
    }
    if (strcmp(cmd, "step") == 0) {
        const char *in = opt(argc, argv, "--in"), *out = opt(argc, argv, "--out");
        const char *ks = opt(argc, argv, "--steps"), *ts = opt(argc, argv, "--threads");
        if (!in) return usage("missing --in");
        if (!ks) return usage("missing --steps");
        if (!out) return usage("missing --out");
        flock_t f;
        if (load(in, &f)) return 2;
        fr_step(f.w, f.n, &f.time, (uint32_t)strtoul(ks, NULL, 10), ts ? atoi(ts) : 1);
        return fr_write_flk(out, f.n, f.time, f.w) ? 1 : 0;
    }
    if (strcmp(cmd, "stats") == 0) {
        if (argc < 3) return usage("missing FILE");
        flock_t f;
        if (load(argv[2], &f)) return 2;
        float s[4];
        fr_stats(f.w, f.n, s);
        printf("birds %u  time %.3f s\n", f.n, f.time);
        printf("mean speed %.4f m/s  mean neighbours %.3f  mean radius %.3f m  polarization %.4f\n",
               s[0], s[1], s[2], s[3]);
        return 0;
    }
    if (strcmp(cmd, "diff") == 0 || strcmp(cmd, "statdiff") == 0) {
        if (argc < 4) return usage("missing A or B");
        flock_t a, b;
        if (load(argv[2], &a) || load(argv[3], &b)) return 2;
        if (a.n != b.n) return usage("files hold different bird counts");
        if (strcmp(cmd, "statdiff") == 0) {
            static const char *names[4] = {"mean speed", "mean neighbours", "mean radius", "polarization"};
            static const char *tols[4] = {"0.02", "0.1", "0.01", "0.005"};
            float s[8];
            int ok = fr_statdiff(a.w, b.w, a.n, s);
            for (int k = 0; k < 4; ++k) {
                float d = s[k] - s[4 + k];
                if (d < 0) d = -d;
                printf("%-16s A %9.4f  B %9.4f  |d| %.4f  (tol %s)\n", names[k], s[k], s[4 + k], d, tols[k]);
            }
            printf("%s\n", ok ? "PASS" : "FAIL");
            return ok ? 0 : 1;
        }
        const char *pt = opt(argc, argv, "--pos-tol"), *vt = opt(argc, argv, "--vel-tol");
        const char *al = opt(argc, argv, "--allow-outliers"), *om = opt(argc, argv, "--outlier-max");
        float pos_tol = pt ? strtof(pt, NULL) : 1e-3f, vel_tol = vt ? strtof(vt, NULL) : 1e-3f;
        uint32_t allow = al ? (uint32_t)strtoul(al, NULL, 10) : 0;
        float omax = om ? strtof(om, NULL) : 1e-2f;
        float d[7];
        uint32_t c[2];
        int ok = fr_diff(a.w, b.w, a.n, pos_tol, vel_tol, allow, omax, d, c);
        printf("birds %u  time A %.4f B %.4f\n", a.n, a.time, b.time);
        printf("max |dpos| %.3e  max |dvel| %.3e  max |dradius| %.3e  max |dcentrality| %.3e\n", d[0], d[1],
               d[2], d[3]);
        printf("neighbour-count mismatches %u of %u\n", c[0], a.n);
        printf("(informational) max |dsteering| %.3e  max |droll| %.3e  max |dspin| %.3e\n", d[4], d[5], d[6]);
        if (allow > 0) printf("birds above tolerance: %u (allowed %u, each below %g)\n", c[1], allow, omax);
        printf("%s\n", ok ? "PASS" : "FAIL");
        return ok ? 0 : 1;
    }
    return usage("missing or unknown command");
}
