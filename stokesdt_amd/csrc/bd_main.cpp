// Command-line BD driver over the C ABI: the options and progress output of
// the reference's bin/main/main.cc:12-158 (--config, --model, --xyz,
// -n/--nsteps, --log), running the simulation box in blocks of 100 steps,
// plus --restart FILE (write a restart frame at the end).
#include <getopt.h>
#include <sys/time.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "../../include/stokesdt_amd.h"

static void usage(const char* call) {
    fprintf(stderr, "Usage: %s [OPTIONS]\n", call);
    fprintf(stderr, "Options:\n");
    fprintf(stderr, "\t-h or --help         Display this information\n");
    fprintf(stderr, "\t-v or --version      Display version information\n");
    fprintf(stderr, "\t--config             Config file\n");
    fprintf(stderr, "\t--model              Model file\n");
    fprintf(stderr, "\t--xyz                XYZ file\n");
    fprintf(stderr, "\t--log                log file (accepted, not used)\n");
    fprintf(stderr, "\t--restart            write a restart XYZ frame at the end\n");
    fprintf(stderr, "\t-n or --nsteps       Number of steps\n");
}

static double now() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

int main(int argc, char** argv) {
    static const struct option opts[] = {
        {"help", 0, nullptr, 'h'},    {"version", 0, nullptr, 'v'}, {"nsteps", 1, nullptr, 'n'},
        {"config", 1, nullptr, 10},   {"model", 1, nullptr, 20},    {"xyz", 1, nullptr, 30},
        {"log", 1, nullptr, 40},      {"restart", 1, nullptr, 50},  {nullptr, 0, nullptr, 0}};
    const char* config = nullptr;
    const char* model = nullptr;
    const char* xyz = nullptr;
    const char* restart = nullptr;
    int nsteps = 1;
    int c;
    while ((c = getopt_long(argc, argv, ":hvn:", opts, nullptr)) != -1) {
        switch (c) {
            case 'h': usage(argv[0]); return 0;
            case 'v': printf("%s version %s\n", argv[0], sdt_version()); return 0;
            case 'n':
                nsteps = atoi(optarg);
                if (nsteps <= 0) {
                    fprintf(stderr, "Invalid argument --nsteps %d\n", nsteps);
                    return -1;
                }
                break;
            case 10: config = optarg; break;
            case 20: model = optarg; break;
            case 30: xyz = optarg; break;
            case 40: break;
            case 50: restart = optarg; break;
            case ':': fprintf(stderr, "Option -%c requires an argument.\n", optopt); return -1;
            default: usage(argv[0]); return -1;
        }
    }
    if (!config || !model || !xyz) {
        usage(argv[0]);
        return -1;
    }
    sdt_simbox_t sb = nullptr;
    if (sdt_simbox_create(config, model, xyz, &sb) != SDT_OK || sdt_simbox_init(sb) != SDT_OK) {
        fprintf(stderr, "Failed to initialize the simulation box: %s\n", sdt_last_error());
        sdt_simbox_destroy(sb);
        return -1;
    }
    printf("Start BD simulation\n-------------------\n");
    const int blk = 100;
    double total = 0.0;
    for (int i = 0; i < nsteps; i += blk) {
        const int n = (i + blk >= nsteps) ? nsteps - i : blk;
        const double t0 = now();
        if (sdt_simbox_advance(sb, n) != SDT_OK) {
            fprintf(stderr, "\nBD step failed: %s\n", sdt_last_error());
            sdt_simbox_destroy(sb);
            return -1;
        }
        total += now() - t0;
        printf("finished steps: %10d (%7.3f%%),    elapsed time: %10.2f secs\r", i + n,
               100.0 * (i + n) / nsteps, total);
        fflush(stdout);
    }
    printf("\n\nBD simulation finished at %.3g secs, %.3g secs/step\n", total, total / nsteps);
    int rc = 0;
    if (restart && sdt_simbox_write_xyz(sb, restart) != SDT_OK) {
        fprintf(stderr, "Failed to write %s: %s\n", restart, sdt_last_error());
        rc = -1;
    }
    sdt_simbox_destroy(sb);
    return rc;
}
