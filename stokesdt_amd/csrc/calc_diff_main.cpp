// calc_diff: diffusion coefficients from an XYZ trajectory, the command-line
// tool of tools/calc_diff/calc_diff.cc with its input conventions and output
// lines; the sums run on the GPU through sdt_diffusion.
//
// Input (calc_diff.cc:11-98): lines holding only white space are skipped; the
// first remaining line holds the particle count np; every frame is np + 2
// lines (count, comment "<frame> <time> ...", np lines "<type> x y z"); the
// frame interval is the second field of the second frame's comment minus
// that of the first. Every particle counts with radius 1 (the tool reads no
// radii), so there is one radius column.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>
#include "../../include/stokesdt_amd.h"

static bool blank(const std::string& s) {
    for (char c : s)
        if (c != ' ' && c != '\t' && c != '\n' && c != '\r' && c != '\v' && c != '\f') return false;
    return true;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: calc_diff <trajectory.xyz>\n");
        return 2;
    }
    std::ifstream in(argv[1]);
    if (!in) {
        fprintf(stderr, "calc_diff: cannot open %s\n", argv[1]);
        return 1;
    }
    std::vector<std::string> lines;
    for (std::string l; std::getline(in, l);)
        if (!blank(l)) lines.push_back(l);
    int np = 0;
    if (lines.empty() || sscanf(lines[0].c_str(), "%d", &np) != 1 || np <= 0) {
        fprintf(stderr, "calc_diff: the first line must hold the particle count\n");
        return 1;
    }
    const size_t per = (size_t)np + 2;
    if (lines.size() % per != 0 || lines.size() / per <= 2) {
        fprintf(stderr, "calc_diff: %zu lines do not form more than two frames of %d particles\n",
                lines.size(), np);
        return 1;
    }
    const int nsteps = (int)(lines.size() / per);
    auto stamp = [&](size_t i) {
        char tag[1024];
        double t = 0.0;
        if (sscanf(lines[i].c_str(), "%1023s %lf", tag, &t) != 2) {
            fprintf(stderr, "calc_diff: no time in comment line %zu\n", i + 1);
            exit(1);
        }
        return t;
    };
    const double tslen = stamp(per + 1) - stamp(1);
    printf("nsteps = %d, step-length = %.3le\n", nsteps, tslen);
    std::vector<double> pos((size_t)nsteps * 3 * np);
    for (int f = 0; f < nsteps; f++)
        for (int k = 0; k < np; k++) {
            char tag[1024];
            double* p = &pos[((size_t)f * np + k) * 3];
            if (sscanf(lines[f * per + 2 + k].c_str(), "%1023s %lf %lf %lf", tag, p, p + 1, p + 2) !=
                4) {
                fprintf(stderr, "calc_diff: bad particle line in frame %d\n", f);
                return 1;
            }
        }
    printf("Have %d different radii\n", 1);
    printf("\ntau, r = %.3f\n", 1.0);
    std::vector<double> D(nsteps / 2 + 1);
    int nlags = 0;
    if (sdt_diffusion(pos.data(), nsteps, np, tslen, D.data(), &nlags) != SDT_OK) {
        fprintf(stderr, "calc_diff: %s\n", sdt_last_error());
        return 1;
    }
    for (int i = 1; i <= nlags; i++)
        printf("%.3f, %.6f, %.6f\n", i * tslen, D[i - 1], D[i - 1] * i * tslen);
    return 0;
}
