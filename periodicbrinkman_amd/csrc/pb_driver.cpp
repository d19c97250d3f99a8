// pb_driver -- the reference's driver program (periodicBrinkman.cu:73-219) on
// this library: reads the parameter file (argv[1], default param.txt), puts
// boxSize^3 targets uniformly at random in the box (this library's generator,
// seed argv[2] or 0x5EED0000) and one source at (0.5, 0.5, 0.5), sweeps the
// Ewald split over d_i = (0.1 + 0.95 i) d, i < 8, against d, and writes the
// maximum matrix differences to log.txt in the reference's format.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/periodicbrinkman.h"

int main(int argc, char **argv)
{
    const char *file = argc > 1 ? argv[1] : "param.txt";
    const unsigned long long seed = argc > 2 ? std::strtoull(argv[2], nullptr, 0) : 0x5EED0000ULL;
    pb_driver_params p;
    if (pb_read_param_file(file, &p) != PB_OK) std::printf("Cannot open file!\n"); // defaults, as the reference
    const int n = p.boxSize > 0 ? p.boxSize : 1;
    const int nt = n * n * n;
    const double L[3] = {p.Lx, p.Ly, p.Lz};
    std::vector<double> x(3 * (size_t)nt), c(nt);
    unsigned long long st[4];
    pb_rng_seed(st, seed);
    for (int a = 0; a < 3; a++) {
        pb_rng_uniform(st, nt, c.data(), L[a]);
        for (int t = 0; t < nt; t++) x[3 * (size_t)t + a] = c[t];
    }
    std::printf("optimal d is %g\n", p.d);
    double dval[8], err[8];
    int rc = pb_split_sweep(&p, x.data(), nt, dval, err);
    if (rc != PB_OK) {
        std::fprintf(stderr, "pb_split_sweep: %s\n", pb_status_string(rc));
        return 1;
    }
    FILE *log = std::fopen("log.txt", "w");
    if (!log) {
        std::fprintf(stderr, "cannot write log.txt\n");
        return 1;
    }
    std::fprintf(log, "For Lx = %g, Ly = %g, Lz = %g, and a = %g, we have: \n", p.Lx, p.Ly, p.Lz, p.alpha);
    for (int i = 0; i < 8; i++) std::fprintf(log, "%.10e : %d : %.10e\n", dval[i], 0, err[i]);
    std::fclose(log);
    std::printf("Done!!!!!!!!!!!!!!\nValues of parameters are \nLx = %g, Ly = %g, Lz = %g\n"
                "d = %g, e = %g, alpha = %g\nM = %d, boxSize = %d, toPrint = %d\n",
                p.Lx, p.Ly, p.Lz, p.d, p.e, p.alpha, p.M, n, p.toPrint);
    return 0;
}
