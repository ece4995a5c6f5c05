/*
 * dropin_run.cpp -- a BSim user program written against the reference's C++
 * API (include/BSim.h): two LIF_brian populations, an all-to-all projection
 * given as weight / delay arrays, SGSim(dt).run(T) with the raster and a V
 * trace logged.  The network's numbers come from a binary file so that a test
 * can run the same network through the CPU oracle:
 *
 *   int32 npop; per population: int32 n, float params[12] (LIFENeuron's 10
 *   then the composite's tau_syn_E, tau_syn_I), float i_offset[n];
 *   int32 nproj; per projection: int32 src, dst, float w[n_src * n_dst],
 *   float d[n_src * n_dst]; float dt, T; int32 nmon, int32 gid[nmon].
 *
 * usage: dropin_run <input> <output dir> [plain | logdata]
 *   default: run(T, log) with the FireInfo keys "count", "Y" and "X", written
 *            to GFire.count, input_Y.txt (hex floats) and input_X.txt;
 *   plain:   run(T), which writes GFire.count itself;
 *   logdata: setLogData(true) and run(T): the reference's -DLOG_DATA files
 *            (GSim.log, g_v.data, g_ie.data, g_ii.data, g_input_e.data,
 *            g_input_i.data for every LIF neuron).
 */
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "BSim.h"

template <class T>
static void rd(FILE *f, T *p, size_t n)
{
    if (fread(p, sizeof(T), n, f) != n) {
        fprintf(stderr, "short read\n");
        exit(2);
    }
}

int main(int argc, char **argv)
{
    if (argc != 3 && argc != 4) {
        fprintf(stderr, "usage: %s <input> <output dir> [plain | logdata]\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) {
        perror(argv[1]);
        return 2;
    }
    Network c;
    std::vector<Population<LIF_brian> *> pops;
    int32_t npop;
    rd(f, &npop, 1);
    for (int p = 0; p < npop; p++) {
        int32_t n;
        float q[12];
        rd(f, &n, 1);
        rd(f, q, 12);
        std::vector<float> ioff((size_t)n);
        rd(f, ioff.data(), (size_t)n);
        LIF_brian templ(LIFENeuron(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], 0.0f), q[10], q[11]);
        Population<LIF_brian> *pop = c.createPopulation(n, templ, true);
        for (int i = 0; i < n; i++)
            pop->addNeuron(LIF_brian(LIFENeuron(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8],
                                                ioff[i]), q[10], q[11]));
        pops.push_back(pop);
    }
    int32_t nproj;
    rd(f, &nproj, 1);
    for (int k = 0; k < nproj; k++) {
        int32_t s, d;
        rd(f, &s, 1);
        rd(f, &d, 1);
        const size_t m = (size_t)pops[s]->getNum() * (size_t)pops[d]->getNum();
        std::vector<real> w(m), dl(m);
        rd(f, w.data(), m);
        rd(f, dl.data(), m);
        c.connect(pops[s], pops[d], w.data(), dl.data(), NULL, (int)m);
    }
    float dt, T;
    rd(f, &dt, 1);
    rd(f, &T, 1);
    int32_t nmon;
    rd(f, &nmon, 1);
    std::vector<int> mon((size_t)nmon);
    rd(f, mon.data(), (size_t)nmon);
    /* optional: Poisson_spikes drive populations, created after the LIF
     * populations, each projecting all-to-all onto one of them */
    int32_t npp = 0;
    if (fread(&npp, sizeof(npp), 1, f) != 1) npp = 0;
    for (int k = 0; k < npp; k++) {
        int32_t n, tgt;
        float par[3];
        uint64_t seed;
        rd(f, &n, 1);
        rd(f, par, 3);
        rd(f, &seed, 1);
        rd(f, &tgt, 1);
        Poisson_spikes templ(PoissonNeuron(par[0], 0.0f, par[1], par[2]), 1e-3f, 1e-3f);
        templ.setSeed(seed);
        Population<Poisson_spikes> *pp = c.createPopulation(n, templ);
        const size_t m = (size_t)n * (size_t)pops[tgt]->getNum();
        std::vector<real> w(m), dl(m);
        rd(f, w.data(), m);
        rd(f, dl.data(), m);
        c.connect(pp, pops[tgt], w.data(), dl.data(), NULL, (int)m);
    }
    fclose(f);

    if (chdir(argv[2]) != 0) {
        perror(argv[2]);
        return 2;
    }
    SGSim sg(&c, dt);
    sg.setLog(true);
    sg.monitor(mon);
    if (argc > 3 && std::string(argv[3]) == "logdata") sg.setLogData(true);
    if (argc > 3) {   /* plain / logdata */
        try {
            if (sg.run(T) != 0) return 1;
        } catch (const bsim_error &e) {
            fprintf(stderr, "%s\n", e.what());
            return 1;
        }
        printf("dropin_run ok\n");
        return 0;
    }
    FireInfo log;
    log["count"];
    log["Y"];
    log["X"];
    try {
        if (sg.run(T, log) != 0) return 1;
    } catch (const bsim_error &e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    FILE *o = fopen("GFire.count", "w+");
    int *cnt = (int *)log["count"].data;
    for (int i = 0; i < log["count"].size; i++) fprintf(o, "%d \t", cnt[i]);
    fprintf(o, "\n");
    fclose(o);
    free(cnt);
    o = fopen("input_Y.txt", "w+");
    real *y = (real *)log["Y"].data;
    for (int i = 0; i < log["Y"].size; i++) fprintf(o, "%a\n", (double)y[i]);
    fclose(o);
    free(y);
    o = fopen("input_X.txt", "w+");
    int *x = (int *)log["X"].data;
    for (int i = 0; i < log["X"].size; i++) fprintf(o, "%d\n", x[i]);
    fclose(o);
    free(x);
    printf("dropin_run ok\n");
    return 0;
}
