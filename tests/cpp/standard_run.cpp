/*
 * standard_run.cpp -- a BSim user program with the topology of the
 * reference's benchmark program (test/gpu/standard_test.cpp:55-138), written
 * against the C++ API of include/BSim.h:
 *
 *   g[1..depth]: LIF_brian populations of n neurons (i_offset 2 nA on g[1]);
 *   for i = 2..depth: connect(g[i-1], g[i]) and, when i % who == 0,
 *                     connect(g[1], g[i]) -- a target fed by a later
 *                     population first and by an earlier one second;
 *   for i = 1, 5, 9, ... while i + 1 <= depth: a side population p[i],
 *                     connect(g[i], p[i]) and connect(p[i], g[i+1]) with
 *                     Inhibitory synapses.
 *
 * All projections are all-to-all with constant weights 1e-9 * w_k / n and a
 * 0.1 ms delay built with getConstArray, as the reference program does.
 *
 * usage: standard_run depth n w1 w2 w3 w4 who T outdir [gid ...]
 *   writes GSim.log (fired ids per cycle), g_v.data (V of the listed
 *   neurons) and GFire.count into outdir.
 */
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include <vector>

#include "BSim.h"

static LIF_brian lif(real i_offset)
{
    const real v_thresh = -54e-3f, v_reset = -60e-3f, cm = 0.2e-9f, v_rest = -74e-3f;
    const real tau_e = 2e-3f, tau_i = 0.6e-3f, tau_m = 10e-3f, v_init = -74e-3f;
    return LIF_brian(LIFENeuron(v_init, v_rest, v_reset, cm, tau_m, 0.0f, tau_e, tau_i, v_thresh,
                                i_offset),
                     tau_e, tau_i);
}

int main(int argc, char **argv)
{
    if (argc < 10) {
        fprintf(stderr, "usage: %s depth n w1 w2 w3 w4 who T outdir [gid ...]\n", argv[0]);
        return 2;
    }
    const int depth = atoi(argv[1]);
    const int n = atoi(argv[2]);
    const real w1 = (real)atof(argv[3]), w2 = (real)atof(argv[4]);
    const real w3 = (real)atof(argv[5]), w4 = (real)atof(argv[6]);
    const int who = atoi(argv[7]);
    const real T = (real)atof(argv[8]);
    std::vector<int> mon;
    for (int a = 10; a < argc; a++) mon.push_back(atoi(argv[a]));
    if (depth < 2 || n < 1 || who < 1) {
        fprintf(stderr, "bad arguments\n");
        return 2;
    }

    Network net;
    std::vector<Population<LIF_brian> *> g((size_t)depth + 1, NULL);
    g[1] = net.createPopulation(1, n, lif(2e-9f));
    for (int i = 2; i <= depth; i++) g[i] = net.createPopulation(i, n, lif(0.0f));

    const int nn = n * n;
    real *w_chain = getConstArray((real)(1e-9) * w1 / n, nn);
    real *w_back = getConstArray((real)(1e-9) * w2 / n, nn);
    real *w_side = getConstArray((real)(1e-9) * w3 / n, nn);
    real *w_inh = getConstArray((real)(-1e-9) * w4 / n, nn);
    real *delay = getConstArray((real)0.1e-3, nn);
    SpikeType *inh = getConstArray(Inhibitory, nn);

    for (int i = 2; i <= depth; i++) {
        net.connect(g[i - 1], g[i], w_chain, delay, NULL, nn);
        if (i % who == 0) net.connect(g[1], g[i], w_back, delay, NULL, nn);
    }
    for (int i = 1; i + 1 <= depth; i += 4) {
        Population<LIF_brian> *side = net.createPopulation(i + depth, n, lif(0.0f));
        net.connect(g[i], side, w_side, delay, NULL, nn);
        net.connect(side, g[i + 1], w_inh, delay, inh, nn);
    }
    delete[] w_chain;
    delete[] w_back;
    delete[] w_side;
    delete[] w_inh;
    delete[] delay;
    delete[] inh;

    if (chdir(argv[9]) != 0) {
        perror(argv[9]);
        return 2;
    }
    SGSim sg(&net, 1e-4f);
    sg.setLog(true);
    sg.monitor(mon);
    try {
        if (sg.run(T) != 0) return 1;
    } catch (const bsim_error &e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    printf("standard_run ok\n");
    return 0;
}
