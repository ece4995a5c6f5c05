/*
 * ref_harness.cpp -- drives the reference's own CPU simulator to produce
 * golden vectors (TEST INFRASTRUCTURE; built by oracle/build_ref.sh into
 * oracle/_ref/, never shipped with the product).
 *
 * It links BSim's CPU translation units straight from the reference tree
 * (Network, Population, LIFENeuron, StaticSynapse, ArrayNeuron, ...) and
 * repeats SingleThreadSimulator::run's loop (src/sim/SingleThreadSimulator.cpp:
 * 74-96): Network::reset once, then Network::update per cycle, reading the
 * per-cycle SimInfo that LIFENeuron::update fills (vm, i_syn_E, i_syn_I, i_E,
 * i_I per LIF neuron, LIFENeuron.cpp:144-155) and the fired flags.
 *
 * Usage: ref_stsim <input.bsnt> <output.bsout>
 *        ref_stsim --time <input.bsnt>   (times STSim::run, LOG_DATA off)
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/time.h>
#include <vector>

#include "net/Network.h"
#include "sim/SingleThreadSimulator.h"
#include "neuron/lif/LIFENeuron.h"
#include "neuron/array/ArrayNeuron.h"
#include "synapse/static/StaticSynapse.h"
#include "base/CompositeNeuron.h"
#include "third_party/json/json.h"

typedef CompositeNeuron<LIFENeuron, StaticSynapse> LIF_brian;       // include/Neuron.h:31
typedef CompositeNeuron<ArrayNeuron, StaticSynapse> Array_spikes;   // include/Neuron.h:27

static FILE *xopen(const char *p, const char *m)
{
    FILE *f = fopen(p, m);
    if (!f) {
        fprintf(stderr, "cannot open %s\n", p);
        exit(2);
    }
    return f;
}

template <typename T>
static void rd(FILE *f, T *p, size_t n)
{
    if (n && fread(p, sizeof(T), n, f) != n) {
        fprintf(stderr, "short read\n");
        exit(3);
    }
}

template <typename T>
static void wr(FILE *f, const T *p, size_t n)
{
    if (n) fwrite(p, sizeof(T), n, f);
}

struct Input {
    int32_t n = 0, steps = 0, n_rec = 0;
    int64_t s = 0;
    real dt = 0;
    std::vector<int32_t> rec;
    std::vector<int8_t> kind;
    std::vector<real> lif;
    std::vector<int64_t> arr_off;
    std::vector<int32_t> arr_t;
    std::vector<int32_t> pre, post;
    std::vector<real> w, d, tau;
    std::vector<int8_t> stype;
};

static Input load(const char *path)
{
    Input in;
    FILE *f = xopen(path, "rb");
    char magic[4];
    uint32_t ver, rs;
    rd(f, magic, 4);
    rd(f, &ver, 1);
    rd(f, &rs, 1);
    if (memcmp(magic, "BSNT", 4) != 0 || ver != 1 || rs != sizeof(real)) {
        fprintf(stderr, "bad input header (real size %u, built with %zu)\n", rs, sizeof(real));
        exit(4);
    }
    rd(f, &in.n, 1);
    rd(f, &in.s, 1);
    rd(f, &in.steps, 1);
    rd(f, &in.dt, 1);
    rd(f, &in.n_rec, 1);
    in.rec.resize(in.n_rec);
    rd(f, in.rec.data(), in.n_rec);
    in.kind.resize(in.n);
    rd(f, in.kind.data(), in.n);
    in.lif.resize((size_t)in.n * 12);
    rd(f, in.lif.data(), in.lif.size());
    in.arr_off.resize((size_t)in.n + 1);
    rd(f, in.arr_off.data(), in.arr_off.size());
    int64_t na;
    rd(f, &na, 1);
    in.arr_t.resize(na);
    rd(f, in.arr_t.data(), na);
    in.pre.resize(in.s);
    in.post.resize(in.s);
    in.w.resize(in.s);
    in.d.resize(in.s);
    in.stype.resize(in.s);
    in.tau.resize(in.s);
    rd(f, in.pre.data(), in.s);
    rd(f, in.post.data(), in.s);
    rd(f, in.w.data(), in.s);
    rd(f, in.d.data(), in.s);
    rd(f, in.stype.data(), in.s);
    rd(f, in.tau.data(), in.s);
    fclose(f);
    return in;
}

/* Builds the network with the reference API: one population per maximal run
 * of neurons of one kind, filled neuron by neuron with addNeuron (each neuron
 * has its own parameters), then every synapse with
 * Network::connect(NeuronBase*, NeuronBase*, w, d, type, tau, false)
 * (src/net/Network.cpp:118-150) in creation order. */
static void build(Network &net, const Input &in, std::vector<NeuronBase *> &neurons)
{
    neurons.resize(in.n);
    int32_t i = 0;
    while (i < in.n) {
        int32_t j = i;
        while (j < in.n && in.kind[j] == in.kind[i]) j++;
        int cnt = j - i;
        if (in.kind[i] == 0) {
            const real *p = &in.lif[(size_t)i * 12];
            Population<LIF_brian> *pop = net.createPopulation(
                cnt, LIF_brian(LIFENeuron(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9]),
                               p[10], p[11]),
                true);
            for (int32_t k = i; k < j; k++) {
                const real *q = &in.lif[(size_t)k * 12];
                pop->addNeuron(LIF_brian(
                    LIFENeuron(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9]), q[10], q[11]));
            }
            for (int32_t k = i; k < j; k++) neurons[k] = pop->getNeuron(k - i);
        } else if (in.kind[i] == 1) {
            Population<Array_spikes> *pop =
                net.createPopulation(cnt, Array_spikes(ArrayNeuron(), 1e-3, 1e-3), true);
            for (int32_t k = i; k < j; k++) {
                std::vector<int> t(in.arr_t.begin() + in.arr_off[k], in.arr_t.begin() + in.arr_off[k + 1]);
                const real *q = &in.lif[(size_t)k * 12];
                pop->addNeuron(Array_spikes(ArrayNeuron(t.data(), (int)t.size()), q[10], q[11]));
            }
            for (int32_t k = i; k < j; k++) neurons[k] = pop->getNeuron(k - i);
        } else {
            fprintf(stderr, "unsupported neuron kind %d\n", in.kind[i]);
            exit(5);
        }
        i = j;
    }
    for (int64_t k = 0; k < in.s; k++) {
        net.connect(neurons[in.pre[k]], neurons[in.post[k]], in.w[k], in.d[k],
                    in.stype[k] ? Inhibitory : Excitatory, in.tau[k], false);
    }
}

int main(int argc, char **argv)
{
    if (argc == 3 && strcmp(argv[1], "--time") == 0) {
        Input in = load(argv[2]);
        Network net;
        std::vector<NeuronBase *> neurons;
        build(net, in, neurons);
        SingleThreadSimulator st(&net, in.dt);
        struct timeval a, b;
        gettimeofday(&a, NULL);
        st.run(in.dt * in.steps);
        gettimeofday(&b, NULL);
        double sec = (b.tv_sec - a.tv_sec) + 1e-6 * (b.tv_usec - a.tv_usec);
        printf("{\"seconds\": %.6f, \"steps\": %d, \"neurons\": %d, \"synapses\": %lld}\n", sec,
               in.steps, in.n, (long long)in.s);
        return 0;
    }
    if (argc != 3) {
        fprintf(stderr, "usage: %s in.bsnt out.bsout | --time in.bsnt\n", argv[0]);
        return 1;
    }
    Input in = load(argv[1]);
    Network net;
    std::vector<NeuronBase *> neurons;
    build(net, in, neurons);

    // LIF column index of each neuron inside SimInfo::input (LIF neurons only push)
    std::vector<int32_t> lif_col(in.n, -1);
    int32_t nl = 0;
    for (int32_t i = 0; i < in.n; i++)
        if (in.kind[i] == 0) lif_col[i] = nl++;

    SimInfo info;
    info.currCycle = 0;
    info.dt = in.dt;
    net.reset(info);

    std::vector<int32_t> rcount(in.steps), rids;
    /* per cycle and recorded neuron, the five values LIFENeuron::update pushes
     * into SimInfo (LIFENeuron.cpp:144-148): V, the E and I input it consumed,
     * i_E and i_I after the update */
    std::vector<real> recv((size_t)in.steps * in.n_rec * 5, 0);
    for (int32_t t = 0; t < in.steps; t++) {
        info.currCycle = t;
        info.fired.clear();
        info.input.clear();
        net.update(info);
        int32_t c = 0;
        for (size_t i = 0; i < info.fired.size(); i++)
            if (info.fired[i]) {
                rids.push_back((int32_t)i);
                c++;
            }
        rcount[t] = c;
        for (int32_t r = 0; r < in.n_rec; r++) {
            int32_t col = lif_col[in.rec[r]];
            for (int q = 0; q < 5; q++)
                recv[((size_t)q * in.steps + t) * in.n_rec + r] =
                    col >= 0 ? info.input[(size_t)col * 5 + q] : 0;
        }
    }

    FILE *f = xopen(argv[2], "wb");
    wr(f, rcount.data(), rcount.size());
    int64_t len = (int64_t)rids.size();
    wr(f, &len, 1);
    wr(f, rids.data(), rids.size());
    wr(f, recv.data(), recv.size());
    // LIF constants (getData: LIFENeuron.cpp:171-190) as the reference holds them
    for (int32_t i = 0; i < in.n; i++) {
        real c9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        int32_t rt = 0;
        if (in.kind[i] == 0) {
            Json::Value v;
            neurons[i]->getData(&v);
            c9[0] = (real)v["Cm"].asDouble();
            c9[1] = (real)v["CE"].asDouble();
            c9[2] = (real)v["CI"].asDouble();
            c9[3] = (real)v["C_E"].asDouble();
            c9[4] = (real)v["C_I"].asDouble();
            c9[5] = (real)v["v_tmp"].asDouble();
            c9[6] = (real)v["v_thresh"].asDouble();
            c9[7] = (real)v["v_reset"].asDouble();
            c9[8] = (real)v["vm"].asDouble();  // final membrane potential
            rt = v["refrac_time"].asInt();
        }
        wr(f, c9, 9);
        wr(f, &rt, 1);
    }
    for (int64_t k = 0; k < in.s; k++) {
        real w = net.pSynapses[k]->getWeight();
        int32_t ds = net.pSynapses[k]->getDelay();
        wr(f, &w, 1);
        wr(f, &ds, 1);
    }
    for (int32_t i = 0; i < in.n; i++) {
        int32_t fc = neurons[i]->getFireCount();
        wr(f, &fc, 1);
    }
    fclose(f);
    return 0;
}
