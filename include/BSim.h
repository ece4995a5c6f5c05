/*
 * BSim.h -- the reference's C++ user API, as a header-only front end over
 * libbsim_amd's C ABI (bsim_amd.h).  A program written against BSim's
 * include/BSim.h --
 *
 *     Network net;
 *     Population<LIF_brian> *p0 = net.createPopulation(1000, LIF_brian(
 *         LIFENeuron(v_init, v_rest, v_reset, cm, tau_m, tau_refrac,
 *                    tau_syn_E, tau_syn_I, v_thresh, i_offset), tau_syn_E, tau_syn_I));
 *     net.connect(p0, p1, weights, delays, NULL, 1000 * 1000);
 *     SGSim sg(&net, 1e-4f);
 *     sg.run(0.1f);
 *
 * -- compiles unchanged against this header and runs its simulation step on
 * the GPU.  Link with -lbsim_amd.
 *
 * Mirrored interfaces (reference file:line):
 *   Network::createPopulation (num / id, num)      src/net/Network.h:120-144
 *   Network::connect, all-to-all scalar / arrays   src/net/Network.h:146-209
 *   Network::connectOne2One                        src/net/Network.h:211-242
 *   Network::connect(NeuronBase*, NeuronBase*, ..) src/net/Network.cpp:118-150
 *   Network::connect(popSrc, nSrc, popDst, nDst)   src/net/Network.cpp:235-284
 *   Population::getNum / getNeuron / addNeuron     src/base/Population.h:86-89, 158-168
 *   LIFENeuron(v_init, ..., i_offset)              src/neuron/lif/LIFENeuron.h:9
 *   CompositeNeuron / LIF_brian / Array_spikes     src/base/CompositeNeuron.h:20-21,
 *                                                  include/Neuron.h:25-31
 *   ArrayNeuron(int *array, int num)               src/neuron/array/ArrayNeuron.h:19-21
 *   PoissonNeuron(rate, refract, startTime,        src/neuron/poisson/PoissonNeuron.cpp:13-63
 *                 duration) / Poisson_spikes       (seeded: setSeed; see the class)
 *   SGSim(Network*, dt), run(T), run(T, FireInfo&) src/sim/SingleGPUSimulator.h:12-16,
 *                                                  src/base/SimulatorBase.cpp:29
 *   FireInfo, ArrayInfo                            src/base/SimInfo.h:26-31
 *   getConstArray / getRandomArray                 src/utils/utils.h:34-58
 *
 * Differences a user can observe:
 *   - Neuron ids are creation order over the whole network (the reference
 *     renumbers by neuron type when it builds the GPU network); with one
 *     neuron type, as in the reference's LIF programs, the two agree.
 *   - Synapses are handed to the library as they are connected; connect
 *     returns 0 / a count as the reference does, and the NeuronBase* overload
 *     returns a handle valid until the next connect.
 *   - Errors (an unsupported network, a HIP failure) throw bsim_error rather
 *     than exiting the process.
 *   - STSim (the reference's CPU simulator) is not provided: constructing
 *     one fails to compile with a message (use SGSim).
 *   - Output files: GSim.info always; GSim.log (fired ids per cycle, ascending)
 *     and g_v.data (V of the monitored neurons per cycle, "%.10lf \t") when
 *     enabled with setLog / monitor; setLogData(true) writes what the
 *     reference writes under -DLOG_DATA (src/sim/SingleGPUSimulator.cu:
 *     150-232): GSim.log and, for every LIF neuron, g_v.data, g_ie.data,
 *     g_ii.data, g_input_e.data and g_input_i.data.  FireInfo "count",
 *     "Y" (summed LIF input) and "X" (delay slots visited) are filled as the
 *     reference GPU simulator fills them (src/sim/SingleGPUSimulator.cu:263-296);
 *     run(T) writes GFire.count (src/base/SimulatorBase.cpp:29-37).
 */
#ifndef BSIM_AMD_BSIM_H
#define BSIM_AMD_BSIM_H

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>
#include <time.h>

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "bsim_amd.h"

typedef float real;

enum SpikeType { Excitatory = 0, Inhibitory = 1 };

struct ArrayInfo {
    int size;
    void *data;
};
typedef std::map<std::string, ArrayInfo> FireInfo;

class bsim_error : public std::runtime_error {
public:
    explicit bsim_error(const std::string &what) : std::runtime_error(what) {}
};

/* Array helpers of the reference's utils (src/utils/utils.h:34-58), which its
 * programs use to build weight / delay / type arrays for connect.  The caller
 * owns the result (delete[]). */
template <typename T>
T *getConstArray(T value, int size)
{
    T *a = new T[size > 0 ? size : 1];
    for (int i = 0; i < size; i++) a[i] = value;
    return a;
}

/* Uniform values in [0, range] from rand(), seeded once from the clock as the
 * reference does (its realRandom, src/utils/utils.cpp:9-13). */
template <typename T>
T *getRandomArray(T range, int size)
{
    static bool seeded = false;
    if (!seeded) {
        srand((unsigned)time(NULL));
        seeded = true;
    }
    T *a = new T[size > 0 ? size : 1];
    for (int i = 0; i < size; i++)
        a[i] = static_cast<T>((double)rand() / (double)RAND_MAX * static_cast<double>(range));
    return a;
}

namespace bsim_cpp {
inline void check(int64_t rc, const char *what)
{
    if (rc < 0) throw bsim_error(std::string(what) + ": " + bsim_last_error());
}
}  // namespace bsim_cpp

class Network;

/* One neuron: a handle carrying its global id once it is in a network. */
class NeuronBase {
public:
    NeuronBase() : gid_(-1) {}
    virtual ~NeuronBase() {}
    int getGid() const { return gid_; }

protected:
    friend class Network;
    template <class N> friend class Population;
    int gid_;
};

class SynapseBase {
public:
    SynapseBase() : index_(-1) {}
    int64_t getIndex() const { return index_; }

private:
    friend class Network;
    int64_t index_;
};

class StaticSynapse {};

class LIFENeuron : public NeuronBase {
public:
    LIFENeuron(real v_init, real v_rest, real v_reset, real cm, real tau_m, real tau_refrac,
               real tau_syn_E, real tau_syn_I, real v_thresh, real i_offset)
    {
        const real p[10] = {v_init, v_rest, v_reset, cm, tau_m, tau_refrac,
                            tau_syn_E, tau_syn_I, v_thresh, i_offset};
        for (int i = 0; i < 10; i++) param_[i] = p[i];
    }
    /* the 10 LIFENeuron parameters in constructor order */
    const real *params() const { return param_; }
    static const int kind = 0;

private:
    real param_[10];
};

class ArrayNeuron : public NeuronBase {
public:
    ArrayNeuron() {}
    /* fire cycles, ascending */
    ArrayNeuron(int *array, int num) : fireTime(array, array + num) {}
    int addFireTime(int cycle)
    {
        fireTime.push_back(cycle);
        return 0;
    }
    std::vector<int> &getFireTimes() { return fireTime; }
    const std::vector<int> &getFireTimes() const { return fireTime; }
    static const int kind = 1;

protected:
    std::vector<int> fireTime;
};

/* PoissonNeuron(rate, refract, startTime, duration)
 * (src/neuron/poisson/PoissonNeuron.cpp:13-63): first spike at
 * int(startTime/dt) + X, then 1 + X cycles after each spike, before
 * int((startTime + duration)/dt); X ~ Poisson(rate).  The reference seeds its
 * engine from the clock; here the draws are a function of the seed (setSeed,
 * default 0), the neuron's id and the cycle, so runs repeat exactly.  refract
 * has no effect on the reference's CPU path and none here. */
class PoissonNeuron : public NeuronBase {
public:
    PoissonNeuron(real rate, real refract, real startTime, real duration)
        : rate_(rate), refract_(refract), start_(startTime), duration_(duration) {}
    PoissonNeuron &setSeed(uint64_t seed)
    {
        seed_ = seed;
        return *this;
    }
    real rate() const { return rate_; }
    real refract() const { return refract_; }
    real startTime() const { return start_; }
    real duration() const { return duration_; }
    uint64_t seed() const { return seed_; }
    static const int kind = 3;

protected:
    real rate_, refract_, start_, duration_;
    uint64_t seed_ = 0;
};

template <class Neuron, class Synapse>
class CompositeNeuron : public Neuron {
public:
    explicit CompositeNeuron(const Neuron &templ, real tau_syn_E = 1e-3f, real tau_syn_I = 1e-3f)
        : Neuron(templ), tau_syn_E_(tau_syn_E), tau_syn_I_(tau_syn_I) {}
    real tauSynE() const { return tau_syn_E_; }
    real tauSynI() const { return tau_syn_I_; }

private:
    real tau_syn_E_, tau_syn_I_;
};

typedef CompositeNeuron<LIFENeuron, StaticSynapse> LIF_brian;
typedef CompositeNeuron<LIFENeuron, StaticSynapse> IF_curr_exp;
typedef CompositeNeuron<ArrayNeuron, StaticSynapse> Array_spikes;
typedef CompositeNeuron<PoissonNeuron, StaticSynapse> Poisson_spikes;

class PopulationBase {
public:
    virtual ~PopulationBase() {}
    virtual int getNum() = 0;
    virtual NeuronBase *getNeuron(int idx) = 0;
};

namespace bsim_cpp {
/* Appends neurons of one kind to the C-ABI network; returns the first id. */
inline int32_t add_neurons(bsim_network *net, const std::vector<LIF_brian> &v, size_t from)
{
    const size_t n = v.size() - from;
    std::vector<float> rows(n * 12);
    for (size_t i = 0; i < n; i++) {
        const LIF_brian &x = v[from + i];
        for (int j = 0; j < 10; j++) rows[i * 12 + j] = x.params()[j];
        rows[i * 12 + 10] = x.tauSynE();
        rows[i * 12 + 11] = x.tauSynI();
    }
    const int32_t id = bsim_net_add_lif(net, (int32_t)n, rows.data());
    check(id, "createPopulation");
    return id;
}

inline int32_t add_neurons(bsim_network *net, const std::vector<Array_spikes> &v, size_t from)
{
    const size_t n = v.size() - from;
    std::vector<int64_t> off(n + 1, 0);
    std::vector<int32_t> cyc;
    std::vector<float> taus(n * 2);
    for (size_t i = 0; i < n; i++) {
        const std::vector<int> &ft = v[from + i].getFireTimes();
        cyc.insert(cyc.end(), ft.begin(), ft.end());
        off[i + 1] = (int64_t)cyc.size();
        taus[i * 2] = v[from + i].tauSynE();
        taus[i * 2 + 1] = v[from + i].tauSynI();
    }
    const int32_t id = bsim_net_add_array(net, (int32_t)n, off.data(),
                                          cyc.empty() ? NULL : cyc.data(), taus.data());
    check(id, "createPopulation");
    return id;
}

inline int32_t add_neurons(bsim_network *net, const std::vector<Poisson_spikes> &v, size_t from)
{
    int32_t first = -1;
    for (size_t a = from; a < v.size();) {          /* one call per run of equal seeds */
        size_t b = a;
        while (b < v.size() && v[b].seed() == v[a].seed()) b++;
        const size_t n = b - a;
        std::vector<float> rate(n), start(n), dur(n), taus(n * 2);
        for (size_t i = 0; i < n; i++) {
            const Poisson_spikes &x = v[a + i];
            rate[i] = x.rate();
            start[i] = x.startTime();
            dur[i] = x.duration();
            taus[i * 2] = x.tauSynE();
            taus[i * 2 + 1] = x.tauSynI();
        }
        const int32_t id = bsim_net_add_poisson(net, (int32_t)n, rate.data(), start.data(), dur.data(),
                                                taus.data(), v[a].seed());
        check(id, "createPopulation");
        if (first < 0) first = id;
        a = b;
    }
    return first;
}
}  // namespace bsim_cpp

template <class Neuron>
class Population : public PopulationBase {
public:
    /* the declared size, or the number of neurons if more were added */
    int getNum() { return (int)(neurons_.size() > declared_ ? neurons_.size() : declared_); }
    NeuronBase *getNeuron(int idx) { return &neurons_[idx]; }
    /* Appends one neuron and returns its index in the population
     * (Population::addNeuron, src/base/Population.h:158-168). */
    int addNeuron(const Neuron &templ)
    {
        neurons_.push_back(templ);
        const int32_t id = bsim_cpp::add_neurons(net_, neurons_, neurons_.size() - 1);
        neurons_.back().gid_ = id;
        return (int)neurons_.size() - 1;
    }

private:
    friend class Network;
    Population(bsim_network *net, int num, const Neuron *templ)
        : net_(net), declared_(num > 0 ? (size_t)num : 0)
    {
        neurons_.reserve(declared_);
        if (templ && num > 0) {
            neurons_.assign((size_t)num, *templ);
            const int32_t id0 = bsim_cpp::add_neurons(net_, neurons_, 0);
            for (int i = 0; i < num; i++) neurons_[i].gid_ = id0 + i;
        }
    }
    bsim_network *net_;
    size_t declared_;
    std::vector<Neuron> neurons_;
};

class Network {
public:
    Network() : net_(bsim_net_create()), last_()
    {
        if (!net_) throw bsim_error("bsim_net_create failed");
    }
    ~Network()
    {
        flush();
        for (size_t i = 0; i < pops_.size(); i++) delete pops_[i];
        bsim_net_destroy(net_);
    }

    template <class Neuron>
    Population<Neuron> *createPopulation(int num, Neuron templ, bool empty = false)
    {
        flush();
        Population<Neuron> *p = new Population<Neuron>(net_, num, empty ? NULL : &templ);
        pops_.push_back(p);
        return p;
    }
    template <class Neuron>
    Population<Neuron> *createPopulation(int id, int num, Neuron templ, bool empty = false)
    {
        (void)id;
        return createPopulation(num, templ, empty);
    }

    /* all-to-all, one weight and delay (src/net/Network.h:146-174) */
    template <class N1, class N2>
    int connect(Population<N1> *pSrc, Population<N2> *pDst, real weight, real delay, SpikeType type)
    {
        const int ns = pSrc->getNum(), nd = pDst->getNum();
        for (int i = 0; i < ns; i++)
            for (int j = 0; j < nd; j++)
                add(pSrc->getNeuron(i)->getGid(), pDst->getNeuron(j)->getGid(), weight, delay, type, 0.0f);
        return ns * nd;
    }
    /* all-to-all with per-synapse arrays, k = iSrc * dstSize + iDst
     * (src/net/Network.h:176-209); type NULL means excitatory */
    template <class N1, class N2>
    int connect(Population<N1> *pSrc, Population<N2> *pDst, real *weight, real *delay, SpikeType *type,
                int size)
    {
        const int ns = pSrc->getNum(), nd = pDst->getNum();
        if (size != ns * nd) throw bsim_error("connect: size must equal src * dst");
        for (int k = 0; k < size; k++)
            add(pSrc->getNeuron(k / nd)->getGid(), pDst->getNeuron(k % nd)->getGid(), weight[k],
                delay[k], type ? type[k] : Excitatory, 0.0f);
        return size;
    }
    /* neuron i to neuron i (src/net/Network.h:211-242) */
    template <class N1, class N2>
    int connectOne2One(Population<N1> *pSrc, Population<N2> *pDst, real *weight, real *delay,
                       SpikeType *type, int size)
    {
        if (size != pSrc->getNum() || size != pDst->getNum())
            throw bsim_error("connectOne2One: size must equal both population sizes");
        for (int i = 0; i < size; i++)
            add(pSrc->getNeuron(i)->getGid(), pDst->getNeuron(i)->getGid(), weight[i], delay[i],
                type ? type[i] : Excitatory, 0.0f);
        return size;
    }
    /* one synapse (src/net/Network.cpp:118-150) */
    SynapseBase *connect(NeuronBase *pSrc, NeuronBase *pDst, real weight, real delay,
                         SpikeType type = Excitatory, real tau = 0, bool store = true)
    {
        (void)store;
        last_.index_ = add(pSrc->getGid(), pDst->getGid(), weight, delay, type, tau);
        return &last_;
    }
    /* by population / neuron index; inhibitory iff weight < 0
     * (src/net/Network.cpp:235-284) */
    int connect(int popSrc, int nSrc, int popDst, int nDst, real weight, real delay, real tau = 0)
    {
        NeuronBase *a = findNeuron(popSrc, nSrc);
        NeuronBase *b = findNeuron(popDst, nDst);
        if (!a) return -1;
        if (!b) return -2;
        add(a->getGid(), b->getGid(), weight, delay, weight < 0.0f ? Inhibitory : Excitatory, tau);
        return 0;
    }

    PopulationBase *findPopulation(int id)
    {
        return (id >= 0 && id < (int)pops_.size()) ? pops_[id] : NULL;
    }
    NeuronBase *findNeuron(int pop, int idx)
    {
        PopulationBase *p = findPopulation(pop);
        if (!p || idx < 0 || idx >= p->getNum()) return NULL;
        return p->getNeuron(idx);
    }

    /* The C-ABI network with every synapse connected so far. */
    const bsim_network *handle()
    {
        flush();
        return net_;
    }

private:
    int64_t add(int pre, int post, real w, real d, SpikeType type, real tau)
    {
        if (pre < 0 || post < 0) throw bsim_error("connect: neuron is not in a network");
        pre_.push_back(pre);
        post_.push_back(post);
        w_.push_back(w);
        d_.push_back(d);
        type_.push_back((int8_t)type);
        tau_.push_back(tau);
        const int64_t idx = bsim_net_num_synapses(net_) + (int64_t)pre_.size() - 1;
        if (pre_.size() >= (size_t)1 << 20) flush();
        return idx;
    }
    void flush()
    {
        if (pre_.empty()) return;
        const int64_t rc = bsim_net_connect(net_, (int64_t)pre_.size(), pre_.data(), post_.data(),
                                            w_.data(), d_.data(), type_.data(), tau_.data());
        pre_.clear(); post_.clear(); w_.clear(); d_.clear(); type_.clear(); tau_.clear();
        bsim_cpp::check(rc, "connect");
    }

    bsim_network *net_;
    std::vector<PopulationBase *> pops_;
    std::vector<int32_t> pre_, post_;
    std::vector<float> w_, d_, tau_;
    std::vector<int8_t> type_;
    SynapseBase last_;
};

/* SingleGPUSimulator (src/sim/SingleGPUSimulator.h:10-20). */
class SingleGPUSimulator {
public:
    SingleGPUSimulator(Network *network, real dt)
        : network_(network), dt_(dt), device_(0), log_raster_(false), log_data_(false) {}
    virtual ~SingleGPUSimulator() {}

    /* Extensions: GSim.log raster and g_v.data trace of the given neurons. */
    void setLog(bool raster) { log_raster_ = raster; }
    void monitor(const std::vector<int> &gids) { monitor_ = gids; }
    void setDevice(int device) { device_ = device; }
    /* The output of the reference built with -DLOG_DATA
     * (src/sim/SingleGPUSimulator.cu:150-232): GSim.log, and per cycle one
     * line of every LIF_brian neuron's V (g_v.data), i_E and i_I after the
     * update (g_ie.data, g_ii.data) and the E / I synaptic input the update
     * consumed (g_input_e.data, g_input_i.data), "%.10lf \t" per neuron.
     * Replaces monitor(). */
    void setLogData(bool on) { log_data_ = on; }

    /* SimulatorBase::run (src/base/SimulatorBase.cpp:29-37): runs with a
     * "count" log and writes it to GFire.count ("%d " per neuron, then a
     * newline, as logFireInfo in src/base/SimInfo.cpp:4-33 does). */
    int run(real time)
    {
        FireInfo log;
        ArrayInfo zero = {0, NULL};
        log["count"] = zero;
        const int ret = run(time, log);
        FILE *f = fopen("GFire.count", "w+");
        if (!f) {
            printf("ERROR: Open file GFire.count failed\n");
            free(log["count"].data);
            return -1;
        }
        const int *c = reinterpret_cast<const int *>(log["count"].data);
        for (int j = 0; j < log["count"].size; j++) fprintf(f, "%d ", c[j]);
        fprintf(f, "\n");
        fclose(f);
        free(log["count"].data);
        return ret;
    }

    int run(real time, FireInfo &log)
    {
        const int cycles = (int)roundf(time / dt_);
        const bsim_network *net = network_->handle();
        if (log_data_) {
            /* every LIF neuron, in id order */
            const int32_t n = bsim_net_num_neurons(net);
            std::vector<int8_t> kind((size_t)(n > 0 ? n : 1));
            bsim_cpp::check(bsim_net_kinds(net, kind.data()), "SGSim");
            monitor_.clear();
            for (int32_t i = 0; i < n; i++)
                if (kind[i] == 0) monitor_.push_back(i);
        }
        const bool raster = log_raster_ || log_data_;
        bsim_sim_config cfg;
        cfg.device = device_;
        cfg.tile = 0;
        cfg.rec_steps = (raster || !monitor_.empty()) ? cycles : 0;
        cfg.n_rec = (int32_t)monitor_.size();
        cfg.rec_ids = monitor_.empty() ? NULL : monitor_.data();
        bsim_sim *sim = bsim_sim_create(net, dt_, &cfg);
        if (!sim) throw bsim_error(std::string("SGSim: ") + bsim_last_error());
        struct Guard {
            bsim_sim *s;
            ~Guard() { bsim_sim_destroy(s); }
        } guard = {sim};
        if (log_data_ && !monitor_.empty())
            bsim_cpp::check(bsim_sim_enable_state_trace(sim), "SGSim state trace");
        bsim_sim_info info;
        bsim_cpp::check(bsim_sim_get_info(sim, &info), "SGSim");
        const bool want_y = log.find("Y") != log.end(), want_x = log.find("X") != log.end();
        if (want_y || want_x) bsim_cpp::check(bsim_sim_enable_input_log(sim), "SGSim input log");

        FILE *inf = fopen("GSim.info", "w+");
        if (!inf) {
            printf("ERROR: Open file GSim.info failed\n");
            return -1;
        }
        fprintf(inf, "NeuronNum: %d, SynapseNum: %lld\n", info.n, (long long)info.s);
        fprintf(inf, "MAX_DELAY: %d\n", info.dmax);
        fprintf(inf, "Start runing for %d cycles\n", cycles);
        struct timeval ts, te;
        gettimeofday(&ts, NULL);
        bsim_cpp::check(bsim_sim_step(sim, cycles), "SGSim step");
        bsim_cpp::check(bsim_sim_finish(sim), "SGSim finish");
        bsim_cpp::check(bsim_sim_sync(sim), "SGSim sync");
        gettimeofday(&te, NULL);
        long us = (te.tv_sec - ts.tv_sec) * 1000000L + (te.tv_usec - ts.tv_usec);
        fprintf(inf, "Simulation finished in %ld:%ld:%ld.%06lds\n", us / 3600000000L,
                (us / 60000000L) % 60, (us / 1000000L) % 60, us % 1000000L);
        fprintf(inf, "Synaptic events: %lld\n", (long long)bsim_sim_events(sim));
        fclose(inf);

        if (raster) write_raster(sim, cycles);
        if (!monitor_.empty()) write_trace(sim, cycles);
        if (log_data_ && !monitor_.empty()) write_state_traces(sim, cycles);

        if (log.find("count") != log.end()) {
            int *count = (int *)malloc(sizeof(int) * (size_t)(info.n > 0 ? info.n : 1));
            bsim_cpp::check(bsim_sim_read_state(sim, NULL, NULL, NULL, NULL, count), "SGSim count");
            log["count"].size = info.n;
            log["count"].data = count;
        }
        if (want_y || want_x) {
            const size_t n = (size_t)(info.n > 0 ? info.n : 1);
            real *Y = (real *)malloc(sizeof(real) * n);
            int *X = (int *)malloc(sizeof(int) * n);
            bsim_cpp::check(bsim_sim_read_input_log(sim, Y, X), "SGSim input log");
            if (want_y) {
                log["Y"].size = info.n;
                log["Y"].data = Y;
            } else {
                free(Y);
            }
            if (want_x) {
                log["X"].size = info.n;
                log["X"].data = X;
            } else {
                free(X);
            }
        }
        return 0;
    }

private:
    void write_raster(bsim_sim *sim, int cycles)
    {
        std::vector<int32_t> counts((size_t)(cycles > 0 ? cycles : 1));
        int64_t need = 0;
        int64_t rc = bsim_sim_read_raster(sim, counts.data(), NULL, 0, &need);
        std::vector<int32_t> ids((size_t)(need > 0 ? need : 1));
        rc = bsim_sim_read_raster(sim, counts.data(), ids.data(), need, &need);
        bsim_cpp::check(rc, "SGSim raster");
        FILE *f = fopen("GSim.log", "w+");
        if (!f) throw bsim_error("Open file GSim.log failed");
        size_t k = 0;
        for (int c = 0; c < cycles; c++) {
            for (int i = 0; i < counts[c]; i++) fprintf(f, "%d ", ids[k++]);
            fprintf(f, "\n");
        }
        fclose(f);
    }
    void write_trace(bsim_sim *sim, int cycles)
    {
        std::vector<float> v((size_t)cycles * monitor_.size() + 1);
        bsim_cpp::check(bsim_sim_read_trace(sim, v.data()), "SGSim trace");
        FILE *f = fopen("g_v.data", "w+");
        if (!f) throw bsim_error("Open file g_v.data failed");
        for (int c = 0; c < cycles; c++) {
            for (size_t i = 0; i < monitor_.size(); i++)
                fprintf(f, "%.10lf \t", (double)v[(size_t)c * monitor_.size() + i]);
            fprintf(f, "\n");
        }
        fclose(f);
    }

    void write_state_traces(bsim_sim *sim, int cycles)
    {
        const size_t m = monitor_.size(), sz = (size_t)cycles * m + 1;
        std::vector<float> a[4] = {std::vector<float>(sz), std::vector<float>(sz),
                                   std::vector<float>(sz), std::vector<float>(sz)};
        bsim_cpp::check(bsim_sim_read_state_trace(sim, a[0].data(), a[1].data(), a[2].data(),
                                                  a[3].data()),
                        "SGSim state trace");
        const char *names[4] = {"g_ie.data", "g_ii.data", "g_input_e.data", "g_input_i.data"};
        for (int k = 0; k < 4; k++) {
            FILE *f = fopen(names[k], "w+");
            if (!f) throw bsim_error(std::string("Open file ") + names[k] + " failed");
            for (int c = 0; c < cycles; c++) {
                for (size_t i = 0; i < m; i++) fprintf(f, "%.10lf \t", (double)a[k][(size_t)c * m + i]);
                fprintf(f, "\n");
            }
            fclose(f);
        }
    }

    Network *network_;
    real dt_;
    int device_;
    bool log_raster_;
    bool log_data_;
    std::vector<int> monitor_;
};

typedef SingleGPUSimulator SGSim;

/* The reference's single-threaded CPU simulator (src/sim/
 * SingleThreadSimulator.cpp) is not part of this library: the step runs on
 * the GPU only.  Constructing an STSim is a compile-time error, so that a
 * program that compares STSim with SGSim (as test/gpu/gpusim_unit_test.cpp
 * does) cannot silently compare the GPU with itself. */
class STSim {
public:
    template <class... A>
    STSim(A &&...)
    {
        static_assert(sizeof...(A) == 0 && sizeof...(A) != 0,
                      "STSim (the reference's CPU simulator) is not provided by bsim_amd; "
                      "use SGSim, the GPU simulator");
    }
};

#endif
