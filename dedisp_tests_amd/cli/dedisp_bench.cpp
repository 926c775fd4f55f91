// dedisp_bench: execution-time benchmark of a plan on uninitialised data
// (reference bin/benchmark/bench.hpp: BenchParameters, parseParameters and
// run<PlanType>, with benchdedisp.cpp / benchtdd.cpp / benchfdd.cpp as the -m
// choice).
//
//   dedisp_bench [-m dedisp|tdd|fdd] [-n ntrials] [-s nsamps] [-c nchans]
//                [-r start[,end]] [-t Tobs] [-i niterations] [-d] [-q] [-h]
//
// Geometry as the reference: dt 64 us, highest channel 1581 MHz, 400 MHz band
// split into nchans channels, 8-bit input, 32-bit output, linear DM trials
// start + 2*i (i < ntrials) or, with -r start,end, the generated list of
// dedisp_generate_dm_list(start, end, 4.0, 1.25). Each iteration creates a
// plan, sets its DM list and times one execute; input and output share one
// host buffer whose contents are never initialised (the execution time does
// not depend on them). Differences from the reference tool: -q and -h take no
// argument; -r start,end works (the reference rejects it while parsing); -t
// accepts fractions of a second; the buffer is sized for the plan's output
// (whole-series FDD spectra are longer than nsamps); -d keeps input and
// output in device memory (DEDISP_DEVICE_POINTERS); each iteration also prints
// its execute time and DM-trials x samples per second.
#include <getopt.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "dedisp.hpp"

namespace {

struct BenchParameters {  // defaults of bench.hpp:144-161
    float sampletime_base = 64.0e-6f;
    float downsamp = 1.0f;
    float dt = 64.0e-6f;
    float f0 = 1581.0f;
    float bw = 400.0f;
    size_t nchans = 1024;
    float df = -400.0f / 1024;
    float dm_start = 0.0f;
    float dm_end = 0.0f;
    float dm_step = 2.0f;
    size_t dm_count = 100;
    size_t nsamps = 120000;
    double Tobs = 0.0;
    bool verbose = true;
    int niterations = 1;
    std::string impl = "dedisp";
    bool device = false;
};

void usage()
{
    std::printf("Usage: dedisp_bench [-m dedisp|tdd|fdd] -n [nr of DM trials] -s [nr of samples] -c [nr of channels]\n\n");
    std::printf("No arguments        Default settings\n");
    std::printf("-m [plan]           Plan: dedisp (legacy), tdd or fdd (default: dedisp)\n");
    std::printf("-n [ntrials]        Number of DM trials\n");
    std::printf("-s [samples]        Number of samples to generate\n");
    std::printf("-c [channels]       Number of channels to generate\n");
    std::printf("-r [DM start(,end)] Alternative Start (and optional end) values of DM range to dedisperse (use end instead of -n)\n");
    std::printf("-t [Tobs]           Alternative observation time to use [seconds] (use instead of -s)\n");
    std::printf("-i [niterations]    Number of iterations for the dedispersion plan and execution\n");
    std::printf("-d                  Input and output in device memory\n");
    std::printf("-q                  Quiet; no verbose information to screen\n");
    std::printf("-h                  This help\n");
}

int parse(int argc, char** argv, BenchParameters& p)
{
    if (argc <= 1) {
        usage();
        std::printf("Using default settings\n");
    }
    int opt;
    while ((opt = getopt(argc, argv, "m:r:s:n:c:t:i:dqh")) != -1) {
        switch (opt) {
        case 'm': p.impl = optarg; break;
        case 'r':
            if (std::strchr(optarg, ',')) {
                std::sscanf(optarg, "%f,%f", &p.dm_start, &p.dm_end);
                p.dm_count = 0;
            } else {
                p.dm_start = (float)std::atof(optarg);
            }
            break;
        case 'n': p.dm_count = (size_t)std::atol(optarg); break;
        case 's': p.nsamps = (size_t)std::atol(optarg); break;
        case 'c': p.nchans = (size_t)std::atol(optarg); break;
        case 't': p.Tobs = std::atof(optarg); break;
        case 'i': p.niterations = std::atoi(optarg); break;
        case 'd': p.device = true; break;
        case 'q': p.verbose = false; break;
        case 'h': usage(); return 1;
        default: usage(); return -1;
        }
    }
    if (p.impl != "dedisp" && p.impl != "tdd" && p.impl != "fdd") {
        usage();
        return -1;
    }
    if (p.dm_count != 0 && p.dm_end == 0.0f) {
        p.dm_end = p.dm_start + (float)p.dm_count * p.dm_step;
    } else if (p.dm_count != 0 || p.dm_end <= p.dm_start) {
        std::fprintf(stderr, "Error parsing DM range. Provide number of dms or start and end of a range.\n");
        return -1;
    }
    p.dt = p.downsamp * p.sampletime_base;
    p.df = -1.0f * p.bw / (float)p.nchans;
    if (p.Tobs != 0.0)
        p.nsamps = (size_t)(p.Tobs / p.dt);
    else
        p.Tobs = (double)p.nsamps * p.dt;
    return 0;
}

using clk = std::chrono::steady_clock;
double secs(clk::duration d) { return std::chrono::duration<double>(d).count(); }

std::unique_ptr<dedisp::Plan> make_plan(const BenchParameters& p)
{
    if (p.impl == "tdd") return std::unique_ptr<dedisp::Plan>(new dedisp::TDDPlan(p.nchans, p.dt, p.f0, p.df));
    if (p.impl == "fdd") return std::unique_ptr<dedisp::Plan>(new dedisp::FDDGPUPlan(p.nchans, p.dt, p.f0, p.df));
    return std::unique_ptr<dedisp::Plan>(new dedisp::DedispPlan(p.nchans, p.dt, p.f0, p.df));
}

int run(const BenchParameters& p)
{
    const size_t in_nbits = 8, out_nbits = 32;
    const float pulse_width = 4.0f, dm_tol = 1.25f;  // bench.hpp:246-247
    clk::duration t_total{}, t_init{}, t_plan{}, t_exec{};
    const auto t0 = clk::now();
    std::printf("Starting benchmark for %d iterations \n", p.niterations);
    if (p.verbose) {
        std::printf("------------------------ INPUT AND OUTPUT DATA ---------------------------\n");
        std::printf("Frequency of highest chanel [MHz]            : %.4f\n", p.f0);
        std::printf("NCHANS (Channel Width [MHz])                 : %zu (%f)\n", p.nchans, p.df);
        std::printf("Sample time (after downsampling by %.0f [us])   : %f\n", p.downsamp, p.dt / 1e-6);
        std::printf("Observation duration [s]                     : %f (%zu samples)\n", p.Tobs, p.nsamps);
        std::printf("Input data                                   : %2zu bits \n", in_nbits);
        std::printf("Output data                                  : %2zu bits \n", out_nbits);
        std::printf("Input data array size                        : %zu MB\n", (p.nsamps * p.nchans) >> 20);
        std::printf("\n----------------------------- INJECTED SIGNAL  ----------------------------\n");
        std::printf("Uninitialized dummy data \n");
    }
    // shared input/output host buffer, never initialised on purpose (bench.hpp:296-307)
    std::unique_ptr<unsigned char, void (*)(void*)> host(nullptr, std::free);
    void* dev_in = nullptr;
    void* dev_out = nullptr;
    size_t host_bytes = 0, dev_out_bytes = 0;
    t_init += clk::now() - t0;

    for (int it = 0; it < p.niterations; ++it) {
        const auto tp = clk::now();
        std::printf("\n------------------------- ITERATION %d out of %d  -------------------------\n", it + 1,
                    p.niterations);
        if (p.verbose) std::printf("Create plan and init GPU\n");
        std::unique_ptr<dedisp::Plan> plan = make_plan(p);
        if (p.verbose) std::printf("Gen DM list\n");
        if (p.dm_count == 0) {
            if (p.verbose) std::printf("Generating optimal DM trials\n");
            plan->generate_dm_list(p.dm_start, p.dm_end, pulse_width, dm_tol);
        } else {
            if (p.verbose) std::printf("Generating linear DM trials\n");
            std::vector<float> dms(p.dm_count);
            for (size_t i = 0; i < p.dm_count; ++i) dms[i] = p.dm_start + p.dm_step * (float)i;
            plan->set_dm_list(dms.data(), dms.size());
        }
        const size_t ndm = plan->get_dm_count();
        const size_t max_delay = (size_t)plan->get_max_delay();
        if (p.nsamps <= max_delay) {
            std::fprintf(stderr, "dedisp_bench: %zu samples, max delay %zu\n", p.nsamps, max_delay);
            return 1;
        }
        const size_t nsc = p.nsamps - max_delay;
        size_t out_len = nsc;
        if (auto* f = dynamic_cast<dedisp::FDDGPUPlan*>(plan.get())) out_len = f->output_length(p.nsamps);
        const float* dms = plan->get_dm_list();
        if (p.verbose) {
            std::printf("\n----------------------------- DM COMPUTATIONS  ----------------------------\n");
            std::printf("Computing %zu DMs from %f to %f pc/cm^3\n", ndm, dms[0], dms[ndm - 1]);
            std::printf("Max DM delay is %zu samples (%.f seconds)\n", max_delay, max_delay * p.dt);
            std::printf("Computing %zu out of %zu total samples (%.2f%% efficiency)\n", nsc, p.nsamps,
                        100.0 * (double)nsc / (double)p.nsamps);
        }
        const size_t in_bytes = p.nsamps * p.nchans * in_nbits / 8, out_bytes = ndm * out_len * out_nbits / 8;
        if (p.device) {
            if (!dev_in && hipMalloc(&dev_in, in_bytes) != hipSuccess) return 1;
            if (dev_out_bytes < out_bytes) {
                if (dev_out) (void)hipFree(dev_out);
                if (hipMalloc(&dev_out, out_bytes) != hipSuccess) return 1;
                dev_out_bytes = out_bytes;
            }
        } else if (host_bytes < std::max(in_bytes, out_bytes)) {
            host_bytes = std::max(in_bytes, out_bytes);
            host.reset(static_cast<unsigned char*>(std::malloc(host_bytes)));
            if (!host) {
                std::fprintf(stderr, "\nERROR: Failed to allocate shared input/output dummy buffer array\n");
                return 1;
            }
        }
        t_plan += clk::now() - tp;

        std::printf("\n--------------------------- PERFORM DEDISPERSION  -------------------------\n");
        const auto te = clk::now();
        unsigned char* in = p.device ? static_cast<unsigned char*>(dev_in) : host.get();
        unsigned char* out = p.device ? static_cast<unsigned char*>(dev_out) : host.get();
        plan->execute(p.nsamps, in, in_nbits, out, out_nbits,
                      p.device ? DEDISP_DEVICE_POINTERS : DEDISP_HOST_POINTERS);
        const double dt_exec = secs(clk::now() - te);
        t_exec += clk::now() - te;
        std::printf("Execute time:              %.6f sec. (%.4g DM*samples/s)\n", dt_exec,
                    (double)ndm * (double)nsc / dt_exec);
        if (p.verbose) {
            t_total = clk::now() - t0;
            std::printf("\n------------- BENCHMARK TIMES (accumulated for %d iteration(s)) -------------\n", it + 1);
            std::printf("Benchmark total time:      %.6f sec.\n", secs(t_total));
            std::printf("Benchmark init time:       %.6f sec.\n", secs(t_init));
            std::printf("Benchmark plan time:       %.6f sec.\n", secs(t_plan));
            std::printf("Benchmark execute time:    %.6f sec.\n", secs(t_exec));
        }
    }
    t_total = clk::now() - t0;
    if (!p.verbose) {
        std::printf("\n------------- BENCHMARK TIMES (accumulated for %d iteration(s)) -------------\n",
                    p.niterations);
        std::printf("Benchmark total time:      %.6f sec.\n", secs(t_total));
        std::printf("Benchmark init time:       %.6f sec.\n", secs(t_init));
        std::printf("Benchmark plan time:       %.6f sec.\n", secs(t_plan));
        std::printf("Benchmark execute time:    %.6f sec.\n", secs(t_exec));
    }
    if (dev_in) (void)hipFree(dev_in);
    if (dev_out) (void)hipFree(dev_out);
    std::printf("Benchmark finished\n");
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    BenchParameters p;
    const int rc = parse(argc, argv, p);
    if (rc != 0) return rc > 0 ? 0 : 1;
    try {
        return run(p);
    } catch (const dedisp::Error& e) {
        std::fprintf(stderr, "dedisp_bench: %s (%s)\n", e.what(), dedisp_get_error_string(e.code()));
        return 1;
    }
}
