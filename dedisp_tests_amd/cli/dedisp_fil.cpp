// dedisp_fil: dedisperse a SIGPROC filterbank file into PRESTO .dat/.inf
// time series, one pair per DM trial (reference bin/fil/fil.hpp run<> with
// dedisp_fil.cpp / tdd_fil.cpp / fdd_fil.cpp as the -i choice).
//
//   dedisp_fil -f <file.fil> [-o prefix] [-D device] [-i dedisp|tdd|fdd]
//              [-r start[,end]] [-s step] [-n ntrials] [-d decimate]
//              [-N nsamples] [-w pulse_width_us] [-t tolerance] [-q]
//
// A linear DM list (start + step*i, i < ntrials; ntrials defaults to
// ceil((end - start)/step)) is used when -s is given, otherwise the generated
// list of dedisp_generate_dm_list(start, end, width, tolerance); the reference
// tool accepts only the linear form. FDD plans write time series.
#include <getopt.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dedisp_amd.h"

namespace {

void usage()
{
    std::printf(
        "Usage: dedisp_fil -f <file> [-o prefix] [-D device] [-i dedisp|tdd|fdd] [-r start[,end]]\n"
        "                  [-s step] [-n ntrials] [-d decimate] [-N nsamples] [-w width_us]\n"
        "                  [-t tolerance] [-q]\n\n"
        "-f <file>          SIGPROC filterbank file to read\n"
        "-o <prefix>        output prefix (default: test)\n"
        "-D <device>        GPU device (default: 0)\n"
        "-i <impl>          plan: dedisp (legacy selector), tdd or fdd (default: dedisp)\n"
        "-r <start[,end]>   DM range (default: 0,50)\n"
        "-s <step>          linear DM step (default: generated DM trials)\n"
        "-n <ntrials>       number of linear DM trials\n"
        "-d <factor>        decimate in time by this factor (default: 1)\n"
        "-N <nsamples>      samples to write per trial (default: all)\n"
        "-w <width>         intrinsic pulse width in us for generated trials (default: 4)\n"
        "-t <tolerance>     smearing tolerance for generated trials (default: 1.25)\n"
        "-q                 quiet\n");
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int fail(const char* what, dedisp_error e)
{
    std::fprintf(stderr, "dedisp_fil: %s: %s\n", what, dedisp_get_error_string(e));
    return 1;
}

}  // namespace

int main(int argc, char** argv)
{
    const char* filename = nullptr;
    std::string prefix = "test";
    int device = 0;
    dedisp_implementation impl = DEDISP_DEDISP;
    float dm_start = 0.f, dm_end = 50.f, dm_step = 0.f, width = 4.f, tol = 1.25f;
    size_t ntrials = 0, ndec = 1, numout = 0;
    bool verbose = true;
    int opt;
    while ((opt = getopt(argc, argv, "f:o:D:i:r:s:n:d:N:w:t:qh")) != -1) {
        switch (opt) {
        case 'f': filename = optarg; break;
        case 'o': prefix = optarg; break;
        case 'D': device = std::atoi(optarg); break;
        case 'i':
            if (!std::strcmp(optarg, "tdd")) impl = DEDISP_TDD;
            else if (!std::strcmp(optarg, "fdd")) impl = DEDISP_FDD;
            else if (!std::strcmp(optarg, "dedisp")) impl = DEDISP_DEDISP;
            else { usage(); return 1; }
            break;
        case 'r':
            if (std::strchr(optarg, ',')) std::sscanf(optarg, "%f,%f", &dm_start, &dm_end);
            else dm_start = (float)std::atof(optarg);
            break;
        case 's': dm_step = (float)std::atof(optarg); break;
        case 'n': ntrials = (size_t)std::atol(optarg); break;
        case 'd': ndec = (size_t)std::atol(optarg); break;
        case 'N': numout = (size_t)std::atol(optarg); break;
        case 'w': width = (float)std::atof(optarg); break;
        case 't': tol = (float)std::atof(optarg); break;
        case 'q': verbose = false; break;
        case 'h': usage(); return 0;
        default: usage(); return 1;
        }
    }
    if (!filename || ndec == 0) {
        usage();
        return 1;
    }

    dedisp_amd_fil_header h;
    dedisp_error e = dedisp_amd_fil_read_header(filename, &h);
    if (e != DEDISP_NO_ERROR) return fail("reading the header", e);
    if (verbose) {
        std::printf("----------------------------- INPUT DATA ---------------------------------\n");
        std::printf("Frequency of highest channel              : %f MHz\n", h.fch1);
        std::printf("Bandwidth                                 : %f MHz\n", std::fabs(h.foff) * h.nchans);
        std::printf("Number of channels (channel width)        : %d (%f MHz)\n", h.nchans, std::fabs(h.foff));
        std::printf("Sample time                               : %f us\n", h.tsamp * 1e6 * (double)ndec);
        std::printf("Observation duration                      : %f s (%d samples)\n", h.tsamp * h.nsamps,
                    (int)(h.nsamps / (int)ndec));
        std::printf("Number of polarizations/bit depth         : %d/%d\n", h.nifs, h.nbits);
        std::printf("Header size                               : %ld bytes\n\n", h.header_bytes);
    }
    if (h.nifs != 1) {
        std::fprintf(stderr, "dedisp_fil: %d polarisations (1 supported)\n", h.nifs);
        return 1;
    }
    if (h.nbits != 8 && (ndec > 1 || (h.nbits != 1 && h.nbits != 2 && h.nbits != 4 &&
                                      !(h.nbits == 16 && impl != DEDISP_FDD)))) {
        std::fprintf(stderr, "dedisp_fil: %d-bit samples are not supported here\n", h.nbits);
        return 1;
    }

    const size_t row = (size_t)h.nchans * h.nbits / 8;
    std::vector<dedisp_byte> input(row * (size_t)h.nsamps);
    auto t0 = std::chrono::steady_clock::now();
    e = dedisp_amd_fil_read_data(filename, &h, input.data(), input.size());
    if (e != DEDISP_NO_ERROR) return fail("reading the samples", e);
    if (verbose) std::printf("Reading input file took %.3f seconds\n", seconds_since(t0));
    size_t nsamps = (size_t)h.nsamps;
    if (ndec > 1) {
        nsamps = dedisp_amd_fil_decimate(input.data(), (size_t)h.nchans, nsamps, ndec);
        h.tsamp *= (double)ndec;
        h.nsamps = (int)nsamps;
    }

    if ((e = dedisp_set_device(device)) != DEDISP_NO_ERROR) return fail("selecting the device", e);
    if ((e = dedisp_select_implementation(impl)) != DEDISP_NO_ERROR) return fail("selecting the plan", e);
    dedisp_plan plan = nullptr;
    if ((e = dedisp_create_plan(&plan, (size_t)h.nchans, (float)h.tsamp, (float)h.fch1, (float)h.foff)) !=
        DEDISP_NO_ERROR)
        return fail("creating the plan", e);
    if (dm_step > 0.f) {
        if (!ntrials) ntrials = (size_t)std::ceil((dm_end - dm_start) / dm_step);
        std::vector<float> dms(ntrials);
        for (size_t i = 0; i < ntrials; ++i) dms[i] = dm_start + dm_step * (float)i;
        e = dedisp_set_dm_list(plan, dms.data(), ntrials);
    } else {
        e = dedisp_generate_dm_list(plan, dm_start, dm_end, width, tol);
    }
    if (e != DEDISP_NO_ERROR) return fail("setting the DM list", e);
    if (impl == DEDISP_FDD) dedisp_amd_fdd_set_output(plan, DEDISP_AMD_FDD_TIMESERIES);

    const size_t ndm = dedisp_get_dm_count(plan), max_delay = dedisp_get_max_delay(plan);
    if (nsamps <= max_delay) {
        std::fprintf(stderr, "dedisp_fil: %zu samples, max delay %zu\n", nsamps, max_delay);
        return 1;
    }
    const size_t nsc = nsamps - max_delay;
    const float* dms = dedisp_get_dm_list(plan);
    if (verbose) {
        std::printf("----------------------------- DM COMPUTATIONS  ----------------------------\n");
        std::printf("Computing %zu DMs from %f to %f pc/cm^3\n", ndm, dms[0], dms[ndm - 1]);
        std::printf("Max DM delay is %zu samples (%f seconds)\n", max_delay, max_delay * h.tsamp);
        std::printf("Computing %zu out of %zu total samples (%.2f%% efficiency)\n\n", nsc, nsamps,
                    100.0 * (double)nsc / (double)nsamps);
    }
    std::vector<float> output(ndm * nsc);
    t0 = std::chrono::steady_clock::now();
    e = dedisp_execute(plan, nsamps, input.data(), (size_t)h.nbits, reinterpret_cast<dedisp_byte*>(output.data()),
                       32, DEDISP_USE_DEFAULT);
    if (e != DEDISP_NO_ERROR) return fail("dedispersing", e);
    if (verbose) std::printf("Dedispersion took %.3f seconds\n", seconds_since(t0));

    if (numout == 0 || numout > nsc) numout = nsc;
    numout &= ~size_t(1);  // even length (fil.hpp:262-266)
    t0 = std::chrono::steady_clock::now();
    for (size_t i = 0; i < ndm; ++i) {
        if ((e = dedisp_amd_write_dat(prefix.c_str(), dms[i], output.data() + i * nsc, numout)) != DEDISP_NO_ERROR)
            return fail("writing a .dat file", e);
        if ((e = dedisp_amd_write_inf(prefix.c_str(), &h, dms[i], numout)) != DEDISP_NO_ERROR)
            return fail("writing an .inf file", e);
    }
    if (verbose) std::printf("Writing DM trials took %.3f seconds\n", seconds_since(t0));
    dedisp_destroy_plan(plan);
    return 0;
}
