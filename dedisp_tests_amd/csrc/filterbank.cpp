// SIGPROC filterbank input and PRESTO .dat/.inf output of the filterbank
// front end (reference bin/fil/fil.hpp).
//
// Reading follows the reference's header scan (fil.hpp:288-350): a sequence of
// length-prefixed keywords, each followed by its value, up to HEADER_END; a
// length outside (1, 80) is skipped four bytes at a time, so values of
// keywords that are not interpreted are stepped over the same way. The sample
// count comes from the file size (fil.hpp:342-346). Differences: the value of
// source_name is read (the reference stores the keyword itself), every field
// starts at zero, and end of file inside the header is an error instead of an
// endless loop. The .inf text is PRESTO's fixed layout, field for field as the
// reference writes it (fil.hpp:468-539).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "dedisp_amd.h"

namespace {

struct File {
    std::FILE* f = nullptr;
    explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

template <typename T>
bool read_value(std::FILE* f, T& v)
{
    return std::fread(&v, sizeof(T), 1, f) == 1;
}

// SIGPROC telescope and machine identifiers (the table of fil.hpp:354-466).
const char* telescope_name(int id)
{
    static const char* names[] = {"Fake",       "Arecibo", "Ooty",  "Nancay", "Parkes",
                                  "Jodrell",    "GBT",     "GMRT",  "Effelsberg", "ATA",
                                  "UTR-2",      "LOFAR",   "FR606", "DE601",  "UK608"};
    return (id >= 0 && id < 15) ? names[id] : "???????";
}

const char* backend_name(int id)
{
    switch (id) {
    case 0: return "FAKE";
    case 1: return "PSPM";
    case 2: return "WAPP";
    case 3: return "AOFTM";
    case 4: return "BPP";
    case 5: return "OOTY";
    case 6: return "SCAMP";
    case 7: return "SPIGOT";
    case 10: return "ARTEMIS";
    case 11: return "Cobalt";
    default: return "????";
    }
}

std::string stem(const char* prefix, float dm)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "_DM%.3f", (double)dm);
    return std::string(prefix) + buf;
}

}  // namespace

extern "C" {

dedisp_error dedisp_amd_fil_read_header(const char* path, dedisp_amd_fil_header* h)
{
    if (!path || !h) return DEDISP_INVALID_POINTER;
    std::memset(h, 0, sizeof *h);
    File file(path, "rb");
    if (!file.f) return DEDISP_UNKNOWN_ERROR;
    std::FILE* f = file.f;
    for (;;) {
        int nchar = 0;
        if (!read_value(f, nchar)) return DEDISP_UNKNOWN_ERROR;  // no HEADER_END
        if (!(nchar > 1 && nchar < 80)) continue;
        char key[80];
        if (std::fread(key, 1, (size_t)nchar, f) != (size_t)nchar) return DEDISP_UNKNOWN_ERROR;
        key[nchar] = '\0';
        if (!std::strcmp(key, "HEADER_END")) break;
        bool ok = true;
        if (!std::strcmp(key, "tsamp")) ok = read_value(f, h->tsamp);
        else if (!std::strcmp(key, "tstart")) ok = read_value(f, h->tstart);
        else if (!std::strcmp(key, "fch1")) ok = read_value(f, h->fch1);
        else if (!std::strcmp(key, "foff")) ok = read_value(f, h->foff);
        else if (!std::strcmp(key, "nchans")) ok = read_value(f, h->nchans);
        else if (!std::strcmp(key, "nifs")) ok = read_value(f, h->nifs);
        else if (!std::strcmp(key, "nbits")) ok = read_value(f, h->nbits);
        else if (!std::strcmp(key, "nsamples")) ok = read_value(f, h->nsamps);
        else if (!std::strcmp(key, "az_start")) ok = read_value(f, h->az_start);
        else if (!std::strcmp(key, "za_start")) ok = read_value(f, h->za_start);
        else if (!std::strcmp(key, "src_raj")) ok = read_value(f, h->src_raj);
        else if (!std::strcmp(key, "src_dej")) ok = read_value(f, h->src_dej);
        else if (!std::strcmp(key, "telescope_id")) ok = read_value(f, h->telescope_id);
        else if (!std::strcmp(key, "machine_id")) ok = read_value(f, h->machine_id);
        else if (!std::strcmp(key, "nbeams")) ok = read_value(f, h->nbeams);
        else if (!std::strcmp(key, "ibeam")) ok = read_value(f, h->ibeam);
        else if (!std::strcmp(key, "source_name")) {
            int n = 0;
            ok = read_value(f, n) && n >= 0 && n < (int)sizeof h->source_name;
            if (ok) {
                ok = std::fread(h->source_name, 1, (size_t)n, f) == (size_t)n;
                h->source_name[ok ? n : 0] = '\0';
            }
        }
        if (!ok) return DEDISP_UNKNOWN_ERROR;
    }
    const long header_end = std::ftell(f);
    if (header_end < 0 || std::fseek(f, 0, SEEK_END) != 0) return DEDISP_UNKNOWN_ERROR;
    const long end = std::ftell(f);
    h->header_bytes = header_end;
    h->data_bytes = end - header_end;
    const long sample_bytes = (long)h->nchans * h->nifs * h->nbits / 8;
    h->nsamps = sample_bytes > 0 ? (int)(h->data_bytes / sample_bytes) : 0;
    return DEDISP_NO_ERROR;
}

dedisp_error dedisp_amd_fil_read_data(const char* path, const dedisp_amd_fil_header* h, dedisp_byte* data,
                                      dedisp_size nbytes)
{
    if (!path || !h || !data) return DEDISP_INVALID_POINTER;
    if (nbytes > (dedisp_size)h->data_bytes) return DEDISP_INVALID_POINTER;
    File file(path, "rb");
    if (!file.f) return DEDISP_UNKNOWN_ERROR;
    if (std::fseek(file.f, h->header_bytes, SEEK_SET) != 0) return DEDISP_UNKNOWN_ERROR;
    if (std::fread(data, 1, nbytes, file.f) != nbytes) return DEDISP_UNKNOWN_ERROR;
    return DEDISP_NO_ERROR;
}

// Averages groups of `factor` consecutive samples per channel in place (8-bit
// samples; the mean truncated toward zero and clamped to 0..255, fil.hpp:31-52).
dedisp_size dedisp_amd_fil_decimate(dedisp_byte* data, dedisp_size nchans, dedisp_size nsamps,
                                    dedisp_size factor)
{
    if (!data || factor == 0) return 0;
    if (factor == 1) return nsamps;
    const dedisp_size nout = nsamps / factor;
    for (dedisp_size l = 0; l < nout; ++l) {
        const dedisp_size j = l * factor;
        for (dedisp_size i = 0; i < nchans; ++i) {
            float acc = 0.f;
            for (dedisp_size k = 0; k < factor; ++k) acc += (float)data[i + nchans * (j + k)];
            acc /= (float)factor;
            data[i + nchans * l] = acc > 255.f ? (dedisp_byte)255 : acc < 0.f ? (dedisp_byte)0 : (dedisp_byte)acc;
        }
    }
    return nout;
}

dedisp_error dedisp_amd_write_dat(const char* prefix, dedisp_float dm, const dedisp_float* series,
                                  dedisp_size numout)
{
    if (!prefix || (!series && numout)) return DEDISP_INVALID_POINTER;
    const std::string name = stem(prefix, dm) + ".dat";
    File file(name.c_str(), "wb");
    if (!file.f) return DEDISP_UNKNOWN_ERROR;
    if (std::fwrite(series, sizeof(dedisp_float), numout, file.f) != numout) return DEDISP_UNKNOWN_ERROR;
    return DEDISP_NO_ERROR;
}

dedisp_error dedisp_amd_write_inf(const char* prefix, const dedisp_amd_fil_header* h, dedisp_float dm,
                                  dedisp_size numout)
{
    if (!prefix || !h) return DEDISP_INVALID_POINTER;
    const std::string base = stem(prefix, dm);
    File file((base + ".inf").c_str(), "w");
    if (!file.f) return DEDISP_UNKNOWN_ERROR;
    std::FILE* f = file.f;
    // src_raj / src_dej are SIGPROC's hhmmss.s / ddmmss.s numbers
    const int ra_h = (int)std::floor(h->src_raj / 10000.0);
    const int ra_m = (int)std::floor((h->src_raj - ra_h * 10000) / 100.0);
    const double ra_s = h->src_raj - ra_h * 10000 - ra_m * 100;
    const double adec = std::fabs(h->src_dej);
    const int dec_d = (int)std::floor(adec / 10000.0);
    const int dec_m = (int)std::floor((adec - dec_d * 10000) / 100.0);
    const double dec_s = adec - dec_d * 10000 - dec_m * 100;
    const char sign = h->src_dej < 0.0 ? '-' : ' ';
    // MJD: integer day, then the fifteen printed digits of the day fraction
    const double day = std::floor(h->tstart);
    char frac[64];
    std::snprintf(frac, sizeof frac, "%.15f", h->tstart - day);
    const char* digits = std::strchr(frac, '.');
    digits = digits ? digits + 1 : "0";
    const double bw = std::fabs(h->foff);

    std::fprintf(f, " Data file name without suffix          =  %s\n", base.c_str());
    std::fprintf(f, " Telescope used                         =  %s\n", telescope_name(h->telescope_id));
    std::fprintf(f, " Instrument used                        =  %s\n", backend_name(h->machine_id));
    std::fprintf(f, " Object being observed                  =  %s\n", h->source_name);
    std::fprintf(f, " J2000 Right Ascension (hh:mm:ss.ssss)  =  %02d:%02d:%02f\n", ra_h, ra_m, ra_s);
    std::fprintf(f, " J2000 Declination     (dd:mm:ss.ssss)  = %c%02d:%02d:%s%f\n", sign, dec_d, dec_m,
                 dec_s < 10 ? "0" : "", dec_s);
    std::fprintf(f, " Data observed by                       =  Unknown\n");
    std::fprintf(f, " Epoch of observation (MJD)             =  %d.%s\n", (int)day, digits);
    std::fprintf(f, " Barycentered?           (1=yes, 0=no)  =  0\n");
    std::fprintf(f, " Number of bins in the time series      =  %d\n", (int)numout);
    std::fprintf(f, " Width of each time series bin (sec)    =  %.15g\n", h->tsamp);
    std::fprintf(f, " Any breaks in the data? (1=yes, 0=no)  =  0\n");
    std::fprintf(f, " Type of observation (EM band)          =  Radio\n");
    std::fprintf(f, " Beam diameter (arcsec)                 =  3600\n");
    std::fprintf(f, " Dispersion measure (cm-3 pc)           =  %.12g\n", (double)dm);
    std::fprintf(f, " Central freq of low channel (Mhz)      =  %.12g\n", h->fch1 - (h->nchans - 1) * bw);
    std::fprintf(f, " Total bandwidth (Mhz)                  =  %.12g\n", bw * h->nchans);
    std::fprintf(f, " Number of channels                     =  %d\n", h->nchans);
    std::fprintf(f, " Channel bandwidth (Mhz)                =  %.12g\n", bw);
    std::fprintf(f, " Data analyzed by                       =  Unknown\n");
    std::fprintf(f, " Any additional notes:\n    Input filterbank samples have %d bits.\n", h->nbits);
    return std::ferror(f) ? DEDISP_UNKNOWN_ERROR : DEDISP_NO_ERROR;
}

}  // extern "C"
