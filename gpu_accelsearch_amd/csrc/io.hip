// io.hip -- DM-trial input files: PRESTO .dat time series and .inf metadata.
//
// readData restates source/bpd_data.c:9-75: the file is raw float32 samples, at
// most 2^26 of them; the buffer carries ACCEL_PADDING (2000) zero floats before
// and after the data, and the returned pointer is the first sample.  Failures
// return -1 with the message in fdas_last_error (the reference exits).  The .inf
// reader takes the sample count, sample time and DM from the PRESTO text header,
// which the reference's driver hard-codes instead (AccelSearch.cu:90).
#include <sys/stat.h>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "fdas_internal.hpp"

using namespace fdas;

namespace {

constexpr long kPadding = 2000;          // headers/bpd_common.h:35 ACCEL_PADDING
constexpr long kMaxSamples = 67108864;   // source/bpd_data.c:34

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

bool has(const std::string& key, const char* what) { return key.find(what) != std::string::npos; }

}  // namespace

extern "C" long readData(char* path, float** output_data)
{
    if (!path || !output_data) {
        set_error("readData: null argument");
        return -1;
    }
    *output_data = nullptr;
    FILE* f = fopen(path, "rb");
    if (!f) {
        set_error("readData: cannot open %s", path);
        return -1;
    }
    struct stat st;
    if (fstat(fileno(f), &st) != 0) {
        fclose(f);
        set_error("readData: cannot stat %s", path);
        return -1;
    }
    const long n = (long)(st.st_size / (off_t)sizeof(float));
    if (n <= 0) {
        fclose(f);
        set_error("readData: %s holds no samples", path);
        return -1;
    }
    if (n > kMaxSamples) {
        fclose(f);
        set_error("readData: %s holds %ld samples (at most %ld)", path, n, kMaxSamples);
        return -1;
    }
    float* buf = (float*)calloc((size_t)(n + 2 * kPadding), sizeof(float));
    if (!buf) {
        fclose(f);
        set_error("readData: cannot allocate %ld samples", n);
        return -1;
    }
    const size_t got = fread(buf + kPadding, sizeof(float), (size_t)n, f);
    const bool bad = got != (size_t)n && ferror(f);
    fclose(f);
    if (bad) {
        free(buf);
        set_error("readData: read error in %s", path);
        return -1;
    }
    *output_data = buf + kPadding;
    return n;
}

extern "C" void fdas_free_data(float* data)
{
    if (data) free(data - kPadding);
}

extern "C" int fdas_read_inf(const char* path, fdas_inf* out)
{
    if (!path || !out) {
        set_error("fdas_read_inf: null argument");
        return FDAS_ERR_INVALID;
    }
    memset(out, 0, sizeof(*out));
    FILE* f = fopen(path, "r");
    if (!f) {
        set_error("fdas_read_inf: cannot open %s", path);
        return FDAS_ERR_INVALID;
    }
    char line[1024];
    int have_n = 0, have_dt = 0;
    while (fgets(line, sizeof(line), f)) {
        const char* eq = strchr(line, '=');
        if (!eq) continue;
        const std::string key = trim(std::string(line, eq - line));
        const std::string val = trim(std::string(eq + 1));
        if (has(key, "Data file name without suffix")) {
            snprintf(out->basename, sizeof(out->basename), "%s", val.c_str());
        } else if (has(key, "Telescope used")) {
            snprintf(out->telescope, sizeof(out->telescope), "%s", val.c_str());
        } else if (has(key, "Number of bins in the time series")) {
            out->n = atol(val.c_str());
            have_n = 1;
        } else if (has(key, "Width of each time series bin")) {
            out->dt = atof(val.c_str());
            have_dt = 1;
        } else if (has(key, "Dispersion measure")) {
            out->dm = atof(val.c_str());
        } else if (has(key, "Epoch of observation")) {
            out->epoch_mjd = atof(val.c_str());
        } else if (has(key, "Barycentered?")) {
            out->barycentred = atoi(val.c_str());
        }
    }
    fclose(f);
    if (!have_n || !have_dt || out->n <= 0 || !(out->dt > 0.0)) {
        set_error("fdas_read_inf: %s lacks a positive sample count and sample time", path);
        return FDAS_ERR_INVALID;
    }
    return FDAS_OK;
}
