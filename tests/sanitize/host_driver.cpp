// host_driver.cpp — the host code of libcubo_hip built with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/sanitize/Makefile; host code only, no GPU: every call below returns before any device work). It drives the
// .npy header parser over well-formed and malformed files, the file checks of cubo_hip_cube_load_file, the probe
// size, the availability rule with injected memory sizes, the phase calls' refusals and the last-error buffer.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cubo_hip.h"

static std::string dir;

static std::string write_file(const char *name, const std::string &bytes) {
    const std::string p = dir + "/" + name;
    FILE *f = fopen(p.c_str(), "wb");
    fwrite(bytes.data(), 1, bytes.size(), f);
    fclose(f);
    return p;
}

// A .npy file: magic, version, header length (2 or 4 bytes), the dict padded with spaces to a multiple of 64, data.
static std::string npy(int major, const std::string &dict, size_t data_bytes, bool pad = true) {
    std::string h = dict;
    const size_t pre = 6 + 2 + (major == 1 ? 2 : 4);
    if (pad)
        while ((pre + h.size() + 1) % 64) h += ' ';
    h += '\n';
    std::string s = "\x93NUMPY";
    s += (char)major;
    s += (char)0;
    const uint32_t n = (uint32_t)h.size();
    if (major == 1) { s += (char)(n & 255); s += (char)(n >> 8); }
    else for (int i = 0; i < 4; ++i) s += (char)((n >> (8 * i)) & 255);
    return s + h + std::string(data_bytes, '\0');
}

int main(int argc, char **argv) {
    dir = argc > 1 ? argv[1] : ".";
    char msg[256];
    int bad = 0;
    auto expect = [&](bool ok, const char *what) {
        if (!ok) { fprintf(stderr, "unexpected result: %s\n", what); ++bad; }
    };
    cubo_hip_npy_header h;
    const char *good = "{'descr': '<u2', 'fortran_order': False, 'shape': (2, 3, 4, 8), }";
    for (int major = 1; major <= 3; ++major) {
        const std::string p = write_file("good.npy", npy(major, good, 2 * 3 * 4 * 8 * 2));
        expect(cubo_hip_npy_read_header(p.c_str(), &h) == CUBO_HIP_OK && h.rank == 4 && h.shape[3] == 8, "good");
    }
    const char *dicts[] = {
        "{'descr': '>f4', 'fortran_order': False, 'shape': (2, 2), }",
        "{'descr': '<f8', 'fortran_order': True, 'shape': (2, 2), }",
        "{'descr': '<c16', 'fortran_order': False, 'shape': (2,), }",
        "{'descr': '<i2', 'fortran_order': False, 'shape': (), }",
        "{'descr': '<f4', 'fortran_order': False, 'shape': (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, "
        "17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34), }",
        "{'descr': '<f4', 'fortran_order': False, 'shape': (99999999999999999999, 2), }",
        "{'descr': '<f4', 'fortran_order': False, 'shape': (-3, 2), }",
        "{'descr': '<f4', 'fortran_order': False, 'shape': (3, 2)",
        "{'descr': '<f4' 'fortran_order': False}",
        "{'descr': '<f4', 'fortran_order': False, 'shape': (3, 2), 'extra': 1}",
        "{'descr': ",
        "{",
        "",
        "{'descr': '<u2', 'fortran_order': Maybe, 'shape': (2,), }",
        "{'shape': (2, 2), 'descr': '<u1', 'fortran_order': False}",
    };
    for (const char *d : dicts) {
        const std::string p = write_file("bad.npy", npy(1, d, 64));
        cubo_hip_npy_read_header(p.c_str(), &h);
        const std::string q = write_file("bad2.npy", npy(2, d, 0, false));
        cubo_hip_npy_read_header(q.c_str(), &h);
    }
    for (size_t cut = 0; cut < 80; cut += 3) {                     // truncated files
        const std::string full = npy(1, good, 16);
        const std::string p = write_file("cut.npy", full.substr(0, cut));
        expect(cubo_hip_npy_read_header(p.c_str(), &h) != CUBO_HIP_OK || cut >= 64, "truncated");
    }
    expect(cubo_hip_npy_read_header((dir + "/missing.npy").c_str(), &h) == CUBO_HIP_ERR_FILE, "missing file");
    expect(cubo_hip_npy_read_header(nullptr, &h) != CUBO_HIP_OK, "NULL path");

    // file checks of the loader, before any device call
    const int64_t shape[4] = {2, 3, 4, 8};
    cubo_hip_cube *c = nullptr;
    const std::string raw = write_file("raw.bin", std::string(2 * 3 * 4 * 8 * 2 - 2, '\0'));
    expect(cubo_hip_cube_load_file(0, raw.c_str(), 1, CUBO_HIP_ELEMENT_U16, shape, &c) == CUBO_HIP_ERR_FILE, "misaligned");
    expect(cubo_hip_cube_load_file(0, raw.c_str(), 0, CUBO_HIP_ELEMENT_U16, shape, &c) == CUBO_HIP_ERR_FILE, "short");
    const int64_t bad_shape[4] = {2, -3, 4, 8};
    expect(cubo_hip_cube_load_file(0, raw.c_str(), 0, CUBO_HIP_ELEMENT_U16, bad_shape, &c) == CUBO_HIP_ERR_FILE, "shape");
    expect(cubo_hip_cube_load_file(0, raw.c_str(), 0, 99, shape, &c) == CUBO_HIP_ERR_INVALID_ARGUMENT, "type");

    // probe size on a disk, a flat pattern and a NaN pattern (host only)
    for (int kind = 0; kind < 3; ++kind) {
        const int rows = 21, cols = 17;
        std::vector<double> p(rows * cols);
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < cols; ++j)
                p[i * cols + j] = kind == 1 ? 5.0 : (std::hypot(i - 10.2, j - 8.1) < 5.5 ? 100.0 : 1.0);
        if (kind == 2) p[7] = NAN;
        cubo_hip_probe_estimate e;
        cubo_hip_probe_size(p.data(), rows, cols, 0.01, 0.99, 100, &e);
        cubo_hip_probe_size(p.data(), rows, cols, 0.01, 0.99, 1, &e);
    }

    // availability with injected memory sizes (the cap is checked before the device)
    int reason = 0;
    uint64_t bytes = 0, cap = 0;
    const int64_t big[4] = {256, 256, 128, 128};
    expect(cubo_hip_availability(0, big, CUBO_HIP_U16, 1ull << 30, &reason, &bytes, &cap) == CUBO_HIP_OK &&
               reason == CUBO_HIP_UNAVAILABLE_EXCEEDS_MEMORY_CAP, "availability cap");
    cubo_hip_availability(0, big, CUBO_HIP_F32, 288ull << 30, &reason, &bytes, &cap);

    // phase refusals before any device work
    std::vector<double> f(20, 1.0), o(20), ang = {0.0, NAN};
    expect(cubo_hip_phase_normalize(0, f.data(), f.data(), 4, 5, 1, 1.0, -1.0, nullptr, nullptr, o.data(), o.data(),
                                    o.data(), o.data()) == CUBO_HIP_ERR_INVALID_ARGUMENT, "tolerance");
    double deg = 0;
    int32_t t = 0, n0 = 0, n1 = 0;
    expect(cubo_hip_phase_rotation(0, f.data(), f.data(), 4, 5, ang.data(), 2, 0, nullptr, -1, nullptr, &deg, &t, &n0,
                                   &n1, o.data(), o.data(), o.data(), o.data()) == CUBO_HIP_ERR_INVALID_ARGUMENT,
           "angles");
    const volatile int32_t cancel = 1;
    const double forced = 3.0;
    expect(cubo_hip_phase_rotation(0, f.data(), f.data(), 4, 5, nullptr, 0, 0, &forced, 1, &cancel, &deg, &t, &n0, &n1,
                                   nullptr, nullptr, o.data(), o.data()) == CUBO_HIP_ERR_CANCELLED, "cancel");

    // NULL handles and the last-error buffer at every small size
    expect(cubo_hip_reduce_rows(nullptr, 0, 1) == CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL reduction");
    expect(cubo_hip_batch_run(nullptr, nullptr) == CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL batch");
    for (size_t n = 0; n < 8; ++n) cubo_hip_last_error(n ? msg : nullptr, n);
    cubo_hip_last_error(msg, sizeof msg);
    printf("host sanitize driver: done (%d unexpected)\n", bad);
    return bad ? 1 : 0;
}
