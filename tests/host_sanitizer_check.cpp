// Host-side C-ABI calls of libcoolbubble run under AddressSanitizer and
// UndefinedBehaviorSanitizer (host code only: the device code is built as
// always).  No GPU is needed: every call here either is host logic
// (parameters, stability check, run-folder names, the HDF5 snapshot writer
// with its deflate threads) or must fail cleanly without a device.  Exit
// status 0 when every check holds; the sanitizers abort on their first
// report.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/coolbubble.h"

static int failures = 0;
#define CHECK(cond)                                                             \
    do {                                                                        \
        if (!(cond)) {                                                          \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
            ++failures;                                                         \
        }                                                                       \
    } while (0)

int main(int argc, char **argv)
{
    const char *dir = argc > 1 ? argv[1] : ".";
    cb_params p;
    CHECK(cb_default_params(&p) == CB_OK);
    CHECK(p.D == 2 && p.nGrids == 512 && p.readxns == 1);
    CHECK(cb_check_stability(&p) == CB_OK);
    cb_params bad = p;
    bad.dt = 1.0;
    CHECK(cb_check_stability(&bad) == CB_CFL_VIOLATION);
    bad = p;
    bad.D = 4;
    CHECK(cb_check_stability(&bad) == CB_INVALID_ARG);
    CHECK(std::strlen(cb_last_error()) > 0);
    bad = p;
    bad.nGrids = 1;
    CHECK(cb_check_stability(&bad) == CB_INVALID_ARG);
    CHECK(cb_check_stability(nullptr) == CB_INVALID_ARG);
    for (int st = 0; st <= 9; ++st) CHECK(cb_status_string((cb_status)st) != nullptr);

    // run folder names (src/main.cpp:243-247) for a few option sets
    cb_run_opts o;
    CHECK(cb_default_run_opts(&o) == CB_OK);
    char name[1024];
    CHECK(cb_run_folder_name(&o, name, sizeof name) == CB_OK);
    CHECK(std::strstr(name, "Sim2_D") == name);
    o.params.D = 3;
    o.params.expansion = 1;
    o.params.Mpl = 10.0;
    o.nt = 3000;
    CHECK(cb_run_folder_name(&o, name, sizeof name) == CB_OK);
    char tiny[8];
    CHECK(cb_run_folder_name(&o, tiny, sizeof tiny) != CB_OK);   // does not fit: refused, not truncated past the end
    CHECK(cb_run_folder_name(&o, nullptr, 0) != CB_OK);

    // without a device the handle-based calls fail cleanly
    cb_sim *s = nullptr;
    cb_params q = p;
    q.nGrids = 64;
    q.L = 8.0;
    q.dt = 0.025;
    cb_status r = cb_create(&q, nullptr, 0, &s);
    CHECK((r == CB_NO_DEVICE || r == CB_HIP_ERROR) && s == nullptr);
    CHECK(cb_step(nullptr, 1) == CB_INVALID_ARG);
    CHECK(cb_energies_row(nullptr, nullptr) == CB_INVALID_ARG);
    CHECK(cb_destroy(nullptr) == CB_OK);

    // the HDF5 snapshot writer: deflate threads, shuffle layout, attributes
    if (cb_hdf5_available() == CB_OK) {
        for (int D = 1; D <= 3; ++D) {
            const int N = D == 1 ? 3000 : (D == 2 ? 700 : 37);
            size_t n = 1;
            for (int d = 0; d < D; ++d) n *= (size_t)N;
            std::vector<std::vector<float>> f(5, std::vector<float>(n));
            for (int k = 0; k < 5; ++k)
                for (size_t i = 0; i < n; ++i) f[k][i] = (float)std::sin(0.001 * (double)(i * (k + 1)));
            const float *ptr[5] = {f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data()};
            std::string path = std::string(dir) + "/snap_" + std::to_string(D) + ".h5";
            CHECK(cb_write_snapshot_h5(path.c_str(), D, N, 16.0, 4.0, 0.5, 1.0, 5, ptr, 9) == CB_OK);
            CHECK(cb_write_snapshot_h5(path.c_str(), D, N, 16.0, 4.0, 0.5, 1.0, 2, ptr, 1) == CB_OK);
        }
        const float *none[2] = {nullptr, nullptr};
        CHECK(cb_write_snapshot_h5((std::string(dir) + "/x.h5").c_str(), 2, 8, 1.0, 4.0, 0.0, 1.0, 2, none, 9) ==
              CB_INVALID_ARG);
        CHECK(cb_write_snapshot_h5((std::string(dir) + "/x.h5").c_str(), 2, 8, 1.0, 4.0, 0.0, 1.0, 3, none, 9) ==
              CB_INVALID_ARG);
    }
    std::printf("host sanitizer check: %d failure(s)\n", failures);
    return failures ? 1 : 0;
}
