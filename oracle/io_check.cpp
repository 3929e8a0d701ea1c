// io_check.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// Drives the library's host-only file readers (des-alm_amd/csrc/io_files.hip,
// compiled with it under AddressSanitizer and UBSan by the Makefile's
// `sanitized` target) on one file, so that damaged inputs can be shown to be
// refused without any out-of-bounds access.
//
// usage: io_check grid|solu <path> <img> <jmg> <kmg> <turbulence>
//        io_check deck <path> <ng>
//        io_check filestat <path> <n_tur>
// exit: 0 read, 1 refused with a message, 2 usage; a sanitizer report ends
// the process with the sanitizer's own exit code.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/desalm.h"
#include "../des-alm_amd/csrc/host_error.h"

namespace desalm {
int host_fail(const std::string& msg) {
    std::fprintf(stderr, "refused: %s\n", msg.c_str());
    return -1;
}
}  // namespace desalm

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string what = argv[1];
    const char* path = argv[2];
    int rc = 0;
    if ((what == "grid" || what == "solu") && argc == 7) {
        const int32_t dims[3] = {std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5])};
        const int32_t turb = std::atoi(argv[6]);
        const long n = (long)dims[0] * dims[1] * dims[2];
        if (what == "grid") {
            std::vector<double> x(n), y(n), z(n), wd(n);
            rc = desalm_read_grid_file(path, 1, dims, turb, 0, x.data(), y.data(), z.data(), wd.data());
        } else {
            std::vector<double> q(5 * n), qn(5 * n), xnut(n);
            rc = desalm_read_solu_file(path, 1, dims, turb, 0, q.data(), qn.data(), xnut.data());
        }
    } else if (what == "deck" && argc == 4) {
        desalm_deck d;
        rc = desalm_read_deck(path, std::atoi(argv[3]), &d);
    } else if (what == "filestat" && argc == 4) {
        const int nt = std::atoi(argv[3]);
        std::vector<double> angle(nt > 0 ? nt : 1);
        double time0 = 0.0;
        int32_t nt1 = 0;
        rc = desalm_read_filestat(path, nt, &time0, &nt1, angle.data());
    } else {
        return 2;
    }
    return rc == 0 ? 0 : 1;
}
