// Robustness driver for the image decoders (shiro_hip_load_image_memory),
// built by the CPU tests with AddressSanitizer and UndefinedBehaviorSanitizer
// on the host: reads a corpus of records (uint32 little-endian length, then
// the file's bytes) and decodes each one.  A decoder may accept or reject a
// record; any out-of-range access or undefined behaviour aborts the run.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "shiro_hip.h"

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: image_fuzz corpus\n");
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<uint8_t> all;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) all.insert(all.end(), buf, buf + n);
    std::fclose(f);
    size_t at = 0, records = 0, decoded = 0;
    while (at + 4 <= all.size()) {
        uint32_t len;
        std::memcpy(&len, all.data() + at, 4);
        at += 4;
        if (len > all.size() - at) return 3;
        // A copy of exactly the record's size, so reads past its end are caught.
        std::vector<uint8_t> rec(all.begin() + static_cast<std::ptrdiff_t>(at),
                                 all.begin() + static_cast<std::ptrdiff_t>(at + len));
        at += len;
        shiro_hip_image img;
        char err[256];
        const int rc = shiro_hip_load_image_memory(rec.data(), rec.size(), &img, err, sizeof(err));
        if (rc == SHIRO_HIP_OK) {
            ++decoded;
            shiro_hip_free_image(&img);
        }
        ++records;
    }
    std::printf("records %zu decoded %zu\n", records, decoded);
    return 0;
}
