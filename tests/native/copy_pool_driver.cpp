// Lifecycle stress of gdp::CopyPool (csrc/gdp_copy_pool.h): the pool is
// started and stopped 200 times, as the staging of bulk copies does when
// problems are created and destroyed, with copies of varied sizes in
// between, each checked byte for byte.  Prints "ok" and exits 0; a wrong
// copy exits 1; a deadlock is caught by the caller's time limit.
#include <cstdio>
#include <cstring>
#include <vector>

#include "gdp_copy_pool.h"

int main() {
    gdp::CopyPool pool;
    std::vector<char> src(40u << 20), dst(40u << 20);
    for (size_t i = 0; i < src.size(); i++) src[i] = static_cast<char>(i * 131 + 7);
    for (int life = 0; life < 200; life++) {
        for (int c = 0; c < 3; c++) {
            const size_t len = (1u << 20) + static_cast<size_t>((life * 7 + c) % 9) * (3u << 20) + 12345;
            std::memset(dst.data(), 0, len);
            pool.copy(dst.data(), src.data(), len, 8);
            if (std::memcmp(dst.data(), src.data(), len) != 0) {
                std::printf("lifecycle %d copy %d: wrong bytes\n", life, c);
                return 1;
            }
        }
        pool.stop();
    }
    std::printf("ok\n");
    return 0;
}
